"""scw_eval_planes (the reference's gpu_cw_evaluate, gpuEvaluator.cu:131-165, at
the C ABI, with its per-plan exec cache) against execs built for the same
planes: bit-identical on first calls, on repeated calls (cache hits) and while
one plane moves through more positions than the cache holds (its execs are
then rebuilt in place), for the GEMM and the direct form; and the planes
against the FP64 oracle under the GPU region rule."""
import numpy as np
import pytest

import oracle
from scwatter_amd import Exec, Plan, axis_gpu, configs, eval_planes, hip

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def stack():
    rng = np.random.default_rng(11)
    n = np.concatenate([[1.0], rng.uniform(1, 2, 5) + 1j * rng.uniform(0, 0.05, 5)])
    z = np.concatenate([[0.0], np.cumsum(rng.uniform(0.1, 2.0, 4))]) - 2.0
    t = configs.layered_structure(n, z, 0.6, 0.7, 11, ("x",)).unpack()
    return Plan(t), t


def _planes(plan, xs0, ys0, zs0, x, y, z, d, N, form):
    bufs = [hip.DeviceArray((N, N, 3), np.complex64) for _ in range(3)]
    eval_planes(plan, xs0, ys0, zs0, x, y, z, d, N, *(b.ptr for b in bufs), form=form)
    hip.synchronize()
    return [b.to_host() for b in bufs]


def _execs(plan, xs0, ys0, zs0, x, y, z, d, N, form):
    ax, ay, az = axis_gpu(xs0, d, N), axis_gpu(ys0, d, N), axis_gpu(zs0, d, N)
    f = lambda v: [float(np.float32(v))]
    out = []
    for g in ((ax, ay, f(z)), (ax, f(y), az), (f(x), ay, az)):
        out.append(Exec(plan, *g, 0.001, form).run_host().reshape(N, N, 3))
    return out


@pytest.mark.parametrize("form", ["auto", "direct"])
def test_eval_planes_matches_execs_and_cache(stack, form):
    plan, t = stack
    N, d = 40, np.float32(0.21)
    xs0, ys0, zs0 = np.float32(-4.1), np.float32(-3.9), np.float32(-4.2)
    x, y, z = 0.3, -0.2, float(t.z_layers[1]) - 0.0005  # z within 1e-3 below a boundary
    first = _planes(plan, xs0, ys0, zs0, x, y, z, d, N, form)
    ref = _execs(plan, xs0, ys0, zs0, x, y, z, d, N, form)
    for a, b, name in zip(first, ref, ("xy", "xz", "yz")):
        assert np.array_equal(a, b), f"{form} {name}: scw_eval_planes differs from its exec"
    again = _planes(plan, xs0, ys0, zs0, x, y, z, d, N, form)
    for a, b in zip(first, again):
        assert np.array_equal(a, b)
    # the xy plane moves through 8 positions: past the cache's six entries
    for k, zk in enumerate(np.linspace(-3.5, 3.0, 8)):
        got = _planes(plan, xs0, ys0, zs0, x, y, float(zk), d, N, form)
        want = _execs(plan, xs0, ys0, zs0, x, y, float(zk), d, N, form)
        for a, b, name in zip(got, want, ("xy", "xz", "yz")):
            assert np.array_equal(a, b), f"{form} move {k} {name}"
    # and against the oracle, GPU region rule on the float boundaries
    ax, ay, az = axis_gpu(xs0, d, N), axis_gpu(ys0, d, N), axis_gpu(zs0, d, N)
    zl = t.z_layers
    f = lambda v: [float(np.float32(v))]
    for E, g, name in ((first[0], (ax, ay, f(z)), "xy"), (first[1], (ax, f(y), az), "xz"),
                       (first[2], (f(x), ay, az), "yz")):
        R = oracle.eval_grid_f64(t.W, zl, t.begin, t.end, *g, tol=0.001).reshape(N, N, 3)
        err = oracle.rel_l2(E, R)
        print(f"eval_planes {form} {name}: rel-L2 {err:.3e}", flush=True)
        assert err <= (1e-5 if form == "direct" else 1e-5)


def test_eval_planes_skips_null_outputs(stack):
    plan, _ = stack
    N = 16
    b = hip.DeviceArray((N, N, 3), np.complex64)
    eval_planes(plan, -1, -1, -1, 0.1, 0.2, 0.3, 0.1, N, 0, b.ptr, 0)
    hip.synchronize()
    ref = _execs(plan, -1, -1, -1, 0.1, 0.2, 0.3, 0.1, N, "auto")[1]
    assert np.array_equal(b.to_host(), ref)
