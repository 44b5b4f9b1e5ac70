"""scw_eval_planes_noalloc (the allocation-free variant of scw_eval_planes'
direct form) against scw_eval_planes itself, through the C ABI, on two calls in
a row (the second with a smaller N): bit-identical planes, which also match
the FP64 oracle.  The structure's internal region has 9,604 waves, so small
planes split their wave range and the partial sums live in the fixed device
array."""
import numpy as np
import pytest

import oracle
from scwatter_amd import Plan, ScwError, _native, axis_gpu, configs, eval_planes, hip
from scwatter_amd._native import check

pytestmark = pytest.mark.gpu


def _noalloc(plan, xs0, ys0, zs0, x, y, z, d, N, outs, tol=0.001):
    """The entry point through the C ABI; returns its status."""
    g = lambda v: float(np.float32(v))
    return _native.lib().scw_eval_planes_noalloc(plan.handle, g(xs0), g(ys0), g(zs0), g(x), g(y), g(z), g(d), int(N),
                                                 float(tol), *outs, None)


@pytest.fixture(scope="module")
def slab():
    t = configs.tile_sample_structure(M=7, Z=2, size=(12.0, 12.0, 2.0), lam=0.8, seed=5, incidence_deg=8.0).unpack()
    assert int(t.end[2] - t.begin[2]) == 4 * 49 * 49
    return Plan(t), t


@pytest.mark.parametrize("calls", [((48, 0.23, -5.4, -0.2), (20, 0.31, -3.0, 0.4))])
def test_noalloc_planes_match_eval_planes(slab, calls):
    plan, t = slab
    for N, d, s0, zfix in calls:
        print(f"noalloc planes N={N} d={d} start={s0} z={zfix}", flush=True)
        a = [hip.DeviceArray((N, N, 3), np.complex64) for _ in range(3)]
        b = [hip.DeviceArray((N, N, 3), np.complex64) for _ in range(3)]
        check(_noalloc(plan, s0, s0 + 0.1, s0 - 0.3, 0.15, -0.25, zfix, d, N, [x.ptr for x in a]))
        eval_planes(plan, s0, s0 + 0.1, s0 - 0.3, 0.15, -0.25, zfix, d, N, *(x.ptr for x in b), form="direct")
        hip.synchronize()
        ax, ay, az = axis_gpu(s0, d, N), axis_gpu(s0 + 0.1, d, N), axis_gpu(s0 - 0.3, d, N)
        f = lambda v: [float(np.float32(v))]
        for k, (g, name) in enumerate((((ax, ay, f(zfix)), "xy"), ((ax, f(-0.25), az), "xz"),
                                       ((f(0.15), ay, az), "yz"))):
            A, B = a[k].to_host(), b[k].to_host()
            assert np.array_equal(A, B), f"N={N} {name}: allocation-free plane differs from scw_eval_planes"
            R = oracle.eval_grid_f64(t.W, t.z_layers, t.begin, t.end, *g, tol=0.001).reshape(N, N, 3)
            err = oracle.rel_l2(A, R)
            print(f"  {name}: rel-L2 vs oracle {err:.3e}", flush=True)
            assert err <= 1e-5


def test_noalloc_rejects_sizes_past_its_array(slab):
    plan, _ = slab
    st = _noalloc(plan, 0, 0, 0, 0, 0, 0, 0.1, 16385, [None, None, None])
    assert st == -1 and b"16384" in _native.lib().scw_last_error()
    with pytest.raises(ScwError, match="ARG"):
        check(_noalloc(plan, 0, 0, 0, 0, 0, 0, 0.1, 0, [None, None, None]))
