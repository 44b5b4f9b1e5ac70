"""GPU parity of the on-device post-processing (scw_post_*, through the C
ABI) with its restatement oracle/post.py (sCWatterView.cpp:194-352): bit
for bit on random fields with NaN, infinities and signed zeros, and on an
evaluated field."""
import numpy as np
import pytest

import oracle
from oracle import post as opost
from scwatter_amd import Exec, Plan, configs, hip, post

pytestmark = pytest.mark.gpu


def _field(n, seed, nan0=0):
    rng = np.random.default_rng(seed)
    E = (rng.standard_normal((n, 3)) * np.exp(rng.uniform(-20, 20, (n, 3)))
         + 1j * rng.standard_normal((n, 3))).astype(np.complex64)
    E.real[nan0::97, 0] = np.nan
    E.imag[::89, 1] = np.inf
    E.real[::83, 2] = -0.0
    E[min(5, n - 1)] = 0
    return E


@pytest.mark.parametrize("nan0", [0, 1])
@pytest.mark.parametrize("n", [1, 1000, 100003])
def test_scalar_minmax_color_bit_exact(n, nan0):
    """nan0 = 0: the x component's real part is NaN at index 0 (the
    reference's seed), so its range is NaN for the whole call sequence; nan0 =
    1: the NaN entries start at index 1 and are skipped."""
    s = hip.Stream()
    slices = [_field(n, 7 + k, nan0) for k in range(3)]
    dS = [hip.DeviceArray((n,), np.complex64) for _ in range(3)]
    mm = hip.DeviceArray((4,), np.float32)
    rgb = hip.DeviceArray((n, 3), np.uint8)
    for mode in range(4):
        refs = [opost.scalar(E, mode) for E in slices]
        for k, E in enumerate(slices):
            dE = hip.DeviceArray.from_host(E)
            post.scalar(dE.ptr, n, mode, dS[k].ptr, s.handle)
            post.minmax(dS[k].ptr, n, mm.ptr, accumulate=k > 0, stream=s.handle)
            s.synchronize()
            np.testing.assert_array_equal(dS[k].to_host(), refs[k])
        rmm = opost.minmax(refs)
        np.testing.assert_array_equal(mm.to_host(), rmm)
        for args in (dict(), dict(show_real=False), dict(real_color=(0.3, 0.9, 0.1), imag_color=(0.8, 0.2, 0.7))):
            post.color(dS[1].ptr, n, mm.ptr, rgb.ptr, stream=s.handle, **args)
            s.synchronize()
            assert np.array_equal(rgb.to_host(), opost.color(refs[1], rmm, **args)), (mode, args)
    s.destroy()


def test_pipeline_on_an_evaluated_slice():
    cw, (xs, ys, zs) = configs.config1()
    plan = Plan(cw.unpack())
    ex = Exec(plan, xs, ys, zs[20:21], form="gemm")
    n = xs.size * ys.size
    dE = hip.DeviceArray((n, 3), np.complex64)
    dS = hip.DeviceArray((n,), np.complex64)
    mm = hip.DeviceArray((4,), np.float32)
    rgb = hip.DeviceArray((n, 3), np.uint8)
    ex.run(dE.ptr)
    post.scalar(dE.ptr, n, "intensity", dS.ptr)
    post.minmax(dS.ptr, n, mm.ptr)
    post.color(dS.ptr, n, mm.ptr, rgb.ptr)
    hip.synchronize()
    E = dE.to_host()
    S = opost.scalar(E, 3)
    rmm = opost.minmax([S])
    assert np.array_equal(dS.to_host(), S) and np.array_equal(mm.to_host(), rmm)
    assert np.array_equal(rgb.to_host(), opost.color(S, rmm))
    assert rmm[0] >= 0 and rmm[2] == 0 and rmm[3] == 0  # intensity is real and non-negative
    assert rgb.to_host()[:, 0].max() == 0 and rgb.to_host()[:, 1].max() == 255
