"""Pins of the oracle (CPU).  The reference ships no tests or golden vectors
for its evaluator, so each oracle function is pinned by what the mathematics
fixes, chosen so a plausible mistake fails at least one check:

  oracle_eval_grid_f64 / oracle_eval_points_f64 (FP64 restatement of
  cpuEvaluator.cpp:294-407):
    - mpmath 40-digit term-by-term sums (tests/golden/mp_cases.npz): catches a
      wrong sign of i, conjugated k or E0, swapped axes, a dropped k component,
      the region rule at z_l - 2e-3, z_l - 5e-4, z_l (tol 0 and 1e-3), empty
      regions, and |E0| ~ 1e-300 times exp(+710) (cases A-E);
    - numpy.fft.ifft2 on a uniform lattice (a textbook routine): catches
      transposed operands on a non-square grid and a wrong sign convention;
    - linearity (superposition) and translation (wind) invariants;
    - closed forms of a single wave: constant |E| and a phase step of k.dr for
      real k, exp(-Im(kz) z) decay in an absorbing medium;
    - a long-double term-by-term sum on tiny inputs (16 waves, 5x4x3 points);
    - region membership on exact integer-valued fields.
  refcpu_eval_{xy,yz,xz,grid} (FP32 reference-semantics restatement):
    - agreement with the pinned FP64 oracle at the same float coordinates
      within FP32 rounding, row/column layout of each plane, and the region
      rule z >= z_l on both sides of a boundary.
"""
import numpy as np
import pytest

import oracle


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("case", ["A", "B", "C", "D"])
def test_grid_matches_mpmath(golden, case):
    g = golden[case]
    E = oracle.eval_grid_f64(g["W"], g["zl"], g["begin"], g["end"], g["xs"], g["ys"], g["zs"], float(g["tol"]))
    assert E.shape == g["E"].shape
    assert _rel(E, g["E"]) < 1e-12


def test_points_match_mpmath(golden):
    g = golden["E"]
    E = oracle.eval_points_f64(g["W"], g["zl"], g["begin"], g["end"], g["pts"], float(g["tol"]))
    assert _rel(E, g["E"]) < 1e-12


def test_exponent_range_is_finite(golden):
    g = golden["D"]
    E = oracle.eval_grid_f64(g["W"], g["zl"], g["begin"], g["end"], g["xs"], g["ys"], g["zs"])
    assert np.isfinite(E).all()
    # |E0| = 1e-300 * exp(+710) ~ 2.2e8 for the first wave: the result is large but finite.
    assert np.abs(E).max() > 1e8


def test_inverse_dft_on_uniform_lattice():
    rng = np.random.default_rng(1)
    Nx, Ny, dx, dy, z, kz = 8, 6, 0.7, 1.3, 1.1, 0.37
    A = rng.standard_normal((3, Nx, Ny)) + 1j * rng.standard_normal((3, Nx, Ny))
    m, n = np.meshgrid(np.arange(Nx), np.arange(Ny), indexing="ij")
    kx = 2 * np.pi * m / (Nx * dx)
    ky = 2 * np.pi * n / (Ny * dy)
    W = np.column_stack([A[0].ravel(), A[1].ravel(), A[2].ravel(), kx.ravel(), ky.ravel(),
                         np.full(Nx * Ny, kz)])
    xs, ys = np.arange(Nx) * dx, np.arange(Ny) * dy
    E = oracle.eval_grid_f64(W, np.zeros(0), [0], [W.shape[0]], xs, ys, [z])[0]  # (Ny, Nx, 3)
    for c in range(3):
        ref = Nx * Ny * np.fft.ifft2(A[c]).T * np.exp(1j * kz * z)  # ifft2 over (m -> i, n -> j), then [j, i]
        assert _rel(E[..., c], ref) < 1e-12


def _rand_waves(rng, n, kimag=0.1):
    E0 = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    k = rng.uniform(-8, 8, (n, 3)) + 1j * rng.uniform(-kimag, kimag, (n, 3))
    return np.concatenate([E0, k], 1)


def test_superposition():
    rng = np.random.default_rng(2)
    W1, W2 = _rand_waves(rng, 7), _rand_waves(rng, 5)
    xs, ys, zs = rng.uniform(-3, 3, 4), rng.uniform(-3, 3, 3), rng.uniform(-3, 3, 2)
    f = lambda W: oracle.eval_grid_f64(W, np.zeros(0), [0], [W.shape[0]], xs, ys, zs)
    assert _rel(f(np.concatenate([W1, W2])), f(W1) + f(W2)) < 1e-13


def test_translation():
    rng = np.random.default_rng(3)
    W = _rand_waves(rng, 9)
    t = np.array([0.37, -1.21, 0.58])
    xs, ys, zs = rng.uniform(-3, 3, 4), rng.uniform(-3, 3, 3), rng.uniform(-3, 3, 2)
    Wt = W.copy()
    Wt[:, :3] *= np.exp(1j * (W[:, 3:] @ t))[:, None]
    f = lambda W, a, b, c: oracle.eval_grid_f64(W, np.zeros(0), [0], [W.shape[0]], a, b, c)
    assert _rel(f(W, xs + t[0], ys + t[1], zs + t[2]), f(Wt, xs, ys, zs)) < 1e-12


def test_single_wave_closed_form():
    """One wave gives E0 exp(i k.r) exactly: for real k, |E_c| = |E0_c| at every
    point and the phase advances by k.dr between points; for an absorbing
    medium (Im kz > 0), |E_c(z)| = |E0_c| exp(-Im(kz) z)."""
    rng = np.random.default_rng(5)
    xs, ys, zs = np.linspace(-4, 4, 7), np.linspace(-3, 5, 5), np.linspace(-2, 3, 6)
    E0 = rng.standard_normal(3) + 1j * rng.standard_normal(3)
    k = np.array([1.3, -0.7, 2.1])
    W = np.concatenate([E0, k]).astype(np.complex128)[None, :]
    E = oracle.eval_grid_f64(W, np.zeros(0), [0], [1], xs, ys, zs)  # [z][y][x][3]
    assert np.allclose(np.abs(E), np.abs(E0)[None, None, None, :], rtol=1e-13, atol=0)
    ratio = E[:, :, 1:, :] / E[:, :, :-1, :]
    assert np.allclose(ratio, np.exp(1j * k[0] * (xs[1] - xs[0])), rtol=1e-12, atol=0)
    kz = 2.1 + 0.35j
    W[0, 5] = kz
    E = oracle.eval_grid_f64(W, np.zeros(0), [0], [1], xs, ys, zs)
    amp = np.abs(E0)[None, :] * np.exp(-kz.imag * zs)[:, None]
    assert np.allclose(np.abs(E[:, 2, 3, :]), amp, rtol=1e-13, atol=0)


def test_brute_force_long_double():
    """Tiny inputs (W <= 16, N <= 8) against a long-double sum written out
    term by term (numpy longdouble), independent of the oracle's code path."""
    rng = np.random.default_rng(6)
    W = _rand_waves(rng, 16)
    xs, ys, zs = rng.uniform(-2, 2, 5), rng.uniform(-2, 2, 4), rng.uniform(-2, 2, 3)
    E = oracle.eval_grid_f64(W, np.zeros(0), [0], [W.shape[0]], xs, ys, zs)
    ld = np.longdouble
    for iz, z in enumerate(zs):
        for iy, y in enumerate(ys):
            for ix, x in enumerate(xs):
                acc = np.zeros((3, 2), ld)
                for w in W:
                    kr = (ld(w[3].real) * ld(x) + ld(w[4].real) * ld(y) + ld(w[5].real) * ld(z))
                    ki = (ld(w[3].imag) * ld(x) + ld(w[4].imag) * ld(y) + ld(w[5].imag) * ld(z))
                    mag = np.exp(-ki)
                    c, s = np.cos(kr) * mag, np.sin(kr) * mag
                    for comp in range(3):
                        er, ei = ld(w[comp].real), ld(w[comp].imag)
                        acc[comp, 0] += er * c - ei * s
                        acc[comp, 1] += er * s + ei * c
                ref = acc[:, 0].astype(np.float64) + 1j * acc[:, 1].astype(np.float64)
                assert np.allclose(E[iz, iy, ix], ref, rtol=1e-12, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("tol", [0.0, 1e-3])
def test_region_membership(tol):
    zl = np.array([-1.0, 0.5, 2.0])
    # region r holds one wave with k = 0 and E0 = (r + 1, 0, 0): the field names its region
    W = np.zeros((4, 6), np.complex128)
    W[:, 0] = np.arange(1, 5)
    begin, end = np.arange(4, dtype=np.int32), np.arange(1, 5, dtype=np.int32)
    zs = np.array([-5, -1.002, -1.0005, -1.0, 0.4995, 0.5, 1.9995, 2.0, 9.0])
    E = oracle.eval_grid_f64(W, zl, begin, end, [0.0], [0.0], zs, tol)[:, 0, 0, 0].real
    expect = [1 + sum(z >= b - tol for b in zl) for z in zs]
    assert list(E.astype(int)) == expect
    assert [oracle.region_of(z, zl, tol) for z in zs] == [e - 1 for e in expect]


def _c1_like(seed=4):
    from scwatter_amd import configs
    cw = configs.layered_structure([1.0, 1.4 + 0.05j], [0.0], 1.0, 0.5, 7, ("x",))
    return cw.unpack()


@pytest.mark.parametrize("kind", ["xy", "yz", "xz"])
def test_refcpu_planes_match_oracle(kind):
    t = _c1_like()
    N, d, a0, b0, fixed = 12, 0.35, -2.0, -1.9, 0.15
    E = oracle.refcpu_plane(kind, t.W, t.z_layers, t.begin, t.end, a0, b0, fixed, d, N)
    f32 = lambda v: np.float32(v)
    a = (f32(a0) + np.arange(N, dtype=np.float32) * f32(d)).astype(np.float64)
    b = (f32(b0) + np.arange(N, dtype=np.float32) * f32(d)).astype(np.float64)
    fx = [float(f32(fixed))]
    # the reference narrows z_layers to float before comparing
    zl = t.z_layers.astype(np.float32).astype(np.float64)
    grid = {"xy": (a, b, fx), "yz": (fx, a, b), "xz": (a, fx, b)}[kind]
    ref = oracle.eval_grid_f64(t.W, zl, t.begin, t.end, *grid).reshape(N, N, 3)
    assert oracle.rel_l2(E, ref) < 2e-6


def test_refcpu_region_rule_at_boundary():
    zl = np.array([0.25])
    W = np.zeros((2, 6), np.complex128)
    W[:, 0] = [1.0, 2.0]
    for z, expect in ((0.25, 2.0), (np.nextafter(np.float32(0.25), np.float32(0)), 1.0), (0.2495, 1.0)):
        E = oracle.refcpu_plane("xy", W, zl, [0, 1], [1, 2], 0.0, 0.0, z, 1.0, 2)
        assert np.all(E[..., 0].real == expect)


def test_refcpu_grid_matches_planes():
    t = _c1_like()
    xs = (np.float32(-1.5) + np.arange(9, dtype=np.float32) * np.float32(0.4))
    ys = (np.float32(-0.7) + np.arange(9, dtype=np.float32) * np.float32(0.4))
    E = oracle.refcpu_grid(t.W, t.z_layers, t.begin, t.end, xs, ys, [0.3], nthreads=2)[0]
    P = oracle.refcpu_plane("xy", t.W, t.z_layers, t.begin, t.end, -1.5, -0.7, 0.3, 0.4, 9)
    assert np.array_equal(E, P)
