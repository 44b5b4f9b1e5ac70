"""Pins of the host sample solver (scwatter_amd/sample_solver.py, sCWatter's
Phi / eigen-solve / boundary pipeline), on the CPU:

* a homogeneous sample reduces every order to a textbook slab: the field of
  its structure equals that of the independent layered-medium solver
  (configs.solve_stack) for the same incident plane wave, at oblique
  incidence and in two sublayers;
* a lossless heterogeneous sample conserves the Poynting flux of the
  propagating orders (reflected + transmitted = incident), which a wrong Phi
  or boundary system breaks;
* tangential E and H are continuous at every boundary, evaluated by the FP64
  oracle at the boundary itself from both sides;
* an absorbing sample loses flux;
* the structure has the wave layout of sCWatter.cpp:890-956 (region sizes
  2 MF, 4 MF^2 per sublayer, MF), each internal region's waves share one kz
  per mode, and the mode order interleaves decaying (even) and growing (odd)
  modes."""
import math

import numpy as np
import pytest

import oracle
from scwatter_amd import configs
from scwatter_amd import sample_solver as S
from scwatter_amd.cw import CoupledWaveStructure, HomogeneousLayer


def _field(t, pts, tol=0.0):
    return oracle.eval_points_f64(t.W, t.z_layers, t.begin, t.end, pts, tol=tol)


def _flux(sol):
    prop0 = np.abs(sol["Sz0"].imag) < 1e-12
    prop1 = np.abs(sol["Sz1"].imag) < 1e-12
    inc = sol["Sz0"][sol["r0"]].real * np.sum(np.abs(sol["E_inc"]) ** 2)
    r = np.sum(sol["Sz0"].real[prop0] * np.sum(np.abs(sol["R"][prop0]) ** 2, 1))
    t = np.sum(sol["Sz1"].real[prop1] * np.sum(np.abs(sol["T"][prop1]) ** 2, 1))
    return r / inc, t / inc


def test_homogeneous_slab_matches_layered_solver():
    ns = 1.5 + 0.02j
    d = np.array([0.3, 0.1, 1.0])
    cw = S.solve_sample(np.full((2, 8, 8), ns), 0.7, (6.0, 6.0, 1.2), 0.1, (3, 3), d, (1.0, 0.5, 0.0), workers=1)
    sol = cw.solution
    k, z = sol["k"], sol["z"]
    d = d / np.linalg.norm(d)
    kx, ky = np.array([d[0] * k]), np.array([d[1] * k])
    kz, D, U = configs.solve_stack(kx, ky, k, np.array([1.0, ns, ns, 1.0]), z, sol["E_inc"][None, :])
    pack = lambda E0, kzv: np.column_stack([E0, kx, ky, kzv])
    layers = [HomogeneousLayer(z[l], pack(U[:, l] * np.exp(1j * kz[:, l] * z[l])[:, None], -kz[:, l]),
                               pack(D[:, l + 1] * np.exp(-1j * kz[:, l + 1] * z[l])[:, None], kz[:, l + 1]))
              for l in range(3)]
    ref = CoupledWaveStructure(pack(D[:, 0], kz[:, 0]), layers).unpack()
    rng = np.random.default_rng(0)
    pts = np.stack([rng.uniform(-3, 3, 60), rng.uniform(-3, 3, 60), rng.uniform(z[0] - 1, z[-1] + 1, 60)], 1)
    E, R = _field(cw.unpack(), pts), _field(ref, pts)
    assert oracle.rel_l2(E, R) < 1e-12


@pytest.fixture(scope="module")
def lossless():
    n = S.tile_sample(3, 16, 16, tile=4, n_background=1.0, n_tile=1.5, seed=3)
    return S.solve_sample(n, 0.8, (5.0, 5.0, 1.5), 0.0, (5, 5), (0.2, 0.0, 1.0), (0.0, 1.0, 0.0), workers=2)


def test_lossless_sample_conserves_flux(lossless):
    r, t = _flux(lossless.solution)
    assert 0.01 < r < 0.5 and abs(r + t - 1.0) < 1e-10


def test_tangential_fields_continuous_at_boundaries(lossless):
    """E and H = (k / k0) x E (the waves' own H, summed by the oracle from a
    table of k x E0 / k0) at each boundary z_l, once as the region below
    (tol < 0) and once as the region above (tol = 0)."""
    t = lossless.unpack()
    k0 = lossless.solution["k"]
    WH = t.W.copy()
    WH[:, :3] = np.cross(t.W[:, 3:], t.W[:, :3]) / k0
    from scwatter_amd.cw import WaveTables
    tH = WaveTables(WH, t.z_layers, t.begin, t.end)
    rng = np.random.default_rng(1)
    xy = np.stack([rng.uniform(-2.5, 2.5, 24), rng.uniform(-2.5, 2.5, 24)], 1)
    for zl in t.z_layers:
        pts = np.column_stack([xy, np.full(24, zl)])
        for tab in (t, tH):
            above, below = _field(tab, pts, 0.0), _field(tab, pts, -1e-9)
            assert oracle.rel_l2(below[:, :2], above[:, :2]) < 1e-10, zl


def test_absorbing_sample_loses_flux():
    n = S.tile_sample(2, 16, 16, tile=4, n_tile=1.4 + 0.05j, seed=9)
    cw = S.solve_sample(n, 0.9, (4.0, 4.0, 2.0), 0.0, (3, 3), (0.0, 0.0, 1.0), (1.0, 0.0, 0.0), workers=1)
    r, t = _flux(cw.solution)
    assert 0.0 < r and 0.0 < t and r + t < 0.99


def test_wave_layout_and_mode_order(lossless):
    t = lossless.unpack()
    MF = 25
    sizes = np.asarray(t.end) - np.asarray(t.begin)
    assert sizes.tolist() == [2 * MF] + [4 * MF * MF] * 3 + [MF]
    for r in (1, 2, 3):
        # Pt of the boundary above (down modes), then Pr of the one below (up
        # modes), each per order its 2 MF modes; a mode's kz is the same for
        # every order
        kz = t.W[t.begin[r]:t.end[r], 5].reshape(2, MF, 2 * MF)
        assert np.array_equal(kz, np.broadcast_to(kz[:, :1], kz.shape))
        assert np.all(kz[0].imag >= kz[1].imag)
    P, _ = S.phi_matrix(S.tile_sample(1, 8, 8, tile=4, n_tile=1.3 + 0.02j, seed=2)[0], 3, 3,
                        np.linspace(-1, 1, 9), np.linspace(-0.5, 0.7, 9), 2 * math.pi)
    g, V = S.sorted_modes(P)
    assert np.all(g[0::2].imag >= g[1::2].imag)
    assert np.allclose(P @ V, V * g[None, :], atol=1e-9)
