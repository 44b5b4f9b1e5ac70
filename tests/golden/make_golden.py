"""Write tests/golden/mp_cases.npz: known-answer vectors for the field sum.

The reference ships no tests, fixtures or golden vectors for its evaluator, so
these vectors come from the mathematics itself: E(r) = sum_w E0_w exp(i k_w . r)
over the waves of the region that holds r (cpuEvaluator.cpp:294-334,
gpuEvaluator.cu:28-68), summed term by term with mpmath at 40 significant
digits.  Inputs are float64 values, which mpmath reads exactly; the region is
chosen with the same double-precision comparison z >= z_l - tol that
cpuEvaluator.cpp:303 (tol = 0) and gpuEvaluator.cu:34 (tol = 0.001) use.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 40
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mp_cases.npz")


def region(z, zl, tol):
    r = 0
    for l, zb in enumerate(zl):
        if z >= float(zb) - tol:
            r = l + 1
    return r


def field(W, r):
    E = [mp.mpc(0), mp.mpc(0), mp.mpc(0)]
    for w in W:
        kr = mp.mpc(0)
        for j in range(3):
            kr += mp.mpc(w[3 + j].real, w[3 + j].imag) * mp.mpf(float(r[j]))
        p = mp.exp(mp.mpc(0, 1) * kr)
        for c in range(3):
            E[c] += mp.mpc(w[c].real, w[c].imag) * p
    return [complex(e) for e in E]


def grid_case(W, zl, begin, end, xs, ys, zs, tol):
    out = np.zeros((len(zs), len(ys), len(xs), 3), np.complex128)
    for iz, z in enumerate(zs):
        r = region(z, zl, tol)
        Wr = W[begin[r]:end[r]]
        for iy, y in enumerate(ys):
            for ix, x in enumerate(xs):
                out[iz, iy, ix] = field(Wr, (x, y, z))
    return out


def points_case(W, zl, begin, end, pts, tol):
    out = np.zeros((len(pts), 3), np.complex128)
    for i, p in enumerate(pts):
        r = region(p[2], zl, tol)
        out[i] = field(W[begin[r]:end[r]], p)
    return out


def rand_waves(rng, n, kmag, kimag):
    E0 = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    k = rng.uniform(-kmag, kmag, (n, 3)) + 1j * rng.uniform(-kimag, kimag, (n, 3))
    return np.concatenate([E0, k], axis=1)


def main():
    rng = np.random.default_rng(20261203)
    cases = {}

    # A/B: three regions (two boundaries), distinct waves per region, grid straddling boundaries.
    W = rand_waves(rng, 14, 6.0, 0.05)
    zl = np.array([-0.5, 0.75])
    begin = np.array([0, 5, 9], np.int32)
    end = np.array([5, 9, 14], np.int32)
    xs = np.array([-1.25, 0.3, 2.0])
    ys = np.array([0.7, -1.9])
    zs = np.array([-2.0, -0.5 - 2e-3, -0.5 - 5e-4, -0.5, 0.1, 0.75 - 5e-4, 0.75, 3.3])
    for name, tol in (("A", 0.0), ("B", 1e-3)):
        cases[name] = dict(W=W, zl=zl, begin=begin, end=end, xs=xs, ys=ys, zs=zs, tol=np.float64(tol),
                           E=grid_case(W, zl, begin, end, xs, ys, zs, tol))

    # C: strongly absorbing / evanescent waves, one boundary, non-square grid.
    W = rand_waves(rng, 9, 9.0, 1.5)
    zl = np.array([0.0])
    begin = np.array([0, 4], np.int32)
    end = np.array([4, 9], np.int32)
    xs = np.linspace(-1.0, 1.0, 4)
    ys = np.linspace(-0.5, 0.5, 3)
    zs = np.array([-1.2, -0.3, 0.0, 0.4, 1.1])
    cases["C"] = dict(W=W, zl=zl, begin=begin, end=end, xs=xs, ys=ys, zs=zs, tol=np.float64(0.0),
                      E=grid_case(W, zl, begin, end, xs, ys, zs, 0.0))

    # D: exponent range — |E0| ~ 1e-300 times exp(+710): each factor alone under/overflows FP64.
    W = np.array([[1e-300, 0.5e-300j, -2e-300 + 1e-300j, 0.3, -0.2, 0.1 - 71.0j],
                  [1.0, 1j, 0.0, 0.0, 0.5, 1.0]], np.complex128)
    zl = np.zeros(0)
    begin = np.array([0], np.int32)
    end = np.array([2], np.int32)
    xs = np.array([0.0, 0.25])
    ys = np.array([0.0])
    zs = np.array([9.95, 10.0])
    cases["D"] = dict(W=W, zl=zl, begin=begin, end=end, xs=xs, ys=ys, zs=zs, tol=np.float64(0.0),
                      E=grid_case(W, zl, begin, end, xs, ys, zs, 0.0))

    # E: scattered points, four regions with an empty one.
    W = rand_waves(rng, 12, 12.0, 0.2)
    zl = np.array([-1.0, 0.0, 2.0])
    begin = np.array([0, 4, 4, 8], np.int32)
    end = np.array([4, 4, 8, 12], np.int32)
    pts = np.column_stack([rng.uniform(-5, 5, 24), rng.uniform(-5, 5, 24), rng.uniform(-3, 4, 24)])
    pts[0, 2] = -1.0
    pts[1, 2] = 0.0
    pts[2, 2] = 2.0
    cases["E"] = dict(W=W, zl=zl, begin=begin, end=end, pts=pts, tol=np.float64(0.0),
                      E=points_case(W, zl, begin, end, pts, 0.0))

    flat = {f"{c}_{k}": v for c, d in cases.items() for k, v in d.items()}
    np.savez_compressed(OUT, **flat)
    print("wrote", OUT, sorted(cases))


if __name__ == "__main__":
    main()
