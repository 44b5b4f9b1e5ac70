"""FP64 generators of the benchmark structures and grids (seeded).

These produce the same kind of wave sets the reference's producers write to
.cw files (sCWatterPlane.cpp:130-270, sCWatterLayer.cpp:133-429,
sCWatter.cpp:890-978) — the evaluator only ever sees (E0, k) per wave and the
region tables.  Conventions:

* Focused beams use a square Cartesian pupil lattice: kx, ky on an M x M
  lattice spanning [-NA k0, NA k0] (k0 = 2 pi / lambda), uniform weights
  1/M^2, polarisation made transverse to k as FourierWave.h:32-39 does.
* Layered media are solved per incident wave from the boundary conditions
  (tangential E and tangential k x E continuous, k . E = 0 per wave), the
  same conditions sCWatterLayer.cpp:143-233 imposes, in a 6(R-1) system per
  wave with each amplitude referenced at its own boundary.
* Every wave is stored as the producers store it, "wound" to the origin
  (planewave::wind(0, 0, -z_l) in sCWatter.cpp:907-948 and
  sCWatterLayer.cpp:409-412): E0 = amplitude at z_l * exp(-i kz z_l).
* The tile-sample structure (sCWatter's heterogeneous slab stack) uses the
  exact order lattice of sCWatter.cpp:893-899 and the exact wave layout of
  its .cw assembly (sCWatter.cpp:902-956); the eigenvalues and eigenvector
  amplitudes are seeded synthetic values, since the eigen-solve that makes
  them is host work outside the evaluator.
"""
import math

import numpy as np

from .cw import CoupledWaveStructure, HomogeneousLayer

BASE_SEED = 20261203


def kz_of(n, k0, kx, ky):
    """Principal branch: Im kz >= 0 (decays along +z), FourierWave.h:332-333."""
    a = (np.asarray(n, np.complex128) * k0) ** 2 - kx ** 2 - ky ** 2
    a = np.asarray(a, np.complex128)
    a = a.real + 1j * np.where(a.imag == 0, 0.0, a.imag)  # keep +0 imaginary parts
    return np.sqrt(a)


def orthogonalize(e, d):
    """FourierWave.h:32-39: E /= sqrt(E.E); E = d x (E x d)."""
    e = np.asarray(e, np.complex128)
    d = np.asarray(d, np.complex128)
    mag = np.sqrt(np.sum(e * e, axis=-1, keepdims=True))
    e = e / mag
    s = np.cross(e, d)
    return np.cross(d, s)


def pupil_lattice(M, NA, k0):
    """kx, ky (each M*M, order index p = b*M + a) on the square lattice."""
    u = np.linspace(-1.0, 1.0, M) * NA * k0 if M > 1 else np.zeros(1)
    kx, ky = np.meshgrid(u, u)  # rows: ky index b, columns: kx index a
    return kx.ravel(), ky.ravel()


def solve_stack(kx, ky, k0, n, z, Einc):
    """Plane waves of a layered medium for each incident wave.

    kx, ky: (W,) real transverse wavenumbers; n: (R,) complex indices of the R
    regions; z: (R-1,) ascending boundaries; Einc: (W, 3) incident amplitudes
    at the origin (region 0, down-going).  Returns kz (W, R), D (W, R, 3)
    down-going amplitudes (D[:, 0] = Einc, D[:, j] referenced at z[j-1]) and
    U (W, R, 3) up-going amplitudes (U[:, j] referenced at z[j], U[:, R-1] = 0).
    """
    W = kx.shape[0]
    R = len(n)
    kz = np.stack([kz_of(n[j], k0, kx, ky) for j in range(R)], axis=1)  # (W, R)
    if R == 1:
        return kz, np.concatenate([Einc[:, None, :]], 1), np.zeros((W, 1, 3), np.complex128)
    nu = 6 * (R - 1)
    A = np.zeros((W, nu, nu), np.complex128)
    b = np.zeros((W, nu), np.complex128)

    def iu(j, c):  # up-going amplitude of region j (j <= R-2)
        return 3 * j + c

    def id_(j, c):  # down-going amplitude of region j (j >= 1)
        return 3 * (R - 1) + 3 * (j - 1) + c

    def tang(kvec):  # rows of [Ex, Ey, (k x E)_x, (k x E)_y] as linear maps of (Ex, Ey, Ez)
        kxv, kyv, kzv = kvec
        o = np.zeros(W)
        one = np.ones(W)
        return [(one, o, o), (o, one, o), (o, -kzv, kyv), (kzv, o, -kxv)]

    row = 0
    for l in range(R - 1):
        zl = z[l]
        # left side (region l)
        fD = np.exp(1j * kz[:, l] * (zl - (0.0 if l == 0 else z[l - 1])))
        tD = tang((kx, ky, kz[:, l]))
        tU = tang((kx, ky, -kz[:, l]))
        # right side (region l+1)
        tD2 = tang((kx, ky, kz[:, l + 1]))
        tU2 = tang((kx, ky, -kz[:, l + 1]))
        fU2 = np.exp(-1j * kz[:, l + 1] * (zl - z[l + 1])) if l + 1 <= R - 2 else None
        for q in range(4):
            for c in range(3):
                # left: down (known for l == 0) + up
                if l == 0:
                    b[:, row] -= tD[q][c] * fD * Einc[:, c]
                else:
                    A[:, row, id_(l, c)] += tD[q][c] * fD
                A[:, row, iu(l, c)] += tU[q][c]
                # minus right: down + up
                A[:, row, id_(l + 1, c)] -= tD2[q][c]
                if fU2 is not None:
                    A[:, row, iu(l + 1, c)] -= tU2[q][c] * fU2
            row += 1
    for j in range(R - 1):  # transversality k . E = 0 of each unknown wave
        for c, kc in enumerate((kx, ky, -kz[:, j])):
            A[:, row, iu(j, c)] = kc
        row += 1
        for c, kc in enumerate((kx, ky, kz[:, j + 1])):
            A[:, row, id_(j + 1, c)] = kc
        row += 1
    x = np.linalg.solve(A, b[..., None])[..., 0]
    D = np.zeros((W, R, 3), np.complex128)
    U = np.zeros((W, R, 3), np.complex128)
    D[:, 0] = Einc
    for j in range(R - 1):
        U[:, j] = x[:, 3 * j:3 * j + 3]
        D[:, j + 1] = x[:, 3 * (R - 1) + 3 * j:3 * (R - 1) + 3 * j + 3]
    return kz, D, U


def layered_structure(n, z, lam, NA, M, pols=("x",)):
    """Focused beam (M x M Cartesian pupil, NA) on a layered medium -> CoupledWaveStructure.

    n: (R,) complex indices (region 0 = incidence medium, must be real),
    z: (R-1,) ascending boundaries.  One incident wave per pupil point and
    polarisation, weight 1/M^2.
    """
    n = np.asarray(n, np.complex128)
    z = np.asarray(z, np.float64)
    R = len(n)
    k0 = 2 * math.pi / lam
    kxp, kyp = pupil_lattice(M, NA, k0)
    waves_by_pol = []
    for pol in pols:
        e = {"x": (1.0, 0.0, 0.0), "y": (0.0, 1.0, 0.0)}[pol]
        kz0 = kz_of(n[0], k0, kxp, kyp)
        d = np.stack([kxp, kyp, kz0.real], 1) / (n[0].real * k0)
        Einc = orthogonalize(np.broadcast_to(np.array(e, np.complex128), d.shape), d) / (M * M)
        waves_by_pol.append((kxp, kyp, Einc))
    kx = np.concatenate([w[0] for w in waves_by_pol])
    ky = np.concatenate([w[1] for w in waves_by_pol])
    Einc = np.concatenate([w[2] for w in waves_by_pol])
    kz, D, U = solve_stack(kx, ky, k0, n, z, Einc)

    def pack(E0, kzv):
        return np.column_stack([E0, kx, ky, kzv])

    Pi = pack(D[:, 0], kz[:, 0])
    layers = []
    for l in range(R - 1):
        Pr = pack(U[:, l] * np.exp(1j * kz[:, l] * z[l])[:, None], -kz[:, l])
        Pt = pack(D[:, l + 1] * np.exp(-1j * kz[:, l + 1] * z[l])[:, None], kz[:, l + 1])
        layers.append(HomogeneousLayer(z[l], Pr, Pt))
    return CoupledWaveStructure(Pi, layers)


def tile_sample_structure(M=21, Z=20, size=(64.0, 64.0, 20.0), lam=1.0, z_center=0.0, seed=None,
                          evanescent_fraction=0.1, incidence_deg=0.0):
    """Eigenmode structure of an sCWatter slab stack (sCWatter.cpp:890-956).

    M x M orders with sx = (j - M/2)/size_x * lam + dir_x (sCWatter.cpp:894-895),
    Z sublayers between Z+1 boundaries z_l = z_center - size_z/2 + l size_z/Z
    (sCWatter.cpp:157-163).  Each sublayer has 2MF down modes and 2MF up modes
    (MF = M^2) with propagation constants gamma shared by every order
    (kz = gamma k, sCWatter.cpp:912, 928); up modes use -gamma.  Waves per
    region: region 0 = Pi + Pr_0 (2 MF), internal regions 4 MF^2, last MF.
    """
    rng = np.random.default_rng(BASE_SEED + 1 if seed is None else seed)
    k = 2 * math.pi / lam
    MF = M * M
    th = math.radians(incidence_deg)
    dirx = math.sin(th)
    j = np.arange(M)
    sx1 = (j - M // 2) / size[0] * lam + dirx
    sy1 = (j - M // 2) / size[1] * lam
    # order p = kk * M + j  (sCWatter.cpp:897)
    sx = np.tile(sx1, M)
    sy = np.repeat(sy1, M)
    kx, ky = sx * k, sy * k
    kz0 = kz_of(1.0, k, kx, ky)
    zb = z_center - size[2] / 2 + np.arange(Z + 1) * (size[2] / Z)

    def gammas():
        g = np.sqrt(rng.uniform(1.0, 1.4, 2 * MF) ** 2 - rng.uniform(0.0, 0.05, 2 * MF)) + \
            1j * rng.uniform(0.0, 0.01, 2 * MF)
        ev = rng.random(2 * MF) < evanescent_fraction
        g[ev] = rng.uniform(0.0, 0.05, ev.sum()) + 1j * rng.uniform(0.05, 0.4, ev.sum())
        return g

    g_down = [gammas() for _ in range(Z)]
    amp_int = 1.0 / math.sqrt(4.0 * MF * MF)
    amp_ext = 1.0 / math.sqrt(2.0 * MF)

    def cplx(shape, scale):
        return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * (scale / math.sqrt(2))

    def modes(slab, z_ref, up):
        """Waves [p][jj] of one slab's 2MF down (or up) modes, wound to z_ref."""
        g = -g_down[slab] if up else g_down[slab]
        kzm = np.tile(g * k, MF)                    # (MF*2MF,)  jj inner
        kxm = np.repeat(kx, 2 * MF)
        kym = np.repeat(ky, 2 * MF)
        Q = cplx((MF * 2 * MF, 3), amp_int)
        E0 = Q * np.exp(-1j * kzm * z_ref)[:, None]
        return np.column_stack([E0, kxm, kym, kzm])

    Pi = np.column_stack([cplx((MF, 3), amp_ext), kx, ky, kz0])
    layers = []
    for l in range(Z + 1):
        if l == 0:
            Pr = np.column_stack([cplx((MF, 3), amp_ext) * np.exp(1j * kz0 * zb[0])[:, None], kx, ky, -kz0])
            Pt = modes(0, zb[0], up=False)
        elif l == Z:
            Pr = modes(Z - 1, zb[Z], up=True)
            Pt = np.column_stack([cplx((MF, 3), amp_ext) * np.exp(-1j * kz0 * zb[Z])[:, None], kx, ky, kz0])
        else:
            Pr = modes(l - 1, zb[l], up=True)
            Pt = modes(l, zb[l], up=False)
        layers.append(HomogeneousLayer(zb[l], Pr, Pt))
    return CoupledWaveStructure(Pi, layers)


def axis(lo, hi, n):
    """x_i = lo + i * (hi - lo) / (n - 1) (the reference's grid convention, sCWatterView.cpp:356-361)."""
    if n == 1:
        return np.array([0.5 * (lo + hi)])
    d = (hi - lo) / (n - 1)
    return lo + np.arange(n) * d


# ---------------------------------------------------------------------------
# The five benchmark configurations.  `scale` < 1 shrinks grids for tests.
# ---------------------------------------------------------------------------

def config1(grid=64):
    """Single interface n1=1.0 / n2=1.4+0.05i at z=0, lambda 1.0, NA 0.5, x-pol, 21x21 pupil; 64^3 over [-10,10]^3."""
    cw = layered_structure([1.0, 1.4 + 0.05j], [0.0], 1.0, 0.5, 21, ("x",))
    g = axis(-10.0, 10.0, grid)
    return cw, (g, g, g)


def config2(nx=256, ny=256, nz=128, M=21, Z=20, seed=None):
    """Tile-sample slab stack, 21x21 orders, 20 sublayers of 1 um; grid 256x256x128 over 64x64x20 um."""
    cw = tile_sample_structure(M=M, Z=Z, size=(64.0, 64.0, 20.0), lam=1.0, seed=seed)
    return cw, (axis(-32.0, 32.0, nx), axis(-32.0, 32.0, ny), axis(-10.0, 10.0, nz))


def config3(nx=512, ny=512, nz=256, half_extent=60.0):
    """Config 1 physics with a 41x41 pupil and x and y polarisations; grid 512x512x256 over [-L, L]^3."""
    cw = layered_structure([1.0, 1.4 + 0.05j], [0.0], 1.0, 0.5, 41, ("x", "y"))
    g = axis(-half_extent, half_extent, nx)
    return cw, (g, axis(-half_extent, half_extent, ny), axis(-half_extent, half_extent, nz))


def config5_structure(seed=None, n_layers=16):
    rng = np.random.default_rng(BASE_SEED + 5 if seed is None else seed)
    t = rng.uniform(0.1, 2.0, n_layers)
    n_mid = rng.uniform(1.0, 2.0, n_layers + 1) + 1j * rng.uniform(0.0, 0.05, n_layers + 1)
    n = np.concatenate([[1.0 + 0j], n_mid])           # incidence medium, 16 layers, substrate
    z = np.concatenate([[0.0], np.cumsum(t)])          # 17 boundaries
    z = z - z[-1] / 2
    return layered_structure(n, z, 0.6, 0.7, 31, ("x",)), z


def config5(nx=1024, ny=1024, nz=512, seed=None):
    """16-layer stack, 31x31 pupil, NA 0.7, lambda 0.6; grid through the stack with 2 um margins."""
    cw, z = config5_structure(seed)
    return cw, (axis(-10.0, 10.0, nx), axis(-10.0, 10.0, ny), axis(z[0] - 2.0, z[-1] + 2.0, nz))


def config4_instances(n=64, seed=None, orders=(11, 21, 31, 41), size=(64.0, 64.0, 4.0)):
    """Parameters of the batched sweep: n instances with lambda ~ U[0.4, 1.0] um,
    incidence ~ U[0, 40] deg and an order count drawn uniformly from `orders`
    (per axis).  Each instance is one tile-sample slab (one internal region of
    4 MF^2 eigenmode waves) of size 64 x 64 x 4 um centred at z = 0."""
    rng = np.random.default_rng(BASE_SEED + 4 if seed is None else seed)
    return [dict(M=int(rng.choice(orders)), lam=float(rng.uniform(0.4, 1.0)), angle=float(rng.uniform(0.0, 40.0)),
                 seed=int(rng.integers(1 << 30)), size=size) for _ in range(n)]


def config4_structure(p):
    return tile_sample_structure(M=p["M"], Z=1, size=p["size"], lam=p["lam"], seed=p["seed"],
                                 incidence_deg=p["angle"])


def config4_grid(nx=256, ny=256, nz=64, size=(64.0, 64.0, 4.0)):
    """256 x 256 x 64 per instance, z inside the slab."""
    h = size[2] / 2
    return (axis(-size[0] / 2, size[0] / 2, nx), axis(-size[1] / 2, size[1] / 2, ny),
            axis(-h + 0.05, h - 0.05, nz))


def map_instances(fn, params, workers=8):
    """fn(i, config4_structure(params[i]).unpack()) for every instance, on a
    pool of threads (numpy's random fills and copies and the library's plan
    builder run without the interpreter lock); results in instance order.
    Each instance's tables are generated from its own seed, so the results do
    not depend on the number of workers."""
    from concurrent.futures import ThreadPoolExecutor

    def one(i):
        return fn(i, config4_structure(params[i]).unpack())
    if workers <= 1:
        return [one(i) for i in range(len(params))]
    with ThreadPoolExecutor(max_workers=workers) as pool:
        return list(pool.map(one, range(len(params))))


# ---------------------------------------------------------------------------
# Physical eigen-tables: the same slab stacks solved by the host pipeline
# (sample_solver, sCWatter's Phi / eigen-solve / boundary chain) instead of
# seeded synthetic eigenvalues and amplitudes.
# ---------------------------------------------------------------------------

def config2_physical(nx=256, ny=256, nz=128, M=21, Z=20, pixels=64, tile=8, seed=None, workers=None):
    """Tile-sample slab stack 64 x 64 x 20 um, Z sublayers of pixels^2 voxels in
    seeded random tile x tile tiles of n in {1.0, 1.4 + 0.01i}, lit at normal
    incidence by a y-polarised plane wave at lambda 1.0 um, M x M orders,
    solved by sample_solver; the grid of config2."""
    from .sample_solver import solve_sample, tile_sample
    n = tile_sample(Z, pixels, pixels, tile=tile, n_background=1.0, n_tile=1.4 + 0.01j,
                    seed=BASE_SEED + 21 if seed is None else seed)
    cw = solve_sample(n, 1.0, (64.0, 64.0, 20.0), 0.0, (M, M), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0), workers=workers)
    return cw, (axis(-32.0, 32.0, nx), axis(-32.0, 32.0, ny), axis(-10.0, 10.0, nz))


def config4_physical_structure(p, pixels=64, tile=8, workers=None):
    """One instance of the batched sweep (its lambda, incidence angle and order
    count) as a physical single-slab tile sample of the sweep's size."""
    from .sample_solver import solve_sample, tile_sample
    n = tile_sample(1, pixels, pixels, tile=tile, n_background=1.0, n_tile=1.4 + 0.01j, seed=p["seed"])
    th = math.radians(p["angle"])
    return solve_sample(n, p["lam"], p["size"], 0.0, (p["M"], p["M"]), (math.sin(th), 0.0, math.cos(th)),
                        (0.0, 1.0, 0.0), workers=workers)
