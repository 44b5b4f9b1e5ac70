"""Host FP64 restatement of sCWatter's internal-field pipeline: the coupled
waves of a heterogeneous slab stack, written as the .cw structure the
evaluator reads.  This is host work in the reference and stays host work
here (numpy / LAPACK); it produces physical eigen-tables for the evaluator's
tile-sample workloads in place of seeded synthetic ones.

Steps, each following the reference:

* sublayer boundaries z_l = z_c - size_z / 2 + l size_z / Z (sCWatter.cpp:157-163);
* orders p (x) and q (y) from -M // 2 to (M - 1) // 2 with transverse
  wavenumbers u_p = 2 pi p / size_x + dir_x k and w_q = 2 pi q / size_y +
  dir_y k (FourierWave.h:319-333);
* per sublayer the Fourier coefficients of n^2 and n^-2: the 2D DFT of the
  slice divided by its pixel count, shifted so that the zero frequency is
  central, and the central M_y x M_x block kept (FourierWave.h:166-194);
* per sublayer the 4 MF x 4 MF property matrix Phi acting on the tangential
  field (Ex, Ey, Hx, Hy) of every order, H = (k_vector / k) x E, with the
  permittivity convolutions taken circularly over the kept block, as the
  reference indexes them (FourierWave.h:336-397);
* its eigen-decomposition (LAPACK zgeev), the eigenvalues gamma = kz / k
  sorted by imaginary part (ties by real part) and interleaved: even
  columns from the largest imaginary part down (modes decaying along +z),
  odd columns from the smallest up (sCWatter.cpp:168-225, 248-272);
* per sublayer the maps from mode amplitudes to the field at its top (Gd)
  and bottom (Gc) boundary, each mode referenced at the boundary it decays
  away from, and their chain across the stack (sCWatter.cpp:280-332);
* the 6 MF system of the outer boundaries: transversality of the reflected
  and transmitted orders, and continuity of (Ex, Ey, Hx, Hy) at the top
  boundary with the stack's transfer of the bottom field, solved for the
  reflected (at the top boundary) and transmitted (at the bottom boundary)
  amplitudes (sCWatter.cpp:334-489, 722);
* the mode amplitudes beta of every sublayer (sCWatter.cpp:749-766), and the
  mode tables Q = V diag(beta) with Ez from the n^-2 convolution of
  (u Hy - w Hx) (sCWatter.cpp:806-872);
* the .cw structure: per order the incident wave, per boundary the waves of
  the regions on either side, each wound to the origin
  (planewave::wind(0, 0, -z_l), sCWatter.cpp:890-956).

Differences from the reference, on purpose:

* One order index r = q M_x + p is used for everything.  The reference
  builds Phi and the .cw assembly in that order but flattens its Sx, Sy and
  Sz vectors column-major (q fastest, sCWatter.cpp:690-693), which agrees
  with it only for lattices symmetric under exchanging x and y (normal
  incidence on a square sample with square M).
* The outer waves' kz are Sz k; the reference multiplies them by the layer
  index once more (sCWatter.cpp:467, 474), which only matters when an outer
  index is not 1.
* Chains and beta use linear solves instead of explicit inverses.
"""
import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from .cw import CoupledWaveStructure, HomogeneousLayer


def orders(M):
    """Order numbers -M // 2 .. (M - 1) // 2 (Eigen setLinSpaced of sCWatter.cpp:794-795)."""
    return np.arange(M) - M // 2


def fourier_block(a, My, Mx):
    """Central My x Mx block of the shifted, normalised 2D DFT of a[y][x] (FourierWave.h:166-194)."""
    Ny, Nx = a.shape
    F = np.fft.fftshift(np.fft.fft2(a)) / (Ny * Nx)
    y0, x0 = Ny // 2 - My // 2, Nx // 2 - Mx // 2
    return F[y0:y0 + My, x0:x0 + Mx]


def conv_matrix(C, My, Mx):
    """MF x MF matrix of a circular convolution over the kept orders:
    entry (r, c), r = (qi, pi), c = (qj, pj), is C[(qi - qj) mod My shifted,
    (pi - pj) mod Mx shifted] with the zero frequency at (My // 2, Mx // 2),
    the reference's indexing of its coefficient block (FourierWave.h:363-368)."""
    q, p = np.divmod(np.arange(My * Mx), Mx)
    dq = (q[:, None] - q[None, :]) % My
    dp = (p[:, None] - p[None, :]) % Mx
    return C[(dq + My // 2) % My, (dp + Mx // 2) % Mx]


def phi_matrix(n_slice, My, Mx, u, w, k):
    """Property matrix of one sublayer (FourierWave.h:336-397) and the n^-2
    convolution matrix (for Ez).  u, w: per order r, the transverse
    wavenumbers."""
    MF = My * Mx
    A0 = conv_matrix(fourier_block(n_slice ** 2, My, Mx), My, Mx)
    Ni = conv_matrix(fourier_block(n_slice ** -2, My, Mx), My, Mx)
    A1 = Ni * u[None, :]
    A2 = Ni * w[None, :]
    kk = 1.0 / (k * k)
    I = np.eye(MF)
    P = np.zeros((4 * MF, 4 * MF), np.complex128)
    s0, s1, s2, s3 = (slice(i * MF, (i + 1) * MF) for i in range(4))
    P[s0, s2] = (u * kk)[:, None] * A2
    P[s0, s3] = -(u * kk)[:, None] * A1 + I
    P[s1, s2] = (w * kk)[:, None] * A2 - I
    P[s1, s3] = -(w * kk)[:, None] * A1
    P[s2, s0] = np.diag(-u * w * kk)
    P[s2, s1] = -A0 + np.diag(u * u * kk)
    P[s3, s0] = A0 + np.diag(-w * w * kk)
    P[s3, s1] = np.diag(u * w * kk)
    return P, Ni


def sorted_modes(P):
    """Eigen-decomposition of Phi with the reference's order: ascending by
    imaginary part (ties by real part), then interleaved, even columns from
    the top of that order (decaying along +z), odd from the bottom
    (sCWatter.cpp:168-225)."""
    lam, V = np.linalg.eig(P)
    order = np.lexsort((lam.real, lam.imag))
    n = lam.size
    idx = np.empty(n, np.int64)
    idx[0::2] = order[::-1][:n // 2]
    idx[1::2] = order[:n // 2]
    return lam[idx], V[:, idx]


def outer_rows(Sx, Sy, Sz):
    """(Ex, Ey, Hx, Hy) of a wave of direction (Sx, Sy, Sz) per unit (Ex, Ey, Ez): 4 MF x 3 MF."""
    MF = Sx.size
    D = lambda v: np.diag(v)
    I = np.eye(MF)
    Z = np.zeros((MF, MF))
    return np.block([[I, Z, Z], [Z, I, Z], [Z, -D(Sz), D(Sy)], [D(Sz), Z, -D(Sx)]])


def solve_sample(n, lam, size, z_center=0.0, M=(5, 5), direction=(0.0, 0.0, 1.0), pol=(0.0, 1.0, 0.0),
                 n_outer=(1.0, 1.0), workers=None):
    """Coupled-wave structure of the sample n[z][y][x] (complex indices,
    Z sublayers of Ny x Nx pixels) of physical size (size_x, size_y, size_z)
    centred at z_center, lit from above by the plane wave of direction
    `direction` (in the upper medium) and polarisation `pol` (made transverse
    as FourierWave.h:32-39 does), with M = (M_x, M_y) orders.  n_outer: the
    real index above and the complex index below the sample."""
    n = np.asarray(n, np.complex128)
    Zs, Ny, Nx = n.shape
    Mx, My = int(M[0]), int(M[1])
    if Mx > Nx or My > Ny:
        raise ValueError("more orders than pixels")
    MF = Mx * My
    k = 2 * math.pi / lam
    n0, n1 = complex(n_outer[0]), complex(n_outer[1])
    d = np.asarray(direction, np.float64)
    d = d / np.linalg.norm(d) * n0.real
    e = np.asarray(pol, np.complex128)
    e = e / np.sqrt(np.sum(e * e))
    e = np.cross(d, np.cross(e, d)) / (n0.real ** 2)  # d x (E x d) with |d| = n0
    z = z_center - size[2] / 2 + np.arange(Zs + 1) * (size[2] / Zs)
    q, p = np.divmod(np.arange(MF), Mx)
    u = 2 * math.pi * orders(Mx)[p] / size[0] + d[0] * k
    w = 2 * math.pi * orders(My)[q] / size[1] + d[1] * k
    Sx, Sy = u / k, w / k
    sq = lambda m: np.sqrt(np.asarray(m * m - Sx ** 2 - Sy ** 2, np.complex128))
    Sz0, Sz1 = sq(n0), sq(n1)

    # per sublayer: Phi, its sorted modes, and the n^-2 convolution
    def layer(i):
        P, Ni = phi_matrix(n[i], My, Mx, u, w, k)
        g, V = sorted_modes(P)
        return g, V, Ni
    workers = workers or min(16, os.cpu_count() or 1)
    if workers > 1 and Zs > 1:
        try:
            from threadpoolctl import threadpool_limits
            limit = threadpool_limits(1)
        except ImportError:
            limit = None
        with ThreadPoolExecutor(max_workers=workers) as pool:
            modes = list(pool.map(layer, range(Zs)))
        if limit is not None:
            limit.restore_original_limits()
    else:
        modes = [layer(i) for i in range(Zs)]

    # Gd / Gc per sublayer: down modes (even) referenced at the top z_i, up
    # modes (odd) at the bottom z_{i+1}
    even = np.arange(4 * MF) % 2 == 0
    Gd, Gc = [], []
    for i, (g, V, _) in enumerate(modes):
        h = z[i + 1] - z[i]
        Cf = np.exp(1j * k * g * h)     # down modes at the bottom
        Df = np.exp(-1j * k * g * h)    # up modes at the top
        Gd.append(V * np.where(even, 1.0, Df)[None, :])
        Gc.append(V * np.where(even, Cf, 1.0)[None, :])
    T = Gc[0]
    for i in range(1, Zs):
        T = Gc[i] @ np.linalg.solve(Gd[i], T)   # field at the bottom from beta_0

    # outer boundaries: unknowns x = [R (3 MF, at z_0), Tr (3 MF, at z_Z)]
    # incident orders referenced at the origin, evaluated at z_0 (their phase per column)
    f1 = outer_rows(Sx, Sy, Sz0) * np.tile(np.exp(1j * k * z[0] * Sz0), 3)[None, :]
    f2 = outer_rows(Sx, Sy, -Sz0)
    f3 = -outer_rows(Sx, Sy, Sz1)
    EF = np.zeros(3 * MF, np.complex128)
    r0 = (My // 2) * Mx + Mx // 2       # the zero order carries the incident plane wave
    EF[r0::MF] = e
    A = np.zeros((6 * MF, 6 * MF), np.complex128)
    b = np.zeros(6 * MF, np.complex128)
    rows = np.arange(MF)
    A[rows, rows], A[rows, MF + rows], A[rows, 2 * MF + rows] = Sx, Sy, -Sz0
    A[MF + rows, 3 * MF + rows], A[MF + rows, 4 * MF + rows], A[MF + rows, 5 * MF + rows] = Sx, Sy, Sz1
    A[2 * MF:, :3 * MF] = f2
    A[2 * MF:, 3 * MF:] = Gd[0] @ np.linalg.solve(T, f3)
    b[2 * MF:] = -f1 @ EF
    x = np.linalg.solve(A, b)
    R, Tr = x[:3 * MF], x[3 * MF:]

    # mode amplitudes and tables
    beta = [np.linalg.solve(Gd[0], f1 @ EF + f2 @ R)]
    for i in range(1, Zs):
        beta.append(np.linalg.solve(Gd[i], Gc[i - 1] @ beta[i - 1]))

    def tables(i, parity):
        g, V, Ni = modes[i]
        sel = slice(0, None, 2) if parity == 0 else slice(1, None, 2)
        Q = V[:, sel] * beta[i][sel][None, :]          # 4 MF x 2 MF: (Ex, Ey, Hx, Hy) per mode
        Ez = -(Ni @ (u[:, None] * Q[3 * MF:] - w[:, None] * Q[2 * MF:3 * MF])) / k
        return g[sel], np.stack([Q[:MF], Q[MF:2 * MF], Ez], axis=-1)   # (2 MF,), (MF, 2 MF, 3)

    kx, ky = Sx * k, Sy * k

    def waves(E0, kz, zl):
        """Records of waves with amplitudes E0 (count, 3) at z_l, wound to the origin."""
        kz = np.broadcast_to(kz, E0.shape[:1])
        E = E0 * np.exp(-1j * kz * zl)[:, None]
        return E, kz

    def pack(E, kxv, kyv, kzv):
        return np.column_stack([E, kxv, kyv, kzv])

    Pi = pack(EF.reshape(3, MF).T, kx, ky, Sz0 * k)
    down = [tables(i, 0) for i in range(Zs)]
    up = [tables(i, 1) for i in range(Zs)]
    layers = []
    for l in range(Zs + 1):
        zl = z[l]
        if l == 0:
            E, kz = waves(R.reshape(3, MF).T, -Sz0 * k, zl)
            Pr = pack(E, kx, ky, kz)
        else:
            g, Qh = up[l - 1]                               # up modes of the sublayer above, at its bottom
            kzm = np.tile(g * k, MF)
            E, _ = waves(Qh.reshape(MF * 2 * MF, 3), kzm, zl)
            Pr = pack(E, np.repeat(kx, 2 * MF), np.repeat(ky, 2 * MF), kzm)
        if l == Zs:
            E, kz = waves(Tr.reshape(3, MF).T, Sz1 * k, zl)
            Pt = pack(E, kx, ky, kz)
        else:
            g, Qc = down[l]                                 # down modes of the sublayer below, at its top
            kzm = np.tile(g * k, MF)
            E, _ = waves(Qc.reshape(MF * 2 * MF, 3), kzm, zl)
            Pt = pack(E, np.repeat(kx, 2 * MF), np.repeat(ky, 2 * MF), kzm)
        layers.append(HomogeneousLayer(zl, Pr, Pt))
    cw = CoupledWaveStructure(Pi, layers)
    cw.solution = {"R": R.reshape(3, MF).T, "T": Tr.reshape(3, MF).T, "Sx": Sx, "Sy": Sy, "Sz0": Sz0, "Sz1": Sz1,
                   "E_inc": e, "r0": r0, "z": z, "k": k}
    return cw


def tile_sample(Z, Ny, Nx, tile=8, n_background=1.0, n_tile=1.4 + 0.01j, fill=0.5, seed=0):
    """Seeded random tile sample n[z][y][x]: each slice is a grid of tile x
    tile pixel squares, each of index n_tile with probability `fill`, else
    n_background."""
    rng = np.random.default_rng(seed)
    ty, tx = -(-Ny // tile), -(-Nx // tile)
    on = rng.random((Z, ty, tx)) < fill
    big = np.repeat(np.repeat(on, tile, axis=1), tile, axis=2)[:, :Ny, :Nx]
    return np.where(big, complex(n_tile), complex(n_background))
