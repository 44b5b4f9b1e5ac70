// Separable GEMM form of the field sum (no counterpart in the reference,
// which only has the direct form gpuEvaluator.cu:28-68).
//
// A region whose waves sit on a lattice kx = kx_a, ky = ky_b (every
// producer emits such sets: sCWatter.cpp:893-899, sCWatterPlane.cpp:254-262,
// sCWatterLayer.cpp:402-414) satisfies, per field component,
//
//     E(x_i, y_j, z) = sum_b Y[j,b] sum_a X[i,a] C_z[a,b],
//     C_z[a,b]       = sum_c A[a,b,c] exp(i kz_c (z - zref)),
//     X[i,a] = exp(i kx_a x_i),  Y[j,b] = exp(i ky_b y_j).
//
// Three stages per grid:
//   coefficients: C_z for every z of the grid, in partial-sum slots
//                 (coef_kernel: VALU ELL contraction of the region's terms;
//                 modal_kernel / modal_wg_kernel: MFMA mode contraction of
//                 dense eigenmode regions).
//   X contraction: U_z[b][x][comp] = sum_a C_z[a,b,comp] X[a][x]; FP32:
//                 u_mfma (MFMA), U_z through HBM; 16-bit dtypes: csum_kernel
//                 adds the slots (and sums the terms of regions of at most 4
//                 per cell, which get no coefficient items), and main_16
//                 contracts X for its own columns.
//   main_*      : E_z[y][x][comp] = sum_b Y[y,b] U_z[b][x][comp] on MFMA.
//
// The main GEMM is a real GEMM with the complex product folded into K:
//   E[y][n] = sum_k Yp[y][k] Bp[k][n],  n = 6x + 2comp + (re|im),
//   Yp[y][2b] = Re Y[y,b], Yp[y][2b+1] = Im Y[y,b],
//   Bp[2b][n]   = U re/im as stored,  Bp[2b+1][n] = (i U) re/im,
// so its output rows are already the evaluator's AoS layout.  The MFMA
// "row" dimension is n (four consecutive n per lane -> 16-byte stores) and
// the "column" dimension is y.  Y fragments are pre-laid out on the host in
// the exact lane order of the MFMA B operand, so each is one coalesced load.
#include "scw_internal.hpp"

#include <mutex>
#include <type_traits>

namespace scw {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef uint32_t uint32x4 __attribute__((ext_vector_type(4)));
typedef float floatx2 __attribute__((ext_vector_type(2)));

// Term amplitudes of long z runs (the mode contraction's MT >= 4 classes) are
// read with non-temporal loads: they are streamed once per run, and keeping
// them out of the caches leaves those to the phase fragments and partial sums
// (12% faster on 64-slice runs).  Short runs use ordinary loads, which were
// faster there: between repeated evaluations part of the table stays in the
// last-level cache.
typedef const __attribute__((address_space(1))) floatx2* global_floatx2_ptr;  // global, not flat, loads
__device__ inline float2 stream_load(const float2* p) {
    const floatx2 v = __builtin_nontemporal_load((global_floatx2_ptr)(p));
    return make_float2(v[0], v[1]);
}
typedef uint32_t uint32x2 __attribute__((ext_vector_type(2)));
typedef const __attribute__((address_space(1))) uint32x2* global_uint32x2_ptr;
__device__ inline uint32x2 stream_load(const uint32x2* p) { return __builtin_nontemporal_load((global_uint32x2_ptr)(p)); }
typedef const __attribute__((address_space(1))) uint32x4* global_uint32x4_ptr;
__device__ inline uint32x4 stream_load(const uint32x4* p) { return __builtin_nontemporal_load((global_uint32x4_ptr)(p)); }
typedef const __attribute__((address_space(1))) uint32_t* global_uint32_ptr;
__device__ inline uint32_t stream_load(const uint32_t* p) { return __builtin_nontemporal_load((global_uint32_ptr)(p)); }

// threadIdx.x through an empty volatile asm: the compiler cannot hoist what is
// computed from it out of a loop.  In the single-launch kernel, whose unit loop
// runs every body, lane-derived addresses hoisted out of that loop stay live
// across all bodies and push them into scratch spills.
__device__ inline int tid_x() {
    int t = (int)threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

// Ordinary load through a pointer read from a device struct (the region
// tables).  The compiler cannot infer the address space of such a pointer and
// emits a FLAT load, which counts against both the vector-memory and the LDS
// wait counters; the first wait for anything then also waits for every
// prefetch in flight.  The cast makes it a global load.
template <typename T>
__device__ inline T gload(const T* p) {
    typedef const __attribute__((address_space(1))) T* global_ptr;
    return *(global_ptr)(p);  // an address-space cast (reinterpret_cast refuses it)
}
// float2's copy constructor takes a generic reference, which would turn the
// load back into a FLAT one: load the two floats as a vector instead.
__device__ inline float2 gload(const float2* p) {
    const floatx2 v = *(global_floatx2_ptr)(p);
    return make_float2(v[0], v[1]);
}

// Partial sums handed from one workgroup to another inside the single-launch
// form (fused16) go through agent-scope relaxed atomic accesses, which gfx950
// issues with sc1: stores write through and loads read past the L2 of their
// XCD, so a consumer on another XCD sees what the producer wrote once the
// producer's stores are complete (s_waitcnt vmcnt(0)) and it has raised the
// counter the consumer polls.  COH = false is the ordinary cached access of
// the stage-by-stage kernels.
template <bool COH>
__device__ inline void put2(float2* p, const float2 v) {
    if constexpr (COH) {
        const uint64_t u = __builtin_bit_cast(uint64_t, make_float2(v.x, v.y));
        __hip_atomic_store(reinterpret_cast<uint64_t*>(p), u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        *p = v;
    }
}
template <bool COH>
__device__ inline float2 get2(const float2* p) {
    if constexpr (COH) {
        const uint64_t u = __hip_atomic_load(reinterpret_cast<uint64_t*>(const_cast<float2*>(p)), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        return __builtin_bit_cast(float2, u);
    } else {
        return *p;
    }
}
// Two consecutive float2 as one floatx4 (16-byte aligned).
template <bool COH>
__device__ inline floatx4 get4(const float2* p) {
    if constexpr (COH) {
        const float2 a = get2<true>(p), b = get2<true>(p + 1);
        return floatx4{a.x, a.y, b.x, b.y};
    } else {
        return *reinterpret_cast<const floatx4*>(p);
    }
}
// Waits until this wave's stores have completed (vmcnt 0): with sc1 stores,
// they are then visible to every XCD.  The memory clobber keeps the compiler
// from moving the stores past it.
__device__ inline void stores_done() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// ---------------------------------------------------------------------------
// C_z[a,b,comp] partial sums over a term slice, for up to 8 z of one region.
// One wave per work item; the item, region and Z-row offsets are wave-uniform
// (scalar registers), so for dense eigenmode regions (kz index = term index)
// the Z values are scalar loads.  Terms are consumed four at a time with all
// twelve amplitude loads issued before the arithmetic, to keep several
// HBM requests in flight per wave.
// ---------------------------------------------------------------------------
__device__ inline void cmac6(float* acc, const float2 a0, const float2 a1, const float2 a2, const float2 f) {
    acc[0] = fmaf(a0.x, f.x, fmaf(-a0.y, f.y, acc[0]));
    acc[1] = fmaf(a0.x, f.y, fmaf(a0.y, f.x, acc[1]));
    acc[2] = fmaf(a1.x, f.x, fmaf(-a1.y, f.y, acc[2]));
    acc[3] = fmaf(a1.x, f.y, fmaf(a1.y, f.x, acc[3]));
    acc[4] = fmaf(a2.x, f.x, fmaf(-a2.y, f.y, acc[4]));
    acc[5] = fmaf(a2.x, f.y, fmaf(a2.y, f.x, acc[5]));
}

// Contraction of one work item for exactly NZ z rows.  DENSE: the kz index
// of term j is j, so the Z address is wave-uniform and Z is read through the
// scalar cache.
template <int NZ, bool DENSE>
__device__ inline void coef_loop(float (&acc)[8][6], const float2* __restrict__ Ab,
                                 const int32_t* __restrict__ Cb, int64_t stride, int64_t arow,
                                 const float2* __restrict__ Z, const int64_t (&zrow)[8], int j0, int j1) {
    typedef float2 Group[4][3];
    auto load = [&](Group& a, int (&c)[4], int jj) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float2* p = Ab + (int64_t)(jj + u) * arow;
            a[u][0] = gload(p);
            a[u][1] = gload(p + 1);
            a[u][2] = gload(p + 2);
            c[u] = DENSE ? jj + u : gload(Cb + (int64_t)(jj + u) * stride);
        }
    };
    auto mac = [&](const Group& a, const int (&c)[4]) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            float2 f[NZ];
#pragma unroll
            for (int zi = 0; zi < NZ; ++zi) f[zi] = Z[zrow[zi] + c[u]];
#pragma unroll
            for (int zi = 0; zi < NZ; ++zi) cmac6(acc[zi], a[u][0], a[u][1], a[u][2], f[zi]);
        }
    };
    // Whole groups of 4 terms, software-pipelined over two named buffers:
    // the next group's loads are in flight while the current one is
    // accumulated.  The prefetch start is clamped to the last whole group, so
    // the loads are unconditional (a clamped group is loaded and not used).
    int j = j0;
    const int ng = (j1 - j0) >> 2;
    if (ng > 0) {
        const int jlast = j0 + 4 * (ng - 1);
        Group a0, a1;
        int c0[4], c1[4];
        load(a0, c0, j);
        for (int g = 0;;) {
            load(a1, c1, min(j + 4, jlast));
            mac(a0, c0);
            j += 4;
            if (++g == ng) break;
            load(a0, c0, min(j + 4, jlast));
            mac(a1, c1);
            j += 4;
            if (++g == ng) break;
        }
    }
    for (; j < j1; ++j) {
        const float2* p = Ab + (int64_t)j * arow;
        const float2 a0 = gload(p), a1 = gload(p + 1), a2 = gload(p + 2);
        const int c = DENSE ? j : gload(Cb + (int64_t)j * stride);
#pragma unroll
        for (int zi = 0; zi < NZ; ++zi) cmac6(acc[zi], a0, a1, a2, Z[zrow[zi] + c]);
    }
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on
// XCD b % 8), each with its own L2.  The mode contraction's short-run class
// (modal_kernel, MT = 1) uses this logical block index instead of
// blockIdx.x: XCD x takes the contiguous logical range [x q + min(x, r), ...)
// (q = n / 8, r = n % 8, a bijection on [0, n)), so the workgroups of one run
// and mode slot, which read the same phase fragments, share an L2 instead of
// each XCD fetching them from HBM.  Tile sample, BF16x3: mode contraction
// 43.0 -> 38.2 us.  (The VALU item kernel and the 64-z workgroup kernel
// measured no gain from it; the 16-bit main GEMM uses the 3D form below.)
__device__ inline int xcd_block() {
    const int nb = (int)gridDim.x, b = (int)blockIdx.x, q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + min(x, r) + (b >> 3);
}
// The same order over a 3D grid (dispatch order: x fastest, then y, then z),
// returned as logical (x, y, z) block coordinates.
__device__ inline void xcd_block3(int& bx, int& by, int& bz) {
    const int gx = (int)gridDim.x, gy = (int)gridDim.y;
    const int nb = gx * gy * (int)gridDim.z;
    const int b = (int)blockIdx.x + gx * ((int)blockIdx.y + gy * (int)blockIdx.z);
    const int q = nb >> 3, r = nb & 7, x = b & 7;
    const int l = x * q + min(x, r) + (b >> 3);
    bx = l % gx;
    by = (l / gx) % gy;
    bz = l / (gx * gy);
}

template <bool COH>
__device__ inline void coef_body(const CoefItem& I, const RegionDev* __restrict__ regions,
                                 const ZMeta* __restrict__ zmeta, const float2* __restrict__ Z, int64_t c_elems,
                                 float2* __restrict__ C) {
    const int lane = tid_x() & 63;
    const RegionDev R = regions[I.region];
    const int cell = I.cell0 + lane;
    const bool active = cell < R.cells;
    const int cl = active ? cell : 0;
    const int nz = I.nz;
    int64_t zrow[8];
#pragma unroll
    for (int zi = 0; zi < 8; ++zi) zrow[zi] = zmeta[I.z0 + min(zi, nz - 1)].zrow_off;
    float acc[8][6];
#pragma unroll
    for (int zi = 0; zi < 8; ++zi)
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[zi][c] = 0.0f;

    const float2* Ab = R.ellA + (int64_t)cl * 3;
    const int32_t* Cb = R.dense ? nullptr : R.ellC + cl;
    const int64_t stride = (int64_t)R.cells, arow = R.arow;
#define SCW_COEF_CASE(N)                                                                   \
    case N:                                                                                \
        if (R.dense) coef_loop<N, true>(acc, Ab, Cb, stride, arow, Z, zrow, I.j0, I.j1);        \
        else coef_loop<N, false>(acc, Ab, Cb, stride, arow, Z, zrow, I.j0, I.j1);               \
        break;
    switch (nz) {
        SCW_COEF_CASE(1)
        SCW_COEF_CASE(2)
        SCW_COEF_CASE(3)
        SCW_COEF_CASE(4)
        SCW_COEF_CASE(5)
        SCW_COEF_CASE(6)
        SCW_COEF_CASE(7)
        default:
            if (R.dense) coef_loop<8, true>(acc, Ab, Cb, stride, arow, Z, zrow, I.j0, I.j1);
            else coef_loop<8, false>(acc, Ab, Cb, stride, arow, Z, zrow, I.j0, I.j1);
            break;
    }
#undef SCW_COEF_CASE
    if (!active) return;
#pragma unroll
    for (int zi = 0; zi < 8; ++zi) {
        if (zi < nz) {
            float2* o = C + (int64_t)I.slot * c_elems + zmeta[I.z0 + zi].c_off + (int64_t)cell * 3;
            put2<COH>(o, make_float2(acc[zi][0], acc[zi][1]));
            put2<COH>(o + 1, make_float2(acc[zi][2], acc[zi][3]));
            put2<COH>(o + 2, make_float2(acc[zi][4], acc[zi][5]));
        }
    }
}

__global__ __launch_bounds__(256) void coef_kernel(const CoefItem* __restrict__ items, int nitems,
                                                   const RegionDev* __restrict__ regions,
                                                   const ZMeta* __restrict__ zmeta, const float2* __restrict__ Z,
                                                   int64_t c_elems, float2* __restrict__ C) {
    const int it = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (tid_x() >> 6)));
    if (it >= nitems) return;
    coef_body<false>(items[it], regions, zmeta, Z, c_elems, C);
}


// ---------------------------------------------------------------------------
// X contraction of the FP32 dtype on MFMA (v_mfma_f32_16x16x4_f32): per z,
// U_z[b][x][comp] = sum_a C_z[a,b,comp] X[a][x] is a real GEMM with rows
// i = (m = 3 b + comp, re|im), K = (a, re|im) and columns x.  The A operand
// holds the signed parts of C (re,re = Re C, re,im = -Im C, im,re = Im C,
// im,im = Re C), taken from LDS after the partial-sum slots are added in a
// fixed order; the B operand is X as stored.  K mapping per 16 a as in the
// mode contraction: MFMA q, lane k = g <-> a = g + 4 (q / 2).  (The 16-bit
// dtypes run their X contraction inside the main GEMM.)
// Workgroup = 8 b (24 rows m) x 256 x of one z; wave = 3 row tiles x 64 x.
// ---------------------------------------------------------------------------
constexpr int UB = 8;  // b per workgroup

// PITCH: LDS row pitch in float2, ka + 2 for the largest ka (mx rounded up to
// 16) it takes (rows stay 16-byte aligned): 130 up to 128 kx, 258 up to 256.
template <int PITCH>
__global__ __launch_bounds__(256) void u_mfma(const ZMeta* __restrict__ zmeta, const RegionDev* __restrict__ regions,
                                              const float2* __restrict__ C, int64_t c_elems,
                                              const float2* __restrict__ X, int nx, float2* __restrict__ U, int zoff) {
    __shared__ float2 cs[3 * UB][PITCH];
    const int z = blockIdx.z + zoff;
    const ZMeta zm = zmeta[z];
    const int b0 = blockIdx.y * UB;
    if (b0 >= zm.my) return;
    const RegionDev R = regions[zm.region];
    const int mx = R.mx, T = (mx + 15) >> 4, ka = 16 * T;
    // C_z[a][b0 + r / 3][r % 3], slots added in order (16 loads in flight);
    // zero past mx and past my.
    for (int e = tid_x(); e < 3 * UB * ka; e += 256) {
        const int r = e / ka, a = e - r * ka;
        const int b = b0 + r / 3;
        float2 sum = make_float2(0.0f, 0.0f);
        if (a < mx && b < zm.my) {
            const int64_t off = zm.c_off + ((int64_t)b * mx + a) * 3 + (r % 3);
            for (int k0 = 0; k0 < R.nslots; k0 += 16) {
                float2 v[16];
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    v[u] = k0 + u < R.nslots ? C[(int64_t)(k0 + u) * c_elems + off] : make_float2(0.0f, 0.0f);
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    if (k0 + u < R.nslots) {
                        sum.x += v[u].x;
                        sum.y += v[u].y;
                    }
            }
        }
        cs[r][a] = sum;
    }
    __syncthreads();
    const int wv = tid_x() >> 6, lane = tid_x() & 63, j = lane & 15, g = lane >> 4;
    const int x0 = blockIdx.x * 256 + wv * 64;
    if (x0 >= nx) return;
    int xc[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) xc[nt] = min(x0 + 16 * nt + j, nx - 1);
    const float2* Xr = X + R.x_off;
    floatx4 acc[4][3];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int mt = 0; mt < 3; ++mt) acc[nt][mt] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    const int o = j & 1;  // row i = lane % 16: m = 8 mt + i / 2, part o
    for (int t = 0; t < T; ++t) {
        float2 xa[4][4], cv[3][4];
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const int a = 16 * t + g + 4 * h;
            const float2* xrow = Xr + (int64_t)min(a, mx - 1) * nx;
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) xa[nt][h] = xrow[xc[nt]];
#pragma unroll
            for (int mt = 0; mt < 3; ++mt) cv[mt][h] = cs[8 * mt + (j >> 1)][a];
        }
        // signed part of C for (h, s): o = 0 -> (Re, -Im), o = 1 -> (Im, Re)
        auto cpart = [&](const float2 c, int s) { return o == 0 ? (s == 0 ? c.x : -c.y) : (s == 0 ? c.y : c.x); };
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const float bv = (q & 1) ? xa[nt][q >> 1].y : xa[nt][q >> 1].x;
#pragma unroll
                for (int mt = 0; mt < 3; ++mt)
                    acc[nt][mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(cpart(cv[mt][q >> 1], q & 1), bv,
                                                                        acc[nt][mt], 0, 0, 0);
            }
    }
    // lane holds D[i = 4 g + r][x]: rows (m = 8 mt + 2 g + r / 2, part r % 2)
#pragma unroll
    for (int mt = 0; mt < 3; ++mt)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int m = 8 * mt + 2 * g + h;
            const int b = b0 + m / 3;
            if (b >= zm.my) continue;
            float2* ub = U + zm.u_off + (int64_t)b * nx * 3 + (m % 3);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const int x = x0 + 16 * nt + j;
                if (x < nx) ub[(int64_t)x * 3] = make_float2(acc[nt][mt][2 * h], acc[nt][mt][2 * h + 1]);
            }
        }
}

constexpr int kBufferRsrcWord3 = 0x00020000;  // raw dword buffer, gfx9 family
constexpr int kStoreNT = 2;                   // non-temporal

// The main GEMM's resident operand Bp[k][n], read from one z-slice of U (my
// rows of ncols floats, under 2 GiB) through a raw buffer.  The hardware range
// check returns zero for rows past my, which covers every k past the region's
// K; a lane past the field (n >= ncols) reads part of the next row or zero,
// which only reaches output columns that the range-checked stores drop.  So
// the loads carry no conditions: a load under a lane-dependent branch is
// waited for before the branch ends, which would serialize the fragment.
__device__ inline __amdgpu_buffer_rsrc_t u_slice_rsrc(const float* Ub, int my, int64_t ncols) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(Ub), 0, (int)(my * ncols * 4), kBufferRsrcWord3);
}
__device__ inline float bp_buffer(__amdgpu_buffer_rsrc_t rs, int64_t ncols, int k, int64_t n) {
    const int off = (int)(((int64_t)(k >> 1) * ncols + (n & ~(int64_t)1)) * 4);
    const uint32x2 w = __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0);
    const uint32_t wx = w.x, wy = w.y;
    const float ux = __builtin_bit_cast(float, wx), uy = __builtin_bit_cast(float, wy);
    const bool odd = (n & 1) != 0;
    return (k & 1) == 0 ? (odd ? uy : ux) : (odd ? ux : -uy);
}

// Epilogue of one 16-row y tile of a workgroup (256 output columns): every
// wave parks its 16 x 64 accumulator tile in LDS, then each store
// instruction writes one full 1 KB row segment (lane l: columns 4l..4l+3).
// T is double-buffered by tile parity, so one barrier per tile suffices: a
// wave can only pass the barrier of tile t once every wave has finished
// reading the buffer of tile t-1, which tile t+1 then reuses.  The field is
// written once and not read back, so the stores are non-temporal.
constexpr int TROW = 256 + 4;

// One 16-byte (VEC) or 4 x 4-byte store per lane into row `orow` through a
// raw buffer resource whose range is the row (empty when the row is past ny):
// the hardware range check drops the lanes and rows outside the field, so
// the stores carry no branches and the wait counters see the same number of
// stores on every path.
// With acc (a ky chunk after the first), the values are added to what the
// row holds (loads through the same range-checked resource).
template <bool VEC>
__device__ inline void row_store(float* orow, bool row_ok, int64_t n, int64_t ncols, const floatx4& v, bool acc) {
    const uint32_t nrec = row_ok ? (uint32_t)(ncols * 4) : 0u;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(orow, 0, (int)nrec, kBufferRsrcWord3);
    const int off = (int)(n * 4);
    if (VEC) {
        floatx4 w = v;
        if (acc) w += __builtin_bit_cast(floatx4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(uint32x4, w), rs, off, 0, kStoreNT);
    } else {
        // Each element gets its own offset, moved far out of range past the
        // row's end, so a partly valid group of 4 columns is written element
        // by element and the stores are never merged into one 16-byte access.
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float f = v[r];
            const int off_r = n + r < ncols ? off + 4 * r : 0x7ffffff0;
            if (acc) f += __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, off_r, 0, 0));
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, f), rs, off_r, 0, kStoreNT);
        }
    }
}

template <bool VEC>
__device__ inline void tile_out(float (*T)[TROW], const floatx4 (&acc)[4], float* __restrict__ oz, int64_t ncols,
                                int ny, int yt, int64_t nwg0, bool add) {
    const int lane = tid_x() & 63, li = lane & 15, lk = lane >> 4;
    const int wv = __builtin_amdgcn_readfirstlane(tid_x() >> 6);
#pragma unroll
    for (int it = 0; it < 4; ++it) *reinterpret_cast<floatx4*>(&T[li][64 * wv + 16 * it + 4 * lk]) = acc[it];
    __syncthreads();
    floatx4 v[4];
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) v[rr] = *reinterpret_cast<const floatx4*>(&T[4 * rr + wv][4 * lane]);
    const int64_t n = nwg0 + 4 * lane;
    const int y0 = yt * 16 + wv;
#pragma unroll
    for (int rr = 0; rr < 4; ++rr)
        row_store<VEC>(oz + (int64_t)(y0 + 4 * rr) * ncols, y0 + 4 * rr < ny, n, ncols, v[rr], add);
}

// Slice of a region with no waves: the workgroup's columns of y tiles
// [yt0, yt1) are zero.
template <bool VEC>
__device__ inline void zero_tiles(float* __restrict__ oz, int64_t ncols, int ny, int yt0, int yt1, int64_t nwg0) {
    const int lane = tid_x() & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid_x() >> 6);
    const int64_t n = nwg0 + 4 * lane;
    const int yend = min(ny, 16 * yt1);
    for (int y = 16 * yt0 + wv; y < yend; y += 4)
        row_store<VEC>(oz + (int64_t)y * ncols, true, n, ncols, floatx4{0.0f, 0.0f, 0.0f, 0.0f}, false);
}

// ---------------------------------------------------------------------------
// Main GEMM, FP32 operands: v_mfma_f32_16x16x4_f32.
// Workgroup = 4 waves, each owns 64 output columns (n) of one z-slice and
// walks y in tiles of 16; its A-operand fragment (from U) stays in registers.
// The B-operand (Y) fragment of a tile is four k-steps per 16-byte load,
// prefetched two tiles ahead.
// ---------------------------------------------------------------------------
template <int KMAX, bool VEC>
__global__ __launch_bounds__(256) void main_f32(const ZMeta* __restrict__ zmeta, const float2* __restrict__ U,
                                                const float* __restrict__ Yfrag, int nx, int ny, int ytiles_per_split,
                                                float* __restrict__ out, int zoff) {
    constexpr int Q = KMAX / 16;  // 16-byte Y loads per tile
    // (plain block order: the XCD-contiguous order of main_16 measured slower
    // here, 3.62 -> 3.94 ms on the 16-layer stack)
    const int bx = (int)blockIdx.x, by = (int)blockIdx.y, bz = (int)blockIdx.z;
    const ZMeta zm = zmeta[by + zoff];  // main-GEMM entry
    const int K16 = zm.kp >> 4;
    const int wv = tid_x() >> 6, lane = tid_x() & 63;
    const int64_t ncols = 6LL * nx;
    const int64_t n0 = (int64_t)bx * 256 + wv * 64;  // waves past ncols compute zeros
    const int li = lane & 15, lk = lane >> 4;
    const __amdgpu_buffer_rsrc_t urs = u_slice_rsrc(reinterpret_cast<const float*>(U + zm.u_off), zm.my, ncols);
    const int ytiles = (ny + 15) >> 4;
    const int yt0 = bz * ytiles_per_split;
    const int yt1 = min(ytiles, yt0 + ytiles_per_split);
    float* oz = out + (int64_t)zm.zout * ny * ncols;
    if (K16 == 0) {
        zero_tiles<VEC>(oz, ncols, ny, yt0, yt1, (int64_t)bx * 256);
        return;
    }
    __shared__ float T[2][16][TROW];

    float uf[4][KMAX / 4];
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        const int64_t n = n0 + 16 * it + li;
#pragma unroll
        for (int s = 0; s < KMAX / 4; ++s) {
            const int k = 4 * s + lk;
            uf[it][s] = bp_buffer(urs, ncols, k, n);
        }
    }

    const floatx4* Yf = reinterpret_cast<const floatx4*>(Yfrag + zm.yfrag_off) + lane;
    // Y fragment loads are unconditional (k-slices past the region's K16
    // re-read its last one and are not used; prefetches past the last tile
    // re-read that tile), so every path issues the same loads.  The
    // loop is unrolled by two over named register sets, so a prefetch is
    // waited for only when its tile is computed, after the previous tile's
    // stores were issued.
    auto yload = [&](int yt, floatx4 (&y)[Q]) {
#pragma unroll
        for (int q = 0; q < Q; ++q) y[q] = Yf[((int64_t)yt * K16 + min(q, K16 - 1)) * 64];
    };
    auto tile = [&](int yt, const floatx4 (&y)[Q]) {
        floatx4 acc[4];
#pragma unroll
        for (int it = 0; it < 4; ++it) acc[it] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int q = 0; q < Q; ++q)
            if (q < K16)  // uniform; k-slices past the region's contribute zero
#pragma unroll
                for (int e = 0; e < 4; ++e)
#pragma unroll
                    for (int it = 0; it < 4; ++it)
                        acc[it] = __builtin_amdgcn_mfma_f32_16x16x4f32(uf[it][4 * q + e], y[q][e], acc[it], 0, 0, 0);
        tile_out<VEC>(T[yt & 1], acc, oz, ncols, ny, yt, (int64_t)bx * 256, zm.acc != 0);
    };
    // The first pair of tiles is peeled so that the loop is entered with the
    // same loads and stores in flight as at its back edge.
    floatx4 y0[Q], y1[Q];
    yload(yt0, y0);
    yload(min(yt0 + 1, yt1 - 1), y1);
    tile(yt0, y0);
    yload(min(yt0 + 2, yt1 - 1), y0);
    if (yt0 + 1 >= yt1) return;
    tile(yt0 + 1, y1);
    yload(min(yt0 + 3, yt1 - 1), y1);
    int yt = yt0 + 2;
    for (; yt + 1 < yt1; yt += 2) {
        tile(yt, y0);
        yload(min(yt + 2, yt1 - 1), y0);
        tile(yt + 1, y1);
        yload(min(yt + 3, yt1 - 1), y1);
    }
    if (yt < yt1) tile(yt, y0);
}

// ---------------------------------------------------------------------------
// Main GEMM, 16-bit operands with FP32 accumulation:
//   MODE 0: bf16      v_mfma_f32_16x16x32_bf16
//   MODE 1: bf16 x3   hi*hi + hi*lo + lo*hi (hi = rn(v), lo = rn(v - hi))
//   MODE 2: fp16      v_mfma_f32_16x16x32_f16
// A-operand lane layout: lane l holds k = 32 s + 8 (l >> 4) + e, e = 0..7.
// ---------------------------------------------------------------------------
template <int MODE>
struct Op16;
template <>
struct Op16<0> {
    typedef bf16x8 vec;
    typedef __bf16 elem;
};
template <>
struct Op16<1> {
    typedef bf16x8 vec;
    typedef __bf16 elem;
};
template <>
struct Op16<2> {
    typedef halfx8 vec;
    typedef _Float16 elem;
};

template <int MODE>
__device__ inline floatx4 mfma16(const typename Op16<MODE>::vec a, const typename Op16<MODE>::vec b, floatx4 c) {
    if constexpr (MODE == 2)
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// X contraction inside the 16-bit main GEMM.  A slot-sum kernel first adds
// the coefficient partials of every z in slot order into
//   Cs_z[m][a] = C_z[a, b = m / 3, comp = m % 3],  m < 3 my, a < ka (zero past mx),
// and each main-GEMM workgroup then computes U_z[b][p] = sum_a Cs_z[3 b + comp][a]
// X[a][x] for its own 128 column pairs p = 3 x + comp on MFMA, into LDS, so
// U_z never goes through HBM.  The GEMM is real, with rows (m, re|im), K =
// (a, re|im) and columns x; the A operand holds the signed parts of C (re,re =
// Re C, re,im = -Im C, im,re = Im C, im,im = Re C), the B operand X, both split
// into bf16 hi + lo (bf16 x3: hi hi + hi lo + lo hi), as the FP32 X
// contraction kernel did before it moved here.
// ---------------------------------------------------------------------------
// Element e = m ka + a of the slot-summed Cs_z of one z (m < 3 my, a < ka).
// Element (m, a) of Cs_z for a region of few terms per cell (nslots = 0), not
// contracted before: C_z[cell][comp] = sum_j A[j][cell][comp] Z_z[kz index of
// term j], b = m / 3, comp = m % 3, cell = b mx + a (zero past mx).
__device__ inline float2 folded_elem(const ZMeta& zm, const RegionDev& R, const float2* __restrict__ Z, int m, int a) {
    float2 sum = make_float2(0.0f, 0.0f);
    if (a >= zm.mx) return sum;
    const int b = m / 3, comp = m - 3 * b, cell = b * zm.mx + a;
    for (int j = 0; j < R.width; ++j) {
        const float2 av = R.ellA[(int64_t)j * R.arow + (int64_t)cell * 3 + comp];
        const float2 zv = Z[zm.zrow_off + R.ellC[(int64_t)j * R.cells + cell]];
        sum.x += av.x * zv.x - av.y * zv.y;
        sum.y += av.x * zv.y + av.y * zv.x;
    }
    return sum;
}

template <bool COH>
__device__ inline float2 csum_elem(const ZMeta& zm, const RegionDev& R, const float2* __restrict__ C,
                                   int64_t c_elems, const float2* __restrict__ Z, int e) {
    const int ka = R.ka, ns = R.nslots;
    const int m = e / ka, a = e - m * ka;
    float2 sum = make_float2(0.0f, 0.0f);
    if (a >= zm.mx) return sum;
    if (ns == 0) return folded_elem(zm, R, Z, m, a);
    const int64_t off = zm.c_off + ((int64_t)(m / 3) * zm.mx + a) * 3 + (m % 3);
    for (int k0 = 0; k0 < ns; k0 += 16) {
        float2 v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u)
            v[u] = k0 + u < ns ? get2<COH>(C + (int64_t)(k0 + u) * c_elems + off) : make_float2(0.0f, 0.0f);
#pragma unroll
        for (int u = 0; u < 16; ++u)
            if (k0 + u < ns) {
                sum.x += v[u].x;
                sum.y += v[u].y;
            }
    }
    return sum;
}

__global__ __launch_bounds__(256) void csum_kernel(const ZMeta* __restrict__ zmeta,
                                                   const RegionDev* __restrict__ regions,
                                                   const float2* __restrict__ C, int64_t c_elems,
                                                   const float2* __restrict__ Z, float2* __restrict__ Cs,
                                                   const int2* __restrict__ blocks) {
    const int2 blk = blocks[blockIdx.x];  // (z, first element): only the elements each z has
    const ZMeta zm = zmeta[blk.x];
    const RegionDev& R = regions[zm.region];
    const int e = blk.y + tid_x();
    if (e >= 3 * zm.my * R.ka) return;
    Cs[zm.u_off + e] = csum_elem<false>(zm, R, C, c_elems, Z, e);
}

constexpr int UPL = 130;  // LDS row pitch (float2) of a workgroup's U_z columns
// K steps (16 a each) whose X fragments a workgroup stages in LDS: 3 (mx <=
// 48) up to KMAX = 96; none for KMAX = 128, whose U_z columns alone take
// 66.5 KB (two workgroups per CU).
template <int KMAX>
constexpr int xt_staged() { return KMAX <= 96 ? 3 : 0; }
typedef uint32x4 XFrag[2];    // B-operand fragment of one lane: bf16 x 8 high, then low

__device__ inline void split_bf16(float v, __bf16& h, __bf16& l) {
    h = (__bf16)v;
    l = (__bf16)(v - (float)h);
}

// A operand of the X contraction for row m and K step t: the signed parts of
// C_z[m][16 t + 4 g + h] (c01 = a, a + 1; c23 = a + 2, a + 3; re, im each),
// row part o = 0 -> (Re, -Im), o = 1 -> (Im, Re), split into bf16 hi / lo.
__device__ inline void xc_a_operand(const floatx4 c01, const floatx4 c23, int o, bf16x8& ah, bf16x8& al) {
    const float cre[4] = {c01[0], c01[2], c23[0], c23[2]}, cim[4] = {c01[1], c01[3], c23[1], c23[3]};
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int h = e >> 1;
        const float v = o == 0 ? ((e & 1) ? -cim[h] : cre[h]) : ((e & 1) ? cre[h] : cim[h]);
        __bf16 hh, ll;
        split_bf16(v, hh, ll);
        ah[e] = hh;
        al[e] = ll;
    }
}
// B operand: X[a][x] for a = 16 t + 4 g + e / 2, part e % 2 (q0 = a, a + 1; q1 = a + 2, a + 3).
__device__ inline void xc_b_operand(const floatx4 q0, const floatx4 q1, bf16x8& bh, bf16x8& bl) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const float v = e < 4 ? q0[e] : q1[e - 4];
        __bf16 hh, ll;
        split_bf16(v, hh, ll);
        bh[e] = hh;
        bl[e] = ll;
    }
}
__device__ inline void xc_mfma(floatx4& acc, const bf16x8& ah, const bf16x8& al, const bf16x8& bh, const bf16x8& bl) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
}

// U_z[b][128 bx + pl] for pl < 128 and b < my into Ul[b][pl] (bx: the
// workgroup's column block).  Wave
// wv takes row tiles wv, wv + 4, ... (8 rows m each) against the 3 column
// tiles of 16 x from x_lo = 128 bx / 3 on (pl = 3 (x - x_lo) + comp -
// r0 covers 0..127 with x - x_lo < 44).  Rows past 3 my and x past nx
// re-read the last row / column (results not stored / only for columns the
// stores drop).  With at most XS K steps (mx <= 16 XS) the workgroup
// first stages the X fragments, already split, in LDS (xs), and each row tile
// issues all its C loads at once: one L2 round trip per row tile instead of
// one per K step; larger mx takes the K steps one at a time.
// Four consecutive elements a0 .. a0 + 3 of row m of Cs_z.  Ordinary loads
// also in the single-launch kernel (COH): its main units read Cs_z only after
// the slot sums of z are complete, and every Cs_z starts on its own 256-byte
// boundary, so no cache holds a line of it from before (caches start clean
// at launch); the lines then stay in the L2 for the other column blocks of z.
__device__ inline void cs_row4(const ZMeta& zm, const RegionDev& R, const float2* __restrict__ Cz, int m, int a0,
                               floatx4& c01, floatx4& c23) {
    const float2* cp = Cz + (int64_t)m * R.ka + a0;
    c01 = get4<false>(cp);
    c23 = get4<false>(cp + 2);
}

template <int XS, bool COH>
__device__ __attribute__((always_inline)) inline void xcontract_columns(const ZMeta& zm, const RegionDev& R, const float2* __restrict__ Cs,
                                         const float2* __restrict__ X, int nx, int bx, float2 (*Ul)[UPL],
                                         XFrag* xs) {
    const int lane = tid_x() & 63, j = lane & 15, g = lane >> 4, o = j & 1;
    const int wv = __builtin_amdgcn_readfirstlane(tid_x() >> 6);
    const int p0 = 128 * bx, x_lo = p0 / 3, r0 = p0 - 3 * x_lo;
    const int M = 3 * zm.my, RT = (M + 7) >> 3, T = R.ka >> 4;
    const float2* Cz = Cs + zm.u_off;
    const float2* Xr = X + R.x_off;
    auto xload = [&](int ct, int t, floatx4& q0, floatx4& q1) {
        const int x = min(x_lo + 16 * ct + j, nx - 1);
        const floatx4* xp = reinterpret_cast<const floatx4*>(Xr + (int64_t)x * R.ka + 16 * t + 4 * g);
        q0 = xp[0];
        q1 = xp[1];
    };
    auto store = [&](const floatx4 (&acc)[3], int rt) {
        // lane holds D[i = 4 g + r][x - x_lo = 16 ct + j]: rows m = 8 rt + 2 g + r / 2, part r % 2
#pragma unroll
        for (int ct = 0; ct < 3; ++ct)
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                const int mm = 8 * rt + 2 * g + hh;
                const int b = mm / 3, comp = mm - 3 * b;
                const int pl = 3 * (16 * ct + j) + comp - r0;
                if (mm < M && pl >= 0 && pl < 128) Ul[b][pl] = make_float2(acc[ct][2 * hh], acc[ct][2 * hh + 1]);
            }
    };
    if (XS > 0 && T <= XS) {  // uniform
        // unrolled to the largest count (3 XS 64 / 256 per thread) so the
        // loads of all iterations are in flight together; the loads are
        // unconditional (index clamped), only the LDS writes are guarded
        const int nf = 3 * T * 64;
#pragma unroll
        for (int it = 0; it < (3 * (XS > 0 ? XS : 1) * 64 + 255) / 256; ++it) {
            const int f = tid_x() + 256 * it, fc = min(f, nf - 1);
            const int l = fc & 63, ct = (fc >> 6) % 3, t = (fc >> 6) / 3;
            const int x = min(x_lo + 16 * ct + (l & 15), nx - 1);
            const floatx4* xp = reinterpret_cast<const floatx4*>(Xr + (int64_t)x * R.ka + 16 * t + 4 * (l >> 4));
            bf16x8 bh, bl;
            xc_b_operand(xp[0], xp[1], bh, bl);
            if (f < nf) {
                xs[f][0] = __builtin_bit_cast(uint32x4, bh);
                xs[f][1] = __builtin_bit_cast(uint32x4, bl);
            }
        }
        __syncthreads();
        // Two row tiles per round trip: a wave issues the C loads of its next
        // two row tiles together (the second clamped to the last tile, its
        // result then unused).  When C_z comes from HBM rather than a cache,
        // as after another kernel has streamed through them, each round trip
        // is long, and this halves their number (64-slab batch, BF16x3 main
        // GEMM after a 256 MiB copy: 1.73 -> 1.51 ms; 1.10 ms warm either way).
        constexpr int NB = 2;
        for (int rb = wv; rb < RT; rb += 4 * NB) {
            floatx4 cv[NB][XS > 0 ? XS : 1][2];
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int m = min(8 * min(rb + 4 * b, RT - 1) + (j >> 1), M - 1);
#pragma unroll
                for (int t = 0; t < XS; ++t)  // unconditional: steps past T re-read the last one
                    cs_row4(zm, R, Cz, m, 16 * min(t, T - 1) + 4 * g, cv[b][t][0], cv[b][t][1]);
            }
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const int rt = rb + 4 * b;
                if (rt >= RT) break;  // uniform
                floatx4 acc[3];
#pragma unroll
                for (int ct = 0; ct < 3; ++ct) acc[ct] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int t = 0; t < XS; ++t) {
                    if (t >= T) break;  // uniform
                    bf16x8 ah, al;
                    xc_a_operand(cv[b][t][0], cv[b][t][1], o, ah, al);
#pragma unroll
                    for (int ct = 0; ct < 3; ++ct) {
                        const XFrag& xf = xs[(t * 3 + ct) * 64 + lane];
                        xc_mfma(acc[ct], ah, al, __builtin_bit_cast(bf16x8, xf[0]), __builtin_bit_cast(bf16x8, xf[1]));
                    }
                }
                store(acc, rt);
            }
        }
    } else {
        for (int rt = wv; rt < RT; rt += 4) {
            const int m = min(8 * rt + (j >> 1), M - 1);
            floatx4 acc[3];
#pragma unroll
            for (int ct = 0; ct < 3; ++ct) acc[ct] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
            for (int t = 0; t < T; ++t) {
                floatx4 xv[3][2];
#pragma unroll
                for (int ct = 0; ct < 3; ++ct) xload(ct, t, xv[ct][0], xv[ct][1]);
                floatx4 c01, c23;
                cs_row4(zm, R, Cz, m, 16 * t + 4 * g, c01, c23);
                bf16x8 ah, al;
                xc_a_operand(c01, c23, o, ah, al);
#pragma unroll
                for (int ct = 0; ct < 3; ++ct) {
                    bf16x8 bh, bl;
                    xc_b_operand(xv[ct][0], xv[ct][1], bh, bl);
                    xc_mfma(acc[ct], ah, al, bh, bl);
                }
            }
            store(acc, rt);
        }
    }
}

// LDS of main_16: the workgroup's U_z columns while the fragments are built,
// then the double-buffered output transpose.
template <int KMAX>
struct Main16Lds {
    static constexpr int kTBytes = 2 * 16 * TROW * 4, kUBytes = KMAX / 2 * UPL * 8;
    static constexpr int kXS = xt_staged<KMAX>(), kXBytes = 3 * kXS * 64 * (int)sizeof(XFrag);
    static constexpr int kBytes = kTBytes > kUBytes + kXBytes ? kTBytes : kUBytes + kXBytes;
};

// One main-GEMM workgroup: column block bx, z-slice z, y split bz.
template <int KMAX, bool VEC, int MODE, bool COH>
__device__ inline void main_16_body(int bx, int z, int bz, const ZMeta* __restrict__ zmeta,
                                    const RegionDev* __restrict__ regions, const float2* __restrict__ Cs,
                                    const float2* __restrict__ X, const uint16_t* __restrict__ Yfrag,
                                    int64_t ylo_delta, int nx, int ny, int ytiles_per_split, float* __restrict__ out,
                                    unsigned char* lds) {
    typedef typename Op16<MODE>::vec vec;
    typedef typename Op16<MODE>::elem elem;
    constexpr int S = KMAX / 32;
    const ZMeta zm = zmeta[z];  // main-GEMM entry
    const int K32 = zm.kp >> 5;
    const int wv = tid_x() >> 6, lane = tid_x() & 63;
    const int64_t ncols = 6LL * nx;
    const int64_t n0 = (int64_t)bx * 256 + wv * 64;  // waves past ncols compute zeros
    const int li = lane & 15, lk = lane >> 4;
    const int ytiles = (ny + 15) >> 4;
    const int yt0 = bz * ytiles_per_split;
    const int yt1 = min(ytiles, yt0 + ytiles_per_split);
    float* oz = out + (int64_t)zm.zout * ny * ncols;
    if (K32 == 0) {
        zero_tiles<VEC>(oz, ncols, ny, yt0, yt1, (int64_t)bx * 256);
        return;
    }
    typedef Main16Lds<KMAX> L;
    float (*T)[16][TROW] = reinterpret_cast<float (*)[16][TROW]>(lds);
    float2 (*Ul)[UPL] = reinterpret_cast<float2 (*)[UPL]>(lds);
    XFrag* xs = reinterpret_cast<XFrag*>(lds + L::kUBytes);
    xcontract_columns<L::kXS, COH>(zm, regions[zm.region], Cs, X, nx, bx, Ul, xs);
    __syncthreads();
    // A fragment: Bp[k][n] from U_z[b = k >> 1][n >> 1]; rows past my are zero
    vec ahi[4][S], alo[4][S];
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        const int64_t n = n0 + 16 * it + li;
        const int pl = (64 * wv + 16 * it + li) >> 1;
        const bool odd = (n & 1) != 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = 32 * s + 8 * lk + e, b = k >> 1;
                const float2 u = Ul[min(b, KMAX / 2 - 1)][pl];
                const float bp = (k & 1) == 0 ? (odd ? u.y : u.x) : (odd ? u.x : -u.y);
                const float v = b < zm.my ? bp : 0.0f;
                const elem h = (elem)v;
                ahi[it][s][e] = h;
                if constexpr (MODE == 1) alo[it][s][e] = (elem)(v - (float)h);
            }
        }
    }
    __syncthreads();  // the transpose buffers reuse the U_z columns

    const uint16_t* Yf = Yfrag + zm.yfrag_off + (int64_t)lane * 8;
    // Unconditional Y loads and a loop unrolled by two (see main_f32).
    auto yload = [&](int yt, vec (&h)[S], vec (&l)[S]) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const uint16_t* p = Yf + ((int64_t)yt * K32 + min(s, K32 - 1)) * 512;
            h[s] = *reinterpret_cast<const vec*>(p);
            if constexpr (MODE == 1) l[s] = *reinterpret_cast<const vec*>(p + ylo_delta);
        }
    };
    auto tile = [&](int yt, const vec (&h)[S], const vec (&l)[S]) {
        floatx4 acc[4];
#pragma unroll
        for (int it = 0; it < 4; ++it) acc[it] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (s >= K32) continue;  // uniform; k-slices past the region's contribute zero
            if constexpr (MODE == 1) {
#pragma unroll
                for (int it = 0; it < 4; ++it) acc[it] = mfma16<MODE>(alo[it][s], h[s], acc[it]);
#pragma unroll
                for (int it = 0; it < 4; ++it) acc[it] = mfma16<MODE>(ahi[it][s], l[s], acc[it]);
            }
#pragma unroll
            for (int it = 0; it < 4; ++it) acc[it] = mfma16<MODE>(ahi[it][s], h[s], acc[it]);
        }
        tile_out<VEC>(T[yt & 1], acc, oz, ncols, ny, yt, (int64_t)bx * 256, zm.acc != 0);
    };
    vec h0[S], l0[S], h1[S], l1[S];
    yload(yt0, h0, l0);
    yload(min(yt0 + 1, yt1 - 1), h1, l1);
    tile(yt0, h0, l0);
    yload(min(yt0 + 2, yt1 - 1), h0, l0);
    if (yt0 + 1 >= yt1) return;
    tile(yt0 + 1, h1, l1);
    yload(min(yt0 + 3, yt1 - 1), h1, l1);
    int yt = yt0 + 2;
    for (; yt + 1 < yt1; yt += 2) {
        tile(yt, h0, l0);
        yload(min(yt + 2, yt1 - 1), h0, l0);
        tile(yt + 1, h1, l1);
        yload(min(yt + 3, yt1 - 1), h1, l1);
    }
    if (yt < yt1) tile(yt, h0, l0);
}


template <int KMAX, bool VEC, int MODE>
__global__ __launch_bounds__(256) void main_16(const ZMeta* __restrict__ zmeta, const RegionDev* __restrict__ regions,
                                               const float2* __restrict__ Cs, const float2* __restrict__ X,
                                               const uint16_t* __restrict__ Yfrag, int64_t ylo_delta, int nx,
                                               int ny, int ytiles_per_split, float* __restrict__ out, int zoff) {
    // logical block (column block bx, z, y split) in XCD-contiguous order: the
    // column blocks of one z and the z of one region share an XCD's L2 for
    // C_z, X and the Y fragments
    int bx, by, bz;
    xcd_block3(bx, by, bz);
    __shared__ __attribute__((aligned(16))) unsigned char lds[Main16Lds<KMAX>::kBytes];
    main_16_body<KMAX, VEC, MODE, false>(bx, by + zoff, bz, zmeta, regions, Cs, X, Yfrag, ylo_delta, nx, ny,
                                         ytiles_per_split, out, lds);
}

template <int KMAX, bool VEC>
static void launch_main(const scw_exec* ex, dim3 grid, hipStream_t st, int ytps, float* out, int zoff) {
    const ZMeta* zm = ex->d_mksort.as<ZMeta>();  // main-GEMM entries [zoff, zoff + grid.y), all of kernel K KMAX
    const float2* U = ex->d_U.as<float2>();
    const int nx = (int)ex->nx, ny = (int)ex->ny;
    if (ex->dtype == SCW_DTYPE_FP32) {
        hipLaunchKernelGGL((main_f32<KMAX, VEC>), grid, dim3(256), 0, st, zm, U, ex->d_Yfrag.as<float>(), nx, ny,
                           ytps, out, zoff);
        return;
    }
    if constexpr (KMAX >= 32) {
        const uint16_t* Y = ex->d_Yfrag.as<uint16_t>();
        const RegionDev* rg = ex->d_regions.as<RegionDev>();
        const float2* X = ex->d_X.as<float2>();
        if (ex->dtype == SCW_DTYPE_BF16)
            hipLaunchKernelGGL((main_16<KMAX, VEC, 0>), grid, dim3(256), 0, st, zm, rg, U, X, Y, ex->ylo_delta, nx,
                               ny, ytps, out, zoff);
        else if (ex->dtype == SCW_DTYPE_BF16X3)
            hipLaunchKernelGGL((main_16<KMAX, VEC, 1>), grid, dim3(256), 0, st, zm, rg, U, X, Y, ex->ylo_delta, nx,
                               ny, ytps, out, zoff);
        else
            hipLaunchKernelGGL((main_16<KMAX, VEC, 2>), grid, dim3(256), 0, st, zm, rg, U, X, Y, ex->ylo_delta, nx,
                               ny, ytps, out, zoff);
    }
}

// ---------------------------------------------------------------------------
// Block floating-point amplitudes for the 16-bit dtypes.  Group q of a dense
// region holds modes 4q .. 4q + 3 of every column m = 3 cell + comp; record
// (q, m) is the 8 values v[e], e = 2 h + (re|im) of mode 4q + h, as 16-bit
// signed mantissas, and the six values of one mode at one cell (3 components,
// re and im) share one exponent byte: v[e] = mant[e] 2^(E[h] - 128), E[h]
// chosen so that the largest of the six has 15 significant bits.  Every value
// then has at most 15 significant bits and splits exactly into bf16 hi (its
// top 16 bits) + bf16 lo (the rest); its error is at most 2^-15 of the
// largest |re| or |im| of its term (half a step of the 15-bit scale), so each
// term keeps its own relative precision whatever the magnitudes of the other
// modes (an exponent shared across modes loses it for weak modes whose phase
// factor is large: rel-L2 3.4e-4 on the 64-slab batch).  Planes: mantissas
// [q][arow] x 16 bytes (a 16-lane load of 16 columns reads 256 contiguous
// bytes), then exponents [q][erow] x 4 bytes (erow = cells, padded: arow / 3
// rounded up).  13 bytes per term (3 complex values) against 24 for FP32
// values.  Modes past the width and columns past 3 cells are zero.
// ---------------------------------------------------------------------------
constexpr int kBfpBias = 128;

__global__ __launch_bounds__(256) void bfp_encode(const float2* __restrict__ A, int64_t arow, int64_t erow, int width,
                                                  int mcols, int64_t groups, uint32x4* __restrict__ mant,
                                                  uint32_t* __restrict__ expo) {
    const int64_t n = groups * erow;
    for (int64_t i = blockIdx.x * 256LL + tid_x(); i < n; i += (int64_t)gridDim.x * 256) {
        const int64_t q = i / erow, cell = i - q * erow;
        float2 a[4][3];
        int s[4];
        uint32_t ex = 0;
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const int64_t c = 4 * q + h;
            float mx = 0.0f;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int64_t m = 3 * cell + k;
                a[h][k] = (c < width && m < mcols) ? A[c * arow + m] : make_float2(0.0f, 0.0f);
                mx = fmaxf(mx, fmaxf(fabsf(a[h][k].x), fabsf(a[h][k].y)));
            }
            // scale 2^s with mx 2^-s in [2^14, 2^15) (one step up if it rounds
            // to 2^15), kept a normal float; a zero term takes s = 0
            s[h] = 0;
            if (mx > 0.0f) {
                s[h] = max(ilogbf(mx) - 14, -126);
                if (rintf(ldexpf(mx, -s[h])) > 32767.0f) ++s[h];
            }
            ex |= (uint32_t)(s[h] + kBfpBias) << (8 * h);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int64_t m = 3 * cell + k;
            if (m >= arow) break;  // the last cell of a padded row
            uint32x4 w;
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const int re = (int)rintf(ldexpf(a[h][k].x, -s[h])), im = (int)rintf(ldexpf(a[h][k].y, -s[h]));
                w[h] = ((uint32_t)re & 0xffffu) | ((uint32_t)im << 16);
            }
            mant[q * arow + m] = w;
        }
        expo[i] = ex;
    }
}

// Value k (0..7) of a block floating-point record: mantissa words w, scale of
// its mode.
__device__ inline float bfp_value(const uint32_t (&w)[4], float scale, int k) {
    const uint32_t x = w[k >> 1];
    const int mant = (k & 1) ? ((int)x >> 16) : (int)(int16_t)(x & 0xffffu);
    return (float)mant * scale;
}
// 2^(E - 128) from an exponent byte E (2 <= E <= 255 by construction).
__device__ inline float bfp_scale(uint32_t e) { return __uint_as_float((e - (kBfpBias - 127)) << 23); }

// One lane's record: the 4 mantissa words and the exponent word.
typedef uint32_t RawBfp[5];
// Record (q, column): mantissas mrow[column] and exponents erow[column / 3]
// of mode group q's rows mrow = mantissas + q arow, erow = exponents + q erow.
template <bool NT>
__device__ inline void bfp_load(RawBfp& r, const uint32x4* mrow, const uint32_t* erow, uint32_t col, uint32_t cell) {
    const uint32x4* mp = mrow + col;
    const uint32x4 w = NT ? stream_load(mp) : gload(mp);
#pragma unroll
    for (int k = 0; k < 4; ++k) r[k] = w[k];
    const uint32_t* ep = erow + cell;
    r[4] = NT ? stream_load(ep) : gload(ep);
}
// 16-bit MFMA operands of a record: MODE 1 the exact hi/lo split, MODE 0 / 2
// the value rounded to bf16 / fp16.
template <int MODE>
__device__ inline void bfp_decode(const RawBfp& r, typename Op16<MODE>::vec& bh, typename Op16<MODE>::vec& bl) {
    typedef typename Op16<MODE>::elem elem;
    const uint32_t w[4] = {r[0], r[1], r[2], r[3]};
#pragma unroll
    for (int e = 0; e < 8; ++e) {  // e = 2 h + (re|im)
        const float v = bfp_value(w, bfp_scale((r[4] >> (8 * (e >> 1))) & 0xffu), e);
        if constexpr (MODE == 1) {
            // exact split: hi = the top 16 bits, lo = the remaining (at most 7) significant bits
            const uint32_t bits = __float_as_uint(v);
            bh[e] = __builtin_bit_cast(elem, (uint16_t)(bits >> 16));
            bl[e] = (elem)(v - __uint_as_float(bits & 0xffff0000u));
        } else {
            bh[e] = (elem)v;
        }
    }
}

int build_bfp(scw_plan* P) {
    // execs of one plan may be created from several threads
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    if (!P->bfp_off.empty()) return SCW_OK;
    std::vector<int64_t> off(P->regions.size(), -1);
    int64_t bytes = 0;
    for (size_t r = 0; r < P->regions.size(); ++r) {
        const RegionPlan& R = P->regions[r];
        if (!R.separable || !R.dense || R.width == 0) continue;
        off[r] = bytes;
        const int64_t groups = 4LL * ((R.width + 15) / 16);
        bytes += groups * (R.arow * 16 + bfp_erow(R.arow) * 4);
        bytes = (bytes + 255) / 256 * 256;  // next region's mantissas 256-byte aligned
    }
    int st;
    if ((st = P->d_bfp.alloc((size_t)std::max<int64_t>(bytes, 4))) != SCW_OK) return st;
    for (size_t r = 0; r < P->regions.size(); ++r) {
        if (off[r] < 0) continue;
        const RegionPlan& R = P->regions[r];
        const int64_t groups = 4LL * ((R.width + 15) / 16);
        const int64_t n = groups * bfp_erow(R.arow);  // one thread per (mode group, cell)
        uint8_t* base = P->d_bfp.as<uint8_t>() + off[r];
        const dim3 g((unsigned)std::min<int64_t>((n + 255) / 256, 65536));
        hipLaunchKernelGGL(bfp_encode, g, dim3(256), 0, nullptr, P->d_ellA.as<float2>() + R.ell_off, R.arow,
                           bfp_erow(R.arow), R.width, (int)(3 * (int64_t)R.mx * R.my), groups,
                           reinterpret_cast<uint32x4*>(base), reinterpret_cast<uint32_t*>(base + groups * R.arow * 16));
        SCW_HIP(hipGetLastError());
    }
    SCW_HIP(hipDeviceSynchronize());
    P->bfp_off = off;
    return SCW_OK;
}

// Exponent plane of a dense region with mode 4 q + h scaled by 2^shift[4 q + h]
// (dst[i] = src[i] with each byte E -> clamp(E + shift, 1, 255); byte 1 reads
// as scale 0).  The FP16 mode contraction moves part of each mode's
// magnitude from its phase factor into its amplitudes this way.
__global__ __launch_bounds__(256) void bfp_shift_exp(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                                     int64_t erow, int64_t n, const int32_t* __restrict__ shift) {
    for (int64_t i = blockIdx.x * 256LL + tid_x(); i < n; i += (int64_t)gridDim.x * 256) {
        const int64_t q = i / erow;
        const uint32_t w = src[i];
        uint32_t o = 0;
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const int e = min(max((int)((w >> (8 * h)) & 0xffu) + shift[4 * q + h], 1), 255);
            o |= (uint32_t)e << (8 * h);
        }
        dst[i] = o;
    }
}

int shift_bfp_exponents(const uint8_t* src, uint8_t* dst, int64_t erow, int64_t n, const int32_t* shift) {
    const dim3 g((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536)));
    hipLaunchKernelGGL(bfp_shift_exp, g, dim3(256), 0, nullptr, reinterpret_cast<const uint32_t*>(src),
                       reinterpret_cast<uint32_t*>(dst), erow, n, shift);
    SCW_HIP(hipGetLastError());
    SCW_HIP(hipDeviceSynchronize());
    return SCW_OK;
}

// ---------------------------------------------------------------------------
// MFMA mode contraction for dense eigenmode regions (every cell holds every
// kz; the term index is the kz index c):
//   C_z[m] = sum_c Z_z[c] A[c][m],   m = complex column (cell, comp)
// as a real GEMM D[i][m] = sum_k Z'[i][k] A'[k][m] with rows i = (z, re|im)
// and k = (c, re|im).  A' is the amplitude table as stored (float2 re, im per
// (c, m), read once per run in 128-byte row segments); the complex product
// lives in the small operand Z' (signed parts of Z, precomputed per run in
// the MFMA A-operand lane layout, split into 16-bit high and low parts for
// the 16-bit modes).  Rows: 16 per tile = 8 z; the wave owns 64 columns
// (4 tiles) and MT row tiles, and walks its mode range in steps of 16 modes:
//   MODE 0..2: v_mfma_f32_16x16x32_{bf16,bf16,f16}, lane k = 8 g + e
//              <-> (c = 4 g + e / 2, part e % 2), g = lane / 16
//   MODE 3:    v_mfma_f32_16x16x4_f32, MFMA q = 0..7, lane k = g
//              <-> (c = g + 4 (q / 2), part q % 2)
// Either way lane (j, g) reads the float2 A[c][m0 + 16 nt + j] of 4 modes.
// ---------------------------------------------------------------------------
// Partial sums of a mode-contraction item into its slot of C: lane holds
// D[i = 4 g + r][m0 + 16 nt + j]: rows 4g..4g+3 are (z, re), (z, im),
// (z + 1, re), (z + 1, im) with z = 8 mt + 2 g.
// An item of a run with a single slot (I.cs_direct) stores into the
// slot-summed C_z instead ([3 my][ka] rows, row 3 b + comp for cell b mx + a),
// which no slot sum then copies.
template <int MT, bool COH>
__device__ inline void modal_store(const floatx4 (&acc)[4][MT], const ModalItem& I, const RegionDev& R,
                                   const ZMeta* __restrict__ zmeta, int64_t c_elems, float2* __restrict__ C,
                                   float2* __restrict__ Cs, int mcols) {
    const int lane = tid_x() & 63, j = lane & 15, g = lane >> 4;
    if (I.cs_direct) {  // uniform
        int off[4];  // Cs_z element of each of the lane's columns
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const int m = min(I.m0 + 16 * nt + j, mcols - 1), cell = m / 3, b = cell / R.mx;
            off[nt] = (3 * b + (m - 3 * cell)) * R.ka + (cell - b * R.mx);
        }
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int zi = 8 * mt + 2 * g + h;
                if (zi >= I.nz) continue;
                float2* o = Cs + zmeta[I.z0 + zi].u_off;
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
                    if (I.m0 + 16 * nt + j < mcols)
                        put2<COH>(o + off[nt], make_float2(acc[nt][mt][2 * h], acc[nt][mt][2 * h + 1]));
            }
        return;
    }
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int zi = 8 * mt + 2 * g + h;
            if (zi >= I.nz) continue;
            float2* o = C + (int64_t)I.slot * c_elems + zmeta[I.z0 + zi].c_off;
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const int m = I.m0 + 16 * nt + j;
                if (m < mcols) put2<COH>(o + m, make_float2(acc[nt][mt][2 * h], acc[nt][mt][2 * h + 1]));
            }
        }
}

template <int MODE, int MT, bool COH>
__device__ inline void modal_body(const ModalItem& I, const RegionDev& R, const ZMeta* __restrict__ zmeta,
                                  const uint32_t* __restrict__ zf, int64_t c_elems, float2* __restrict__ C,
                                  float2* __restrict__ Cs) {
    const int lane = tid_x() & 63, j = lane & 15, g = lane >> 4;
    const int mcols = 3 * R.cells;
    const int W = R.width;
    const float2* A = R.ellA;
    int col[4], ecol[4];  // column and its cell (block floating-point exponent word)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        col[nt] = min(I.m0 + 16 * nt + j, mcols - 1);
        ecol[nt] = col[nt] / 3;
    }
    floatx4 acc[4][MT];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) acc[nt][mt] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};

    // Raw operands of one 16-mode step.  Loads are unconditional: FP32 modes
    // past the width re-read the last mode (their Z' entries are zero), the
    // block floating-point table is zero-padded to whole steps, and columns
    // past the table re-read the last column (not stored).  FP32 (MODE 3)
    // reads the float2 amplitudes; the 16-bit modes read the block
    // floating-point table: lane (j, g) of step t takes record (group 4 t + g,
    // column), i.e. modes 16 t + 4 g .. + 3.
    typedef float2 RawA[4][4];   // MODE 3: [nt][h] float2
    typedef RawBfp RawB[4];      // 16-bit modes: [nt] record
    typedef typename std::conditional<MODE == 3, RawA, RawB>::type RawT;
    typedef uint32x4 RawZ[MT][2];
    const uint32_t* zb = zf + I.zf_off + lane * 8;
    auto load = [&](RawT& ra, RawZ& rz, int t) {
        if constexpr (MODE == 3) {
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const int c = min(16 * t + g + 4 * h, W - 1);
                const float2* row = A + (int64_t)c * R.arow;
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) ra[nt][h] = MT >= 4 ? stream_load(row + col[nt]) : gload(row + col[nt]);
            }
        } else {
            const uint32x4* mrow = reinterpret_cast<const uint32x4*>(R.bfp) + (int64_t)(4 * t + g) * R.arow;
            const uint32_t* erow = reinterpret_cast<const uint32_t*>(R.bfp_exp) + (int64_t)(4 * t + g) * R.erow;
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) bfp_load<(MT >= 4)>(ra[nt], mrow, erow, col[nt], ecol[nt]);
        }
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const uint32_t* p = zb + ((int64_t)mt * I.T + t) * 512;
            rz[mt][0] = *reinterpret_cast<const uint32x4*>(p);
            if constexpr (MODE == 1 || MODE == 3) rz[mt][1] = *reinterpret_cast<const uint32x4*>(p + 4);
        }
    };
    auto step = [&](const RawT& ra, const RawZ& rz) {
        if constexpr (MODE == 3) {
#pragma unroll
            for (int q = 0; q < 8; ++q)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) {
                    const float2 u = ra[nt][q >> 1];
                    const float bv = (q & 1) ? u.y : u.x;
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt) {
                        const uint32_t w = rz[mt][q >> 2][q & 3];  // element copy first: bit-casting a
                        const float av = __builtin_bit_cast(float, w);  // vector element reads element 0
                        acc[nt][mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[nt][mt], 0, 0, 0);
                    }
                }
        } else {
            typedef typename Op16<MODE>::vec vec;
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                vec bh, bl;
                bfp_decode<MODE>(ra[nt], bh, bl);
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) {
                    const vec zh = __builtin_bit_cast(vec, rz[mt][0]);
                    if constexpr (MODE == 1) {
                        const vec zl = __builtin_bit_cast(vec, rz[mt][1]);
                        acc[nt][mt] = mfma16<MODE>(zh, bl, acc[nt][mt]);
                        acc[nt][mt] = mfma16<MODE>(zl, bh, acc[nt][mt]);
                    }
                    acc[nt][mt] = mfma16<MODE>(zh, bh, acc[nt][mt]);
                }
            }
        }
    };
    // Two named buffers, the next step's loads in flight while the current
    // step computes; prefetch indices are clamped to the last step.
    const int t0 = I.c0 >> 4, t1 = I.c1 >> 4;
    RawT a0, a1;
    RawZ z0, z1;
    load(a0, z0, t0);
    for (int t = t0;;) {
        load(a1, z1, min(t + 1, t1 - 1));
        step(a0, z0);
        if (++t == t1) break;
        load(a0, z0, min(t + 1, t1 - 1));
        step(a1, z1);
        if (++t == t1) break;
    }
    modal_store<MT, COH>(acc, I, R, zmeta, c_elems, C, Cs, mcols);
}

// Runs of 64 z (MT = 8 row tiles) with the 16-bit modes.  All four waves of
// a workgroup hold items of the same run and mode slot (the host pads each
// run's column chunks to whole workgroups with items past the table, which
// store nothing), so they need the same phase fragments: each step's 16 KB of
// fragments is loaded once per workgroup, a quarter per wave, into
// double-buffered LDS and read from there by every wave.  That takes three
// quarters of the fragment traffic off the vector-memory return path, which
// the fragments would otherwise occupy longer than the amplitudes.  One
// barrier per step: a buffer is refilled only after every wave has passed the
// barrier that follows its last read.  Amplitude loads run two steps ahead
// (each step's raw records are decoded once into hi/lo operands for all four
// column tiles, and their registers take the load of the step after next).
template <int MODE>
struct ModalWgLds {  // the double-buffered phase fragments of a step
    static constexpr int MT = 8, NH = MODE == 1 ? 2 : 1;
    static constexpr int kBytes = 2 * MT * NH * 64 * (int)sizeof(uint32x4);
};

template <int MODE, bool COH>
__device__ inline void modal_body_wg(const ModalItem& I, const ModalItem& L, const RegionDev& R,
                                     const ZMeta* __restrict__ zmeta, const uint32_t* __restrict__ zf,
                                     int64_t c_elems, float2* __restrict__ C, float2* __restrict__ Cs,
                                     unsigned char* lds) {
    static_assert(MODE != 3, "16-bit modes");
    if (L.nz == 0) return;  // an empty workgroup (uniform: every wave reads the same leader)
    constexpr int MT = 8, NH = MODE == 1 ? 2 : 1;  // fragment halves: high (and low)
    typedef typename Op16<MODE>::vec vec;
    uint32x4 (*zs)[MT][NH][64] = reinterpret_cast<uint32x4 (*)[MT][NH][64]>(lds);
    const int lane = tid_x() & 63, j = lane & 15, g = lane >> 4;
    const int wv = __builtin_amdgcn_readfirstlane(tid_x() >> 6);
    const int mcols = 3 * R.cells;
    int col[4], ecol[4];  // column and its cell (block floating-point exponent word)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        col[nt] = min(I.m0 + 16 * nt + j, mcols - 1);
        ecol[nt] = col[nt] / 3;
    }
    floatx4 acc[4][MT];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) acc[nt][mt] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};

    typedef RawBfp RawB[4];
    // Unconditional loads, as in modal_body: the table is zero-padded to whole
    // steps, columns past the table re-read the last column, and prefetches
    // past the last step re-read it.
    auto loadA = [&](RawB& ra, int t) {
        const uint32x4* mrow = reinterpret_cast<const uint32x4*>(R.bfp) + (int64_t)(4 * t + g) * R.arow;
        const uint32_t* erow = reinterpret_cast<const uint32_t*>(R.bfp_exp) + (int64_t)(4 * t + g) * R.erow;
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) bfp_load<true>(ra[nt], mrow, erow, col[nt], ecol[nt]);
    };
    // This wave's quarter of a step's fragments: row tiles 2 wv, 2 wv + 1.
    // The steps and fragments are the workgroup leader's L (the same run and
    // slot as this wave's item I; I supplies the columns).
    constexpr int RW = 2;
    typedef uint32x4 Stage[RW][NH];
    const uint32_t* zb = zf + L.zf_off + lane * 8;
    auto loadZ = [&](Stage& st, int t) {
#pragma unroll
        for (int u = 0; u < RW; ++u) {
            const uint32_t* p = zb + ((int64_t)(RW * wv + u) * L.T + t) * 512;
#pragma unroll
            for (int h = 0; h < NH; ++h) st[u][h] = *reinterpret_cast<const uint32x4*>(p + 4 * h);
        }
    };
    auto storeZ = [&](const Stage& st, int buf) {
#pragma unroll
        for (int u = 0; u < RW; ++u)
#pragma unroll
            for (int h = 0; h < NH; ++h) zs[buf][RW * wv + u][h][lane] = st[u][h];
    };
    const int t0 = L.c0 >> 4, t1 = L.c1 >> 4;
    Stage stg;
    loadZ(stg, t0);
    RawB a0, a1;
    loadA(a0, t0);
    loadA(a1, min(t0 + 1, t1 - 1));
    storeZ(stg, 0);
    __syncthreads();
    auto step = [&](RawB& ra, int t, int buf) {
        vec bh[4], bl[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) bfp_decode<MODE>(ra[nt], bh[nt], bl[nt]);
        // next step's fragments first, then the amplitudes of the step after
        // next: the fragments are waited for at the end of this step, the
        // amplitudes only two steps later
        loadZ(stg, min(t + 1, t1 - 1));
        loadA(ra, min(t + 2, t1 - 1));
        // keep the loads ahead of the MFMAs: left to itself the scheduler
        // sinks them to the end of the step, where the fragment wait before
        // the barrier then exposes their whole latency
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const vec zh = __builtin_bit_cast(vec, zs[buf][mt][0][lane]);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                if constexpr (MODE == 1) {
                    const vec zl = __builtin_bit_cast(vec, zs[buf][mt][NH - 1][lane]);
                    acc[nt][mt] = mfma16<MODE>(zh, bl[nt], acc[nt][mt]);
                    acc[nt][mt] = mfma16<MODE>(zl, bh[nt], acc[nt][mt]);
                }
                acc[nt][mt] = mfma16<MODE>(zh, bh[nt], acc[nt][mt]);
            }
        }
        storeZ(stg, buf ^ 1);
        __syncthreads();
    };
    for (int t = t0;;) {
        step(a0, t, 0);
        if (++t == t1) break;
        step(a1, t, 1);
        if (++t == t1) break;
    }
    modal_store<MT, COH>(acc, I, R, zmeta, c_elems, C, Cs, mcols);
}

template <int MODE, int MT>
__global__ __launch_bounds__(256) void modal_kernel(const ModalItem* __restrict__ items, int nitems,
                                                    const RegionDev* __restrict__ regions,
                                                    const ZMeta* __restrict__ zmeta, const uint32_t* __restrict__ zf,
                                                    int64_t c_elems, float2* __restrict__ C, float2* __restrict__ Cs) {
    // XCD-contiguous order for the short-run class only: the FP32 64-z class
    // measured slower with it (64-slab batch, 3.65 -> 4.29 ms)
    const int lb = MT == 1 ? xcd_block() : (int)blockIdx.x;
    const int it = __builtin_amdgcn_readfirstlane(lb * 4 + (int)(tid_x() >> 6));
    if (it >= nitems) return;
    const ModalItem I = items[it];
    modal_body<MODE, MT, false>(I, regions[I.region], zmeta, zf, c_elems, C, Cs);
}

// The 64-z contraction in a ping-pong schedule: a workgroup of eight waves,
// two per SIMD, all running the same loop of two halves per step (decode the
// step's amplitudes and issue the next loads | barrier | the step's MFMAs and
// the next fragments into LDS | barrier), with waves 0-3 one half behind
// waves 4-7 (an extra barrier before the loop, and one after it for waves
// 4-7).  While one wave of a SIMD runs its MFMAs, the other decodes.  The
// phase fragments stay double-buffered: the buffer a wave refills after its
// MFMAs of step t was last read by the other half's MFMAs of step t - 1, one
// half earlier.
template <int MODE, bool COH>
__device__ inline void modal_body_wg_pp(const ModalItem& I, const ModalItem& L, const RegionDev& R,
                                        const ZMeta* __restrict__ zmeta, const uint32_t* __restrict__ zf,
                                        int64_t c_elems, float2* __restrict__ C, float2* __restrict__ Cs,
                                        unsigned char* lds) {
    static_assert(MODE != 3, "16-bit modes");
    if (L.nz == 0) return;  // an empty workgroup (uniform: every wave reads the same leader)
    constexpr int MT = 8, NH = MODE == 1 ? 2 : 1;
    typedef typename Op16<MODE>::vec vec;
    uint32x4 (*zs)[MT][NH][64] = reinterpret_cast<uint32x4 (*)[MT][NH][64]>(lds);
    const int lane = tid_x() & 63, j = lane & 15, g = lane >> 4;
    const int wv = __builtin_amdgcn_readfirstlane(tid_x() >> 6);
    const int mcols = 3 * R.cells;
    int col[4], ecol[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        col[nt] = min(I.m0 + 16 * nt + j, mcols - 1);
        ecol[nt] = col[nt] / 3;
    }
    floatx4 acc[4][MT];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) acc[nt][mt] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    typedef RawBfp RawB[4];
    auto loadA = [&](RawB& ra, int t) {
        const uint32x4* mrow = reinterpret_cast<const uint32x4*>(R.bfp) + (int64_t)(4 * t + g) * R.arow;
        const uint32_t* erow = reinterpret_cast<const uint32_t*>(R.bfp_exp) + (int64_t)(4 * t + g) * R.erow;
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) bfp_load<true>(ra[nt], mrow, erow, col[nt], ecol[nt]);
    };
    // this wave's eighth of a step's fragments: row tile wv
    typedef uint32x4 Stage[NH];
    const uint32_t* zb = zf + L.zf_off + lane * 8 + (int64_t)wv * L.T * 512;
    auto loadZ = [&](Stage& st, int t) {
#pragma unroll
        for (int h = 0; h < NH; ++h) st[h] = *reinterpret_cast<const uint32x4*>(zb + (int64_t)t * 512 + 4 * h);
    };
    auto storeZ = [&](const Stage& st, int buf) {
#pragma unroll
        for (int h = 0; h < NH; ++h) zs[buf][wv][h][lane] = st[h];
    };
    const int t0 = L.c0 >> 4, t1 = L.c1 >> 4;
    Stage stg;
    loadZ(stg, t0);
    RawB a0, a1;
    loadA(a0, t0);
    loadA(a1, min(t0 + 1, t1 - 1));
    storeZ(stg, 0);
    __syncthreads();
    if (wv < 4) __syncthreads();  // waves 0-3 trail by half a step
    auto step = [&](RawB& ra, int t, int buf) {
        vec bh[4], bl[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) bfp_decode<MODE>(ra[nt], bh[nt], bl[nt]);
        loadZ(stg, min(t + 1, t1 - 1));
        loadA(ra, min(t + 2, t1 - 1));
        __syncthreads();
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const vec zh = __builtin_bit_cast(vec, zs[buf][mt][0][lane]);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                if constexpr (MODE == 1) {
                    const vec zl = __builtin_bit_cast(vec, zs[buf][mt][NH - 1][lane]);
                    acc[nt][mt] = mfma16<MODE>(zh, bl[nt], acc[nt][mt]);
                    acc[nt][mt] = mfma16<MODE>(zl, bh[nt], acc[nt][mt]);
                }
                acc[nt][mt] = mfma16<MODE>(zh, bh[nt], acc[nt][mt]);
            }
        }
        storeZ(stg, buf ^ 1);
        __syncthreads();
    };
    for (int t = t0;;) {
        step(a0, t, 0);
        if (++t == t1) break;
        step(a1, t, 1);
        if (++t == t1) break;
    }
    if (wv >= 4) __syncthreads();  // the barrier count of waves 0-3
    modal_store<MT, COH>(acc, I, R, zmeta, c_elems, C, Cs, mcols);
}

// The 64-z class of the 16-bit modes, two waves per SIMD: workgroups of NW
// waves (NW items), two workgroups of 4 per CU (modal_body_wg) or one of 8
// in the ping-pong schedule (modal_body_wg_pp).
template <int MODE, int NW>
__global__ __launch_bounds__(64 * NW, 8 / NW) void modal_wg_kernel(const ModalItem* __restrict__ items, int nitems,
                                                                  const RegionDev* __restrict__ regions,
                                                                  const ZMeta* __restrict__ zmeta,
                                                                  const uint32_t* __restrict__ zf, int64_t c_elems,
                                                                  float2* __restrict__ C, float2* __restrict__ Cs) {
    // Whole workgroups only (the host pads this class to them), and the step
    // range comes from the workgroup's first item, so every wave passes the
    // same barriers.
    const int b0 = (int)blockIdx.x * NW;
    if (b0 + NW - 1 >= nitems || items[b0].nz == 0) return;  // nz = 0: empty workgroup
    const int it = __builtin_amdgcn_readfirstlane(b0 + (int)(tid_x() >> 6));
    const ModalItem I = items[it];
    __shared__ __attribute__((aligned(16))) unsigned char lds[ModalWgLds<MODE>::kBytes];
    if constexpr (NW == 8)
        modal_body_wg_pp<MODE, false>(I, items[b0], regions[I.region], zmeta, zf, c_elems, C, Cs, lds);
    else
        modal_body_wg<MODE, false>(I, items[b0], regions[I.region], zmeta, zf, c_elems, C, Cs, lds);
}

template <int MODE>
static void launch_modal(const scw_exec* ex, hipStream_t st) {
    const ModalItem* mi = ex->d_mitems.as<ModalItem>();
    const RegionDev* rg = ex->d_regions.as<RegionDev>();
    const ZMeta* zm = ex->d_zmeta.as<ZMeta>();
    const uint32_t* zf = ex->d_Zf.as<uint32_t>();
    float2* C = ex->d_C.as<float2>();
    float2* Cs = ex->d_U.as<float2>();  // slot-summed C_z (16-bit dtypes), for runs stored in place
    for (int k = 0; k < 4; ++k) {
        const int64_t off = ex->mclass_off[k], n = ex->mclass_off[k + 1] - off;
        if (n <= 0) continue;
        const dim3 g((unsigned)((n + 3) / 4));
        switch (k) {
            case 0: hipLaunchKernelGGL((modal_kernel<MODE, 1>), g, dim3(256), 0, st, mi + off, (int)n, rg, zm, zf, ex->c_elems, C, Cs); break;
            case 1: hipLaunchKernelGGL((modal_kernel<MODE, 2>), g, dim3(256), 0, st, mi + off, (int)n, rg, zm, zf, ex->c_elems, C, Cs); break;
            case 2: hipLaunchKernelGGL((modal_kernel<MODE, 4>), g, dim3(256), 0, st, mi + off, (int)n, rg, zm, zf, ex->c_elems, C, Cs); break;
            default:
                if constexpr (MODE == 3)
                    hipLaunchKernelGGL((modal_kernel<MODE, 8>), g, dim3(256), 0, st, mi + off, (int)n, rg, zm, zf, ex->c_elems, C, Cs);
                else if (ex->modal_wg_items == 8)
                    hipLaunchKernelGGL((modal_wg_kernel<MODE, 8>), dim3((unsigned)(n / 8)), dim3(512), 0, st, mi + off,
                                       (int)n, rg, zm, zf, ex->c_elems, C, Cs);
                else
                    hipLaunchKernelGGL((modal_wg_kernel<MODE, 4>), g, dim3(256), 0, st, mi + off, (int)n, rg, zm, zf, ex->c_elems, C, Cs);
                break;
        }
    }
}

// ---------------------------------------------------------------------------
// Single-launch form of the 16-bit dtypes: one persistent kernel runs every
// stage.  Its work units (a table built on the host, in dependency order) are
//   coefficient units: four wave items of the VALU contraction or of one
//     MFMA mode-contraction class (a 64-z workgroup for the 64-z class);
//   slot-sum units: one slot-sum workgroup (z, 256 elements of C_z);
//   main units: one main-GEMM workgroup (column block, z, y split).
// Workgroups take units in table order from a device counter (the first one
// is blockIdx.x), so a unit only ever waits for units taken before it, by
// workgroups that are running: the slot sum of z waits until every wave item
// covering z has raised done[z] to need[z]; a main unit of z until the slot
// sums of z have raised csd[z] to their count.  Data handed between units
// goes through sc1 accesses (put2 / get2 with COH), and a producer raises its
// counter only after its stores have completed.  Every wait is bounded by a
// wall-clock limit; one that expires sets the error word (the unit then runs
// on whatever it reads, so the launch still ends) for the host to check.  The
// last workgroup to finish resets the counters, so the next launch (or graph
// replay) starts from zero without a memset.
// ---------------------------------------------------------------------------
enum : int { kUnitCoef = 0, kUnitModal1 = 1, kUnitModal2 = 2, kUnitModal4 = 3, kUnitModalWg = 4, kUnitCsum = 5, kUnitMain = 6 };
constexpr int kCtlHead = 0, kCtlExit = 1, kCtlErr = 2, kCtlWords = 4;  // then done[nz], csd[nz]

struct FusedArgs {
    const int4* units;  // (kind, first item | cblk index | column block, item end | z, y split)
    int nunits;
    int nz;
    const CoefItem* items;
    const ModalItem* mitems;
    const RegionDev* regions;
    const ZMeta* zmeta;
    const ZMeta* mzmeta;  // main-GEMM entries (one ky chunk per region: one per z)
    const float2* Z;
    const uint32_t* zf;
    int64_t c_elems;
    float2* C;
    float2* Cs;
    const int2* cblk;
    const float2* X;
    const uint16_t* Yfrag;
    int64_t ylo_delta;
    int nx, ny, ytiles_per_split;
    float* out;
    int* ctl;          // kCtlWords control words, then done[nz], csd[nz]
    const int* need;   // need[nz] wave items per z, then the slot-sum units per z
    uint64_t wait_ticks;  // wall-clock ticks a wait may last
    uint64_t wait_polls;  // and its most polls
    int modal_wg_items;   // items per 64-z mode-contraction unit (4 or 8)
};

__device__ inline int load_count(const int* p) {
    return __hip_atomic_load(const_cast<int*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void raise_count(int* p) { __hip_atomic_fetch_add(p, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Waits (one thread) until *p >= target, at most `ticks` of the wall clock
// and at most `polls` polls (a bound that holds whatever the clock does).
__device__ inline bool wait_count(const int* p, int target, uint64_t ticks, uint64_t polls) {
    if (load_count(p) >= target) return true;
    const uint64_t t0 = wall_clock64();
    for (uint64_t n = 0; n < polls; ++n) {
        __builtin_amdgcn_s_sleep(2);
        if (load_count(p) >= target) return true;
        if (wall_clock64() - t0 > ticks) return false;
    }
    return false;
}

// A wave item's z run is complete: its stores first, then done[z] per z.
__device__ inline void items_done(int* done, int z0, int nz) {
    stores_done();
    const int lane = tid_x() & 63;
    if (lane < nz) raise_count(done + z0 + lane);
}

template <int KMAX, int MODE>
struct FusedLds {
    static constexpr int kA = Main16Lds<KMAX>::kBytes, kB = ModalWgLds<MODE>::kBytes;
    static constexpr int kBytes = kA > kB ? kA : kB;
};

template <int KMAX, bool VEC, int MODE>
__global__ __launch_bounds__(256, 2) void fused16(const FusedArgs a) {
    __shared__ __attribute__((aligned(16))) unsigned char lds[FusedLds<KMAX, MODE>::kBytes];
    __shared__ int s_unit[2], s_last;
    int* const done = a.ctl + kCtlWords;
    int* const csd = done + a.nz;
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    int next = (int)blockIdx.x;  // thread 0: the unit taken for the next iteration
    for (int it = 0;; ++it) {
        // Thread 0 waits for what its unit needs; the barrier hands the unit
        // to the other waves (double-buffered, so one barrier per unit
        // suffices).
        if (threadIdx.x == 0) {
            const int u = next;
            if (u < a.nunits) {
                const int4 U = a.units[u];
                const int* p = nullptr;
                int target = 0;
                if (U.x == kUnitCsum) {
                    const int z = U.z > 0 ? a.cblk[U.y].x : U.w;
                    p = done + z;
                    target = a.need[z];
                } else if ((U.x & 0xff) == kUnitMain) {
                    p = csd + U.z;
                    target = a.need[a.nz + U.z];
                }
                if (p && !wait_count(p, target, a.wait_ticks, a.wait_polls))
                    __hip_atomic_fetch_or(a.ctl + kCtlErr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            s_unit[it & 1] = u;
        }
        __syncthreads();
        const int u = __builtin_amdgcn_readfirstlane(s_unit[it & 1]);
        if (u >= a.nunits) break;
        // take the next unit now: the counter's round trip overlaps this one
        if (threadIdx.x == 0) next = (int)gridDim.x + atomicAdd(a.ctl + kCtlHead, 1);
        const int4 U = a.units[u];
        switch (U.x & 0xff) {
            case kUnitCoef: {
                const int i = U.y + wv;
                if (i < U.z) {
                    const CoefItem I = a.items[i];
                    coef_body<true>(I, a.regions, a.zmeta, a.Z, a.c_elems, a.C);
                    items_done(done, I.z0, I.nz);
                }
                break;
            }
            case kUnitModal1:
            case kUnitModal2:
            case kUnitModal4: {
                const int i = U.y + wv;
                if (i < U.z) {
                    const ModalItem I = a.mitems[i];
                    const RegionDev& R = a.regions[I.region];
                    if (U.x == kUnitModal1) modal_body<MODE, 1, true>(I, R, a.zmeta, a.zf, a.c_elems, a.C, a.Cs);
                    else if (U.x == kUnitModal2) modal_body<MODE, 2, true>(I, R, a.zmeta, a.zf, a.c_elems, a.C, a.Cs);
                    else modal_body<MODE, 4, true>(I, R, a.zmeta, a.zf, a.c_elems, a.C, a.Cs);
                    items_done(done, I.z0, I.nz);
                }
                break;
            }
            case kUnitModalWg:  // a.modal_wg_items items of one run and slot, four at a time
                for (int h = 0; h < a.modal_wg_items; h += 4) {
                    const ModalItem I = a.mitems[U.y + h + wv];
                    modal_body_wg<MODE, true>(I, a.mitems[U.y + h], a.regions[I.region], a.zmeta, a.zf, a.c_elems,
                                              a.C, a.Cs, lds);
                    items_done(done, I.z0, I.nz);
                }
                break;
            case kUnitCsum: {  // U.z consecutive slot-sum blocks of one z (none: z = U.w, stored in place)
                const int z = U.z > 0 ? a.cblk[U.y].x : U.w;
                const ZMeta zm = a.zmeta[z];
                const RegionDev& R = a.regions[zm.region];
                for (int k = 0; k < U.z; ++k) {
                    const int e = a.cblk[U.y + k].y + tid_x();
                    if (e < 3 * zm.my * R.ka)
                        put2<true>(a.Cs + zm.u_off + e, csum_elem<true>(zm, R, a.C, a.c_elems, a.Z, e));
                }
                stores_done();
                __syncthreads();
                if (threadIdx.x == 0) raise_count(csd + z);
                break;
            }
            default:  // U.x >> 8 consecutive column blocks from U.y of z-slice U.z, y split U.w
                for (int t = 0; t < (U.x >> 8); ++t) {
                    if (t > 0) __syncthreads();  // the previous block's last transpose is read
                    main_16_body<KMAX, VEC, MODE, true>(U.y + t, U.z, U.w, a.mzmeta, a.regions, a.Cs, a.X,
                                                        a.Yfrag, a.ylo_delta, a.nx, a.ny, a.ytiles_per_split, a.out,
                                                        lds);
                }
                break;
        }
    }
    // Every workgroup ends here once the units are taken; the last one
    // resets the counters for the next launch.  Each wave first waits for its
    // counter increments (issued without a return value, still in flight
    // otherwise), so no increment can land after the reset.
    stores_done();
    __syncthreads();
    if (threadIdx.x == 0) s_last = atomicAdd(a.ctl + kCtlExit, 1) == (int)gridDim.x - 1;
    __syncthreads();
    if (s_last) {
        for (int i = (int)threadIdx.x; i < 2 * a.nz; i += 256)
            __hip_atomic_store(done + i, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (threadIdx.x == 0) {
            __hip_atomic_store(a.ctl + kCtlHead, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(a.ctl + kCtlExit, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

template <int KMAX, bool VEC, int MODE>
static int fused_kernel_occupancy(int* per_cu) {
    SCW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, fused16<KMAX, VEC, MODE>, 256, 0));
    return SCW_OK;
}

static FusedArgs fused_args(const scw_exec* ex, float* out) {
    FusedArgs a;
    a.units = ex->d_units.as<int4>();
    a.nunits = (int)ex->nunits;
    a.nz = (int)ex->nz;
    a.items = ex->d_items.as<CoefItem>();
    a.mitems = ex->d_mitems.as<ModalItem>();
    a.regions = ex->d_regions.as<RegionDev>();
    a.zmeta = ex->d_zmeta.as<ZMeta>();
    a.mzmeta = ex->d_mzmeta.as<ZMeta>();
    a.Z = ex->d_Z.as<float2>();
    a.zf = ex->d_Zf.as<uint32_t>();
    a.c_elems = ex->c_elems;
    a.C = ex->d_C.as<float2>();
    a.Cs = ex->d_U.as<float2>();
    a.cblk = ex->d_cblk.as<int2>();
    a.X = ex->d_X.as<float2>();
    a.Yfrag = ex->d_Yfrag.as<uint16_t>();
    a.ylo_delta = ex->ylo_delta;
    a.nx = (int)ex->nx;
    a.ny = (int)ex->ny;
    a.ytiles_per_split = ex->ytiles_per_split;
    a.out = out;
    a.ctl = ex->d_ctl.as<int>();
    a.need = ex->d_need.as<int>();
    a.wait_ticks = ex->wait_ticks;
    a.wait_polls = ex->wait_polls;
    a.modal_wg_items = ex->modal_wg_items;
    return a;
}

template <int MODE>
static int dispatch_fused(const scw_exec* ex, float* out, hipStream_t st, int* per_cu) {
    const bool vec = (6 * ex->nx) % 4 == 0;
#define SCW_FUSED(K)                                                                                        \
    if (per_cu) return vec ? fused_kernel_occupancy<K, true, MODE>(per_cu) : fused_kernel_occupancy<K, false, MODE>(per_cu); \
    if (vec) hipLaunchKernelGGL((fused16<K, true, MODE>), dim3((unsigned)ex->pgrid), dim3(256), 0, st, fused_args(ex, out)); \
    else hipLaunchKernelGGL((fused16<K, false, MODE>), dim3((unsigned)ex->pgrid), dim3(256), 0, st, fused_args(ex, out)); \
    break;
    switch (ex->kmax) {
        case 32: SCW_FUSED(32)
        case 64: SCW_FUSED(64)
        case 96: SCW_FUSED(96)
        default: SCW_FUSED(128)
    }
#undef SCW_FUSED
    SCW_HIP(hipGetLastError());
    return SCW_OK;
}

// Resident workgroups per CU of the single-launch kernel for this exec, or
// the launch itself (per_cu == nullptr).
int fused_launch(const scw_exec* ex, float* out, hipStream_t st, int* per_cu) {
    switch (ex->dtype) {
        case SCW_DTYPE_BF16: return dispatch_fused<0>(ex, out, st, per_cu);
        case SCW_DTYPE_BF16X3: return dispatch_fused<1>(ex, out, st, per_cu);
        case SCW_DTYPE_FP16: return dispatch_fused<2>(ex, out, st, per_cu);
        default: return fail(SCW_ERR_UNSUPPORTED, "single-launch form: 16-bit dtypes only");
    }
}

// Stages: 1 = coefficient contraction (VALU ELL items and MFMA mode items),
// 2 = X contraction, 3 = main MFMA GEMM; 0 = all three.
int launch_gemm(const scw_exec* ex, float* out, hipStream_t st, int stage) {
    if (stage == 0 && ex->persistent) {
        TraceRange tr("gemm single launch");
        return fused_launch(ex, out, st, nullptr);
    }
    const int nitems = (int)ex->items.size();
    const int nmitems = (int)ex->mitems.size();
    if (stage == 0 || stage == 1) {
        TraceRange tr("gemm stage 1: coefficient contraction");
        if (nitems > 0) {
            hipLaunchKernelGGL(coef_kernel, dim3((unsigned)((nitems + 3) / 4)), dim3(256), 0, st,
                               ex->d_items.as<CoefItem>(), nitems, ex->d_regions.as<RegionDev>(),
                               ex->d_zmeta.as<ZMeta>(), ex->d_Z.as<float2>(), ex->c_elems, ex->d_C.as<float2>());
            SCW_HIP(hipGetLastError());
        }
        if (nmitems > 0) {
            if (ex->dtype == SCW_DTYPE_FP32) launch_modal<3>(ex, st);
            else if (ex->dtype == SCW_DTYPE_BF16) launch_modal<0>(ex, st);
            else if (ex->dtype == SCW_DTYPE_BF16X3) launch_modal<1>(ex, st);
            else launch_modal<2>(ex, st);
            SCW_HIP(hipGetLastError());
        }
    }
    if (ex->my_max > 0 && (stage == 0 || stage == 2)) {
        TraceRange tr("gemm stage 2: X contraction / slot sum");
        const dim3 g((unsigned)((ex->nx + 255) / 256), (unsigned)((ex->my_max + UB - 1) / UB), (unsigned)ex->nz);
        if (ex->dtype == SCW_DTYPE_FP32) {
            if (ex->ka_max <= 128)
                hipLaunchKernelGGL(u_mfma<130>, g, dim3(256), 0, st, ex->d_zmeta.as<ZMeta>(),
                                   ex->d_regions.as<RegionDev>(), ex->d_C.as<float2>(), ex->c_elems,
                                   ex->d_X.as<float2>(), (int)ex->nx, ex->d_U.as<float2>(), 0);
            else
                hipLaunchKernelGGL(u_mfma<258>, g, dim3(256), 0, st, ex->d_zmeta.as<ZMeta>(),
                                   ex->d_regions.as<RegionDev>(), ex->d_C.as<float2>(), ex->c_elems,
                                   ex->d_X.as<float2>(), (int)ex->nx, ex->d_U.as<float2>(), 0);
        }
        else if (ex->ncblk > 0)
            hipLaunchKernelGGL(csum_kernel, dim3((unsigned)ex->ncblk), dim3(256), 0, st,
                               ex->d_zmeta.as<ZMeta>(), ex->d_regions.as<RegionDev>(), ex->d_C.as<float2>(),
                               ex->c_elems, ex->d_Z.as<float2>(), ex->d_U.as<float2>(), ex->d_cblk.as<int2>());
        SCW_HIP(hipGetLastError());
    }
    if (stage != 0 && stage != 3) return SCW_OK;
    TraceRange tr("gemm stage 3: main GEMM");
    const int64_t ncols = 6 * ex->nx;
    const int ytps = ex->ytiles_per_split;
    const bool vec = (ncols % 4) == 0;
    // per ky-chunk level (a level adds to what the previous wrote) one launch
    // per kernel K, over that K's entries
    for (size_t i = 0; i + 2 < ex->mk_launch.size(); i += 3) {
        const int zoff = (int)ex->mk_launch[i];
        const dim3 grid((unsigned)((ncols + 255) / 256), (unsigned)ex->mk_launch[i + 1], (unsigned)ex->ysplit);
#define SCW_MAIN(K)                                                        \
    if (vec) launch_main<K, true>(ex, grid, st, ytps, out, zoff);          \
    else launch_main<K, false>(ex, grid, st, ytps, out, zoff);
        if (ex->dtype == SCW_DTYPE_FP32) {
            switch (ex->mk_launch[i + 2]) {
                case 16: SCW_MAIN(16) break;
                case 32: SCW_MAIN(32) break;
                case 64: SCW_MAIN(64) break;
                default: SCW_MAIN(128) break;
            }
        } else {
            switch (ex->mk_launch[i + 2]) {
                case 32: SCW_MAIN(32) break;
                case 64: SCW_MAIN(64) break;
                case 96: SCW_MAIN(96) break;
                default: SCW_MAIN(128) break;
            }
        }
#undef SCW_MAIN
        SCW_HIP(hipGetLastError());
    }
    return SCW_OK;
}

}  // namespace scw
