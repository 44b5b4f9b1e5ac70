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
// Three launches per grid:
//   coef_kernel : C_z for every z of the grid (ELL contraction over the
//                 region's terms; wave-level work items, FP32 VALU).
//   u_kernel    : U_z[b][x][comp] = sum_a C_z[a,b,comp] X[a][x]  (FP32 VALU).
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

namespace scw {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef uint32_t uint32x4 __attribute__((ext_vector_type(4)));

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
                                 const int32_t* __restrict__ Cb, int64_t stride, const float2* __restrict__ Z,
                                 const int64_t (&zrow)[8], int j0, int j1) {
    int j = j0;
    for (; j + 4 <= j1; j += 4) {
        float2 a[4][3];
        int c[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float2* p = Ab + (int64_t)(j + u) * stride * 3;
            a[u][0] = p[0];
            a[u][1] = p[1];
            a[u][2] = p[2];
            c[u] = DENSE ? j + u : Cb[(int64_t)(j + u) * stride];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            float2 f[NZ];
#pragma unroll
            for (int zi = 0; zi < NZ; ++zi) f[zi] = Z[zrow[zi] + c[u]];
#pragma unroll
            for (int zi = 0; zi < NZ; ++zi) cmac6(acc[zi], a[u][0], a[u][1], a[u][2], f[zi]);
        }
    }
    for (; j < j1; ++j) {
        const float2* p = Ab + (int64_t)j * stride * 3;
        const float2 a0 = p[0], a1 = p[1], a2 = p[2];
        const int c = DENSE ? j : Cb[(int64_t)j * stride];
#pragma unroll
        for (int zi = 0; zi < NZ; ++zi) cmac6(acc[zi], a0, a1, a2, Z[zrow[zi] + c]);
    }
}

__global__ __launch_bounds__(256) void coef_kernel(const CoefItem* __restrict__ items, int nitems,
                                                   const RegionDev* __restrict__ regions,
                                                   const ZMeta* __restrict__ zmeta, const float2* __restrict__ Z,
                                                   int64_t c_elems, float2* __restrict__ C) {
    const int lane = threadIdx.x & 63;
    const int it = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
    if (it >= nitems) return;
    const CoefItem I = items[it];
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
    const int64_t stride = (int64_t)R.cells;
#define SCW_COEF_CASE(N)                                                                   \
    case N:                                                                                \
        if (R.dense) coef_loop<N, true>(acc, Ab, Cb, stride, Z, zrow, I.j0, I.j1);        \
        else coef_loop<N, false>(acc, Ab, Cb, stride, Z, zrow, I.j0, I.j1);               \
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
            if (R.dense) coef_loop<8, true>(acc, Ab, Cb, stride, Z, zrow, I.j0, I.j1);
            else coef_loop<8, false>(acc, Ab, Cb, stride, Z, zrow, I.j0, I.j1);
            break;
    }
#undef SCW_COEF_CASE
    if (!active) return;
#pragma unroll
    for (int zi = 0; zi < 8; ++zi) {
        if (zi < nz) {
            float2* o = C + (int64_t)I.slot * c_elems + zmeta[I.z0 + zi].c_off + (int64_t)cell * 3;
            o[0] = make_float2(acc[zi][0], acc[zi][1]);
            o[1] = make_float2(acc[zi][2], acc[zi][3]);
            o[2] = make_float2(acc[zi][4], acc[zi][5]);
        }
    }
}

// ---------------------------------------------------------------------------
// U_z[b][x][comp] = sum_a (sum_slots C_z[a,b,comp]) X[a][x]
// The block (z, b) first adds the partial-sum slots of C_z[:, b, :] into LDS
// (fixed slot order), then each thread contracts one x against them.
// ---------------------------------------------------------------------------
constexpr int MX_MAX = 128;

__global__ __launch_bounds__(256) void u_kernel(const ZMeta* __restrict__ zmeta, const RegionDev* __restrict__ regions,
                                                const float2* __restrict__ C, int64_t c_elems,
                                                const float2* __restrict__ X, int nx, float2* __restrict__ U, int zoff) {
    __shared__ float2 sc[MX_MAX * 3];
    __shared__ float2 part[256];
    const int z = blockIdx.z + zoff, b = blockIdx.y;
    const ZMeta zm = zmeta[z];
    if (b >= zm.my) return;
    const RegionDev R = regions[zm.region];
    // Sum the partial-sum slots of C_z[:, b, :]: entry t = (a, comp) of m = 3 mx
    // entries; when m < 128 the threads form G = 256 / m groups that each add
    // every G-th slot, and the groups are then added in group order.
    const int m = R.mx * 3;
    const int G = m < 128 ? 256 / m : 1;
    for (int t0 = 0; t0 < m; t0 += 256) {
        const int t = t0 + (int)threadIdx.x % (G > 1 ? m : 256);
        const int g = G > 1 ? (int)threadIdx.x / m : 0;
        float2 s = make_float2(0.0f, 0.0f);
        if (t < m && g < G) {
            const int64_t off = zm.c_off + ((int64_t)b * R.mx + t / 3) * 3 + (t % 3);
            for (int k = g; k < R.nslots; k += G) {
                const float2 v = C[(int64_t)k * c_elems + off];
                s.x += v.x;
                s.y += v.y;
            }
        }
        if (G > 1) {
            if (g < G) part[threadIdx.x] = s;
            __syncthreads();
            if ((int)threadIdx.x < m) {
                float2 tot = part[threadIdx.x];
                for (int q = 1; q < G; ++q) {
                    const float2 v = part[q * m + threadIdx.x];
                    tot.x += v.x;
                    tot.y += v.y;
                }
                sc[threadIdx.x] = tot;
            }
        } else if (t < m) {
            sc[t] = s;
        }
    }
    __syncthreads();
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x >= nx) return;
    float acc[6] = {0, 0, 0, 0, 0, 0};
    const float2* Xr = X + R.x_off + x;
    for (int a = 0; a < R.mx; ++a)
        cmac6(acc, sc[3 * a], sc[3 * a + 1], sc[3 * a + 2], Xr[(int64_t)a * nx]);
    float2* o = U + zm.u_off + ((int64_t)b * nx + x) * 3;
    o[0] = make_float2(acc[0], acc[1]);
    o[1] = make_float2(acc[2], acc[3]);
    o[2] = make_float2(acc[4], acc[5]);
}

// Value of the main GEMM's B operand Bp[k][n] read from U row b = k >> 1.
__device__ inline float bp_value(const float* __restrict__ Ub, int64_t ncols, int k, int64_t n) {
    const float2 u = *reinterpret_cast<const float2*>(Ub + (int64_t)(k >> 1) * ncols + (n & ~(int64_t)1));
    const bool odd = (n & 1) != 0;
    return (k & 1) == 0 ? (odd ? u.y : u.x) : (odd ? u.x : -u.y);
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
constexpr int kBufferRsrcWord3 = 0x00020000;  // raw dword buffer, gfx9 family
constexpr int kStoreNT = 2;                   // non-temporal

template <bool VEC>
__device__ inline void row_store(float* orow, bool row_ok, int64_t n, int64_t ncols, const floatx4& v) {
    const uint32_t nrec = row_ok ? (uint32_t)(ncols * 4) : 0u;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(orow, 0, (int)nrec, kBufferRsrcWord3);
    const int off = (int)(n * 4);
    if (VEC) {
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(uint32x4, v), rs, off, 0, kStoreNT);
    } else {
        // Each element gets its own offset, moved far out of range past the
        // row's end, so a partly valid group of 4 columns is written element
        // by element and the stores are never merged into one 16-byte access.
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float f = v[r];
            const int off_r = n + r < ncols ? off + 4 * r : 0x7ffffff0;
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, f), rs, off_r, 0, kStoreNT);
        }
    }
}

template <bool VEC>
__device__ inline void tile_out(float (*T)[TROW], const floatx4 (&acc)[4], float* __restrict__ oz, int64_t ncols,
                                int ny, int yt, int64_t nwg0) {
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
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
        row_store<VEC>(oz + (int64_t)(y0 + 4 * rr) * ncols, y0 + 4 * rr < ny, n, ncols, v[rr]);
}

// Slice of a region with no waves: the workgroup's columns of y tiles
// [yt0, yt1) are zero.
template <bool VEC>
__device__ inline void zero_tiles(float* __restrict__ oz, int64_t ncols, int ny, int yt0, int yt1, int64_t nwg0) {
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int64_t n = nwg0 + 4 * lane;
    const int yend = min(ny, 16 * yt1);
    for (int y = 16 * yt0 + wv; y < yend; y += 4)
        row_store<VEC>(oz + (int64_t)y * ncols, true, n, ncols, floatx4{0.0f, 0.0f, 0.0f, 0.0f});
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
    const int z = blockIdx.y + zoff;
    const ZMeta zm = zmeta[z];
    const int K16 = zm.kp >> 4;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t ncols = 6LL * nx;
    const int64_t n0 = (int64_t)blockIdx.x * 256 + wv * 64;  // waves past ncols compute zeros
    const int li = lane & 15, lk = lane >> 4;
    const float* Ub = reinterpret_cast<const float*>(U + zm.u_off);
    const int ytiles = (ny + 15) >> 4;
    const int yt0 = blockIdx.z * ytiles_per_split;
    const int yt1 = min(ytiles, yt0 + ytiles_per_split);
    float* oz = out + (int64_t)z * ny * ncols;
    if (K16 == 0) {
        zero_tiles<VEC>(oz, ncols, ny, yt0, yt1, (int64_t)blockIdx.x * 256);
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
            uf[it][s] = ((s >> 2) < K16 && (k >> 1) < zm.my && n < ncols) ? bp_value(Ub, ncols, k, n) : 0.0f;
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
        tile_out<VEC>(T[yt & 1], acc, oz, ncols, ny, yt, (int64_t)blockIdx.x * 256);
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

template <int KMAX, bool VEC, int MODE>
__global__ __launch_bounds__(256) void main_16(const ZMeta* __restrict__ zmeta, const float2* __restrict__ U,
                                               const uint16_t* __restrict__ Yfrag, int64_t ylo_delta, int nx,
                                               int ny, int ytiles_per_split, float* __restrict__ out, int zoff) {
    typedef typename Op16<MODE>::vec vec;
    typedef typename Op16<MODE>::elem elem;
    constexpr int S = KMAX / 32;
    const int z = blockIdx.y + zoff;
    const ZMeta zm = zmeta[z];
    const int K32 = zm.kp >> 5;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t ncols = 6LL * nx;
    const int64_t n0 = (int64_t)blockIdx.x * 256 + wv * 64;  // waves past ncols compute zeros
    const int li = lane & 15, lk = lane >> 4;
    const float* Ub = reinterpret_cast<const float*>(U + zm.u_off);
    const int ytiles = (ny + 15) >> 4;
    const int yt0 = blockIdx.z * ytiles_per_split;
    const int yt1 = min(ytiles, yt0 + ytiles_per_split);
    float* oz = out + (int64_t)z * ny * ncols;
    if (K32 == 0) {
        zero_tiles<VEC>(oz, ncols, ny, yt0, yt1, (int64_t)blockIdx.x * 256);
        return;
    }
    __shared__ float T[2][16][TROW];

    vec ahi[4][S], alo[4][S];
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        const int64_t n = n0 + 16 * it + li;
#pragma unroll
        for (int s = 0; s < S; ++s) {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = 32 * s + 8 * lk + e;
                const float v = (s < K32 && (k >> 1) < zm.my && n < ncols) ? bp_value(Ub, ncols, k, n) : 0.0f;
                const elem h = (elem)v;
                ahi[it][s][e] = h;
                if constexpr (MODE == 1) alo[it][s][e] = (elem)(v - (float)h);
            }
        }
    }

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
        tile_out<VEC>(T[yt & 1], acc, oz, ncols, ny, yt, (int64_t)blockIdx.x * 256);
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

template <int KMAX, bool VEC>
static void launch_main(const scw_exec* ex, dim3 grid, hipStream_t st, int ytps, float* out, int zoff) {
    const ZMeta* zm = ex->d_zmeta.as<ZMeta>();
    const float2* U = ex->d_U.as<float2>();
    const int nx = (int)ex->nx, ny = (int)ex->ny;
    if (ex->dtype == SCW_DTYPE_FP32) {
        hipLaunchKernelGGL((main_f32<KMAX, VEC>), grid, dim3(256), 0, st, zm, U, ex->d_Yfrag.as<float>(), nx, ny,
                           ytps, out, zoff);
        return;
    }
    if constexpr (KMAX >= 32) {
        const uint16_t* Y = ex->d_Yfrag.as<uint16_t>();
        if (ex->dtype == SCW_DTYPE_BF16)
            hipLaunchKernelGGL((main_16<KMAX, VEC, 0>), grid, dim3(256), 0, st, zm, U, Y, ex->ylo_delta, nx, ny,
                               ytps, out, zoff);
        else if (ex->dtype == SCW_DTYPE_BF16X3)
            hipLaunchKernelGGL((main_16<KMAX, VEC, 1>), grid, dim3(256), 0, st, zm, U, Y, ex->ylo_delta, nx, ny,
                               ytps, out, zoff);
        else
            hipLaunchKernelGGL((main_16<KMAX, VEC, 2>), grid, dim3(256), 0, st, zm, U, Y, ex->ylo_delta, nx, ny,
                               ytps, out, zoff);
    }
}

// ---------------------------------------------------------------------------
// MFMA mode contraction for dense eigenmode regions (16-bit operand modes):
//   C_z[n] = sum_c Z[z][c] A[c][n],   n = real column (cell, comp, re|im)
// as a real GEMM with the complex product folded into K (k = 2c + s):
//   A'[2c][n] = A re|im as stored, A'[2c+1][n] = (i A) re|im,
//   Z'[z][2c] = Re Z, Z'[z][2c+1] = Im Z.
// MFMA rows are n (4 consecutive per lane), columns z (16 per tile, up to 4
// tiles), K 32 per step = 16 modes.  Every term is read once per run.
// ---------------------------------------------------------------------------
template <int MODE, int NZT>
__device__ inline void modal_body(const ModalItem& I, const RegionDev& R, const ZMeta* __restrict__ zmeta,
                                  const float2* __restrict__ Z, int64_t c_elems, float2* __restrict__ C) {
    typedef typename Op16<MODE>::vec vec;
    typedef typename Op16<MODE>::elem elem;
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const int64_t ncols = 6LL * R.cells;
    const float* Ab = reinterpret_cast<const float*>(R.ellA);
    int64_t zr[NZT];
    bool zok[NZT];
#pragma unroll
    for (int zt = 0; zt < NZT; ++zt) {
        const int zz = 16 * zt + li;
        zok[zt] = zz < I.nz;
        zr[zt] = zmeta[I.z0 + min(zz, I.nz - 1)].zrow_off;
    }
    floatx4 acc[4][NZT];
#pragma unroll
    for (int it = 0; it < 4; ++it)
#pragma unroll
        for (int zt = 0; zt < NZT; ++zt) acc[it][zt] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};

    // Raw operands of one 16-mode step; the next step's loads are issued
    // before the current step's conversions and MFMAs.
    float2 ra[4][4], rz[NZT][4];
    auto load_step = [&](int cb) {
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            const int64_t n = I.n0 + 16 * it + li;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int c = cb + 4 * lk + q;
                ra[it][q] = (n < ncols && c < I.c1)
                                ? *reinterpret_cast<const float2*>(Ab + (int64_t)c * ncols + (n & ~(int64_t)1))
                                : make_float2(0.0f, 0.0f);
            }
        }
#pragma unroll
        for (int zt = 0; zt < NZT; ++zt)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int c = cb + 4 * lk + q;
                rz[zt][q] = (zok[zt] && c < I.c1) ? Z[zr[zt] + c] : make_float2(0.0f, 0.0f);
            }
    };
    if (I.c0 < I.c1) load_step(I.c0);
    for (int cb = I.c0; cb < I.c1; cb += 16) {
        vec ah[4], al[4], zh[NZT], zl[NZT];
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            const bool odd = ((I.n0 + 16 * it + li) & 1) != 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float2 u = ra[it][q];
                const float v0 = odd ? u.y : u.x;   // k = 2c:   A re (n even) / A im (n odd)
                const float v1 = odd ? u.x : -u.y;  // k = 2c+1: -A im (n even) / A re (n odd)
                const elem h0 = (elem)v0, h1 = (elem)v1;
                ah[it][2 * q] = h0;
                ah[it][2 * q + 1] = h1;
                if constexpr (MODE == 1) {
                    al[it][2 * q] = (elem)(v0 - (float)h0);
                    al[it][2 * q + 1] = (elem)(v1 - (float)h1);
                }
            }
        }
#pragma unroll
        for (int zt = 0; zt < NZT; ++zt) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float2 f = rz[zt][q];
                const elem h0 = (elem)f.x, h1 = (elem)f.y;
                zh[zt][2 * q] = h0;
                zh[zt][2 * q + 1] = h1;
                if constexpr (MODE == 1) {
                    zl[zt][2 * q] = (elem)(f.x - (float)h0);
                    zl[zt][2 * q + 1] = (elem)(f.y - (float)h1);
                }
            }
        }
        if (cb + 16 < I.c1) load_step(cb + 16);
#pragma unroll
        for (int it = 0; it < 4; ++it)
#pragma unroll
            for (int zt = 0; zt < NZT; ++zt) {
                if constexpr (MODE == 1) {
                    acc[it][zt] = mfma16<MODE>(al[it], zh[zt], acc[it][zt]);
                    acc[it][zt] = mfma16<MODE>(ah[it], zl[zt], acc[it][zt]);
                }
                acc[it][zt] = mfma16<MODE>(ah[it], zh[zt], acc[it][zt]);
            }
    }
    // lane holds D[n = n0 + 16 it + 4 lk + r][z = 16 zt + li]
#pragma unroll
    for (int zt = 0; zt < NZT; ++zt) {
        if (!zok[zt]) continue;
        float* o = reinterpret_cast<float*>(C + (int64_t)I.slot * c_elems + zmeta[I.z0 + 16 * zt + li].c_off);
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            const int64_t n = I.n0 + 16 * it + 4 * lk;
            if (n + 1 < ncols) *reinterpret_cast<float2*>(o + n) = make_float2(acc[it][zt][0], acc[it][zt][1]);
            if (n + 3 < ncols) *reinterpret_cast<float2*>(o + n + 2) = make_float2(acc[it][zt][2], acc[it][zt][3]);
        }
    }
}

template <int MODE>
__global__ __launch_bounds__(256) void modal_kernel(const ModalItem* __restrict__ items, int nitems,
                                                    const RegionDev* __restrict__ regions,
                                                    const ZMeta* __restrict__ zmeta, const float2* __restrict__ Z,
                                                    int64_t c_elems, float2* __restrict__ C) {
    const int it = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
    if (it >= nitems) return;
    const ModalItem I = items[it];
    const RegionDev R = regions[I.region];
    if (I.nz <= 16) modal_body<MODE, 1>(I, R, zmeta, Z, c_elems, C);
    else if (I.nz <= 32) modal_body<MODE, 2>(I, R, zmeta, Z, c_elems, C);
    else modal_body<MODE, 4>(I, R, zmeta, Z, c_elems, C);
}

// Stages: 1 = coefficient contraction (VALU ELL items and MFMA mode items),
// 2 = X contraction, 3 = main MFMA GEMM; 0 = all three.
int launch_gemm(const scw_exec* ex, float* out, hipStream_t st, int stage) {
    const int nitems = (int)ex->items.size();
    const int nmitems = (int)ex->mitems.size();
    if (stage == 0 || stage == 1) {
        if (nitems > 0) {
            hipLaunchKernelGGL(coef_kernel, dim3((unsigned)((nitems + 3) / 4)), dim3(256), 0, st,
                               ex->d_items.as<CoefItem>(), nitems, ex->d_regions.as<RegionDev>(),
                               ex->d_zmeta.as<ZMeta>(), ex->d_Z.as<float2>(), ex->c_elems, ex->d_C.as<float2>());
            SCW_HIP(hipGetLastError());
        }
        if (nmitems > 0) {
            const dim3 g((unsigned)((nmitems + 3) / 4));
            const ModalItem* mi = ex->d_mitems.as<ModalItem>();
            const RegionDev* rg = ex->d_regions.as<RegionDev>();
            const ZMeta* zm = ex->d_zmeta.as<ZMeta>();
            if (ex->dtype == SCW_DTYPE_BF16)
                hipLaunchKernelGGL(modal_kernel<0>, g, dim3(256), 0, st, mi, nmitems, rg, zm, ex->d_Z.as<float2>(),
                                   ex->c_elems, ex->d_C.as<float2>());
            else if (ex->dtype == SCW_DTYPE_BF16X3)
                hipLaunchKernelGGL(modal_kernel<1>, g, dim3(256), 0, st, mi, nmitems, rg, zm, ex->d_Z.as<float2>(),
                                   ex->c_elems, ex->d_C.as<float2>());
            else
                hipLaunchKernelGGL(modal_kernel<2>, g, dim3(256), 0, st, mi, nmitems, rg, zm, ex->d_Z.as<float2>(),
                                   ex->c_elems, ex->d_C.as<float2>());
            SCW_HIP(hipGetLastError());
        }
    }
    if (ex->my_max > 0 && (stage == 0 || stage == 2)) {
        hipLaunchKernelGGL(u_kernel, dim3((unsigned)((ex->nx + 255) / 256), (unsigned)ex->my_max, (unsigned)ex->nz),
                           dim3(256), 0, st, ex->d_zmeta.as<ZMeta>(), ex->d_regions.as<RegionDev>(),
                           ex->d_C.as<float2>(), ex->c_elems, ex->d_X.as<float2>(), (int)ex->nx, ex->d_U.as<float2>(),
                           0);
        SCW_HIP(hipGetLastError());
    }
    if (stage != 0 && stage != 3) return SCW_OK;
    const int64_t ncols = 6 * ex->nx;
    dim3 grid((unsigned)((ncols + 255) / 256), (unsigned)ex->nz, (unsigned)ex->ysplit);
    const int ytps = ex->ytiles_per_split;
    const bool vec = (ncols % 4) == 0;
#define SCW_MAIN(K)                                                        \
    if (vec) launch_main<K, true>(ex, grid, st, ytps, out, 0);             \
    else launch_main<K, false>(ex, grid, st, ytps, out, 0);
    if (ex->dtype == SCW_DTYPE_FP32) {
        switch (ex->kmax) {
            case 16: SCW_MAIN(16) break;
            case 32: SCW_MAIN(32) break;
            case 64: SCW_MAIN(64) break;
            default: SCW_MAIN(128) break;
        }
    } else {
        switch (ex->kmax) {
            case 32: SCW_MAIN(32) break;
            case 64: SCW_MAIN(64) break;
            case 96: SCW_MAIN(96) break;
            default: SCW_MAIN(128) break;
        }
    }
#undef SCW_MAIN
    SCW_HIP(hipGetLastError());
    return SCW_OK;
}

}  // namespace scw
