// Direct form: per-point summation E = sum_w E0_w exp(i k_w . r) over the
// waves of the point's region (gpuEvaluator.cu:28-68, cpuEvaluator.cpp:316-333).
//
// Layout and schedule (gfx950):
//  * blockIdx.y = one z-slice, so the region — and the wave loop bounds — are
//    uniform over the workgroup (the reference's 32x32 xz/yz blocks spanned
//    two z rows per 64-lane wavefront).
//  * Each of the 256 threads owns PPT consecutive points of the slice's
//    flattened (y, x) plane; outputs are written once, AoS, 24 bytes/point.
//  * The region's waves are staged through LDS in chunks of 256.  Staging is
//    done in FP64, once per wave per workgroup: the wave's amplitude at the
//    workgroup's phase origin (xc, yc, z),
//        E0_eff = (E0/s) exp(i k.(xc, yc, z) + ln s),
//    which carries the large part of k.r and any extreme exponent exactly.
//    The per-point remainder exp(i (kx dx + ky dy)) with dx = x - xc,
//    dy = y - yc is done in FP32: the phase in revolutions, reduced with
//    v_fract_f32, then v_sin_f32 / v_cos_f32 (which take revolutions), and
//    the decay with v_exp_f32 (2^x) only in regions whose transverse k has an
//    imaginary part.
//  * blockIdx.z splits a region's wave range for grids too small to fill the
//    GPU; partial sums then go to a scratch buffer and a second kernel adds
//    them in a fixed order.
#include "scw_internal.hpp"

namespace scw {

struct DirectRegion {  // device copy of a region's wave range
    int64_t begin, end;
    int32_t real_transverse, pad;
};

constexpr int DT = 256;      // threads per workgroup
constexpr int CHUNK = 256;   // waves staged per LDS round

template <int PPT>
__global__ __launch_bounds__(DT) void direct_kernel(const double* __restrict__ k64, const double* __restrict__ e64,
                                                    const double* __restrict__ lns, const float4* __restrict__ kf,
                                                    const DirectRegion* __restrict__ regions,
                                                    const int32_t* __restrict__ zreg, const double* __restrict__ zs,
                                                    const float* __restrict__ dxs, const float* __restrict__ dys,
                                                    int nx, int ny, double xc, double yc, int nsplit,
                                                    float* __restrict__ out) {
    __shared__ float4 s_k[CHUNK];    // Re kx/2pi, Re ky/2pi, -Im kx log2e, -Im ky log2e
    __shared__ float4 s_e01[CHUNK];  // E0x, E0y (re, im) at the phase origin
    __shared__ float2 s_e2[CHUNK];   // E0z

    const int z = blockIdx.y;
    const int split = blockIdx.z;
    const int64_t P = (int64_t)nx * ny;
    const DirectRegion R = regions[zreg[z]];
    const double zz = zs[z];

    // Wave range of this split.
    const int64_t nw = R.end - R.begin;
    const int64_t per = (nw + nsplit - 1) / nsplit;
    const int64_t wb = R.begin + (int64_t)split * per;
    const int64_t we = min(R.end, wb + per);

    const int64_t p0 = ((int64_t)blockIdx.x * DT + threadIdx.x) * PPT;
    float dx[PPT], dy[PPT];
    float acc[PPT][6];
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
        const int64_t f = min(p0 + p, P - 1);
        const int iy = (int)(f / nx), ix = (int)(f - (int64_t)iy * nx);
        dx[p] = dxs[ix];
        dy[p] = dys[iy];
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[p][c] = 0.0f;
    }

    const bool decay = R.real_transverse == 0;
    for (int64_t c0 = wb; c0 < we; c0 += CHUNK) {
        const int n = (int)min((int64_t)CHUNK, we - c0);
        __syncthreads();
        if ((int)threadIdx.x < n) {
            const int64_t w = c0 + threadIdx.x;
            const double* k = k64 + 6 * w;
            const double* e = e64 + 6 * w;
            const double th = k[0] * xc + k[2] * yc + k[4] * zz;
            const double ps = lns[w] - (k[1] * xc + k[3] * yc + k[5] * zz);
            const double mag = exp(ps);
            double sn, cs;
            sincos(th, &sn, &cs);
            const double pr = mag * cs, pi = mag * sn;
            s_k[threadIdx.x] = kf[w];
            s_e01[threadIdx.x] = make_float4((float)(e[0] * pr - e[1] * pi), (float)(e[0] * pi + e[1] * pr),
                                             (float)(e[2] * pr - e[3] * pi), (float)(e[2] * pi + e[3] * pr));
            s_e2[threadIdx.x] = make_float2((float)(e[4] * pr - e[5] * pi), (float)(e[4] * pi + e[5] * pr));
        }
        __syncthreads();
        if (decay) {
            for (int j = 0; j < n; ++j) {
                const float4 kk = s_k[j];
                const float4 a = s_e01[j];
                const float2 b = s_e2[j];
#pragma unroll
                for (int p = 0; p < PPT; ++p) {
                    const float t = __builtin_amdgcn_fractf(fmaf(kk.x, dx[p], kk.y * dy[p]));
                    const float m = __builtin_amdgcn_exp2f(fmaf(kk.z, dx[p], kk.w * dy[p]));
                    const float cr = m * __builtin_amdgcn_cosf(t);
                    const float ci = m * __builtin_amdgcn_sinf(t);
                    acc[p][0] = fmaf(a.x, cr, fmaf(-a.y, ci, acc[p][0]));
                    acc[p][1] = fmaf(a.x, ci, fmaf(a.y, cr, acc[p][1]));
                    acc[p][2] = fmaf(a.z, cr, fmaf(-a.w, ci, acc[p][2]));
                    acc[p][3] = fmaf(a.z, ci, fmaf(a.w, cr, acc[p][3]));
                    acc[p][4] = fmaf(b.x, cr, fmaf(-b.y, ci, acc[p][4]));
                    acc[p][5] = fmaf(b.x, ci, fmaf(b.y, cr, acc[p][5]));
                }
            }
        } else {
            for (int j = 0; j < n; ++j) {
                const float4 kk = s_k[j];
                const float4 a = s_e01[j];
                const float2 b = s_e2[j];
#pragma unroll
                for (int p = 0; p < PPT; ++p) {
                    const float t = __builtin_amdgcn_fractf(fmaf(kk.x, dx[p], kk.y * dy[p]));
                    const float cr = __builtin_amdgcn_cosf(t);
                    const float ci = __builtin_amdgcn_sinf(t);
                    acc[p][0] = fmaf(a.x, cr, fmaf(-a.y, ci, acc[p][0]));
                    acc[p][1] = fmaf(a.x, ci, fmaf(a.y, cr, acc[p][1]));
                    acc[p][2] = fmaf(a.z, cr, fmaf(-a.w, ci, acc[p][2]));
                    acc[p][3] = fmaf(a.z, ci, fmaf(a.w, cr, acc[p][3]));
                    acc[p][4] = fmaf(b.x, cr, fmaf(-b.y, ci, acc[p][4]));
                    acc[p][5] = fmaf(b.x, ci, fmaf(b.y, cr, acc[p][5]));
                }
            }
        }
    }

    // Output: out[split][z][point][6] (split > 0 only with a partial buffer).
    float* o = out + ((int64_t)split * gridDim.y + z) * P * 6;
#pragma unroll
    for (int p = 0; p < PPT; ++p) {
        const int64_t f = p0 + p;
        if (f < P) {
            float2* q = reinterpret_cast<float2*>(o + f * 6);
            q[0] = make_float2(acc[p][0], acc[p][1]);
            q[1] = make_float2(acc[p][2], acc[p][3]);
            q[2] = make_float2(acc[p][4], acc[p][5]);
        }
    }
}

// Adds nsplit partial fields in split order: out[i] = sum_s part[s][i].
__global__ __launch_bounds__(256) void direct_reduce(const float* __restrict__ part, int nsplit, int64_t n,
                                                     float* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int k = 0; k < nsplit; ++k) s += part[(int64_t)k * n + i];
    out[i] = s;
}

int direct_ppt(int64_t pts) { return pts >= 1024 ? 4 : 1; }

// Split the wave range until the grid has about 2048 workgroups, keeping at
// least 1024 waves per split and at most 64 splits.
int direct_split(int64_t pts, int64_t nz, int64_t maxw, int ppt) {
    const int64_t bps = (pts + 256 * ppt - 1) / (256 * ppt);
    const int64_t blocks = bps * nz;
    int split = 1;
    while (blocks * split < 2048 && maxw / (2 * split) >= 1024 && split < 64) split *= 2;
    return split;
}

// One direct-form evaluation of an nx x ny x nz grid from device tables:
// partial sums of the nsplit wave slices go to part (nsplit > 1) and are
// added into out by a second kernel.
int launch_direct_raw(const scw_plan* P, const void* region_table, const int32_t* zreg, const double* zs,
                      const float* dxs, const float* dys, int64_t nx, int64_t ny, int64_t nz, double xc, double yc,
                      int ppt, int nsplit, float* part, float* out, hipStream_t st) {
    const DirectRegion* regions = static_cast<const DirectRegion*>(region_table);
    const int64_t pts = nx * ny;
    const int64_t bps = (pts + (int64_t)DT * ppt - 1) / ((int64_t)DT * ppt);
    float* dst = nsplit > 1 ? part : out;
    dim3 grid((unsigned)bps, (unsigned)nz, (unsigned)nsplit);
    if (ppt == 4)
        hipLaunchKernelGGL(direct_kernel<4>, grid, dim3(DT), 0, st, P->d_k64.as<double>(), P->d_e64.as<double>(),
                           P->d_lns.as<double>(), P->d_kf.as<float4>(), regions, zreg, zs, dxs, dys, (int)nx, (int)ny,
                           xc, yc, nsplit, dst);
    else
        hipLaunchKernelGGL(direct_kernel<1>, grid, dim3(DT), 0, st, P->d_k64.as<double>(), P->d_e64.as<double>(),
                           P->d_lns.as<double>(), P->d_kf.as<float4>(), regions, zreg, zs, dxs, dys, (int)nx, (int)ny,
                           xc, yc, nsplit, dst);
    SCW_HIP(hipGetLastError());
    if (nsplit > 1) {
        const int64_t n = pts * nz * 6;
        hipLaunchKernelGGL(direct_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (const float*)dst,
                           nsplit, n, out);
        SCW_HIP(hipGetLastError());
    }
    return SCW_OK;
}

int launch_direct(const scw_exec* ex, float* out, hipStream_t st) {
    TraceRange tr("direct form");
    return launch_direct_raw(ex->plan, ex->plan->d_dreg.p, ex->d_zslab.as<int32_t>(),
                             ex->d_z.as<double>(), ex->d_dx.as<float>(), ex->d_dy.as<float>(), ex->nx, ex->ny, ex->nz,
                             ex->xc, ex->yc, ex->direct_ppt, ex->direct_split, ex->d_C.as<float>(), out, st);
}

size_t direct_region_bytes() { return sizeof(DirectRegion); }

void direct_region_fill(void* dst, int64_t begin, int64_t end, int real_transverse) {
    DirectRegion* d = static_cast<DirectRegion*>(dst);
    d->begin = begin;
    d->end = end;
    d->real_transverse = real_transverse;
    d->pad = 0;
}

}  // namespace scw

