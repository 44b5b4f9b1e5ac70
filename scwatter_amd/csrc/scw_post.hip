// On-device post-processing of evaluated fields, the viewer's per-slice
// pipeline (sCWatterView.cpp:222-352) without moving the complex field to the
// host:
//   scalar  S = E[comp] or S = sum_c E_c conj(E_c)      (EvaluateScalarSlices :299-352)
//   min/max of Re S and Im S over any number of slices (CalculateMinMax :264-296)
//   colour  rgb = 255 (clamp(rn) real_color + clamp(in) imag_color)
//                                                     (EvaluateColorValue :200-220,
//                                                      EvaluateColorSlices :222-262)
// Arithmetic follows the reference's float operations one rounding at a time
// (no contraction into FMAs), so results are bit-identical to it.
#include "scw_internal.hpp"

namespace scw {

// Order-preserving map of a float to uint32 (larger float -> larger key).
__device__ inline uint32_t fkey(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float fkey_inv(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__global__ __launch_bounds__(256) void post_scalar_kernel(const float2* __restrict__ E, int64_t n, int mode,
                                                          float2* __restrict__ S) {
#pragma clang fp contract(off)  // every product and sum rounded on its own, as in the reference
    for (int64_t i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const float2* e = E + 3 * i;
        if (mode < 3) {
            S[i] = e[mode];
        } else {
            // std::complex<float> products v * conj(v) = (a a - b (-b)) + (a (-b) + b a) i,
            // then summed over the components in order
            float re = 0.0f, im = 0.0f;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float a = e[c].x, b = e[c].y;
                const float pr = a * a - b * (-b);
                const float pi = a * (-b) + b * a;
                if (c == 0) {
                    re = pr;
                    im = pi;
                } else {
                    re = re + pr;
                    im = im + pi;
                }
            }
            S[i] = make_float2(re, im);
        }
    }
}

// mm holds 4 keys (min Re, max Re, min Im, max Im) while reducing; floats
// before and after.  CalculateMinMax seeds its four values from S_xy[0] and
// then compares with < and >, so a NaN seed never changes again (every
// comparison with it is false) while a NaN met later is skipped.  Here the
// seed is scalar[0] (or, with accumulate, the previous contents), and a NaN
// seed is kept as the key no data value can reach: 0 for a minimum slot,
// 0xffffffff for a maximum slot (keys of NaNs only, which the reduction skips).
__global__ void post_mm_prepare(uint32_t* mm, const float2* __restrict__ S, int64_t n, int accumulate) {
    const int t = threadIdx.x;
    if (t >= 4) return;
    const bool is_max = (t & 1) != 0;
    const uint32_t init = is_max ? 0u : 0xffffffffu, sticky = is_max ? 0xffffffffu : 0u;
    uint32_t k = init;
    if (accumulate || n > 0) {
        const float seed = accumulate ? __uint_as_float(mm[t]) : (t < 2 ? S[0].x : S[0].y);
        k = isnan(seed) ? sticky : fkey(seed);
    }
    mm[t] = k;
}

__global__ __launch_bounds__(256) void post_mm_reduce(const float2* __restrict__ S, int64_t n, uint32_t* mm) {
    uint32_t k[4] = {0xffffffffu, 0u, 0xffffffffu, 0u};
    for (int64_t i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const float2 s = S[i];
        if (!isnan(s.x)) {
            const uint32_t u = fkey(s.x);
            k[0] = min(k[0], u);
            k[1] = max(k[1], u);
        }
        if (!isnan(s.y)) {
            const uint32_t u = fkey(s.y);
            k[2] = min(k[2], u);
            k[3] = max(k[3], u);
        }
    }
    // wave reduction, then one atomic per wave and value (min/max are exact
    // and order-independent, so the result does not depend on scheduling)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t o = __shfl_xor(k[q], off);
            k[q] = (q & 1) ? max(k[q], o) : min(k[q], o);
        }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&mm[0], k[0]);
        atomicMax(&mm[1], k[1]);
        atomicMin(&mm[2], k[2]);
        atomicMax(&mm[3], k[3]);
    }
}

// Keys back to floats; a NaN seed, or no entry at all, gives the quiet NaN.
__global__ void post_mm_finish(uint32_t* mm) {
    const int t = threadIdx.x;
    if (t >= 4) return;
    const float f = fkey_inv(mm[t]);
    mm[t] = isnan(f) ? 0x7fc00000u : __float_as_uint(f);
}

__device__ inline float clamp01(float v) {  // sCWatterView.cpp:194-199
    if (isnan(v)) return 0.0f;
    if (v < 0.0f) return 0.0f;
    if (v > 1.0f) return 1.0f;
    return v;
}

struct ColorArgs {
    float rc[3], ic[3];
    int show_real, show_imag;
};

__global__ __launch_bounds__(256) void post_color_kernel(const float2* __restrict__ S, int64_t n,
                                                         const float* __restrict__ lh, ColorArgs a,
                                                         uint8_t* __restrict__ rgb) {
#pragma clang fp contract(off)  // every product and sum rounded on its own, as in the reference
    const float real_low = lh[0], real_high = lh[1], imag_low = lh[2], imag_high = lh[3];
    const float r_interval = real_high - real_low;
    const float i_interval = imag_high - imag_low;
    for (int64_t i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const float2 s = S[i];
        const float rn = clamp01((s.x - real_low) / r_interval);
        const float in = clamp01((s.y - imag_low) / i_interval);
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            float c = 0.0f;
            if (a.show_real) c = c + rn * a.rc[ch];
            if (a.show_imag) c = c + in * a.ic[ch];
            const float v = c * 255.0f;
            // float -> unsigned char truncates; values outside [0, 255] (colours
            // adding up past 1, negative or NaN colours) saturate
            rgb[3 * i + ch] = (uint8_t)(v >= 255.0f ? 255.0f : (v > 0.0f ? v : 0.0f));
        }
    }
}

static dim3 post_grid(int64_t n) { return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096))); }

}  // namespace scw

using namespace scw;

extern "C" {

int scw_post_scalar(const void* field_device, size_t n, int mode, void* scalar_device, void* stream) {
    if (n == 0) return SCW_OK;
    if (!field_device || !scalar_device) return fail(SCW_ERR_ARG, "NULL buffer");
    if (mode < 0 || mode > 3) return fail(SCW_ERR_ARG, "mode %d (0 x, 1 y, 2 z, 3 intensity)", mode);
    hipLaunchKernelGGL(post_scalar_kernel, post_grid((int64_t)n), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const float2*>(field_device), (int64_t)n, mode, static_cast<float2*>(scalar_device));
    SCW_HIP(hipGetLastError());
    return SCW_OK;
}

int scw_post_minmax(const void* scalar_device, size_t n, float* minmax_device, int accumulate, void* stream) {
    if (!minmax_device || (n > 0 && !scalar_device)) return fail(SCW_ERR_ARG, "NULL buffer");
    hipStream_t st = static_cast<hipStream_t>(stream);
    uint32_t* mm = reinterpret_cast<uint32_t*>(minmax_device);
    hipLaunchKernelGGL(post_mm_prepare, dim3(1), dim3(64), 0, st, mm, static_cast<const float2*>(scalar_device),
                       (int64_t)n, accumulate ? 1 : 0);
    if (n > 0)
        hipLaunchKernelGGL(post_mm_reduce, post_grid((int64_t)n), dim3(256), 0, st,
                           static_cast<const float2*>(scalar_device), (int64_t)n, mm);
    hipLaunchKernelGGL(post_mm_finish, dim3(1), dim3(64), 0, st, mm);
    SCW_HIP(hipGetLastError());
    return SCW_OK;
}

int scw_post_color(const void* scalar_device, size_t n, const float* lowhigh_device, const float real_color[3],
                   const float imag_color[3], int show_real, int show_imag, void* rgb_device, void* stream) {
    if (n == 0) return SCW_OK;
    if (!scalar_device || !lowhigh_device || !rgb_device || !real_color || !imag_color)
        return fail(SCW_ERR_ARG, "NULL argument");
    ColorArgs a;
    for (int c = 0; c < 3; ++c) {
        a.rc[c] = real_color[c];
        a.ic[c] = imag_color[c];
    }
    a.show_real = show_real ? 1 : 0;
    a.show_imag = show_imag ? 1 : 0;
    hipLaunchKernelGGL(post_color_kernel, post_grid((int64_t)n), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const float2*>(scalar_device), (int64_t)n, lowhigh_device, a,
                       static_cast<uint8_t*>(rgb_device));
    SCW_HIP(hipGetLastError());
    return SCW_OK;
}

}  // extern "C"
