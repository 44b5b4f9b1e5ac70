// Allocation-free evaluation of the reference's three planes (no counterpart
// in the reference: gpu_cw_evaluate, gpuEvaluator.cu:131-165, allocates and
// frees a device frame on every call).  The per-plane tables an exec would
// upload are built here on the device from the call's scalars, into a fixed
// device array, which also holds the partial sums of the wave-range split:
//   [3 planes][dx float[N] | dy float[N] | z double[N] | region int32[N],
//   padded to 256 bytes] (a plane uses the first nx, ny, nz entries of its
//   tables), then the
//   partial sums of the plane being evaluated (the planes run one after the
//   other on the caller's stream, so they share that space).
// Coordinates follow gpuEvaluator.cu:78-79 (start + d * i fused into one
// multiply-add) and the region of z is 1 + the last l with
// z >= z_layers[l] - tol in double, as scw_exec_create does, so the tables are
// bit-identical to the ones scw_eval_planes uploads, and the wave-range split
// is the same unless the array cannot hold its partial sums.
// Every table entry and partial sum is written before it is read, so nothing
// needs clearing between calls; calls that overlap on one device race for the
// array.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "scw_internal.hpp"

namespace scw {

constexpr int64_t kNoallocMaxN = 16384;
constexpr size_t kNoallocScratchBytes = (size_t)16 << 20;  // 16 MiB per device
constexpr size_t kPlaneTableBytes = 20;                    // per grid line: dx, dy float, z double, region int32

// Bytes of one plane's tables, padded so that every plane starts 256-byte
// aligned (the doubles at 8 N bytes from its start stay aligned for odd N).
__host__ __device__ inline size_t plane_stride(int64_t N) { return (kPlaneTableBytes * (size_t)N + 255) / 256 * 256; }

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

struct PlaneGeom {  // one plane's grid: nx x ny x nz with one extent equal to 1
    int32_t nx, ny, nz, pad;
    float x_start, y_start, z_start, d;  // axis i of extent N: start + d * i (fused)
    float x, y, z, pad2;                 // the fixed coordinate of an extent-1 axis
    double xc, yc;                       // phase origin of the x and y axes
};

__device__ inline float plane_coord(int n, float start, float d, float fixed, int i) {
    return n == 1 ? fixed : fmaf(d, (float)i, start);
}

__global__ __launch_bounds__(256) void plane_tables_kernel(PlaneGeom g0, PlaneGeom g1, PlaneGeom g2, int64_t N,
                                                           const double* __restrict__ zl, int layers, double tol,
                                                           unsigned char* __restrict__ base) {
    const PlaneGeom g = blockIdx.y == 0 ? g0 : (blockIdx.y == 1 ? g1 : g2);
    unsigned char* t = base + (size_t)blockIdx.y * plane_stride(N);
    float* dx = reinterpret_cast<float*>(t);
    float* dy = dx + N;
    double* zs = reinterpret_cast<double*>(dy + N);
    int32_t* zr = reinterpret_cast<int32_t*>(zs + N);
    for (int64_t i = blockIdx.x * 256LL + threadIdx.x; i < N; i += (int64_t)gridDim.x * 256) {
        if (i < g.nx) dx[i] = (float)((double)plane_coord(g.nx, g.x_start, g.d, g.x, (int)i) - g.xc);
        if (i < g.ny) dy[i] = (float)((double)plane_coord(g.ny, g.y_start, g.d, g.y, (int)i) - g.yc);
        if (i < g.nz) {
            const double z = (double)plane_coord(g.nz, g.z_start, g.d, g.z, (int)i);
            int r = 0;
            for (int l = 0; l < layers; ++l)
                if (z >= zl[l] - tol) r = l + 1;
            zs[i] = z;
            zr[i] = r;
        }
    }
}

// Device address of the scratch array on each device, looked up once per
// device with that device current (the array's host-side name is only a
// stand-in for it).
static std::mutex g_scratch_mu;
static std::vector<unsigned char*> g_scratch_ptr;

static int scratch_ptr(int device, unsigned char** ptr) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[device] = static_cast<unsigned char*>(p);
    }
    *ptr = g_scratch_ptr[device];
    return SCW_OK;
}

}  // namespace scw

extern "C" int scw_eval_planes_noalloc(const scw_plan* plan, float x_start, float y_start, float z_start, float x,
                                       float y, float z, float d, int N, double region_tol, void* E_xy, void* E_xz,
                                       void* E_yz, void* stream) {
    using namespace scw;
    if (!plan) return fail(SCW_ERR_ARG, "plan is NULL");
    if (N < 1) return fail(SCW_ERR_ARG, "N = %d (must be at least 1)", N);
    if ((int64_t)N > kNoallocMaxN)
        return fail(SCW_ERR_ARG, "N = %d exceeds the allocation-free maximum of %lld (its per-plane tables and "
                    "partial sums live in a fixed %zu-byte device array)", N, (long long)kNoallocMaxN,
                    kNoallocScratchBytes);
    if (!std::isfinite(x_start) || !std::isfinite(y_start) || !std::isfinite(z_start) || !std::isfinite(x) ||
        !std::isfinite(y) || !std::isfinite(z) || !std::isfinite(d) || !std::isfinite(region_tol))
        return fail(SCW_ERR_ARG, "non-finite coordinate, spacing or region_tol");
    DeviceGuard g(plan->device);
    if (g.st != SCW_OK) return g.st;
    unsigned char* base = nullptr;
    int st = scratch_ptr(plan->device, &base);
    if (st != SCW_OK) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t n = N;
    // the three planes (gpuEvaluator.cu:153-160): xy at z, xz at y, yz at x
    auto axis_center = [&](float start) {
        return 0.5 * ((double)std::fmaf(d, 0.0f, start) + (double)std::fmaf(d, (float)(unsigned)(N - 1), start));
    };
    PlaneGeom geo[3];
    const int ext[3][3] = {{N, N, 1}, {N, 1, N}, {1, N, N}};
    void* outs[3] = {E_xy, E_xz, E_yz};
    for (int p = 0; p < 3; ++p) {
        PlaneGeom& G = geo[p];
        G.nx = ext[p][0];
        G.ny = ext[p][1];
        G.nz = ext[p][2];
        G.pad = 0;
        G.x_start = x_start;
        G.y_start = y_start;
        G.z_start = z_start;
        G.d = d;
        G.x = x;
        G.y = y;
        G.z = z;
        G.pad2 = 0.0f;
        G.xc = G.nx == 1 ? (double)x : axis_center(x_start);
        G.yc = G.ny == 1 ? (double)y : axis_center(y_start);
    }
    const size_t table_bytes = 3 * plane_stride(n);
    float* part = reinterpret_cast<float*>(base + table_bytes);
    const size_t part_bytes = kNoallocScratchBytes - table_bytes;
    hipLaunchKernelGGL(plane_tables_kernel, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 64), 3), dim3(256), 0, s,
                       geo[0], geo[1], geo[2], n, plan->d_zl.as<double>(), plan->layers, region_tol, base);
    SCW_HIP(hipGetLastError());
    for (int p = 0; p < 3; ++p) {
        if (!outs[p]) continue;
        const PlaneGeom& G = geo[p];
        unsigned char* t = base + (size_t)p * plane_stride(n);
        const float* dx = reinterpret_cast<const float*>(t);
        const float* dy = dx + n;
        const double* zs = reinterpret_cast<const double*>(dy + n);
        const int32_t* zr = reinterpret_cast<const int32_t*>(zs + n);
        // the split of an exec of the same plane (its largest region over the
        // plane's z is bounded by the largest region of the structure only
        // when the plane crosses it; take the regions the plane's z meet)
        int64_t maxw = 0;
        for (int i = 0; i < G.nz; ++i) {
            const double zz = (double)(G.nz == 1 ? z : std::fmaf(d, (float)(unsigned)i, z_start));
            int r = 0;
            for (int l = 0; l < plan->layers; ++l)
                if (zz >= plan->zl[l] - region_tol) r = l + 1;
            maxw = std::max(maxw, plan->regions[r].wend - plan->regions[r].wbegin);
        }
        const int64_t pts = (int64_t)G.nx * G.ny;
        const int ppt = direct_ppt(pts);
        int split = direct_split(pts, G.nz, maxw, ppt);
        while (split > 1 && (size_t)split * pts * G.nz * 24 > part_bytes) split /= 2;  // what the array holds
        st = launch_direct_raw(plan, plan->d_dreg.p, zr, zs, dx, dy, G.nx, G.ny, G.nz, G.xc, G.yc,
                               ppt, split, part, static_cast<float*>(outs[p]), s);
        if (st != SCW_OK) return st;
    }
    return SCW_OK;
}
