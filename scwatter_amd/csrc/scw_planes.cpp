// The reference's per-event evaluation, gpu_cw_evaluate (gpuEvaluator.cu:131-165),
// at the C ABI: the xy plane at z, the xz plane at y and the yz plane at x, each
// N x N with the fused coordinates start + d * i (gpuEvaluator.cu:78-79), into
// caller-owned device buffers on the caller's stream.
//
// The reference allocates a device frame, launches and copies back three
// times on every call.  Here each plane is an exec (scw_exec_create), kept in a
// small per-plan cache keyed by everything that defines its tables; a GUI that
// redraws the same view, or moves one plane, rebuilds only the planes whose
// key changed, and once the cache is full a changed plane rebuilds the least
// recently used exec in place, reusing its device buffers where they are
// large enough (no hipMalloc for a plane moved within the same regions).
#include <cmath>

#include "scw_internal.hpp"

namespace scw {

static const size_t kPlaneCacheSize = 6;  // two views' worth of planes

static std::vector<double> fused_axis(float start, float d, int N) {
    std::vector<double> v((size_t)N);
    for (int i = 0; i < N; ++i) v[(size_t)i] = (double)std::fmaf(d, (float)(unsigned)i, start);
    return v;
}

// The exec of one plane, from the cache or newly built (evicting the least
// recently used entry when the cache is full).
static int plane_exec(scw_plan* P, int kind, int N, float a0, float b0, float fixed, float d, double tol, int form,
                      int dtype, scw_exec** out) {
    auto& C = P->plane_cache;
    for (size_t i = 0; i < C.size(); ++i) {
        const auto& e = C[i];
        if (e.kind == kind && e.n == N && e.form == form && e.dtype == dtype && e.start_a == a0 && e.start_b == b0 &&
            e.fixed == fixed && e.d == d && e.tol == tol) {
            auto hit = e;
            C.erase(C.begin() + (long)i);
            C.push_back(hit);  // most recent last
            *out = hit.ex;
            return SCW_OK;
        }
    }
    const std::vector<double> a = fused_axis(a0, d, N), b = fused_axis(b0, d, N), f = {(double)fixed};
    const std::vector<double>* g[3];
    if (kind == 0) g[0] = &a, g[1] = &b, g[2] = &f;       // xy: x, y vary, z fixed
    else if (kind == 1) g[0] = &a, g[1] = &f, g[2] = &b;  // xz: x, z vary, y fixed
    else g[0] = &f, g[1] = &a, g[2] = &b;                 // yz: y, z vary, x fixed
    scw_exec* ex = nullptr;
    int st;
    if (C.size() >= kPlaneCacheSize) {
        // rebuild the least recently used exec in place: its device buffers
        // are reused where they are large enough
        ex = C.front().ex;
        C.erase(C.begin());
        // its last run may still be in flight on the caller's stream
        DeviceGuard dg(P->device);
        if (dg.st != SCW_OK || hipDeviceSynchronize() != hipSuccess) {
            scw_exec_destroy(ex);
            return fail(SCW_ERR_DEVICE, "hipDeviceSynchronize before rebuilding a plane");
        }
        st = exec_init(ex, P, g[0]->data(), g[0]->size(), g[1]->data(), g[1]->size(), g[2]->data(), g[2]->size(),
                       tol, form, dtype);
        if (st != SCW_OK) {
            scw_exec_destroy(ex);
            return st;
        }
    } else {
        st = scw_exec_create(P, g[0]->data(), g[0]->size(), g[1]->data(), g[1]->size(), g[2]->data(), g[2]->size(),
                             tol, form, dtype, &ex);
        if (st != SCW_OK) return st;
    }
    C.push_back({kind, N, form, dtype, a0, b0, fixed, d, tol, ex});
    *out = ex;
    return SCW_OK;
}

}  // namespace scw

using namespace scw;

extern "C" int scw_eval_planes(scw_plan* plan, float x_start, float y_start, float z_start, float x, float y, float z,
                               float d, int N, double region_tol, int form, int dtype, void* E_xy, void* E_xz,
                               void* E_yz, void* stream) {
    TraceRange tr("scw_eval_planes");
    if (!plan) return fail(SCW_ERR_ARG, "plan is NULL");
    if (N < 1 || N > (1 << 16)) return fail(SCW_ERR_ARG, "N = %d out of range [1, 65536]", N);
    if (!std::isfinite(x_start) || !std::isfinite(y_start) || !std::isfinite(z_start) || !std::isfinite(x) ||
        !std::isfinite(y) || !std::isfinite(z) || !std::isfinite(d) || !std::isfinite(region_tol))
        return fail(SCW_ERR_ARG, "non-finite coordinate, spacing or region_tol");
    // (plane, in-plane starts, fixed coordinate): gpuEvaluator.cu:153-160
    const float a0[3] = {x_start, x_start, y_start}, b0[3] = {y_start, z_start, z_start}, fx[3] = {z, y, x};
    void* outs[3] = {E_xy, E_xz, E_yz};
    for (int k = 0; k < 3; ++k) {
        if (!outs[k]) continue;
        scw_exec* ex = nullptr;
        int st;
        try {
            st = plane_exec(plan, k, N, a0[k], b0[k], fx[k], d, region_tol, form, dtype, &ex);
        } catch (const std::bad_alloc&) {
            st = fail(SCW_ERR_NOMEM, "host allocation failed");
        }
        if (st != SCW_OK) return st;
        if ((st = scw_exec_run(ex, outs[k], stream)) != SCW_OK) return st;
    }
    return SCW_OK;
}
