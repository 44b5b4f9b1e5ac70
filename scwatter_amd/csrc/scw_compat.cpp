// The reference's evaluator entry points (cpuEvaluator.h:12-36,
// gpuEvaluator.h:8-18) on top of the C ABI.  Every evaluation runs on the GPU;
// the cpu_* names are kept so existing callers link unchanged, and they keep
// the reference CPU evaluator's semantics: region rule z >= z_l
// (cpuEvaluator.cpp:303) and coordinates start + i*d rounded like its float
// arithmetic (cpuEvaluator.cpp:317-319).  gpu_cw_evaluate keeps the GPU rule
// z >= z_l - 0.001 (gpuEvaluator.cu:34) and its fused start + d*i
// coordinates (gpuEvaluator.cu:78-79).
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/scw_evaluator.hpp"
#include "scw_internal.hpp"

int layers = 0;
std::vector<float> z_layers;
std::vector<int> waves_begin;
std::vector<int> waves_end;
std::vector<UnpackedWave<float>> W;
// The reference viewer defines these (sCWatterView.cpp:102-103); an
// executable that defines them too takes precedence over these.
__attribute__((weak)) size_t free_gpu_memory = 0;
__attribute__((weak)) size_t total_gpu_memory = 0;

namespace {

std::vector<scw_planewave64> g_w64;  // FP64 copy of W (what cw_unpack narrowed)
scw_plan* g_plan = nullptr;
scw::DevBuf g_frame;  // gpu_cw_evaluate's three device planes, kept between calls
int g_frame_dev = -1;  // the device g_frame lives on
int g_status = SCW_OK;
int g_form = SCW_FORM_AUTO;
int g_dtype = SCW_DTYPE_FP32;

int ensure_plan() {
    if (g_plan) return SCW_OK;
    scw::hip_runtime_once();
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    std::vector<double> zl(z_layers.begin(), z_layers.end());  // the float boundaries the reference compares with
    std::vector<int32_t> b(waves_begin.begin(), waves_begin.end()), e(waves_end.begin(), waves_end.end());
    if ((int)b.size() != layers + 1 || (int)e.size() != layers + 1)
        return scw::fail(SCW_ERR_ARG, "cw_allocate / cw_unpack have not been called");
    return scw_plan_create(dev, g_w64.data(), g_w64.size(), zl.data(), layers, b.data(), e.data(), &g_plan);
}

// One plane through the exec API, copied to caller-owned host memory.
int eval_plane(void* out_host, const std::vector<double>& xs, const std::vector<double>& ys,
               const std::vector<double>& zs, double tol) {
    int st = ensure_plan();
    if (st != SCW_OK) return st;
    scw_exec* ex = nullptr;
    st = scw_exec_create(g_plan, xs.data(), xs.size(), ys.data(), ys.size(), zs.data(), zs.size(), tol, g_form,
                         g_form == SCW_FORM_DIRECT ? SCW_DTYPE_FP32 : g_dtype, &ex);
    if (st != SCW_OK) return st;
    st = scw_exec_run_host(ex, out_host, nullptr);
    scw_exec_destroy(ex);
    return st;
}

// cpuEvaluator.cpp:317-319: unsigned index converted to float, multiplied,
// then added — two roundings.
std::vector<double> cpu_axis(float start, float d, size_t N) {
    std::vector<double> v(N);
    for (size_t i = 0; i < N; ++i) {
        volatile float t = (float)(unsigned)i * d;
        v[i] = (double)(float)(start + t);
    }
    return v;
}

}  // namespace

void cw_allocate(CoupledWaveStructure<double>* cw) {
    layers = (int)cw->Layers.size();
    z_layers.resize(layers);
    waves_begin.clear();
    waves_end.clear();
    waves_begin.push_back(0);
    int total = (int)cw->Pi.size();
    for (int l = 0; l < layers; ++l) {
        z_layers[l] = (float)cw->Layers[l].z;
        total += (int)cw->Layers[l].Pr.size();
        waves_end.push_back(total);
        waves_begin.push_back(total);
        total += (int)cw->Layers[l].Pt.size();
    }
    waves_end.push_back(total);
    W.resize(total);
    g_w64.resize(total);
    if (g_plan) {
        scw_plan_destroy(g_plan);
        g_plan = nullptr;
    }
}

void cw_unpack(CoupledWaveStructure<double>* cw) {
    size_t idx = 0;
    auto put = [&](const scw::planewave<double>& p) {
        if (idx >= W.size()) return;
        for (int c = 0; c < 3; ++c) {
            W[idx].E0[c] = std::complex<float>((float)p.E0[c].real(), (float)p.E0[c].imag());
            W[idx].k[c] = std::complex<float>((float)p.k[c].real(), (float)p.k[c].imag());
            g_w64[idx].E0[c][0] = p.E0[c].real();
            g_w64[idx].E0[c][1] = p.E0[c].imag();
            g_w64[idx].k[c][0] = p.k[c].real();
            g_w64[idx].k[c][1] = p.k[c].imag();
        }
        ++idx;
    };
    for (const auto& p : cw->Pi) put(p);
    for (const auto& L : cw->Layers) {
        for (const auto& p : L.Pr) put(p);
        for (const auto& p : L.Pt) put(p);
    }
    if (g_plan) {
        scw_plan_destroy(g_plan);
        g_plan = nullptr;
    }
}

void gpu_initialize() {
    if (g_plan) {
        scw_plan_destroy(g_plan);
        g_plan = nullptr;
    }
    g_status = ensure_plan();
    if (g_status != SCW_OK) fprintf(stderr, "gpu_initialize: %s\n", scw_last_error());
}

void gpu_cw_evaluate(std::complex<float>* E_xy, std::complex<float>* E_xz, std::complex<float>* E_yz, float x_start,
                     float y_start, float z_start, float x, float y, float z, float d, int N, int device) {
    g_status = SCW_OK;
    if (N <= 0 || !E_xy || !E_xz || !E_yz) {
        g_status = scw::fail(SCW_ERR_ARG, "gpu_cw_evaluate: invalid arguments");
        return;
    }
    if (g_plan && g_plan->device != device) {  // the reference ignores `device`; this honours it
        scw_plan_destroy(g_plan);
        g_plan = nullptr;
    }
    if (!g_plan) {
        int cur = -1;
        scw::hip_runtime_once();
        (void)hipGetDevice(&cur);
        if (hipSetDevice(device) != hipSuccess) {
            g_status = scw::fail(SCW_ERR_ARG, "gpu_cw_evaluate: device %d unavailable", device);
            return;
        }
        g_status = ensure_plan();
        if (cur >= 0) (void)hipSetDevice(cur);
        if (g_status != SCW_OK) return;
    }
    const size_t n = (size_t)N;
    if (scw_device_mem_info(device, &free_gpu_memory, &total_gpu_memory) != SCW_OK) {
        g_status = SCW_ERR_DEVICE;
        return;
    }
    if (free_gpu_memory <= n * n * 24) {  // gpuEvaluator.cu:140-145
        printf("Insufficient GPU memory for an %dx%dframe\n", N, N);
        g_status = scw::fail(SCW_ERR_NOMEM, "insufficient GPU memory for an %dx%d frame", N, N);
        return;
    }
    // The three planes through scw_eval_planes (its per-plan cache keeps the
    // execs of unchanged planes) into one device frame kept between calls,
    // then copied to the caller's host arrays as the reference does
    // (gpuEvaluator.cu:153-160).
    const size_t plane = n * n * 24;
    if (g_frame_dev != g_plan->device && g_frame.p) {  // a frame left on another device
        scw::DeviceGuard g0(g_frame_dev);
        g_frame.release();
    }
    scw::DeviceGuard g(g_plan->device);
    if ((g_status = g.st) != SCW_OK) return;
    if ((g_status = g_frame.alloc(3 * plane)) != SCW_OK) return;
    g_frame_dev = g_plan->device;
    unsigned char* f = g_frame.as<unsigned char>();
    if ((g_status = scw_eval_planes(g_plan, x_start, y_start, z_start, x, y, z, d, N, 0.001, g_form,
                                    g_form == SCW_FORM_DIRECT ? SCW_DTYPE_FP32 : g_dtype, f, f + plane, f + 2 * plane,
                                    nullptr)) != SCW_OK)
        return;
    void* outs[3] = {E_xy, E_xz, E_yz};
    for (int k = 0; k < 3; ++k) {
        const hipError_t e = hipMemcpy(outs[k], f + k * plane, plane, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            g_status = scw::hip_fail(e, "hipMemcpy");
            return;
        }
    }
}

void cpu_cw_evaluate_xy(scw_vec3c32* E_xy, float x_start, float y_start, float z, float d, size_t N) {
    g_status = eval_plane(E_xy, cpu_axis(x_start, d, N), cpu_axis(y_start, d, N), {(double)z}, 0.0);
}

void cpu_cw_evaluate_yz(scw_vec3c32* E_yz, float y_start, float z_start, float x, float d, size_t N) {
    g_status = eval_plane(E_yz, {(double)x}, cpu_axis(y_start, d, N), cpu_axis(z_start, d, N), 0.0);
}

void cpu_cw_evaluate_xz(scw_vec3c32* E_xz, float x_start, float z_start, float y, float d, size_t N) {
    g_status = eval_plane(E_xz, cpu_axis(x_start, d, N), {(double)y}, cpu_axis(z_start, d, N), 0.0);
}

int scw_compat_last_status() { return g_status; }

void scw_compat_set_form(int form, int dtype) {
    g_form = form;
    g_dtype = dtype;
}
