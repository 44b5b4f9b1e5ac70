// Drop-in declarations for the reference's evaluator module: the names,
// argument order, units and layouts of cpuEvaluator.h:6-36 and
// gpuEvaluator.h:8-18, implemented by libscwatter_amd.so on the HIP kernels.
//
// Source-level deltas against the reference headers:
//  * gpu_cw_evaluate takes std::complex<float>* where the reference takes
//    thrust::complex<float>* (same 8-byte layout, no thrust on the host).
//  * glm::vec<3, std::complex<float>> is replaced by the layout-identical
//    POD scw_vec3c32 (24 bytes), so glm is not needed.
//  * EvaluateSample / cpu_cw_evaluate_sample (cpuEvaluator.h:21-24) are not
//    provided: the reference never populates their state or calls them.
#pragma once

#include <complex>
#include <cstddef>
#include <vector>

#include "scw_coupledwave.hpp"

template <typename T>
struct UnpackedWave {  // cpuEvaluator.h:6-10
    std::complex<T> E0[3];
    std::complex<T> k[3];
};

struct scw_vec3c32 {   // glm::vec<3, std::complex<float>>
    std::complex<float> v[3];
    std::complex<float>& operator[](int i) { return v[i]; }
    const std::complex<float>& operator[](int i) const { return v[i]; }
};
static_assert(sizeof(scw_vec3c32) == 24, "one field point is 24 bytes");

// Module state, as in cpuEvaluator.cpp:3-7 (filled by cw_allocate / cw_unpack).
extern int layers;
extern std::vector<float> z_layers;
extern std::vector<int> waves_begin;
extern std::vector<int> waves_end;
extern std::vector<UnpackedWave<float>> W;
// Written by gpu_cw_evaluate (gpuEvaluator.cu:11-12, 139).
extern size_t free_gpu_memory;
extern size_t total_gpu_memory;

void cw_allocate(CoupledWaveStructure<double>* cw);
void cw_unpack(CoupledWaveStructure<double>* cw);

void cpu_cw_evaluate_xy(scw_vec3c32* E_xy, float x_start, float y_start, float z, float d, size_t N);
void cpu_cw_evaluate_yz(scw_vec3c32* E_yz, float y_start, float z_start, float x, float d, size_t N);
void cpu_cw_evaluate_xz(scw_vec3c32* E_xz, float x_start, float z_start, float y, float d, size_t N);

void gpu_initialize();
void gpu_cw_evaluate(std::complex<float>* E_xy, std::complex<float>* E_xz, std::complex<float>* E_yz, float x_start,
                     float y_start, float z_start, float x, float y, float z, float d, int N, int device);

// Additions (no counterpart in the reference): status of the last call above
// (SCW_OK or a negative scw_status; the reference aborts or returns silently),
// and the evaluation form / GEMM dtype those calls use (default SCW_FORM_AUTO,
// SCW_DTYPE_FP32).
int scw_compat_last_status();
void scw_compat_set_form(int form, int dtype);
