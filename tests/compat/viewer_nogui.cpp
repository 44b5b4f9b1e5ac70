// A caller written against the reference's evaluator API, the way
// sCWatterView.cpp drives it in --nogui mode (sCWatterView.cpp:741-789):
// load a .cw file, cw_allocate + cw_unpack, gpu_initialize, then
// gpu_cw_evaluate for the three planes and the cpu_cw_evaluate_* trio.
// Only the header names change (scw_evaluator.hpp for cpuEvaluator.h /
// gpuEvaluator.h).  Writes raw complex64 planes for the GPU test to check.
//
// usage: viewer_nogui in.cw out_prefix N extent cx cy cz px py pz device
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "scw_evaluator.hpp"

size_t free_gpu_memory;   // owned by the caller, as in sCWatterView.cpp:102-103
size_t total_gpu_memory;

static bool dump(const std::string& path, const void* p, size_t bytes) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = fwrite(p, 1, bytes, f) == bytes;
    return fclose(f) == 0 && ok;
}

int main(int argc, char** argv) {
    if (argc != 12) {
        fprintf(stderr, "usage: %s in.cw out_prefix N extent cx cy cz px py pz device\n", argv[0]);
        return 2;
    }
    const std::string in = argv[1], out = argv[2];
    const unsigned N = (unsigned)atoi(argv[3]);
    const float extent = (float)atof(argv[4]);
    const float center[3] = {(float)atof(argv[5]), (float)atof(argv[6]), (float)atof(argv[7])};
    const float plane[3] = {(float)atof(argv[8]), (float)atof(argv[9]), (float)atof(argv[10])};
    const int device = atoi(argv[11]);

    CoupledWaveStructure<double> cw;
    if (!cw.load(in)) {
        printf("ERROR: file %s not found\n", in.c_str());
        return 1;
    }
    cw_allocate(&cw);
    cw_unpack(&cw);
    gpu_initialize();
    if (scw_compat_last_status() != SCW_OK) return 3;

    // EvaluateVectorSlices (sCWatterView.cpp:354-370)
    const float d = extent / (N - 1);
    const float x_start = center[0] - extent / 2, y_start = center[1] - extent / 2, z_start = center[2] - extent / 2;
    std::vector<std::complex<float>> Exy(3 * N * N), Exz(3 * N * N), Eyz(3 * N * N);
    gpu_cw_evaluate(Exy.data(), Exz.data(), Eyz.data(), x_start, y_start, z_start, plane[0], plane[1], plane[2], d,
                    (int)N, device);
    if (scw_compat_last_status() != SCW_OK) return 4;
    std::vector<scw_vec3c32> Cxy(N * N), Cxz(N * N), Cyz(N * N);
    cpu_cw_evaluate_xy(Cxy.data(), x_start, y_start, plane[2], d, N);
    cpu_cw_evaluate_xz(Cxz.data(), x_start, z_start, plane[1], d, N);
    cpu_cw_evaluate_yz(Cyz.data(), y_start, z_start, plane[0], d, N);
    if (scw_compat_last_status() != SCW_OK) return 5;

    const size_t b = (size_t)N * N * 24;
    bool ok = dump(out + "_gpu_xy.bin", Exy.data(), b) && dump(out + "_gpu_xz.bin", Exz.data(), b) &&
              dump(out + "_gpu_yz.bin", Eyz.data(), b) && dump(out + "_cpu_xy.bin", Cxy.data(), b) &&
              dump(out + "_cpu_xz.bin", Cxz.data(), b) && dump(out + "_cpu_yz.bin", Cyz.data(), b);
    printf("layers %d waves %zu free %zu total %zu\n", layers, W.size(), free_gpu_memory, total_gpu_memory);
    return ok ? 0 : 6;
}
