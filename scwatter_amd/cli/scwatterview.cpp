// scwatterview --nogui on this library: the batch path of the reference
// viewer (sCWatterView.cpp:679-796).  Options, defaults, grid and output
// follow the reference:
//
//   scwatterview [--input] c.cw [--cuda|-c 0] [--nogui] [--verbose|-v]
//                [--size 100] [--resolution 8] [--output xz.npy] [--axis 1]
//                [--center 0 0 0] [--slice 0]
//
// N = 2^resolution, d = size / (N - 1), start = center - size / 2 per axis
// (EvaluateVectorSlices, sCWatterView.cpp:354-370); the plane along `axis`
// sits at `slice` (an absolute coordinate, as the reference uses it at
// :720-725), the other two at 0.  The three planes come from gpu_cw_evaluate
// (device >= 0) or from the cpu_cw_evaluate_* names (device -1; they run on
// the GPU with the reference CPU evaluator's region rule and coordinate
// rounding), and the plane of `axis` is written as an (N, N, 3) complex64
// .npy (axis 0 -> yz, 1 -> xz, 2 -> xy; :773-797).
//
// Differences, each deliberate:
//   - there is no GUI: the program always runs as with --nogui;
//   - it exits 0 after a successful save (the reference exits 1, :795);
//   - --sample (a voxel sample for the GUI overlay) is rejected;
//   - additions: --volume writes the whole N^3 cube as (N, N, N, 3) complex64
//     [z][y][x]; --form auto|direct|gemm and --dtype fp32|bf16x3|bf16|fp16
//     choose the evaluation form.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "scw_evaluator.hpp"
#include "scwatter_amd.h"

size_t free_gpu_memory;  // owned by the caller, as in sCWatterView.cpp:102-103
size_t total_gpu_memory;

namespace {

struct Options {
    std::string input = "c.cw", output = "xz.npy";
    int device = 0, resolution = 8, axis = 1;
    float size = 100.0f, slice = 0.0f, center[3] = {0.0f, 0.0f, 0.0f};
    bool verbose = false, help = false, volume = false;
    int form = SCW_FORM_AUTO, dtype = SCW_DTYPE_FP32;
};

const char* kHelp =
    "Allowed options:\n"
    "  --input arg (=c.cw)        coupled wave structure (.cw) to evaluate (also positional)\n"
    "  --help                     produce help message\n"
    "  -c [ --cuda ] arg (=0)     device number (-1: the CPU evaluator's rules, still on the GPU)\n"
    "  --nogui                    save an output file (this program always runs without a GUI)\n"
    "  -v [ --verbose ]           produce verbose output\n"
    "  --size arg (=100)          size of the sampled volume (extent)\n"
    "  --resolution arg (=8)      resolution of the sampled field: N = 2^resolution\n"
    "  --output arg (=xz.npy)     output file\n"
    "  --axis arg (=1)            axis to cut (0 = X, 1 = Y, 2 = Z)\n"
    "  --center arg (={0, 0, 0})  center position of the sampled volume\n"
    "  --slice arg (=0)           coordinate of the cut along the axis\n"
    "  --volume                   write the whole N^3 volume (N, N, N, 3) instead of one plane\n"
    "  --form arg (=auto)         evaluation form: auto, direct or gemm\n"
    "  --dtype arg (=fp32)        GEMM-form arithmetic: fp32, bf16x3, bf16 or fp16\n";

bool parse_float(const char* s, float* v) {
    char* end = nullptr;
    const double x = std::strtod(s, &end);
    if (end == s || *end != '\0' || !std::isfinite(x)) return false;
    *v = (float)x;
    return true;
}

bool parse_int(const char* s, int* v) {
    char* end = nullptr;
    const long x = std::strtol(s, &end, 10);
    if (end == s || *end != '\0') return false;
    *v = (int)x;
    return true;
}

// Returns an empty string on success, else the error message.
std::string parse(int argc, char** argv, Options& o) {
    std::vector<std::string> args(argv + 1, argv + argc);
    for (size_t i = 0; i < args.size(); ++i) {
        std::string a = args[i], val;
        bool has_val = false;
        const size_t eq = a.find('=');
        if (a.rfind("--", 0) == 0 && eq != std::string::npos) {
            val = a.substr(eq + 1);
            a = a.substr(0, eq);
            has_val = true;
        }
        auto next = [&](std::string& v) {
            if (has_val) {
                v = val;
                return true;
            }
            if (i + 1 >= args.size()) return false;
            v = args[++i];
            return true;
        };
        std::string v;
        if (a == "--help") {
            o.help = true;
        } else if (a == "--nogui") {
        } else if (a == "--verbose" || a == "-v") {
            o.verbose = true;
        } else if (a == "--volume") {
            o.volume = true;
        } else if (a == "--sample") {
            return "--sample (voxel sample overlay of the GUI) is not supported";
        } else if (a == "--input") {
            if (!next(o.input)) return "--input needs a value";
        } else if (a == "--output") {
            if (!next(o.output)) return "--output needs a value";
        } else if (a == "--cuda" || a == "-c") {
            if (!next(v) || !parse_int(v.c_str(), &o.device)) return "--cuda needs an integer";
        } else if (a == "--resolution") {
            if (!next(v) || !parse_int(v.c_str(), &o.resolution)) return "--resolution needs an integer";
        } else if (a == "--axis") {
            if (!next(v) || !parse_int(v.c_str(), &o.axis)) return "--axis needs an integer";
        } else if (a == "--size") {
            if (!next(v) || !parse_float(v.c_str(), &o.size)) return "--size needs a number";
        } else if (a == "--slice") {
            if (!next(v) || !parse_float(v.c_str(), &o.slice)) return "--slice needs a number";
        } else if (a == "--center") {
            // multitoken: up to three numbers
            int k = 0;
            if (has_val) {
                if (!parse_float(val.c_str(), &o.center[k++])) return "--center needs numbers";
            }
            while (k < 3 && i + 1 < args.size() && parse_float(args[i + 1].c_str(), &o.center[k])) {
                ++k;
                ++i;
            }
            if (k == 0) return "--center needs numbers";
        } else if (a == "--form") {
            if (!next(v)) return "--form needs a value";
            if (v == "auto") o.form = SCW_FORM_AUTO;
            else if (v == "direct") o.form = SCW_FORM_DIRECT;
            else if (v == "gemm") o.form = SCW_FORM_GEMM;
            else return "--form must be auto, direct or gemm";
        } else if (a == "--dtype") {
            if (!next(v)) return "--dtype needs a value";
            if (v == "fp32") o.dtype = SCW_DTYPE_FP32;
            else if (v == "bf16x3") o.dtype = SCW_DTYPE_BF16X3;
            else if (v == "bf16") o.dtype = SCW_DTYPE_BF16;
            else if (v == "fp16") o.dtype = SCW_DTYPE_FP16;
            else return "--dtype must be fp32, bf16x3, bf16 or fp16";
        } else if (a.rfind("-", 0) == 0) {
            return "unknown option " + a;
        } else {
            o.input = a;  // positional input
        }
    }
    if (o.resolution < 1 || o.resolution > 14) return "--resolution must be in 1..14";
    if (!(o.size > 0.0f)) return "--size must be positive";
    return "";
}

// .npy version 1.0, little-endian complex64, C order.
bool save_npy(const std::string& path, const std::vector<size_t>& shape, const void* data, size_t bytes) {
    std::string dict = "{'descr': '<c8', 'fortran_order': False, 'shape': (";
    for (size_t i = 0; i < shape.size(); ++i) dict += std::to_string(shape[i]) + (shape.size() == 1 || i + 1 < shape.size() ? ", " : "");
    dict += "), }";
    size_t total = 10 + dict.size() + 1;
    const size_t pad = (64 - total % 64) % 64;
    dict += std::string(pad, ' ') + "\n";
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    const uint16_t hlen = (uint16_t)dict.size();
    const unsigned char hl[2] = {(unsigned char)(hlen & 0xff), (unsigned char)(hlen >> 8)};
    bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(hl, 1, 2, f) == 2 &&
              std::fwrite(dict.data(), 1, dict.size(), f) == dict.size() &&
              std::fwrite(data, 1, bytes, f) == bytes;
    return std::fclose(f) == 0 && ok;
}

// The whole cube through the C ABI: grid start + d i per axis with the
// coordinate rounding and region rule of the evaluator the device selects.
int evaluate_volume(const Options& o, unsigned N, float d, const float start[3], std::vector<std::complex<float>>& E) {
    scw_cw* cw = nullptr;
    int st = scw_cw_load(o.input.c_str(), &cw);
    if (st != SCW_OK) return st;
    size_t ni = 0, nl = 0, nw = 0;
    scw_cw_counts(cw, &ni, &nl, &nw);
    std::vector<scw_planewave64> waves(nw);
    std::vector<double> zl(nl);
    std::vector<int32_t> wb(nl + 1), we(nl + 1);
    st = scw_cw_unpack(cw, waves.data(), zl.data(), wb.data(), we.data());
    scw_cw_destroy(cw);
    if (st != SCW_OK) return st;
    scw_plan* plan = nullptr;
    if ((st = scw_plan_create(std::max(o.device, 0), waves.data(), nw, zl.data(), (int)nl, wb.data(), we.data(),
                              &plan)) != SCW_OK)
        return st;
    std::vector<double> ax[3];
    for (int a = 0; a < 3; ++a) {
        ax[a].resize(N);
        for (unsigned i = 0; i < N; ++i) {
            volatile float p = d * (float)i;  // cpu_cw_evaluate_*: product then sum, each rounded
            ax[a][i] = o.device >= 0 ? (double)std::fmaf(d, (float)i, start[a]) : (double)(float)(start[a] + p);
        }
    }
    scw_exec* ex = nullptr;
    st = scw_exec_create(plan, ax[0].data(), N, ax[1].data(), N, ax[2].data(), N, o.device >= 0 ? 0.001 : 0.0,
                         o.form, o.dtype, &ex);
    if (st == SCW_OK) {
        E.resize((size_t)N * N * N * 3);
        st = scw_exec_run_host(ex, E.data(), nullptr);
        scw_exec_destroy(ex);
    }
    scw_plan_destroy(plan);
    return st;
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    const std::string err = parse(argc, argv, o);
    if (!err.empty()) {
        std::fprintf(stderr, "scwatterview: %s\n", err.c_str());
        return 2;
    }
    if (o.help) {
        std::printf("%s", kHelp);
        return 1;  // as the reference (sCWatterView.cpp:710-713)
    }
    const auto start_all = std::chrono::system_clock::now();
    float plane_position[3] = {0.0f, 0.0f, 0.0f};
    if (o.axis >= 0 && o.axis <= 2) plane_position[o.axis] = o.slice;

    CoupledWaveStructure<double> cw;
    if (!cw.load(o.input)) {
        std::printf("ERROR: file %s not found\n", o.input.c_str());
        return 1;
    }
    const unsigned N = 1u << o.resolution;
    const float d = o.size / (N - 1);
    const float start[3] = {o.center[0] - o.size / 2, o.center[1] - o.size / 2, o.center[2] - o.size / 2};
    scw_compat_set_form(o.form, o.dtype);

    if (o.volume) {
        std::vector<std::complex<float>> E;
        const int st = evaluate_volume(o, N, d, start, E);
        if (st != SCW_OK) {
            std::fprintf(stderr, "scwatterview: %s\n", scw_last_error());
            return 3;
        }
        const std::chrono::duration<double> el = std::chrono::system_clock::now() - start_all;
        std::printf("Evaluation (no gui) takes:%gs\n", el.count());
        if (!save_npy(o.output, {N, N, N, 3}, E.data(), E.size() * sizeof(E[0]))) {
            std::fprintf(stderr, "scwatterview: cannot write %s\n", o.output.c_str());
            return 4;
        }
        return 0;
    }

    cw_allocate(&cw);
    cw_unpack(&cw);
    if (o.device >= 0) {
        gpu_initialize();
        if (scw_compat_last_status() != SCW_OK) {
            std::fprintf(stderr, "scwatterview: %s\n", scw_last_error());
            return 3;
        }
        if (o.verbose) std::printf("Using Device %d for data processing\n", o.device);
    }
    // planes as 24-byte points: E[row * N + col]
    const size_t np = (size_t)N * N;
    std::vector<std::complex<float>> Gxy(3 * np), Gxz(3 * np), Gyz(3 * np);
    std::vector<scw_vec3c32> Cxy, Cxz, Cyz;
    if (o.device >= 0) {
        gpu_cw_evaluate(Gxy.data(), Gxz.data(), Gyz.data(), start[0], start[1], start[2], plane_position[0],
                        plane_position[1], plane_position[2], d, (int)N, o.device);
    } else {
        Cxy.resize(np);
        Cxz.resize(np);
        Cyz.resize(np);
        cpu_cw_evaluate_xy(Cxy.data(), start[0], start[1], plane_position[2], d, N);
        cpu_cw_evaluate_xz(Cxz.data(), start[0], start[2], plane_position[1], d, N);
        cpu_cw_evaluate_yz(Cyz.data(), start[1], start[2], plane_position[0], d, N);
    }
    if (scw_compat_last_status() != SCW_OK) {
        std::fprintf(stderr, "scwatterview: %s\n", scw_last_error());
        return 3;
    }
    const std::chrono::duration<double> el = std::chrono::system_clock::now() - start_all;
    std::printf("Evaluation (no gui) takes:%gs\n", el.count());
    const bool gpu = o.device >= 0;
    const void* plane = nullptr;
    if (o.axis == 1) plane = gpu ? (const void*)Gxz.data() : (const void*)Cxz.data();
    else if (o.axis == 2) plane = gpu ? (const void*)Gxy.data() : (const void*)Cxy.data();
    else if (o.axis == 0) plane = gpu ? (const void*)Gyz.data() : (const void*)Cyz.data();
    if (!plane) {
        std::printf("Wrong click at the wrong region. \n");
        return 1;
    }
    if (!save_npy(o.output, {N, N, 3}, plane, np * 24)) {
        std::fprintf(stderr, "scwatterview: cannot write %s\n", o.output.c_str());
        return 4;
    }
    return 0;
}
