// ORACLE — test infrastructure, not product code.
//
// CPU restatement of scwatter's plane-wave field evaluator, used ONLY by the
// tests, by the package smoke check and by the CPU-baseline leg of bench.py,
// as the checker.  Nothing in the shipped library links, loads or calls this.
//
// Two families of functions live here:
//
//  * oracle_*  — FP64 restatement of cpu_cw_evaluate_{xy,yz,xz}
//    (cpuEvaluator.cpp:294-407) generalised to rectilinear 3D grids and to
//    point lists.  Region selection follows cpuEvaluator.cpp:301-306 with an
//    optional tolerance (gpuEvaluator.cu:32-37 uses 0.001).  Each term
//    E0_w * exp(i k_w . r) is formed as  (E0_w / s_w) * exp(i k_w . r + ln s_w)
//    with s_w = max_c |E0_w,c|, so that a huge |E0| times a tiny exponential
//    (or the reverse) does not overflow to inf*0 = NaN; mathematically the
//    term is unchanged.
//
//  * refcpu_*  — FP32 restatement with the reference CPU evaluator's own
//    semantics: UnpackedWave<float> tables narrowed from FP64 the way
//    cw_unpack does (cpuEvaluator.cpp:61-102), std::complex<float> phase and
//    std::exp (cpuEvaluator.cpp:312-329), the same loop order, the same
//    coordinate formula x = x_start + i*d in float, the same region rule and
//    the same AoS output (24 bytes per point).  cpu_cw_evaluate_xz
//    accumulates in complex<double> (cpuEvaluator.cpp:395); so does
//    refcpu_eval_xz.  Single-threaded, as the reference is; the grid variant
//    takes a thread count for the all-core baseline.
//
// Wave record (12 doubles, CoupledWaveStructure.h:7-55 record order):
//   E0x.re E0x.im E0y.re E0y.im E0z.re E0z.im kx.re kx.im ky.re ky.im kz.re kz.im
//
// Parity status of each function is stated in tests/test_oracle_pins.py.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef std::complex<double> cd;
typedef std::complex<float> cf;

// cpuEvaluator.cpp:301-306 — region = 1 + the largest l with z >= z_l (0 if none).
// gpuEvaluator.cu:34 compares against z_l - 0.001 in double precision.
inline int region_of(double z, const double* zl, int L, double tol) {
    int r = 0;
    for (int l = 0; l < L; ++l)
        if (z >= zl[l] - tol) r = l + 1;
    return r;
}

struct Wave64 {
    cd e[3];     // E0 / s
    cd k[3];
    double lns;  // ln s, s = max_c |E0_c|
    bool zero;
};

void prepare_one(const double* r, Wave64& o) {
    cd e0[3] = {cd(r[0], r[1]), cd(r[2], r[3]), cd(r[4], r[5])};
    double s = std::max(std::abs(e0[0]), std::max(std::abs(e0[1]), std::abs(e0[2])));
    o.zero = !(s > 0.0);
    o.lns = o.zero ? 0.0 : std::log(s);
    for (int c = 0; c < 3; ++c) o.e[c] = o.zero ? cd(0, 0) : e0[c] / s;
    o.k[0] = cd(r[6], r[7]);
    o.k[1] = cd(r[8], r[9]);
    o.k[2] = cd(r[10], r[11]);
}

std::vector<Wave64> prepare(const double* w, int64_t n) {
    std::vector<Wave64> out((size_t)n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) prepare_one(w + 12 * i, out[(size_t)i]);
    return out;
}

// Only the waves of the regions flagged in `used` (the others stay unset).
std::vector<Wave64> prepare_regions(const double* w, int64_t n, const int32_t* begin, const int32_t* end, int layers,
                                    const std::vector<char>& used) {
    std::vector<Wave64> out((size_t)n);
    for (int r = 0; r <= layers; ++r) {
        if (!used[(size_t)r]) continue;
#pragma omp parallel for schedule(static)
        for (int64_t i = begin[r]; i < end[r]; ++i) prepare_one(w + 12 * i, out[(size_t)i]);
    }
    return out;
}

// Sum over the waves [b, e) of a region at one point (cpuEvaluator.cpp:322-329).
// Complex products are written out in real arithmetic (std::complex's
// operator* calls the C99 NaN-recovering multiply, a library call per
// product); for finite operands the values are the same textbook products.
inline void term(const Wave64& w, double x, double y, double z, double* E) {
    // k . r = (kx x + ky y) + kz z, real and imaginary parts
    const double krr = (w.k[0].real() * x + w.k[1].real() * y) + w.k[2].real() * z;
    const double kri = (w.k[0].imag() * x + w.k[1].imag() * y) + w.k[2].imag() * z;
    // exp(i kr + ln s) = exp(-Im kr + ln s) * (cos Re kr + i sin Re kr)
    const double mag = std::exp(w.lns - kri);
    double sn, cs;
    sincos(krr, &sn, &cs);
    const double pr = mag * cs, pi = mag * sn;
    for (int c = 0; c < 3; ++c) {
        const double er = w.e[c].real(), ei = w.e[c].imag();
        E[2 * c] += er * pr - ei * pi;
        E[2 * c + 1] += er * pi + ei * pr;
    }
}

inline void sum_point(const std::vector<Wave64>& W, int64_t b, int64_t e, double x, double y, double z,
                      double* out6) {
    double E[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t i = b; i < e; ++i) {
        const Wave64& w = W[(size_t)i];
        if (!w.zero) term(w, x, y, z, E);
    }
    for (int c = 0; c < 6; ++c) out6[c] = E[c];
}

// The same sum for a block of up to PB points of one region, each wave read
// once for the whole block (the tables of large regions do not fit in cache).
constexpr int PB = 16;
inline void sum_block(const std::vector<Wave64>& W, int64_t b, int64_t e, const double* const* p, int np,
                      double* const* out6) {
    double E[PB][6] = {};
    for (int64_t i = b; i < e; ++i) {
        const Wave64& w = W[(size_t)i];
        if (w.zero) continue;
        for (int q = 0; q < np; ++q) term(w, p[q][0], p[q][1], p[q][2], E[q]);
    }
    for (int q = 0; q < np; ++q)
        for (int c = 0; c < 6; ++c) out6[q][c] = E[q][c];
}

#ifdef _OPENMP
inline void set_threads(int nthreads) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
}
#else
inline void set_threads(int) {}
#endif

}  // namespace

extern "C" {

int oracle_region_of(double z, const double* z_layers, int layers, double tol) {
    return region_of(z, z_layers, layers, tol);
}

// FP64 field on the rectilinear grid xs[nx] x ys[ny] x zs[nz].
// out: double[nz][ny][nx][3][2]  (the evaluator's AoS layout, in FP64).
// begin/end: int32[layers+1] wave ranges per region (cpuEvaluator.cpp:34-53).
void oracle_eval_grid_f64(const double* waves, int64_t n_waves, const double* z_layers, int layers,
                          const int32_t* begin, const int32_t* end, const double* xs, int64_t nx,
                          const double* ys, int64_t ny, const double* zs, int64_t nz, double tol,
                          double* out, int nthreads) {
    set_threads(nthreads);
    std::vector<Wave64> W = prepare(waves, n_waves);
    const int64_t rows = nz * ny;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t row = 0; row < rows; ++row) {
        const int64_t iz = row / ny, iy = row % ny;
        const int r = region_of(zs[iz], z_layers, layers, tol);
        for (int64_t ix = 0; ix < nx; ++ix)
            sum_point(W, begin[r], end[r], xs[ix], ys[iy], zs[iz], out + 6 * (row * nx + ix));
    }
}

// FP64 field at explicit points pts[n][3] (x, y, z); out double[n][3][2].
// Points are grouped by region and taken PB at a time.
void oracle_eval_points_f64(const double* waves, int64_t n_waves, const double* z_layers, int layers,
                            const int32_t* begin, const int32_t* end, const double* pts, int64_t n,
                            double tol, double* out, int nthreads) {
    set_threads(nthreads);
    std::vector<std::vector<int64_t>> by_region((size_t)layers + 1);
    for (int64_t i = 0; i < n; ++i) by_region[(size_t)region_of(pts[3 * i + 2], z_layers, layers, tol)].push_back(i);
    std::vector<char> used((size_t)layers + 1);
    for (int r = 0; r <= layers; ++r) used[(size_t)r] = !by_region[(size_t)r].empty();
    std::vector<Wave64> W = prepare_regions(waves, n_waves, begin, end, layers, used);
    std::vector<std::pair<int, int64_t>> blocks;  // (region, first index into by_region[region])
    for (int r = 0; r <= layers; ++r)
        for (size_t k = 0; k < by_region[(size_t)r].size(); k += PB) blocks.push_back({r, (int64_t)k});
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t bi = 0; bi < (int64_t)blocks.size(); ++bi) {
        const int r = blocks[(size_t)bi].first;
        const std::vector<int64_t>& idx = by_region[(size_t)r];
        const int64_t k0 = blocks[(size_t)bi].second;
        const int np = (int)std::min<int64_t>(PB, (int64_t)idx.size() - k0);
        const double* p[PB];
        double* o[PB];
        for (int q = 0; q < np; ++q) {
            p[q] = pts + 3 * idx[(size_t)(k0 + q)];
            o[q] = out + 6 * idx[(size_t)(k0 + q)];
        }
        sum_block(W, begin[r], end[r], p, np, o);
    }
}

// ---------------------------------------------------------------------------
// FP32 reference-semantics restatement (cpuEvaluator.cpp:61-102, 294-407).
// ---------------------------------------------------------------------------

struct UnpackedWaveF {  // cpuEvaluator.h:6-10 with T = float
    cf E0[3];
    cf k[3];
};

// cw_unpack narrowing (cpuEvaluator.cpp:63-73): complex<double> -> complex<float> per part.
void refcpu_unpack_f32(const double* waves, int64_t n, float* out12) {
    for (int64_t i = 0; i < 12 * n; ++i) out12[i] = (float)waves[i];
}

// cpu_cw_evaluate_xy (cpuEvaluator.cpp:294-334).  W: float[n][12]; z_layers float[layers].
void refcpu_eval_xy(const float* Wf, const float* z_layers, int layers, const int32_t* begin,
                    const int32_t* end, float* E_xy, float x_start, float y_start, float z, float d,
                    int64_t N) {
    const UnpackedWaveF* W = reinterpret_cast<const UnpackedWaveF*>(Wf);
    float x, y;
    size_t l = 0;
    for (size_t li = 0; li < (size_t)layers; li++)
        if (z >= z_layers[li]) l = li + 1;
    size_t b = begin[l], e = end[l];
    cf phase, k_dot_r = 0, i(0.0f, 1.0f);
    for (unsigned int iy = 0; iy < (unsigned)N; iy++) {
        y = y_start + iy * d;
        for (unsigned int ix = 0; ix < (unsigned)N; ix++) {
            x = x_start + ix * d;
            cf E[3] = {0, 0, 0};
            for (size_t w = b; w < e; w++) {
                k_dot_r = x * W[w].k[0] + y * W[w].k[1] + z * W[w].k[2];
                phase = std::exp(i * k_dot_r);
                E[0] += W[w].E0[0] * phase;
                E[1] += W[w].E0[1] * phase;
                E[2] += W[w].E0[2] * phase;
            }
            float* o = E_xy + 6 * ((size_t)iy * N + ix);
            for (int c = 0; c < 3; ++c) { o[2 * c] = E[c].real(); o[2 * c + 1] = E[c].imag(); }
        }
    }
}

// cpu_cw_evaluate_yz (cpuEvaluator.cpp:336-371): rows z, columns y, at fixed x.
void refcpu_eval_yz(const float* Wf, const float* z_layers, int layers, const int32_t* begin,
                    const int32_t* end, float* E_yz, float y_start, float z_start, float x, float d,
                    int64_t N) {
    const UnpackedWaveF* W = reinterpret_cast<const UnpackedWaveF*>(Wf);
    float y, z;
    cf phase, k_dot_r = 0, i(0.0f, 1.0f);
    size_t l = 0;
    for (unsigned int iz = 0; iz < (unsigned)N; iz++) {
        z = z_start + iz * d;
        for (size_t li = 0; li < (size_t)layers; li++)
            if (z >= z_layers[li]) l = li + 1;
        size_t b = begin[l], e = end[l];
        for (unsigned int iy = 0; iy < (unsigned)N; iy++) {
            y = y_start + iy * d;
            cf E[3] = {0, 0, 0};
            for (size_t w = b; w < e; w++) {
                k_dot_r = x * W[w].k[0] + y * W[w].k[1] + z * W[w].k[2];
                phase = std::exp(i * k_dot_r);
                E[0] += W[w].E0[0] * phase;
                E[1] += W[w].E0[1] * phase;
                E[2] += W[w].E0[2] * phase;
            }
            float* o = E_yz + 6 * ((size_t)iz * N + iy);
            for (int c = 0; c < 3; ++c) { o[2 * c] = E[c].real(); o[2 * c + 1] = E[c].imag(); }
        }
    }
}

// cpu_cw_evaluate_xz (cpuEvaluator.cpp:373-407): rows z, columns x, at fixed y;
// accumulates in complex<double> (cpuEvaluator.cpp:395), phase in complex<float>.
void refcpu_eval_xz(const float* Wf, const float* z_layers, int layers, const int32_t* begin,
                    const int32_t* end, float* E_xz, float x_start, float z_start, float y, float d,
                    int64_t N) {
    const UnpackedWaveF* W = reinterpret_cast<const UnpackedWaveF*>(Wf);
    float x, z;
    cf phase, k_dot_r = 0, i(0.0f, 1.0f);
    size_t l = 0;
    for (unsigned int iz = 0; iz < (unsigned)N; iz++) {
        z = z_start + iz * d;
        for (size_t li = 0; li < (size_t)layers; li++)
            if (z >= z_layers[li]) l = li + 1;
        size_t b = begin[l], e = end[l];
        for (unsigned int ix = 0; ix < (unsigned)N; ix++) {
            x = x_start + ix * d;
            cd E[3] = {0, 0, 0};
            for (size_t w = b; w < e; w++) {
                k_dot_r = x * W[w].k[0] + y * W[w].k[1] + z * W[w].k[2];
                phase = std::exp(i * k_dot_r);
                E[0] += cd(W[w].E0[0] * phase);
                E[1] += cd(W[w].E0[1] * phase);
                E[2] += cd(W[w].E0[2] * phase);
            }
            float* o = E_xz + 6 * ((size_t)iz * N + ix);
            for (int c = 0; c < 3; ++c) {
                o[2 * c] = (float)E[c].real();
                o[2 * c + 1] = (float)E[c].imag();
            }
        }
    }
}

// The xy loop body of cpu_cw_evaluate_xy applied to a rectilinear grid of float
// coordinates (one xy plane per z).  Used to time the reference-semantics CPU
// path on a bounded sample of a 3D workload; nthreads > 1 parallelises rows.
void refcpu_eval_grid(const float* Wf, const float* z_layers, int layers, const int32_t* begin,
                      const int32_t* end, const float* xs, int64_t nx, const float* ys, int64_t ny,
                      const float* zs, int64_t nz, float* out, int nthreads) {
    const UnpackedWaveF* W = reinterpret_cast<const UnpackedWaveF*>(Wf);
    set_threads(nthreads);
    const int64_t rows = nz * ny;
#pragma omp parallel for schedule(dynamic, 1) if (nthreads != 1)
    for (int64_t row = 0; row < rows; ++row) {
        const int64_t iz = row / ny, iy = row % ny;
        const float z = zs[iz], y = ys[iy];
        size_t l = 0;
        for (size_t li = 0; li < (size_t)layers; li++)
            if (z >= z_layers[li]) l = li + 1;
        const size_t b = begin[l], e = end[l];
        const cf i(0.0f, 1.0f);
        for (int64_t ix = 0; ix < nx; ++ix) {
            const float x = xs[ix];
            cf E[3] = {0, 0, 0};
            for (size_t w = b; w < e; w++) {
                cf k_dot_r = x * W[w].k[0] + y * W[w].k[1] + z * W[w].k[2];
                cf phase = std::exp(i * k_dot_r);
                E[0] += W[w].E0[0] * phase;
                E[1] += W[w].E0[1] * phase;
                E[2] += W[w].E0[2] * phase;
            }
            float* o = out + 6 * (row * nx + ix);
            for (int c = 0; c < 3; ++c) { o[2 * c] = E[c].real(); o[2 * c + 1] = E[c].imag(); }
        }
    }
}

}  // extern "C"
