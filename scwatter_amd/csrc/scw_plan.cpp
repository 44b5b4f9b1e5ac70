// Structure plan: validation, FP64 planning and device upload.
//
// Replaces gpu_initialize (gpuEvaluator.cu:15-26), which copied the FP32
// UnpackedWave table (cpuEvaluator.cpp:61-102) and the region tables
// (cpuEvaluator.cpp:34-53) to the device.  Here the FP64 waves are kept, and
// each region is additionally decomposed for the separable GEMM form:
// distinct kx (a < mx), ky (b < my) and kz (c < mc) values, and terms
// A[a,b,c] = sum of E0 over the waves with those values, re-referenced in
// FP64 to the region's phase origin zref.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <array>
#include <atomic>
#include <mutex>
#include <unordered_map>
#include <new>

#include "scw_internal.hpp"

namespace scw {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int fail(int status, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return status;
}

void hip_runtime_once() {
    static std::once_flag once;
    std::call_once(once, [] {
        int n = 0;
        (void)hipGetDeviceCount(&n);
    });
}

int hip_fail(hipError_t e, const char* what) {
    return fail(e == hipErrorOutOfMemory ? SCW_ERR_NOMEM : SCW_ERR_DEVICE, "%s: %s", what,
                hipGetErrorString(e));
}

// Allocation-failure injection for tests of the error paths: after
// scw_test_fail_alloc_after(k), the (k+1)-th non-empty buffer request from
// then on fails as out-of-memory (k < 0 turns it off).
static std::atomic<int64_t> g_fail_after{-1};

int DevBuf::alloc(size_t n) {
    if (n == 0) {
        release();
        return SCW_OK;
    }
    // every non-empty request counts for the injected failures, reused or not
    if (g_fail_after.load() >= 0 && g_fail_after.fetch_sub(1) == 0) {
        release();
        return fail(SCW_ERR_NOMEM, "hipMalloc: injected allocation failure (%zu bytes)", n);
    }
    if (p && n <= cap) {
        bytes = n;
        return SCW_OK;
    }
    release();
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
        p = nullptr;
        return hip_fail(e, "hipMalloc");
    }
    bytes = cap = n;
    return SCW_OK;
}

int DevBuf::upload(const void* host, size_t n) {
    if (n > 0 && !host) return fail(SCW_ERR_ARG, "internal: upload of %zu bytes from NULL", n);
    int st = alloc(n);
    if (st != SCW_OK || n == 0) return st;
    SCW_HIP(hipMemcpy(p, host, n, hipMemcpyHostToDevice));
    return SCW_OK;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = cap = 0;
}

// Group values that agree to within tol in both parts; returns distinct
// representatives (in sorted order) and a group index per input.  Exact
// duplicates (the common case: producers compute kx, ky, kz with one formula
// per order or mode) are merged by hashing first, so only the distinct values
// are sorted.
struct PairHash {
    size_t operator()(const std::pair<uint64_t, uint64_t>& p) const {
        return std::hash<uint64_t>()(p.first * 0x9E3779B97F4A7C15ull ^ p.second);
    }
};

static void group_values(const std::vector<cd>& v, double tol, std::vector<cd>& reps, std::vector<int>& idx) {
    const size_t n = v.size();
    std::unordered_map<std::pair<uint64_t, uint64_t>, int, PairHash> exact;
    std::vector<cd> uniq;
    std::vector<int> uid(n);
    for (size_t i = 0; i < n; ++i) {
        double re = v[i].real() + 0.0, im = v[i].imag() + 0.0;  // -0 == +0
        uint64_t a, b;
        std::memcpy(&a, &re, 8);
        std::memcpy(&b, &im, 8);
        auto it = exact.find({a, b});
        if (it == exact.end()) {
            it = exact.emplace(std::make_pair(a, b), (int)uniq.size()).first;
            uniq.push_back(cd(re, im));
        }
        uid[i] = it->second;
    }
    const size_t u = uniq.size();
    std::vector<int> order(u);
    for (size_t i = 0; i < u; ++i) order[i] = (int)i;
    std::sort(order.begin(), order.end(), [&](int x, int y) {
        if (uniq[x].real() != uniq[y].real()) return uniq[x].real() < uniq[y].real();
        return uniq[x].imag() < uniq[y].imag();
    });
    reps.clear();
    std::vector<int> grp(u, -1);
    for (size_t t = 0; t < u; ++t) {
        const cd x = uniq[order[t]];
        int g = -1;
        // Look back over representatives whose real part is within tol.
        for (int k = (int)reps.size() - 1; k >= 0; --k) {
            if (x.real() - reps[k].real() > tol) break;
            if (std::fabs(x.imag() - reps[k].imag()) <= tol) {
                g = k;
                break;
            }
        }
        if (g < 0) {
            reps.push_back(x);
            g = (int)reps.size() - 1;
        }
        grp[order[t]] = g;
    }
    idx.resize(n);
    for (size_t i = 0; i < n; ++i) idx[i] = grp[uid[i]];
}

// Largest ky count of the GEMM form: the main GEMM takes K = 2 my <= 128 per
// launch, and a region of more ky runs it in chunks of at most 64 ky rows
// whose launches after the first add to the output.
static const int MY_MAX = 1024;
static const int MX_MAX = 256;  // largest kx count (LDS staging of the FP32 X contraction)

static void plan_region(RegionPlan& R, const scw_planewave64* W, const std::vector<double>& lns,
                        const std::vector<cd>& ehat, std::vector<float>& ellA, std::vector<int32_t>& ellC,
                        int64_t& ellA_slots, int64_t& ellC_slots) {  // ellA_slots: float2 count
    const int64_t n = R.wend - R.wbegin;
    R.separable = false;
    R.dense = false;
    R.ellc_off = -1;
    if (n == 0) {  // empty region: the field is zero; trivially separable
        R.separable = true;
        R.mx = R.my = R.mc = R.width = 0;
        R.n_terms = 0;
        R.ell_off = ellA_slots;
        R.arow = 0;
        return;
    }
    std::vector<cd> kx(n), ky(n), kz(n);
    double kscale = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const scw_planewave64& w = W[R.wbegin + i];
        kx[i] = cd(w.k[0][0], w.k[0][1]);
        ky[i] = cd(w.k[1][0], w.k[1][1]);
        kz[i] = cd(w.k[2][0], w.k[2][1]);
        kscale = std::max(kscale, std::max(std::abs(kx[i]), std::max(std::abs(ky[i]), std::abs(kz[i]))));
    }
    const double tol = 1e-11 * std::max(kscale, 1e-300);
    std::vector<int> ia, ib, ic;
    group_values(kx, tol, R.kx, ia);
    group_values(ky, tol, R.ky, ib);
    group_values(kz, tol, R.kz, ic);
    R.mx = (int)R.kx.size();
    R.my = (int)R.ky.size();
    R.mc = (int)R.kz.size();
    const int64_t cells = (int64_t)R.mx * R.my;

    // Per distinct kz: the re-referencing factor exp(i kz zref) as a real
    // exponent and a unit rotation, so that every wave costs one real exp.
    std::vector<double> zexp(R.mc);
    std::vector<cd> zrot(R.mc);
    for (int c = 0; c < R.mc; ++c) {
        const cd ph = cd(0, 1) * R.kz[c] * R.zref;
        zexp[c] = ph.real();
        zrot[c] = cd(std::cos(ph.imag()), std::sin(ph.imag()));
    }
    // Terms keyed by (cell, c); E0 summed in FP64 after re-referencing to zref.
    struct Term {
        int64_t cell;
        int c;
        cd e[3];
    };
    std::vector<Term> terms;
    const int64_t keyspace = cells * (int64_t)R.mc;
    std::vector<int64_t> slot_of;
    std::unordered_map<int64_t, int64_t> slot_map;
    const bool direct_keys = keyspace <= (int64_t(1) << 26);
    if (direct_keys) slot_of.assign((size_t)keyspace, -1);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t g = R.wbegin + i;
        const int64_t cell = (int64_t)ib[i] * R.mx + ia[i];  // b-major: C_z[:, b, :] is contiguous
        const int64_t key = cell * R.mc + ic[i];
        int64_t t;
        if (direct_keys) {
            t = slot_of[(size_t)key];
            if (t < 0) {
                t = (int64_t)terms.size();
                slot_of[(size_t)key] = t;
                terms.push_back({cell, ic[i], {}});
            }
        } else {
            auto it = slot_map.find(key);
            if (it == slot_map.end()) {
                t = (int64_t)terms.size();
                slot_map[key] = t;
                terms.push_back({cell, ic[i], {}});
            } else {
                t = it->second;
            }
        }
        // E0 exp(i kz zref) = (E0/s) exp(ln s - Im(kz) zref) (cos, sin)(Re(kz) zref)
        const double mag = std::exp(lns[g] + zexp[ic[i]]);
        const cd f = zrot[ic[i]] * mag;
        for (int c = 0; c < 3; ++c) terms[(size_t)t].e[c] += ehat[3 * g + c] * f;
    }
    // Terms keep their first-appearance order: dense regions index the ELL by
    // kz index, sparse ones fill each cell's columns in that order.
    R.n_terms = (int64_t)terms.size();
    std::vector<int> count(cells, 0);
    for (auto& t : terms) count[t.cell]++;
    int width = 0;
    for (int64_t c = 0; c < cells; ++c) width = std::max(width, count[c]);
    R.width = width;
    bool dense = (int64_t)width == R.mc;
    for (int64_t c = 0; c < cells && dense; ++c) dense = count[c] == R.mc;
    R.dense = dense;

    if (R.my > MY_MAX) {
        R.why_not = "more than 1024 distinct ky values";
        return;
    }
    if (R.mx > MX_MAX) {
        R.why_not = "more than 256 distinct kx values";
        return;
    }
    const int64_t slots = (int64_t)width * cells;
    if (cells > std::max<int64_t>(4 * R.n_terms, 4096) || slots > 4 * R.n_terms + 4096) {
        R.why_not = "wave set is not a separable kx x ky lattice";
        return;
    }
    for (auto& t : terms)
        for (int c = 0; c < 3; ++c)
            if (!std::isfinite(t.e[c].real()) || !std::isfinite(t.e[c].imag()) || std::abs(t.e[c]) > 1e30) {
                R.why_not = "re-referenced amplitude outside FP32 range";
                return;
            }
    R.separable = true;
    // Term rows padded to whole 128-byte lines, so every row starts on a line
    // and the contractions' 128-byte row segments never straddle two.
    R.arow = (3 * cells + 15) / 16 * 16;
    R.ell_off = ellA_slots;
    const size_t baseA = ellA.size();
    ellA.resize(baseA + (size_t)width * R.arow * 2, 0.0f);
    if (!dense) {
        R.ellc_off = ellC_slots;
        ellC.resize(ellC.size() + (size_t)slots, 0);
    }
    std::vector<int> fill(cells, 0);
    for (auto& t : terms) {
        const int j = dense ? t.c : fill[t.cell]++;
        const int64_t slot = (int64_t)j * cells + t.cell;
        float* a = &ellA[baseA + ((size_t)j * R.arow + (size_t)t.cell * 3) * 2];
        for (int c = 0; c < 3; ++c) {
            a[2 * c] = (float)t.e[c].real();
            a[2 * c + 1] = (float)t.e[c].imag();
        }
        if (!dense) ellC[(size_t)(R.ellc_off + slot)] = t.c;
    }
    ellA_slots += (int64_t)width * R.arow;
    if (!dense) ellC_slots += slots;
}

}  // namespace scw

using namespace scw;

extern "C" {

const char* scw_last_error(void) { return g_err; }
void scw_test_fail_alloc_after(int64_t k) { g_fail_after.store(k < 0 ? -1 : k); }
int scw_abi_version(void) { return SCW_ABI_VERSION; }

int scw_device_count(int* count) {
    if (!count) return fail(SCW_ERR_ARG, "count is NULL");
    hip_runtime_once();
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return hip_fail(e, "hipGetDeviceCount");
    }
    *count = n;
    return SCW_OK;
}

int scw_device_mem_info(int device, size_t* free_bytes, size_t* total_bytes) {
    if (!free_bytes || !total_bytes) return fail(SCW_ERR_ARG, "NULL output");
    DeviceGuard g(device);
    if (g.st != SCW_OK) return g.st;
    SCW_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return SCW_OK;
}

}  // extern "C"

namespace {

// Argument checks of scw_plan_create that need no device.
int validate_tables(const scw_planewave64* waves, size_t n_waves, const double* z_layers, int layers,
                    const int32_t* waves_begin, const int32_t* waves_end) {
    if (layers < 0) return fail(SCW_ERR_ARG, "layers < 0");
    if (n_waves > 0 && !waves) return fail(SCW_ERR_ARG, "waves is NULL");
    if (layers > 0 && !z_layers) return fail(SCW_ERR_ARG, "z_layers is NULL");
    if (!waves_begin || !waves_end) return fail(SCW_ERR_ARG, "waves_begin / waves_end is NULL");
    if (n_waves > (size_t)INT32_MAX) return fail(SCW_ERR_ARG, "more than 2^31-1 waves");
    for (int l = 0; l < layers; ++l)
        if (!std::isfinite(z_layers[l])) return fail(SCW_ERR_ARG, "z_layers[%d] is not finite", l);
    for (int r = 0; r <= layers; ++r) {
        if (waves_begin[r] < 0 || waves_end[r] < waves_begin[r] || (size_t)waves_end[r] > n_waves)
            return fail(SCW_ERR_ARG, "region %d wave range [%d, %d) invalid for %zu waves", r, waves_begin[r],
                        waves_end[r], n_waves);
    }
    for (size_t i = 0; i < n_waves; ++i) {
        const double* v = &waves[i].E0[0][0];
        for (int j = 0; j < 12; ++j)
            if (!std::isfinite(v[j])) return fail(SCW_ERR_ARG, "wave %zu has a non-finite value", i);
    }
    return SCW_OK;
}

// Host tables of a plan, before their upload.
struct HostTables {
    std::vector<double> k64, e64, lns;
    std::vector<float> kf, ellA;
    std::vector<int32_t> ellC;
    std::vector<unsigned char> dreg;
};

// The host half of planning (FP64): the direct form's tables, every
// region's separable decomposition and the GEMM form's amplitude tables.
// Fills P's host fields (layers, boundaries, regions) and H; touches no
// device.
int plan_host(scw_plan* P, const scw_planewave64* waves, size_t n_waves, const double* z_layers, int layers,
              const int32_t* waves_begin, const int32_t* waves_end, HostTables& H) {
    P->layers = layers;
    P->n_waves = (int64_t)n_waves;
    P->zl.assign(z_layers, z_layers + layers);
    P->regions.resize(layers + 1);

    // Direct-form tables.
    const double log2e = 1.4426950408889634;
    const double inv2pi = 0.15915494309189535;
    std::vector<double>& k64 = H.k64;
    std::vector<double>& e64 = H.e64;
    std::vector<double>& lns = H.lns;
    std::vector<float>& kf = H.kf;
    k64.assign(6 * n_waves, 0.0);
    e64.assign(6 * n_waves, 0.0);
    lns.assign(n_waves, 0.0);
    kf.assign(4 * n_waves, 0.0f);
    std::vector<cd> ehat(3 * n_waves);
    for (size_t i = 0; i < n_waves; ++i) {
        const scw_planewave64& w = waves[i];
        double s = 0.0;
        for (int c = 0; c < 3; ++c) s = std::max(s, std::hypot(w.E0[c][0], w.E0[c][1]));
        lns[i] = s > 0.0 ? std::log(s) : 0.0;
        for (int c = 0; c < 3; ++c) {
            ehat[3 * i + c] = s > 0.0 ? cd(w.E0[c][0] / s, w.E0[c][1] / s) : cd(0, 0);
            e64[6 * i + 2 * c] = ehat[3 * i + c].real();
            e64[6 * i + 2 * c + 1] = ehat[3 * i + c].imag();
            k64[6 * i + 2 * c] = w.k[c][0];
            k64[6 * i + 2 * c + 1] = w.k[c][1];
        }
        kf[4 * i + 0] = (float)(w.k[0][0] * inv2pi);
        kf[4 * i + 1] = (float)(w.k[1][0] * inv2pi);
        kf[4 * i + 2] = (float)(-w.k[0][1] * log2e);
        kf[4 * i + 3] = (float)(-w.k[1][1] * log2e);
    }

    // Region planning.
    int64_t slotsA = 0, slotsC = 0;
    for (int r = 0; r <= layers; ++r) {
        RegionPlan& R = P->regions[r];
        R.wbegin = waves_begin[r];
        R.wend = waves_end[r];
        if (layers == 0) R.zref = 0.0;
        else if (r == 0) R.zref = z_layers[0];
        else if (r == layers) R.zref = z_layers[layers - 1];
        else R.zref = 0.5 * (z_layers[r - 1] + z_layers[r]);
        R.real_transverse = true;
        for (int64_t i = R.wbegin; i < R.wend; ++i)
            if (waves[i].k[0][1] != 0.0 || waves[i].k[1][1] != 0.0) R.real_transverse = false;
        plan_region(R, waves, lns, ehat, H.ellA, H.ellC, slotsA, slotsC);
    }
    P->ell_elems = slotsA;
    H.dreg.assign(direct_region_bytes() * (layers + 1), 0);
    for (int r = 0; r <= layers; ++r)
        direct_region_fill(&H.dreg[direct_region_bytes() * r], P->regions[r].wbegin, P->regions[r].wend,
                           P->regions[r].real_transverse ? 1 : 0);
    return SCW_OK;
}

}  // namespace

extern "C" {

int scw_plan_create(int device, const scw_planewave64* waves, size_t n_waves, const double* z_layers,
                    int layers, const int32_t* waves_begin, const int32_t* waves_end, scw_plan** out) {
    if (!out) return fail(SCW_ERR_ARG, "plan output is NULL");
    *out = nullptr;
    int st = validate_tables(waves, n_waves, z_layers, layers, waves_begin, waves_end);
    if (st != SCW_OK) return st;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(SCW_ERR_DEVICE, "no HIP device");
    if (device < 0 || device >= ndev) return fail(SCW_ERR_ARG, "device %d out of range [0, %d)", device, ndev);

    TraceRange tr("scw_plan_create");
    DeviceGuard g(device);
    if (g.st != SCW_OK) return g.st;

    scw_plan* P = new (std::nothrow) scw_plan();
    if (!P) return fail(SCW_ERR_NOMEM, "host allocation failed");
    P->device = device;
    HostTables H;
    try {
        st = plan_host(P, waves, n_waves, z_layers, layers, waves_begin, waves_end, H);
    } catch (const std::bad_alloc&) {
        delete P;
        return fail(SCW_ERR_NOMEM, "host allocation failed while planning");
    }
    if (st != SCW_OK) {
        delete P;
        return st;
    }
    const std::vector<double>& k64 = H.k64;
    const std::vector<double>& e64 = H.e64;
    const std::vector<double>& lns = H.lns;
    const std::vector<float>& kf = H.kf;
    const std::vector<float>& ellA = H.ellA;
    const std::vector<int32_t>& ellC = H.ellC;
    const std::vector<unsigned char>& dreg = H.dreg;
    if ((st = P->d_k64.upload(k64.data(), k64.size() * sizeof(double))) != SCW_OK ||
        (st = P->d_e64.upload(e64.data(), e64.size() * sizeof(double))) != SCW_OK ||
        (st = P->d_lns.upload(lns.data(), lns.size() * sizeof(double))) != SCW_OK ||
        (st = P->d_kf.upload(kf.data(), kf.size() * sizeof(float))) != SCW_OK ||
        (st = P->d_ellA.upload(ellA.data(), ellA.size() * sizeof(float))) != SCW_OK ||
        (st = P->d_ellC.upload(ellC.data(), ellC.size() * sizeof(int32_t))) != SCW_OK ||
        (st = P->d_dreg.upload(dreg.data(), dreg.size())) != SCW_OK ||
        (st = P->d_zl.upload(P->zl.data(), P->zl.size() * sizeof(double))) != SCW_OK) {
        scw_plan_destroy(P);
        return st;
    }
    *out = P;
    return SCW_OK;
}

int scw_plan_destroy(scw_plan* plan) {
    if (!plan) return SCW_OK;
    {
        DeviceGuard g(plan->device);
        for (auto& e : plan->plane_cache) scw_exec_destroy(e.ex);
        plan->plane_cache.clear();
        plan->d_dreg.release();
        plan->d_zl.release();
        plan->d_k64.release();
        plan->d_e64.release();
        plan->d_lns.release();
        plan->d_kf.release();
        plan->d_ellA.release();
        plan->d_ellC.release();
        plan->d_bfp.release();
    }
    delete plan;
    return SCW_OK;
}

int scw_plan_layers(const scw_plan* plan, int* layers) {
    if (!plan || !layers) return fail(SCW_ERR_ARG, "NULL argument");
    *layers = plan->layers;
    return SCW_OK;
}

int scw_plan_region_info(const scw_plan* plan, int region, int64_t* n_waves, int32_t* mx, int32_t* my,
                         int32_t* mc, int64_t* n_terms, int32_t* width, int32_t* separable) {
    if (!plan) return fail(SCW_ERR_ARG, "plan is NULL");
    if (region < 0 || region > plan->layers) return fail(SCW_ERR_ARG, "region %d out of range", region);
    const RegionPlan& R = plan->regions[region];
    if (n_waves) *n_waves = R.wend - R.wbegin;
    if (mx) *mx = R.mx;
    if (my) *my = R.my;
    if (mc) *mc = R.mc;
    if (n_terms) *n_terms = R.n_terms;
    if (width) *width = R.width;
    if (separable) *separable = R.separable ? 1 : 0;
    return SCW_OK;
}

int scw_plan_device_bytes(const scw_plan* plan, size_t* bytes) {
    if (!plan || !bytes) return fail(SCW_ERR_ARG, "NULL argument");
    *bytes = plan->d_k64.bytes + plan->d_e64.bytes + plan->d_lns.bytes + plan->d_kf.bytes +
             plan->d_ellA.bytes + plan->d_ellC.bytes + plan->d_bfp.bytes + plan->d_dreg.bytes + plan->d_zl.bytes;
    return SCW_OK;
}

int scw_test_plan_host(const scw_planewave64* waves, size_t n_waves, const double* z_layers, int layers,
                       const int32_t* waves_begin, const int32_t* waves_end, int64_t* info) {
    if (!info) return fail(SCW_ERR_ARG, "info is NULL");
    int st = validate_tables(waves, n_waves, z_layers, layers, waves_begin, waves_end);
    if (st != SCW_OK) return st;
    scw_plan P;  // host fields only: its device buffers stay empty
    HostTables H;
    try {
        st = plan_host(&P, waves, n_waves, z_layers, layers, waves_begin, waves_end, H);
    } catch (const std::bad_alloc&) {
        return fail(SCW_ERR_NOMEM, "host allocation failed while planning");
    }
    if (st != SCW_OK) return st;
    for (int r = 0; r <= layers; ++r) {
        const RegionPlan& R = P.regions[r];
        int64_t* o = info + 8 * r;
        o[0] = R.separable ? 1 : 0;
        o[1] = R.dense ? 1 : 0;
        o[2] = R.mx;
        o[3] = R.my;
        o[4] = R.mc;
        o[5] = R.width;
        o[6] = R.n_terms;
        o[7] = R.wend - R.wbegin;
    }
    return SCW_OK;
}

}  // extern "C"
