// Grid execution: binds a plan to one rectilinear grid, form and dtype.
//
// Replaces the per-call work of gpu_cw_evaluate (gpuEvaluator.cu:131-165):
// instead of cudaMemGetInfo + cudaMalloc + three launches + three blocking
// D2H copies + cudaFree on every call, all grid tables and scratch are built
// here once (FP64 on the host, rounded once), and scw_exec_run only launches
// kernels on the caller's stream — so a run can be captured in a hipGraph.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>

#include "scw_internal.hpp"

namespace scw {

static inline int region_of(double z, const std::vector<double>& zl, double tol) {
    int r = 0;
    for (size_t l = 0; l < zl.size(); ++l)
        if (z >= zl[l] - tol) r = (int)l + 1;
    return r;
}

static inline uint16_t bf16_rne(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

static inline float bf16_f32(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

static inline uint16_t f16_rne(float f) {
    _Float16 h = (_Float16)f;
    uint16_t r;
    std::memcpy(&r, &h, 2);
    return r;
}

static inline int round_up(int v, int m) { return (v + m - 1) / m * m; }

// Launch-shape defaults, overridable through the environment for tuning sweeps.
static int64_t tunable(const char* name, int64_t dflt) {
    const char* v = std::getenv(name);
    if (!v) return dflt;
    const long long x = std::atoll(v);
    return x > 0 ? (int64_t)x : dflt;
}

static int prep_direct(scw_exec* ex, const double* xs, const double* ys, const double* zs) {
    const scw_plan* P = ex->plan;
    const double xmin = *std::min_element(xs, xs + ex->nx), xmax = *std::max_element(xs, xs + ex->nx);
    const double ymin = *std::min_element(ys, ys + ex->ny), ymax = *std::max_element(ys, ys + ex->ny);
    ex->xc = 0.5 * (xmin + xmax);
    ex->yc = 0.5 * (ymin + ymax);
    std::vector<float> dx(ex->nx), dy(ex->ny);
    for (int64_t i = 0; i < ex->nx; ++i) dx[i] = (float)(xs[i] - ex->xc);
    for (int64_t i = 0; i < ex->ny; ++i) dy[i] = (float)(ys[i] - ex->yc);
    std::vector<int32_t> zr(ex->zregion.begin(), ex->zregion.end());
    int64_t maxw = 0;
    for (int64_t z = 0; z < ex->nz; ++z) {
        const RegionPlan& R = P->regions[ex->zregion[z]];
        maxw = std::max(maxw, R.wend - R.wbegin);
    }
    const int64_t pts = ex->nx * ex->ny;
    ex->direct_ppt = direct_ppt(pts);
    const int split = direct_split(pts, ex->nz, maxw, ex->direct_ppt);
    ex->direct_split = split;
    int st;
    if ((st = ex->d_dx.upload(dx.data(), dx.size() * 4)) != SCW_OK) return st;
    if ((st = ex->d_dy.upload(dy.data(), dy.size() * 4)) != SCW_OK) return st;
    if ((st = ex->d_z.upload(zs, ex->nz * sizeof(double))) != SCW_OK) return st;
    if ((st = ex->d_zslab.upload(zr.data(), zr.size() * 4)) != SCW_OK) return st;
    if (split > 1 && (st = ex->d_C.alloc((size_t)split * pts * ex->nz * 24)) != SCW_OK) return st;
    ex->flops = 36.0 * ex->units;
    ex->bytes = 24.0 * (double)pts * ex->nz;
    ex->stage_flops[0] = ex->stage_flops[1] = ex->flops;
    ex->stage_bytes[0] = ex->stage_bytes[1] = ex->bytes;
    ex->launches = split > 1 ? 2 : 1;
    return SCW_OK;
}

// One grid z of a GEMM-form run: the plan and region it reads and its z.
struct ZEntry {
    const scw_plan* plan;
    int region;
    double z;
};

static inline bool region_ok(const cd v) {
    return std::isfinite(v.real()) && std::isfinite(v.imag()) && std::abs(v) <= 1e30;
}

// Kernel launches of a stage-by-stage GEMM-form run.
static int stage_launches(const scw_exec* ex) {
    int n = ex->items.empty() ? 0 : 1;
    for (int k = 0; k < 4; ++k) n += ex->mclass_off[k + 1] > ex->mclass_off[k] ? 1 : 0;
    const bool xstage = ex->dtype == SCW_DTYPE_FP32 ? ex->my_max > 0 : ex->ncblk > 0;  // u_mfma / slot sum
    return n + (xstage ? 1 : 0) + (int)std::max<size_t>(1, ex->mk_launch.size() / 3);
}

// Work units of the single-launch kernel (kinds and fields as fused16 in
// scw_gemm.hip reads them), the counts its waits compare against (per z: the
// wave items covering z, then the slot-sum units of z), its zeroed control
// words, its grid (the resident capacity, or fewer) and its wait limit.
//
// Order: the coefficient units sorted by their first z; the slot-sum units
// of z placed `lag` coefficient units after the last one covering z, and the
// main units of z `lag` units after that, so that by the time a workgroup
// takes a unit, what it waits for was taken about `lag` units earlier and is
// most likely done (a smaller lag makes waiting workgroups pile up behind
// unfinished items; an unbounded one is the stage order).  Every unit comes
// after the units it depends on.
static int build_units(scw_exec* ex) {
    enum { kCoef = 0, kModal1 = 1, kModalWg = 4, kCsum = 5, kMain = 6, kCtlWords = 4 };
    const int64_t nz = ex->nz;
    const int64_t gx = (6 * ex->nx + 255) / 256;
    const int tiles = (int)std::max<int64_t>(1, std::min<int64_t>(255, ex->main_tiles > 0 ? ex->main_tiles
                                                                                       : tunable("SCW_SINGLE_MAIN_TILES", 2)));
    const int cgroup = (int)std::max<int64_t>(1, tunable("SCW_SINGLE_CSUM_BLOCKS", 8));
    const int64_t n = ((int64_t)ex->items.size() + 3) / 4 + ((int64_t)ex->mitems.size() + 12) / 4 +
                      (int64_t)ex->hcblk.size() + nz + gx * nz * ex->ysplit;
    if (n > (int64_t)INT32_MAX - (1 << 20))
        return fail(SCW_ERR_UNSUPPORTED, "single-launch form: %lld work units exceed 32-bit indices", (long long)n);
    std::vector<int> need((size_t)(2 * nz), 0);
    // coefficient units with their first z
    std::vector<std::pair<int, int4>> cu;
    const int nitems = (int)ex->items.size();
    for (int i = 0; i < nitems; i += 4) cu.push_back({ex->items[i].z0, make_int4(kCoef, i, nitems, 0)});
    for (const CoefItem& it : ex->items)
        for (int z = it.z0; z < it.z0 + it.nz; ++z) ++need[z];
    // items per unit: four (a wave each), or the 64-z class's workgroup
    const int ipw3 = ex->dtype != SCW_DTYPE_FP32 ? ex->modal_wg_items : 4;
    for (int k = 0; k < 4; ++k) {
        const int off = (int)ex->mclass_off[k], end = (int)ex->mclass_off[k + 1];
        const int per = k == 3 ? ipw3 : 4;
        if (k == 3 && (end - off) % per != 0 && ex->dtype != SCW_DTYPE_FP32)
            return fail(SCW_ERR_ARG, "internal: 64-z mode items not in whole workgroups");
        for (int i = off; i < end; i += per) {
            if (ex->mitems[i].nz == 0) continue;  // an empty workgroup of the XCD-grouped order
            int z0 = ex->mitems[i].z0;
            for (int w = 1; w < per && i + w < end; ++w) z0 = std::min(z0, ex->mitems[i + w].z0);
            cu.push_back({z0, make_int4(k < 3 ? kModal1 + k : kModalWg, i, end, 0)});
        }
    }
    for (const ModalItem& it : ex->mitems)
        for (int z = it.z0; z < it.z0 + it.nz; ++z) ++need[z];
    std::stable_sort(cu.begin(), cu.end(), [](const std::pair<int, int4>& a, const std::pair<int, int4>& b) {
        return a.first < b.first;
    });
    // position of the last coefficient unit covering each z
    std::vector<int64_t> last((size_t)nz, -1);
    for (size_t p = 0; p < cu.size(); ++p) {
        const int4 U = cu[p].second;
        const int end = U.z;
        for (int w = 0; w < (U.x == kModalWg ? ipw3 : 4) && U.y + w < end; ++w) {
            int z0, zn;
            if (U.x == kCoef) {
                z0 = ex->items[U.y + w].z0;
                zn = ex->items[U.y + w].nz;
            } else {
                z0 = ex->mitems[U.y + w].z0;
                zn = ex->mitems[U.y + w].nz;
            }
            for (int z = z0; z < z0 + zn; ++z) last[z] = std::max<int64_t>(last[z], (int64_t)p);
        }
    }
    const int64_t lag = std::max<int64_t>(0, ex->order_lag >= 0 ? ex->order_lag : tunable("SCW_SINGLE_LAG", 1 << 30));
    // (key, unit): key = 4 x position in the coefficient stream + kind rank
    std::vector<std::pair<int64_t, int4>> all;
    all.reserve((size_t)n);
    for (size_t p = 0; p < cu.size(); ++p) all.push_back({4 * (int64_t)p, cu[p].second});
    std::vector<int64_t> ckey((size_t)nz, -4);
    for (size_t c = 0; c < ex->hcblk.size();) {
        const int z = ex->hcblk[c].x;
        size_t e = c;
        while (e < ex->hcblk.size() && ex->hcblk[e].x == z && (int)(e - c) < cgroup) ++e;
        const int64_t pos = last[z] < 0 ? -1 : last[z] + lag;
        all.push_back({4 * pos + 1, make_int4(kCsum, (int)c, (int)(e - c), 0)});
        ckey[z] = std::max(ckey[z], pos);
        ++need[nz + z];
        c = e;
    }
    // z whose mode contraction stores in place: one slot-sum unit without
    // blocks (z in its last field), which waits for the z's items and passes
    // the count on to the main units
    std::vector<char> zin((size_t)nz, 0);
    for (const ModalItem& it : ex->mitems)
        if (it.cs_direct)
            for (int z = it.z0; z < it.z0 + it.nz; ++z) zin[z] = 1;
    for (int z = 0; z < nz; ++z)
        if (zin[z]) {
            const int64_t pos = last[z] < 0 ? -1 : last[z] + lag;
            all.push_back({4 * pos + 1, make_int4(kCsum, -1, 0, z)});
            ckey[z] = std::max(ckey[z], pos);
            ++need[nz + z];
        }
    for (int z = 0; z < nz; ++z) {
        const int64_t pos = ckey[z] < -1 ? -1 : ckey[z] + lag;
        for (int bz = 0; bz < ex->ysplit; ++bz)
            for (int bx = 0; bx < gx; bx += tiles)
                all.push_back({4 * pos + 2, make_int4(kMain | ((int)std::min<int64_t>(tiles, gx - bx) << 8), bx, z, bz)});
    }
    std::stable_sort(all.begin(), all.end(), [](const std::pair<int64_t, int4>& a, const std::pair<int64_t, int4>& b) {
        return a.first < b.first;
    });
    std::vector<int4> units(all.size());
    for (size_t k = 0; k < all.size(); ++k) units[k] = all[k].second;
    if (ex->stall) ++need[nz];  // the main units of z-slice 0 wait for a slot sum that never comes
    int st;
    if ((st = ex->d_units.upload(units.data(), units.size() * sizeof(int4))) != SCW_OK) return st;
    if ((st = ex->d_need.upload(need.data(), need.size() * sizeof(int))) != SCW_OK) return st;
    const size_t ctl = (size_t)(kCtlWords + 2 * nz) * sizeof(int);
    if ((st = ex->d_ctl.alloc(ctl)) != SCW_OK) return st;
    SCW_HIP(hipMemset(ex->d_ctl.p, 0, ctl));
    // complete before the first run, whatever stream the caller runs on
    SCW_HIP(hipStreamSynchronize(nullptr));
    ex->nunits = (int64_t)units.size();
    int per_cu = 0, cus = 0, khz = 0, dev = 0;
    if ((st = fused_launch(ex, nullptr, nullptr, &per_cu)) != SCW_OK) return st;
    SCW_HIP(hipGetDevice(&dev));
    SCW_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    SCW_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    const int64_t cap = (int64_t)per_cu * cus;
    if (cap < 1) return fail(SCW_ERR_DEVICE, "single-launch kernel: no workgroup fits a CU");
    // at most the resident capacity: the first unit of workgroup b is unit b,
    // which must not wait behind workgroups that cannot start
    ex->pgrid = std::min<int64_t>(ex->nunits, ex->tune_grid > 0 ? std::min(ex->tune_grid, cap) : cap);
    ex->wait_ticks = (uint64_t)std::max(1.0, ex->wait_ms * (double)std::max(khz, 1));
    ex->wait_polls = (uint64_t)std::max(1000.0, ex->wait_ms * 1.0e5);
    ex->launches = 1;
    return SCW_OK;
}

// Dispatch order of workgroup groups over nxcd XCDs (workgroup b of a launch
// runs on XCD b mod nxcd, each XCD with its own L2).  Group g has
// groups[g].first workgroups of groups[g].second steps each, numbered
// consecutively in group order.  Every group goes whole to one XCD: groups
// sorted longest first go to the least-loaded XCD (longest processing time
// first), and the XCD queues are interleaved back into dispatch order,
// position k * nxcd + x holding the k-th workgroup of XCD x's queue, or -1
// past the end of a shorter queue.
static std::vector<int64_t> xcd_group_order(const std::vector<std::pair<int64_t, int64_t>>& groups, int nxcd) {
    std::vector<int64_t> first(groups.size() + 1, 0), gi(groups.size());
    for (size_t g = 0; g < groups.size(); ++g) {
        first[g + 1] = first[g] + groups[g].first;
        gi[g] = (int64_t)g;
    }
    std::stable_sort(gi.begin(), gi.end(),
                     [&](int64_t a, int64_t b) { return groups[(size_t)a].second > groups[(size_t)b].second; });
    std::vector<std::vector<int64_t>> qx((size_t)nxcd);
    std::vector<int64_t> load((size_t)nxcd, 0);
    for (const int64_t g : gi) {
        const size_t x = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
        for (int64_t k = 0; k < groups[(size_t)g].first; ++k) qx[x].push_back(first[(size_t)g] + k);
        load[x] += groups[(size_t)g].first * groups[(size_t)g].second;
    }
    size_t depth = 0;
    for (const auto& q : qx) depth = std::max(depth, q.size());
    std::vector<int64_t> order;
    order.reserve(depth * (size_t)nxcd);
    for (size_t k = 0; k < depth; ++k)
        for (int x = 0; x < nxcd; ++x) order.push_back(k < qx[(size_t)x].size() ? qx[(size_t)x][k] : -1);
    return order;
}

// Builds every GEMM-form table of a run over the z entries ze (one per output
// z-slice, in output order).  Regions are identified by (plan, region) pairs,
// so a run may mix structures (batches).
static int prep_gemm(scw_exec* ex, const double* xs, const double* ys, const std::vector<ZEntry>& ze) {
    const int nx = (int)ex->nx, ny = (int)ex->ny, nz = (int)ze.size();
    const bool is16 = ex->dtype != SCW_DTYPE_FP32;
    const int kq = is16 ? 32 : 16;  // K granule: one MFMA (16-bit) or one 16-byte Y load of four k-steps (FP32)
    const int ytiles = (ny + 15) / 16;

    // Virtual regions: distinct (plan, region) pairs in order of first use.
    std::vector<std::pair<const scw_plan*, int>> vreg;
    std::vector<int> zv(nz);
    for (int z = 0; z < nz; ++z) {
        int v = -1;
        for (int k = (int)vreg.size() - 1; k >= 0; --k)
            if (vreg[k].first == ze[z].plan && vreg[k].second == ze[z].region) {
                v = k;
                break;
            }
        if (v < 0) {
            v = (int)vreg.size();
            vreg.push_back({ze[z].plan, ze[z].region});
        }
        zv[z] = v;
    }
    const int NV = (int)vreg.size();
    std::vector<RegionDev> rd(NV);
    std::memset(rd.data(), 0, rd.size() * sizeof(RegionDev));
    std::vector<float> X;  // float2 pairs
    std::vector<int64_t> yfrag_off(NV, 0);
    struct Chunk {
        int b0, my, kp;
        int64_t yfrag_off;
    };
    std::vector<std::vector<Chunk>> chunks(NV);
    std::vector<float> yf32;
    std::vector<uint16_t> yhi, ylo;
    int kmax = 0;
    ex->my_max = 0;
    ex->ka_max = 0;
    for (int v = 0; v < NV; ++v) {
        const scw_plan* P = vreg[v].first;
        const int r = vreg[v].second;
        const RegionPlan& R = P->regions[r];
        if (!R.separable)
            return fail(SCW_ERR_UNSUPPORTED, "GEMM form: region %d is not separable (%s)", r, R.why_not.c_str());
        RegionDev& d = rd[v];
        d.ellA = P->d_ellA.as<float2>() + R.ell_off;
        d.arow = (int32_t)R.arow;
        d.bfp = d.bfp_exp = nullptr;
        if (is16 && R.dense && R.width > 0) {  // the 16-bit mode contraction reads block floating point
            int st2 = build_bfp(const_cast<scw_plan*>(P));
            if (st2 != SCW_OK) return st2;
            d.bfp = P->d_bfp.as<uint8_t>() + P->bfp_off[r];
            d.bfp_exp = d.bfp + 4LL * ((R.width + 15) / 16) * R.arow * 16;
        }
        d.ellC = R.ellc_off >= 0 ? P->d_ellC.as<int32_t>() + R.ellc_off : nullptr;
        d.cells = R.mx * R.my;
        d.width = R.width;
        d.mc = R.mc;
        d.dense = R.dense ? 1 : 0;
        d.mx = R.mx;
        d.my = R.my;
        d.nslots = 1;
        ex->my_max = std::max(ex->my_max, R.my);
        if (R.my > 0) ex->ka_max = std::max(ex->ka_max, (R.mx + 15) / 16 * 16);
        // X_r[a][x] = exp(i kx_a x): [mx][nx] for the FP32 X-contraction
        // kernel; x-major [nx][ka] (zero past mx) for the 16-bit dtypes, whose
        // main GEMM reads four consecutive a of one x per lane
        d.ka = (R.mx + 15) / 16 * 16;
        d.erow = (int32_t)bfp_erow(R.arow);
        d.x_off = (int64_t)X.size() / 2;
        X.resize(X.size() + (size_t)2 * (is16 ? (int64_t)d.ka : R.mx) * nx, 0.0f);
        for (int a = 0; a < R.mx; ++a)
            for (int i = 0; i < nx; ++i) {
                const cd val = std::exp(cd(0, 1) * R.kx[a] * xs[i]);
                if (!region_ok(val)) return fail(SCW_ERR_UNSUPPORTED, "region %d: exp(i kx x) outside FP32 range", r);
                const int64_t e = d.x_off + (is16 ? (int64_t)i * d.ka + a : (int64_t)a * nx + i);
                X[2 * e] = (float)val.real();
                X[2 * e + 1] = (float)val.imag();
            }
        // ky chunks of the main GEMM: K = 2 my <= 128 per launch, so a region
        // of more than 64 ky runs it over balanced chunks of ky rows
        const int nch = R.my > 64 ? (R.my + 63) / 64 : 1;
        const int myc = (R.my + nch - 1) / nch;
        for (int c = 0; c < nch; ++c) {
            const int b0 = c * myc, bc = std::max(0, std::min(myc, R.my - b0));
            chunks[v].push_back({b0, bc, bc > 0 ? round_up(2 * bc, kq) : 0, 0});
        }
        // Y fragments per chunk: Yp[y][2b] = Re exp(i ky_(b0 + b) y), Yp[y][2b+1] = Im.
        for (Chunk& ch : chunks[v]) {
            const int kp = ch.kp;
            kmax = std::max(kmax, kp);
            std::vector<double> Yp((size_t)ny * std::max(kp, 1), 0.0);
            for (int j = 0; j < ny; ++j)
                for (int b = 0; b < ch.my; ++b) {
                    const cd val = std::exp(cd(0, 1) * R.ky[ch.b0 + b] * ys[j]);
                    if (!region_ok(val))
                        return fail(SCW_ERR_UNSUPPORTED, "region %d: exp(i ky y) outside FP32 range", r);
                    Yp[(size_t)j * kp + 2 * b] = val.real();
                    Yp[(size_t)j * kp + 2 * b + 1] = val.imag();
                }
            auto yval = [&](int y, int k) { return (y < ny && k < kp) ? Yp[(size_t)y * kp + k] : 0.0; };
            if (!is16) {
                // [yt][q][lane][e]: lane l holds k = 4 (4q + e) + (l >> 4) of row 16 yt + (l & 15)
                ch.yfrag_off = (int64_t)yf32.size();
                const int K16 = kp / 16;
                for (int yt = 0; yt < ytiles; ++yt)
                    for (int q = 0; q < K16; ++q)
                        for (int l = 0; l < 64; ++l)
                            for (int e = 0; e < 4; ++e)
                                yf32.push_back((float)yval(16 * yt + (l & 15), 4 * (4 * q + e) + (l >> 4)));
            } else {
                // [yt][s][lane][e]: lane l holds k = 32 s + 8 (l >> 4) + e of row 16 yt + (l & 15)
                ch.yfrag_off = (int64_t)yhi.size();
                const int K32 = kp / 32;
                for (int yt = 0; yt < ytiles; ++yt)
                    for (int s2 = 0; s2 < K32; ++s2)
                        for (int l = 0; l < 64; ++l)
                            for (int e = 0; e < 8; ++e) {
                                const double val = yval(16 * yt + (l & 15), 32 * s2 + 8 * (l >> 4) + e);
                                if (ex->dtype == SCW_DTYPE_FP16) {
                                    yhi.push_back(f16_rne((float)val));
                                    ylo.push_back(0);
                                } else {
                                    const uint16_t h = bf16_rne((float)val);
                                    yhi.push_back(h);
                                    ylo.push_back(bf16_rne((float)(val - (double)bf16_f32(h))));
                                }
                            }
            }
        }
        yfrag_off[v] = chunks[v][0].yfrag_off;
    }
    if (kmax > 128) return fail(SCW_ERR_ARG, "internal: a ky chunk of the main GEMM past K = 128");
    if (!is16)
        ex->kmax = kmax <= 16 ? 16 : kmax <= 32 ? 32 : kmax <= 64 ? 64 : 128;
    else
        ex->kmax = kmax <= 32 ? 32 : kmax <= 64 ? 64 : kmax <= 96 ? 96 : 128;
    // The FP32 MFMA GEMM reads one z-slice of U (up to kmax / 2 rows of 6 nx
    // floats, plus a lane overhang past the last column) with 32-bit offsets.
    if (!is16 && ((double)ex->kmax / 2 * 6.0 * nx + 512.0) * 4.0 >= 2147483648.0)
        return fail(SCW_ERR_UNSUPPORTED, "GEMM form: nx = %d is too large for %d distinct ky (U slice over 2 GiB)",
                    nx, kmax / 2);

    // FP16 (5 exponent bits) cannot hold both factors of an evanescent mode:
    // across a slab of height h, |exp(i kz (z - zref))| spans e^(+-Im(kz) h / 2)
    // and the amplitude re-referenced to zref carries the inverse factor (e^12
    // and e^-12 for a 4 um slab at 0.4 um).  Mode c of a dense region is
    // therefore scaled by 2^-s_c in Z' and by 2^s_c in its amplitudes (an
    // exec-owned copy of the plan's exponent plane), with s_c chosen so that
    // the largest |Z_c| over the grid's z lies in [1/2, 1).
    std::vector<std::vector<int32_t>> zshift(NV);
    if (ex->dtype == SCW_DTYPE_FP16) {
        int64_t words = 0, nshift = 0;
        std::vector<int64_t> woff(NV, -1), soff(NV, -1);
        for (int v = 0; v < NV; ++v) {
            if (!rd[v].bfp) continue;
            const RegionPlan& R = vreg[v].first->regions[vreg[v].second];
            const int64_t groups = 4LL * ((R.width + 15) / 16);
            std::vector<double> lmax(R.width, -HUGE_VAL);  // ln max |Z_c|
            for (int z = 0; z < nz; ++z)
                if (zv[z] == v)
                    for (int c = 0; c < R.width; ++c)
                        lmax[c] = std::max(lmax[c], -R.kz[c].imag() * (ze[z].z - R.zref));
            zshift[v].assign((size_t)groups * 4, 0);
            for (int c = 0; c < R.width; ++c)
                zshift[v][c] = (int32_t)std::min(120.0, std::max(-120.0, std::floor(lmax[c] / M_LN2) + 1.0));
            woff[v] = words;
            words += groups * bfp_erow(R.arow);
            soff[v] = nshift;
            nshift += groups * 4;
        }
        if (words > 0) {
            int st2;
            if ((st2 = ex->d_bfpx.alloc((size_t)(words + nshift) * 4)) != SCW_OK) return st2;
            std::vector<int32_t> shifts((size_t)nshift);
            for (int v = 0; v < NV; ++v)
                if (soff[v] >= 0) std::copy(zshift[v].begin(), zshift[v].end(), shifts.begin() + soff[v]);
            int32_t* dshift = ex->d_bfpx.as<int32_t>() + words;
            SCW_HIP(hipMemcpy(dshift, shifts.data(), shifts.size() * 4, hipMemcpyHostToDevice));
            for (int v = 0; v < NV; ++v) {
                if (woff[v] < 0) continue;
                const RegionPlan& R = vreg[v].first->regions[vreg[v].second];
                const int64_t n = 4LL * ((R.width + 15) / 16) * bfp_erow(R.arow);
                uint8_t* dst = ex->d_bfpx.as<uint8_t>() + woff[v] * 4;
                if ((st2 = shift_bfp_exponents(rd[v].bfp_exp, dst, bfp_erow(R.arow), n, dshift + soff[v])) != SCW_OK)
                    return st2;
                rd[v].bfp_exp = dst;
            }
        } else {
            ex->d_bfpx.release();
        }
    }

    // Runs: consecutive z of one virtual region.  Dense eigenmode regions
    // (every cell holds every kz, term index = kz index) contract on MFMA in
    // runs of up to 64 z, every term read once per run; other regions use the
    // VALU ELL contraction in runs of up to 8 z.  A maximal stretch longer
    // than 64 z becomes several runs of one group, whose items interleave so
    // that the runs of a stretch read the same terms at the same time.
    struct Run {
        int z0, z1, v;
        bool modal;
        int group;
    };
    std::vector<Run> runs;
    // Shortest stretch contracted on MFMA: every stretch for the 16-bit
    // dtypes; for FP32 (MFMA at the VALU's FP32 rate) only stretches the VALU
    // path would have to read more than once.
    const int64_t modal_min = tunable("SCW_MODAL_MIN_RUN", is16 ? 1 : 9);
    int ngroups = 0;
    for (int z = 0; z < nz;) {
        const int v = zv[z];
        int zend = z + 1;  // maximal stretch of consecutive z in region v
        while (zend < nz && zv[zend] == v) ++zend;
        const bool modal = rd[v].dense && rd[v].width > 0 && zend - z >= modal_min;
        // z-run tile of a work item: up to 64 z for the mode contraction
        // (8 row tiles of 8 z), up to 8 z for the VALU contraction
        const int cap = modal ? (ex->tune_modal_run > 0 ? (int)ex->tune_modal_run : 64)
                              : (ex->tune_coef_run > 0 ? (int)ex->tune_coef_run : 8);
        const int nrun = (zend - z + cap - 1) / cap;  // balanced runs
        for (int k = 0; k < nrun; ++k)
            runs.push_back({z + (int)((int64_t)(zend - z) * k / nrun), z + (int)((int64_t)(zend - z) * (k + 1) / nrun),
                            v, modal, ngroups});
        ++ngroups;
        z = zend;
    }
    // Slices of the term (or mode) range: enough wave items to cover every
    // SIMD several times, with at least 16 terms per slice.
    std::vector<int> mper(NV, 16);  // mode-slice width of dense regions
    {
        int64_t base = 0;
        for (const Run& rn : runs) {
            const RegionDev& d = rd[rn.v];
            base += rn.modal ? (3LL * d.cells + 63) / 64 : (d.cells + 63) / 64;
        }
        const int64_t target = ex->tune_coef_items > 0 ? ex->tune_coef_items : tunable("SCW_COEF_ITEMS", 4096);
        const int want = base > 0 ? (int)std::min<int64_t>(64, (target + base - 1) / base) : 1;
        for (int v = 0; v < NV; ++v) {
            int n = std::max(1, std::min(want, rd[v].width / 16));
            if (rd[v].dense && rd[v].width > 0) {
                // mode slices are whole 16-mode steps; every slot gets modes
                mper[v] = round_up((rd[v].width + n - 1) / n, 16);
                n = (rd[v].width + mper[v] - 1) / mper[v];
            }
            rd[v].nslots = n;
            ex->nslots_max = std::max(ex->nslots_max, n);
        }
    }
    // 16-bit dtypes: regions of at most 4 terms per cell (the half-spaces and
    // layer stacks: one wave per direction and order) get no coefficient
    // items; the slot-sum kernel sums their terms itself (nslots = 0 marks
    // them), which saves a launch and a write and read of C_z.
    std::vector<char> fold(NV, 0);
    for (int v = 0; v < NV; ++v)
        if (is16 && !rd[v].dense && rd[v].width > 0 && rd[v].width <= 4 && rd[v].cells > 0) {
            fold[v] = 1;
            rd[v].nslots = 0;
        }
    // 16-bit dtypes: a mode-contraction run of a region with one mode slice
    // stores its sums straight into the slot-summed layout the main GEMM
    // reads ([3 my][ka], zero past mx), so its z need no slot sum, which
    // would only copy them (SCW_CS_DIRECT=0 keeps the copy).
    const char* csd_env = std::getenv("SCW_CS_DIRECT");
    const bool cs_direct_ok = is16 && (csd_env == nullptr || std::atoi(csd_env) != 0);
    std::vector<char> zdirect((size_t)nz, 0);
    for (const Run& rn : runs)
        if (cs_direct_ok && rn.modal && rd[rn.v].nslots == 1 && !fold[rn.v])
            for (int z = rn.z0; z < rn.z1; ++z) zdirect[z] = 1;

    // Per-z metadata, Z rows, C and U offsets.
    std::vector<ZMeta> zm(nz);
    std::vector<float> Z;
    int64_t u_off = 0, c_off = 0;
    double f_coef = 0, f_u = 0, f_main = 0;
    double b_coef = 0, b_u = 0, b_main = 24.0 * (double)nx * ny * nz;
    std::vector<char> x_counted(NV, 0);  // X_r is needed once per run, whatever the number of its z
    for (int z = 0; z < nz; ++z) {
        const int v = zv[z];
        const RegionPlan& R = vreg[v].first->regions[vreg[v].second];
        ZMeta& m = zm[z];
        m.region = v;
        m.my = R.my;
        m.mx = R.mx;
        m.kp = chunks[v][0].kp;
        m.u_off = u_off;
        m.zout = z;
        m.acc = 0;
        m.yfrag_off = yfrag_off[v];
        m.zrow_off = (int64_t)Z.size() / 2;
        m.c_off = c_off;
        // 16-bit dtypes: every Cs_z on its own 256-byte lines (the
        // single-launch kernel reads them through the caches once final)
        u_off += is16 ? (3LL * R.my * rd[v].ka + 31) / 32 * 32 : (int64_t)R.my * nx * 3;
        c_off += (int64_t)R.mx * R.my * 3;
        for (int c = 0; c < R.mc; ++c) {
            const cd val = std::exp(cd(0, 1) * R.kz[c] * (ze[z].z - R.zref));
            if (!region_ok(val))
                return fail(SCW_ERR_UNSUPPORTED, "region %d: exp(i kz (z - zref)) outside FP32 range at z=%g",
                            vreg[v].second, ze[z].z);
            Z.push_back((float)val.real());
            Z.push_back((float)val.imag());
        }
        (fold[v] ? f_u : f_coef) += 24.0 * (double)R.n_terms;
        f_main += 24.0 * (double)R.my * nx * ny;
        const double cbytes = 24.0 * (double)R.mx * R.my;  // C_z, one slot
        b_coef += cbytes * rd[v].nslots;
        if (!is16) {
            // X contraction kernel: slots and X in, U_z out; the GEMM reads U_z
            const double ubytes = 24.0 * (double)R.my * nx;
            f_u += 24.0 * (double)R.mx * R.my * nx;
            b_u += cbytes * rd[v].nslots + ubytes + (x_counted[v] ? 0.0 : 8.0 * R.mx * nx);
            b_main += ubytes;
            x_counted[v] = 1;
        } else {
            // slot sum: slots in (or, folded regions, their terms: 3 float2
            // and a kz index per term, and the Z row), C_z out; the GEMM reads
            // C_z and X and runs the X contraction of its columns itself
            f_main += 24.0 * (double)R.mx * R.my * nx;
            if (!zdirect[z]) b_u += cbytes * rd[v].nslots + cbytes;
            if (fold[v]) b_u += 28.0 * (double)rd[v].cells * rd[v].width + 8.0 * R.mc;
            b_main += cbytes + (x_counted[v] ? 0.0 : 8.0 * R.mx * nx);
            x_counted[v] = 1;
        }
    }
    ex->c_elems = c_off;
    // Main-GEMM entries: level L holds chunk L of every z whose region has
    // one (level 0: every z, in z order); each level is one launch, and the
    // levels after the first add to the output.
    std::vector<ZMeta> mzm;
    ex->mlevel_off.assign(1, 0);
    size_t levels = 1;
    for (int v = 0; v < NV; ++v) levels = std::max(levels, chunks[v].size());
    for (size_t L = 0; L < levels; ++L) {
        for (int z = 0; z < nz; ++z) {
            const int v = zv[z];
            if (chunks[v].size() <= L) continue;
            const Chunk& ch = chunks[v][L];
            ZMeta e = zm[z];
            e.my = ch.my;
            e.kp = ch.kp;
            e.yfrag_off = ch.yfrag_off;
            e.u_off = zm[z].u_off + (is16 ? 3LL * ch.b0 * rd[v].ka : (int64_t)ch.b0 * nx * 3);
            e.acc = L > 0 ? 1 : 0;
            mzm.push_back(e);
        }
        ex->mlevel_off.push_back((int64_t)mzm.size());
    }
    // The FP32 stage-form main GEMM runs each level as one launch per kernel
    // K (largest first), over the entries regrouped by the K their chunk
    // needs: a batch mixes order counts, and the FP32 kernels of smaller K
    // hold fewer A-operand registers, so more of their workgroups fit a CU
    // than under the batch's largest K (64-slab batch: main GEMM 2.39 ->
    // 2.15 ms).  The 16-bit main GEMM, which runs at the write bandwidth
    // either way, measured slower split (BF16x3 1.20 -> 1.26 ms, FP16 1.12 ->
    // 1.18 ms) and stays one launch per level at the largest K, as does the
    // single-launch kernel over the z-ordered entries.  SCW_MAIN_KSPLIT=0 / 1
    // forces either way.
    const bool ksplit = std::getenv("SCW_MAIN_KSPLIT") ? std::atoi(std::getenv("SCW_MAIN_KSPLIT")) != 0 : !is16;
    auto kclass = [&](int kp) {
        if (!ksplit) return ex->kmax;  // one launch per level, the largest K
        if (!is16) return kp <= 16 ? 16 : kp <= 32 ? 32 : kp <= 64 ? 64 : 128;
        return kp <= 32 ? 32 : kp <= 64 ? 64 : kp <= 96 ? 96 : 128;
    };
    std::vector<ZMeta> mks;
    mks.reserve(mzm.size());
    ex->mk_launch.clear();
    for (size_t L = 0; L + 1 < ex->mlevel_off.size(); ++L)
        for (const int K : {128, 96, 64, 32, 16}) {
            const int64_t first = (int64_t)mks.size();
            for (int64_t e = ex->mlevel_off[L]; e < ex->mlevel_off[L + 1]; ++e)
                if (kclass(mzm[e].kp) == K) mks.push_back(mzm[e]);
            if ((int64_t)mks.size() > first) ex->mk_launch.insert(ex->mk_launch.end(), {first, (int64_t)mks.size() - first, K});
        }
    // Slot-sum workgroups: 256 elements of one z's C_z each, per z only as
    // many as its region needs (a batch mixes regions of very different size).
    std::vector<int2> cblk;
    if (is16)
        for (int z = 0; z < nz; ++z) {
            if (zdirect[z]) continue;  // stored in place by the mode contraction
            const int n = 3 * zm[z].my * rd[zm[z].region].ka;
            for (int e0 = 0; e0 < n; e0 += 256) cblk.push_back(make_int2(z, e0));
        }
    ex->ncblk = (int64_t)cblk.size();

    // Work items.
    ex->items.clear();
    ex->mitems.clear();
    std::vector<uint32_t> zf;  // phase fragments of the modal runs
    const bool f32 = ex->dtype == SCW_DTYPE_FP32;
    for (size_t r0 = 0; r0 < runs.size();) {
        size_t r1 = r0 + 1;  // runs of one group: [r0, r1)
        while (r1 < runs.size() && runs[r1].group == runs[r0].group) ++r1;
        const RegionDev& d = rd[runs[r0].v];
        if (d.cells == 0 || d.width == 0 || fold[runs[r0].v]) {
            r0 = r1;
            continue;
        }
        // terms once per run: 24 bytes (3 complex FP32), or 13 for the 16-bit
        // mode contraction's block floating-point amplitudes (3 columns of 2
        // 16-bit mantissas and one exponent byte); plus the Z rows
        const double term_bytes = runs[r0].modal && is16 ? 13.0 : 24.0;
        for (size_t r = r0; r < r1; ++r)
            b_coef += term_bytes * (double)d.cells * d.width + 8.0 * d.mc * (runs[r].z1 - runs[r].z0);
        if (runs[r0].modal) {
            const RegionPlan& R = vreg[runs[r0].v].first->regions[vreg[runs[r0].v].second];
            const int T = (d.width + 15) / 16;
            std::vector<ModalItem> proto(r1 - r0);
            for (size_t r = r0; r < r1; ++r) {
                const Run& rn = runs[r];
                const int nzr = rn.z1 - rn.z0;
                const int MT = nzr <= 8 ? 1 : nzr <= 16 ? 2 : nzr <= 32 ? 4 : 8;
                ModalItem& it = proto[r - r0];
                it.region = rn.v;
                it.z0 = rn.z0;
                it.nz = nzr;
                it.T = T;
                it.cs_direct = zdirect[rn.z0];
                it.pad = 0;
                it.zf_off = (int64_t)zf.size();
                // Z'[i][k] with rows i = (z, re|im), k = (mode c, re|im):
                //   (re, re) = Re Z, (re, im) = -Im Z, (im, re) = Im Z, (im, im) = Re Z,
                // Z = exp(i kz_c (z - zref)).  Per (mt, t, lane) 8 words: FP32
                // values for MFMA q = 0..7 (c = 16t + g + 4(q/2), part q%2), or
                // 8 high then 8 low 16-bit halves (c = 16t + 4g + e/2, part e%2),
                // g = lane / 16, row i = lane % 16.
                std::vector<cd> zc((size_t)nzr * d.width);
                for (int zi = 0; zi < nzr; ++zi)
                    for (int c = 0; c < d.width; ++c) {
                        const cd val = std::exp(cd(0, 1) * R.kz[c] * (ze[rn.z0 + zi].z - R.zref));
                        if (!region_ok(val))
                            return fail(SCW_ERR_UNSUPPORTED, "region %d: exp(i kz (z - zref)) outside FP32 range at z=%g",
                                        vreg[rn.v].second, ze[rn.z0 + zi].z);
                        zc[(size_t)zi * d.width + c] =
                            zshift[rn.v].empty() ? val : val * std::ldexp(1.0, -zshift[rn.v][c]);
                    }
                zf.resize(zf.size() + (size_t)MT * T * 64 * 8);
                uint32_t* out = zf.data() + it.zf_off;
                for (int mt = 0; mt < MT; ++mt)
                    for (int t = 0; t < T; ++t)
                        for (int l = 0; l < 64; ++l) {
                            const int i = l & 15, g = l >> 4, zi = 8 * mt + (i >> 1), o = i & 1;
                            double v8[8];
                            for (int w = 0; w < 8; ++w) {
                                const int c = f32 ? 16 * t + g + 4 * (w >> 1) : 16 * t + 4 * g + (w >> 1);
                                const int part = w & 1;
                                double val = 0.0;
                                if (zi < nzr && c < d.width) {
                                    const cd zv2 = zc[(size_t)zi * d.width + c];
                                    val = o == 0 ? (part == 0 ? zv2.real() : -zv2.imag())
                                                 : (part == 0 ? zv2.imag() : zv2.real());
                                }
                                v8[w] = val;
                            }
                            uint32_t* wd = out + (((int64_t)mt * T + t) * 64 + l) * 8;
                            if (f32) {
                                for (int w = 0; w < 8; ++w) {
                                    const float fv = (float)v8[w];
                                    std::memcpy(&wd[w], &fv, 4);
                                }
                            } else {
                                uint16_t hi[8], lo[8];
                                for (int w = 0; w < 8; ++w) {
                                    if (ex->dtype == SCW_DTYPE_FP16) {
                                        hi[w] = f16_rne((float)v8[w]);
                                        lo[w] = 0;
                                    } else {
                                        hi[w] = bf16_rne((float)v8[w]);
                                        lo[w] = bf16_rne((float)(v8[w] - (double)bf16_f32(hi[w])));
                                    }
                                }
                                for (int w = 0; w < 4; ++w) {
                                    wd[w] = (uint32_t)hi[2 * w] | ((uint32_t)hi[2 * w + 1] << 16);
                                    wd[4 + w] = (uint32_t)lo[2 * w] | ((uint32_t)lo[2 * w + 1] << 16);
                                }
                            }
                        }
            }
            // Slot-major: the waves of a workgroup share a mode slice, so they
            // stream the same phase fragments at the same time.
            const int per = mper[runs[r0].v];
            for (int s2 = 0; s2 < d.nslots; ++s2)
                for (int m0 = 0; m0 < 3 * d.cells; m0 += 64) {
                    const int c0 = s2 * per, c1 = std::min(16 * T, (s2 + 1) * per);  // c0 < width
                    for (ModalItem it : proto) {
                        it.m0 = m0;
                        it.c0 = c0;
                        it.c1 = c1;
                        it.slot = s2;
                        ex->mitems.push_back(it);
                    }
                }
        } else {
            const int per = (d.width + d.nslots - 1) / d.nslots;
            for (size_t r = r0; r < r1; ++r)
                for (int c0 = 0; c0 < d.cells; c0 += 64)
                    for (int s2 = 0; s2 < d.nslots; ++s2) {
                        CoefItem it;
                        it.region = runs[r].v;
                        it.cell0 = c0;
                        it.j0 = std::min(d.width, s2 * per);
                        it.j1 = std::min(d.width, (s2 + 1) * per);
                        it.z0 = runs[r].z0;
                        it.nz = runs[r].z1 - runs[r].z0;
                        it.slot = s2;
                        it.pad = 0;
                        ex->items.push_back(it);
                    }
        }
        r0 = r1;
    }
    // Modal items grouped by row-tile class, one launch per class.
    {
        auto cls = [](const ModalItem& it) { return it.nz <= 8 ? 0 : it.nz <= 16 ? 1 : it.nz <= 32 ? 2 : 3; };
        std::stable_sort(ex->mitems.begin(), ex->mitems.end(),
                         [&](const ModalItem& a, const ModalItem& b) { return cls(a) < cls(b); });
        for (int k = 0; k <= 4; ++k) ex->mclass_off[k] = 0;
        for (const ModalItem& it : ex->mitems) ++ex->mclass_off[cls(it) + 1];
        for (int k = 1; k <= 4; ++k) ex->mclass_off[k] += ex->mclass_off[k - 1];
        // The 64-z class of the 16-bit modes runs one workgroup per IPW column
        // chunks of one run and mode slot (the waves share the phase
        // fragments in LDS): order it by (run, slot), keeping the column
        // order, and pad every group to a multiple of IPW with items past the
        // table (their waves read clamped columns and store nothing).  IPW = 8
        // (one 512-thread workgroup per CU, its two halves in ping-pong)
        // halves the fragment loads per MFMA of IPW = 4 (two 256-thread
        // workgroups): 64-slab batch, FP16 0.715 -> 0.656 ms; BF16x3 went
        // 0.953 -> 0.971 ms and keeps IPW = 4 (SCW_MODAL_WG_ITEMS = 4 or 8
        // overrides).
        const int ipw = ex->modal_wg_items =
            tunable("SCW_MODAL_WG_ITEMS", ex->dtype == SCW_DTYPE_BF16X3 ? 4 : 8) == 8 ? 8 : 4;
        if (is16 && ex->mclass_off[4] > ex->mclass_off[3]) {
            std::vector<ModalItem> out(ex->mitems.begin(), ex->mitems.begin() + ex->mclass_off[3]);
            std::vector<ModalItem> cl(ex->mitems.begin() + ex->mclass_off[3], ex->mitems.end());
            std::stable_sort(cl.begin(), cl.end(), [](const ModalItem& a, const ModalItem& b) {
                return a.zf_off != b.zf_off ? a.zf_off < b.zf_off : a.slot < b.slot;
            });
            std::vector<ModalItem> wg;  // whole workgroups of this class
            for (size_t i = 0; i < cl.size();) {
                size_t k = i;
                while (k < cl.size() && cl[k].zf_off == cl[i].zf_off && cl[k].slot == cl[i].slot) wg.push_back(cl[k++]);
                ModalItem pad = cl[k - 1];
                for (size_t n = k - i; n % ipw; ++n) {
                    pad.m0 += 64;
                    wg.push_back(pad);
                }
                i = k;
            }
            // Longest workgroups first (workgroups are dispatched in order):
            // a batch mixes mode counts (a 41^2-order slab's workgroups take
            // about 13x the steps of an 11^2 one), and in batch order the last
            // long ones set the end of the launch while most CUs idle; taken
            // first, the short ones fill the end.  64-slab sweep: mode
            // contraction 1.14 -> 1.04 ms (BF16x3), 0.99 -> 0.80 ms (FP16).
            const int nxcd = (int)std::min<int64_t>(64, tunable("SCW_MODAL_XCD", 8));
            if (nxcd > 1) {
                // each (run, slot) group of workgroups whole on one XCD (see
                // xcd_group_order); padding slots become empty workgroups
                // (nz = 0: the kernel returns at once)
                std::vector<std::pair<int64_t, int64_t>> gs;  // (workgroups, steps) per group
                for (size_t q = 0; q < wg.size() / ipw;) {
                    size_t e = q;
                    while (e < wg.size() / ipw && wg[ipw * e].zf_off == wg[ipw * q].zf_off &&
                           wg[ipw * e].slot == wg[ipw * q].slot)
                        ++e;
                    gs.push_back({(int64_t)(e - q), wg[ipw * q].c1 - wg[ipw * q].c0});
                    q = e;
                }
                ModalItem none = wg[0];
                none.nz = 0;
                none.c0 = none.c1 = 0;
                for (const int64_t q : xcd_group_order(gs, nxcd)) {
                    if (q >= 0) out.insert(out.end(), wg.begin() + ipw * q, wg.begin() + ipw * q + ipw);
                    else out.insert(out.end(), ipw, none);
                }
            } else {
                std::vector<size_t> gi(wg.size() / ipw);
                for (size_t q = 0; q < gi.size(); ++q) gi[q] = q;
                std::stable_sort(gi.begin(), gi.end(), [&](size_t a, size_t b) {
                    return wg[ipw * a].c1 - wg[ipw * a].c0 > wg[ipw * b].c1 - wg[ipw * b].c0;
                });
                for (size_t q : gi) out.insert(out.end(), wg.begin() + ipw * q, wg.begin() + ipw * q + ipw);
            }
            ex->mitems.swap(out);
            ex->mclass_off[4] = (int64_t)ex->mitems.size();
        } else if (ex->mclass_off[4] > ex->mclass_off[3]) {
            // FP32's 64-z class (one wave per item): longest items first too
            std::stable_sort(ex->mitems.begin() + ex->mclass_off[3], ex->mitems.end(),
                             [](const ModalItem& a, const ModalItem& b) { return a.c1 - a.c0 > b.c1 - b.c0; });
        }
    }
    ex->stage_flops[1] = f_coef;
    ex->stage_flops[2] = f_u;
    ex->stage_flops[3] = f_main;
    ex->stage_bytes[1] = b_coef;
    ex->stage_bytes[2] = b_u;
    ex->stage_bytes[3] = b_main;
    ex->stage_flops[0] = ex->flops = f_coef + f_u + f_main;
    ex->stage_bytes[0] = ex->bytes = b_coef + b_u + b_main;
    ex->launches = stage_launches(ex);

    const int gx = (int)((6LL * nx + 255) / 256);
    // Each wave keeps its A-operand fragment (from U_z) in registers for all
    // of its y tiles, and the 16-bit workgroups compute their U_z columns
    // themselves, so split y across workgroups only as far as needed to give
    // every CU a few workgroups (about 768 in all).
    const int64_t mtarget = ex->tune_main_blocks > 0 ? ex->tune_main_blocks : tunable("SCW_MAIN_BLOCKS", 768);
    const int ysplit =
        (int)std::min<int64_t>(ytiles, std::max<int64_t>(1, (mtarget + (int64_t)gx * nz - 1) / ((int64_t)gx * nz)));
    ex->ytiles_per_split = (ytiles + ysplit - 1) / ysplit;
    ex->ysplit = (ytiles + ex->ytiles_per_split - 1) / ex->ytiles_per_split;

    int st;
    if ((st = ex->d_regions.upload(rd.data(), rd.size() * sizeof(RegionDev))) != SCW_OK) return st;
    if ((st = ex->d_zmeta.upload(zm.data(), zm.size() * sizeof(ZMeta))) != SCW_OK) return st;
    if ((st = ex->d_mzmeta.upload(mzm.data(), mzm.size() * sizeof(ZMeta))) != SCW_OK) return st;
    if ((st = ex->d_mksort.upload(mks.data(), mks.size() * sizeof(ZMeta))) != SCW_OK) return st;
    if ((st = ex->d_items.upload(ex->items.data(), ex->items.size() * sizeof(CoefItem))) != SCW_OK) return st;
    if ((st = ex->d_mitems.upload(ex->mitems.data(), ex->mitems.size() * sizeof(ModalItem))) != SCW_OK) return st;
    if ((st = ex->d_Z.upload(Z.data(), Z.size() * sizeof(float))) != SCW_OK) return st;
    if ((st = ex->d_Zf.upload(zf.data(), zf.size() * sizeof(uint32_t))) != SCW_OK) return st;
    if ((st = ex->d_cblk.upload(cblk.data(), cblk.size() * sizeof(int2))) != SCW_OK) return st;
    if ((st = ex->d_X.upload(X.data(), X.size() * sizeof(float))) != SCW_OK) return st;
    if ((st = ex->d_C.alloc((size_t)ex->nslots_max * std::max<int64_t>(c_off, 1) * 8)) != SCW_OK) return st;
    if ((st = ex->d_U.alloc((size_t)std::max<int64_t>(u_off, 1) * 8)) != SCW_OK) return st;
    // runs stored in place never write the zero columns past mx of their C_z;
    // the fill (on the null stream) is complete before any run, whatever
    // stream the caller runs on
    if (std::find(zdirect.begin(), zdirect.end(), 1) != zdirect.end()) {
        SCW_HIP(hipMemset(ex->d_U.p, 0, (size_t)std::max<int64_t>(u_off, 1) * 8));
        SCW_HIP(hipStreamSynchronize(nullptr));
    }
    if (!is16) {
        if (yf32.empty()) yf32.push_back(0.0f);  // every region on the grid is empty: keep a valid pointer
        if ((st = ex->d_Yfrag.upload(yf32.data(), yf32.size() * 4)) != SCW_OK) return st;
        ex->ylo_delta = 0;
    } else {
        std::vector<uint16_t> both(yhi);
        both.insert(both.end(), ylo.begin(), ylo.end());
        if (both.empty()) both.push_back(0);
        ex->ylo_delta = (int64_t)yhi.size();
        if ((st = ex->d_Yfrag.upload(both.data(), both.size() * 2)) != SCW_OK) return st;
    }
    ex->hcblk.swap(cblk);
    if (ex->persistent) return build_units(ex);
    return SCW_OK;
}

// The z entries of an exec: plan, region and z of every output slice.
static std::vector<ZEntry> entries(const scw_exec* ex) {
    std::vector<ZEntry> ze(ex->hz.size());
    for (size_t z = 0; z < ze.size(); ++z) {
        const scw_plan* P = ex->batch.empty() ? ex->plan : ex->batch[ex->hzp[z]];
        ze[z] = {P, ex->zregion[z], ex->hz[z]};
    }
    return ze;
}

// Drops the GEMM-form tables and work lists (before a fall back to the direct form).
static void release_gemm(scw_exec* ex) {
    for (DevBuf* b : {&ex->d_items, &ex->d_mitems, &ex->d_regions, &ex->d_zmeta, &ex->d_Z, &ex->d_Zf, &ex->d_X, &ex->d_C,
                      &ex->d_U, &ex->d_Yfrag, &ex->d_bfpx, &ex->d_cblk, &ex->d_units, &ex->d_need, &ex->d_ctl,
                      &ex->d_mzmeta, &ex->d_mksort})
        b->release();
    ex->mlevel_off.clear();
    ex->mk_launch.clear();
    ex->persistent = false;
    ex->nunits = ex->pgrid = 0;
    ex->hcblk.clear();
    ex->items.clear();
    ex->mitems.clear();
    for (int k = 0; k <= 4; ++k) ex->mclass_off[k] = 0;
    ex->ncblk = 0;
    ex->my_max = ex->ka_max = ex->kmax = 0;
    ex->nslots_max = 1;
    for (int k = 0; k < 4; ++k) ex->stage_flops[k] = ex->stage_bytes[k] = 0.0;
}

static void exec_free(scw_exec* ex) {
    if (ex->graph_exec) (void)hipGraphExecDestroy(ex->graph_exec);
    if (ex->graph) (void)hipGraphDestroy(ex->graph);
    if (ex->graph_stream) (void)hipStreamDestroy(ex->graph_stream);
    for (DevBuf* b : {&ex->d_dx, &ex->d_dy, &ex->d_z, &ex->d_zslab, &ex->d_items, &ex->d_mitems, &ex->d_regions, &ex->d_zmeta,
                      &ex->d_Z, &ex->d_Zf, &ex->d_X, &ex->d_C, &ex->d_U, &ex->d_Yfrag, &ex->d_bfpx, &ex->d_cblk,
                      &ex->d_units, &ex->d_need, &ex->d_ctl, &ex->d_mzmeta, &ex->d_mksort, &ex->d_out})
        b->release();
    delete ex;
}

// (Re)builds ex for a single plan on the grid xs x ys x zs: validation, form
// choice, tables.  An exec being rebuilt keeps its device buffers where they
// are large enough (DevBuf capacity), so a rebuild of a similar grid needs no
// allocation; a captured graph is dropped.  On failure the exec is left
// invalid (runs refused) and the caller destroys or rebuilds it.
int exec_init(scw_exec* ex, scw_plan* plan, const double* xs, size_t nx, const double* ys, size_t ny,
              const double* zs, size_t nz, double region_tol, int form, int dtype) {
    if (!plan) return fail(SCW_ERR_ARG, "plan is NULL");
    if (!xs || !ys || !zs) return fail(SCW_ERR_ARG, "coordinate array is NULL");
    if (nx == 0 || ny == 0 || nz == 0) return fail(SCW_ERR_ARG, "empty grid (%zu x %zu x %zu)", nx, ny, nz);
    if (nx > (1u << 24) || ny > (1u << 24) || nz > 65535)
        return fail(SCW_ERR_ARG, "grid too large (nx, ny <= 2^24, nz <= 65535)");
    if ((double)nx * ny * nz * 24.0 > 1.0e12) return fail(SCW_ERR_ARG, "grid output exceeds 1 TB");
    if (!std::isfinite(region_tol)) return fail(SCW_ERR_ARG, "region_tol is not finite");
    if (form < SCW_FORM_AUTO || form > SCW_FORM_GEMM) return fail(SCW_ERR_ARG, "unknown form %d", form);
    if (dtype < SCW_DTYPE_FP32 || dtype > SCW_DTYPE_FP16) return fail(SCW_ERR_ARG, "unknown dtype %d", dtype);
    for (size_t i = 0; i < nx; ++i)
        if (!std::isfinite(xs[i])) return fail(SCW_ERR_ARG, "xs[%zu] is not finite", i);
    for (size_t i = 0; i < ny; ++i)
        if (!std::isfinite(ys[i])) return fail(SCW_ERR_ARG, "ys[%zu] is not finite", i);
    for (size_t i = 0; i < nz; ++i)
        if (!std::isfinite(zs[i])) return fail(SCW_ERR_ARG, "zs[%zu] is not finite", i);
    if (ex->plan && ex->plan != plan) return fail(SCW_ERR_ARG, "internal: exec rebuilt for another plan");

    DeviceGuard g(plan->device);
    if (g.st != SCW_OK) return g.st;
    ex->invalid = true;  // until the tables below are complete
    if (ex->graph_exec) {
        (void)hipGraphExecDestroy(ex->graph_exec);
        ex->graph_exec = nullptr;
    }
    if (ex->graph) {
        (void)hipGraphDestroy(ex->graph);
        ex->graph = nullptr;
    }
    ex->plan = plan;
    ex->batch.clear();
    if (dtype == SCW_DTYPE_FP32) ex->persistent = false;  // the single-launch form takes 16-bit dtypes
    ex->nx = (int64_t)nx;
    ex->ny = (int64_t)ny;
    ex->nz = (int64_t)nz;
    ex->dtype = dtype;
    int st;
    try {
        ex->zregion.resize(nz);
        bool all_sep = true;
        int first_bad = -1;
        double units = 0.0;
        for (size_t z = 0; z < nz; ++z) {
            const int r = region_of(zs[z], plan->zl, region_tol);
            ex->zregion[z] = r;
            const RegionPlan& R = plan->regions[r];
            units += (double)nx * ny * (double)(R.wend - R.wbegin);
            if (!R.separable && all_sep) {
                all_sep = false;
                first_bad = r;
            }
        }
        ex->units = units;
        const bool auto_form = form == SCW_FORM_AUTO;
        if (form == SCW_FORM_AUTO) form = all_sep ? SCW_FORM_GEMM : SCW_FORM_DIRECT;
        if (form == SCW_FORM_GEMM && !all_sep)
            return fail(SCW_ERR_UNSUPPORTED, "GEMM form: region %d is not separable (%s)", first_bad,
                        plan->regions[first_bad].why_not.c_str());
        ex->form = form;
        if (form == SCW_FORM_DIRECT) ex->dtype = SCW_DTYPE_FP32;
        ex->hx.assign(xs, xs + nx);
        ex->hy.assign(ys, ys + ny);
        ex->hz.assign(zs, zs + nz);
        ex->hzp.assign(nz, 0);
        ex->nslots_max = 1;
        if (form == SCW_FORM_DIRECT) {
            release_gemm(ex);
            st = prep_direct(ex, xs, ys, zs);
        } else {
            for (DevBuf* b : {&ex->d_dx, &ex->d_dy, &ex->d_z, &ex->d_zslab}) b->release();
            st = prep_gemm(ex, xs, ys, entries(ex));
            if (st == SCW_ERR_UNSUPPORTED && auto_form) {
                // Separable, but this grid takes a phase table past FP32's
                // range (a point far from the region's phase origin in a
                // strongly evanescent mode) or a U slice past 32-bit offsets:
                // AUTO falls back to the direct form, which carries such
                // exponents in FP64.
                release_gemm(ex);
                ex->form = SCW_FORM_DIRECT;
                ex->dtype = SCW_DTYPE_FP32;
                st = prep_direct(ex, xs, ys, zs);
            }
        }
    } catch (const std::bad_alloc&) {
        st = fail(SCW_ERR_NOMEM, "host allocation failed while building grid tables");
    }
    if (st == SCW_OK) ex->invalid = false;
    return st;
}

}  // namespace scw

using namespace scw;

extern "C" {

int scw_exec_create(scw_plan* plan, const double* xs, size_t nx, const double* ys, size_t ny, const double* zs,
                    size_t nz, double region_tol, int form, int dtype, scw_exec** out) {
    if (!out) return fail(SCW_ERR_ARG, "exec output is NULL");
    *out = nullptr;
    TraceRange tr("scw_exec_create");
    scw_exec* ex = new (std::nothrow) scw_exec();
    if (!ex) return fail(SCW_ERR_NOMEM, "host allocation failed");
    const int st = exec_init(ex, plan, xs, nx, ys, ny, zs, nz, region_tol, form, dtype);
    if (st != SCW_OK) {
        exec_free(ex);
        return st;
    }
    *out = ex;
    return SCW_OK;
}

int scw_batch_exec_create(int n, scw_plan* const* plans, const double* xs, size_t nx, const double* ys, size_t ny,
                          const double* const* zs, const size_t* nz, double region_tol, int dtype, scw_exec** out) {
    if (!out) return fail(SCW_ERR_ARG, "exec output is NULL");
    *out = nullptr;
    if (n <= 0 || !plans || !zs || !nz) return fail(SCW_ERR_ARG, "empty or NULL batch");
    if (!xs || !ys || nx == 0 || ny == 0) return fail(SCW_ERR_ARG, "empty x / y axes");
    if (nx > (1u << 24) || ny > (1u << 24)) return fail(SCW_ERR_ARG, "grid too large (nx, ny <= 2^24)");
    if (dtype < SCW_DTYPE_FP32 || dtype > SCW_DTYPE_FP16) return fail(SCW_ERR_ARG, "unknown dtype %d", dtype);
    if (!std::isfinite(region_tol)) return fail(SCW_ERR_ARG, "region_tol is not finite");
    size_t total = 0;
    for (int i = 0; i < n; ++i) {
        if (!plans[i] || !zs[i] || nz[i] == 0) return fail(SCW_ERR_ARG, "instance %d: NULL plan or empty z axis", i);
        if (plans[i]->device != plans[0]->device) return fail(SCW_ERR_ARG, "instance %d: plans on different devices", i);
        for (size_t k = 0; k < nz[i]; ++k)
            if (!std::isfinite(zs[i][k])) return fail(SCW_ERR_ARG, "instance %d: z[%zu] is not finite", i, k);
        total += nz[i];
    }
    if (total > 65535) return fail(SCW_ERR_ARG, "batch has more than 65535 z-slices");
    if ((double)nx * ny * total * 24.0 > 1.0e12) return fail(SCW_ERR_ARG, "batch output exceeds 1 TB");
    for (size_t i = 0; i < nx; ++i)
        if (!std::isfinite(xs[i])) return fail(SCW_ERR_ARG, "xs[%zu] is not finite", i);
    for (size_t i = 0; i < ny; ++i)
        if (!std::isfinite(ys[i])) return fail(SCW_ERR_ARG, "ys[%zu] is not finite", i);

    TraceRange tr("scw_batch_exec_create");
    scw_exec* ex = new (std::nothrow) scw_exec();
    if (!ex) return fail(SCW_ERR_NOMEM, "host allocation failed");
    int st = SCW_OK;
    try {
        ex->plan = plans[0];
        ex->batch.assign(plans, plans + n);
        ex->nx = (int64_t)nx;
        ex->ny = (int64_t)ny;
        ex->nz = (int64_t)total;
        ex->dtype = dtype;
        ex->form = SCW_FORM_GEMM;
        ex->hx.assign(xs, xs + nx);
        ex->hy.assign(ys, ys + ny);
        double units = 0.0;
        for (int i = 0; i < n; ++i)
            for (size_t k = 0; k < nz[i]; ++k) {
                const int r = region_of(zs[i][k], plans[i]->zl, region_tol);
                ex->hz.push_back(zs[i][k]);
                ex->hzp.push_back(i);
                ex->zregion.push_back(r);
                const RegionPlan& R = plans[i]->regions[r];
                units += (double)nx * ny * (double)(R.wend - R.wbegin);
            }
        ex->units = units;
    } catch (const std::bad_alloc&) {
        delete ex;
        return fail(SCW_ERR_NOMEM, "host allocation failed");
    }
    DeviceGuard g(plans[0]->device);
    if (g.st != SCW_OK) {
        delete ex;
        return g.st;
    }
    try {
        st = prep_gemm(ex, xs, ys, entries(ex));
    } catch (const std::bad_alloc&) {
        st = fail(SCW_ERR_NOMEM, "host allocation failed while building grid tables");
    }
    if (st != SCW_OK) {
        exec_free(ex);
        return st;
    }
    *out = ex;
    return SCW_OK;
}

// Rebuilds the GEMM-form work lists and tables after a launch-shape or tile
// change (the grid and plans are those of the exec).
static int rebuild_gemm(scw_exec* ex) {
    TraceRange tr("scw_exec rebuild");
    DeviceGuard g(ex->plan->device);
    if (g.st != SCW_OK) return g.st;
    // a run of this exec may still be in flight on any stream (also one that
    // does not synchronise with the copies below): let it finish before its
    // tables are replaced
    SCW_HIP(hipDeviceSynchronize());
    if (ex->graph_exec) {
        (void)hipGraphExecDestroy(ex->graph_exec);
        ex->graph_exec = nullptr;
    }
    if (ex->graph) {
        (void)hipGraphDestroy(ex->graph);
        ex->graph = nullptr;
    }
    ex->nslots_max = 1;
    // prep_gemm replaces the device tables one by one; a failure part way
    // leaves some of them released next to rebuilt work lists, so the exec
    // stays refused until a later rebuild succeeds.
    ex->invalid = true;
    int st;
    try {
        st = prep_gemm(ex, ex->hx.data(), ex->hy.data(), entries(ex));
    } catch (const std::bad_alloc&) {
        st = fail(SCW_ERR_NOMEM, "host allocation failed while building grid tables");
    }
    if (st == SCW_OK) ex->invalid = false;
    return st;
}

int scw_exec_set_launch(scw_exec* ex, int64_t coef_items, int64_t main_blocks) {
    if (!ex) return fail(SCW_ERR_ARG, "exec is NULL");
    if (coef_items < 0 || main_blocks < 0) return fail(SCW_ERR_ARG, "negative launch parameter");
    if (ex->form != SCW_FORM_GEMM) return SCW_OK;  // the direct form has no tunable launch shape
    ex->tune_coef_items = coef_items;
    ex->tune_main_blocks = main_blocks;
    return rebuild_gemm(ex);
}

int scw_exec_set_tiles(scw_exec* ex, int modal_run, int coef_run) {
    if (!ex) return fail(SCW_ERR_ARG, "exec is NULL");
    if (modal_run < 0 || modal_run > 64 || coef_run < 0 || coef_run > 8)
        return fail(SCW_ERR_ARG, "tile: modal_run in [0, 64], coef_run in [0, 8] (0 = default)");
    if (ex->form != SCW_FORM_GEMM) return SCW_OK;
    ex->tune_modal_run = modal_run;
    ex->tune_coef_run = coef_run;
    return rebuild_gemm(ex);
}

static int refuse_invalid(const scw_exec* ex);

static void drop_graph(scw_exec* ex) {
    if (ex->graph_exec) {
        (void)hipGraphExecDestroy(ex->graph_exec);
        ex->graph_exec = nullptr;
    }
    if (ex->graph) {
        (void)hipGraphDestroy(ex->graph);
        ex->graph = nullptr;
    }
}

int scw_exec_set_persistent(scw_exec* ex, int on, int64_t workgroups, double wait_ms) {
    if (!ex) return fail(SCW_ERR_ARG, "exec is NULL");
    if (workgroups < 0) return fail(SCW_ERR_ARG, "negative workgroup count");
    if (!std::isfinite(wait_ms) || wait_ms < 0 || wait_ms > 3.6e6)
        return fail(SCW_ERR_ARG, "wait limit must be in [0, 3.6e6] ms (0 = default)");
    if (on && (ex->form != SCW_FORM_GEMM || ex->dtype == SCW_DTYPE_FP32))
        return fail(SCW_ERR_UNSUPPORTED, "single-launch form: GEMM form with a 16-bit dtype only");
    if (on && ex->mlevel_off.size() > 2)
        return fail(SCW_ERR_UNSUPPORTED, "single-launch form: a region has more than 64 ky (chunked main GEMM)");
    if (ex->invalid) return refuse_invalid(ex);
    DeviceGuard g(ex->plan->device);
    if (g.st != SCW_OK) return g.st;
    SCW_HIP(hipDeviceSynchronize());  // a run in flight still reads the unit table and counters
    drop_graph(ex);
    for (DevBuf* b : {&ex->d_units, &ex->d_need, &ex->d_ctl}) b->release();
    ex->persistent = false;
    ex->nunits = ex->pgrid = 0;
    ex->launches = ex->form == SCW_FORM_GEMM ? stage_launches(ex) : ex->launches;
    if (!on) return SCW_OK;
    ex->tune_grid = workgroups;
    ex->wait_ms = wait_ms > 0 ? wait_ms : 10000.0;
    ex->persistent = true;
    int st;
    try {
        st = build_units(ex);
    } catch (const std::bad_alloc&) {
        st = fail(SCW_ERR_NOMEM, "host allocation failed while building work units");
    }
    if (st != SCW_OK) {
        // the stage-by-stage tables are untouched: the exec stays usable in that form
        for (DevBuf* b : {&ex->d_units, &ex->d_need, &ex->d_ctl}) b->release();
        ex->persistent = false;
        ex->nunits = ex->pgrid = 0;
        ex->launches = stage_launches(ex);
    }
    return st;
}

int scw_exec_persistent_info(const scw_exec* ex, int* on, int64_t* units, int64_t* workgroups) {
    if (!ex) return fail(SCW_ERR_ARG, "exec is NULL");
    if (on) *on = ex->persistent ? 1 : 0;
    if (units) *units = ex->nunits;
    if (workgroups) *workgroups = ex->pgrid;
    return SCW_OK;
}

int scw_exec_check(scw_exec* ex, void* stream) {
    if (!ex) return fail(SCW_ERR_ARG, "exec is NULL");
    if (!ex->persistent || !ex->d_ctl.p) return SCW_OK;
    DeviceGuard g(ex->plan->device);
    if (g.st != SCW_OK) return g.st;
    SCW_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    int* err = ex->d_ctl.as<int>() + 2;  // the error word (kCtlErr)
    int v = 0;
    SCW_HIP(hipMemcpy(&v, err, sizeof(int), hipMemcpyDeviceToHost));
    if (v == 0) return SCW_OK;
    const int zero = 0;
    SCW_HIP(hipMemcpy(err, &zero, sizeof(int), hipMemcpyHostToDevice));
    return fail(SCW_ERR_TIMEOUT, "single-launch run: a dependency wait exceeded %.0f ms; its output is invalid",
                ex->wait_ms);
}

int scw_test_xcd_order(int ngroups, const int64_t* group_wgs, const int64_t* group_steps, int nxcd, int64_t* order,
                       int64_t capacity, int64_t* count) {
    if (ngroups < 0 || nxcd < 1 || nxcd > 64 || !count || (ngroups > 0 && (!group_wgs || !group_steps)))
        return fail(SCW_ERR_ARG, "xcd order: bad arguments");
    std::vector<std::pair<int64_t, int64_t>> gs((size_t)ngroups);
    for (int g = 0; g < ngroups; ++g) {
        if (group_wgs[g] < 0 || group_steps[g] < 0) return fail(SCW_ERR_ARG, "xcd order: negative group size");
        gs[(size_t)g] = {group_wgs[g], group_steps[g]};
    }
    std::vector<int64_t> o;
    try {
        o = xcd_group_order(gs, nxcd);
    } catch (const std::bad_alloc&) {
        return fail(SCW_ERR_NOMEM, "host allocation failed");
    }
    *count = (int64_t)o.size();
    if ((int64_t)o.size() > capacity || (!order && !o.empty()))
        return fail(SCW_ERR_ARG, "xcd order: %zu slots exceed the capacity %lld", o.size(), (long long)capacity);
    std::copy(o.begin(), o.end(), order);
    return SCW_OK;
}

int scw_test_exec_stall(scw_exec* ex, int on) {
    if (!ex) return fail(SCW_ERR_ARG, "exec is NULL");
    ex->stall = on != 0;
    if (!ex->persistent) return SCW_OK;
    DeviceGuard g(ex->plan->device);
    if (g.st != SCW_OK) return g.st;
    drop_graph(ex);
    int st;
    try {
        st = build_units(ex);
    } catch (const std::bad_alloc&) {
        st = fail(SCW_ERR_NOMEM, "host allocation failed while building work units");
    }
    // a half-built unit table must never be launched: refuse runs until a
    // later rebuild succeeds
    if (st != SCW_OK) ex->invalid = true;
    return st;
}

static int refuse_invalid(const scw_exec* ex) {
    return fail(SCW_ERR_ARG, "exec is unusable: its last scw_exec_set_launch failed (%s)",
                ex->form == SCW_FORM_GEMM ? "rebuild the launch shape or destroy it" : "destroy it");
}

int scw_exec_run(scw_exec* ex, void* out_device, void* stream) {
    if (!ex) return fail(SCW_ERR_ARG, "exec is NULL");
    if (!out_device) return fail(SCW_ERR_ARG, "output pointer is NULL");
    if (ex->invalid) return refuse_invalid(ex);
    TraceRange tr(ex->form == SCW_FORM_DIRECT ? "scw_exec_run direct" : "scw_exec_run gemm");
    DeviceGuard g(ex->plan->device);
    if (g.st != SCW_OK) return g.st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    return ex->form == SCW_FORM_DIRECT ? launch_direct(ex, static_cast<float*>(out_device), s)
                                       : launch_gemm(ex, static_cast<float*>(out_device), s);
}

int scw_exec_run_stage(scw_exec* ex, int stage, void* out_device, void* stream) {
    if (!ex) return fail(SCW_ERR_ARG, "exec is NULL");
    if (!out_device) return fail(SCW_ERR_ARG, "output pointer is NULL");
    if (stage < 0 || stage > 3) return fail(SCW_ERR_ARG, "stage %d out of range", stage);
    if (ex->invalid) return refuse_invalid(ex);
    DeviceGuard g(ex->plan->device);
    if (g.st != SCW_OK) return g.st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    return ex->form == SCW_FORM_DIRECT ? launch_direct(ex, static_cast<float*>(out_device), s)
                                       : launch_gemm(ex, static_cast<float*>(out_device), s, stage);
}

int scw_exec_run_host(scw_exec* ex, void* out_host, void* stream) {
    if (!ex) return fail(SCW_ERR_ARG, "exec is NULL");
    if (!out_host) return fail(SCW_ERR_ARG, "output pointer is NULL");
    if (ex->invalid) return refuse_invalid(ex);
    DeviceGuard g(ex->plan->device);
    if (g.st != SCW_OK) return g.st;
    const size_t bytes = (size_t)ex->nx * ex->ny * ex->nz * 24;
    if (ex->d_out.bytes < bytes) {
        const int st = ex->d_out.alloc(bytes);
        if (st != SCW_OK) return st;
    }
    int st = scw_exec_run(ex, ex->d_out.p, stream);
    if (st != SCW_OK) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    SCW_HIP(hipMemcpyAsync(out_host, ex->d_out.p, bytes, hipMemcpyDeviceToHost, s));
    SCW_HIP(hipStreamSynchronize(s));
    return SCW_OK;
}

int scw_exec_destroy(scw_exec* ex) {
    if (!ex) return SCW_OK;
    DeviceGuard g(ex->plan->device);
    exec_free(ex);
    return SCW_OK;
}

int scw_exec_info(const scw_exec* ex, int* form, int* dtype, double* units, double* flops, double* bytes,
                  int* launches) {
    if (!ex) return fail(SCW_ERR_ARG, "exec is NULL");
    if (form) *form = ex->form;
    if (dtype) *dtype = ex->dtype;
    if (units) *units = ex->units;
    if (flops) *flops = ex->flops;
    if (bytes) *bytes = ex->bytes;
    if (launches) *launches = ex->launches;
    return SCW_OK;
}

int scw_exec_stage_info(const scw_exec* ex, int stage, double* flops, double* bytes) {
    if (!ex) return fail(SCW_ERR_ARG, "exec is NULL");
    if (stage < 0 || stage > 3) return fail(SCW_ERR_ARG, "stage %d out of range", stage);
    if (flops) *flops = ex->stage_flops[stage];
    if (bytes) *bytes = ex->stage_bytes[stage];
    return SCW_OK;
}

int scw_exec_graph_capture(scw_exec* ex, void* out_device, void* stream) {
    if (!ex) return fail(SCW_ERR_ARG, "exec is NULL");
    if (!out_device) return fail(SCW_ERR_ARG, "output pointer is NULL");
    if (ex->invalid) return refuse_invalid(ex);
    DeviceGuard g(ex->plan->device);
    if (g.st != SCW_OK) return g.st;
    if (ex->graph_exec) {
        (void)hipGraphExecDestroy(ex->graph_exec);
        ex->graph_exec = nullptr;
    }
    if (ex->graph) {
        (void)hipGraphDestroy(ex->graph);
        ex->graph = nullptr;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (!s) {  // capture needs a non-legacy stream
        if (!ex->graph_stream) SCW_HIP(hipStreamCreateWithFlags(&ex->graph_stream, hipStreamNonBlocking));
        s = ex->graph_stream;
    }
    SCW_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int st = ex->form == SCW_FORM_DIRECT ? launch_direct(ex, static_cast<float*>(out_device), s)
                                         : launch_gemm(ex, static_cast<float*>(out_device), s);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(s, &graph);
    if (st != SCW_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    ex->graph = graph;
    SCW_HIP(hipGraphInstantiate(&ex->graph_exec, graph, nullptr, nullptr, 0));
    return SCW_OK;
}

int scw_exec_graph_launch(scw_exec* ex, void* stream) {
    if (!ex || !ex->graph_exec) return fail(SCW_ERR_ARG, "no captured graph");
    if (ex->invalid) return refuse_invalid(ex);
    DeviceGuard g(ex->plan->device);
    if (g.st != SCW_OK) return g.st;
    SCW_HIP(hipGraphLaunch(ex->graph_exec, static_cast<hipStream_t>(stream)));
    return SCW_OK;
}

}  // extern "C"
