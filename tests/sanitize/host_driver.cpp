// Host paths of libscwatter_amd under AddressSanitizer and
// UndefinedBehaviorSanitizer (built by tests/sanitize/Makefile, run by
// tests/test_sanitize_host.py, no GPU): the .cw loader and writer
// (scw_cw.cpp) on valid structures, on every truncation of a valid file and
// on corrupted headers and counts; scw_cw_unpack; and the planner's host half
// (scw_test_plan_host: validation and the FP64 separable decomposition of
// scw_plan.cpp) on lattices, dense slabs, scattered sets, degenerate regions
// and invalid tables.  Exits non-zero on a wrong status; the sanitizers abort
// on a memory or UB error.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/scwatter_amd.h"

static int g_failures = 0;
#define CHECK(cond)                                                              \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
            ++g_failures;                                                        \
        }                                                                        \
    } while (0)

static std::vector<scw_planewave64> random_waves(std::mt19937_64& rng, size_t n) {
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    std::vector<scw_planewave64> w(n);
    for (auto& x : w)
        for (int c = 0; c < 3; ++c) {
            x.E0[c][0] = u(rng);
            x.E0[c][1] = u(rng);
            x.k[c][0] = 6.0 * u(rng);
            x.k[c][1] = 0.1 * u(rng);
        }
    return w;
}

static std::vector<unsigned char> read_file(const std::string& p) {
    std::vector<unsigned char> b;
    FILE* f = std::fopen(p.c_str(), "rb");
    if (!f) return b;
    int c;
    while ((c = std::fgetc(f)) != EOF) b.push_back((unsigned char)c);
    std::fclose(f);
    return b;
}

static void write_file(const std::string& p, const unsigned char* d, size_t n) {
    FILE* f = std::fopen(p.c_str(), "wb");
    if (n) std::fwrite(d, 1, n, f);
    std::fclose(f);
}

static void test_cw(const std::string& dir) {
    std::mt19937_64 rng(7);
    scw_cw* cw = nullptr;
    CHECK(scw_cw_create(&cw) == SCW_OK);
    auto Pi = random_waves(rng, 5);
    CHECK(scw_cw_set_incident(cw, Pi.data(), Pi.size()) == SCW_OK);
    const size_t nr[3] = {4, 0, 7}, nt[3] = {3, 6, 0};
    for (int l = 0; l < 3; ++l) {
        auto r = random_waves(rng, nr[l]), t = random_waves(rng, nt[l]);
        CHECK(scw_cw_add_layer(cw, 0.5 * l, r.data(), nr[l], t.data(), nt[l]) == SCW_OK);
    }
    const std::string path = dir + "/ok.cw";
    CHECK(scw_cw_save(cw, path.c_str()) == SCW_OK);
    scw_cw* back = nullptr;
    CHECK(scw_cw_load(path.c_str(), &back) == SCW_OK);
    size_t ni = 0, nl = 0, nw = 0;
    CHECK(scw_cw_counts(back, &ni, &nl, &nw) == SCW_OK);
    CHECK(ni == 5 && nl == 3 && nw == 5 + 4 + 3 + 0 + 6 + 7 + 0);
    for (size_t l = 0; l < 3; ++l) {
        double z = 0;
        size_t a = 0, b = 0;
        CHECK(scw_cw_layer(back, l, &z, &a, &b) == SCW_OK);
        CHECK(z == 0.5 * l && a == nr[l] && b == nt[l]);
    }
    double zl = 0;
    size_t a = 0, b = 0;
    CHECK(scw_cw_layer(back, 3, &zl, &a, &b) == SCW_ERR_ARG);
    std::vector<scw_planewave64> W(nw);
    std::vector<double> z(nl);
    std::vector<int32_t> wb(nl + 1), we(nl + 1);
    CHECK(scw_cw_unpack(back, W.data(), z.data(), wb.data(), we.data()) == SCW_OK);
    for (size_t r = 0; r <= nl; ++r) CHECK(wb[r] <= we[r] && (size_t)we[r] <= nw);
    // the unpacked tables through the host planner
    std::vector<int64_t> info(8 * (nl + 1));
    CHECK(scw_test_plan_host(W.data(), nw, z.data(), (int)nl, wb.data(), we.data(), info.data()) == SCW_OK);
    scw_cw_destroy(back);

    // every truncation of the file is refused
    const std::vector<unsigned char> bytes = read_file(path);
    CHECK(bytes.size() > 16);
    const std::string cut = dir + "/cut.cw";
    for (size_t n = 0; n < bytes.size(); ++n) {
        write_file(cut, bytes.data(), n);
        scw_cw* c = nullptr;
        CHECK(scw_cw_load(cut.c_str(), &c) != SCW_OK);
        CHECK(c == nullptr);
    }
    // corrupted precision field and counts
    const std::string bad = dir + "/bad.cw";
    std::vector<unsigned char> v = bytes;
    v[0] = 3;  // precision field (first 8 bytes): neither 4 nor 8
    write_file(bad, v.data(), v.size());
    scw_cw* c = nullptr;
    CHECK(scw_cw_load(bad.c_str(), &c) == SCW_ERR_FORMAT);
    for (size_t off = 8; off + 8 <= std::min<size_t>(bytes.size(), 64); off += 8) {
        v = bytes;
        std::memset(&v[off], 0xff, 8);  // a huge count or a NaN where the header is
        write_file(bad, v.data(), v.size());
        c = nullptr;
        const int st = scw_cw_load(bad.c_str(), &c);
        if (st == SCW_OK) scw_cw_destroy(c);  // a NaN field value is data, not structure
    }
    // random files
    std::uniform_int_distribution<int> byte(0, 255);
    for (int k = 0; k < 200; ++k) {
        std::vector<unsigned char> r(1 + k * 3);
        for (auto& x : r) x = (unsigned char)byte(rng);
        if (k % 2 == 0 && r.size() >= 8) {
            const uint64_t t = 8;
            std::memcpy(r.data(), &t, 8);
        }
        write_file(bad, r.data(), r.size());
        c = nullptr;
        if (scw_cw_load(bad.c_str(), &c) == SCW_OK) scw_cw_destroy(c);
    }
    CHECK(scw_cw_load((dir + "/missing.cw").c_str(), &c) == SCW_ERR_IO);
    CHECK(scw_cw_load(nullptr, &c) == SCW_ERR_ARG);
    scw_cw_destroy(cw);
}

// M x M lattice region plus its reflected copy (an interface), then a dense
// eigenmode slab of nm modes per cell, then scattered waves.
static void test_planner() {
    std::mt19937_64 rng(11);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (int M : {1, 3, 8, 65}) {
        std::vector<scw_planewave64> W;
        for (int i = 0; i < M; ++i)
            for (int j = 0; j < M; ++j)
                for (int dir = 0; dir < 2; ++dir) {
                    scw_planewave64 w{};
                    w.E0[0][0] = u(rng);
                    w.E0[1][1] = u(rng);
                    w.k[0][0] = 0.3 * (i - M / 2);
                    w.k[1][0] = 0.2 * (j - M / 2);
                    w.k[2][0] = dir ? -5.0 : 5.0;
                    W.push_back(w);
                }
        const int nm = 6;  // dense slab: every cell holds the same nm kz values
        const size_t slab0 = W.size();
        std::vector<double> kz(nm);
        for (auto& x : kz) x = 3.0 + u(rng);
        for (int i = 0; i < M; ++i)
            for (int j = 0; j < M; ++j)
                for (int m = 0; m < nm; ++m) {
                    scw_planewave64 w{};
                    w.E0[2][0] = u(rng);
                    w.k[0][0] = 0.3 * (i - M / 2);
                    w.k[1][0] = 0.2 * (j - M / 2);
                    w.k[2][0] = kz[m];
                    w.k[2][1] = 0.01 * m;
                    W.push_back(w);
                }
        const size_t scat0 = W.size();
        for (int k = 0; k < 300; ++k) {
            scw_planewave64 w{};
            w.E0[0][0] = 1.0;
            w.k[0][0] = u(rng);
            w.k[1][0] = u(rng);
            w.k[2][0] = 1.0;
            W.push_back(w);
        }
        const double z[3] = {-1.0, 0.0, 2.0};
        const int32_t b[4] = {0, (int32_t)slab0, (int32_t)scat0, (int32_t)W.size()};
        const int32_t e[4] = {(int32_t)slab0, (int32_t)scat0, (int32_t)W.size(), (int32_t)W.size()};
        std::vector<int64_t> info(8 * 4);
        CHECK(scw_test_plan_host(W.data(), W.size(), z, 3, b, e, info.data()) == SCW_OK);
        CHECK(info[0] == 1 && info[2] == M && info[3] == M && info[5] == 2);  // lattice, 2 terms per cell
        CHECK(info[8] == 1 && info[9] == 1 && info[12] == nm && info[13] == nm);  // dense slab
        CHECK(info[16] == 0);                                                // scattered: no lattice
        CHECK(info[24 + 7] == 0);                                            // empty last region
        // invalid tables
        std::vector<scw_planewave64> Wn = W;
        Wn[1].k[2][1] = NAN;
        CHECK(scw_test_plan_host(Wn.data(), Wn.size(), z, 3, b, e, info.data()) == SCW_ERR_ARG);
        const int32_t e2[4] = {(int32_t)slab0, (int32_t)scat0, (int32_t)W.size(), (int32_t)W.size() + 1};
        CHECK(scw_test_plan_host(W.data(), W.size(), z, 3, b, e2, info.data()) == SCW_ERR_ARG);
        const double zn[3] = {-1.0, INFINITY, 2.0};
        CHECK(scw_test_plan_host(W.data(), W.size(), zn, 3, b, e, info.data()) == SCW_ERR_ARG);
        CHECK(scw_test_plan_host(W.data(), W.size(), z, -1, b, e, info.data()) == SCW_ERR_ARG);
    }
    // random small tables: whatever the planner decides, it must not fault
    for (int k = 0; k < 300; ++k) {
        const size_t n = 1 + (size_t)(rng() % 200);
        auto W = random_waves(rng, n);
        for (auto& w : W)  // snap to a coarse lattice now and then, so both outcomes occur
            if (k % 3 == 0) {
                w.k[0][0] = std::round(w.k[0][0]);
                w.k[1][0] = std::round(w.k[1][0]);
                w.k[0][1] = w.k[1][1] = 0.0;
            }
        const int layers = (int)(rng() % 3);
        std::vector<double> z(layers);
        for (int l = 0; l < layers; ++l) z[l] = l;
        std::vector<int32_t> b(layers + 1), e(layers + 1);
        size_t at = 0;
        for (int r = 0; r <= layers; ++r) {
            b[r] = (int32_t)at;
            at = r == layers ? n : std::min(n, at + (size_t)(rng() % (n + 1)));
            e[r] = (int32_t)at;
        }
        std::vector<int64_t> info(8 * (layers + 1));
        CHECK(scw_test_plan_host(W.data(), n, z.data(), layers, b.data(), e.data(), info.data()) == SCW_OK);
    }
}

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    test_cw(dir);
    test_planner();
    if (g_failures) {
        std::fprintf(stderr, "%d check(s) failed\n", g_failures);
        return 1;
    }
    std::printf("host_driver: all checks passed\n");
    return 0;
}
