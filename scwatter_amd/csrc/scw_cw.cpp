// .cw coupled-wave structure files (format: CoupledWaveStructure.h:7-55).
//
//   [u64 sizeof_T] [u64 nPi] nPi x record
//   [u64 nLayers]  per layer: [T z] [u64 nPr] nPr x record [u64 nPt] nPt x record
//   record = E0.x E0.y E0.z k.x k.y k.z, each (re, im), 12 values of T
//
// save() writes T = double like CoupledWaveStructure<double>::save
// (CoupledWaveStructure.h:72-101).  load() differs from the reference's
// (CoupledWaveStructure.h:103-134), which ignores the precision field and
// trusts every count: it accepts sizeof_T 4 or 8 (widening floats), checks
// every count against the bytes left in the file before allocating, and
// reports a truncated or malformed file as an error instead of returning
// partly filled vectors.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "scw_internal.hpp"

struct scw_cw {
    struct Layer {
        double z = 0.0;
        std::vector<scw_planewave64> Pr, Pt;
    };
    std::vector<scw_planewave64> Pi;
    std::vector<Layer> layers;
};

namespace scw {

namespace {

struct Reader {
    FILE* f = nullptr;
    uint64_t left = 0;
    bool read(void* dst, size_t n) {
        if (n > left) return false;
        if (fread(dst, 1, n, f) != n) return false;
        left -= n;
        return true;
    }
};

bool read_records(Reader& rd, size_t elem, uint64_t count, std::vector<scw_planewave64>& out) {
    if (count > rd.left / (12 * elem)) return false;
    out.resize((size_t)count);
    if (elem == 8) return rd.read(out.data(), (size_t)count * sizeof(scw_planewave64));
    std::vector<float> tmp((size_t)count * 12);
    if (!rd.read(tmp.data(), tmp.size() * 4)) return false;
    for (size_t i = 0; i < (size_t)count; ++i) {
        double* d = &out[i].E0[0][0];
        for (int j = 0; j < 12; ++j) d[j] = tmp[12 * i + j];
    }
    return true;
}

}  // namespace
}  // namespace scw

using namespace scw;

extern "C" {

int scw_cw_create(scw_cw** cw) {
    if (!cw) return fail(SCW_ERR_ARG, "NULL output");
    *cw = new (std::nothrow) scw_cw();
    return *cw ? SCW_OK : fail(SCW_ERR_NOMEM, "host allocation failed");
}

int scw_cw_destroy(scw_cw* cw) {
    delete cw;
    return SCW_OK;
}

int scw_cw_load(const char* path, scw_cw** out) {
    if (!path || !out) return fail(SCW_ERR_ARG, "NULL argument");
    *out = nullptr;
    FILE* f = fopen(path, "rb");
    if (!f) return fail(SCW_ERR_IO, "cannot open %s", path);
    Reader rd;
    rd.f = f;
    if (fseek(f, 0, SEEK_END) != 0) {
        fclose(f);
        return fail(SCW_ERR_IO, "cannot seek %s", path);
    }
    const long sz = ftell(f);
    rewind(f);
    rd.left = sz > 0 ? (uint64_t)sz : 0;
    scw_cw* cw = new (std::nothrow) scw_cw();
    if (!cw) {
        fclose(f);
        return fail(SCW_ERR_NOMEM, "host allocation failed");
    }
    int st = SCW_OK;
    try {
        uint64_t tsize = 0, n = 0, nl = 0;
        if (!rd.read(&tsize, 8)) st = fail(SCW_ERR_IO, "%s: truncated header", path);
        else if (tsize != 8 && tsize != 4) st = fail(SCW_ERR_FORMAT, "%s: precision field %llu (expected 4 or 8)", path,
                                                     (unsigned long long)tsize);
        else if (!rd.read(&n, 8) || !read_records(rd, (size_t)tsize, n, cw->Pi))
            st = fail(SCW_ERR_IO, "%s: truncated or invalid incident wave block", path);
        else if (!rd.read(&nl, 8) || nl > rd.left / (tsize + 16))
            st = fail(SCW_ERR_IO, "%s: truncated or invalid layer count", path);
        else {
            cw->layers.resize((size_t)nl);
            for (uint64_t l = 0; l < nl && st == SCW_OK; ++l) {
                scw_cw::Layer& L = cw->layers[(size_t)l];
                uint64_t nr = 0, nt = 0;
                if (tsize == 8) {
                    if (!rd.read(&L.z, 8)) st = fail(SCW_ERR_IO, "%s: truncated layer %llu", path, (unsigned long long)l);
                } else {
                    float zf;
                    if (!rd.read(&zf, 4)) st = fail(SCW_ERR_IO, "%s: truncated layer %llu", path, (unsigned long long)l);
                    L.z = zf;
                }
                if (st == SCW_OK && (!rd.read(&nr, 8) || !read_records(rd, (size_t)tsize, nr, L.Pr) ||
                                     !rd.read(&nt, 8) || !read_records(rd, (size_t)tsize, nt, L.Pt)))
                    st = fail(SCW_ERR_IO, "%s: truncated or invalid wave block in layer %llu", path,
                              (unsigned long long)l);
            }
        }
    } catch (const std::bad_alloc&) {
        st = fail(SCW_ERR_NOMEM, "host allocation failed while reading %s", path);
    }
    fclose(f);
    if (st != SCW_OK) {
        delete cw;
        return st;
    }
    *out = cw;
    return SCW_OK;
}

int scw_cw_save(const scw_cw* cw, const char* path) {
    if (!cw || !path) return fail(SCW_ERR_ARG, "NULL argument");
    FILE* f = fopen(path, "wb");
    if (!f) return fail(SCW_ERR_IO, "cannot create %s", path);
    bool ok = true;
    auto put = [&](const void* p, size_t n) {
        if (ok && n) ok = fwrite(p, 1, n, f) == n;
    };
    const uint64_t tsize = 8, n = cw->Pi.size(), nl = cw->layers.size();
    put(&tsize, 8);
    put(&n, 8);
    put(cw->Pi.data(), n * sizeof(scw_planewave64));
    put(&nl, 8);
    for (const auto& L : cw->layers) {
        const uint64_t nr = L.Pr.size(), nt = L.Pt.size();
        put(&L.z, 8);
        put(&nr, 8);
        put(L.Pr.data(), nr * sizeof(scw_planewave64));
        put(&nt, 8);
        put(L.Pt.data(), nt * sizeof(scw_planewave64));
    }
    if (fclose(f) != 0) ok = false;
    return ok ? SCW_OK : fail(SCW_ERR_IO, "write failed: %s", path);
}

int scw_cw_set_incident(scw_cw* cw, const scw_planewave64* Pi, size_t n) {
    if (!cw || (n && !Pi)) return fail(SCW_ERR_ARG, "NULL argument");
    try {
        cw->Pi.assign(Pi, Pi + n);
    } catch (const std::bad_alloc&) {
        return fail(SCW_ERR_NOMEM, "host allocation failed");
    }
    return SCW_OK;
}

int scw_cw_add_layer(scw_cw* cw, double z, const scw_planewave64* Pr, size_t n_r, const scw_planewave64* Pt,
                     size_t n_t) {
    if (!cw || (n_r && !Pr) || (n_t && !Pt)) return fail(SCW_ERR_ARG, "NULL argument");
    try {
        scw_cw::Layer L;
        L.z = z;
        L.Pr.assign(Pr, Pr + n_r);
        L.Pt.assign(Pt, Pt + n_t);
        cw->layers.push_back(std::move(L));
    } catch (const std::bad_alloc&) {
        return fail(SCW_ERR_NOMEM, "host allocation failed");
    }
    return SCW_OK;
}

int scw_cw_counts(const scw_cw* cw, size_t* n_incident, size_t* n_layers, size_t* n_waves) {
    if (!cw) return fail(SCW_ERR_ARG, "cw is NULL");
    size_t total = cw->Pi.size();
    for (const auto& L : cw->layers) total += L.Pr.size() + L.Pt.size();
    if (n_incident) *n_incident = cw->Pi.size();
    if (n_layers) *n_layers = cw->layers.size();
    if (n_waves) *n_waves = total;
    return SCW_OK;
}

int scw_cw_layer(const scw_cw* cw, size_t layer, double* z, size_t* n_r, size_t* n_t) {
    if (!cw) return fail(SCW_ERR_ARG, "cw is NULL");
    if (layer >= cw->layers.size()) return fail(SCW_ERR_ARG, "layer %zu out of range", layer);
    if (z) *z = cw->layers[layer].z;
    if (n_r) *n_r = cw->layers[layer].Pr.size();
    if (n_t) *n_t = cw->layers[layer].Pt.size();
    return SCW_OK;
}

// cw_allocate (cpuEvaluator.cpp:34-53) + cw_unpack (cpuEvaluator.cpp:61-102):
// W = Pi, then for each layer l: Pr_l, Pt_l.  Region 0 = Pi + Pr_0,
// region l = Pt_{l-1} + Pr_l, region L = Pt_{L-1}.
int scw_cw_unpack(const scw_cw* cw, scw_planewave64* waves, double* z_layers, int32_t* waves_begin,
                  int32_t* waves_end) {
    if (!cw || !waves_begin || !waves_end) return fail(SCW_ERR_ARG, "NULL argument");
    size_t total = 0;
    scw_cw_counts(cw, nullptr, nullptr, &total);
    if (total > 0 && !waves) return fail(SCW_ERR_ARG, "waves is NULL");
    if (!cw->layers.empty() && !z_layers) return fail(SCW_ERR_ARG, "z_layers is NULL");
    if (total > (size_t)INT32_MAX) return fail(SCW_ERR_ARG, "more than 2^31-1 waves");
    size_t idx = 0;
    auto copy = [&](const std::vector<scw_planewave64>& v) {
        if (!v.empty()) std::memcpy(waves + idx, v.data(), v.size() * sizeof(scw_planewave64));
        idx += v.size();
    };
    waves_begin[0] = 0;
    copy(cw->Pi);
    for (size_t l = 0; l < cw->layers.size(); ++l) {
        z_layers[l] = cw->layers[l].z;
        copy(cw->layers[l].Pr);
        waves_end[l] = (int32_t)idx;
        waves_begin[l + 1] = (int32_t)idx;
        copy(cw->layers[l].Pt);
    }
    waves_end[cw->layers.size()] = (int32_t)idx;
    return SCW_OK;
}

}  // extern "C"
