// C++ data model of a coupled-wave structure, in the shape of the
// reference's CoupledWaveStructure<T> (CoupledWaveStructure.h:57-148), with a
// POD plane-wave record instead of tira::planewave<T> (whose byte layout the
// .cw format documents as E0.xyz then k.xyz, each (re, im);
// CoupledWaveStructure.h:7-55).  File I/O goes through the C ABI
// (scw_cw_load / scw_cw_save), which validates what it reads.
#pragma once

#include <array>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

#include "scwatter_amd.h"

namespace scw {

template <typename T>
struct planewave {
    std::complex<T> E0[3];
    std::complex<T> k[3];
    planewave() = default;
    planewave(std::complex<T> kx, std::complex<T> ky, std::complex<T> kz, std::complex<T> ex, std::complex<T> ey,
              std::complex<T> ez)
        : E0{ex, ey, ez}, k{kx, ky, kz} {}
    std::array<std::complex<T>, 3> getE0() const { return {E0[0], E0[1], E0[2]}; }
    std::array<std::complex<T>, 3> getK() const { return {k[0], k[1], k[2]}; }
};
static_assert(sizeof(planewave<double>) == sizeof(scw_planewave64), "planewave<double> must match the .cw record");

}  // namespace scw

template <typename T>
struct HomogeneousLayer {
    T z;                                   // z coordinate of the boundary
    std::vector<scw::planewave<T>> Pr;     // waves reflected off the boundary (along -z)
    std::vector<scw::planewave<T>> Pt;     // waves transmitted through it (along +z)
};

template <typename T>
class CoupledWaveStructure {
   public:
    std::vector<scw::planewave<T>> Pi;     // incident waves
    std::vector<HomogeneousLayer<T>> Layers;

    bool save(std::string filename) const {
        static_assert(sizeof(T) == 8, ".cw files are written in double precision");
        scw_cw* cw = nullptr;
        if (scw_cw_create(&cw) != SCW_OK) return false;
        bool ok = scw_cw_set_incident(cw, rec(Pi), Pi.size()) == SCW_OK;
        for (size_t l = 0; ok && l < Layers.size(); ++l)
            ok = scw_cw_add_layer(cw, Layers[l].z, rec(Layers[l].Pr), Layers[l].Pr.size(), rec(Layers[l].Pt),
                                  Layers[l].Pt.size()) == SCW_OK;
        ok = ok && scw_cw_save(cw, filename.c_str()) == SCW_OK;
        scw_cw_destroy(cw);
        return ok;
    }

    bool load(std::string filename) {
        static_assert(sizeof(T) == 8, "load into CoupledWaveStructure<double>");
        scw_cw* cw = nullptr;
        if (scw_cw_load(filename.c_str(), &cw) != SCW_OK) return false;
        size_t n_i = 0, n_l = 0;
        scw_cw_counts(cw, &n_i, &n_l, nullptr);
        std::vector<scw_planewave64> W;
        std::vector<double> zl(n_l);
        std::vector<int32_t> b(n_l + 1), e(n_l + 1);
        size_t total = 0;
        scw_cw_counts(cw, nullptr, nullptr, &total);
        W.resize(total);
        scw_cw_unpack(cw, W.data(), zl.data(), b.data(), e.data());
        Pi.assign(reinterpret_cast<const scw::planewave<T>*>(W.data()),
                  reinterpret_cast<const scw::planewave<T>*>(W.data()) + n_i);
        Layers.resize(n_l);
        for (size_t l = 0; l < n_l; ++l) {
            size_t nr = 0, nt = 0;
            scw_cw_layer(cw, l, nullptr, &nr, &nt);
            const scw::planewave<T>* p = reinterpret_cast<const scw::planewave<T>*>(W.data()) + b[l];
            // waves of region l are [Pt_{l-1}, Pr_l); Pr_l is the tail of that range.
            const scw::planewave<T>* pr = p + (e[l] - b[l]) - nr;
            Layers[l].z = zl[l];
            Layers[l].Pr.assign(pr, pr + nr);
            const scw::planewave<T>* pt = reinterpret_cast<const scw::planewave<T>*>(W.data()) + b[l + 1];
            Layers[l].Pt.assign(pt, pt + nt);
        }
        scw_cw_destroy(cw);
        return true;
    }

    // CoupledWaveStructure.h:141-147, kept as the reference has it: l + 1 for
    // the FIRST boundary l with z >= Layers[l].z, else 0.  With ascending
    // boundaries that is 1 for every z at or above the first boundary, which
    // is not the evaluator's region rule (1 + the LAST such l, scwatter_amd.h);
    // the reference never calls it, and neither does this library.
    size_t getPlaneIndex(T z) const {
        for (size_t l = 0; l < Layers.size(); l++)
            if (z >= Layers[l].z) return l + 1;
        return 0;
    }

   private:
    static const scw_planewave64* rec(const std::vector<scw::planewave<T>>& v) {
        return reinterpret_cast<const scw_planewave64*>(v.data());
    }
};
