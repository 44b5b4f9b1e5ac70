// Internal structures shared by the host planner, the exec builder and the
// kernel launchers.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/scwatter_amd.h"

namespace scw {

typedef std::complex<double> cd;

void set_error(const char* fmt, ...);
int fail(int status, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);

#define SCW_HIP(call)                                            \
    do {                                                         \
        hipError_t e_ = (call);                                  \
        if (e_ != hipSuccess) return ::scw::hip_fail(e_, #call); \
    } while (0)

// roctx range over a scope (plan and exec builds, runs, stage launches):
// what `rocprofv3 --marker-trace` shows of the library.  Without a profiler
// attached a push / pop is a call that returns at once.
struct TraceRange {
    explicit TraceRange(const char* name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;
};

// Makes `dev` current for the scope and restores the caller's device after.
// Initialises the HIP runtime once per process, on the first thread that asks
// (the others wait): the runtime's lazy initialisation is not safe to enter
// from several threads at once, and callers create plans from thread pools.
void hip_runtime_once();

struct DeviceGuard {
    int prev = -1;
    int st = 0;
    explicit DeviceGuard(int dev) {
        hip_runtime_once();
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        hipError_t e = hipSetDevice(dev);
        if (e != hipSuccess) st = hip_fail(e, "hipSetDevice");
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// Device buffer owned by a plan or exec.  alloc / upload keep the current
// allocation when it is large enough (a rebuilt table of the same or smaller
// size then needs no hipMalloc); release frees it.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;  // size of the current contents
    size_t cap = 0;    // size of the allocation
    int alloc(size_t n);
    int upload(const void* host, size_t n);
    void release();
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Separable decomposition of one region (the GEMM form's planning result).
struct RegionPlan {
    int64_t wbegin = 0, wend = 0;  // wave range in the unpacked table
    double zref = 0.0;             // FP64 phase reference for kz
    bool real_transverse = true;   // every wave has Im kx = Im ky = 0
    // separable form
    bool separable = false;
    bool dense = false;            // every cell holds every kz value (eigenmode slabs)
    int mx = 0, my = 0, mc = 0, width = 0;
    int64_t n_terms = 0;
    std::vector<cd> kx, ky, kz;    // distinct values
    int64_t ell_off = 0;           // float2 offset of this region's ELL amplitudes
    int64_t arow = 0;              // float2 per term row: 3 cells rounded up to 16 (128-byte rows)
    int64_t ellc_off = -1;         // slot offset of its kz indices (-1: dense, index = j)
    std::string why_not;           // reason when not separable
};

}  // namespace scw

struct scw_plan {
    int device = 0;
    int layers = 0;
    int64_t n_waves = 0;
    std::vector<double> zl;
    std::vector<scw::RegionPlan> regions;
    // direct-form tables, one entry per wave (unpacked order)
    scw::DevBuf d_k64;   // double[n][6]  kx, ky, kz (re, im)
    scw::DevBuf d_e64;   // double[n][6]  E0 / s
    scw::DevBuf d_lns;   // double[n]     ln s  (s = max_c |E0_c|, 0 -> wave skipped)
    scw::DevBuf d_kf;    // float4[n]     Re kx/2pi, Re ky/2pi, -Im kx log2e, -Im ky log2e
    // GEMM-form tables, concatenated over separable regions
    scw::DevBuf d_ellA;  // float2[...][width][arow]: [cells][3] E0 re-referenced to zref, padded rows
    scw::DevBuf d_ellC;  // int32[...][width][cells]      kz index (sparse regions only)
    int64_t ell_elems = 0;
    // Dense-region amplitudes for the 16-bit GEMM dtypes in block floating
    // point: per group of 4 modes and column, 8 16-bit mantissas; per group
    // and cell, an exponent byte per mode, [region]{[mode / 4][arow] x 16
    // bytes, then [mode / 4][bfp_erow(arow)] x 4 bytes}.  Built on the first
    // 16-bit exec that needs it (scw::build_bfp).
    scw::DevBuf d_bfp;
    std::vector<int64_t> bfp_off;  // byte offset per region, -1 if none
    // direct form, per region: wave range and whether the transverse k is real
    // (scw::direct_region_fill records), and the boundaries as double[layers]
    scw::DevBuf d_dreg, d_zl;
    // execs of recent scw_eval_planes calls, most recent last (scw_planes.cpp)
    struct PlaneEntry {
        int kind, n, form, dtype;
        float start_a, start_b, fixed, d;
        double tol;
        scw_exec* ex;
    };
    std::vector<PlaneEntry> plane_cache;
};

namespace scw {

// Per-z metadata of the GEMM form (device struct).
struct ZMeta {
    int32_t region;     // region index
    int32_t my;         // distinct ky of the region
    int32_t mx;         // distinct kx of the region
    int32_t kp;         // padded K of the main GEMM (2*my rounded up to the MFMA K)
    int64_t u_off;      // float2 offset of U_z ([my][nx][3], FP32), or of the slot-summed
                        // C_z ([3 my][ka], 16-bit dtypes, zero past mx)
    int64_t yfrag_off;  // element offset of the region's Y fragment table
    int64_t zrow_off;   // float2 offset of this z's row of exp(i kz_c (z - zref))
    int64_t c_off;      // float2 offset of C_z ([cells][3]) within one partial-sum slot
    // main-GEMM entries (scw_exec::d_mzmeta): a region of more than 64 ky
    // runs its main GEMM in chunks of ky rows, one entry per (z, chunk) with
    // my, kp, u_off and yfrag_off of the chunk; entries of chunks after the
    // first add to the output slice instead of writing it
    int32_t zout;       // output z-slice
    int32_t acc;        // 1: add to the output
};

// One wave-sized work item of the coefficient contraction.
struct CoefItem {
    int32_t region;
    int32_t cell0;      // first cell (lane = cell0 + lane id)
    int32_t j0, j1;     // term slice [j0, j1)
    int32_t z0, nz;     // first z index in the grid, count (<= 8), all in `region`
    int32_t slot;       // partial-sum slot
    int32_t pad;
};

// One wave-sized work item of the MFMA mode contraction (dense regions):
// C_z[n] += sum_{c in [c0, c1)} Z[z][c] A[c][n] for real columns n in
// [n0, n0 + 64) and the run z0 .. z0 + nz - 1 (nz <= 64).
struct ModalItem {       // one wave of the MFMA mode contraction
    int32_t region;      // virtual region
    int32_t z0, nz;      // z run (nz <= 32)
    int32_t m0;          // first complex column (cell * 3 + comp); the wave covers 64
    int32_t c0, c1;      // mode range, multiples of 16 (c1 may pass the width)
    int32_t slot;        // partial-sum slot
    int32_t T;           // 16-mode steps in the run's phase-fragment table
    int64_t zf_off;      // 32-bit word offset of the run's phase fragments
    int32_t cs_direct;   // 1: the run's only slot, stored straight into the slot-summed C_z layout
    int32_t pad;
};

struct RegionDev {      // per-region data the GEMM-form kernels read (device)
    const float2* ellA; // the region's terms: [width][arow], row j = [cells][3] float2
    const int32_t* ellC;// kz index per term slot (nullptr when dense)
    int64_t x_off;      // float2 offset of X_r: [mx][nx] (FP32) or [nx][ka] (16-bit dtypes, zero past mx)
    int32_t cells, width, mc, dense;
    int32_t mx, my, nslots, arow;
    int32_t ka;         // mx rounded up to 16 (the X contraction's K in complex steps)
    int32_t erow;       // exponent words per mode group of the block floating-point table: bfp_erow(arow)
    const uint8_t* bfp;      // block floating-point mantissas of a dense region (16-bit dtypes), else nullptr
    const uint8_t* bfp_exp;  // their exponents: one word per (mode group, cell), one byte per mode
};

// Cells per mode group of the block floating-point exponent plane (the
// columns of a term row, arow, are 3 per cell, padded).
inline int64_t bfp_erow(int64_t arow) { return (arow + 2) / 3; }

// Direct form helpers (scw_direct.hip).
size_t direct_region_bytes();
void direct_region_fill(void* dst, int64_t begin, int64_t end, int real_transverse);

}  // namespace scw

struct scw_exec {
    scw_plan* plan = nullptr;
    int form = SCW_FORM_DIRECT;
    int dtype = SCW_DTYPE_FP32;
    int64_t nx = 0, ny = 0, nz = 0;
    double units = 0, flops = 0, bytes = 0;
    double stage_flops[4] = {0, 0, 0, 0}, stage_bytes[4] = {0, 0, 0, 0};
    int launches = 0;
    // Set while the tables are being rebuilt (scw_exec_set_launch) and left
    // set when a rebuild fails: the device tables may then be missing or
    // inconsistent with the work lists, so every launch entry point refuses.
    bool invalid = false;
    std::vector<int> zregion;
    // direct form
    double xc = 0, yc = 0;
    scw::DevBuf d_dx, d_dy;   // float[nx], float[ny]   x - xc, y - yc
    scw::DevBuf d_z;          // double[nz]
    scw::DevBuf d_zslab;      // int32[nz]: region per z
    int direct_ppt = 4;       // points per thread
    int direct_split = 1;     // wave-range split (partials summed by a second kernel)
    // GEMM form
    int kmax = 0;             // largest padded K
    int my_max = 0;
    int ka_max = 0;           // largest mx rounded up to 16 (regions with waves)
    int nslots_max = 1;
    std::vector<scw::CoefItem> items;
    std::vector<scw::ModalItem> mitems;   // grouped by row-tile class (MT = 1, 2, 4, 8)
    int64_t mclass_off[5] = {0, 0, 0, 0, 0};
    int modal_wg_items = 4;   // 16-bit 64-z class: items (waves) per workgroup, 4 or 8
    scw::DevBuf d_items, d_mitems, d_regions, d_zmeta, d_Z, d_Zf, d_X, d_C, d_U, d_Yfrag;
    scw::DevBuf d_mzmeta;     // main-GEMM entries, by chunk level (level 0: one per z, in z order)
    std::vector<int64_t> mlevel_off;  // first entry of each chunk level, then the entry count
    // Stage-form main launches: the same entries regrouped, within each
    // level, by the kernel K their chunk needs; launch i covers d_mksort
    // entries [mk_launch[3 i], + mk_launch[3 i + 1]) with kernel K mk_launch[3 i + 2]
    scw::DevBuf d_mksort;
    std::vector<int64_t> mk_launch;
    scw::DevBuf d_bfpx;       // FP16: shifted exponent planes of the dense regions, then the shifts
    scw::DevBuf d_cblk;       // 16-bit dtypes: slot-sum workgroups, int2 (z, first element of C_z)
    int64_t ncblk = 0;
    int64_t tune_coef_items = 0;   // launch-shape choices (0 = default)
    int64_t tune_main_blocks = 0;
    int64_t tune_modal_run = 0;    // z-run tiles (0 = default: 64 mode-contraction, 8 VALU)
    int64_t tune_coef_run = 0;
    std::vector<double> hx, hy;    // grid kept for re-planning launch shapes
    std::vector<scw_plan*> batch;  // plans of a batch exec (empty for a single plan)
    std::vector<double> hz;        // z of every output slice
    std::vector<int> hzp;          // plan index (into batch, or 0) of every output slice
    int64_t c_elems = 0;      // float2 per slot of the C buffer
    int64_t ylo_delta = 0;    // element offset of the lo Y fragments (bf16 x3)
    int ysplit = 1;           // grid.z of the main GEMM
    int ytiles_per_split = 1;
    // single-launch form (16-bit dtypes, scw_exec_set_persistent)
    bool persistent = false;
    int64_t nunits = 0;       // work units of the persistent kernel
    int64_t pgrid = 0;        // its workgroups
    int64_t tune_grid = 0;    // requested workgroups (0 = resident capacity)
    int64_t order_lag = -1;   // unit-order lag (-1 = default)
    int64_t main_tiles = 0;   // main-GEMM column blocks per unit (0 = default)
    double wait_ms = 0;       // limit of one dependency wait
    uint64_t wait_ticks = 0;  // the same in wall-clock ticks
    uint64_t wait_polls = 0;  // poll bound of a wait (10^5 per ms, each poll sleeps >= 128 cycles)
    bool stall = false;       // test hook: one wait can never be met
    std::vector<int2> hcblk;  // slot-sum workgroups (z, first element), host copy
    scw::DevBuf d_units, d_need, d_ctl;
    scw::DevBuf d_out;        // staging for scw_exec_run_host
    // graph
    hipGraph_t graph = nullptr;
    hipStream_t graph_stream = nullptr;
    hipGraphExec_t graph_exec = nullptr;
};

namespace scw {
// kernel launchers (scw_direct.hip, scw_gemm.hip)
int launch_direct(const scw_exec* ex, float* out, hipStream_t s);
// the direct form's wave-range split for a grid of pts points per slice, nz
// slices and at most maxw waves per region (shared by execs and the
// allocation-free plane evaluation, so both sum in the same order)
int direct_split(int64_t pts, int64_t nz, int64_t maxw, int ppt);
int direct_ppt(int64_t pts);
// One direct-form evaluation of an nx x ny x nz grid from device tables
// (region_table: the plan's d_dreg records): partial sums of the nsplit wave
// slices go to part (nsplit > 1) and are added into out by a second kernel.
int launch_direct_raw(const scw_plan* P, const void* region_table, const int32_t* zreg, const double* zs,
                      const float* dxs, const float* dys, int64_t nx, int64_t ny, int64_t nz, double xc, double yc,
                      int ppt, int nsplit, float* part, float* out, hipStream_t st);
int launch_gemm(const scw_exec* ex, float* out, hipStream_t s, int stage = 0);
// the single-launch kernel: launch (per_cu == nullptr) or its resident
// workgroups per CU (scw_gemm.hip)
int fused_launch(const scw_exec* ex, float* out, hipStream_t s, int* per_cu);
int build_bfp(scw_plan* plan);  // fills plan->d_bfp / bfp_off once (scw_gemm.hip)
// (re)builds a single-plan exec in place (scw_exec.cpp); on failure it is left invalid
int exec_init(scw_exec* ex, scw_plan* plan, const double* xs, size_t nx, const double* ys, size_t ny,
              const double* zs, size_t nz, double region_tol, int form, int dtype);
// dst = the exponent plane src (n words, erow per mode group) with the byte
// of mode c raised by shift[c] (device pointers; scw_gemm.hip)
int shift_bfp_exponents(const uint8_t* src, uint8_t* dst, int64_t erow, int64_t n, const int32_t* shift);
}  // namespace scw
