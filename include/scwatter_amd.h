/*
 * scwatter_amd — MI355X (gfx950) plane-wave field evaluator, C ABI.
 *
 * Evaluates  E(r) = sum_w E0_w exp(i k_w . r)  over the plane waves of the
 * layer region that holds r, on rectilinear 3D grids, with hand-written HIP
 * kernels: a direct per-point summation form and a separable MFMA GEMM form
 * E_z = X . (A o e^{i kz z}) . Y^T.
 *
 * These entry points replace the reference's evaluator module:
 *   cw_allocate / cw_unpack          cpuEvaluator.cpp:34-102   -> scw_cw_unpack
 *   gpu_initialize                   gpuEvaluator.cu:15-26     -> scw_plan_create
 *   gpu_cw_evaluate, kernel_{xy,xz,yz} gpuEvaluator.cu:70-165  -> scw_exec_create + scw_exec_run
 *   cpu_cw_evaluate_{xy,yz,xz}       cpuEvaluator.cpp:294-407  -> scw_exec_create + scw_exec_run
 *   CoupledWaveStructure::load/save  CoupledWaveStructure.h:72-134 -> scw_cw_load / scw_cw_save
 * The reference's own C++ signatures are kept in scw_evaluator.hpp, which is
 * implemented on top of this ABI.
 *
 * Conventions
 *  - Every function returns SCW_OK (0) or a negative status; it never exits
 *    or aborts.  scw_last_error() gives a message for the calling thread.
 *  - Output layout (the reference's AoS, gpuEvaluator.cu:76): complex64
 *    [nz][ny][nx][3] = per point Ex.re Ex.im Ey.re Ey.im Ez.re Ez.im (float32).
 *    A reference xy plane is nz = 1, an xz plane ny = 1, a yz plane nx = 1.
 *  - Region of a point: 1 + the largest l with z >= z_layers[l] - region_tol
 *    (0 if none), compared in double precision.  region_tol = 0 is the CPU
 *    rule (cpuEvaluator.cpp:303), region_tol = 0.001 the GPU rule
 *    (gpuEvaluator.cu:34).
 *  - Waves of region r are waves[waves_begin[r] .. waves_end[r]) (the
 *    cw_allocate tables, cpuEvaluator.cpp:34-53).
 */
#ifndef SCWATTER_AMD_H
#define SCWATTER_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCW_ABI_VERSION 1

/* One plane wave as stored in a .cw file (CoupledWaveStructure.h:7-55):
 * E0.x E0.y E0.z then k.x k.y k.z, each (re, im), FP64.  96 bytes. */
typedef struct scw_planewave64 {
    double E0[3][2];
    double k[3][2];
} scw_planewave64;

typedef struct scw_plan scw_plan; /* device-resident wave tables of one structure */
typedef struct scw_exec scw_exec; /* a plan bound to one grid, form and dtype     */
typedef struct scw_cw scw_cw;     /* an in-memory .cw coupled-wave structure       */

enum scw_status {
    SCW_OK = 0,
    SCW_ERR_ARG = -1,         /* invalid argument or table                     */
    SCW_ERR_NOMEM = -2,       /* host or device allocation failed              */
    SCW_ERR_DEVICE = -3,      /* HIP runtime error / no device                 */
    SCW_ERR_IO = -4,          /* file missing, unreadable or truncated         */
    SCW_ERR_FORMAT = -5,      /* .cw content invalid                           */
    SCW_ERR_UNSUPPORTED = -6, /* form/dtype not available for this structure  */
    SCW_ERR_TIMEOUT = -7      /* a single-launch run's dependency wait expired */
};

enum scw_form {
    SCW_FORM_AUTO = 0,   /* GEMM when every region on the grid is separable  */
    SCW_FORM_DIRECT = 1, /* per-point direct summation, FP32                  */
    SCW_FORM_GEMM = 2    /* separable MFMA form                               */
};

enum scw_dtype {         /* operand type of the GEMM form's MFMA stage        */
    SCW_DTYPE_FP32 = 0,  /* v_mfma_f32_16x16x4_f32                            */
    SCW_DTYPE_BF16 = 1,  /* v_mfma_f32_16x16x32_bf16                          */
    SCW_DTYPE_BF16X3 = 2,/* split hi+lo bf16: hi*hi + hi*lo + lo*hi           */
    SCW_DTYPE_FP16 = 3   /* v_mfma_f32_16x16x32_f16.  Each mode's phase factor  */
                         /* and amplitude are rescaled by a power of two, but  */
                         /* the field and its partial sums must stay within    */
                         /* FP16's range (|value| < 65504).                    */
};

/* ---- errors / device --------------------------------------------------- */
const char* scw_last_error(void);
int scw_abi_version(void);
int scw_device_count(int* count);
/* hipMemGetInfo of a device (fills the reference's free/total_gpu_memory). */
int scw_device_mem_info(int device, size_t* free_bytes, size_t* total_bytes);

/* ---- structure plan (replaces gpu_initialize, gpuEvaluator.cu:15-26) ---- */
/* waves: n_waves records in cw_unpack order; z_layers: layers boundaries;
 * waves_begin / waves_end: layers + 1 entries each.  All host pointers; the
 * plan copies what it needs to `device` and keeps no reference to them. */
int scw_plan_create(int device, const scw_planewave64* waves, size_t n_waves, const double* z_layers,
                    int layers, const int32_t* waves_begin, const int32_t* waves_end, scw_plan** plan);
int scw_plan_destroy(scw_plan* plan);
int scw_plan_layers(const scw_plan* plan, int* layers);
/* Per-region structure: wave count, distinct kx / ky / kz counts, terms after
 * merging identical k, ELL width, and whether the GEMM form applies. */
int scw_plan_region_info(const scw_plan* plan, int region, int64_t* n_waves, int32_t* mx, int32_t* my,
                         int32_t* mc, int64_t* n_terms, int32_t* width, int32_t* separable);
/* Device memory held by the plan; grows once when the first exec with a
 * 16-bit dtype adds the block floating-point amplitude table of its dense
 * regions. */
int scw_plan_device_bytes(const scw_plan* plan, size_t* bytes);

/* ---- grid execution (replaces gpu_cw_evaluate / kernel_{xy,xz,yz}) ------ */
/* Rectilinear grid: xs[nx], ys[ny], zs[nz] (host, double).  Builds the
 * per-grid tables (FP64 on the host, rounded once) and allocates scratch.
 * With SCW_FORM_AUTO, a separable structure whose grid the GEMM form cannot
 * take (a phase factor exp(i kz (z - zref)) past FP32's range, far from the
 * region's phase origin in a strongly evanescent mode) gets the direct form;
 * scw_exec_info reports the form chosen. */
int scw_exec_create(scw_plan* plan, const double* xs, size_t nx, const double* ys, size_t ny,
                    const double* zs, size_t nz, double region_tol, int form, int dtype, scw_exec** exec);
/* Launch the evaluation into out_device (nz*ny*nx*3 complex64) on `stream`
 * (a hipStream_t, NULL = default stream).  Asynchronous; allocates nothing. */
int scw_exec_run(scw_exec* exec, void* out_device, void* stream);
/* Launch one stage of a run (profiling hook): 0 = all, 1 = coefficient
 * contraction, 2 = X contraction (16-bit dtypes: the slot sum of the
 * coefficients, which also contracts regions of at most 4 terms per cell;
 * their X contraction runs in stage 3), 3 = main MFMA GEMM
 * (GEMM form); the direct form has a single stage.  Later stages read what
 * earlier stages wrote. */
int scw_exec_run_stage(scw_exec* exec, int stage, void* out_device, void* stream);
/* Same as scw_exec_run, then copy to host memory out_host and wait. */
int scw_exec_run_host(scw_exec* exec, void* out_host, void* stream);
int scw_exec_destroy(scw_exec* exec);
/* form actually used (after AUTO), field pts x waves units, algorithmic FLOPs
 * and HBM bytes of one run, and the number of kernel launches per run. */
int scw_exec_info(const scw_exec* exec, int* form, int* dtype, double* units, double* flops, double* bytes,
                  int* launches);
/* Algorithmic FLOPs and HBM bytes of one stage (numbering as in
 * scw_exec_run_stage; 0 = the whole run).  FLOPs count 8 per complex
 * multiply-add whatever the dtype; bytes count each table read and each
 * result written once. */
int scw_exec_stage_info(const scw_exec* exec, int stage, double* flops, double* bytes);
/* Batch of independent structures in one run (GEMM form): instance i is
 * plans[i] on the grid xs x ys x zs[i] (nz[i] slices; shared x and y axes).
 * Output: the instances' fields concatenated along z,
 * complex64 [sum nz[i]][ny][nx][3].  Every stage is one launch over all
 * instances (uneven instance sizes are balanced by the work-item split), so
 * a captured run is one graph launch for the whole batch.  All plans must
 * live on the same device. */
int scw_batch_exec_create(int n, scw_plan* const* plans, const double* xs, size_t nx, const double* ys, size_t ny,
                          const double* const* zs, const size_t* nz, double region_tol, int dtype, scw_exec** exec);
/* Launch shape of the GEMM form: the target number of coefficient work
 * items (wave-sized) and of main-GEMM workgroups; 0 keeps the default.
 * Rebuilds the work lists (drops a captured graph), after waiting for the
 * device: runs of the exec still in flight finish on the old tables.  Used by
 * the autotuner.
 * If the rebuild fails (e.g. SCW_ERR_NOMEM), every run, stage, graph capture
 * and graph launch of the exec returns SCW_ERR_ARG until a later call
 * succeeds; the exec can always be destroyed. */
int scw_exec_set_launch(scw_exec* exec, int64_t coef_items, int64_t main_blocks);
/* Tile shape of the GEMM form's coefficient contraction: the z runs of one
 * work item, at most modal_run slices for the MFMA mode contraction (1..64;
 * its row-tile class is 1, 2, 4 or 8 tiles of 8 z) and at most coef_run for
 * the VALU term contraction (1..8); 0 keeps the default (64, 8).  Rebuilds
 * and fails like scw_exec_set_launch.  Used by the autotuner. */
int scw_exec_set_tiles(scw_exec* exec, int modal_run, int coef_run);
/* Single-launch form of a GEMM-form exec with a 16-bit dtype (no
 * counterpart in the reference, whose gpu_cw_evaluate launches one kernel per
 * plane, gpuEvaluator.cu:149-157): with on != 0, every later run (and graph
 * capture) is ONE persistent kernel launch that takes the coefficient,
 * slot-sum and main-GEMM work units of the run from a device queue in
 * dependency order; a unit waits for the units it reads on device counters,
 * each wait at most wait_ms (0 = 10000 ms).  workgroups = 0 launches the
 * resident capacity (workgroups per CU x CUs); a larger request is clamped
 * to it.  Results are bit-identical to the stage-by-stage runs.  An expired
 * wait does not stop the launch (it still ends) but leaves its output
 * invalid and sets an error word that scw_exec_check reports.  on = 0
 * returns to stage-by-stage launches.  Drops a captured graph; the form
 * survives scw_exec_set_launch.  SCW_ERR_UNSUPPORTED for FP32 or the direct
 * form.  Runs of one exec must not overlap (they share its counters). */
int scw_exec_set_persistent(scw_exec* exec, int on, int64_t workgroups, double wait_ms);
/* on: whether the single-launch form is active; units / workgroups: its work
 * units and launch size (0 when off).  Pointers may be NULL. */
int scw_exec_persistent_info(const scw_exec* exec, int* on, int64_t* units, int64_t* workgroups);
/* Waits for `stream`, then reports (and clears) an expired wait of the
 * single-launch runs since the last check: SCW_ERR_TIMEOUT, else SCW_OK
 * (always SCW_OK for stage-by-stage execs). */
int scw_exec_check(scw_exec* exec, void* stream);
/* Capture one run into a hipGraph and replay it with scw_exec_graph_launch. */
int scw_exec_graph_capture(scw_exec* exec, void* out_device, void* stream);
int scw_exec_graph_launch(scw_exec* exec, void* stream);

/* ---- the reference's three planes (gpu_cw_evaluate, gpuEvaluator.cu:131-165) ---
 * Layouts as gpu_cw_evaluate: E_xy[iy N + ix] (plane
 * z), E_xz[iz N + ix] (plane y), E_yz[iz N + iy] (plane x), each N x N x 3
 * complex64 in device memory (NULL skips that plane), with the axis
 * coordinates start + d * i rounded as one fused multiply-add
 * (gpuEvaluator.cu:78-79) and the region rule z >= z_l - region_tol (0.001 is
 * the reference GPU rule, gpuEvaluator.cu:34).  form / dtype as for
 * scw_exec_create.  Asynchronous on `stream`.  The plan keeps the execs of
 * the last six planes (keyed by plane, N, d, starts, fixed coordinate,
 * region_tol, form and dtype): a repeated plane launches at once, a changed
 * one is rebuilt (host tables and device uploads, into the buffers of the
 * least recently used exec once six are held).  Calls on one plan must not
 * overlap. */
int scw_eval_planes(scw_plan* plan, float x_start, float y_start, float z_start, float x, float y, float z,
                    float d, int N, double region_tol, int form, int dtype, void* E_xy, void* E_xz, void* E_yz,
                    void* stream);

/* Allocation-free variant of scw_eval_planes' direct form.  No counterpart in
 * the reference (its gpu_cw_evaluate allocates and frees a device frame on
 * every call, gpuEvaluator.cu:137-163); it is for callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving or GUI
 * loops, harnesses that forbid hipMalloc.  Same arguments, coordinates,
 * region rule and layouts as scw_eval_planes with form SCW_FORM_DIRECT; the
 * planes are bit-identical to it.  The per-plane tables and the partial sums
 * of the wave-range split live in one fixed 16 MiB device array per device,
 * held until the process ends; N is limited to 16384 (larger N returns
 * SCW_ERR_ARG), and the split is reduced where its partial sums would not fit
 * (the sums are then still correct, but no longer bit-identical).  Calls on
 * one device share that array: calls that overlap on one device race. */
int scw_eval_planes_noalloc(const scw_plan* plan, float x_start, float y_start, float z_start, float x, float y,
                            float z, float d, int N, double region_tol, void* E_xy, void* E_xz, void* E_yz,
                            void* stream);

/* ---- On-device post-processing of fields (sCWatterView.cpp:194-352) -----
 * Device buffers, asynchronous on `stream`, no allocation.  Scalars and
 * colours are bit-identical to the reference's float arithmetic; min/max
 * follow its comparisons with two stated exceptions (scw_post_minmax).
 * scw_post_scalar: EvaluateScalarSlices (sCWatterView.cpp:299-352).  field:
 *   complex64 [n][3]; scalar: complex64 [n] = E[mode] (0 x, 1 y, 2 z) or
 *   mode 3: sum_c E_c conj(E_c) (the intensity, imaginary part 0). */
int scw_post_scalar(const void* field_device, size_t n, int mode, void* scalar_device, void* stream);
/* scw_post_minmax: CalculateMinMax (sCWatterView.cpp:264-296).  minmax: float
 * [4] on the device = (min Re, max Re, min Im, max Im) of scalar[0..n); with
 * accumulate != 0 the previous contents are included, so several slices give
 * one range.  As in the reference, the seed is scalar[0] (or, with
 * accumulate, the previous contents): a NaN seed makes that result NaN, a NaN
 * anywhere else is skipped, and no entry at all gives NaN.  Exceptions: a NaN
 * result is the quiet NaN 0x7fc00000 (not the seed's bits), and -0.0 counts
 * as below +0.0 (the reference keeps whichever of the two it met first). */
int scw_post_minmax(const void* scalar_device, size_t n, float* minmax_device, int accumulate, void* stream);
/* scw_post_color: EvaluateColorValue / EvaluateColorSlices
 * (sCWatterView.cpp:194-262).  lowhigh: float[4] on the device = (real_low,
 * real_high, imag_low, imag_high), the layout scw_post_minmax writes; rgb:
 * uint8 [n][3] = 255 (clamp((Re - real_low) / (real_high - real_low))
 * real_color + clamp((Im - imag_low) / (imag_high - imag_low)) imag_color),
 * truncated; sums past 255 saturate (the reference's conversion is
 * undefined there). */
int scw_post_color(const void* scalar_device, size_t n, const float* lowhigh_device, const float real_color[3],
                   const float imag_color[3], int show_real, int show_imag, void* rgb_device, void* stream);

/* ---- test hook (no counterpart in the reference) -------------------------
 * Makes the (k+1)-th device buffer request from now on (a new allocation or
 * the reuse of a large enough one) fail as SCW_ERR_NOMEM
 * (k < 0: off), so that tests can drive the allocation-failure paths without
 * exhausting the device.  Process-wide; not for production use. */
void scw_test_fail_alloc_after(int64_t k);
/* With on != 0, the single-launch form of `exec` counts one slot-sum unit too
 * many for z-slice 0, so its main units' wait can never be met and must
 * expire (a host-side condition: every kernel still ends normally); on = 0
 * restores the counts.  Not for production use. */
int scw_test_exec_stall(scw_exec* exec, int on);
/* The host half of scw_plan_create (the same validation and FP64 region
 * planning, no device): per region r, info[8 r ..] = separable, dense, mx,
 * my, mc, width, n_terms, waves.  For tests of the planner without a GPU
 * (the host sanitizer build).  Not for production use. */
int scw_test_plan_host(const scw_planewave64* waves, size_t n_waves, const double* z_layers, int layers,
                       const int32_t* waves_begin, const int32_t* waves_end, int64_t* info);
/* Test hook (no counterpart in the reference; device-free): the dispatch
 * order the BF16x3/BF16/FP16 64-z mode contraction gives its workgroup
 * groups over nxcd XCDs.  Group g has group_wgs[g] workgroups of
 * group_steps[g] steps, numbered consecutively; order[k] receives the
 * workgroup of dispatch slot k, or -1 for an empty slot; *count the slot
 * count (SCW_ERR_ARG if it exceeds capacity). */
int scw_test_xcd_order(int ngroups, const int64_t* group_wgs, const int64_t* group_steps, int nxcd, int64_t* order,
                       int64_t capacity, int64_t* count);

/* ---- .cw files (CoupledWaveStructure.h:7-134) --------------------------- */
int scw_cw_load(const char* path, scw_cw** cw);
int scw_cw_save(const scw_cw* cw, const char* path);
int scw_cw_create(scw_cw** cw);
int scw_cw_destroy(scw_cw* cw);
int scw_cw_set_incident(scw_cw* cw, const scw_planewave64* Pi, size_t n);
int scw_cw_add_layer(scw_cw* cw, double z, const scw_planewave64* Pr, size_t n_r, const scw_planewave64* Pt,
                     size_t n_t);
int scw_cw_counts(const scw_cw* cw, size_t* n_incident, size_t* n_layers, size_t* n_waves);
int scw_cw_layer(const scw_cw* cw, size_t layer, double* z, size_t* n_r, size_t* n_t);
/* cw_allocate + cw_unpack (cpuEvaluator.cpp:34-102) in FP64: waves[n_waves],
 * z_layers[n_layers], waves_begin/end[n_layers + 1]. */
int scw_cw_unpack(const scw_cw* cw, scw_planewave64* waves, double* z_layers, int32_t* waves_begin,
                  int32_t* waves_end);

#ifdef __cplusplus
}
#endif
#endif /* SCWATTER_AMD_H */
