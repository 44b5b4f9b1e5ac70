"""Host-side evaluator API over the C ABI.

Plan  ~ gpu_initialize (gpuEvaluator.cu:15-26): device-resident wave tables.
Exec  ~ one grid of gpu_cw_evaluate / kernel_{xy,xz,yz} (gpuEvaluator.cu:70-165),
        generalised to rectilinear 3D grids with independent axes.
gpu_cw_evaluate / cpu_cw_evaluate_{xy,yz,xz} keep the reference's argument
meaning, coordinate rounding and region rules (see each docstring); every
evaluation runs on the GPU.
"""
import ctypes

import numpy as np

from . import _native
from .cw import CoupledWaveStructure, WaveTables
from ._native import check

_vp = ctypes.c_void_p


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Plan:
    """Wave tables of one structure, planned and resident on `device`."""

    def __init__(self, tables: WaveTables, device=0):
        L = _native.lib()
        self.tables = tables
        rec = np.ascontiguousarray(tables.records())
        zl = np.ascontiguousarray(tables.z_layers, np.float64)
        b = np.ascontiguousarray(tables.begin, np.int32)
        e = np.ascontiguousarray(tables.end, np.int32)
        h = _vp()
        check(L.scw_plan_create(int(device), rec.ctypes.data if rec.size else None, rec.shape[0], _dptr(zl),
                                zl.shape[0], b.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                e.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.byref(h)))
        self.handle = h.value
        self.device = device
        self.layers = zl.shape[0]

    @classmethod
    def from_cw(cls, cw: CoupledWaveStructure, device=0):
        return cls(cw.unpack(), device)

    def region_info(self, r):
        n = ctypes.c_int64()
        mx, my, mc, width, sep = (ctypes.c_int32() for _ in range(5))
        nt = ctypes.c_int64()
        check(_native.lib().scw_plan_region_info(self.handle, int(r), ctypes.byref(n), ctypes.byref(mx),
                                                 ctypes.byref(my), ctypes.byref(mc), ctypes.byref(nt),
                                                 ctypes.byref(width), ctypes.byref(sep)))
        return dict(n_waves=n.value, mx=mx.value, my=my.value, mc=mc.value, n_terms=nt.value, width=width.value,
                    separable=bool(sep.value))

    def device_bytes(self):
        b = ctypes.c_size_t()
        check(_native.lib().scw_plan_device_bytes(self.handle, ctypes.byref(b)))
        return b.value

    def destroy(self):
        if getattr(self, "handle", None):
            _native.lib().scw_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Exec:
    """A plan bound to the grid xs x ys x zs; output complex64 (nz, ny, nx, 3)."""

    def __init__(self, plan: Plan, xs, ys, zs, region_tol=0.0, form="auto", dtype="fp32"):
        self.plan = plan
        self.xs = np.ascontiguousarray(np.atleast_1d(xs), np.float64)
        self.ys = np.ascontiguousarray(np.atleast_1d(ys), np.float64)
        self.zs = np.ascontiguousarray(np.atleast_1d(zs), np.float64)
        h = _vp()
        check(_native.lib().scw_exec_create(plan.handle, _dptr(self.xs), self.xs.size, _dptr(self.ys), self.ys.size,
                                            _dptr(self.zs), self.zs.size, float(region_tol), _native.FORMS[form],
                                            _native.DTYPES[dtype], ctypes.byref(h)))
        self.handle = h.value
        self.shape = (self.zs.size, self.ys.size, self.xs.size, 3)

    def info(self):
        f, d, l = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        u, fl, by = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(_native.lib().scw_exec_info(self.handle, ctypes.byref(f), ctypes.byref(d), ctypes.byref(u),
                                          ctypes.byref(fl), ctypes.byref(by), ctypes.byref(l)))
        return dict(form=_native.FORM_NAMES[f.value], dtype=_native.DTYPE_NAMES[d.value], units=u.value,
                    flops=fl.value, bytes=by.value, launches=l.value)

    def stage_info(self, stage):
        """Algorithmic (flops, bytes) of one stage (0 = whole run)."""
        f, b = ctypes.c_double(), ctypes.c_double()
        check(_native.lib().scw_exec_stage_info(self.handle, int(stage), ctypes.byref(f), ctypes.byref(b)))
        return f.value, b.value

    @property
    def out_bytes(self):
        return int(np.prod(self.shape)) * 8

    def run(self, out_ptr, stream=None):
        """Launch into device memory out_ptr (asynchronous, on `stream`)."""
        check(_native.lib().scw_exec_run(self.handle, out_ptr, stream))

    def run_stage(self, stage, out_ptr, stream=None):
        """Launch one stage (1 coefficients, 2 X contraction, 3 main GEMM; 0 all)."""
        check(_native.lib().scw_exec_run_stage(self.handle, int(stage), out_ptr, stream))

    def run_host(self, stream=None):
        out = np.empty(self.shape, np.complex64)
        check(_native.lib().scw_exec_run_host(self.handle, out.ctypes.data, stream))
        return out

    def set_launch(self, coef_items=0, main_blocks=0):
        """GEMM-form launch shape (0 = default); rebuilds the work lists."""
        check(_native.lib().scw_exec_set_launch(self.handle, int(coef_items), int(main_blocks)))

    def set_tiles(self, modal_run=0, coef_run=0):
        """GEMM-form z-run tiles of the coefficient contraction (0 = default 64 / 8)."""
        check(_native.lib().scw_exec_set_tiles(self.handle, int(modal_run), int(coef_run)))

    def set_persistent(self, on=True, workgroups=0, wait_ms=0.0):
        """Single-launch form (GEMM form, 16-bit dtypes): every run is one
        persistent kernel; workgroups 0 = resident capacity, wait_ms 0 = 10 s."""
        check(_native.lib().scw_exec_set_persistent(self.handle, int(bool(on)), int(workgroups), float(wait_ms)))

    def persistent_info(self):
        on, units, wg = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64()
        check(_native.lib().scw_exec_persistent_info(self.handle, ctypes.byref(on), ctypes.byref(units),
                                                     ctypes.byref(wg)))
        return dict(on=bool(on.value), units=units.value, workgroups=wg.value)

    def check(self, stream=None):
        """Wait for `stream`; raise ScwError(SCW_ERR_TIMEOUT) if a wait of a
        single-launch run expired since the last check."""
        check(_native.lib().scw_exec_check(self.handle, stream))

    def graph_capture(self, out_ptr, stream=None):
        check(_native.lib().scw_exec_graph_capture(self.handle, out_ptr, stream))

    def graph_launch(self, stream=None):
        check(_native.lib().scw_exec_graph_launch(self.handle, stream))

    def destroy(self):
        if getattr(self, "handle", None):
            _native.lib().scw_exec_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class BatchExec(Exec):
    """Independent structures in one GEMM-form run: instance i is plans[i] on
    xs x ys x zs_list[i]; output complex64 (sum nz_i, ny, nx, 3), instances
    concatenated along z."""

    def __init__(self, plans, xs, ys, zs_list, region_tol=0.0, dtype="bf16x3"):
        self.plans = list(plans)
        self.xs = np.ascontiguousarray(np.atleast_1d(xs), np.float64)
        self.ys = np.ascontiguousarray(np.atleast_1d(ys), np.float64)
        self.zs_list = [np.ascontiguousarray(np.atleast_1d(z), np.float64) for z in zs_list]
        n = len(self.plans)
        handles = (_vp * n)(*[p.handle for p in self.plans])
        zptrs = (ctypes.POINTER(ctypes.c_double) * n)(*[_dptr(z) for z in self.zs_list])
        nzs = (ctypes.c_size_t * n)(*[z.size for z in self.zs_list])
        h = _vp()
        check(_native.lib().scw_batch_exec_create(n, handles, _dptr(self.xs), self.xs.size, _dptr(self.ys),
                                                  self.ys.size, zptrs, nzs, float(region_tol),
                                                  _native.DTYPES[dtype], ctypes.byref(h)))
        self.handle = h.value
        self.zs = np.concatenate(self.zs_list)
        self.shape = (self.zs.size, self.ys.size, self.xs.size, 3)
        self.offsets = np.concatenate([[0], np.cumsum([z.size for z in self.zs_list])])

    def instance(self, E, i):
        """Slice instance i's field out of a batch output."""
        return E[self.offsets[i]:self.offsets[i + 1]]


# ---- coordinate rules of the reference ---------------------------------------

def axis_cpu(start, d, N):
    """cpuEvaluator.cpp:317-319: x = x_start + ix * d in float (product rounded, then sum)."""
    i = np.arange(N, dtype=np.float32)
    return (np.float32(start) + i * np.float32(d)).astype(np.float64)


def axis_gpu(start, d, N):
    """gpuEvaluator.cu:78-79: x = x_start + d * xi, fused into one multiply-add (one rounding)."""
    i = np.arange(N, dtype=np.float64)
    return (np.float64(np.float32(d)) * i + np.float64(np.float32(start))).astype(np.float32).astype(np.float64)


def eval_planes(plan: Plan, x_start, y_start, z_start, x, y, z, d, N, out_xy, out_xz, out_yz, region_tol=0.001,
                form="auto", dtype="fp32", stream=None):
    """scw_eval_planes: the three N x N planes of gpu_cw_evaluate into device
    memory (out_* device pointers, 0/None skips a plane), asynchronously."""
    f = lambda v: float(np.float32(v))
    check(_native.lib().scw_eval_planes(plan.handle, f(x_start), f(y_start), f(z_start), f(x), f(y), f(z), f(d),
                                        int(N), float(region_tol), _native.FORMS[form], _native.DTYPES[dtype],
                                        out_xy or None, out_xz or None, out_yz or None, stream))


def eval_planes_noalloc(plan: Plan, x_start, y_start, z_start, x, y, z, d, N, out_xy, out_xz, out_yz,
                        region_tol=0.001, stream=None):
    """scw_eval_planes_noalloc: eval_planes' direct form with no device
    allocation (tables and partial sums in the library's fixed device array;
    N <= 16384)."""
    f = lambda v: float(np.float32(v))
    check(_native.lib().scw_eval_planes_noalloc(plan.handle, f(x_start), f(y_start), f(z_start), f(x), f(y), f(z),
                                                f(d), int(N), float(region_tol), out_xy or None, out_xz or None,
                                                out_yz or None, stream))


def gpu_cw_evaluate(plan: Plan, x_start, y_start, z_start, x, y, z, d, N, form="auto", dtype="fp32"):
    """gpu_cw_evaluate (gpuEvaluator.cu:131-165): xy plane at z, xz plane at y, yz plane at x.

    Region rule z >= z_l - 0.001 (gpuEvaluator.cu:34), coordinates start + d i
    fused (gpuEvaluator.cu:78-79); through scw_eval_planes.  Returns three
    complex64 (N, N, 3) arrays: E_xy[iy, ix], E_xz[iz, ix], E_yz[iz, iy].
    """
    from . import hip
    bufs = [hip.DeviceArray((N, N, 3), np.complex64) for _ in range(3)]
    eval_planes(plan, x_start, y_start, z_start, x, y, z, d, N, *(b.ptr for b in bufs), form=form, dtype=dtype)
    out = tuple(b.to_host() for b in bufs)
    for b in bufs:
        b.free()
    return out


def cpu_cw_evaluate(kind, plan: Plan, a_start, b_start, fixed, d, N, form="auto", dtype="fp32"):
    """cpu_cw_evaluate_{xy,yz,xz} (cpuEvaluator.cpp:294-407) semantics on the GPU.

    kind 'xy': (x_start, y_start, z); 'yz': (y_start, z_start, x); 'xz': (x_start, z_start, y).
    Region rule z >= z_l (cpuEvaluator.cpp:303).  Returns complex64 (N, N, 3).
    """
    a, b = axis_cpu(a_start, d, N), axis_cpu(b_start, d, N)
    f = [float(np.float32(fixed))]
    grids = {"xy": (a, b, f), "yz": (f, a, b), "xz": (a, f, b)}
    ex = Exec(plan, *grids[kind], 0.0, form, dtype)
    out = ex.run_host().reshape(N, N, 3)
    ex.destroy()
    return out
