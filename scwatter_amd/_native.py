"""ctypes bindings of libscwatter_amd.so (the C ABI in include/scwatter_amd.h).

The library is loaded from this package directory only; if it is missing the
import fails loudly — there is no CPU fallback.
"""
import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libscwatter_amd.so")

SCW_OK = 0
FORMS = {"auto": 0, "direct": 1, "gemm": 2}
FORM_NAMES = {v: k for k, v in FORMS.items()}
DTYPES = {"fp32": 0, "bf16": 1, "bf16x3": 2, "fp16": 3}
DTYPE_NAMES = {v: k for k, v in DTYPES.items()}
STATUS = {0: "SCW_OK", -1: "SCW_ERR_ARG", -2: "SCW_ERR_NOMEM", -3: "SCW_ERR_DEVICE", -4: "SCW_ERR_IO",
          -5: "SCW_ERR_FORMAT", -6: "SCW_ERR_UNSUPPORTED", -7: "SCW_ERR_TIMEOUT"}

# Every extern "C" symbol declared in include/scwatter_amd.h.
EXPORTS = (
    "scw_last_error", "scw_abi_version", "scw_device_count", "scw_device_mem_info",
    "scw_plan_create", "scw_plan_destroy", "scw_plan_layers", "scw_plan_region_info", "scw_plan_device_bytes",
    "scw_exec_create", "scw_exec_run", "scw_exec_run_stage", "scw_exec_run_host", "scw_exec_destroy", "scw_exec_info",
    "scw_exec_stage_info", "scw_batch_exec_create", "scw_exec_set_launch", "scw_exec_set_tiles",
    "scw_exec_graph_capture", "scw_exec_graph_launch", "scw_exec_set_persistent", "scw_exec_persistent_info",
    "scw_exec_check", "scw_test_exec_stall", "scw_test_plan_host", "scw_test_xcd_order",
    "scw_cw_load", "scw_cw_save", "scw_cw_create", "scw_cw_destroy", "scw_cw_set_incident", "scw_cw_add_layer",
    "scw_cw_counts", "scw_cw_layer", "scw_cw_unpack",
    "scw_post_scalar", "scw_post_minmax", "scw_post_color", "scw_test_fail_alloc_after", "scw_eval_planes",
    "scw_eval_planes_noalloc",
)


class ScwError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


_lib = None
_lib_lock = threading.Lock()
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_szp = ctypes.POINTER(ctypes.c_size_t)
_intp = ctypes.POINTER(ctypes.c_int)


def lib():
    """Load libscwatter_amd.so (raises OSError if it has not been built).
    Thread-safe: plans are created from thread pools, and the first caller
    loads the library while the others wait."""
    if _lib is not None:
        return _lib
    with _lib_lock:
        return _load()


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} is missing: build it with `make -C scwatter_amd` "
                      "or __graft_entry__.build(); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    sig = {
        "scw_last_error": (ctypes.c_char_p, []),
        "scw_abi_version": (ctypes.c_int, []),
        "scw_device_count": (ctypes.c_int, [_intp]),
        "scw_device_mem_info": (ctypes.c_int, [ctypes.c_int, _szp, _szp]),
        "scw_plan_create": (ctypes.c_int, [ctypes.c_int, _vp, ctypes.c_size_t, _dp, ctypes.c_int, _ip, _ip,
                                           ctypes.POINTER(_vp)]),
        "scw_plan_destroy": (ctypes.c_int, [_vp]),
        "scw_plan_layers": (ctypes.c_int, [_vp, _intp]),
        "scw_plan_region_info": (ctypes.c_int, [_vp, ctypes.c_int, _i64p, _ip, _ip, _ip, _i64p, _ip, _ip]),
        "scw_plan_device_bytes": (ctypes.c_int, [_vp, _szp]),
        "scw_exec_create": (ctypes.c_int, [_vp, _dp, ctypes.c_size_t, _dp, ctypes.c_size_t, _dp, ctypes.c_size_t,
                                           ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
        "scw_exec_run": (ctypes.c_int, [_vp, _vp, _vp]),
        "scw_exec_run_stage": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp]),
        "scw_exec_run_host": (ctypes.c_int, [_vp, _vp, _vp]),
        "scw_exec_destroy": (ctypes.c_int, [_vp]),
        "scw_exec_info": (ctypes.c_int, [_vp, _intp, _intp, _dp, _dp, _dp, _intp]),
        "scw_exec_stage_info": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp]),
        "scw_batch_exec_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), _dp, ctypes.c_size_t, _dp,
                                                 ctypes.c_size_t, ctypes.POINTER(_dp), _szp, ctypes.c_double,
                                                 ctypes.c_int, ctypes.POINTER(_vp)]),
        "scw_exec_set_launch": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int64]),
        "scw_exec_set_tiles": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
        "scw_exec_graph_capture": (ctypes.c_int, [_vp, _vp, _vp]),
        "scw_exec_graph_launch": (ctypes.c_int, [_vp, _vp]),
        "scw_exec_set_persistent": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int64, ctypes.c_double]),
        "scw_exec_persistent_info": (ctypes.c_int, [_vp, _intp, _i64p, _i64p]),
        "scw_exec_check": (ctypes.c_int, [_vp, _vp]),
        "scw_test_exec_stall": (ctypes.c_int, [_vp, ctypes.c_int]),
        "scw_test_plan_host": (ctypes.c_int, [_vp, ctypes.c_size_t, _dp, ctypes.c_int, _ip, _ip, _i64p]),
        "scw_test_xcd_order": (ctypes.c_int, [ctypes.c_int, _i64p, _i64p, ctypes.c_int, _i64p, ctypes.c_int64, _i64p]),
        "scw_cw_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_vp)]),
        "scw_cw_save": (ctypes.c_int, [_vp, ctypes.c_char_p]),
        "scw_cw_create": (ctypes.c_int, [ctypes.POINTER(_vp)]),
        "scw_cw_destroy": (ctypes.c_int, [_vp]),
        "scw_cw_set_incident": (ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
        "scw_cw_add_layer": (ctypes.c_int, [_vp, ctypes.c_double, _vp, ctypes.c_size_t, _vp, ctypes.c_size_t]),
        "scw_cw_counts": (ctypes.c_int, [_vp, _szp, _szp, _szp]),
        "scw_cw_layer": (ctypes.c_int, [_vp, ctypes.c_size_t, _dp, _szp, _szp]),
        "scw_cw_unpack": (ctypes.c_int, [_vp, _vp, _dp, _ip, _ip]),
        "scw_test_fail_alloc_after": (None, [ctypes.c_int64]),
        "scw_eval_planes": (ctypes.c_int, [_vp] + [ctypes.c_float] * 7 + [ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                                                          ctypes.c_int, _vp, _vp, _vp, _vp]),
        "scw_eval_planes_noalloc": (ctypes.c_int, [_vp] + [ctypes.c_float] * 7 + [ctypes.c_int, ctypes.c_double, _vp,
                                                                                  _vp, _vp, _vp]),
        "scw_post_scalar": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.c_int, _vp, _vp]),
        "scw_post_minmax": (ctypes.c_int, [_vp, ctypes.c_size_t, _vp, ctypes.c_int, _vp]),
        "scw_post_color": (ctypes.c_int, [_vp, ctypes.c_size_t, _vp, ctypes.POINTER(ctypes.c_float),
                                          ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.c_int, _vp, _vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(status):
    if status != SCW_OK:
        msg = lib().scw_last_error()
        raise ScwError(status, msg.decode() if msg else "")
    return status
