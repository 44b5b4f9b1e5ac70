"""On-device post-processing of evaluated fields: the viewer's scalar
selection, min/max range and two-colour map (sCWatterView.cpp:194-352),
through scw_post_* in include/scwatter_amd.h.  Buffers are device pointers;
calls are asynchronous on `stream` (None = the null stream)."""
import ctypes

from . import _native
from ._native import check

MODES = {"x": 0, "y": 1, "z": 2, "intensity": 3}       # DisplayMode (sCWatterView.cpp:59)
REAL_COLOR = (0.0, 1.0, 0.0)                            # defaults, sCWatterView.cpp:68
IMAG_COLOR = (0.0, 0.0, 1.0)                            # sCWatterView.cpp:75


def scalar(field_ptr, n, mode, scalar_ptr, stream=None):
    """S[i] = E[i][mode] or the intensity sum_c E_c conj(E_c); field complex64 [n][3], S complex64 [n]."""
    m = MODES[mode] if isinstance(mode, str) else int(mode)
    check(_native.lib().scw_post_scalar(field_ptr, n, m, scalar_ptr, stream))


def minmax(scalar_ptr, n, minmax_ptr, accumulate=False, stream=None):
    """(min Re, max Re, min Im, max Im) of S into a device float[4], optionally merged with its contents."""
    check(_native.lib().scw_post_minmax(scalar_ptr, n, minmax_ptr, 1 if accumulate else 0, stream))


def color(scalar_ptr, n, lowhigh_ptr, rgb_ptr, real_color=REAL_COLOR, imag_color=IMAG_COLOR, show_real=True,
          show_imag=True, stream=None):
    """uint8 [n][3] colours of S with the range (real_low, real_high, imag_low, imag_high) in device float[4]."""
    rc = (ctypes.c_float * 3)(*real_color)
    ic = (ctypes.c_float * 3)(*imag_color)
    check(_native.lib().scw_post_color(scalar_ptr, n, lowhigh_ptr, rc, ic, int(bool(show_real)),
                                       int(bool(show_imag)), rgb_ptr, stream))
