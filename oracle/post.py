"""CPU restatement of the viewer's post-processing (sCWatterView.cpp:194-352)
in numpy float32, one IEEE rounding per reference operation.  Test
infrastructure only: the library never imports it.

The reference multiplies std::complex<float> values with the textbook formula
(re = a c - b d, im = a d + b c, each product and sum rounded; no FMA on the
default x86-64 target), so the intensity of one component is
(a a - b (-b)) + (a (-b) + b a) i."""
import numpy as np

F = np.float32


def scalar(E, mode):
    """EvaluateScalarSlices (sCWatterView.cpp:299-352): E complex64 [n][3] -> complex64 [n]."""
    E = np.asarray(E, np.complex64).reshape(-1, 3)
    if mode < 3:
        return E[:, mode].copy()
    a, b = E.real.astype(F), E.imag.astype(F)
    with np.errstate(all="ignore"):
        pr = a * a - b * (-b)
        pi = a * (-b) + b * a
        re = (pr[:, 0] + pr[:, 1]) + pr[:, 2]
        im = (pi[:, 0] + pi[:, 1]) + pi[:, 2]
    out = np.empty(E.shape[0], np.complex64)
    out.real, out.imag = re, im
    return out


def minmax(slices):
    """CalculateMinMax (sCWatterView.cpp:264-296) over a list of complex64 arrays:
    (min Re, max Re, min Im, max Im), NaN entries skipped (NaN if none is left)."""
    re = np.concatenate([np.asarray(s, np.complex64).ravel().real for s in slices]).astype(F)
    im = np.concatenate([np.asarray(s, np.complex64).ravel().imag for s in slices]).astype(F)

    def mm(v):
        v = v[~np.isnan(v)]
        return (F(np.nan), F(np.nan)) if v.size == 0 else (v.min(), v.max())
    (a, b), (c, d) = mm(re), mm(im)
    return np.array([a, b, c, d], F)


def _clamp(v):  # sCWatterView.cpp:194-199
    v = np.where(np.isnan(v), F(0), v)
    return np.minimum(np.maximum(v, F(0)), F(1)).astype(F)


def color(S, lowhigh, real_color=(0.0, 1.0, 0.0), imag_color=(0.0, 0.0, 1.0), show_real=True, show_imag=True):
    """EvaluateColorValue / EvaluateColorSlices (sCWatterView.cpp:200-262): uint8 [n][3];
    values past 255 saturate (undefined in the reference's conversion)."""
    S = np.asarray(S, np.complex64).ravel()
    rl, rh, il, ih = (F(v) for v in lowhigh)
    with np.errstate(all="ignore"):
        ri, ii = F(rh - rl), F(ih - il)
        rn = _clamp((S.real.astype(F) - rl) / ri)
        inn = _clamp((S.imag.astype(F) - il) / ii)
        out = np.zeros((S.size, 3), np.uint8)
        for ch in range(3):
            c = np.zeros(S.size, F)
            if show_real:
                c = c + rn * F(real_color[ch])
            if show_imag:
                c = c + inn * F(imag_color[ch])
            v = c * F(255)
            v = np.where(np.isnan(v), F(0), np.minimum(np.maximum(v, F(0)), F(255)))
            out[:, ch] = v.astype(np.uint8)
    return out
