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
    (min Re, max Re, min Im, max Im).  The four values are seeded from
    slices[0][0] and updated with strict < and > (sCWatterView.cpp:271-289), so
    a NaN seed stays NaN and every other NaN is skipped; NaN if there are no
    entries.  A NaN result is the quiet NaN; among -0.0 and +0.0 the minimum is
    -0.0 (the reference keeps the first it met; those ties are left out of the
    tests)."""
    re = np.concatenate([np.asarray(s, np.complex64).ravel().real for s in slices]).astype(F)
    im = np.concatenate([np.asarray(s, np.complex64).ravel().imag for s in slices]).astype(F)

    def mm(v):
        if v.size == 0 or np.isnan(v[0]):
            return F(np.nan), F(np.nan)
        v = v[~np.isnan(v)]
        return v.min(), v.max()
    (a, b), (c, d) = mm(re), mm(im)
    return np.array([a, b, c, d], F)


def minmax_loop(slices):
    """The same as minmax, as the reference's loop itself: seed from
    slices[0][0], then for every index i the slices in order, each compared
    with > and <.  Pure Python; for small inputs (it pins minmax)."""
    first = np.asarray(slices[0], np.complex64).ravel()
    if first.size == 0:
        return np.array([np.nan] * 4, F)
    rmax = rmin = F(first[0].real)
    imax = imin = F(first[0].imag)
    n = first.size
    flat = [np.asarray(s, np.complex64).ravel() for s in slices]
    for i in range(n):
        for s in flat:
            r, m = F(s[i].real), F(s[i].imag)
            if r > rmax:
                rmax = r
            if r < rmin:
                rmin = r
            if m > imax:
                imax = m
            if m < imin:
                imin = m
    return np.array([rmin, rmax, imin, imax], F)


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
