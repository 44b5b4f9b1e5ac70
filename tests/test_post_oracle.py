"""Pins of the post-processing restatement (oracle/post.py) against values
worked out by hand from sCWatterView.cpp:194-352."""
import numpy as np

from oracle import post


def test_intensity_and_components():
    E = np.array([[3 + 4j, 1 - 2j, 0j], [0.5j, -2, 1 + 1j]], np.complex64)
    S = post.scalar(E, 3)
    assert S.dtype == np.complex64
    assert np.array_equal(S, np.array([30, 0.25 + 4 + 2], np.complex64))  # |3+4i|^2 + |1-2i|^2, ...
    assert np.all(S.imag == 0)
    for m in range(3):
        assert np.array_equal(post.scalar(E, m), E[:, m])


def test_minmax_skips_nan_and_spans_slices():
    a = np.array([1 + 2j, np.nan + 0j, -3 + 5j], np.complex64)
    b = np.array([7 - 1j, complex(0, np.nan)], np.complex64)
    assert np.array_equal(post.minmax([a, b]), np.array([-3, 7, -1, 5], np.float32))
    assert np.all(np.isnan(post.minmax([np.array([complex(np.nan, np.nan)], np.complex64)])))


def test_minmax_nan_seed_sticks():
    """CalculateMinMax seeds from S_xy[0]: a NaN there makes that part NaN for
    good (every later comparison is false), the other part is unaffected."""
    a = np.array([complex(np.nan, 2), 5 - 1j, -3 + 5j], np.complex64)
    r = post.minmax([a, np.array([9 + 9j, 0j, 1j], np.complex64)])
    assert np.isnan(r[0]) and np.isnan(r[1]) and r[2] == -1 and r[3] == 9


def test_minmax_matches_the_reference_loop():
    """The vectorised restatement against the reference's own loop (seed, then
    strict comparisons over the interleaved slices) on small random slices
    with NaN and infinities, NaN at index 0 or not; zeros of both signs are
    left out (the reference keeps the first of -0.0 / +0.0 it meets)."""
    rng = np.random.default_rng(3)
    for trial in range(200):
        n = int(rng.integers(1, 9))
        sl = []
        for k in range(3):
            v = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
            re, im = v.real.copy(), v.imag.copy()
            for part in (re, im):
                m = rng.random(n)
                part[m < 0.2] = np.nan
                part[(m >= 0.2) & (m < 0.25)] = np.inf
                part[(m >= 0.25) & (m < 0.3)] = -np.inf
            sl.append((re + 1j * im).astype(np.complex64))
        if trial % 4 == 0:
            sl[0][0] = complex(np.nan, sl[0][0].imag)
        got, want = post.minmax(sl), post.minmax_loop(sl)
        assert np.array_equal(np.isnan(got), np.isnan(want)), (sl, got, want)
        assert np.array_equal(got[~np.isnan(got)], want[~np.isnan(want)]), (sl, got, want)


def test_color_hand_values():
    S = np.array([0.5 + 0.25j, 2 - 1j, complex(np.nan, 0.5), 0j], np.complex64)
    rgb = post.color(S, (0, 1, 0, 1))
    # green = floor(255 rn), blue = floor(255 in), red 0 with the default colours
    assert rgb.tolist() == [[0, 127, 63], [0, 255, 0], [0, 0, 127], [0, 0, 0]]
    assert post.color(S, (0, 1, 0, 1), show_imag=False)[:, 2].tolist() == [0, 0, 0, 0]
    # equal low and high: 0/0 is NaN -> 0, x/0 is inf -> 1
    assert post.color(np.array([1 + 0j, 2 + 0j], np.complex64), (1, 1, 0, 1))[:, 1].tolist() == [0, 255]
    # both colours on one channel add up and saturate at 255
    assert post.color(np.array([1 + 1j], np.complex64), (0, 1, 0, 1), (1, 0, 0), (1, 0, 0))[0, 0] == 255
