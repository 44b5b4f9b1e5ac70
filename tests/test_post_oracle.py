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
