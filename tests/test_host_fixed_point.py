"""The fixed-point conversion of the deterministic paints (fixed_point_q in
jaxpm_amd/csrc/common.hpp): adding 1.5 * 2^52 to an exactly scaled float32
deposit rounds it to the nearest integer, ties to even, and the sum's bit
pattern minus that of 1.5 * 2^52 is that integer, for every |value| < 2^51.
Checked here in float64 NumPy (the product by a power of two is exact, so a
separate multiply and add round as the kernel's single fma does), against
np.rint, including ties, negative values, values at the 2^51 limit and the
scale bound of fixed_point_scale."""
import numpy as np

MAGIC = 1.5 * 2.0 ** 52
MAGIC_BITS = np.float64(MAGIC).view(np.int64)


def fixed_q(v32, scale):
    t = v32.astype(np.float64) * scale + MAGIC
    return t.view(np.int64) - MAGIC_BITS


def scale_of(wmax, n):
    """fixed_point_scale: 2^min(62 - e, 51 - ew), n wmax < 2^e, wmax < 2^ew."""
    _, e = np.frexp(np.float64(wmax) * max(n, 1))
    _, ew = np.frexp(np.float64(wmax))
    return 2.0 ** min(62 - int(e), 51 - int(ew))


def test_magic_conversion_is_rint():
    r = np.random.default_rng(3)
    for n, wmax in [(1, 1.0), (100, 3.5), (2047, 1.0), (2048, 1.0), (256 ** 3, 1.0), (512 ** 3, 1.0),
                    (10 ** 6, 1e-3), (40000, 7.25e4)]:
        s = scale_of(wmax, n)
        assert wmax * s < 2.0 ** 51
        v = (r.uniform(-1, 1, 200000) * wmax).astype(np.float32)
        v = np.concatenate([v, np.float32([0.0, -0.0, wmax, -wmax, np.float32(wmax) * np.float32(0.999999)])])
        assert np.array_equal(fixed_q(v, s), np.rint(v.astype(np.float64) * s).astype(np.int64))


def test_ties_round_to_even():
    # values whose scaled form ends in .5 exactly
    v = np.float32([0.5, 1.5, 2.5, -0.5, -1.5, -2.5, 3.5, 1e-3 + 0.5])
    got = fixed_q(v, 1.0)
    want = np.rint(v.astype(np.float64)).astype(np.int64)
    assert np.array_equal(got, want)
    assert list(got[:7]) == [0, 2, 2, 0, -2, -2, 4]


def test_limit_of_the_range():
    # just inside 2^51 the conversion is exact; the scale keeps deposits there
    big = np.float64(2.0 ** 51 - 1.0)
    t = np.float64(big) + MAGIC
    assert t.view(np.int64) - MAGIC_BITS == int(big)
    t = np.float64(-big) + MAGIC
    assert t.view(np.int64) - MAGIC_BITS == -int(big)
    # every cell total n * wmax * scale stays below 2^62
    for n, wmax in [(1, 1.0), (2048, 2.0), (256 ** 3, 1.0), (1024 ** 3, 0.5)]:
        assert n * wmax * scale_of(wmax, n) < 2.0 ** 62


def test_planted_mistake_is_caught():
    """A magic constant of 2^52 (no half binade of headroom) breaks negative
    values: the check above would fail on it."""
    v = np.float32([-3.0, -0.75, 2.0])
    t = v.astype(np.float64) + 2.0 ** 52
    wrong = t.view(np.int64) - np.float64(2.0 ** 52).view(np.int64)
    assert not np.array_equal(wrong, np.rint(v.astype(np.float64)).astype(np.int64))
