"""The oracle's glare filter (rt/camera.cc:145-462, 1504-1630) pinned by
independent computations:

* the random generator (common/common.cc:30-105) against numpy's MT19937:
  the reference's refill loop stops at index 622, so every 624th word is the
  tempered initial mt[623], and each later block is numpy's standard block of
  the state with that stale word put back;
* the filter's normalisation (1e-4 of the frame's pixel count per channel),
  its grey corona and bloom at the photopic level (no halo there: every
  channel equal), and the light levels' thresholds on log10 of the adapted
  luminance;
* the convolution against numpy's FFT of the same shifted filter, and against
  the closed form of one bright pixel (the filter, moved to that pixel, scaled
  by its value / 255), which fixes the direction of the shift;
* the tone-map hook: glare after the mapping equals the mapping then glare at
  the adapted luminance it reports.
"""
import numpy as np
import pytest


def temper(y):
    y = np.asarray(y, np.uint32).copy()
    y ^= y >> np.uint32(11)
    y ^= (y << np.uint32(7)) & np.uint32(0x9D2C5680)
    y ^= (y << np.uint32(15)) & np.uint32(0xEFC60000)
    y ^= y >> np.uint32(18)
    return y


def numpy_quirk_stream(count):
    """The reference's stream from numpy's generator: block b is numpy's first
    623 outputs of a state whose word 623 is the initial one, then that word
    tempered."""
    rs = np.random.RandomState(5489)
    init = np.array(rs.get_state()[1], np.uint32)
    stale = init[623]
    out = []
    while len(out) < count:
        block = rs._bit_generator.random_raw(624).astype(np.uint32)
        out.extend(block[:623].tolist())
        out.append(int(temper(np.array([stale], np.uint32))[0]))
        key = np.array(rs.get_state()[1], np.uint32)
        key[623] = stale
        rs.set_state(("MT19937", key, 624))
    return np.array(out[:count], np.uint32)


def test_temper_matches_numpy():
    rs = np.random.RandomState(5489)
    first = rs._bit_generator.random_raw(624).astype(np.uint32)
    key = np.array(rs.get_state()[1], np.uint32)  # the refilled state the first block came from
    assert np.array_equal(temper(key), first)


def test_generator_against_numpy(O):
    got = O.mt_stream(3 * 624 + 5)
    assert np.array_equal(got[:623], np.random.RandomState(5489)._bit_generator.random_raw(623).astype(np.uint32))
    assert np.array_equal(got, numpy_quirk_stream(got.size))
    # The stale word is not what a standard generator gives there.
    assert got[623] != np.random.RandomState(5489)._bit_generator.random_raw(624)[623]


@pytest.mark.parametrize("xr,yr,level", [(64, 48, 0), (63, 47, 1), (80, 45, 2), (17, 33, 1)])
def test_filter_normalised(O, xr, yr, level):
    gf = O.glare_filter(xr, yr, 0.9, level)
    assert np.all(np.isfinite(gf)) and np.all(gf >= 0)
    total = gf.reshape(-1, 3).astype(np.float64).sum(0)
    np.testing.assert_allclose(total, 1e-4 * xr * yr, rtol=2e-5)


def test_filter_scale_uses_the_float_running_sum(O):
    """At 1080p the float running sum's rounding bias shows: the double total
    of the filter is short of 1e-4 per pixel by about 6e-4, which a double
    normalisation would not give."""
    gf = O.glare_filter(1920, 1080, 1.2, 1)
    total = gf.reshape(-1, 3).astype(np.float64).sum(0)
    rel = total / (1e-4 * 1920 * 1080) - 1.0
    assert np.all(rel < -3e-4) and np.all(rel > -1e-3), rel


def test_photopic_filter_is_grey_and_the_halo_needs_a_wide_view(O):
    gp = O.glare_filter(64, 48, 0.9, 2)
    assert np.array_equal(gp[..., 0], gp[..., 1]) and np.array_equal(gp[..., 1], gp[..., 2])
    # The spreading functions get the angle in radians (pixdeg is atan(..) * 2
    # / x_res) where Spencer's take degrees, so the halo's ring at about 3.2
    # lies outside an ordinary field of view and underflows to 0: grey filters.
    for level in (0, 1):
        g = O.glare_filter(64, 48, 0.9, level)
        assert np.array_equal(g[..., 0], g[..., 2])
    # A tall frame with a field of view near 180 degrees reaches the ring,
    # and its colours (611, 549, 467 nm) part the channels.
    for level in (0, 1):
        g = O.glare_filter(32, 64, 20.0, level)
        assert np.abs(g[..., 0] - g[..., 2]).max() > 1e-5 * g[..., 1].max()


def test_filter_depends_on_field_of_view_and_level(O):
    xr, yr = 64, 48
    a, b = O.glare_filter(xr, yr, 0.5, 1), O.glare_filter(xr, yr, 1.5, 1)
    assert not np.allclose(a, b)
    assert not np.allclose(O.glare_filter(xr, yr, 0.9, 0), O.glare_filter(xr, yr, 0.9, 2))


def shifted_filter(gf):
    """glare_in of rt/camera.cc:1571-1587 in the image's linear pixel order."""
    yr, xr = gf.shape[:2]
    n = xr * yr
    off = (xr >> 1) + xr * ((yr - 1) >> 1)
    flat = gf.reshape(n, 3)
    i = np.arange(n)
    return flat[np.where(i < off, i + off, i - off)].reshape(yr, xr, 3)


def numpy_glare_increment(img, gf):
    """The glare the reference adds, from numpy's FFTs in double."""
    yr, xr = img.shape[:2]
    bright = (img > 255.0).any(-1, keepdims=True)
    src = np.where(bright, img, 0.0).astype(np.float64)
    gin = shifted_filter(gf).astype(np.float64)
    conv = np.fft.ifft2(np.fft.fft2(src, axes=(0, 1)) * np.fft.fft2(gin, axes=(0, 1)), axes=(0, 1)).real
    return conv / 255.0


def hdr_image(xr, yr, seed, nbright=20, scale=100.0):
    rng = np.random.default_rng(seed)
    img = (rng.random((yr, xr, 3)) * scale).astype(np.float32)
    idx = rng.choice(xr * yr, nbright, replace=False)
    img.reshape(-1, 3)[idx] = (rng.random((nbright, 3)) * 4000.0 + 256.0).astype(np.float32)
    return img


@pytest.mark.parametrize("xr,yr,yi", [(48, 32, 0.005), (47, 31, 1.0), (40, 27, 50.0)])
def test_convolution_against_numpy_fft(O, xr, yr, yi):
    img = hdr_image(xr, yr, seed=xr)
    out, level = O.glare(img, yi, 0.9)
    assert level == (0 if yi < 0.01 else (1 if yi < 10 ** 0.6 else 2))
    gf = O.glare_filter(xr, yr, 0.9, level)
    want = numpy_glare_increment(img, gf)
    got = out.astype(np.float64) - img
    # float32 rounding of the sums onto values up to about 4000.
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=np.float64(np.spacing(np.float32(4300.0))))
    assert np.abs(want).max() > 1e-3


def test_one_bright_pixel_spreads_the_moved_filter(O):
    xr, yr = 36, 24
    img = np.zeros((yr, xr, 3), np.float32)
    qy, qx, v = 5, 30, np.float32(1000.0)
    img[qy, qx] = v
    out, level = O.glare(img, 1.0, 0.9)
    gin = shifted_filter(O.glare_filter(xr, yr, 0.9, level)).astype(np.float64)
    # Pixel (py, px) gets v * glare_in[(py - qy) mod yr, (px - qx) mod xr] / 255.
    want = v * np.roll(gin, (qy, qx), axis=(0, 1)) / 255.0
    want[qy, qx] += v
    np.testing.assert_allclose(out, want, rtol=1e-5, atol=1e-9)


def test_no_bright_pixel_no_glare(O):
    img = hdr_image(30, 20, seed=3, nbright=0, scale=255.0)
    out, _ = O.glare(img, 1.0, 0.9)
    assert np.array_equal(out, img)


@pytest.mark.parametrize("yi,level", [(0.0099, 0), (0.01, 1), (3.9, 1), (3.99, 2)])
def test_light_level_thresholds(O, yi, level):
    img = hdr_image(16, 12, seed=1)
    assert O.glare(img, yi, 0.9)[1] == level


def test_tone_map_then_glare(O):
    img = hdr_image(40, 30, seed=9, scale=300.0)
    plain, _, yi = O.tone_map(img, mode=0)
    both, _, yi2 = O.tone_map(img, mode=0, glare=True, x_angle=0.9)
    assert yi == yi2
    want, _ = O.glare(plain, yi, 0.9)
    assert np.array_equal(both, want)
    with pytest.raises(ValueError):
        O.tone_map(img, mode=0, glare=True)
