"""The GPU glare filter (camera::tone_map with gf, rt/camera.cc:145-462,
1455-1458, 1504-1630) through the C ABI against the oracle.

* The filter (rt_glare_filter) agrees with the oracle's within 1e-5 of each
  value plus 1e-7 of the largest: the lines, the corona's sum (the bloom's
  level) and the normalising sums are the same float operations in the same
  order; the spreading functions' expf may differ from the host's by an ulp.
* The glare a tone map adds agrees with the oracle's (a direct double sum of
  the circular convolution) within 1e-4 of the largest increment plus one
  float spacing of the pixel it is added to: the GPU convolves through
  zero-padded single-precision FFTs, as the reference does through fftw3f.
* At 1920 x 1080 the increment agrees with numpy's double FFTs of the GPU's
  own filter, and the filter with the oracle's.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _filter_close(g, o):
    err = np.abs(g.astype(np.float64) - o)
    tol = 1e-5 * np.abs(o) + 1e-7 * np.abs(o).max()
    return np.all(err <= tol), (err.max(), np.abs(o).max(), (err / np.maximum(np.abs(o), 1e-30)).max())


@pytest.mark.parametrize("xr,yr,x_angle,level", [
    (64, 48, 0.9, 0), (64, 48, 0.9, 1), (64, 48, 0.9, 2), (17, 33, 0.9, 1), (320, 180, 1.2, 1),
    (32, 64, 20.0, 0), (32, 64, 20.0, 1), (1, 1, 0.9, 1), (2, 1, 0.9, 2), (5, 3, 0.5, 0)])
def test_filter_matches_oracle(gpu_ctx, rt, O, xr, yr, x_angle, level):
    g = rt.glare_filter(gpu_ctx, xr, yr, x_angle, level)
    o = O.glare_filter(xr, yr, x_angle, level)
    ok, info = _filter_close(g, o)
    assert ok, info
    # 1e-4 per pixel as the reference's float running sum counts it; in
    # double the total is short by that sum's rounding bias (6e-4 at 1080p).
    np.testing.assert_allclose(g.reshape(-1, 3).astype(np.float64).sum(0), 1e-4 * xr * yr, rtol=1e-3)


def test_filter_cache_follows_size_angle_and_level(gpu_ctx, rt, O):
    a = rt.glare_filter(gpu_ctx, 64, 48, 0.9, 1)
    b = rt.glare_filter(gpu_ctx, 80, 45, 0.9, 1)      # new size: the cache is rebuilt
    c = rt.glare_filter(gpu_ctx, 64, 48, 0.7, 1)      # new angle at the old size
    d = rt.glare_filter(gpu_ctx, 64, 48, 0.9, 1)
    assert np.array_equal(a, d)
    assert _filter_close(b, O.glare_filter(80, 45, 0.9, 1))[0]
    assert _filter_close(c, O.glare_filter(64, 48, 0.7, 1))[0]


def hdr_image(xr, yr, seed, scale, nbright=20):
    """Values in [0, scale) (scale at most 250) and nbright pixels of 300-4300."""
    rng = np.random.default_rng(seed)
    img = (rng.random((yr, xr, 3)) * scale).astype(np.float32)
    idx = rng.choice(xr * yr, nbright, replace=False)
    img.reshape(-1, 3)[idx] = (rng.random((nbright, 3)) * 4000.0 + 300.0).astype(np.float32)
    return img


def _increment_close(g_plain, g_glare, o_plain, o_glare):
    dg = g_glare.astype(np.float64) - g_plain
    do = o_glare.astype(np.float64) - o_plain
    tol = 1e-4 * np.abs(do).max() + np.spacing(np.abs(o_plain).astype(np.float32)).astype(np.float64) + 1e-9
    err = np.abs(dg - do)
    return np.all(err <= tol), (err.max(), np.abs(do).max())


def light_level(yi):
    return 0 if np.log10(np.float32(yi)) < -2.0 else (1 if np.log10(np.float32(yi)) < 0.6 else 2)


# Non-bright scales that put the log-average luminance (683 * Y / 255) in
# each light level: scotopic, mesopic, photopic.
@pytest.mark.parametrize("scale,level", [(0.002, 0), (1.0, 1), (200.0, 2)])
@pytest.mark.parametrize("xr,yr", [(96, 64), (33, 17), (160, 90)])
def test_tone_map_glare_matches_oracle(gpu_ctx, rt, O, xr, yr, scale, level):
    img = hdr_image(xr, yr, seed=xr + yr, scale=scale)
    gp, _, gyi, _ = rt.tone_map(gpu_ctx, img, mode=0)
    gg, _, gyi2, _ = rt.tone_map(gpu_ctx, img, mode=0, glare=True, x_angle=0.9)
    op, _, oyi = O.tone_map(img, mode=0)
    og, _, _ = O.tone_map(img, mode=0, glare=True, x_angle=0.9)
    assert gyi == gyi2 and light_level(gyi) == light_level(oyi) == level
    ok, info = _increment_close(gp, gg, op, og)
    assert ok, info
    assert info[1] > 0.0


def test_glare_on_device_tensor_matches_host_path(gpu_ctx, rt):
    import torch

    img = hdr_image(128, 72, seed=5, scale=50.0, nbright=40)
    host, _, _, _ = rt.tone_map(gpu_ctx, img, mode=4, glare=True, x_angle=1.1)
    dev = torch.from_numpy(img).to("cuda:0")
    out, _, _, ms = rt.tone_map(gpu_ctx, dev, mode=4, glare=True, x_angle=1.1)
    assert out is dev and ms > 0.0
    assert np.array_equal(dev.cpu().numpy(), host)


def test_no_bright_pixel_adds_nothing(gpu_ctx, rt):
    img = hdr_image(64, 40, seed=8, scale=250.0, nbright=0)
    gp, _, _, _ = rt.tone_map(gpu_ctx, img, mode=0)
    gg, _, _, _ = rt.tone_map(gpu_ctx, img, mode=0, glare=True, x_angle=0.9)
    assert np.array_equal(gp, gg)


def test_glare_1080p_against_numpy_fft(gpu_ctx, rt, O):
    xr, yr = 1920, 1080
    rng = np.random.default_rng(11)
    img = (rng.random((yr, xr, 3)) * 120.0).astype(np.float32)
    # Scattered highlights and one bright disc near a corner (the circular
    # wrap of the convolution reaches the opposite edges).
    idx = rng.choice(xr * yr, 300, replace=False)
    img.reshape(-1, 3)[idx] = (rng.random((300, 3)) * 3000.0 + 300.0).astype(np.float32)
    yy, xx = np.mgrid[0:yr, 0:xr]
    img[(yy - 40) ** 2 + (xx - 1880) ** 2 < 400] = np.float32(2000.0)
    gp, _, gyi, _ = rt.tone_map(gpu_ctx, img, mode=0)
    gg, _, _, _ = rt.tone_map(gpu_ctx, img, mode=0, glare=True, x_angle=1.2)
    level = light_level(gyi)
    gf = rt.glare_filter(gpu_ctx, xr, yr, 1.2, level)
    ok, info = _filter_close(gf, O.glare_filter(xr, yr, 1.2, level))
    assert ok, info
    # numpy: the bright pixels of the mapped image convolved with the moved filter.
    n = xr * yr
    off = (xr >> 1) + xr * ((yr - 1) >> 1)
    i = np.arange(n)
    gin = gf.reshape(n, 3)[np.where(i < off, i + off, i - off)].reshape(yr, xr, 3).astype(np.float64)
    src = np.where((gp > 255.0).any(-1, keepdims=True), gp, 0.0).astype(np.float64)
    want = np.fft.ifft2(np.fft.fft2(src, axes=(0, 1)) * np.fft.fft2(gin, axes=(0, 1)), axes=(0, 1)).real / 255.0
    got = gg.astype(np.float64) - gp
    tol = 1e-4 * np.abs(want).max() + np.spacing(np.abs(gp).astype(np.float32)).astype(np.float64) + 1e-9
    err = np.abs(got - want)
    assert np.all(err <= tol), (err.max(), np.abs(want).max())
    assert np.abs(want).max() > 1e-2


def test_errors(gpu_ctx, rt):
    img = hdr_image(16, 8, seed=1, scale=10.0)
    with pytest.raises(rt.RTError, match="x_angle"):
        rt.tone_map(gpu_ctx, img, mode=0, glare=True)
    with pytest.raises(rt.RTError, match="4096"):
        rt.tone_map(gpu_ctx, np.zeros((1, 4097, 3), np.float32), mode=0, glare=True, x_angle=0.9)
    with pytest.raises(rt.RTError):
        rt.glare_filter(gpu_ctx, 16, 8, 0.9, 3)
    with pytest.raises(rt.RTError):
        rt.glare_filter(gpu_ctx, 16, 8, 0.0, 1)
    # The context still works.
    assert rt.glare_filter(gpu_ctx, 16, 8, 0.9, 1).shape == (8, 16, 3)
