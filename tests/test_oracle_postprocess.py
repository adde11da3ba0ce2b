"""Oracle pins for the camera post-process (rt/camera.cc:1124-1501,
rt/camera.h:289-351) and checks of the host image writers.

The pins use what the mathematics fixes: the none mode is the identity up to
the rounded conversion constants; the global modes scale the none image's
luminance by their closed-form factors (Ward's contrast factor, key / max Y,
1 - exp(-key Y / Yi)); the desaturation leaves bright pixels alone and gives
dark ones the white point's chroma; a long adaptation time reaches the frame's own level; gamma
follows x^(1/g) or the sRGB curve; the clip is a box filter, a clamp and a
truncation. Ferwerda's mapping is pinned only as a linear map of luminance
(its response constants are unpinned).
"""
import struct
import zlib

import numpy as np
import pytest

from oracle import oracle as O

from helpers import rel_err


def _image(h=48, w=64, seed=0, scale=300.0):
    rng = np.random.default_rng(seed)
    return (rng.uniform(0.0, 1.0, (h, w, 3)) ** 2 * scale + 1.0).astype(np.float32)


def _lum683(img):
    rgb = img.astype(np.float64) / 255.0
    return 683.0 * (0.2126 * rgb[..., 0] + 0.7151 * rgb[..., 1] + 0.0721 * rgb[..., 2])


def test_none_mode_is_the_identity():
    img = _image()
    out, _, yi = O.tone_map(img, mode=0)
    # The two conversion matrices are inverses up to their printed digits.
    assert np.abs(out.astype(np.float64) - img).max() < 1e-3 * img.max()
    # The log-average luminance of the frame.
    y = _lum683(img)
    assert abs(yi / (10.0 ** np.mean(np.log10(y))) - 1.0) < 1e-5


@pytest.mark.parametrize("key", [0.18, 0.5])
def test_max_luminance_scales_the_brightest_pixel_to_key(key):
    img = _image(seed=1)
    base, _, _ = O.tone_map(img, mode=0)
    out, _, _ = O.tone_map(img, mode=7, key=key)
    factor = 683.0 * key / _lum683(img).max()
    assert np.abs(out - base * factor).max() < 1e-5 * np.abs(out).max()


def test_contrast_mode_is_wards_factor():
    img = _image(seed=2)
    base, _, yi = O.tone_map(img, mode=0)
    out, _, _ = O.tone_map(img, mode=1)
    sf = (1.0 / 300.0) * ((1.219 + 150.0 ** 0.4) / (1.219 + yi ** 0.4)) ** 2.5
    assert np.abs(out - base * 683.0 * sf).max() < 1e-5 * np.abs(out).max()


def test_exposure_mode_compresses_luminance():
    img = _image(seed=3)
    base, _, yi = O.tone_map(img, mode=0)
    out, _, _ = O.tone_map(img, mode=5, key=0.3)
    y = _lum683(img)
    expect = 1.0 - np.exp(-y * 0.3 / yi)
    # Chroma is kept, so each pixel is the none pixel scaled by Y' / (Y / 683).
    factor = (expect / (y / 683.0))[..., None]
    assert np.abs(out - base * factor).max() < 1e-4 * np.abs(out).max()


def test_avg_luminance_mode_keeps_the_683_scale():
    """The reference averages rgb only in global_exposure, so in
    global_avg_luminance the average is 0 and the scale stays 1: the image
    keeps the 683 lm/W factor."""
    img = _image(seed=4, scale=0.2)
    base, _, _ = O.tone_map(img, mode=0)
    out, _, _ = O.tone_map(img, mode=6, key=0.7)
    assert np.abs(out - base * 683.0).max() < 1e-5 * np.abs(out).max()


def test_ferwerda_is_linear_in_luminance():
    img = _image(seed=5)
    base, _, _ = O.tone_map(img, mode=0)
    out, _, _ = O.tone_map(img, mode=9)
    ratio = out[..., 1].astype(np.float64) / base[..., 1]
    ok = np.abs(base[..., 1]) > 1.0
    assert np.ptp(ratio[ok]) < 1e-4 * np.abs(ratio[ok]).mean()


def test_desaturation():
    bright = _image(seed=6, scale=3000.0) + 800.0
    a, _, _ = O.tone_map(bright, mode=0)
    b, _, _ = O.tone_map(bright, mode=0, desaturate=True)
    assert rel_err(b, a, floor=1.0).max() < 1e-5
    # Below log10 Y = -2 the chroma becomes the white point (1/3, 1/3), so
    # every pixel is X = Y = Z times the XYZ -> rgb matrix's row sums.
    dark = _image(seed=7, scale=0.002) * 0.001
    d, _, _ = O.tone_map(dark, mode=7, key=1.0, desaturate=True)
    white = np.array([826.353 - 391.986 - 127.143, -247.0695 + 478.329 + 10.5825, 14.2035 - 52.02 + 269.535])
    chroma = d.astype(np.float64) / d[..., 1:2]
    assert np.abs(chroma - white / white[1]).max() < 1e-4


def test_long_adaptation_reaches_the_frame_level():
    img = _image(seed=8)
    _, lvl0, yi0 = O.tone_map(img, mode=7)
    for start in (0.01, 1e5):
        _, lvl, yi = O.tone_map(img, mode=7, adapt=True, time_step=1e4, adaptation_level=start)
        assert abs(lvl / yi0 - 1.0) < 1e-5 and yi == lvl
    _, lvl, _ = O.tone_map(img, mode=7, adapt=True, time_step=0.01, adaptation_level=1e5)
    assert yi0 < lvl < 1e5  # a short step moves part of the way


def test_stretch_maps_the_luminance_range_to_unit():
    img = _image(seed=9) + 50.0
    out, _, _ = O.tone_map(img, mode=7, key=3.0, stretch=True)
    # Mapped luminance runs from m = 3 Ymin / Ymax to 3; the stretch maps Y'
    # to (Y' - m) / (min(1, 3) - m), which scales each pixel's rgb by
    # (Y' - m) / ((1 - m) Y').
    y = _lum683(img)
    m = 3.0 * y.min() / y.max()
    yp = 3.0 * y / y.max()
    base, _, _ = O.tone_map(img, mode=7, key=3.0)
    expect = base * ((yp - m) / ((1.0 - m) * yp))[..., None]
    assert np.abs(out - expect).max() < 1e-4 * np.abs(out).max()


@pytest.mark.parametrize("gamma", [1.0, 1.8, 2.2, 2.4])
def test_gamma(gamma):
    img = _image(seed=10, scale=400.0) - 30.0
    out = O.gamma_correct(img, gamma).astype(np.float64)
    v = np.clip(img.astype(np.float64) / 255.0, 0.0, 1.0)
    if abs(gamma - 2.2) > 0.001:
        expect = v ** (1.0 / gamma)
    else:
        expect = np.where(v <= 0.0031308, 12.92 * v, 1.055 * v ** (1.0 / 2.4) - 0.055)
    assert np.abs(out - 255.0 * expect).max() < 2e-4


@pytest.mark.parametrize("xs,ys,bgr", [(1, 1, False), (2, 2, False), (3, 2, True), (1, 4, True)])
def test_clip_is_a_box_filter(xs, ys, bgr):
    img = _image(h=24, w=36, seed=11, scale=320.0) - 20.0
    out = O.clip(img, xs, ys, bgr)
    h, w = 24 // ys, 36 // xs
    box = img[: h * ys, : w * xs].astype(np.float64).reshape(h, ys, w, xs, 3).mean(axis=(1, 3))
    expect = np.floor(np.clip(box, 0.0, 255.0)).astype(np.int64)
    if bgr:
        expect = expect[..., ::-1]
    near = np.abs(np.clip(box, 0, 255) - np.round(np.clip(box, 0, 255)))
    near = (near < 1e-4)[..., ::-1] if bgr else near < 1e-4
    assert np.array_equal(out[~near], expect[~near])


def test_png_writer_round_trips(tmp_path):
    from raptor_amd import rt

    rng = np.random.default_rng(12)
    img = rng.integers(0, 256, (37, 70, 3), dtype=np.uint8)
    path = tmp_path / "f.png"
    rt.write_png(path, img)
    data = path.read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, ihdr = 8, b"", None
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert crc == zlib.crc32(kind + body) & 0xFFFFFFFF
        if kind == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    assert ihdr == (70, 37, 8, 2, 0, 0, 0)
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(37, 70 * 3 + 1)
    assert np.all(raw[:, 0] == 0)
    # The file's first row is the image's last (rt/camera.cc:2194-2200).
    assert np.array_equal(raw[:, 1:].reshape(37, 70, 3), img[::-1])


def test_tga_writer(tmp_path):
    from raptor_amd import rt

    rng = np.random.default_rng(13)
    img = rng.integers(0, 256, (5, 300, 3), dtype=np.uint8)
    path = tmp_path / "f.tga"
    rt.write_tga(path, img)
    data = path.read_bytes()
    assert data[:18] == bytes([0, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 300 & 255, 300 >> 8, 5, 0, 24, 0])
    assert data[18:-8] == img.tobytes() and data[-8:] == bytes(8)


def test_bilateral_grid_blur_is_the_fft_circular_convolution():
    # rt/camera.cc:1697-1734 convolves with r2c / c2r FFTs and a 1 / N scale:
    # a circular convolution with the periodic Gaussian the reference builds.
    rng = np.random.default_rng(0)
    for shape, ss, sr in [((2, 9, 11, 14), 1.0, 1.0), ((2, 6, 8, 5), 1.3, 0.7)]:
        g = rng.uniform(0, 1, shape).astype(np.float32)
        out = O.gauss_circular3d(g, ss, sr)
        _, D, H, W = shape

        def wrap(n):
            i = np.arange(n)
            return i - np.where(i > (n >> 1), n, 0)

        X, Y, Z = np.meshgrid(wrap(W), wrap(H), wrap(D), indexing="ij")
        k = np.exp(-0.5 * ((X * X + Y * Y) / ss ** 2 + Z * Z / sr ** 2)).transpose(2, 1, 0)
        kf = np.fft.rfftn(k, axes=(0, 1, 2))
        for c in range(2):
            ref = np.fft.irfftn(np.fft.rfftn(g[c].astype(np.float64), axes=(0, 1, 2)) * kf, s=(D, H, W),
                                axes=(0, 1, 2))
            assert np.abs(out[c] - ref).max() <= 1e-6 * np.abs(ref).max()


def _hdr(h=90, w=140, seed=3, detail=0.0):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    logl = 2.5 * x / w - 0.5 + 0.4 * np.sin(y / 9.0) + 0.05 * rng.uniform(-1, 1, (h, w))
    logl = logl + detail * np.where((x + y) % 2 == 0, 1.0, -1.0)
    return (10.0 ** logl)[..., None] * np.array([1.0, 0.9, 0.7])


def test_bilateral_scales_as_a_power_of_the_exposure():
    # Scaling the image by c shifts every log intensity, the grid's origin
    # and the base by log10 c, so the output scales by c^gamma, one factor
    # for every pixel.
    img = _hdr().astype(np.float32)
    a, _, _ = O.tone_map(img, mode=8)
    b, _, _ = O.tone_map((img * np.float32(4.0)).astype(np.float32), mode=8)
    ratio = (b[..., 1] / a[..., 1]).ravel()
    assert ratio.min() > 1.0 and (ratio.max() - ratio.min()) / ratio.mean() < 2e-4


def test_bilateral_keeps_fine_detail():
    # A checkerboard of +-0.1 decades sits inside one depth cell (0.4) and
    # below the grid's spatial scale, so it is detail: the output keeps its
    # phase and its amplitude (a log ratio of about 0.2 between neighbours),
    # while the smooth part is compressed.
    img = _hdr(detail=0.1).astype(np.float32)
    out, _, _ = O.tone_map(img, mode=8)
    lum = out[..., 1].astype(np.float64)
    lin = img[..., 1].astype(np.float64)
    d_out = np.log10(lum[20:70, 21:121] / lum[20:70, 20:120])
    d_in = np.log10(lin[20:70, 21:121] / lin[20:70, 20:120])
    assert np.all(np.sign(d_out) == np.sign(d_in))
    assert np.abs(d_out - d_in).max() < 0.02 and abs(np.abs(d_in).mean() - 0.2) < 0.01


def _lum(img):
    f = np.float64
    return 0.2126 * img[..., 0].astype(f) + 0.7151 * img[..., 1].astype(f) + 0.0721 * img[..., 2].astype(f)


def test_histogram_flat_image_takes_the_linear_branch():
    # A flat image averaged over whole blocks (40 x 30 samples of 5 x 4
    # pixels sharing their edges: each block sums count = extent pixels) gives
    # equal samples, so the log range is bumped by 1e-4 |ln Lw|, below the
    # display range: Y is scaled by ln 300 / (1e-4 |ln Lw|)
    # (rt/camera.cc:1846-1937), and rgb (= 255 M^-1 XYZ) by 683 times that.
    deg = 360.0 / (2.0 * np.pi)
    img = np.full((90, 160, 3), 100.0, np.float32)
    out, _, _ = O.tone_map(img, mode=2, x_angle=40.5 / deg, y_angle=30.5 / deg)
    f = np.float32
    y = f(683.0) * ((f(0.2126) * f(100.0 / 255.0)) + (f(0.7151) * f(100.0 / 255.0)) + (f(0.0721) * f(100.0 / 255.0)))
    sf = np.log(300.0) / (1e-4 * abs(np.log(float(y))))
    ratio = _lum(out) / _lum(img)
    assert ratio.std() / ratio.mean() < 1e-5
    assert abs(ratio.mean() / (683.0 * sf) - 1.0) < 2e-3


@pytest.mark.parametrize("mode", [2, 3])
def test_histogram_mapping_is_monotone_into_the_display_range(mode):
    # The distribution mapping is non-decreasing in luminance and lands in
    # [1, 300] / 300 (rt/camera.cc:2038-2057).
    rng = np.random.default_rng(5)
    h, w = 120, 200
    y, x = np.mgrid[0:h, 0:w]
    lum = 10.0 ** (4.0 * x / w - 1.5 + 0.5 * np.sin(y / 11.0)) * (1.0 + 0.1 * rng.uniform(-1, 1, (h, w)))
    img = np.repeat(lum[..., None], 3, axis=2).astype(np.float32)  # grey: Y maps straight through
    out, _, _ = O.tone_map(img, mode=mode, x_angle=1.2, y_angle=0.72)
    yin = (683.0 / 255.0) * lum.ravel()
    yout = out[..., 1].ravel().astype(np.float64)  # grey in, grey out: g channel follows Y
    order = np.argsort(yin, kind="stable")
    assert np.all(np.diff(yout[order]) >= -1e-3 * np.abs(yout).max())
    ymap = _lum(out).ravel() / 255.0  # rgb = 255 M^-1 XYZ
    assert ymap.min() >= (1.0 / 300.0) * 0.999 and ymap.max() <= 1.001


def test_human_histogram_differs_from_the_plain_one():
    rng = np.random.default_rng(6)
    lum = 10.0 ** rng.uniform(-2.0, 3.0, (100, 160))
    img = np.repeat(lum[..., None], 3, axis=2).astype(np.float32)
    a, _, _ = O.tone_map(img, mode=2, x_angle=1.0, y_angle=0.625)
    b, _, _ = O.tone_map(img, mode=3, x_angle=1.0, y_angle=0.625)
    assert np.abs(a - b).max() > 1.0
