"""GPU camera post-process (rt/camera.cc:1124-1501, rt/camera.h:289-351)
through the C-ABI against the oracle.

Tone mapping agrees within tol(n) = 1e-4 + 1e-9 n of the image's largest
value (n pixels): the GPU sums log10 Y in double over fixed blocks, the
reference (and the oracle) in float pixel by pixel, and the log-average
luminance they give differs by the float sum's rounding, which grows with n.
The GPU's log-average is checked against a float64 evaluation to 1e-5.
Gamma agrees within 1e-4 (powf on either side). The clip is bit-exact on the
same float image.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODES = [0, 1, 4, 5, 6, 7, 9]


def _frame(rt, ctx, res=(320, 180)):
    from raptor_amd import workloads

    w = workloads.terrain(grid=200, x_res=res[0], y_res=res[1], bounce=True)
    scene = rt.Scene(ctx, w.abc, w.albedo)
    accel = rt.Accel(scene)
    return accel.render(w.camera, w.lights, bounce=True, seed=w.seed)["rgb"]


def _tol(img):
    return 1e-4 + 1e-9 * img.shape[0] * img.shape[1]


def _close(a, b, tol=1e-4):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b).max() <= tol * max(1.0, np.abs(b).max())


def _log_average(img):
    """10^mean(log10 Y) in float64 over the float32 luminance the shader
    computes (rt/camera.cc:1137-1162), zero-luminance pixels left out."""
    f = np.float32
    rgb = img.astype(np.float32) / f(255.0)
    y = f(683.0) * ((f(0.2126) * rgb[..., 0]) + (f(0.7151) * rgb[..., 1]) + (f(0.0721) * rgb[..., 2]))
    y = y[y >= f(3.18e-7)].astype(np.float64)
    return 10.0 ** np.mean(np.log10(y))


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("ds,sc", [(False, False), (True, False), (False, True)])
def test_tone_map_matches_oracle(gpu_ctx, rt, O, mode, ds, sc):
    img = _frame(rt, gpu_ctx)
    g, glvl, gyi, _ = rt.tone_map(gpu_ctx, img, mode=mode, key=0.3, desaturate=ds, stretch=sc)
    o, olvl, oyi = O.tone_map(img, mode=mode, key=0.3, desaturate=ds, stretch=sc)
    assert abs(gyi / _log_average(img) - 1.0) < 1e-5
    assert abs(gyi / oyi - 1.0) < _tol(img) and abs(glvl / olvl - 1.0) < _tol(img)
    assert _close(g, o, 2 * _tol(img)), (np.abs(g - o).max(), np.abs(o).max())


def test_tone_map_random_images_and_adaptation(gpu_ctx, rt, O):
    rng = np.random.default_rng(3)
    for k, (h, w) in enumerate([(1, 1), (7, 13), (257, 129), (600, 900)]):
        img = (rng.uniform(0, 1, (h, w, 3)) ** 3 * 500).astype(np.float32)
        for mode in (4, 9):
            g, glvl, _, _ = rt.tone_map(gpu_ctx, img, mode=mode, adapt=True, time_step=0.05,
                                        adaptation_level=10.0 ** (k - 1))
            o, olvl, _ = O.tone_map(img, mode=mode, adapt=True, time_step=0.05, adaptation_level=10.0 ** (k - 1))
            assert abs(glvl / olvl - 1.0) < _tol(img)
            assert _close(g, o, 2 * _tol(img))


def test_tone_map_on_device_tensors(gpu_ctx, rt, O):
    import torch

    img = _frame(rt, gpu_ctx, res=(256, 144))
    dev = torch.from_numpy(img).to("cuda:0")
    out, _, _, ms = rt.tone_map(gpu_ctx, dev, mode=4, key=0.18)
    assert out is dev and ms > 0.0
    o, _, _ = O.tone_map(img, mode=4, key=0.18)
    assert _close(dev.cpu().numpy(), o, 2 * _tol(img))


@pytest.mark.parametrize("gamma", [1.0, 1.8, 2.2])
def test_gamma_matches_oracle(gpu_ctx, rt, O, gamma):
    rng = np.random.default_rng(4)
    img = (rng.uniform(-0.1, 1.2, (90, 160, 3)) * 255).astype(np.float32)
    g, _ = rt.gamma_correct(gpu_ctx, img, gamma)
    assert np.abs(g - O.gamma_correct(img, gamma)).max() < 1e-4 * 255


@pytest.mark.parametrize("xs,ys,bgr", [(1, 1, False), (2, 2, False), (2, 3, True), (4, 1, True)])
def test_clip_is_bit_exact(gpu_ctx, rt, O, xs, ys, bgr):
    rng = np.random.default_rng(5)
    img = (rng.uniform(-0.2, 1.3, (180, 320, 3)) * 255).astype(np.float32)
    g, _ = rt.clip(gpu_ctx, img, xs, ys, bgr)
    assert np.array_equal(g, O.clip(img, xs, ys, bgr))


def test_frame_to_png_pipeline(gpu_ctx, rt, O, tmp_path):
    """render -> tone map -> gamma -> clip -> PNG on the device, against the
    oracle's pipeline on the same frame: bytes equal except where a float
    lands within rounding of an integer (at most one step apart)."""
    import torch
    import zlib

    img = _frame(rt, gpu_ctx, res=(480, 270))
    dev = torch.from_numpy(img).to("cuda:0")
    rt.tone_map(gpu_ctx, dev, mode=4, key=0.25)
    rt.gamma_correct(gpu_ctx, dev, 2.2)
    b, _ = rt.clip(gpu_ctx, dev, 2, 2)
    b = b.cpu().numpy()
    o, _, _ = O.tone_map(img, mode=4, key=0.25)
    ob = O.clip(O.gamma_correct(o, 2.2), 2, 2)
    d = np.abs(b.astype(int) - ob.astype(int))
    assert d.max() <= 1 and (d == 0).mean() > 0.95
    path = tmp_path / "frame.png"
    rt.write_png(path, b)
    assert path.stat().st_size > b.size


def test_post_process_errors(gpu_ctx, rt):
    img = np.zeros((4, 4, 3), np.float32)
    for mode in (2, 3):
        with pytest.raises(rt.RTError, match="field of view"):
            rt.tone_map(gpu_ctx, img, mode=mode)  # no x_angle / y_angle
    with pytest.raises(rt.RTError, match="bilateral filter needs"):
        rt.tone_map(gpu_ctx, img, mode=8)  # no bright pixel: no log-intensity range
    with pytest.raises(rt.RTError, match="unknown mode"):
        rt.tone_map(gpu_ctx, img, mode=12)
    with pytest.raises(rt.RTError, match="sample counts"):
        rt.clip(gpu_ctx, img, 5, 1)


def test_bilateral_refuses_an_extremely_wide_frame(gpu_ctx, rt):
    """One row of 50M pixels: the grid's x extent (49,999,999 / 0.02 cells)
    is beyond any int, so it is sized in double and refused before any
    narrowing conversion."""
    import torch

    img = torch.full((1, 50_000_000, 3), 100.0, dtype=torch.float32, device="cuda:0")
    img[0, ::7] = 3.0  # a range of bright pixels, so the size check is reached
    with pytest.raises(rt.RTError, match="2\\^27"):
        rt.tone_map(gpu_ctx, img, mode=8)
    del img
    torch.cuda.empty_cache()


@pytest.mark.parametrize("ds,sc", [(False, False), (True, False), (False, True)])
@pytest.mark.parametrize("res", [(320, 180), (256, 256)])
def test_bilateral_tone_map_matches_oracle(gpu_ctx, rt, O, ds, sc, res):
    # global_bilateral_filter (rt/camera.cc:1354-1358, 1632-1775). The grid
    # is blurred in double on both sides (the reference: single-precision
    # FFTs); the splat and the float pow round differently, so the bound is
    # relative to the image's range.
    img = _frame(rt, gpu_ctx, res)
    g, _, gyi, _ = rt.tone_map(gpu_ctx, img, mode=8, desaturate=ds, stretch=sc)
    o, _, oyi = O.tone_map(img, mode=8, desaturate=ds, stretch=sc)
    assert np.isfinite(g).all()
    assert abs(gyi / oyi - 1.0) < _tol(img)
    err = np.abs(g.astype(np.float64) - o) / max(1.0, np.abs(o).max())
    assert err.max() <= 2e-4, (err.max(), np.abs(o).max())
    g2, _, _, _ = rt.tone_map(gpu_ctx, img, mode=8, desaturate=ds, stretch=sc)
    assert np.array_equal(g, g2)  # fixed-order sums: the frame repeats bit for bit


def test_bilateral_on_hdr_image(gpu_ctx, rt, O):
    rng = np.random.default_rng(4)
    h, w = 150, 230
    y, x = np.mgrid[0:h, 0:w]
    lum = 10.0 ** (3.0 * x / w - 1.0 + 0.3 * np.sin(y / 7.0)) * (1.0 + 0.2 * rng.uniform(-1, 1, (h, w)))
    img = (lum[..., None] * np.array([1.0, 0.8, 0.6])).astype(np.float32)
    img[60:70, 100:140] = 0.0  # a dark patch: no log intensity
    g, _, _, _ = rt.tone_map(gpu_ctx, img, mode=8)
    o, _, _ = O.tone_map(img, mode=8)
    err = np.abs(g.astype(np.float64) - o) / max(1.0, np.abs(o).max())
    assert err.max() <= 2e-4, err.max()


@pytest.mark.parametrize("mode", [2, 3])
@pytest.mark.parametrize("sc", [False, True])
def test_histogram_tone_map_matches_oracle(gpu_ctx, rt, O, mode, sc):
    # local_histogram / local_human_histogram (rt/camera.cc:1843-2062) with
    # the frame's field of view (camera half width 0.5 at distance 1). The
    # one-per-degree samples are summed in the reference's order on both
    # sides; logf / expf may round a sample across a bin edge, so the bound
    # is relative to the image's range.
    img = _frame(rt, gpu_ctx, (320, 180))
    xa, ya = 1.0, 2.0 * 0.5 * 180.0 / 320.0
    g, glvl, _, _ = rt.tone_map(gpu_ctx, img, mode=mode, stretch=sc, x_angle=xa, y_angle=ya)
    o, olvl, _ = O.tone_map(img, mode=mode, stretch=sc, x_angle=xa, y_angle=ya)
    assert np.isfinite(g).all()
    err = np.abs(g.astype(np.float64) - o) / max(1.0, np.abs(o).max())
    assert err.max() <= 2e-3 and np.median(err) <= 1e-5, (err.max(), np.median(err))


def test_histogram_hdr_and_linear_branches(gpu_ctx, rt, O):
    rng = np.random.default_rng(8)
    lum = 10.0 ** rng.uniform(-2.0, 3.0, (150, 260))
    hdr = np.repeat(lum[..., None], 3, axis=2).astype(np.float32)
    flat = np.full((90, 160, 3), 100.0, np.float32)
    deg = 360.0 / (2.0 * np.pi)
    for img, xa, ya in [(hdr, 1.1, 0.63), (flat, 40.5 / deg, 30.5 / deg)]:
        for mode in (2, 3):
            g, _, _, _ = rt.tone_map(gpu_ctx, img, mode=mode, x_angle=xa, y_angle=ya)
            o, _, _ = O.tone_map(img, mode=mode, x_angle=xa, y_angle=ya)
            err = np.abs(g.astype(np.float64) - o) / max(1.0, np.abs(o).max())
            assert err.max() <= 2e-3 and np.median(err) <= 1e-5, (mode, err.max())


def _psnr(a, b):
    mse = np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)
    return 10.0 * np.log10(255.0 ** 2 / max(mse, 1e-12))


@pytest.mark.parametrize("quality", [50, 75, 95])
@pytest.mark.parametrize("res", [(320, 180), (101, 67)])
def test_jpeg_writer(gpu_ctx, rt, tmp_path, quality, res):
    # camera::write_jpeg_file (rt/camera.cc:2305-2366): libjpeg defaults at
    # the given quality. Checked by decoding with an independent decoder
    # (Pillow's libjpeg) and against Pillow's own encoder with the same
    # settings: identical quantisation tables (the IJG quality rule), and the
    # same fidelity and size within a small margin.
    pil = pytest.importorskip("PIL.Image")
    img = _frame(rt, gpu_ctx, res)
    b, _ = rt.clip(gpu_ctx, img)
    p = tmp_path / "a.jpg"
    rt.write_jpeg(gpu_ctx, p, b, quality)
    ours = pil.open(p)
    assert ours.size == (res[0], res[1]) and ours.mode == "RGB"
    dec = np.asarray(ours.convert("RGB"))
    src = b[::-1]  # written last row first
    ref_path = tmp_path / "ref.jpg"
    pil.fromarray(np.ascontiguousarray(src)).save(ref_path, quality=quality, subsampling=2)
    ref = pil.open(ref_path)
    assert ours.quantization == ref.quantization
    p_ours, p_ref = _psnr(dec, src), _psnr(np.asarray(ref.convert("RGB")), src)
    assert p_ours > 15.0 and p_ours > p_ref - 0.3, (p_ours, p_ref)  # the frame is high-frequency noise
    assert 0.85 < os.path.getsize(p) / os.path.getsize(ref_path) < 1.15
