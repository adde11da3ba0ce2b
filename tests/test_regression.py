"""The image-regression harness on the host: the reference checker's rule
(rt/regression_tests/regression_checker.h:44-170) on hand-built images, the
PNG reader against the library's writer and an independent encoder, and the
case table read from the reference's test sources."""
import json
import os

import numpy as np
import pytest

from raptor_amd import regression as R

REF_TESTS = "/root/reference/RAPTor/raytracer/regression_tests"


def blank(x_res, y_res, v=100):
    return np.full((y_res, x_res, 3), v, np.uint8)


def test_identical_images_pass():
    a = blank(16, 9)
    assert R.checker(a, a.copy(), 16, 9) == (0, 0)


def test_pixel_tolerance_is_five():
    a, e = blank(16, 9), blank(16, 9)
    a[4, 5, 1] = 105  # off by 5: not a mismatch
    assert R.checker(a, e, 16, 9) == (0, 0)
    a[4, 5, 1] = 106  # off by 6: a mismatch, but alone
    assert R.checker(a, e, 16, 9) == (0, 1)


def checker_view_block(x_res, y_res, i0, j0):
    """Pixels forming a 3 x 3 block in the checker's (x_res, y_res) view."""
    a = blank(x_res, y_res).reshape(-1, 3)
    for i in range(i0, i0 + 3):
        for j in range(j0, j0 + 3):
            a[i * y_res + j] = 0
    return a.reshape(y_res, x_res, 3)


def test_block_of_mismatches_fails_five_pixels():
    # In a 3 x 3 block of mismatches the centre sees 9 and each edge middle 6
    # mismatching neighbours (> 4), each corner 4 (not > 4): 5 failures.
    e = blank(12, 10)
    a = checker_view_block(12, 10, 4, 3)
    assert R.checker(a, e, 12, 10) == (5, 9)


def test_border_pixels_are_not_checked():
    e = blank(12, 10)
    a = checker_view_block(12, 10, 0, 3)  # rows 0..2 of the checker's view
    # As in the interior block, but the top edge middle (0, 4), which sees
    # 6, lies on the border row 0: 4 failures instead of 5.
    assert R.checker(a, e, 12, 10) == (4, 9)


def test_checker_reads_the_buffer_as_x_res_rows():
    # An image-space 3 x 3 block at x 1..3, y 1..3 of a 6 x 4 image: flat
    # indices 7..9, 13..15, 19..21 are, as (i, j) = divmod(k, 4), the cells
    # (1,3) (2,0) (2,1) (3,1) (3,2) (3,3) (4,3) (5,0) (5,1). Interior
    # mismatches: (2,1) sees 4, (3,1) sees 4, (3,2) sees 5: one failure.
    a, e = blank(6, 4), blank(6, 4)
    a[1:4, 1:4] = 0
    assert R.checker(a, e, 6, 4) == (1, 9)


def test_png_round_trip_through_the_writer(rt, tmp_path):
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (37, 53, 3), dtype=np.uint8)
    p = tmp_path / "a.png"
    rt.write_png(p, img)
    # write_png_file stores the last row first; read_png_file flips it back.
    assert np.array_equal(R.read_png(p), img[::-1])
    assert np.array_equal(R.expected_image(p), img)


def test_png_reader_decodes_filtered_rows(tmp_path):
    pil = pytest.importorskip("PIL.Image")
    rng = np.random.default_rng(2)
    # smooth gradients plus noise make an encoder use every filter type
    y, x = np.mgrid[0:41, 0:67]
    img = np.stack([(x * 3 + y) % 256, (y * 5) % 256, rng.integers(0, 256, x.shape)], -1).astype(np.uint8)
    p = tmp_path / "b.png"
    pil.fromarray(img).save(p, optimize=True)
    assert np.array_equal(R.read_png(p), img)
    pil.fromarray(np.dstack([img, img[:, :, :1]])).save(tmp_path / "c.png")
    assert np.array_equal(R.read_png(tmp_path / "c.png"), img)


def test_case_table():
    cases = R.load_cases()
    assert len(cases) == 194
    assert {s: len(R.load_cases(s)) for s in ("off", "ply", "obj", "nff")} == {"off": 68, "ply": 23, "obj": 59,
                                                                               "nff": 44}
    assert len({c["name"] for c in cases}) == 194
    for c in cases:
        assert c["format"] == c["suite"] and not c["unsupported"]
        if c["suite"] == "nff":  # viewpoint, lights and background come from the file
            assert c["camera_from_file"] and not c["lights"] and c["input"].endswith(".nff")
            continue
        assert c["lights"] and c["resolution"][0] > 0 and c["input"].startswith("/" + c["suite"] + "_scenes/")
    torus = [c for c in cases if c["name"] == "off_tst_torus_model_test"][0]
    assert torus["camera"]["pos"] == [75.0, -100.0, -11.0] and torus["resolution"] == [1920, 1080]
    assert torus["lights"] == [dict(rgb=[255.0] * 3, pos=[50.0, -200.0, -70.0], d=0.0, r=10.0)]


@pytest.mark.skipif(not os.path.isdir(REF_TESTS), reason="reference checkout absent")
def test_case_table_matches_the_reference_sources(tmp_path, monkeypatch):
    import importlib.util

    spec = importlib.util.spec_from_file_location(
        "make_regression_cases", os.path.join(os.path.dirname(__file__), "golden", "make_regression_cases.py"))
    M = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(M)

    monkeypatch.setattr(M, "OUT", str(tmp_path / "cases.json"))
    M.main("/root/reference/RAPTor")
    assert json.load(open(tmp_path / "cases.json")) == json.load(open(R.CASES))
    # every case but one names an expected image the reference holds
    missing = [c["name"] for c in R.load_cases()
               if not os.path.exists(os.path.join(REF_TESTS, "test_data", c["suite"] + "_tests", c["name"] + "_exp.png"))]
    assert missing == ["obj_crytek_sponza_test"]


def test_missing_model_is_reported(tmp_path):
    case = R.load_cases("off")[0]
    assert R.run_case(None, case, str(tmp_path))["status"] == "missing-input"
