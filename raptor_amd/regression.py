"""The reference's image-regression harness with this library as the
accelerator under test.

The reference renders each case of rt/regression_tests/{off,ply,obj,nff}_tests.cc
with its bvh, kd-tree and BIH and compares every image with an expected PNG
(rt/regression_tests/regression_checker.h:44-170). This module runs the same
cases through the GPU path: parse the model under $RAPTOR_DATA
(rt_mesh_load), build the BVH, render with the fixture's camera, lights and
background (rt/regression_tests/regression_fixture.h:86-241), clip to bytes,
write the actual PNG, and apply the reference's checker to the expected one.

The checker, as the reference writes it: a pixel differs when any channel
differs by more than 5; a difference counts as a failure when, in a 3 x 3
neighbourhood away from the border, more than 4 of the 9 pixels differ. Its
neighbourhood indexes the image buffer as x_res rows of y_res pixels
(index (i * y_res + j) for i < x_res, j < y_res), not as the image's own
rows; that reading is kept, so verdicts match the reference's.

    python -m raptor_amd.regression --data $RAPTOR_DATA \\
        --expected RAPTor/raytracer/regression_tests/test_data --suite off --out /tmp/regress
"""
import argparse
import json
import os
import struct
import sys
import time
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = os.path.join(HERE, "regression_cases.json")
PIXEL_ERROR = 5   # regression_checker.h:25
COUNT_ERROR = 4   # regression_checker.h:24
FAILURE_LIMIT = 1000  # regression_checker.h:26


def load_cases(suite=None):
    with open(CASES) as fh:
        cases = json.load(fh)["cases"]
    return [c for c in cases if suite in (None, "all", c["suite"])]


def pixel_mismatch(actual, expected):
    """(n,) bool: any channel differs by more than PIXEL_ERROR."""
    a = np.asarray(actual, np.int32).reshape(-1, 3)
    e = np.asarray(expected, np.int32).reshape(-1, 3)
    return (np.abs(a - e) > PIXEL_ERROR).any(axis=1)


def checker(actual, expected, x_res, y_res):
    """The reference checker's verdict on two x_res * y_res RGB byte images in
    camera order: (perceivable failures, mismatching pixels). Failures are
    the pixels whose 3 x 3 neighbourhood (the checker's indexing, see the
    module text) holds more than COUNT_ERROR mismatches, border excluded."""
    bad = pixel_mismatch(actual, expected).reshape(x_res, y_res)
    if x_res < 3 or y_res < 3:
        return 0, int(bad.sum())
    b = bad.astype(np.int32)
    win = np.zeros((x_res - 2, y_res - 2), np.int32)
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            win += b[1 + di:x_res - 1 + di, 1 + dj:y_res - 1 + dj]
    fails = bad[1:-1, 1:-1] & (win > COUNT_ERROR)
    return int(fails.sum()), int(bad.sum())


def read_png(path):
    """An 8-bit RGB or RGBA PNG as (height, width, 3) bytes in file row order."""
    data = open(path, "rb").read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError(f"{path}: not a PNG")
    pos, idat, w = 8, b"", None
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        pos += 12 + n
        if kind == b"IHDR":
            w, h, depth, ctype, _, _, interlace = struct.unpack(">IIBBBBB", body)
            if depth != 8 or ctype not in (2, 6) or interlace:
                raise ValueError(f"{path}: only 8-bit RGB/RGBA non-interlaced PNGs are read")
            bpp = 3 if ctype == 2 else 4
        elif kind == b"IDAT":
            idat += body
        elif kind == b"IEND":
            break
    raw = np.frombuffer(zlib.decompress(idat), np.uint8)
    stride = w * bpp
    out = np.zeros((h, stride), np.uint8)
    prev = np.zeros(stride, np.int32)
    for y in range(h):
        f = raw[y * (stride + 1)]
        line = raw[y * (stride + 1) + 1:(y + 1) * (stride + 1)].astype(np.int32)
        cur = np.zeros(stride, np.int32)
        if f == 0:
            cur = line
        elif f == 2:
            cur = (line + prev) & 255
        elif f in (1, 3, 4):
            for x in range(stride):  # left-dependent filters, byte by byte
                a = cur[x - bpp] if x >= bpp else 0
                b = prev[x]
                c = prev[x - bpp] if x >= bpp else 0
                if f == 1:
                    p = a
                elif f == 3:
                    p = (a + b) >> 1
                else:
                    pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                    p = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
                cur[x] = (line[x] + p) & 255
        else:
            raise ValueError(f"{path}: bad PNG filter {f}")
        out[y] = cur
        prev = cur
    return out.reshape(h, w, bpp)[:, :, :3]


def expected_image(path):
    """read_png_file (rt/camera.cc:2227-2300): the file's first row is the
    image's last, so flip back to camera order."""
    return read_png(path)[::-1]


def case_camera(case):
    from .workloads import fixture_camera

    c = case["camera"]
    xr, yr = case["resolution"]
    return fixture_camera(c["pos"], c["x"], c["y"], c["z"], xr, yr)


def case_lights(case):
    from .workloads import Light, f32

    return [Light(f32(l["pos"]), f32(l["rgb"]), float(np.float32(l["d"]))) for l in case["lights"]]


def render_case(ctx, case, model_path, width=4, mesh=None):
    """The case's frame on the GPU: (rgb float image, clipped bytes, stats)."""
    from . import rt

    if case["unsupported"]:
        raise rt.RTError("unsupported case: " + ", ".join(case["unsupported"]))
    mesh = mesh or rt.Mesh(model_path)
    scene = rt.Scene.from_mesh(ctx, mesh)
    accel = rt.Accel(scene, width=width)
    if case.get("camera_from_file"):  # NFF: the viewpoint, background and lights of the file
        from .workloads import mesh_camera_lights

        cam, background, lights = mesh_camera_lights(mesh)
        lights = lights + case_lights(case)
    else:
        cam, background, lights = case_camera(case), tuple(case["background"]), case_lights(case)
    out = accel.render(cam, lights, want_hits=False, background=background)
    img, _ = rt.clip(ctx, out["rgb"])
    return out["rgb"], img, out["stats"], accel.info()


def run_case(ctx, case, data_dir, expected_dir=None, out_dir=None, width=4):
    """Renders one case; with expected_dir, applies the checker. Returns a
    result dict (status: pass, fail, missing-input, missing-expected,
    unsupported)."""
    from . import rt

    res = dict(name=case["name"], suite=case["suite"])
    if case["unsupported"]:
        return dict(res, status="unsupported", why=case["unsupported"])
    model = os.path.join(data_dir, case["input"].lstrip("/"))
    if not os.path.exists(model):
        return dict(res, status="missing-input", path=model)
    t0 = time.perf_counter()
    mesh = rt.Mesh(model)
    t1 = time.perf_counter()
    _, img, stats, info = render_case(ctx, case, model, width, mesh)
    t2 = time.perf_counter()
    res.update(triangles=mesh.n, parse_s=t1 - t0, frame_s=t2 - t1, frame_ms=stats["ms"][0], build_ms=info["build_ms"])
    if out_dir:
        os.makedirs(os.path.join(out_dir, case["suite"]), exist_ok=True)
        rt.write_png(os.path.join(out_dir, case["suite"], case["name"] + "_gpu_act.png"), img)
    if expected_dir is None:
        return dict(res, status="rendered")
    exp_path = os.path.join(expected_dir, case["suite"] + "_tests", case["name"] + "_exp.png")
    if not os.path.exists(exp_path):
        return dict(res, status="missing-expected", path=exp_path)
    exp = expected_image(exp_path)
    xr, yr = img.shape[1], img.shape[0]
    if exp.shape != (yr, xr, 3):
        return dict(res, status="fail", why=f"expected image is {exp.shape[1]}x{exp.shape[0]}")
    fails, mism = checker(img, exp, xr, yr)
    return dict(res, status="pass" if fails == 0 else "fail", failures=min(fails, FAILURE_LIMIT),
                mismatched_pixels=mism)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--data", default=os.environ.get("RAPTOR_DATA"), help="the models ($RAPTOR_DATA)")
    ap.add_argument("--expected", help="the reference's regression_tests/test_data directory")
    ap.add_argument("--suite", default="all", choices=["all", "off", "ply", "obj", "nff"])
    ap.add_argument("--out", help="write the actual images here")
    ap.add_argument("--width", type=int, default=4)
    ap.add_argument("--limit", type=int, default=0)
    a = ap.parse_args(argv)
    if not a.data:
        ap.error("--data (or $RAPTOR_DATA) is required")
    from . import rt

    ctx = rt.Context(0)
    cases = load_cases(a.suite)
    if a.limit:
        cases = cases[:a.limit]
    tally = {}
    for c in cases:
        r = run_case(ctx, c, a.data, a.expected, a.out, a.width)
        tally[r["status"]] = tally.get(r["status"], 0) + 1
        print(json.dumps(r), flush=True)
    print(json.dumps(dict(summary=tally, cases=len(cases))), flush=True)
    return 0 if tally.get("fail", 0) == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
