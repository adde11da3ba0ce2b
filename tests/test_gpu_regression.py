"""The image-regression harness end to end with the GPU as the accelerator
under test: a case in the reference's table layout (model under a data
directory, fixture camera, lights, background) is parsed, built, rendered,
clipped and checked with the reference's checker against an expected PNG
written the way the reference writes them (here: the oracle's frame of the
same case, clipped to bytes)."""
import os
import shutil

import numpy as np
import pytest

from helpers import oracle_material_rows

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = os.path.join(HERE, "golden", "meshes")


def _case(rt, name, mesh_file, background, direction, res=(320, 180)):
    from raptor_amd import workloads

    mesh = rt.Mesh(os.path.join(MESHES, mesh_file))
    w = workloads.mesh_view(mesh.abc, res[0], res[1], direction=direction)
    c = w.camera
    lights = [dict(rgb=list(l.rgb), pos=list(l.pos), d=l.d, r=10.0) for l in w.lights]
    lights.append(dict(rgb=[120.0, 120.0, 160.0], pos=[-5.0, 6.0, -5.0], d=0.0, r=10.0))
    return dict(suite="off", name=name, input="/off_scenes/" + mesh_file, format="off",
                camera=dict(pos=list(c.pos), x=list(c.x_axis), y=list(c.y_axis), z=list(c.z_axis)),
                background=list(background), rotation=[0.0, 0.0, 0.0], resolution=list(res), lights=lights,
                unsupported=[]), mesh


def _expected(rt, O, R, case, mesh, path, lights=None):
    cam = R.case_camera(case)
    albedo = (mesh.materials["kd"][mesh.tri_material] / np.float32(255.0)).astype(np.float32)
    ref = O.OracleScene(mesh.abc).render(albedo, cam, lights or R.case_lights(case), nthreads=8,
                                         materials=oracle_material_rows(mesh.materials),
                                         tri_material=mesh.tri_material, background=tuple(case["background"]))
    img = O.clip(ref["rgb"].reshape(cam.y_res, cam.x_res, 3))
    os.makedirs(os.path.dirname(path), exist_ok=True)
    rt.write_png(path, img)  # last row first, as write_png_file stores it
    return img


@pytest.mark.parametrize("mesh_file,background,direction", [
    ("tstTorusModel.off", (0.0, 0.0, 0.0), (0.45, 0.35, -0.82)),
    ("helmet.off", (255.0, 255.0, 255.0), (-0.3, -0.5, 0.8)),
    ("hornbug.off", (0.0, 0.0, 200.0), (0.7, -0.2, 0.5)),
])
def test_harness_passes_the_gpu_frame(gpu_ctx, rt, O, tmp_path, mesh_file, background, direction):
    from raptor_amd import regression as R

    name = "off_" + mesh_file.split(".")[0].lower() + "_test"
    case, mesh = _case(rt, name, mesh_file, background, direction)
    data, expected, out = tmp_path / "data", tmp_path / "expected", tmp_path / "out"
    os.makedirs(data / "off_scenes")
    shutil.copy(os.path.join(MESHES, mesh_file), data / "off_scenes" / mesh_file)
    _expected(rt, O, R, case, mesh, str(expected / "off_tests" / (name + "_exp.png")))
    res = R.run_case(gpu_ctx, case, str(data), str(expected), str(out))
    assert res["status"] == "pass" and res["failures"] == 0, res
    assert res["mismatched_pixels"] <= 4, res  # exact-t ties at most
    act = R.expected_image(str(out / "off" / (name + "_gpu_act.png")))
    assert act.shape == (180, 320, 3)

    # A different light makes the same case fail the checker.
    moved = [R.case_lights(case)[0], R.case_lights(case)[0]]
    _expected(rt, O, R, case, mesh, str(expected / "off_tests" / (name + "_exp.png")), lights=moved)
    res = R.run_case(gpu_ctx, case, str(data), str(expected))
    assert res["status"] == "fail" and res["failures"] > 0, res
