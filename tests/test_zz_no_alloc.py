"""rt_render_static (no device allocation; the frame workspace is the
library's static scratch) against rt_render on the same inputs, through the
C ABI: images and hit records equal bit for bit on two calls in a row, the
second with a smaller frame (the array keeps the first call's contents).
The first frame's workspace (1920x1080 with a bounce, ~0.5 GB) exceeds the
64 MiB scratch, so it runs in chunks of tiles; the second fits in one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _frames(rt, accel, w, res, spp, static):
    import torch

    cam = w.camera
    cam = type(cam)(cam.pos, cam.x_axis, cam.y_axis, cam.z_axis, cam.w, cam.w * res[1] / res[0], cam.t, res[0],
                    res[1])
    nout = res[0] * res[1] * spp
    rgb = torch.full((res[1], res[0], 3), -1.0, dtype=torch.float32, device="cuda:0")
    prim = torch.full((nout, 4), -1.0, dtype=torch.float32, device="cuda:0")
    bnc = torch.full((nout, 4), -1.0, dtype=torch.float32, device="cuda:0")
    st = accel.render_device(cam, w.lights, rgb, bounce=True, spp=spp, seed=w.seed, primary_dev=prim,
                             bounce_dev=bnc, static=static)
    torch.cuda.synchronize()
    return rgb.cpu().numpy(), prim.cpu().numpy(), bnc.cpu().numpy(), st


@pytest.mark.parametrize("width", [4, 8])
def test_static_render_equals_render_on_two_calls(gpu_ctx, rt, width):
    from raptor_amd import workloads

    w = workloads.terrain(grid=256, x_res=1920, y_res=1080, bounce=True, seed=3)
    scene = rt.Scene(gpu_ctx, w.abc, w.albedo)
    accel = rt.Accel(scene, width=width)
    for res, spp in (((1920, 1080), 1), ((320, 200), 2)):
        ref = _frames(rt, accel, w, res, spp, static=False)
        got = _frames(rt, accel, w, res, spp, static=True)
        for a, b in zip(got[:3], ref[:3]):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        assert got[3]["rays"] == ref[3]["rays"]
        assert got[3]["rays"][0] == res[0] * res[1] * spp and got[3]["rays"][3] > 0


def test_static_render_refuses_what_it_cannot_hold(gpu_ctx, rt):
    import torch

    from raptor_amd import workloads

    w = workloads.terrain(grid=32, x_res=16, y_res=16, bounce=True)
    scene = rt.Scene(gpu_ctx, w.abc, w.albedo)
    accel = rt.Accel(scene)
    rgb = torch.empty((16, 16, 3), dtype=torch.float32, device="cuda:0")
    # One 8x8 tile at 8192 samples per pixel needs ~0.5 GB of workspace.
    with pytest.raises(rt.RTError, match="static scratch holds"):
        accel.render_device(w.camera, w.lights, rgb, bounce=True, spp=8192, static=True)
    mats = rt.material_table([{"kd": (200.0, 200.0, 200.0)}])
    scene.set_materials(mats, np.zeros(scene.n, np.int32))
    with pytest.raises(rt.RTError, match="materials"):
        accel.render_device(w.camera, w.lights, rgb, bounce=False, static=True)
