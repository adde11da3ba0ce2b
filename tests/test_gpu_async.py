"""Frames enqueued without waiting (RT_ASYNC): each image and hit record
equals the same frame rendered synchronously, bit for bit, while several
frames are in flight on the context's stream; the device's ray sums equal
the synchronous frames' stats, also for a frame that runs in two chunks and
through rt_render_static; the entry points refuse what an asynchronous frame
cannot report."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _buffers(res, spp, n):
    import torch

    nout = res[0] * res[1] * spp
    return [(torch.full((res[1], res[0], 3), -1.0, dtype=torch.float32, device="cuda:0"),
             torch.full((nout, 4), -1.0, dtype=torch.float32, device="cuda:0"),
             torch.full((nout, 4), -1.0, dtype=torch.float32, device="cuda:0")) for _ in range(n)]


def _camera(w, res):
    cam = w.camera
    return type(cam)(cam.pos, cam.x_axis, cam.y_axis, cam.z_axis, cam.w, cam.w * res[1] / res[0], cam.t, res[0],
                     res[1])


@pytest.mark.parametrize("res,spp,static", [((640, 360), 1, False), ((1920, 1080), 5, False),
                                            ((800, 600), 2, True)])
def test_async_frames_equal_synchronous_ones(gpu_ctx, rt, res, spp, static):
    import torch

    from raptor_amd import workloads

    w = workloads.terrain(grid=128, x_res=res[0], y_res=res[1], bounce=True, seed=5)
    scene = rt.Scene(gpu_ctx, w.abc, w.albedo)
    accel = rt.Accel(scene)
    cam = _camera(w, res)
    seeds = [11, 12, 13, 11]
    sync_bufs, async_bufs = _buffers(res, spp, len(seeds)), _buffers(res, spp, len(seeds))
    want = [0, 0, 0, 0]
    for seed, (rgb, prim, bnc) in zip(seeds, sync_bufs):
        st = accel.render_device(cam, w.lights, rgb, bounce=True, spp=spp, seed=seed, primary_dev=prim,
                                 bounce_dev=bnc, static=static)
        want = [a + b for a, b in zip(want, st["rays"])]
    gpu_ctx.frame_rays()  # clears the sums
    for seed, (rgb, prim, bnc) in zip(seeds, async_bufs):
        assert accel.render_device(cam, w.lights, rgb, bounce=True, spp=spp, seed=seed, primary_dev=prim,
                                   bounce_dev=bnc, static=static, asynchronous=True) is None
    assert gpu_ctx.frame_rays() == want
    assert want[0] == len(seeds) * res[0] * res[1] * spp and want[3] > 0
    torch.cuda.synchronize()
    for a, b in zip(sync_bufs, async_bufs):
        for x, y in zip(a, b):
            assert torch.equal(x.view(torch.int32), y.view(torch.int32))
    assert gpu_ctx.frame_rays() == [0, 0, 0, 0]


def test_async_frames_refuse_what_they_cannot_report(gpu_ctx, rt):
    import torch

    from raptor_amd import workloads

    w = workloads.terrain(grid=32, x_res=16, y_res=16, bounce=True)
    scene = rt.Scene(gpu_ctx, w.abc, w.albedo)
    accel = rt.Accel(scene)
    lib = rt.lib()
    cam = w.camera.to_c()
    nl = len(w.lights)
    larr = (rt.rt_light * nl)(*[l.to_c() for l in w.lights])
    host_rgb = np.zeros((16, 16, 3), np.float32)
    dev_rgb = torch.zeros((16, 16, 3), dtype=torch.float32, device="cuda:0")
    st = rt.rt_stats()
    # Host outputs.
    p = rt.rt_render_params(1, 1, 0, rt.RT_ASYNC, 0)
    assert lib.rt_render(accel.h, C.byref(cam), larr, nl, C.byref(p), host_rgb.ctypes.data_as(C.c_void_p), None,
                         None, None) == rt.RT_ERR_INVALID
    # Stats would need the frame done.
    p = rt.rt_render_params(1, 1, 0, rt.RT_ASYNC | rt.RT_DEVICE_PTRS, 0)
    assert lib.rt_render(accel.h, C.byref(cam), larr, nl, C.byref(p), C.c_void_p(dev_rgb.data_ptr()), None, None,
                         C.byref(st)) == rt.RT_ERR_INVALID
    assert lib.rt_render_static(accel.h, C.byref(cam), larr, nl, C.byref(p), C.c_void_p(dev_rgb.data_ptr()), None,
                                None, C.byref(st)) == rt.RT_ERR_INVALID
    # Visit counting reads its counters back after the frame.
    gpu_ctx.set_counting(True)
    try:
        with pytest.raises(rt.RTError, match="counting"):
            accel.render_device(w.camera, w.lights, dev_rgb, bounce=True, asynchronous=True)
    finally:
        gpu_ctx.set_counting(False)
    # Material frames size their generations on the host.
    mats = rt.material_table([{"kd": (200.0, 200.0, 200.0)}])
    scene.set_materials(mats, np.zeros(scene.n, np.int32))
    with pytest.raises(rt.RTError, match="materials"):
        accel.render_device(w.camera, w.lights, dev_rgb, bounce=False, asynchronous=True)
    assert gpu_ctx.frame_rays() == [0, 0, 0, 0]
