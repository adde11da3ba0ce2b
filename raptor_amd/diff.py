"""torch.autograd bridge over the differentiable frame of libraptor_rt.

render(accel, camera, lights_static, albedo, light_params, ...) returns the
image (H, W, 3) as a device tensor whose backward calls the HIP adjoint
(rt_render_adjoint). albedo is (n, 3) float32 on the device; light_params is
(L, 6) float32: position xyz and rgb on the 0-255 scale (the falloff d of each
light is taken from lights_static and is not differentiated). Hit points and
visibility are held at the forward pass's.
"""
import torch

from . import rt, workloads


class _RenderFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, albedo, light_params, accel, camera, falloff, bounce, spp, seed, max_record_bytes):
        accel.scene.set_albedo(albedo)
        lp = light_params.detach().to("cpu", torch.float32).numpy()
        lights = [workloads.Light(tuple(float(v) for v in lp[k, :3]), tuple(float(v) for v in lp[k, 3:6]),
                                  float(falloff[k])) for k in range(lp.shape[0])]
        img = torch.empty((camera.y_res, camera.x_res, 3), dtype=torch.float32, device=albedo.device)
        tape = accel.render_tape(camera, lights, img, bounce=bounce, spp=spp, seed=seed,
                                 max_record_bytes=max_record_bytes)
        ctx.tape = tape
        ctx.shapes = (albedo.shape, light_params.shape, albedo.device)
        return img

    @staticmethod
    def backward(ctx, grad_img):
        a_shape, l_shape, dev = ctx.shapes
        d_albedo = torch.zeros(a_shape, dtype=torch.float32, device=dev)
        d_lights = torch.zeros(l_shape, dtype=torch.float32, device=dev)
        ctx.tape.adjoint(grad_img.contiguous(), d_albedo, d_lights)
        ctx.tape.close()
        ctx.tape = None
        return d_albedo, d_lights, None, None, None, None, None, None, None


def render(accel, camera, lights_static, albedo, light_params, bounce=True, spp=1, seed=0, max_record_bytes=0):
    falloff = [l.d for l in lights_static]
    return _RenderFn.apply(albedo, light_params, accel, camera, falloff, bounce, spp, seed, max_record_bytes)


def light_params(lights, device="cuda:0"):
    """(L, 6) tensor of the lights' positions and rgb."""
    rows = [list(l.pos) + list(l.rgb) for l in lights]
    return torch.tensor(rows, dtype=torch.float32, device=device)
