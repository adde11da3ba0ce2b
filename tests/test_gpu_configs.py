"""The batched and differentiable configurations at their full sizes, against
the oracle (sampled rays and pixels keep the CPU side to seconds):

* C3 batched: the 64 scenes of 1k-2M triangles (terrain and random
  triangles) with 64^2-1024^2 incoherent surface rays each. For BVH4 and
  BVH8 the autotuned persistent launch (hipGraph) returns the hits of one
  launch per scene bit for bit, and sampled rays of every scene match the
  oracle's closest hits (ids bit-exact except exact-t ties, t/u/v
  bit-exact).
* C4 differentiable: the 3,993,138-triangle terrain at 1024^2 x 4 spp with
  four lights. The taped forward matches the oracle's radiance on a pixel
  stride; the adjoint of an L2 loss over those pixels matches the oracle's
  float64 CPU adjoint and float64 central differences on 256 parameters
  (232 albedo entries, all 24 light parameters) within 1e-3; path replay
  returns the stored-record gradients.
* C5 batched + differentiable: the configuration's largest scene (15.6M
  triangles, terrain) through the path megakernel at 512^2 x 2 spp with two
  bounces: radiance against the oracle on a pixel stride, the in-lane
  adjoint against the CPU adjoint and 64 float64 central differences.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from helpers import compare_hits, rel_err, tuv_exact
from oracle import adjoint as A

pytestmark = pytest.mark.gpu


def _hits_np(t, O):
    g = t.cpu().numpy()
    h = np.zeros(g.shape[0], O.HIT_DTYPE)
    h["t"], h["u"], h["v"] = g[:, 0], g[:, 1], g[:, 2]
    h["prim"] = g[:, 3].view(np.int32)
    return h


# ---------------------------------------------------------------- C3

def test_c3_batched_config_autotuned_matches_plain_and_oracle(gpu_ctx, rt, O):
    import torch

    from raptor_amd import workloads

    scenes = workloads.batched()  # 64 scenes, N log-uniform in [1e3, 2e6], batch sides in [64, 1024]
    ns = [sc.abc.shape[0] for sc in scenes]
    sides = [int(round(np.sqrt(sc.rays))) for sc in scenes]
    assert len(scenes) == 64 and max(ns) > 1.9e6 and min(ns) < 2e3 and max(sides) > 800 and min(sides) < 100
    gscenes = [rt.Scene(gpu_ctx, sc.abc, sc.albedo) for sc in scenes]
    rays = [rt.surface_rays(g, sc.seed, sc.rays) for g, sc in zip(gscenes, scenes)]
    sampled = {}  # width -> per-scene sampled hits
    picks = [np.linspace(0, sc.rays - 1, min(sc.rays, 3000)).astype(np.int64) for sc in scenes]
    for width in (4, 8):
        accels = [rt.Accel(g, width=width) for g in gscenes]
        scratch = [torch.empty((sc.rays, 4), device="cuda:0") for sc in scenes]
        params, ms = rt.Batch.autotune(gpu_ctx, accels, rays, scratch, reps=2)
        assert params["lds_stack"] in (4, 8, 16) and params["waves_per_cu"] in (8, 16, 32) and params["grab"] >= 1
        assert ms > 0
        hits_p = [torch.full((sc.rays, 4), -7.0, device="cuda:0") for sc in scenes]
        hits_g = [torch.full((sc.rays, 4), -7.0, device="cuda:0") for sc in scenes]
        plain = rt.Batch(gpu_ctx, accels, rays, hits_p, use_graph=False)
        plain.run(plain=True)
        tuned = rt.Batch(gpu_ctx, accels, rays, hits_g, use_graph=True, **params)
        for _ in range(2):
            tuned.run()
            torch.cuda.synchronize()
            for hp, hg in zip(hits_p, hits_g):
                assert torch.equal(hp.view(torch.int32), hg.view(torch.int32))
        sampled[width] = [_hits_np(h[torch.from_numpy(p).cuda()], O) for h, p in zip(hits_p, picks)]
        plain.close()
        tuned.close()
        for a in accels:
            a.close()
    # The oracle builds the reference's AAC tree of every scene (threads:
    # separate scenes share nothing) and traces the sampled rays.
    with ThreadPoolExecutor(8) as ex:
        oscenes = list(ex.map(lambda sc: O.OracleScene(sc.abc), scenes))
    checked = hits = 0
    for k, (sc, osc, r, p) in enumerate(zip(scenes, oscenes, rays, picks)):
        ref, _ = osc.trace_closest(r[torch.from_numpy(p).cuda()].cpu().numpy())
        rtuv = np.stack([ref["t"], ref["u"], ref["v"]], 1)
        for width in (4, 8):
            g = sampled[width][k]
            bad, ties = compare_hits(g, ref["prim"], rtuv)
            assert not bad.any(), f"scene {k} ({sc.kind}, {ns[k]} triangles), BVH{width}: {bad.sum()} ids differ"
            assert tuv_exact(g, rtuv, (ref["prim"] >= 0) & ~ties)
        checked += len(p)
        hits += int((ref["prim"] >= 0).sum())
        osc.close()
    assert checked > 100_000 and hits > 0.2 * checked  # many incoherent surface rays leave the scene


# ---------------------------------------------------------------- C4

def _compact(rec, sel, spp):
    idx = (sel[:, None] * spp + np.arange(spp)[None, :]).ravel()
    return {k: v[:, idx] for k, v in rec.items()}, idx


def test_c4_differentiable_config_forward_adjoint_and_finite_differences(gpu_ctx, rt, O):
    import torch

    from raptor_amd import workloads

    w = workloads.diff_terrain()  # 1414^2 grid, 1024^2 x 4 spp, 4 ring lights, falloff 1
    assert w.abc.shape[0] == 3_993_138 and w.spp == 4
    spp, npix = w.spp, w.camera.x_res * w.camera.y_res
    scene = rt.Scene(gpu_ctx, w.abc, w.albedo)
    accel = rt.Accel(scene)
    out = accel.render(w.camera, w.lights, bounce=True, spp=spp, seed=w.seed)
    img_dev = torch.empty((w.camera.y_res, w.camera.x_res, 3), dtype=torch.float32, device="cuda:0")
    tape = accel.render_tape(w.camera, w.lights, img_dev, bounce=True, spp=spp, seed=w.seed)
    img = img_dev.cpu().numpy().reshape(-1, 3)
    assert np.array_equal(img, out["rgb"].reshape(-1, 3))
    ref = O.OracleScene(w.abc).render(w.albedo, w.camera, w.lights, bounce=True, spp=spp, seed=w.seed,
                                      pixel_stride=97, nthreads=16, records=True)
    sel = np.flatnonzero(ref["prim"] != -2)
    assert len(sel) > 10000
    crec, idx = _compact(ref["records"], sel, spp)
    # Pixels with a sample whose GPU primary or bounce id differs from the
    # oracle's (exact-t ties: another path) leave the sampled set, so neither
    # adjoint sees them.
    diff = ((out["primary"]["prim"][idx] != crec["prim"][0]) | (out["bounce"]["prim"][idx] != crec["prim"][1]))
    ok = ~diff.reshape(-1, spp).any(1)
    assert ok.mean() > 0.995
    sel = sel[ok]
    crec, idx = _compact(ref["records"], sel, spp)
    e = rel_err(img[sel], ref["rgb"][sel])
    assert e.max() <= 1e-5, e.max()

    lights = [(l.pos, l.rgb, l.d) for l in w.lights]
    target = A.shade(crec, w.meta["target_albedo"], lights, spp, True)
    g = A.loss_grad(ref["rgb"][sel].astype(np.float64), target)
    g_full = np.zeros((npix, 3), np.float32)
    g_full[sel] = g
    g_dev = torch.from_numpy(g_full.reshape(w.camera.y_res, w.camera.x_res, 3)).cuda()

    def grads(t):
        d_alb = torch.zeros((w.abc.shape[0], 3), dtype=torch.float32, device="cuda:0")
        d_l = torch.zeros((len(w.lights), 6), dtype=torch.float32, device="cuda:0")
        t.adjoint(g_dev, d_alb, d_l)
        return d_alb.cpu().numpy().astype(np.float64), d_l.cpu().numpy().astype(np.float64)

    d_alb, d_l = grads(tape)
    info = tape.info()
    assert not info["replay"]
    # Path replay at this size: a budget of a fifth of the records.
    img_r = torch.empty_like(img_dev)
    tape_r = accel.render_tape(w.camera, w.lights, img_r, bounce=True, spp=spp, seed=w.seed,
                               max_record_bytes=info["record_bytes"] // 5)
    assert tape_r.info()["replay"]
    assert torch.equal(img_r.view(torch.int32), img_dev.view(torch.int32))
    d_alb_r, d_l_r = grads(tape_r)
    tape_r.close()
    tape.close()
    assert np.allclose(d_alb_r, d_alb, rtol=1e-5, atol=1e-7 * np.abs(d_alb).max())
    assert np.allclose(d_l_r, d_l, rtol=1e-5, atol=1e-7 * np.abs(d_l).max())

    # The CPU adjoint over the same records.
    r_alb, r_l = A.adjoint(crec, w.albedo, lights, spp, True, g)
    scale = np.abs(r_alb).max()
    err = np.abs(d_alb - r_alb)
    assert (err <= 1e-4 * np.abs(r_alb) + 1e-6 * scale).all(), err.max()
    assert (np.abs(r_alb) > 0).sum() > 5000
    assert (np.abs(d_l - r_l) <= 1e-3 * np.abs(r_l) + 1e-5 * np.abs(r_l).max()).all()

    # float64 central differences on 256 parameters.
    alb = w.albedo.astype(np.float64)

    def loss_at(a, ls):
        return A.loss(A.shade(crec, a, ls, spp, True), target)

    rng = np.random.default_rng(4)
    seen = np.unique(crec["prim"][crec["prim"] >= 0])
    n_alb = 256 - 6 * len(lights)
    got, fd = [], []
    for t, c in zip(rng.choice(seen, n_alb, replace=False), rng.integers(0, 3, n_alb)):
        ap, am = alb.copy(), alb.copy()
        ap[t, c] += 1e-4
        am[t, c] -= 1e-4
        fd.append((loss_at(ap, lights) - loss_at(am, lights)) / 2e-4)
        got.append(d_alb[t, c])
    lp = [list(map(float, l[0])) + list(map(float, l[1])) for l in lights]
    for l in range(len(lights)):
        for k in range(6):
            hk = 1e-5 if k < 3 else 1e-2

            def moved(dv):
                q = [row[:] for row in lp]
                q[l][k] += dv
                return [(tuple(q[i][:3]), tuple(q[i][3:]), lights[i][2]) for i in range(len(lights))]

            fd.append((loss_at(alb, moved(hk)) - loss_at(alb, moved(-hk))) / (2 * hk))
            got.append(d_l[l, k])
    got, fd = np.array(got), np.array(fd)
    assert len(got) == 256
    err = np.abs(got - fd)
    ok = err <= 1e-3 * np.abs(fd) + 1e-6 * np.abs(fd).max()
    assert ok.all(), f"max rel {(err / np.maximum(np.abs(fd), 1e-30)).max():.3g}"


# ---------------------------------------------------------------- C5

def test_c5_largest_scene_path_kernel_forward_and_gradients(gpu_ctx, rt, O):
    import torch

    from raptor_amd import workloads

    seed = 5
    (w,) = workloads.invrender_scenes(pick=[6])  # the configuration's largest scene
    assert w.abc.shape[0] > 15_000_000 and w.meta["kind"] == "terrain"
    scene = rt.Scene(gpu_ctx, w.abc, w.albedo)
    accel = rt.Accel(scene)
    n, shape = w.abc.shape[0], (w.camera.y_res, w.camera.x_res, 3)
    alb = torch.from_numpy(w.albedo.copy()).cuda()
    dalb = torch.zeros((n, 3), device="cuda:0")
    tgt = torch.zeros(shape, device="cuda:0")
    img = torch.zeros(shape, device="cuda:0")
    inv = rt.InvRender(gpu_ctx, [(accel, w.camera, alb, dalb, tgt, img)], w.lights, spp=2, bounces=2, seed=seed,
                       use_graph=False)
    inv.run(0)
    gimg = img.cpu().numpy().reshape(-1, 3).astype(np.float64)
    ref = O.OracleScene(w.abc).render(w.albedo, w.camera, w.lights, bounce=2, spp=2, seed=seed, pixel_stride=17,
                                      nthreads=16, records=True)
    sel = np.flatnonzero(ref["prim"] != -2)
    assert len(sel) > 10000
    e = rel_err(gimg[sel], ref["rgb"][sel])
    # Exact-t ties can pick another triangle (and another path) in rare
    # pixels; those leave the sampled set, so neither adjoint sees them.
    assert (e > 1e-5).mean() < 0.002, (e > 1e-5).sum()
    assert np.median(e) < 1e-7
    sel = sel[e.max(1) <= 1e-5] if e.ndim == 2 else sel[e <= 1e-5]
    crec, _ = _compact(ref["records"], sel, 2)
    lights = [(l.pos, l.rgb, l.d) for l in w.lights]
    rng = np.random.default_rng(7)
    target_albedo = np.clip(w.albedo * (1.0 + 0.2 * rng.standard_normal(w.albedo.shape)), 0.0, 1.0)
    t_sel = A.shade(crec, target_albedo, lights, 2, 2)
    # Target: the GPU image itself away from the sampled pixels (so they add
    # nothing to the gradient), the perturbed-albedo frame at them.
    t_full = gimg.copy()
    t_full[sel] = t_sel
    tgt.copy_(torch.from_numpy(t_full.astype(np.float32).reshape(shape)))
    loss, _, _ = inv.run(1)
    assert abs(loss[0] - A.loss(gimg[sel], t_sel)) <= 1e-4 * A.loss(gimg[sel], t_sel)
    d_alb = dalb.cpu().numpy().astype(np.float64)
    g = A.loss_grad(gimg[sel], t_sel)
    r_alb, _ = A.adjoint(crec, w.albedo, lights, 2, 2, g)
    scale = np.abs(r_alb).max()
    err = np.abs(d_alb - r_alb)
    assert (err <= 1e-4 * np.abs(r_alb) + 1e-6 * scale).all(), err.max()
    assert (np.abs(r_alb) > 0).sum() > 5000
    seen = np.unique(crec["prim"][crec["prim"] >= 0])
    alb64 = w.albedo.astype(np.float64)
    for t, c in zip(rng.choice(seen, 64, replace=False), rng.integers(0, 3, 64)):
        ap, am = alb64.copy(), alb64.copy()
        ap[t, c] += 1e-4
        am[t, c] -= 1e-4
        fd = (A.loss(A.shade(crec, ap, lights, 2, 2), t_sel) - A.loss(A.shade(crec, am, lights, 2, 2), t_sel)) / 2e-4
        assert abs(d_alb[t, c] - fd) <= 1e-3 * abs(fd) + 1e-6 * scale, (t, c, d_alb[t, c], fd)
    inv.close()
