#!/usr/bin/env python3
"""Throughput of the ray-tracing hot path on one MI355X.

Workload (one "step" = one frame): the 999,698-triangle fBm terrain (708 x 708
grid), 1920x1080 primary rays, one shadow ray per hit, one cosine-weighted
diffuse bounce per hit and one shadow ray per bounce hit, FP32. The scene and
its BVH are resident in HBM before the timed region; the BVH build time is
reported separately. value = rays traced (counted on the device) / wall time.

The CPU baseline is the oracle's restatement of the reference's own CPU
program for this frame, the SSE packet tracer (SIMD_PACKET_TRACING) over the
reference's AAC BVH, on the host cores of the same machine, timed on a bounded
sample of whole frames; the scalar tracer is timed beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s at 1M tris (primary+shadow+diffuse); batched & adjoint vs plain on GPU"
PEAK_HBM_GBS = 8000.0
STAGES = ["primary_closest", "shadow_anyhit", "bounce_closest", "bounce_shadow_anyhit"]
# Kernel of each trace stage: the bounce closest-hit launch uses the library's
# block-local form with ray refills, the others one lane per ray.
KERNELS = ["k_trace", "k_trace", "k_trace_pers", "k_trace"]


def algorithmic_bytes(width, rays, node_visits, tri_tests, closest):
    """Per-visit model of a launch's bytes: each ray in (32 B), each result
    out (16 B hit or 1 B flag), each node visit reads width child boxes (24 B)
    and child codes (4 B), each triangle test reads a, c-a, b-a (36 B) and the
    id (4 B), with no reuse through the caches."""
    return rays * (32 + (16 if closest else 1)) + node_visits * width * 28 + tri_tests * 40


def roofline_entry(args, dom, launch_ms, traffic, model_bytes, model_gbs, compulsory, counted):
    """The dominant trace launch against the HBM roofline. achieved / frac:
    the HBM bytes one launch moves, measured by rocprofv3 counters
    (FETCH_SIZE x 2 + WRITE_SIZE x 1 per dispatch, both scales from a copy
    kernel of known size; profiles/trace_terrain1m.json), over the launch
    time measured here with HIP events on the launch's stream. Two labelled
    models sit beside it: the per-visit model prices every node and triangle
    fetch as HBM traffic (no cache reuse; the tree and triangles, 81 MB, live
    in L2 and the 256 MB Infinity Cache, so it exceeds the traffic ~4x and is
    no bound), and the compulsory bytes (rays, results and the tree once) are
    the floor of the launch's traffic. The kernel is limited by VALU issue
    under divergence, not by bandwidth."""
    t_s = launch_ms * 1e-3
    entry = {
        "bound": "hbm",
        "kernel": f"{KERNELS[dom]}<{args.width}, {dom}, false, {args.lds_stack}> ({STAGES[dom]})",
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "traffic": traffic,
        "launch_ms": round(launch_ms, 4),
        "limiter": "VALU issue under divergence (VALUBusy / VALUUtilization: profiles/counters_terrain1m.md)",
        "visit_model": {
            "bytes_per_launch": int(model_bytes),
            "gb_s": round(model_gbs, 2),
            "frac": round(model_gbs / PEAK_HBM_GBS, 4),
            "node_visits_per_ray": round(counted["visits"][2 * dom] / max(1, counted["rays"][dom]), 3),
            "tri_tests_per_ray": round(counted["visits"][2 * dom + 1] / max(1, counted["rays"][dom]), 3),
            "note": "every node and triangle fetch priced as HBM traffic (no cache reuse): an upper bound on "
                    "demand, not a roofline",
        },
        "compulsory": {
            "bytes_per_launch": int(compulsory),
            "frac": round(compulsory / t_s / 1e9 / PEAK_HBM_GBS, 4),
        },
    }
    if traffic:
        achieved = traffic / t_s / 1e9
        entry.update(achieved=round(achieved, 2), frac=round(achieved / PEAK_HBM_GBS, 4),
                     achieved_source="rocprofv3 counter traffic per dispatch / HIP-event launch time")
    else:
        entry.update(achieved=round(model_gbs, 2), frac=round(model_gbs / PEAK_HBM_GBS, 4),
                     achieved_source="visit model (no counter profile matches this configuration)")
    return entry


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, target_s, gpu_hits=None, gpu_frame_ms=None):
    """The reference's own CPU program for this frame, restated in the oracle:
    the SSE packet tracer (SIMD_PACKET_TRACING: 8x8 tiles of 16 four-ray
    packets, packet traversal and rcp-based Moller-Trumbore, shadow rays
    packed four at a time) over the reference's AAC BVH, (b) on all host
    cores with a tile work queue for about target_s seconds of whole frames
    and (a) on one core over a sample of tiles. The scalar IEEE tracer (the
    parity oracle) is timed beside it for one frame on all cores; its
    traversal counters price the frame with the per-ray bytes 32 (ray) + 16
    (closest hit) or 1 (any hit) + 32 per child-box test + 36 per triangle
    test on the reference's tree, which against the GPU frame time is the
    reference-tree demand model (it exceeds HBM peak: most of those reads
    are cache hits, so it is no roofline)."""
    import numpy as np

    import oracle.oracle as O

    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    t0 = time.perf_counter()
    scene = O.OracleScene(w.abc)
    build_s = time.perf_counter() - t0

    frames_out = {}

    def run(stride, threads, packet, keep=None):
        t = time.perf_counter()
        r = scene.render(w.albedo, w.camera, w.lights, bounce=w.bounce, seed=w.seed, pixel_stride=stride,
                         nthreads=threads, packet=packet)
        dt = time.perf_counter() - t
        if keep:
            frames_out[keep] = r
        return int(r["counters"][:4].sum()), dt

    total_rays, total_s, frames = 0, 0.0, 0
    while total_s < target_s or frames == 0:
        rays, dt = run(1, cores, True, keep="packet")
        total_rays += rays
        total_s += dt
        frames += 1
    rate1 = total_rays / total_s / cores
    tiles = ((w.camera.x_res + 7) // 8) * ((w.camera.y_res + 7) // 8)
    per_tile = total_rays / frames / tiles
    stride1 = max(1, int(tiles * per_tile / max(rate1 * target_s / 4, 1.0)))
    rays1, dt1 = run(stride1, 1, True)
    srays, sdt = run(1, cores, False, keep="scalar")
    # The packet path (rcp + one Newton step for 1/det, no |det| or t > EPSILON
    # test, rt/triangle.h:242-305) against the scalar IEEE path and the GPU:
    # primary and bounce hit ids that differ, per frame.
    pk, sc = frames_out["packet"], frames_out["scalar"]
    disagree = {
        "primary_packet_vs_scalar": int((pk["prim"] != sc["prim"]).sum()),
        "bounce_packet_vs_scalar": int((pk["bounce_prim"] != sc["bounce_prim"]).sum()),
        "pixels": int(pk["prim"].size),
    }
    cnt = sc["counters"].astype(np.int64)
    ref_bytes = (32 * int(cnt[:4].sum()) + 16 * int(cnt[0] + cnt[2]) + int(cnt[1] + cnt[3]) + 32 * int(cnt[4])
                 + 36 * int(cnt[5]))
    ref_model = {"box_tests_per_ray": float(cnt[4]) / float(cnt[:4].sum()),
                 "tri_tests_per_ray": float(cnt[5]) / float(cnt[:4].sum()),
                 "bytes_per_frame": ref_bytes}
    if gpu_frame_ms:
        ref_model["gb_s_at_gpu_frame_time"] = round(ref_bytes / (gpu_frame_ms * 1e-3) / 1e9, 1)
        ref_model["frac_of_hbm_peak"] = round(ref_bytes / (gpu_frame_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 3)
    if gpu_hits is not None:
        disagree["primary_packet_vs_gpu"] = int((pk["prim"] != gpu_hits[0]).sum())
        disagree["bounce_packet_vs_gpu"] = int((pk["bounce_prim"] != gpu_hits[1]).sum())
        disagree["primary_scalar_vs_gpu"] = int((sc["prim"] != gpu_hits[0]).sum())
        disagree["bounce_scalar_vs_gpu"] = int((sc["bounce_prim"] != gpu_hits[1]).sum())
    return {
        "value": total_rays / total_s / 1e6,
        "unit": "Mrays/s",
        "cores": cores,
        "kind": "port",
        "cpu_model": cpu_model(),
        "sample": f"{frames} whole 1920x1080 frames ({total_rays} rays, {total_s:.1f} s) on {cores} threads "
                  f"({cpu_model()}); the oracle's restatement of the reference's SSE packet tracer over the "
                  f"reference's AAC BVH (built in {build_s:.2f} s), whose traversal keeps the reference's "
                  f"absolute 100*EPSILON slack (ssd/bvh.cc:620,627,705)",
        "single_thread_value": rays1 / dt1 / 1e6,
        "single_thread_sample": f"every {stride1}th 8x8 tile ({rays1} rays, {dt1:.1f} s) on 1 thread",
        "scalar_value": srays / sdt / 1e6,
        "scalar_sample": f"one frame of the scalar IEEE tracer on {cores} threads ({srays} rays, {sdt:.1f} s)",
        "id_disagreements": disagree,
        "reference_tree_model": ref_model,
    }


def batched_section(ctx, n_scenes, reps):
    """Batched workload: n_scenes independent scenes (triangle counts
    log-uniform in [1e3, 2e6], terrain and random triangles 50/50), each with
    side x side incoherent surface rays (side log-uniform in [64, 1024]),
    closest-hit queries. The library's autotuner picks node width (BVH4 or
    BVH8, building the other from the scenes), LDS stack depth, resident
    wavefronts per CU, wavefronts per workgroup, the refill threshold and
    chunks per grab; the tuned persistent launch (hipGraph replay) is then
    timed against one plain launch per scene at the faster width for those,
    median of reps frames each."""
    import statistics

    import torch

    from raptor_amd import rt, workloads

    t0 = time.perf_counter()
    specs = workloads.batched(n_scenes=n_scenes)
    gen_s = time.perf_counter() - t0
    scenes, rays, hits_a, hits_b = [], [], [], []
    for sp in specs:
        sc = rt.Scene(ctx, sp.abc, sp.albedo)
        scenes.append(sc)
        rays.append(rt.surface_rays(sc, sp.seed, sp.rays))
        hits_a.append(torch.empty((sp.rays, 4), dtype=torch.float32, device="cuda:0"))
        hits_b.append(torch.empty((sp.rays, 4), dtype=torch.float32, device="cuda:0"))
    total_rays = sum(sp.rays for sp in specs)
    total_tris = sum(int(sp.abc.shape[0]) for sp in specs)
    del specs

    def timed(batch, plain):
        for _ in range(2):
            batch.run(plain=plain)
        return statistics.median(batch.run(plain=plain) for _ in range(reps))

    # The library's autotuner picks the node width too (it builds the other
    # width's BVHs from the scenes); both widths' per-scene launches are timed
    # and the faster is the baseline.
    per_width, by_width = {}, {}
    for w in (4, 8):
        by_width[w] = [rt.Accel(sc, width=w, leaf_size=2, lds_stack=8) for sc in scenes]
        plain = rt.Batch(ctx, by_width[w], rays, hits_b, use_graph=False)
        per_width[f"bvh{w}"] = {"build_ms_total": round(sum(a.info()["build_ms"] for a in by_width[w]), 2),
                                "plain_ms": round(timed(plain, True), 4)}
        plain.close()
    params, tuned_ms = rt.Batch.autotune(ctx, by_width[4], rays, hits_a, reps=3)
    width = params["width"]
    accels = by_width[width]
    plain_width = min((4, 8), key=lambda w: per_width[f"bvh{w}"]["plain_ms"])
    pers = rt.Batch(ctx, accels, rays, hits_a, use_graph=True, **params)
    plain = rt.Batch(ctx, by_width[plain_width], rays, hits_b, use_graph=False)
    pers_ms = timed(pers, False)
    plain_ms = timed(plain, True)
    same = all(torch.equal(a.view(torch.int32), b.view(torch.int32)) for a, b in zip(hits_a, hits_b))
    pers.close()
    plain.close()
    return {
        "workload": f"{n_scenes} scenes, {total_tris} triangles, {total_rays} incoherent surface rays "
                    "(closest hit), one frame = every scene's batch",
        "generate_s": round(gen_s, 2),
        "width": width,
        "params": params,
        "autotune_ms": round(tuned_ms, 4),
        "plain_width": plain_width,
        "persistent_graph_ms": round(pers_ms, 4),
        "plain_per_scene_ms": round(plain_ms, 4),
        "persistent_mrays_s": round(total_rays / pers_ms / 1e3, 2),
        "plain_mrays_s": round(total_rays / plain_ms / 1e3, 2),
        "speedup": round(plain_ms / pers_ms, 3),
        "hits_identical": same,
        "per_width": per_width,
    }


def adjoint_section(ctx, reps, grid=1414, res=1024, spp=4):
    """Differentiable frame: grid x grid fBm terrain, res x res at spp samples,
    direct light + one diffuse bounce, four point lights (falloff 1). Loss =
    sum (I - I_target)^2 against the frame rendered with perturbed albedo. The
    forward (taped) frame and the adjoint are timed with device events, median
    of reps; once with stored path records and once with path replay forced by
    a record budget of a quarter of the frame's records."""
    import statistics

    import torch

    from raptor_amd import rt, workloads

    w = workloads.diff_terrain(grid=grid, x_res=res, y_res=res, spp=spp)
    scene = rt.Scene(ctx, w.abc, w.albedo)
    accel = rt.Accel(scene, width=4, leaf_size=2, lds_stack=8)
    n = int(w.abc.shape[0])
    target_alb = w.meta["target_albedo"]
    del w.abc
    img = torch.empty((res, res, 3), dtype=torch.float32, device="cuda:0")
    tgt = torch.empty_like(img)
    scene.set_albedo(target_alb)
    accel.render_device(w.camera, w.lights, tgt, bounce=True, spp=spp, seed=w.seed)
    scene.set_albedo(w.albedo)
    d_alb = torch.zeros((n, 3), dtype=torch.float32, device="cuda:0")
    d_l = torch.zeros((len(w.lights), 6), dtype=torch.float32, device="cuda:0")

    def run(budget):
        fwd, bwd, rays, info = [], [], 0, None
        for k in range(reps + 2):
            tape = accel.render_tape(w.camera, w.lights, img, bounce=True, spp=spp, seed=w.seed,
                                     max_record_bytes=budget)
            g = 2.0 * (img - tgt)
            d_alb.zero_()
            d_l.zero_()
            torch.cuda.synchronize()
            ms_b = tape.adjoint(g, d_alb, d_l)
            if k >= 2:
                fwd.append(tape.stats["ms"][0])
                bwd.append(ms_b)
            rays = sum(tape.stats["rays"])
            info = tape.info()
            tape.close()
        return statistics.median(fwd), statistics.median(bwd), rays, info

    f_ms, b_ms, rays, info = run(0)
    loss = float(((img - tgt) ** 2).sum())
    grad_norm = float(d_alb.norm())
    rf_ms, rb_ms, _, rinfo = run(info["record_bytes"] // 4)
    chunks = -(-((res + 7) // 8) ** 2 // rinfo["chunk_tiles"])
    return {
        "workload": f"{n}-triangle fBm terrain ({grid}x{grid}), {res}x{res} at {spp} spp, direct + 1 bounce, "
                    f"4 point lights (falloff 1); loss = sum (I - I_target)^2, target albedo * (1 + 0.2 N(0,1))",
        "rays_per_frame": rays,
        "forward_ms": round(f_ms, 4),
        "adjoint_ms": round(b_ms, 4),
        "adjoint_over_forward": round(b_ms / f_ms, 4),
        "forward_mrays_s": round(rays / f_ms / 1e3, 2),
        "record_bytes": info["record_bytes"],
        "replay": {"record_budget": info["record_bytes"] // 4, "chunks": chunks, "forward_ms": round(rf_ms, 4),
                   "adjoint_ms": round(rb_ms, 4), "adjoint_over_forward": round(rb_ms / rf_ms, 4),
                   "adjoint_mrays_s": round(rays / rb_ms / 1e3, 2)},
        "loss": loss,
        "albedo_grad_norm": grad_norm,
    }


def invrender_section(ctx, steps, n_scenes=16, res=512, spp=2, bounces=2):
    """Batched inverse rendering: n_scenes scenes (0.5M-16M triangles,
    terrain / random 50/50), res x res at spp samples, `bounces` diffuse
    bounces, four lights. Targets are rendered from the generated albedo
    (ground truth); the albedo starts at 0.5 and `steps` Adam steps (lr 0.02)
    run, each one persistent path launch (forward + in-lane adjoint) plus the
    update, replayed from a hipGraph. Also timed (median of 5): the forward
    frames alone, forward + adjoint, and the same as one launch per scene."""
    import numpy as np
    import torch

    from raptor_amd import rt, workloads

    t0 = time.perf_counter()
    ws = workloads.invrender_scenes(n_scenes=n_scenes, res=res)
    gen_s = time.perf_counter() - t0
    dev = "cuda:0"
    scenes, accels, items, truth = [], [], [], []
    build_ms = 0.0
    for w in ws:
        sc = rt.Scene(ctx, w.abc, w.albedo)
        acc = rt.Accel(sc)
        build_ms += acc.info()["build_ms"]
        n = int(w.abc.shape[0])
        gt = torch.from_numpy(w.albedo).to(dev)
        alb = gt.clone()
        shape = (res, res, 3)
        items.append((acc, w.camera, alb, torch.zeros((n, 3), device=dev), torch.zeros(shape, device=dev),
                      torch.zeros(shape, device=dev)))
        scenes.append(sc)
        accels.append(acc)
        truth.append(gt)
        w.abc = None
    tris = sum(int(t.shape[0]) for t in truth)
    inv = rt.InvRender(ctx, items, ws[0].lights, spp=spp, bounces=bounces, seed=ws[0].seed, lr=0.02)
    # Targets from the ground truth.
    inv.run(0)
    for it in items:
        it[4].copy_(it[5])
    for it in items:
        it[2].fill_(0.5)

    def median_run(mode, plain=False, reps=5):
        runs = [inv.run(mode, plain=plain) for _ in range(reps)]
        return runs[0][1], sorted(r[2] for r in runs)[reps // 2]

    rays_f, ms_f = median_run(0)
    rays_fb, ms_fb = median_run(1)
    _, ms_plain = median_run(1, plain=True)
    seen = [it[3].abs().sum(1) > 0 for it in items]  # triangles with a gradient at the start
    for it in items:
        it[3].zero_()
    loss0, _, _ = inv.run(2, steps=1)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    loss1, rays_s, ms_s = inv.run(2, steps=steps - 1)
    wall = time.perf_counter() - t1

    def rmse(mask_list):
        num, den = 0.0, 0
        for it, gt, m in zip(items, truth, mask_list):
            d = (it[2] - gt)[m] if m is not None else (it[2] - gt)
            num += float((d.double() ** 2).sum())
            den += d.numel()
        return (num / max(den, 1)) ** 0.5

    init_rmse = (sum(float(((0.5 - gt).double() ** 2).sum()) for gt in truth) / (3 * tris)) ** 0.5
    # Identifiability: the loss constrains a triangle's albedo through the
    # path vertices that land on it. The same samples' primary and first
    # bounce vertices (rt_render with the same seed, spp and bounce draws
    # as the path launch) count them per triangle.
    well = []
    verts_seen, n_seen = 0, 0
    for acc, w, m in zip(accels, ws, seen):
        out = acc.render(w.camera, ws[0].lights, bounce=True, spp=spp, seed=ws[0].seed, stats=False)
        ids = np.concatenate([out["primary"]["prim"], out["bounce"]["prim"]])
        cnt = np.bincount(ids[ids >= 0], minlength=int(acc.scene.n))
        c = torch.from_numpy(cnt).to(dev)
        verts_seen += int(c[m].sum())
        n_seen += int(m.sum())
        well.append(c >= 4)
    res_out = {
        "workload": f"{n_scenes} scenes, {tris} triangles (terrain and random 50/50, 0.5M-16M each), {res}x{res} "
                    f"at {spp} spp, {bounces} diffuse bounces, 4 point lights; loss sum (I - I_target)^2, targets "
                    f"from the generated albedo, Adam lr 0.02 from albedo 0.5",
        "generate_s": round(gen_s, 1),
        "build_ms_total": round(build_ms, 1),
        "forward_ms": round(ms_f, 3),
        "forward_mrays_s": round(rays_f / ms_f / 1e3, 1),
        "forward_adjoint_ms": round(ms_fb, 3),
        "forward_adjoint_mrays_s": round(rays_fb / ms_fb / 1e3, 1),
        "adjoint_over_forward": round((ms_fb - ms_f) / ms_f, 4),
        "plain_per_scene_forward_adjoint_ms": round(ms_plain, 3),
        "persistent_speedup": round(ms_plain / ms_fb, 3),
        "steps": steps,
        "step_ms": round(ms_s / (steps - 1), 3),
        "steps_wall_s": round(wall, 2),
        "loss_first_step": sum(loss0),
        "loss_last_step": sum(loss1),
        "albedo_rmse_start": round(init_rmse, 5),
        "albedo_rmse_all": round(rmse([None] * len(items)), 5),
        "albedo_rmse_seen": round(rmse(seen), 5),
        "seen_triangles": int(sum(int(m.sum()) for m in seen)),
        "albedo_rmse_seen_start": round((sum(float(((0.5 - gt)[m].double() ** 2).sum()) for gt, m in zip(truth, seen))
                                         / max(1, 3 * sum(int(m.sum()) for m in seen))) ** 0.5, 5),
        "path_record_bytes": 0,
        "path_vertices_per_seen_triangle": round(verts_seen / max(1, n_seen), 2),
        "well_sampled_rule": "at least 4 primary or first-bounce vertices of the step's samples on the triangle",
        "well_sampled_triangles": int(sum(int(m.sum()) for m in well)),
        "albedo_rmse_well_sampled_start": round((sum(float(((0.5 - gt)[m].double() ** 2).sum())
                                                     for gt, m in zip(truth, well))
                                                 / max(1, 3 * sum(int(m.sum()) for m in well))) ** 0.5, 5),
        "albedo_rmse_well_sampled": round(rmse(well), 5),
    }
    inv.close()
    return res_out


def aac_section(ctx, w, scene, accel, reps):
    """The reference builder's tree (rt_build_aac) on the headline terrain:
    GPU build time (median of reps), and synchronous frames on the accel
    built over it (RT_BUILDER_AAC) against the default tree's."""
    import numpy as np

    from raptor_amd import rt

    builds = [scene.build_aac()["build_ms"] for _ in range(reps)]

    def frames(acc):
        for _ in range(2):
            acc.render(w.camera, w.lights, bounce=True, seed=w.seed, want_hits=False)
        runs = [acc.render(w.camera, w.lights, bounce=True, seed=w.seed, want_hits=False) for _ in range(reps + 2)]
        st = np.median(np.array([r["stats"]["ms"] for r in runs]), axis=0)
        ctx.set_counting(True)
        c = acc.render(w.camera, w.lights, bounce=True, seed=w.seed, want_hits=False)["stats"]
        ctx.set_counting(False)
        prim = max(1, c["rays"][0])
        return {"frame_gpu_ms": round(float(st[0]), 4), "primary_ms": round(float(st[1]), 4),
                "shadow_ms": round(float(st[2]), 4), "bounce_ms": round(float(st[3]), 4),
                "node_visits_per_primary_ray": round(c["visits"][0] / prim, 2),
                "tri_tests_per_primary_ray": round(c["visits"][1] / prim, 2)}

    acc = rt.Accel(scene, builder=rt.BUILDER_AAC)
    info = acc.info()
    out = {
        "workload": "the headline terrain's AAC tree (alpha 0.4, delta 20), built on the GPU node for node as "
                    "ssd/bvh_builder.cc builds it, and frames traced on it",
        "build_ms": round(sorted(builds)[reps // 2], 1),
        "accel_build_ms": round(info["build_ms"], 1),
        "wide_nodes": info["nodes"], "leaves": info["leaves"], "depth": info["depth"],
        "on_aac_tree": frames(acc),
        "on_default_tree": frames(accel),
    }
    acc.close()
    return out


def dynamic_section(ctx, w, accel, scene, frames, args):
    """Per-frame update + refit + frame for a moving terrain (a travelling wave
    of amplitude 0.01, one phase step per frame), against rebuilding the
    accelerator every frame as the reference's physics render loop does
    (RAPTor/physics_engine/simulation_environment.cc:53-58). Vertex positions
    for every frame are made on the device beforehand (the simulation's part)."""
    import torch

    from raptor_amd import rt, workloads

    dev = torch.from_numpy(w.abc).to("cuda:0")
    pos = [workloads.terrain_wave(dev, k / frames, amplitude=0.01).contiguous() for k in range(frames)]
    rgb = torch.empty((w.camera.y_res, w.camera.x_res, 3), dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()

    def loop(rebuild):  # rebuild: None refits, else the sah_bins of a new build
        upd = fit = frame = gpu = 0.0
        rays = 0
        for k in range(frames):
            t0 = time.perf_counter()
            scene.update(pos[k])
            t1 = time.perf_counter()
            if rebuild is not None:
                a = rt.Accel(scene, width=args.width, leaf_size=args.leaf_size, lds_stack=args.lds_stack,
                             sah_bins=rebuild)
            else:
                accel.refit()
                a = accel
            t2 = time.perf_counter()
            st = a.render_device(w.camera, w.lights, rgb, bounce=w.bounce, seed=w.seed)
            torch.cuda.synchronize()
            t3 = time.perf_counter()
            if rebuild is not None:
                a.close()
            upd += t1 - t0
            fit += t2 - t1
            frame += t3 - t2
            gpu += st["ms"][0] * 1e-3  # event-timed stages (no workspace allocation)
            rays += sum(st["rays"])
        n = frames
        return {"update_ms": round(upd / n * 1e3, 3), "refit_or_build_ms": round(fit / n * 1e3, 3),
                "frame_ms": round(frame / n * 1e3, 3), "total_ms": round((upd + fit + frame) / n * 1e3, 3),
                "frame_gpu_ms": round(gpu / n * 1e3, 3), "trace_mrays_s": round(rays / gpu / 1e6, 1)}

    loop(None)  # warm-up
    refit = loop(None)
    lbvh = loop(0)
    sah = loop(args.sah_bins) if args.sah_bins > 0 else None
    scene.update(dev)  # back to the rest pose for what follows
    accel.refit()
    torch.cuda.synchronize()
    return {"workload": f"terrain1m under a travelling wave (amplitude 0.01, wavelength 0.25), {frames} frames, "
                        "update + refit + frame vs update + rebuild + frame",
            "refit": refit, "rebuild_lbvh": lbvh, "rebuild_sah": sah}


def post_section(ctx, rgb, reps, cpu_baseline=True):
    """Camera post-process of the 1920x1080 headline frame on the device:
    tone map (global_non_linear, key 0.18), sRGB gamma, clip to bytes, as a
    RAPTor driver does before writing the image (rt/camera.cc:1124-1501,
    rt/camera.h:289-351). Algorithmic HBM bytes per pixel: tone map 2 passes
    x (12 read + 12 written), gamma 12 + 12, clip 12 read + 3 written = 87.
    The oracle's single-threaded pipeline on the same frame is timed beside
    it."""
    import torch

    from raptor_amd import rt

    h, w = rgb.shape[0], rgb.shape[1]
    src = rgb.clone()
    work = torch.empty_like(src)
    out = torch.empty((h, w, 3), dtype=torch.uint8, device=src.device)
    tone = gam = clp = 0.0
    for k in range(reps + 1):
        work.copy_(src)
        torch.cuda.synchronize()
        _, _, _, t0 = rt.tone_map(ctx, work, mode=rt.TONE_NON_LINEAR, key=0.18)
        _, t1 = rt.gamma_correct(ctx, work, 2.2)
        _, t2 = rt.clip(ctx, work, 1, 1, out=out)
        if k > 0:  # the first round warms up
            tone, gam, clp = tone + t0, gam + t1, clp + t2
    tone, gam, clp = tone / reps, gam / reps, clp / reps
    total = tone + gam + clp
    # Glare (gf): the same tone map plus perform_glare_filter; the first call
    # builds the frame's glare filter and its spectrum, later calls reuse them.
    work.copy_(src)
    torch.cuda.synchronize()
    t = time.perf_counter()
    rt.tone_map(ctx, work, mode=rt.TONE_NON_LINEAR, key=0.18, glare=True, x_angle=1.0)
    glare_build_ms = (time.perf_counter() - t) * 1e3
    glare = []
    for k in range(reps):
        work.copy_(src)
        torch.cuda.synchronize()
        glare.append(rt.tone_map(ctx, work, mode=rt.TONE_NON_LINEAR, key=0.18, glare=True, x_angle=1.0)[3])
    bil, hist = [], []
    for k in range(reps + 1):  # the local tone mappers: bilateral filter, histogram adjustment
        work.copy_(src)
        torch.cuda.synchronize()
        t = rt.tone_map(ctx, work, mode=8)[3]
        work.copy_(src)
        torch.cuda.synchronize()
        # the frame's field of view: half width 0.5 at distance 1, 16:9
        t2 = rt.tone_map(ctx, work, mode=3, x_angle=1.0, y_angle=h / w)[3]
        if k > 0:
            bil.append(t)
            hist.append(t2)
    nbytes = 87 * h * w
    res = {"workload": f"{w}x{h} frame: tone map (global_non_linear) + gamma 2.2 + clip to bytes",
           "tone_map_ms": round(tone, 4), "gamma_ms": round(gam, 4), "clip_ms": round(clp, 4),
           "total_ms": round(total, 4), "algorithmic_bytes": nbytes,
           "bilateral_tone_map_ms": round(sorted(bil)[len(bil) // 2], 4),
           "human_histogram_tone_map_ms": round(sorted(hist)[len(hist) // 2], 4),
           "glare_tone_map_ms": round(sorted(glare)[len(glare) // 2], 4),
           "glare_first_call_wall_ms": round(glare_build_ms, 2),
           "achieved_gb_s": round(nbytes / (total * 1e-3) / 1e9, 1), "peak_gb_s": PEAK_HBM_GBS}
    if not cpu_baseline:
        return res
    try:
        from oracle import oracle as O  # the CPU baseline only

        img = src.cpu().numpy()
        t = time.perf_counter()
        o, _, _ = O.tone_map(img, mode=4, key=0.18)
        O.clip(O.gamma_correct(o, 2.2), 1, 1)
        res["cpu_baseline_ms"] = round((time.perf_counter() - t) * 1e3, 2)
        res["cpu_baseline_cores"] = 1
    except Exception as e:  # the oracle is test infrastructure; its absence only drops the baseline
        res["cpu_baseline_ms"] = None
        res["cpu_baseline_error"] = str(e)[:120]
    return res


def materials_section(ctx, frames, args):
    """Frames of the terrain with phong materials (a mirror wall, a glass pane,
    wet rock): primary rays, one shadow ray per hit, and mirror and refraction
    rays traced while their power stays above 0.1, each with its shadow ray.
    Rays per second count every ray traced; the wall time includes the
    per-generation read-back of the ray counts."""
    import torch

    from raptor_amd import rt, workloads

    w, mats, kinds, flags = workloads.material_terrain()
    scene = rt.Scene(ctx, w.abc, w.albedo, flags)
    scene.set_materials(mats, kinds)
    accel = rt.Accel(scene, width=args.width, leaf_size=args.leaf_size, lds_stack=args.lds_stack,
                     sah_bins=args.sah_bins)
    rgb = torch.empty((w.camera.y_res, w.camera.x_res, 3), dtype=torch.float32, device="cuda:0")
    accel.render_device(w.camera, w.lights, rgb)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    per = [accel.render_device(w.camera, w.lights, rgb) for _ in range(frames)]
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / frames
    rays = per[-1]["rays"]
    gpu_ms = sum(s["ms"][0] for s in per) / frames
    out = {
        "workload": "terrain1m with phong materials (4 terrain materials incl. wet rock rf 0.25, mirror wall rf 0.8, "
                    "glass pane tran 0.6 ri 1.5), 1920x1080, 1 light, recursion to MIN_REFLECTIVE_POWER",
        "frames": frames,
        "rays_per_frame": {"primary": rays[0], "shadow": rays[1], "secondary": rays[2], "secondary_shadow": rays[3]},
        "frame_ms": round(wall * 1e3, 3),
        "frame_gpu_ms": round(gpu_ms, 3),
        "closest_hit_ms": round(sum(s["ms"][1] for s in per) / frames, 3),
        "any_hit_ms": round(sum(s["ms"][2] for s in per) / frames, 3),
        "mrays_s": round(sum(rays) / wall / 1e6, 1),
    }
    accel.close()
    scene.close()
    # The terrain generator's textured terrain: one coloured_mapper_shader
    # whose kd is a tiled planar map (primary and shadow rays only).
    w, mats, kinds, flags, maps, images = workloads.textured_terrain()
    scene = rt.Scene(ctx, w.abc, w.albedo, flags)
    scene.set_textures(maps, images)
    scene.set_materials(mats, kinds)
    accel = rt.Accel(scene, width=args.width, leaf_size=args.leaf_size, lds_stack=args.lds_stack,
                     sah_bins=args.sah_bins)
    accel.render_device(w.camera, w.lights, rgb)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    per = [accel.render_device(w.camera, w.lights, rgb) for _ in range(frames)]
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / frames
    rays = per[-1]["rays"]
    out["textured"] = {
        "workload": "terrain1m shaded as the reference's terrain generator: coloured_mapper_shader, white kd "
                    "through a planar tile/tile map of a 256x256 image every 0.1 units, 1920x1080, 1 light",
        "rays_per_frame": {"primary": rays[0], "shadow": rays[1]},
        "frame_ms": round(wall * 1e3, 3),
        "frame_gpu_ms": round(sum(s["ms"][0] for s in per) / frames, 3),
        "mrays_s": round(sum(rays) / wall / 1e6, 1),
    }
    accel.close()
    scene.close()
    return out


def ingest_section(ctx, w, args, reps=3):
    """Scene ingestion of the 1M-triangle terrain from a model file: the
    terrain as a binary big-endian PLY (shared vertices, one triangle per
    face, as scanned models ship), parsed on the host (rt_mesh_load), uploaded
    with its material (rt_scene_create_from_mesh), built and rendered. The
    parsed triangles must give the generated scene's primary hits bit for
    bit. Host timings are medians of reps."""
    import tempfile

    import numpy as np
    import torch

    from raptor_amd import rt

    v = np.ascontiguousarray(np.asarray(w.abc, np.float32).reshape(-1, 3))
    keys, inv = np.unique(v.view(np.dtype((np.void, 12))).reshape(-1), return_inverse=True)
    verts = keys.view(np.float32).reshape(-1, 3)
    faces = inv.reshape(-1, 3).astype(">i4")
    nf = faces.shape[0]
    head = ("ply\nformat binary_big_endian 1.0\nelement vertex %d\nproperty float x\nproperty float y\n"
            "property float z\nelement face %d\nproperty list uchar int vertex_indices\nend_header\n"
            % (verts.shape[0], nf)).encode()
    body = np.empty((nf, 13), np.uint8)
    body[:, 0] = 3
    body[:, 1:] = faces.view(np.uint8).reshape(nf, 12)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "terrain1m.ply")
        with open(path, "wb") as fh:
            fh.write(head + verts.astype(">f4").tobytes() + body.tobytes())
        size = os.path.getsize(path)
        parse, upload = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            mesh = rt.Mesh(path)
            t1 = time.perf_counter()
            sc = rt.Scene.from_mesh(ctx, mesh)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            parse.append(t1 - t0)
            upload.append(t2 - t1)
            sc.close()
    sc = rt.Scene.from_mesh(ctx, mesh)
    t0 = time.perf_counter()
    acc = rt.Accel(sc, width=args.width, leaf_size=args.leaf_size, lds_stack=args.lds_stack, sah_bins=args.sah_bins)
    build_wall = time.perf_counter() - t0
    cam = w.camera
    t0 = time.perf_counter()
    first = acc.render(cam, w.lights, want_hits=True)
    first_wall = time.perf_counter() - t0
    ms = sorted(acc.render(cam, w.lights, want_hits=False)["stats"]["ms"][0] for _ in range(5))[2]
    gen_scene = rt.Scene(ctx, w.abc, w.albedo)
    gen_acc = rt.Accel(gen_scene, width=args.width, leaf_size=args.leaf_size, lds_stack=args.lds_stack,
                       sah_bins=args.sah_bins)
    gen = gen_acc.render(cam, w.lights, want_hits=True)
    same = bool(np.array_equal(first["primary"]["prim"], gen["primary"]["prim"]))
    med = lambda xs: sorted(xs)[len(xs) // 2]  # noqa: E731
    out = {
        "workload": "terrain1m as a binary big-endian PLY (%d vertices, %d faces, %.1f MB): parse, upload with "
                    "materials, BVH build, 1920x1080 frame (primary + shadow rays, phong kd 200)"
                    % (verts.shape[0], nf, size / 1e6),
        "triangles": int(mesh.n),
        "parse_ms": round(med(parse) * 1e3, 2),
        "parse_mb_s": round(size / med(parse) / 1e6, 1),
        "parse_mtris_s": round(mesh.n / med(parse) / 1e6, 2),
        "scene_upload_ms": round(med(upload) * 1e3, 2),
        "build_ms": round(acc.info()["build_ms"], 2),
        "build_wall_ms": round(build_wall * 1e3, 2),
        "first_frame_wall_ms": round(first_wall * 1e3, 2),
        "frame_gpu_ms": round(ms, 3),
        "file_to_first_frame_ms": round((med(parse) + med(upload) + build_wall + first_wall) * 1e3, 1),
        "primary_ids_identical_to_generated_scene": same,
    }
    for h in (gen_acc, gen_scene, acc, sc, mesh):
        h.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=4)
    ap.add_argument("--leaf-size", type=int, default=2)
    ap.add_argument("--sah-bins", type=int, default=16, help="binned-SAH rebuild of the LBVH (0 = keep the LBVH)")
    ap.add_argument("--lds-stack", type=int, default=8)
    ap.add_argument("--regroup", type=int, default=0, help="regroup bounce rays before tracing (0/1)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batched-scenes", type=int, default=64, help="0 skips the batched workload")
    ap.add_argument("--batched-reps", type=int, default=10)
    ap.add_argument("--adjoint-reps", type=int, default=10, help="0 skips the differentiable workload")
    ap.add_argument("--invrender-steps", type=int, default=100, help="0 skips batched inverse rendering")
    ap.add_argument("--dynamic-frames", type=int, default=8, help="0 skips the moving-scene workload")
    ap.add_argument("--material-frames", type=int, default=5, help="0 skips the phong-material workload")
    ap.add_argument("--post-reps", type=int, default=10, help="0 skips the camera post-process workload")
    ap.add_argument("--ingest-reps", type=int, default=3, help="0 skips the scene-ingestion workload")
    ap.add_argument("--aac-reps", type=int, default=3, help="0 skips the reference builder's tree")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU would trace its own share "
              "of the rays against a replicated BVH). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch

    from raptor_amd import rt, workloads

    print("bench: generating terrain and building the BVH", file=sys.stderr, flush=True)
    w = workloads.terrain1m()
    ctx = rt.Context(0)
    scene = rt.Scene(ctx, w.abc, w.albedo)
    accel = rt.Accel(scene, width=args.width, leaf_size=args.leaf_size, lds_stack=args.lds_stack,
                     sah_bins=args.sah_bins)
    info = accel.info()
    rgb = torch.empty((w.camera.y_res, w.camera.x_res, 3), dtype=torch.float32, device="cuda:0")

    # Frames are enqueued back to back (RT_ASYNC): the host does not wait
    # for a frame before enqueueing the next, and each frame's rays are summed
    # on the device from its own hit counts.
    def frame(asynchronous):
        return accel.render_device(w.camera, w.lights, rgb, bounce=w.bounce, seed=w.seed,
                                   regroup=bool(args.regroup), asynchronous=asynchronous)

    for _ in range(args.warmup):
        frame(True)
    print("bench: timing", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    ctx.frame_rays()  # clears the sums
    t0 = time.perf_counter()
    for _ in range(args.steps):
        frame(True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    stage_rays = ctx.frame_rays()

    total_rays = sum(stage_rays)
    value = total_rays / elapsed / 1e6
    # Per-stage device times from frames that report them (each waits for
    # itself), after the timed region.
    per_step, sync_wall = [], []
    for _ in range(5):
        t1 = time.perf_counter()
        per_step.append(frame(False))
        sync_wall.append(time.perf_counter() - t1)
    rays = per_step[-1]["rays"]
    if [r * args.steps for r in rays] != stage_rays:
        raise RuntimeError(f"device ray sums {stage_rays} differ from {args.steps} x {rays}")
    # A frame that first rebuilds its BVH from the scene (a scene that changed
    # since the last frame): wall time of build plus frame, median of 3.
    rebuilds, rebuild_build = [], []
    for _ in range(3):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        acc2 = rt.Accel(scene, width=args.width, leaf_size=args.leaf_size, lds_stack=args.lds_stack,
                        sah_bins=args.sah_bins)
        acc2.render_device(w.camera, w.lights, rgb, bounce=w.bounce, seed=w.seed, regroup=bool(args.regroup))
        rebuilds.append(time.perf_counter() - t1)
        rebuild_build.append(acc2.info()["build_ms"])
        acc2.close()

    # Dominant kernel and its roofline: one extra frame with visit counting.
    ctx.set_counting(True)
    counted = accel.render_device(w.camera, w.lights, rgb, bounce=w.bounce, seed=w.seed, regroup=bool(args.regroup))
    ctx.set_counting(False)
    # The GPU's hit ids of the frame, for the CPU baseline's packet-path report.
    gpu_hits = None
    if not args.no_cpu_baseline:
        hf = accel.render(w.camera, w.lights, bounce=w.bounce, seed=w.seed, stats=False)
        gpu_hits = (hf["primary"]["prim"].copy(), hf["bounce"]["prim"].copy())
        del hf
    stage_ms = [sum(s["ms"][1 + k] for s in per_step) / len(per_step) for k in range(4)]
    dom = max(range(4), key=lambda k: stage_ms[k])
    closest = dom in (0, 2)
    bytes_per_launch = algorithmic_bytes(args.width, counted["rays"][dom], counted["visits"][2 * dom],
                                         counted["visits"][2 * dom + 1], closest)
    model_gbs = bytes_per_launch / (stage_ms[dom] * 1e-3) / 1e9
    # Compulsory traffic: each ray in and result out once, the tree and the
    # leaf triangles once.
    compulsory = counted["rays"][dom] * (32 + (16 if closest else 1)) + info["bytes"]
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "trace_terrain1m.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            t = json.load(f)
        same = (t.get("kernel_stage") == STAGES[dom] and t.get("width") == args.width
                and t.get("kernel", "").startswith(KERNELS[dom] + "<")
                and t.get("leaf_size") == args.leaf_size and t.get("lds_stack") == args.lds_stack
                and t.get("sah_bins", 0) == args.sah_bins)
        if same:
            traffic = t.get("bytes_per_launch")

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "terrain1m: 999,698-triangle fBm terrain (708x708 grid), 1920x1080 primary + 1 shadow ray "
                        "per hit + 1 cosine-weighted bounce + 1 bounce shadow ray, FP32",
            "triangles": int(scene.n),
            "resolution": [w.camera.x_res, w.camera.y_res],
            "bvh_width": args.width,
            "leaf_size": args.leaf_size,
            "sah_bins": args.sah_bins,
            "lds_stack": args.lds_stack,
            "regroup_bounce": bool(args.regroup),
            "frames": "enqueued back to back (RT_ASYNC), rays summed on the device",
        },
        "rays_per_step": {"primary": rays[0], "shadow": rays[1], "bounce": rays[2], "bounce_shadow": rays[3]},
        "frame_waiting_for_itself_ms": round(sorted(sync_wall)[2] * 1e3, 4),
        "stage_ms": {STAGES[k]: round(stage_ms[k], 4) for k in range(4)},
        "bvh": {"nodes": info["nodes"], "leaves": info["leaves"], "depth": info["depth"],
                "bytes": info["bytes"], "build_ms": round(info["build_ms"], 2),
                "rebuild_ms": round(sorted(rebuild_build)[1], 2)},
        "frame_with_rebuild_ms": round(sorted(rebuilds)[1] * 1e3, 3),
        "roofline": roofline_entry(args, dom, stage_ms[dom], traffic, bytes_per_launch, model_gbs, compulsory,
                                   counted),
    }
    if args.dynamic_frames > 0:
        print("bench: moving scene", file=sys.stderr, flush=True)
        line["dynamic"] = dynamic_section(ctx, w, accel, scene, args.dynamic_frames, args)
    if args.post_reps > 0:
        print("bench: camera post-process", file=sys.stderr, flush=True)
        line["postprocess"] = post_section(ctx, rgb, args.post_reps, cpu_baseline=not args.no_cpu_baseline)
    if args.material_frames > 0:
        print("bench: phong materials", file=sys.stderr, flush=True)
        line["materials"] = materials_section(ctx, args.material_frames, args)
    if args.ingest_reps > 0:
        print("bench: scene ingestion", file=sys.stderr, flush=True)
        line["ingest"] = ingest_section(ctx, w, args, args.ingest_reps)
    if args.aac_reps > 0:
        print("bench: the reference builder's tree", file=sys.stderr, flush=True)
        line["aac"] = aac_section(ctx, w, scene, accel, args.aac_reps)
    if args.batched_scenes > 0:
        del rgb, accel, scene
        print("bench: batched workload", file=sys.stderr, flush=True)
        line["batched"] = batched_section(ctx, args.batched_scenes, args.batched_reps)
    if args.adjoint_reps > 0:
        print("bench: differentiable workload", file=sys.stderr, flush=True)
        line["adjoint"] = adjoint_section(ctx, args.adjoint_reps)
    if args.invrender_steps > 1:
        print("bench: batched inverse rendering", file=sys.stderr, flush=True)
        line["invrender"] = invrender_section(ctx, args.invrender_steps)
    if not args.no_cpu_baseline:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        cb = cpu_baseline(w, args.cpu_seconds, gpu_hits, line["ms_per_step"])
        line["cpu_baseline"] = cb
        line["gpu_over_cpu"] = round(value / cb["value"], 2)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
