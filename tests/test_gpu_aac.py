"""The reference builder's tree built on the GPU (rt_build_aac) against the
oracle's restatement of ssd/bvh_builder.cc, node for node:

* the Morton codes (per-axis cells of the scene extent * 1.00001 / 1024) and
  the radix order equal the oracle's (ssd/bvh_builder.cc:33-158);
* the final triangle order, after the builder's spatial-midpoint partitions
  and leaf packing swaps, is the same permutation;
* the trees are the same: in preorder (left child first, the reference's
  left_b), the same sequence of interior and leaf nodes, every box equal bit
  for bit, and every leaf holding the same triangles in the same order.

Scenes cover random clouds, terrain grids (many exactly equal surface areas,
so the reference's SSE tie rule decides merges), duplicate Morton codes (the
partition fallback), scenes below and at the leaf-range size, one triangle,
and the full 1M-triangle terrain of the headline frame.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def preorder(lohi, child, leaf, indirect, root):
    out = []
    st = [int(root)]
    while st:
        x = st.pop()
        box = np.asarray(lohi[x], np.float32).tobytes()
        if leaf[x]:
            b, e = int(child[x][0]), int(child[x][1])
            out.append(("L", box, tuple(int(v) for v in indirect[b:e])))
        else:
            out.append(("N", box))
            st.append(int(child[x][1]))
            st.append(int(child[x][0]))
    return out


def preorder_arrays(lohi, child, leaf, indirect, root):
    """The same comparison as preorder() for large trees, vectorised: node
    boxes and leaf flags in preorder, leaf sizes and the concatenated leaf
    triangles in preorder."""
    child = np.asarray(child, np.int64)
    leaf = np.asarray(leaf).astype(bool)
    levels = []
    cur = np.array([int(root)], np.int64)
    while cur.size:
        levels.append(cur)
        inner = cur[~leaf[cur]]
        cur = np.concatenate([child[inner, 0], child[inner, 1]])
    size = np.ones(len(leaf), np.int64)
    for lv in reversed(levels):
        inner = lv[~leaf[lv]]
        size[inner] = 1 + size[child[inner, 0]] + size[child[inner, 1]]
    pre = np.full(len(leaf), -1, np.int64)
    pre[int(root)] = 0
    for lv in levels:
        inner = lv[~leaf[lv]]
        pre[child[inner, 0]] = pre[inner] + 1
        pre[child[inner, 1]] = pre[inner] + 1 + size[child[inner, 0]]
    reach = np.nonzero(pre >= 0)[0]
    order = reach[np.argsort(pre[reach])]
    boxes = np.ascontiguousarray(np.asarray(lohi, np.float32)[order]).view(np.uint32)
    lf = order[leaf[order]]
    b, e = child[lf, 0], child[lf, 1]
    tris = np.concatenate([np.asarray(indirect[x:y]) for x, y in zip(b, e)]) if lf.size else np.zeros(0)
    return boxes, leaf[order], e - b, tris


def compare(rt, O, ctx, abc):
    scene = rt.Scene(ctx, abc, np.full((abc.shape[0], 3), 0.5, np.float32))
    g = scene.build_aac()
    ref = O.OracleScene(abc)
    t = ref.tree()
    codes, order = ref.aac_morton()
    assert np.array_equal(g["codes"].astype(np.int64), codes.astype(np.int64)), "Morton codes differ"
    assert np.array_equal(g["order"].astype(np.int64), order.astype(np.int64)), "radix order differs"
    assert np.array_equal(g["indirect"], t["indirect"]), "triangle order differs"
    assert len(g["leaf"]) == len(t["leaf"]), (len(g["leaf"]), len(t["leaf"]))
    a = preorder(g["lohi"], g["child"], g["leaf"], g["indirect"], g["root"])
    b = preorder(t["lohi"], t["child"], t["leaf"], t["indirect"], t["root"])
    assert len(a) == len(b)
    for k, (x, y) in enumerate(zip(a, b)):
        assert x == y, f"preorder node {k} differs: {x[0]} vs {y[0]}"
    return g, t


@pytest.mark.parametrize("n,spread,seed", [(2000, 0.01, 1), (7000, 0.05, 3), (50000, 0.004, 7)])
def test_random_clouds(gpu_ctx, rt, O, n, spread, seed):
    abc, _ = rt.gen_random_triangles(seed, n, spread)
    compare(rt, O, gpu_ctx, abc)


@pytest.mark.parametrize("grid", [40, 64, 200])
def test_terrain_grids(gpu_ctx, rt, O, grid):
    from raptor_amd import workloads

    abc, _ = workloads.terrain_arrays(grid)
    compare(rt, O, gpu_ctx, abc)


def test_duplicate_codes_take_the_partition(gpu_ctx, rt, O):
    rng = np.random.default_rng(5)
    base, _ = rt.gen_random_triangles(9, 3000, 0.01)
    c1 = np.array([0.3, 0.6, 0.2], np.float32)
    c2 = np.array([0.71, 0.12, 0.9], np.float32)
    tiny1 = (c1 + rng.uniform(-2e-4, 2e-4, (400, 3, 3))).astype(np.float32).reshape(-1, 9)
    tiny2 = (c2 + rng.uniform(-1e-4, 1e-4, (60, 3, 3))).astype(np.float32).reshape(-1, 9)
    abc = np.concatenate([base, tiny1, tiny2]).astype(np.float32)
    g, _ = compare(rt, O, gpu_ctx, abc)
    _, counts = np.unique(g["codes"], return_counts=True)
    assert counts.max() > 100  # the partition path ran


@pytest.mark.parametrize("n", [1, 2, 19, 20, 21, 97])
def test_small_scenes(gpu_ctx, rt, O, n):
    abc, _ = rt.gen_random_triangles(11, n, 0.05)
    compare(rt, O, gpu_ctx, abc)


def _flat_grid(n_side, y=0.25):
    """A flat square grid in the plane y: the scene box has zero extent on y,
    so that axis's cell width is 0 and its Morton bits come from 0 * inf."""
    g = np.linspace(0.0, 1.0, n_side + 1, dtype=np.float32)
    tris = []
    for i in range(n_side):
        for j in range(n_side):
            a = (g[i], y, g[j])
            b = (g[i + 1], y, g[j])
            c = (g[i + 1], y, g[j + 1])
            d = (g[i], y, g[j + 1])
            tris.append(a + b + c)
            tris.append(a + c + d)
    return np.asarray(tris, np.float32)


@pytest.mark.parametrize("case", ["flat_grid", "identical", "degenerate", "mixed_scales"])
def test_edge_scenes(gpu_ctx, rt, O, case):
    """Scenes the reference can meet: a flat model (one axis of zero extent),
    many identical triangles (one Morton code, partitions down to the depth
    limit), zero-area triangles among ordinary ones, and a few large
    triangles over many small ones."""
    rng = np.random.default_rng(17)
    if case == "flat_grid":
        abc = _flat_grid(30)
    elif case == "identical":
        one = np.array([0.1, 0.2, 0.3, 0.4, 0.2, 0.3, 0.1, 0.6, 0.35], np.float32)
        abc = np.tile(one, (300, 1))
    elif case == "degenerate":
        abc, _ = rt.gen_random_triangles(21, 2000, 0.02)
        abc = abc.copy()
        k = rng.choice(abc.shape[0], 300, replace=False)
        abc[k, 6:9] = abc[k, 0:3] + 0.5 * (abc[k, 3:6] - abc[k, 0:3])  # collinear
        abc[k[:100], 3:9] = np.tile(abc[k[:100], 0:3], (1, 2))         # a point
    else:
        small, _ = rt.gen_random_triangles(23, 3000, 0.002)
        big = rng.uniform(-3.0, 4.0, (12, 9)).astype(np.float32)
        abc = np.concatenate([small, big]).astype(np.float32)
    compare(rt, O, gpu_ctx, abc)


def test_empty_scene(gpu_ctx, rt):
    scene = rt.Scene(gpu_ctx, np.zeros((0, 9), np.float32), np.zeros((0, 3), np.float32))
    g = scene.build_aac()
    assert g["root"] == -1 and len(g["leaf"]) == 0


def test_terrain_1m_headline_scene(gpu_ctx, rt, O):
    """The 999,698-triangle terrain of the headline frame."""
    from raptor_amd import workloads

    abc, _ = workloads.terrain_arrays(708)
    g, t = compare(rt, O, gpu_ctx, abc)
    print(f"AAC tree of {abc.shape[0]} triangles: {len(g['leaf'])} nodes, GPU build {g['build_ms']:.1f} ms")


def test_terrain_beyond_lds_lists(gpu_ctx, rt, O):
    """A 9.7M-triangle terrain: the root's two children keep more than 1024
    clusters each, so the root's merge list (over 2048 clusters) lives in
    global memory instead of LDS."""
    from raptor_amd import workloads

    abc, _ = workloads.terrain_arrays(2200)
    scene = rt.Scene(gpu_ctx, abc, np.full((abc.shape[0], 3), 0.5, np.float32))
    g = scene.build_aac()
    t = O.OracleScene(abc).tree()
    assert np.array_equal(g["indirect"], t["indirect"]), "triangle order differs"
    a = preorder_arrays(g["lohi"], g["child"], g["leaf"], g["indirect"], g["root"])
    b = preorder_arrays(t["lohi"], t["child"], t["leaf"], t["indirect"], t["root"])
    for x, y, what in zip(a, b, ("boxes", "leaf flags", "leaf sizes", "leaf triangles")):
        assert x.shape == y.shape and np.array_equal(x, y), what
    print(f"AAC tree of {abc.shape[0]} triangles: {len(g['leaf'])} nodes, GPU build {g['build_ms']:.1f} ms")


# ------------------------------------------- frames on the reference's tree

@pytest.mark.parametrize("width", [4, 8])
def test_accel_on_the_aac_tree(gpu_ctx, rt, O, width):
    """rt_accel_build with RT_BUILDER_AAC: the wide tree keeps the AAC
    triangle order and leaves (split only where a leaf exceeds 16 triangles),
    satisfies the tree invariants, and traces the oracle's hits."""
    from test_gpu_build import check_tree
    from raptor_amd import workloads

    from helpers import compare_hits, tuv_exact

    w = workloads.random65k()
    scene = rt.Scene(gpu_ctx, w.abc, w.albedo)
    acc = rt.Accel(scene, width=width, builder=rt.BUILDER_AAC)
    check_tree(acc, w.abc, 16)
    g = scene.build_aac()
    t = acc.download()
    assert np.array_equal(t["tri_ids"], g["indirect"])
    # every wide leaf lies inside one AAC leaf
    bounds = np.zeros(w.abc.shape[0] + 1, np.int64)
    for x in np.nonzero(g["leaf"])[0]:
        bounds[g["child"][x][0]] = 1
    owner = np.cumsum(bounds[:-1]) - 1
    codes = t["child"][(t["child"] != -1) & (t["child"] < 0)].view(np.uint32)
    first = ((codes & np.uint32(0x7FFFFFFF)) >> np.uint32(4)).astype(np.int64)
    count = (codes & np.uint32(15)).astype(np.int64) + 1
    assert np.all(owner[first] == owner[first + count - 1])
    rays = O.primary_rays(w.camera)
    gh = acc.trace_closest(rays)
    r, _ = O.OracleScene(w.abc).trace_closest(rays)
    rtuv = np.stack([r["t"], r["u"], r["v"]], 1)
    bad, ties = compare_hits(gh, r["prim"], rtuv)
    assert not bad.any() and tuv_exact(gh, rtuv, (r["prim"] >= 0) & ~ties)


def test_frame_and_refit_on_the_aac_tree(gpu_ctx, rt, O):
    """A terrain frame with a bounce on the AAC-built accel, then the terrain
    moved and refit, both against the oracle (ids bit-exact, radiance 1e-5)."""
    from raptor_amd import workloads

    from helpers import compare_hits, rel_err

    w = workloads.terrain(grid=200, x_res=320, y_res=180, bounce=True)
    scene = rt.Scene(gpu_ctx, w.abc, w.albedo)
    acc = rt.Accel(scene, builder=rt.BUILDER_AAC)
    for step in range(2):
        abc = w.abc if step == 0 else workloads.terrain_wave(w.abc, 0.7)
        if step == 1:
            scene.update(abc)
            acc.refit()
        out = acc.render(w.camera, w.lights, bounce=True, seed=w.seed)
        ref = O.OracleScene(abc).render(w.albedo, w.camera, w.lights, bounce=True, seed=w.seed, nthreads=8)
        bad, ties = compare_hits(out["primary"], ref["prim"], ref["tuv"])
        assert not bad.any()
        e = rel_err(out["rgb"].reshape(-1, 3)[~ties], ref["rgb"][~ties])
        assert e.max() <= 1e-5, e.max()
