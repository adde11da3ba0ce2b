"""The GPU build against restatements on the host, through the C ABI:

* Morton codes (rt_build_morton) equal a numpy restatement of the build's
  quantisation: box centroids (lo + hi) / 2 in float32, one cell size for
  all axes (the largest centroid extent / 1024), truncation, clamping to
  [0, 1023] and bit interleaving x | y << 1 | z << 2 done bit by bit. (The
  reference quantises each axis by its own extent, ssd/bvh_builder.cc:45-48;
  the GPU's cubic cells are a deliberate change, so the codes are pinned by
  this restatement and the order by the reference's sort.)
* The radix sort's permutation equals a stable argsort of the codes and the
  oracle's restatement of the reference's 3-pass LSD sort
  (ssd/bvh_builder.cc:50-158) applied to the same codes.
* The downloaded tree is a tree over the triangles: every interior node is
  reached exactly once from the root, leaf ranges are disjoint and cover the
  BVH order, the BVH order is a permutation of the emplace ids, every child
  box contains the boxes of that child's children (or the vertices of the
  leaf's triangles), leaves hold at most leaf_size triangles, and the depth
  is the one rt_accel_info reports.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def spread10(q, axis):
    out = np.zeros_like(q, dtype=np.uint32)
    for b in range(10):
        out |= ((q >> np.uint32(b)) & np.uint32(1)) << np.uint32(3 * b + axis)
    return out


def cubic_morton(abc):
    v = np.ascontiguousarray(abc, np.float32).reshape(-1, 3, 3)
    lo, hi = v.min(1), v.max(1)
    c = (lo + hi) * np.float32(0.5)
    cmin, cmax = c.min(0), c.max(0)
    ext = np.float32((cmax - cmin).max())
    f = (c - cmin) / ext if ext > 0 else np.zeros_like(c)
    q = np.minimum(np.maximum(f * np.float32(1024.0), np.float32(0.0)), np.float32(1023.0)).astype(np.uint32)
    return spread10(q[:, 0], 0) | spread10(q[:, 1], 1) | spread10(q[:, 2], 2)


def _scenes(rt):
    from raptor_amd import workloads

    return {
        "random65k": workloads.random65k().abc,
        "terrain300": workloads.terrain(grid=300, x_res=8, y_res=8).abc,
        "terrain1m": workloads.terrain1m().abc,
        "duplicates": np.repeat(workloads.random65k(n=500).abc, 8, axis=0),  # equal codes in runs of 8
    }


@pytest.fixture(scope="module")
def scenes(rt):
    return _scenes(rt)


@pytest.mark.parametrize("name", ["random65k", "terrain300", "terrain1m", "duplicates"])
def test_morton_codes_and_sort(gpu_ctx, rt, O, scenes, name):
    abc = scenes[name]
    scene = rt.Scene(gpu_ctx, abc)
    codes, order = scene.morton()
    exp = cubic_morton(abc)
    assert np.array_equal(codes, exp), f"{(codes != exp).sum()} codes differ"
    assert codes.max() < (1 << 30)
    assert np.array_equal(order, np.argsort(codes, kind="stable").astype(np.uint32))
    _, ref_order = O.morton_sort(codes.astype(np.int32))
    assert np.array_equal(order, ref_order.astype(np.uint32))


def check_tree(acc, abc, leaf_size):
    t = acc.download()
    info = acc.info()
    n = abc.shape[0]
    W, child, lo, hi = t["width"], t["child"], t["lo"], t["hi"]
    nn = child.shape[0]
    assert nn == info["nodes"] and t["root"] == 0
    # BVH order is a permutation of the emplace ids.
    assert np.array_equal(np.sort(t["tri_ids"]), np.arange(n, dtype=np.int32))
    leaf = (child != -1) & (child < 0)
    inner = child >= 0
    # Every interior node but the root is the child of exactly one slot.
    refs = np.bincount(child[inner], minlength=nn)
    assert refs[0] == 0 and np.all(refs[1:] == 1)
    # Leaves: disjoint ranges covering [0, n), at most leaf_size each.
    codes = child[leaf].view(np.uint32)
    first = ((codes & np.uint32(0x7FFFFFFF)) >> np.uint32(4)).astype(np.int64)
    count = (codes & np.uint32(15)).astype(np.int64) + 1
    assert count.max() <= leaf_size
    # BVH positions of every leaf's triangles, leaf by leaf.
    pos = np.repeat(first, count) + (np.arange(count.sum()) - np.repeat(np.cumsum(count) - count, count))
    cover = np.bincount(pos, minlength=n)
    assert len(cover) == n and np.all(cover == 1), f"{(cover != 1).sum()} BVH positions not in exactly one leaf"
    # An interior child's box contains its children's boxes.
    p, k = np.nonzero(inner)
    c = child[p, k]
    valid = child[c] != -1                              # (m, W) slots of each child node
    big = np.float32(np.finfo(np.float32).max)
    cl = np.where(valid[:, None, :], lo[c], big).min(2)     # (m, 3)
    ch = np.where(valid[:, None, :], hi[c], -big).max(2)
    assert np.all(lo[p, :, k] <= cl) and np.all(hi[p, :, k] >= ch)
    # A leaf child's box contains the vertices of its triangles.
    lp, lk = np.nonzero(leaf)
    owner = np.repeat(np.arange(len(first)), count)
    v = abc.reshape(-1, 3, 3)[t["tri_ids"][pos]]          # (T, 3 vertices, 3 axes)
    blo, bhi = lo[lp[owner], :, lk[owner]], hi[lp[owner], :, lk[owner]]
    assert np.all(v.min(1) >= blo) and np.all(v.max(1) <= bhi)
    # Depth: interior levels below the root.
    depth = np.zeros(nn, np.int64)
    for node in range(nn):  # nodes are numbered level by level, parents first
        kids = child[node][child[node] >= 0]
        depth[kids] = depth[node] + 1
    assert depth.max() + 1 == info["depth"]
    return info


@pytest.mark.parametrize("width,leaf,sah", [(4, 2, 16), (4, 1, 0), (8, 4, 16), (8, 2, 0)])
@pytest.mark.parametrize("name", ["random65k", "terrain300", "duplicates"])
def test_tree_invariants(gpu_ctx, rt, scenes, name, width, leaf, sah):
    abc = scenes[name]
    scene = rt.Scene(gpu_ctx, abc)
    acc = rt.Accel(scene, width=width, leaf_size=leaf, sah_bins=sah)
    check_tree(acc, abc, leaf)


def test_tree_invariants_1m_terrain_default_build(gpu_ctx, rt, scenes):
    abc = scenes["terrain1m"]
    acc = rt.Accel(rt.Scene(gpu_ctx, abc))
    info = check_tree(acc, abc, 2)
    assert info["leaves"] > abc.shape[0] // 2


def test_single_triangle_and_refit_keep_the_tree(gpu_ctx, rt, scenes):
    one = scenes["random65k"][:1]
    acc = rt.Accel(rt.Scene(gpu_ctx, one))
    t = acc.download()
    assert acc.info()["nodes"] == 0 and t["root"] < 0 and list(t["tri_ids"]) == [0]
    # A refit after moving the vertices keeps the topology and contains the new geometry.
    from raptor_amd import workloads

    abc = scenes["terrain300"]
    scene = rt.Scene(gpu_ctx, abc)
    acc = rt.Accel(scene)
    before = acc.download()
    moved = workloads.terrain_wave(abc, 0.3, amplitude=0.05)
    scene.update(moved)
    acc.refit()
    after = acc.download()
    assert np.array_equal(before["child"], after["child"]) and np.array_equal(before["tri_ids"], after["tri_ids"])
    check_tree(acc, moved, 2)
