"""Comparison helpers shared by the parity tests."""
import numpy as np


def compare_hits(gpu, ref_prim, ref_tuv):
    """Hit ids bit-exact, except where two triangles give the same t bit for
    bit (an exact tie, which the reference breaks by visit order and the GPU by
    the lowest id). Returns (bad, ties) boolean masks."""
    gprim = np.asarray(gpu["prim"])
    gt = np.asarray(gpu["t"], np.float32)
    diff = gprim != ref_prim
    ties = diff & (gt == ref_tuv[:, 0]) & (gprim >= 0) & (ref_prim >= 0)
    bad = diff & ~ties
    return bad, ties


def tuv_exact(gpu, ref_tuv, mask):
    """t, u, v bit-identical where mask holds (hit, not a tie)."""
    g = np.stack([gpu["t"], gpu["u"], gpu["v"]], axis=1).astype(np.float32)[mask]
    r = ref_tuv[mask]
    return np.array_equal(g.view(np.uint32), r.view(np.uint32))


def rel_err(a, b, floor=1e-3):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def oracle_material_rows(materials):
    """rt_material rows (the library's MATERIAL_DTYPE) as the oracle's
    40-float rows: ka, kd, ks, s, tran, ri, rf, kind, cx, cy, rd, sr, ri_i,
    six texture slots, mapper kd, ks, clrh, clrf, layer firsts, counts,
    iterations."""
    m = materials
    col = lambda k: m[k][:, None].astype(np.float32)  # noqa: E731
    return np.concatenate([m["ka"], m["kd"], m["ks"], col("shininess"), col("transmittance"),
                           col("refractive_index"), col("reflectance"), col("kind"), m["chroma"], col("diffuse"),
                           col("roughness"), col("extinction"), m["tex"].astype(np.float32), col("mapper_kd"),
                           col("mapper_ks"), col("clrh"), col("clrf"), m["layer_first"].astype(np.float32),
                           m["layer_count"].astype(np.float32), col("iterations")], axis=1).astype(np.float32)
