"""Oracle — TEST INFRASTRUCTURE ONLY.

Double-precision restatement of a frame's shading as a function of the
per-triangle albedo and the point lights, over fixed path records, and the
project's CPU adjoint of it. Only tests/ import this module.

The records come from the oracle's float frame (OracleScene.render with
records=True): per sample the path vertices (primary hit, then each diffuse
bounce hit: triangle, shadow-ray origin, shading normal, visibility per
light). Over them the radiance is (rt/raytracer.cc:20-87,
rt/materials/phong_shader.cc:21-64 with kd = 255 * albedo and ks = ka = 0,
rt/light.h:44-59; the bounces are this project's definition)

    c = sum_k T_k D_k,   T_0 = 1,  T_k = T_{k-1} * kd_{k-1} / 255,
    D_k = kd_k * sum_l V_kl S_kl I_kl,   S = dot(dir_l, n) (lights with S < 0
    skipped), dir_l = (pos_l - o) / len_l,  I = rgb_l / 255 / ((len_l d_l)^2 + 1),
    pixel = mean over samples,

which this module evaluates in float64 so that central finite differences of
it are accurate to ~1e-9; the adjoint below is checked against them.
"""
import numpy as np


def _direct(o, n, vis, pos, rgb, d):
    """Light sums at surfaces o (S,3) with normals n, visibility vis (S,L):
    returns s (S,3) = sum_l V S I (so D = kd * s) and the terms."""
    v = pos[None, :, :] - o[:, None, :]                 # (S, L, 3)
    ln = np.sqrt((v * v).sum(-1))                         # (S, L)
    dirn = v / ln[..., None]
    shade = (dirn * n[:, None, :]).sum(-1)                # (S, L)
    act = (vis > 0) & (shade >= 0.0)
    D = (ln * d[None, :]) ** 2 + 1.0                      # (S, L)
    inten = (rgb[None, :, :] / 255.0) / D[..., None]      # (S, L, 3)
    w = np.where(act, shade, 0.0)                         # (S, L)
    s = (w[..., None] * inten).sum(1)                     # (S, 3)
    return s, dict(ln=ln, dirn=dirn, shade=shade, act=act, D=D, inten=inten)


def _as_f64(albedo, lights):
    pos = np.array([l[0] for l in lights], np.float64).reshape(-1, 3)
    rgb = np.array([l[1] for l in lights], np.float64).reshape(-1, 3)
    d = np.array([l[2] for l in lights], np.float64).reshape(-1)
    return np.asarray(albedo, np.float64), pos, rgb, d


def _levels(rec, bounces):
    return min(int(bounces) + 1, rec["prim"].shape[0])


def _walk(rec, kd, pos, rgb, d, bounces):
    """Per level: (mask of samples with a vertex, kd, light sums, terms)."""
    out = []
    alive = np.ones(rec["prim"].shape[1], bool)
    for k in range(_levels(rec, bounces)):
        pk = rec["prim"][k]
        alive = alive & (pk >= 0)
        if not alive.any():
            break
        o = rec["o"][k][alive].astype(np.float64)
        n = rec["n"][k][alive].astype(np.float64)
        s, t = _direct(o, n, rec["vis"][k][alive], pos, rgb, d)
        out.append(dict(mask=alive.copy(), prim=pk[alive], kd=kd[pk[alive]], s=s, t=t, n=n))
    return out


def shade(rec, albedo, lights, spp, bounces):
    """Pixel radiance (npix, 3) in float64. lights: [(pos3, rgb3, d), ...]."""
    alb, pos, rgb, d = _as_f64(albedo, lights)
    kd = 255.0 * alb
    ns = rec["prim"].shape[1]
    c = np.zeros((ns, 3))
    T = np.ones((ns, 3))
    for lv in _walk(rec, kd, pos, rgb, d, bounces):
        m = lv["mask"]
        c[m] += T[m] * lv["kd"] * lv["s"]
        T[m] = T[m] * lv["kd"] / 255.0
    return c.reshape(-1, spp, 3).sum(1) / spp


def _light_grads(t, w, kd, n, d):
    """Light gradients of sum_s w_s . D_s for one level (w already carries T)."""
    act = t["act"]
    shade = np.where(act, t["shade"], 0.0)
    D, ln, dirn, inten = t["D"], t["ln"], t["dirn"], t["inten"]
    # dD/drgb (0-255 scale) = V S kd / (255 D)
    d_rgb = ((shade / (255.0 * D))[..., None] * (w * kd)[:, None, :]).sum(0)
    # dD/dpos = V kd I [(n - S dir) / len - S (2 d^2 len / D) dir]
    T = ((w * kd)[:, None, :] * inten).sum(-1) * act      # (S, L)
    fall = shade * (2.0 * d[None, :] ** 2 * ln) / D
    g = (n[:, None, :] - shade[..., None] * dirn) / ln[..., None] - fall[..., None] * dirn
    d_pos = (T[..., None] * g).sum(0)
    return d_pos, d_rgb


def adjoint(rec, albedo, lights, spp, bounces, grad):
    """Vector-Jacobian product of shade() for grad = dL/d pixel (npix, 3).
    Returns (d_albedo (n,3), d_lights (L,6): position xyz then rgb).

    With R_k = D_k + T'_k R_{k+1} (T'_k = kd_k / 255) the radiance is
    c = R_0, and dc/dkd_k = T_k (s_k + R_{k+1} / 255)."""
    alb, pos, rgb, d = _as_f64(albedo, lights)
    kd = 255.0 * alb
    nl = pos.shape[0]
    g = np.repeat(np.asarray(grad, np.float64).reshape(-1, 3), spp, axis=0) / spp
    ns = g.shape[0]
    levels = _walk(rec, kd, pos, rgb, d, bounces)
    # Throughput T_k into each level.
    T = np.ones((ns, 3))
    Ts = []
    for lv in levels:
        m = lv["mask"]
        Ts.append(T[m].copy())
        T[m] = T[m] * lv["kd"] / 255.0
    d_kd = np.zeros_like(kd)
    d_l = np.zeros((nl, 6))
    R_next = np.zeros((ns, 3))   # R_{k+1}, indexed by sample
    for k in range(len(levels) - 1, -1, -1):
        lv = levels[k]
        m = lv["mask"]
        w = g[m] * Ts[k]
        np.add.at(d_kd, lv["prim"], w * (lv["s"] + R_next[m] / 255.0))
        dp, dr = _light_grads(lv["t"], w, lv["kd"], lv["n"], d)
        d_l[:, :3] += dp
        d_l[:, 3:] += dr
        R = np.zeros((ns, 3))
        R[m] = lv["kd"] * lv["s"] + (lv["kd"] / 255.0) * R_next[m]
        R_next = R
    return 255.0 * d_kd, d_l


def loss(img, target):
    """L = sum over pixels and channels of (img - target)^2."""
    r = np.asarray(img, np.float64) - np.asarray(target, np.float64)
    return float((r * r).sum())


def loss_grad(img, target):
    return 2.0 * (np.asarray(img, np.float64) - np.asarray(target, np.float64))
