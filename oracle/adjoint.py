"""Oracle — TEST INFRASTRUCTURE ONLY.

Double-precision restatement of a frame's shading as a function of the
per-triangle albedo and the point lights, over fixed path records, and the
project's CPU adjoint of it. Only tests/ import this module.

The records come from the oracle's float frame (OracleScene.render with
records=True): per sample the primary and bounce hit (triangle, shadow-ray
origin, shading normal, visibility per light). Over them the radiance is
(rt/raytracer.cc:20-87, rt/materials/phong_shader.cc:21-64 with kd = 255 *
albedo and ks = ka = 0, rt/light.h:44-59)

    c = sum_l V1 S1 kd1 I1 + (kd1 / 255) * sum_l V2 S2 kd2 I2,
    S = dot(dir_l, n) (lights with S < 0 skipped), dir_l = (pos_l - o) / len_l,
    I = rgb_l / 255 / ((len_l d_l)^2 + 1),    pixel = mean over samples,

which this module evaluates in float64 so that central finite differences of
it are accurate to ~1e-9; the adjoint below is checked against them.
"""
import numpy as np


def _direct(o, n, vis, kd, pos, rgb, d):
    """Direct light at surfaces o (S,3) with normals n, visibility vis (S,L),
    kd (S,3). Returns radiance (S,3) and the intermediate terms."""
    v = pos[None, :, :] - o[:, None, :]                 # (S, L, 3)
    ln = np.sqrt((v * v).sum(-1))                         # (S, L)
    dirn = v / ln[..., None]
    shade = (dirn * n[:, None, :]).sum(-1)                # (S, L)
    act = (vis > 0) & (shade >= 0.0)
    D = (ln * d[None, :]) ** 2 + 1.0                      # (S, L)
    inten = (rgb[None, :, :] / 255.0) / D[..., None]      # (S, L, 3)
    w = np.where(act, shade, 0.0)                         # (S, L)
    c = (w[..., None] * inten * kd[:, None, :]).sum(1)    # (S, 3)
    return c, dict(ln=ln, dirn=dirn, shade=shade, act=act, D=D, inten=inten)


def _as_f64(albedo, lights):
    pos = np.array([l[0] for l in lights], np.float64).reshape(-1, 3)
    rgb = np.array([l[1] for l in lights], np.float64).reshape(-1, 3)
    d = np.array([l[2] for l in lights], np.float64).reshape(-1)
    return np.asarray(albedo, np.float64), pos, rgb, d


def shade(rec, albedo, lights, spp, bounce):
    """Pixel radiance (npix, 3) in float64. lights: [(pos3, rgb3, d), ...]."""
    alb, pos, rgb, d = _as_f64(albedo, lights)
    kd = 255.0 * alb
    p1, p2 = rec["prim1"], rec["prim2"]
    ns = p1.shape[0]
    c = np.zeros((ns, 3))
    m1 = p1 >= 0
    c1, _ = _direct(rec["o1"][m1].astype(np.float64), rec["n1"][m1].astype(np.float64), rec["vis1"][m1],
                    kd[p1[m1]], pos, rgb, d)
    c[m1] = c1
    if bounce:
        m2 = m1 & (p2 >= 0)
        c2, _ = _direct(rec["o2"][m2].astype(np.float64), rec["n2"][m2].astype(np.float64), rec["vis2"][m2],
                        kd[p2[m2]], pos, rgb, d)
        c[m2] += (kd[p1[m2]] / 255.0) * c2
    return c.reshape(-1, spp, 3).sum(1) / spp


def _light_grads(t, w, kd, n, d):
    """Light gradients of sum_s w_s . c_s for one surface level."""
    act = t["act"]
    shade = np.where(act, t["shade"], 0.0)
    D, ln, dirn, inten = t["D"], t["ln"], t["dirn"], t["inten"]
    # dc/drgb (0-255 scale) = V S kd / (255 D)
    d_rgb = ((shade / (255.0 * D))[..., None] * (w * kd)[:, None, :]).sum(0)
    # dc/dpos = V kd I [(n - S dir) / len - S (2 d^2 len / D) dir]
    T = ((w * kd)[:, None, :] * inten).sum(-1) * act      # (S, L)
    fall = shade * (2.0 * d[None, :] ** 2 * ln) / D
    g = (n[:, None, :] - shade[..., None] * dirn) / ln[..., None] - fall[..., None] * dirn
    d_pos = (T[..., None] * g).sum(0)
    return d_pos, d_rgb


def adjoint(rec, albedo, lights, spp, bounce, grad):
    """Vector-Jacobian product of shade() for grad = dL/d pixel (npix, 3).
    Returns (d_albedo (n,3), d_lights (L,6): position xyz then rgb)."""
    alb, pos, rgb, d = _as_f64(albedo, lights)
    kd = 255.0 * alb
    nl = pos.shape[0]
    g = np.repeat(np.asarray(grad, np.float64).reshape(-1, 3), spp, axis=0) / spp
    p1, p2 = rec["prim1"], rec["prim2"]
    d_kd = np.zeros_like(kd)
    d_l = np.zeros((nl, 6))
    m1 = p1 >= 0
    o1, n1 = rec["o1"][m1].astype(np.float64), rec["n1"][m1].astype(np.float64)
    kd1 = kd[p1[m1]]
    w1 = g[m1]
    c1, t1 = _direct(o1, n1, rec["vis1"][m1], kd1, pos, rgb, d)
    s1 = (np.where(t1["act"], t1["shade"], 0.0)[..., None] * t1["inten"]).sum(1)   # (S, 3)
    dk1 = w1 * s1
    if bounce:
        sel = p2[m1] >= 0
        m2 = m1.copy()
        m2[m1] = sel
        o2, n2 = rec["o2"][m2].astype(np.float64), rec["n2"][m2].astype(np.float64)
        kd2 = kd[p2[m2]]
        c2, t2 = _direct(o2, n2, rec["vis2"][m2], kd2, pos, rgb, d)
        dk1[sel] += w1[sel] * c2 / 255.0
        w2 = w1[sel] * kd1[sel] / 255.0
        s2 = (np.where(t2["act"], t2["shade"], 0.0)[..., None] * t2["inten"]).sum(1)
        np.add.at(d_kd, p2[m2], w2 * s2)
        dp, dr = _light_grads(t2, w2, kd2, n2, d)
        d_l[:, :3] += dp
        d_l[:, 3:] += dr
    np.add.at(d_kd, p1[m1], dk1)
    dp, dr = _light_grads(t1, w1, kd1, n1, d)
    d_l[:, :3] += dp
    d_l[:, 3:] += dr
    return 255.0 * d_kd, d_l


def loss(img, target):
    """L = sum over pixels and channels of (img - target)^2."""
    r = np.asarray(img, np.float64) - np.asarray(target, np.float64)
    return float((r * r).sum())


def loss_grad(img, target):
    return 2.0 * (np.asarray(img, np.float64) - np.asarray(target, np.float64))
