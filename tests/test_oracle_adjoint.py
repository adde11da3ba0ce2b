"""The CPU adjoint of the oracle (oracle/adjoint.py), pinned on the CPU:

* its float64 shading over the path records reproduces the oracle's float
  frame (which the parity tests pin to the reference's arithmetic);
* its gradients equal float64 central finite differences of that shading for
  albedo entries, light positions and light intensities;
* the finite-difference check fails when a term of the adjoint is dropped.
"""
import numpy as np
import pytest

from oracle import adjoint as A


@pytest.fixture(scope="module")
def frame(O):
    from raptor_amd import workloads

    w = workloads.diff_terrain(grid=24, x_res=32, y_res=24, spp=2, seed=5, zoom=1.0, amplitude=0.3)
    sc = O.OracleScene(w.abc)
    out = sc.render(w.albedo, w.camera, w.lights, bounce=True, spp=w.spp, seed=w.seed, nthreads=4, records=True)
    lights = [(l.pos, l.rgb, l.d) for l in w.lights]
    return w, out, lights


def test_records_cover_the_frame(frame):
    w, out, _ = frame
    rec = out["records"]
    ns = w.camera.x_res * w.camera.y_res * w.spp
    assert rec["prim"].shape == (2, ns)
    assert (rec["prim"][0] >= -1).all() and (rec["prim"][1] >= -1).all()
    # The first sample's ids are the frame's primary and bounce ids.
    assert np.array_equal(rec["prim"][0][:: w.spp], out["prim"])
    assert np.array_equal(rec["prim"][1][:: w.spp], out["bounce_prim"])
    assert (rec["prim"][0] >= 0).mean() > 0.5
    assert (rec["prim"][1] >= 0).sum() > 50
    assert rec["vis"][0][rec["prim"][0] >= 0].mean() > 0.3


def test_float64_shading_reproduces_float_frame(frame):
    w, out, lights = frame
    img = A.shade(out["records"], w.albedo, lights, w.spp, True)
    ref = out["rgb"].astype(np.float64)
    assert np.abs(img - ref).max() <= 2e-6 * max(1.0, np.abs(ref).max())
    assert np.abs(ref).max() > 1.0


def _fd_check(w, out, lights, adjoint_fn, n_albedo=40, seed=0, bounces=1):
    rec = out["records"]
    target = A.shade(rec, w.meta["target_albedo"], lights, w.spp, bounces)
    alb = w.albedo.astype(np.float64)

    def loss_of(a, ls):
        return A.loss(A.shade(rec, a, ls, w.spp, bounces), target)

    g = A.loss_grad(A.shade(rec, alb, lights, w.spp, bounces), target)
    d_alb, d_l = adjoint_fn(rec, alb, lights, w.spp, bounces, g)
    rng = np.random.default_rng(seed)
    seen = np.unique(rec["prim"][rec["prim"] >= 0])
    picks = [(int(t), int(c)) for t, c in zip(rng.choice(seen, n_albedo), rng.integers(0, 3, n_albedo))]
    got, fd = [], []
    h = 1e-4
    for t, c in picks:
        ap, am = alb.copy(), alb.copy()
        ap[t, c] += h
        am[t, c] -= h
        fd.append((loss_of(ap, lights) - loss_of(am, lights)) / (2 * h))
        got.append(d_alb[t, c])
    for l in range(len(lights)):
        for k in range(6):
            hk = 1e-5 if k < 3 else 1e-2
            lp = [list(map(float, lights[i][0])) + list(map(float, lights[i][1])) for i in range(len(lights))]

            def with_delta(dv):
                q = [row[:] for row in lp]
                q[l][k] += dv
                return [(tuple(q[i][:3]), tuple(q[i][3:]), lights[i][2]) for i in range(len(lights))]

            fd.append((loss_of(alb, with_delta(hk)) - loss_of(alb, with_delta(-hk))) / (2 * hk))
            got.append(d_l[l, k])
    return np.array(got), np.array(fd)


def test_adjoint_matches_finite_differences(frame):
    w, out, lights = frame
    got, fd = _fd_check(w, out, lights, A.adjoint)
    scale = np.abs(fd).max()
    assert scale > 0
    err = np.abs(got - fd)
    assert (err <= 1e-5 * np.abs(fd) + 1e-7 * scale).all(), (err / np.maximum(np.abs(fd), 1e-30)).max()
    assert (np.abs(fd) > 1e-3 * scale).sum() > 40  # most checked parameters matter


def test_finite_differences_catch_a_dropped_term(frame, monkeypatch):
    """Plant a mistake in the oracle's adjoint (the falloff term of the light
    position gradient dropped): the check must fail."""
    w, out, lights = frame
    orig = A._light_grads

    def dropped(t, wv, kd, n, d):
        return orig(t, wv, kd, n, np.zeros_like(d))

    monkeypatch.setattr(A, "_light_grads", dropped)
    got, fd = _fd_check(w, out, lights, A.adjoint, n_albedo=4)
    err = np.abs(got - fd)
    assert (err > 1e-3 * np.abs(fd)).any()


@pytest.fixture(scope="module")
def frame2(O):
    from raptor_amd import workloads

    w = workloads.diff_terrain(grid=24, x_res=32, y_res=24, spp=2, seed=6, zoom=1.0, amplitude=0.3)
    out = O.OracleScene(w.abc).render(w.albedo, w.camera, w.lights, bounce=2, spp=w.spp, seed=w.seed, nthreads=4,
                                      records=True)
    lights = [(l.pos, l.rgb, l.d) for l in w.lights]
    return w, out, lights


def test_two_bounce_frame(frame2, frame):
    """Two bounces: the first bounce's vertex is the one-bounce frame's, the
    float64 shading reproduces the float frame, and the second vertex adds
    light."""
    w, out, lights = frame2
    rec = out["records"]
    assert rec["prim"].shape[0] == 3
    assert (rec["prim"][2] >= 0).sum() > 20
    assert ((rec["prim"][2] >= 0) <= (rec["prim"][1] >= 0)).all()
    img = A.shade(rec, w.albedo, lights, w.spp, 2)
    ref = out["rgb"].astype(np.float64)
    assert np.abs(img - ref).max() <= 2e-6 * max(1.0, np.abs(ref).max())
    one = A.shade(rec, w.albedo, lights, w.spp, 1)
    assert (img >= one - 1e-9).all() and (img > one + 1e-6).any()


def test_two_bounce_adjoint_matches_finite_differences(frame2):
    w, out, lights = frame2
    got, fd = _fd_check(w, out, lights, A.adjoint, bounces=2, seed=3)
    scale = np.abs(fd).max()
    err = np.abs(got - fd)
    assert (err <= 1e-5 * np.abs(fd) + 1e-7 * scale).all(), (err / np.maximum(np.abs(fd), 1e-30)).max()
