"""The FP64 differentiable renderer (oracle/adjoint.c): the adjoint of the
expected-marching image against central finite differences of its own
forward, for positions, radii, attributes and all three transfer
functions, and the forward against a direct Monte Carlo average of the
stochastic marching estimator over its threshold."""
import numpy as np
import pytest

from aarbf import scenes
from helpers import default_camera
import oracle


def _scene(seed=3, n=40):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.3, 0.3, (n, 3))
    attr = rng.uniform(0.05, 0.95, n)
    radii = rng.uniform(0.1, 0.2, n)
    cm = scenes.ramp_table(16).astype(np.float64)
    cm[:, 0] = np.linspace(0.2, 0.9, 16)
    cm[:, 3] = np.linspace(0.3, 0.9, 16)
    dm = scenes.ramp_table(8).astype(np.float64)
    rm = np.ones((4, 4))
    rm[:, 0] = [0.8, 0.9, 1.0, 1.1]
    return dict(pos=pos, attr=attr, radii=radii, cm=cm, dm=dm, rm=rm)


def _diff(s, W=8, H=8, spp=2, estimator=0):
    cam = default_camera(W, H, look_from=(0.2, 0.1, 1.3))
    aabb = np.concatenate([s["pos"].min(0), s["pos"].max(0)])
    return oracle.Diff(s["pos"], s["attr"], s["radii"], s["cm"], s["dm"], s["rm"], cam, W, H, 0.02, 0.2, aabb,
                       spp=spp, frame_id=3, estimator=estimator)


@pytest.mark.parametrize("estimator", [0, 1])
def test_adjoint_matches_central_differences(estimator):
    """Estimator 0 (expected marching, path replay) and 1 (ratio tracking,
    stored points and a reverse recursion) against central differences of
    their own forward."""
    s = _scene()
    _d = lambda t: _diff(t, estimator=estimator)
    img = _d(s).forward()
    assert (img > 0).mean() > 0.3
    rng = np.random.default_rng(9)
    gimg = rng.normal(size=img.shape)
    g = _d(s).adjoint(gimg)
    keymap = {"pos": "pos", "radii": "radii", "attr": "attr", "cm": "colormap", "dm": "densitymap",
              "rm": "radiusmap"}
    checked = 0
    for kind, count in (("pos", 6), ("radii", 4), ("attr", 4), ("cm", 6), ("dm", 3), ("rm", 2)):
        for _ in range(count):
            a = s[kind]
            idx = tuple(int(rng.integers(0, d)) for d in a.shape)
            if kind in ("dm", "rm"):
                idx = (idx[0], 0)
            h = 1e-6
            t = {k: v.copy() for k, v in s.items()}
            t[kind][idx] += h
            lp = float((gimg * _d(t).forward()).sum())
            t[kind][idx] -= 2 * h
            lm = float((gimg * _d(t).forward()).sum())
            fd = (lp - lm) / (2 * h)
            adj = g[keymap[kind]][idx if kind not in ("dm", "rm") else idx[0]]
            assert abs(adj - fd) <= 1e-5 * max(1.0, abs(fd)), (kind, idx, adj, fd)
            checked += abs(fd) > 1e-6
    assert checked > 10


def test_forward_is_the_marching_expectation():
    """Average the stochastic marching sample over many thresholds xi: the
    mean converges to the deterministic expected-marching forward."""
    s = _scene(seed=4, n=25)
    W = H = 4
    cam = default_camera(W, H, look_from=(0.2, 0.1, 1.3))
    aabb = np.concatenate([s["pos"].min(0), s["pos"].max(0)]).astype(np.float32)
    img = _diff(s, W, H, spp=1).forward()
    # the renderer's marching estimator with explicit thresholds, through the
    # oracle's render: a uniform 'STBN' value per frame
    xyzw = np.concatenate([s["pos"], s["attr"][:, None]], 1).astype(np.float32)
    acc = np.zeros((H, W, 3))
    K = 512
    for k in range(K):
        v = int((k + 0.5) / K * 255.999)
        atlas = np.full((16 * 256, 8 * 256, 4), v, np.uint8)
        r = oracle.Render(xyzw, s["radii"].astype(np.float32), np.concatenate([aabb, [0, 1]]), s["cm"], s["rm"],
                          s["dm"], cam, W, H, 0.02, 0.2, frame_id=3, accum_id=1, blue_noise=True, stbn=atlas,
                          light=(0.0, 0.0, 1.0))
        im, _, _ = r.run(spp=1, threads=8)
        acc += im[..., :3]
    acc /= K
    # the 8-bit threshold grid and float/double differences bound the agreement
    np.testing.assert_allclose(acc, img, atol=2e-2)


def test_ratio_tracking_forward_is_the_tracking_expectation():
    """The ratio-tracking forward averaged over many tentative-point
    sequences converges to the exact expectation of the reference's delta
    tracking (the FP64 quadrature of the same scene, ambient 1)."""
    s = _scene(seed=4, n=25)
    W = H = 4
    K = 1024
    cam = default_camera(W, H, look_from=(0.2, 0.1, 1.3))
    aabb = np.concatenate([s["pos"].min(0), s["pos"].max(0)]).astype(np.float32)
    img = _diff(s, W, H, spp=K, estimator=1).forward()
    xyzw = np.concatenate([s["pos"], s["attr"][:, None]], 1).astype(np.float32)
    e = oracle.Render(xyzw, s["radii"].astype(np.float32), np.concatenate([aabb, [0, 1]]),
                      s["cm"].astype(np.float32), s["rm"].astype(np.float32), s["dm"].astype(np.float32), cam, W, H,
                      0.02, 0.2, light=(0.0, 0.0, 1.0)).expected()
    assert (e > 0.05).mean() > 0.3
    # measured: max 2.2e-3, mean 5.9e-4 of a mean colour 0.20 at K = 1024
    np.testing.assert_allclose(img, e, atol=8e-3)
    assert float(np.abs(img - e).mean()) < 2e-3


def test_per_pixel_evaluation_on_the_particles_near_the_ray():
    """The per-pixel entry points over only the particles whose spheres can
    meet the pixel's ray give the full-scene values bit for bit: the
    forward, and the adjoint of one pixel (per-particle blocks on those
    particles, zero elsewhere; transfer-function blocks)."""
    from helpers import default_camera, near_ray, pixel_ray
    rng = np.random.default_rng(8)
    n, W, H = 3000, 24, 20
    pos = rng.uniform(0.0, 1.0, (n, 3))
    attr = rng.uniform(0.05, 0.95, n)
    radii = rng.uniform(0.02, 0.06, n)
    cm, dm, rm = scenes.ramp_table(64), scenes.ramp_table(32), scenes.constant_table(n=4)
    cam = default_camera(W, H, look_from=(0.5, 0.6, 2.2), look_at=(0.5, 0.5, 0.5))
    aabb = np.concatenate([pos.min(0), pos.max(0)]).astype(np.float32)
    kw = dict(spp=2, frame_id=4)
    full = oracle.Diff(pos, attr, radii, cm, dm, rm, cam, W, H, 0.004, 0.06, aabb, **kw)
    img = full.forward()
    for x, y in [(3, 4), (12, 10), (20, 15), (7, 18)]:
        o, d, t0, t1 = pixel_ray(cam, W, H, x, y, aabb, 0.06)
        idx = near_ray(pos, radii, o, d, t0, t1)
        assert 0 < len(idx) < n // 4
        sub = oracle.Diff(pos[idx], attr[idx], radii[idx], cm, dm, rm, cam, W, H, 0.004, 0.06, aabb, **kw)
        px = np.array([[x, y]], np.int32)
        assert np.array_equal(sub.forward(pixels=px)[0], img[y, x]) and img[y, x].sum() > 0
        g = np.array([[0.3, -1.2, 0.7]])
        gi = np.zeros((H, W, 3))
        gi[y, x] = g[0]
        a_full = full.adjoint(gi)
        a_sub = sub.adjoint(g, pixels=px)
        for k in ("pos", "radii", "attr"):
            assert np.array_equal(a_full[k][idx], a_sub[k])
            rest = np.delete(a_full[k], idx, axis=0)
            assert not np.any(rest), k
        for k in ("colormap", "densitymap", "radiusmap"):
            assert np.array_equal(a_full[k], a_sub[k]), k
