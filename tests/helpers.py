"""Scene builders shared by the oracle and GPU tests."""
import numpy as np

import aarbf
from aarbf import scenes

import oracle


def random_scene(n=4096, seed=5, extent=1.0, r=(0.02, 0.06), per_particle=True):
    rng = np.random.default_rng(seed)
    xyzw = np.empty((n, 4), np.float32)
    xyzw[:, :3] = rng.uniform(-extent / 2, extent / 2, (n, 3))
    xyzw[:, 3] = rng.normal(size=n)
    radii = rng.uniform(r[0], r[1], n).astype(np.float32) if per_particle else None
    return xyzw, radii


def oracle_render(p, radii, aabb, tfs, cam, W, H, unit, rbf, **kw):
    cm, rm, dm = tfs
    return oracle.Render(p, radii, aabb, cm, rm, dm, cam, W, H, unit, rbf, **kw)


def default_camera(W, H, look_from=(0.0, 0.0, 3.0), look_at=(0.0, 0.0, 0.0), fovy=45.0):
    return aarbf.camera_basis(look_from, look_at, (0.0, 1.0, 0.0), scenes.fovy_scale(fovy), W, H)


def near_ray(pos, r_eff, o, d, t0, t1, margin=1e-3):
    """Indices of particles whose sphere can meet the segment o + t d,
    t in [t0, t1] (float64 geometry, radii widened by `margin` relative
    plus 1e-6)."""
    rel = pos - o
    t = rel @ d
    perp = np.linalg.norm(rel - t[:, None] * d[None, :], axis=1)
    rr = r_eff * (1.0 + margin) + 1e-6
    return np.nonzero((perp < rr) & (t > t0 - rr) & (t < t1 + rr))[0]


def pixel_ray(cam, W, H, px, py, aabb, rbf):
    s = np.array([(px + 0.5) / W, (py + 0.5) / H])
    d = cam.dir_00.astype(np.float64) + s[0] * cam.dir_du + s[1] * cam.dir_dv
    d /= np.linalg.norm(d)
    o = cam.pos.astype(np.float64)
    lo, hi = np.asarray(aabb[:3], np.float64) - 2 * rbf, np.asarray(aabb[3:6], np.float64) + 2 * rbf
    with np.errstate(divide="ignore"):
        t1, t2 = (lo - o) / d, (hi - o) / d
    return o, d, max(0.0, np.minimum(t1, t2).max()), np.maximum(t1, t2).min()
