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
