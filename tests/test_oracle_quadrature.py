"""The FP64 quadrature of the delta-tracking expectation (what the
renderer's images converge to), pinned by a closed form (one sphere of
constant density: E = albedo (1 - exp(-c L / unit)) for a chord of length
L) and by an independent brute-force midpoint integration in numpy for
overlapping particles with NEE shadow rays."""
import numpy as np

from aarbf import scenes
from helpers import default_camera
import oracle


def _dirs(cam, W, H):
    out = np.empty((H, W, 3), np.float32)
    for py in range(H):
        for px in range(W):
            sx = np.float32((px + 0.5) / W)
            sy = np.float32((py + 0.5) / H)
            d = (np.float32(cam.dir_00) + sx * np.float32(cam.dir_du)) + sy * np.float32(cam.dir_dv)
            d = d.astype(np.float32)
            n = np.sqrt(np.float32(d[0] * d[0] + d[1] * d[1]) + np.float32(d[2] * d[2]))
            out[py, px] = d / n
    return out


def test_constant_density_sphere_closed_form():
    W = H = 15
    cam = default_camera(W, H, look_from=(0.0, 0.0, 1.2))
    c, unit, R = 0.3, 0.01, 0.25
    xyzw = np.array([[0.02, -0.01, 0.0, 0.5]], np.float32)
    tfs = (scenes.constant_table((0.7, 0.4, 0.2, 0.9), n=1), scenes.constant_table(n=4),
           scenes.constant_table((c, 0, 0, 0), n=1))
    aabb = np.array([0.02, -0.01, 0, 0.02, -0.01, 0, 0, 1], np.float32)
    r = oracle.Render(xyzw, None, aabb, *tfs, cam, W, H, unit, R, light=(0, 0, 1.0))
    e = r.expected()
    dirs = _dirs(cam, W, H).astype(np.float64)
    o = np.array(cam.pos, np.float64)
    v = xyzw[0, :3].astype(np.float64) - o
    tc = dirs @ v
    h2 = v @ v - tc ** 2
    L = 2 * np.sqrt(np.maximum(np.float64(np.float32(R)) ** 2 - h2, 0.0))
    cm = oracle.tf_sample(tfs[0], 0.5).astype(np.float64)
    alb = (np.float64(np.float32(cm[:3] ** 2.2))) ** (1 / 2.2)
    exp = alb[None, None, :] * (1 - np.exp(-c * L / unit))[..., None]
    assert (L > 0).sum() > 10
    np.testing.assert_allclose(e, exp, rtol=2e-6, atol=1e-9)


def _brute(xyzw, radii, tfs, cam, W, H, unit, rbf, amb, light_dir, aabb, pix):
    cm, rm, dm = tfs
    dirs = _dirs(cam, W, H)
    o = np.array(cam.pos, np.float64)
    cols = np.array([oracle.tf_sample(cm, w) for w in xyzw[:, 3]], np.float64)
    rad = radii.astype(np.float64) * np.array([oracle.tf_sample(rm, w)[0] for w in xyzw[:, 3]])
    a3 = cols[:, 3] ** 3
    c22 = cols[:, :3] ** 2.2
    P = xyzw[:, :3].astype(np.float64)
    lo = aabb[:3] - 2 * rbf
    hi = aabb[3:6] + 2 * rbf

    def tf(table, u):
        n = len(table)
        x = np.clip(u * n - 0.5, -1, n)
        f = np.floor(x)
        i0 = f.astype(int)
        a = np.clip(i0, 0, n - 1)
        b = np.clip(i0 + 1, 0, n - 1)
        return table[a, 0] + (x - f) * (table[b, 0] - table[a, 0])

    def field(X):
        d = np.linalg.norm(X[:, None, :] - P[None], axis=2)
        cov = d < rad[None]
        w = np.where(cov, np.exp(-0.5 * (3 * d / rad[None]) ** 2) * a3[None], 0.0)
        rho = w.sum(1)
        sig = np.where(cov.any(1), np.clip(tf(np.asarray(dm, np.float64).reshape(-1, 4), rho), 0, 1) / unit, 0)
        alb = (w @ c22) / np.maximum(rho, 1e-300)[:, None]
        return sig, alb ** (1 / 2.2)

    def exit_t(x, d):
        with np.errstate(divide="ignore"):
            t1, t2 = (lo - x) / d, (hi - x) / d
        return np.min(np.maximum(t1, t2))

    out = []
    for px, py in pix:
        d = dirs[py, px].astype(np.float64)
        tx = exit_t(o, d)
        n = 12000
        t = (np.arange(n) + 0.5) * tx / n
        sig, alb = field(o[None] + t[:, None] * d[None])
        tau = np.cumsum(sig) * (tx / n) - 0.5 * sig * (tx / n)
        wgt = sig * np.exp(-tau) * (tx / n)
        shade = np.full(n, amb)
        for k in np.nonzero(wgt > 1e-12)[0]:
            x = o + t[k] * d
            sx = exit_t(x, light_dir)
            m = 1500
            s = (np.arange(m) + 0.5) * sx / m
            ss, _ = field(x[None] + s[:, None] * light_dir[None])
            shade[k] = amb + (1 - amb) * np.exp(-ss.sum() * sx / m)
        out.append((wgt[:, None] * alb * shade[:, None]).sum(0))
    return np.array(out)


def test_overlapping_particles_with_shadows_match_brute_force():
    W, H = 16, 12
    cam = default_camera(W, H, look_from=(0.3, 0.2, 1.0))
    xyzw = np.array([[0.0, 0.0, 0.0, 0.3], [0.08, 0.05, -0.1, 0.8], [-0.05, 0.1, 0.15, 0.6]], np.float32)
    radii = np.array([0.25, 0.2, 0.18], np.float32)
    tfs = (scenes.ramp_table(256), scenes.constant_table(n=4), scenes.ramp_table(64))
    aabb = np.concatenate([xyzw[:, :3].min(0), xyzw[:, :3].max(0), [0, 1]]).astype(np.float32)
    light = (0.3, 0.2, 0.4)
    unit = 0.02
    r = oracle.Render(xyzw, radii, aabb, *tfs, cam, W, H, unit, 0.25, light=light)
    e = r.expected()
    ld = oracle.light_direction(light[0], light[1]).astype(np.float64)
    lit = np.argwhere(e[..., 0] > 0.01)
    assert len(lit) > 20
    pix = [(int(x), int(y)) for y, x in lit[np.linspace(0, len(lit) - 1, 3).astype(int)]]
    b = _brute(xyzw, radii, tfs, cam, W, H, unit, 0.25, light[2], ld, aabb, pix)
    got = np.array([e[py, px] for px, py in pix])
    assert got.min() > 0.01
    np.testing.assert_allclose(got, b, rtol=2e-3, atol=2e-5)
