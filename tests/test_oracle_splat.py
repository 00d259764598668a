"""The oracle's restatement of ParticleSplatRayGen (deviceCodeSplat.hlsl:
29-194), pinned by closed forms: a single on-axis particle, a miss, an
empty scene, and a slab with more than 15 hits (only the first 15 in
traversal order are composited, sorted by t)."""
import numpy as np

from aarbf import scenes
from helpers import default_camera
import oracle


def _splat(xyzw, radii, rbf, unit, tfs, W=5, H=5, vis=True, ppl=1, aabb=None):
    xyzw = np.ascontiguousarray(xyzw, np.float32)
    cm, rm, dm = tfs
    if aabb is None:
        lo = xyzw[:, :3].min(0) if len(xyzw) else np.zeros(3, np.float32)
        hi = xyzw[:, :3].max(0) if len(xyzw) else np.zeros(3, np.float32)
        aabb = np.concatenate([lo, hi, [0.0, 1.0]]).astype(np.float32)
    cam = default_camera(W, H, look_from=(0.0, 0.0, 3.0))
    r = oracle.Render(xyzw, radii, aabb, cm, rm, dm, cam, W, H, unit, rbf, visualize_attributes=vis)
    leaves = oracle.gen_bounds(xyzw, ppl, rbf, scenes.constant_table(n=4), radii=radii)
    return r.splat(leaves, ppl)


def _over(a, b):
    ra = a[3] + b[3] * (1 - a[3])
    if ra == 0:
        return a
    return np.array([*((a[:3] * a[3] + b[:3] * b[3] * (1 - a[3])) / ra), ra])


def test_single_particle_on_axis():
    w = 0.6
    xyzw = np.array([[0.0, 0.0, 0.0, w]], np.float32)
    tfs = (scenes.ramp_table(256), scenes.constant_table(n=4), scenes.ramp_table(64))
    unit = 0.01
    img = _splat(xyzw, None, 0.1, unit, tfs)
    cm = oracle.tf_sample(tfs[0], w).astype(np.float64)
    a = cm[3] * oracle.tf_sample(tfs[2], 1.0)[0] * np.exp(-unit)  # drbf = 1 at the centre
    exp = _over(_over(np.zeros(4), np.array([*cm[:3], a])), np.array([0.0, 0.0, 0.0, 1.0]))
    np.testing.assert_allclose(img[2, 2], exp, rtol=2e-5)
    # the corner pixels look past the particle: background only
    np.testing.assert_array_equal(img[0, 0], np.array([0, 0, 0, 1], np.float32))


def test_density_colour_mode_on_axis():
    xyzw = np.array([[0.0, 0.0, 0.0, 0.3]], np.float32)
    tfs = (scenes.reference_tf("colormap"), scenes.reference_tf("radiusmap"), scenes.reference_tf("densitymap"))
    img = _splat(xyzw, None, 0.1, 0.01, tfs, vis=False)
    cm = oracle.tf_sample(tfs[0], 1.0).astype(np.float64)
    a = oracle.tf_sample(tfs[2], 1.0)[0] * np.exp(-0.01)
    exp = _over(_over(np.zeros(4), np.array([*cm[:3], a])), np.array([0.0, 0.0, 0.0, 1.0]))
    np.testing.assert_allclose(img[2, 2], exp, rtol=2e-5)


def test_miss_and_empty_scene_are_background():
    tfs = (scenes.ramp_table(256), scenes.constant_table(n=4), scenes.ramp_table(64))
    far = np.array([[5.0, 5.0, 0.0, 0.5]], np.float32)  # off every pixel ray of the 5x5 view
    img = _splat(far, None, 0.1, 0.01, tfs, aabb=np.array([-1, -1, -1, 6, 6, 1, 0, 1], np.float32))
    assert np.all(img == np.array([0, 0, 0, 1], np.float32))
    img = _splat(np.zeros((0, 4), np.float32), None, 0.1, 0.01, tfs,
                 aabb=np.array([-1, -1, -1, 1, 1, 1, 0, 1], np.float32))
    assert np.all(img == np.array([0, 0, 0, 1], np.float32))


def test_slab_keeps_first_fifteen_hits():
    n = 20
    xyzw = np.zeros((n, 4), np.float32)
    xyzw[:, 2] = -0.005 - 1e-5 * np.arange(n)  # t = 3.005 + 1e-5 i, all in one 0.01 slab
    xyzw[:, 3] = 0.95 - 0.03 * np.arange(n)
    tfs = (scenes.ramp_table(256), scenes.constant_table(n=4), scenes.constant_table((0.05, 0, 0, 0), n=1))
    unit = 0.01
    img = _splat(xyzw, None, 0.1, unit, tfs)

    def composite(k):
        res = np.zeros(4)
        for i in range(k):  # t increases with i: the sort keeps index order
            cm = oracle.tf_sample(tfs[0], xyzw[i, 3]).astype(np.float64)
            res = _over(res, np.array([*cm[:3], cm[3] * 0.05 * np.exp(-unit)]))
        return _over(res, np.array([0.0, 0.0, 0.0, 1.0]))

    np.testing.assert_allclose(img[2, 2], composite(15), rtol=2e-5)
    assert np.abs(composite(20) - composite(15)).max() > 1e-3
