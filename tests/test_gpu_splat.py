"""GPU parity of the splat path (ParticleSplatRayGen,
deviceCodeSplat.hlsl:29-194) against the oracle, through the C ABI.  The
bar is bit-exact; the 1e-4 relative tolerance is reported as well."""
import numpy as np
import pytest

import aarbf
from aarbf import scenes
from helpers import default_camera, random_scene
import oracle

pytestmark = pytest.mark.gpu

RTOL = 1e-4


def _compare(got, exp, what):
    rel = np.abs(got - exp) / np.maximum(np.abs(exp), 1e-6)
    bad = np.argwhere(rel > RTOL)
    assert bad.size == 0, f"{what}: {len(bad)} values differ, first at {bad[:3].tolist()} " \
                          f"got {got[tuple(bad[0])]} expected {exp[tuple(bad[0])]}"
    assert np.array_equal(got, exp), f"{what}: within {RTOL} but not bit-exact"


def _splat_both(ctx, xyzw, radii, tfs, rbf, unit, W, H, cam, ppl=1, K=8, vis=True, hits=False):
    import torch
    cm, rm, dm = tfs
    sc = aarbf.Scene.prepare(ctx, xyzw, cm, rm, dm, rbf, radii=radii, ppl=ppl, branching=K)
    img = torch.full((H, W, 4), -1.0, dtype=torch.float32, device="cuda")
    acc = torch.zeros_like(img)
    pc = sc.push_constants(cam, unit, visualize_attributes=vis)
    ctx.render_splat(sc.raygen_data(img, acc, W, H), pc, W, H)
    ctx.sync()
    op = sc.particles.cpu().numpy()
    orad = sc.radii.cpu().numpy() if sc.radii is not None else None
    leaves = oracle.gen_bounds(op, ppl, rbf, rm, radii=orad)
    assert np.array_equal(leaves, sc.aabbs.cpu().numpy(), equal_nan=True)
    r = oracle.Render(op, orad, sc.aabb, cm, rm, dm, cam, W, H, unit, rbf, visualize_attributes=vis)
    if hits:
        oimg, h = r.splat(leaves, ppl, hits=True)
        return img.cpu().numpy(), oimg, h
    return img.cpu().numpy(), r.splat(leaves, ppl)


CASES = [
    # name, n, ppl, K, per-particle radii, visualize_attributes
    ("ppl1-K8", 3000, 1, 8, True, True),
    ("ppl3-K4", 3000, 3, 4, True, True),
    ("global-radius-K2", 2000, 1, 2, False, True),
    ("density-colour-ppl2-K16", 2500, 2, 16, True, False),
    ("ppl1-K16", 2800, 1, 16, True, True),
    # rbfRadius at least the largest radius: the splat clip box holds every sphere
    ("rbf-covers-radii", 2600, 1, 8, True, True),
]


@pytest.mark.parametrize("name,n,ppl,K,per,vis", CASES)
def test_splat_matches_oracle(gpu_ctx, name, n, ppl, K, per, vis):
    W, H = 40, 32
    xyzw, radii = random_scene(n, seed=n + 7 * ppl, extent=1.0, r=(0.03, 0.08), per_particle=per)
    tfs = (scenes.ramp_table(256), scenes.constant_table(n=4), scenes.ramp_table(64))
    if not vis:
        tfs = (scenes.reference_tf("colormap"), scenes.reference_tf("radiusmap"), scenes.reference_tf("densitymap"))
    cam = default_camera(W, H, look_from=(0.4, 0.3, 2.2))
    rbf = float(radii.max()) if name == "rbf-covers-radii" else 0.06
    img, oimg = _splat_both(gpu_ctx, xyzw, radii, tfs, rbf, 0.005, W, H, cam, ppl, K, vis)
    assert (oimg[..., :3] > 0).mean() > 0.05, "scene too empty to be a meaningful check"
    _compare(img, oimg, name)


def test_splat_crowded_slabs(gpu_ctx):
    """A tight cluster where slabs hold more than 15 hits (the first-15
    rule), some single slab holds more than the kernel's 32-entry hit list
    (a one-slab chunk keeps its first 32), and runs of 16 slabs hold more
    than 32 (a walk over several slabs stops at its first hit past the
    list and the chunk shrinks to a quarter): the oracle's hit counts show
    each case is reached."""
    W, H = 32, 32
    rng = np.random.default_rng(21)
    n = 5000
    xyzw = np.empty((n, 4), np.float32)
    xyzw[:, :3] = rng.normal(0.0, 0.04, (n, 3))
    xyzw[:, 3] = rng.uniform(0, 1, n)
    radii = rng.uniform(0.05, 0.1, n).astype(np.float32)
    tfs = (scenes.ramp_table(256), scenes.constant_table(n=4), scenes.constant_table((0.05, 0, 0, 0), n=1))
    cam = default_camera(W, H, look_from=(0.0, 0.2, 1.2))
    img, oimg, hits = _splat_both(gpu_ctx, xyzw, radii, tfs, 0.1, 0.01, W, H, cam, hits=True)
    assert (oimg[..., :3] > 0).mean() > 0.1
    assert hits[:, 0].max() > 32, f"no slab with more than 32 hits (largest {hits[:, 0].max()})"
    assert ((hits[:, 1] > 32) & (hits[:, 0] <= 32)).any(), "no run of slabs overflowing the list without one slab doing so"
    assert (hits[:, 0] > 15).mean() > 0.05
    _compare(img, oimg, "crowded")


def test_splat_rejects_bad_unit(gpu_ctx):
    import torch
    xyzw, radii = random_scene(100, seed=1)
    tfs = (scenes.ramp_table(256), scenes.constant_table(n=4), scenes.ramp_table(64))
    sc = aarbf.Scene.prepare(gpu_ctx, xyzw, *tfs, 0.05, radii=radii)
    img = torch.zeros((8, 8, 4), dtype=torch.float32, device="cuda")
    pc = sc.push_constants(default_camera(8, 8), 0.0)
    with pytest.raises(aarbf.AarbfError):
        gpu_ctx.render_splat(sc.raygen_data(img, img, 8, 8), pc, 8, 8)
