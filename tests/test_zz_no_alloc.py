"""aarbf_render_rbf_spp_noalloc (the render entry point that allocates
nothing; its baked particle payload lives in a static per-device array)
against aarbf_render_rbf_spp on the same inputs, bit for bit: two calls in a
row, the second on a smaller scene and image, for both payload formats; and
the rejection of scenes above its particle limit."""
import numpy as np
import pytest

import aarbf
from aarbf import scenes
from helpers import default_camera, random_scene

pytestmark = pytest.mark.gpu


def _render(ctx, fn, sc, cam, W, H, spp, blue, stbn):
    import torch
    img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    acc = torch.full((H, W, 4), 0.25, dtype=torch.float32, device="cuda")  # accumID 2 blends it in
    pc = sc.push_constants(cam, 0.004, frame_id=9, accum_id=2, blue_noise=blue, light=(0.3, 0.2, 0.4))
    fn(sc.raygen_data(img, acc, W, H, stbn=stbn), pc, W, H, spp)
    ctx.sync()
    return img.cpu().numpy(), acc.cpu().numpy()


@pytest.mark.parametrize("payload", [0, 1])
def test_noalloc_equals_render_two_calls(gpu_ctx, payload):
    tfs = (scenes.ramp_table(256), scenes.constant_table(n=4), scenes.ramp_table(64))
    atlas = scenes.synthetic_stbn(seed=6)
    stbn = gpu_ctx.stbn(atlas)
    gpu_ctx.set_payload(payload)
    try:
        # first call: 6,000 particles, 64x48, 2 spp delta tracking; second:
        # 1,500 particles (a smaller payload over the first call's), 40x32,
        # 1 spp stochastic marching
        for n, W, H, spp, blue in ((6000, 64, 48, 2, False), (1500, 40, 32, 1, True)):
            print(f"noalloc call: {n} particles, {W}x{H}, {spp} spp, payload {payload}", flush=True)
            xyzw, radii = random_scene(n, seed=n, r=(0.03, 0.08))
            sc = aarbf.Scene.prepare(gpu_ctx, xyzw, *tfs, 0.05, radii=radii)
            cam = default_camera(W, H, look_from=(0.3, 0.25, 2.2))
            got = _render(gpu_ctx, gpu_ctx.render_noalloc, sc, cam, W, H, spp, blue, stbn)
            want = _render(gpu_ctx, gpu_ctx.render, sc, cam, W, H, spp, blue, stbn)
            assert (want[1][..., :3] > 0.25).mean() > 0.05, "scene too empty to be a check"
            assert np.array_equal(got[1], want[1]), f"accum differs ({n} particles)"
            assert np.array_equal(got[0], want[0]), f"image differs ({n} particles)"
            gpu_ctx.bvh_destroy(sc.bvh)
    finally:
        gpu_ctx.set_payload(0)


def test_noalloc_rejects_too_many_particles(gpu_ctx):
    import torch
    xyzw, radii = random_scene(100, seed=2)
    tfs = (scenes.ramp_table(16), scenes.constant_table(n=1), scenes.ramp_table(16))
    sc = aarbf.Scene.prepare(gpu_ctx, xyzw, *tfs, 0.05, radii=radii)
    img = torch.zeros((8, 8, 4), dtype=torch.float32, device="cuda")
    pc = sc.push_constants(default_camera(8, 8), 0.004)
    pc.numParticles = gpu_ctx.NOALLOC_MAX_PARTICLES + 1
    with pytest.raises(aarbf.AarbfError, match="AARBF_NOALLOC_MAX_PARTICLES"):
        gpu_ctx.render_noalloc(sc.raygen_data(img, img, 8, 8), pc, 8, 8, 1)
    gpu_ctx.bvh_destroy(sc.bvh)
