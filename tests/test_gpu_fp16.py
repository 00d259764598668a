"""FP16 leaf-local payload (aarbf_set_payload(ctx, 1)): the decoded payload
keeps every sphere inside the original one (so the FP32-built BVH stays
valid) with the stated precision (positions within half a chunk step, the
radius within one step of its chunk's radius code, alpha within half a
16-bit unorm step, rgb^2.2 within 2^-11 relative), the renderer's images
equal the oracle's bit for bit on the decoded scene, and converged images
stay within the stated RMSE of the FP64 quadrature of the original scene."""
import numpy as np
import pytest

import aarbf
from aarbf import scenes
from helpers import default_camera, random_scene
import oracle

pytestmark = pytest.mark.gpu


def _render(ctx, sc, cam, W, H, unit, spp=1, frame=0, light=(0.3, 0.2, 0.4), mode=0):
    import torch
    img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    acc = torch.zeros_like(img)
    ctx.set_render_mode(mode)
    try:
        ctx.render(sc.raygen_data(img, acc, W, H), sc.push_constants(cam, unit, frame_id=frame, light=light), W, H,
                   spp)
        ctx.sync()
    finally:
        ctx.set_render_mode(0)
    return img.cpu().numpy(), acc.cpu().numpy()


def _setup(ctx, n=3000, seed=6):
    xyzw, radii = random_scene(n, seed=seed, extent=1.0, r=(0.03, 0.09))
    tfs = (scenes.ramp_table(256), scenes.constant_table(n=4), scenes.ramp_table(64))
    return aarbf.Scene.prepare(ctx, xyzw, *tfs, 0.09, radii=radii), tfs


def test_fp16_payload_bounds(gpu_ctx):
    sc, tfs = _setup(gpu_ctx)
    W = H = 16
    cam = default_camera(W, H, look_from=(0.4, 0.3, 2.2))
    try:
        _render(gpu_ctx, sc, cam, W, H, 0.004)
        ref = gpu_ctx.payload_read(sc.n)
        gpu_ctx.set_payload(gpu_ctx.PAYLOAD_FP16)
        _render(gpu_ctx, sc, cam, W, H, 0.004)
        q = gpu_ctx.payload_read(sc.n)
    finally:
        gpu_ctx.set_payload(gpu_ctx.PAYLOAD_FP32)
    shift = np.linalg.norm(q[:, :3].astype(np.float64) - ref[:, :3], axis=1)
    # chunk steps: extent of 32 consecutive centres (the last chunk ragged) / 65535 per axis
    bound = np.empty(sc.n)
    for c0 in range(0, sc.n, 32):
        c = ref[c0:c0 + 32, :3].astype(np.float64)
        bound[c0:c0 + 32] = 0.5 * np.linalg.norm((c.max(0) - c.min(0)) / 65535.0) * 1.01 + 1e-7
    assert sc.n % 32 != 0 and np.all(shift <= bound)
    live = ref[:, 3] > 0
    assert np.all(q[live, 3] + shift[live] <= ref[live, 3])          # decoded sphere inside the original
    assert np.all(q[~live, 3] == 0)
    # radius: 16-bit code over the chunk's range of shrunk radii, rounded down
    shrunk = np.where(live, ref[:, 3].astype(np.float64) - shift, np.nan)
    lo = np.empty(sc.n)
    for c0 in range(0, sc.n, 32):
        c = shrunk[c0:c0 + 32]
        step = (np.nanmax(c) - np.nanmin(c)) / 65534.0 if np.isfinite(c).any() else 0.0
        lo[c0:c0 + 32] = c - 1.01 * step - 1e-6 * c
    assert np.all(q[live, 3] >= lo[live]) and np.all(q[live, 3] > 0)
    # alpha as 16-bit unorm (alpha^3 = its cube), rgb^2.2 as binary16 with
    # subnormals (the decode must not flush them: absolute error 2^-25 below 2^-14)
    np.testing.assert_allclose(np.cbrt(q[:, 4]), np.cbrt(ref[:, 4]), rtol=1e-6, atol=0.51 / 65535)
    np.testing.assert_allclose(q[:, 5:], ref[:, 5:], rtol=2 ** -11, atol=2 ** -25)
    dark = (ref[:, 5:] > 0) & (ref[:, 5:] < 2 ** -14)
    assert np.all(q[:, 5:][dark] > 0)
    assert np.all((q[:, 4] > 0) == (ref[:, 4] > 0))  # no weight rounds to zero


@pytest.mark.parametrize("mode", [0, 1])
def test_fp16_render_matches_oracle_on_decoded_scene(gpu_ctx, mode):
    sc, tfs = _setup(gpu_ctx)
    W, H = 48, 40
    cam = default_camera(W, H, look_from=(0.4, 0.3, 2.2))
    try:
        gpu_ctx.set_payload(gpu_ctx.PAYLOAD_FP16)
        img, acc = _render(gpu_ctx, sc, cam, W, H, 0.004, spp=2, frame=5, mode=mode)
        q = gpu_ctx.payload_read(sc.n)
    finally:
        gpu_ctx.set_payload(gpu_ctx.PAYLOAD_FP32)
    op = sc.particles.cpu().numpy()
    r = oracle.Render(op, sc.radii.cpu().numpy(), sc.aabb, *tfs, cam, W, H, 0.004, 0.09, frame_id=5, accum_id=1,
                      light=(0.3, 0.2, 0.4), baked=q)
    oimg, oacc, _ = r.run(spp=2, threads=8)
    assert (oacc[..., :3] > 0).mean() > 0.05
    assert np.array_equal(acc, oacc) and np.array_equal(img, oimg)


def test_fp16_converges_to_quadrature_of_the_fp32_scene(gpu_ctx):
    W = H = 20
    xyzw, radii = random_scene(400, seed=31, extent=0.8, r=(0.05, 0.12))
    tfs = (scenes.ramp_table(256), scenes.constant_table(n=4), scenes.ramp_table(64))
    sc = aarbf.Scene.prepare(gpu_ctx, xyzw, *tfs, 0.12, radii=radii)
    cam = default_camera(W, H, look_from=(0.3, 0.2, 1.6))
    try:
        gpu_ctx.set_payload(gpu_ctx.PAYLOAD_FP16)
        _, a = _render(gpu_ctx, sc, cam, W, H, 0.01, spp=2048, frame=0, light=(0.3, 0.2, 1.0))
        _, b = _render(gpu_ctx, sc, cam, W, H, 0.01, spp=2048, frame=40000, light=(0.3, 0.2, 1.0))
    finally:
        gpu_ctx.set_payload(gpu_ctx.PAYLOAD_FP32)
    e = oracle.Render(sc.particles.cpu().numpy(), sc.radii.cpu().numpy(), sc.aabb, *tfs, cam, W, H, 0.01, 0.12,
                      light=(0.3, 0.2, 1.0)).expected()
    a, b = a[..., :3].astype(np.float64), b[..., :3].astype(np.float64)
    rmse = float(np.sqrt(np.mean((a - e) ** 2)))
    noise = float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(2.0))
    print(f"FP16 payload: RMSE vs FP64 quadrature of the FP32 scene {rmse:.2e}, Monte Carlo noise {noise:.2e}")
    assert np.isfinite(a).all() and np.isfinite(b).all()
    assert rmse < 1.5 * noise + 1e-3


@pytest.mark.parametrize("n", [1, 33])
def test_fp16_small_and_ragged_scenes(gpu_ctx, n):
    xyzw, radii = random_scene(n, seed=40 + n, extent=0.3, r=(0.1, 0.2))
    tfs = (scenes.ramp_table(256), scenes.constant_table(n=4), scenes.ramp_table(64))
    sc = aarbf.Scene.prepare(gpu_ctx, xyzw, *tfs, 0.2, radii=radii)
    W, H = 24, 20
    cam = default_camera(W, H, look_from=(0.2, 0.1, 1.2))
    try:
        gpu_ctx.set_payload(gpu_ctx.PAYLOAD_FP16)
        img, acc = _render(gpu_ctx, sc, cam, W, H, 0.005, spp=4, frame=3)
        q = gpu_ctx.payload_read(sc.n)
    finally:
        gpu_ctx.set_payload(gpu_ctx.PAYLOAD_FP32)
    r = oracle.Render(sc.particles.cpu().numpy(), sc.radii.cpu().numpy(), sc.aabb, *tfs, cam, W, H, 0.005, 0.2,
                      frame_id=3, accum_id=1, light=(0.3, 0.2, 0.4), baked=q)
    oimg, oacc, _ = r.run(spp=4, threads=8)
    assert np.isfinite(acc).all()
    assert np.array_equal(acc, oacc) and np.array_equal(img, oimg)


def test_fp16_long_queues_with_and_without_statistics(gpu_ctx):
    """A 160x128 view (20,480 rays, above the 16,384 from which the
    one-lane gather with persistent lanes takes the queue) with the FP16
    payload: the payload-specific kernels (statistics off) and the
    statistics kernels (which read the payload kind at run time) render the
    same image bit for bit, equal to the oracle on the decoded scene on a
    pixel grid."""
    sc, tfs = _setup(gpu_ctx, n=6000, seed=12)
    W, H = 160, 128
    cam = default_camera(W, H, look_from=(0.35, 0.25, 2.0))
    try:
        gpu_ctx.set_payload(gpu_ctx.PAYLOAD_FP16)
        img, acc = _render(gpu_ctx, sc, cam, W, H, 0.004, spp=2, frame=9)
        gpu_ctx.set_stats(True)
        try:
            img_s, acc_s = _render(gpu_ctx, sc, cam, W, H, 0.004, spp=2, frame=9)
        finally:
            gpu_ctx.set_stats(False)
            gpu_ctx.read_stats(reset=True)
            gpu_ctx.read_histogram(reset=True)
        q = gpu_ctx.payload_read(sc.n)
    finally:
        gpu_ctx.set_payload(gpu_ctx.PAYLOAD_FP32)
    assert (acc[..., :3] > 0).mean() > 0.1
    assert np.array_equal(acc, acc_s) and np.array_equal(img, img_s)
    px = np.array([(x, y) for y in range(1, H, 11) for x in range(2, W, 13)], np.int32)
    r = oracle.Render(sc.particles.cpu().numpy(), sc.radii.cpu().numpy(), sc.aabb, *tfs, cam, W, H, 0.004, 0.09,
                      frame_id=9, accum_id=1, light=(0.3, 0.2, 0.4), baked=q)
    oimg, oacc, _ = r.run(pixels=px, spp=2, threads=8)
    sel = (px[:, 1], px[:, 0])
    assert np.array_equal(acc[sel], oacc[sel]) and np.array_equal(img[sel], oimg[sel])
