"""Error of the FP16 leaf-local payload against the FP64 quadrature oracle
on the large configurations.

The renderer's image on the FP16 payload equals the oracle's on the decoded
payload bit for bit (test_gpu_fp16.py), and the oracle's Monte Carlo image
converges to its FP64 quadrature (test_oracle_quadrature.py).  So the
converged FP16 image is the quadrature of the decoded scene, and
    quadrature(decoded FP16 payload) - quadrature(FP32 scene)
is exactly what the payload adds to the converged image error, free of
Monte Carlo noise.  It is evaluated on a crop of pixels (a grid over the
middle half of the frame) with the light's ambient term at 1, so that the
quadrature needs no shadow rays (their optical depth is an O(N) sum per
quadrature node, out of reach at these particle counts).

Configurations: the 8,388,608-particle mixture of the main workload
(1920x1080, orbit frame 1, a 6 x 4 pixel grid over the middle of the
frame) and the 67,108,864-particle hierarchical clusters (3 levels of 16
children, power-law radii in [5e-4, 1e-2], 3840x2160, 8 spp delta tracking
with the FP16 payload), whose full frame is also rendered and timed, and
whose 8-spp image is compared with the quadrature on 24 of the pixels an
independent 8-spp image reached, against its own Monte Carlo noise."""
import time

import numpy as np
import pytest

import aarbf
from aarbf import scenes
import oracle

pytestmark = pytest.mark.gpu


def _crop(W, H, nx, ny):
    xs = np.linspace(0.25 * W, 0.75 * W, nx).astype(np.int32)
    ys = np.linspace(0.25 * H, 0.75 * H, ny).astype(np.int32)
    return np.array([(x, y) for y in ys for x in xs], np.int32)


def _render(ctx, sc, cam, W, H, unit, spp, frame, light):
    import torch
    img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    acc = torch.zeros_like(img)
    torch.cuda.synchronize()
    t0 = time.time()
    ctx.render(sc.raygen_data(img, acc, W, H), sc.push_constants(cam, unit, frame_id=frame, light=light), W, H, spp)
    ctx.sync()
    return acc, time.time() - t0


def _payloads(ctx, sc, cam, unit):
    """Decoded FP32 and FP16 payloads of the scene (baked by a tiny render)."""
    out = []
    try:
        for fmt in (ctx.PAYLOAD_FP32, ctx.PAYLOAD_FP16):
            ctx.set_payload(fmt)
            _render(ctx, sc, cam, 8, 8, unit, 1, 0, (0.0, 0.0, 1.0))
            out.append(ctx.payload_read(sc.n))
    finally:
        ctx.set_payload(ctx.PAYLOAD_FP32)
    return out


def _payload_error(sc, tfs, cam, W, H, unit, rbf, q32, q16, px, threads=16):
    op = sc.particles.cpu().numpy()
    orad = sc.radii.cpu().numpy() if sc.radii is not None else None
    kw = dict(light=(0.0, 0.0, 1.0))
    e32 = oracle.Render(op, orad, sc.aabb, *tfs, cam, W, H, unit, rbf, baked=q32, **kw).expected(px, threads)
    e16 = oracle.Render(op, orad, sc.aabb, *tfs, cam, W, H, unit, rbf, baked=q16, **kw).expected(px, threads)
    d = e16 - e32
    return e32, e16, float(np.sqrt(np.mean(d ** 2))), float(np.abs(d).max())


def test_fp16_payload_error_mixture_8m(gpu_ctx):
    wl = scenes.mixture_8m()
    W, H = wl.width, wl.height
    tfs = (wl.colormap, wl.radiusmap, wl.densitymap)
    sc = aarbf.Scene.prepare(gpu_ctx, wl.xyzw, *tfs, wl.rbf_radius, radii=wl.radii, ppl=1, branching=8)
    lf, la, lu = scenes.orbit_camera(1, wl.frames, 1.6)
    cam = aarbf.camera_basis(lf, la, lu, wl.cos_fovy, W, H)
    q32, q16 = _payloads(gpu_ctx, sc, cam, wl.unit)
    px = _crop(W, H, 6, 4)
    e32, e16, rmse, emax = _payload_error(sc, tfs, cam, W, H, wl.unit, wl.rbf_radius, q32, q16, px)
    lit = float((e32.sum(1) > 0).mean())
    print(f"\n8M mixture, FP16 payload vs FP32 (FP64 quadrature, {len(px)} pixels, {lit:.0%} lit): "
          f"RMSE {rmse:.2e}, max {emax:.2e}, mean pixel {e32.mean():.3f}")
    assert lit > 0.5
    assert np.isfinite(e16).all()
    assert rmse <= 1e-3 and emax <= 5e-3


def test_precision_config_hierarchical_67m(gpu_ctx):
    import torch
    n, W, H, spp, unit = 67108864, 3840, 2160, 8, 0.0004
    xyzw, radii = scenes.hierarchical_clusters(n=n, seed=13, device="cuda")
    tfs = (scenes.reference_tf("colormap"), scenes.reference_tf("radiusmap"), scenes.reference_tf("densitymap"))
    rbf = float(radii.max())
    sc = aarbf.Scene.prepare(gpu_ctx, xyzw, *tfs, rbf, radii=radii, ppl=1, branching=8)
    del xyzw, radii
    torch.cuda.empty_cache()
    lf, la, lu = scenes.orbit_camera(1, 32, 1.4)
    cam = aarbf.camera_basis(lf, la, lu, scenes.fovy_scale(45.0), W, H)
    q32, q16 = _payloads(gpu_ctx, sc, cam, unit)
    gpu_ctx.set_payload(gpu_ctx.PAYLOAD_FP16)
    try:
        _render(gpu_ctx, sc, cam, W, H, unit, spp, 0, (0.0, 0.0, 0.5))  # warm-up
        acc, sec = _render(gpu_ctx, sc, cam, W, H, unit, spp, 8, (0.0, 0.0, 0.5))
        assert torch.isfinite(acc).all()
        print(f"\n67M hierarchical, FP16 payload, {W}x{H} {spp} spp delta tracking (NEE, ambient 0.5): "
              f"{sec * 1e3:.0f} ms, {W * H * spp / sec / 1e6:.1f} Msamples/s")
        a, _ = _render(gpu_ctx, sc, cam, W, H, unit, spp, 0, (0.0, 0.0, 1.0))
        b, _ = _render(gpu_ctx, sc, cam, W, H, unit, spp, 40000, (0.0, 0.0, 1.0))
        c, _ = _render(gpu_ctx, sc, cam, W, H, unit, spp, 20000, (0.0, 0.0, 1.0))
    finally:
        gpu_ctx.set_payload(gpu_ctx.PAYLOAD_FP32)
    # the clusters cover a small part of the frame: 24 pixels spread over
    # those a third, independent 8-spp image reached (so the selection does
    # not favour the samples of the two compared images)
    hit = torch.nonzero(c.reshape(-1, 4)[:, :3].sum(1) > 0).flatten().cpu().numpy()
    del c
    assert hit.size >= 24
    pick = hit[np.linspace(0, hit.size - 1, 24).astype(np.int64)]
    px = np.stack([pick % W, pick // W], axis=1).astype(np.int32)
    e32, e16, rmse, emax = _payload_error(sc, tfs, cam, W, H, unit, rbf, q32, q16, px)
    lit = float((e32.sum(1) > 0).mean())
    ix = torch.from_numpy(px[:, 1].astype(np.int64) * W + px[:, 0])
    a = a.reshape(-1, 4)[ix.cuda(), :3].double().cpu().numpy()
    b = b.reshape(-1, 4)[ix.cuda(), :3].double().cpu().numpy()
    img_rmse = float(np.sqrt(np.mean((a - e32) ** 2)))
    img_max = float(np.abs(a - e32).max())
    noise = float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(2.0))
    print(f"67M hierarchical, FP16 payload vs FP32 (FP64 quadrature, {len(px)} pixels, {lit:.0%} lit): "
          f"RMSE {rmse:.2e}, max {emax:.2e}; {spp}-spp image vs the FP32 quadrature: RMSE {img_rmse:.2e}, "
          f"max {img_max:.2e}, Monte Carlo noise {noise:.2e}")
    assert lit > 0.5
    assert rmse <= 1e-3 and emax <= 5e-3
    assert img_rmse <= 2.0 * noise + 1e-3
