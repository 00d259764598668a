"""Converged delta-tracking images against the FP64 quadrature of their
expectation (oracle.Render.expected).  Two independent GPU estimates at N
spp give the Monte Carlo noise level; the error against the quadrature
must be of that size (an estimator bias would show as a larger error), and
below the stated absolute RMSE."""
import numpy as np
import pytest

import aarbf
from aarbf import scenes
from helpers import default_camera, random_scene
import oracle

pytestmark = pytest.mark.gpu

SPP = 2048
RMSE_BOUND = 4e-3  # stated RMSE vs the FP64 quadrature at 2048 spp on these scenes


def _render(ctx, sc, cam, W, H, unit, frame, light, mode=0):
    import torch
    ctx.set_render_mode(mode)
    img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    acc = torch.zeros_like(img)
    pc = sc.push_constants(cam, unit, frame_id=frame, accum_id=1, light=light)
    try:
        ctx.render(sc.raygen_data(img, acc, W, H), pc, W, H, SPP)
        ctx.sync()
    finally:
        ctx.set_render_mode(0)
    return acc.cpu().numpy()[..., :3].astype(np.float64)


@pytest.mark.parametrize("amb,W,mode", [(1.0, 24, 0), (0.5, 14, 0), (1.0, 24, 2), (0.5, 14, 2)])
def test_tracking_converges_to_quadrature(gpu_ctx, amb, W, mode):
    """mode 0: the reference's estimator; mode 2: local-majorant tracking."""
    H = W
    xyzw, radii = random_scene(400, seed=31, extent=0.8, r=(0.05, 0.12))
    tfs = (scenes.ramp_table(256), scenes.constant_table(n=4), scenes.ramp_table(64))
    sc = aarbf.Scene.prepare(gpu_ctx, xyzw, *tfs, 0.12, radii=radii)
    cam = default_camera(W, H, look_from=(0.3, 0.2, 1.6))
    unit, light = 0.01, (0.3, 0.2, amb)
    a = _render(gpu_ctx, sc, cam, W, H, unit, 0, light, mode)
    b = _render(gpu_ctx, sc, cam, W, H, unit, 40000, light, mode)
    op = sc.particles.cpu().numpy()
    orad = sc.radii.cpu().numpy()
    e = oracle.Render(op, orad, sc.aabb, *tfs, cam, W, H, unit, 0.12, light=light).expected()
    assert (e > 0.01).mean() > 0.3
    rmse = float(np.sqrt(np.mean((a - e) ** 2)))
    noise = float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(2.0))
    print(f"mode {mode} amb {amb}: RMSE vs FP64 quadrature {rmse:.2e}, Monte Carlo noise {noise:.2e} at {SPP} spp")
    assert rmse < 1.5 * noise + 1e-4
    assert rmse < RMSE_BOUND
