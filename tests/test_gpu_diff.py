"""The HIP differentiable renderer (aarbf_diff_forward / _backward) against
the FP64 oracle (oracle/adjoint.c, itself pinned by central finite
differences), and the torch.autograd.Function wrapper.  Tolerances: the
forward within 1e-4 absolute per pixel; each gradient block within 1e-3
relative L2 of the oracle's (float32 kernels against float64)."""
import time

import numpy as np
import pytest

import aarbf
from aarbf import scenes
from aarbf.diff import DiffRenderer
from helpers import default_camera
import oracle

pytestmark = pytest.mark.gpu


def _scene(seed=3, n=300):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.3, 0.3, (n, 3)).astype(np.float32)
    attr = rng.uniform(0.05, 0.95, n).astype(np.float32)
    radii = rng.uniform(0.06, 0.12, n).astype(np.float32)
    cm = scenes.ramp_table(16)
    cm[:, 0] = np.linspace(0.2, 0.9, 16)
    cm[:, 3] = np.linspace(0.3, 0.9, 16)
    dm = scenes.ramp_table(8)
    rm = np.ones((4, 4), np.float32)
    rm[:, 0] = [0.8, 0.9, 1.0, 1.1]
    return pos, attr, radii, cm.astype(np.float32), dm.astype(np.float32), rm


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def test_diff_forward_and_adjoint_match_oracle(gpu_ctx):
    import torch
    W = H = 24
    spp, unit = 2, 0.01
    pos, attr, radii, cm, dm, rm = _scene()
    cam = default_camera(W, H, look_from=(0.2, 0.1, 1.3))
    R = DiffRenderer(gpu_ctx, cam, W, H, unit, spp=spp, frame_id=3, rbf_radius=0.2)
    t = lambda a: torch.tensor(a, device="cuda", requires_grad=True)
    P, A, Rd, CM, DM, RM = (t(a) for a in (pos, attr, radii, cm, dm, rm))
    img = R(P, A, Rd, CM, DM, RM)
    rng = np.random.default_rng(11)
    gimg = rng.normal(size=(H, W, 3)).astype(np.float32)
    (img * torch.tensor(gimg, device="cuda")).sum().backward()
    torch.cuda.synchronize()
    aabb = np.concatenate([pos.min(0), pos.max(0)])
    D = oracle.Diff(pos.astype(np.float64), attr.astype(np.float64), radii.astype(np.float64), cm, dm, rm, cam, W, H,
                    unit, 0.2, aabb, spp=spp, frame_id=3)
    oimg = D.forward()
    assert (oimg > 0.01).mean() > 0.3
    np.testing.assert_allclose(img.detach().cpu().numpy(), oimg, atol=1e-4)
    og = D.adjoint(gimg.astype(np.float64))
    got = {"pos": P.grad, "attr": A.grad, "radii": Rd.grad, "colormap": CM.grad, "densitymap": DM.grad[:, 0],
           "radiusmap": RM.grad[:, 0]}
    for k, v in got.items():
        e = _rel(v.cpu().numpy(), og[k])
        print(f"{k}: relative L2 error {e:.2e}")
        assert e < 1e-3, k


def test_diff_backward_cost(gpu_ctx):
    """forward + backward per view against forward alone (<= 3x)."""
    import torch
    W = H = 128
    pos, attr, radii, cm, dm, rm = _scene(seed=5, n=20000)
    pos *= 1.5
    radii *= 0.3
    cam = default_camera(W, H, look_from=(0.2, 0.1, 1.4))
    R = DiffRenderer(gpu_ctx, cam, W, H, 0.004, spp=4, frame_id=1, rbf_radius=0.05)
    t = lambda a: torch.tensor(a, device="cuda", requires_grad=True)
    args = [t(a) for a in (pos, attr, radii, cm, dm, rm)]
    img = R(*args)
    img.sum().backward()  # warm-up
    torch.cuda.synchronize()
    t0 = time.time()
    for _ in range(3):
        img = R(*args)
    torch.cuda.synchronize()
    f = (time.time() - t0) / 3
    t0 = time.time()
    for _ in range(3):
        img = R(*args)
        img.sum().backward()
    torch.cuda.synchronize()
    fb = (time.time() - t0) / 3
    print(f"forward {f * 1e3:.2f} ms, forward + backward {fb * 1e3:.2f} ms per view (ratio {fb / f:.2f})")
    assert fb / f < 3.0
