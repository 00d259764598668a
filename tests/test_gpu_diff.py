"""The HIP differentiable renderer (aarbf_diff_forward / _backward) against
the FP64 oracle (oracle/adjoint.c, itself pinned by central finite
differences), and the torch.autograd.Function wrapper.  Tolerances: the
forward within 1e-4 absolute per pixel; each gradient block within 1e-3
relative L2 of the oracle's (float32 kernels against float64).  Expected
marching runs in both schedules (one wave per sample, the default, and
one lane per sample); ratio tracking in the lane schedule."""
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


@pytest.fixture
def schedule(gpu_ctx, request):
    """AARBF_TUNE_DIFF_WAVES for the test, back to the default after it."""
    gpu_ctx.set_tuning(gpu_ctx.TUNE_DIFF_WAVES, request.param)
    yield request.param
    gpu_ctx.set_tuning(gpu_ctx.TUNE_DIFF_WAVES, 1)


@pytest.mark.parametrize("estimator,schedule", [(0, 1), (0, 0), (1, 0)], indirect=["schedule"])
def test_diff_forward_and_adjoint_match_oracle(gpu_ctx, estimator, schedule):
    """estimator 0: expected marching (path replay, total - prefix), in the
    wave and the lane schedule; 1: ratio tracking (path replay with the
    remaining-colour recursion, against the oracle's stored-points reverse
    recursion)."""
    import torch
    W = H = 24
    spp, unit = 2, 0.01
    pos, attr, radii, cm, dm, rm = _scene()
    cam = default_camera(W, H, look_from=(0.2, 0.1, 1.3))
    R = DiffRenderer(gpu_ctx, cam, W, H, unit, spp=spp, frame_id=3, rbf_radius=0.2, estimator=estimator)
    t = lambda a: torch.tensor(a, device="cuda", requires_grad=True)
    P, A, Rd, CM, DM, RM = (t(a) for a in (pos, attr, radii, cm, dm, rm))
    img = R(P, A, Rd, CM, DM, RM)
    rng = np.random.default_rng(11)
    gimg = rng.normal(size=(H, W, 3)).astype(np.float32)
    (img * torch.tensor(gimg, device="cuda")).sum().backward()
    torch.cuda.synchronize()
    aabb = np.concatenate([pos.min(0), pos.max(0)])
    D = oracle.Diff(pos.astype(np.float64), attr.astype(np.float64), radii.astype(np.float64), cm, dm, rm, cam, W, H,
                    unit, 0.2, aabb, spp=spp, frame_id=3, estimator=estimator)
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


@pytest.mark.parametrize("estimator,schedule", [(0, 1), (0, 0), (1, 0)], indirect=["schedule"])
def test_diff_backward_cost(gpu_ctx, estimator, schedule):
    """forward + backward per view against forward alone (<= 3x): the
    library's forward and adjoint calls, each timed to the end of its work
    (the per-view scene rebuild, common to both, is not part of either),
    best of three views, and the torch.autograd.Function round trip once."""
    import torch
    from aarbf.diff import _Render
    W = H = 256
    pos, attr, radii, cm, dm, rm = _scene(seed=5, n=100000)
    pos *= 1.5
    radii *= 0.15
    cam = default_camera(W, H, look_from=(0.2, 0.1, 1.4))
    T = lambda a: torch.tensor(a, device="cuda")
    r = _Render(gpu_ctx, cam, W, H, 0.002, 4, 1, 0.05, 1, 8, T(pos), T(attr), T(radii), T(cm), T(dm), T(rm),
                estimator=estimator)
    gpu_ctx.set_diff_estimator(estimator)
    n = len(pos)
    g = torch.full((H, W, 4), 0.01, dtype=torch.float32, device="cuda")
    grads = [torch.zeros(k, dtype=torch.float32, device="cuda") for k in (3 * n, n, n, cm.size, len(dm), len(rm))]
    best = None
    for rep in range(3):
        torch.cuda.synchronize()
        gpu_ctx.sync()
        t0 = time.perf_counter()
        gpu_ctx.diff_forward(r.rg, r.pc, W, H, 4, r.img, r.path)
        gpu_ctx.sync()
        t1 = time.perf_counter()
        gpu_ctx.diff_backward(r.rg, r.pc, W, H, 4, g, *grads, path=r.path)
        gpu_ctx.sync()
        t2 = time.perf_counter()
        f, b = (t1 - t0) * 1e3, (t2 - t1) * 1e3
        if rep and (best is None or f + b < best[0] + best[1]):
            best = (f, b)
    r.close()
    f, b = best
    print(f"forward {f:.2f} ms, forward + backward {f + b:.2f} ms per view (ratio {(f + b) / f:.2f})")
    assert (f + b) / f < 3.0
    # the autograd wrapper end to end
    R = DiffRenderer(gpu_ctx, cam, W, H, 0.002, spp=4, frame_id=1, rbf_radius=0.05, estimator=estimator)
    args = [torch.tensor(a, device="cuda", requires_grad=True) for a in (pos, attr, radii, cm, dm, rm)]
    img = R(*args)
    img.sum().backward()
    assert all(a.grad is not None and torch.isfinite(a.grad).all() for a in args)


def test_diff_renderer_releases_its_handles(gpu_ctx):
    """Each forward builds a BVH and three tables on the context: they are
    released right away when no backward can follow (no_grad, no input
    needing a gradient), after backward otherwise, and when an output is
    dropped without a backward."""
    import gc
    import torch
    W = H = 12
    pos, attr, radii, cm, dm, rm = _scene(n=60)
    cam = default_camera(W, H, look_from=(0.2, 0.1, 1.3))
    R = DiffRenderer(gpu_ctx, cam, W, H, 0.01, spp=1, rbf_radius=0.2)
    t = lambda a, g: torch.tensor(a, device="cuda", requires_grad=g)
    base = gpu_ctx.handle_counts()
    with torch.no_grad():
        R(*(t(a, True) for a in (pos, attr, radii, cm, dm, rm)))
    assert gpu_ctx.handle_counts() == base, "forward under no_grad"
    R(*(t(a, False) for a in (pos, attr, radii, cm, dm, rm)))
    assert gpu_ctx.handle_counts() == base, "forward without inputs that need gradients"
    P = t(pos, True)
    img = R(P, *(t(a, False) for a in (attr, radii, cm, dm, rm)))
    assert gpu_ctx.handle_counts()["bvh"] == base["bvh"] + 1
    img.sum().backward(retain_graph=True)
    g1 = P.grad.clone()
    img.sum().backward()  # a second backward through the retained graph
    assert torch.allclose(P.grad, 2 * g1)
    del img
    gc.collect()
    assert gpu_ctx.handle_counts() == base, "after backward, graph freed"
    img = R(P, *(t(a, False) for a in (attr, radii, cm, dm, rm)))
    del img
    gc.collect()
    assert gpu_ctx.handle_counts() == base, "output dropped without backward"


@pytest.mark.parametrize("spp", [1, 4])
def test_diff_schedules_agree(gpu_ctx, spp):
    """Expected marching in the wave schedule against the lane schedule on a
    crowded scene (points covered by more particles than a stretch holds
    take the serial query): the same forward to float rounding, the same
    path vectors and gradients to 1e-4 relative L2."""
    import torch
    from aarbf.diff import _Render
    W, H = 64, 48
    pos, attr, radii, cm, dm, rm = _scene(seed=9, n=20000)
    radii *= 0.6
    cam = default_camera(W, H, look_from=(0.25, 0.15, 1.2))
    T = lambda a: torch.tensor(a, device="cuda")
    r = _Render(gpu_ctx, cam, W, H, 0.004, spp, 2, 0.2, 1, 8, T(pos), T(attr), T(radii), T(cm), T(dm), T(rm))
    g = torch.tensor(np.random.default_rng(4).normal(size=(H, W, 3)).astype(np.float32), device="cuda")
    out = {}
    try:
        for waves in (0, 1):
            gpu_ctx.set_tuning(gpu_ctx.TUNE_DIFF_WAVES, waves)
            img = r.forward()
            path = r.path.clone()
            grads = [x.clone() for x in r.backward(g)]
            out[waves] = (img, path, grads)
    finally:
        gpu_ctx.set_tuning(gpu_ctx.TUNE_DIFF_WAVES, 1)
        r.close()
    (i0, p0, g0), (i1, p1, g1) = out[0], out[1]
    assert float(i0.mean()) > 0.05
    assert float((i0 - i1).abs().max()) < 1e-5
    assert float(torch.linalg.norm((p0 - p1).flatten()) / torch.linalg.norm(p0.flatten())) < 1e-4
    for k, (a, b) in enumerate(zip(g0, g1)):
        e = float(torch.linalg.norm((a - b).flatten()) / max(float(torch.linalg.norm(a.flatten())), 1e-30))
        print(f"gradient block {k}: relative L2 {e:.2e}")
        assert e < 1e-4, k
