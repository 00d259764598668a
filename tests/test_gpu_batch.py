"""Batched rendering (aarbf_render_batch): views of uneven size, spp,
camera, counters, light and tracker (delta tracking / stochastic marching)
rendered in one wavefront schedule give, view by view, the images of
separate aarbf_render_rbf_spp calls bit for bit -- and so the oracle's,
which those calls match (test_gpu_render.py).  One small batch is also
checked against the oracle directly.  The latency record gives one time per
view within the batch's wall time."""
import time

import numpy as np
import pytest

import aarbf
from aarbf import scenes
from helpers import default_camera, random_scene
import oracle

pytestmark = pytest.mark.gpu

# W, H, spp, marching, frame, accum, look_from, light, show_noise
VIEWS = [
    (37, 23, 3, False, 5, 1, (0.4, 0.3, 2.2), (0.3, 0.2, 0.4), False),
    (64, 64, 1, True, 9, 1, (-0.5, 0.2, 2.0), (0.1, 0.6, 0.5), False),
    (8, 8, 5, False, 0, 3, (0.0, 1.8, 1.0), (0.7, 0.3, 0.2), False),
    (50, 17, 2, True, 77, 2, (1.5, -0.4, 1.5), (0.2, 0.4, 1.0), True),
    (1, 1, 4, False, 3, 1, (0.2, 0.1, 2.5), (0.3, 0.2, 0.4), False),
    (70, 9, 2, False, 65535, 1, (2.0, 0.5, -0.8), (0.5, 0.5, 0.3), False),
]


def _scene(ctx, n=3000, seed=11):
    xyzw, radii = random_scene(n, seed=seed, extent=1.0, r=(0.03, 0.09))
    tfs = (scenes.ramp_table(256), scenes.constant_table(n=4), scenes.ramp_table(64))
    return aarbf.Scene.prepare(ctx, xyzw, *tfs, 0.05, radii=radii), tfs


def _views(ctx, sc, views, stbn, seed_acc=0):
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed_acc)
    out = []
    for (W, H, spp, march, frame, accum, lf, light, noise) in views:
        cam = default_camera(W, H, look_from=lf)
        pc = sc.push_constants(cam, 0.004, frame_id=frame, accum_id=accum, blue_noise=march, show_noise=noise,
                               light=light)
        img = torch.zeros((H, W, 4), dtype=torch.float32)
        acc = torch.rand((H, W, 4), generator=g) if accum > 1 else torch.zeros((H, W, 4))
        img, acc = img.cuda(), acc.cuda()
        out.append([pc, img, acc, W, H, spp, march])
    return out


def _run_batch(ctx, sc, vs, stbn):
    reqs = [(sc.raygen_data(img, acc, W, H, stbn=stbn if m else 0), pc, W, H, spp)
            for (pc, img, acc, W, H, spp, m) in vs]
    ctx.render_batch(reqs)
    ctx.sync()


def test_batch_equals_separate_calls(gpu_ctx):
    sc, _ = _scene(gpu_ctx)
    stbn = gpu_ctx.stbn(scenes.synthetic_stbn(seed=3))
    a = _views(gpu_ctx, sc, VIEWS, stbn)
    b = _views(gpu_ctx, sc, VIEWS, stbn)
    for (pc, img, acc, W, H, spp, m) in a:
        gpu_ctx.render(sc.raygen_data(img, acc, W, H, stbn=stbn if m else 0), pc, W, H, spp)
    gpu_ctx.sync()
    t0 = time.perf_counter()
    _run_batch(gpu_ctx, sc, b, stbn)
    wall_ms = (time.perf_counter() - t0) * 1e3
    lit = 0
    for k, (va, vb) in enumerate(zip(a, b)):
        ia, aa = va[1].cpu().numpy(), va[2].cpu().numpy()
        ib, ab = vb[1].cpu().numpy(), vb[2].cpu().numpy()
        assert np.array_equal(aa, ab), f"view {k}: accum differs"
        assert np.array_equal(ia, ib), f"view {k}: image differs"
        lit += (ab[..., :3] > 0).mean() > 0.05
    assert lit >= 4
    lat = gpu_ctx.batch_latency(len(VIEWS))
    print("latency ms per view:", np.round(lat, 3).tolist(), f"batch wall {wall_ms:.2f} ms")
    assert np.all(lat >= 0) and np.all(lat <= wall_ms + 1.0)


def test_batch_repeat_and_stats(gpu_ctx):
    """Two batches in a row (the second smaller) and a batch under the
    statistics kernels give the same images as separate calls."""
    sc, _ = _scene(gpu_ctx, n=2000, seed=12)
    stbn = gpu_ctx.stbn(scenes.synthetic_stbn(seed=4))
    for views, stats in ((VIEWS, False), (VIEWS[:3], False), (VIEWS[1:4], True)):
        a = _views(gpu_ctx, sc, views, stbn)
        b = _views(gpu_ctx, sc, views, stbn)
        for (pc, img, acc, W, H, spp, m) in a:
            gpu_ctx.render(sc.raygen_data(img, acc, W, H, stbn=stbn if m else 0), pc, W, H, spp)
        gpu_ctx.set_stats(stats)
        try:
            _run_batch(gpu_ctx, sc, b, stbn)
        finally:
            gpu_ctx.set_stats(False)
        for va, vb in zip(a, b):
            assert np.array_equal(va[2].cpu().numpy(), vb[2].cpu().numpy())
            assert np.array_equal(va[1].cpu().numpy(), vb[1].cpu().numpy())


def test_batch_matches_oracle(gpu_ctx):
    sc, tfs = _scene(gpu_ctx, n=1500, seed=13)
    stbn_img = scenes.synthetic_stbn(seed=5)
    stbn = gpu_ctx.stbn(stbn_img)
    views = [VIEWS[0], VIEWS[1], VIEWS[3]]
    vs = _views(gpu_ctx, sc, views, stbn, seed_acc=7)
    acc0 = [v[2].cpu().numpy().copy() for v in vs]
    _run_batch(gpu_ctx, sc, vs, stbn)
    op, orad = sc.particles.cpu().numpy(), sc.radii.cpu().numpy()
    for k, ((W, H, spp, march, frame, accum, lf, light, noise), v) in enumerate(zip(views, vs)):
        cam = default_camera(W, H, look_from=lf)
        r = oracle.Render(op, orad, sc.aabb, *tfs, cam, W, H, 0.004, 0.05, frame_id=frame, accum_id=accum,
                          blue_noise=march, stbn=stbn_img if march else None, show_noise=noise, light=light)
        oimg, oacc, _ = r.run(spp=spp, threads=8, accum=acc0[k])
        assert np.array_equal(v[2].cpu().numpy(), oacc), f"view {k}: accum differs from the oracle"
        assert np.array_equal(v[1].cpu().numpy(), oimg), f"view {k}: image differs from the oracle"


def test_batch_rejects_mixed_scenes(gpu_ctx):
    import torch
    sc1, _ = _scene(gpu_ctx, n=500, seed=14)
    sc2, _ = _scene(gpu_ctx, n=600, seed=15)
    W = H = 8
    cam = default_camera(W, H)
    img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    acc = torch.zeros_like(img)
    reqs = [(sc1.raygen_data(img, acc, W, H), sc1.push_constants(cam, 0.004), W, H, 1),
            (sc2.raygen_data(img, acc, W, H), sc2.push_constants(cam, 0.004), W, H, 1)]
    with pytest.raises(aarbf.AarbfError, match="share the scene"):
        gpu_ctx.render_batch(reqs)


@pytest.mark.parametrize("rounds", [2, 8])
def test_graph_rounds_equal_host_rounds(gpu_ctx, rounds):
    """Rounds captured into a hipGraph and replayed (the graph is cached and
    replayed again on the next call) give the host-driven images; eager
    renders run between the graph calls."""
    sc, _ = _scene(gpu_ctx, n=2500, seed=16)
    stbn = gpu_ctx.stbn(scenes.synthetic_stbn(seed=6))
    ref = _views(gpu_ctx, sc, VIEWS, stbn)
    _run_batch(gpu_ctx, sc, ref, stbn)
    single = _views(gpu_ctx, sc, VIEWS[:2], stbn)
    for (pc, img, acc, W, H, spp, m) in single:
        gpu_ctx.render(sc.raygen_data(img, acc, W, H, stbn=stbn if m else 0), pc, W, H, spp)
    gpu_ctx.sync()
    try:
        for rep in range(3):
            gpu_ctx.set_tuning(gpu_ctx.TUNE_GRAPH_ROUNDS, rounds)
            got = _views(gpu_ctx, sc, VIEWS, stbn)
            _run_batch(gpu_ctx, sc, got, stbn)
            for k, (va, vb) in enumerate(zip(ref, got)):
                assert np.array_equal(va[2].cpu().numpy(), vb[2].cpu().numpy()), f"replay {rep} view {k}"
                assert np.array_equal(va[1].cpu().numpy(), vb[1].cpu().numpy()), f"replay {rep} view {k}"
            one = _views(gpu_ctx, sc, VIEWS[:2], stbn)
            for (pc, img, acc, W, H, spp, m) in one:
                gpu_ctx.render(sc.raygen_data(img, acc, W, H, stbn=stbn if m else 0), pc, W, H, spp)
            gpu_ctx.sync()
            for va, vb in zip(single, one):
                assert np.array_equal(va[2].cpu().numpy(), vb[2].cpu().numpy())
            # eager (host-driven) work between the graph calls
            gpu_ctx.set_tuning(gpu_ctx.TUNE_GRAPH_ROUNDS, 0)
            eager = _views(gpu_ctx, sc, VIEWS[2:4], stbn)
            _run_batch(gpu_ctx, sc, eager, stbn)
    finally:
        gpu_ctx.set_tuning(gpu_ctx.TUNE_GRAPH_ROUNDS, 0)


def test_autotune_keeps_images_and_caller_buffers(gpu_ctx):
    """The schedule autotuner leaves the callers' buffers alone, keeps a
    setting from its candidate sets, and images rendered under the kept
    setting equal those under the default one."""
    from aarbf import tune
    sc, tfs = _scene(gpu_ctx, n=2000, seed=18)
    stbn = gpu_ctx.stbn(scenes.synthetic_stbn(seed=8))
    ref = _views(gpu_ctx, sc, VIEWS, stbn)
    _run_batch(gpu_ctx, sc, ref, stbn)
    probe = _views(gpu_ctx, sc, VIEWS, stbn)
    reqs = [(sc.raygen_data(img, acc, W, H, stbn=stbn if m else 0), pc, W, H, spp)
            for (pc, img, acc, W, H, spp, m) in probe]
    before = [(v[1].cpu().numpy().copy(), v[2].cpu().numpy().copy()) for v in probe]
    try:
        r = gpu_ctx.autotune(reqs, reps=1)
        print("autotune:", r)
        assert r["step_cap"] in (2048, 4096, 8192, 16384) and r["step_waves"] in (0, 4, 5, 6)
        assert r["graph_rounds"] in (0, 8) and r["walk_budget"] in (256, 1024, 65536)
        assert r["step_min_live"] in (0, 16)
        assert r["candidates"] >= 10 and 0 < r["best_ms"] <= r["baseline_ms"]
        for v, (i0, a0) in zip(probe, before):  # untouched
            assert np.array_equal(v[1].cpu().numpy(), i0) and np.array_equal(v[2].cpu().numpy(), a0)
        got = _views(gpu_ctx, sc, VIEWS, stbn)
        _run_batch(gpu_ctx, sc, got, stbn)
        for va, vb in zip(ref, got):
            assert np.array_equal(va[2].cpu().numpy(), vb[2].cpu().numpy())
            assert np.array_equal(va[1].cpu().numpy(), vb[1].cpu().numpy())
        # leaf size around it: rebuilds the scene per candidate
        xyzw, radii = random_scene(2000, seed=18, extent=1.0, r=(0.03, 0.09))

        def make(s):
            return [(s.raygen_data(img, acc, W, H, stbn=stbn if m else 0),
                     s.push_constants(default_camera(W, H, look_from=VIEWS[k][6]), 0.004, frame_id=VIEWS[k][4],
                                      blue_noise=m), W, H, spp)
                    for k, (pc, img, acc, W, H, spp, m) in enumerate(probe)]
        best, info = tune.tune_leaf_size(gpu_ctx, xyzw, *tfs, 0.05, make, radii=radii, leaf_sizes=(1, 2))
        print("leaf size:", info)
        assert info["particles_per_leaf"] in (1, 2) and best.ppl == info["particles_per_leaf"]
    finally:
        for k in (gpu_ctx.TUNE_STEP_CAP, gpu_ctx.TUNE_STEP_WAVES, gpu_ctx.TUNE_GRAPH_ROUNDS):
            gpu_ctx.set_tuning(k, 8192 if k == gpu_ctx.TUNE_STEP_CAP else 0)
        gpu_ctx.set_tuning(gpu_ctx.TUNE_WALK_BUDGET, 256)
        gpu_ctx.set_tuning(gpu_ctx.TUNE_STEP_MIN_LIVE, 16)


@pytest.mark.parametrize("rounds", [0, 4])
def test_batch_views_spanning_several_passes(gpu_ctx, rounds):
    """Views whose spp ends in different 64-sample passes (70, 3 and 130
    spp): every view equals its own call, host-driven and captured rounds."""
    sc, _ = _scene(gpu_ctx, n=1200, seed=19)
    stbn = gpu_ctx.stbn(scenes.synthetic_stbn(seed=9))
    views = [(12, 9, 70, False, 5, 1, (0.4, 0.3, 2.2), (0.3, 0.2, 0.4), False),
             (16, 16, 3, True, 9, 2, (-0.5, 0.2, 2.0), (0.1, 0.6, 0.5), False),
             (7, 13, 130, False, 65500, 1, (0.0, 1.8, 1.0), (0.7, 0.3, 0.2), False)]
    a = _views(gpu_ctx, sc, views, stbn, seed_acc=3)
    b = _views(gpu_ctx, sc, views, stbn, seed_acc=3)
    for (pc, img, acc, W, H, spp, m) in a:
        gpu_ctx.render(sc.raygen_data(img, acc, W, H, stbn=stbn if m else 0), pc, W, H, spp)
    gpu_ctx.sync()
    try:
        gpu_ctx.set_tuning(gpu_ctx.TUNE_GRAPH_ROUNDS, rounds)
        _run_batch(gpu_ctx, sc, b, stbn)
    finally:
        gpu_ctx.set_tuning(gpu_ctx.TUNE_GRAPH_ROUNDS, 0)
    for k, (va, vb) in enumerate(zip(a, b)):
        assert np.array_equal(va[2].cpu().numpy(), vb[2].cpu().numpy()), f"view {k}: accum"
        assert np.array_equal(va[1].cpu().numpy(), vb[1].cpu().numpy()), f"view {k}: image"
    lat = gpu_ctx.batch_latency(len(views))
    assert np.all(lat >= 0)
