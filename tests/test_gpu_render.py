"""GPU parity of ParticleRBFRayGen (delta tracking and stochastic marching)
and the TAA resolve against the oracle at fixed RNG streams, through the C
ABI.  The bar is bit-exact output; the stated tolerance (1e-4 relative) is
checked as well so that a failure reports its size."""
import numpy as np
import pytest

import aarbf
from aarbf import scenes
from helpers import default_camera, random_scene
import oracle

pytestmark = pytest.mark.gpu

RTOL = 1e-4


def _torch():
    import torch
    return torch


def _compare(got, exp, what):
    got = np.asarray(got, np.float32)
    exp = np.asarray(exp, np.float32)
    rel = np.abs(got - exp) / np.maximum(np.abs(exp), 1e-6)
    bad = np.argwhere(rel > RTOL)
    assert bad.size == 0, f"{what}: {len(bad)} values differ, first at {bad[:3].tolist()} " \
                          f"got {got[tuple(bad[0])]} expected {exp[tuple(bad[0])]}"
    assert np.array_equal(got, exp), f"{what}: within {RTOL} but not bit-exact"


def _setup(ctx, xyzw, radii, tfs, rbf, ppl=1, K=8):
    cm, rm, dm = tfs
    sc = aarbf.Scene.prepare(ctx, xyzw, cm, rm, dm, rbf, radii=radii, ppl=ppl, branching=K)
    op, _, operm, oaabb, _, orad = oracle.preprocess(xyzw, radii)
    assert np.array_equal(sc.particles.cpu().numpy(), op)
    assert np.array_equal(sc.aabb, oaabb)
    return sc, op, orad


def _gpu_render(ctx, sc, cam, W, H, unit, spp=1, frame=0, accum=1, stbn=0, **kw):
    torch = _torch()
    img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    acc = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    pc = sc.push_constants(cam, unit, frame_id=frame, accum_id=accum, **kw)
    rg = sc.raygen_data(img, acc, W, H, stbn=stbn)
    ctx.render(rg, pc, W, H, spp)
    ctx.sync()
    return img.cpu().numpy(), acc.cpu().numpy()


CASES = [
    # name, n, blue_noise, ppl, K, visualize_attributes, show_noise, spp
    ("tracking", 3000, False, 1, 8, True, False, 2),
    ("marching", 3000, True, 1, 8, True, False, 2),
    ("tracking-ppl4-K4", 3000, False, 4, 4, True, False, 1),
    ("marching-ppl3-K16", 2000, True, 3, 16, True, False, 1),
    ("tracking-density-colour", 3000, False, 1, 2, False, False, 1),
    ("marching-shownoise", 1500, True, 2, 8, True, True, 1),
]


@pytest.mark.parametrize("name,n,blue,ppl,K,vis,noise,spp", CASES)
def test_render_matches_oracle(gpu_ctx, name, n, blue, ppl, K, vis, noise, spp):
    W, H = 48, 40
    xyzw, radii = random_scene(n, seed=n + ppl, extent=1.0, r=(0.03, 0.09))
    tfs = (scenes.ramp_table(256), scenes.constant_table(n=4), scenes.ramp_table(64))
    if not vis:
        tfs = (scenes.reference_tf("colormap"), scenes.reference_tf("radiusmap"), scenes.reference_tf("densitymap"))
    sc, op, orad = _setup(gpu_ctx, xyzw, radii, tfs, 0.05, ppl, K)
    cam = default_camera(W, H, look_from=(0.4, 0.3, 2.2))
    unit = 0.004
    stbn_img = scenes.synthetic_stbn(seed=3) if blue else None
    stbn = gpu_ctx.stbn(stbn_img) if blue else 0
    light = (0.3, 0.2, 0.4)
    r = oracle.Render(op, orad, sc.aabb, *tfs, cam, W, H, unit, 0.05, frame_id=5, accum_id=1, blue_noise=blue,
                      stbn=stbn_img, show_noise=noise, visualize_attributes=vis, light=light)
    oimg, oacc, _ = r.run(spp=spp, threads=8)
    assert (oacc[..., :3] > 0).mean() > 0.05, "scene too empty to be a meaningful check"
    for mode in (0, 1):  # persistent segment-gather kernel, per-pixel point-query kernel
        gpu_ctx.set_render_mode(mode)
        try:
            img, acc = _gpu_render(gpu_ctx, sc, cam, W, H, unit, spp=spp, frame=5, accum=1, stbn=stbn,
                                   blue_noise=blue, show_noise=noise, visualize_attributes=vis, light=light)
        finally:
            gpu_ctx.set_render_mode(0)
        _compare(acc, oacc, f"{name} mode {mode} accum")
        _compare(img, oimg, f"{name} mode {mode} image")


def test_fused_spp_equals_consecutive_launches(gpu_ctx):
    W, H = 32, 24
    xyzw, radii = random_scene(2000, seed=3)
    tfs = (scenes.ramp_table(256), scenes.constant_table(n=4), scenes.ramp_table(64))
    sc, _, _ = _setup(gpu_ctx, xyzw, radii, tfs, 0.05)
    cam = default_camera(W, H, look_from=(0.0, 0.5, 2.0))
    torch = _torch()
    img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    acc = torch.zeros_like(img)
    for s in range(3):
        pc = sc.push_constants(cam, 0.004, frame_id=10 + s, accum_id=1 + s)
        gpu_ctx.render(sc.raygen_data(img, acc, W, H), pc, W, H, 1)
    gpu_ctx.sync()
    fimg, facc = _gpu_render(gpu_ctx, sc, cam, W, H, 0.004, spp=3, frame=10, accum=1)
    assert np.array_equal(acc.cpu().numpy(), facc) and np.array_equal(img.cpu().numpy(), fimg)


def test_uniform_cube_full_frame(gpu_ctx):
    """The CPU-reference workload (65,536 particles, 512x512, 1 spp
    marching) rendered in full on both sides."""
    wl = scenes.uniform_cube()
    sc, op, orad = _setup(gpu_ctx, wl.xyzw, wl.radii, (wl.colormap, wl.radiusmap, wl.densitymap), wl.rbf_radius)
    cam = aarbf.camera_basis(wl.look_from, wl.look_at, wl.look_up, wl.cos_fovy, wl.width, wl.height)
    stbn_img = scenes.synthetic_stbn()
    stbn = gpu_ctx.stbn(stbn_img)
    img, acc = _gpu_render(gpu_ctx, sc, cam, wl.width, wl.height, wl.unit, spp=1, stbn=stbn, blue_noise=True)
    r = oracle.Render(op, orad, sc.aabb, wl.colormap, wl.radiusmap, wl.densitymap, cam, wl.width, wl.height,
                      wl.unit, wl.rbf_radius, blue_noise=True, stbn=stbn_img)
    oimg, oacc, _ = r.run(spp=1, threads=16)
    _compare(img, oimg, "uniform cube image")


def test_dense_cluster_uses_point_query_fallback(gpu_ctx):
    """A cluster far denser than the per-segment candidate capacity: the
    persistent kernel falls back to point queries there and must still
    match the oracle bit for bit."""
    W, H = 40, 40
    rng = np.random.default_rng(12)
    n = 6000
    xyzw = np.empty((n, 4), np.float32)
    xyzw[:, :3] = rng.normal(0.0, 0.05, (n, 3))
    xyzw[:, 3] = rng.uniform(0, 1, n)
    radii = rng.uniform(0.05, 0.1, n).astype(np.float32)
    tfs = (scenes.ramp_table(256), scenes.constant_table(n=4), scenes.constant_table((0.02, 0, 0, 0), n=1))
    sc, op, orad = _setup(gpu_ctx, xyzw, radii, tfs, 0.1)
    cam = default_camera(W, H, look_from=(0.0, 0.2, 1.2))
    r = oracle.Render(op, orad, sc.aabb, *tfs, cam, W, H, 0.01, 0.1, frame_id=1, accum_id=1, light=(0.1, 0.6, 0.3))
    oimg, oacc, _ = r.run(spp=2, threads=8)
    img, acc = _gpu_render(gpu_ctx, sc, cam, W, H, 0.01, spp=2, frame=1, accum=1, light=(0.1, 0.6, 0.3))
    _compare(acc, oacc, "dense accum")


def test_stats_count_samples(gpu_ctx):
    W, H = 16, 16
    xyzw, radii = random_scene(1000, seed=1)
    tfs = (scenes.ramp_table(16), scenes.constant_table(n=1), scenes.ramp_table(16))
    sc, _, _ = _setup(gpu_ctx, xyzw, radii, tfs, 0.05)
    cam = default_camera(W, H)
    gpu_ctx.set_stats(True)
    gpu_ctx.read_stats(reset=True)
    _gpu_render(gpu_ctx, sc, cam, W, H, 0.004, spp=2)
    st = gpu_ctx.read_stats(reset=True)
    gpu_ctx.set_stats(False)
    assert st["samples"] == W * H * 2
    assert st["steps"] > 0 and st["nodes"] >= st["queries"] > 0 and st["covering"] <= st["particles"]


@pytest.mark.parametrize("gui_layer,disable", [(False, False), (True, False), (True, True)])
def test_taa_matches_oracle(gpu_ctx, gui_layer, disable):
    """CompositeGui: TAA resolve, history copy, the GUI layer composited with
    over() and the BGRA8 frame-buffer words, all compared bit for bit."""
    torch = _torch()
    rng = np.random.default_rng(0)
    H, W = 37, 53
    image = rng.uniform(0, 1, (H, W, 4)).astype(np.float32)
    image[rng.uniform(size=(H, W)) < 0.2, 3] = 1.0
    image[0, :3] = [[-0.25, 1.5, 0.5, 1.0], [0.2, 2.0, 0.7, 0.5], [0.1, 0.2, 0.3, 0.0]]
    prev = rng.uniform(-0.5, 1.5, (H, W, 4)).astype(np.float32)
    taa = rng.uniform(0, 1, (H, W, 4)).astype(np.float32)
    gui = rng.integers(0, 2 ** 32, (H, W), dtype=np.uint64).astype(np.uint32) if gui_layer else None
    if gui is not None:
        gui[:4] &= 0x00FFFFFF  # transparent texels
        gui[4:8] |= 0xFF000000  # opaque texels
    otaa, oprev = taa.copy(), prev.copy()
    oframe = np.zeros((H, W), np.uint32)
    oracle.taa_resolve(image, otaa, oprev, disable=disable, gui=gui, frame=oframe)
    d_img = torch.from_numpy(image).cuda()
    d_taa = torch.from_numpy(taa).cuda()
    d_prev = torch.from_numpy(prev).cuda()
    d_frame = torch.zeros((H, W), dtype=torch.int32, device="cuda")
    d_gui = torch.from_numpy(gui.view(np.int32)).cuda() if gui is not None else None
    tc = aarbf.TAAConstants(1 if disable else 0, 0, aarbf._lib.Int2(W, H), d_img.data_ptr(),
                            d_gui.data_ptr() if d_gui is not None else 0, d_frame.data_ptr(),
                            d_taa.data_ptr(), d_prev.data_ptr())
    gpu_ctx.taa_resolve(tc)
    gpu_ctx.sync()
    got_taa = d_taa.cpu().numpy()
    assert np.array_equal(got_taa, otaa, equal_nan=True), "taa"
    assert np.array_equal(d_prev.cpu().numpy(), oprev, equal_nan=True), "history copy"
    f = d_frame.cpu().numpy().view(np.uint32)
    bad = np.argwhere(f != oframe)
    assert bad.size == 0, f"frame words differ at {bad[:3].tolist()}: {f[tuple(bad[0])]:08x} vs " \
                          f"{oframe[tuple(bad[0])]:08x}"


@pytest.mark.parametrize("graph_rounds", [0, 4])
def test_sparse_waves_leaving_early_is_exact(gpu_ctx, graph_rounds):
    """Step-kernel waves that leave their loop once fewer than
    AARBF_TUNE_STEP_MIN_LIVE lanes still step (their rays go on in the next
    round) render bit for bit the same image as waves that run until their
    last lane leaves.  The view has 81,920 pixels, so the step queue passes
    the 65,536 rays from which the early exit applies; thresholds 0 (never),
    16 (the default) and 48 (waves leave when 17 lanes are gone); with
    host-driven rounds and with rounds captured in a hipGraph (fixed grids,
    queue lengths read on the device)."""
    W, H = 512, 160
    xyzw, radii = random_scene(20000, seed=91, extent=1.0, r=(0.02, 0.06))
    tfs = (scenes.ramp_table(256), scenes.constant_table(n=4), scenes.ramp_table(64))
    sc, op, orad = _setup(gpu_ctx, xyzw, radii, tfs, 0.06)
    cam = default_camera(W, H, look_from=(0.2, 0.3, 2.2))
    kw = dict(light=(0.3, 0.2, 0.4))
    out = {}
    try:
        gpu_ctx.set_tuning(gpu_ctx.TUNE_GRAPH_ROUNDS, graph_rounds)
        for live in (0, 16, 48):
            gpu_ctx.set_tuning(gpu_ctx.TUNE_STEP_MIN_LIVE, live)
            out[live] = _gpu_render(gpu_ctx, sc, cam, W, H, 0.004, spp=2, frame=5, **kw)
    finally:
        gpu_ctx.set_tuning(gpu_ctx.TUNE_STEP_MIN_LIVE, 16)
        gpu_ctx.set_tuning(gpu_ctx.TUNE_GRAPH_ROUNDS, 0)
    img0, acc0 = out[0]
    assert (acc0[..., :3] > 0).mean() > 0.1
    for live in (16, 48):
        _compare(out[live][1], acc0, f"accum, min-live {live} vs 0")
        _compare(out[live][0], img0, f"image, min-live {live} vs 0")
    px = np.array([(x, y) for y in range(2, H, 13) for x in range(3, W, 29)], np.int32)
    r = oracle.Render(op, orad, sc.aabb, *tfs, cam, W, H, 0.004, 0.06, frame_id=5, accum_id=1, **kw)
    oimg, oacc, _ = r.run(pixels=px, spp=2, threads=8)
    _compare(acc0[px[:, 1], px[:, 0]], oacc[px[:, 1], px[:, 0]], "accum vs oracle")


@pytest.mark.parametrize("cap", [16, 48])
def test_flight_blocks_cut_by_the_step_budget_are_exact(gpu_ctx, cap):
    """Delta tracking's empty-space steps run to the next flight anchor (32
    steps); a per-launch step budget of 16 or 48 cuts such blocks in the
    middle, so the rays leave with a partial product and its anchor saved
    and go on in the next round.  The image equals the default budget's bit
    for bit, and the oracle's on a pixel grid."""
    W, H = 160, 96
    xyzw, radii = random_scene(3000, seed=17, extent=1.0, r=(0.02, 0.07))
    tfs = (scenes.ramp_table(256), scenes.constant_table(n=4), scenes.ramp_table(64))
    sc, op, orad = _setup(gpu_ctx, xyzw, radii, tfs, 0.07)
    cam = default_camera(W, H, look_from=(0.1, 0.4, 2.4))
    kw = dict(light=(0.3, 0.2, 0.5))
    try:
        gpu_ctx.set_tuning(gpu_ctx.TUNE_STEP_CAP, cap)
        img, acc = _gpu_render(gpu_ctx, sc, cam, W, H, 0.004, spp=2, frame=3, **kw)
    finally:
        gpu_ctx.set_tuning(gpu_ctx.TUNE_STEP_CAP, 8192)
    img0, acc0 = _gpu_render(gpu_ctx, sc, cam, W, H, 0.004, spp=2, frame=3, **kw)
    assert (acc0[..., :3] > 0).mean() > 0.1
    _compare(acc, acc0, f"accum, step budget {cap} vs 8192")
    _compare(img, img0, f"image, step budget {cap} vs 8192")
    px = np.array([(x, y) for y in range(1, H, 7) for x in range(2, W, 9)], np.int32)
    r = oracle.Render(op, orad, sc.aabb, *tfs, cam, W, H, 0.004, 0.07, frame_id=3, accum_id=1, **kw)
    oimg, oacc, _ = r.run(pixels=px, spp=2, threads=8)
    _compare(acc[px[:, 1], px[:, 0]], oacc[px[:, 1], px[:, 0]], "accum vs oracle")


@pytest.mark.parametrize("blue_noise", [False, True])
def test_walk_budget_handoff_is_exact(gpu_ctx, blue_noise):
    """Gather walks that open more nodes than the budget hand their ray to
    the one-wave-per-ray gather in the same round.  An image whose first
    rounds' gather queues exceed the wave gather's own limit (16,384 rays)
    renders bit for bit the same with every walk handed over (budget 1) as
    with the default budget, and equals the oracle on a pixel sample."""
    W, H = 176, 128
    xyzw, radii = random_scene(12000, seed=77, extent=1.0, r=(0.02, 0.07))
    tfs = (scenes.ramp_table(256), scenes.constant_table(n=4), scenes.ramp_table(64))
    sc, op, orad = _setup(gpu_ctx, xyzw, radii, tfs, 0.07)
    cam = default_camera(W, H, look_from=(0.3, 0.25, 2.3))
    stbn = gpu_ctx.stbn(scenes.synthetic_stbn(seed=5)) if blue_noise else 0
    kw = dict(blue_noise=blue_noise, light=(0.3, 0.2, 0.4))
    img0, acc0 = _gpu_render(gpu_ctx, sc, cam, W, H, 0.004, spp=2, frame=3, stbn=stbn, **kw)
    try:
        gpu_ctx.set_tuning(gpu_ctx.TUNE_WALK_BUDGET, 1)
        img1, acc1 = _gpu_render(gpu_ctx, sc, cam, W, H, 0.004, spp=2, frame=3, stbn=stbn, **kw)
    finally:
        gpu_ctx.set_tuning(gpu_ctx.TUNE_WALK_BUDGET, 256)
    assert (acc0[..., :3] > 0).mean() > 0.2
    _compare(acc1, acc0, "accum, budget 1 vs default")
    _compare(img1, img0, "image, budget 1 vs default")
    px = np.array([(x, y) for y in range(3, H, 17) for x in range(5, W, 19)], np.int32)
    r = oracle.Render(op, orad, sc.aabb, *tfs, cam, W, H, 0.004, 0.07, frame_id=3, accum_id=1,
                      blue_noise=blue_noise, stbn=scenes.synthetic_stbn(seed=5) if blue_noise else None,
                      light=(0.3, 0.2, 0.4))
    oimg, oacc, _ = r.run(pixels=px, spp=2, threads=8)
    _compare(acc0[px[:, 1], px[:, 0]], oacc[px[:, 1], px[:, 0]], "accum vs oracle")
