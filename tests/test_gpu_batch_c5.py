"""The batching workload at its own size: 16,777,216 particles from the
64-component mixture, 256 camera requests of uneven resolution (side
log-uniform in [128, 1024], aspect in [0.5, 2], 1-8 spp, half stochastic
marching, half delta tracking) rendered by one aarbf_render_batch call with
the FP16 leaf-local payload.

  - pre-processing equals the oracle's at 16M particles;
  - every view is finite and lit somewhere;
  - every view's latency is positive and within the batch's own GPU time;
  - on 8 views (4 of each tracker, the largest and smallest among them),
    64 pixels each equal the oracle's ParticleRBFRayGen
    (deviceCodeRBF.hlsl:206-336) on the decoded payload bit for bit."""
import numpy as np
import pytest

import aarbf
from aarbf import scenes
import oracle

pytestmark = pytest.mark.gpu


def test_batching_workload_256_requests(gpu_ctx):
    import torch
    xyzw, radii = scenes.mixture_16m()
    views = scenes.batch_views(n=256)
    tfs = (scenes.reference_tf("colormap"), scenes.reference_tf("radiusmap"), scenes.reference_tf("densitymap"))
    rbf, unit = float(radii.max()), 0.0004
    sc = aarbf.Scene.prepare(gpu_ctx, xyzw, *tfs, rbf, radii=radii, ppl=1, branching=8)
    op, _, _, oaabb, _, orad = oracle.preprocess(xyzw, radii)
    assert np.array_equal(sc.aabb, oaabb) and np.array_equal(sc.particles.cpu().numpy(), op)
    assert np.array_equal(sc.radii.cpu().numpy(), orad)
    del xyzw, radii
    atlas = scenes.synthetic_stbn(seed=3)
    stbn = gpu_ctx.stbn(atlas)
    bufs, reqs, cams = [], [], []
    for v in views:
        W, H = v["width"], v["height"]
        img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        acc = torch.zeros_like(img)
        cam = aarbf.camera_basis(v["look_from"], v["look_at"], v["up"], scenes.fovy_scale(45.0), W, H)
        pc = sc.push_constants(cam, unit, frame_id=v["frame"], accum_id=1, blue_noise=v["marching"])
        bufs.append((img, acc))
        cams.append(cam)
        reqs.append((sc.raygen_data(img, acc, W, H, stbn=stbn if v["marching"] else 0), pc, W, H, v["spp"]))
    samples = sum(W * H * spp for (_, _, W, H, spp) in reqs)
    gpu_ctx.set_payload(gpu_ctx.PAYLOAD_FP16)
    try:
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        gpu_ctx.render_batch(reqs)
        e1.record()
        torch.cuda.synchronize()
        batch_ms = e0.elapsed_time(e1)
        lat = gpu_ctx.batch_latency(len(reqs))
        decoded = gpu_ctx.payload_read(sc.n)
    finally:
        gpu_ctx.set_payload(gpu_ctx.PAYLOAD_FP32)
    print(f"\n256 requests, {samples} samples in {batch_ms:.0f} ms ({samples / batch_ms / 1e3:.1f} Msamples/s); "
          f"latency p50 {np.percentile(lat, 50):.0f} ms, p99 {np.percentile(lat, 99):.0f} ms")
    assert samples > 2e8
    for k, (img, acc) in enumerate(bufs):
        assert torch.isfinite(acc).all() and torch.isfinite(img).all(), f"view {k}"
        assert (acc[..., :3] > 0).any(), f"view {k} is black"
    assert np.all(lat > 0) and np.all(lat <= batch_ms * 1.01), (lat.min(), lat.max(), batch_ms)

    order = sorted(range(len(views)), key=lambda k: views[k]["width"] * views[k]["height"])
    march = [k for k in order if views[k]["marching"]]
    track = [k for k in order if not views[k]["marching"]]
    pick = [march[0], march[len(march) // 3], march[2 * len(march) // 3], march[-1],
            track[0], track[len(track) // 3], track[2 * len(track) // 3], track[-1]]
    rng = np.random.default_rng(9)
    for k in pick:
        v = views[k]
        W, H = v["width"], v["height"]
        acc = bufs[k][1].cpu().numpy()
        lit = np.argwhere(acc[..., :3].max(2) > 0)
        # 48 lit pixels and 16 anywhere
        sel = np.concatenate([lit[rng.choice(len(lit), size=min(48, len(lit)), replace=False)],
                              np.stack([rng.integers(0, H, 16), rng.integers(0, W, 16)], 1)])
        px = np.stack([sel[:, 1], sel[:, 0]], 1).astype(np.int32)
        r = oracle.Render(op, orad, oaabb, *tfs, cams[k], W, H, unit, rbf, frame_id=v["frame"], accum_id=1,
                          blue_noise=v["marching"], stbn=atlas if v["marching"] else None, baked=decoded)
        oimg, oacc, _ = r.run(pixels=px, spp=v["spp"], threads=16)
        got, want = acc[px[:, 1], px[:, 0]], oacc[px[:, 1], px[:, 0]]
        bad = np.argwhere(np.any(got != want, axis=1))
        assert bad.size == 0, f"view {k} ({W}x{H}, {v['spp']} spp, marching {v['marching']}): pixel " \
                              f"{px[bad[0][0]].tolist()} {got[bad[0][0]]} vs {want[bad[0][0]]}"
        assert np.array_equal(bufs[k][0].cpu().numpy()[px[:, 1], px[:, 0]], oimg[px[:, 1], px[:, 0]])
