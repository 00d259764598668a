"""Benchmark: Msamples/s of RBF delta tracking at 8,388,608 particles,
1920x1080, 16 spp per frame, TAA over an orbit of frames, one MI355X.

A step is one frame of the workload: one fused 16-spp launch of
ParticleRBFRayGen (deviceCodeRBF.hlsl:206-336) with the camera at the
frame's orbit position, followed by the TAA resolve
(deviceCodeCommon.hlsl:67-115).  Particles, leaf bounds and the BVH are
resident in HBM before the timed region; pre-processing is timed and
reported separately.

Prints ONE JSON line (rank 0).  Usage:
    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "attribute-aware-rbfs_amd")]

HBM_PEAK_GBS = 8000.0
VALU_PEAK_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16, help="timed frames")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=16)
    ap.add_argument("--orbit-frames", type=int, default=32)
    ap.add_argument("--particles", type=int, default=8388608)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mode", type=int, default=0, help="0 wavefront gather/step kernels, 1 point queries")
    ap.add_argument("--payload", type=int, default=0, help="0 FP32 particle payload, 1 FP16 leaf-local quantised")
    ap.add_argument("--step-waves", type=int, default=0)
    ap.add_argument("--step-cap", type=int, default=8192)
    ap.add_argument("--graph-rounds", type=int, default=0, help="rounds per captured hipGraph (0 = host-driven)")
    ap.add_argument("--step-min-live", type=int, default=16,
                    help="step-kernel waves leave their loop below this many live lanes (0 = never)")
    ap.add_argument("--walk-budget", type=int, default=256,
                    help="nodes a one-lane gather walk opens before its ray goes to a one-wave gather")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline's all-core run (0 = the cores this process may use, at most 16)")
    ap.add_argument("--prep-67m", type=int, default=1,
                    help="time pre-processing and the first frame of the 67,108,864-particle precision scene (0 = skip)")
    ap.add_argument("--diff-views", type=int, default=2,
                    help="views of the differentiable-rendering leg timed (0 = skip the leg)")
    ap.add_argument("--splat-frames", type=int, default=2, help="full-size splat frames timed (0 = no splat leg)")
    ap.add_argument("--diff-particles", type=int, default=1048576)
    ap.add_argument("--batch-views", type=int, default=256,
                    help="camera requests of the batching leg (0 = skip the leg)")
    ap.add_argument("--batch-particles", type=int, default=16777216)
    ap.add_argument("--batch-payload", type=int, default=1)
    ap.add_argument("--frame-batch", type=int, default=32,
                    help="the frame-batch leg renders the timed frames (at most this many) in one "
                         "aarbf_render_batch call, FP16 and FP32 payload (0 = skip)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "render_traffic.json"),
                    help="HBM bytes per render launch measured with rocprofv3 --pmc")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one plain device-to-device copy of BYTES (the reference figure that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes)")
    a = ap.parse_args()
    if a.copy_calibration:
        return copy_calibration(a.copy_calibration)

    if a.gpus != 1:
        print(f"bench.py: this renderer is measured on one GPU only (asked for --gpus {a.gpus}); the render "
              "path would shard by image tiles / requests with a full particle replica per GPU and no exchange",
              file=sys.stderr)
        return 2

    import torch
    import aarbf
    from aarbf import scenes

    if not torch.cuda.is_available():
        log("bench.py: no HIP device")
        return 1

    t0 = time.time()
    wl = scenes.mixture_8m(n=a.particles, width=a.width, height=a.height, spp=a.spp, frames=a.orbit_frames)
    log(f"workload generated in {time.time() - t0:.1f} s")
    W, H, S = wl.width, wl.height, wl.spp

    ctx = aarbf.Context(0)
    # one stream for the library's kernels and for the timing events
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    ctx.set_stream(stream)
    ctx.set_render_mode(a.mode)
    ctx.set_payload(a.payload)
    ctx.set_tuning(ctx.TUNE_STEP_WAVES, a.step_waves)
    ctx.set_tuning(ctx.TUNE_STEP_CAP, a.step_cap)
    ctx.set_tuning(ctx.TUNE_GRAPH_ROUNDS, a.graph_rounds)
    ctx.set_tuning(ctx.TUNE_WALK_BUDGET, a.walk_budget)
    ctx.set_tuning(ctx.TUNE_STEP_MIN_LIVE, a.step_min_live)
    torch.cuda.synchronize()
    tp = time.time()
    sc = aarbf.Scene.prepare(ctx, wl.xyzw, wl.colormap, wl.radiusmap, wl.densitymap, wl.rbf_radius,
                             radii=wl.radii, ppl=1, branching=8)
    torch.cuda.synchronize()
    prep_s = time.time() - tp
    log(f"pre-process + bounds + BVH: {prep_s * 1e3:.1f} ms (includes host->device copy)")

    # a timestep change (leaf bounds + refit, hostCode.cpp:790-794) and a
    # radius / transfer-function change (leaf bounds + a new tree, :781-789)
    def upd_ms(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        t = time.time()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.time() - t) * 1e3 / reps

    refit_ms, rebuild_ms = upd_ms(sc.refit), upd_ms(sc.rebuild)
    log(f"bounds + refit {refit_ms:.2f} ms, bounds + rebuild {rebuild_ms:.2f} ms")

    image = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    accum = torch.zeros_like(image)
    taa = torch.zeros_like(image)
    taa_prev = torch.zeros_like(image)
    frame_buf = torch.zeros((H, W), dtype=torch.int32, device="cuda")
    rg = sc.raygen_data(image, accum, W, H)
    tc = aarbf.TAAConstants(0, 0, aarbf._lib.Int2(W, H), image.data_ptr(), 0, frame_buf.data_ptr(),
                            taa.data_ptr(), taa_prev.data_ptr())
    cams = []
    for f in range(wl.frames):
        lf, la, lu = scenes.orbit_camera(f, wl.frames, 1.6)
        cams.append(aarbf.camera_basis(lf, la, lu, wl.cos_fovy, W, H))

    ev = []

    def frame(f, timed_kernel=False):
        cam = cams[f % len(cams)]
        pc = sc.push_constants(cam, wl.unit, frame_id=(f * S) & 0xFFFF, accum_id=1)
        if timed_kernel:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
        ctx.render(rg, pc, W, H, S)
        if timed_kernel:
            e1.record(stream)
            ev.append((e0, e1))
        ctx.taa_resolve(tc)

    # instrumented launch (statistics kernel) for the algorithmic byte count
    ctx.set_stats(True)
    ctx.read_stats(reset=True)
    frame(0)
    torch.cuda.synchronize()
    st = ctx.read_stats(reset=True)
    hist = ctx.read_histogram(reset=True)
    ctx.set_stats(False)
    log("gather walk cost histogram (log2 nodes -> count): " + str({b: v for b, v in enumerate(hist[:18]) if v}))
    for name, o in (("step", 18), ("dense", 21)):
        if hist[o]:
            log(f"{name} kernel: {hist[o]} wave iterations, lane utilisation {hist[o + 1] / (64 * hist[o]):.3f}, "
                f"iterations testing candidates {hist[o + 2] / hist[o]:.3f}")
    samples = max(1, st["samples"])
    # Algorithmic bytes of one render call (per sample: the BVH nodes the
    # segment gathers test, at the quantised group's 80 B per 8 children;
    # 16 B position + radius per particle tested, by a gather or a step;
    # 16 B colour payload per covering particle; 32 B for the sample's
    # colour write and read-back into the running mean), plus the accum
    # read + write and the image write, 48 B per pixel.
    node_bytes = 80 / 8
    bytes_launch = (node_bytes * st["nodes"] + 32 * st["leaves"] + 16 * st["particles"] + 16 * st["covering"]
                    + 32 * samples + 48 * W * H)
    log(f"stats per sample: steps {st['steps'] / samples:.0f} queries {st['queries'] / samples:.0f} "
        f"nodes {st['nodes'] / samples:.0f} particles {st['particles'] / samples:.0f} "
        f"covering {st['covering'] / samples:.1f} exhausted {st['exhausted'] / samples:.4f}")

    for f in range(a.warmup):
        frame(f)
    torch.cuda.synchronize()

    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    ctx.set_timing(False)  # the timed frames carry no per-kernel events
    torch.cuda.synchronize()
    w0 = time.time()
    t_start.record(stream)
    for k in range(a.steps):
        frame(a.warmup + k, timed_kernel=True)
        if (k + 1) % 4 == 0:
            torch.cuda.synchronize()
            log(f"frame {k + 1}/{a.steps} done, {time.time() - w0:.1f} s")
    t_end.record(stream)
    torch.cuda.synchronize()
    wall = time.time() - w0
    gpu_ms = t_start.elapsed_time(t_end)
    kern_ms = float(np.mean([e0.elapsed_time(e1) for e0, e1 in ev]))
    # per-kernel split of the render rounds (mode 0): the library's HIP
    # events on the same stream, in an untimed pass over the first timed
    # frames (events and per-round waits make it a little slower)
    kernels = None
    if a.mode == 0:
        nk = min(2, a.steps)
        ctx.set_timing(True)
        ctx.read_timing(reset=True)
        for k in range(nk):
            frame(a.warmup + k)
        torch.cuda.synchronize()
        ktime = ctx.read_timing(reset=True)
        ctx.set_timing(False)
        kernels = {k: {"ms_per_frame": round(ms / nk, 3), "launches_per_frame": round(n / nk, 2),
                       "ms_per_launch": round(ms / n, 4) if n else None}
                   for k, (ms, n) in ktime.items()}
        kernels["frames"] = nk
    total_samples = W * H * S * a.steps
    msps = total_samples / (gpu_ms * 1e-3) / 1e6
    achieved = bytes_launch / (kern_ms * 1e-3) / 1e9

    traffic = None
    if a.traffic and os.path.exists(a.traffic):
        try:
            traffic = json.load(open(a.traffic)).get("bytes_per_launch")
        except Exception:
            traffic = None
    # the step kernel's VALU issue rate (fraction of the nominal peak), from
    # the profiled SQ counters of this workload
    valu = None
    vpath = os.path.join(ROOT, "profiles", "render_valu.json")
    if os.path.exists(vpath):
        try:
            v = json.load(open(vpath))
            valu = {"kernel": v["kernel"], "busy": v["valu_busy"], "source": v["source"]}
        except Exception:
            valu = None

    prep = preprocess_leg(ctx, stream, wl, sc, cams, prep_s)
    if a.prep_67m:
        prep["hierarchical_67m"] = preprocess_67m_leg(ctx, stream)

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(sc, wl, cams[0], a.cpu_threads)
    ctx.set_timing(False)
    ctx.set_render_mode(0)
    ctx.set_payload(0)
    fbatch = frame_batch_leg(ctx, stream, a, sc, wl, cams, msps if a.payload == 0 and a.mode == 0 else None) \
        if a.frame_batch > 0 else None
    splat = splat_leg(ctx, stream, a, sc, wl, cams) if a.splat_frames > 0 else None
    diff = diff_leg(ctx, stream, a) if a.diff_views > 0 else None
    batch = batch_leg(ctx, stream, a) if a.batch_views > 0 else None

    out = {
        "metric": "Msamples/s at 8M RBFs 1080p delta tracking; fwd+bwd ms/view; 1 MI355X",
        "value": round(msps, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(gpu_ms / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 32-component Gaussian mixture, no dataset)",
        "config": {
            "workload": f"mixture-{wl.xyzw.shape[0]} delta tracking {W}x{H} {S} spp/frame, TAA, orbit of "
                        f"{wl.frames} frames",
            "particles": int(wl.xyzw.shape[0]), "width": W, "height": H, "spp": S, "particles_per_leaf": 1,
            "bvh_branching": 8, "unit": wl.unit, "render_mode": a.mode,
            "payload": ("fp32", "fp16 leaf-local quantised")[a.payload], "max_depth": aarbf.MAX_DEPTH,
            "exhausted_fraction": st["exhausted"] / samples, "preprocess_ms": round(prep_s * 1e3, 1),
            "refit_ms": round(refit_ms, 2), "rebuild_ms": round(rebuild_ms, 2),
            "step_cap": a.step_cap, "step_min_live": a.step_min_live,
            "parallelism": "single GPU", "timed_schedule": "host-driven rounds, no per-kernel events",
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": "aarbf_render_rbf_spp (wf_gather + wf_step + dense wf_step rounds, wf_fold)",
            "kernel_ms": round(kern_ms, 3),
            "algorithmic_bytes_per_launch": int(bytes_launch),
            "algorithmic_bytes_per_sample": round(bytes_launch / samples, 1),
            "kernels": kernels,
            "valu": valu,
        },
        "cpu_baseline": cpu,
        "preprocess": prep,
        "frame_batch": fbatch,
        "splat": splat,
        "diff": diff,
        "batch": batch,
        "wall_s": round(wall, 3),
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


def frame_batch_leg(ctx, stream, a, sc, wl, cams, plain_msps):
    """The frames the line's value is timed on (the main workload's orbit
    frames warmup .. warmup + steps - 1, 1920x1080, 16 spp delta tracking
    each; at most --frame-batch of them) rendered by one aarbf_render_batch
    call, with the FP16 leaf-local payload and with the FP32 payload,
    against the plain FP32 path (one aarbf_render_rbf_spp call per frame,
    the line's value; over fewer frames than it times, the same frames are
    also timed one call each).  Each view's image equals its own call's on
    the same payload bit for bit."""
    import torch
    W, H, S = wl.width, wl.height, wl.spp
    K = min(a.steps, a.frame_batch)
    f0 = a.warmup
    bufs = [(torch.zeros((H, W, 4), dtype=torch.float32, device="cuda"),
             torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")) for _ in range(K)]
    reqs = []
    for k in range(K):
        f = f0 + k
        pc = sc.push_constants(cams[f % len(cams)], wl.unit, frame_id=(f * S) & 0xFFFF, accum_id=1)
        reqs.append((sc.raygen_data(bufs[k][0], bufs[k][1], W, H), pc, W, H, S))
    out = {"workload": f"orbit frames {f0}..{f0 + K - 1} of the main workload (the timed frames), one "
                       f"aarbf_render_batch call", "frames": K}
    for payload, name in ((1, "fp16"), (0, "fp32")):
        ctx.set_payload(payload)
        ctx.render_batch(reqs)  # warm-up (slot pool, payload)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        ctx.render_batch(reqs)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        out[f"{name}_batched_ms"] = round(ms, 2)
        out[f"{name}_batched_msamples_per_s"] = round(W * H * S * K / (ms * 1e-3) / 1e6, 3)
        log(f"frame batch ({name} payload): {K} frames in {ms:.1f} ms")
        # the same frames one call each
        e0.record(stream)
        for rq in reqs:
            ctx.render(*rq)
        e1.record(stream)
        torch.cuda.synchronize()
        ms1 = e0.elapsed_time(e1)
        out[f"{name}_one_call_msamples_per_s"] = round(W * H * S * K / (ms1 * 1e-3) / 1e6, 3)
    ctx.set_payload(0)
    if plain_msps and K < a.steps:
        plain_msps = out["fp32_one_call_msamples_per_s"]
    if plain_msps:
        out["plain_fp32_msamples_per_s"] = round(plain_msps, 3)
        out["fp16_batched_over_plain_fp32"] = round(out["fp16_batched_msamples_per_s"] / plain_msps, 3)
    del bufs
    torch.cuda.synchronize()
    return out


def splat_leg(ctx, stream, a, sc, wl, cams):
    """The splat path (ParticleSplatRayGen, deviceCodeSplat.hlsl:29-194,
    one ray per pixel) on the main workload's scene and orbit cameras.
    One frame at a quarter of each side is timed first; the full-size
    frames follow only when that frame predicts at most 20 s per full
    frame, so the leg stays bounded whatever the scene."""
    import torch
    import aarbf
    from aarbf import scenes
    W, H = wl.width, wl.height
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(w, h, cam):
        img = torch.zeros((h, w, 4), dtype=torch.float32, device="cuda")
        rg = sc.raygen_data(img, img, w, h)
        e0.record(stream)
        ctx.render_splat(rg, sc.push_constants(cam, wl.unit), w, h)
        e1.record(stream)
        torch.cuda.synchronize()
        # every pixel ends over an opaque black background, so lit = some colour
        return e0.elapsed_time(e1), float((img[..., :3].amax(-1) > 0).float().mean())

    qw, qh = (W + 3) // 4, (H + 3) // 4
    lf, la, lu = scenes.orbit_camera(1, len(cams), 1.6)
    qcam = aarbf.camera_basis(lf, la, lu, wl.cos_fovy, qw, qh)
    q_ms, _ = timed(qw, qh, qcam)
    log(f"splat {qw}x{qh}: {q_ms:.1f} ms")
    out = {"workload": f"splat, one ray per pixel, scene of the main workload, orbit frames 1..{a.splat_frames}",
           "quarter_size": [qw, qh], "quarter_ms": round(q_ms, 2)}
    if q_ms * (W * H) / (qw * qh) > 20e3:
        out["skipped_full_size"] = "predicted over 20 s per frame"
        return out
    ms = []
    lit = 0.0
    for k in range(a.splat_frames):
        t, lit = timed(W, H, cams[(1 + k) % len(cams)])
        ms.append(t)
        log(f"splat {W}x{H} frame {k + 1}/{a.splat_frames}: {t:.1f} ms")
    m = float(np.mean(ms))
    out.update({"width": W, "height": H, "frames": a.splat_frames, "ms_per_frame": round(m, 2),
                "mrays_per_s": round(W * H / (m * 1e-3) / 1e6, 3), "lit_fraction": round(lit, 3)})
    return out


def diff_leg(ctx, stream, a):
    """Differentiable rendering: forward and forward + adjoint ms per view on
    the differentiation workload (1,048,576 particles from the 32-component
    mixture with attributes normalised to [0, 1], 512x512, 4 spp expected
    stochastic marching, 256-entry colormap and densitymap, L2 image loss
    against a perturbed copy of the scene), library calls timed with HIP
    events on the library's stream; the per-view scene rebuild (sort,
    bounds, BVH) is timed separately."""
    import torch
    from aarbf import scenes
    from aarbf.diff import _Render
    t0 = time.time()
    dw = scenes.diff_workload(n=a.diff_particles)
    xyzw, radii, tgt, tgt_r = dw.xyzw, dw.radii, dw.target_xyzw, dw.target_radii
    cm, dm, rm = dw.colormap, dw.densitymap, dw.radiusmap
    dev = torch.device("cuda")
    T = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    W, H = dw.width, dw.height
    spp, unit, rbf = dw.spp, dw.unit, dw.rbf_radius
    log(f"diff leg: workload generated in {time.time() - t0:.1f} s")
    fw, bw, rebuild = [], [], []
    for v in range(a.diff_views):
        lf, la, up = dw.view(v)
        cam = aarbf_camera(lf, la, up, W, H)
        torch.cuda.synchronize()
        tr = time.time()
        target = _Render(ctx, cam, W, H, unit, spp, v, rbf, 1, 8, T(tgt[:, :3]), T(tgt[:, 3]), T(tgt_r), T(cm),
                         T(dm), T(rm))
        timg = target.forward()
        target.close()
        torch.cuda.synchronize()
        tr = time.time()
        r = _Render(ctx, cam, W, H, unit, spp, v, rbf, 1, 8, T(xyzw[:, :3]), T(xyzw[:, 3]), T(radii), T(cm),
                    T(dm), T(rm))
        torch.cuda.synchronize()
        rebuild.append((time.time() - tr) * 1e3)
        n = len(xyzw)
        g = torch.zeros((H, W, 4), dtype=torch.float32, device=dev)
        grads = [torch.zeros(k, dtype=torch.float32, device=dev) for k in (3 * n, n, n, 256 * 4, 256, 4)]
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        torch.cuda.synchronize()
        e[0].record(stream)
        ctx.diff_forward(r.rg, r.pc, W, H, spp, r.img, r.path)
        e[1].record(stream)
        torch.cuda.synchronize()
        g[..., :3] = 2.0 * (r.img[..., :3] - timg)  # dL/dimage of the L2 loss
        torch.cuda.synchronize()
        e2 = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        e2[0].record(stream)
        ctx.diff_backward(r.rg, r.pc, W, H, spp, g, *grads, path=r.path)
        e2[1].record(stream)
        torch.cuda.synchronize()
        fw.append(e[0].elapsed_time(e[1]))
        bw.append(e2[0].elapsed_time(e2[1]))
        r.close()
        log(f"diff view {v}: forward {fw[-1]:.2f} ms, adjoint {bw[-1]:.2f} ms, rebuild {rebuild[-1]:.1f} ms")
    f, b = float(np.mean(fw)), float(np.mean(bw))
    return {"workload": f"mixture-{a.diff_particles} expected marching 512x512 4 spp, L2 loss vs a perturbed copy, "
                        f"{a.diff_views} views on a seeded sphere",
            "fwd_ms_per_view": round(f, 3), "fwd_bwd_ms_per_view": round(f + b, 3),
            "ratio": round((f + b) / f, 3), "rebuild_ms_per_view": round(float(np.mean(rebuild)), 2),
            "unit": unit, "spp": spp}


def aarbf_camera(look_from, look_at, up, W, H):
    import aarbf
    from aarbf import scenes
    return aarbf.camera_basis(look_from, look_at, up, scenes.fovy_scale(45.0), W, H)


def batch_leg(ctx, stream, a):
    """Batching: camera requests of uneven resolution (side log-uniform in
    [128, 1024], aspect in [0.5, 2], 1-8 spp, half stochastic marching with
    the synthetic blue-noise atlas, half delta tracking) over the
    64-component mixture, FP16 payload by default; the schedule autotuned on
    the first 32 requests; one aarbf_render_batch call timed against the
    same requests rendered one call each (latency of a request = time from
    the start of the set to its last sample)."""
    import torch
    import aarbf
    from aarbf import scenes
    t0 = time.time()
    xyzw, radii = scenes.mixture_16m(n=a.batch_particles)
    views = scenes.batch_views(n=a.batch_views)
    log(f"batch leg: workload generated in {time.time() - t0:.1f} s")
    ctx.set_payload(a.batch_payload)
    sc = aarbf.Scene.prepare(ctx, xyzw, scenes.reference_tf("colormap"), scenes.reference_tf("radiusmap"),
                             scenes.reference_tf("densitymap"), float(radii.max()), radii=radii, ppl=1, branching=8)
    stbn = ctx.stbn(scenes.synthetic_stbn(seed=3))
    unit = 0.0004
    bufs, reqs = [], []
    for v in views:
        W, H = v["width"], v["height"]
        img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
        acc = torch.zeros_like(img)
        bufs.append((img, acc))
        cam = aarbf_camera(v["look_from"], v["look_at"], v["up"], W, H)
        pc = sc.push_constants(cam, unit, frame_id=v["frame"], accum_id=1, blue_noise=v["marching"])
        reqs.append((sc.raygen_data(img, acc, W, H, stbn=stbn if v["marching"] else 0), pc, W, H, v["spp"]))
    torch.cuda.synchronize()
    tuned = ctx.autotune(reqs[:32], reps=1)
    log(f"batch leg: autotuned {tuned}")
    samples = sum(W * H * spp for (_, _, W, H, spp) in reqs)
    ctx.render_batch(reqs)  # warm-up (graph capture, buffers)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    ctx.render_batch(reqs)
    e1.record(stream)
    torch.cuda.synchronize()
    batch_ms = e0.elapsed_time(e1)
    lat = ctx.batch_latency(len(reqs))
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(len(reqs) + 1)]
    ev[0].record(stream)
    for k, (rg, pc, W, H, spp) in enumerate(reqs):
        ctx.render(rg, pc, W, H, spp)
        ev[k + 1].record(stream)
    torch.cuda.synchronize()
    seq = np.array([ev[0].elapsed_time(e) for e in ev[1:]])
    ctx.set_payload(0)
    for k in (ctx.TUNE_STEP_CAP, ctx.TUNE_STEP_WAVES, ctx.TUNE_GRAPH_ROUNDS, ctx.TUNE_WALK_BUDGET):
        ctx.set_tuning(k, {ctx.TUNE_STEP_CAP: a.step_cap, ctx.TUNE_STEP_WAVES: a.step_waves,
                           ctx.TUNE_GRAPH_ROUNDS: a.graph_rounds, ctx.TUNE_WALK_BUDGET: a.walk_budget}[k])
    return {"workload": f"mixture-{a.batch_particles} (64 components), {len(reqs)} requests, side log-U[128,1024], "
                        "aspect U[0.5,2], spp U{1..8}, half marching / half delta tracking",
            "payload": ("fp32", "fp16 leaf-local quantised")[a.batch_payload],
            "samples": int(samples), "batch_ms": round(batch_ms, 2),
            "msamples_per_s": round(samples / (batch_ms * 1e-3) / 1e6, 3),
            "latency_ms_p50": round(float(np.percentile(lat, 50)), 2),
            "latency_ms_p99": round(float(np.percentile(lat, 99)), 2),
            "sequential_ms": round(float(seq[-1]), 2),
            "sequential_msamples_per_s": round(samples / (seq[-1] * 1e-3) / 1e6, 3),
            "sequential_latency_ms_p50": round(float(np.percentile(seq, 50)), 2),
            "sequential_latency_ms_p99": round(float(np.percentile(seq, 99)), 2),
            "autotuned": tuned}


def copy_calibration(nbytes: int) -> int:
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    log(f"copied {src.numel() * 4} bytes")
    return 0


def cpu_baseline(sc, wl, cam, threads):
    """The oracle (scalar C restatement of the reference shaders, pthreads
    over pixels) on frame 0 at 2 spp: on all the cores this process may use
    (at most 16, the host share of one GPU) over every 4th pixel in x and y,
    and on one core over every 16th pixel."""
    import oracle
    p = sc.particles.cpu().numpy()
    r = sc.radii.cpu().numpy() if sc.radii is not None else None
    W, H = wl.width, wl.height
    nproc = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = nproc
    threads = threads or min(16, avail)
    R = oracle.Render(p, r, sc.aabb, wl.colormap, wl.radiusmap, wl.densitymap, cam, W, H, wl.unit, wl.rbf_radius,
                      frame_id=0, accum_id=1)
    t0 = time.time()
    prep = R.prepared()
    setup = time.time() - t0
    spp = 2

    def run(step, nthreads):
        ys, xs = np.mgrid[0:H:step, 0:W:step]
        pix = np.stack([xs.ravel(), ys.ravel()], 1).astype(np.int32)
        t = time.time()
        prep.render(pix, spp=spp, threads=nthreads)
        dt = time.time() - t
        n = len(pix) * spp
        log(f"cpu baseline: {n} samples in {dt:.2f} s on {nthreads} thread(s) (scene set-up {setup:.1f} s)")
        return n / dt / 1e6, n

    v_all, n_all = run(4, threads)
    v_one, n_one = run(16, 1)
    return {"value": round(v_all, 6), "unit": "Msamples/s", "cores": threads, "kind": "port",
            "sample": f"frame 0, {spp} spp, every 4th pixel in x and y ({n_all} samples of the {W}x{H} frame); "
                      f"oracle/oracle.c -O2 -ffp-contract=off, {threads} threads, grid set-up {setup:.1f} s untimed",
            "nproc": nproc, "cores_available": avail,
            "single_core": {"value": round(v_one, 6), "unit": "Msamples/s", "cores": 1,
                            "sample": f"frame 0, {spp} spp, every 16th pixel in x and y ({n_one} samples)"}}


# Bytes per particle that each pre-processing stage moves as implemented
# (every read and write of a pass counted once): space-time AABB reads the
# float4; the Hilbert pass reads it and writes the 8-byte code; the sort
# copies the codes (16), writes the index (4), runs six 8-bit radix passes
# over 48-bit keys with 32-bit values (histogram read 8 + scatter read 12 +
# write 12 each), gathers the particles (36), checks runs of equal codes
# (24) and copies particles, codes and permutation back (56); the radii
# follow the permutation (12); attribute range 16, normalisation 32; leaf
# bounds read particle + radius and write the 24-byte box; the BVH reads the
# leaf boxes and writes 32-byte level-0 nodes plus the internal levels and
# quantised groups (about 12 more at 8 children per node).
PREP_STAGE_BYTES = {"aabb": 16, "hilbert": 24, "sort": 16 + 4 + 6 * 32 + 36 + 24 + 56, "radii": 12, "range": 16,
                    "normalize": 32, "bounds": 44, "bvh": 24 + 32 + 12}


def prep_stages(ctx, stream, p_raw, r_raw, radiusmap, rbf, ppl=1, K=8):
    """The pre-processing stages (hostCode.cpp:183-250, then GenRBFBounds
    and the BVH build) on device-resident input, each bracketed by HIP
    events on the library's stream.  Returns {stage: ms}."""
    import torch
    p = p_raw.clone()
    r = r_raw.clone() if r_raw is not None else None
    torch.cuda.synchronize()
    ev = [("start", torch.cuda.Event(enable_timing=True))]
    ev[0][1].record(stream)

    def mark(name):
        e = torch.cuda.Event(enable_timing=True)
        e.record(stream)
        ev.append((name, e))

    aabb = ctx.spacetime_aabb(p)
    mark("aabb")
    codes = ctx.hilbert_codes(p, aabb)
    mark("hilbert")
    perm = ctx.sort_particles(p, codes)
    mark("sort")
    if r is not None:
        r = r[perm.long()].contiguous()
        mark("radii")
    rng = ctx.attribute_range(p)
    mark("range")
    ctx.normalize_attributes(p, rng)
    mark("normalize")
    aabbs = ctx.gen_bounds(p, ppl, rbf, radiusmap, r)
    mark("bounds")
    bvh = ctx.bvh_build(aabbs, K)
    mark("bvh")
    torch.cuda.synchronize()
    out = {name: ev[i][1].elapsed_time(e) for i, (name, e) in enumerate(ev[1:])}
    ctx.bvh_destroy(bvh)
    del p, r, codes, perm, aabbs
    return out


def _prep_summary(stages, n):
    total = sum(stages.values())
    nbytes = sum(PREP_STAGE_BYTES[k] for k in stages) * n
    return {"stage_ms": {k: round(v, 3) for k, v in stages.items()}, "total_ms": round(total, 3),
            "bytes": int(nbytes), "gbs": round(nbytes / (total * 1e-3) / 1e9, 1),
            "frac": round(nbytes / (total * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}


def preprocess_leg(ctx, stream, wl, sc, cams, prep_s):
    """Pre-processing of the main workload: per-stage device time on
    device-resident input (GB/s), and the time to the first frame from host
    arrays on a fresh context (context creation, host-to-device copy,
    pre-processing, bounds, BVH, one 16-spp 1080p frame)."""
    import torch
    import aarbf
    n = int(wl.xyzw.shape[0])
    p_raw = torch.from_numpy(wl.xyzw).cuda()
    r_raw = torch.from_numpy(wl.radii).cuda()
    prep_stages(ctx, stream, p_raw, r_raw, sc.radiusmap, wl.rbf_radius)  # warm-up
    out = _prep_summary(prep_stages(ctx, stream, p_raw, r_raw, sc.radiusmap, wl.rbf_radius), n)
    del p_raw, r_raw
    log(f"pre-process stages at {n} particles: {out}")
    # time to first frame, fresh context
    torch.cuda.synchronize()
    t0 = time.time()
    c2 = aarbf.Context(0)
    s2 = aarbf.Scene.prepare(c2, wl.xyzw, wl.colormap, wl.radiusmap, wl.densitymap, wl.rbf_radius, radii=wl.radii,
                             ppl=1, branching=8)
    img = torch.zeros((wl.height, wl.width, 4), dtype=torch.float32, device="cuda")
    acc = torch.zeros_like(img)
    c2.render(s2.raygen_data(img, acc, wl.width, wl.height), s2.push_constants(cams[0], wl.unit), wl.width,
              wl.height, wl.spp)
    c2.sync()
    ttff = time.time() - t0
    c2.close()
    del s2, img, acc
    torch.cuda.empty_cache()
    out.update({"workload": f"mixture-{n}", "prepare_ms_host_arrays": round(prep_s * 1e3, 1),
                "time_to_first_frame_ms": round(ttff * 1e3, 1),
                "time_to_first_frame": "fresh context: host arrays -> device, pre-process, bounds, BVH, first "
                                       f"{wl.spp}-spp {wl.width}x{wl.height} frame"})
    log(f"time to first frame (8M, fresh context): {ttff * 1e3:.0f} ms")
    out.update(file_load_leg(wl, cams))
    return out


def file_load_leg(wl, cams):
    """The particle file (importParticles.h format, written to a temporary
    directory just before, so it is read from the page cache) streamed to
    the device by aarbf_import_step_device on a fresh context, then
    pre-processed and rendered: load GB/s and the time to the first frame
    from the file (radii, which the format does not hold, from memory)."""
    import shutil
    import tempfile
    import torch
    import aarbf
    from aarbf import series
    d = tempfile.mkdtemp(prefix="aarbf_bench_")
    try:
        path = os.path.join(d, "particles.txt")
        series.export_particles(path, [("t0", wl.xyzw)])
        torch.cuda.synchronize()
        c2 = aarbf.Context(0)
        t0 = time.time()
        dev = series.import_step_device(c2, path, 0)
        t_load = time.time() - t0
        s2 = aarbf.Scene.prepare(c2, dev, wl.colormap, wl.radiusmap, wl.densitymap, wl.rbf_radius, radii=wl.radii,
                                 ppl=1, branching=8)
        img = torch.zeros((wl.height, wl.width, 4), dtype=torch.float32, device="cuda")
        acc = torch.zeros_like(img)
        c2.render(s2.raygen_data(img, acc, wl.width, wl.height), s2.push_constants(cams[0], wl.unit), wl.width,
                  wl.height, wl.spp)
        c2.sync()
        ttff = time.time() - t0
        c2.close()
        del s2, img, acc, dev
        torch.cuda.empty_cache()
    finally:
        shutil.rmtree(d, ignore_errors=True)
    nbytes = 16 * wl.xyzw.shape[0]
    log(f"file -> device: {nbytes / t_load / 1e9:.2f} GB/s; time to first frame from the file {ttff * 1e3:.0f} ms")
    return {"file_load_ms": round(t_load * 1e3, 1), "file_load_gbs": round(nbytes / t_load / 1e9, 2),
            "time_to_first_frame_from_file_ms": round(ttff * 1e3, 1)}


def preprocess_67m_leg(ctx, stream):
    """The 67,108,864-particle precision scene (3-level hierarchical
    clusters, generated on the device): per-stage pre-processing time and
    the time from device-resident particles to the first 3840x2160 8-spp
    frame with the FP16 payload."""
    import torch
    import aarbf
    from aarbf import scenes
    n, W, H, spp, unit = 67108864, 3840, 2160, 8, 0.0004
    t0 = time.time()
    xyzw, radii = scenes.hierarchical_clusters(n=n, seed=13, device="cuda")
    torch.cuda.synchronize()
    gen_s = time.time() - t0
    tfs = (scenes.reference_tf("colormap"), scenes.reference_tf("radiusmap"), scenes.reference_tf("densitymap"))
    rbf = float(radii.max())
    rm = ctx.tf(tfs[1])
    stages = prep_stages(ctx, stream, xyzw, radii, rm, rbf)
    ctx.tf_destroy(rm)
    out = _prep_summary(stages, n)
    torch.cuda.synchronize()
    e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    e[0].record(stream)
    sc = aarbf.Scene.prepare(ctx, xyzw, *tfs, rbf, radii=radii, ppl=1, branching=8)
    e[1].record(stream)
    lf, la, lu = scenes.orbit_camera(1, 32, 1.4)
    cam = aarbf.camera_basis(lf, la, lu, scenes.fovy_scale(45.0), W, H)
    img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    acc = torch.zeros_like(img)
    ctx.set_payload(ctx.PAYLOAD_FP16)
    try:
        ctx.render(sc.raygen_data(img, acc, W, H), sc.push_constants(cam, unit, frame_id=8), W, H, spp)
        e[2].record(stream)
        torch.cuda.synchronize()
    finally:
        ctx.set_payload(ctx.PAYLOAD_FP32)
    prep_ms, frame_ms = e[0].elapsed_time(e[1]), e[1].elapsed_time(e[2])
    for h in (sc.colormap, sc.radiusmap, sc.densitymap):
        ctx.tf_destroy(h)
    ctx.bvh_destroy(sc.bvh)
    del sc, xyzw, radii, img, acc
    torch.cuda.empty_cache()
    out.update({"workload": "hierarchical-67108864 (3 levels x 16^3 clusters, power-law radii), generated on the "
                            "device", "generate_s": round(gen_s, 2), "prepare_ms": round(prep_ms, 1),
                "first_frame_ms": round(frame_ms, 1), "time_to_first_frame_ms": round(prep_ms + frame_ms, 1),
                "first_frame": f"{W}x{H}, {spp} spp delta tracking, FP16 payload"})
    log(f"67M: {out}")
    return out


if __name__ == "__main__":
    sys.exit(main())
