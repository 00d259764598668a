"""The path a reference caller takes, checked against the oracle:

  - the reference's own radius model (pc.radii = 0: radius = rbfRadius *
    radiusmap(w), deviceCodeRBF.hlsl:349-351) with a non-constant
    radiusmap, the transfer functions uploaded as 64 RGBA8 sRGB texels
    (hostCode.cpp:355-365,524-572) through aarbf_tf_create_from_marks from
    the reference's default marks and through aarbf_tf_create_rgba8, the
    default light (0, 0, 0.5) and unit = 0.1 rbfRadius (:481-486), one and
    four particles per leaf, both render schedules, both trackers;
  - tests/cpp/dropin_snippet: a C++ host program written the way
    INTEGRATION.md tells a maintainer to bind the library (include
    aarbf.h, link libaarbf.so), from the reference's particle file to the
    BGRA8 frame buffer over three frames of the render loop.
The bar is bit-exact output."""
import os
import subprocess

import numpy as np
import pytest

import aarbf
from aarbf import scenes, series
import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SNIPPET = os.path.join(ROOT, "tests", "cpp", "dropin_snippet")
F32 = np.float32


def _texels():
    """The reference's colormap and densitymap presets and a radiusmap
    ramp from half to full radius, as the RGBA8 bytes written into the
    R8G8B8A8_SRGB textures."""
    cm = scenes.reference_tf_texels("colormap")
    dm = scenes.reference_tf_texels("densitymap")
    rm = np.zeros((64, 4), np.uint8)
    rm[:, 0] = [scenes.make_8bit(0.6 + 0.4 * i / 63.0) for i in range(64)]
    return cm, rm, dm


def _scene(n, seed):
    rng = np.random.default_rng(seed)
    xyzw = np.empty((n, 4), F32)
    xyzw[:, :3] = rng.uniform(-0.5, 0.5, (n, 3))
    xyzw[:, 3] = rng.normal(size=n)
    return xyzw


CASES = [
    # tracker, ppl
    ("tracking", 1), ("tracking", 4), ("marching", 1), ("marching", 4),
]


@pytest.mark.parametrize("tracker,ppl", CASES)
def test_reference_radius_model_and_rgba8_tfs(gpu_ctx, tracker, ppl):
    import torch
    W, H, spp, rbf = 48, 40, 2, F32(0.06)
    unit = F32(rbf * F32(0.1))
    xyzw = _scene(2500, 31 + ppl)
    cm8, rm8, dm8 = _texels()
    # colormap and densitymap from the reference's default marks through
    # the library's mark pipeline, the radiusmap from its RGBA8 bytes
    handles = [gpu_ctx.tf_from_marks(scenes.REFERENCE_COLORMAP_MARKS, scenes.TF_COLORMAP),
               gpu_ctx.tf_rgba8(rm8, srgb=True),
               gpu_ctx.tf_from_marks(scenes.REFERENCE_DENSITYMAP_MARKS, scenes.TF_SCALAR)]
    tables = [scenes.srgb_decode_table(t) for t in (cm8, rm8, dm8)]
    try:
        sc = aarbf.Scene.prepare(gpu_ctx, xyzw, *handles, float(rbf), radii=None, ppl=ppl)
        op, _, _, oaabb, _, _ = oracle.preprocess(xyzw)
        assert sc.radii is None
        cam = aarbf.camera_basis((0.4, 0.3, 2.2), (0, 0, 0), (0, 1, 0), scenes.fovy_scale(45.0), W, H)
        blue = tracker == "marching"
        atlas = scenes.synthetic_stbn(seed=4) if blue else None
        stbn = gpu_ctx.stbn(atlas) if blue else 0
        r = oracle.Render(op, None, oaabb, *tables, cam, W, H, float(unit), float(rbf), frame_id=7, accum_id=1,
                          blue_noise=blue, stbn=atlas)
        oimg, oacc, _ = r.run(spp=spp, threads=8)
        assert (oacc[..., :3] > 0).mean() > 0.1
        for mode in (0, 1):
            img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
            acc = torch.zeros_like(img)
            gpu_ctx.set_render_mode(mode)
            try:
                pc = sc.push_constants(cam, float(unit), frame_id=7, accum_id=1, blue_noise=blue)
                assert pc.radii == 0 and (pc.light.azimuth, pc.light.elevation, pc.light.ambient) == (0, 0, 0.5)
                gpu_ctx.render(sc.raygen_data(img, acc, W, H, stbn=stbn), pc, W, H, spp)
            finally:
                gpu_ctx.set_render_mode(0)
            g_acc, g_img = acc.cpu().numpy(), img.cpu().numpy()
            bad = np.argwhere(np.any(g_acc != oacc, axis=2))
            assert bad.size == 0, f"mode {mode}: {len(bad)} pixels differ, first {bad[:3].tolist()}"
            assert np.array_equal(g_img, oimg)
    finally:
        for h in handles:
            gpu_ctx.tf_destroy(h)


def test_cpp_dropin_program(tmp_path):
    """The C++ program renders three frames of the reference's render loop
    (render, CompositeGui, accumID / frameID + 1) from a particle file; its
    image, accumulation, TAA and frame-buffer words equal the oracle's."""
    assert os.path.exists(SNIPPET), "tests/cpp/dropin_snippet not built (python -c '__graft_entry__.build()')"
    W, H, frames, ppl, frame0 = 32, 32, 3, 2, 11
    rbf_s, from_s, at_s, up_s, fov_s = "0.05", ("0.5", "0.35", "2.0"), ("0", "0", "0"), ("0", "1", "0"), "0.66"
    xyzw = _scene(3000, 8)
    header = os.path.join(str(tmp_path), "particles.txt")
    series.export_particles(header, [("t0", xyzw)])
    cm8, rm8, dm8 = _texels()
    tex = os.path.join(str(tmp_path), "texels.bin")
    np.concatenate([cm8, rm8, dm8]).tofile(tex)
    out = os.path.join(str(tmp_path), "out")
    cmd = [SNIPPET, header, tex, out, str(W), str(H), str(frames), str(ppl), rbf_s, str(frame0), *from_s, *at_s,
           *up_s, fov_s]
    print("running", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    img = np.fromfile(out + "_image.f32", F32).reshape(H, W, 4)
    acc = np.fromfile(out + "_accum.f32", F32).reshape(H, W, 4)
    taa = np.fromfile(out + "_taa.f32", F32).reshape(H, W, 4)
    fb = np.fromfile(out + "_frame.u32", np.uint32).reshape(H, W)

    f = lambda s: float(F32(float(s)))
    rbf = F32(f(rbf_s))
    unit = float(F32(rbf * F32(0.1)))
    cam = aarbf.camera_basis([f(v) for v in from_s], [f(v) for v in at_s], [f(v) for v in up_s], f(fov_s), W, H)
    op, _, _, oaabb, _, _ = oracle.preprocess(xyzw)
    tables = [scenes.srgb_decode_table(t) for t in (cm8, rm8, dm8)]
    o_img = np.zeros((H, W, 4), F32)
    o_acc = np.zeros_like(o_img)
    o_taa = np.zeros_like(o_img)
    o_prev = np.zeros_like(o_img)
    o_fb = np.zeros((H, W), np.uint32)
    for k in range(frames):
        rr = oracle.Render(op, None, oaabb, *tables, cam, W, H, unit, float(rbf), frame_id=frame0 + k,
                           accum_id=1 + k)
        rr.run(spp=1, threads=4, image=o_img, accum=o_acc)
        oracle.taa_resolve(o_img, o_taa, o_prev, frame=o_fb)
    assert (o_acc[..., :3] > 0).mean() > 0.1
    assert np.array_equal(acc, o_acc), f"accum: {np.argwhere(np.any(acc != o_acc, axis=2))[:3].tolist()}"
    assert np.array_equal(img, o_img)
    assert np.array_equal(taa, o_taa)
    assert np.array_equal(fb, o_fb)
