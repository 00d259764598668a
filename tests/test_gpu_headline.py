"""The headline workload checked at its own size: 8,388,608 particles from
the 32-component Gaussian mixture, 1920x1080, 16 spp delta tracking per
frame, TAA along the orbit (the configuration the benchmark times).

  - pre-processing (space-time AABB, Hilbert sort, normalisation) equals the
    oracle's bit for bit at 8M particles;
  - one orbit frame rendered by aarbf_render_rbf_spp equals the oracle's
    ParticleRBFRayGen (deviceCodeRBF.hlsl:206-336) bit for bit on every
    16th pixel in x and y (8,160 pixels x 16 samples);
  - two consecutive frames resolved by aarbf_taa_resolve equal the oracle's
    CompositeGui (deviceCodeCommon.hlsl:67-115) on the GPU's images over the
    whole frame, and the images the TAA reads equal the oracle's on the 3x3
    neighbourhoods of a sparser grid."""
import numpy as np
import pytest

import aarbf
from aarbf import scenes
import oracle

pytestmark = pytest.mark.gpu

THREADS = 16


@pytest.fixture(scope="module")
def headline(gpu_ctx):
    import torch
    wl = scenes.mixture_8m()
    sc = aarbf.Scene.prepare(gpu_ctx, wl.xyzw, wl.colormap, wl.radiusmap, wl.densitymap, wl.rbf_radius,
                             radii=wl.radii, ppl=1, branching=8)
    op, _, _, oaabb, _, orad = oracle.preprocess(wl.xyzw, wl.radii)
    yield wl, sc, op, orad, oaabb
    del sc
    torch.cuda.empty_cache()


def _frame(wl, f):
    lf, la, lu = scenes.orbit_camera(f, wl.frames, 1.6)
    return aarbf.camera_basis(lf, la, lu, wl.cos_fovy, wl.width, wl.height), (f * wl.spp) & 0xFFFF


def _gpu_frame(ctx, sc, wl, f):
    import torch
    W, H = wl.width, wl.height
    img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    acc = torch.zeros_like(img)
    cam, fid = _frame(wl, f)
    ctx.render(sc.raygen_data(img, acc, W, H), sc.push_constants(cam, wl.unit, frame_id=fid, accum_id=1), W, H,
               wl.spp)
    ctx.sync()
    return img, acc


def _oracle_pixels(wl, op, orad, oaabb, f, px):
    cam, fid = _frame(wl, f)
    r = oracle.Render(op, orad, oaabb, wl.colormap, wl.radiusmap, wl.densitymap, cam, wl.width, wl.height, wl.unit,
                      wl.rbf_radius, frame_id=fid, accum_id=1)
    img, acc, _ = r.run(pixels=px, spp=wl.spp, threads=THREADS)
    return img, acc


def test_headline_preprocess_matches_oracle(headline):
    wl, sc, op, orad, oaabb = headline
    assert np.array_equal(sc.aabb, oaabb)
    assert np.array_equal(sc.particles.cpu().numpy(), op), "Hilbert-sorted, normalised particles"
    assert np.array_equal(sc.radii.cpu().numpy(), orad)


def test_headline_frame_matches_oracle(gpu_ctx, headline):
    wl, sc, op, orad, oaabb = headline
    W, H = wl.width, wl.height
    img, acc = _gpu_frame(gpu_ctx, sc, wl, 1)
    ys, xs = np.mgrid[0:H:16, 0:W:16]
    px = np.stack([xs.ravel(), ys.ravel()], 1).astype(np.int32)
    oimg, oacc = _oracle_pixels(wl, op, orad, oaabb, 1, px)
    g_img, g_acc = img.cpu().numpy()[px[:, 1], px[:, 0]], acc.cpu().numpy()[px[:, 1], px[:, 0]]
    e_img, e_acc = oimg[px[:, 1], px[:, 0]], oacc[px[:, 1], px[:, 0]]
    lit = (e_acc[:, :3].max(1) > 0).mean()
    assert lit > 0.3, f"only {lit:.2f} of the checked pixels are lit"
    rel = np.abs(g_acc - e_acc) / np.maximum(np.abs(e_acc), 1e-6)
    assert rel.max() <= 1e-4, f"{(rel > 1e-4).sum()} values beyond 1e-4 relative"
    bad = np.argwhere(np.any(g_acc != e_acc, axis=1))
    assert bad.size == 0, f"{len(bad)} of {len(px)} pixels differ, first {px[bad[0][0]].tolist()}"
    assert np.array_equal(g_img, e_img)


def test_headline_taa_matches_oracle(gpu_ctx, headline):
    import torch
    wl, sc, op, orad, oaabb = headline
    W, H = wl.width, wl.height
    taa = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    prev = torch.zeros_like(taa)
    frame = torch.zeros((H, W), dtype=torch.int32, device="cuda")
    o_taa = np.zeros((H, W, 4), np.float32)
    o_prev = np.zeros_like(o_taa)
    o_frame = np.zeros((H, W), np.uint32)
    # the centres of a grid every 128 pixels and their 3x3 neighbourhoods
    cy, cx = np.mgrid[5:H:128, 7:W:128]
    nb = np.array([(x + dx, y + dy) for x, y in zip(cx.ravel(), cy.ravel()) for dy in (-1, 0, 1)
                   for dx in (-1, 0, 1)], np.int32)
    for f in (2, 3):
        img, _ = _gpu_frame(gpu_ctx, sc, wl, f)
        tc = aarbf.TAAConstants(0, 0, aarbf._lib.Int2(W, H), img.data_ptr(), 0, frame.data_ptr(), taa.data_ptr(),
                                prev.data_ptr())
        gpu_ctx.taa_resolve(tc)
        gpu_ctx.sync()
        g_img = img.cpu().numpy()
        oimg, _ = _oracle_pixels(wl, op, orad, oaabb, f, nb)
        assert np.array_equal(g_img[nb[:, 1], nb[:, 0]], oimg[nb[:, 1], nb[:, 0]]), f"frame {f}: TAA inputs"
        oracle.taa_resolve(g_img, o_taa, o_prev, frame=o_frame)
        assert np.array_equal(taa.cpu().numpy(), o_taa), f"frame {f}: TAA"
        assert np.array_equal(prev.cpu().numpy(), o_prev), f"frame {f}: history"
        assert np.array_equal(frame.cpu().numpy().view(np.uint32), o_frame), f"frame {f}: BGRA8"
