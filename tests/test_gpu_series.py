"""Time series (hostCode.cpp:133-250, 751-799): timesteps of different
sizes written in the importer's format and read back, pre-processed across
steps (one space-time box, a Hilbert sort per step against it, one
attribute range), then rendered step by step with the leaf bounds
regenerated and the BVH refitted.  Each step's prepared particles equal the
oracle's pre-processing of that step with the series-wide box and range;
each step's image equals the oracle's bit for bit, and equals the image of
a tree built from scratch for the step."""
import numpy as np
import pytest

import aarbf
from aarbf import scenes, series
from helpers import default_camera, random_scene
import oracle

pytestmark = pytest.mark.gpu


def _render(ctx, sc, cam, W, H, frame):
    import torch
    img = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    acc = torch.zeros_like(img)
    ctx.render(sc.raygen_data(img, acc, W, H), sc.push_constants(cam, 0.004, frame_id=frame, light=(0.3, 0.2, 0.4)),
               W, H, 2)
    ctx.sync()
    return img.cpu().numpy(), acc.cpu().numpy()


@pytest.mark.parametrize("ppl", [1, 3])
def test_time_series_refit(gpu_ctx, tmp_path, ppl):
    steps = []
    for k, n in enumerate((3000, 2201, 2600)):
        xyzw, radii = random_scene(n, seed=50 + k, extent=1.0, r=(0.03, 0.08))
        xyzw[:, :3] += np.float32(0.05 * k)  # the cloud drifts
        xyzw[:, 3] += np.float32(k)          # the attribute range spans the steps
        steps.append((str(10 * k), xyzw, radii))
    path = str(tmp_path / "series.txt")
    series.export_particles(path, [(nm, x) for nm, x, _ in steps])
    back = series.import_particles(path)
    assert [nm for nm, _ in back] == [nm for nm, _, _ in steps]
    for (_, a), (_, x, _) in zip(back, steps):
        assert np.array_equal(a, x)

    tfs = (scenes.ramp_table(256), scenes.constant_table(n=4), scenes.ramp_table(64))
    ts = series.TimeSeries.prepare(gpu_ctx, [a for _, a in back], *tfs, 0.08, radii=[r for _, _, r in steps],
                                   ppl=ppl)
    # series-wide box and attribute range, per-step sort (the oracle's pre-process)
    aabb = None
    for _, x, _ in steps:
        aabb = oracle.spacetime_aabb(x, aabb)
    assert np.array_equal(ts.aabb, aabb)
    prepared = []
    for _, x, r in steps:
        p, _, perm, = oracle.sort_particles(x, oracle.hilbert_codes(x, aabb))
        prepared.append((p, r[perm]))
    rng = None
    for p, _ in prepared:
        rng = oracle.attribute_range(p, rng)
    assert np.array_equal(ts.attr_range, rng)
    prepared = [(oracle.normalize_attributes(p, rng), r) for p, r in prepared]
    for k, (p, r) in enumerate(prepared):
        assert np.array_equal(ts.steps[k].cpu().numpy(), p)
        assert np.array_equal(ts.radii[k].cpu().numpy(), r)

    W, H = 40, 32
    cam = default_camera(W, H, look_from=(0.45, 0.35, 2.4), look_at=(0.05, 0.05, 0.0))
    for k in (1, 2, 0, 2):  # refits, including to a larger step than the last
        sc = ts.set_step(k)
        assert sc.n == prepared[k][0].shape[0]
        img, acc = _render(gpu_ctx, sc, cam, W, H, frame=k)
        p, r = prepared[k]
        o = oracle.Render(p, r, aabb, *tfs, cam, W, H, 0.004, 0.08, frame_id=k, accum_id=1, light=(0.3, 0.2, 0.4))
        oimg, oacc, _ = o.run(spp=2, threads=8)
        assert (oacc[..., :3] > 0).mean() > 0.05
        assert np.array_equal(acc, oacc) and np.array_equal(img, oimg)
        ts.set_step(k, rebuild=True)
        img2, acc2 = _render(gpu_ctx, ts.scene, cam, W, H, frame=k)
        assert np.array_equal(acc2, acc) and np.array_equal(img2, img)


def test_import_step_device_streams_the_file(gpu_ctx, tmp_path):
    """aarbf_import_step_device reads a timestep straight into device memory
    (chunks larger than one staging buffer included): the same particles as
    the host importer; a buffer too small for the step is refused with the
    count reported; a step the header does not list is refused."""
    import torch
    from aarbf import series
    rng = np.random.default_rng(4)
    steps = [("a", rng.normal(size=(1000, 4)).astype(np.float32)),
             ("b", rng.normal(size=(4_500_000, 4)).astype(np.float32))]  # 72 MB: two staging chunks
    path = str(tmp_path / "particles.txt")
    series.export_particles(path, steps)
    for k, (_, a) in enumerate(steps):
        got = series.import_step_device(gpu_ctx, path, k)
        assert np.array_equal(got.cpu().numpy(), a), f"step {k}"
    buf = torch.zeros((5_000_000, 4), dtype=torch.float32, device="cuda")
    got = series.import_step_device(gpu_ctx, path, 1, out=buf)
    assert got.data_ptr() == buf.data_ptr() and np.array_equal(got.cpu().numpy(), steps[1][1])
    small = torch.zeros((999, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(aarbf.AarbfError, match="more particles than the buffer"):
        series.import_step_device(gpu_ctx, path, 0, out=small)
    with pytest.raises(aarbf.AarbfError, match="fewer timesteps"):
        series.import_step_device(gpu_ctx, path, 2, out=buf)
