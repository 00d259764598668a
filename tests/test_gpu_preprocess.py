"""GPU parity of the pre-process, bounds and BVH passes against the oracle
(bit-exact), through the C ABI."""
import json
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_hilbert_codes_golden(gpu_ctx):
    G = json.load(open(os.path.join(HERE, "golden", "hilbert.json")))["particles"]
    p = _dev(np.array(G["xyzw"], np.float32))
    codes = gpu_ctx.hilbert_codes(p, G["aabb"]).cpu().numpy().astype(np.uint64)
    assert codes.tolist() == G["codes"]


@pytest.mark.parametrize("n", [1, 1000, 300001])
def test_hilbert_codes_random(gpu_ctx, n):
    rng = np.random.default_rng(n)
    xyzw = rng.normal(size=(n, 4)).astype(np.float32)
    aabb = oracle.spacetime_aabb(xyzw)
    got = gpu_ctx.hilbert_codes(_dev(xyzw), aabb).cpu().numpy().astype(np.uint64)
    assert np.array_equal(got, oracle.hilbert_codes(xyzw, aabb))


def test_spacetime_aabb_and_range(gpu_ctx):
    rng = np.random.default_rng(7)
    steps = [rng.normal(size=(n, 4)).astype(np.float32) for n in (1, 4097, 100000)]
    a_gpu, a_orc = None, None
    r_gpu, r_orc = None, None
    for s in steps:
        a_gpu = gpu_ctx.spacetime_aabb(_dev(s), a_gpu)
        a_orc = oracle.spacetime_aabb(s, a_orc)
        r_gpu = gpu_ctx.attribute_range(_dev(s), r_gpu)
        r_orc = oracle.attribute_range(s, r_orc)
    assert np.array_equal(a_gpu, a_orc) and np.array_equal(r_gpu, r_orc)
    p = _dev(steps[2])
    gpu_ctx.normalize_attributes(p, r_orc)
    assert np.array_equal(p.cpu().numpy(), oracle.normalize_attributes(steps[2], r_orc))


def test_nan_particles_rejected(gpu_ctx):
    import aarbf
    x = np.zeros((10, 4), np.float32)
    x[3, 1] = np.nan
    with pytest.raises(aarbf.AarbfError):
        gpu_ctx.spacetime_aabb(_dev(x))


def _sort_case(gpu_ctx, xyzw):
    aabb = oracle.spacetime_aabb(xyzw)
    codes = oracle.hilbert_codes(xyzw, aabb)
    p = _dev(xyzw)
    c = gpu_ctx.hilbert_codes(p, aabb)
    perm = gpu_ctx.sort_particles(p, c)
    op, oc, operm = oracle.sort_particles(xyzw, codes)
    assert np.array_equal(perm.cpu().numpy().astype(np.uint32), operm)
    assert np.array_equal(p.cpu().numpy(), op)
    assert np.array_equal(c.cpu().numpy().astype(np.uint64), oc)


@pytest.mark.parametrize("n", [1, 2, 2047, 2049, 250000])
def test_sort_random(gpu_ctx, n):
    rng = np.random.default_rng(n + 11)
    _sort_case(gpu_ctx, rng.uniform(0, 1, (n, 4)).astype(np.float32))


def test_sort_short_tie_runs(gpu_ctx):
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 1, (20000, 4)).astype(np.float32)
    x[100:140] = x[7]                    # identical particles
    x[200:230, :3] = x[9, :3]            # same position, different attribute
    x[300:310, :3] = x[11, :3] + np.float32(1e-7)
    x[500:520, 0] = -0.0                 # signed zeros on an axis boundary
    x[520:540, 0] = 0.0
    _sort_case(gpu_ctx, x)


def test_sort_long_tie_run_fallback(gpu_ctx):
    rng = np.random.default_rng(6)
    x = rng.uniform(0, 1, (30000, 4)).astype(np.float32)
    x[1000:1500, :3] = x[3, :3]          # a 500-long run of equal codes
    x[1000:1500, 3] = rng.uniform(0, 1, 500)
    _sort_case(gpu_ctx, x)


@pytest.mark.parametrize("ppl,per_particle,n", [(1, False, 5000), (1, True, 5000), (3, True, 10001),
                                                (8, False, 777), (16, True, 100000)])
def test_gen_bounds(gpu_ctx, ppl, per_particle, n):
    rng = np.random.default_rng(ppl * 31 + n)
    xyzw = rng.uniform(-1, 1, (n, 4)).astype(np.float32)
    xyzw[:, 3] = rng.uniform(0, 1, n)
    rm = np.stack([np.linspace(0, 1, 64)] * 4, 1).astype(np.float32)
    rm[:6, 0] = 0.0
    radii = rng.uniform(0.01, 0.05, n).astype(np.float32) if per_particle else None
    h = gpu_ctx.tf(rm)
    got = gpu_ctx.gen_bounds(_dev(xyzw), ppl, 0.03, h, _dev(radii) if radii is not None else None)
    exp = oracle.gen_bounds(xyzw, ppl, 0.03, rm, radii)
    assert np.array_equal(got.cpu().numpy(), exp, equal_nan=True)


def test_gen_bounds_fewer_particles_than_leaves(gpu_ctx):
    # numAABBs sized for a bigger timestep (hostCode.cpp:383): extra leaves are NaN
    xyzw = np.random.default_rng(1).uniform(0, 1, (100, 4)).astype(np.float32)
    rm = np.ones((1, 4), np.float32)
    h = gpu_ctx.tf(rm)
    got = gpu_ctx.gen_bounds(_dev(xyzw), 4, 0.01, h, num_particles=62)
    exp = oracle.gen_bounds(xyzw, 4, 0.01, rm, num_particles=62)
    assert np.array_equal(got.cpu().numpy(), exp, equal_nan=True)
    assert np.isnan(exp[6 * 16:]).all()


@pytest.mark.parametrize("K", [2, 4, 8, 16])
def test_bvh_build_and_refit(gpu_ctx, K):
    rng = np.random.default_rng(K)
    n = 20000 + K
    a = rng.uniform(-1, 1, (n, 6)).astype(np.float32)
    a[:, 3:] = a[:, :3] + rng.uniform(0, 0.05, (n, 3)).astype(np.float32)
    a[rng.integers(0, n, 50)] = np.nan
    d = _dev(a.reshape(-1))
    h = gpu_ctx.bvh_build(d, K)
    lv = oracle.bvh_levels(a.reshape(-1), K)
    assert gpu_ctx.bvh_levels(h) == [len(l) for l in lv]
    for l in range(len(lv)):
        assert np.array_equal(gpu_ctx.bvh_read_level(h, l), lv[l])
    a2 = a + np.float32(0.25)
    d.copy_(_dev(a2.reshape(-1)))
    gpu_ctx.bvh_refit(h, d)
    lv2 = oracle.bvh_levels(a2.reshape(-1), K)
    for l in range(len(lv2)):
        assert np.array_equal(gpu_ctx.bvh_read_level(h, l), lv2[l])
    gpu_ctx.bvh_destroy(h)


def test_single_hip_runtime(gpu_ctx):
    """torch and the library share one HIP runtime in the process."""
    maps = open("/proc/self/maps").read()
    libs = {l.split()[-1] for l in maps.splitlines() if "libamdhip64" in l}
    assert len(libs) == 1, libs


def _dequantise(groups):
    """Child boxes [n, 8, 2, 3] of raw groups (lo + q * step in float32)."""
    f = groups[:, :8].view(np.float32)
    lo = f[:, 0:3]
    step = np.stack([f[:, 3], f[:, 4], f[:, 5]], 1)
    qb = groups[:, 8:20].copy().view(np.uint8).reshape(-1, 6, 8).astype(np.float32)
    qlo, qhi = qb[:, 0:3, :], qb[:, 3:6, :]
    dlo = (lo[:, :, None] + (qlo * step[:, :, None]).astype(np.float32)).astype(np.float32)
    dhi = (lo[:, :, None] + (qhi * step[:, :, None]).astype(np.float32)).astype(np.float32)
    empty = qb[:, 0, :] > qb[:, 3, :]
    return dlo.transpose(0, 2, 1), dhi.transpose(0, 2, 1), empty


@pytest.mark.parametrize("K", [2, 4, 8])
def test_quantised_groups_contain_children(gpu_ctx, K):
    rng = np.random.default_rng(40 + K)
    n = 30000
    a = rng.normal(0, 1, (n, 6)).astype(np.float32)
    a[:, 3:] = a[:, :3] + rng.uniform(0, 0.01, (n, 3)).astype(np.float32)
    a[rng.integers(0, n, 30)] = np.nan
    h = gpu_ctx.bvh_build(_dev(a.reshape(-1)), K)
    counts = gpu_ctx.bvh_levels(h)
    for l in range(1, len(counts)):
        child = gpu_ctx.bvh_read_level(h, l - 1)
        dlo, dhi, empty = _dequantise(gpu_ctx.bvh_read_groups(h, l))
        for j in range(counts[l]):
            for k in range(K):
                c = j * K + k
                if c >= counts[l - 1]:
                    assert empty[j, k]
                    continue
                clo, chi = child[c, :3], child[c, 4:7]
                if not (clo[0] <= chi[0]):
                    assert empty[j, k]
                    continue
                assert not empty[j, k]
                assert np.all(dlo[j, k] <= clo) and np.all(dhi[j, k] >= chi), (l, j, k)
    gpu_ctx.bvh_destroy(h)
