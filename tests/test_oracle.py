"""Pins the CPU oracle before it is trusted as the checker: golden vectors
from the reference's own hilbert.h, published RNG test vectors, libm
accuracy bounds, and closed forms of the estimators."""
import ctypes as C
import json
import math
import os
import struct

import numpy as np
import pytest

import oracle
from helpers import default_camera

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "hilbert.json")))


# ---- Hilbert (hilbert.h:124-225, hostCode.cpp:196-209) ------------------------

def test_hilbert_edge_cases_match_reference():
    for x, y, z, code in GOLD["edge"]:
        assert oracle.c2i(x, y, z) == code, (x, y, z)
    # the 17-bit spill of the reference's packing (hilbert.h:189-193)
    assert oracle.c2i(65536, 0, 0) == oracle.c2i(0, 1, 0) == 3
    assert oracle.c2i(0, 0, 65536) == 0


def test_hilbert_random_match_reference():
    for x, y, z, code in GOLD["random"]:
        assert oracle.c2i(x, y, z) == code


def test_hilbert_blocks_checksums():
    for lo, hi, s, x in GOLD["blocks"]:
        g = np.arange(lo, hi)
        acc, xo = 0, 0
        for a in g:
            for b in g:
                for c in g:
                    v = oracle.c2i(int(a), int(b), int(c))
                    acc = (acc + v) & 0xFFFFFFFFFFFFFFFF
                    xo ^= v
        assert (acc, xo) == (s, x)


def test_hilbert_i2c_match_reference():
    for i, x, y, z in GOLD["i2c"]:
        assert oracle.i2c(i) == (x, y, z)


def test_hilbert_particle_codes_match_reference():
    P = GOLD["particles"]
    xyzw = np.array(P["xyzw"], np.float32)
    codes = oracle.hilbert_codes(xyzw, P["aabb"])
    assert [int(c) for c in codes] == P["codes"]


def test_hilbert_invariants():
    # i2c(c2i(c)) = c and consecutive indices are unit steps
    rng = np.random.default_rng(3)
    for _ in range(2000):
        c = tuple(int(v) for v in rng.integers(0, 65536, 3))
        assert oracle.i2c(oracle.c2i(*c)) == c
    start = int(rng.integers(0, 2 ** 48 - 5000))
    prev = oracle.i2c(start)
    for i in range(start + 1, start + 4000):
        cur = oracle.i2c(i)
        assert sum(abs(a - b) for a, b in zip(prev, cur)) == 1
        prev = cur


def test_hilbert_matches_compiled_reference_when_available():
    R = oracle.ref_lib()
    if R is None:
        pytest.skip("reference harness not built here (make -C oracle ref)")
    rng = np.random.default_rng(11)
    A = C.c_uint64 * 3
    for _ in range(20000):
        c = [int(v) for v in rng.integers(0, 65537, 3)]
        assert oracle.c2i(*c) == R.ref_hilbert_c2i(A(*c))


# ---- RNG (rng.h:20-70) -----------------------------------------------------------

def _murmur3_x86_32(data: bytes, seed: int) -> int:
    L = oracle.lib()
    L.orc_murmur_mix.restype = C.c_uint32
    L.orc_murmur_fmix.restype = C.c_uint32
    L.orc_murmur_mix.argtypes = [C.c_uint32, C.c_uint32]
    L.orc_murmur_fmix.argtypes = [C.c_uint32]
    h = seed
    n = len(data) // 4
    for i in range(n):
        h = L.orc_murmur_mix(h, struct.unpack("<I", data[4 * i:4 * i + 4])[0])
    tail = data[4 * n:]
    k = 0
    for j in reversed(range(len(tail))):
        k ^= tail[j] << (8 * j)
    if tail:
        k = (k * 0xCC9E2D51) & 0xFFFFFFFF
        k = ((k << 15) | (k >> 17)) & 0xFFFFFFFF
        k = (k * 0x1B873593) & 0xFFFFFFFF
        h ^= k
    return L.orc_murmur_fmix(h ^ len(data))


@pytest.mark.parametrize("data,seed,expect", [
    (b"", 0, 0), (b"", 1, 0x514E28B7), (b"", 0xFFFFFFFF, 0x81F16F39), (b"\0\0\0\0", 0, 0x2362F9DE),
    (b"\xff\xff\xff\xff", 0, 0x76293B50), (b"\x21\x43\x65\x87", 0, 0xF55B516B),
    (b"\x21\x43\x65\x87", 0x5082EDEE, 0x2362F9DE), (b"aaaa", 0x9747B28C, 0x5A97808A),
    (b"abcd", 0x9747B28C, 0xF0478627), (b"Hello, world!", 0x9747B28C, 0x24884CBA),
])
def test_murmur3_published_vectors(data, seed, expect):
    """rng.h's mix/finalize are MurmurHash3_x86_32's block mix and fmix32."""
    assert _murmur3_x86_32(data, seed) == expect


def _stream(frame, px, py, w, k):
    L = oracle.lib()
    L.orc_rng_stream.restype = C.c_uint32
    L.orc_rng_stream.argtypes = [C.c_uint32] * 4 + [C.POINTER(C.c_float), C.c_uint32]
    d = (C.c_float * k)()
    s = L.orc_rng_stream(frame, px, py, w, d, k)
    return s, np.array(d[:], np.float32)


def test_rng_seed_and_first_draw():
    # get_rng(0, (0,0), 512x512): the value rng.h (rng.h:20-70) gives for pixel (0, 0) of a 512-wide frame
    s, d = _stream(0, 0, 0, 512, 1)
    assert s == 256347020
    assert abs(float(d[0]) - 0.145855993) < 1e-9


def test_rng_seed_composition():
    L = oracle.lib()
    L.orc_murmur_mix.restype = C.c_uint32
    L.orc_murmur_fmix.restype = C.c_uint32
    for frame, px, py, w in [(0, 5, 7, 640), (17, 639, 479, 640), (65535, 1, 2, 3)]:
        expect = L.orc_murmur_fmix(L.orc_murmur_mix(L.orc_murmur_mix(0, px + py * w), frame))
        assert _stream(frame, px, py, w, 1)[0] == expect


def test_lcg_sequence_and_float_conversion():
    # Numerical Recipes LCG from state 0
    L = oracle.lib()
    L.orc_math.argtypes = [C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                           C.c_uint64]
    states = [0]
    for _ in range(5):
        states.append((states[-1] * 1664525 + 1013904223) & 0xFFFFFFFF)
    assert states[1:4] == [1013904223, 1196435762, 3519870697]
    x = np.array(states[:5], np.uint32).view(np.float32).copy()
    out = np.empty(5, np.float32)
    L.orc_math(4, x.ctypes.data_as(C.POINTER(C.c_float)), x.ctypes.data_as(C.POINTER(C.c_float)),
               out.ctypes.data_as(C.POINTER(C.c_float)), 5)
    expect = (np.array(states[1:6], np.uint32).astype(np.float32) * np.float32(2.0 ** -32))
    assert np.array_equal(out, expect)
    # states >= 0xFFFFFF80 round to 2^32: lcg_randomf returns exactly 1.0
    s_prev = ((0xFFFFFF90 - 1013904223) * pow(1664525, -1, 2 ** 32)) & 0xFFFFFFFF
    x = np.array([s_prev], np.uint32).view(np.float32).copy()
    L.orc_math(4, x.ctypes.data_as(C.POINTER(C.c_float)), x.ctypes.data_as(C.POINTER(C.c_float)),
               out.ctypes.data_as(C.POINTER(C.c_float)), 1)
    assert out[0] == np.float32(1.0)


# ---- deterministic math vs libm -------------------------------------------------------

def _math(which, x, y=None):
    L = oracle.lib()
    L.orc_math.argtypes = [C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                           C.c_uint64]
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(x if y is None else y, np.float32)
    out = np.empty_like(x)
    fp = C.POINTER(C.c_float)
    L.orc_math(which, x.ctypes.data_as(fp), y.ctypes.data_as(fp), out.ctypes.data_as(fp), x.size)
    return out


def _ulp_err(got, ref):
    ref32 = ref.astype(np.float32)
    ulp = np.spacing(np.abs(ref32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - ref) / ulp


def test_logf_within_one_ulp():
    x = np.concatenate([np.linspace(2 ** -24, 1.0, 200001), np.geomspace(1e-30, 1e30, 100001)]).astype(np.float32)
    err = _ulp_err(_math(0, x), np.log(x.astype(np.float64)))
    assert err.max() <= 1.0
    assert _math(0, np.array([0.0], np.float32))[0] == -np.inf
    assert np.isnan(_math(0, np.array([-1.0], np.float32))[0])


def test_expf_mid_equals_expf_on_its_domain():
    L = oracle.lib()
    L.orc_expf_mid_mismatches.restype = C.c_uint64
    assert L.orc_expf_mid_mismatches() == 0


def test_log_step_within_one_ulp_on_its_domain():
    # exhaustive: all 2e8 binary32 values in [2^-24, 1] against the double
    # logarithm; ln 1 = 0 exactly, ln 0 = -inf
    L = oracle.lib()
    L.orc_log_step_max_ulp.restype = C.c_double
    at = C.c_float(0.0)
    err = L.orc_log_step_max_ulp(C.byref(at))
    assert 0.0 <= err <= 0.9, (err, at.value)
    assert _math(5, np.array([0.0, 1.0], np.float32)).tolist() == [-np.inf, 0.0]


def test_log_unit_within_one_ulp():
    # every 61st binary32 in (0, 1), subnormals included, against the double
    # logarithm; ln 0 = -inf and ln 1 = 0 exactly
    L = oracle.lib()
    L.orc_log_unit_max_ulp.restype = C.c_double
    L.orc_log_unit_max_ulp.argtypes = [C.c_uint32, C.POINTER(C.c_float)]
    at = C.c_float(0.0)
    err = L.orc_log_unit_max_ulp(61, C.byref(at))
    assert 0.0 <= err <= 1.0, (err, at.value)


def test_flight_bound_is_sound():
    # products above aa_flight_bound place the position below lim (400k
    # trials, 128 products each, near the bound and random); the same
    # trials find products above the margin-free bound that do not, so
    # they reach the boundary
    L = oracle.lib()
    L.orc_flight_bound_violations.restype = C.c_uint64
    L.orc_flight_bound_violations.argtypes = [C.c_uint64, C.c_uint32, C.c_int]
    assert L.orc_flight_bound_violations(400000, 12345, 0) == 0
    assert L.orc_flight_bound_violations(400000, 12345, 1) > 0


def test_flight_bound_special_limits():
    # lim = +inf: every step is in empty space (bound 0); a NaN lim: none is
    # (NaN fails every compare); lim at or behind the anchor, or -inf (a
    # segment end not yet gathered): the bound exceeds 1, so no product
    # passes; ahead: at least exp(-(lim - tb) / unit), and above it by at
    # most twice the margin (|c| + |lim| / unit) 2^-19
    tb = np.array([0.5, 0.5, 0.5, 0.5, 0.5, 0.5], np.float32)
    lim = np.array([np.inf, np.nan, 0.5, 0.4, -np.inf, 0.5 + 0.004], np.float32)
    b = _math(6, tb, lim)
    assert b[0] == 0.0 and np.isnan(b[1])
    assert (b[2:5] > 1.0).all(), b[2:5]
    c = (0.5 - float(lim[5])) * 2500.0
    m = (abs(c) + float(lim[5]) * 2500.0) * 2.0 ** -18
    assert np.exp(c) <= b[5] <= np.exp(c + m) < 1.0, (b[5], np.exp(c))


def test_expf_within_one_ulp():
    x = np.linspace(-87.0, 88.0, 400001).astype(np.float32)
    err = _ulp_err(_math(1, x), np.exp(x.astype(np.float64)))
    assert err.max() <= 1.0
    sp = _math(1, np.array([-np.inf, 0.0, 100.0, -110.0], np.float32))
    assert list(sp) == [0.0, 1.0, np.inf, 0.0]


def test_powf_bounded_error():
    x = np.linspace(1e-4, 1.0, 100001).astype(np.float32)
    for y in (np.float32(2.2), np.float32(1.0) / np.float32(2.2)):
        got = _math(2, x, np.full_like(x, y))
        err = _ulp_err(got, np.power(x.astype(np.float64), float(y)))
        assert err.max() <= 64.0
    assert _math(2, np.array([0.0, 1.0], np.float32), np.array([2.2, 2.2], np.float32)).tolist() == [0.0, 1.0]


def test_rbf_closed_form():
    d = np.linspace(0.0, 0.999, 1001).astype(np.float32)
    got = _math(3, d, np.ones_like(d))
    ref = np.exp(-0.5 * (3.0 * d.astype(np.float64)) ** 2)
    assert np.max(np.abs(got - ref) / ref) < 2e-6


# ---- pre-process (hostCode.cpp:183-250) ------------------------------------------------

def test_sort_order_and_ties():
    rng = np.random.default_rng(9)
    xyzw = rng.uniform(0, 1, (3000, 4)).astype(np.float32)
    xyzw[100:140] = xyzw[7]                      # exact duplicates: one code
    xyzw[200:230, :3] = xyzw[9, :3]              # same position, different w
    xyzw[300:310, 0] = xyzw[11, 0] + np.float32(1e-7)  # same cell, different x
    xyzw[300:310, 1:3] = xyzw[11, 1:3]
    aabb = oracle.spacetime_aabb(xyzw)
    codes = oracle.hilbert_codes(xyzw, aabb)
    p, c, perm = oracle.sort_particles(xyzw, codes)
    keyed = sorted(range(len(xyzw)), key=lambda i: (int(codes[i]), *xyzw[i].tolist(), i))
    assert perm.tolist() == keyed
    assert np.array_equal(p, xyzw[perm])


def test_spacetime_aabb_and_normalisation():
    rng = np.random.default_rng(1)
    steps = [rng.normal(size=(500, 4)).astype(np.float32) for _ in range(3)]
    a = None
    for s in steps:
        a = oracle.spacetime_aabb(s, a)
    allp = np.concatenate(steps)
    assert np.array_equal(a[:3], allp[:, :3].min(0)) and np.array_equal(a[3:], allp[:, :3].max(0))
    r = oracle.attribute_range(allp)
    assert r[0] == allp[:, 3].min() and r[1] == allp[:, 3].max()
    n = oracle.normalize_attributes(allp, r)
    expect = ((allp[:, 3] - r[0]) / (r[1] - r[0])).astype(np.float32)
    assert np.array_equal(n[:, 3], expect)
    assert n[:, 3].min() == 0.0 and n[:, 3].max() == 1.0


# ---- transfer functions ------------------------------------------------------------------

def test_tf_sampling_semantics():
    t = np.random.default_rng(2).uniform(0, 1, (64, 4)).astype(np.float32)
    for i in (0, 5, 63):
        assert np.array_equal(oracle.tf_sample(t, (i + 0.5) / 64), t[i])    # texel centres exact
    assert np.array_equal(oracle.tf_sample(t, -3.0), t[0])                  # clamp to edge
    assert np.array_equal(oracle.tf_sample(t, 1e30), t[63])
    assert np.array_equal(oracle.tf_sample(t, float("nan")), t[0])
    mid = oracle.tf_sample(t, 10.0 / 64)                                     # halfway 9..10
    assert np.allclose(mid, 0.5 * (t[9] + t[10]), atol=1e-6)
    const = np.tile(np.float32([0.3, 0.2, 0.1, 0.7]), (16, 1))
    for u in np.linspace(-1, 2, 301):
        assert np.array_equal(oracle.tf_sample(const, float(u)), const[0])


# ---- bounds (deviceCodeBounds.hlsl:29-82) ------------------------------------------------

@pytest.mark.parametrize("ppl", [1, 3, 8])
def test_gen_bounds_closed_form(ppl):
    rng = np.random.default_rng(ppl)
    n = 1001
    xyzw = rng.uniform(-1, 1, (n, 4)).astype(np.float32)
    xyzw[:, 3] = rng.uniform(0, 1, n)
    rm = np.stack([np.linspace(0, 1, 64)] * 4, 1).astype(np.float32)
    rm[:8, 0] = 0.0  # attributes below ~0.125 get radius 0 and are skipped
    radii = rng.uniform(0.01, 0.05, n).astype(np.float32)
    got = oracle.gen_bounds(xyzw, ppl, 0.0, rm, radii).reshape(-1, 6)
    for leaf in range(got.shape[0]):
        lo = np.full(3, np.inf, np.float32)
        hi = np.full(3, -np.inf, np.float32)
        for i in range(leaf * ppl, min(n, leaf * ppl + ppl)):
            rp = oracle.tf_sample(rm, float(xyzw[i, 3]))[0]
            if rp == 0:
                continue
            e = np.float32(radii[i] * rp)
            lo = np.minimum(lo, xyzw[i, :3] - e)
            hi = np.maximum(hi, xyzw[i, :3] + e)
        if np.any(lo > hi):
            assert np.all(np.isnan(got[leaf]))
        else:
            assert np.array_equal(got[leaf, :3], lo) and np.array_equal(got[leaf, 3:], hi)


def test_bvh_levels_contain_children():
    rng = np.random.default_rng(4)
    a = rng.uniform(-1, 1, (1000, 6)).astype(np.float32)
    a[:, 3:] = a[:, :3] + rng.uniform(0, 0.1, (1000, 3)).astype(np.float32)
    a[17] = np.nan
    lv = oracle.bvh_levels(a.reshape(-1), 4)
    assert [len(l) for l in lv] == [1000, 250, 63, 16, 4, 1]
    assert np.all(lv[0][17, :3] == np.inf)
    ok = np.arange(1000) != 17
    assert np.all(lv[0][ok, :3] < a[ok, :3]) and np.all(lv[0][ok, 4:7] > a[ok, 3:])
    for l in range(1, len(lv)):
        for j in range(len(lv[l])):
            ch = lv[l - 1][4 * j:4 * j + 4]
            assert np.array_equal(lv[l][j, :3], ch[:, :3].min(0))
            assert np.array_equal(lv[l][j, 4:7], ch[:, 4:7].max(0))


# ---- estimator closed forms (deviceCodeRBF.hlsl) ------------------------------------------

def _one_particle(W, H, d0=None, unit=0.01, R=0.5, spp=32, ambient=1.0, ramp=False):
    xyzw = np.array([[0.0, 0.0, 0.0, 0.5]], np.float32)
    radii = np.array([R], np.float32)
    cm = np.tile(np.float32([1, 1, 1, 1]), (4, 1))
    rm = np.ones((1, 4), np.float32)
    dm = (np.stack([(np.arange(256) + 0.5) / 256] * 4, 1).astype(np.float32) if ramp
          else np.full((1, 4), d0, np.float32))
    cam = default_camera(W, H, fovy=25.0)
    aabb = np.zeros(6, np.float32)
    r = oracle.Render(xyzw, radii, aabb, cm, rm, dm, cam, W, H, unit, R, light=(0.0, 0.0, ambient))
    img, acc, counts = r.run(spp=spp, threads=8)
    return img, cam, counts


def _chords(cam, W, H, R):
    """(impact parameter, chord length) of each pixel ray (float64)."""
    py, px = np.mgrid[0:H, 0:W]
    sx = (px + 0.5) / W
    sy = (py + 0.5) / H
    d = (cam.dir_00[None, None, :] + sx[..., None] * cam.dir_du + sy[..., None] * cam.dir_dv).astype(np.float64)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    o = cam.pos.astype(np.float64)
    tc = -(d @ o)
    b2 = np.maximum(0.0, (o * o).sum() - tc * tc)
    chord = 2.0 * np.sqrt(np.maximum(0.0, R * R - b2))
    return np.sqrt(b2), chord


def test_delta_tracking_homogeneous_sphere_escape_probability():
    """Constant densitymap d0 inside one sphere: P(hit) = 1 - exp(-d0 L / unit)."""
    W = H = 32
    d0, unit, R, spp = 0.25, 0.02, 0.5, 64
    img, cam, _ = _one_particle(W, H, d0=d0, unit=unit, R=R, spp=spp)
    _, L = _chords(cam, W, H, R)
    p = 1.0 - np.exp(-d0 * L / unit)
    mean = img[..., 0].astype(np.float64)  # ambient 1: pixel = hit fraction
    z = (mean - p).sum() * spp / math.sqrt((p * (1 - p)).sum() * spp + 1e-12)
    assert abs(z) < 5.0
    assert np.all(img[L == 0, 0] == 0.0)   # misses stay black
    assert np.all(img[..., 3] == 1.0)


def test_delta_tracking_gaussian_particle_optical_depth():
    """Ramp densitymap (D(rho) = rho): optical depth of the truncated
    Gaussian exp(-4.5 d^2/R^2) along a chord is closed form (erf)."""
    W = H = 32
    unit, R, spp = 0.05, 0.5, 64
    img, cam, _ = _one_particle(W, H, unit=unit, R=R, spp=spp, ramp=True)
    b, L = _chords(cam, W, H, R)
    k = math.sqrt(4.5) / R
    erf = np.vectorize(math.erf)
    tau = np.exp(-4.5 * b * b / (R * R)) * math.sqrt(math.pi) / k * erf(k * L / 2) / unit
    p = 1.0 - np.exp(-tau)
    mean = img[..., 0].astype(np.float64)
    z = (mean - p).sum() * spp / math.sqrt((p * (1 - p)).sum() * spp + 1e-12)
    assert abs(z) < 5.0


def test_empty_scene_is_black_and_exhaustion_counted():
    W, H = 16, 8
    xyzw = np.array([[0.0, 0.0, 0.0, 0.5]], np.float32)
    cm = np.zeros((4, 4), np.float32)  # alpha 0: every particle is transparent
    cam = default_camera(W, H)
    r = oracle.Render(xyzw, np.array([0.5], np.float32), np.zeros(6, np.float32), cm, np.ones((1, 4), np.float32),
                      np.ones((1, 4), np.float32), cam, W, H, 0.01, 0.5)
    img, _, (steps, exhausted) = r.run(spp=2, threads=2)
    assert np.all(img == np.float32([0, 0, 0, 1]))
    # with a tiny unit every path runs out of its 10,000-step budget
    r2 = oracle.Render(xyzw, np.array([0.5], np.float32), np.zeros(6, np.float32), cm, np.ones((1, 4), np.float32),
                       np.ones((1, 4), np.float32), cam, W, H, 1e-6, 0.5, max_depth=500)
    _, _, (steps2, exhausted2) = r2.run(spp=1, threads=2)
    assert exhausted2 > 0 and steps2 >= 500 * exhausted2


def test_taa_fixed_point_and_clamp():
    rng = np.random.default_rng(0)
    img = np.tile(np.float32([0.2, 0.4, 0.6, 1.0]), (6, 5, 1))
    prev = img.copy()
    taa = np.zeros_like(img)
    oracle.taa_resolve(img, taa, prev)
    assert np.allclose(taa, img, atol=1e-7) and np.array_equal(prev, taa)
    # history far outside the neighbourhood is clamped to it before blending
    img = rng.uniform(0.4, 0.6, (6, 5, 4)).astype(np.float32)
    prev = np.full_like(img, 10.0)
    taa = np.zeros_like(img)
    oracle.taa_resolve(img, taa, prev)
    assert taa[..., :3].max() < 0.7


def test_detmath_coefficients_are_the_committed_fits():
    """The polynomial coefficients in include/aarbf_detmath.h are exactly
    what tests/golden/fit_detmath.py computes."""
    import importlib.util
    import os
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("fit_detmath", os.path.join(root, "tests", "golden", "fit_detmath.py"))
    fit = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(fit)
    text = open(os.path.join(root, "include", "aarbf_detmath.h")).read()

    def horner(fn):  # the literals of the Horner chain in the function's body
        body = text[text.index(f"AA_FN float {fn}("):]
        body = body[:body.index("\n}\n")]
        start = re.search(r"float (?:q|p) = (-?[0-9.e-]+)f;", body)
        rest = re.findall(r"AA_FMAF\((?:q|p), (?:f|r), (-?[0-9.e-]+)f\)", body)
        return np.array([float(start.group(1))] + [float(v) for v in rest], np.float32)

    log_q, exp_p = fit.log_coefficients()[::-1], fit.exp_coefficients()[::-1]
    for fn in ("aa_logf", "aa_log_step_nz"):
        assert np.array_equal(horner(fn), log_q), fn
    for fn in ("aa_expf", "aa_expf_mid"):
        assert np.array_equal(horner(fn), exp_p), fn
