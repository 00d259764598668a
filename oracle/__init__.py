"""Python loader of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by the tests, the smoke check and the
benchmark's CPU-baseline leg, never by the product package.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libhilbert_ref.so")

f32p = C.POINTER(C.c_float)
u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)


class OrcParams(C.Structure):
    _fields_ = [
        ("accumID", C.c_uint32), ("frameID", C.c_uint32), ("enableBlueNoise", C.c_uint32),
        ("showNoise", C.c_uint32), ("visualizeAttributes", C.c_uint32),
        ("light_azimuth", C.c_float), ("light_elevation", C.c_float), ("light_ambient", C.c_float),
        ("cam_pos", C.c_float * 3), ("cam_d00", C.c_float * 3), ("cam_du", C.c_float * 3),
        ("cam_dv", C.c_float * 3),
        ("rbfRadius", C.c_float), ("unit", C.c_float),
        ("width", C.c_int32), ("height", C.c_int32),
        ("aabb_min", C.c_float * 3), ("aabb_max", C.c_float * 3),
        ("xyzw", f32p), ("radii", f32p), ("numParticles", C.c_uint32),
        ("colormap", f32p), ("radiusmap", f32p), ("densitymap", f32p),
        ("cm_n", C.c_uint32), ("rm_n", C.c_uint32), ("dm_n", C.c_uint32),
        ("stbn", C.POINTER(C.c_uint8)), ("stbn_w", C.c_int32), ("stbn_h", C.c_int32),
        ("max_depth", C.c_uint32),
        ("baked", f32p),
    ]


def build(ref: bool = False) -> None:
    """make -C oracle (and the reference hilbert.h harness when asked)."""
    targets = ["all"] + (["ref"] if ref else [])
    subprocess.run(["make", "-s", "-C", HERE, *targets], check=True)


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = _bind(C.CDLL(LIB))
    return _lib


@contextlib.contextmanager
def using(path: str):
    """Routes every wrapper of this module through another build of
    oracle.c (the tests build copies with a planted mistake to show that a
    check catches it), then restores the normal build."""
    global _lib
    saved = _lib
    _lib = _bind(C.CDLL(path))
    try:
        yield _lib
    finally:
        _lib = saved


def _bind(L: C.CDLL) -> C.CDLL:
    L.orc_hilbert_c2i.restype = C.c_uint64
    L.orc_hilbert_c2i.argtypes = [u64p]
    L.orc_hilbert_i2c.argtypes = [C.c_uint64, u64p]
    L.orc_hilbert_codes.argtypes = [f32p, C.c_uint64, f32p, u64p]
    L.orc_spacetime_aabb.argtypes = [f32p, C.c_uint64, f32p]
    L.orc_sort_particles.argtypes = [f32p, u64p, C.c_uint64, u32p]
    L.orc_attribute_range.argtypes = [f32p, C.c_uint64, f32p]
    L.orc_normalize_attributes.argtypes = [f32p, C.c_uint64, f32p]
    L.orc_tf_sample.argtypes = [f32p, C.c_uint32, C.c_float, f32p]
    L.orc_gen_bounds.argtypes = [f32p, C.c_uint32, C.c_uint32, C.c_float, f32p, C.c_uint32, f32p, f32p,
                                 C.c_uint32]
    L.orc_bvh_num_levels.restype = C.c_uint32
    L.orc_bvh_num_levels.argtypes = [C.c_uint32, C.c_uint32]
    L.orc_bvh_build_level.argtypes = [f32p, C.c_uint32, C.c_uint32, C.c_uint32, f32p]
    L.orc_render.restype = C.c_int
    L.orc_render.argtypes = [C.POINTER(OrcParams), C.POINTER(C.c_int32), C.c_uint64, C.c_uint32, f32p, f32p,
                             C.c_int, u64p]
    L.orc_scene_new.restype = C.c_void_p
    L.orc_scene_new.argtypes = [C.POINTER(OrcParams)]
    L.orc_scene_render.restype = C.c_int
    L.orc_scene_render.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.c_uint64, C.c_uint32, f32p, f32p,
                                   C.c_int, u64p]
    L.orc_scene_free.argtypes = [C.c_void_p]
    L.orc_point_queries.restype = C.c_int
    L.orc_point_queries.argtypes = [C.POINTER(OrcParams), f32p, C.c_uint64, f32p]
    L.orc_light_direction.argtypes = [C.c_float, C.c_float, f32p]
    L.orc_taa_resolve.argtypes = [f32p, f32p, f32p, C.c_int, C.c_int, C.c_int, u32p, u32p]
    L.orc_shadow_march.restype = C.c_int
    L.orc_shadow_march.argtypes = [C.POINTER(OrcParams), f32p, C.c_float, f32p]
    L.orc_expected.restype = C.c_int
    L.orc_expected.argtypes = [C.POINTER(OrcParams), C.POINTER(C.c_int32), C.c_uint64,
                               C.POINTER(C.c_double), C.c_int]
    L.orc_splat.restype = C.c_int
    L.orc_splat.argtypes = [C.POINTER(OrcParams), f32p, C.c_uint32, C.c_uint32, C.POINTER(C.c_int32),
                            C.c_uint64, f32p, u32p]
    L.orc_rng_stream.restype = C.c_uint32
    L.orc_rng_stream.argtypes = [C.c_uint32] * 4 + [f32p, C.c_uint32]
    return L


def ref_lib():
    """The reference's own hilbert.h compiled by `make -C oracle ref`, or None."""
    if not os.path.exists(REF_LIB):
        return None
    L = C.CDLL(REF_LIB)
    L.ref_hilbert_c2i.restype = C.c_uint64
    L.ref_hilbert_c2i.argtypes = [u64p]
    L.ref_hilbert_i2c.argtypes = [C.c_uint64, u64p]
    L.ref_hilbert_codes.argtypes = [f32p, C.c_uint64, f32p, u64p]
    return L


def _f(a):
    return a.ctypes.data_as(f32p)


def c2i(x: int, y: int, z: int) -> int:
    return int(lib().orc_hilbert_c2i((C.c_uint64 * 3)(x, y, z)))


def i2c(index: int):
    out = (C.c_uint64 * 3)()
    lib().orc_hilbert_i2c(C.c_uint64(index), out)
    return tuple(int(v) for v in out)


def hilbert_codes(xyzw: np.ndarray, aabb) -> np.ndarray:
    x = np.ascontiguousarray(xyzw, dtype=np.float32)
    a = np.asarray(aabb, dtype=np.float32)
    out = np.empty(x.shape[0], dtype=np.uint64)
    lib().orc_hilbert_codes(_f(x), x.shape[0], _f(a), out.ctypes.data_as(u64p))
    return out


def spacetime_aabb(xyzw: np.ndarray, aabb=None) -> np.ndarray:
    x = np.ascontiguousarray(xyzw, dtype=np.float32)
    a = np.array(aabb if aabb is not None else [1e38, 1e38, 1e38, -1e38, -1e38, -1e38], dtype=np.float32)
    lib().orc_spacetime_aabb(_f(x), x.shape[0], _f(a))
    return a


def sort_particles(xyzw: np.ndarray, codes: np.ndarray):
    x = np.ascontiguousarray(xyzw, dtype=np.float32).copy()
    c = np.ascontiguousarray(codes, dtype=np.uint64).copy()
    perm = np.empty(x.shape[0], dtype=np.uint32)
    lib().orc_sort_particles(_f(x), c.ctypes.data_as(u64p), x.shape[0], perm.ctypes.data_as(u32p))
    return x, c, perm


def attribute_range(xyzw: np.ndarray, rng=None) -> np.ndarray:
    x = np.ascontiguousarray(xyzw, dtype=np.float32)
    r = np.array(rng if rng is not None else [1e20, -1e20], dtype=np.float32)
    lib().orc_attribute_range(_f(x), x.shape[0], _f(r))
    return r


def normalize_attributes(xyzw: np.ndarray, rng) -> np.ndarray:
    x = np.ascontiguousarray(xyzw, dtype=np.float32).copy()
    r = np.asarray(rng, dtype=np.float32)
    lib().orc_normalize_attributes(_f(x), x.shape[0], _f(r))
    return x


def preprocess(xyzw: np.ndarray, radii=None):
    """hostCode.cpp:183-250 for one timestep: returns sorted+normalised
    particles, sorted codes, permutation, aabb, attribute range, radii."""
    aabb = spacetime_aabb(xyzw)
    codes = hilbert_codes(xyzw, aabb)
    p, c, perm = sort_particles(xyzw, codes)
    rng = attribute_range(p)
    p = normalize_attributes(p, rng)
    r = None if radii is None else np.ascontiguousarray(np.asarray(radii, np.float32)[perm])
    return p, c, perm, aabb, rng, r


def tf_sample(table: np.ndarray, u: float) -> np.ndarray:
    t = np.ascontiguousarray(table, dtype=np.float32)
    out = np.empty(4, dtype=np.float32)
    lib().orc_tf_sample(_f(t), t.shape[0], C.c_float(u), _f(out))
    return out


def gen_bounds(xyzw, ppl: int, rbf_radius: float, radiusmap, radii=None, num_aabbs=None, num_particles=None):
    x = np.ascontiguousarray(xyzw, dtype=np.float32)
    n = x.shape[0] if num_particles is None else num_particles
    na = (x.shape[0] + ppl - 1) // ppl if num_aabbs is None else num_aabbs
    rm = np.ascontiguousarray(radiusmap, dtype=np.float32)
    r = None if radii is None else np.ascontiguousarray(radii, dtype=np.float32)
    out = np.empty(na * 6, dtype=np.float32)
    lib().orc_gen_bounds(_f(x), n, ppl, C.c_float(rbf_radius), _f(rm), rm.shape[0],
                         _f(r) if r is not None else None, _f(out), na)
    return out


def bvh_levels(aabbs: np.ndarray, K: int):
    """Node boxes of every level of the Hilbert-packed BVH (level 0 = the
    widened leaf boxes), as float32 [n, 8] arrays."""
    L = lib()
    a = np.ascontiguousarray(aabbs, dtype=np.float32)
    n0 = a.size // 6
    levels = []
    # level 0: widened leaves (each leaf is its own 1-child node)
    lv = np.empty((n0, 8), dtype=np.float32)
    L.orc_bvh_build_level(_f(a), 1, n0, 1, _f(lv))
    levels.append(lv)
    n = n0
    nl = L.orc_bvh_num_levels(n0, K)
    while len(levels) < max(2, nl):
        m = (n + K - 1) // K
        nxt = np.empty((m, 8), dtype=np.float32)
        L.orc_bvh_build_level(_f(levels[-1]), 0, n, K, _f(nxt))
        levels.append(nxt)
        n = m
    return levels


def rng_stream(frame: int, px: int, py: int, width: int, k: int):
    """get_rng(frameID, pixel, width) and its first k lcg_randomf draws."""
    d = np.empty(max(1, k), np.float32)
    s = lib().orc_rng_stream(frame, px, py, width, _f(d), k)
    return int(s), d[:k]


def light_direction(azimuth: float, elevation: float) -> np.ndarray:
    out = np.empty(3, dtype=np.float32)
    lib().orc_light_direction(C.c_float(azimuth), C.c_float(elevation), _f(out))
    return out


class Render:
    """Keeps the numpy arrays referenced by an OrcParams alive."""

    def __init__(self, xyzw, radii, aabb, colormap, radiusmap, densitymap, camera, width, height, unit,
                 rbf_radius, frame_id=0, accum_id=1, blue_noise=False, stbn=None, show_noise=False,
                 visualize_attributes=True, light=(0.0, 0.0, 0.5), max_depth=0, baked=None):
        self.keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            self.keep.append(a)
            return a

        p = OrcParams()
        p.accumID, p.frameID = accum_id, frame_id
        p.enableBlueNoise, p.showNoise, p.visualizeAttributes = int(blue_noise), int(show_noise), int(
            visualize_attributes)
        p.light_azimuth, p.light_elevation, p.light_ambient = light
        for name, v in (("cam_pos", camera.pos), ("cam_d00", camera.dir_00), ("cam_du", camera.dir_du),
                        ("cam_dv", camera.dir_dv)):
            getattr(p, name)[:] = [float(np.float32(x)) for x in v]
        p.rbfRadius, p.unit = rbf_radius, unit
        p.width, p.height = width, height
        p.aabb_min[:] = [float(x) for x in aabb[:3]]
        p.aabb_max[:] = [float(x) for x in aabb[3:6]]
        x = arr(xyzw, np.float32)
        p.xyzw = _f(x)
        p.numParticles = x.shape[0]
        if radii is not None:
            p.radii = _f(arr(radii, np.float32))
        for nm, t in (("colormap", colormap), ("radiusmap", radiusmap), ("densitymap", densitymap)):
            a = arr(np.asarray(t, np.float32).reshape(-1, 4), np.float32)
            setattr(p, nm, _f(a))
            setattr(p, {"colormap": "cm_n", "radiusmap": "rm_n", "densitymap": "dm_n"}[nm], a.shape[0])
        if stbn is not None:
            s = arr(stbn, np.uint8)
            p.stbn = s.ctypes.data_as(C.POINTER(C.c_uint8))
            p.stbn_h, p.stbn_w = s.shape[0], s.shape[1]
        p.max_depth = max_depth
        if baked is not None:  # decoded payload [N, 8] used instead of baking
            p.baked = _f(arr(np.asarray(baked, np.float32).reshape(-1, 8), np.float32))
        self.p = p

    def run(self, pixels=None, spp: int = 1, threads: int = 8, image=None, accum=None):
        W, H = self.p.width, self.p.height
        image = np.zeros((H, W, 4), np.float32) if image is None else image
        accum = np.zeros((H, W, 4), np.float32) if accum is None else accum
        counts = (C.c_uint64 * 2)()
        if pixels is not None:
            px = np.ascontiguousarray(np.asarray(pixels, dtype=np.int32).reshape(-1, 2))
            st = lib().orc_render(C.byref(self.p), px.ctypes.data_as(C.POINTER(C.c_int32)), px.shape[0], spp,
                                  _f(image), _f(accum), threads, counts)
        else:
            st = lib().orc_render(C.byref(self.p), None, 0, spp, _f(image), _f(accum), threads, counts)
        if st != 0:
            raise RuntimeError("orc_render failed")
        return image, accum, (int(counts[0]), int(counts[1]))

    def prepared(self):
        """Bakes and grids the scene once; returns an object whose
        render(pixels, spp, threads) only renders (for timing)."""
        outer = self

        class _Prepared:
            def __init__(self):
                self.h = lib().orc_scene_new(C.byref(outer.p))
                if not self.h:
                    raise RuntimeError("orc_scene_new failed")

            def render(self, pixels, spp=1, threads=8):
                W, H = outer.p.width, outer.p.height
                image = np.zeros((H, W, 4), np.float32)
                accum = np.zeros((H, W, 4), np.float32)
                px = np.ascontiguousarray(np.asarray(pixels, dtype=np.int32).reshape(-1, 2))
                counts = (C.c_uint64 * 2)()
                lib().orc_scene_render(self.h, px.ctypes.data_as(C.POINTER(C.c_int32)), px.shape[0], spp,
                                       _f(image), _f(accum), threads, counts)
                return image, accum, (int(counts[0]), int(counts[1]))

            def __del__(self):
                if getattr(self, "h", None):
                    lib().orc_scene_free(self.h)
                    self.h = None

        return _Prepared()

    def expected(self, pixels=None, threads: int = 8) -> np.ndarray:
        """FP64 quadrature of the delta-tracking expectation: [H, W, 3] for
        the whole image, or [n, 3] for the given (x, y) pixels (pixels
        shared among `threads` workers; the result does not depend on it)."""
        W, H = self.p.width, self.p.height
        if pixels is not None:
            px = np.ascontiguousarray(np.asarray(pixels, dtype=np.int32).reshape(-1, 2))
            out = np.zeros((px.shape[0], 3), np.float64)
            st = lib().orc_expected(C.byref(self.p), px.ctypes.data_as(C.POINTER(C.c_int32)), px.shape[0],
                                    out.ctypes.data_as(C.POINTER(C.c_double)), threads)
        else:
            out = np.zeros((H, W, 3), np.float64)
            st = lib().orc_expected(C.byref(self.p), None, 0, out.ctypes.data_as(C.POINTER(C.c_double)), threads)
        if st != 0:
            raise RuntimeError("orc_expected failed (delta tracking only)")
        return out

    def splat(self, leaf_aabbs: np.ndarray, ppl: int = 1, pixels=None, hits: bool = False):
        """ParticleSplatRayGen over all pixels (or the given (x, y) pixels);
        leaf_aabbs = the GenRBFBounds output the BVH was built from.  With
        hits=True also returns, per rendered pixel, the largest hit count of
        one slab and of 16 consecutive slabs (uint32 [n, 2])."""
        W, H = self.p.width, self.p.height
        image = np.zeros((H, W, 4), np.float32)
        a = np.ascontiguousarray(leaf_aabbs, np.float32).reshape(-1)
        nl = a.size // 6
        n = W * H if pixels is None else len(pixels)
        hbuf = np.zeros((n, 2), np.uint32) if hits else None
        hp = hbuf.ctypes.data_as(u32p) if hits else None
        if pixels is not None:
            px = np.ascontiguousarray(np.asarray(pixels, dtype=np.int32).reshape(-1, 2))
            st = lib().orc_splat(C.byref(self.p), _f(a), nl, ppl, px.ctypes.data_as(C.POINTER(C.c_int32)),
                                 px.shape[0], _f(image), hp)
        else:
            st = lib().orc_splat(C.byref(self.p), _f(a), nl, ppl, None, 0, _f(image), hp)
        if st != 0:
            raise RuntimeError("orc_splat failed")
        return (image, hbuf) if hits else image

    def shadow_visibility(self, origin, t0: float) -> float:
        """Visibility of the marching tracker's NEE shadow ray from `origin`
        toward the light, started at t = t0."""
        o = np.ascontiguousarray(np.asarray(origin, np.float32).reshape(3))
        v = C.c_float()
        if lib().orc_shadow_march(C.byref(self.p), _f(o), C.c_float(t0), C.byref(v)) != 0:
            raise RuntimeError("orc_shadow_march failed")
        return float(v.value)

    def point_queries(self, pts: np.ndarray) -> np.ndarray:
        q = np.ascontiguousarray(pts, dtype=np.float32).reshape(-1, 3)
        out = np.empty((q.shape[0], 5), dtype=np.float32)
        if lib().orc_point_queries(C.byref(self.p), _f(q), q.shape[0], _f(out)) != 0:
            raise RuntimeError("orc_point_queries failed")
        return out


def taa_resolve(image: np.ndarray, taa: np.ndarray, taa_prev: np.ndarray, disable: bool = False, gui=None,
                frame=None):
    """CompositeGui + the history copy: updates taa and taa_prev in place;
    `gui` (uint32 RGBA8 words [H, W]) and `frame` (uint32 [H, W], receives
    the BGRA8 words) are optional."""
    H, W = image.shape[:2]
    im = np.ascontiguousarray(image, np.float32)
    for a in (taa, taa_prev):
        assert a.dtype == np.float32 and a.flags.c_contiguous and a.shape == im.shape
    g = None if gui is None else np.ascontiguousarray(gui, np.uint32).reshape(H, W)
    if frame is not None:
        assert frame.dtype == np.uint32 and frame.flags.c_contiguous and frame.size == W * H
    lib().orc_taa_resolve(_f(im), _f(taa), _f(taa_prev), W, H, int(disable),
                          g.ctypes.data_as(u32p) if g is not None else None,
                          frame.ctypes.data_as(u32p) if frame is not None else None)


# ---- differentiable renderer (adjoint.c) -----------------------------------------------

class OrcDiff(C.Structure):
    _fields_ = [
        ("pos", C.POINTER(C.c_double)), ("attr", C.POINTER(C.c_double)), ("radii", C.POINTER(C.c_double)),
        ("n", C.c_uint32),
        ("cm", C.POINTER(C.c_double)), ("dm", C.POINTER(C.c_double)), ("rm", C.POINTER(C.c_double)),
        ("cm_n", C.c_uint32), ("dm_n", C.c_uint32), ("rm_n", C.c_uint32),
        ("cam_pos", C.c_float * 3), ("cam_d00", C.c_float * 3), ("cam_du", C.c_float * 3), ("cam_dv", C.c_float * 3),
        ("width", C.c_int32), ("height", C.c_int32),
        ("spp", C.c_uint32), ("frameID", C.c_uint32), ("max_depth", C.c_uint32),
        ("unit", C.c_float), ("rbfRadius", C.c_float),
        ("aabb", C.c_float * 6),
        ("estimator", C.c_uint32),
    ]


class Diff:
    """FP64 differentiable image (estimator 0: expected stochastic marching,
    1: ratio tracking of delta tracking) and its adjoint (oracle/adjoint.c).
    pos [N,3], attr [N], radii [N], TF tables [n,4]; all float64."""

    def __init__(self, pos, attr, radii, colormap, densitymap, radiusmap, camera, width, height, unit, rbf_radius,
                 aabb, spp=1, frame_id=0, max_depth=0, estimator=0):
        self.keep = []

        def arr(a):
            a = np.ascontiguousarray(a, np.float64)
            self.keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_double))

        p = OrcDiff()
        p.pos, p.attr, p.radii = arr(pos), arr(attr), arr(radii)
        p.n = len(attr)
        cm, dm, rm = (np.asarray(t, np.float64).reshape(-1, 4) for t in (colormap, densitymap, radiusmap))
        p.cm, p.dm, p.rm = arr(cm), arr(dm), arr(rm)
        p.cm_n, p.dm_n, p.rm_n = len(cm), len(dm), len(rm)
        for name, v in (("cam_pos", camera.pos), ("cam_d00", camera.dir_00), ("cam_du", camera.dir_du),
                        ("cam_dv", camera.dir_dv)):
            getattr(p, name)[:] = [float(np.float32(x)) for x in v]
        p.width, p.height, p.spp, p.frameID, p.max_depth = width, height, spp, frame_id, max_depth
        p.unit, p.rbfRadius = unit, rbf_radius
        p.aabb[:] = [float(x) for x in aabb[:6]]
        p.estimator = estimator
        self.p = p
        self.sizes = (len(attr), len(cm), len(dm), len(rm))
        L = lib()
        L.orc_diff_forward.restype = C.c_int
        L.orc_diff_forward.argtypes = [C.POINTER(OrcDiff), C.POINTER(C.c_double)]
        L.orc_diff_adjoint.restype = C.c_int
        L.orc_diff_adjoint.argtypes = [C.POINTER(OrcDiff)] + [C.POINTER(C.c_double)] * 7
        L.orc_diff_forward_pixels.restype = C.c_int
        L.orc_diff_forward_pixels.argtypes = [C.POINTER(OrcDiff), C.POINTER(C.c_int32), C.c_uint64,
                                              C.POINTER(C.c_double)]
        L.orc_diff_adjoint_pixels.restype = C.c_int
        L.orc_diff_adjoint_pixels.argtypes = [C.POINTER(OrcDiff), C.POINTER(C.c_int32), C.c_uint64] + \
            [C.POINTER(C.c_double)] * 7

    def forward(self, pixels=None) -> np.ndarray:
        """[H, W, 3] image, or [n, 3] for the given (x, y) pixels only."""
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        if pixels is not None:
            px = np.ascontiguousarray(np.asarray(pixels, np.int32).reshape(-1, 2))
            out = np.zeros((px.shape[0], 3), np.float64)
            if lib().orc_diff_forward_pixels(C.byref(self.p), px.ctypes.data_as(C.POINTER(C.c_int32)), px.shape[0],
                                             dp(out)) != 0:
                raise RuntimeError("orc_diff_forward_pixels failed")
            return out
        img = np.zeros((self.p.height, self.p.width, 3), np.float64)
        if lib().orc_diff_forward(C.byref(self.p), dp(img)) != 0:
            raise RuntimeError("orc_diff_forward failed")
        return img

    def adjoint(self, gimage: np.ndarray, pixels=None) -> dict:
        """Gradients of sum(gimage * image); with `pixels`, gimage is [n, 3]
        for those (x, y) pixels and the other pixels do not contribute."""
        n, ncm, ndm, nrm = self.sizes
        g = np.ascontiguousarray(gimage, np.float64)
        out = {"pos": np.zeros((n, 3)), "radii": np.zeros(n), "attr": np.zeros(n), "colormap": np.zeros((ncm, 4)),
               "densitymap": np.zeros(ndm), "radiusmap": np.zeros(nrm)}
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        grads = (dp(out["pos"]), dp(out["radii"]), dp(out["attr"]), dp(out["colormap"]), dp(out["densitymap"]),
                 dp(out["radiusmap"]))
        if pixels is not None:
            px = np.ascontiguousarray(np.asarray(pixels, np.int32).reshape(-1, 2))
            assert g.shape == (px.shape[0], 3)
            st = lib().orc_diff_adjoint_pixels(C.byref(self.p), px.ctypes.data_as(C.POINTER(C.c_int32)), px.shape[0],
                                               dp(g), *grads)
        else:
            st = lib().orc_diff_adjoint(C.byref(self.p), dp(g), *grads)
        if st != 0:
            raise RuntimeError("orc_diff_adjoint failed")
        return out
