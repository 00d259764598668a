"""aarbf -- Python host side of the MI355X attribute-aware RBF renderer.

A thin layer over the C ABI in include/aarbf.h (libaarbf.so).  It mirrors
the reference's host program (hostCode.cpp): pre-process particles
(space-time AABB, Hilbert sort, attribute normalisation), generate leaf
bounds, build the BVH, render with ParticleRBFRayGen, resolve TAA.  Device
memory is held in torch tensors on the HIP device; all compute runs in the
library's HIP kernels -- there is no host fallback.
"""
from __future__ import annotations

import ctypes as C
import functools
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import (BoundsConstants, Float3, Int2, PushConstants, RayGenData, RenderStats,
                   TAAConstants)

MAX_DEPTH = 10000  # sharedCode.h:27

__all__ = [
    "AarbfError", "Context", "Scene", "Camera", "camera_basis", "light_direction", "CameraController",
    "PushConstants", "RayGenData", "BoundsConstants", "TAAConstants", "RenderStats", "MAX_DEPTH", "load_library",
]


class AarbfError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status


def load_library():
    return _lib.load()


def _f3(v) -> Float3:
    return Float3(float(v[0]), float(v[1]), float(v[2]))


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _torch():
    import torch
    return torch


@dataclass
class Camera:
    """PushConstants.camera (sharedCode.h:50-55)."""
    pos: np.ndarray
    dir_00: np.ndarray
    dir_du: np.ndarray
    dir_dv: np.ndarray


def _camera_of(pc: PushConstants) -> Camera:
    c = pc.camera
    v = lambda f: np.array([f.x, f.y, f.z], dtype=np.float32)
    return Camera(v(c.pos), v(c.dir_00), v(c.dir_du), v(c.dir_dv))


def camera_basis(look_from, look_at, look_up, cos_fovy: float, width: int, height: int) -> Camera:
    """Camera basis of hostCode.cpp:669-682 (aarbf_camera_basis, binary32).

    cos_fovy is the reference's image-plane scale (its 10th --camera value);
    for a true vertical field of view f it is 2*tan(f/2).
    """
    lib = _lib.load()
    f = lambda v: np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(3))
    a, b, c = f(look_from), f(look_at), f(look_up)
    pc = PushConstants()
    st = lib.aarbf_camera_basis(_fptr(a), _fptr(b), _fptr(c), float(cos_fovy), int(width), int(height),
                                C.byref(pc))
    if st != 0:
        raise AarbfError(st, "aarbf_camera_basis: invalid arguments")
    return _camera_of(pc)


def light_direction(azimuth: float, elevation: float) -> np.ndarray:
    """getLightDirection (sharedCode.h:127-132), as the renderer evaluates it."""
    out = np.empty(3, dtype=np.float32)
    st = _lib.load().aarbf_light_direction(float(azimuth), float(elevation), _fptr(out))
    if st != 0:
        raise AarbfError(st, "aarbf_light_direction")
    return out


class CameraController:
    """The reference's interactive camera (hostCode.cpp:593-731) over
    aarbf_camera_update: call update() once per frame with the cursor
    movement and the buttons / keys held."""

    def __init__(self, look_from, look_at, look_up, cos_fovy: float, width: int, height: int,
                 diagonal: float = 1.0):
        self.state = _lib.CameraState(_f3(look_from), _f3(look_at), _f3(look_up), float(cos_fovy), 1)
        self.width, self.height, self.diagonal = int(width), int(height), float(diagonal)
        self.pc = PushConstants()

    def update(self, dx: float = 0.0, dy: float = 0.0, left=False, right=False, middle=False, key_x=False,
               key_y=False, key_z=False, left_shift=False) -> bool:
        """One frame of input; returns True when the camera moved (accumID reset to 1)."""
        inp = _lib.CameraInput(float(dx), float(dy), int(bool(left)), int(bool(right)), int(bool(middle)),
                               int(bool(key_x)), int(bool(key_y)), int(bool(key_z)), int(bool(left_shift)))
        changed = C.c_int()
        st = _lib.load().aarbf_camera_update(C.byref(self.state), C.byref(inp), self.width, self.height,
                                             self.diagonal, C.byref(self.pc), C.byref(changed))
        if st != 0:
            raise AarbfError(st, "aarbf_camera_update: invalid arguments")
        return bool(changed.value)

    @property
    def camera(self) -> Camera:
        return _camera_of(self.pc)

    @property
    def look(self):
        s = self.state
        v = lambda f: np.array([f.x, f.y, f.z], dtype=np.float32)
        return v(s.lookFrom), v(s.lookAt), v(s.lookUp)


def _ordered(fn):
    """Orders a call that reads or writes torch tensors with torch's own
    work.  On the context's own stream torch's pending work finishes before
    the library starts and the library's before the call returns; a context
    given a torch stream (set_stream) is ordered by that stream instead."""
    @functools.wraps(fn)
    def wrap(self, *a, **k):
        if self._own_stream:
            _torch().cuda.current_stream(self.device).synchronize()
        r = fn(self, *a, **k)
        if self._own_stream:
            self.sync()
        return r
    return wrap


class Context:
    """One aarbf_context: a device, a HIP stream and the handles made on it."""

    def __init__(self, device: int = 0, stream=None):
        self.lib = _lib.load()
        h = C.c_void_p()
        st = self.lib.aarbf_context_create(device, C.byref(h))
        if st != 0:
            raise AarbfError(st, "aarbf_context_create: " + self.lib.aarbf_status_string(st).decode())
        self.h = h
        self.device = device
        self._own_stream = True
        if stream is not None:
            self.set_stream(stream)

    # -- plumbing ---------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.aarbf_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, st: int, what: str = ""):
        if st != 0:
            msg = self.lib.aarbf_last_error(self.h).decode() if self.h else ""
            raise AarbfError(st, f"{what}: {self.lib.aarbf_status_string(st).decode()}: {msg}")

    def set_stream(self, stream):
        """Use a torch.cuda.Stream (or a raw hipStream_t int) for all work."""
        ptr = stream.cuda_stream if hasattr(stream, "cuda_stream") else stream
        if not ptr:
            raise ValueError("pass a non-default stream (torch.cuda.Stream()); 0 selects the context's own stream")
        self.check(self.lib.aarbf_set_stream(self.h, C.c_void_p(ptr)), "aarbf_set_stream")
        self._own_stream = False

    def sync(self):
        self.check(self.lib.aarbf_sync(self.h), "aarbf_sync")

    def handle_counts(self) -> dict:
        """Live transfer-function, blue-noise and BVH handles on the context."""
        c = (C.c_uint32 * 3)()
        self.check(self.lib.aarbf_handle_counts(self.h, c), "aarbf_handle_counts")
        return {"tf": int(c[0]), "stbn": int(c[1]), "bvh": int(c[2])}

    def set_render_mode(self, mode: int):
        """0: persistent segment-gather kernel (default); 1: per-pixel point queries."""
        self.check(self.lib.aarbf_set_render_mode(self.h, int(mode)), "aarbf_set_render_mode")

    TUNE_STEP_WAVES = 1
    TUNE_STEP_CAP = 2
    TUNE_GRAPH_ROUNDS = 3
    TUNE_WALK_BUDGET = 4
    TUNE_DIFF_WAVES = 5
    TUNE_STEP_MIN_LIVE = 6

    def set_tuning(self, param: int, value: int):
        """Schedule parameters of the default renderer (images do not change)."""
        self.check(self.lib.aarbf_set_tuning(self.h, int(param), int(value)), "aarbf_set_tuning")

    PAYLOAD_FP32, PAYLOAD_FP16 = 0, 1

    def set_payload(self, fmt: int):
        """Particle payload format of the renders (PAYLOAD_FP32 / PAYLOAD_FP16)."""
        self.check(self.lib.aarbf_set_payload(self.h, int(fmt)), "aarbf_set_payload")

    @_ordered
    def payload_read(self, n: int):
        """Decoded payload of the last render: float32 [n, 8] (x, y, z, r, a^3, rgb^2.2)."""
        torch = _torch()
        out = torch.empty((n, 8), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        self.check(self.lib.aarbf_payload_read(self.h, out.data_ptr(), n), "aarbf_payload_read")
        self.sync()
        return out.cpu().numpy()

    def set_stats(self, enable: bool):
        self.check(self.lib.aarbf_set_stats(self.h, 1 if enable else 0), "aarbf_set_stats")

    def read_stats(self, reset: bool = True) -> dict:
        s = RenderStats()
        self.check(self.lib.aarbf_render_stats_read(self.h, C.byref(s), 1 if reset else 0), "stats")
        return {n: int(getattr(s, n)) for n, _ in RenderStats._fields_}

    def read_histogram(self, reset: bool = True) -> list:
        """Gather walk-cost histogram: entry b = gathers testing [2^b, 2^(b+1)) nodes."""
        h = (C.c_uint64 * 24)()
        self.check(self.lib.aarbf_render_histogram(self.h, h, 1 if reset else 0), "aarbf_render_histogram")
        return [int(v) for v in h]

    TIMING_KERNELS = ("gather", "step", "dense_step", "setup_fold")

    def set_timing(self, enable: bool):
        """Bracket the render rounds' kernels with HIP events (see read_timing)."""
        self.check(self.lib.aarbf_set_timing(self.h, 1 if enable else 0), "aarbf_set_timing")

    def read_timing(self, reset: bool = True) -> dict:
        """{kernel: (total ms, launches)} since the last reset."""
        ms = (C.c_double * 4)()
        n = (C.c_uint64 * 4)()
        self.check(self.lib.aarbf_read_timing(self.h, ms, n, 1 if reset else 0), "aarbf_read_timing")
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(self.TIMING_KERNELS)}

    # -- pre-process (hostCode.cpp:183-250) ---------------------------------
    @_ordered
    def spacetime_aabb(self, particles, aabb=None) -> np.ndarray:
        a = np.array(aabb if aabb is not None else [1e38, 1e38, 1e38, -1e38, -1e38, -1e38], dtype=np.float32)
        self.check(self.lib.aarbf_spacetime_aabb(self.h, particles.data_ptr(), particles.shape[0], _fptr(a)),
                   "aarbf_spacetime_aabb")
        return a

    @_ordered
    def hilbert_codes(self, particles, aabb):
        torch = _torch()
        a = np.asarray(aabb, dtype=np.float32)
        codes = torch.empty(particles.shape[0], dtype=torch.int64, device=particles.device)
        self.check(self.lib.aarbf_hilbert_codes(self.h, particles.data_ptr(), particles.shape[0], _fptr(a),
                                                codes.data_ptr()), "aarbf_hilbert_codes")
        return codes

    @_ordered
    def sort_particles(self, particles, codes, want_perm: bool = True):
        torch = _torch()
        n = particles.shape[0]
        perm = torch.empty(n, dtype=torch.int32, device=particles.device) if want_perm else None
        self.check(self.lib.aarbf_sort_particles(self.h, particles.data_ptr(), codes.data_ptr(), n,
                                                 perm.data_ptr() if perm is not None else 0),
                   "aarbf_sort_particles")
        return perm

    @_ordered
    def attribute_range(self, particles, rng=None) -> np.ndarray:
        r = np.array(rng if rng is not None else [1e20, -1e20], dtype=np.float32)
        self.check(self.lib.aarbf_attribute_range(self.h, particles.data_ptr(), particles.shape[0], _fptr(r)),
                   "aarbf_attribute_range")
        return r

    @_ordered
    def normalize_attributes(self, particles, rng):
        r = np.asarray(rng, dtype=np.float32)
        self.check(self.lib.aarbf_normalize_attributes(self.h, particles.data_ptr(), particles.shape[0],
                                                       _fptr(r)), "aarbf_normalize_attributes")

    # -- tables ----------------------------------------------------------------
    def tf(self, rgba) -> int:
        t = np.ascontiguousarray(np.asarray(rgba, dtype=np.float32).reshape(-1, 4))
        h = C.c_uint64()
        self.check(self.lib.aarbf_tf_create(self.h, _fptr(t), t.shape[0], C.byref(h)), "aarbf_tf_create")
        return h.value

    def tf_destroy(self, handle: int):
        self.check(self.lib.aarbf_tf_destroy(self.h, handle), "aarbf_tf_destroy")

    def tf_rgba8(self, texels, srgb: bool = True) -> int:
        t = np.ascontiguousarray(np.asarray(texels, dtype=np.uint8).reshape(-1, 4))
        h = C.c_uint64()
        self.check(self.lib.aarbf_tf_create_rgba8(self.h, t.ctypes.data_as(C.POINTER(C.c_uint8)), t.shape[0],
                                                  1 if srgb else 0, C.byref(h)), "aarbf_tf_create_rgba8")
        return h.value

    def tf_from_marks(self, marks, kind: int) -> int:
        """Table from a gradient of (position, rgba) marks, as the reference
        builds it after a widget edit (aarbf_tf_create_from_marks); kind 0
        colormap, 1 radius / density map."""
        arr = (_lib.TFMark * len(marks))()
        for k, (pos, rgba) in enumerate(marks):
            arr[k].position = float(pos)
            arr[k].rgba[:] = [float(v) for v in rgba]
        h = C.c_uint64()
        self.check(self.lib.aarbf_tf_create_from_marks(self.h, arr, len(marks), int(kind), C.byref(h)),
                   "aarbf_tf_create_from_marks")
        return h.value

    def stbn(self, atlas) -> int:
        a = np.ascontiguousarray(np.asarray(atlas, dtype=np.uint8))
        assert a.ndim == 3 and a.shape[2] == 4
        h = C.c_uint64()
        self.check(self.lib.aarbf_stbn_create(self.h, a.ctypes.data_as(C.POINTER(C.c_uint8)), a.shape[1],
                                              a.shape[0], C.byref(h)), "aarbf_stbn_create")
        return h.value

    # -- bounds + BVH --------------------------------------------------------------
    @_ordered
    def gen_bounds(self, particles, ppl: int, rbf_radius: float, radiusmap: int, radii=None, aabbs=None,
                   num_particles: int | None = None):
        torch = _torch()
        n = particles.shape[0] if num_particles is None else num_particles
        num_aabbs = (particles.shape[0] + ppl - 1) // ppl
        if aabbs is None:
            aabbs = torch.empty(num_aabbs * 6, dtype=torch.float32, device=particles.device)
        bc = BoundsConstants(aabbs.data_ptr(), num_aabbs, particles.data_ptr(), n, ppl, float(rbf_radius),
                             radiusmap, radii.data_ptr() if radii is not None else 0)
        self.check(self.lib.aarbf_gen_bounds(self.h, C.byref(bc)), "aarbf_gen_bounds")
        return aabbs

    @_ordered
    def bvh_build(self, aabbs, branching: int = 8) -> int:
        h = C.c_uint64()
        self.check(self.lib.aarbf_bvh_build(self.h, aabbs.data_ptr(), aabbs.numel() // 6, branching,
                                            C.byref(h)), "aarbf_bvh_build")
        return h.value

    @_ordered
    def bvh_refit(self, bvh: int, aabbs):
        self.check(self.lib.aarbf_bvh_refit(self.h, bvh, aabbs.data_ptr()), "aarbf_bvh_refit")

    def bvh_destroy(self, bvh: int):
        self.check(self.lib.aarbf_bvh_destroy(self.h, bvh), "aarbf_bvh_destroy")

    def bvh_levels(self, bvh: int) -> list[int]:
        n = C.c_uint32()
        counts = (C.c_uint32 * 32)()
        self.check(self.lib.aarbf_bvh_levels(self.h, bvh, C.byref(n), counts, 32), "aarbf_bvh_levels")
        return list(counts[: n.value])

    def bvh_read_level(self, bvh: int, level: int) -> np.ndarray:
        counts = self.bvh_levels(bvh)
        out = np.empty((counts[level], 8), dtype=np.float32)
        self.check(self.lib.aarbf_bvh_read_level(self.h, bvh, level, _fptr(out)), "aarbf_bvh_read_level")
        return out

    def bvh_read_groups(self, bvh: int, level: int) -> np.ndarray:
        """Raw quantised groups of `level` as uint32 [n, 20]."""
        counts = self.bvh_levels(bvh)
        out = np.empty((counts[level], 20), dtype=np.uint32)
        self.check(self.lib.aarbf_bvh_read_groups(self.h, bvh, level, out.ctypes.data_as(C.POINTER(C.c_uint32))),
                   "aarbf_bvh_read_groups")
        return out

    # -- render --------------------------------------------------------------------
    @_ordered
    def render(self, rg: RayGenData, pc: PushConstants, width: int, height: int, spp: int = 1):
        self.check(self.lib.aarbf_render_rbf_spp(self.h, C.byref(rg), C.byref(pc), width, height, spp),
                   "aarbf_render_rbf_spp")

    NOALLOC_MAX_PARTICLES = 1048576

    @_ordered
    def render_noalloc(self, rg: RayGenData, pc: PushConstants, width: int, height: int, spp: int = 1):
        """aarbf_render_rbf_spp_noalloc: the same image as render(), no device
        allocation (one-lane-per-pixel schedule, payload in a static per-device
        array of NOALLOC_MAX_PARTICLES particles)."""
        self.check(self.lib.aarbf_render_rbf_spp_noalloc(self.h, C.byref(rg), C.byref(pc), width, height, spp),
                   "aarbf_render_rbf_spp_noalloc")

    @_ordered
    def render_batch(self, requests):
        """Renders views of one scene in one schedule (aarbf_render_batch);
        `requests`: sequence of (rg, pc, width, height, spp)."""
        self.check(self.lib.aarbf_render_batch(self.h, self._requests(requests), len(requests)),
                   "aarbf_render_batch")

    @staticmethod
    def _requests(requests):
        arr = (_lib.Request * len(requests))()
        for i, (rg, pc, w, h, spp) in enumerate(requests):
            arr[i].rg = rg
            arr[i].pc = pc
            arr[i].width, arr[i].height, arr[i].spp = int(w), int(h), int(spp)
        return arr

    @_ordered
    def autotune(self, requests, reps: int = 1) -> dict:
        """Schedule autotuner (aarbf_autotune) over the views `requests`;
        keeps the fastest setting in the context and returns it."""
        res = _lib.TuneResult()
        self.check(self.lib.aarbf_autotune(self.h, self._requests(requests), len(requests), reps, C.byref(res)),
                   "aarbf_autotune")
        return {"step_cap": res.step_cap, "step_waves": res.step_waves, "graph_rounds": res.graph_rounds,
                "walk_budget": res.walk_budget, "step_min_live": res.step_min_live, "candidates": res.candidates,
                "baseline_ms": res.baseline_ms,
                "best_ms": res.best_ms}

    def batch_latency(self, n: int) -> np.ndarray:
        """Per view of the last batch: ms from its start to the view's last sample."""
        out = np.empty(n, dtype=np.float64)
        self.check(self.lib.aarbf_batch_latency(self.h, out.ctypes.data_as(C.POINTER(C.c_double)), n),
                   "aarbf_batch_latency")
        return out

    DIFF_MARCHING = 0
    DIFF_RATIO_TRACKING = 1

    def set_diff_estimator(self, estimator: int):
        """0: expected stochastic marching; 1: ratio tracking of delta tracking."""
        self.check(self.lib.aarbf_set_diff_estimator(self.h, estimator), "aarbf_set_diff_estimator")

    @_ordered
    def diff_forward(self, rg: RayGenData, pc: PushConstants, width: int, height: int, spp: int, image, path=None):
        """Expected stochastic-marching image into `image` (float32 [H, W, 4]);
        `path` (float32 [H, W, spp, 4]) optionally keeps what the backward needs."""
        self.check(self.lib.aarbf_diff_forward(self.h, C.byref(rg), C.byref(pc), width, height, spp,
                                               image.data_ptr(), path.data_ptr() if path is not None else 0),
                   "aarbf_diff_forward")

    @_ordered
    def diff_backward(self, rg: RayGenData, pc: PushConstants, width: int, height: int, spp: int, grad_image,
                      g_pos, g_radii, g_attr, g_colormap, g_densitymap, g_radiusmap, path=None):
        """Adds the adjoint of diff_forward into the (zeroed) gradient tensors."""
        self.check(self.lib.aarbf_diff_backward(self.h, C.byref(rg), C.byref(pc), width, height, spp,
                                                path.data_ptr() if path is not None else 0,
                                                grad_image.data_ptr(), g_pos.data_ptr(), g_radii.data_ptr(),
                                                g_attr.data_ptr(), g_colormap.data_ptr(), g_densitymap.data_ptr(),
                                                g_radiusmap.data_ptr()), "aarbf_diff_backward")

    @_ordered
    def render_splat(self, rg: RayGenData, pc: PushConstants, width: int, height: int):
        self.check(self.lib.aarbf_render_splat(self.h, C.byref(rg), C.byref(pc), width, height),
                   "aarbf_render_splat")

    @_ordered
    def taa_resolve(self, tc: TAAConstants):
        self.check(self.lib.aarbf_taa_resolve(self.h, C.byref(tc)), "aarbf_taa_resolve")


@dataclass
class Scene:
    """Particles prepared the way hostCode.cpp prepares one timestep, held on
    the device: Hilbert-sorted, attributes normalised, leaf bounds and BVH
    built.  `particles` is float32 [N, 4]; `radii` optional float32 [N]
    (permuted with the particles)."""
    ctx: Context
    particles: object
    radii: object
    aabb: np.ndarray
    colormap: int
    radiusmap: int
    densitymap: int
    rbf_radius: float
    ppl: int = 1
    branching: int = 8
    aabbs: object = None
    bvh: int = 0
    perm: object = None
    attr_range: np.ndarray = field(default_factory=lambda: np.array([0, 1], np.float32))

    @staticmethod
    def prepare(ctx: Context, xyzw: np.ndarray, colormap, radiusmap, densitymap, rbf_radius: float,
                radii: np.ndarray | None = None, ppl: int = 1, branching: int = 8, normalize: bool = True,
                device: str = "cuda"):
        torch = _torch()
        dev = torch.device(device, ctx.device) if device == "cuda" else torch.device(device)
        # numpy arrays or tensors (a tensor already on the device is copied there)
        def dev_f32(a, shape):
            if isinstance(a, torch.Tensor):
                return a.detach().to(device=dev, dtype=torch.float32).reshape(shape).contiguous().clone()
            return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32).reshape(shape)).to(dev)

        p = dev_f32(xyzw, (-1, 4))
        cm = colormap if isinstance(colormap, int) else ctx.tf(colormap)
        rm = radiusmap if isinstance(radiusmap, int) else ctx.tf(radiusmap)
        dm = densitymap if isinstance(densitymap, int) else ctx.tf(densitymap)
        aabb = ctx.spacetime_aabb(p)
        codes = ctx.hilbert_codes(p, aabb)
        perm = ctx.sort_particles(p, codes)
        r = None
        if radii is not None:
            r = dev_f32(radii, (-1,))
            r = r[perm.long()].contiguous()
        rng = ctx.attribute_range(p)
        if normalize:
            ctx.normalize_attributes(p, rng)
        s = Scene(ctx, p, r, aabb, cm, rm, dm, float(rbf_radius), ppl, branching, perm=perm, attr_range=rng)
        s.rebuild()
        return s

    def rebuild(self):
        """Bounds + BVH build (hostCode.cpp:770-789)."""
        self.aabbs = self.ctx.gen_bounds(self.particles, self.ppl, self.rbf_radius, self.radiusmap, self.radii)
        if self.bvh:
            self.ctx.bvh_destroy(self.bvh)
        self.bvh = self.ctx.bvh_build(self.aabbs, self.branching)

    def refit(self):
        """Bounds + BVH refit (hostCode.cpp:770-779,790-794)."""
        self.aabbs = self.ctx.gen_bounds(self.particles, self.ppl, self.rbf_radius, self.radiusmap, self.radii,
                                         aabbs=self.aabbs)
        self.ctx.bvh_refit(self.bvh, self.aabbs)

    @property
    def n(self) -> int:
        return int(self.particles.shape[0])

    def push_constants(self, camera: Camera, unit: float, frame_id: int = 0, accum_id: int = 1,
                       blue_noise: bool = False, show_noise: bool = False, visualize_attributes: bool = True,
                       light=(0.0, 0.0, 0.5)) -> PushConstants:
        pc = PushConstants()
        pc.accumID = accum_id & 0xFFFF
        pc.frameID = frame_id & 0xFFFF
        pc.enableBlueNoise = 1 if blue_noise else 0
        pc.showNoise = 1 if show_noise else 0
        pc.visualizeAttributes = 1 if visualize_attributes else 0
        pc.particlesPerLeaf = self.ppl
        pc.light.azimuth, pc.light.elevation, pc.light.ambient = (float(v) for v in light)
        pc.camera.pos = _f3(camera.pos)
        pc.camera.dir_00 = _f3(camera.dir_00)
        pc.camera.dir_du = _f3(camera.dir_du)
        pc.camera.dir_dv = _f3(camera.dir_dv)
        pc.rbfRadius = self.rbf_radius
        pc.unit = float(unit)
        pc.world = self.bvh
        pc.particles = self.particles.data_ptr()
        pc.numParticles = self.n
        pc.colormap, pc.radiusmap, pc.densitymap = self.colormap, self.radiusmap, self.densitymap
        pc.radii = self.radii.data_ptr() if self.radii is not None else 0
        return pc

    def raygen_data(self, image, accum, width: int, height: int, stbn: int = 0) -> RayGenData:
        rg = RayGenData()
        rg.stbnTexture = stbn
        rg.imageBuffer = image.data_ptr()
        rg.accumBuffer = accum.data_ptr()
        rg.fbSize = Int2(width, height)
        rg.globalAABBMin = _f3(self.aabb[:3])
        rg.globalAABBMax = _f3(self.aabb[3:])
        return rg
