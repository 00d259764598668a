"""ctypes bindings of libaarbf.so (include/aarbf.h).

The Structure classes mirror the C structs field by field; load() checks
their sizes and offsets against aarbf_abi_layout() exported by the library.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libaarbf.so")

u16, u32, u64, i32, f32 = C.c_uint16, C.c_uint32, C.c_uint64, C.c_int32, C.c_float


class Float3(C.Structure):
    _fields_ = [("x", f32), ("y", f32), ("z", f32)]


class Int2(C.Structure):
    _fields_ = [("x", i32), ("y", i32)]


class Light(C.Structure):
    _fields_ = [("azimuth", f32), ("elevation", f32), ("ambient", f32)]


class Camera(C.Structure):
    _fields_ = [("pos", Float3), ("dir_00", Float3), ("dir_du", Float3), ("dir_dv", Float3)]


class PushConstants(C.Structure):
    """sharedCode.h:33-73."""
    _fields_ = [
        ("accumID", u16), ("frameID", u16), ("enableBlueNoise", u16), ("showNoise", u16),
        ("visualizeAttributes", u16), ("particlesPerLeaf", u16),
        ("light", Light), ("camera", Camera),
        ("rbfRadius", f32), ("unit", f32),
        ("world", u64), ("particles", u64), ("numParticles", u32),
        ("colormap", u64), ("radiusmap", u64), ("densitymap", u64),
        ("radii", u64),
    ]


class BoundsConstants(C.Structure):
    """sharedCode.h:76-86."""
    _fields_ = [
        ("aabbs", u64), ("numAABBs", u32), ("particles", u64), ("numParticles", u32),
        ("particlesPerLeaf", u16), ("rbfRadius", f32), ("radiusmap", u64), ("radii", u64),
    ]


class TAAConstants(C.Structure):
    """sharedCode.h:88-105."""
    _fields_ = [
        ("disableTAA", u16), ("showNoise", u16), ("fbSize", Int2),
        ("imageTexture", u64), ("guiTexture", u64), ("frameBuffer", u64),
        ("taaBuffer", u64), ("taaPrevBuffer", u64),
    ]


class RayGenData(C.Structure):
    """sharedCode.h:108-117."""
    _fields_ = [
        ("stbnTexture", u64), ("imageBuffer", u64), ("accumBuffer", u64), ("fbSize", Int2),
        ("globalAABBMin", Float3), ("globalAABBMax", Float3),
    ]


class Request(C.Structure):
    """One view of aarbf_render_batch."""
    _fields_ = [("rg", RayGenData), ("pc", PushConstants), ("width", C.c_int), ("height", C.c_int), ("spp", u32)]


class TuneResult(C.Structure):
    _fields_ = [("step_cap", C.c_int32), ("step_waves", C.c_int32), ("graph_rounds", C.c_int32),
                ("candidates", u32), ("baseline_ms", C.c_double), ("best_ms", C.c_double),
                ("walk_budget", C.c_int32), ("step_min_live", C.c_int32)]


class CameraState(C.Structure):
    """The reference's interactive camera (hostCode.cpp:253-271)."""
    _fields_ = [("lookFrom", Float3), ("lookAt", Float3), ("lookUp", Float3), ("cosFovy", f32),
                ("firstFrame", i32)]


class CameraInput(C.Structure):
    """Mouse and key input of one frame of the camera controller."""
    _fields_ = [("dx", f32), ("dy", f32), ("left", i32), ("right", i32), ("middle", i32),
                ("key_x", i32), ("key_y", i32), ("key_z", i32), ("left_shift", i32)]


class TFMark(C.Structure):
    """One mark of a transfer-function gradient."""
    _fields_ = [("position", f32), ("rgba", f32 * 4)]


class RenderStats(C.Structure):
    _fields_ = [(n, u64) for n in ("samples", "steps", "queries", "nodes", "leaves", "particles",
                                   "covering", "exhausted")]


P = C.c_void_p
FP = C.POINTER(C.c_float)

# name -> (restype, argtypes)
SIGNATURES = {
    "aarbf_context_create": (C.c_int, [C.c_int, C.POINTER(P)]),
    "aarbf_context_destroy": (C.c_int, [P]),
    "aarbf_set_stream": (C.c_int, [P, P]),
    "aarbf_get_stream": (P, [P]),
    "aarbf_sync": (C.c_int, [P]),
    "aarbf_last_error": (C.c_char_p, [P]),
    "aarbf_status_string": (C.c_char_p, [C.c_int]),
    "aarbf_set_stats": (C.c_int, [P, C.c_int]),
    "aarbf_set_render_mode": (C.c_int, [P, C.c_int]),
    "aarbf_set_tuning": (C.c_int, [P, C.c_int, C.c_int]),
    "aarbf_render_stats_read": (C.c_int, [P, C.POINTER(RenderStats), C.c_int]),
    "aarbf_render_histogram": (C.c_int, [P, C.POINTER(u64), C.c_int]),
    "aarbf_set_timing": (C.c_int, [P, C.c_int]),
    "aarbf_set_payload": (C.c_int, [P, C.c_int]),
    "aarbf_payload_read": (C.c_int, [P, u64, C.c_uint32]),
    "aarbf_diff_forward": (C.c_int, [P, C.POINTER(RayGenData), C.POINTER(PushConstants), C.c_int, C.c_int,
                                     C.c_uint32, u64, u64]),
    "aarbf_diff_backward": (C.c_int, [P, C.POINTER(RayGenData), C.POINTER(PushConstants), C.c_int, C.c_int,
                                      C.c_uint32] + [u64] * 8),
    "aarbf_read_timing": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(u64), C.c_int]),
    "aarbf_import_particles": (C.c_int, [C.c_char_p, C.POINTER(P)]),
    "aarbf_timesteps_count": (u32, [P]),
    "aarbf_timesteps_get": (C.c_int, [P, u32, C.POINTER(FP), C.POINTER(u64), C.POINTER(C.c_char_p)]),
    "aarbf_timesteps_destroy": (None, [P]),
    "aarbf_import_step_device": (C.c_int, [P, C.c_char_p, u32, u64, u64, C.POINTER(u64)]),
    "aarbf_export_particles": (C.c_int, [C.c_char_p, u32, C.POINTER(C.c_char_p), C.POINTER(FP),
                                         C.POINTER(u64)]),
    "aarbf_spacetime_aabb": (C.c_int, [P, u64, u64, FP]),
    "aarbf_hilbert_codes": (C.c_int, [P, u64, u64, FP, u64]),
    "aarbf_sort_particles": (C.c_int, [P, u64, u64, u64, u64]),
    "aarbf_attribute_range": (C.c_int, [P, u64, u64, FP]),
    "aarbf_normalize_attributes": (C.c_int, [P, u64, u64, FP]),
    "aarbf_tf_create": (C.c_int, [P, FP, u32, C.POINTER(u64)]),
    "aarbf_tf_create_rgba8": (C.c_int, [P, C.POINTER(C.c_uint8), u32, C.c_int, C.POINTER(u64)]),
    "aarbf_tf_destroy": (C.c_int, [P, u64]),
    "aarbf_tf_texels_from_marks": (C.c_int, [C.POINTER(TFMark), u32, C.c_int, C.POINTER(C.c_uint8)]),
    "aarbf_tf_create_from_marks": (C.c_int, [P, C.POINTER(TFMark), u32, C.c_int, C.POINTER(u64)]),
    "aarbf_stbn_create": (C.c_int, [P, C.POINTER(C.c_uint8), C.c_int, C.c_int, C.POINTER(u64)]),
    "aarbf_stbn_destroy": (C.c_int, [P, u64]),
    "aarbf_gen_bounds": (C.c_int, [P, C.POINTER(BoundsConstants)]),
    "aarbf_bvh_build": (C.c_int, [P, u64, u32, u32, C.POINTER(u64)]),
    "aarbf_bvh_refit": (C.c_int, [P, u64, u64]),
    "aarbf_bvh_destroy": (C.c_int, [P, u64]),
    "aarbf_bvh_levels": (C.c_int, [P, u64, C.POINTER(u32), C.POINTER(u32), u32]),
    "aarbf_bvh_read_level": (C.c_int, [P, u64, u32, FP]),
    "aarbf_bvh_read_groups": (C.c_int, [P, u64, u32, C.POINTER(u32)]),
    "aarbf_render_rbf": (C.c_int, [P, C.POINTER(RayGenData), C.POINTER(PushConstants), C.c_int, C.c_int]),
    "aarbf_render_rbf_spp": (C.c_int, [P, C.POINTER(RayGenData), C.POINTER(PushConstants), C.c_int,
                                       C.c_int, u32]),
    "aarbf_render_splat": (C.c_int, [P, C.POINTER(RayGenData), C.POINTER(PushConstants), C.c_int,
                                     C.c_int]),
    "aarbf_render_rbf_spp_noalloc": (C.c_int, [P, C.POINTER(RayGenData), C.POINTER(PushConstants), C.c_int,
                                               C.c_int, u32]),
    "aarbf_render_batch": (C.c_int, [P, C.POINTER(Request), u32]),
    "aarbf_batch_latency": (C.c_int, [P, C.POINTER(C.c_double), u32]),
    "aarbf_autotune": (C.c_int, [P, C.POINTER(Request), u32, u32, C.POINTER(TuneResult)]),
    "aarbf_set_diff_estimator": (C.c_int, [P, C.c_int]),
    "aarbf_taa_resolve": (C.c_int, [P, C.POINTER(TAAConstants)]),
    "aarbf_abi_layout": (u32, [C.POINTER(u32), u32]),
    "aarbf_handle_counts": (C.c_int, [P, C.POINTER(u32)]),
    "aarbf_light_direction": (C.c_int, [f32, f32, FP]),
    "aarbf_camera_basis": (C.c_int, [FP, FP, FP, f32, C.c_int, C.c_int, C.POINTER(PushConstants)]),
    "aarbf_camera_update": (C.c_int, [C.POINTER(CameraState), C.POINTER(CameraInput), C.c_int, C.c_int, f32,
                                      C.POINTER(PushConstants), C.POINTER(C.c_int)]),
}


def _expected_layout():
    return [
        C.sizeof(PushConstants), PushConstants.light.offset, PushConstants.camera.offset,
        PushConstants.rbfRadius.offset, PushConstants.world.offset, PushConstants.numParticles.offset,
        PushConstants.colormap.offset, PushConstants.radii.offset,
        C.sizeof(BoundsConstants), BoundsConstants.particlesPerLeaf.offset,
        BoundsConstants.radiusmap.offset,
        C.sizeof(TAAConstants), TAAConstants.imageTexture.offset,
        C.sizeof(RayGenData), RayGenData.fbSize.offset, RayGenData.globalAABBMax.offset,
        C.sizeof(RenderStats),
        C.sizeof(Request), Request.pc.offset, Request.spp.offset,
        C.sizeof(TuneResult), TuneResult.walk_budget.offset, TuneResult.step_min_live.offset,
        C.sizeof(CameraState), CameraState.firstFrame.offset,
        C.sizeof(CameraInput), CameraInput.left_shift.offset,
    ]


_lib = None


def load(path: str | None = None) -> C.CDLL:
    """Loads libaarbf.so; raises if it is missing (there is no fallback)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libaarbf.so not built at {p}: run attribute-aware-rbfs_amd/build.py")
    # torch ships its own libamdhip64 / libhsa-runtime64 (same sonames as
    # ROCm's).  Loading torch first makes the library bind to that already
    # loaded runtime, so one HIP runtime owns the device and torch tensors,
    # streams and events are valid handles for the library.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    buf = (u32 * 32)()
    m = lib.aarbf_abi_layout(buf, 32)
    got = list(buf[:m])
    exp = _expected_layout()
    if got != exp:
        raise RuntimeError(f"ABI layout mismatch between include/aarbf.h and _lib.py: {got} != {exp}")
    if path is None:
        _lib = lib
    return lib
