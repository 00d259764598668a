"""Closed-form pins of the oracle functions that the GPU is compared with, and
of the host helpers of the library (camera basis, camera controller, light
direction).  Each check is also run on a build with one plausible planted
mistake and must fail there: a check a slip passes pins nothing.

  TAA resolve       deviceCodeCommon.hlsl:27-43,67-115 (lerp .2f, YCoCg clamp)
  BGRA8 + GUI over  deviceCodeCommon.hlsl:110-114, sharedCode.h:142-149
  marching NEE      deviceCodeRBF.hlsl:131-137 (alpha += 1 - e^-D, clamp at 1)
                    and its composition in the raygen, :287-320
  light direction   sharedCode.h:127-132 (3.14, not pi)
  camera basis      hostCode.cpp:669-682
  camera controller hostCode.cpp:593-731
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from aarbf import _lib as L
from aarbf import scenes
from mutants import host_mutant, oracle_mutant

F32 = np.float32


def _fails(check, *a):
    try:
        check(*a)
    except AssertionError:
        return True
    return False


# ---- TAA --------------------------------------------------------------------------

def _taa_inputs(seed=0, H=7, W=9):
    """Colours on a 1/16 grid, so the YCoCg transform and its inverse are
    exact in binary32 and only the blend rounds."""
    rng = np.random.default_rng(seed)
    img = (rng.integers(0, 17, (H, W, 4)) / 16.0).astype(F32)
    return img


def check_taa_history_inside_box():
    """History inside the neighbourhood's YCoCg box is not clamped:
    out = old + 0.2 * (center - old), bit for bit in binary32."""
    img = _taa_inputs(1)
    H, W = img.shape[:2]
    # history = mean of the pixel and its right neighbour: inside the box
    # componentwise (YCoCg is linear), on a 1/32 grid (still exact)
    right = img[:, np.minimum(np.arange(W) + 1, W - 1)]
    prev = ((img + right) / F32(2)).astype(F32)
    prev[..., 3] = (img[..., 3] + F32(0.25)).astype(F32)  # alpha is blended, never clamped
    taa = np.zeros_like(img)
    p = prev.copy()
    oracle.taa_resolve(img, taa, p)
    expect = (prev + F32(0.2) * (img - prev)).astype(F32)
    assert np.array_equal(taa, expect), f"max diff {np.abs(taa - expect).max()}"
    assert np.array_equal(p, taa), "history copy"


def _ycocg(c):
    r, g, b = c[..., 0], c[..., 1], c[..., 2]
    return np.stack([r / 4 + g / 2 + b / 4, r - b, g - (r + b) / 2], -1)


def _rgb(y):
    Y, co, cg = y[..., 0], y[..., 1], y[..., 2]
    return np.stack([Y + co / 2 - cg / 2, Y + cg / 2, Y - co / 2 - cg / 2], -1)


def check_taa_history_outside_box():
    """History outside the box on one YCoCg component is moved to the box's
    face: the clamp of YCoCg(old) to [min, max] of the 3x3 neighbourhood,
    mapped back to RGB, then blended -- all evaluated here in float64 from
    the definition of the YCoCg-R transform."""
    img = _taa_inputs(2)
    H, W = img.shape[:2]
    ys = _ycocg(img.astype(np.float64))
    pad = np.pad(ys, ((1, 1), (1, 1), (0, 0)), mode="edge")
    win = np.stack([pad[1 + dy:1 + dy + H, 1 + dx:1 + dx + W] for dy in (-1, 0, 1) for dx in (-1, 0, 1)])
    lo, hi = win.min(0), win.max(0)
    for comp, off in ((0, 0.75), (1, -1.5), (2, 1.25)):
        yo = ys.copy()
        yo[..., comp] = np.where(off > 0, hi[..., comp], lo[..., comp]) + off
        prev = np.concatenate([_rgb(yo), img[..., 3:4].astype(np.float64) - 0.5], -1).astype(F32)
        taa = np.zeros_like(img)
        oracle.taa_resolve(img, taa, prev.copy())
        clamped = np.clip(_ycocg(prev.astype(np.float64)), lo, hi)
        old = np.concatenate([_rgb(clamped), prev[..., 3:4].astype(np.float64)], -1)
        expect = old + 0.2 * (img.astype(np.float64) - old)
        err = np.abs(taa - expect).max()
        assert err < 2e-6, f"component {comp}: max error {err}"


TAA_MUTANTS = [
    ("out[j] = old[j] + 0.2f * (center[j] - old[j]);", "out[j] = old[j] + 0.8f * (center[j] - old[j]);"),
    ("out[j] = old[j] + 0.2f * (center[j] - old[j]);", "out[j] = center[j] + 0.2f * (old[j] - center[j]);"),
    ("for (int j = 0; j < 3; ++j) y[j] = fminf(y[j], mx[j]);", "for (int j = 0; j < 3; ++j) y[j] = y[j];"),
    ("for (int j = 0; j < 3; ++j) y[j] = fmaxf(y[j], mn[j]);", "for (int j = 0; j < 3; ++j) y[j] = y[j];"),
]


def test_taa_pins():
    check_taa_history_inside_box()
    check_taa_history_outside_box()


@pytest.mark.parametrize("k", range(len(TAA_MUTANTS)))
def test_taa_pins_catch_planted_mistakes(tmp_path, k):
    with oracle.using(oracle_mutant(tmp_path, *TAA_MUTANTS[k])):
        assert _fails(check_taa_history_inside_box) or _fails(check_taa_history_outside_box)


def test_taa_disabled_passes_the_frame_and_copies_history():
    img = _taa_inputs(3)
    img[..., 3] = 1.0
    img[0, 0, 3] = 0.0  # over(no GUI, transparent pixel) = transparent black
    taa = np.full_like(img, 0.125)
    prev = np.zeros_like(img)
    frame = np.zeros(img.shape[:2], np.uint32)
    oracle.taa_resolve(img, taa, prev, disable=True, frame=frame)
    assert np.all(taa == F32(0.125)) and np.array_equal(prev, taa)
    exp = _pack_expect(img)
    exp[0, 0] = 0
    assert np.array_equal(frame, exp)


# ---- BGRA8 packing and the GUI layer ------------------------------------------------

def _pack_expect(c):
    """B | G << 8 | R << 16 | A << 24 of channels on a 1/16 grid: v * 255
    is then k * 15.9375, never within 1/16 of a rounding boundary."""
    q = lambda v: np.floor(np.clip(v.astype(np.float64), 0, 1) * 255 + 0.5).astype(np.uint32)
    return q(c[..., 2]) | (q(c[..., 1]) << 8) | (q(c[..., 0]) << 16) | (q(c[..., 3]) << 24)


def _srgb(b):
    c = b / 255.0
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def check_frame_packing_and_gui_over():
    H, W = 16, 16
    img = _taa_inputs(4, H, W)
    img[..., 3] = 1.0
    img[0, :4] = [[-0.5, 2.0, 0.5, 1.0], [np.nan, 0.25, 1.5, 1.0], [0.0, 0.0, 1.0, 1.0], [1.0, 0.5, 0.0, 1.0]]
    taa = np.zeros_like(img)
    prev = img.copy()  # history = frame: TAA keeps the frame
    frame = np.zeros((H, W), np.uint32)
    oracle.taa_resolve(img, taa, prev, frame=frame)
    exp = _pack_expect(np.nan_to_num(taa, nan=0.0))
    assert np.array_equal(frame, exp), "BGRA8 without a GUI layer"
    # an opaque GUI texel replaces the pixel; its sRGB bytes are decoded:
    # each byte b appears once per channel over the 256 pixels
    b = np.arange(256, dtype=np.uint32).reshape(H, W)
    gui = b | (((b * 7) & 255) << 8) | (((b * 13) & 255) << 16) | (np.uint32(255) << 24)
    oracle.taa_resolve(img, taa, img.copy(), gui=gui, frame=frame)
    lin = np.stack([_srgb(b), _srgb((b * 7) & 255), _srgb((b * 13) & 255)], -1)
    q = np.floor(lin * 255 + 0.5).astype(np.uint32)
    near = np.abs(lin * 255 + 0.5 - np.round(lin * 255 + 0.5)) < 1e-3  # at a rounding boundary
    got = np.stack([(frame >> 16) & 255, (frame >> 8) & 255, frame & 255], -1)
    assert np.all((got == q) | near), "opaque GUI layer"
    assert np.all(frame >> 24 == 255)
    # a GUI texel with alpha a over an opaque pixel p: a * g + (1 - a) * p
    a8 = 96
    gui = np.full((H, W), 0xC0 | (0x40 << 8) | (0x10 << 16) | (a8 << 24), np.uint32)
    oracle.taa_resolve(img, taa, img.copy(), gui=gui, frame=frame)
    a = a8 / 255.0
    g = _srgb(np.array([0xC0, 0x40, 0x10], np.float64))
    mix = a * g + (1 - a) * taa[..., :3].astype(np.float64)  # over the TAA result
    got = np.stack([(frame >> 16) & 255, (frame >> 8) & 255, frame & 255], -1).astype(np.float64)
    ok = ~np.isnan(mix)  # a NaN channel packs as 0 (checked above)
    assert np.abs(got - np.clip(mix, 0, 1) * 255)[ok].max() <= 0.5 + 1e-3, "translucent GUI layer"


PACK_MUTANTS = [
    ("frame[ofs] = to_unorm8(c[2]) | (to_unorm8(c[1]) << 8) | (to_unorm8(c[0]) << 16) |",
     "frame[ofs] = to_unorm8(c[0]) | (to_unorm8(c[1]) << 8) | (to_unorm8(c[2]) << 16) |"),
    ("over(g, out, c);", "over(out, g, c);"),
    ("return (uint32_t)(v * 255.0f + 0.5f);", "return (uint32_t)(v * 255.0f);"),
    ("return c <= 0.04045f ? c / 12.92f : aa_powf((c + 0.055f) / 1.055f, 2.4f);",
     "return c <= 0.04045f ? c / 12.92f : aa_powf((c + 0.055f) / 1.055f, 2.2f);"),
]


def test_frame_packing_pins():
    check_frame_packing_and_gui_over()


@pytest.mark.parametrize("k", range(len(PACK_MUTANTS)))
def test_frame_packing_pins_catch_planted_mistakes(tmp_path, k):
    with oracle.using(oracle_mutant(tmp_path, *PACK_MUTANTS[k])):
        assert _fails(check_frame_packing_and_gui_over)


# ---- stochastic marching: the NEE shadow ray ---------------------------------------

def _sphere_scene(c, R, P=(0.0, 0.0, 0.0), light=(0.1, 0.3, 0.3), unit=0.01, cam=None, W=8, H=8,
                  colour=(0.8, 0.5, 0.25, 1.0), blue=False, stbn=None):
    xyzw = np.array([[*P, 0.5]], F32)
    radii = np.array([R], F32)
    aabb = np.array([*P, *P], F32)
    cam = cam or scenes_camera((0.0, 0.0, 1.0), P, W, H, 0.5)
    return oracle.Render(xyzw, radii, aabb, scenes.constant_table(colour), scenes.constant_table(n=1),
                         scenes.constant_table((c, c, c, c)), cam, W, H, unit, R, light=light,
                         blue_noise=blue, stbn=stbn)


def scenes_camera(look_from, look_at, W, H, s):
    import aarbf
    return aarbf.camera_basis(look_from, look_at, (0.0, 1.0, 0.0), s, W, H)


def _chord_steps(o, d, t0, unit, P, R, tmax):
    """Number of shadow-march points t_k = t0 + k * unit (k >= 1, accumulated
    in binary32 as the tracker adds unit) inside the sphere, or None when a
    point lies within 1e-4 R of the sphere's surface (left undecided)."""
    o32, d32 = np.asarray(o, F32), np.asarray(d, F32)
    t = F32(t0)
    n = 0
    while True:
        t = F32(t + F32(unit))
        if t >= tmax:
            return n
        x = (o32 + t * d32).astype(np.float64)
        dist = math.dist(x, P)
        if abs(dist - R) < 1e-4 * R:
            return None
        n += dist < R


def _box_exit(o, d, lo, hi):
    t1 = (np.asarray(lo, np.float64) - o) / d
    t2 = (np.asarray(hi, np.float64) - o) / d
    return float(np.minimum(np.maximum(t1, t2), np.inf).min())


def check_shadow_visibility():
    """Visibility of the shadow ray through one sphere of constant density D
    (constant densitymap): n unit steps inside the chord add 1 - e^-D each,
    clamped at 1, so visibility = max(0, 1 - n (1 - e^-D))."""
    c, R, unit = 0.05, 0.25, 0.01
    r = _sphere_scene(c, R, unit=unit)
    L = oracle.light_direction(0.1, 0.3).astype(np.float64)
    rng = np.random.default_rng(5)
    box_lo, box_hi = np.full(3, -2 * R), np.full(3, 2 * R)
    checked = saturated = 0
    for k in range(120):
        # origins on the light's line through the sphere, offset sideways,
        # starting below it (the chord is 0 .. 2R long)
        side = np.cross(L, [0.0, 1.0, 0.0])
        side /= np.linalg.norm(side)
        o = (-0.45 * L + rng.uniform(-1.1, 1.1) * R * side).astype(F32)
        t0 = -unit * float(rng.uniform(0, 1))
        tmax = _box_exit(o.astype(np.float64), L, box_lo, box_hi)
        n = _chord_steps(o, oracle.light_direction(0.1, 0.3), t0, unit, (0, 0, 0), R, tmax)
        if n is None:
            continue
        want = max(0.0, 1.0 - n * (1.0 - math.exp(-c)))
        got = r.shadow_visibility(o, t0)
        assert abs(got - want) < 2e-6, f"origin {o} t0 {t0}: visibility {got}, expected {want} ({n} steps)"
        checked += 1
        saturated += want == 0.0
    assert checked > 100 and saturated > 5


def check_marching_nee_render():
    """ParticleRBFRayGen under stochastic marching, ambient 0.3, one small
    sphere of density 5 (> -ln(1 - 0.99 xi) for every xi, so the primary
    march stops at its first point inside the sphere): each pixel is
    albedo * (vis * 0.7 + 0.3) with vis = max(0, 1 - n (1 - e^-5)) for the n
    shadow points (from t = -unit * random.z) inside the sphere, or black
    when no primary point falls inside."""
    W = H = 24
    unit, R, c, amb = 0.05, 0.04, 5.0, 0.3
    colour = (0.8, 0.5, 0.25, 1.0)
    cam = scenes_camera((0.02, 0.01, 0.2), (0.0, 0.0, 0.0), W, H, 0.35)
    stbn = scenes.synthetic_stbn(seed=9)
    r = _sphere_scene(c, R, unit=unit, cam=cam, W=W, H=H, colour=colour, blue=True, stbn=stbn,
                      light=(0.15, 0.35, amb))
    img, _, _ = r.run(spp=1, threads=2)
    L = oracle.light_direction(0.15, 0.35)
    lo, hi = np.full(3, -2 * R), np.full(3, 2 * R)
    hits = checked = 0
    for py in range(H):
        for px in range(W):
            _, draws = oracle.rng_stream(0, px, py, W, 2)
            ry, rz = float(draws[0]), float(draws[1])
            s = np.array([(px + 0.5) / W, (py + 0.5) / H], np.float64)
            d = cam.dir_00.astype(np.float64) + s[0] * cam.dir_du + s[1] * cam.dir_dv
            d /= np.linalg.norm(d)
            o = cam.pos.astype(np.float64)
            t1, t2 = (lo - o) / d, (hi - o) / d
            tin, tout = np.minimum(t1, t2).max(), np.maximum(t1, t2).min()
            want = np.zeros(3)
            ok = True
            if tin < tout:
                t = F32(F32(tin) + F32(unit) * F32(ry))
                hit = None
                while True:
                    t = F32(t + F32(unit))
                    if t >= tout:
                        break
                    x = (cam.pos + t * d.astype(F32)).astype(F32)
                    dist = np.linalg.norm(x.astype(np.float64))
                    if abs(dist - R) < 1e-4 * R:
                        ok = False
                        break
                    if dist < R:
                        hit = x
                        break
                if ok and hit is not None:
                    n = _chord_steps(hit, L, -(F32(unit) * F32(rz)), unit, (0, 0, 0), R,
                                     _box_exit(hit.astype(np.float64), L.astype(np.float64), lo, hi))
                    if n is None:
                        ok = False
                    else:
                        vis = max(0.0, 1.0 - n * (1.0 - math.exp(-c)))
                        want = np.asarray(colour[:3]) * (vis * (1 - amb) + amb)
                        hits += 1
            if not ok:
                continue
            checked += 1
            assert np.abs(img[py, px, :3] - want).max() < 3e-6, f"pixel {(px, py)}: {img[py, px]} vs {want}"
            assert img[py, px, 3] == 1.0
    assert checked > 0.9 * W * H and hits > 0.3 * W * H


NEE_MUTANTS = [
    ("tr->albedo[3] += 1.0f - aa_expf(-density);", "tr->albedo[3] += aa_expf(-density);"),
    ("if (tr->albedo[3] >= 1.0f) {", "if (tr->albedo[3] >= 2.0f) {"),
    ("tr.t = p->enableBlueNoise ? -p->unit * random[2] : 0.0f;", "tr.t = 0.0f;"),
    ("color[j] = (albedo[j] * visibility) * (1.0f - p->light_ambient) + albedo[j] * p->light_ambient;",
     "color[j] = (albedo[j] * visibility) * p->light_ambient + albedo[j] * (1.0f - p->light_ambient);"),
]


def test_marching_nee_pins():
    check_shadow_visibility()
    check_marching_nee_render()


@pytest.mark.parametrize("k", range(len(NEE_MUTANTS)))
def test_marching_nee_pins_catch_planted_mistakes(tmp_path, k):
    with oracle.using(oracle_mutant(tmp_path, *NEE_MUTANTS[k])):
        assert _fails(check_shadow_visibility) or _fails(check_marching_nee_render)


# ---- host helpers: light direction, camera basis, camera controller ---------------------

def _host(lib_path=None):
    lib = C.CDLL(lib_path) if lib_path else L.load()
    for name in ("aarbf_light_direction", "aarbf_camera_basis", "aarbf_camera_update"):
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = L.SIGNATURES[name]
    return lib


def _light(lib, az, el):
    out = np.empty(3, F32)
    assert lib.aarbf_light_direction(az, el, out.ctypes.data_as(C.POINTER(C.c_float))) == 0
    return out


# getLightDirection uses 3.14 for pi: (azimuth, elevation) = (0, 0) gives
# (sin(-1.57), 0, cos(-1.57)); (0.25, 1) gives a = b = 1.57; (0.5, 0.75)
# gives a = 0.785, b = 3.14 (sin 3.14 = 0.0015927, not 0)
LIGHT_KNOWN = [((0.0, 0.0), (-0.99999968, 0.0, 0.00079627)), ((0.25, 0.5), (0.0, 0.0, 1.0)),
               ((0.25, 1.0), (0.00079627, 0.99999937, 0.00079627)),
               ((0.5, 0.75), (-0.70682426, 0.00112565, 0.70738829))]


def check_light_direction(fn):
    for (az, el), want in LIGHT_KNOWN:
        got = fn(az, el)
        assert np.abs(np.asarray(got, np.float64) - want).max() < 2e-7, f"({az}, {el}): {got} vs {want}"


def test_light_direction_known_answers():
    lib = _host()
    check_light_direction(lambda a, e: _light(lib, a, e))
    check_light_direction(oracle.light_direction)


def test_light_direction_catches_planted_mistakes(tmp_path):
    lib = _host(host_mutant(tmp_path, "camera.cpp", "float a = elevation * 3.14f - 0.5f * 3.14f;",
                            "float a = elevation * 3.14159265f - 0.5f * 3.14159265f;"))
    assert _fails(check_light_direction, lambda a, e: _light(lib, a, e))
    with oracle.using(oracle_mutant(tmp_path, "float b = azimuth * 2.0f * 3.14f;", "float b = azimuth * 3.14f;")):
        assert _fails(check_light_direction, oracle.light_direction)


def _basis(lib, lf, la, lu, s, W, H):
    f = lambda v: np.ascontiguousarray(np.asarray(v, F32))
    a, b, c = f(lf), f(la), f(lu)
    pc = L.PushConstants()
    fp = lambda x: x.ctypes.data_as(C.POINTER(C.c_float))
    assert lib.aarbf_camera_basis(fp(a), fp(b), fp(c), s, W, H, C.byref(pc)) == 0
    v = lambda q: np.array([q.x, q.y, q.z], np.float64)
    return v(pc.camera.pos), v(pc.camera.dir_00), v(pc.camera.dir_du), v(pc.camera.dir_dv)


def check_camera_basis(lib):
    s = 2 * math.tan(math.radians(22.5))  # 45 degree field of view: 0.828427
    pos, d00, du, dv = _basis(lib, (0, 0, 3), (0, 0, 0), (0, 1, 0), s, 512, 512)
    assert np.allclose(pos, (0, 0, 3), atol=0) and np.allclose(du, (s, 0, 0), atol=1e-7)
    assert np.allclose(dv, (0, s, 0), atol=1e-7) and np.allclose(d00, (-s / 2, -s / 2, -1), atol=1e-7)
    # 2:1 aspect widens du only; looking down -x with z up
    _, d00, du, dv = _basis(lib, (3, 0, 0), (0, 0, 0), (0, 0, 1), s, 1024, 512)
    assert np.allclose(du, (0, 2 * s, 0), atol=1e-6) and np.allclose(dv, (0, 0, s), atol=1e-7)
    assert np.allclose(d00, (-1, -s, -s / 2), atol=1e-6)


CAMERA_MUTANTS = [
    ("V3 ddu = scale(cos_fovy * aspect, normalize(cross(d00, up)));", "V3 ddu = scale(cos_fovy, normalize(cross(d00, up)));"),
    ("V3 ddu = scale(cos_fovy * aspect, normalize(cross(d00, up)));",
     "V3 ddu = scale(cos_fovy * aspect, normalize(cross(up, d00)));"),
    ("d00 = sub(d00, scale(0.5f, ddv));", "d00 = add(d00, scale(0.5f, ddv));"),
]


def test_camera_basis_closed_form():
    check_camera_basis(_host())


@pytest.mark.parametrize("k", range(len(CAMERA_MUTANTS)))
def test_camera_basis_catches_planted_mistakes(tmp_path, k):
    assert _fails(check_camera_basis, _host(host_mutant(tmp_path, "camera.cpp", *CAMERA_MUTANTS[k])))


def _ctl(lib, lf=(0, 0, 3), la=(0, 0, 0), lu=(0, 1, 0), s=0.8, W=512, H=256):
    st = L.CameraState(L.Float3(*lf), L.Float3(*la), L.Float3(*lu), s, 1)
    pc = L.PushConstants()

    def step(diag=2.0, **kw):
        inp = L.CameraInput(float(kw.get("dx", 0)), float(kw.get("dy", 0)),
                            *(int(kw.get(k, 0)) for k in ("left", "right", "middle", "key_x", "key_y", "key_z",
                                                            "left_shift")))
        ch = C.c_int()
        pc.accumID = 7
        assert lib.aarbf_camera_update(C.byref(st), C.byref(inp), W, H, diag, C.byref(pc), C.byref(ch)) == 0
        v = lambda q: np.array([q.x, q.y, q.z], np.float64)
        return bool(ch.value), v(st.lookFrom), v(st.lookAt), v(st.lookUp), pc
    return step


def check_camera_controller(lib):
    W, H = 512, 256
    step = _ctl(lib, W=W, H=H)
    # first frame: no movement, the basis is set, accumID reset
    ch, lf, la, lu, pc = step(dx=40, dy=9)
    assert ch and pc.accumID == 1 and np.allclose(lf, (0, 0, 3), atol=0)
    _, d00, du, _ = _basis(lib, (0, 0, 3), (0, 0, 0), (0, 1, 0), 0.8, W, H)
    assert np.allclose([pc.camera.dir_00.x, pc.camera.dir_00.y, pc.camera.dir_00.z], d00, atol=0)
    # no button: nothing moves
    ch, lf, _, _, pc = step(dx=5, dy=5)
    assert not ch and pc.accumID == 7
    # drag a quarter of the width: -pi/2 about +y takes (0, 0, 3) to (-3, 0, 0)
    ch, lf, _, _, pc = step(dx=W / 4, left=1)
    assert ch and pc.accumID == 1 and np.allclose(lf, (-3, 0, 0), atol=2e-6), lf
    # back, then a quarter of the height down: about lookRight = (-1, 0, 0) by -pi/4
    step(dx=-W / 4, left=1)
    ch, lf, _, _, _ = step(dy=H / 4, left=1)
    assert np.allclose(lf, (0, -3 / math.sqrt(2), 3 / math.sqrt(2)), atol=3e-6), lf
    # zoom: dy > 0 scales the view vector by 0.95, dy < 0 by 1.05; only pos changes
    step = _ctl(lib, W=W, H=H)
    _, _, _, _, pc = step()
    d00_before = np.array([pc.camera.dir_00.x, pc.camera.dir_00.y, pc.camera.dir_00.z])
    _, lf, _, _, pc = step(dy=2, right=1)
    assert np.allclose(lf, (0, 0, 2.85), atol=1e-6) and abs(pc.camera.pos.z - 2.85) < 1e-6
    assert np.allclose([pc.camera.dir_00.x, pc.camera.dir_00.y, pc.camera.dir_00.z], d00_before, atol=0)
    _, lf, _, _, _ = step(dy=-2, right=1)
    assert np.allclose(lf, (0, 0, 2.85 * 1.05), atol=1e-6)
    # pan: (lookRight * dx - up * dy) * 0.001 * diagonal, lookRight = cross(up, view) = (-1, 0, 0)
    step = _ctl(lib, W=W, H=H)
    step()
    _, lf, la, _, pc = step(dx=10, dy=5, middle=1, diag=2.0)
    assert np.allclose(lf, (-0.02, -0.01, 3), atol=1e-6) and np.allclose(la, (-0.02, -0.01, 0), atol=1e-6)
    # up keys: Z, and shift + Y
    step = _ctl(lib, W=W, H=H)
    step()
    _, _, _, lu, _ = step(key_z=1)
    assert np.array_equal(lu, (0, 0, 1))
    _, _, _, lu, _ = step(key_y=1, left_shift=1)
    assert np.array_equal(lu, (0, -1, 0))


CONTROLLER_MUTANTS = [
    ("float x_angle = -dx * delta_x;", "float x_angle = dx * delta_x;"),
    ("float delta_y = (float)(M_PI / height);", "float delta_y = (float)(2.0 * M_PI / height);"),
    ("if (dy > 0.0f) {", "if (dy < 0.0f) {"),
    ("V3 t = add(scale(dx, right), scale(-dy, up));", "V3 t = add(scale(dx, right), scale(dy, up));"),
    ("if (in->left_shift) up = scale(-1.0f, up);", "if (in->left_shift) up = up;"),
]


def test_camera_controller_closed_forms():
    check_camera_controller(_host())


@pytest.mark.parametrize("k", range(len(CONTROLLER_MUTANTS)))
def test_camera_controller_catches_planted_mistakes(tmp_path, k):
    assert _fails(check_camera_controller, _host(host_mutant(tmp_path, "camera.cpp", *CONTROLLER_MUTANTS[k])))


# ---- transfer-function editing: marks -> 64 RGBA8 texels -------------------------------

def _texels(lib, marks, kind):
    arr = (L.TFMark * len(marks))()
    for k, (pos, rgba) in enumerate(marks):
        arr[k].position = pos
        arr[k].rgba[:] = list(rgba)
    out = np.zeros((64, 4), np.uint8)
    fn = lib.aarbf_tf_texels_from_marks
    fn.restype, fn.argtypes = L.SIGNATURES["aarbf_tf_texels_from_marks"]
    assert fn(arr, len(marks), kind, out.ctypes.data_as(C.POINTER(C.c_uint8))) == 0
    return out


def check_tf_texels(lib):
    """Known answers of hostCode.cpp:524-572's texel writes."""
    # a grey ramp: texel i samples i / 63; make_8bit truncates f * 256 and
    # clamps (1.0 -> 256 -> 255); the colormap gamma-encodes RGB
    ramp = [(0.0, (0, 0, 0, 0)), (1.0, (1, 1, 1, 1))]
    t = _texels(lib, ramp, 0)
    assert t[0].tolist() == [0, 0, 0, 0] and t[63].tolist() == [255, 255, 255, 255]
    # i = 21: 1/3 -> alpha int(85.33) = 85; rgb (1/3)^(1/2.2) = 0.6069 -> 155
    assert t[21].tolist() == [155, 155, 155, 85], t[21]
    # i = 63 * 0.5 is not a texel; i = 9: 1/7 -> alpha 36, rgb 0.4129 -> 105
    assert t[9].tolist() == [105, 105, 105, 36], t[9]
    s = _texels(lib, ramp, 1)  # a scalar map: red only, linear
    assert s[21].tolist() == [85, 0, 0, 0] and s[9].tolist() == [36, 0, 0, 0] and s[63].tolist() == [255, 0, 0, 0]
    # the reference's default colormap: blue at 0, cyan-ish at 0.25, ..., red at 1
    cm = _texels(lib, [(0.0, (0, 0, 1, 1)), (0.25, (0, 0.972549, 1, 1)), (0.5, (0, 1, 0.00784314, 1)),
                       (0.75, (0.996078, 1, 0, 1)), (1.0, (1, 0, 0, 1))], 0)
    assert cm[0].tolist() == [0, 0, 255, 255] and cm[63].tolist() == [255, 0, 0, 255]
    # marks in any order; a single mark is a constant table; outside the
    # marks the end colours hold
    shuffled = _texels(lib, [(1.0, (1, 0, 0, 1)), (0.0, (0, 0, 1, 1)), (0.75, (0.996078, 1, 0, 1)),
                             (0.5, (0, 1, 0.00784314, 1)), (0.25, (0, 0.972549, 1, 1))], 0)
    assert np.array_equal(shuffled, cm)
    one = _texels(lib, [(0.3, (0.75, 0.5, 0.5, 0.5))], 1)
    assert np.all(one[:, 0] == 192) and not one[:, 1:].any()  # int(0.75 * 256), not round(0.75 * 255) = 191
    step = _texels(lib, [(0.5, (0, 0, 0, 0)), (0.6, (1, 1, 1, 1))], 1)
    assert np.all(step[:32, 0] == 0) and np.all(step[38:, 0] == 255)
    # i = 35: 35/63 = 0.5556 -> t = 0.5556 -> int(142.2) = 142
    assert step[35, 0] == 142, step[35]


TF_MUTANTS = [
    ("for (int k = 0; k < 3; ++k) t[k] = make_8bit(std::pow(c[k], 1.0f / 2.2f));",
     "for (int k = 0; k < 3; ++k) t[k] = make_8bit(std::pow(c[k], 2.2f));"),
    ("uint8_t make_8bit(float f) { return (uint8_t)std::min(255, std::max(0, (int)(f * 256.0f))); }",
     "uint8_t make_8bit(float f) { return (uint8_t)std::min(255, std::max(0, (int)(f * 255.0f + 0.5f))); }"),
    ("gradient_at(m, (float)i / 63.0f, c);", "gradient_at(m, (float)i / 64.0f, c);"),
    ("for (int c = 0; c < 4; ++c) out[c] = a.rgba[c] + t * (b.rgba[c] - a.rgba[c]);",
     "for (int c = 0; c < 4; ++c) out[c] = b.rgba[c] + t * (a.rgba[c] - b.rgba[c]);"),
]


def test_tf_texels_known_answers():
    check_tf_texels(_host())
    # and the Python presets are these texels
    assert np.array_equal(scenes.reference_tf_texels("colormap"),
                          _texels(_host(), scenes.REFERENCE_COLORMAP_MARKS, 0))


@pytest.mark.parametrize("k", range(len(TF_MUTANTS)))
def test_tf_texels_catch_planted_mistakes(tmp_path, k):
    assert _fails(check_tf_texels, C.CDLL(host_mutant(tmp_path, "tf_marks.cpp", *TF_MUTANTS[k])))
