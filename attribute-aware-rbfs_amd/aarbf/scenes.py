"""Synthetic workloads, transfer-function presets and the blue-noise stand-in.

Random streams come from splitmix64 on explicit integer seeds (numpy uint64
arithmetic), so every workload is reproducible independently of the numpy
version; uniforms are the top 53 bits / 2^53 and Gaussians use Box-Muller.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """n outputs of splitmix64 started at seed + stream * 2^32."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64((seed + (stream << 32)) & 0xFFFFFFFFFFFFFFFF) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, n: int, stream: int = 0) -> np.ndarray:
    return (splitmix64(seed, n, stream) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, n: int, stream: int = 0) -> np.ndarray:
    u1 = uniform(seed, n, 2 * stream + 1)
    u2 = uniform(seed, n, 2 * stream + 2)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


# ---- transfer functions -----------------------------------------------------

def ramp_table(n: int = 256, channels=(True, True, True, True)) -> np.ndarray:
    """Linear ramp: texel i = (i + 0.5) / n, so a linearly filtered lookup at
    u returns u for u in [0.5/n, 1 - 0.5/n] and clamps outside."""
    v = (np.arange(n, dtype=np.float64) + 0.5) / n
    t = np.zeros((n, 4), dtype=np.float32)
    for c in range(4):
        t[:, c] = v if channels[c] else 1.0
    return t


def constant_table(value=(1.0, 1.0, 1.0, 1.0), n: int = 1) -> np.ndarray:
    return np.tile(np.asarray(value, dtype=np.float32), (n, 1))


def make_8bit(f: float) -> int:
    """hostCode.cpp:524-526."""
    return min(255, max(0, int(np.float32(f) * np.float32(256.0))))


# hostCode.cpp:420-436 default marks
REFERENCE_COLORMAP_MARKS = [(0.0, (0.0, 0.0, 1.0, 1.0)), (0.25, (0.0, 0.972549, 1.0, 1.0)),
                            (0.5, (0.0, 1.0, 0.00784314, 1.0)), (0.75, (0.996078, 1.0, 0.0, 1.0)),
                            (1.0, (1.0, 0.0, 0.0, 1.0))]
REFERENCE_RADIUSMAP_MARKS = [(0.0, (1.0, 1.0, 1.0, 1.0)), (1.0, (1.0, 1.0, 1.0, 1.0))]
REFERENCE_DENSITYMAP_MARKS = [(0.0, (0.0, 0.0, 0.0, 0.0)), (1.0, (1.0, 1.0, 1.0, 1.0))]

TF_COLORMAP, TF_SCALAR = 0, 1


def tf_texels_from_marks(marks, kind: int) -> np.ndarray:
    """The 64 RGBA8 texels of a gradient of (position, rgba) marks
    (aarbf_tf_texels_from_marks, hostCode.cpp:524-572): kind TF_COLORMAP
    gamma-encodes RGB with pow(c, 1/2.2) and stores alpha linearly;
    TF_SCALAR (radius / density maps) writes only R."""
    import ctypes as C
    from . import _lib
    arr = (_lib.TFMark * len(marks))()
    for k, (pos, rgba) in enumerate(marks):
        arr[k].position = float(pos)
        arr[k].rgba[:] = [float(v) for v in rgba]
    out = np.zeros((64, 4), dtype=np.uint8)
    st = _lib.load().aarbf_tf_texels_from_marks(arr, len(marks), int(kind), out.ctypes.data_as(C.POINTER(C.c_uint8)))
    if st != 0:
        raise ValueError("aarbf_tf_texels_from_marks: marks need positions in [0, 1]")
    return out


def reference_tf_texels(which: str) -> np.ndarray:
    """The 64 RGBA8 texels the reference writes into its R8G8B8A8_SRGB TF
    textures for its default marks (hostCode.cpp:416-436,528-572)."""
    marks = {"colormap": REFERENCE_COLORMAP_MARKS, "radiusmap": REFERENCE_RADIUSMAP_MARKS,
             "densitymap": REFERENCE_DENSITYMAP_MARKS}[which]
    return tf_texels_from_marks(marks, TF_COLORMAP if which == "colormap" else TF_SCALAR)


def srgb_decode_table(texels: np.ndarray) -> np.ndarray:
    """Float RGBA of R8G8B8A8_SRGB texels: RGB through the sRGB EOTF."""
    t = texels.astype(np.float64) / 255.0
    rgb = np.where(t[:, :3] <= 0.04045, t[:, :3] / 12.92, ((t[:, :3] + 0.055) / 1.055) ** 2.4)
    out = np.concatenate([rgb, t[:, 3:4]], axis=1)
    return out.astype(np.float32)


def reference_tf(which: str) -> np.ndarray:
    return srgb_decode_table(reference_tf_texels(which))


# ---- blue noise ------------------------------------------------------------------

def synthetic_stbn(seed: int = 7, tiles_x: int = 8, tiles_y: int = 16, tile: int = 256) -> np.ndarray:
    """Deterministic stand-in for stbn.png (absent from the reference tree):
    an 8x16 atlas of 256x256 RGBA8 tiles of hashed white noise, indexed the
    same way (deviceCodeRBF.hlsl:245-249)."""
    h, w = tiles_y * tile, tiles_x * tile
    v = splitmix64(seed, h * w * 4)
    return (v >> np.uint64(56)).astype(np.uint8).reshape(h, w, 4)


# ---- workloads -----------------------------------------------------------------------

@dataclass
class Workload:
    name: str
    xyzw: np.ndarray            # float32 [N, 4] raw attributes
    radii: np.ndarray | None    # float32 [N] or None
    rbf_radius: float           # clip-box margin / global radius
    unit: float
    width: int
    height: int
    spp: int
    blue_noise: bool
    look_from: tuple
    look_at: tuple
    look_up: tuple
    cos_fovy: float
    colormap: np.ndarray
    radiusmap: np.ndarray
    densitymap: np.ndarray
    frames: int = 1
    taa: bool = False


def fovy_scale(deg: float) -> float:
    """Image-plane scale 2 tan(fovy/2) for the reference's cosFovy slot."""
    return 2.0 * math.tan(math.radians(deg) / 2.0)


def uniform_cube(n: int = 65536, seed: int = 1, width: int = 512, height: int = 512) -> Workload:
    """CPU reference workload: N particles uniform in the unit cube centred
    on the origin, radii U[0.01, 0.03], attribute = sum of 3 random-phase
    sinusoids + N(0, 0.05); 256-entry RGBA ramp TF; 1 spp stochastic
    marching from (0, 0, 3) looking at the origin with a 45 degree field of
    view."""
    pos = np.stack([uniform(seed, n, s) for s in range(3)], axis=1) - 0.5
    radii = 0.01 + 0.02 * uniform(seed, n, 3)
    f = (splitmix64(seed, 9, 4) % np.uint64(4)).astype(np.float64).reshape(3, 3) + 1.0
    phase = 2.0 * np.pi * uniform(seed, 3, 5)
    attr = sum(np.sin(2.0 * np.pi * (pos @ f[k]) + phase[k]) for k in range(3)) + 0.05 * normal(seed, n, 6)
    xyzw = np.concatenate([pos, attr[:, None]], axis=1).astype(np.float32)
    return Workload("uniform-cube-65536", xyzw, radii.astype(np.float32), 0.03, 0.002, width, height, 1, True,
                    (0.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), fovy_scale(45.0),
                    ramp_table(256), constant_table(n=256), ramp_table(256))


def gaussian_mixture(n: int, components: int, seed: int, sigma_range=(0.02, 0.15), r_median: float = 0.004,
                     r_sigma: float = 0.5, noise_octaves: int = 4):
    """Particles from an equal-weight Gaussian mixture: centres U[0,1]^3,
    sigma log-uniform in sigma_range; log-normal radii; attribute = distance
    to the nearest centre plus 1/f value noise (octave o has amplitude
    2^-o)."""
    centres = np.stack([uniform(seed, components, s) for s in range(3)], axis=1)
    lo, hi = np.log(sigma_range[0]), np.log(sigma_range[1])
    sig = np.exp(lo + (hi - lo) * uniform(seed, components, 3))
    comp = (splitmix64(seed, n, 4) % np.uint64(components)).astype(np.int64)
    g = np.stack([normal(seed, n, 5 + s) for s in range(3)], axis=1)
    pos = centres[comp] + g * sig[comp, None]
    radii = r_median * np.exp(r_sigma * normal(seed, n, 9))
    # distance to the nearest centre, in chunks
    dmin = np.empty(n, dtype=np.float64)
    for a in range(0, n, 1 << 20):
        b = min(n, a + (1 << 20))
        d = np.sqrt(((pos[a:b, None, :] - centres[None, :, :]) ** 2).sum(-1))
        dmin[a:b] = d.min(1)
    noise = np.zeros(n)
    for o in range(noise_octaves):
        freq = 4.0 * 2 ** o
        cell = np.floor(pos * freq).astype(np.int64)
        h = (cell[:, 0] * 73856093) ^ (cell[:, 1] * 19349663) ^ (cell[:, 2] * 83492791) ^ (o * 2654435761)
        hv = (h.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)) >> np.uint64(40)
        noise += (hv.astype(np.float64) / float(1 << 24) - 0.5) * 2.0 ** (-o) * 0.1
    attr = dmin + noise
    xyzw = np.concatenate([pos, attr[:, None]], axis=1).astype(np.float32)
    return xyzw, radii.astype(np.float32)


def hierarchical_clusters(n: int = 67108864, seed: int = 13, levels: int = 3, children: int = 16,
                          sigmas=(0.12, 0.035, 0.01), leaf_sigma: float = 0.004, r_range=(0.0005, 0.01),
                          r_power: float = 2.0, device: str = "cuda"):
    """Precision workload scene: a `levels`-level Gaussian mixture with
    `children` children per node (16 x 16 x 16 = 4,096 leaf clusters at the
    defaults), the root at the centre of the unit cube, level l offsets
    N(0, sigmas[l]); particles N(leaf centre, leaf_sigma); radii power-law
    p(r) ~ r^-r_power on r_range; attribute = distance to the particle's
    top-level centre plus N(0, 0.02).  Drawn with torch's seeded generator
    on `device` (67M particles in seconds); returns (xyzw [n, 4] float32,
    radii [n] float32) tensors on `device`."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    f64 = torch.float64
    centres = torch.full((1, 3), 0.5, dtype=f64, device=device)
    tops = None
    for lvl in range(levels):
        k = centres.shape[0]
        off = torch.randn((k, children, 3), generator=g, device=device, dtype=f64) * sigmas[lvl]
        centres = (centres[:, None, :] + off).reshape(-1, 3)
        if lvl == 0:
            tops = centres
    leaf_per_top = centres.shape[0] // tops.shape[0]
    comp = torch.randint(0, centres.shape[0], (n,), generator=g, device=device)
    pos = centres[comp] + torch.randn((n, 3), generator=g, device=device, dtype=f64) * leaf_sigma
    a, b = r_range
    u = torch.rand((n,), generator=g, device=device, dtype=f64)
    if r_power == 1.0:
        radii = a * (b / a) ** u
    else:
        e = 1.0 - r_power
        radii = (a ** e + u * (b ** e - a ** e)) ** (1.0 / e)
    attr = torch.linalg.vector_norm(pos - tops[comp // leaf_per_top], dim=1) + 0.02 * torch.randn(
        (n,), generator=g, device=device, dtype=f64)
    xyzw = torch.cat([pos, attr[:, None]], dim=1).to(torch.float32)
    return xyzw, radii.to(torch.float32)


def orbit_camera(frame: int, frames: int, radius: float, centre=(0.5, 0.5, 0.5), seed: int = 2,
                 elevation_deg: float = 20.0):
    """Seeded orbit: azimuth starts at a seeded angle and sweeps one full
    circle over `frames` frames at a fixed elevation."""
    a0 = 2.0 * np.pi * float(uniform(seed, 1, 0)[0])
    a = a0 + 2.0 * np.pi * frame / max(1, frames)
    el = np.radians(elevation_deg)
    c = np.asarray(centre, np.float64)
    pos = c + radius * np.array([np.cos(el) * np.cos(a), np.sin(el), np.cos(el) * np.sin(a)])
    return tuple(pos), tuple(c), (0.0, 1.0, 0.0)


def mixture_8m(n: int = 8388608, seed: int = 3, width: int = 1920, height: int = 1080, spp: int = 16,
               frames: int = 32) -> Workload:
    """GPU FP32 workload: 8,388,608 particles from a 32-component Gaussian
    mixture, log-normal radii (median 0.004, sigma 0.5); delta tracking at
    1920x1080, 16 spp, TAA over 32 frames along a seeded orbit.  The
    reference's default transfer functions (hostCode.cpp:420-436)."""
    xyzw, radii = gaussian_mixture(n, 32, seed)
    lf, la, lu = orbit_camera(0, frames, 1.6)
    return Workload(f"mixture-{n}", xyzw, radii, float(radii.max()), 0.0004, width, height, spp, False, lf, la, lu,
                    fovy_scale(45.0), reference_tf("colormap"), reference_tf("radiusmap"),
                    reference_tf("densitymap"), frames=frames, taa=True)


@dataclass
class DiffWorkload:
    """Differentiation workload: the scene, its perturbed target copy, the
    tables and the view set."""
    xyzw: np.ndarray         # float32 [N, 4], attributes normalised to [0, 1]
    radii: np.ndarray        # float32 [N]
    target_xyzw: np.ndarray  # positions + N(0, 0.002), attributes + N(0, 0.05) clipped to [0, 1]
    target_radii: np.ndarray  # radii * U[0.9, 1.1]
    colormap: np.ndarray     # 256-entry RGBA ramp
    densitymap: np.ndarray   # 256-entry ramp
    radiusmap: np.ndarray    # constant 1
    unit: float              # 0.1 x the median radius
    rbf_radius: float        # the largest radius of scene and target (clip-box margin)
    width: int = 512
    height: int = 512
    spp: int = 4
    views: int = 64

    def view(self, v: int):
        """(look_from, look_at, up) of view v: a seeded sphere of radius 1.6
        about the centre of the unit cube."""
        z = 2.0 * float(uniform(23, self.views, 0)[v]) - 1.0
        ph = 2.0 * np.pi * float(uniform(23, self.views, 1)[v])
        rr = np.sqrt(max(0.0, 1.0 - z * z))
        lf = (0.5 + 1.6 * rr * np.cos(ph), 0.5 + 1.6 * z, 0.5 + 1.6 * rr * np.sin(ph))
        return lf, (0.5, 0.5, 0.5), ((0.0, 1.0, 0.0) if abs(z) < 0.95 else (1.0, 0.0, 0.0))


def diff_workload(n: int = 1048576, seed: int = 7, views: int = 64) -> DiffWorkload:
    """Differentiation workload: n particles from the 32-component mixture
    (as mixture_8m) with attributes normalised to [0, 1]; the target is the
    same set with positions + N(0, 0.002), radii x U[0.9, 1.1] and
    attributes + N(0, 0.05); 64 views of 512 x 512 at 4 spp on a seeded
    sphere of cameras."""
    xyzw, radii = gaussian_mixture(n, 32, seed)
    attr = xyzw[:, 3].astype(np.float64)
    xyzw[:, 3] = ((attr - attr.min()) / max(attr.max() - attr.min(), 1e-12)).astype(np.float32)
    rng = np.random.default_rng(17)
    tgt = xyzw.copy()
    tgt[:, :3] += rng.normal(0.0, 0.002, (n, 3)).astype(np.float32)
    tgt[:, 3] = np.clip(tgt[:, 3] + rng.normal(0.0, 0.05, n), 0.0, 1.0).astype(np.float32)
    tgt_r = (radii * rng.uniform(0.9, 1.1, n)).astype(np.float32)
    return DiffWorkload(xyzw, radii, tgt, tgt_r, ramp_table(256), ramp_table(256), constant_table(n=4),
                        0.1 * float(np.median(radii)), float(max(radii.max(), tgt_r.max())), views=views)


def mixture_16m(n: int = 16777216, seed: int = 5):
    """Batching workload scene: 16,777,216 particles from a 64-component
    Gaussian mixture (as mixture_8m otherwise)."""
    xyzw, radii = gaussian_mixture(n, 64, seed)
    return xyzw, radii


def batch_views(n: int = 256, seed: int = 11, side=(128, 1024), aspect=(0.5, 2.0), spp=(1, 8),
                radius: float = 1.6, centre=(0.5, 0.5, 0.5)):
    """Camera requests of uneven resolution: side log-uniform in `side`,
    aspect uniform in `aspect` (W = side sqrt(aspect), H = side / sqrt(aspect)),
    spp uniform in {spp[0]..spp[1]}, stochastic marching or delta tracking
    with probability 1/2 each, cameras on a seeded sphere of `radius` about
    `centre` looking at it.  Returns a list of dicts."""
    u = [uniform(seed, n, s) for s in range(6)]
    lo, hi = np.log(side[0]), np.log(side[1])
    out = []
    for i in range(n):
        s = float(np.exp(lo + (hi - lo) * u[0][i]))
        a = aspect[0] + (aspect[1] - aspect[0]) * float(u[1][i])
        W = max(1, int(round(s * math.sqrt(a))))
        H = max(1, int(round(s / math.sqrt(a))))
        k = spp[0] + int(u[2][i] * (spp[1] - spp[0] + 1))
        z = 2.0 * float(u[3][i]) - 1.0
        phi = 2.0 * math.pi * float(u[4][i])
        rr = math.sqrt(max(0.0, 1.0 - z * z))
        pos = tuple(np.asarray(centre) + radius * np.array([rr * math.cos(phi), z, rr * math.sin(phi)]))
        up = (0.0, 1.0, 0.0) if abs(z) < 0.95 else (1.0, 0.0, 0.0)
        out.append({"width": W, "height": H, "spp": min(k, spp[1]), "marching": bool(u[5][i] < 0.5),
                    "look_from": pos, "look_at": tuple(centre), "up": up, "frame": i})
    return out
