"""Time series of particle sets, handled the way hostCode.cpp handles its
timesteps.

Pre-processing (hostCode.cpp:183-250): one space-time bounding box over all
steps, Hilbert codes and a sort per step against that box, one attribute
range over all steps and every step normalised with it.  The device keeps
every prepared step.  The render buffers are sized for the largest step
(hostCode.cpp:383-385: numAABBs = ceil(maxNumParticles / ppl); the leaves a
smaller step leaves unused are inactive).  Changing the step
(hostCode.cpp:751-799) copies its particles into the render buffer,
regenerates the leaf bounds for its particle count and refits the BVH (the
reference's gprtAccelUpdate); `rebuild=True` builds a new tree instead (the
reference does so when the radius or a transfer function changes).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import AarbfError, Context, Scene, _lib, _torch


def import_particles(path: str) -> list[tuple[str, np.ndarray]]:
    """importParticles.h:8-32 through aarbf_import_particles: one
    (name, float32 [N, 4]) per timestep.  Raises on a missing or truncated
    file."""
    lib = _lib.load()
    h = C.c_void_p()
    st = lib.aarbf_import_particles(path.encode(), C.byref(h))
    if st:
        raise AarbfError(st, f"aarbf_import_particles({path!r}): {lib.aarbf_status_string(st).decode()}")
    try:
        out = []
        for k in range(lib.aarbf_timesteps_count(h)):
            p = C.POINTER(C.c_float)()
            n = C.c_uint64()
            nm = C.c_char_p()
            st = lib.aarbf_timesteps_get(h, k, C.byref(p), C.byref(n), C.byref(nm))
            if st:
                raise AarbfError(st, "aarbf_timesteps_get")
            a = (np.ctypeslib.as_array(p, shape=(n.value * 4,)).reshape(-1, 4).copy() if n.value
                 else np.zeros((0, 4), np.float32))
            out.append((nm.value.decode(), a))
        return out
    finally:
        lib.aarbf_timesteps_destroy(h)


def import_step_device(ctx: Context, path: str, step: int, out=None, capacity: int | None = None):
    """Timestep `step` of the file set `path` read straight into a device
    float32 [N, 4] tensor (aarbf_import_step_device: chunked reads through
    pinned staging buffers overlapped with the copies).  `out` (a device
    tensor with room for `capacity` particles) is filled in place when
    given; returns the tensor holding the step's particles."""
    import torch
    lib = _lib.load()
    if out is None:
        # the count first, then a buffer of that size
        n = C.c_uint64()
        st = lib.aarbf_import_step_device(ctx.h, path.encode(), step, 0, 0, C.byref(n))
        if st not in (0, 1):
            ctx.check(st, "aarbf_import_step_device")
        out = torch.empty((max(1, n.value), 4), dtype=torch.float32, device=f"cuda:{ctx.device}")
        capacity = n.value
    cap = out.shape[0] if capacity is None else capacity
    n = C.c_uint64()
    torch.cuda.current_stream(ctx.device).synchronize()
    ctx.check(lib.aarbf_import_step_device(ctx.h, path.encode(), step, out.data_ptr(), cap, C.byref(n)),
              "aarbf_import_step_device")
    return out[: n.value]


def export_particles(path: str, steps) -> None:
    """Writes [(name, float32 [N, 4])] in the importer's format."""
    lib = _lib.load()
    names = (C.c_char_p * len(steps))(*[n.encode() for n, _ in steps])
    arrs = [np.ascontiguousarray(a, np.float32).reshape(-1, 4) for _, a in steps]
    ptrs = (C.POINTER(C.c_float) * len(steps))(*[a.ctypes.data_as(C.POINTER(C.c_float)) for a in arrs])
    counts = (C.c_uint64 * len(steps))(*[a.shape[0] for a in arrs])
    st = lib.aarbf_export_particles(path.encode(), len(steps), names, ptrs, counts)
    if st:
        raise AarbfError(st, f"aarbf_export_particles({path!r})")


class TimeSeries:
    """Prepared timesteps plus one render scene showing the current step.

    `steps[k]` / `radii[k]`: the k-th step's particles (float32 [N_k, 4],
    Hilbert-sorted, attributes normalised) and radii (or None) on the
    device.  `scene` is the Scene to render; after `set_step` it shows that
    step (its particle count, leaf bounds and refitted BVH)."""

    def __init__(self, ctx: Context, steps, radii, aabb, attr_range, colormap, radiusmap, densitymap,
                 rbf_radius: float, ppl: int, branching: int):
        torch = _torch()
        self.ctx = ctx
        self.steps = steps
        self.radii = radii
        self.aabb = aabb
        self.attr_range = attr_range
        self.max_n = max(int(s.shape[0]) for s in steps)
        dev = steps[0].device
        self.buf = torch.zeros((max(self.max_n, 1), 4), dtype=torch.float32, device=dev)
        self.rbuf = torch.zeros(max(self.max_n, 1), dtype=torch.float32, device=dev) if radii[0] is not None else None
        self.scene = Scene(ctx, self.buf[:0], None if self.rbuf is None else self.rbuf[:0], aabb, colormap,
                           radiusmap, densitymap, float(rbf_radius), ppl, branching, attr_range=attr_range)
        self.step = -1

    @staticmethod
    def prepare(ctx: Context, steps, colormap, radiusmap, densitymap, rbf_radius: float, radii=None,
                ppl: int = 1, branching: int = 8, device: str = "cuda") -> "TimeSeries":
        """steps: list of float32 [N_k, 4] (numpy or tensors), raw attributes;
        radii: None or a list of float32 [N_k].  Shows step 0."""
        torch = _torch()
        dev = torch.device(device, ctx.device) if device == "cuda" else torch.device(device)

        def on_dev(a, shape):
            if isinstance(a, torch.Tensor):
                return a.detach().to(device=dev, dtype=torch.float32).reshape(shape).contiguous().clone()
            return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32).reshape(shape)).to(dev)

        ps = [on_dev(s, (-1, 4)) for s in steps]
        rs = [None] * len(ps) if radii is None else [on_dev(r, (-1,)) for r in radii]
        if not ps or any(p.shape[0] == 0 for p in ps):
            raise ValueError("TimeSeries.prepare: every step needs at least one particle")
        aabb = None
        for p in ps:  # hostCode.cpp:186-193
            aabb = ctx.spacetime_aabb(p, aabb)
        for k, p in enumerate(ps):  # :196-221
            perm = ctx.sort_particles(p, ctx.hilbert_codes(p, aabb))
            if rs[k] is not None:
                rs[k] = rs[k][perm.long()].contiguous()
        rng = None
        for p in ps:  # :224-239
            rng = ctx.attribute_range(p, rng)
        for p in ps:  # :241-250
            ctx.normalize_attributes(p, rng)
        cm = colormap if isinstance(colormap, int) else ctx.tf(colormap)
        rm = radiusmap if isinstance(radiusmap, int) else ctx.tf(radiusmap)
        dm = densitymap if isinstance(densitymap, int) else ctx.tf(densitymap)
        ts = TimeSeries(ctx, ps, rs, aabb, rng, cm, rm, dm, rbf_radius, ppl, branching)
        ts.set_step(0, rebuild=True)
        return ts

    def __len__(self) -> int:
        return len(self.steps)

    def set_step(self, k: int, rebuild: bool = False) -> Scene:
        """hostCode.cpp:755-799: upload step k, regenerate the leaf bounds
        for its count, refit the BVH (or build a new one)."""
        sc = self.scene
        n = int(self.steps[k].shape[0])
        self.buf[:n].copy_(self.steps[k])
        if self.rbuf is not None:
            self.rbuf[:n].copy_(self.radii[k])
        sc.particles = self.buf[:n]
        sc.radii = None if self.rbuf is None else self.rbuf[:n]
        # bounds for every leaf of the largest step; the ones past n inactive
        sc.aabbs = self.ctx.gen_bounds(self.buf, sc.ppl, sc.rbf_radius, sc.radiusmap, self.rbuf, aabbs=sc.aabbs,
                                       num_particles=n)
        if rebuild or not sc.bvh:
            if sc.bvh:
                self.ctx.bvh_destroy(sc.bvh)
            sc.bvh = self.ctx.bvh_build(sc.aabbs, sc.branching)
        else:
            self.ctx.bvh_refit(sc.bvh, sc.aabbs)
        self.step = k
        return sc
