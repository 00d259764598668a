"""torch.autograd.Function over the library's differentiable renderer
(aarbf_diff_forward / aarbf_diff_backward).

    R = DiffRenderer(ctx, camera, width, height, unit, spp=4)
    image = R(positions, attributes, radii, colormap, densitymap, radiusmap)
    loss(image).backward()

positions [N, 3], attributes [N] (in [0, 1]), radii [N] and the RGBA
tables [n, 4] are float32 tensors on the HIP device; image is [H, W, 3].
Each call Hilbert-sorts the particles, builds bounds and the BVH for the
current values and renders; gradients come back in the caller's particle
order (densitymap / radiusmap gradients in their red channel, where the
renderer reads them).  The image is the expectation of stochastic marching
over its threshold (primary ray, ambient 1), see include/aarbf.h.
"""
from __future__ import annotations

import weakref

import numpy as np


def _torch():
    import torch
    return torch


class _Render:
    """Per-call device state: sorted particles, tables, BVH."""

    def __init__(self, ctx, camera, W, H, unit, spp, frame_id, rbf_radius, ppl, branching, positions, attributes,
                 radii, colormap, densitymap, radiusmap, estimator=0):
        torch = _torch()
        from . import PushConstants, RayGenData, Int2, _f3
        self.ctx, self.W, self.H, self.spp = ctx, W, H, spp
        self.estimator = estimator
        # The library works on its own stream: torch's work is finished before
        # each library call reads a tensor, and the library's before torch reads.
        p4 = torch.cat([positions.detach().float(), attributes.detach().float()[:, None]], 1).contiguous()
        torch.cuda.synchronize()
        self.aabb = ctx.spacetime_aabb(p4)
        codes = ctx.hilbert_codes(p4, self.aabb)
        self.perm = ctx.sort_particles(p4, codes).long()  # sorted[i] = original[perm[i]]
        ctx.sync()
        self.p4 = p4
        self.radii = radii.detach().float()[self.perm].contiguous()
        torch.cuda.synchronize()
        self.tfs = [ctx.tf(t.detach().float().cpu().numpy()) for t in (colormap, densitymap, radiusmap)]
        self.n_tex = [t.shape[0] for t in (colormap, densitymap, radiusmap)]
        rbf = float(rbf_radius if rbf_radius is not None else (radii.detach().max().item() *
                                                               radiusmap.detach()[:, 0].max().item()))
        self.aabbs = ctx.gen_bounds(p4, ppl, rbf, self.tfs[2], self.radii)
        self.bvh = ctx.bvh_build(self.aabbs, branching)
        pc = PushConstants()
        pc.accumID, pc.frameID = 1, frame_id & 0xFFFF
        pc.visualizeAttributes = 1
        pc.particlesPerLeaf = ppl
        pc.camera.pos, pc.camera.dir_00 = _f3(camera.pos), _f3(camera.dir_00)
        pc.camera.dir_du, pc.camera.dir_dv = _f3(camera.dir_du), _f3(camera.dir_dv)
        pc.rbfRadius, pc.unit = rbf, float(unit)
        pc.world, pc.particles, pc.numParticles = self.bvh, p4.data_ptr(), p4.shape[0]
        pc.colormap, pc.densitymap, pc.radiusmap = self.tfs
        pc.radii = self.radii.data_ptr()
        self.pc = pc
        self.img = torch.zeros((H, W, 4), dtype=torch.float32, device=p4.device)
        self.path = torch.zeros((H, W, spp, 4), dtype=torch.float32, device=p4.device)
        torch.cuda.synchronize()
        rg = RayGenData()
        rg.imageBuffer = rg.accumBuffer = self.img.data_ptr()
        rg.fbSize = Int2(W, H)
        rg.globalAABBMin, rg.globalAABBMax = _f3(self.aabb[:3]), _f3(self.aabb[3:])
        self.rg = rg
        # the BVH and the tables live in the context until closed: a forward
        # that never reaches backward (no_grad, validation, an output left
        # out of the loss) releases them when this object is collected
        self._finalizer = weakref.finalize(self, _Render._release, ctx, self.bvh, list(self.tfs))

    def forward(self):
        self.ctx.set_diff_estimator(self.estimator)
        self.ctx.diff_forward(self.rg, self.pc, self.W, self.H, self.spp, self.img, self.path)
        self.ctx.sync()
        return self.img[..., :3].clone()

    def backward(self, grad):
        torch = _torch()
        dev = self.p4.device
        n = self.p4.shape[0]
        g = torch.zeros((self.H, self.W, 4), dtype=torch.float32, device=dev)
        g[..., :3] = grad
        gp = torch.zeros(n * 3, dtype=torch.float32, device=dev)
        gr = torch.zeros(n, dtype=torch.float32, device=dev)
        ga = torch.zeros(n, dtype=torch.float32, device=dev)
        gcm = torch.zeros(self.n_tex[0] * 4, dtype=torch.float32, device=dev)
        gdm = torch.zeros(self.n_tex[1], dtype=torch.float32, device=dev)
        grm = torch.zeros(self.n_tex[2], dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        self.ctx.set_diff_estimator(self.estimator)
        self.ctx.diff_backward(self.rg, self.pc, self.W, self.H, self.spp, g, gp, gr, ga, gcm, gdm, grm,
                               path=self.path)
        self.ctx.sync()
        inv = lambda t: torch.zeros_like(t).index_copy_(0, self.perm, t)  # back to the caller's order
        gdm4 = torch.zeros((self.n_tex[1], 4), dtype=torch.float32, device=dev)
        gdm4[:, 0] = gdm
        grm4 = torch.zeros((self.n_tex[2], 4), dtype=torch.float32, device=dev)
        grm4[:, 0] = grm
        return inv(gp.view(n, 3)), inv(ga), inv(gr), gcm.view(-1, 4), gdm4, grm4

    @staticmethod
    def _release(ctx, bvh, tfs):
        if getattr(ctx, "h", None) is None:  # the context is gone, and its handles with it
            return
        ctx.bvh_destroy(bvh)
        for h in tfs:
            ctx.tf_destroy(h)

    def close(self):
        """Releases the BVH and tables now; later calls do nothing."""
        self._finalizer()


class _Fn:
    @staticmethod
    def build():
        torch = _torch()

        class ParticleRender(torch.autograd.Function):
            @staticmethod
            def forward(fctx, settings, positions, attributes, radii, colormap, densitymap, radiusmap):
                *args, estimator, keep = settings
                r = _Render(*args, positions, attributes, radii, colormap, densitymap, radiusmap,
                            estimator=estimator)
                img = r.forward()
                if keep:
                    fctx.r = r  # released with the graph (several backwards with retain_graph are fine)
                else:
                    r.close()  # no backward can follow
                return img

            @staticmethod
            def backward(fctx, grad):
                r = fctx.r
                gp, ga, gr, gcm, gdm, grm = r.backward(grad.float().contiguous())
                return None, gp, ga, gr, gcm, gdm, grm

        return ParticleRender


class DiffRenderer:
    def __init__(self, ctx, camera, width: int, height: int, unit: float, spp: int = 4, frame_id: int = 0,
                 rbf_radius: float | None = None, ppl: int = 1, branching: int = 8, estimator: int = 0):
        """estimator 0: expected stochastic marching; 1: ratio tracking of
        delta tracking (include/aarbf.h, aarbf_set_diff_estimator)."""
        self.settings = (ctx, camera, width, height, unit, spp, frame_id, rbf_radius, ppl, branching, estimator)
        self.fn = _Fn.build()

    def __call__(self, positions, attributes, radii, colormap, densitymap, radiusmap):
        torch = _torch()
        inputs = (positions, attributes, radii, colormap, densitymap, radiusmap)
        keep = torch.is_grad_enabled() and any(t.requires_grad for t in inputs)
        return self.fn.apply(self.settings + (keep,), *inputs)
