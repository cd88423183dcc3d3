"""Batched inverse lighting: per-light RGB emission and per-triangle RGB albedo of many
scenes recovered by Adam on the photon-map radiance loss
    loss = sum over scenes, hits x of |L(x) - L_target(x)|^2,
where every step runs one batched kNN radiance launch (MNRT_PHASE_KNN) and one batched
adjoint launch (MNRT_PHASE_KNN_BWD) for all scenes.

Photon paths are traced once.  A photon's power is its trace-time power reweighted
linearly: times the current emission multiplier of its light (1 = the traced light power)
and the ratio of the current to the trace-time albedos of the triangles it scattered from
(`adjoint.reweight_photons`).  A hit's diffuse colour is the current albedo of its
triangle.  Both are torch ops, so the chain from dLoss/dpower (per photon) and
dLoss/dcolour (per hit) to the parameters is torch autograd; the kNN passes and their
adjoint are the library's.

The reference has no inverse problem (MNRT renders only); the kNN radiance it
differentiates is KernelPMPhotonGather (photon_knn.cu:487-552).
"""

from __future__ import annotations

import numpy as np

from . import adjoint, api
from . import native as N


def _torch():
    import torch

    return torch


class BatchedInverseLighting:
    """scenes: list of Scene; photon_counts[s]: photons traced in scene s; dims[s]: (W, H)
    pinhole camera of scene s (its primary hits are the queries); k neighbours per query."""

    def __init__(self, scenes, photon_counts, dims, k: int = 50, seed: int = 0):
        torch = _torch()
        self.S, self.k = len(scenes), k
        cat = lambda a: api.dev(np.concatenate([getattr(sc, a) for sc in scenes]).astype(np.float32))
        v0, v1, v2, alb = cat("v0"), cat("v1"), cat("v2"), cat("albedo")
        self.toff = np.zeros(self.S + 1, np.int64)
        self.toff[1:] = np.cumsum([sc.num_tris for sc in scenes])
        self.loff = np.zeros(self.S + 1, np.int64)
        self.loff[1:] = np.cumsum([sc.lights.shape[0] for sc in scenes])
        tsl = [slice(int(self.toff[s]), int(self.toff[s + 1])) for s in range(self.S)]
        self.TB = api.build_tri_batch(v0, v1, v2, self.toff.astype(np.uint32))
        eps = np.array([sc.ray_epsilon() for sc in scenes], np.float32)
        rmax = np.array([sc.r_max() for sc in scenes], np.float32)
        # photon paths, traced once; light and triangle indices made global
        pos, pw, dr, light, tris = [], [], [], [], []
        for s, sc in enumerate(scenes):
            g = api.photons_generate(self.TB[s], v0[tsl[s]], v1[tsl[s]], v2[tsl[s]], alb[tsl[s]], sc.lights,
                                     int(photon_counts[s]), 4, seed + s, float(eps[s]), want_meta=True)
            pos.append(g["pos"])
            pw.append(g["power"])
            dr.append(g["dir"])
            light.append((g["meta"] & 0xFF).long() + int(self.loff[s]))
            t = g["tris"].long()
            tris.append(torch.where(t >= 0, t + int(self.toff[s]), t))
        self.ppos, self.pdir = torch.cat(pos).contiguous(), torch.cat(dr).contiguous()
        self.light, self.tris = torch.cat(light), torch.cat(tris)
        self.poff = np.zeros(self.S + 1, np.int64)
        self.poff[1:] = np.cumsum([p.shape[0] for p in pos])
        self.albedo0 = alb[:, :3].clone()
        valid = self.tris >= 0
        prod0 = torch.where(valid[..., None], self.albedo0[torch.where(valid, self.tris, 0)],
                            torch.ones(1, device="cuda")).prod(1)
        self.base = torch.cat(pw).clone()
        self.base[:, :3] = self.base[:, :3] / prod0
        boxes = np.stack([np.concatenate(sc.bounds()) for sc in scenes]).astype(np.float32)
        self.PB = api.build_points_batch(self.ppos, self.poff.astype(np.uint32), root_boxes=boxes,
                                         max_query_radius=rmax)
        api.attach_photon_records_batch(self.PB, self.ppos, torch.cat(pw).contiguous(), self.pdir)
        # queries: primary hits of every scene's camera
        self.roff = np.zeros(self.S + 1, np.int64)
        self.roff[1:] = np.cumsum([w * h for w, h in dims])
        R = max(int(self.roff[-1]), 1)
        o = torch.empty((R, 4), dtype=torch.float32, device="cuda")
        d = torch.empty((R, 4), dtype=torch.float32, device="cuda")
        for s, (w, h) in enumerate(dims):
            if w * h:
                api.primary_rays(w, h, out_orig=o[self.roff[s]:self.roff[s + 1]],
                                 out_dir=d[self.roff[s]:self.roff[s + 1]])
        self.Q = api.QueryBatch(self.TB, self.toff.astype(np.uint32), self.roff.astype(np.uint32), eps, self.PB, rmax,
                                k)
        f4 = lambda: torch.zeros((R, 4), dtype=torch.float32, device="cuda")
        self.bufs = dict(orig4=o, dir4=d, v0=v0, v1=v1, v2=v2, albedo4=alb,
                         hit_tri=torch.empty(R, dtype=torch.int32, device="cuda"),
                         hit_t=torch.empty(R, dtype=torch.float32, device="cuda"),
                         sp_pos4=f4(), sp_nrm4=f4(), sp_dir4=f4(), sp_clr4=f4(), sp_infl4=f4(),
                         sp_src=torch.empty(R, dtype=torch.int32, device="cuda"),
                         sp_count=torch.zeros(self.S, dtype=torch.int32, device="cuda"), io_rad4=f4(),
                         out_r2=torch.empty(R, dtype=torch.float32, device="cuda"),
                         out_count=torch.empty(R, dtype=torch.int32, device="cuda"),
                         grad_rad4=f4(), grad_pow4=torch.empty((max(int(self.poff[-1]), 1), 4), device="cuda"),
                         grad_clr4=f4())
        self.Q.make_io(**self.bufs)
        self.Q.run(N.PHASE_TRACE | N.PHASE_SHADE)
        torch.cuda.synchronize()
        # valid slots (each scene's first count[s]) and the global triangle each one hit
        cnt = self.bufs["sp_count"].cpu().numpy().astype(np.int64)
        scene_of = np.repeat(np.arange(self.S), np.diff(self.roff))
        local = np.arange(int(self.roff[-1])) - self.roff[scene_of]
        vmask = local < cnt[scene_of]
        self.valid = torch.from_numpy(vmask).cuda()
        vidx = np.nonzero(vmask)[0]
        src = self.bufs["sp_src"].cpu().numpy().astype(np.int64)[vidx]
        hit = self.bufs["hit_tri"].cpu().numpy().astype(np.int64)
        self.vidx = torch.from_numpy(vidx).cuda()
        self.hit_tri = torch.from_numpy(hit[self.roff[scene_of[vidx]] + src] + self.toff[scene_of[vidx]]).cuda()
        assert torch.equal(self.bufs["sp_clr4"][self.vidx, :3], self.albedo0[self.hit_tri]), "hit colours"
        self.slot_scene = torch.from_numpy(scene_of[vidx]).cuda()
        self.num_queries = int(vidx.shape[0])

    def radiance(self, emission, albedo):
        """Radiance at every valid slot (V, 3) for emission multipliers (total lights, 3) and
        albedo (total triangles, 3); no gradient."""
        torch = _torch()
        with torch.no_grad():
            power = adjoint.reweight_photons(self.base, self.light, self.tris, emission, albedo)
            self._forward(power, albedo[self.hit_tri])
            return self.bufs["io_rad4"][self.vidx, :3].clone()

    def _forward(self, power, clr):
        api.attach_photon_records_batch(self.PB, self.ppos, power.detach().contiguous(), self.pdir)
        self.bufs["sp_clr4"][self.vidx, :3] = clr.detach()
        self.bufs["io_rad4"].zero_()
        self.Q.run(N.PHASE_KNN)

    def loss_and_grad(self, emission, albedo, target):
        """Loss and its gradients: accumulates into emission.grad / albedo.grad."""
        torch = _torch()
        power = adjoint.reweight_photons(self.base, self.light, self.tris, emission, albedo)
        clr = albedo[self.hit_tri]
        self._forward(power, clr)
        diff = self.bufs["io_rad4"][self.vidx, :3] - target
        loss = (diff.double() ** 2).sum()
        g = self.bufs["grad_rad4"]
        g.zero_()
        g[self.vidx, :3] = 2.0 * diff
        self.Q.run(N.PHASE_KNN_BWD)
        gp = self.bufs["grad_pow4"][: power.shape[0]]
        gc = self.bufs["grad_clr4"][self.vidx, :3]
        torch.autograd.backward([power, clr], [gp, gc])
        return float(loss)

    def fit(self, target, steps: int = 200, lr: float = 1e-2, emission0=None, albedo0=None, fit_albedo=True,
            log=None):
        """Adam from emission multipliers 1 and the trace-time albedo (or the given starts);
        fit_albedo=False holds the albedo.  Returns (emission, albedo, loss per step)."""
        torch = _torch()
        L = int(self.loff[-1])
        e = (emission0.clone() if emission0 is not None else torch.ones((L, 3), device="cuda")).requires_grad_(True)
        a = (albedo0.clone() if albedo0 is not None else self.albedo0.clone()).requires_grad_(True)
        opt = torch.optim.Adam([e, a] if fit_albedo else [e], lr=lr)
        losses = []
        for it in range(steps):
            opt.zero_grad(set_to_none=False)
            losses.append(self.loss_and_grad(e, a, target))
            opt.step()
            with torch.no_grad():
                e.clamp_(min=0.0)
                a.clamp_(0.0, 1.0)
            if log and (it % 50 == 0 or it == steps - 1):
                log(f"step {it}: loss {losses[-1]:.6e}")
        return e.detach(), a.detach(), losses

    def release(self):
        self.Q.close()
        self.TB.release()
        self.PB.release()
