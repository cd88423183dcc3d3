"""Batched inverse lighting: per-light RGB emission and per-triangle RGB albedo of many
scenes recovered by Adam on the photon-map radiance loss
    loss = sum over scenes, hits x of |L(x) - L_target(x)|^2,
where every step runs one batched kNN radiance launch (MNRT_PHASE_KNN) and one batched
adjoint launch (MNRT_PHASE_KNN_BWD) for all scenes.

Photon paths are traced once.  With the paths fixed a photon's power is its base power
(its light's share of the emitted power, Russian-roulette doublings, 1 / paths shot) times
the emission multiplier of its light times the albedo of every triangle it scattered from;
a hit's diffuse colour is the albedo of its triangle.  Both maps and their adjoint are the
library's (mnrt_photons_reweight / _bwd: fixed-point sums, the same bits on every run), so
a step is a fixed sequence of device work with no host synchronisation; the loss of each
step stays on the device until the fit ends.

The reference has no inverse problem (MNRT renders only); the kNN radiance it
differentiates is KernelPMPhotonGather (photon_knn.cu:487-552).
"""

from __future__ import annotations

import numpy as np

from . import api
from . import native as N


def _torch():
    import torch

    return torch


def base_powers(meta, lights, paths: int):
    """Emission-only photon powers (P, 4) of a traced map: the power the tracer gives a
    photon before any albedo, i.e. (light power * light count) doubled once per Russian-
    roulette survival (bounce index > 3), times 1 / paths shot (mnrt_photons_generate)."""
    torch = _torch()
    lt = (meta & 0xFF).long()
    bounce = ((meta >> 8) & 0xFF).long()
    lp = torch.from_numpy(np.asarray(lights, np.float32)[:, 3, :].copy()).cuda()
    nl = float(np.asarray(lights).shape[0])
    rr = torch.clamp(bounce - 3, min=0).to(torch.float32)
    p = (lp[lt] * nl) * torch.exp2(rr)[:, None] * (1.0 / float(paths))
    return torch.cat([p, torch.zeros_like(p[:, :1])], dim=1).contiguous()


class BatchedInverseLighting:
    """scenes: list of Scene; photon_counts[s]: photons traced in scene s; dims[s]: (W, H)
    pinhole camera of scene s (its primary hits are the queries); k neighbours per query."""

    def __init__(self, scenes, photon_counts, dims, k: int = 50, seed: int = 0):
        torch = _torch()
        self.S, self.k = len(scenes), k
        self.scenes = scenes
        cat = lambda a: api.dev(np.concatenate([getattr(sc, a) for sc in scenes]).astype(np.float32))
        v0, v1, v2, alb = cat("v0"), cat("v1"), cat("v2"), cat("albedo")
        self.toff = np.zeros(self.S + 1, np.int64)
        self.toff[1:] = np.cumsum([sc.num_tris for sc in scenes])
        self.loff = np.zeros(self.S + 1, np.int64)
        self.loff[1:] = np.cumsum([sc.lights.shape[0] for sc in scenes])
        tsl = [slice(int(self.toff[s]), int(self.toff[s + 1])) for s in range(self.S)]
        self.TB = api.build_tri_batch(v0, v1, v2, self.toff.astype(np.uint32))
        self.eps = np.array([sc.ray_epsilon() for sc in scenes], np.float32)
        self.rmax = np.array([sc.r_max() for sc in scenes], np.float32)
        # photon paths, traced once; light and triangle indices made global
        pos, base, dr, light, tris = [], [], [], [], []
        for s, sc in enumerate(scenes):
            g = api.photons_generate(self.TB[s], v0[tsl[s]], v1[tsl[s]], v2[tsl[s]], alb[tsl[s]], sc.lights,
                                     int(photon_counts[s]), 4, seed + s, float(self.eps[s]), want_meta=True)
            pos.append(g["pos"])
            base.append(base_powers(g["meta"], sc.lights, g["paths"]))
            dr.append(g["dir"])
            light.append((g["meta"] & 0xFF) + int(self.loff[s]))
            t = g["tris"]
            tris.append(torch.where(t >= 0, t + int(self.toff[s]), t))
        self.ppos, self.pdir = torch.cat(pos).contiguous(), torch.cat(dr).contiguous()
        self.base = torch.cat(base).contiguous()
        self.light = torch.cat(light).to(torch.int32).contiguous()
        self.tris = torch.cat(tris).to(torch.int32).contiguous()
        self.poff = np.zeros(self.S + 1, np.int64)
        self.poff[1:] = np.cumsum([p.shape[0] for p in pos])
        self.albedo0 = alb[:, :3].clone().contiguous()
        L = int(self.loff[-1])
        self.power = api.photons_reweight(self.base, self.light, self.tris, torch.ones((L, 3), device="cuda"),
                                          self.albedo0).clone()
        boxes = np.stack([np.concatenate(sc.bounds()) for sc in scenes]).astype(np.float32)
        self.PB = api.build_points_batch(self.ppos, self.poff.astype(np.uint32), root_boxes=boxes,
                                         max_query_radius=self.rmax)
        api.attach_photon_records_batch(self.PB, self.ppos, self.power, self.pdir)
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
        self.Q = api.QueryBatch(self.TB, self.toff.astype(np.uint32), self.roff.astype(np.uint32), self.eps, self.PB,
                                self.rmax, k)
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
        # the hit triangle of every slot (-1: no shading point), for the colour reweighting
        self.slot_tri = torch.full((R,), -1, dtype=torch.int32, device="cuda")
        self.slot_tri[self.vidx] = self.hit_tri.to(torch.int32)
        self.slot_scene = torch.from_numpy(scene_of[vidx]).cuda()
        self.num_queries = int(vidx.shape[0])

    def _forward(self, emission, albedo):
        """Radiance at every slot (io_rad4) for the given parameters."""
        api.photons_reweight(self.base, self.light, self.tris, emission, albedo, hit_tri=self.slot_tri,
                             io_clr=self.bufs["sp_clr4"], out_pow=self.power)
        api.attach_photon_records_batch(self.PB, self.ppos, self.power, self.pdir)
        self.bufs["io_rad4"].zero_()
        self.Q.run(N.PHASE_KNN)

    def radiance(self, emission, albedo):
        """Radiance at every valid slot (V, 3) for emission multipliers (total lights, 3) and
        albedo (total triangles, 3); no gradient."""
        torch = _torch()
        with torch.no_grad():
            self._forward(emission, albedo)
            return self.bufs["io_rad4"][self.vidx, :3].clone()

    def full_target(self, target):
        """A (V, 3) target at the valid slots as a (slots, 4) array, zero elsewhere."""
        torch = _torch()
        t = torch.zeros_like(self.bufs["io_rad4"])
        t[self.vidx, :3] = target.to(torch.float32)
        return t

    def loss_and_grad(self, emission, albedo, target_full):
        """Loss (a 0-d float64 device tensor) and the gradients (dL/demission, dL/dalbedo),
        for a target from full_target.  No host synchronisation."""
        torch = _torch()
        self._forward(emission, albedo)
        diff = self.bufs["io_rad4"] - target_full  # 0 at slots without a shading point
        loss = torch.square(diff[:, :3].to(torch.float64)).sum()
        torch.mul(diff, 2.0, out=self.bufs["grad_rad4"])
        self.Q.run(N.PHASE_KNN_BWD)
        gp = self.bufs["grad_pow4"][: self.power.shape[0]]
        ge, ga = api.photons_reweight_bwd(self.base, self.light, self.tris, emission, albedo, gp,
                                          hit_tri=self.slot_tri, grad_clr=self.bufs["grad_clr4"])
        return loss, ge, ga

    def fit(self, target, steps: int = 200, lr: float = 1e-2, emission0=None, albedo0=None, fit_albedo=True,
            log=None):
        """Adam from emission multipliers 1 and the trace-time albedo (or the given starts);
        fit_albedo=False holds the albedo.  Returns (emission, albedo, loss per step)."""
        torch = _torch()
        L = int(self.loff[-1])
        e = (emission0.clone() if emission0 is not None else torch.ones((L, 3), device="cuda")).contiguous()
        a = (albedo0.clone() if albedo0 is not None else self.albedo0.clone()).contiguous()
        e.requires_grad_(True)
        a.requires_grad_(True)
        opt = torch.optim.Adam([e, a] if fit_albedo else [e], lr=lr)
        tf = self.full_target(target)
        losses = []
        for it in range(steps):
            loss, ge, ga = self.loss_and_grad(e.detach(), a.detach(), tf)
            e.grad = ge
            a.grad = ga if fit_albedo else None
            opt.step()
            with torch.no_grad():
                e.clamp_(min=0.0)
                a.clamp_(0.0, 1.0)
            losses.append(loss)
            if log and (it % 50 == 0 or it == steps - 1):
                log(f"step {it}: loss {float(loss):.6e}")
        losses = [float(v) for v in torch.stack(losses).cpu()] if losses else []
        return e.detach(), a.detach(), losses

    def scene_arrays(self, s: int):
        """Host copies of scene s for checks, as the last forward left them (radiance() or
        a fit step): photon positions / powers / directions / paths and the scene's shading
        points (pos, nrm, dir, infl, clr) with the GPU radiance at them."""
        torch = _torch()
        torch.cuda.synchronize()
        ps = slice(int(self.poff[s]), int(self.poff[s + 1]))
        cnt = int(self.bufs["sp_count"][s].item())
        qs = slice(int(self.roff[s]), int(self.roff[s]) + cnt)
        h = lambda t: t.detach().cpu().numpy()
        return dict(ppos=h(self.ppos[ps]), ppow=h(self.power[ps]), pdir=h(self.pdir[ps]),
                    qpos=h(self.bufs["sp_pos4"][qs]), qnrm=h(self.bufs["sp_nrm4"][qs]),
                    qdir=h(self.bufs["sp_dir4"][qs]), qinfl=h(self.bufs["sp_infl4"][qs]),
                    qclr=h(self.bufs["sp_clr4"][qs]), rad=h(self.bufs["io_rad4"][qs]),
                    base=h(self.base[ps]), light=h(self.light[ps]), tris=h(self.tris[ps]),
                    k=self.k, r_max=float(self.rmax[s]), root=np.concatenate(self.scenes[s].bounds()))

    def release(self):
        self.Q.close()
        self.TB.release()
        self.PB.release()
