"""Differentiable photon-map radiance estimate: torch.autograd.Function over the C ABI.

forward : mnrt_knn_gather with the photon records rebuilt from the current photon powers
          (positions and the point tree stay fixed).
backward: mnrt_knn_gather_bwd recomputes every query's neighbour set (identical to the
          forward's) and returns dLoss/dpower per photon and dLoss/d(diffuse colour) per
          query; the per-photon sums are order-independent (fixed-point), so repeated
          backward passes give identical gradients.

Chain rules to light emission and per-triangle albedo are ordinary differentiable torch
ops on top (photon power = emission of its light x product of the albedos it scattered
from; query colour = albedo of the hit triangle), see `reweight_photons`.
"""

from __future__ import annotations

import torch

from . import api


class KnnRadiance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, power, clr, pmap, ppos, pdir, pos, nrm, dirs, infl, k, r_max):
        api.attach_photon_records(pmap, ppos, power.detach().contiguous(), pdir)
        out = api.knn_gather(pmap, k, r_max, pos, nrm, dirs, infl, clr.detach().contiguous(), want_r2=True,
                             want_count=True)
        ctx.pmap, ctx.k, ctx.r_max = pmap, k, r_max
        ctx.save_for_backward(pos, nrm, dirs, infl, clr.detach(), out["r2"], out["count"])
        return out["rad"]

    @staticmethod
    def backward(ctx, grad_rad):
        pos, nrm, dirs, infl, clr, r2, count = ctx.saved_tensors
        gp, gc = api.knn_gather_bwd(ctx.pmap, ctx.k, ctx.r_max, pos, nrm, dirs, infl, clr, grad_rad, r2, count,
                                    want_grad_clr=ctx.needs_input_grad[1])
        return gp, gc, None, None, None, None, None, None, None, None, None


def knn_radiance(power, clr, pmap, ppos, pdir, q, k, r_max):
    """Radiance (Q, 4) at queries q (dict with pos/nrm/dir/infl) from photon powers (P, 4)."""
    return KnnRadiance.apply(power, clr, pmap, ppos, pdir, q["pos"], q["nrm"], q["dir"], q["infl"], k, r_max)


def reweight_photons(base_power, light, path_tris, emission, albedo):
    """Photon powers as a differentiable function of per-light emission (L, 3) and
    per-triangle albedo (T, 3): base_power / (emission, albedo products at trace time)
    times the current ones.  `path_tris` (P, 4) lists the triangles each path scattered
    from before the photon (-1 padded)."""
    torch.use_deterministic_algorithms(True, warn_only=False)
    e = emission[light.long()]  # (P, 3)
    valid = path_tris >= 0
    idx = torch.where(valid, path_tris, torch.zeros_like(path_tris)).long()
    a = albedo[idx]  # (P, 4, 3)
    a = torch.where(valid[..., None], a, torch.ones_like(a))
    prod = a.prod(dim=1)
    p = base_power[:, :3] * e * prod
    return torch.cat([p, torch.zeros_like(p[:, :1])], dim=1)
