"""Differentiable photon-map radiance estimate: torch.autograd.Function over the C ABI.

forward : mnrt_knn_gather with the photon records rebuilt from the current photon powers
          (positions and the point tree stay fixed).
backward: mnrt_knn_gather_bwd recomputes every query's neighbour set (identical to the
          forward's) and returns dLoss/dpower per photon and dLoss/d(diffuse colour) per
          query; the per-photon sums are order-independent (fixed-point), so repeated
          backward passes give identical gradients.

Chain rules to light emission and per-triangle albedo: `reweight_photons` gives photon
powers as a function of per-light emission and per-triangle albedo with the photon paths
fixed (mnrt_photons_reweight), its backward the parameter gradients as fixed-point sums
(mnrt_photons_reweight_bwd), so the whole chain is deterministic and needs no global
torch setting.
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


class PhotonReweight(torch.autograd.Function):
    """power (P, 4) = base * emission[light] * albedo products along each photon's path."""

    @staticmethod
    def forward(ctx, emission, albedo, base, light, path_tris):
        light, path_tris = api._i32(light), api._i32(path_tris)
        ctx.save_for_backward(emission.detach(), albedo.detach(), base, light, path_tris)
        return api.photons_reweight(base, light, path_tris, emission, albedo)

    @staticmethod
    def backward(ctx, grad_pow):
        emission, albedo, base, light, path_tris = ctx.saved_tensors
        ge, ga = api.photons_reweight_bwd(base, light, path_tris, emission, albedo, grad_pow.contiguous())
        return ge, ga, None, None, None


def reweight_photons(base_power, light, path_tris, emission, albedo):
    """Photon powers as a differentiable function of per-light emission (L, 3) and
    per-triangle albedo (T, 3): base_power times the emission of the photon's light times
    the albedo of every triangle in `path_tris` (P, 4; -1 padded), the triangles the path
    scattered from before the photon."""
    return PhotonReweight.apply(emission, albedo, base_power.contiguous(), light, path_tris)


class HitColour(torch.autograd.Function):
    """Diffuse colours (Q, 4) of hits: rgb = albedo[hit_tri] where hit_tri >= 0, the rest
    (w, and rows without a triangle) from clr_in; dLoss/dalbedo by fixed-point sums."""

    @staticmethod
    def forward(ctx, albedo, hit_tri, clr_in):
        hit_tri = api._i32(hit_tri)
        out = clr_in.detach().clone().contiguous()
        api.hit_colours(albedo, hit_tri, out)
        ctx.save_for_backward(albedo.detach(), hit_tri)
        return out

    @staticmethod
    def backward(ctx, grad_clr):
        albedo, hit_tri = ctx.saved_tensors
        return api.hit_colours_bwd(albedo, hit_tri, grad_clr.contiguous()), None, None


def hit_colours(albedo, hit_tri, clr_in):
    """Differentiable hit colours: albedo (T, 3) gathered at each hit's triangle."""
    return HitColour.apply(albedo, hit_tri, clr_in)

