"""Batched inverse lighting (emission and albedo of several scenes at once through the
batched kNN forward and adjoint launches): the gradient of the total loss equals the sum
of the single-scene adjoints, matches central differences of the loss, and Adam recovers
the target emission."""

import numpy as np
import pytest

from mnrt_amd import adjoint, api, scene as S
from mnrt_amd.inverse import BatchedInverseLighting

pytestmark = pytest.mark.gpu

K = 50


def _problem(cuda, sizes=(2000, 6000, 900), photons=(8000, 20000, 5000), dims=((32, 32), (40, 30), (24, 24))):
    scs = [S.scene_sized(n, seed=200 + i, num_lights=4) for i, n in enumerate(sizes)]
    print(f"inverse problem: {len(scs)} scenes, {sum(s.num_tris for s in scs)} triangles", flush=True)
    return scs, BatchedInverseLighting(scs, photons, dims, k=K, seed=3)


def test_gradient_equals_single_scene_adjoints(cuda):
    torch = cuda
    scs, P = _problem(cuda)
    L = int(P.loff[-1])
    g = torch.Generator(device="cpu").manual_seed(1)
    e = (0.5 + 1.5 * torch.rand((L, 3), generator=g)).cuda().requires_grad_(True)
    a = (P.albedo0 * 0.9).requires_grad_(True)
    target = P.radiance(torch.ones((L, 3), device="cuda"), P.albedo0)
    loss, ge, ga = P.loss_and_grad(e.detach(), a.detach(), P.full_target(target))
    loss = float(loss)
    # the same loss scene by scene through the single-scene autograd Function
    e2 = e.detach().clone().requires_grad_(True)
    a2 = a.detach().clone().requires_grad_(True)
    power = adjoint.reweight_photons(P.base, P.light, P.tris, e2, a2)
    clr_all = P.bufs["sp_clr4"].clone()
    total = torch.zeros((), dtype=torch.float64, device="cuda")
    cnt = P.bufs["sp_count"].cpu().numpy()
    vpos, pms = 0, []
    for s in range(P.S):
        a0, m = int(P.roff[s]), int(cnt[s])
        sl = slice(a0, a0 + m)
        psl = slice(int(P.poff[s]), int(P.poff[s + 1]))
        clr = torch.cat([a2[P.hit_tri[vpos:vpos + m]], clr_all[sl, 3:]], dim=1)
        q = {n: P.bufs[b][sl].contiguous() for n, b in (("pos", "sp_pos4"), ("nrm", "sp_nrm4"),
                                                        ("dir", "sp_dir4"), ("infl", "sp_infl4"))}
        rmax = float(scs[s].r_max())
        pm = api.build_points(P.ppos[psl], api.default_params(1, np.concatenate(scs[s].bounds()).astype(np.float32),
                                                              rmax))
        rad = adjoint.knn_radiance(power[psl], clr, pm, P.ppos[psl], P.pdir[psl], q, K, rmax)
        pms.append(pm)
        total = total + ((rad[:, :3] - target[vpos:vpos + m]).double() ** 2).sum()
        vpos += m
    total.backward()
    for pm in pms:
        pm.release()
    tot = float(total.detach())
    print(f"loss {loss:.6e} vs {tot:.6e}", flush=True)
    assert abs(loss - tot) <= 1e-6 * abs(tot)
    torch.testing.assert_close(ge, e2.grad, rtol=1e-4, atol=1e-6 * float(e2.grad.abs().max()))
    torch.testing.assert_close(ga, a2.grad, rtol=1e-4, atol=1e-6 * float(a2.grad.abs().max()))
    assert float(ge.abs().sum()) > 0 and float(ga.abs().sum()) > 0
    # the same bits on a second evaluation (fixed-point sums, no global torch setting)
    assert not torch.are_deterministic_algorithms_enabled()
    _, ge2, ga2 = P.loss_and_grad(e.detach(), a.detach(), P.full_target(target))
    assert torch.equal(ge.view(torch.int32), ge2.view(torch.int32))
    assert torch.equal(ga.view(torch.int32), ga2.view(torch.int32))
    P.release()


def test_gradient_matches_central_differences(cuda):
    """With the neighbour sets fixed the radiance is linear in each light's emission, so the
    loss is quadratic along an emission direction and central differences are exact up to
    fp32 rounding of the radiance."""
    torch = cuda
    scs, P = _problem(cuda)
    L = int(P.loff[-1])
    g = torch.Generator(device="cpu").manual_seed(2)
    target = P.radiance((0.5 + 1.5 * torch.rand((L, 3), generator=g)).cuda(), P.albedo0)
    e = torch.ones((L, 3), device="cuda")
    _, ge, _ = P.loss_and_grad(e, P.albedo0, P.full_target(target))
    v = torch.randn((L, 3), generator=g).cuda()
    h = 0.05

    def loss_at(x):
        r = P.radiance(x, P.albedo0)
        return float(((r - target).double() ** 2).sum())

    fd = (loss_at(e + h * v) - loss_at(e - h * v)) / (2 * h)
    an = float((ge.double() * v.double()).sum())
    print(f"directional derivative {an:.6e}, central difference {fd:.6e}", flush=True)
    assert abs(an - fd) <= 2e-3 * abs(fd)
    P.release()


def test_adam_recovers_emission(cuda):
    torch = cuda
    scs, P = _problem(cuda)
    L = int(P.loff[-1])
    g = torch.Generator(device="cpu").manual_seed(4)
    e_true = (0.5 + 1.5 * torch.rand((L, 3), generator=g)).cuda()
    target = P.radiance(e_true, P.albedo0)
    e, a, losses = P.fit(target, steps=120, lr=5e-2, fit_albedo=False, log=lambda s: print(s, flush=True))
    err0 = float((1.0 - e_true).abs().mean())
    err = float((e - e_true).abs().mean())
    print(f"loss {losses[0]:.4e} -> {losses[-1]:.4e}; mean emission error {err0:.4f} -> {err:.4f}", flush=True)
    assert losses[-1] < 0.05 * losses[0]
    assert err < 0.5 * err0
    P.release()


def test_adam_on_emission_and_albedo_lr_1e2(cuda):
    """The batched inverse-lighting settings (Adam lr 1e-2, 200 steps, emission and albedo
    both free) on a small batch: the loss falls and the emission moves toward the target."""
    torch = cuda
    scs, P = _problem(cuda)
    L = int(P.loff[-1])
    g = torch.Generator(device="cpu").manual_seed(6)
    e_true = (0.5 + 1.5 * torch.rand((L, 3), generator=g)).cuda()
    target = P.radiance(e_true, P.albedo0)
    e, a, losses = P.fit(target, steps=200, lr=1e-2, fit_albedo=True)
    err0 = float((1.0 - e_true).abs().mean())
    err = float((e - e_true).abs().mean())
    print(f"loss {losses[0]:.4e} -> {losses[-1]:.4e}; emission error {err0:.4f} -> {err:.4f}", flush=True)
    assert len(losses) == 200 and all(np.isfinite(losses))
    assert losses[-1] < 0.05 * losses[0]
    assert err < 0.7 * err0
    assert bool(((a >= 0) & (a <= 1)).all())
    P.release()
