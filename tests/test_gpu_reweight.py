"""GPU photon reweighting (mnrt_photons_reweight / _bwd) vs the oracle's fp64 restatement:
powers within fp32 rounding of the fp64 products, hit colours exact, parameter gradients
within the fp32 / fixed-point rounding bound of the fp64 sums, the same bits on every run;
a non-finite upstream gradient gives NaN gradients; an index out of range gives a NaN
power without any out-of-range read."""

import numpy as np
import pytest

from mnrt_amd import api, inverse, scene as S

pytestmark = pytest.mark.gpu


def _setup(cuda, lights=4, n=60000, seed=9):
    torch = cuda
    sc = S.scene_small(subdiv=3, num_lights=lights)
    eps = sc.ray_epsilon()
    T = api.build_tri(sc.v0, sc.v1, sc.v2)
    ph = api.photons_generate(T, sc.v0, sc.v1, sc.v2, sc.albedo, sc.lights, n, 4, seed, eps, want_meta=True)
    base = inverse.base_powers(ph["meta"], sc.lights, ph["paths"])
    rng = np.random.default_rng(seed)
    e = rng.uniform(0.5, 2.0, (lights, 3)).astype(np.float32)
    alb = rng.uniform(0.1, 0.95, (sc.num_tris, 3)).astype(np.float32)
    T.release()
    return sc, ph, base, (ph["meta"] & 0xFF).contiguous(), ph["tris"].contiguous(), e, alb


def _h(t):
    return t.detach().cpu().numpy()


def test_reweight_forward_vs_oracle(orc, cuda):
    torch = cuda
    sc, ph, base, light, tris, e, alb = _setup(cuda)
    nq = 5000
    rng = np.random.default_rng(1)
    hit = rng.integers(-1, sc.num_tris, nq).astype(np.int32)
    clr = torch.rand((nq, 4), device="cuda")
    clr0 = clr.clone()
    pw = api.photons_reweight(base, light, tris, api.dev(e), api.dev(alb), hit_tri=api.dev(hit), io_clr=clr)
    o = orc.reweight(_h(base), _h(light), _h(tris), e, alb)
    rel = np.abs(_h(pw)[:, :3] - o) / np.abs(o)
    print(f"reweight: {pw.shape[0]} photons, max rel err vs fp64 {rel.max():.2e}", flush=True)
    assert rel.max() < 5e-7
    assert (_h(pw)[:, 3] == 0).all()
    c = _h(clr)
    ok = hit >= 0
    assert np.array_equal(c[ok, :3].view(np.uint32), alb[hit[ok]].view(np.uint32))
    assert np.array_equal(c[:, 3].view(np.uint32), _h(clr0)[:, 3].view(np.uint32))
    assert np.array_equal(c[~ok].view(np.uint32), _h(clr0)[~ok].view(np.uint32))
    # identity parameters: the traced powers within fp32 rounding
    p1 = api.photons_reweight(base, light, tris, torch.ones((e.shape[0], 3), device="cuda"),
                              api.dev(sc.albedo[:, :3].copy()))
    rel1 = np.abs(_h(p1)[:, :3] - _h(ph["power"])[:, :3]) / np.abs(_h(ph["power"])[:, :3])
    assert rel1.max() < 1e-6


def test_reweight_backward_vs_oracle_and_reproducible(orc, cuda):
    torch = cuda
    sc, ph, base, light, tris, e, alb = _setup(cuda, seed=10)
    rng = np.random.default_rng(2)
    P = base.shape[0]
    g = rng.normal(size=(P, 4)).astype(np.float32)
    nq = 20000
    hit = rng.integers(-1, sc.num_tris, nq).astype(np.int32)
    hit[:5000] = 7  # many hits on one triangle: the wave reduction path
    gc = rng.normal(size=(nq, 4)).astype(np.float32)
    args = (base, light, tris, api.dev(e), api.dev(alb), api.dev(g))
    ge, ga = api.photons_reweight_bwd(*args, hit_tri=api.dev(hit), grad_clr=api.dev(gc))
    ge2, ga2 = api.photons_reweight_bwd(*args, hit_tri=api.dev(hit), grad_clr=api.dev(gc))
    assert np.array_equal(_h(ge).view(np.uint32), _h(ge2).view(np.uint32))
    assert np.array_equal(_h(ga).view(np.uint32), _h(ga2).view(np.uint32))
    oe, oa = orc.reweight_grad(_h(base), _h(light), _h(tris), e, alb, g, hit, gc)
    # magnitude of each sum's terms (fp32 rounding of a term is relative to the term)
    me, ma = orc.reweight_grad(_h(base), _h(light), _h(tris), e, alb, np.abs(g), hit, np.abs(gc))
    # fixed-point rounding: at most n_terms per sum, each within half a unit, the unit
    # about n_terms * (largest term) * 2^-62 (largest term <= mag.max())
    for got, want, mag, nt in ((_h(ge), oe, me, P), (_h(ga), oa, ma, 4 * P + nq)):
        err = np.abs(got.astype(np.float64) - want)
        bound = 1e-6 * mag + max(1e-12, float(nt) ** 2 * 2.0 ** -61) * mag.max()
        print(f"reweight bwd: max err / bound {np.max(err / (bound + 1e-300)):.3f}", flush=True)
        assert np.all(err <= bound)
    assert np.abs(oa).max() > 0 and np.abs(oe).max() > 0


def test_reweight_nonfinite_and_out_of_range(cuda):
    torch = cuda
    sc, ph, base, light, tris, e, alb = _setup(cuda, n=8000, seed=11)
    P = base.shape[0]
    g = torch.ones((P, 4), device="cuda")
    g[17, 1] = float("nan")
    ge, ga = api.photons_reweight_bwd(base, light, tris, api.dev(e), api.dev(alb), g)
    assert torch.isnan(ge).all() and torch.isnan(ga).all()
    bad = light.clone()
    bad[5] = e.shape[0]  # one past the last light
    pw = api.photons_reweight(base, bad, tris, api.dev(e), api.dev(alb))
    assert torch.isnan(pw[5, :3]).all()
    assert torch.isfinite(pw[torch.arange(P, device="cuda") != 5, :3]).all()
    ge, ga = api.photons_reweight_bwd(base, bad, tris, api.dev(e), api.dev(alb), torch.ones((P, 4), device="cuda"))
    assert torch.isfinite(ge).all() and torch.isfinite(ga).all()
