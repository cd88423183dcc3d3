"""GPU kNN adjoint (mnrt_knn_gather_bwd behind torch.autograd) vs the oracle's fp64
adjoint and vs fp64 central finite differences of the oracle's loss on 1000 random
coordinates; gradients reproducible bit for bit; chain rule to emission and albedo."""

import numpy as np
import pytest

from mnrt_amd import adjoint, api, scene as S

pytestmark = pytest.mark.gpu

K = 50


def _scene(orc, cuda, lights=3):
    sc = S.scene_small(subdiv=4, num_lights=lights)
    eps, rmax = sc.ray_epsilon(), sc.r_max()
    T = api.build_tri(sc.v0, sc.v1, sc.v2)
    ph = api.photons_generate(T, sc.v0, sc.v1, sc.v2, sc.albedo, sc.lights, 30000, 4, 2, eps, want_meta=True)
    o, d, _ = S.camera_rays(96, 96)
    dv = [api.dev(a) for a in (sc.v0, sc.v1, sc.v2, sc.albedo)]
    tri, t, _, _ = api.trace_closest(T, o, d, eps)
    q = api.shading_points(api.dev(o), api.dev(d), tri, t, *dv, want_src=True)
    root = np.concatenate(sc.bounds()).astype(np.float32)
    P = api.build_points(ph["pos"], api.default_params(1, root, rmax))
    cuda.cuda.synchronize()
    return sc, T, ph, q, P, rmax


def _host(t):
    return t.detach().cpu().numpy()


def test_adjoint_vs_oracle_and_fd(orc, cuda):
    torch = cuda
    sc, T, ph, q, P, rmax = _scene(orc, cuda)
    power = ph["power"].clone().requires_grad_(True)
    clr = q["clr"].clone().requires_grad_(True)
    rad = adjoint.knn_radiance(power, clr, P, ph["pos"], ph["dir"], q, K, rmax)
    rng = np.random.default_rng(0)
    target = torch.from_numpy(rng.uniform(0, 2, size=tuple(rad.shape)).astype(np.float32)).cuda() * rad.detach()
    loss = ((rad - target)[:, :3] ** 2).sum()
    loss.backward()
    gp, gc = _host(power.grad), _host(clr.grad)
    # oracle adjoint on the same sets and the same fp32 upstream gradient
    ppos, ppow, pdir = _host(ph["pos"]), _host(ph["power"]), _host(ph["dir"])
    qh = {k: _host(v) for k, v in q.items() if k != "count"}
    op = orc.build(1, ppos, max_qr=rmax, root=np.concatenate(sc.bounds()))
    nbr, cnt, _ = orc.knn(op, ppos, qh["pos"], K, rmax)
    G = (2 * (_host(rad) - _host(target))).astype(np.float32)
    G[:, 3] = 0
    ogp, ogc = orc.radiance_grad(nbr, cnt, K, rmax, ppos, ppow, pdir, qh["pos"], qh["nrm"], qh["dir"], qh["infl"],
                                 qh["clr"], G)
    assert np.abs(gp - ogp).max() <= 1e-5 * np.abs(ogp).max()
    assert np.all(np.abs(gp - ogp) <= 1e-5 * np.abs(ogp) + 1e-7 * np.abs(ogp).max())
    assert np.all(np.abs(gc - ogc) <= 1e-5 * np.abs(ogc) + 1e-7 * np.abs(ogc).max())
    # fp64 central differences on 1000 random coordinates of a linear functional
    # loss = sum W * L: L is linear in each photon power and in each query's colour, so
    # the difference quotient of the oracle's fp64 radiance is the exact gradient up to
    # fp64 rounding, and the GPU's fp32 adjoint (upstream gradient W) must match it
    # positive weights: every gradient entry is a sum of same-sign terms, so its fp32
    # rounding stays relative to the entry (mixed signs can cancel to far below the
    # rounding of the terms)
    W = torch.from_numpy(rng.uniform(0.5, 1.5, size=tuple(rad.shape)).astype(np.float32)).cuda()
    W[:, 3] = 0
    power2 = ph["power"].clone().requires_grad_(True)
    clr2 = q["clr"].clone().requires_grad_(True)
    rad2 = adjoint.knn_radiance(power2, clr2, P, ph["pos"], ph["dir"], q, K, rmax)
    (rad2 * W).sum().backward()
    gp, gc = _host(power2.grad), _host(clr2.grad)
    Wh = _host(W).astype(np.float64)

    def loss64(pw, cl):
        _, o64, _, _, _ = orc.radiance(nbr, cnt, K, rmax, ppos, pw, pdir, qh["pos"], qh["nrm"], qh["dir"],
                                       qh["infl"], cl)
        return float((o64[:, :3] * Wh[:, :3]).sum())

    # the functional is linear in each coordinate, so the step only has to keep the loss
    # change far above the fp64 rounding of the loss: 2^-10 of the coordinate's scale
    # (its own magnitude, or the typical one for a zero such as a light's albedo)
    p_scale = float(np.median(ppow[:, :3]))

    def step(x, typical):
        m = max(abs(float(x)), typical)
        return float(2.0 ** (np.floor(np.log2(m)) - 10))

    used = np.unique(nbr[nbr != 0xFFFFFFFF])
    coords = [("p", int(p), int(c)) for p, c in zip(rng.choice(used, 800), rng.integers(0, 3, 800))]
    coords += [("c", int(i), int(c)) for i, c in zip(rng.choice(qh["pos"].shape[0], 200), rng.integers(0, 3, 200))]
    worst, worst_at = 0.0, None
    scale_p, scale_c = np.abs(gp).max(), np.abs(gc).max()
    for kind, i, c in coords:
        if kind == "p":
            h = step(ppow[i, c], p_scale)
            up, dn = ppow.copy(), ppow.copy()
            up[i, c] += np.float32(h)
            dn[i, c] -= np.float32(h)
            fd = (loss64(up, qh["clr"]) - loss64(dn, qh["clr"])) / (float(up[i, c]) - float(dn[i, c]))
            err = abs(gp[i, c] - fd) / (abs(fd) + 1e-6 * scale_p)
        else:
            h = step(qh["clr"][i, c], 0.5)
            up, dn = qh["clr"].copy(), qh["clr"].copy()
            up[i, c] += np.float32(h)
            dn[i, c] -= np.float32(h)
            fd = (loss64(ppow, up) - loss64(ppow, dn)) / (float(up[i, c]) - float(dn[i, c]))
            err = abs(gc[i, c] - fd) / (abs(fd) + 1e-6 * scale_c)
        if err > worst:
            worst = err
            worst_at = (kind, i, c, float(gp[i, c] if kind == "p" else gc[i, c]), fd)
    print(f"adjoint FD worst relative error over 1000 coordinates: {worst:.2e} at {worst_at}", flush=True)
    assert worst < 1e-5, worst_at


def test_adjoint_is_reproducible(orc, cuda):
    torch = cuda
    sc, T, ph, q, P, rmax = _scene(orc, cuda, lights=1)
    grads = []
    for _ in range(2):
        power = ph["power"].clone().requires_grad_(True)
        rad = adjoint.knn_radiance(power, q["clr"], P, ph["pos"], ph["dir"], q, K, rmax)
        (rad[:, :3] ** 2).sum().backward()
        grads.append(_host(power.grad).view(np.uint32).copy())
    assert np.array_equal(grads[0], grads[1])


def test_chain_to_emission_and_albedo(orc, cuda):
    torch = cuda
    sc, T, ph, q, P, rmax = _scene(orc, cuda)
    L = sc.lights.shape[0]
    emission = torch.ones((L, 3), device="cuda", dtype=torch.float32, requires_grad=True)
    albedo = torch.from_numpy(sc.albedo[:, :3].copy()).cuda().requires_grad_(True)
    light = (ph["meta"] & 0xFF)
    tris = ph["tris"]
    # base power divided by the trace-time albedo products (emission was 1)
    a0 = torch.from_numpy(sc.albedo[:, :3]).cuda()
    valid = tris >= 0
    prod0 = torch.where(valid[..., None], a0[torch.where(valid, tris, 0).long()], torch.ones(1, device="cuda")).prod(1)
    base = ph["power"].clone()
    base[:, :3] = base[:, :3] / prod0
    power = adjoint.reweight_photons(base, light, tris, emission, albedo)
    assert torch.allclose(power, ph["power"], rtol=1e-5)
    rad = adjoint.knn_radiance(power, q["clr"], P, ph["pos"], ph["dir"], q, K, rmax)
    loss = (rad[:, :3] ** 2).sum()
    loss.backward()
    # dLoss/demission_l = sum over photons of light l of dLoss/dpower * power / emission
    pw = power.detach().clone().requires_grad_(True)
    rad2 = adjoint.knn_radiance(pw, q["clr"], P, ph["pos"], ph["dir"], q, K, rmax)
    (rad2[:, :3] ** 2).sum().backward()
    gp = _host(pw.grad)[:, :3].astype(np.float64)
    lh = _host(light)
    ph_pow = _host(power)[:, :3].astype(np.float64)
    for l in range(L):
        ref = (gp[lh == l] * ph_pow[lh == l]).sum(0)
        np.testing.assert_allclose(_host(emission.grad)[l], ref, rtol=1e-4)
    assert np.isfinite(_host(albedo.grad)).all() and np.abs(_host(albedo.grad)).sum() > 0
