"""Oracle photon reweighting pins: emission 1 and the trace-time albedo give back the
powers the photon tracer stored (an independent code path: the tracer multiplies the
power bounce by bounce); the fp64 gradient equals central differences of the forward and
a plain loop over every (photon, path position) and hit; a planted dropped term fails."""

import numpy as np

from mnrt_amd import scene as S


def _traced(orc, lights=3, n=6000, seed=5):
    sc = S.scene_small(subdiv=2, num_lights=lights)
    T = orc.build(0, sc.v0, sc.v1, sc.v2)
    ph = orc.photons(T, sc.v0, sc.v1, sc.v2, sc.albedo, sc.lights, n, 4, seed, sc.ray_epsilon())
    return sc, ph


def test_identity_parameters_give_traced_powers(orc):
    sc, ph = _traced(orc)
    base = orc.base_powers(ph["meta"], sc.lights, ph["paths"])
    light = ph["meta"] & 0xFF
    p = orc.reweight(base, light, ph["tris"], np.ones((sc.lights.shape[0], 3)), sc.albedo[:, :3])
    traced = ph["power"][:, :3].astype(np.float64)
    rel = np.abs(p - traced) / np.abs(traced)
    assert rel.max() < 4e-7, rel.max()  # the tracer rounds to fp32 once per bounce
    # a different emission scales each light's photons exactly
    e = np.array([[2.0, 0.5, 1.0], [1.0, 1.0, 3.0], [0.25, 4.0, 1.0]])
    p2 = orc.reweight(base, light, ph["tris"], e, sc.albedo[:, :3])
    np.testing.assert_allclose(p2, p * e[light], rtol=1e-15)


def _loop_grad(base, light, tris, e, alb, g, hit, gc):
    ge = np.zeros_like(e)
    ga = np.zeros_like(alb)
    for p in range(base.shape[0]):
        ts = [int(t) for t in tris[p] if t != 0xFFFFFFFF]
        for c in range(3):
            prod = 1.0
            for t in ts:
                prod *= alb[t, c]
            ge[light[p], c] += g[p, c] * base[p, c] * prod
            for k, t in enumerate(ts):
                o = base[p, c] * e[light[p], c]
                for j, u in enumerate(ts):
                    if j != k:
                        o *= alb[u, c]
                ga[t, c] += g[p, c] * o
    for q in range(hit.shape[0]):
        if hit[q] >= 0:
            ga[hit[q]] += gc[q, :3]
    return ge, ga


def test_gradient_loop_and_central_differences(orc):
    sc, ph = _traced(orc, n=1500, seed=6)
    rng = np.random.default_rng(0)
    L, T = sc.lights.shape[0], sc.num_tris
    base = orc.base_powers(ph["meta"], sc.lights, ph["paths"])
    light = (ph["meta"] & 0xFF).astype(np.int64)
    e = rng.uniform(0.5, 2.0, (L, 3))
    alb = rng.uniform(0.2, 0.9, (T, 3))
    g = rng.normal(size=(base.shape[0], 3))
    hit = rng.integers(-1, T, 400)
    gc = rng.normal(size=(400, 3))
    ge, ga = orc.reweight_grad(base, light, ph["tris"], e, alb, g, hit, gc)
    le, la = _loop_grad(base, light, ph["tris"], e, alb, g, hit, gc)
    np.testing.assert_allclose(ge, le, rtol=1e-12, atol=1e-14 * np.abs(le).max())
    np.testing.assert_allclose(ga, la, rtol=1e-12, atol=1e-14 * np.abs(la).max())

    def loss(e_, a_):
        return float((g * orc.reweight(base, light, ph["tris"], e_, a_)).sum())

    # central differences (the loss is a polynomial of degree <= 4 in one albedo entry and
    # linear in one emission entry: h = 1e-4 leaves a relative error ~1e-8)
    h = 1e-4
    for l, c in [(0, 0), (1, 2), (2, 1)]:
        up, dn = e.copy(), e.copy()
        up[l, c] += h
        dn[l, c] -= h
        fd = (loss(up, alb) - loss(dn, alb)) / (2 * h)
        assert abs(fd - ge[l, c]) <= 1e-7 * max(1.0, abs(fd))
    used = np.unique(ph["tris"][ph["tris"] != 0xFFFFFFFF])
    for t in rng.choice(used, 6, replace=False):
        c = int(rng.integers(0, 3))
        up, dn = alb.copy(), alb.copy()
        up[t, c] += h
        dn[t, c] -= h
        fd = (loss(e, up) - loss(e, dn)) / (2 * h)
        photon_part = ga[t, c] - gc[hit == t, c].sum()
        assert abs(fd - photon_part) <= 1e-6 * max(1.0, abs(fd)), (t, c, fd, photon_part)


def test_planted_dropped_term_is_caught(orc):
    """A gradient that forgets the repeated-triangle term (a path scattering twice from one
    triangle contributes 2 a) differs from the loop."""
    sc, ph = _traced(orc, n=3000, seed=7)
    tris = ph["tris"]
    base = orc.base_powers(ph["meta"], sc.lights, ph["paths"])
    light = (ph["meta"] & 0xFF).astype(np.int64)
    e = np.ones((sc.lights.shape[0], 3))
    alb = np.full((sc.num_tris, 3), 0.7)
    g = np.ones((base.shape[0], 3))
    _, ga = orc.reweight_grad(base, light, tris, e, alb, g)
    rep = [p for p in range(tris.shape[0]) if len({int(t) for t in tris[p] if t != 0xFFFFFFFF})
           < sum(1 for t in tris[p] if t != 0xFFFFFFFF)]
    assert rep, "the scene should have paths that hit one triangle twice"
    wrong = np.zeros_like(ga)
    for p in range(tris.shape[0]):
        ts = sorted({int(t) for t in tris[p] if t != 0xFFFFFFFF})
        n = sum(1 for t in tris[p] if t != 0xFFFFFFFF)
        for t in ts:  # counts each triangle once: d(a^2)/da taken as a
            wrong[t] += g[p] * base[p] * 0.7 ** (n - 1)
    assert not np.allclose(wrong, ga, rtol=1e-9)
