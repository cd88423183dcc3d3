"""Oracle adjoint of the radiance estimate vs fp64 central finite differences of the
fp64 loss computed by the forward estimator (sets fixed from the forward pass)."""

import numpy as np

from mnrt_amd import scene as S


def _setup(orc, nq=600, k=30, seed=0):
    sc = S.scene_small(subdiv=3)
    ppos, ppow, pdir = S.random_surface_photons(sc, 4000, seed=seed)
    rng = np.random.default_rng(seed)
    q = ppos[rng.choice(4000, nq, replace=False)].copy()
    nrm = np.zeros_like(q)
    nrm[:, :3] = rng.normal(size=(nq, 3))
    nrm[:, :3] /= np.linalg.norm(nrm[:, :3], axis=1, keepdims=True)
    qd = np.zeros_like(q)
    qd[:, :3] = -nrm[:, :3] + 0.3 * rng.normal(size=(nq, 3))
    infl = np.ones_like(q)
    clr = rng.uniform(0.2, 0.9, size=(nq, 4)).astype(np.float32)
    clr[:, 3] = 1.0
    rmax = sc.r_max()
    T = orc.build(1, ppos, max_qr=rmax, root=np.concatenate(sc.bounds()))
    nbr, cnt, _ = orc.knn(T, ppos, q, k, rmax)
    return dict(sc=sc, ppos=ppos, ppow=ppow, pdir=pdir, q=q, nrm=nrm, qd=qd, infl=infl, clr=clr, rmax=rmax, k=k,
                nbr=nbr, cnt=cnt, target=rng.uniform(0, 2e-4, size=(nq, 4)))


def _loss64(orc, d, ppow, clr):
    _, o64, _, _, _ = orc.radiance(d["nbr"], d["cnt"], d["k"], d["rmax"], d["ppos"], ppow, d["pdir"], d["q"],
                                   d["nrm"], d["qd"], d["infl"], clr)
    return float(((o64[:, :3] - d["target"][:, :3]) ** 2).sum()), o64


def _step(x):
    """A float32-exact perturbation: 2^-10 of the value's binade."""
    e = np.floor(np.log2(abs(float(x)))) if x != 0 else -30
    return float(2.0 ** (e - 10))


def test_adjoint_matches_finite_differences(orc):
    d = _setup(orc)
    _, o64 = _loss64(orc, d, d["ppow"], d["clr"])
    G = (2 * (o64 - d["target"])).astype(np.float32)
    G[:, 3] = 0
    gp, gc = orc.radiance_grad(d["nbr"], d["cnt"], d["k"], d["rmax"], d["ppos"], d["ppow"], d["pdir"], d["q"],
                               d["nrm"], d["qd"], d["infl"], d["clr"], G)
    rng = np.random.default_rng(1)
    used = np.unique(d["nbr"][d["nbr"] != 0xFFFFFFFF])
    checked = 0
    for p in rng.choice(used, 60, replace=False):
        for c in range(3):
            h = _step(d["ppow"][p, c])
            up, dn = d["ppow"].copy(), d["ppow"].copy()
            up[p, c] += np.float32(h)
            dn[p, c] -= np.float32(h)
            fd = (_loss64(orc, d, up, d["clr"])[0] - _loss64(orc, d, dn, d["clr"])[0]) / (2 * h)
            # G is the fp32 rounding of the fp64 upstream gradient
            assert abs(gp[p, c] - fd) <= 1e-6 * max(abs(fd), 1e-12) + 1e-9 * np.abs(gp).max(), (p, c, gp[p, c], fd)
            checked += 1
    assert checked == 180
    for qi in rng.choice(d["q"].shape[0], 40, replace=False):
        for c in range(3):
            h = _step(d["clr"][qi, c])
            up, dn = d["clr"].copy(), d["clr"].copy()
            up[qi, c] += np.float32(h)
            dn[qi, c] -= np.float32(h)
            fd = (_loss64(orc, d, d["ppow"], up)[0] - _loss64(orc, d, d["ppow"], dn)[0]) / (2 * h)
            assert abs(gc[qi, c] - fd) <= 1e-6 * max(abs(fd), 1e-12) + 1e-9 * np.abs(gc).max()


def test_adjoint_back_facing_and_gated_photons_get_zero(orc):
    d = _setup(orc, nq=200, k=5)  # k < 8: every query is gated, no photon gradient
    G = np.ones((200, 4), np.float32)
    gp, gc = orc.radiance_grad(d["nbr"], d["cnt"], d["k"], d["rmax"], d["ppos"], d["ppow"], d["pdir"], d["q"],
                               d["nrm"], d["qd"], d["infl"], d["clr"], G)
    assert (gp == 0).all() and (gc == 0).all()
