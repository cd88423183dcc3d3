"""GPU exact-kNN radiance estimate vs the oracle on the same point tree: neighbour
sets bit-exact (sorted by (d2, photon index)) and equal to the brute force; radiance
bit-exact to the oracle's fp32 path and within 1e-5 relative of its fp64 path."""

import numpy as np
import pytest

from mnrt_amd import api, scene as S
from helpers import hit_queries

pytestmark = pytest.mark.gpu

RTOL64 = 1e-5


def _run(orc, cuda, sc, ppos, ppow, pdir, q, k, rmax, brute=True, clr=True, build_qr=None, root=None):
    root = np.concatenate(sc.bounds()) if root is None else root
    P = orc.build(1, ppos, max_qr=rmax if build_qr is None else build_qr, root=root)
    nbr, cnt, _ = orc.knn(P, ppos, q["pos"], k, rmax)
    if brute:
        bn, bc = orc.knn_brute(ppos, q["pos"], k, rmax)
        assert np.array_equal(nbr, bn) and np.array_equal(cnt, bc)
    o32, o64, i32, i64, r2 = orc.radiance(nbr, cnt, k, rmax, ppos, ppow, pdir, q["pos"], q["nrm"], q["dir"],
                                          q["infl"], q["clr"] if clr else None)
    DP = api.tree_from_host(1, P.nodes, P.refs, P.root, ppos.shape[0], P.max_level)
    api.attach_photon_records(DP, ppos, ppow, pdir)
    out = api.knn_gather(DP, k, rmax, q["pos"], q["nrm"], q["dir"], q["infl"], q["clr"] if clr else None,
                         want_nbr=True, want_r2=True, want_irr=True, want_count=True)
    cuda.cuda.synchronize()
    gn = out["nbr"].cpu().numpy().view(np.uint32)
    gc = out["count"].cpu().numpy().view(np.uint32)
    assert np.array_equal(gc, cnt), f"{(gc != cnt).sum()} counts differ"
    assert np.array_equal(gn, nbr), f"{(gn != nbr).any(axis=1).sum()} neighbour sets differ"
    grad = out["rad"].cpu().numpy()
    assert np.array_equal(grad.view(np.uint32), o32.view(np.uint32)), "radiance not bit-exact to fp32 oracle"
    assert np.array_equal(out["r2"].cpu().numpy().view(np.uint32), r2.view(np.uint32))
    scale = np.abs(o64).max(axis=1, keepdims=True) + 1e-30
    err = np.abs(grad.astype(np.float64) - o64) / scale
    assert err.max() <= RTOL64, f"max rel err vs fp64 {err.max():.3e}"
    # the list-free path (no final sort: the selection alone decides the set) gives the same
    # radiance, r2 and counts
    free = api.knn_gather(DP, k, rmax, q["pos"], q["nrm"], q["dir"], q["infl"], q["clr"] if clr else None,
                          want_r2=True, want_count=True)
    cuda.cuda.synchronize()
    assert np.array_equal(free["count"].cpu().numpy().view(np.uint32), cnt), "list-free counts differ"
    assert np.array_equal(free["r2"].cpu().numpy().view(np.uint32), r2.view(np.uint32)), "list-free r2 differs"
    assert np.array_equal(free["rad"].cpu().numpy().view(np.uint32), o32.view(np.uint32)), \
        "list-free radiance not bit-exact to fp32 oracle"
    return gn, gc


def test_knn_small_scene_k50(orc, cuda):
    sc = S.scene_small(subdiv=5)
    ppos, ppow, pdir = S.random_surface_photons(sc, 65536, seed=0)
    o, d, _ = S.camera_rays(128, 128)
    q = hit_queries(orc, sc, o, d)
    print(f"knn: {len(q['pos'])} queries k=50", flush=True)
    _run(orc, cuda, sc, ppos, ppow, pdir, q, 50, sc.r_max())


@pytest.mark.parametrize("k", [1, 8, 64, 100, 200, 256])
def test_knn_k_sweep(orc, cuda, k):
    sc = S.scene_small(subdiv=3)
    ppos, ppow, pdir = S.random_surface_photons(sc, 20000, seed=k)
    o, d, _ = S.camera_rays(48, 48)
    q = hit_queries(orc, sc, o, d)
    print(f"knn: k={k}", flush=True)
    _run(orc, cuda, sc, ppos, ppow, pdir, q, k, sc.r_max())


def test_knn_small_radius_and_few_photons(orc, cuda):
    # many queries find fewer than k (and fewer than 8) photons: r2 = r_max^2, E = 0
    sc = S.scene_small(subdiv=2)
    ppos, ppow, pdir = S.random_surface_photons(sc, 500, seed=3)
    o, d, _ = S.camera_rays(40, 40)
    q = hit_queries(orc, sc, o, d)
    _, gc = _run(orc, cuda, sc, ppos, ppow, pdir, q, 50, 0.05)
    assert (gc < 8).any() and (gc > 0).any()
    _run(orc, cuda, sc, ppos, ppow, pdir, q, 50, 0.05, clr=False)


def test_knn_coincident_photons_tie_by_index(orc, cuda):
    # clusters of identical photon positions: equal d2, order and cut by photon index
    sc = S.scene_small(subdiv=2)
    ppos, ppow, pdir = S.random_surface_photons(sc, 4000, seed=4)
    ppos[1000:3000] = ppos[np.repeat(np.arange(100), 20)]
    o, d, _ = S.camera_rays(32, 32)
    q = hit_queries(orc, sc, o, d)
    q["pos"][:50] = ppos[:50]  # queries exactly on photons (d2 = 0)
    _run(orc, cuda, sc, ppos, ppow, pdir, q, 30, sc.r_max())


def test_knn_unbounded_radius_far_queries_and_offset_scene(orc, cuda):
    """r_max = inf (every photon is a candidate), queries far outside the photon map (large
    d2: the window and final selections search high exponents), and the scene moved to
    coordinates near 1e4 (where the fp32 photon distances and the kernel's box distances,
    summed in another order, round most): sets, radii and radiance equal the oracle's."""
    sc = S.scene_small(subdiv=2)
    ppos, ppow, pdir = S.random_surface_photons(sc, 3000, seed=21)
    o, d, _ = S.camera_rays(32, 32)
    q = hit_queries(orc, sc, o, d)
    print("knn: r_max = inf", flush=True)
    _run(orc, cuda, sc, ppos, ppow, pdir, q, 40, float("inf"), build_qr=sc.r_max())
    far = dict(q)
    far["pos"] = q["pos"].copy()
    far["pos"][:, :3] += np.float32(40.0)
    print("knn: far queries", flush=True)
    _run(orc, cuda, sc, ppos, ppow, pdir, far, 40, float("inf"), build_qr=sc.r_max())
    shift = np.float32(1.0e4)
    pp = ppos.copy()
    pp[:, :3] += shift
    qs = dict(q)
    qs["pos"] = q["pos"].copy()
    qs["pos"][:, :3] += shift
    root = (np.concatenate(sc.bounds()) + shift).astype(np.float32)
    print("knn: scene near 1e4", flush=True)
    _run(orc, cuda, sc, pp, ppow, pdir, qs, 40, sc.r_max(), root=root)


def test_knn_per_query_radius_and_inverse_distance(orc, cuda):
    """Per-query radius caps equal separate calls at each cap (the oracle pins those);
    the inverse-distance sums equal fp32 1/|x - o| in the same operation order, and with
    shared output slots both sums run in query order."""
    torch = cuda
    sc = S.scene_small(subdiv=3)
    ppos, ppow, pdir = S.random_surface_photons(sc, 30000, seed=4)
    o, d, _ = S.camera_rays(48, 48)
    q = hit_queries(orc, sc, o, d)
    n = q["pos"].shape[0]
    P = orc.build(1, ppos, max_qr=sc.r_max(), root=np.concatenate(sc.bounds()))
    DP = api.tree_from_host(1, P.nodes, P.refs, P.root, ppos.shape[0], P.max_level)
    api.attach_photon_records(DP, ppos, ppow, pdir)
    caps = np.where(np.arange(n) % 3 == 0, 0.02, 0.15).astype(np.float32)
    args = (q["pos"], q["nrm"], q["dir"], q["infl"], q["clr"])
    orig = q["pos"].copy()
    orig[:, :3] -= q["dir"][:, :3] * 0.5
    inv = torch.zeros(n, dtype=torch.float32, device="cuda")
    per = api.knn_gather(DP, 50, 0.15, *args, query_radius=caps, orig=orig, io_inv_dist=inv, want_nbr=True)
    cuda.cuda.synchronize()
    for cap in (0.02, 0.15):
        sel = np.nonzero(caps == np.float32(cap))[0]
        one = api.knn_gather(DP, 50, float(cap), *(a[sel] for a in args), want_nbr=True)
        nbr, _, _ = orc.knn(P, ppos, q["pos"][sel], 50, float(cap))
        assert np.array_equal(one["nbr"].cpu().numpy().view(np.uint32), nbr)
        assert torch.equal(per["nbr"][sel], one["nbr"])
        assert torch.equal(per["rad"][sel], one["rad"])
    od = (q["pos"][:, :3] - orig[:, :3]).astype(np.float32)
    d2 = (od[:, 0] * od[:, 0] + od[:, 1] * od[:, 1]) + od[:, 2] * od[:, 2]
    want = np.float32(1.0) / np.sqrt(d2)
    assert np.array_equal(inv.cpu().numpy().view(np.uint32), want.view(np.uint32))
    # shared slots (sorted, as final gathering produces them): run sums in query order
    slots = (np.arange(n) // 4).astype(np.uint32)
    rad = torch.zeros((n // 4 + 1, 4), dtype=torch.float32, device="cuda")
    inv2 = torch.zeros(n // 4 + 1, dtype=torch.float32, device="cuda")
    api.knn_gather(DP, 50, 0.15, *args, idx_sp=slots, io_rad=rad, query_radius=caps, orig=orig, io_inv_dist=inv2)
    cuda.cuda.synchronize()
    pr = per["rad"].cpu().numpy()
    er = np.zeros((n // 4 + 1, 4), np.float32)
    ei = np.zeros(n // 4 + 1, np.float32)
    for s in range(n // 4 + 1):
        acc = np.zeros(3, np.float32)
        ai = np.float32(0)
        for j in np.nonzero(slots == s)[0]:
            acc = acc + pr[j, :3]
            ai = ai + want[j]
        er[s, :3] = np.float32(0) + acc
        ei[s] = np.float32(0) + ai
    assert np.array_equal(rad.cpu().numpy().view(np.uint32), er.view(np.uint32))
    assert np.array_equal(inv2.cpu().numpy().view(np.uint32), ei.view(np.uint32))


def test_knn_many_incoherent_queries(orc, cuda):
    # more queries than the reordering threshold, in random order (as hits of diffuse
    # bounce rays come): the launch processes them grouped by spatial cell; every set,
    # radius and radiance is still bit-exact to the oracle, and radii agree
    sc = S.scene_small(subdiv=4)
    ppos, ppow, pdir = S.random_surface_photons(sc, 65536, seed=11)
    qp, _, qd = S.random_surface_photons(sc, 70000, seed=12)
    rng = np.random.default_rng(13)
    nrm = qd.copy()
    qdir = qd.copy()
    qdir[:, :3] = -qd[:, :3]
    infl = np.ones_like(qp)
    infl[:, 3] = 0
    clr = np.tile(np.array([0.7, 0.6, 0.5, 1.0], np.float32), (len(qp), 1))
    q = dict(pos=qp, nrm=nrm, dir=qdir, infl=infl, clr=clr)
    print(f"knn: {len(qp)} incoherent queries k=100", flush=True)
    _run(orc, cuda, sc, ppos, ppow, pdir, q, 100, sc.r_max(), brute=False)
    # radii of the same queries through mnrt_knn_radii
    P = orc.build(1, ppos, max_qr=sc.r_max(), root=np.concatenate(sc.bounds()))
    DP = api.tree_from_host(1, P.nodes, P.refs, P.root, ppos.shape[0], P.max_level)
    api.attach_photon_records(DP, ppos, ppow, pdir)
    sel = rng.permutation(len(qp))
    r = api.knn_radii(DP, 100, sc.r_max(), qp[sel]).cpu().numpy()
    nbr, cnt, _ = orc.knn(P, ppos, qp[sel], 100, sc.r_max())
    full = cnt >= 100
    d = ppos[nbr[full, 99], :3] - qp[sel][full, :3]
    d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    assert np.array_equal(r[full], np.sqrt(d2.astype(np.float32)))
    assert np.all(r[~full] == np.float32(sc.r_max()))
