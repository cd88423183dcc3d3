"""GPU kd-tree build vs the oracle build: the packed preorder node words, leaf
references, root box, level and leaf counts are identical; the SAH/VVH cost therefore
equals the oracle's (the bar is 0.5 %)."""

import numpy as np
import pytest

from mnrt_amd import api, scene as S
from helpers import f4

pytestmark = pytest.mark.gpu


def _cmp(orc, kind, dt, ot, max_qr=1.0):
    gn, gr = dt.nodes_host(), dt.refs_host()
    assert dt.num_nodes == ot.num_nodes, (dt.num_nodes, ot.num_nodes)
    assert dt.num_refs == ot.refs.size, (dt.num_refs, ot.refs.size)
    assert np.array_equal(dt.root.view(np.uint32), ot.root.view(np.uint32)), (dt.root, ot.root)
    if not np.array_equal(gn, ot.nodes):
        bad = np.nonzero(gn != ot.nodes)[0]
        raise AssertionError(f"{bad.size} node words differ, first at word {bad[0]}")
    assert np.array_equal(gr, ot.refs), "leaf references differ"
    assert dt.max_level == ot.max_level
    assert dt.t.num_leaves == ot.num_leaves
    gc = dt.cost(max_qr)
    oc = orc.tree_cost(ot.nodes, ot.root, kind, max_qr)
    assert abs(gc - oc) <= 0.005 * oc
    assert orc.check_structure(gn, gr, dt.t.num_elems) == 0


def _tri(orc, cuda, sc, root=None):
    print(f"build tri: {sc.num_tris}", flush=True)
    params = api.default_params(0, root)
    dt = api.build_tri(sc.v0, sc.v1, sc.v2, params)
    cuda.cuda.synchronize()
    ot = orc.build(0, sc.v0, sc.v1, sc.v2, root=root)
    _cmp(orc, 0, dt, ot)
    return dt, ot


def _scene(tris):
    tris = np.asarray(tris, np.float32)
    return S.Scene(f4(tris[:, 0]), f4(tris[:, 1]), f4(tris[:, 2]), np.ones((len(tris), 4), np.float32),
                   np.zeros((0, 4, 3), np.float32))


def test_build_tri_small_scene(orc, cuda):
    _tri(orc, cuda, S.scene_small(subdiv=5))


def test_build_tri_tiny_and_single_small_root(orc, cuda):
    _tri(orc, cuda, S.scene_small(subdiv=0))  # 34 triangles: the root is a small root
    _tri(orc, cuda, _scene(S.box_tris()[:1]))


def test_build_tri_random_soup_with_root_box(orc, cuda):
    sc = S.scene_sized(30000, seed=9)
    _tri(orc, cuda, sc, root=np.array([-1.5, -1.5, -1.5, 1.5, 1.5, 1.5], np.float32))


def test_build_tri_degenerate(orc, cuda):
    # 200 coincident triangles (no-progress leaf), planar triangles lying in split
    # planes, zero-area triangles, axis-aligned walls
    base = S.icosphere(1, 0.4)
    same = np.repeat(S.icosphere(0, 0.3)[:1], 200, axis=0)
    planar = np.array([[[x, 0.0, -0.5], [x + 0.1, 0.0, -0.4], [x, 0.0, -0.3]] for x in np.linspace(-0.9, 0.9, 90)])
    zero = base[:20].copy()
    zero[:, 2] = zero[:, 1]
    sc = _scene(np.concatenate([S.box_tris(), base, same, planar, zero]))
    _tri(orc, cuda, sc)


def test_build_tri_large_overlapping_triangles(orc, cuda):
    # triangles spanning most of the scene: nearly every element straddles every split, so
    # the clip queue holds up to two jobs per element, and no-progress nodes (whose queued
    # sides are skipped) are common; mixed with small triangles that rarely straddle
    rng = np.random.default_rng(17)
    big = rng.uniform(-1.0, 1.0, size=(2500, 3, 3))
    small = rng.uniform(-1.0, 1.0, size=(4000, 1, 3)) + rng.normal(scale=0.01, size=(4000, 3, 3))
    _tri(orc, cuda, _scene(np.concatenate([big, small])))


def test_build_points_vvh(orc, cuda):
    sc = S.scene_small(subdiv=3)
    ppos, _, _ = S.random_surface_photons(sc, 50000, seed=2)
    root = np.concatenate(sc.bounds()).astype(np.float32)
    rmax = sc.r_max()
    print("build points", flush=True)
    params = api.default_params(1, root, rmax)
    dt = api.build_points(ppos, params)
    cuda.cuda.synchronize()
    ot = orc.build(1, ppos, max_qr=rmax, root=root)
    _cmp(orc, 1, dt, ot, rmax)


def test_build_points_coincident(orc, cuda):
    # 300 photons at one position + a cloud: level cap / parity split of on-plane points
    rng = np.random.default_rng(3)
    p = rng.uniform(-1, 1, size=(5000, 3))
    p[:300] = [0.25, -0.125, 0.5]
    pos = f4(p)
    root = np.array([-1, -1, -1, 1, 1, 1], np.float32)
    dt = api.build_points(pos, api.default_params(1, root, 0.1))
    cuda.cuda.synchronize()
    ot = orc.build(1, pos, max_qr=0.1, root=root)
    _cmp(orc, 1, dt, ot, 0.1)


def test_build_empty_input(orc, cuda):
    v = np.zeros((0, 4), np.float32)
    dt = api.build_tri(v, v, v)
    cuda.cuda.synchronize()
    assert dt.num_nodes == 1 and dt.num_refs == 0


@pytest.mark.parametrize("cap", ["0,0", "1,2", "2,8"])
def test_build_small_stage_fallback(orc, cuda, monkeypatch, cap):
    # small subtrees that outgrow their staging area (forced by a tiny capacity: none,
    # some or most fit) are emitted by the second pass: the trees stay identical
    monkeypatch.setenv("MNRT_SMALL_STAGE_CAP", cap)
    print(f"build with small staging capacity {cap}", flush=True)
    sc = S.scene_small(subdiv=4)
    _tri(orc, cuda, sc)
    ppos, _, _ = S.random_surface_photons(sc, 20000, seed=4)
    root = np.concatenate(sc.bounds()).astype(np.float32)
    dt = api.build_points(ppos, api.default_params(1, root, sc.r_max()))
    cuda.cuda.synchronize()
    ot = orc.build(1, ppos, max_qr=sc.r_max(), root=root)
    _cmp(orc, 1, dt, ot, sc.r_max())
    # a batch of scenes through the same stage
    scenes = [S.scene_sized(n, seed=30 + i) for i, n in enumerate((400, 3000, 90))]
    v0 = np.concatenate([s.v0 for s in scenes])
    v1 = np.concatenate([s.v1 for s in scenes])
    v2 = np.concatenate([s.v2 for s in scenes])
    off = np.cumsum([0] + [s.num_tris for s in scenes]).astype(np.uint32)
    TB = api.build_tri_batch(v0, v1, v2, off)
    cuda.cuda.synchronize()
    for i, s in enumerate(scenes):
        _cmp(orc, 0, TB[i], orc.build(0, s.v0, s.v1, s.v2))


def test_builds_from_two_host_threads(orc, cuda):
    # a triangle tree and a photon tree built at the same time from two host threads on two
    # streams (each thread has its own build workspace) equal the trees built one by one
    import concurrent.futures

    torch = cuda
    sc = S.scene_small(subdiv=4)
    ppos, _, _ = S.random_surface_photons(sc, 30000, seed=8)
    root = np.concatenate(sc.bounds()).astype(np.float32)
    params = api.default_params(1, root, sc.r_max())
    v0, v1, v2, pp = (api.dev(a) for a in (sc.v0, sc.v1, sc.v2, ppos))
    torch.cuda.synchronize()
    side = torch.cuda.Stream()

    def points():
        with torch.cuda.stream(side):
            P = api.build_points(pp, params)
            side.synchronize()
        return P

    for _ in range(3):
        with concurrent.futures.ThreadPoolExecutor(max_workers=1) as ex:
            fut = ex.submit(points)
            T = api.build_tri(v0, v1, v2)
            P = fut.result()
        torch.cuda.synchronize()
        _cmp(orc, 0, T, orc.build(0, sc.v0, sc.v1, sc.v2))
        _cmp(orc, 1, P, orc.build(1, ppos, max_qr=sc.r_max(), root=root), sc.r_max())
        T.release()
        P.release()
