"""GPU any-hit segments, shadow rays and area-light samples vs the oracle
(raytracing.cu:503-645 with needClosest = false, raytracing.cu:963-1114).

Visibility flags are bit-exact against the oracle's tree walk on the same tree and
against the oracle's test over every triangle; rectangle samples are bit-exact, disc
samples (cosf / sinf of the device vs the host) within 1e-6 of the disc's scale."""

import numpy as np
import pytest

from mnrt_amd import api, native as N, scene as S
from helpers import f4

pytestmark = pytest.mark.gpu


def _unit(v):
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _trees(orc, sc):
    T = orc.build(0, sc.v0, sc.v1, sc.v2)
    DT = api.tree_from_host(0, T.nodes, T.refs, T.root, sc.num_tris, T.max_level)
    api.attach_tri_records(DT, sc.v0, sc.v1, sc.v2)
    return T, DT


def _hit_points(sc, DT, W, H):
    o, d, _ = S.camera_rays(W, H)
    tri, t, _, pt = api.trace_closest(DT, o, d, sc.ray_epsilon(), want_pt=True)
    tri, pt = tri.cpu().numpy(), pt.cpu().numpy()
    return pt[tri >= 0]


def _lights(orc):
    """(api light, oracle light16, radius) for the four light types."""
    pos, dirn = (0.1, 0.6, 0.2), tuple(_unit(np.array([[0.2, -1.0, 0.1]]))[0])
    out = []
    for kind in (N.LIGHT_POINT, N.LIGHT_DIRECTIONAL, N.LIGHT_AREA_RECT, N.LIGHT_AREA_DISC):
        out.append((api.make_light(kind, pos, dirn, v1=(0.3, 0, 0), v2=(0, 0, 0.3), radius=0.3),
                    orc.light16(kind, pos, dirn, v1=(0.3, 0, 0), v2=(0, 0, 0.3)), 0.3))
    return out


def test_trace_any_random_segments(orc, cuda):
    sc = S.scene_small(subdiv=4)
    T, DT = _trees(orc, sc)
    rng = np.random.default_rng(21)
    n = 60000
    o = f4(rng.uniform(-0.95, 0.95, size=(n, 3)))
    d = f4(_unit(rng.normal(size=(n, 3))))
    tmax = rng.uniform(0.0, 2.0, size=n).astype(np.float32)
    tmax[::5] = orc.FLT_MAX
    eps = sc.ray_epsilon()
    print(f"trace_any: {n} random segments, {sc.num_tris} tris", flush=True)
    g = api.trace_any(DT, o, d, eps, tmax).cpu().numpy().astype(np.uint32)
    h, _ = orc.trace_any(T, T.root, sc.v0, sc.v1, sc.v2, o, d, eps, tmax)
    assert np.array_equal(g, h), f"{(g != h).sum()} segments differ from the oracle"
    sel = slice(0, n, 13)
    b, _ = orc.trace_any(None, T.root, sc.v0, sc.v1, sc.v2, o[sel], d[sel], eps, tmax[sel], brute=True)
    assert np.array_equal(g[sel], b)
    # unlimited segments: hit iff the closest-hit kernel finds a hit; limited at the hit: no hit
    gu = api.trace_any(DT, o, d, eps).cpu().numpy()
    tri, t, _, _ = api.trace_closest(DT, o, d, eps)
    tri, t = tri.cpu().numpy(), t.cpu().numpy()
    assert np.array_equal(gu.astype(bool), tri >= 0)
    hit = tri >= 0
    at = api.trace_any(DT, o[hit], d[hit], eps, t[hit]).cpu().numpy()
    past = api.trace_any(DT, o[hit], d[hit], eps, np.nextafter(t[hit], np.float32(np.inf))).cpu().numpy()
    assert not at.any() and past.all()


@pytest.mark.parametrize("li", range(4))
def test_shadow_small_scene(orc, cuda, li):
    sc = S.scene_small(subdiv=5)
    T, DT = _trees(orc, sc)
    pts = _hit_points(sc, DT, 256, 256)
    eps = sc.ray_epsilon()
    light, L16, radius = _lights(orc)[li]
    smp = None
    if light.type in (N.LIGHT_AREA_RECT, N.LIGHT_AREA_DISC):
        rng = np.random.default_rng(40 + li)
        r1 = rng.random(len(pts), dtype=np.float32)
        r2 = rng.random(len(pts), dtype=np.float32)
        smp = api.sample_area_light(light, r1, r2, 1.0, 0.5, 0.0, 0.5).cpu().numpy()
        hs = orc.sample_area_light(L16, radius, r1, r2, 1.0, 0.5, 0.0, 0.5)
        if light.type == N.LIGHT_AREA_RECT:
            assert np.array_equal(smp.view(np.uint32), hs.view(np.uint32))
        else:
            assert np.max(np.abs(smp - hs)) <= 1e-6 * (1.0 + radius)
    print(f"shadow: light type {light.type}, {len(pts)} points", flush=True)
    g = api.trace_shadow(DT, light, pts, eps, smp).cpu().numpy().astype(np.uint32)
    h = orc.trace_shadow(T, T.root, sc.v0, sc.v1, sc.v2, L16, radius, pts, eps, smp)
    assert np.array_equal(g, h), f"{(g != h).sum()} shadow flags differ from the oracle"
    sel = slice(0, len(pts), 11)
    b = orc.trace_shadow(None, T.root, sc.v0, sc.v1, sc.v2, L16, radius, pts[sel], eps,
                         None if smp is None else smp[sel], brute=True)
    assert np.array_equal(g[sel], b)
    assert 0.02 < g.mean() < 0.999


def test_shadow_large_scene_gpu_tree(orc, cuda):
    # GPU-built tree of the 1M-triangle scene, full-HD hit points, point light; oracle walk
    # of the same (downloaded) tree on a sample
    sc = S.scene_large()
    DT = api.build_tri(sc.v0, sc.v1, sc.v2)
    pts = _hit_points(sc, DT, 1920, 1080)
    eps = sc.ray_epsilon()
    lq = sc.lights[0]
    centre = lq[0] + 0.5 * lq[1] + 0.5 * lq[2] - np.array([0, 0.01, 0], np.float32)
    light = api.make_light(N.LIGHT_POINT, centre)
    print(f"shadow: large scene {sc.num_tris} tris, {len(pts)} points", flush=True)
    g = api.trace_shadow(DT, light, pts, eps).cpu().numpy().astype(np.uint32)
    T = orc.Tree(0, DT.nodes_host(), DT.refs_host(), DT.root, DT.max_level, 0, sc.num_tris)
    sel = slice(0, len(pts), 17)
    h = orc.trace_shadow(T, T.root, sc.v0, sc.v1, sc.v2, orc.light16(0, centre), 0.0, pts[sel], eps)
    assert np.array_equal(g[sel], h)
    assert 0.05 < g.mean() < 0.999
    DT.release()


def test_shadow_edge_cases(orc, cuda):
    sc = S.scene_small(subdiv=2)
    T, DT = _trees(orc, sc)
    eps = sc.ray_epsilon()
    lib = N.load()
    light = api.make_light(N.LIGHT_POINT, (0.0, 0.5, 0.0))
    # empty input
    assert lib.mnrt_trace_shadow(DT.ref(), light, 0, None, None, eps, None, N.stream_handle()) == 0
    assert lib.mnrt_trace_any(DT.ref(), 0, None, None, None, eps, None, N.stream_handle()) == 0
    # a point at the light position is unoccluded; one at an area sample is not lit
    pts = f4(np.array([[0.0, 0.5, 0.0]], np.float32))
    assert api.trace_shadow(DT, light, pts, eps).cpu().tolist() == [1]
    area = api.make_light(N.LIGHT_AREA_RECT, (0.0, 0.5, 0.0), (0, -1, 0), v1=(0.1, 0, 0), v2=(0, 0, 0.1))
    assert api.trace_shadow(DT, area, pts, eps, pts).cpu().tolist() == [0]
    # argument errors: area light without samples, unknown light type, point tree
    dummy = cuda.zeros((1, 4), device="cuda")
    out = cuda.zeros(1, dtype=cuda.int32, device="cuda")
    assert lib.mnrt_trace_shadow(DT.ref(), area, 1, N.ptr(dummy), None, eps, N.ptr(out),
                                 N.stream_handle()) == N.ERR_ARG
    bad = api.make_light(7)
    assert lib.mnrt_trace_shadow(DT.ref(), bad, 1, N.ptr(dummy), None, eps, N.ptr(out),
                                 N.stream_handle()) == N.ERR_ARG
    with pytest.raises(N.MnrtError):
        api.sample_area_light(light, np.zeros(4, np.float32), np.zeros(4, np.float32))
    P = api.build_points(sc.v0, api.default_params(1, np.concatenate(sc.bounds()).astype(np.float32), 0.1))
    assert lib.mnrt_trace_shadow(P.ref(), light, 1, N.ptr(dummy), None, eps, N.ptr(out),
                                 N.stream_handle()) == N.ERR_STATE
