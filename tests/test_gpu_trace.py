"""GPU closest-hit traversal vs the oracle (same tree, same rays): hit triangle index,
lambda and barycentrics bit-exact; both equal the brute force over all triangles."""

import numpy as np
import pytest

from mnrt_amd import api, scene as S

pytestmark = pytest.mark.gpu


def _check(orc, sc, o, d, cuda, brute_stride=1):
    eps = sc.ray_epsilon()
    T = orc.build(0, sc.v0, sc.v1, sc.v2)
    tri, t, uv, _ = orc.trace(T, sc.v0, sc.v1, sc.v2, o, d, eps)
    DT = api.tree_from_host(0, T.nodes, T.refs, T.root, sc.num_tris, T.max_level)
    api.attach_tri_records(DT, sc.v0, sc.v1, sc.v2)
    gtri, gt, guv, gpt = api.trace_closest(DT, o, d, eps, want_pt=True)
    cuda.cuda.synchronize()
    gtri, gt, guv = gtri.cpu().numpy(), gt.cpu().numpy(), guv.cpu().numpy()
    assert np.array_equal(gtri, tri), f"{(gtri != tri).sum()} hit indices differ"
    assert np.array_equal(gt.view(np.uint32), t.view(np.uint32)), "lambda not bit-exact"
    assert np.array_equal(guv.view(np.uint32), uv.view(np.uint32)), "barycentrics not bit-exact"
    sel = np.arange(0, o.shape[0], brute_stride)
    btri, bt, _ = orc.trace_brute(T.root, sc.v0, sc.v1, sc.v2, o[sel], d[sel], eps)
    assert np.array_equal(gtri[sel], btri)
    hit = btri >= 0
    assert np.array_equal(gt[sel][hit].view(np.uint32), bt[hit].view(np.uint32))
    return gtri, gt


def test_trace_small_scene_primary_rays(orc, cuda):
    sc = S.scene_small(subdiv=5)
    o, d, _ = S.camera_rays(512, 512)
    print(f"trace: small scene {sc.num_tris} tris, {o.shape[0]} rays", flush=True)
    tri, t = _check(orc, sc, o, d, cuda, brute_stride=7)
    assert (tri >= 0).mean() > 0.99


def test_trace_random_rays_inside_box(orc, cuda):
    sc = S.scene_small(subdiv=3)
    rng = np.random.default_rng(5)
    n = 50000
    o = np.zeros((n, 4), np.float32)
    o[:, :3] = rng.uniform(-0.95, 0.95, size=(n, 3))
    d = np.zeros((n, 4), np.float32)
    v = rng.normal(size=(n, 3))
    d[:, :3] = v / np.linalg.norm(v, axis=1, keepdims=True)
    print(f"trace: random rays {n}", flush=True)
    _check(orc, sc, o, d, cuda)


def test_trace_axis_aligned_and_grazing(orc, cuda):
    # rays parallel to split planes, starting in planes, and along box walls
    sc = S.scene_small(subdiv=2)
    dirs = []
    for ax in range(3):
        for s in (-1.0, 1.0):
            v = np.zeros(3)
            v[ax] = s
            dirs.append(v)
    rng = np.random.default_rng(11)
    o = []
    d = []
    for _ in range(3000):
        p = rng.choice([-0.5, 0.0, 0.25, 0.5, rng.uniform(-0.9, 0.9)], size=3)
        for v in dirs:
            o.append(p)
            d.append(v)
    o4 = np.zeros((len(o), 4), np.float32)
    o4[:, :3] = np.array(o)
    d4 = np.zeros((len(d), 4), np.float32)
    d4[:, :3] = np.array(d)
    print(f"trace: axis-aligned {len(o)}", flush=True)
    _check(orc, sc, o4, d4, cuda)


def test_trace_degenerate_and_coincident(orc, cuda):
    # coincident triangles (ties resolved by index), zero-area triangles, tiny scene
    base = S.icosphere(1, 0.5)
    tris = np.concatenate([base, base, base[:10] * np.array([1, 1, 1]), base[:5, [0, 0, 1]]]).astype(np.float32)
    box = S.box_tris()
    tris = np.concatenate([box, tris]).astype(np.float32)
    f4 = lambda a: np.concatenate([a, np.zeros((a.shape[0], 1), np.float32)], axis=1)
    sc = S.Scene(f4(tris[:, 0]), f4(tris[:, 1]), f4(tris[:, 2]), np.ones((len(tris), 4), np.float32),
                 np.zeros((0, 4, 3), np.float32))
    o, d, _ = S.camera_rays(64, 64)
    print("trace: degenerate", flush=True)
    _check(orc, sc, o, d, cuda)


def test_trace_empty_ray_set(orc, cuda):
    sc = S.scene_small(subdiv=1)
    o = np.zeros((0, 4), np.float32)
    T = orc.build(0, sc.v0, sc.v1, sc.v2)
    DT = api.tree_from_host(0, T.nodes, T.refs, T.root, sc.num_tris, T.max_level)
    api.attach_tri_records(DT, sc.v0, sc.v1, sc.v2)
    tri, t, uv, _ = api.trace_closest(DT, o, o, sc.ray_epsilon())
    assert tri.numel() == 0


def test_trace_misses_and_ragged_counts(orc, cuda):
    # rays that miss the root box (started and finished at the fetch) mixed with hits,
    # in counts that are not multiples of a wave or a workgroup; every lane of the
    # persistent kernel takes several rays
    sc = S.scene_small(subdiv=3)
    rng = np.random.default_rng(17)
    for n in (1, 63, 65, 1000, 70001):
        o = np.zeros((n, 4), np.float32)
        d = np.zeros((n, 4), np.float32)
        inside = rng.random(n) < 0.5
        o[:, :3] = np.where(inside[:, None], rng.uniform(-0.9, 0.9, size=(n, 3)), rng.uniform(2.0, 3.0, size=(n, 3)))
        v = rng.normal(size=(n, 3))
        v = v / np.linalg.norm(v, axis=1, keepdims=True)
        v[~inside] = np.abs(v[~inside])  # outside rays point away from the box
        d[:, :3] = v
        print(f"trace: ragged n={n}", flush=True)
        _check(orc, sc, o, d, cuda, brute_stride=max(1, n // 2000))
