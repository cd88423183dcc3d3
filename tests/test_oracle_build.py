"""Oracle kd-tree build pins: hand-computed trees (empty-space cut, SAH split of two
clusters, no-progress leaf), structural and coverage invariants (the reference's
TEST_KDTREE checks, KDTreeGPU.cpp:327-442), an independent SAH cost evaluation, and
determinism."""

import numpy as np
import pytest

from mnrt_amd import scene as S
from helpers import f4

LEAF = 3


def _tris(t):
    t = np.asarray(t, np.float32)
    return f4(t[:, 0]), f4(t[:, 1]), f4(t[:, 2])


def _words(T, i):
    return int(T.nodes[2 * i]), int(T.nodes[2 * i + 1])


def _f(w):
    return np.array([w], np.uint32).view(np.float32)[0]


def test_empty_space_cut(orc):
    rng = np.random.default_rng(0)
    base = np.stack([rng.uniform(0, 0.08, 100), rng.uniform(0, 0.9, 100), rng.uniform(0, 0.9, 100)], 1)
    tris = base[:, None, :] + rng.uniform(0, 0.02, size=(100, 3, 3)) * np.array([1, 5, 5])
    v0, v1, v2 = _tris(tris)
    T = orc.build(0, v0, v1, v2, root=np.array([0, 0, 0, 1, 1, 1], np.float32))
    w0, w1 = _words(T, 0)
    thi = max(v0[:, 0].max(), v1[:, 0].max(), v2[:, 0].max())
    # min sides have no empty space > 25 %; the max side of x does: cut at the tight max
    assert w1 >> 30 == 0 and _f(w0) == thi
    right = w1 & 0x3FFFFFFF
    rw0, rw1 = _words(T, right)
    assert rw1 >> 30 == LEAF and rw0 == 0  # the empty leaf


def test_sah_two_clusters(orc):
    a = np.array([[0, 0, 0], [0.1, 0.1, 0], [0, 0.1, 0.1]], np.float64)
    b = a + [0.9, 0, 0]
    tris = np.concatenate([np.repeat(a[None], 10, 0), np.repeat(b[None], 10, 0)])
    v0, v1, v2 = _tris(tris)
    T = orc.build(0, v0, v1, v2)
    assert T.num_nodes == 3
    w0, w1 = _words(T, 0)
    assert w1 >> 30 == 0 and _f(w0) == np.float32(0.9)
    (l0, l1), (r0, r1) = _words(T, 1), _words(T, 2)
    assert l1 >> 30 == LEAF and r1 >> 30 == LEAF and l0 == 10 and r0 == 10
    assert list(T.refs) == list(range(20))
    lo, hi = T.root[:3].astype(np.float64), T.root[3:].astype(np.float64)
    e = hi - lo
    area = 2 * (e[0] * e[1] + e[0] * e[2] + e[1] * e[2])
    s = float(np.float32(0.9))
    al = 2 * (e[1] * e[2] + (s - lo[0]) * (e[1] + e[2]))
    ar = 2 * (e[1] * e[2] + (hi[0] - s) * (e[1] + e[2]))
    expect = 3.0 + (10 * al + 10 * ar) / area
    assert abs(orc.tree_cost(T.nodes, T.root, 0) - expect) < 1e-9


def test_no_progress_leaf(orc):
    big = np.array([[-1, -1, -1], [1, -1, 1], [-1, 1, 1]], np.float64)
    v0, v1, v2 = _tris(np.repeat(big[None], 100, 0))
    T = orc.build(0, v0, v1, v2)
    assert T.num_nodes == 1 and list(T.refs) == list(range(100))


def _py_cost(nodes, root, kind, qr=1.0):
    def meas(lo, hi):
        e = hi - lo
        if kind == 1:
            return np.prod(e + 2 * qr)
        return 2 * (e[0] * e[1] + e[0] * e[2] + e[1] * e[2])
    st = [(0, root[:3].astype(np.float64), root[3:].astype(np.float64))]
    m0 = meas(*st[0][1:])
    c = 0.0
    while st:
        n, lo, hi = st.pop()
        w0, w1 = int(nodes[2 * n]), int(nodes[2 * n + 1])
        if w1 >> 30 == LEAF:
            c += w0 * meas(lo, hi) / m0
            continue
        c += 3.0 * meas(lo, hi) / m0
        a, s = w1 >> 30, float(_f(w0))
        lh, rl = hi.copy(), lo.copy()
        lh[a], rl[a] = s, s
        st += [(n + 1, lo, lh), (w1 & 0x3FFFFFFF, rl, hi)]
    return c


@pytest.mark.parametrize("which", ["small", "soup"])
def test_tri_invariants(orc, which):
    sc = S.scene_small(subdiv=3) if which == "small" else S.scene_sized(8000, seed=3)
    T = orc.build(0, sc.v0, sc.v1, sc.v2)
    assert orc.check_structure(T.nodes, T.refs, sc.num_tris) == 0
    assert orc.check_coverage(T.nodes, T.refs, T.root, sc.v0, sc.v1, sc.v2) == 0
    assert set(np.unique(T.refs)) == set(range(sc.num_tris))
    assert T.max_level <= 48
    assert abs(orc.tree_cost(T.nodes, T.root, 0) - _py_cost(T.nodes, T.root, 0)) < 1e-9 * T.num_nodes
    T2 = orc.build(0, sc.v0, sc.v1, sc.v2)
    assert np.array_equal(T.nodes, T2.nodes) and np.array_equal(T.refs, T2.refs)
    # a planted mistake in the structure is caught
    bad = T.nodes.copy()
    bad[3] = (bad[3] & 0xC0000000) | 0x3FFFFFF0
    assert orc.check_structure(bad, T.refs, sc.num_tris) != 0


def test_point_invariants_and_vvh(orc):
    rng = np.random.default_rng(7)
    p = rng.uniform(-1, 1, size=(20000, 3))
    p[:300] = [0.25, -0.125, 0.5]  # coincident photons
    p[300:400, 0] = 0.0  # photons on a likely split plane
    pos = f4(p)
    root = np.array([-1, -1, -1, 1, 1, 1], np.float32)
    T = orc.build(1, pos, max_qr=0.1, root=root)
    assert orc.check_structure(T.nodes, T.refs, 20000) == 0
    # points are never duplicated
    assert np.array_equal(np.sort(T.refs), np.arange(20000))
    assert T.max_level <= 48
    assert abs(orc.tree_cost(T.nodes, T.root, 1, 0.1) - _py_cost(T.nodes, T.root, 1, 0.1)) < 1e-9 * T.num_nodes
    # leaves hold at most small_max points except capped or no-progress leaves
    n = T.nodes.reshape(-1, 2)
    leaf = (n[:, 1] >> 30) == LEAF
    assert (n[leaf, 0] <= 32).mean() > 0.99


def test_small_node_max_parameter(orc):
    sc = S.scene_small(subdiv=2)
    for sm in (1, 8, 64):
        T = orc.build(0, sc.v0, sc.v1, sc.v2, small_max=sm)
        assert orc.check_structure(T.nodes, T.refs, sc.num_tris) == 0
    with pytest.raises(ValueError):
        orc.build(0, sc.v0, sc.v1, sc.v2, small_max=65)


def test_build_stream_counts_invariants(orc):
    """The element-stream counts behind the build's algorithmic bytes: every distinct point
    goes to exactly one child per large level and ends in exactly one small root; a
    small-enough input goes straight to the small stage; triangle clips come in pairs."""
    rng = np.random.default_rng(3)
    pts = np.zeros((50000, 4), np.float32)
    pts[:, :3] = rng.uniform(-1, 1, (50000, 3))
    P = orc.build(1, pts)
    large, child, clips, small = (int(c) for c in P.counts)
    assert large > 50000 and child == large and clips == 0 and small == 50000
    few = orc.build(1, pts[:20])
    assert [int(c) for c in few.counts] == [0, 0, 0, 20]
    sc = S.scene_small(subdiv=3)
    T = orc.build(0, sc.v0, sc.v1, sc.v2)
    large, child, clips, small = (int(c) for c in T.counts)
    assert clips % 2 == 0 and child >= large and small >= sc.num_tris
    b = orc.build_bytes(T)
    assert b > sc.num_tris * (36 + 28) + 8 * T.num_nodes


def test_parallel_build_equals_single_thread(orc):
    """The level-parallel large stage and the per-root parallel small stage give the same
    tree (node words, references, counts) as one thread, for triangles and points."""
    sc = S.scene_small(subdiv=4)
    rng = np.random.default_rng(8)
    pts = np.zeros((120000, 4), np.float32)
    pts[:, :3] = rng.uniform(-1, 1, (120000, 3))
    for kind, args in ((0, (sc.v0, sc.v1, sc.v2)), (1, (pts,))):
        one = orc.build(kind, *args, nthreads=1)
        many = orc.build(kind, *args, nthreads=7)
        assert np.array_equal(one.nodes, many.nodes) and np.array_equal(one.refs, many.refs)
        assert np.array_equal(one.counts, many.counts) and one.max_level == many.max_level
