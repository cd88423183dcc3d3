"""Shared input builders for the tests (numpy only)."""

import numpy as np

from mnrt_amd import scene as S


def f4(a):
    a = np.asarray(a, np.float32)
    return np.concatenate([a, np.zeros((a.shape[0], 1), np.float32)], axis=1)


def hit_queries(orc, sc, o, d, tree=None):
    """Query inputs at the first hits of rays o, d: position, geometric normal, ray
    direction, influence 1, diffuse colour of the hit triangle (alpha in w)."""
    eps = sc.ray_epsilon()
    if tree is None:
        tree = orc.build(0, sc.v0, sc.v1, sc.v2)
    tri, t, uv, _ = orc.trace(tree, sc.v0, sc.v1, sc.v2, o, d, eps)
    hit = tri >= 0
    tri, t = tri[hit], t[hit]
    oo, dd = o[hit], d[hit]
    pos = oo.copy()
    pos[:, :3] = oo[:, :3] + dd[:, :3] * t[:, None]
    e1 = sc.v1[tri, :3] - sc.v0[tri, :3]
    e2 = sc.v2[tri, :3] - sc.v0[tri, :3]
    n = np.cross(e1, e2)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    infl = np.ones((len(tri), 4), np.float32)
    infl[:, 3] = 0
    clr = sc.albedo[tri].copy()
    return dict(pos=pos.astype(np.float32), nrm=f4(n), dir=dd.copy(), infl=infl, clr=clr, tri=tri)


def same_tree(orc, kind, dt, ot, max_qr=1.0):
    """A device tree equals an oracle tree: node words, references, root box, level and
    leaf counts; its SAH / VVH cost is the oracle's within 0.5 %."""
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


def host(t):
    return t.detach().cpu().numpy()
