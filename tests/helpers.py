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
