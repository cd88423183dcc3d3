"""Writes tests/golden/*.npz: inputs and expected outputs computed with numpy float64
by methods independent of the oracle's formulas.

  hits.npz : a fixed small scene + rays; closest hit found by solving, per triangle,
             the 3x3 system  o + lam d = a + u (b - a) + v (c - a)  with np.linalg.solve
             (not Moller-Trumbore), keeping lam > eps, u, v >= 0, u + v <= 1.
  knn.npz  : photons + queries; the k nearest photons by float64 distance (ties by index).

The reference (jonike/mnrt) ships no golden vectors or tests; these fixtures pin the
oracle to what the mathematics fixes.  Run: python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from mnrt_amd import scene as S  # noqa: E402


def solve_hits(v0, v1, v2, o, d, eps):
    a, b, c = v0[:, :3].astype(np.float64), v1[:, :3].astype(np.float64), v2[:, :3].astype(np.float64)
    n = o.shape[0]
    tri = np.full(n, -1, np.int64)
    lam = np.zeros(n)
    uv = np.zeros((n, 2))
    for i in range(n):
        oo, dd = o[i, :3].astype(np.float64), d[i, :3].astype(np.float64)
        M = np.stack([-np.broadcast_to(dd, a.shape), b - a, c - a], axis=2)  # (t, 3, 3)
        rhs = oo[None, :] - a
        det = np.linalg.det(M)
        ok = np.abs(det) > 1e-14
        sol = np.full((a.shape[0], 3), np.nan)
        sol[ok] = np.linalg.solve(M[ok], rhs[ok][..., None])[..., 0]
        l, u, v = sol[:, 0], sol[:, 1], sol[:, 2]
        good = ok & (l > eps) & (u >= 0) & (v >= 0) & (u + v <= 1)
        if good.any():
            j = np.nonzero(good)[0][np.argmin(l[good])]
            tri[i], lam[i], uv[i] = j, l[j], (u[j], v[j])
    return tri, lam, uv


def main():
    sc = S.scene_small(subdiv=2)
    o, d, _ = S.camera_rays(24, 24)
    rng = np.random.default_rng(17)
    ro = np.zeros((300, 4), np.float32)
    ro[:, :3] = rng.uniform(-0.9, 0.9, size=(300, 3))
    rd = np.zeros((300, 4), np.float32)
    v = rng.normal(size=(300, 3))
    rd[:, :3] = v / np.linalg.norm(v, axis=1, keepdims=True)
    o = np.concatenate([o, ro])
    d = np.concatenate([d, rd])
    eps = sc.ray_epsilon()
    tri, lam, uv = solve_hits(sc.v0, sc.v1, sc.v2, o, d, eps)
    np.savez_compressed(os.path.join(HERE, "hits.npz"), v0=sc.v0, v1=sc.v1, v2=sc.v2, orig=o, dirs=d, eps=eps,
                        tri=tri, lam=lam, uv=uv)

    ppos, ppow, pdir = S.random_surface_photons(sc, 3000, seed=21)
    qpos = ppos[rng.choice(3000, 200, replace=False)].copy()
    qpos[:, :3] += rng.normal(scale=0.02, size=(200, 3)).astype(np.float32)
    k = 24
    d2 = ((qpos[:, None, :3].astype(np.float64) - ppos[None, :, :3].astype(np.float64)) ** 2).sum(-1)
    order = np.lexsort((np.broadcast_to(np.arange(3000), d2.shape), d2), axis=1)[:, :k]
    np.savez_compressed(os.path.join(HERE, "knn.npz"), ppos=ppos, qpos=qpos, k=k,
                        nbr=order.astype(np.int64), d2=np.take_along_axis(d2, order, 1))
    print("wrote hits.npz, knn.npz")


if __name__ == "__main__":
    main()
