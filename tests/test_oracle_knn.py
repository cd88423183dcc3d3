"""Oracle exact kNN and radiance pins: brute force, fp64 neighbour sets (golden
fixture), closed-form radiance cases."""

import os

import numpy as np
import pytest

from mnrt_amd import scene as S
from helpers import f4

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _np_knn32(ppos, qpos, k, rmax):
    """numpy float32 distances (same operation order), sorted by (d2, index)."""
    dx = ppos[None, :, 0] - qpos[:, None, 0]
    dy = ppos[None, :, 1] - qpos[:, None, 1]
    dz = ppos[None, :, 2] - qpos[:, None, 2]
    d2 = (dx * dx + dy * dy) + dz * dz
    r2 = np.float32(rmax) * np.float32(rmax)
    out = np.full((qpos.shape[0], k), 0xFFFFFFFF, np.uint32)
    cnt = np.zeros(qpos.shape[0], np.uint32)
    for i in range(qpos.shape[0]):
        idx = np.nonzero(d2[i] < r2)[0]
        o = idx[np.lexsort((idx, d2[i, idx]))][:k]
        out[i, :len(o)] = o
        cnt[i] = len(o)
    return out, cnt


@pytest.mark.parametrize("k", [1, 5, 50, 128])
def test_knn_tree_equals_brute_force(orc, k):
    sc = S.scene_small(subdiv=2)
    ppos, _, _ = S.random_surface_photons(sc, 6000, seed=k)
    rng = np.random.default_rng(k)
    q = ppos[rng.choice(6000, 400, replace=False)].copy()
    q[:, :3] += rng.normal(scale=0.03, size=(400, 3)).astype(np.float32)
    rmax = sc.r_max()
    T = orc.build(1, ppos, max_qr=rmax, root=np.concatenate(sc.bounds()))
    nbr, cnt, ctr = orc.knn(T, ppos, q, k, rmax)
    bn, bc = orc.knn_brute(ppos, q, k, rmax)
    assert np.array_equal(nbr, bn) and np.array_equal(cnt, bc)
    nn, nc = _np_knn32(ppos, q, k, rmax)
    assert np.array_equal(nbr, nn) and np.array_equal(cnt, nc)
    assert ctr[3] < 400 * 6000 * 0.5  # pruning happens


def test_knn_ties_and_coincident(orc):
    rng = np.random.default_rng(2)
    p = rng.uniform(-1, 1, size=(2000, 3)).astype(np.float32)
    p[500:900] = p[:400]  # duplicates: equal d2, order by index
    p[900:1000] = p[0]
    ppos = f4(p)
    q = ppos[:60].copy()
    T = orc.build(1, ppos, max_qr=0.5, root=np.array([-1, -1, -1, 1, 1, 1], np.float32))
    nbr, cnt, _ = orc.knn(T, ppos, q, 40, 0.5)
    nn, nc = _np_knn32(ppos, q, 40, 0.5)
    assert np.array_equal(nbr, nn) and np.array_equal(cnt, nc)
    assert nbr[0, 0] == 0  # d2 = 0, lowest index first


def test_golden_knn_fp64(orc):
    g = np.load(os.path.join(GOLD, "knn.npz"))
    ppos, qpos, k = g["ppos"], g["qpos"], int(g["k"])
    T = orc.build(1, ppos, max_qr=2.0, root=np.array([-1, -1, -1, 1, 1, 1], np.float32))
    nbr, cnt, _ = orc.knn(T, ppos, qpos, k, 4.0)
    assert (cnt == k).all()
    for i in range(qpos.shape[0]):
        a, b = set(nbr[i].tolist()), set(g["nbr"][i].tolist())
        if a != b:
            # a swap is only allowed at the k-th distance, within fp32 rounding
            dk = g["d2"][i, -1]
            for j in a ^ b:
                d = ((ppos[j, :3].astype(np.float64) - qpos[i, :3]) ** 2).sum()
                assert abs(d - dk) <= 1e-6 * dk


def _rad(orc, nbr, cnt, k, rmax, ppos, ppow, pdir, qpos, qnrm, qdir, clr=None):
    infl = np.ones_like(qpos)
    return orc.radiance(nbr, cnt, k, rmax, ppos, ppow, pdir, qpos, qnrm, qdir, infl, clr)


def test_radiance_closed_forms(orc):
    ppos = f4([[0.1, 0.0, 0.0], [0.0, 0.3, 0.0], [0.0, 0.0, 2.0]])
    ppow = f4([[1.0, 2.0, 3.0], [0.5, 0.5, 0.5], [9.0, 9.0, 9.0]])
    pdir = f4([[0, 0, -1.0], [0, 0, -1.0], [0, 0, 1.0]])  # photons travel -z: front for n=+z
    qpos = f4([[0.0, 0.0, 0.0]])
    qnrm = f4([[0, 0, 1.0]])
    qdir = f4([[0, 0, -1.0]])
    # k = 1: E = power / (pi d^2)
    nbr = np.array([[0]], np.uint32)
    o32, o64, i32, i64, r2 = _rad(orc, nbr, np.array([1], np.uint32), 1, 1.0, ppos, ppow, pdir, qpos, qnrm, qdir)
    d2_64 = float(np.float32(0.1)) ** 2  # the fp64 path squares the stored fp32 coordinate
    np.testing.assert_allclose(i64[0, :3], np.array([1, 2, 3]) / (np.pi * d2_64), rtol=1e-12)
    np.testing.assert_allclose(o64[0, :3], i64[0, :3] / np.pi, rtol=1e-12)
    assert r2[0] == np.float32(0.1) * np.float32(0.1)
    # k = 2 of 2 in range: sum of both front-facing powers over pi d_2^2
    nbr = np.array([[0, 1]], np.uint32)
    _, _, _, i64, r2 = _rad(orc, nbr, np.array([2], np.uint32), 2, 1.0, ppos, ppow, pdir, qpos, qnrm, qdir)
    np.testing.assert_allclose(i64[0, :3], np.array([1.5, 2.5, 3.5]) / (np.pi * float(np.float32(0.3)) ** 2),
                               rtol=1e-12)
    # fewer than k found: r2 = r_max^2
    _, _, _, i64, r2 = _rad(orc, nbr, np.array([2], np.uint32), 5, 1.0, ppos, ppow, pdir, qpos, qnrm, qdir)
    assert r2[0] == 1.0
    np.testing.assert_allclose(i64[0, :3], np.array([1.5, 2.5, 3.5]) / np.pi, rtol=1e-6)
    # back-facing photon contributes nothing; normal flips toward the ray
    nbr = np.array([[2]], np.uint32)
    _, _, _, i64, _ = _rad(orc, nbr, np.array([1], np.uint32), 1, 3.0, ppos, ppow, pdir, qpos, qnrm, qdir)
    assert (i64[0, :3] == 0).all()
    _, _, _, i64, _ = _rad(orc, nbr, np.array([1], np.uint32), 1, 3.0, ppos, ppow, pdir, qpos, qnrm,
                           f4([[0, 0, 1.0]]))
    assert (i64[0, :3] > 0).all()
    # with a colour, fewer than 8 photons gives zero (photon_knn.cu:518-519)
    clr = np.array([[0.5, 0.5, 0.5, 1.0]], np.float32)
    o32, o64, _, _, _ = _rad(orc, np.array([[0, 1]], np.uint32), np.array([2], np.uint32), 2, 1.0, ppos, ppow,
                             pdir, qpos, qnrm, qdir, clr)
    assert (o64 == 0).all() and (o32 == 0).all()


def test_radiance_fp32_vs_fp64(orc):
    sc = S.scene_small(subdiv=3)
    ppos, ppow, pdir = S.random_surface_photons(sc, 20000, seed=5)
    rng = np.random.default_rng(4)
    q = ppos[rng.choice(20000, 500, replace=False)].copy()
    nrm = f4(rng.normal(size=(500, 3)))
    nrm[:, :3] /= np.linalg.norm(nrm[:, :3], axis=1, keepdims=True)
    qd = f4(rng.normal(size=(500, 3)))
    rmax = sc.r_max()
    T = orc.build(1, ppos, max_qr=rmax, root=np.concatenate(sc.bounds()))
    nbr, cnt, _ = orc.knn(T, ppos, q, 100, rmax)
    clr = np.tile(np.array([0.7, 0.6, 0.5, 1.0], np.float32), (500, 1))
    o32, o64, _, _, _ = _rad(orc, nbr, cnt, 100, rmax, ppos, ppow, pdir, q, nrm, qd, clr)
    scale = np.abs(o64).max(axis=1, keepdims=True) + 1e-30
    assert (np.abs(o32 - o64) / scale).max() < 1e-5
    # independent fp64 evaluation of the estimator from the set
    codes = orc.oct_encode(pdir)
    for i in range(0, 500, 37):
        n = nrm[i, :3].astype(np.float64)
        if np.dot(-qd[i, :3], n) < 0:
            n = -n
        s = np.zeros(3)
        for j in nbr[i, :cnt[i]]:
            if np.dot(n, orc.oct_decode(codes[j])) < 0:
                s += ppow[j, :3]
        dk = ((ppos[nbr[i, 99], :3].astype(np.float64) - q[i, :3]) ** 2).sum()
        ref = clr[i, :3].astype(np.float64) / np.pi * s / (np.pi * dk)
        np.testing.assert_allclose(o64[i, :3], ref, rtol=1e-9, atol=1e-30)


def test_oct_code_round_trip(orc):
    rng = np.random.default_rng(0)
    d = rng.normal(size=(5000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    codes = orc.oct_encode(f4(d))
    for i in range(0, 5000, 7):
        v = orc.oct_decode(codes[i]).astype(np.float64)
        v /= np.linalg.norm(v)
        assert np.dot(v, d[i]) > 1 - 1e-7  # < 0.03 degrees
    for axis in np.eye(3):
        for s in (1, -1):
            v = orc.oct_decode(orc.oct_encode(f4([axis * s]))[0])
            assert np.allclose(v / np.linalg.norm(v), axis * s)


def test_fixed_sum_is_the_rounded_exact_sum(orc):
    # terms within 2^17 of the largest are exact multiples of 2^(e-40): the total is the
    # exact rational sum rounded to float32 once, whatever the order of the terms
    from fractions import Fraction

    rng = np.random.default_rng(12)
    for trial in range(200):
        n = int(rng.integers(1, 257))
        v = (rng.uniform(0.5, 1.0, size=n) * 2.0 ** rng.integers(-16, 1, size=n)).astype(np.float32)
        v *= np.float32(2.0 ** rng.integers(-60, 60))
        if trial % 3 == 0:
            v[rng.random(n) < 0.3] *= -1
        exact = sum(Fraction(float(x)) for x in v)
        want = np.float32(float(exact))  # exact sum fits a double: one rounding to float32
        got = orc.fixed_sum(v)
        assert got == want, (trial, got, want)
        assert orc.fixed_sum(rng.permutation(v)) == got


def test_fixed_sum_edge_cases(orc):
    assert orc.fixed_sum(np.zeros(0, np.float32)) == 0.0
    assert orc.fixed_sum(np.zeros(5, np.float32)) == 0.0
    assert orc.fixed_sum([1.0, -1.0, 2.0 ** -10]) == np.float32(2.0 ** -10)
    # terms below 2^-41 of the largest are rounded to multiples of 2^(e-40)
    assert orc.fixed_sum([1.0, 2.0 ** -45]) == 1.0
    assert orc.fixed_sum([1.0] + [2.0 ** -41] * 4) == np.float32(1.0 + 4 * 2.0 ** -41)
    # denormal and huge magnitudes
    tiny = np.float32(1e-44)
    assert orc.fixed_sum([tiny, tiny]) == np.float32(2 * float(tiny))
    big = np.float32(3e38)
    e = np.frexp(big)[1]
    q = 2.0 ** (e - 40)  # a term 2^-28 of the largest is rounded to a multiple of q
    assert orc.fixed_sum([big, -big, 1e30]) == np.float32(round(float(np.float32(1e30)) / q) * q)
    # close to a sequential float32 sum for positive terms (within the sequential sum's error)
    rng = np.random.default_rng(2)
    v = rng.uniform(0, 1, size=100).astype(np.float32)
    seq = np.float32(0)
    for x in v:
        seq = np.float32(seq + x)
    assert abs(float(orc.fixed_sum(v)) - float(seq)) <= 100 * np.spacing(seq)
    assert abs(float(orc.fixed_sum(v)) - float(np.sum(v.astype(np.float64)))) <= np.spacing(np.float32(v.sum()))


def test_oct_quotient_by_reciprocal_product():
    """The kernels decode q / 32767 as the fp64 product with 1/32767 rounded to fp32 (three
    instructions instead of an IEEE division); for every 16-bit code it is the IEEE quotient
    the oracle computes, so decoded directions and front-facing tests are unchanged."""
    q = np.arange(-32767, 32768, dtype=np.int32)
    ieee = q.astype(np.float32) / np.float32(32767.0)
    prod = (q.astype(np.float64) * (1.0 / 32767.0)).astype(np.float32)
    assert np.array_equal(ieee.view(np.uint32), prod.view(np.uint32))


def test_fixed_point_rounding_by_magic_constant():
    """The kernels' fixed-point terms round x (|x| < 2^41 there) to an int64 by adding
    1.5 * 2^52 and subtracting the constant's bits; this equals round-half-even (what the
    oracle's llrint does) for every |x| < 2^51, ties and negative values included."""
    rng = np.random.default_rng(5)
    x = np.concatenate([
        rng.uniform(-2.0**41, 2.0**41, 200000),
        np.arange(-1000, 1000) + 0.5,  # exact ties
        np.array([0.0, -0.0, 0.49999999999999994, -0.5, 1.5, 2.5, 2.0**50 + 0.5, -(2.0**51) + 1.5]),
    ])
    magic = np.float64(1.5 * 2.0**52)
    got = (x + magic).view(np.int64) - magic.view(np.int64)
    assert np.array_equal(got, np.rint(x).astype(np.int64))
