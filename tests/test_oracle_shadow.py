"""Oracle pins for any-hit segments and shadow rays (raytracing.cu:503-645 with
needClosest = false, kernel_TraceShadowRaysDelta/_Area raytracing.cu:1019-1114,
kernel_SampleAreaLight raytracing.cu:963-1000).

The tree walk is pinned to the test over every triangle, to the closest-hit oracle
(a segment is blocked iff the closest hit lies before its end), and to hand-built
scenes whose visibility is known; area-light samples to closed forms of the rectangle
and the concentric disc map."""

import numpy as np
import pytest

from mnrt_amd import scene as S
from helpers import f4, hit_queries

POINT, DIRECTIONAL, DISC, RECT = 0, 1, 2, 3


def _random_scene(seed, n):
    rng = np.random.default_rng(seed)
    base = rng.uniform(-1, 1, size=(n, 1, 3))
    tris = base + rng.normal(scale=0.15, size=(n, 3, 3))
    return f4(tris[:, 0]), f4(tris[:, 1]), f4(tris[:, 2])


def _unit(v):
    return v / np.linalg.norm(v, axis=1, keepdims=True)


@pytest.mark.parametrize("seed", range(3))
def test_any_hit_equals_brute_force_and_closest(orc, seed):
    v0, v1, v2 = _random_scene(seed, 3000)
    T = orc.build(0, v0, v1, v2)
    rng = np.random.default_rng(300 + seed)
    n = 4000
    o = f4(rng.uniform(-1.2, 1.2, size=(n, 3)))
    d = f4(_unit(rng.normal(size=(n, 3))))
    tmax = rng.uniform(0.0, 2.5, size=n).astype(np.float32)
    tmax[::7] = orc.FLT_MAX
    eps = 1e-4
    hit, cnt = orc.trace_any(T, T.root, v0, v1, v2, o, d, eps, tmax)
    bhit, _ = orc.trace_any(None, T.root, v0, v1, v2, o, d, eps, tmax, brute=True)
    assert np.array_equal(hit, bhit)
    # closest hit before the segment end <=> some hit before it
    tri, t, _, _ = orc.trace(T, v0, v1, v2, o, d, eps)
    assert np.array_equal(hit.astype(bool), (tri >= 0) & (t < tmax))
    assert 0.2 < hit.mean() < 0.9
    assert cnt[3] < 0.05 * n * 3000


def test_any_hit_segment_end_exact(orc):
    # the segment ending exactly at the closest hit is not blocked by it (lambda < tmax)
    sc = S.scene_small(subdiv=3)
    T = orc.build(0, sc.v0, sc.v1, sc.v2)
    o, d, _ = S.camera_rays(64, 64)
    eps = sc.ray_epsilon()
    tri, t, _, _ = orc.trace(T, sc.v0, sc.v1, sc.v2, o, d, eps)
    assert np.all(tri >= 0)
    at, _ = orc.trace_any(T, T.root, sc.v0, sc.v1, sc.v2, o, d, eps, t)
    beyond, _ = orc.trace_any(T, T.root, sc.v0, sc.v1, sc.v2, o, d, eps, np.nextafter(t, np.inf))
    assert not at.any()
    assert beyond.all()


def test_shadow_hand_scene(orc):
    # floor quad at y = 0 (two triangles), occluder square at y = 0.5 over x, z in [-0.25, 0.25]
    floor = np.array([[[-1, 0, -1], [1, 0, -1], [1, 0, 1]], [[-1, 0, -1], [1, 0, 1], [-1, 0, 1]]], np.float32)
    occ = np.array([[[-.25, .5, -.25], [.25, .5, -.25], [.25, .5, .25]],
                    [[-.25, .5, -.25], [.25, .5, .25], [-.25, .5, .25]]], np.float32)
    tris = np.concatenate([floor, occ])
    v0, v1, v2 = f4(tris[:, 0]), f4(tris[:, 1]), f4(tris[:, 2])
    T = orc.build(0, v0, v1, v2)
    root = np.array([-1, 0, -1, 1, 1, 1], np.float32)
    eps = 1e-4
    pts = f4(np.array([[0, 0, 0], [0.9, 0, 0.9], [0.1, 0, -0.1], [-0.6, 0, 0.0]], np.float32))
    # point light above the occluder: points under it are shadowed
    L = orc.light16(POINT, position=(0, 0.9, 0))
    vis = orc.trace_shadow(T, root, v0, v1, v2, L, 0.0, pts, eps)
    assert vis.tolist() == [0, 1, 0, 1]
    # point light between occluder and floor: nothing blocks
    L = orc.light16(POINT, position=(0, 0.4, 0))
    assert orc.trace_shadow(T, root, v0, v1, v2, L, 0.0, pts, eps).tolist() == [1, 1, 1, 1]
    # directional light shining straight down (direction = travel direction)
    L = orc.light16(DIRECTIONAL, direction=(0, -1, 0))
    assert orc.trace_shadow(T, root, v0, v1, v2, L, 0.0, pts, eps).tolist() == [0, 1, 0, 1]
    # light at the point itself: unoccluded
    L = orc.light16(POINT, position=(0, 0, 0))
    assert orc.trace_shadow(T, root, v0, v1, v2, L, 0.0, pts[:1], eps).tolist() == [1]
    # area light facing down at y = 0.9; samples above the occluder / off to the side
    L = orc.light16(RECT, position=(-0.1, 0.9, -0.1), direction=(0, -1, 0), v1=(0.2, 0, 0), v2=(0, 0, 0.2))
    smp = f4(np.array([[0, 0.9, 0], [0.9, 0.9, 0.9], [0, 0.9, 0], [-0.6, 0.9, 0.0]], np.float32))
    assert orc.trace_shadow(T, root, v0, v1, v2, L, 0.0, pts, eps, smp).tolist() == [0, 1, 0, 1]
    # the same light facing up: samples face away from every point
    L[4:7] = (0, 1, 0)
    assert orc.trace_shadow(T, root, v0, v1, v2, L, 0.0, pts, eps, smp).tolist() == [0, 0, 0, 0]
    # occluder at distance - eps / 2 does not shadow an area sample (lambda + eps >= distance)
    L = orc.light16(RECT, position=(-0.1, 0.5, -0.1), direction=(0, -1, 0), v1=(0.2, 0, 0), v2=(0, 0, 0.2))
    smp2 = f4(np.array([[0, 0.5 + 0.5 * eps, 0]], np.float32))
    assert orc.trace_shadow(T, root, v0, v1, v2, L, 0.0, pts[:1], eps, smp2).tolist() == [1]
    smp3 = f4(np.array([[0, 0.5 + 4 * eps, 0]], np.float32))
    assert orc.trace_shadow(T, root, v0, v1, v2, L, 0.0, pts[:1], eps, smp3).tolist() == [0]


@pytest.mark.parametrize("kind", [POINT, DIRECTIONAL, RECT, DISC])
def test_shadow_tree_equals_brute_force(orc, kind):
    sc = S.scene_small(subdiv=3)
    T = orc.build(0, sc.v0, sc.v1, sc.v2)
    o, d, _ = S.camera_rays(48, 48)
    q = hit_queries(orc, sc, o, d, T)
    pts = q["pos"]
    eps = sc.ray_epsilon()
    rng = np.random.default_rng(7 + kind)
    L = orc.light16(kind, position=(0.1, 0.6, 0.2), direction=_unit(np.array([[0.2, -1.0, 0.1]]))[0],
                    v1=(0.3, 0, 0), v2=(0, 0, 0.3))
    smp = None
    if kind in (RECT, DISC):
        smp = orc.sample_area_light(L, 0.3, rng.random(len(pts), dtype=np.float32),
                                    rng.random(len(pts), dtype=np.float32))
    vis = orc.trace_shadow(T, T.root, sc.v0, sc.v1, sc.v2, L, 0.3, pts, eps, smp)
    bvis = orc.trace_shadow(T, T.root, sc.v0, sc.v1, sc.v2, L, 0.3, pts, eps, smp, brute=True)
    assert np.array_equal(vis, bvis)
    assert 0.05 < vis.mean() < 0.99


def test_sample_rect_closed_form(orc):
    L = orc.light16(RECT, position=(1, 2, 3), v1=(0.5, 0, 0), v2=(0, 0, 0.25))
    r1 = np.array([0, 1, 0.5, 0.25], np.float32)
    r2 = np.array([0, 1, 0.5, 0.75], np.float32)
    p = orc.sample_area_light(L, 0.0, r1, r2)
    exp = np.array([[1, 2, 3], [1.5, 2, 3.25], [1.25, 2, 3.125], [1.125, 2, 3.1875]], np.float32)
    assert np.array_equal(p[:, :3], exp) and np.all(p[:, 3] == 0)
    # stratification: (idx + rnd) * inv
    p2 = orc.sample_area_light(L, 0.0, r1 * 0 + 0.5, r2 * 0 + 0.5, ix=1.0, inv_x=0.5, iy=0.0, inv_y=0.5)
    assert np.allclose(p2[:, :3], [[1.375, 2, 3.0625]] * 4)


def test_sample_disc_closed_form(orc):
    c, R = np.array([0.5, 1.0, -0.5], np.float32), 0.4
    n = np.array([0.0, -1.0, 0.0], np.float32)
    L = orc.light16(DISC, position=c, direction=n)
    # centre, the four edge midpoints of the square (angle 0, pi/2, pi, 3pi/2) and a corner
    r1 = np.array([0.5, 1.0, 0.5, 0.0, 0.5, 1.0], np.float32)
    r2 = np.array([0.5, 0.5, 1.0, 0.5, 0.0, 1.0], np.float32)
    p = orc.sample_area_light(L, R, r1, r2)[:, :3].astype(np.float64)
    # basis for n = (0,-1,0): |n.x| > |n.y| is false -> s = (0, n.z, -n.y)/|.| = (0,0,1), t = n x s
    s = np.array([0, 0, 1.0])
    t = np.cross(n.astype(np.float64), s)
    ang = np.array([0, np.pi / 2, np.pi, 1.5 * np.pi, np.pi / 4])
    exp = c + R * (np.cos(ang)[:, None] * s + np.sin(ang)[:, None] * t)
    assert np.array_equal(p[0], c.astype(np.float64))
    assert np.allclose(p[1:], exp, atol=1e-6)


def test_sample_disc_uniform(orc):
    rng = np.random.default_rng(5)
    n = _unit(np.array([[0.3, -0.8, 0.5]]))[0].astype(np.float32)
    c = np.array([0.1, 0.2, 0.3], np.float32)
    R = 0.7
    L = orc.light16(DISC, position=c, direction=n)
    m = 200000
    p = orc.sample_area_light(L, R, rng.random(m, dtype=np.float32), rng.random(m, dtype=np.float32))[:, :3]
    rel = p.astype(np.float64) - c
    # in the disc's plane, inside the radius, uniform over the area: E[r^2] = R^2 / 2
    assert np.max(np.abs(rel @ n.astype(np.float64))) < 1e-6
    r2 = np.sum(rel * rel, axis=1)
    assert np.max(r2) <= R * R * (1 + 1e-6)
    assert abs(r2.mean() / (R * R) - 0.5) < 0.005
    # every quarter of the disc gets a quarter of the samples
    e1 = _unit(np.array([[0, n[2], -n[1]]], np.float64))[0]
    e2 = np.cross(n.astype(np.float64), e1)
    ang = np.arctan2(rel @ e2, rel @ e1)
    counts = np.histogram(ang, bins=4, range=(-np.pi, np.pi))[0] / m
    assert np.all(np.abs(counts - 0.25) < 0.005)
