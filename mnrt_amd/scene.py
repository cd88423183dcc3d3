"""Synthetic scene, camera and light generators for the kd-tree hot path.

The reference renders OBJ/3DS scenes loaded through ASSIMP (AssimpScene.cpp:51) with a
Morton-ordered pinhole camera (raygen.cu:171-240).  The workloads here are synthetic and
seeded, so every run (CPU tests, GPU tests, benchmark) sees the same bytes:

* a closed box [-1, 1]^3 (12 triangles, inward facing), diffuse albedo 0.7;
* icospheres (subdivided icosahedra) of 20 * 4**subdiv triangles;
* random small triangles (v0 uniform in the box, v1/v2 = v0 + uniform unit-ball offset
  scaled by U(0, 0.05]);
* square area lights 1e-3 below the ceiling (y = 1) facing -y;
* a pinhole camera at (0, 0, 0.9) looking at the origin, vertical field of view 90
  degrees, one ray through each pixel centre, rays ordered in 8x8 pixel tiles (one
  wave64 per tile) as the reference orders primary rays along a Morton curve.

Arrays are float32 with a float4 (16-byte) stride as the reference's TriangleData and
RayChunk (KernelDefs.h:461-513, RayPool.h:162-171).
"""

from __future__ import annotations

import dataclasses

import numpy as np

ALBEDO_DEFAULT = 0.7


def _f4(a: np.ndarray) -> np.ndarray:
    """(n, 3) -> (n, 4) float32 with w = 0."""
    out = np.zeros((a.shape[0], 4), dtype=np.float32)
    out[:, :3] = a
    return out


def icosphere(subdiv: int, radius: float = 1.0, center=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Triangles (n, 3, 3) of an icosphere with 20 * 4**subdiv faces, outward winding."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = np.array(
        [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
         [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    faces = np.array(
        [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
         [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
         [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
         [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    verts /= np.linalg.norm(verts, axis=1, keepdims=True)
    for _ in range(subdiv):
        v = [tuple(x) for x in verts]
        index = {}
        vlist = list(verts)

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in index:
                m = (vlist[a] + vlist[b]) * 0.5
                m = m / np.linalg.norm(m)
                index[key] = len(vlist)
                vlist.append(m)
            return index[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        del v
        verts = np.array(vlist)
        faces = np.array(nf, dtype=np.int64)
    tris = verts[faces] * radius + np.asarray(center, dtype=np.float64)
    return tris


def box_tris(lo=-1.0, hi=1.0) -> np.ndarray:
    """12 triangles of the closed box [lo, hi]^3, wound so normals face inward."""
    c = np.array([[x, y, z] for x in (lo, hi) for y in (lo, hi) for z in (lo, hi)], dtype=np.float64)
    # corner index = 4*xbit + 2*ybit + zbit
    quads = [
        (0, 1, 3, 2),  # x = lo, normal +x (inward)
        (4, 6, 7, 5),  # x = hi, normal -x
        (0, 4, 5, 1),  # y = lo, normal +y
        (2, 3, 7, 6),  # y = hi, normal -y
        (0, 2, 6, 4),  # z = lo, normal +z
        (1, 5, 7, 3),  # z = hi, normal -z
    ]
    tris = []
    for a, b, cc, d in quads:
        tris.append([c[a], c[b], c[cc]])
        tris.append([c[a], c[cc], c[d]])
    tris = np.array(tris)
    # orient inward: normal must point toward the box centre
    centre = np.full(3, 0.5 * (lo + hi))
    for t in tris:
        n = np.cross(t[1] - t[0], t[2] - t[0])
        if np.dot(n, centre - t.mean(axis=0)) < 0:
            t[[1, 2]] = t[[2, 1]]
    return tris


def light_quad(center_xz=(0.0, 0.0), size=0.5, y=1.0 - 1e-3) -> np.ndarray:
    """Two triangles of a size x size quad at height y facing -y."""
    cx, cz = center_xz
    h = size / 2.0
    p = np.array([[cx - h, y, cz - h], [cx + h, y, cz - h], [cx + h, y, cz + h], [cx - h, y, cz + h]])
    tris = np.array([[p[0], p[1], p[2]], [p[0], p[2], p[3]]])
    for t in tris:
        n = np.cross(t[1] - t[0], t[2] - t[0])
        if n[1] > 0:
            t[[1, 2]] = t[[2, 1]]
    return tris


def random_small_tris(n: int, seed: int, max_edge=0.05, lo=-1.0, hi=1.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    v0 = rng.uniform(lo, hi, size=(n, 3))

    def ball(m):
        d = rng.normal(size=(m, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        r = rng.uniform(0.0, 1.0, size=(m, 1)) ** (1.0 / 3.0)
        return d * r

    s1 = max_edge * (1.0 - rng.uniform(0.0, 1.0, size=(n, 1)))  # U(0, max]
    s2 = max_edge * (1.0 - rng.uniform(0.0, 1.0, size=(n, 1)))
    v1 = v0 + ball(n) * s1
    v2 = v0 + ball(n) * s2
    tris = np.stack([v0, v1, v2], axis=1)
    return np.clip(tris, lo, hi)


@dataclasses.dataclass
class Scene:
    """Triangle soup with per-triangle RGB albedo and quad lights."""

    v0: np.ndarray  # (n, 4) float32
    v1: np.ndarray
    v2: np.ndarray
    albedo: np.ndarray  # (n, 4) float32, rgb + alpha
    lights: np.ndarray  # (m, 4, 3) float32: corner, edge u, edge v, emitted power rgb
    name: str = ""

    @property
    def num_tris(self) -> int:
        return int(self.v0.shape[0])

    def bounds(self):
        allv = np.concatenate([self.v0[:, :3], self.v1[:, :3], self.v2[:, :3]])
        return allv.min(axis=0), allv.max(axis=0)

    def diag(self) -> float:
        lo, hi = self.bounds()
        return float(np.linalg.norm(hi.astype(np.float64) - lo.astype(np.float64)))

    def ray_epsilon(self) -> float:
        """SceneConfig::AssignDefault: epsilon = 5e-5 * |diagonal| (SceneConfig.cpp:111)."""
        return float(np.float32(self.diag() * 0.00005))

    def r_max(self) -> float:
        """SceneConfig::AssignDefault: maximum query radius 0.1 * |diagonal| (SceneConfig.cpp:86)."""
        return float(np.float32(self.diag() * 0.1))


def _assemble(parts, albedos, lights, name) -> Scene:
    tris = np.concatenate(parts).astype(np.float32)
    alb = np.concatenate([np.broadcast_to(np.asarray(a, np.float32), (p.shape[0], 4)) for p, a in zip(parts, albedos)])
    return Scene(_f4(tris[:, 0]), _f4(tris[:, 1]), _f4(tris[:, 2]), np.ascontiguousarray(alb, np.float32),
                 np.asarray(lights, np.float32), name)


def _light_record(quad: np.ndarray, power) -> np.ndarray:
    corner = quad[0][0]
    eu = quad[0][1] - quad[0][0]
    ev = quad[0][2] - quad[0][1]
    return np.array([corner, eu, ev, np.asarray(power, np.float64)])


def _light_positions(num_lights: int):
    if num_lights == 1:
        return [(0.0, 0.0)]
    ring = [(-0.5, -0.5), (0.5, 0.5), (0.5, -0.5), (-0.5, 0.5)]
    return ring[:num_lights]


def scene_small(subdiv: int = 5, num_lights: int = 1, light_power=(10.0, 10.0, 10.0)) -> Scene:
    """Box + centred icosphere (radius 0.5) + quad light(s): the plain small scene."""
    box = box_tris()
    sph = icosphere(subdiv, 0.5)
    parts, albs, lights = [box, sph], [(0.7, 0.7, 0.7, 1.0)] * 2, []
    for cxz in _light_positions(num_lights):
        q = light_quad(cxz)
        parts.append(q)
        albs.append((0.0, 0.0, 0.0, 1.0))
        lights.append(_light_record(q, light_power))
    return _assemble(parts, albs, lights, f"small-ico{subdiv}-l{num_lights}")


def scene_large(num_spheres=48, subdiv=5, num_random=50000, num_lights=1,
                light_power=(10.0, 10.0, 10.0)) -> Scene:
    """Box + 48 random icospheres + 50k small random triangles + light (~1.03M triangles)."""
    rng = np.random.default_rng(1)
    centres = rng.uniform(-1.0, 1.0, size=(num_spheres, 3))
    radii = rng.uniform(0.05, 0.3, size=num_spheres)
    unit = icosphere(subdiv, 1.0)
    parts = [box_tris()] + [unit * r + c for c, r in zip(centres, radii)]
    parts.append(random_small_tris(num_random, seed=2))
    albs = [(0.7, 0.7, 0.7, 1.0)] * len(parts)
    lights = []
    for cxz in _light_positions(num_lights):
        q = light_quad(cxz)
        parts.append(q)
        albs.append((0.0, 0.0, 0.0, 1.0))
        lights.append(_light_record(q, light_power))
    return _assemble(parts, albs, lights, "large")


def scene_sized(num_tris: int, seed: int, num_lights: int = 1) -> Scene:
    """A small-scene variant with about num_tris triangles (batched workloads): box,
    light, and an icosphere plus random triangles making up the count."""
    rng = np.random.default_rng(seed)
    sub = 0
    while 20 * 4 ** (sub + 1) <= num_tris // 2 and sub < 6:
        sub += 1
    sph = icosphere(sub, float(rng.uniform(0.3, 0.6)), rng.uniform(-0.3, 0.3, size=3))
    rest = max(0, num_tris - 12 - 2 * num_lights - sph.shape[0])
    parts = [box_tris(), sph]
    if rest:
        parts.append(random_small_tris(rest, seed=seed + 7919))
    albs = [(0.7, 0.7, 0.7, 1.0)] * len(parts)
    lights = []
    for cxz in _light_positions(num_lights):
        q = light_quad(cxz)
        parts.append(q)
        albs.append((0.0, 0.0, 0.0, 1.0))
        lights.append(_light_record(q, (10.0, 10.0, 10.0)))
    return _assemble(parts, albs, lights, f"sized-{num_tris}-{seed}")


def camera_basis(width: int, height: int, eye=(0.0, 0.0, 0.9), target=(0.0, 0.0, 0.0), vfov_deg=90.0):
    """Camera frame of camera_rays as 14 doubles: eye, unit forward, unit right, up,
    tan(vfov / 2), aspect (the h_cam argument of mnrt_primary_rays)."""
    eye = np.asarray(eye, np.float64)
    fwd = np.asarray(target, np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    up0 = np.array([0.0, 1.0, 0.0])
    right = np.cross(fwd, up0)
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    th = np.tan(np.radians(vfov_deg) / 2.0)
    return np.concatenate([eye, fwd, right, up, [th, width / height]]).astype(np.float64)


def camera_rays(width: int, height: int, eye=(0.0, 0.0, 0.9), target=(0.0, 0.0, 0.0), vfov_deg=90.0,
                tile: int = 8):
    """Pinhole primary rays through pixel centres (raygen.cu:171-240), tile-ordered.

    Returns (origins (n,4), dirs (n,4), pixel index (n,) uint32)."""
    cam = camera_basis(width, height, eye, target, vfov_deg)
    eye, fwd, right, up = cam[0:3], cam[3:6], cam[6:9], cam[9:12]
    th, aspect = cam[12], cam[13]
    ty, tx = np.divmod(np.arange(((height + tile - 1) // tile) * ((width + tile - 1) // tile)),
                       (width + tile - 1) // tile)
    iy, ix = np.divmod(np.arange(tile * tile), tile)
    px = (tx[:, None] * tile + ix[None, :]).ravel()
    py = (ty[:, None] * tile + iy[None, :]).ravel()
    keep = (px < width) & (py < height)
    px, py = px[keep], py[keep]
    sx = (2.0 * (px + 0.5) / width - 1.0) * th * aspect
    sy = (1.0 - 2.0 * (py + 0.5) / height) * th
    d = fwd[None, :] + sx[:, None] * right[None, :] + sy[:, None] * up[None, :]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    n = d.shape[0]
    o = np.broadcast_to(eye, (n, 3))
    return _f4(o.astype(np.float32)), _f4(d.astype(np.float32)), (py * width + px).astype(np.uint32)


def random_surface_photons(scene: Scene, n: int, seed: int):
    """Photons scattered uniformly over the scene's surface area with random powers and
    incoming directions: an input set for kNN tests that needs no photon tracing."""
    rng = np.random.default_rng(seed)
    v0 = scene.v0[:, :3].astype(np.float64)
    e1 = scene.v1[:, :3] - v0
    e2 = scene.v2[:, :3] - v0
    area = 0.5 * np.linalg.norm(np.cross(e1, e2), axis=1)
    p = area / area.sum()
    t = rng.choice(scene.num_tris, size=n, p=p)
    u = rng.uniform(size=n)
    v = rng.uniform(size=n)
    flip = u + v > 1
    u[flip], v[flip] = 1 - u[flip], 1 - v[flip]
    pos = v0[t] + u[:, None] * e1[t] + v[:, None] * e2[t]
    power = rng.uniform(0.01, 1.0, size=(n, 3)) / n
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return _f4(pos.astype(np.float32)), _f4(power.astype(np.float32)), _f4(d.astype(np.float32))


def batched_workload(num_scenes: int, seed: int, query_hi: int = 1 << 20, num_lights: int = 1,
                     scene_seed: int = 4):
    """A many-scene workload: per scene a scene_sized scene with a triangle count
    log-uniform in [1k, 50k], a photon count log-uniform in [4k, 256k] and a pinhole camera
    of about a query count log-uniform in [16k, query_hi] pixels (width a multiple of 8).
    Returns (scenes, photon counts, (W, H) per scene), all drawn from `seed`; scene i uses
    seed scene_seed + i."""
    rng = np.random.default_rng(seed)

    def lu(lo, hi):
        return np.exp(rng.uniform(np.log(lo), np.log(hi), num_scenes))

    ntri, nph, nq = lu(1000, 50000).astype(int), lu(4096, 262144).astype(int), lu(16384, query_hi).astype(int)
    scenes = [scene_sized(int(n), seed=scene_seed + i, num_lights=num_lights) for i, n in enumerate(ntri)]
    dims = []
    for q in nq:
        w = 8 * max(1, int(round(np.sqrt(q) / 8)))
        dims.append((w, max(1, int(round(q / w)))))
    return scenes, [int(p) for p in nph], dims
