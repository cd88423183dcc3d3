"""GPU photon tracing vs the oracle's photon tracer (same RNG, same sampling, same
traversal): photon positions, powers, directions, light/bounce tags, source paths and
path triangles bit-exact; shading points and bounce rays vs a numpy restatement."""

import numpy as np
import pytest

from mnrt_amd import api, scene as S

pytestmark = pytest.mark.gpu


def _u(a):
    return a.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("lights", [1, 3])
def test_photons_bit_exact(orc, cuda, lights):
    sc = S.scene_small(subdiv=4, num_lights=lights)
    eps = sc.ray_epsilon()
    T = api.build_tri(sc.v0, sc.v1, sc.v2)
    ot = orc.build(0, sc.v0, sc.v1, sc.v2)
    print(f"photons: lights={lights}", flush=True)
    g = api.photons_generate(T, sc.v0, sc.v1, sc.v2, sc.albedo, sc.lights, 65536, 4, 7, eps, want_meta=True)
    o = orc.photons(ot, sc.v0, sc.v1, sc.v2, sc.albedo, sc.lights, 65536, 4, 7, eps)
    cuda.cuda.synchronize()
    assert g["count"] == o["count"] == 65536 and g["paths"] == o["paths"]
    for key in ("pos", "power", "dir"):
        assert np.array_equal(_u(g[key]), o[key].view(np.uint32)), key
    assert np.array_equal(_u(g["meta"]), o["meta"])
    assert np.array_equal(_u(g["path"]), o["path"])
    assert np.array_equal(_u(g["tris"]), o["tris"])
    bounces = o["meta"] >> 8
    assert bounces.max() == 4 and (bounces == 0).sum() > 0
    assert set(np.unique(o["meta"] & 0xFF)) == set(range(lights))


def test_shading_points_and_bounce(orc, cuda):
    sc = S.scene_small(subdiv=3)
    eps = sc.ray_epsilon()
    T = api.build_tri(sc.v0, sc.v1, sc.v2)
    o, d, _ = S.camera_rays(96, 64)
    # rays that escape through a hole: remove the front wall (z = +1) so some rays miss
    tri, t, _, _ = api.trace_closest(T, o, d, eps)
    dv = [api.dev(a) for a in (sc.v0, sc.v1, sc.v2, sc.albedo)]
    q = api.shading_points(api.dev(o), api.dev(d), tri, t, *dv, want_src=True)
    cuda.cuda.synchronize()
    tri_h, t_h = tri.cpu().numpy(), t.cpu().numpy()
    hit = np.nonzero(tri_h >= 0)[0]
    assert q["count"] == hit.size
    assert np.array_equal(q["src"].cpu().numpy(), hit)
    pos = o[hit, :3] + d[hit, :3] * t_h[hit, None]
    assert np.array_equal(q["pos"].cpu().numpy()[:, :3].view(np.uint32), pos.astype(np.float32).view(np.uint32))
    tr = tri_h[hit]
    e1 = (sc.v1[tr, :3] - sc.v0[tr, :3]).astype(np.float64)
    e2 = (sc.v2[tr, :3] - sc.v0[tr, :3]).astype(np.float64)
    n = np.cross(e1, e2)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    assert np.abs(q["nrm"].cpu().numpy()[:, :3] - n).max() < 1e-6
    assert np.array_equal(q["clr"].cpu().numpy(), sc.albedo[tr])
    bo, bd, binfl = api.bounce_rays(q, 5)
    bd_h = bd.cpu().numpy()[:, :3]
    assert np.abs(np.linalg.norm(bd_h, axis=1) - 1).max() < 1e-5
    # bounce directions leave on the side the primary ray came from
    nf = q["nrm"].cpu().numpy()[:, :3]
    side = np.sign(np.sum(-q["dir"].cpu().numpy()[:, :3] * nf, axis=1))
    assert (np.sum(bd_h * nf, axis=1) * side >= -1e-6).all()
    assert np.array_equal(binfl.cpu().numpy()[:, :3], q["clr"].cpu().numpy()[:, :3])
