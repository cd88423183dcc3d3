"""Oracle photon tracing pins: exact power bookkeeping (emission x light count x albedo
per bounce, Russian roulette doubling, 1 / paths scaling), cosine emission statistics,
Russian-roulette survival, tags."""

import numpy as np

from mnrt_amd import scene as S


def test_photon_power_bookkeeping(orc):
    sc = S.scene_small(subdiv=3, num_lights=2)
    eps = sc.ray_epsilon()
    T = orc.build(0, sc.v0, sc.v1, sc.v2)
    ph = orc.photons(T, sc.v0, sc.v1, sc.v2, sc.albedo, sc.lights, 20000, 4, 3, eps)
    assert ph["count"] == 20000
    bounce = ph["meta"] >> 8
    light = ph["meta"] & 0xFF
    assert np.array_equal(light, ph["path"] % 2)
    scale = np.float32(1.0) / np.float32(ph["paths"])
    for b in range(5):
        sel = bounce == b
        p = np.float32(sc.lights[0, 3, 0]) * np.float32(2.0)
        for _ in range(b):
            p = p * np.float32(0.7)
        if b == 4:
            p = p * np.float32(2.0)
        p = p * scale
        assert np.all(ph["power"][sel, :3] == p), b
    d = ph["dir"][:, :3].astype(np.float64)
    assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 1e-5
    lo, hi = sc.bounds()
    assert (ph["pos"][:, :3] >= lo - 1e-5).all() and (ph["pos"][:, :3] <= hi + 1e-5).all()
    # path triangles: the b-th photon of a path lists b earlier scattering triangles
    ntri = (ph["tris"] != 0xFFFFFFFF).sum(axis=1)
    assert np.array_equal(ntri, np.minimum(bounce, 4))


def test_photon_emission_cosine_and_roulette(orc):
    sc = S.scene_small(subdiv=3)
    T = orc.build(0, sc.v0, sc.v1, sc.v2)
    ph = orc.photons(T, sc.v0, sc.v1, sc.v2, sc.albedo, sc.lights, 100000, 4, 1, sc.ray_epsilon())
    bounce = ph["meta"] >> 8
    first = ph["dir"][bounce == 0, :3].astype(np.float64)
    cos = -first[:, 1]  # light faces -y
    assert (cos > 0).all()
    n = cos.size
    assert abs(cos.mean() - 2.0 / 3.0) < 4 * np.sqrt(1 / 18 / n)
    assert abs((cos ** 2).mean() - 0.5) < 4 * np.sqrt(1 / 12 / n)
    c3, c4 = (bounce == 3).sum(), (bounce == 4).sum()
    assert 0.44 < c4 / c3 < 0.53  # Russian roulette keeps half (the light absorbs a few)
