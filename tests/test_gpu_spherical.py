"""The reference's spherical photon-direction convention (PhotonData keeps phi in
position.w and theta in power.w; dev_Spherical2Direction / dev_Direction2Spherical,
mncudautil_dev.h:345-375) against fp64 closed forms of the same fp32 inputs:
d = (sin t cos p, sin t sin p, cos t); p = atan2(y, x) (+ 2 pi when negative), t = acos(z).
Bounds: a few ulps of the results (HIP's sinf / cosf / atan2f / acosf, one product).
Axis directions and poles come out exact, the azimuth wraps into [0, 2 pi] (a direction
just below the +x axis gives fp32(2 pi)), and a round trip returns the direction."""

import numpy as np
import pytest

from mnrt_amd import api

pytestmark = pytest.mark.gpu

ULP1 = 2.0 ** -23  # ulp of numbers in [1, 2)


def _angles(phi, theta):
    n = len(phi)
    pos = np.zeros((n, 4), np.float32)
    pw = np.zeros((n, 4), np.float32)
    pos[:, 3], pw[:, 3] = phi, theta
    return pos, pw


def test_spherical_to_direction_closed_form(cuda):
    rng = np.random.default_rng(0)
    n = 200000
    phi = rng.uniform(0, 2 * np.pi, n).astype(np.float32)
    theta = rng.uniform(0, np.pi, n).astype(np.float32)
    # axes and poles: phi 0, pi/2, pi, 3pi/2; theta 0, pi/2, pi
    ax_p = np.array([0, np.pi / 2, np.pi, 1.5 * np.pi] * 3, np.float32)
    ax_t = np.repeat(np.array([0, np.pi / 2, np.pi], np.float32), 4)
    phi, theta = np.concatenate([phi, ax_p]), np.concatenate([theta, ax_t])
    pos, pw = _angles(phi, theta)
    d = api.photons_spherical_to_dir(pos, pw).cpu().numpy()
    p64, t64 = phi.astype(np.float64), theta.astype(np.float64)
    want = np.stack([np.sin(t64) * np.cos(p64), np.sin(t64) * np.sin(p64), np.cos(t64)], axis=1)
    err = np.abs(d[:, :3].astype(np.float64) - want).max()
    print(f"spherical -> direction: max abs err {err:.2e} ({err / ULP1:.2f} ulp of 1)", flush=True)
    assert err <= 4 * ULP1
    assert (d[:, 3] == 0).all()
    # theta = 0 is the +z pole exactly, whatever phi
    pole = theta == 0
    assert np.array_equal(d[pole, :3], np.tile(np.array([0, 0, 1], np.float32), (pole.sum(), 1)))
    # phi = 0 gives y = 0 exactly
    assert (d[phi == 0, 1] == 0).all()


def test_direction_to_spherical_closed_form_and_wrap(cuda):
    rng = np.random.default_rng(1)
    n = 200000
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    extra = np.array([[1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                      [1, -1e-7, 0], [1, -0.0, 0], [-1, -1e-7, 0]], np.float64)
    v = np.concatenate([v, extra]).astype(np.float32)
    dirs = np.zeros((len(v), 4), np.float32)
    dirs[:, :3] = v
    pos = np.full((len(v), 4), 7.0, np.float32)
    pw = np.full((len(v), 4), 9.0, np.float32)
    import torch
    gpos, gpw = torch.from_numpy(pos).cuda(), torch.from_numpy(pw).cuda()
    api.photons_dir_to_spherical(dirs, gpos, gpw)
    p, t = gpos.cpu().numpy(), gpw.cpu().numpy()
    assert (p[:, :3] == 7.0).all() and (t[:, :3] == 9.0).all()  # only w written
    phi, theta = p[:, 3].astype(np.float64), t[:, 3].astype(np.float64)
    v64 = v.astype(np.float64)
    p64 = np.arctan2(v64[:, 1], v64[:, 0])
    p64 = np.where(p64 < 0, p64 + 2 * np.pi, p64)
    t64 = np.arccos(np.clip(v64[:, 2], -1, 1))
    ep = np.abs(phi - p64).max()
    et = np.abs(theta - t64).max()
    print(f"direction -> spherical: phi err {ep:.2e}, theta err {et:.2e}", flush=True)
    assert ep <= 8 * ULP1 and et <= 4 * ULP1
    assert (phi >= 0).all() and (phi <= np.float32(2 * np.pi)).all()
    assert (theta >= 0).all() and (theta <= np.float32(np.pi)).all()
    k = len(v) - len(extra)
    assert phi[k + 0] == 0 and theta[k + 4] == 0  # +x: phi 0; +z: theta 0
    assert abs(phi[k + 1] - np.pi / 2) <= 2 * ULP1 and abs(phi[k + 3] - 1.5 * np.pi) <= 4 * ULP1
    assert abs(theta[k + 5] - np.pi) <= 2 * ULP1
    assert phi[k + 6] == np.float32(2 * np.pi)  # just below +x: wraps to fp32(2 pi)
    assert phi[k + 7] == 0  # y = -0: atan2 gives -0, not negative, no wrap
    assert abs(phi[k + 8] - np.pi) <= 2 * ULP1
    # round trip to the original direction.  theta = acos(z) reads only z, so the part of
    # the fp32 input's length error (|v|^2 - 1, a few 1e-8) that lies across the axis comes
    # back divided by sin(theta) = |(x, y)|: ill-conditioned only near the poles
    back = api.photons_spherical_to_dir(gpos, gpw).cpu().numpy()[:, :3].astype(np.float64)
    err = np.abs(back - v64).max(axis=1)
    lerr = np.abs((v64 ** 2).sum(axis=1) - 1.0)
    bound = 64 * ULP1 + 4.0 * lerr / np.maximum(np.hypot(v64[:, 0], v64[:, 1]), 1e-30)
    print(f"round trip: max abs err {err.max():.2e}, max err / bound {np.max(err / bound):.3f}", flush=True)
    assert np.all(err <= bound)
    away = np.abs(v64[:, 2]) < 0.99
    assert err[away].max() <= 64 * ULP1
