"""Device primary rays (mnrt_primary_rays) equal the host generator bit for bit,
including clipped edge tiles."""

import numpy as np
import pytest

from mnrt_amd import api, scene as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("w,h,tile", [(37, 21, 8), (512, 512, 8), (50, 7, 4), (1, 1, 8), (1920, 1080, 16)])
def test_primary_rays_equal_host(cuda, w, h, tile):
    print(f"primary rays {w}x{h} tile {tile}", flush=True)
    o, d, pix = S.camera_rays(w, h, tile=tile)
    go, gd, gp = api.primary_rays(w, h, tile=tile, want_pixel=True)
    assert np.array_equal(go.cpu().numpy().view(np.uint32), o.view(np.uint32))
    assert np.array_equal(gd.cpu().numpy().view(np.uint32), d.view(np.uint32))
    assert np.array_equal(gp.cpu().numpy().view(np.uint32), pix)


def test_primary_rays_other_camera(cuda):
    kw = dict(eye=(0.3, -0.2, 0.5), target=(-0.1, 0.4, -0.7), vfov_deg=60.0)
    o, d, _ = S.camera_rays(64, 40, **kw)
    go, gd, _ = api.primary_rays(64, 40, **kw)
    assert np.array_equal(gd.cpu().numpy().view(np.uint32), d.view(np.uint32))
    assert np.array_equal(go.cpu().numpy().view(np.uint32), o.view(np.uint32))
