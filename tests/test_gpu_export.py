"""Reference-format export on the GPU equals the oracle's restatement word for word (and
the extents bit for bit); custom bits land in inner nodes' parent words only."""

import numpy as np
import pytest

from mnrt_amd import api, scene as S

pytestmark = pytest.mark.gpu


def _host(t):
    return t.cpu().numpy()


@pytest.mark.parametrize("kind", ["tri", "points"])
def test_export_equals_oracle(orc, cuda, kind):
    if kind == "tri":
        sc = S.scene_sized(20000, seed=3)
        T = api.build_tri(sc.v0, sc.v1, sc.v2)
    else:
        rng = np.random.default_rng(1)
        pos = np.concatenate([rng.uniform(-1, 1, (50000, 3)), np.zeros((50000, 1))], 1).astype(np.float32)
        T = api.build_points(pos, api.default_params(1, np.array([-1, -1, -1, 1, 1, 1], np.float32), 0.1))
    print(f"export {kind}: {T.num_nodes} nodes", flush=True)
    X = api.export_preorder(T)
    cuda.cuda.synchronize()
    O = orc.export_preorder(T.nodes_host(), T.refs_host(), T.root, T.t.num_leaves)
    assert np.array_equal(_host(X["tree"]).view(np.uint32), O["tree"])
    for k in ("num_elems", "addr", "left", "right"):
        assert np.array_equal(_host(X[k]).view(np.uint32), O[k]), k
    assert np.array_equal(_host(X["extent"]).view(np.uint32), O["extent"].view(np.uint32))
    # custom bit 1 = node index parity on inner nodes
    vals = (np.arange(T.num_nodes) & 1).astype(np.uint32)
    api.set_custom_bits(X, 1, vals)
    cuda.cuda.synchronize()
    tr = _host(X["tree"]).view(np.uint32)
    addr = O["addr"]
    leaf = (O["tree"][addr] & 0x80000000) != 0
    inner = np.nonzero(~leaf)[0]
    assert np.array_equal((tr[addr[inner] + 1] >> 29) & 1, vals[inner])
    assert np.array_equal(tr[addr[inner] + 1] & ~np.uint32(1 << 29), O["tree"][addr[inner] + 1])
    mask = np.ones(tr.size, bool)
    mask[addr[inner] + 1] = False
    assert np.array_equal(tr[mask], O["tree"][mask])
    T.release()
