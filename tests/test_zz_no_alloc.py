"""mnrt_knn_gather_noalloc (the entry point that allocates nothing: scratch in a fixed
device array) against mnrt_knn_gather through the C ABI: radiance, r2, counts, neighbour
lists and, with slot indices, the summed slots and inverse distances equal bit for bit, on
two calls in a row (the second with fewer queries, so stale scratch from the first would
show); sizes beyond the array are refused before any launch."""

import ctypes

import numpy as np
import pytest

from mnrt_amd import api, native as N, scene as S

pytestmark = pytest.mark.gpu

K = 50


def _outs(torch, n, k, slots):
    f = lambda *s: torch.full(s, 3.0, dtype=torch.float32, device="cuda")
    i = lambda *s: torch.full(s, 7, dtype=torch.int32, device="cuda")
    return dict(rad=f(slots, 4), inv=f(slots), r2=f(n), irr=f(n, 4), nbr=i(n * k), cnt=i(n))


def _call(fn, P, rmax, q, o, idx_sp=None, orig=None):
    n = q["pos"].shape[0]
    p = N.ptr
    rc = fn(P.ref(), K, rmax, n, p(q["pos"]), p(q["nrm"]), p(q["dir"]), p(q["infl"]), p(q["clr"]), p(idx_sp), None,
            p(orig), p(o["rad"]), p(o["inv"]) if orig is not None else None, p(o["r2"]), p(o["irr"]),
            p(o["nbr"]) if idx_sp is None else None, p(o["cnt"]), N.stream_handle())
    return rc


def _same(torch, a, b, keys):
    torch.cuda.synchronize()
    for k in keys:
        assert torch.equal(a[k].view(torch.int32), b[k].view(torch.int32)), k


def test_noalloc_equals_gather_twice(cuda):
    torch = cuda
    L = N.load()
    sc = S.scene_small(subdiv=4)
    eps, rmax = sc.ray_epsilon(), sc.r_max()
    v = [api.dev(a) for a in (sc.v0, sc.v1, sc.v2, sc.albedo)]
    T = api.build_tri(*v[:3])
    ph = api.photons_generate(T, *v, sc.lights, 100000, 4, 0, eps)
    o, d, _ = api.primary_rays(256, 256)
    tri, t, _, _ = api.trace_closest(T, o, d, eps)
    qa = api.shading_points(o, d, tri, t, *v)
    root = np.concatenate(sc.bounds()).astype(np.float32)
    P = api.build_points(ph["pos"], api.default_params(1, root, rmax))
    api.attach_photon_records(P, ph["pos"], ph["power"], ph["dir"])
    n1 = qa["count"]
    assert n1 > 40000  # both calls above the size where queries are reordered
    for call, (lo, hi) in enumerate(((0, n1), (1000, 41000))):
        q = {k: qa[k][lo:hi].contiguous() for k in ("pos", "nrm", "dir", "infl", "clr")}
        n = hi - lo
        print(f"call {call}: {n} queries", flush=True)
        # per-query outputs
        a, b = _outs(torch, n, K, n), _outs(torch, n, K, n)
        assert _call(L.mnrt_knn_gather, P, rmax, q, a) == N.OK
        assert _call(L.mnrt_knn_gather_noalloc, P, rmax, q, b) == N.OK
        _same(torch, a, b, ("rad", "r2", "irr", "nbr", "cnt"))
        assert bool((b["cnt"] > 0).any())
        # slots shared by runs of consecutive queries, inverse distances accumulated
        slots = n // 3 + 1
        idx = torch.from_numpy((np.arange(n) // 3).astype(np.int32)).cuda()
        orig = o[lo:hi].contiguous()
        a, b = _outs(torch, n, K, slots), _outs(torch, n, K, slots)
        assert _call(L.mnrt_knn_gather, P, rmax, q, a, idx, orig) == N.OK
        assert _call(L.mnrt_knn_gather_noalloc, P, rmax, q, b, idx, orig) == N.OK
        _same(torch, a, b, ("rad", "inv", "r2", "irr", "cnt"))
    # beyond the fixed array: refused before any launch
    q = {k: qa[k].contiguous() for k in ("pos", "nrm", "dir", "infl", "clr")}
    b = _outs(torch, 16, K, 16)
    p = N.ptr
    big = 2097114
    rc = L.mnrt_knn_gather_noalloc(P.ref(), K, rmax, big, p(q["pos"]), p(q["nrm"]), p(q["dir"]), p(q["infl"]),
                                   p(q["clr"]), p(q["pos"]), None, None, p(b["rad"]), None, None, None, None, None,
                                   N.stream_handle())
    assert rc == N.ERR_CAPACITY
    rc = L.mnrt_knn_gather_noalloc(P.ref(), K, rmax, 7339898, p(q["pos"]), p(q["nrm"]), p(q["dir"]), p(q["infl"]),
                                   p(q["clr"]), None, None, None, p(b["rad"]), None, None, None, None, None,
                                   N.stream_handle())
    assert rc == N.ERR_CAPACITY
    torch.cuda.synchronize()
    T.release()
    P.release()
