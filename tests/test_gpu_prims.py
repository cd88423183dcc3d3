"""The parallel primitives (mnrt_scan_u32, mnrt_scan_u32_counted, mnrt_compact_u32; MNCudaPrimitives::Scan /
::Compact, MNCudaPrimitives.cpp:160-190) against plain sequential loops, at the sizes where
a tiled scan can go wrong: empty, one element, around a wave (63 / 64 / 65), around a
2048-element tile, the one-workgroup scans of 2049..65536 elements and their upper end,
many tiles (a second scan level) and more than 2048 tiles (a third), with sums that wrap
modulo 2^32."""

import numpy as np
import pytest

from mnrt_amd import api

pytestmark = pytest.mark.gpu

TILE = 2048
SIZES = [0, 1, 63, 64, 65, TILE - 1, TILE, TILE + 1, 4 * TILE + 7, 32 * TILE - 1, 32 * TILE, 32 * TILE + 1,
         64 * TILE + 5, TILE * TILE + 3]


def _seq_scan(a, inclusive):
    out = np.zeros(len(a), np.uint32)
    run = 0
    for i, v in enumerate(a.tolist()):
        if inclusive:
            run = (run + v) & 0xFFFFFFFF
            out[i] = run
        else:
            out[i] = run
            run = (run + v) & 0xFFFFFFFF
    return out, run


def _seq_compact(a, valid):
    out = []
    for v, f in zip(a.tolist(), valid.tolist()):
        if f:
            out.append(v)
    return np.array(out, np.uint32)


@pytest.mark.parametrize("n", SIZES)
def test_scan_vs_sequential_loop(cuda, n):
    torch = cuda
    rng = np.random.default_rng(n)
    # large words near the top of the range make the running sums wrap
    a = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32) if n < 70000 else \
        rng.integers(0, 1 << 20, n, dtype=np.uint64).astype(np.uint32)
    if n > 3:
        a[-3:] = 0xFFFFFFF0
    d = torch.from_numpy(a.view(np.int32)).cuda()
    for inclusive in (False, True):
        want, tot = _seq_scan(a, inclusive)
        got, gt = api.scan_u32(d, inclusive)
        assert np.array_equal(got.cpu().numpy().view(np.uint32), want), (n, inclusive)
        assert int(gt.cpu().numpy().view(np.uint32)[0]) == tot
        inplace = d.clone()
        api.scan_u32(inplace, inclusive, out=inplace)
        assert np.array_equal(inplace.cpu().numpy().view(np.uint32), want), (n, inclusive, "in place")
    print(f"scan n={n}: exclusive, inclusive, in place equal", flush=True)


@pytest.mark.parametrize("n", SIZES)
def test_compact_vs_sequential_loop(cuda, n):
    torch = cuda
    rng = np.random.default_rng(1000 + n)
    a = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    for density in (0.0, 0.013, 0.5, 1.0):
        valid = (rng.uniform(size=n) < density).astype(np.uint32)
        want = _seq_compact(a, valid)
        out, cnt = api.compact_u32(torch.from_numpy(a.view(np.int32)).cuda(),
                                   torch.from_numpy(valid.view(np.int32)).cuda())
        c = int(cnt.cpu().numpy()[0])
        assert c == len(want), (n, density)
        assert np.array_equal(out[:c].cpu().numpy().view(np.uint32), want), (n, density)
    print(f"compact n={n}: four densities equal", flush=True)


# (capacity, device count): the one-workgroup form (capacity <= 65536) and the two-launch
# form above it, counts of 0, 1, a tile and its neighbours, counts far below the capacity
# (blocks with no tile), more tiles than blocks (1024), and counts beyond the capacity
COUNTED = [(0, 0), (1, 0), (1, 1), (TILE, TILE), (TILE + 1, 1), (65536, 65536), (65536, 2049), (65537, 65537),
           (65537, 0), (65537, 2048), (65537, 2049), (64 * TILE + 5, 64 * TILE + 5), (TILE * TILE + 3, TILE * TILE + 3),
           (TILE * TILE + 3, 1500 * TILE + 17), (5000, 0xFFFFFFFF), (70000, 70001)]


@pytest.mark.parametrize("cap,count", COUNTED)
def test_counted_scan_vs_sequential_loop(cuda, cap, count):
    torch = cuda
    n = min(count, cap)
    rng = np.random.default_rng(cap + 7 * n)
    a = rng.integers(0, 1 << 32, cap, dtype=np.uint64).astype(np.uint32) if cap < 70000 else \
        rng.integers(0, 1 << 20, cap, dtype=np.uint64).astype(np.uint32)
    if n > 3:
        a[n - 3:n] = 0xFFFFFFF0
    d = torch.from_numpy(a.view(np.int32)).cuda()
    want, tot = _seq_scan(a[:n], False)
    for end_total in (False, True):
        got, gt = api.scan_u32_counted(d, count, cap, end_total)
        g = got.cpu().numpy().view(np.uint32)
        assert np.array_equal(g[:n], want), (cap, count, end_total)
        assert int(gt.cpu().numpy().view(np.uint32)[0]) == tot
        rest = g[n + 1:] if end_total else g[n:]
        assert np.all(rest == 0xFFFFFFFF), (cap, count, end_total, "past the count")
        if end_total:
            assert int(g[n]) == tot
        inplace = torch.cat([d, torch.full((1,), -1, dtype=torch.int32, device="cuda")])
        api.scan_u32_counted(inplace, count, cap, end_total, out=inplace)
        g = inplace.cpu().numpy().view(np.uint32)
        assert np.array_equal(g[:n], want), (cap, count, end_total, "in place")
        assert np.array_equal(g[n + 1 if end_total else n:cap], a[n + 1 if end_total else n:cap])
    print(f"counted scan cap={cap} count={count}: equal, nothing written past the count", flush=True)
