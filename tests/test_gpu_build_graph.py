"""Batched builds recorded into a hipGraph (mnrt_build_graph_*): the device-driven build
gives the trees, records and query-time nodes of the host-driven batched build word for
word, and the oracle's trees; graph replays rebuild them in place, also after the input
changed (trees that move, root boxes that change); a replay that outgrows the graph's
buffers reports it instead of writing past them; a query batch over the recorded trees,
with the builds and the query phases all replayed from graphs, equals the eager path."""

import numpy as np
import pytest

from mnrt_amd import api, native as N, scene as S

pytestmark = pytest.mark.gpu


def _concat(parts, width=4):
    parts = [p if p is not None else np.zeros((0, width), np.float32) for p in parts]
    off = np.zeros(len(parts) + 1, np.uint32)
    off[1:] = np.cumsum([len(p) for p in parts])
    return np.concatenate(parts).astype(np.float32), off


def _tri_input():
    """Scenes of these sizes (empty, one triangle, small-root only, at and above the
    small-node limit, large), two large ones of equal size (swapped by a test)."""
    sizes = [3000, 0, 1, 40, 20000, 64, 65, 7000, 12, 7000]
    parts = [[], [], []]
    for i, n in enumerate(sizes):
        if n == 0:
            for p in parts:
                p.append(None)
            continue
        sc = S.scene_sized(max(n, 100), seed=70 + i)
        for p, a in zip(parts, (sc.v0, sc.v1, sc.v2)):
            p.append(a[:n])
    v = [_concat(p) for p in parts]
    return sizes, v[0][0], v[1][0], v[2][0], v[0][1]


def _same(a, b, what):
    assert a.num_nodes == b.num_nodes and a.num_refs == b.num_refs, what
    assert np.array_equal(a.nodes_host(), b.nodes_host()), what
    assert np.array_equal(a.refs_host(), b.refs_host()), what
    assert np.array_equal(a.root.view(np.uint32), b.root.view(np.uint32)), what
    assert a.max_level == b.max_level and a.t.num_leaves == b.t.num_leaves and a.t.num_elems == b.t.num_elems, what
    assert np.array_equal(a.recs_host().view(np.uint32), b.recs_host().view(np.uint32)), what
    assert np.array_equal(a.qnodes_host(), b.qnodes_host()), what


def _same_batch(G, B, what):
    for s in range(len(B)):
        _same(G.trees[s], B[s], f"{what}, scene {s}")


def test_graph_tri_equals_batch_and_oracle(orc, cuda):
    torch = cuda
    sizes, v0, v1, v2, off = _tri_input()
    print(f"recorded tri build: {len(sizes)} scenes, {off[-1]} triangles", flush=True)
    d0, d1, d2 = api.dev(v0), api.dev(v1), api.dev(v2)
    G = api.BuildGraph(N.KIND_TRI, off, d0, d1, d2)
    B = api.build_tri_batch(d0, d1, d2, off)
    torch.cuda.synchronize()
    _same_batch(G, B, "first build")
    for s, n in enumerate(sizes):
        if n == 0:
            continue
        a, b = off[s], off[s + 1]
        ot = orc.build(0, v0[a:b], v1[a:b], v2[a:b])
        assert np.array_equal(G.trees[s].nodes_host(), ot.nodes), s
        assert np.array_equal(G.trees[s].refs_host(), ot.refs), s
    # replays with other work in between rebuild the same trees in place
    nodes0 = [G.trees[s].t.d_nodes for s in range(len(sizes))]
    for rep in range(3):
        x = torch.randn(1 << 20, device="cuda")
        (x * 2.0).sum().item()
        G.replay()
        tmp = api.build_tri_batch(d0, d1, d2, off)
        assert G.status() == 0, rep
        assert [G.trees[s].t.d_nodes for s in range(len(sizes))] == nodes0
        _same_batch(G, B, f"replay {rep}")
        tmp.release()
    G.close()
    B.release()


def test_graph_tri_follows_changed_input(orc, cuda):
    """Vertices scaled by 2 (exact: same topology, doubled root boxes) and two equal-sized
    scenes swapped (same level sizes, so the buffers fit with no margin; the trees move)."""
    torch = cuda
    sizes, v0, v1, v2, off = _tri_input()
    d = [api.dev(a) for a in (v0, v1, v2)]
    G = api.BuildGraph(N.KIND_TRI, off, *d, margin=0.0)
    for t in d:
        t.mul_(2.0)
    G.replay()
    assert G.status() == 0
    B = api.build_tri_batch(*d, off)
    torch.cuda.synchronize()
    _same_batch(G, B, "scaled input")
    B.release()
    i, j = 7, 9  # both 7000 triangles
    assert sizes[i] == sizes[j]
    for t in d:
        a = t[int(off[i]):int(off[i + 1])].clone()
        t[int(off[i]):int(off[i + 1])] = t[int(off[j]):int(off[j + 1])]
        t[int(off[j]):int(off[j + 1])] = a
    G.replay()
    assert G.status() == 0
    B = api.build_tri_batch(*d, off)
    torch.cuda.synchronize()
    _same_batch(G, B, "swapped scenes")
    h = [x.cpu().numpy() for x in d]
    for s in (i, j):
        a, b = off[s], off[s + 1]
        ot = orc.build(0, h[0][a:b], h[1][a:b], h[2][a:b])
        assert np.array_equal(G.trees[s].nodes_host(), ot.nodes), s
    B.release()
    G.close()


def _photon_input(rng, sizes):
    pos, off = _concat([np.concatenate([rng.uniform(-1, 1, (n, 3)), np.zeros((n, 1))], 1) if n else None
                        for n in sizes])
    pw = rng.uniform(0, 1, (len(pos), 4)).astype(np.float32)
    dr = np.concatenate([rng.normal(size=(len(pos), 3)), np.zeros((len(pos), 1))], 1).astype(np.float32)
    dr[:, :3] /= np.linalg.norm(dr[:, :3], axis=1, keepdims=True)
    return pos, off, pw, dr


def test_graph_points_equals_batch_and_oracle(orc, cuda):
    torch = cuda
    rng = np.random.default_rng(8)
    sizes = [20000, 3, 0, 32, 33, 4000, 100000]
    pos, off, pw, dr = _photon_input(rng, sizes)
    boxes = np.tile(np.array([-1, -1, -1, 1, 1, 1], np.float32), (len(sizes), 1))
    qr = np.array([0.1, 0.2, 0.05, 0.1, 0.3, 0.02, 0.1], np.float32)
    print(f"recorded point build: {len(sizes)} maps, {off[-1]} photons", flush=True)
    dp, dw, dd = api.dev(pos), api.dev(pw), api.dev(dr)
    G = api.BuildGraph(N.KIND_POINT, off, dp, power=dw, dirs=dd, root_boxes=boxes, max_query_radius=qr)
    B = api.build_points_batch(dp, off, root_boxes=boxes, max_query_radius=qr)
    api.attach_photon_records_batch(B, dp, dw, dd)
    torch.cuda.synchronize()
    _same_batch(G, B, "first build")
    for s, n in enumerate(sizes):
        if n == 0:
            continue
        a, b = off[s], off[s + 1]
        ot = orc.build(1, pos[a:b], max_qr=float(qr[s]), root=boxes[s])
        assert np.array_equal(G.trees[s].nodes_host(), ot.nodes), s
    for rep in range(3):
        dw.mul_(0.5)  # new powers: the replay rewrites the records from the same addresses
        G.replay()
        api.attach_photon_records_batch(B, dp, dw, dd)
        assert G.status() == 0, rep
        _same_batch(G, B, f"replay {rep}")
    # the in-place record refresh of the batch API also works on recorded trees
    dw.mul_(3.0)
    api.attach_photon_records_batch(G.trees, dp, dw, dd)
    api.attach_photon_records_batch(B, dp, dw, dd)
    torch.cuda.synchronize()
    _same_batch(G, B, "records refreshed")
    G.close()
    # recorded without records (no powers or directions): the trees alone
    G2 = api.BuildGraph(N.KIND_POINT, off, dp, root_boxes=boxes, max_query_radius=qr)
    B2 = api.build_points_batch(dp, off, root_boxes=boxes, max_query_radius=qr)
    G2.replay()
    assert G2.status() == 0
    torch.cuda.synchronize()
    for s in range(len(sizes)):
        assert not G2.trees[s].t.d_recs and not G2.trees[s].t.d_qnodes
        _same(G2.trees[s], B2[s], f"no records, scene {s}")
    G2.close()
    B2.release()
    B.release()


def test_graph_batched_step_equals_eager(orc, cuda):
    """Both builds and the trace / compaction / kNN phases replayed from graphs, the
    query batch made over the recorded trees: results equal the eager path over
    host-built trees."""
    torch = cuda
    sizes = [1200, 5000, 300, 20000, 2500]
    photons = [4096, 20000, 3000, 60000, 8000]
    scs = [S.scene_sized(n, seed=90 + i) for i, n in enumerate(sizes)]
    cat = lambda a: api.dev(np.concatenate([getattr(s, a) for s in scs]).astype(np.float32))
    v0, v1, v2, alb = cat("v0"), cat("v1"), cat("v2"), cat("albedo")
    toff = np.zeros(len(scs) + 1, np.uint32)
    toff[1:] = np.cumsum([s.num_tris for s in scs])
    TB = api.build_tri_batch(v0, v1, v2, toff)
    ph = [api.photons_generate(TB[s], sc.v0, sc.v1, sc.v2, sc.albedo, sc.lights, photons[s], 4, s,
                               sc.ray_epsilon()) for s, sc in enumerate(scs)]
    poff = np.zeros(len(scs) + 1, np.uint32)
    poff[1:] = np.cumsum([g["count"] for g in ph])
    ppos, ppow, pdir = (torch.cat([g[k] for g in ph]).contiguous() for k in ("pos", "power", "dir"))
    boxes = np.stack([np.concatenate(sc.bounds()) for sc in scs]).astype(np.float32)
    rmax = np.array([sc.r_max() for sc in scs], np.float32)
    PB = api.build_points_batch(ppos, poff, root_boxes=boxes, max_query_radius=rmax)
    api.attach_photon_records_batch(PB, ppos, ppow, pdir)
    GT = api.BuildGraph(N.KIND_TRI, toff, v0, v1, v2)
    GP = api.BuildGraph(N.KIND_POINT, poff, ppos, power=ppow, dirs=pdir, root_boxes=boxes, max_query_radius=rmax)
    res = [(32, 32), (48, 32), (16, 16), (40, 40), (24, 24)]
    rays = [S.camera_rays(w, h) for (w, h) in res]
    roff = np.zeros(len(scs) + 1, np.uint32)
    roff[1:] = np.cumsum([len(r[0]) for r in rays])
    o = api.dev(np.concatenate([r[0] for r in rays]))
    d = api.dev(np.concatenate([r[1] for r in rays]))
    eps = np.array([sc.ray_epsilon() for sc in scs], np.float32)
    k, n = 50, int(roff[-1])

    def io(Q):
        f4 = lambda: torch.zeros((n, 4), dtype=torch.float32, device="cuda")
        b = dict(orig4=o, dir4=d, v0=v0, v1=v1, v2=v2, albedo4=alb,
                 hit_tri=torch.empty(n, dtype=torch.int32, device="cuda"),
                 hit_t=torch.empty(n, dtype=torch.float32, device="cuda"),
                 sp_pos4=f4(), sp_nrm4=f4(), sp_dir4=f4(), sp_clr4=f4(), sp_infl4=f4(),
                 sp_count=torch.zeros(len(scs), dtype=torch.int32, device="cuda"), io_rad4=f4())
        Q.make_io(**b)
        return b

    QE = api.QueryBatch(TB, toff, roff, eps, PB, rmax, k)
    BE = io(QE)
    QE.run(7)
    QG = api.QueryBatch(GT.trees, toff, roff, eps, GP.trees, rmax, k)
    BG = io(QG)
    for ph_ in (1, 2, 4):
        QG.capture(ph_)
    for rep in range(3):
        BG["io_rad4"].zero_()
        BG["hit_tri"].fill_(-2)
        GT.replay()
        GP.replay()
        for ph_ in (1, 2, 4):
            QG.replay(ph_)
        assert GT.status() == 0 and GP.status() == 0
        torch.cuda.synchronize()
        assert torch.equal(BG["hit_tri"], BE["hit_tri"]), rep
        assert torch.equal(BG["sp_count"], BE["sp_count"]), rep
        cnt = BE["sp_count"].cpu().numpy()
        for s in range(len(scs)):
            a = int(roff[s])
            m = int(cnt[s])
            assert torch.equal(BG["io_rad4"][a:a + m].view(torch.int32), BE["io_rad4"][a:a + m].view(torch.int32)), \
                (rep, s)
    QG.close()
    QE.close()
    GT.close()
    GP.close()
    TB.release()
    PB.release()


def test_graph_reports_outgrown_buffers(cuda):
    """Recorded with no margin on small scattered triangles, then replayed on large
    overlapping ones (many straddle every split, so the level loop's element arrays
    outgrow the recorded buffers): status reports it and the replay wrote nothing past
    the buffers (a later replay on the first input gives its trees again)."""
    torch = cuda
    rng = np.random.default_rng(3)
    n, ns = 3000, 4
    c = rng.uniform(-1, 1, (n * ns, 3))
    small = [c + rng.normal(scale=0.01, size=c.shape) for _ in range(3)]
    big = [rng.uniform(-1, 1, (n * ns, 3)) for _ in range(3)]
    w = lambda a: np.concatenate([a, np.zeros((len(a), 1))], 1).astype(np.float32)
    off = (np.arange(ns + 1) * n).astype(np.uint32)
    d = [api.dev(w(a)) for a in small]
    G = api.BuildGraph(N.KIND_TRI, off, *d, margin=0.0)
    B = api.build_tri_batch(*d, off)
    torch.cuda.synchronize()
    _same_batch(G, B, "small triangles")
    for t, a in zip(d, big):
        t.copy_(torch.from_numpy(w(a)).cuda())
    G.replay()
    flags = G.status()
    print(f"flags after the outgrowing replay: {flags:#x}", flush=True)
    assert flags != 0
    for t, a in zip(d, small):
        t.copy_(torch.from_numpy(w(a)).cuda())
    G.replay()
    assert G.status() == 0
    _same_batch(G, B, "small triangles again")
    G.close()
    B.release()


def test_graph_reports_deeper_tree(cuda):
    """Recorded on uniformly scattered points, with room to spare in every buffer, then
    replayed on Cantor dust (sums of cube corners shrinking 4x per step: each middle split
    halves the points, as for the uniform points, but leaves an empty slab that is cut off,
    so the tree is many levels deeper than the recorded layout covers): status reports the
    depth bit (512) instead of a partly laid out tree, and a later replay on the first
    input gives its trees again."""
    torch = cuda
    rng = np.random.default_rng(5)
    n, ns = 4000, 2
    uni = rng.uniform(-1, 1, (n * ns, 3))
    geo = np.zeros((n * ns, 3))
    for k in range(8):
        geo += rng.integers(0, 2, (n * ns, 3)) * 0.25 ** k
    geo = geo / geo.max() * 1.8 - 0.9 + rng.uniform(0, 1e-4, (n * ns, 3))
    w = lambda a: np.concatenate([a, np.zeros((len(a), 1))], 1).astype(np.float32)
    off = (np.arange(ns + 1) * n).astype(np.uint32)
    boxes = np.tile(np.array([-1, -1, -1, 1, 1, 1], np.float32), (ns, 1))
    qr = np.full(ns, 0.01, np.float32)
    d = api.dev(w(uni))
    G = api.BuildGraph(N.KIND_POINT, off, d, root_boxes=boxes, max_query_radius=qr, margin=16.0)
    B = api.build_points_batch(d, off, root_boxes=boxes, max_query_radius=qr)
    torch.cuda.synchronize()
    _same_batch(G, B, "uniform points")
    Bg = api.build_points_batch(api.dev(w(geo)), off, root_boxes=boxes, max_query_radius=qr)
    print(f"max level: uniform {max(B[s].max_level for s in range(ns))}, "
          f"Cantor dust {max(Bg[s].max_level for s in range(ns))}", flush=True)
    d.copy_(torch.from_numpy(w(geo)).cuda())
    G.replay()
    flags = G.status()
    print(f"flags after the deeper replay: {flags:#x}", flush=True)
    assert flags & 512
    d.copy_(torch.from_numpy(w(uni)).cuda())
    G.replay()
    assert G.status() == 0
    _same_batch(G, B, "uniform points again")
    G.close()
    B.release()
    Bg.release()
