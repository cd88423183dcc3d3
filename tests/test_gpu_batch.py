"""Batched builds: every tree of a batch equals the single-scene build of that scene
(node words, references, root box, leaf and level counts) and the oracle's; queries
through a batch view give the single-scene results bit for bit."""

import numpy as np
import pytest

from mnrt_amd import api, native as N, scene as S

pytestmark = pytest.mark.gpu


def _same_tree(a, b):
    assert a.num_nodes == b.num_nodes and a.num_refs == b.num_refs
    assert np.array_equal(a.nodes_host(), b.nodes_host())
    assert np.array_equal(a.refs_host(), b.refs_host())
    assert np.array_equal(a.root.view(np.uint32), b.root.view(np.uint32))
    assert a.max_level == b.max_level and a.t.num_leaves == b.t.num_leaves and a.t.num_elems == b.t.num_elems


def _tri_scenes():
    """Scenes of exactly these sizes (the first n triangles of a sized scene): empty,
    single-triangle, small-root-only, at and just above the small-node limit."""
    sizes = [3000, 0, 1, 40, 20000, 64, 65, 7000, 12]
    out = []
    for i, n in enumerate(sizes):
        if n == 0:
            out.append(None)
            continue
        sc = S.scene_sized(max(n, 100), seed=20 + i)
        sc.v0, sc.v1, sc.v2 = sc.v0[:n], sc.v1[:n], sc.v2[:n]
        out.append(sc)
    return out


def _concat(parts, width=4):
    parts = [p if p is not None else np.zeros((0, width), np.float32) for p in parts]
    off = np.zeros(len(parts) + 1, np.uint32)
    off[1:] = np.cumsum([len(p) for p in parts])
    return np.concatenate(parts).astype(np.float32), off


def test_batch_tri_equals_single(orc, cuda):
    scs = _tri_scenes()
    v0, off = _concat([s.v0 if s else None for s in scs])
    v1, _ = _concat([s.v1 if s else None for s in scs])
    v2, _ = _concat([s.v2 if s else None for s in scs])
    print(f"batch tri build: {len(scs)} scenes, {off[-1]} triangles", flush=True)
    B = api.build_tri_batch(v0, v1, v2, off)
    cuda.cuda.synchronize()
    for s, sc in enumerate(scs):
        a, b = off[s], off[s + 1]
        print(f"scene {s}: {b - a} triangles", flush=True)
        single = api.build_tri(v0[a:b], v1[a:b], v2[a:b])
        cuda.cuda.synchronize()
        _same_tree(B[s], single)
        if b - a > 0:
            ot = orc.build(0, v0[a:b], v1[a:b], v2[a:b])
            assert np.array_equal(B[s].nodes_host(), ot.nodes)
            assert np.array_equal(B[s].refs_host(), ot.refs)
            o, d, _ = S.camera_rays(32, 32)
            tri_b, t_b, _, _ = api.trace_closest(B[s], o, d, sc.ray_epsilon())
            tri_s, t_s, _, _ = api.trace_closest(single, o, d, sc.ray_epsilon())
            assert np.array_equal(tri_b.cpu().numpy(), tri_s.cpu().numpy())
            assert np.array_equal(t_b.cpu().numpy().view(np.uint32), t_s.cpu().numpy().view(np.uint32))
    B.release()


def test_batch_tri_root_boxes(orc, cuda):
    scs = [S.scene_sized(n, seed=40 + n) for n in (500, 9000, 30)]
    v0, off = _concat([s.v0 for s in scs])
    v1, _ = _concat([s.v1 for s in scs])
    v2, _ = _concat([s.v2 for s in scs])
    boxes = np.array([[-1.5, -1.5, -1.5, 1.5, 1.5, 1.5], [-1, -1, -1, 1, 1, 1], [-2, -1, -1, 2, 1, 1]], np.float32)
    B = api.build_tri_batch(v0, v1, v2, off, root_boxes=boxes)
    cuda.cuda.synchronize()
    for s in range(len(scs)):
        a, b = off[s], off[s + 1]
        single = api.build_tri(v0[a:b], v1[a:b], v2[a:b], api.default_params(0, boxes[s]))
        cuda.cuda.synchronize()
        _same_tree(B[s], single)


def test_batch_points_equals_single(orc, cuda):
    rng = np.random.default_rng(5)
    sizes = [20000, 3, 0, 32, 33, 4000, 100000]
    pos, off = _concat([np.concatenate([rng.uniform(-1, 1, (n, 3)), np.zeros((n, 1))], 1) if n else None
                        for n in sizes])
    boxes = np.tile(np.array([-1, -1, -1, 1, 1, 1], np.float32), (len(sizes), 1))
    qr = np.array([0.1, 0.2, 0.05, 0.1, 0.3, 0.02, 0.1], np.float32)
    print(f"batch point build: {len(sizes)} maps, {off[-1]} photons", flush=True)
    B = api.build_points_batch(pos, off, root_boxes=boxes, max_query_radius=qr)
    pw = rng.uniform(0, 1, (len(pos), 4)).astype(np.float32)
    dr = np.concatenate([rng.normal(size=(len(pos), 3)), np.zeros((len(pos), 1))], 1).astype(np.float32)
    dr[:, :3] /= np.linalg.norm(dr[:, :3], axis=1, keepdims=True)
    api.attach_photon_records_batch(B, pos, pw, dr)
    cuda.cuda.synchronize()
    for s, n in enumerate(sizes):
        a, b = off[s], off[s + 1]
        single = api.build_points(pos[a:b], api.default_params(1, boxes[s], float(qr[s])))
        cuda.cuda.synchronize()
        _same_tree(B[s], single)
        if n == 0:
            continue
        ot = orc.build(1, pos[a:b], max_qr=float(qr[s]), root=boxes[s])
        assert np.array_equal(B[s].nodes_host(), ot.nodes)
        api.attach_photon_records(single, pos[a:b], pw[a:b], dr[a:b])
        q = np.concatenate([rng.uniform(-1, 1, (500, 3)), np.zeros((500, 1))], 1).astype(np.float32)
        nrm = np.tile(np.array([0, 1, 0, 0], np.float32), (500, 1))
        dirs = np.tile(np.array([0, -1, 0, 0], np.float32), (500, 1))
        infl = np.ones((500, 4), np.float32)
        rb = api.knn_gather(B[s], 16, 0.3, q, nrm, dirs, infl, want_nbr=True, want_r2=True)
        rs = api.knn_gather(single, 16, 0.3, q, nrm, dirs, infl, want_nbr=True, want_r2=True)
        assert np.array_equal(rb["nbr"].cpu().numpy(), rs["nbr"].cpu().numpy())
        assert np.array_equal(rb["rad"].cpu().numpy().view(np.uint32), rs["rad"].cpu().numpy().view(np.uint32))
    B.release()


def _query_setup(cuda, sizes, photon_counts, res, seed=60):
    """Scenes with photon maps from the GPU photon tracer, built as batches, and camera
    rays per scene (res[s] = (W, H), (0, 0) for none)."""
    torch = cuda
    scs = [S.scene_sized(n, seed=seed + i) for i, n in enumerate(sizes)]
    cat = lambda a: np.concatenate([getattr(s, a) for s in scs]).astype(np.float32)
    v0, v1, v2, alb = cat("v0"), cat("v1"), cat("v2"), cat("albedo")
    toff = np.zeros(len(scs) + 1, np.uint32)
    toff[1:] = np.cumsum([s.num_tris for s in scs])
    TB = api.build_tri_batch(v0, v1, v2, toff)
    ph = []
    for s, sc in enumerate(scs):
        g = api.photons_generate(TB[s], sc.v0, sc.v1, sc.v2, sc.albedo, sc.lights, photon_counts[s], 4, s,
                                 sc.ray_epsilon())
        ph.append(g)
    poff = np.zeros(len(scs) + 1, np.uint32)
    poff[1:] = np.cumsum([g["count"] for g in ph])
    ppos = torch.cat([g["pos"] for g in ph]).contiguous()
    ppow = torch.cat([g["power"] for g in ph]).contiguous()
    pdir = torch.cat([g["dir"] for g in ph]).contiguous()
    boxes = np.stack([np.concatenate(sc.bounds()) for sc in scs]).astype(np.float32)
    rmax = np.array([sc.r_max() for sc in scs], np.float32)
    PB = api.build_points_batch(ppos, poff, root_boxes=boxes, max_query_radius=rmax)
    api.attach_photon_records_batch(PB, ppos, ppow, pdir)
    rays = [S.camera_rays(w, h) if w * h else (np.zeros((0, 4), np.float32),) * 2 for (w, h) in res]
    roff = np.zeros(len(scs) + 1, np.uint32)
    roff[1:] = np.cumsum([len(r[0]) for r in rays])
    o = api.dev(np.concatenate([r[0] for r in rays]))
    d = api.dev(np.concatenate([r[1] for r in rays]))
    eps = np.array([sc.ray_epsilon() for sc in scs], np.float32)
    return dict(scs=scs, v=[api.dev(a) for a in (v0, v1, v2, alb)], toff=toff, TB=TB, PB=PB, poff=poff, roff=roff,
                o=o, d=d, eps=eps, rmax=rmax, photons=(ppos, ppow, pdir))


def _batch_io(cuda, Q, X, k):
    torch = cuda
    n = int(X["roff"][-1])
    f4z = lambda: torch.zeros((max(n, 1), 4), dtype=torch.float32, device="cuda")
    bufs = dict(orig4=X["o"], dir4=X["d"], v0=X["v"][0], v1=X["v"][1], v2=X["v"][2], albedo4=X["v"][3],
                hit_tri=torch.empty(max(n, 1), dtype=torch.int32, device="cuda"),
                hit_t=torch.empty(max(n, 1), dtype=torch.float32, device="cuda"),
                hit_bary2=torch.empty((max(n, 1), 2), dtype=torch.float32, device="cuda"),
                sp_pos4=f4z(), sp_nrm4=f4z(), sp_dir4=f4z(), sp_clr4=f4z(), sp_infl4=f4z(),
                sp_src=torch.empty(max(n, 1), dtype=torch.int32, device="cuda"),
                sp_count=torch.zeros(Q.S, dtype=torch.int32, device="cuda"),
                io_rad4=f4z(), out_r2=torch.empty(max(n, 1), dtype=torch.float32, device="cuda"),
                out_nbr=torch.empty(max(n, 1) * k, dtype=torch.int32, device="cuda"))
    Q.make_io(**bufs)
    return bufs


def _check_against_plain(cuda, X, B, k):
    """Per scene: trace, compaction and kNN radiance equal the single-scene entry points."""
    torch = cuda
    cnt = B["sp_count"].cpu().numpy()
    for s, sc in enumerate(X["scs"]):
        a, b = int(X["roff"][s]), int(X["roff"][s + 1])
        if b == a:
            assert cnt[s] == 0
            continue
        print(f"scene {s}: {b - a} rays", flush=True)
        tri, t, bary, _ = api.trace_closest(X["TB"][s], X["o"][a:b].contiguous(), X["d"][a:b].contiguous(),
                                            float(X["eps"][s]))
        assert torch.equal(tri, B["hit_tri"][a:b])
        assert torch.equal(t.view(torch.int32), B["hit_t"][a:b].view(torch.int32))
        assert torch.equal(bary.view(torch.int32), B["hit_bary2"][a:b].view(torch.int32))
        ta, tb = int(X["toff"][s]), int(X["toff"][s + 1])
        q = api.shading_points(X["o"][a:b].contiguous(), X["d"][a:b].contiguous(), tri, t,
                               *[x[ta:tb].contiguous() for x in X["v"]], want_src=True)
        m = q["count"]
        assert cnt[s] == m
        for key, bk in (("pos", "sp_pos4"), ("nrm", "sp_nrm4"), ("dir", "sp_dir4"), ("clr", "sp_clr4"),
                        ("infl", "sp_infl4")):
            assert torch.equal(q[key].view(torch.int32), B[bk][a:a + m].view(torch.int32)), key
        assert torch.equal(q["src"], B["sp_src"][a:a + m])
        r = api.knn_gather(X["PB"][s], k, float(X["rmax"][s]), q["pos"], q["nrm"], q["dir"], q["infl"], q["clr"],
                           want_nbr=True, want_r2=True)
        assert torch.equal(r["rad"].view(torch.int32), B["io_rad4"][a:a + m].view(torch.int32))
        assert torch.equal(r["nbr"].reshape(-1), B["out_nbr"][a * k:(a + m) * k])
        assert torch.equal(r["r2"].view(torch.int32), B["out_r2"][a:a + m].view(torch.int32))


def test_query_batch_equals_plain(orc, cuda):
    torch = cuda
    sizes = [1200, 5000, 300, 20000, 2500, 800]
    photons = [4096, 20000, 3000, 60000, 8000, 5000]
    res = [(32, 32), (64, 48), (0, 0), (40, 40), (16, 16), (8, 8)]
    k = 50
    X = _query_setup(cuda, sizes, photons, res)
    Q = api.QueryBatch(X["TB"], X["toff"], X["roff"], X["eps"], X["PB"], X["rmax"], k)
    B = _batch_io(cuda, Q, X, k)
    print("batched trace + shade + kNN", flush=True)
    Q.run(7)
    cuda.cuda.synchronize()
    _check_against_plain(cuda, X, B, k)
    # graph replays equal the eager run (radiance accumulates: cleared between replays)
    # slots a phase writes: every ray slot for the hits; each scene's first count[s] slots
    # for shading points and kNN outputs
    cnt = B["sp_count"].cpu().numpy()
    valid = np.zeros(int(X["roff"][-1]), bool)
    for s in range(len(cnt)):
        valid[int(X["roff"][s]):int(X["roff"][s]) + int(cnt[s])] = True
    vt = torch.from_numpy(valid).cuda()
    vk = vt.repeat_interleave(k)
    ref = {n: B[n].clone() for n in ("hit_tri", "hit_t", "sp_pos4", "sp_count", "io_rad4", "out_nbr")}
    Q.capture(7)
    for rep in range(3):
        for n in ("hit_tri", "hit_t", "sp_pos4", "sp_count", "out_nbr"):
            B[n].fill_(-1)
        B["io_rad4"].zero_()
        cuda.cuda.synchronize()
        Q.replay(7)
        cuda.cuda.synchronize()
        for n, v in ref.items():
            got, want = B[n].view(torch.int32), v.view(torch.int32)
            if n in ("sp_pos4", "io_rad4"):
                got, want = got[:len(valid)][vt], want[:len(valid)][vt]
            elif n == "out_nbr":
                got, want = got[:len(vk)][vk], want[:len(vk)][vk]
            assert torch.equal(got, want), (rep, n)
    Q.close()


def test_query_batch_autotuned_shapes(orc, cuda):
    """Every launch shape the tuner may pick gives the same results."""
    torch = cuda
    k = 100
    X = _query_setup(cuda, [3000, 900, 12000], [10000, 4096, 30000], [(48, 48), (24, 24), (64, 64)], seed=80)
    Q = api.QueryBatch(X["TB"], X["toff"], X["roff"], X["eps"], X["PB"], X["rmax"], k)
    B = _batch_io(cuda, Q, X, k)
    Q.run(7)
    cuda.cuda.synchronize()
    ref = {n: B[n].clone() for n in ("hit_tri", "hit_t", "io_rad4", "out_nbr")}
    shapes = [dict(trace_block=64, trace_tile=64, trace_grid=256, knn_waves=1, knn_chunk=1, knn_grid=512, parts=1),
              dict(trace_block=128, trace_tile=2048, trace_grid=4096, trace_stack=50, knn_waves=2, knn_chunk=64,
                   knn_cap=512, knn_grid=2048, parts=8)]
    for sh in shapes:
        Q.tuning = sh
        B["io_rad4"].zero_()
        Q.run(7)
        cuda.cuda.synchronize()
        for n, v in ref.items():
            assert torch.equal(B[n].view(torch.int32), v.view(torch.int32)), (sh, n)
    tun = Q.autotune(max_slots=2000)
    print("tuned", tun, flush=True)
    B["io_rad4"].zero_()
    Q.run(7)
    cuda.cuda.synchronize()
    for n, v in ref.items():
        assert torch.equal(B[n].view(torch.int32), v.view(torch.int32)), n
    Q.close()


def test_query_batch_bwd_equals_plain(orc, cuda):
    """The batched kNN adjoint phase gives, per scene, the single-scene adjoint's photon and
    colour gradients bit for bit (each scene in its own fixed-point unit), an empty scene
    gets zero photon gradients, and graph replays equal the eager run."""
    torch = cuda
    sizes = [1500, 300, 8000, 2500]
    photons = [6000, 3000, 30000, 9000]
    res = [(40, 32), (0, 0), (48, 48), (24, 24)]
    k = 50
    X = _query_setup(cuda, sizes, photons, res, seed=90)
    Q = api.QueryBatch(X["TB"], X["toff"], X["roff"], X["eps"], X["PB"], X["rmax"], k)
    B = _batch_io(cuda, Q, X, k)
    n = max(int(X["roff"][-1]), 1)
    poff = Q.photon_offsets
    g = torch.Generator(device="cpu").manual_seed(5)
    # per-scene gradient magnitudes far apart: a shared fixed-point unit would round the small ones
    scale = torch.ones(n, 1)
    for s in range(len(sizes)):
        scale[int(X["roff"][s]):int(X["roff"][s + 1])] = 10.0 ** (-3 * s)
    B["grad_rad4"] = ((torch.rand((n, 4), generator=g) - 0.3) * scale).cuda()
    B["out_count"] = torch.empty(n, dtype=torch.int32, device="cuda")
    B["grad_pow4"] = torch.full((max(int(poff[-1]), 1), 4), 7.0, device="cuda")
    B["grad_clr4"] = torch.zeros((n, 4), device="cuda")
    Q.make_io(**B)
    print("batched forward + adjoint", flush=True)
    Q.run(7)
    Q.run(8)
    cuda.cuda.synchronize()
    cnt = B["sp_count"].cpu().numpy()
    for s in range(len(sizes)):
        a, m = int(X["roff"][s]), int(cnt[s])
        sl = slice(a, a + m)
        print(f"scene {s}: {m} queries, {int(poff[s + 1] - poff[s])} photons", flush=True)
        gp, gc = api.knn_gather_bwd(X["PB"][s], k, float(X["rmax"][s]), B["sp_pos4"][sl].contiguous(),
                                    B["sp_nrm4"][sl].contiguous(), B["sp_dir4"][sl].contiguous(),
                                    B["sp_infl4"][sl].contiguous(), B["sp_clr4"][sl].contiguous(),
                                    B["grad_rad4"][sl].contiguous(), B["out_r2"][sl].contiguous(),
                                    B["out_count"][sl].contiguous())
        got = B["grad_pow4"][int(poff[s]):int(poff[s + 1])]
        assert torch.equal(got.view(torch.int32), gp.view(torch.int32)), s
        if m:
            assert torch.equal(B["grad_clr4"][sl].view(torch.int32), gc.view(torch.int32)), s
            assert bool((gp[:, :3] != 0).any()), s
        else:
            assert not bool(got.any()), s
    ref = {nm: B[nm].clone() for nm in ("grad_pow4", "grad_clr4")}
    Q.capture(8)
    for rep in range(3):
        B["grad_pow4"].fill_(3.0)
        B["grad_clr4"].fill_(5.0)
        cuda.cuda.synchronize()
        Q.replay(8)
        cuda.cuda.synchronize()
        assert torch.equal(B["grad_pow4"].view(torch.int32), ref["grad_pow4"].view(torch.int32)), rep
        for s in range(len(sizes)):
            sl = slice(int(X["roff"][s]), int(X["roff"][s]) + int(cnt[s]))
            assert torch.equal(B["grad_clr4"][sl].view(torch.int32), ref["grad_clr4"][sl].view(torch.int32)), rep
    Q.close()


def test_bwd_graph_dropped_when_photon_total_changes(orc, cuda):
    """A captured adjoint phase clears and converts the photon total it was captured with:
    after update_trees to maps with fewer photons its replay is refused (MNRT_ERR_STATE)
    instead of writing past the new gradient buffer; eager runs and a new capture write
    exactly the new total (guard rows behind it stay untouched)."""
    torch = cuda
    sizes, photons, res, k = [1200, 2500, 900], [5000, 9000, 4000], [(24, 24), (32, 24), (16, 16)], 30
    X = _query_setup(cuda, sizes, photons, res, seed=120)
    Q = api.QueryBatch(X["TB"], X["toff"], X["roff"], X["eps"], X["PB"], X["rmax"], k)
    B = _batch_io(cuda, Q, X, k)
    n = max(int(X["roff"][-1]), 1)
    B["grad_rad4"] = torch.rand((n, 4), device="cuda")
    B["out_count"] = torch.empty(n, dtype=torch.int32, device="cuda")
    B["grad_pow4"] = torch.zeros((int(Q.photon_offsets[-1]), 4), device="cuda")
    B["grad_clr4"] = torch.zeros((n, 4), device="cuda")
    Q.make_io(**B)
    Q.run(7)
    Q.capture(8)
    Q.replay(8)
    cuda.cuda.synchronize()
    # the same scenes with the first half of each photon map
    ppos, ppow, pdir = X["photons"]
    keep = [np.arange(int(X["poff"][s]), int(X["poff"][s]) + (int(X["poff"][s + 1] - X["poff"][s]) // 2))
            for s in range(len(sizes))]
    idx = torch.from_numpy(np.concatenate(keep)).cuda()
    p2 = [t[idx].contiguous() for t in (ppos, ppow, pdir)]
    poff2 = np.zeros(len(sizes) + 1, np.uint32)
    poff2[1:] = np.cumsum([len(kk) for kk in keep])
    boxes = np.stack([np.concatenate(sc.bounds()) for sc in X["scs"]]).astype(np.float32)
    PB2 = api.build_points_batch(p2[0], poff2, root_boxes=boxes, max_query_radius=X["rmax"])
    api.attach_photon_records_batch(PB2, *p2)
    Q.update_trees(X["TB"], PB2)
    with pytest.raises(N.MnrtError) as ex:
        Q.replay(8)
    assert "(-6)" in str(ex.value)
    total = int(poff2[-1])
    guard = 16
    gp = torch.full((total + guard, 4), 7.0, device="cuda")
    B["grad_pow4"] = gp
    Q.make_io(**B)
    Q.run(4)
    Q.run(8)
    cuda.cuda.synchronize()
    assert bool((gp[total:] == 7.0).all())
    eager = gp[:total].clone()
    Q.capture(8)
    gp.fill_(7.0)
    Q.replay(8)
    cuda.cuda.synchronize()
    assert bool((gp[total:] == 7.0).all())
    assert torch.equal(gp[:total].view(torch.int32), eager.view(torch.int32))
    Q.close()
    PB2.release()
