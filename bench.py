#!/usr/bin/env python3
"""Benchmark of the MNRT kd-tree hot path on one MI355X (gfx950).

One step = one pass of the hot path over the plain large workload (about 1.03M
triangles, 1M traced photons, 1920x1080 primary rays plus one cosine bounce ray per hit,
k = 100 kNN radiance estimate at every hit):

  build the triangle kd-tree            mnrt_build_tri        (KDTreeTri + BuildTree)
  trace primary rays                    mnrt_trace_closest    (KernelRTTraceRays)
  shading points + bounce rays + trace  mnrt_shading_points, mnrt_bounce_rays, mnrt_trace_closest
  build the photon kd-tree              mnrt_build_points     (KDTreePoint + BuildTree)
  kNN radiance at all hits              mnrt_knn_gather       (KernelPMPhotonGather)

The photon tree does not depend on the first three stages: by default a second host thread
builds it on a second stream while they run (--serial: after them, on one stream); the
kNN stage waits for both.

Inputs are resident in HBM before the timed region (the photon map is traced once during
set-up with mnrt_photons_generate).  `value` counts rays traced plus kNN queries answered
per second of step time.  The roofline entry is for the kNN gather kernel; its
algorithmic bytes come from the oracle's instrumented replay of a deterministic sample
of the same queries on the same tree.  The CPU baseline is the oracle (C++ restatement of
the reference algorithms) on the host cores, on a bounded sample.
"""

from __future__ import annotations

import argparse
import concurrent.futures
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mrays/s + M kNN queries/s (k=50) at %HBM peak; batched vs plain; bwd/fwd ratio"


def trace_bytes(counts, n_rays):
    """Traversal algorithmic bytes: 24 B ray in + 16 B hit out per ray, 8 B per inner and
    per leaf visit, 4 B per leaf reference, 36 B per triangle test."""
    inner, leaves, refs, tests = [float(c) for c in counts]
    return 40.0 * n_rays + 8.0 * inner + 8.0 * leaves + 4.0 * refs + 36.0 * tests


def knn_bytes(counts, n_q, n_set):
    """kNN algorithmic bytes: 24 B query in + 12 B radiance out, 8 B per node visit,
    4 B per leaf reference, 12 B per photon distance test, 16 B per neighbour kept."""
    inner, leaves, refs, tests, _ = [float(c) for c in counts]
    return 36.0 * n_q + 8.0 * (inner + leaves) + 4.0 * refs + 12.0 * tests + 16.0 * float(n_set)


def shadow_bytes(counts, n_rays):
    """Any-hit algorithmic bytes: 16 B point in + 4 B flag out per ray, then as traversal."""
    inner, leaves, refs, tests = [float(c) for c in counts]
    return 20.0 * n_rays + 8.0 * inner + 8.0 * leaves + 4.0 * refs + 36.0 * tests


def _progress(msg):
    """One progress line on stderr (the JSON result line stays alone on stdout)."""
    print(f"# {msg}", file=sys.stderr, flush=True)


def shadow_leg(torch, api, N, O, T, ttree, sc, pts, eps, sample, ncpu, reps=5):
    """Shadow rays from the primary hits of the large workload (mnrt_trace_shadow): one
    to a point light just below the ceiling light, one to a stratified sample of the
    ceiling area light (mnrt_sample_area_light); median of `reps` launches each.  Bytes
    per ray from the oracle's any-hit replay of a 1-in-`sample` subset of the point-light
    segments on the same tree."""
    lq = sc.lights[0]
    centre = (lq[0] + 0.5 * lq[1] + 0.5 * lq[2] - np.array([0.0, 0.01, 0.0], np.float32)).astype(np.float32)
    point = api.make_light(N.LIGHT_POINT, centre)
    rect = api.make_light(N.LIGHT_AREA_RECT, lq[0] - np.array([0.0, 0.001, 0.0], np.float32), (0.0, -1.0, 0.0),
                          v1=lq[1], v2=lq[2])
    n = pts.shape[0]
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    r1 = torch.rand(n, device="cuda", generator=g)
    r2 = torch.rand(n, device="cuda", generator=g)
    stream = torch.cuda.current_stream()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ms = []
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            r = fn()
            b.record(stream)
            torch.cuda.synchronize()
            ms.append(a.elapsed_time(b))
        return r, float(np.median(ms))

    vis_p, ms_p = timed(lambda: api.trace_shadow(T, point, pts, eps))
    smp, ms_s = timed(lambda: api.sample_area_light(rect, r1, r2, 0.0, 1.0, 0.0, 1.0))
    vis_a, ms_a = timed(lambda: api.trace_shadow(T, rect, pts, eps, smp))
    # algorithmic bytes: oracle any-hit walk of the point-light segments on a sample
    p = pts[::sample].cpu().numpy()
    v = centre[None, :] - p[:, :3]
    dist = np.sqrt(np.sum(v * v, axis=1)).astype(np.float32)
    d = np.zeros_like(p)
    d[:, :3] = v / dist[:, None]
    _, ctr = O.trace_any(ttree, ttree.root, sc.v0, sc.v1, sc.v2, p, d, eps, dist, nthreads=ncpu)
    per_ray = shadow_bytes(ctr, len(p)) / len(p)
    gbs = per_ray * n / (ms_p * 1e-3) / 1e9
    return {"workload": f"shadow rays from {n} primary hits of the large workload to a point light and to "
                        "ceiling area-light samples",
            "rays": int(n), "point_ms": round(ms_p, 4), "area_ms": round(ms_a, 4), "sample_ms": round(ms_s, 4),
            "point_mrays_per_s": round(n / (ms_p * 1e-3) / 1e6, 1),
            "area_mrays_per_s": round(n / (ms_a * 1e-3) / 1e6, 1),
            "unoccluded_point": round(float(vis_p.float().mean()), 4),
            "unoccluded_area": round(float(vis_a.float().mean()), 4),
            "roofline": {"kernel": "k_any_hit<kSegPoint>", "alg_bytes_per_ray": round(per_ray, 1),
                         "achieved": round(gbs, 2), "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 5)}}


def _median_time(fn, reps):
    """(result of a first, untimed call, median seconds of `reps` more calls); reps = 0 times
    the single call."""
    t0 = time.perf_counter()
    r = fn()
    first = time.perf_counter() - t0
    if reps <= 0:
        return r, first
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return r, float(np.median(ts))


def build_leg(args, torch, api, O, v0, v1, v2, ppos, pparams, otri, opts):
    """The reference's kd-tree construction benchmark (RTCore::BenchmarkKDTreeCon,
    RTCore.cpp:287-327, as the UI runs it: 10 warm-up and 200 timed KDTreeTri builds,
    MainFrame.cpp:755-792) on the large workload's triangles, and the same for its photon
    tree.  Each build is timed by the host around the call and a device synchronisation
    (as the reference's host timer) and by HIP events on the build stream.  Algorithmic
    bytes: the oracle build's element streams (oracle.build_bytes) plus, for triangles, the
    leaf-ordered records the build attaches (per reference: 4 B read, 48 B of vertices read,
    48 B written)."""
    stream = torch.cuda.current_stream()
    res = {}
    for name, fn, otree, extra in (
            ("tri", lambda: api.build_tri(v0, v1, v2), otri, 100.0 * otri.refs.size),
            ("points", lambda: api.build_points(ppos, pparams), opts, 0.0)):
        host_ms, ev_ms = [], []
        for i in range(args.build_warmup + args.build_runs):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            a.record(stream)
            Tb = fn()
            b.record(stream)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            if i >= args.build_warmup:
                host_ms.append((t1 - t0) * 1e3)
                ev_ms.append(a.elapsed_time(b))
            Tb.release()
        alg = O.build_bytes(otree) + extra
        ms = float(np.mean(ev_ms))
        gbs = alg / (ms * 1e-3) / 1e9
        res[name] = {"elements": int(otree.num_elems), "nodes": int(otree.num_nodes), "runs": args.build_runs,
                     "warmup": args.build_warmup, "avg_ms_host": round(float(np.mean(host_ms)), 3),
                     "avg_ms_events": round(ms, 3), "min_ms_events": round(float(np.min(ev_ms)), 3),
                     "alg_bytes": int(alg),
                     "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": round(gbs / HBM_PEAK_GBS, 5)}}
    return res


def _offsets(counts):
    off = np.zeros(len(counts) + 1, np.uint32)
    off[1:] = np.cumsum(np.asarray(counts, np.int64))
    return off


def batched_leg(args, torch, api, S, O, ncpu):
    """Batched workload: 256 independent scenes like the small one (box, icosphere,
    random triangles, one light), triangle / photon / query counts log-uniform in
    [1k, 50k] / [4k, 256k] / [16k, 1M] (seed 4).  One step builds every triangle tree and
    every photon tree (one batched build each, device-driven and replayed from a hipGraph:
    mnrt_build_graph, records and query-time nodes included), then traces all primary
    rays, compacts the hits and answers kNN (k = 50) at all hits, each phase one
    persistent launch replayed from a hipGraph.  The same work through the single-scene
    entry points, scene after scene, is timed beside it."""
    from mnrt_amd import native as N

    t0 = time.time()
    NS, K = args.batched_scenes, 50
    scs, nph, dims = S.batched_workload(NS, seed=4)
    v0, v1, v2, alb = (api.dev(np.concatenate([getattr(sc, a) for sc in scs]).astype(np.float32))
                       for a in ("v0", "v1", "v2", "albedo"))
    toff = _offsets([sc.num_tris for sc in scs])
    roff = _offsets([w * h for w, h in dims])
    R = int(roff[-1])
    go = torch.empty((R, 4), dtype=torch.float32, device="cuda")
    gd = torch.empty((R, 4), dtype=torch.float32, device="cuda")
    for s, (w, h) in enumerate(dims):
        api.primary_rays(w, h, out_orig=go[roff[s]:roff[s + 1]], out_dir=gd[roff[s]:roff[s + 1]])
    eps = np.array([sc.ray_epsilon() for sc in scs], np.float32)
    rmax = np.array([sc.r_max() for sc in scs], np.float32)
    boxes = np.stack([np.concatenate(sc.bounds()) for sc in scs]).astype(np.float32)
    tsl = [slice(int(toff[s]), int(toff[s + 1])) for s in range(NS)]
    rsl = [slice(int(roff[s]), int(roff[s + 1])) for s in range(NS)]
    TB = api.build_tri_batch(v0, v1, v2, toff)
    ph = [api.photons_generate(TB[s], v0[tsl[s]], v1[tsl[s]], v2[tsl[s]], alb[tsl[s]], scs[s].lights, int(nph[s]),
                               4, s, float(eps[s])) for s in range(NS)]
    poff = _offsets([g["count"] for g in ph])
    ppos, ppow, pdir = (torch.cat([g[k] for g in ph]).contiguous() for k in ("pos", "power", "dir"))
    psl = [slice(int(poff[s]), int(poff[s + 1])) for s in range(NS)]
    del ph

    TB.release()
    GT = api.BuildGraph(N.KIND_TRI, toff, v0, v1, v2)
    GP = api.BuildGraph(N.KIND_POINT, poff, ppos, power=ppow, dirs=pdir, root_boxes=boxes, max_query_radius=rmax)
    TB, PB = GT.trees, GP.trees
    Q = api.QueryBatch(TB, toff, roff, eps, PB, rmax, K)
    f4 = lambda: torch.empty((R, 4), dtype=torch.float32, device="cuda")
    bufs = dict(orig4=go, dir4=gd, v0=v0, v1=v1, v2=v2, albedo4=alb,
                hit_tri=torch.empty(R, dtype=torch.int32, device="cuda"),
                hit_t=torch.empty(R, dtype=torch.float32, device="cuda"),
                sp_pos4=f4(), sp_nrm4=f4(), sp_dir4=f4(), sp_clr4=f4(), sp_infl4=f4(),
                sp_count=torch.zeros(NS, dtype=torch.int32, device="cuda"), io_rad4=f4())
    Q.make_io(**bufs)
    tt = time.time()
    tuning = Q.autotune(max_slots=min(R, 8 << 20))
    tune_s = time.time() - tt
    for ph_ in (1, 2, 4):
        Q.capture(ph_)
    torch.cuda.synchronize()
    setup_s = time.time() - t0
    stream = torch.cuda.current_stream()
    names = ["build_tri", "build_points", "trace", "shade", "knn"]

    def batched_step():
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
        torch.cuda.synchronize()
        w0 = time.time()
        ev[0].record(stream)
        GT.replay()
        ev[1].record(stream)
        GP.replay()
        bufs["io_rad4"].zero_()
        ev[2].record(stream)
        for i, ph_ in enumerate((1, 2, 4)):
            Q.replay(ph_)
            ev[3 + i].record(stream)
        torch.cuda.synchronize()
        wall = time.time() - w0
        # the rebuilt trees fit the graphs' buffers (same input: same trees, same addresses)
        if GT.status() != 0 or GP.status() != 0:
            raise RuntimeError("a recorded build outgrew its buffers")
        return {n: ev[i].elapsed_time(ev[i + 1]) for i, n in enumerate(names)}, wall

    def plain_step():
        keep = []
        torch.cuda.synchronize()
        w0 = time.time()
        for s in range(NS):
            T = api.build_tri(v0[tsl[s]], v1[tsl[s]], v2[tsl[s]])
            P = api.build_points(ppos[psl[s]], api.default_params(1, boxes[s], float(rmax[s])))
            api.attach_photon_records(P, ppos[psl[s]], ppow[psl[s]], pdir[psl[s]])
            tri, t, _, _ = api.trace_closest(T, go[rsl[s]], gd[rsl[s]], float(eps[s]))
            q = api.shading_points(go[rsl[s]], gd[rsl[s]], tri, t, v0[tsl[s]], v1[tsl[s]], v2[tsl[s]], alb[tsl[s]])
            api.knn_gather(P, K, float(rmax[s]), q["pos"], q["nrm"], q["dir"], q["infl"], q["clr"])
            keep.append((T, P))
        torch.cuda.synchronize()
        return keep, time.time() - w0

    steps = max(1, args.batched_steps)
    phase, walls = [], []
    for it in range(steps + 1):  # the first is a warm-up
        ms, wall = batched_step()
        if it > 0:
            phase.append(ms)
            walls.append(wall)
    batched_s = float(np.mean(walls))
    hits = int(bufs["sp_count"].sum().item())
    pwalls = []
    for it in range(steps + 1):
        keep, wall = plain_step()
        if it > 0:
            pwalls.append(wall)
        for T, P in keep:
            T.release()
            P.release()
    plain_s = float(np.mean(pwalls))
    # algorithmic bytes of k_knn_batch: the oracle's replay of 1 in `sample` queries of
    # every 8th scene (its own photon tree of the scene), weighted by query count
    cnt = bufs["sp_count"].cpu().numpy().astype(np.int64)
    tot_b, tot_q = 0.0, 0
    # the CPU baseline times the oracle on the same sample: both tree builds of every 8th
    # scene, traversal of 1 in `sample` of its rays, kNN sets + estimate at 1 in `sample` of
    # its hits (all host threads, 1 warm-up + median of 5), each scaled to the whole batch
    reps = 0 if args.no_cpu else 5
    ct = dict(build=0.0, elems=0, trace=0.0, rays=0, knn=0.0, queries=0)
    for s in range(0, NS, 8):
        if cnt[s] == 0:
            continue
        pp = ppos[psl[s]].cpu().numpy()
        op, pb_s = _median_time(lambda: O.build(1, pp, max_qr=float(rmax[s]), root=boxes[s], nthreads=ncpu), reps)
        idx = torch.arange(int(roff[s]), int(roff[s]) + int(cnt[s]), args.sample, device="cuda")
        qp = bufs["sp_pos4"][idx].cpu().numpy()
        (nbr, c, ctr), knn_s = _median_time(lambda: O.knn(op, pp, qp, K, float(rmax[s]), nthreads=ncpu), reps)
        tot_b += knn_bytes(ctr, len(qp), int(np.minimum(c, K).sum())) / len(qp) * float(cnt[s])
        tot_q += int(cnt[s])
        if args.no_cpu:
            continue
        sc = scs[s]
        ot, tb_s = _median_time(lambda: O.build(0, sc.v0, sc.v1, sc.v2, nthreads=ncpu), reps)
        ridx = torch.arange(int(roff[s]), int(roff[s + 1]), args.sample, device="cuda")
        ro, rd = go[ridx].cpu().numpy(), gd[ridx].cpu().numpy()
        _, tr_s = _median_time(lambda: O.trace(ot, sc.v0, sc.v1, sc.v2, ro, rd, float(eps[s]), nthreads=ncpu), reps)
        qh = {k_: bufs["sp_" + k_ + "4"][idx].cpu().numpy() for k_ in ("nrm", "dir", "infl", "clr")}
        pw, pdr = ppow[psl[s]].cpu().numpy(), pdir[psl[s]].cpu().numpy()
        _, rad_s = _median_time(lambda: O.radiance(nbr, c, K, float(rmax[s]), pp, pw, pdr, qp, qh["nrm"], qh["dir"],
                                                   qh["infl"], qh["clr"], nthreads=ncpu), reps)
        ct["build"] += tb_s + pb_s
        ct["elems"] += sc.num_tris + len(pp)
        ct["trace"] += tr_s
        ct["rays"] += len(ro)
        ct["knn"] += knn_s + rad_s
        ct["queries"] += len(qp)
    per_q = tot_b / max(tot_q, 1)
    cpu = None
    if not args.no_cpu and ct["queries"] > 0:
        cpu_s = (ct["build"] * (int(toff[-1]) + int(poff[-1])) / ct["elems"] + ct["trace"] * R / ct["rays"]
                 + ct["knn"] * hits / ct["queries"])
        cpu = {"value": round(hits / cpu_s / 1e6, 4), "unit": "M kNN queries/s (step incl. builds and traversal)",
               "ms_per_step": round(cpu_s * 1e3, 1), "cores": ncpu, "kind": "port",
               "gpu_speedup": round(cpu_s / batched_s, 1),
               "sample": (f"oracle on every 8th scene: both tree builds ({ct['elems']} elements, {ct['build']:.2f}s), "
                          f"traversal of 1 in {args.sample} rays ({ct['rays']}, {ct['trace']:.3f}s), kNN sets + "
                          f"estimate at 1 in {args.sample} hits ({ct['queries']}, {ct['knn']:.3f}s), all on {ncpu} "
                          "threads, each part 1 warm-up + median of 5, scaled by elements / rays / hits")}
    knn_ms = float(np.mean([p["knn"] for p in phase]))
    gbs = per_q * hits / (knn_ms * 1e-3) / 1e9
    Q.close()
    GT.close()
    GP.close()
    return {
        "knn_roofline": {"kernel": "k_knn_batch (MNRT_PHASE_KNN)", "bound": "hbm", "achieved": round(gbs, 2),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 5),
                         "alg_bytes_per_query": round(per_q, 1), "phase_ms": round(knn_ms, 3),
                         "sample": f"1 in {args.sample} queries of every 8th scene"},
        "workload": (f"{NS} scenes: triangles log-U[1k,50k], photons log-U[4k,256k], queries log-U[16k,1M] "
                     "(seed 4), k=50; one step = all tree builds + all traversals + compaction + all kNN"),
        "builds": "device-driven, one hipGraph replay per build (mnrt_build_graph: levels, small stage, layout, "
                  "emission, records, query-time nodes; no host synchronisation)",
        "scenes": NS, "triangles": int(toff[-1]), "photons": int(poff[-1]), "rays": R, "knn_queries": hits,
        "value": round(hits / batched_s / 1e6, 3), "unit": "M kNN queries/s (step incl. builds and traversal)",
        "ms_per_step": round(batched_s * 1e3, 3), "plain_ms_per_step": round(plain_s * 1e3, 3),
        "plain_value": round(hits / plain_s / 1e6, 3), "speedup_vs_plain": round(plain_s / batched_s, 2),
        "phase_ms": {n: round(float(np.mean([p[n] for p in phase])), 3) for n in names},
        "tuning": tuning, "autotune_s": round(tune_s, 2), "setup_s": round(setup_s, 1), "steps": steps,
        "cpu_baseline": cpu}


def adjoint_leg(args, torch, api, S, O, ncpu):
    """Differentiable workload: the small scene with 3 lights, 1M traced photons,
    1920x1080 primary rays, k = 50 at all hits; forward = radiance with r2 and counts
    kept (mnrt_knn_gather), backward = the adjoint (mnrt_knn_gather_bwd: neighbour sets
    recomputed, dLoss/dpower per photon and dLoss/dcolour per query)."""
    K = 50
    sc = S.scene_small(subdiv=5, num_lights=3)
    eps, rmax = sc.ray_epsilon(), sc.r_max()
    v = [api.dev(a) for a in (sc.v0, sc.v1, sc.v2, sc.albedo)]
    T = api.build_tri(*v[:3])
    ph = api.photons_generate(T, *v, sc.lights, 1_000_000, 4, 0, eps)
    o, d, _ = api.primary_rays(1920, 1080)
    tri, t, _, _ = api.trace_closest(T, o, d, eps)
    q = api.shading_points(o, d, tri, t, *v)
    P = api.build_points(ph["pos"], api.default_params(1, np.concatenate(sc.bounds()), rmax))
    api.attach_photon_records(P, ph["pos"], ph["power"], ph["dir"])
    grad = torch.rand((q["count"], 4), dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    fwd, bwd = [], []
    for it in range(4):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e[0].record(stream)
        out = api.knn_gather(P, K, rmax, q["pos"], q["nrm"], q["dir"], q["infl"], q["clr"], want_r2=True,
                             want_count=True)
        e[1].record(stream)
        api.knn_gather_bwd(P, K, rmax, q["pos"], q["nrm"], q["dir"], q["infl"], q["clr"], grad, out["r2"],
                           out["count"])
        e[2].record(stream)
        torch.cuda.synchronize()
        if it > 0:
            fwd.append(e[0].elapsed_time(e[1]))
            bwd.append(e[1].elapsed_time(e[2]))
    f, b = float(np.median(fwd)), float(np.median(bwd))
    # algorithmic bytes: the forward's from the oracle's replay of 1 in `sample` queries;
    # the adjoint's = the forward's + 12 B of gradient writes per neighbour
    root = np.concatenate(sc.bounds()).astype(np.float32)
    pp = ph["pos"].cpu().numpy()
    op = O.build(1, pp, max_qr=rmax, root=root)
    qp = q["pos"][:: args.sample].cpu().numpy()
    reps = 0 if args.no_cpu else 5
    (nbr, c, ctr), knn_s = _median_time(lambda: O.knn(op, pp, qp, K, rmax, nthreads=ncpu), reps)
    found = int(np.minimum(c, K).sum())
    fwd_q = knn_bytes(ctr, len(qp), found) / len(qp)
    bwd_q = fwd_q + 12.0 * found / len(qp)
    nq = q["count"]
    fg, bg = fwd_q * nq / (f * 1e-3) / 1e9, bwd_q * nq / (b * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        # the oracle on the same sample: forward = sets + fp32 estimate (all host threads),
        # backward = sets + fp64 adjoint (one thread), scaled to every query
        pw, pdr = ph["power"].cpu().numpy(), ph["dir"].cpu().numpy()
        qh = {k_: q[k_][:: args.sample].cpu().numpy() for k_ in ("nrm", "dir", "infl", "clr")}
        g = grad[:: args.sample].cpu().numpy()
        _, rad_s = _median_time(lambda: O.radiance(nbr, c, K, rmax, pp, pw, pdr, qp, qh["nrm"], qh["dir"],
                                                   qh["infl"], qh["clr"], nthreads=ncpu), reps)
        _, grad_s = _median_time(lambda: O.radiance_grad(nbr, c, K, rmax, pp, pw, pdr, qp, qh["nrm"], qh["dir"],
                                                         qh["infl"], qh["clr"], g), reps)
        scale = nq / len(qp)
        cf, cb = (knn_s + rad_s) * scale * 1e3, (knn_s + grad_s) * scale * 1e3
        cpu = {"fwd_ms": round(cf, 1), "bwd_ms": round(cb, 1), "cores": ncpu, "kind": "port",
               "gpu_speedup_fwd": round(cf / f, 1), "gpu_speedup_bwd": round(cb / b, 1),
               "sample": (f"oracle on 1 in {args.sample} queries ({len(qp)}), each part 1 warm-up + median of 5: "
                          f"kNN sets {knn_s:.3f}s and fp32 estimate {rad_s:.3f}s on {ncpu} threads, fp64 adjoint "
                          f"{grad_s:.3f}s on one thread, scaled to {nq} queries")}
    T.release()
    P.release()
    return {"workload": "small scene + 3 lights, 1M photons, 1920x1080 primary hits, k=50",
            "photons": ph["count"], "queries": nq, "fwd_ms": round(f, 3), "bwd_ms": round(b, 3),
            "bwd_fwd_ratio": round(b / f, 3),
            "fwd_roofline": {"kernel": "k_knn (mnrt_knn_gather, r2 and counts kept)", "achieved": round(fg, 2),
                             "unit": "GB/s", "frac": round(fg / HBM_PEAK_GBS, 5), "alg_bytes_per_query": round(fwd_q, 1)},
            "bwd_roofline": {"kernel": "k_knn_bwd_chunks (mnrt_knn_gather_bwd)", "bound": "hbm",
                             "achieved": round(bg, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": round(bg / HBM_PEAK_GBS, 5), "alg_bytes_per_query": round(bwd_q, 1),
                             "sample": f"1 in {args.sample} queries"},
            "cpu_baseline": cpu}


def inverse_leg(args, torch, api, S, O, ncpu):
    """Batched inverse lighting: 64 scenes (4 lights each; triangles log-U[1k,50k], photons
    log-U[4k,256k], primary-hit queries log-U[16k,256k], seed 5), target radiance from random
    per-light RGB emission in [0.5,2]; Adam (lr 1e-2, 200 steps) on per-light emission and
    per-triangle albedo from emission 1 and the traced albedo.  A step = photon and colour
    reweighting (mnrt_photons_reweight), record refresh, one batched kNN launch for every
    scene, the loss, one batched adjoint launch, the parameter adjoint
    (mnrt_photons_reweight_bwd) and the Adam update, with no host synchronisation.  After the
    fit every scene is checked against the oracle on a 1-in-`sample` subset of its hits: kNN
    sets by the oracle's own tree, radiance against its fp32 path (bits) and fp64 path
    (relative error), photon powers against the fp64 reweighting."""
    from mnrt_amd.inverse import BatchedInverseLighting

    t0 = time.time()
    NS, K = args.inverse_scenes, 50
    scs, nph, dims = S.batched_workload(NS, seed=5, query_hi=262144, num_lights=4, scene_seed=1000)
    P = BatchedInverseLighting(scs, nph, dims, k=K, seed=7)
    L = int(P.loff[-1])
    g = torch.Generator(device="cpu").manual_seed(5)
    e_true = (0.5 + 1.5 * torch.rand((L, 3), generator=g)).cuda()
    target = P.radiance(e_true, P.albedo0)
    torch.cuda.synchronize()
    setup_s = time.time() - t0
    # the two batched launches of a step, timed with events
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    fwd, bwd = [], []
    for it in range(3):
        P.bufs["io_rad4"].zero_()
        ev[0].record(stream)
        P.Q.run(4)
        ev[1].record(stream)
        P.Q.run(8)
        ev[2].record(stream)
        torch.cuda.synchronize()
        if it:
            fwd.append(ev[0].elapsed_time(ev[1]))
            bwd.append(ev[1].elapsed_time(ev[2]))
    torch.cuda.synchronize()
    t1 = time.time()
    _progress(f"inverse lighting set up in {setup_s:.1f}s: {P.num_queries} queries, {int(P.poff[-1])} photons")
    e_fit, a_fit, losses = P.fit(target, steps=args.inverse_steps, lr=1e-2, log=_progress)
    torch.cuda.synchronize()
    fit_s = time.time() - t1
    err0 = float((1.0 - e_true).abs().mean())
    err = float((e_fit - e_true).abs().mean())
    alb_dev = float((a_fit - P.albedo0).abs().mean())
    # per-scene check against the oracle at the fitted parameters
    t2 = time.time()
    P.radiance(e_fit, a_fit)
    eh, ah, et = e_fit.cpu().numpy(), a_fit.cpu().numpy(), e_true.cpu().numpy()
    per = {"rad_rel_err_fp64": [], "rad_fp32_bits_equal": [], "power_rel_err_fp64": [], "emission_mean_abs_err": []}
    # CPU baseline of one Adam step, timed on every 4th scene (1 warm-up + median of 3): the
    # oracle's kNN sets + fp32 estimate and sets + fp64 adjoint on the check's query sample,
    # the fp64 reweighting and its adjoint over all of the scene's photons; scaled by queries
    # and photons to the whole batch
    cst = dict(q=0.0, nq=0, ph=0.0, nph=0)
    for s in range(NS):
        A = P.scene_arrays(s)
        ls, ts = slice(int(P.loff[s]), int(P.loff[s + 1])), slice(int(P.toff[s]), int(P.toff[s + 1]))
        lt = A["light"].astype(np.int64) - int(P.loff[s])
        tr = A["tris"].astype(np.int64)
        tr = np.where(tr >= 0, tr - int(P.toff[s]), -1)
        p64 = O.reweight(A["base"], lt, tr, eh[ls], ah[ts])
        # relative to the fp64 power; a zero power (an albedo channel at 0) must be exactly 0
        rel = np.abs(A["ppow"][:, :3] - p64) / np.where(p64 != 0, np.abs(p64), 1.0)
        per["power_rel_err_fp64"].append(float(np.max(np.where((p64 == 0) & (A["ppow"][:, :3] != 0), np.inf, rel))))
        tree = O.build(1, A["ppos"], max_qr=A["r_max"], root=A["root"])
        sl = slice(None, None, args.sample // 4 or 1)
        q = {n: A[n][sl] for n in ("qpos", "qnrm", "qdir", "qinfl", "qclr", "rad")}
        nbr, cnt, _ = O.knn(tree, A["ppos"], q["qpos"], K, A["r_max"], nthreads=ncpu)
        o32, o64, _, _, _ = O.radiance(nbr, cnt, K, A["r_max"], A["ppos"], A["ppow"], A["pdir"], q["qpos"],
                                       q["qnrm"], q["qdir"], q["qinfl"], q["qclr"], nthreads=ncpu)
        per["rad_fp32_bits_equal"].append(bool(np.array_equal(q["rad"].view(np.uint32), o32.view(np.uint32))))
        scale = np.abs(o64).max(axis=1) + 1e-30
        per["rad_rel_err_fp64"].append(float(np.max(np.abs(q["rad"].astype(np.float64) - o64).max(axis=1) / scale)))
        per["emission_mean_abs_err"].append(float(np.abs(eh[ls] - et[ls]).mean()))
        if args.no_cpu or s % 4:
            continue
        _, knn_s = _median_time(lambda: O.knn(tree, A["ppos"], q["qpos"], K, A["r_max"], nthreads=ncpu), 3)
        _, rad_s = _median_time(lambda: O.radiance(nbr, cnt, K, A["r_max"], A["ppos"], A["ppow"], A["pdir"],
                                                   q["qpos"], q["qnrm"], q["qdir"], q["qinfl"], q["qclr"],
                                                   nthreads=ncpu), 3)
        gq = np.ones((len(q["qpos"]), 4), np.float32)
        (gp, gc), grad_s = _median_time(lambda: O.radiance_grad(nbr, cnt, K, A["r_max"], A["ppos"], A["ppow"],
                                                                A["pdir"], q["qpos"], q["qnrm"], q["qdir"], q["qinfl"],
                                                                q["qclr"], gq), 3)
        _, rw_s = _median_time(lambda: O.reweight(A["base"], lt, tr, eh[ls], ah[ts]), 3)
        _, rwg_s = _median_time(lambda: O.reweight_grad(A["base"], lt, tr, eh[ls], ah[ts], gp[:, :3]), 3)
        nqs = len(A["qpos"])
        cst["q"] += (2 * knn_s + rad_s + grad_s) * nqs / len(q["qpos"])
        cst["nq"] += nqs
        cst["ph"] += rw_s + rwg_s
        cst["nph"] += len(A["ppos"])
    cpu = None
    if cst["nq"] > 0:
        cpu_ms = (cst["q"] * P.num_queries / cst["nq"] + cst["ph"] * int(P.poff[-1]) / cst["nph"]) * 1e3
        cpu = {"ms_per_step": round(cpu_ms, 1), "cores": ncpu, "kind": "port",
               "gpu_speedup": round(cpu_ms / (fit_s / max(1, args.inverse_steps) * 1e3), 1),
               "sample": (f"oracle on every 4th scene ({cst['nq']} queries, {cst['nph']} photons): kNN sets + fp32 "
                          f"estimate and sets + fp64 adjoint at 1 in {args.sample // 4 or 1} hits (sets and estimate "
                          f"on {ncpu} threads, adjoint on one), fp64 reweighting and its adjoint over all photons "
                          "(numpy), each 1 warm-up + median of 3, scaled by queries and photons")}
    check_s = time.time() - t2
    _progress(f"inverse lighting oracle check of {NS} scenes: {check_s:.1f}s")
    steps = max(1, args.inverse_steps)
    out = {"workload": (f"{NS} scenes x 4 lights: triangles log-U[1k,50k], photons log-U[4k,256k], "
                        "queries log-U[16k,256k] (seed 5), k=50; Adam lr 1e-2 on emission + albedo"),
           "scenes": NS, "triangles": int(P.toff[-1]), "photons": int(P.poff[-1]), "queries": P.num_queries,
           "steps": args.inverse_steps, "ms_per_step": round(fit_s / steps * 1e3, 3),
           "knn_fwd_ms": round(float(np.median(fwd)), 3), "knn_bwd_ms": round(float(np.median(bwd)), 3),
           "bwd_fwd_ratio": round(float(np.median(bwd)) / float(np.median(fwd)), 3),
           "step_vs_knn_fwd_bwd": round(fit_s / steps * 1e3 / (float(np.median(fwd)) + float(np.median(bwd))), 3),
           "loss_first": losses[0] if losses else None, "loss_last": losses[-1] if losses else None,
           "emission_mean_abs_err": {"start": round(err0, 4), "end": round(err, 4)},
           "albedo_mean_abs_change": round(alb_dev, 5), "setup_s": round(setup_s, 1),
           "oracle_check": {
               "sample": f"1 in {args.sample // 4 or 1} hits of every scene, at the fitted parameters",
               "max_rad_rel_err_fp64": max(per["rad_rel_err_fp64"]),
               "all_rad_fp32_bits_equal": all(per["rad_fp32_bits_equal"]),
               "max_power_rel_err_fp64": max(per["power_rel_err_fp64"]),
               "per_scene": {k_: ([round(v, 9) for v in vs] if isinstance(vs[0], float) else vs)
                             for k_, vs in per.items()}},
           "cpu_baseline": cpu}
    P.release()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", choices=["large", "small"], default="large")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--sample", type=int, default=64, help="1-in-N subsample for the oracle replay")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-scenes leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--no-shadow", action="store_true", help="skip the shadow-ray leg")
    ap.add_argument("--no-build", action="store_true", help="skip the build benchmark leg")
    ap.add_argument("--build-runs", type=int, default=200)
    ap.add_argument("--build-warmup", type=int, default=10)
    ap.add_argument("--serial", action="store_true",
                    help="build the photon tree after the traces on the same stream (default: concurrently)")
    ap.add_argument("--host-builds", action="store_true",
                    help="build both trees with the host-driven level loop every step (default: replay builds "
                         "recorded at set-up)")
    ap.add_argument("--batched-scenes", type=int, default=256)
    ap.add_argument("--batched-steps", type=int, default=2)
    ap.add_argument("--no-inverse", action="store_true", help="skip the batched inverse-lighting leg")
    ap.add_argument("--inverse-scenes", type=int, default=64)
    ap.add_argument("--inverse-steps", type=int, default=200)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run one 256 MiB device-to-device copy (calibrates FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the hot path shards rays/queries across GPUs by replication, see the design notes)",
              file=sys.stderr)
        sys.exit(3)

    import torch

    from mnrt_amd import api, native, scene as S

    native.check(native.load().mnrt_device_init(0), "mnrt_device_init")
    torch.cuda.set_device(0)
    if args.copy_calibration:
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 4}))
        return

    t_setup = time.time()
    if args.config == "large":
        sc = S.scene_large()
        W, H, K, NPH = 1920, 1080, 100, 1_000_000
        bounce = True
        workload = "plain-large: box + 48 icospheres (subdiv 5) + 50k random tris, 1M photons, 1920x1080 + 1 bounce, k=100"
    else:
        sc = S.scene_small(subdiv=5)
        W, H, K, NPH = 512, 512, 50, 65536
        bounce = False
        workload = "plain-small: box + icosphere subdiv 5, 65536 photons, 512x512 primary, k=50"
    eps, rmax = sc.ray_epsilon(), sc.r_max()
    lo, hi = sc.bounds()
    root = np.concatenate([lo, hi]).astype(np.float32)
    v0, v1, v2, alb = (api.dev(a) for a in (sc.v0, sc.v1, sc.v2, sc.albedo))
    o_np, d_np, _ = S.camera_rays(W, H)
    go, gd = api.dev(o_np), api.dev(d_np)
    # photon map (set-up, not timed)
    T0 = api.build_tri(v0, v1, v2)
    ph = api.photons_generate(T0, v0, v1, v2, alb, sc.lights, NPH, 4, 0, eps)
    T0.release()
    torch.cuda.synchronize()
    ppos, ppow, pdir = ph["pos"].contiguous(), ph["power"].contiguous(), ph["dir"].contiguous()
    pparams = api.default_params(1, root, rmax)
    print(f"# setup {time.time() - t_setup:.1f}s: {sc.num_tris} tris, {ph['count']} photons from {ph['paths']} paths",
          file=sys.stderr, flush=True)

    stream = torch.cuda.current_stream()
    phases = ["build_tri", "trace_primary", "shade_bounce", "trace_bounce", "build_points", "knn"]

    # Both trees are rebuilt from their inputs every step.  By default each build replays a
    # recorded build (mnrt_build_graph_*: the level loop captured once at set-up into a
    # hipGraph whose kernels read every count on the device, so a replay is one launch call
    # and no host synchronisation; replays equal host-driven builds word for word).  The
    # photon tree does not depend on the triangle tree or the rays: unless --serial it is
    # built on a second stream, concurrently with the triangle tree and the traces.  With
    # --host-builds both run the host-driven level loop, the photon tree from a second host
    # thread (the library keeps its build workspace per host thread).
    side = torch.cuda.Stream()
    graphs = not args.host_builds
    worker = None if graphs or args.serial else concurrent.futures.ThreadPoolExecutor(max_workers=1)
    if graphs:
        GT = api.BuildGraph(native.KIND_TRI, np.array([0, v0.shape[0]], np.uint32), v0, v1, v2)
        GP = api.BuildGraph(native.KIND_POINT, np.array([0, ppos.shape[0]], np.uint32), ppos, power=ppow, dirs=pdir,
                            params=pparams, root_boxes=root[None, :], max_query_radius=np.array([rmax], np.float32))
        torch.cuda.synchronize()

    def replayed(G, st, what):
        flags = G.status(st)
        if flags != 0:
            raise RuntimeError(f"recorded {what} build outgrew its buffers (flags {flags:#x})")
        return G.trees[0]

    def photon_tree():
        with torch.cuda.stream(side):
            P = api.build_points(ppos, pparams)
            api.attach_photon_records(P, ppos, ppow, pdir)
            done = torch.cuda.Event()
            done.record(side)
        return P, done

    def step(record):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(len(phases) + 1)]
        kev = []
        ev[0].record(stream)
        pending = pdone = None
        if graphs:
            # the triangle tree first: it is the one the next stages wait for (submitted
            # second, its launches queue behind the photon tree's: 0.1 ms per step slower)
            GT.replay(stream)
            if not args.serial:
                side.wait_event(ev[0])  # the last step's queries of the photon tree are done
                GP.replay(side)
                pdone = torch.cuda.Event()
                pdone.record(side)
            T = replayed(GT, stream, "triangle")
        else:
            pending = None if args.serial else worker.submit(photon_tree)
            T = api.build_tri(v0, v1, v2)
        ev[1].record(stream)
        tri, t, _, _ = api.trace_closest(T, go, gd, eps)
        ev[2].record(stream)
        q1 = api.shading_points(go, gd, tri, t, v0, v1, v2, alb)
        queries = [q1]
        n_b = 0
        if bounce:
            bo, bd, binfl = api.bounce_rays(q1, 3)
            ev[3].record(stream)
            tri2, t2, _, _ = api.trace_closest(T, bo, bd, eps)
            ev[4].record(stream)
            q2 = api.shading_points(bo, bd, tri2, t2, v0, v1, v2, alb, infl=binfl)
            queries.append(q2)
            n_b = bo.shape[0]
        else:
            ev[3].record(stream)
            ev[4].record(stream)
        if graphs and pdone is not None:
            stream.wait_event(pdone)
            P = replayed(GP, side, "photon")
        elif graphs:
            GP.replay(stream)
            P = replayed(GP, stream, "photon")
        elif pending is None:
            P = api.build_points(ppos, pparams)
            api.attach_photon_records(P, ppos, ppow, pdir)
        else:
            P, done = pending.result()
            stream.wait_event(done)
        ev[5].record(stream)
        outs = []
        for q in queries:
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            outs.append(api.knn_gather(P, K, rmax, q["pos"], q["nrm"], q["dir"], q["infl"], q["clr"]))
            b.record(stream)
            kev.append((a, b, q["count"]))
        ev[6].record(stream)
        torch.cuda.synchronize()
        res = dict(rays=go.shape[0] + n_b, queries=sum(q["count"] for q in queries),
                   ms={p: ev[i].elapsed_time(ev[i + 1]) for i, p in enumerate(phases)},
                   knn_launch_ms=[(a.elapsed_time(b), n) for a, b, n in kev],
                   trace_ms=ev[1].elapsed_time(ev[2]) + (ev[3].elapsed_time(ev[4]) if bounce else 0.0))
        if record is not None:
            record.update(T=T, P=P, queries=queries, tri=tri, tri2=tri2 if bounce else None,
                          bounce_rays=(bo, bd) if bounce else None)
        else:
            T.release()
            P.release()
        return res

    for _ in range(args.warmup):
        step(None)
    torch.cuda.synchronize()
    t0 = time.time()
    results = [step(None) for _ in range(args.steps)]
    torch.cuda.synchronize()
    wall = time.time() - t0
    _progress(f"{args.steps} timed steps: {wall * 1e3 / args.steps:.2f} ms per step")
    ms_step = wall * 1e3 / args.steps
    rays, queries = results[0]["rays"], results[0]["queries"]
    phase_ms = {p: float(np.mean([r["ms"][p] for r in results])) for p in phases}
    knn_launch = [x for r in results for x in r["knn_launch_ms"]]
    knn_ms = float(np.sum([m for m, _ in knn_launch])) / args.steps
    trace_ms = float(np.mean([r["trace_ms"] for r in results]))
    value = (rays + queries) / (ms_step * 1e-3) / 1e6

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "M(rays+kNN queries)/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator, traced photons)",
        "config": {"workload": workload, "triangles": sc.num_tris, "photons": ph["count"], "rays": rays,
                   "tree_builds": "host-driven level loop" if args.host_builds else
                   "replays of builds recorded at set-up (hipGraph, counts on the device)",
                   "photon_tree": "after the traces, one stream" if args.serial else
                   "concurrent with the triangle tree and traces (second stream)",
                   "knn_queries": queries, "k": K, "parallelism": "single GPU"},
        "phase_ms": {p: round(v, 3) for p, v in phase_ms.items()},
        "trace_mrays_per_s": round(rays / (trace_ms * 1e-3) / 1e6, 2),
        "knn_mqueries_per_s": round(queries / (knn_ms * 1e-3) / 1e6, 3),
    }

    # ---- oracle replay on a sample: algorithmic bytes, then the CPU baseline ----
    rec = {}
    step(rec)
    from oracle import oracle as O

    T, P = rec["T"], rec["P"]
    pn, pr = P.nodes_host(), P.refs_host()
    ptree = O.Tree(1, pn, pr, P.root, P.max_level, 0, ppos.shape[0])
    ppos_h = ppos.cpu().numpy()
    qpos = np.concatenate([q["pos"].cpu().numpy() for q in rec["queries"]])
    qs = qpos[:: args.sample]
    # host threads for the oracle: the cores this process may use, capped by OMP_NUM_THREADS
    ncpu = O.host_threads()
    reps = 0 if args.no_cpu else 5
    (nbr, cnt, kctr), cpu_knn_s = _median_time(lambda: O.knn(ptree, ppos_h, qs, K, rmax, nthreads=ncpu), reps)
    _progress(f"oracle kNN replay of {len(qs)} queries: {cpu_knn_s:.2f}s")
    per_q = knn_bytes(kctr, len(qs), int(np.minimum(cnt, K).sum())) / len(qs)
    tn, tr = T.nodes_host(), T.refs_host()
    ttree = O.Tree(0, tn, tr, T.root, T.max_level, 0, sc.num_tris)
    ray_o = [o_np]
    ray_d = [d_np]
    if rec["bounce_rays"] is not None:
        ray_o.append(rec["bounce_rays"][0].cpu().numpy())
        ray_d.append(rec["bounce_rays"][1].cpu().numpy())
    ro = np.concatenate(ray_o)[:: args.sample]
    rd = np.concatenate(ray_d)[:: args.sample]
    (_, _, _, tctr), cpu_trace_s = _median_time(lambda: O.trace(ttree, sc.v0, sc.v1, sc.v2, ro, rd, eps,
                                                                nthreads=ncpu), reps)
    _progress(f"oracle traversal replay of {len(ro)} rays: {cpu_trace_s:.3f}s")
    per_ray = trace_bytes(tctr, len(ro)) / len(ro)
    if not args.no_shadow:
        out["shadow"] = shadow_leg(torch, api, native, O, T, ttree, sc, rec["queries"][0]["pos"], eps, args.sample,
                                   ncpu)
    T.release()
    P.release()

    launches = [(m, n) for m, n in knn_launch]
    avg_ms = float(np.mean([m for m, _ in launches]))
    avg_q = float(np.mean([n for _, n in launches]))
    achieved = per_q * avg_q / (avg_ms * 1e-3) / 1e9
    traffic = None
    prof = os.path.join(ROOT, "profiles", f"k_knn_{args.config}_traffic.json")
    if os.path.exists(prof):
        with open(prof) as f:
            traffic = json.load(f).get("bytes_per_launch")
    out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                       "kernel": "k_knn (mnrt_knn_gather)", "alg_bytes_per_query": round(per_q, 1),
                       "avg_launch_ms": round(avg_ms, 4), "queries_per_launch": avg_q}
    trace_gbs = per_ray * rays / (trace_ms * 1e-3) / 1e9
    out["trace_roofline"] = {"kernel": "k_trace_closest", "achieved": round(trace_gbs, 2), "unit": "GB/s",
                             "frac": round(trace_gbs / HBM_PEAK_GBS, 5), "alg_bytes_per_ray": round(per_ray, 1)}
    out["knn_pct_hbm"] = round(100.0 * achieved / HBM_PEAK_GBS, 3)
    out["trace_pct_hbm"] = round(100.0 * trace_gbs / HBM_PEAK_GBS, 3)

    # oracle builds: the CPU baseline's build time and the build roofline's bytes
    (otri, cpu_build_tri) = _median_time(lambda: O.build(0, sc.v0, sc.v1, sc.v2, nthreads=ncpu), reps)
    _progress(f"oracle triangle build: {cpu_build_tri:.2f}s")
    (opts, cpu_build_pts) = _median_time(lambda: O.build(1, ppos_h, max_qr=rmax, root=root, nthreads=ncpu), reps)
    if not args.no_build:
        _progress("build leg")
        out["build"] = build_leg(args, torch, api, O, v0, v1, v2, ppos, pparams, otri, opts)
    if not args.no_cpu:
        cpu_step = cpu_build_tri + cpu_build_pts + rays * (cpu_trace_s / len(ro)) + queries * (cpu_knn_s / len(qs))
        out["cpu_baseline"] = {
            "value": round((rays + queries) / cpu_step / 1e6, 4), "unit": "M(rays+kNN queries)/s", "cores": ncpu,
            "kind": "port", "cpu_model": O.cpu_model(), "host_cpus": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0)),
            "sample": (f"oracle on the host, each part 1 warm-up + median of 5: both tree builds in full "
                       f"(level-parallel large stage, per-root parallel small stage, {ncpu} threads: "
                       f"{cpu_build_tri:.2f}s + {cpu_build_pts:.2f}s), traversal of 1 in {args.sample} "
                       f"rays ({len(ro)} rays, {cpu_trace_s:.3f}s) and kNN of 1 in {args.sample} queries "
                       f"({len(qs)}, {cpu_knn_s:.3f}s) on {ncpu} threads, scaled to the full step ({cpu_step:.1f}s)"),
            "gpu_speedup": round(cpu_step / (ms_step * 1e-3), 1)}
    if not args.no_batched:
        _progress("batched leg")
        out["batched"] = batched_leg(args, torch, api, S, O, ncpu)
        out["batched_vs_plain"] = out["batched"]["speedup_vs_plain"]
    if not args.no_adjoint:
        _progress("adjoint leg")
        out["adjoint"] = adjoint_leg(args, torch, api, S, O, ncpu)
        out["bwd_fwd_ratio"] = out["adjoint"]["bwd_fwd_ratio"]
    if not args.no_inverse:
        _progress("inverse-lighting leg")
        out["inverse"] = inverse_leg(args, torch, api, S, O, ncpu)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
