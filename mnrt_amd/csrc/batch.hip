// batch.hip — many scenes traced, shaded and queried together (mnrt_batch_*).
//
// No reference counterpart: the reference binds one scene at a time into global device
// state and launches KernelRTTraceRays / KernelPMPhotonGather per scene
// (RTCore.cpp:404-439, PhotonMap.cpp:195-232).  Here the per-scene trees and parameters
// live in device descriptor tables; each phase is one persistent launch whose
// workgroups take tiles from device work counters (one counter per XCD share), so one
// launch covers every scene however uneven their sizes, and the phases enqueue
// kernels only and can be captured in hipGraphs.
//
// The autotuner is a coordinate search: for each launch parameter in turn it times
// every candidate with HIP events on the caller's buffers (the other parameters held)
// and keeps the fastest, then moves to the next parameter.
#include <algorithm>
#include <cstring>
#include <vector>

#include "batch_dev.h"
#include "scan.h"

struct mnrt_batch {
    uint32_t ns = 0;
    uint32_t k = 0;
    bool has_knn = false;
    uint32_t total_slots = 0;
    uint32_t max_level = 0;
    uint32_t cus = 256;
    std::vector<mnrt::TraceScene> h_trace;
    std::vector<mnrt::KnnScene> h_knn;
    std::vector<uint32_t> h_slot;  // ns + 1
    std::vector<uint32_t> h_tri;   // ns
    mnrt_batch_tuning tun{};
    uint32_t limit_slots = 0;      // 0: all (autotuning runs on a prefix)
    uint32_t total_tiles = 0;
    // device tables (one allocation)
    void* d_buf = nullptr;
    mnrt::TraceScene* d_trace = nullptr;
    mnrt::KnnScene* d_knn = nullptr;
    uint32_t* d_tile_begin = nullptr;  // ns + 1
    uint32_t* d_slot_begin = nullptr;  // ns + 1
    uint32_t* d_tri_begin = nullptr;   // ns
    uint32_t* d_tparts = nullptr;      // 2 * kMaxParts: trace part begin / end (tiles)
    uint32_t* d_kparts = nullptr;      // 2 * kMaxParts: kNN part begin / end (slots)
    uint32_t* d_work = nullptr;        // 2 * kMaxParts
    uint32_t* d_flags = nullptr;       // total_slots + 1
    uint32_t* d_pos = nullptr;         // total_slots + 1
    uint32_t* d_scan = nullptr;
    size_t tiles_cap = 0;
    // kNN adjoint state (allocated on the first MNRT_PHASE_KNN_BWD run or capture)
    std::vector<uint32_t> h_poff;  // ns + 1: photon offsets of the scenes
    void* d_bwd = nullptr;
    size_t bwd_cap = 0;            // photons the accumulators hold
    size_t bwd_graph_photons = 0;  // total photons baked into captured adjoint graphs
    unsigned long long* d_acc = nullptr;
    uint32_t* d_wmax = nullptr;
    double* d_scale = nullptr;
    uint32_t* d_poff = nullptr;
    hipGraph_t graph[16] = {};
    hipGraphExec_t exec[16] = {};
    hipStream_t cap_stream = nullptr;
};

namespace {

using namespace mnrt;

bool valid_tuning(const mnrt_batch& b, const mnrt_batch_tuning& t) {
    if (t.trace_block != 64 && t.trace_block != 128 && t.trace_block != 256) return false;
    if (t.trace_tile == 0 || t.trace_tile % t.trace_block != 0 || t.trace_tile > (1u << 20)) return false;
    if (t.trace_grid == 0 || t.trace_grid > 65536) return false;
    if (t.trace_stack != 16 && t.trace_stack != 24 && t.trace_stack != 32 && t.trace_stack != (uint32_t)kTraceStackAbi)
        return false;
    if (t.trace_stack < b.max_level) return false;
    if (t.knn_waves != 1 && t.knn_waves != 2 && t.knn_waves != 4) return false;
    if (t.knn_chunk == 0 || t.knn_chunk > 4096 || t.knn_grid == 0 || t.knn_grid > 65536) return false;
    if (t.knn_cap != 128 && t.knn_cap != 256 && t.knn_cap != 512) return false;
    if (b.has_knn && t.knn_cap < b.k + 64) return false;
    if (t.knn_cap == 512 && t.knn_waves > 4) return false;
    if (t.parts != 1 && t.parts != kMaxParts) return false;
    if (t.knn_select > 2) return false;
    return true;
}

void drop_graphs(mnrt_batch& b, uint32_t with_phase = 0) {
    for (uint32_t i = 0; i < 16; ++i) {
        if (with_phase && !(i & with_phase)) continue;
        if (b.exec[i]) (void)hipGraphExecDestroy(b.exec[i]);
        if (b.graph[i]) (void)hipGraphDestroy(b.graph[i]);
        b.exec[i] = nullptr;
        b.graph[i] = nullptr;
    }
}

// Tile tables and work shares for the current tuning (synchronous upload).
int upload_tables(mnrt_batch& b) {
    const uint32_t ns = b.ns;
    const uint32_t lim = b.limit_slots ? std::min(b.limit_slots, b.total_slots) : b.total_slots;
    std::vector<uint32_t> tile(ns + 1, 0);
    uint32_t nt = 0, nt_lim = 0;
    for (uint32_t s = 0; s < ns; ++s) {
        tile[s] = nt;
        const uint32_t rays = b.h_slot[s + 1] - b.h_slot[s];
        const uint32_t tiles = (rays + b.tun.trace_tile - 1) / b.tun.trace_tile;
        for (uint32_t j = 0; j < tiles; ++j) {
            if (b.h_slot[s] + j * b.tun.trace_tile < lim) nt_lim = nt + j + 1;
        }
        nt += tiles;
    }
    tile[ns] = nt;
    b.total_tiles = nt;
    const uint32_t P = b.tun.parts;
    std::vector<uint32_t> parts(4 * kMaxParts, 0);
    for (uint32_t p = 0; p < kMaxParts; ++p) {
        uint32_t tb = p < P ? (uint32_t)((uint64_t)nt_lim * p / P) : nt_lim;
        uint32_t te = p < P ? (uint32_t)((uint64_t)nt_lim * (p + 1) / P) : nt_lim;
        uint32_t sb = p < P ? (uint32_t)((uint64_t)lim * p / P) : lim;
        uint32_t se = p < P ? (uint32_t)((uint64_t)lim * (p + 1) / P) : lim;
        parts[p] = tb;
        parts[kMaxParts + p] = te;
        parts[2 * kMaxParts + p] = sb;
        parts[3 * kMaxParts + p] = se;
    }
    MNRT_HIP_CHECK(hipMemcpy(b.d_tile_begin, tile.data(), (ns + 1) * 4, hipMemcpyHostToDevice));
    MNRT_HIP_CHECK(hipMemcpy(b.d_tparts, parts.data(), 4 * kMaxParts * 4, hipMemcpyHostToDevice));
    return MNRT_OK;
}

int run_trace(mnrt_batch& b, const mnrt_batch_io& io, hipStream_t st) {
    if (!io.d_orig4 || !io.d_dir4 || !io.d_hit_tri) return MNRT_ERR_ARG;
    TraceBatchArgs A{};
    A.scenes = b.d_trace;
    A.tile_begin = b.d_tile_begin;
    A.num_scenes = b.ns;
    A.tile_rays = b.tun.trace_tile;
    A.num_parts = b.tun.parts;
    A.part_begin = b.d_tparts;
    A.part_end = b.d_tparts + kMaxParts;
    A.work = b.d_work;
    A.orig = (const float4*)io.d_orig4;
    A.dir = (const float4*)io.d_dir4;
    A.out_tri = io.d_hit_tri;
    A.out_t = io.d_hit_t;
    A.out_bary = (float2*)io.d_hit_bary2;
    return launch_trace_batch(A, (int)b.tun.trace_stack, b.tun.trace_grid, b.tun.trace_block, st);
}

int run_shade(mnrt_batch& b, const mnrt_batch_io& io, hipStream_t st) {
    if (!io.d_orig4 || !io.d_dir4 || !io.d_hit_tri || !io.d_hit_t || !io.d_v0 || !io.d_v1 || !io.d_v2 ||
        !io.d_albedo4 || !io.d_sp_pos4 || !io.d_sp_nrm4 || !io.d_sp_dir4 || !io.d_sp_clr4 || !io.d_sp_infl4 ||
        !io.d_sp_count)
        return MNRT_ERR_ARG;
    ShadeBatchArgs A{};
    A.n = b.total_slots;
    A.num_scenes = b.ns;
    A.slot_begin = b.d_slot_begin;
    A.tri_begin = b.d_tri_begin;
    A.orig = (const float4*)io.d_orig4;
    A.dir = (const float4*)io.d_dir4;
    A.tri = io.d_hit_tri;
    A.t = io.d_hit_t;
    A.v0 = (const float4*)io.d_v0;
    A.v1 = (const float4*)io.d_v1;
    A.v2 = (const float4*)io.d_v2;
    A.albedo = (const float4*)io.d_albedo4;
    A.infl = (const float4*)io.d_infl4;
    A.pos = b.d_pos;
    A.out_pos = (float4*)io.d_sp_pos4;
    A.out_nrm = (float4*)io.d_sp_nrm4;
    A.out_dir = (float4*)io.d_sp_dir4;
    A.out_clr = (float4*)io.d_sp_clr4;
    A.out_infl = (float4*)io.d_sp_infl4;
    A.out_src = io.d_sp_src;
    A.out_count = io.d_sp_count;
    return launch_shade_batch(A, b.d_flags, b.d_scan, st);
}

int run_knn(mnrt_batch& b, const mnrt_batch_io& io, hipStream_t st) {
    if (!b.has_knn) return MNRT_ERR_STATE;
    if (!io.d_sp_pos4 || !io.d_sp_nrm4 || !io.d_sp_dir4 || !io.d_sp_infl4 || !io.d_sp_clr4 || !io.d_io_rad4)
        return MNRT_ERR_ARG;
    KnnArgs G{};
    G.k = b.k;
    G.n = b.total_slots;
    G.pos = (const float4*)io.d_sp_pos4;
    G.nrm = (const float4*)io.d_sp_nrm4;
    G.dir = (const float4*)io.d_sp_dir4;
    G.infl = (const float4*)io.d_sp_infl4;
    G.clr = (const float4*)io.d_sp_clr4;
    G.rad_out = (float4*)io.d_io_rad4;
    G.accumulate = 1;
    G.out_r2 = io.d_out_r2;
    G.out_nbr = io.d_out_nbr;
    G.out_count = io.d_out_count;
    return launch_knn_batch(G, b.d_knn, b.d_slot_begin, b.ns, io.d_sp_count, b.tun.knn_chunk, b.tun.parts,
                            b.d_tparts + 2 * kMaxParts, b.d_tparts + 3 * kMaxParts, b.d_work + kMaxParts,
                            b.tun.knn_cap, b.tun.knn_waves, b.tun.knn_select, b.tun.knn_grid, st);
}

// Accumulators of the adjoint phase for the current photon counts (synchronous; never
// called while a stream is being captured).  Graphs of the adjoint phase are dropped when
// the accumulators move or the total photon count changes (a captured adjoint clears and
// converts exactly that many photons' sums: launch arguments fixed at capture).
int ensure_bwd(mnrt_batch& b) {
    if (!b.has_knn) return MNRT_ERR_STATE;
    const uint32_t ns = b.ns;
    const size_t np = b.h_poff[ns];
    if (np != b.bwd_graph_photons) drop_graphs(b, MNRT_PHASE_KNN_BWD);
    if (!b.d_bwd || np > b.bwd_cap) {
        MNRT_HIP_CHECK(hipDeviceSynchronize());
        drop_graphs(b, MNRT_PHASE_KNN_BWD);
        if (b.d_bwd) MNRT_HIP_CHECK(hipFree(b.d_bwd));
        b.d_bwd = nullptr;
        const size_t cap = std::max<size_t>(np, 1);
        const size_t o_wmax = (24 * cap + 255) & ~(size_t)255;
        const size_t o_scale = o_wmax + ((4 * (size_t)ns + 255) & ~(size_t)255);
        const size_t o_poff = o_scale + ((8 * (size_t)ns + 255) & ~(size_t)255);
        const size_t bytes = o_poff + 4 * ((size_t)ns + 1);
        if (hipMalloc(&b.d_bwd, bytes) != hipSuccess) {
            b.d_bwd = nullptr;
            b.bwd_cap = 0;
            return MNRT_ERR_NOMEM;
        }
        char* base = (char*)b.d_bwd;
        b.d_acc = (unsigned long long*)base;
        b.d_wmax = (uint32_t*)(base + o_wmax);
        b.d_scale = (double*)(base + o_scale);
        b.d_poff = (uint32_t*)(base + o_poff);
        b.bwd_cap = cap;
    }
    MNRT_HIP_CHECK(hipMemcpy(b.d_poff, b.h_poff.data(), 4 * ((size_t)ns + 1), hipMemcpyHostToDevice));
    return MNRT_OK;
}

int run_knn_bwd(mnrt_batch& b, const mnrt_batch_io& io, hipStream_t st) {
    if (!b.has_knn) return MNRT_ERR_STATE;
    if (!b.d_bwd) return MNRT_ERR_STATE;
    if (!io.d_sp_pos4 || !io.d_sp_nrm4 || !io.d_sp_dir4 || !io.d_sp_infl4 || !io.d_sp_clr4 || !io.d_out_r2 ||
        !io.d_out_count || !io.d_grad_rad4 || !io.d_grad_pow4)
        return MNRT_ERR_ARG;
    KnnArgs G{};
    G.k = b.k;
    G.n = b.total_slots;
    G.pos = (const float4*)io.d_sp_pos4;
    G.nrm = (const float4*)io.d_sp_nrm4;
    G.dir = (const float4*)io.d_sp_dir4;
    G.infl = (const float4*)io.d_sp_infl4;
    G.clr = (const float4*)io.d_sp_clr4;
    G.grad = (const float4*)io.d_grad_rad4;
    G.acc = b.d_acc;
    G.out_grad_clr = (float4*)io.d_grad_clr4;
    return launch_knn_bwd_batch(G, b.d_knn, b.d_slot_begin, b.ns, io.d_sp_count, b.d_poff, b.h_poff[b.ns],
                                io.d_out_r2, io.d_out_count, b.d_wmax, b.d_scale, b.tun.parts,
                                b.d_tparts + 2 * kMaxParts, b.d_tparts + 3 * kMaxParts, b.d_work + kMaxParts,
                                b.tun.knn_grid, io.d_grad_pow4, st);
}

int run_phases(mnrt_batch& b, const mnrt_batch_io& io, uint32_t phases, hipStream_t st) {
    int rc = MNRT_OK;
    if (phases & MNRT_PHASE_TRACE) rc = run_trace(b, io, st);
    if (rc == MNRT_OK && (phases & MNRT_PHASE_SHADE)) rc = run_shade(b, io, st);
    if (rc == MNRT_OK && (phases & MNRT_PHASE_KNN)) rc = run_knn(b, io, st);
    if (rc == MNRT_OK && (phases & MNRT_PHASE_KNN_BWD)) rc = run_knn_bwd(b, io, st);
    return rc;
}

// Milliseconds of one run of the phases (after one untimed run).
int time_phases(mnrt_batch& b, const mnrt_batch_io& io, uint32_t phases, hipStream_t st, float* ms) {
    hipEvent_t e0, e1;
    MNRT_HIP_CHECK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return MNRT_ERR_HIP;
    }
    int rc = run_phases(b, io, phases, st);
    if (rc == MNRT_OK && hipEventRecord(e0, st) != hipSuccess) rc = MNRT_ERR_HIP;
    if (rc == MNRT_OK) rc = run_phases(b, io, phases, st);
    if (rc == MNRT_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                          hipEventElapsedTime(ms, e0, e1) != hipSuccess))
        rc = MNRT_ERR_HIP;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

uint32_t smallest_stack(uint32_t max_level) {
    for (uint32_t s : {16u, 24u, 32u}) {
        if (max_level <= s) return s;
    }
    return (uint32_t)kTraceStackAbi;
}

uint32_t min_cap(uint32_t k) {
    uint32_t c = 128;
    while (c < k + 64) c <<= 1;
    return c;
}

// Per-scene descriptors from the trees (host side); b.h_slot must be set.
int fill_scenes(mnrt_batch& b, const mnrt_tree* tri, const mnrt_tree* pm, const float* eps, const float* r_max) {
    const uint32_t ns = b.ns;
    const bool knn = b.has_knn;
    uint32_t max_level = 0;
    b.h_poff.assign(ns + 1, 0);
    for (uint32_t s = 0; s < ns; ++s) {
        const mnrt_tree& t = tri[s];
        if (t.kind != MNRT_KIND_TRI || !t.d_nodes || !t.d_recs || t.num_nodes == 0) return MNRT_ERR_STATE;
        if (t.max_level > MNRT_MAX_LEVEL) return MNRT_ERR_CAPACITY;
        TraceScene& ts = b.h_trace[s];
        ts.nodes = (const uint2*)t.d_nodes;
        ts.recs = (const float4*)t.d_recs;
        for (int a = 0; a < 3; ++a) {
            ts.lo[a] = t.root_min[a];
            ts.hi[a] = t.root_max[a];
        }
        if (eps) ts.eps = eps[s];
        ts.ray_begin = b.h_slot[s];
        ts.ray_end = b.h_slot[s + 1];
        max_level = std::max(max_level, t.max_level);
        if (knn) {
            const mnrt_tree& p = pm[s];
            if (p.kind != MNRT_KIND_POINT || !p.d_nodes || !p.d_recs || p.num_nodes == 0) return MNRT_ERR_STATE;
            if (p.max_level > MNRT_MAX_LEVEL || p.num_nodes >= kUniform8MaxIndex) return MNRT_ERR_CAPACITY;
            KnnScene& kq = b.h_knn[s];
            kq.nodes = (const uint2*)(p.d_qnodes ? p.d_qnodes : p.d_nodes);
            kq.recs = (const float4*)p.d_recs;
            if (r_max) {
                if (!(r_max[s] > 0.0f)) return MNRT_ERR_ARG;
                kq.r_max = r_max[s];
                kq.r2max = r_max[s] * r_max[s];
            }
            kq.slot_begin = b.h_slot[s];
            kq.slot_end = b.h_slot[s + 1];
            if ((uint64_t)b.h_poff[s] + p.num_elems > 0xFFFFFFFFull) return MNRT_ERR_CAPACITY;
            b.h_poff[s + 1] = b.h_poff[s] + p.num_elems;
        }
    }
    b.max_level = max_level;
    return MNRT_OK;
}

}  // namespace

extern "C" int mnrt_batch_create(const mnrt_batch_desc* d, mnrt_batch** out) {
    if (!d || !out || d->num_scenes == 0 || !d->tri_trees || !d->tri_offsets || !d->ray_offsets || !d->eps)
        return MNRT_ERR_ARG;
    *out = nullptr;
    const uint32_t ns = d->num_scenes;
    const bool knn = d->photon_maps != nullptr;
    if (knn && (!d->r_max || d->k == 0 || d->k > MNRT_KNN_MAX_K)) return MNRT_ERR_ARG;
    auto* b = new (std::nothrow) mnrt_batch;
    if (!b) return MNRT_ERR_NOMEM;
    b->ns = ns;
    b->k = d->k;
    b->has_knn = knn;
    b->h_trace.resize(ns);
    b->h_knn.resize(ns);
    b->h_slot.assign(d->ray_offsets, d->ray_offsets + ns + 1);
    b->h_tri.assign(d->tri_offsets, d->tri_offsets + ns);
    for (uint32_t s = 0; s < ns; ++s) {
        if (d->ray_offsets[s + 1] < d->ray_offsets[s] || d->tri_offsets[s + 1] < d->tri_offsets[s]) {
            delete b;
            return MNRT_ERR_ARG;
        }
    }
    int rc = fill_scenes(*b, d->tri_trees, d->photon_maps, d->eps, d->r_max);
    b->total_slots = d->ray_offsets[ns];
    int dev = 0;
    hipDeviceProp_t prop;
    if (rc == MNRT_OK && (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess))
        rc = MNRT_ERR_HIP;
    if (rc == MNRT_OK) {
        b->cus = (uint32_t)prop.multiProcessorCount;
        const size_t n1 = (size_t)b->total_slots + 1;
        size_t bytes = 0;
        auto carve = [&](size_t n) {
            size_t o = bytes;
            bytes += (n + 255) & ~(size_t)255;
            return o;
        };
        size_t o_trace = carve(sizeof(TraceScene) * ns), o_knn = carve(sizeof(KnnScene) * ns);
        size_t o_tile = carve(4 * (ns + 1)), o_slot = carve(4 * (ns + 1)), o_tri = carve(4 * ns);
        size_t o_parts = carve(4 * 4 * kMaxParts), o_work = carve(4 * 2 * kMaxParts);
        size_t o_flags = carve(4 * n1), o_pos = carve(4 * n1), o_scan = carve(4 * scan_scratch_words((uint32_t)n1));
        if (hipMalloc(&b->d_buf, bytes) != hipSuccess) rc = MNRT_ERR_NOMEM;
        if (rc == MNRT_OK) {
            char* base = (char*)b->d_buf;
            b->d_trace = (TraceScene*)(base + o_trace);
            b->d_knn = (KnnScene*)(base + o_knn);
            b->d_tile_begin = (uint32_t*)(base + o_tile);
            b->d_slot_begin = (uint32_t*)(base + o_slot);
            b->d_tri_begin = (uint32_t*)(base + o_tri);
            b->d_tparts = (uint32_t*)(base + o_parts);
            b->d_work = (uint32_t*)(base + o_work);
            b->d_flags = (uint32_t*)(base + o_flags);
            b->d_pos = (uint32_t*)(base + o_pos);
            b->d_scan = (uint32_t*)(base + o_scan);
            if (hipMemcpy(b->d_trace, b->h_trace.data(), sizeof(TraceScene) * ns, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(b->d_knn, b->h_knn.data(), sizeof(KnnScene) * ns, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(b->d_slot_begin, b->h_slot.data(), 4 * (ns + 1), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(b->d_tri_begin, b->h_tri.data(), 4 * ns, hipMemcpyHostToDevice) != hipSuccess)
                rc = MNRT_ERR_HIP;
        }
    }
    if (rc == MNRT_OK) {
        mnrt_batch_tuning& t = b->tun;
        t.trace_block = 256;
        t.trace_tile = 1024;
        t.trace_grid = 8 * b->cus;
        t.trace_stack = smallest_stack(b->max_level);
        t.knn_waves = 4;
        t.knn_chunk = 16;
        t.knn_grid = 8 * b->cus;
        t.knn_cap = knn ? min_cap(b->k) : 128;
        t.parts = kMaxParts;
        t.knn_select = 2;
        rc = upload_tables(*b);
    }
    if (rc == MNRT_OK && hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking) != hipSuccess) rc = MNRT_ERR_HIP;
    if (rc != MNRT_OK) {
        mnrt_batch_destroy(b);
        return rc;
    }
    *out = b;
    return MNRT_OK;
}

extern "C" int mnrt_batch_destroy(mnrt_batch* b) {
    if (!b) return MNRT_ERR_ARG;
    drop_graphs(*b);
    int rc = MNRT_OK;
    if (b->cap_stream && hipStreamDestroy(b->cap_stream) != hipSuccess) rc = MNRT_ERR_HIP;
    if (b->d_buf && hipFree(b->d_buf) != hipSuccess) rc = MNRT_ERR_HIP;
    if (b->d_bwd && hipFree(b->d_bwd) != hipSuccess) rc = MNRT_ERR_HIP;
    delete b;
    return rc;
}

extern "C" int mnrt_batch_get_tuning(const mnrt_batch* b, mnrt_batch_tuning* out) {
    if (!b || !out) return MNRT_ERR_ARG;
    *out = b->tun;
    return MNRT_OK;
}

extern "C" int mnrt_batch_set_tuning(mnrt_batch* b, const mnrt_batch_tuning* t) {
    if (!b || !t || !valid_tuning(*b, *t)) return MNRT_ERR_ARG;
    MNRT_HIP_CHECK(hipDeviceSynchronize());
    drop_graphs(*b);
    b->tun = *t;
    return upload_tables(*b);
}

extern "C" int mnrt_batch_update_trees(mnrt_batch* b, const mnrt_tree* tri_trees, const mnrt_tree* photon_maps) {
    if (!b || !tri_trees || (b->has_knn && !photon_maps)) return MNRT_ERR_ARG;
    MNRT_HIP_CHECK(hipDeviceSynchronize());
    int rc = fill_scenes(*b, tri_trees, photon_maps, nullptr, nullptr);
    if (rc != MNRT_OK) return rc;
    if (b->tun.trace_stack < b->max_level) {
        drop_graphs(*b);
        b->tun.trace_stack = smallest_stack(b->max_level);
    }
    const uint32_t ns = b->ns;
    MNRT_HIP_CHECK(hipMemcpy(b->d_trace, b->h_trace.data(), sizeof(TraceScene) * ns, hipMemcpyHostToDevice));
    MNRT_HIP_CHECK(hipMemcpy(b->d_knn, b->h_knn.data(), sizeof(KnnScene) * ns, hipMemcpyHostToDevice));
    if (b->d_bwd) return ensure_bwd(*b);  // new photon offsets (and room, if they grew)
    return MNRT_OK;
}

extern "C" int mnrt_batch_run(mnrt_batch* b, const mnrt_batch_io* io, uint32_t phases, void* stream) {
    if (!b || !io || phases == 0 || phases > 15) return MNRT_ERR_ARG;
    if ((phases & MNRT_PHASE_KNN_BWD) && !b->d_bwd) {
        int rc = ensure_bwd(*b);
        if (rc != MNRT_OK) return rc;
    }
    // no slots: only the adjoint has work (zero photon gradients)
    if (b->total_slots == 0) phases &= MNRT_PHASE_KNN_BWD;
    if (phases == 0) return MNRT_OK;
    return run_phases(*b, *io, phases, (hipStream_t)stream);
}

extern "C" int mnrt_batch_capture(mnrt_batch* b, const mnrt_batch_io* io, uint32_t phases, void* stream) {
    if (!b || !io || phases == 0 || phases > 15) return MNRT_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    MNRT_HIP_CHECK(hipStreamSynchronize(st));
    if ((phases & MNRT_PHASE_KNN_BWD) && !b->d_bwd) {
        int rc = ensure_bwd(*b);
        if (rc != MNRT_OK) return rc;
    }
    if (b->exec[phases]) (void)hipGraphExecDestroy(b->exec[phases]);
    if (b->graph[phases]) (void)hipGraphDestroy(b->graph[phases]);
    b->exec[phases] = nullptr;
    b->graph[phases] = nullptr;
    MNRT_HIP_CHECK(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
    const uint32_t run = b->total_slots ? phases : (phases & MNRT_PHASE_KNN_BWD);
    int rc = run ? run_phases(*b, *io, run, b->cap_stream) : MNRT_OK;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
    if (rc != MNRT_OK || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rc != MNRT_OK ? rc : MNRT_ERR_HIP;
    }
    hipGraphExec_t x = nullptr;
    if (hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g);
        return MNRT_ERR_HIP;
    }
    b->graph[phases] = g;
    b->exec[phases] = x;
    if (phases & MNRT_PHASE_KNN_BWD) b->bwd_graph_photons = b->h_poff[b->ns];
    return MNRT_OK;
}

extern "C" int mnrt_batch_replay(mnrt_batch* b, uint32_t phases, void* stream) {
    if (!b || phases == 0 || phases > 15) return MNRT_ERR_ARG;
    if (!b->exec[phases]) return MNRT_ERR_STATE;
    MNRT_HIP_CHECK(hipGraphLaunch(b->exec[phases], (hipStream_t)stream));
    return MNRT_OK;
}

extern "C" int mnrt_batch_autotune(mnrt_batch* b, const mnrt_batch_io* io, uint32_t max_slots, void* stream) {
    if (!b || !io) return MNRT_ERR_ARG;
    if (b->total_slots == 0) return MNRT_OK;
    hipStream_t st = (hipStream_t)stream;
    MNRT_HIP_CHECK(hipStreamSynchronize(st));
    drop_graphs(*b);
    b->limit_slots = max_slots;
    mnrt_batch_tuning best = b->tun;
    int rc = upload_tables(*b);
    // shade once so the kNN phase has its shading points
    if (rc == MNRT_OK) rc = run_phases(*b, *io, MNRT_PHASE_TRACE | MNRT_PHASE_SHADE, st);
    auto search = [&](uint32_t phase, uint32_t mnrt_batch_tuning::*field, std::vector<uint32_t> cands) {
        float best_ms = 3.4e38f;
        uint32_t best_v = best.*field;
        for (uint32_t v : cands) {
            if (rc != MNRT_OK) return;
            mnrt_batch_tuning t = best;
            t.*field = v;
            if (field == &mnrt_batch_tuning::trace_block && t.trace_tile % v != 0) t.trace_tile = v * 4;
            if (!valid_tuning(*b, t)) continue;
            b->tun = t;
            rc = upload_tables(*b);
            float ms = 0.0f;
            if (rc == MNRT_OK) rc = time_phases(*b, *io, phase, st, &ms);
            if (rc == MNRT_OK && ms < best_ms) {
                best_ms = ms;
                best_v = v;
                best = t;
            }
        }
        best.*field = best_v;
    };
    const uint32_t C = b->cus;
    search(MNRT_PHASE_TRACE, &mnrt_batch_tuning::trace_stack, {16, 24, 32, (uint32_t)kTraceStackAbi});
    search(MNRT_PHASE_TRACE, &mnrt_batch_tuning::trace_block, {64, 128, 256});
    search(MNRT_PHASE_TRACE, &mnrt_batch_tuning::trace_tile,
           {best.trace_block, 2 * best.trace_block, 4 * best.trace_block, 16 * best.trace_block});
    search(MNRT_PHASE_TRACE, &mnrt_batch_tuning::trace_grid, {2 * C, 4 * C, 8 * C, 16 * C, 32 * C});
    if (b->has_knn) {
        const uint32_t c0 = min_cap(b->k);
        search(MNRT_PHASE_KNN, &mnrt_batch_tuning::knn_select, {0, 1, 2});
        search(MNRT_PHASE_KNN, &mnrt_batch_tuning::knn_cap, {c0, 2 * c0});
        search(MNRT_PHASE_KNN, &mnrt_batch_tuning::knn_waves, {1, 2, 4});
        search(MNRT_PHASE_KNN, &mnrt_batch_tuning::knn_chunk, {1, 4, 16, 64});
        search(MNRT_PHASE_KNN, &mnrt_batch_tuning::knn_grid, {2 * C, 4 * C, 8 * C, 16 * C});
    }
    search(b->has_knn ? (MNRT_PHASE_TRACE | MNRT_PHASE_KNN) : MNRT_PHASE_TRACE, &mnrt_batch_tuning::parts,
           {1, kMaxParts});
    b->tun = best;
    b->limit_slots = 0;
    int rc2 = upload_tables(*b);
    return rc != MNRT_OK ? rc : rc2;
}
