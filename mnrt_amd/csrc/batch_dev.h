// batch_dev.h — descriptors of the batched (many-scene) launches shared by the kernels
// (trace.hip, knn.hip, shade.hip) and the batch object (batch.hip).
//
// A batch concatenates the per-scene arrays: scene s owns ray / query slots
// slot_begin[s] .. slot_begin[s+1]-1 of every per-ray and per-query array, and its
// triangles are entries tri_begin[s] .. of the vertex and albedo arrays.  Work is cut
// into tiles of consecutive slots that never straddle two scenes.
#pragma once

#include "mnrt_common.h"

namespace mnrt {

// Segment of entry g in an offset table (last s < n with off[s] <= g; empty segments
// are skipped because the segment that holds g is the last one starting at or below g).
__device__ __forceinline__ uint32_t seg_of(const uint32_t* off, uint32_t n, uint32_t g) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

constexpr uint32_t kMaxParts = 8;  // work-counter shares (one per XCD)
constexpr int kTraceStackAbi = MNRT_MAX_LEVEL + 2;  // the full-depth traversal stack

struct TraceScene {
    const uint2* nodes;
    const float4* recs;
    float lo[3], hi[3];
    float eps;
    uint32_t ray_begin, ray_end;
};

struct TraceBatchArgs {
    const TraceScene* scenes;
    const uint32_t* tile_begin;  // num_scenes + 1
    uint32_t num_scenes;
    uint32_t tile_rays;
    uint32_t num_parts;
    const uint32_t* part_begin;  // first tile of each share
    const uint32_t* part_end;
    uint32_t* work;              // num_parts counters, cleared by the launch
    const float4* orig;
    const float4* dir;
    int32_t* out_tri;
    float* out_t;
    float2* out_bary;
};

// Arguments of one kNN launch (plain: one tree; batched: the per-slot arrays, the tree
// fields coming from the slot's scene).
struct KnnArgs {
    const uint2* nodes;
    const float4* recs;  // 2 float4 per reference
    uint32_t k;
    float r2max;
    float r_max;
    uint32_t n;
    const float4* pos;
    const float4* nrm;
    const float4* dir;
    const float4* infl;
    const float4* clr;
    float4* rad_out;      // accumulate target (slot = query), or per-query temp when idx_sp
    uint32_t accumulate;  // 1: rad_out[q] += rad, 0: rad_out[q] = rad
    float* out_r2;
    float4* out_irr;
    uint32_t* out_nbr;
    uint32_t* out_count;
    float* out_radius;
    const float* r_query;  // per-query radius cap (null: r_max)
    const float4* orig;    // source ray origins (inverse-distance sums)
    float* inv_dist;       // 1 / |hit - origin| per query (null: not wanted)
    uint32_t inv_accumulate;
    // backward
    const float4* grad;            // dLoss/dradiance per query
    unsigned long long* acc;       // 3 fixed-point accumulators per photon (original index)
    double fix_scale;              // accumulator units per unit gradient (a power of two)
    float4* out_grad_clr;          // dLoss/d(diffuse colour) per query
    const uint32_t* order;         // processing order of the queries (null: 0..n-1)
    const uint32_t* order_far;     // the order's far-pair count (order_coherent), or null
};

struct KnnScene {
    const uint2* nodes;  // query-time node words
    const float4* recs;
    float r2max, r_max;
    uint32_t slot_begin, slot_end;
};

struct ShadeBatchArgs {
    uint32_t n;                   // total ray slots
    uint32_t num_scenes;
    const uint32_t* slot_begin;   // num_scenes + 1
    const uint32_t* tri_begin;    // num_scenes
    const float4* orig;
    const float4* dir;
    const int32_t* tri;
    const float* t;
    const float4* v0;
    const float4* v1;
    const float4* v2;
    const float4* albedo;
    const float4* infl;
    const uint32_t* pos;          // exclusive scan of the hit flags over all slots
    float4* out_pos;
    float4* out_nrm;
    float4* out_dir;
    float4* out_clr;
    float4* out_infl;
    uint32_t* out_src;            // source ray slot
    uint32_t* out_count;          // hits per scene
};

int launch_trace_batch(const TraceBatchArgs& B, int stack, uint32_t grid, uint32_t block, hipStream_t st);
int launch_knn_batch(const KnnArgs& G, const KnnScene* scenes, const uint32_t* slot_begin, uint32_t ns,
                     const uint32_t* count, uint32_t chunk, uint32_t num_parts, const uint32_t* part_begin,
                     const uint32_t* part_end, uint32_t* work, uint32_t cap, uint32_t waves, uint32_t sel,
                     uint32_t grid, hipStream_t st);
int launch_knn_bwd_batch(const KnnArgs& G, const KnnScene* scenes, const uint32_t* slot_begin, uint32_t ns,
                         const uint32_t* count, const uint32_t* photon_begin, uint32_t total_photons,
                         const float* fwd_r2, const uint32_t* fwd_count, uint32_t* wmax, double* scale,
                         uint32_t num_parts, const uint32_t* part_begin, const uint32_t* part_end, uint32_t* work,
                         uint32_t grid, float* grad_pow4, hipStream_t st);
int launch_shade_batch(const ShadeBatchArgs& A, uint32_t* flags, uint32_t* scan_tmp, hipStream_t st);
int launch_records_batch(int kind, const uint32_t* refs, DN nrefs, const uint32_t* rbase, const uint32_t* ebase,
                         uint32_t ns, const float4* a, const float4* b, const float4* c, float4* recs,
                         const uint32_t* err, hipStream_t st);
int launch_collapse_batch(const uint2* nodes, DN nnodes, const uint32_t* nbase, uint32_t ns, uint32_t max_refs,
                          uint2* qnodes, const uint32_t* err, hipStream_t st);

}  // namespace mnrt

namespace {

// Clears counters inside a launch sequence (a fill kernel rather than a memset node, so
// the sequence can be captured in a graph and replayed).
__global__ __launch_bounds__(64) __attribute__((unused)) void k_fill_u32(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

}  // namespace
