// trace.hip — closest-hit ray traversal of the packed preorder kd-tree on gfx950.
//
// Replaces kernel_FindIntersections + dev_FindNextIntersectionKDWhileWhile<true>
// (raytracing.cu:503-645, 854-879) and its wrapper KernelRTTraceRays
// (raytracing.cu:1617-1635).  One lane per ray, 256-lane workgroups.  Differences from
// the reference, all deliberate:
//   * the tree and triangles are kernel arguments (no __constant__ / texture binding),
//     so several trees can be traced at once;
//   * leaves address contiguous leaf-ordered triangle records (v0, e1, e2: 48 B), not
//     texture fetches of three vertices through an index list;
//   * interval decisions are widened by a relative 2^-16 and equal lambdas resolve to
//     the lower triangle index, so the hit equals the brute-force closest hit over all
//     triangles whatever the tree;
//   * the traversal stack is bounded by the build's level cap instead of KD_MAX_HEIGHT
//     local-memory arrays with no overflow check, and its top entries live in LDS.
#include "batch_dev.h"
#include "trace_dev.h"

namespace mnrt {

__global__ __launch_bounds__(256) void k_trace_closest(
    const uint2* __restrict__ nodes, const float4* __restrict__ recs, RootBox root, uint32_t n,
    const float4* __restrict__ orig, const float4* __restrict__ dir, float eps,
    int32_t* __restrict__ out_tri, float* __restrict__ out_t, float2* __restrict__ out_bary,
    float4* __restrict__ out_pt, uint32_t* __restrict__ out_valid) {
    extern __shared__ uint32_t s_stack[];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f3 o = xyz(orig[i]);
    f3 d = xyz(dir[i]);
    Hit h = trace_ray<kTraceStack, kTraceLds>(nodes, recs, root, o, d, eps, s_stack + threadIdx.x, blockDim.x);
    out_tri[i] = h.tri;
    if (out_t) out_t[i] = h.t;
    if (out_bary) out_bary[i] = make_float2(h.u, h.v);
    if (out_pt) out_pt[i] = make_float4(o.x + d.x * h.t, o.y + d.y * h.t, o.z + d.z * h.t, 0.0f);
    if (out_valid) out_valid[i] = h.tri >= 0 ? 1u : 0u;
}

}  // namespace mnrt

extern "C" int mnrt_trace_closest(const mnrt_tree* tree, uint32_t n, const float* d_orig4,
                                  const float* d_dir4, float eps, int32_t* d_out_tri,
                                  float* d_out_t, float* d_out_bary2, float* d_out_pt4,
                                  uint32_t* d_out_valid, void* stream) {
    if (!tree || (n > 0 && (!d_out_tri || !d_orig4 || !d_dir4))) return MNRT_ERR_ARG;
    if (tree->kind != MNRT_KIND_TRI || !tree->d_nodes || !tree->d_recs || tree->num_nodes == 0)
        return MNRT_ERR_STATE;
    if (tree->max_level > MNRT_MAX_LEVEL) return MNRT_ERR_CAPACITY;
    if (n == 0) return MNRT_OK;
    mnrt::RootBox rb;
    for (int a = 0; a < 3; ++a) {
        rb.lo[a] = tree->root_min[a];
        rb.hi[a] = tree->root_max[a];
    }
    dim3 block(256), grid((n + 255) / 256);
    hipLaunchKernelGGL(mnrt::k_trace_closest, grid, block, mnrt::trace_lds_bytes(mnrt::kTraceLds, 256),
                       (hipStream_t)stream,
                       (const uint2*)tree->d_nodes, (const float4*)tree->d_recs, rb, n,
                       (const float4*)d_orig4, (const float4*)d_dir4, eps, d_out_tri, d_out_t,
                       (float2*)d_out_bary2, (float4*)d_out_pt4, d_out_valid);
    MNRT_HIP_CHECK(hipGetLastError());
    return MNRT_OK;
}

// ---- batches: many trees, one persistent launch ----------------------------------------

namespace mnrt {

// Workgroups take tiles of consecutive rays from a work counter until every tile is
// done; tile t belongs to the scene whose tile range holds it.  Each XCD (workgroup id
// mod 8 on MI355X) has its own counter over a contiguous share of the tiles, so an
// XCD's L2 holds the trees of the scenes it traces; a workgroup whose share is done
// takes tiles from the next XCD's share.
template <int kStack>
__global__ __launch_bounds__(256) void k_trace_batch(TraceBatchArgs B) {
    __shared__ uint32_t s_tile;
    extern __shared__ uint32_t s_stack[];
    const uint32_t xcd = blockIdx.x % B.num_parts;
    for (uint32_t part = 0; part < B.num_parts; ++part) {
        const uint32_t p = (xcd + part) % B.num_parts;
        const uint32_t end = B.part_end[p];
        for (;;) {
            if (threadIdx.x == 0) s_tile = B.part_begin[p] + atomicAdd(&B.work[p], 1u);
            __syncthreads();
            const uint32_t t = s_tile;
            __syncthreads();
            if (t >= end) break;
            const uint32_t s = seg_of(B.tile_begin, B.num_scenes, t);
            const TraceScene& d = B.scenes[s];
            const uint32_t first = d.ray_begin + (t - B.tile_begin[s]) * B.tile_rays;
            const uint32_t last = min(first + B.tile_rays, d.ray_end);
            RootBox root;
            for (int a = 0; a < 3; ++a) {
                root.lo[a] = d.lo[a];
                root.hi[a] = d.hi[a];
            }
            for (uint32_t i = first + threadIdx.x; i < last; i += blockDim.x) {
                f3 o = xyz(B.orig[i]);
                f3 dd = xyz(B.dir[i]);
                Hit h = trace_ray<kStack, kTraceLds>(d.nodes, d.recs, root, o, dd, d.eps, s_stack + threadIdx.x,
                                                     blockDim.x);
                B.out_tri[i] = h.tri;
                if (B.out_t) B.out_t[i] = h.t;
                if (B.out_bary) B.out_bary[i] = make_float2(h.u, h.v);
            }
        }
    }
}

int launch_trace_batch(const TraceBatchArgs& B, int stack, uint32_t grid, uint32_t block, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, B.work, B.num_parts, 0u);
    const uint32_t lds = trace_lds_bytes(kTraceLds, block);
    switch (stack) {
        case 16: hipLaunchKernelGGL(k_trace_batch<16>, dim3(grid), dim3(block), lds, st, B); break;
        case 24: hipLaunchKernelGGL(k_trace_batch<24>, dim3(grid), dim3(block), lds, st, B); break;
        case 32: hipLaunchKernelGGL(k_trace_batch<32>, dim3(grid), dim3(block), lds, st, B); break;
        case kTraceStack: hipLaunchKernelGGL(k_trace_batch<kTraceStack>, dim3(grid), dim3(block), lds, st, B); break;
        default: return MNRT_ERR_ARG;
    }
    MNRT_HIP_CHECK(hipGetLastError());
    return MNRT_OK;
}

}  // namespace mnrt
