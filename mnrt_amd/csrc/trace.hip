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
//     local-memory arrays with no overflow check.
#include "trace_dev.h"

namespace mnrt {

__global__ __launch_bounds__(256) void k_trace_closest(
    const uint2* __restrict__ nodes, const float4* __restrict__ recs, RootBox root, uint32_t n,
    const float4* __restrict__ orig, const float4* __restrict__ dir, float eps,
    int32_t* __restrict__ out_tri, float* __restrict__ out_t, float2* __restrict__ out_bary,
    float4* __restrict__ out_pt, uint32_t* __restrict__ out_valid) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f3 o = xyz(orig[i]);
    f3 d = xyz(dir[i]);
    Hit h = trace_ray(nodes, recs, root, o, d, eps);
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
    hipLaunchKernelGGL(mnrt::k_trace_closest, grid, block, 0, (hipStream_t)stream,
                       (const uint2*)tree->d_nodes, (const float4*)tree->d_recs, rb, n,
                       (const float4*)d_orig4, (const float4*)d_dir4, eps, d_out_tri, d_out_t,
                       (float2*)d_out_bary2, (float4*)d_out_pt4, d_out_valid);
    MNRT_HIP_CHECK(hipGetLastError());
    return MNRT_OK;
}
