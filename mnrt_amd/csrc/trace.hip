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
#include "mnrt_common.h"

namespace mnrt {

struct RootBox {
    float lo[3], hi[3];
};

constexpr int kTraceStack = MNRT_MAX_LEVEL + 2;

__global__ __launch_bounds__(256) void k_trace_closest(
    const uint2* __restrict__ nodes, const float4* __restrict__ recs, RootBox root, uint32_t n,
    const float4* __restrict__ orig, const float4* __restrict__ dir, float eps,
    int32_t* __restrict__ out_tri, float* __restrict__ out_t, float2* __restrict__ out_bary,
    float4* __restrict__ out_pt, uint32_t* __restrict__ out_valid) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f3 o = xyz(orig[i]);
    f3 d = xyz(dir[i]);
    f3 inv = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    float t0, t1;
    int32_t best_tri = -1;
    float best_t = 0.0f, bu = 0.0f, bv = 0.0f;
    if (ray_box(root.lo, root.hi, o, inv, eps, kInf, t0, t1)) {
        t1 = t1 + eps;
        best_t = t1;
        uint32_t snode[kTraceStack];
        float stmin[kTraceStack], stmax[kTraceStack];
        int sp = 0;
        uint32_t node = 0;
        float tmin = t0, tmax = t1;
        for (;;) {
            uint2 nd = nodes[node];
            while (!is_leaf(nd.y)) {
                uint32_t a = nd.y >> 30;
                float s = __uint_as_float(nd.x);
                uint32_t left = node + 1, right = nd.y & MNRT_INDEX_MASK;
                float oa = comp(o, a), da = comp(d, a), ia = comp(inv, a);
                float tsplit = (s - oa) * ia;
                bool below = (oa < s) || (oa == s && da > 0.0f);
                uint32_t first = below ? left : right, second = below ? right : left;
                if (tsplit != tsplit) {
                    if (sp < kTraceStack) {
                        snode[sp] = second;
                        stmin[sp] = tmin;
                        stmax[sp] = tmax;
                        ++sp;
                    }
                    node = first;
                } else if (tsplit < 0.0f || tsplit > tmax * kTravTolHi) {
                    node = first;
                } else if (tsplit < tmin * kTravTolLo) {
                    node = second;
                } else {
                    if (sp < kTraceStack) {
                        snode[sp] = second;
                        stmin[sp] = fmaxf(tmin, tsplit * kTravTolLo);
                        stmax[sp] = tmax;
                        ++sp;
                    }
                    tmax = fminf(tmax, tsplit * kTravTolHi);
                    node = first;
                }
                nd = nodes[node];
            }
            uint32_t count = nd.x, begin = nd.y & MNRT_INDEX_MASK;
            for (uint32_t r = 0; r < count; ++r) {
                const float4* rec = recs + 3ull * (begin + r);
                float4 a4 = rec[0];
                float4 e14 = rec[1];
                float4 e24 = rec[2];
                float lam, u, v;
                if (ray_tri(xyz(a4), xyz(e14), xyz(e24), o, d, lam, u, v) && lam > eps) {
                    int32_t tri = (int32_t)__float_as_uint(a4.w);
                    if (lam < best_t || (lam == best_t && best_tri >= 0 && tri < best_tri)) {
                        best_t = lam;
                        best_tri = tri;
                        bu = u;
                        bv = v;
                    }
                }
            }
            // Next stacked node that may hold a hit no farther than the best one.
            while (sp > 0 && best_tri >= 0 && best_t < stmin[sp - 1]) --sp;
            if (sp == 0) break;
            --sp;
            node = snode[sp];
            tmin = stmin[sp];
            tmax = stmax[sp];
        }
    }
    float t = (best_tri >= 0) ? best_t : 0.0f;
    if (best_tri < 0) bu = bv = 0.0f;
    out_tri[i] = best_tri;
    if (out_t) out_t[i] = t;
    if (out_bary) out_bary[i] = make_float2(bu, bv);
    if (out_pt) out_pt[i] = make_float4(o.x + d.x * t, o.y + d.y * t, o.z + d.z * t, 0.0f);
    if (out_valid) out_valid[i] = best_tri >= 0 ? 1u : 0u;
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
