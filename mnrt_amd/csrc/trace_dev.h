// trace_dev.h — closest-hit traversal of the packed preorder kd-tree (device function),
// shared by the ray kernel (trace.hip) and the photon tracer (photon.hip).
//
// dev_FindNextIntersectionKDWhileWhile<true> (raytracing.cu:503-645) restated for the
// SoA node words and leaf-ordered triangle records:
//   * interval decisions are widened by a relative 2^-16 (split parameter NaN for a ray
//     lying in the plane: both children; tsplit < 0 or beyond tmax: near child; before
//     tmin: far child; else both), and equal lambdas resolve to the lower triangle
//     index, so the hit equals the brute-force closest hit whatever the tree;
//   * a popped node is skipped when the best hit lies before its tmin (the reference
//     ends the search at the first leaf whose tmax exceeds the best hit, which misses
//     hits in a sibling pushed for a ray lying in the split plane);
//   * the stack is bounded by the build's level cap.
#pragma once

#include "mnrt_common.h"

namespace mnrt {

struct RootBox {
    float lo[3], hi[3];
};

struct Hit {
    int32_t tri;
    float t, u, v;
};

constexpr int kTraceStack = MNRT_MAX_LEVEL + 2;

// kStack: traversal stack entries.  Entries on the stack are far children of distinct
// levels of the current path, so a tree of max_level L never needs more than L; a
// smaller stack (fewer registers / scratch bytes) is exact for trees with
// max_level <= kStack.
template <int kStack = kTraceStack>
__device__ __forceinline__ Hit trace_ray(const uint2* __restrict__ nodes, const float4* __restrict__ recs,
                                         const RootBox& root, f3 o, f3 d, float eps) {
    f3 inv = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    float t0, t1;
    int32_t best_tri = -1;
    float best_t = 0.0f, bu = 0.0f, bv = 0.0f;
    if (ray_box(root.lo, root.hi, o, inv, eps, kInf, t0, t1)) {
        t1 = t1 + eps;
        best_t = t1;
        uint32_t snode[kStack];
        float stmin[kStack], stmax[kStack];
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
                    if (sp < kStack) {
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
                    if (sp < kStack) {
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
            while (sp > 0 && best_tri >= 0 && best_t < stmin[sp - 1]) --sp;
            if (sp == 0) break;
            --sp;
            node = snode[sp];
            tmin = stmin[sp];
            tmax = stmax[sp];
        }
    }
    Hit h;
    h.tri = best_tri;
    h.t = (best_tri >= 0) ? best_t : 0.0f;
    h.u = (best_tri >= 0) ? bu : 0.0f;
    h.v = (best_tri >= 0) ? bv : 0.0f;
    return h;
}

}  // namespace mnrt
