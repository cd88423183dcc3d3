// trace_dev.h — closest-hit and any-hit traversal of the packed preorder kd-tree (device
// functions), shared by the ray kernels (trace.hip, shadow.hip) and the photon tracer
// (photon.hip).
//
// dev_FindNextIntersectionKDWhileWhile (raytracing.cu:503-645) restated for the SoA node
// words and leaf-ordered triangle records:
//   * interval decisions are widened by a relative 2^-16 (split parameter NaN for a ray
//     lying in the plane: both children; tsplit < 0 or beyond tmax: near child; before
//     tmin: far child; else both), and equal lambdas resolve to the lower triangle
//     index, so the hit equals the brute-force closest hit whatever the tree;
//   * a popped node is skipped when the best hit lies before its tmin (the reference
//     ends the search at the first leaf whose tmax exceeds the best hit, which misses
//     hits in a sibling pushed for a ray lying in the split plane);
//   * the stack is bounded by the build's level cap; its top kLds entries per lane live
//     in LDS, deeper ones in the lane's private array (the reference keeps the whole
//     stack in local memory).
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

// LDS entries per lane of the traversal stack.  On the 1M-triangle workload about 0.5 %
// of pushes land deeper than 8 entries and 5 % deeper than 6.
#ifndef MNRT_TRACE_LDS
#define MNRT_TRACE_LDS 6
#endif
constexpr int kTraceLds = MNRT_TRACE_LDS;

// LDS bytes a workgroup of `block` lanes needs for kLds stack entries per lane.
__host__ __device__ constexpr uint32_t trace_lds_bytes(int lds, uint32_t block) {
    return (uint32_t)lds * 3u * 4u * block;
}

// Per-lane traversal stack of kStack entries (node, tmin, tmax).  Entry e < kLds is in
// LDS at words (3e + c) * stride + lane (c = node, tmin, tmax: lanes of a wave touch
// consecutive banks); deeper entries in a private array.  kStack entries suffice for a
// tree of max level <= kStack: the entries are far children of distinct levels of the
// current path.
template <int kStack, int kLds>
struct LaneStack {
    static constexpr int kPriv = kStack > kLds ? kStack - kLds : 1;
    __attribute__((address_space(3))) uint32_t* lds;  // this lane's column: lds_base + lane
    uint32_t stride;
    uint32_t pnode[kPriv];
    float pmin[kPriv], pmax[kPriv];

    __device__ __forceinline__ void put(int e, uint32_t node, float a, float b) {
        if (kLds > 0 && e < kLds) {
            lds[(3 * e) * stride] = node;
            lds[(3 * e + 1) * stride] = __float_as_uint(a);
            lds[(3 * e + 2) * stride] = __float_as_uint(b);
        } else {
            pnode[e - kLds] = node;
            pmin[e - kLds] = a;
            pmax[e - kLds] = b;
        }
    }
    __device__ __forceinline__ float tmin_of(int e) const {
        if (kLds > 0 && e < kLds) return __uint_as_float(lds[(3 * e + 1) * stride]);
        return pmin[e - kLds];
    }
    __device__ __forceinline__ void get(int e, uint32_t& node, float& a, float& b) const {
        if (kLds > 0 && e < kLds) {
            node = lds[(3 * e) * stride];
            a = __uint_as_float(lds[(3 * e + 1) * stride]);
            b = __uint_as_float(lds[(3 * e + 2) * stride]);
        } else {
            node = pnode[e - kLds];
            a = pmin[e - kLds];
            b = pmax[e - kLds];
        }
    }
};

// Descends from `node` over [tmin, tmax] to a leaf, pushing far children; returns the
// leaf's node word.
template <int kStack, int kLds>
__device__ __forceinline__ uint2 descend(const uint2* __restrict__ nodes, LaneStack<kStack, kLds>& st, int& sp,
                                         uint32_t& node, float& tmin, float& tmax, f3 o, f3 d, f3 inv) {
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
            if (sp < kStack) st.put(sp++, second, tmin, tmax);
            node = first;
        } else if (tsplit < 0.0f || tsplit > tmax * kTravTolHi) {
            node = first;
        } else if (tsplit < tmin * kTravTolLo) {
            node = second;
        } else {
            if (sp < kStack) st.put(sp++, second, fmaxf(tmin, tsplit * kTravTolLo), tmax);
            tmax = fminf(tmax, tsplit * kTravTolHi);
            node = first;
        }
        nd = nodes[node];
    }
    return nd;
}

// Calls test(a, e1, e2) on the count triangle records from rec (3 float4 each) until it
// returns true; returns whether it did.  Two record buffers alternate, so the next record
// loads while the current one is tested and no registers are copied between tests; the
// loads are unconditional (past the leaf's end they re-read its last record, which is
// not tested again), so the wait before a test leaves the next record's loads in flight.
template <typename F>
__device__ __forceinline__ bool for_leaf_tris(const float4* rec, uint32_t count, F&& test) {
    if (count == 0) return false;
    float4 a0 = rec[0], b0 = rec[1], c0 = rec[2];
    const uint32_t last = count - 1u;
    for (uint32_t r = 0;; r += 2) {
        const float4* p1 = rec + 3u * min(r + 1u, last);
        const float4 a1 = p1[0], b1 = p1[1], c1 = p1[2];
        if (test(a0, b0, c0)) return true;
        if (r + 1 >= count) return false;
        const float4* p2 = rec + 3u * min(r + 2u, last);
        a0 = p2[0];
        b0 = p2[1];
        c0 = p2[2];
        if (test(a1, b1, c1)) return true;
        if (r + 2 >= count) return false;
    }
}

// Closest hit.  lds_lane: this lane's column of the workgroup's stack area (kLds > 0),
// stride = lanes in the workgroup.
template <int kStack = kTraceStack, int kLds = 0>
__device__ __forceinline__ Hit trace_ray(const uint2* __restrict__ nodes, const float4* __restrict__ recs,
                                         const RootBox& root, f3 o, f3 d, float eps,
                                         uint32_t* lds_lane = nullptr, uint32_t stride = 0) {
    f3 inv = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    float t0, t1;
    int32_t best_tri = -1;
    float best_t = 0.0f, bu = 0.0f, bv = 0.0f;
    if (ray_box(root.lo, root.hi, o, inv, eps, kInf, t0, t1)) {
        t1 = t1 + eps;
        best_t = t1;
        LaneStack<kStack, kLds> st;
        st.lds = (__attribute__((address_space(3))) uint32_t*)lds_lane;
        st.stride = stride;
        int sp = 0;
        uint32_t node = 0;
        float tmin = t0, tmax = t1;
        for (;;) {
            uint2 nd = descend(nodes, st, sp, node, tmin, tmax, o, d, inv);
            const uint32_t count = nd.x, begin = nd.y & MNRT_INDEX_MASK;
            for_leaf_tris(recs + 3ull * begin, count, [&](const float4& a4, const float4& e14, const float4& e24) {
                float lam, u, v;
                if (ray_tri(xyz(a4), xyz(e14), xyz(e24), o, d, lam, u, v) && lam > eps) {
                    const int32_t tri = (int32_t)__float_as_uint(a4.w);
                    if (lam < best_t || (lam == best_t && best_tri >= 0 && tri < best_tri)) {
                        best_t = lam;
                        best_tri = tri;
                        bu = u;
                        bv = v;
                    }
                }
                return false;
            });
            while (sp > 0 && best_tri >= 0 && best_t < st.tmin_of(sp - 1)) --sp;
            if (sp == 0) break;
            --sp;
            st.get(sp, node, tmin, tmax);
        }
    }
    Hit h;
    h.tri = best_tri;
    h.t = (best_tri >= 0) ? best_t : 0.0f;
    h.u = (best_tri >= 0) ? bu : 0.0f;
    h.v = (best_tri >= 0) ? bv : 0.0f;
    return h;
}

// Any-hit traversal (dev_FindNextIntersectionKDWhileWhile<false>, raytracing.cu:503-645,
// as the shadow-ray kernels use it, raytracing.cu:1019-1114).  Returns true iff some
// triangle is hit at lambda with
//     lambda > eps,  lambda < min(t_exit, dist) + eps,  lambda + slack < dist
// (t_exit: where the ray leaves the root box).  The reference tests only the first
// hit it meets against the light distance (`lambda >= distance`, `lambda + eps >=
// distance`), so which triangle decides depends on the order of a leaf's list; here
// any occluding triangle decides, so the result equals the brute-force predicate
// whatever the tree.  Same interval widening and stack as trace_ray; the search ends
// at the first occluder.
template <int kStack = kTraceStack, int kLds = 0>
__device__ __forceinline__ bool trace_any(const uint2* __restrict__ nodes, const float4* __restrict__ recs,
                                          const RootBox& root, f3 o, f3 d, float eps, float dist, float slack,
                                          uint32_t* lds_lane = nullptr, uint32_t stride = 0) {
    f3 inv = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    float t0, t1;
    if (!ray_box(root.lo, root.hi, o, inv, eps, kInf, t0, t1)) return false;
    const float tlim = fminf(t1, dist) + eps;
    LaneStack<kStack, kLds> st;
    st.lds = (__attribute__((address_space(3))) uint32_t*)lds_lane;
    st.stride = stride;
    int sp = 0;
    uint32_t node = 0;
    float tmin = t0, tmax = tlim;
    for (;;) {
        uint2 nd = descend(nodes, st, sp, node, tmin, tmax, o, d, inv);
        const uint32_t count = nd.x, begin = nd.y & MNRT_INDEX_MASK;
        if (for_leaf_tris(recs + 3ull * begin, count, [&](const float4& a4, const float4& e14, const float4& e24) {
                float lam, u, v;
                return ray_tri(xyz(a4), xyz(e14), xyz(e24), o, d, lam, u, v) && lam > eps && lam < tlim &&
                       lam + slack < dist;
            }))
            return true;
        if (sp == 0) return false;
        --sp;
        st.get(sp, node, tmin, tmax);
    }
}

}  // namespace mnrt
