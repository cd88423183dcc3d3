// mnrt_common.h — device helpers shared by the gfx950 kernels of the kd-tree hot path.
//
// All fp32 arithmetic here is written in a fixed operation order and compiled with
// -ffp-contract=off and IEEE division / square root, so a result is bit-identical to any
// other implementation that performs the same operations in the same order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mnrt_kd.h"

namespace mnrt {

constexpr float kInf = 3.402823466e+38f;              // FLT_MAX, MN_INFINITY (MNMath.h:60)
constexpr float kPi = 3.14159265358979323846f;        // MN_PI (MNMath.h:53)
constexpr float kInvPi = 0.31830988618379067154f;     // MN_INV_PI (MNMath.h:55)
constexpr float kCostTraverse = 3.0f;                 // KD_COST_TRAVERSE (kdtree.cu:51)
// Relative widening of traversal interval decisions (2^-16).
constexpr float kTravTolLo = 1.0f - 1.0f / 65536.0f;
constexpr float kTravTolHi = 1.0f + 1.0f / 65536.0f;

struct f3 {
    float x, y, z;
};

// An element count that the host knows (p null: v), or that a kernel reads from device
// memory (min(p[0] << sh, v): v is then the capacity the buffers were sized for, so a
// count that outgrew them is clamped rather than overrun).  Device counts let a recorded
// launch sequence work on whatever sizes the kernels before it produced.
struct DN {
    uint32_t v;
    const uint32_t* p;
    uint32_t sh;
    __device__ __forceinline__ uint32_t get() const { return p ? min(p[0] << sh, v) : v; }
};
inline DN hn(uint32_t v) { return DN{v, nullptr, 0u}; }
inline DN dn(const uint32_t* p, uint32_t cap, uint32_t sh = 0u) { return DN{cap, p, sh}; }

__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 sub3(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float comp(f3 v, uint32_t a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
__device__ __forceinline__ f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }

__device__ __forceinline__ bool is_leaf(uint32_t w1) { return (w1 >> 30) == MNRT_LEAF_TAG; }

// Moller-Trumbore with precomputed edges (dev_RayTriIntersect, intersect_dev.h:77-99):
// det == 0 misses, barycentric edges inclusive, lambda may be negative.
__device__ __forceinline__ bool ray_tri(f3 a, f3 e1, f3 e2, f3 o, f3 d, float& lam, float& u,
                                        float& v) {
    f3 pvec = cross3(d, e2);
    float det = dot3(e1, pvec);
    if (det == 0.0f) return false;
    float inv_det = 1.0f / det;
    f3 tvec = sub3(o, a);
    u = dot3(tvec, pvec) * inv_det;
    f3 qvec = cross3(tvec, e1);
    v = dot3(d, qvec) * inv_det;
    lam = dot3(e2, qvec) * inv_det;
    return u >= 0.0f && v >= 0.0f && (u + v) <= 1.0f;
}

// Slab test (dev_RayBoxIntersect, intersect_dev.h:124-164), no early out.
__device__ __forceinline__ bool ray_box(const float* bmin, const float* bmax, f3 o, f3 inv,
                                        float t0, float t1, float& tn, float& tf) {
    bool hit = true;
    float oo[3] = {o.x, o.y, o.z}, ii[3] = {inv.x, inv.y, inv.z};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float v1 = (bmin[i] - oo[i]) * ii[i];
        float v2 = (bmax[i] - oo[i]) * ii[i];
        float tnear = v1, tfar = v2;
        if (v1 > v2) tnear = v2;
        if (v1 > v2) tfar = v1;
        t0 = (tnear > t0) ? tnear : t0;
        t1 = (tfar < t1) ? tfar : t1;
        if (t0 > t1) hit = false;
    }
    tn = t0;
    tf = t1;
    return hit;
}

__device__ __forceinline__ float sgn1(float x) { return x >= 0.0f ? 1.0f : -1.0f; }

// Octahedral direction code: two 16-bit signed coordinates.  Decoding does not
// normalise; callers only use the sign of a dot product.
__device__ __forceinline__ uint32_t oct_encode(float x, float y, float z) {
    float s = fabsf(x) + fabsf(y) + fabsf(z);
    float u = x / s, v = y / s;
    if (z < 0.0f) {
        float ou = u, ov = v;
        u = (1.0f - fabsf(ov)) * sgn1(ou);
        v = (1.0f - fabsf(ou)) * sgn1(ov);
    }
    float qu = rintf(u * 32767.0f), qv = rintf(v * 32767.0f);
    qu = fminf(32767.0f, fmaxf(-32767.0f, qu));
    qv = fminf(32767.0f, fmaxf(-32767.0f, qv));
    int32_t iu = (int32_t)qu, iv = (int32_t)qv;
    return ((uint32_t)(uint16_t)(int16_t)iu) | (((uint32_t)(uint16_t)(int16_t)iv) << 16);
}

// The fp32 quotient q / 32767 of an integer |q| <= 32767 lies at least 2^-39 (relative) from
// any fp32 rounding boundary (its reduced denominator divides the odd 32767, a boundary's
// is a power of two), so the fp64 product with the rounded reciprocal (relative error
// below 2^-52), rounded to fp32, is the correctly rounded quotient: the same value as the
// IEEE division, in three instructions (checked for all 65535 codes on the host).
__device__ __forceinline__ float div32767(int32_t q) { return (float)((double)q * (1.0 / 32767.0)); }

__device__ __forceinline__ f3 oct_decode(uint32_t c) {
    int16_t iu = (int16_t)(uint16_t)(c & 0xFFFFu), iv = (int16_t)(uint16_t)(c >> 16);
    float u = div32767(iu), v = div32767(iv);
    float z = (1.0f - fabsf(u)) - fabsf(v);
    if (z < 0.0f) {
        float ou = u;
        u = (1.0f - fabsf(v)) * sgn1(ou);
        v = (1.0f - fabsf(ou)) * sgn1(v);
    }
    return f3{u, v, z};
}

// Wave-local ordering of LDS accesses (all lanes of one wave, no workgroup barrier).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave-wide reductions; every lane of the wave must be active.  DPP moves within rows of
// 16 lanes (swap in pairs, swap pairs in quads, half-row mirror, row mirror) leave the row's
// reduction in each of its lanes, and the four rows (lanes 0, 16, 32, 48) are combined by
// readlane, so the result is wave-uniform.  For associative, commutative operations (max,
// min, integer sums, lexicographic (cost, index) minimum) the result does not depend on
// the combining order.
constexpr int kDppPairSwap = 0xB1;    // quad_perm [1, 0, 3, 2]
constexpr int kDppQuadSwap = 0x4E;    // quad_perm [2, 3, 0, 1]
constexpr int kDppHalfMirror = 0x141;
constexpr int kDppRowMirror = 0x140;

template <int kCtrl>
__device__ __forceinline__ uint32_t dpp_mov(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kCtrl, 0xF, 0xF, false);
}

__device__ __forceinline__ uint32_t row_lane(uint32_t v, int row) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 16 * row);
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    v = max(v, dpp_mov<kDppPairSwap>(v));
    v = max(v, dpp_mov<kDppQuadSwap>(v));
    v = max(v, dpp_mov<kDppHalfMirror>(v));
    v = max(v, dpp_mov<kDppRowMirror>(v));
    return max(max(row_lane(v, 0), row_lane(v, 1)), max(row_lane(v, 2), row_lane(v, 3)));
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    v = min(v, dpp_mov<kDppPairSwap>(v));
    v = min(v, dpp_mov<kDppQuadSwap>(v));
    v = min(v, dpp_mov<kDppHalfMirror>(v));
    v = min(v, dpp_mov<kDppRowMirror>(v));
    return min(min(row_lane(v, 0), row_lane(v, 1)), min(row_lane(v, 2), row_lane(v, 3)));
}

// Wrapping 32-bit sum (callers keep the true sum below 2^32).
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
    v += dpp_mov<kDppPairSwap>(v);
    v += dpp_mov<kDppQuadSwap>(v);
    v += dpp_mov<kDppHalfMirror>(v);
    v += dpp_mov<kDppRowMirror>(v);
    return (row_lane(v, 0) + row_lane(v, 1)) + (row_lane(v, 2) + row_lane(v, 3));
}

template <int kCtrl>
__device__ __forceinline__ long long dpp_mov64(long long v) {
    const uint32_t lo = dpp_mov<kCtrl>((uint32_t)(unsigned long long)v);
    const uint32_t hi = dpp_mov<kCtrl>((uint32_t)((unsigned long long)v >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ long long wave_sum_i64(long long v) {
    v += dpp_mov64<kDppPairSwap>(v);
    v += dpp_mov64<kDppQuadSwap>(v);
    v += dpp_mov64<kDppHalfMirror>(v);
    v += dpp_mov64<kDppRowMirror>(v);
    long long t = 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t lo = row_lane((uint32_t)(unsigned long long)v, r);
        const uint32_t hi = row_lane((uint32_t)((unsigned long long)v >> 32), r);
        t += (long long)(((unsigned long long)hi << 32) | lo);
    }
    return t;
}

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Load of read-only data at a wave-uniform address through the constant address space, so
// it becomes a scalar load (scalar cache, result in SGPRs).  Only for data no kernel of the
// launch writes (tree node words, descriptor tables).
template <typename T>
__device__ __forceinline__ T ld_uniform(const T* p, uint32_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
    const __attribute__((address_space(4))) T* c = (const __attribute__((address_space(4))) T*)p;
    return c[i];
#else
    return p[i];
#endif
}
// ld_uniform of an 8-byte node word pair at a 32-bit byte offset: one scalar load with an
// SGPR offset, no 64-bit address arithmetic.  Valid for i < kUniform8MaxIndex.
constexpr uint32_t kUniform8MaxIndex = 1u << 29;
__device__ __forceinline__ uint2 ld_uniform8(const uint2* p, uint32_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
    const __attribute__((address_space(4))) char* c = (const __attribute__((address_space(4))) char*)p;
    return *(const __attribute__((address_space(4))) uint2*)(c + (i << 3));
#else
    return p[i];
#endif
}
__device__ __forceinline__ float uniformf(float v) {
    return __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(v)));
}

// Mask of the lanes where p holds.  The builtin takes the predicate as a bool, so a
// comparison's lane mask is used as it is (HIP's int-argument __ballot first turns it into
// a 0/1 vector value and compares that against 0: two more vector instructions).
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Lanes below this one whose predicate is set (exclusive prefix of a 64-bit ballot).
__device__ __forceinline__ uint32_t mask_prefix(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
// base + mask_prefix(m), the addition folded into the count instructions.
__device__ __forceinline__ uint32_t mask_prefix_from(uint64_t m, uint32_t base) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, base));
}

// Query-time leaves of photon trees: subtrees with at most this many references (the kNN
// kernel tests a leaf 64 photons per step; measured on 1M photons, k = 100, with the
// warm-start chains of eight queries: 256 -> 512 saves 5-9 %, 1024 is slower again).
constexpr uint32_t kQueryLeafRefs = 512;

// Device memory of the library: a stream-ordered pool per device that keeps freed blocks
// for reuse (release threshold: never), so repeated builds and queries do not map and
// unmap pages.  Tree arrays come from it too; mnrt_tree_release returns them on the
// null stream.
hipError_t pool_alloc(void** p, size_t bytes, hipStream_t st);
hipError_t pool_free(void* p, hipStream_t st);
// Pinned host staging buffer of at least `bytes`, one per host thread and slot, kept for
// the thread's lifetime.
void* host_staging(int slot, size_t bytes);
// Device address of the fixed kNN scratch array on `device` (which must be current),
// looked up once per device.
int knn_scratch_init(int device);

}  // namespace mnrt

// Status plumbing shared by the C-ABI translation units.
#define MNRT_HIP_CHECK(expr)                       \
    do {                                           \
        hipError_t _e = (expr);                    \
        if (_e != hipSuccess) return MNRT_ERR_HIP; \
    } while (0)
