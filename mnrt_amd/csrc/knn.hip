// knn.hip — exact k-nearest-photon radiance estimate on gfx950.
//
// Replaces the reference's query-radius estimate + range search + gather:
//   KDTreePoint::ComputeQueryRadii (KDTreePoint.cpp:229-248, kdtree_point.cu:229-546),
//   dev_RangeSearch (photon_knn.cu:113-235) and kernel_PhotonGather
//   (photon_knn.cu:487-552, wrapper KernelPMPhotonGather :833-844).
//
// Design (one wave64 per query):
//   * the point tree is walked by the whole wave with wave-uniform control flow (node
//     words are scalar loads) over the query-time node words, in which every subtree of
//     at most kQueryLeafRefs photons is one leaf; far children go onto a per-wave LDS stack with
//     their box distance (dev_RangeSearch pushes them by plane distance,
//     photon_knn.cu:146-171).  The fp32 box distance (the new axis' squared offset plus
//     the sum of the other two) can exceed a photon's fp32 distance only by a few
//     roundings, so pruning against the limit times (1 + 2^-20) drops only nodes none of
//     whose photons can be taken: exact;
//   * a leaf's photon records are contiguous, so its photons are tested one per lane
//     with one coalesced 16-byte buffer load each (bounded at the leaf's end);
//   * candidates (key = d2 bits << 32 | photon index, so integer order is (d2, index)
//     order) are appended to a per-wave LDS buffer with ballot + mbcnt compaction; when
//     k are known or the buffer would overflow, the k best are kept (radix selection of
//     the k-th key, or a bitonic sort) and the pruning threshold tightens;
//   * the final set is sorted only when the caller asks for the neighbour lists; each
//     lane fetches the records of its slots and the front-facing powers are summed per
//     channel as exact 64-bit fixed-point integers (the reference adds floats in its
//     range-search order, photon_knn.cu:506-516, so its sums depend on that order).
#include <algorithm>
#include <type_traits>
#include <cmath>
#include <cstring>
#include <mutex>

#include "batch_dev.h"
#include "order.h"

// Device scratch of mnrt_knn_gather_noalloc (the query processing order and, with slot
// indices, the per-query values before their runs are summed): 64 MiB per device, held
// for the process's lifetime.  Host code never passes this name to a kernel or a copy: it
// is a host-side stand-in; the device address comes from hipGetSymbolAddress, once per
// device (g_scratch_addr).
static __device__ __align__(256) uint32_t scratch[(64u << 20) / 4];

namespace mnrt {

constexpr int kWavesPerBlock = 4;
constexpr int kKnnStack = 64;
// relative margin of the node pruning limit over the photon limit (see knn_query)
constexpr float kPruneMargin = 1.0f + 0x1p-20f;
// word 3 of a raw (untyped, stride 0) buffer descriptor on gfx9-family parts: 32-bit data
// format; reads at or past num_records bytes return zero
constexpr int kBufferWord3 = 0x00020000;

enum : int { kModeRadii = 0, kModeRadiance = 1, kModeBackward = 2 };

// Sorts keys[0..size) ascending (positions follow their keys), size a power of two.
template <int CAP>
__device__ __forceinline__ void wave_bitonic(uint64_t* keys, uint32_t* pos, uint32_t size, uint32_t lane) {
    for (uint32_t k2 = 2; k2 <= size; k2 <<= 1) {
        for (uint32_t j = k2 >> 1; j > 0; j >>= 1) {
            for (uint32_t t = lane; t < (size >> 1); t += 64) {
                uint32_t i = 2 * j * (t / j) + (t % j);
                uint32_t l = i + j;
                uint64_t a = keys[i], b = keys[l];
                bool asc = (i & k2) == 0;
                if ((a > b) == asc) {
                    keys[i] = b;
                    keys[l] = a;
                    uint32_t pa = pos[i];
                    pos[i] = pos[l];
                    pos[l] = pa;
                }
            }
            wave_sync();
        }
    }
}

// x rounded to the nearest integer, ties to even (as __double2ll_rn), for |x| < 2^51: in
// [2^52, 2^53) doubles are spaced 1 apart, so adding 1.5 * 2^52 rounds x itself, and the
// sum's bits differ from the constant's by exactly that integer.
__device__ __forceinline__ long long round_to_i64(double x) {
    constexpr double kMagic = 0x1.8p52;
    return (long long)(__double_as_longlong(x + kMagic) - __double_as_longlong(kMagic));
}

// Exact sum of one float per lane slot (RK <= 4 slots per lane, unused slots 0) as a 64-bit
// fixed-point integer: with M the largest |v| (M = f 2^e, f in [0.5, 1)), every term is
// rounded to a multiple of 2^(e-40) (exact for terms down to 2^-17 M), the integers are
// added in any order (|sum| < 2^48) and the total is rounded to float once.  Each term
// q (|q| <= 2^40) is added as q = hi 2^24 + lo, lo in [0, 2^24): over at most 256 terms both
// parts' sums fit 32 bits, so the wave adds two 32-bit words instead of one 64-bit word.
// Finite inputs; wave-uniform result, the same in every lane.
template <int RK>
__device__ __forceinline__ float wave_fixed_sum(const float (&v)[RK]) {
    static_assert(RK <= 4, "32-bit part sums hold at most 256 terms");
    uint32_t mb = 0;  // max |v| as bits: non-negative floats order as their bits
#pragma unroll
    for (int t = 0; t < RK; ++t) mb = max(mb, __float_as_uint(v[t]) & 0x7FFFFFFFu);
    const float m = __uint_as_float(wave_max_u32(mb));
    if (!(m > 0.0f)) return 0.0f;
    int ex;
    (void)frexpf(m, &ex);
    const int sh = 40 - ex;
    uint32_t lo = 0, hi = 0;
#pragma unroll
    for (int t = 0; t < RK; ++t) {
        const long long q = round_to_i64(ldexp((double)v[t], sh));
        lo += (uint32_t)q & 0xFFFFFFu;
        hi += (uint32_t)(int32_t)(q >> 24);
    }
    const long long acc = (long long)(int32_t)wave_sum_u32(hi) * (1ll << 24) + (long long)wave_sum_u32(lo);
    return (float)ldexp((double)acc, -sh);
}

__device__ __forceinline__ uint32_t pow2_at_least(uint32_t v) {
    uint32_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Selection of the k best candidates when the buffer fills up:
//   kSelSort  : wave-wide bitonic sort of the buffer, keep the first k;
//   kSelRadix : the k-th smallest key by a wave-wide binary search over its bits (count of
//               keys below a probe: one ballot per key slot), then compaction of the keys
//               up to it, unsorted; the final set is sorted once at the end.
//   kSelWindow: intermediate selections keep any set of at least k and at most CAP - 64
//               keys below a high-word threshold (a short binary search); the exact k
//               best are selected once, at the end, as kSelRadix does.
enum : int { kSelSort = 0, kSelRadix = 1, kSelWindow = 2 };

// k-th smallest (1-based) of the keys held R per lane (~0 for empty slots), wave-uniform;
// the largest v with count(key < v) < k.
template <int R>
__device__ __forceinline__ uint64_t wave_kth_key(const uint64_t (&key)[R], uint32_t k) {
    // high word (d2 bits) first, then the low word (photon index) among the ties.  Every
    // key's high word lies in [hmin, hmax], so the answer shares their common leading
    // bits and the search starts below the highest bit in which they differ.
    uint32_t hmin = 0xFFFFFFFFu, hmax = 0u;
#pragma unroll
    for (int t = 0; t < R; ++t) {
        if (key[t] != ~0ull) {
            hmin = min(hmin, (uint32_t)(key[t] >> 32));
            hmax = max(hmax, (uint32_t)(key[t] >> 32));
        }
    }
    hmin = wave_min_u32(hmin);
    hmax = wave_max_u32(hmax);
    const uint32_t differ = hmin ^ hmax;
    const int top = differ ? 31 - __clz(differ) : -1;
    uint32_t hi = (top < 0) ? hmin : (top >= 31) ? 0u : (hmin & ~((2u << top) - 1u));
    for (int b = top; b >= 0; --b) {
        const uint32_t probe = hi | (1u << b);
        uint32_t c = 0;
#pragma unroll
        for (int t = 0; t < R; ++t) c += (uint32_t)__popcll(ballot((uint32_t)(key[t] >> 32) < probe));
        if (c < k) hi = probe;
    }
    uint32_t below = 0, tie = 0;
#pragma unroll
    for (int t = 0; t < R; ++t) {
        const uint32_t h = (uint32_t)(key[t] >> 32);
        below += (uint32_t)__popcll(ballot(h < hi));
        tie += (uint32_t)__popcll(ballot(h == hi));
    }
    const uint32_t need = k - below;  // 1 <= need <= tie
    if (need == tie) {
        // every tie is kept: the threshold is the largest tie
        uint32_t lo = 0;
#pragma unroll
        for (int t = 0; t < R; ++t) lo = max(lo, ((uint32_t)(key[t] >> 32) == hi) ? (uint32_t)key[t] : 0u);
        return ((uint64_t)hi << 32) | wave_max_u32(lo);
    }
    uint64_t eq[R];  // slots whose high word ties (separate ballots: a ballot of a && b costs two more)
#pragma unroll
    for (int t = 0; t < R; ++t) eq[t] = ballot((uint32_t)(key[t] >> 32) == hi);
    uint32_t lo = 0;
    for (int b = 31; b >= 0; --b) {
        const uint32_t probe = lo | (1u << b);
        uint32_t c = 0;
#pragma unroll
        for (int t = 0; t < R; ++t) c += (uint32_t)__popcll(eq[t] & ballot((uint32_t)key[t] < probe));
        if (c < need) lo = probe;
    }
    return ((uint64_t)hi << 32) | lo;
}

// A pruning threshold on the high word (d2 bits) that keeps between k and hi_k of the cnt
// keys held R per lane (~0 for empty slots): *upper = count(h <= T).  Binary search over
// the high word that stops as soon as the keys at or below the range top number at most
// hi_k (far fewer steps than the exact k-th); *upper > hi_k means ties made it impossible.
template <int R>
__device__ __forceinline__ uint32_t wave_window_hi(const uint64_t (&key)[R], uint32_t k, uint32_t hi_k, uint32_t cnt,
                                                   uint32_t* upper_out) {
    if (cnt <= hi_k) {
        uint32_t m = 0;
#pragma unroll
        for (int t = 0; t < R; ++t) {
            const uint32_t h = (uint32_t)(key[t] >> 32);
            m = max(m, h != 0xFFFFFFFFu ? h : 0u);
        }
        *upper_out = cnt;
        return wave_max_u32(m);
    }
    // [prefix, prefix + width) holds the k-th high word; upper = count(h < prefix + width)
    uint32_t prefix = 0, upper = cnt;
    uint64_t width = 1ull << 32;
    while (width > 1 && upper > hi_k) {
        width >>= 1;
        const uint32_t probe = prefix + (uint32_t)width;
        uint32_t c = 0;
#pragma unroll
        for (int t = 0; t < R; ++t) c += (uint32_t)__popcll(ballot((uint32_t)(key[t] >> 32) < probe));
        if (c >= k) upper = c;
        else prefix = probe;
    }
    *upper_out = upper;
    return prefix + (uint32_t)(width - 1);
}

// The tree a query walks: from the launch arguments (one tree) or the query's scene.
struct KnnTree {
    const uint2* nodes;
    const float4* recs;
    float r2max;
    float r_max;
};

// Backward mode, pair output: instead of one atomic per (query, neighbour) the wave appends
// (photon << 8 | slot) to a per-wave LDS list and keeps the query's fixed-point weights in
// wtab[slot]; the caller merges a chunk of consecutive queries before touching memory.
struct PairSink {
    uint64_t* pairs;
    uint32_t* npairs;  // LDS word, wave-uniform
    long long* wtab;   // 3 per slot
    uint32_t slot;
    uint32_t pbase;    // added to the photon index (offset of the scene's photons in a batch)
    double fix_scale;  // accumulator units per unit gradient for this query's scene
};

// Warm start for consecutive queries of one wave on one tree: every one of the previous
// query's k neighbours lies within d_k(prev) + |prev - q| of q, so that radius (rounded
// up) bounds this query's k-th distance and the search starts with it instead of r_max.
struct WarmHint {
    float x, y, z;
    float r2;  // previous query's k-th d2; < 0: none (fewer than k found, or first query)
    // Queries in spatial-cell order (incoherent hits) lie farther apart than the previous
    // query's k-th distance more often, so that the bound above is loose: they search first
    // within d_k(prev) + |prev - q| / 2, a guess, and within the bound only when it holds
    // fewer than k.  Camera hits in tile order lie close together and use the bound.
    bool guess;
};

// Warm-start bound at q from a query at h with k-th distance sqrt(h.r2): every one of
// that query's k neighbours lies within sqrt(h.r2) + |h - q| of q.  Computed in fp32 with
// the hardware square root; margins of 2^-16 per step are far above its and the fp32
// distances' roundings (a few units of 2^-24), and the bound is exclusive, hence the next
// float up.  r2max when the hint has no distance or the bound is not below it.  The search
// stays exact whatever this returns: a bound holding fewer than k photons is searched again
// with r2max, and one holding k or more holds the k nearest.
__device__ __forceinline__ float warm_bound(const WarmHint& h, float qx, float qy, float qz, float r2max) {
    if (!(h.r2 >= 0.0f)) return r2max;
    constexpr float kMargin = 1.0f + 0x1p-16f;
    const float dx = qx - h.x, dy = qy - h.y, dz = qz - h.z;
    const float b = (__builtin_amdgcn_sqrtf(h.r2) + __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz)) * kMargin;
    const float b2 = (b * b) * kMargin;
    return b2 < r2max ? __uint_as_float(__float_as_uint(b2) + 1u) : r2max;
}

// One query by one wave; keys / kpos: CAP LDS entries, snode / soff: kKnnStack.
template <int CAP, int kMode, int kSel = kSelSort, bool kPairs = false>
__device__ __forceinline__ void knn_query(const KnnArgs& A, const KnnTree& Tr, uint32_t q, uint32_t lane,
                                          uint64_t* keys, uint32_t* kpos, uint32_t* snode, float4* soff,
                                          const PairSink* sink = nullptr, WarmHint* hint = nullptr) {
    const float4 qp4 = A.pos[q];
    const float qx = uniformf(qp4.x), qy = uniformf(qp4.y), qz = uniformf(qp4.z);
    const f3 qp = mk3(qx, qy, qz);
    const uint32_t lane32 = lane * 32u;  // this lane's byte offset in a 64-record leaf batch
    const uint32_t k = A.k;
    const float rq = A.r_query ? uniformf(A.r_query[q]) : Tr.r_max;
    const float r2max = A.r_query ? rq * rq : Tr.r2max;
    // search radius: r_max, or the warm-start bound when smaller (warm_bound), and for
    // queries in spatial-cell order first the guess of WarmHint::guess
    float r2lim = r2max;
    float r2next = r2max;  // the next pass's limit when this one holds fewer than k
    if (hint && k > 0) {
        r2lim = r2next = warm_bound(*hint, qx, qy, qz, r2max);
        if (hint->guess && hint->r2 >= 0.0f) {
            const float dx = qx - hint->x, dy = qy - hint->y, dz = qz - hint->z;
            const float g =
                __builtin_amdgcn_sqrtf(hint->r2) + 0.5f * __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz);
            r2lim = fminf(r2next, g * g);
        }
    }

    uint32_t cnt = 0;        // candidates in the buffer
    bool full = false;       // the buffer holds the k best so far (after a window: more)
    float thr_d2 = kInf;     // with full: the k best so far lie at d2 <= thr_d2
    int sp = 0;
    uint32_t node = 0;
    // Squared per-axis distances from the query to the current node's box (0 inside) and,
    // per split axis, the sum of the other two: a far child's box distance is
    // diff^2 + rest[a].  A photon's d2 sums its terms in x, y, z order; the box distance's
    // other order can round a few units in the last place differently, so the pruning
    // limit plim carries a relative margin of 2^-20 (far above those roundings): a node is
    // dropped only when every photon in it lies beyond the limit.
    float ox2 = 0.0f, oy2 = 0.0f, oz2 = 0.0f;
    // lanes 0-2 hold the query's x, y, z and the three rest sums: a split axis picks its
    // value with one lane read
    const float qlane = lane == 0 ? qx : (lane == 1 ? qy : qz);
    float rlane = 0.0f;
    // a photon is taken when d2 < tlim (r2lim, and with full also d2 bits <= thr_d2 bits:
    // a superset of the keys below the k-th; the exact selections trim it)
    float tlim = r2lim, plim = r2lim * kPruneMargin;
    auto set_limits = [&]() {
        tlim = r2lim;
        if (full) {
            const uint32_t tb = __float_as_uint(thr_d2);
            tlim = fminf(r2lim, __uint_as_float(tb < 0x7F800000u ? tb + 1u : tb));
        }
        plim = (full ? fminf(r2lim, thr_d2) : r2lim) * kPruneMargin;
    };
    auto set_rest = [&]() { rlane = lane == 0 ? oy2 + oz2 : (lane == 1 ? ox2 + oz2 : ox2 + oy2); };

    // sorts the buffer and keeps the first `keep`
    auto sort_keep = [&](uint32_t keep) {
        uint32_t size = pow2_at_least(cnt);
        for (uint32_t t = cnt + lane; t < size; t += 64) {
            keys[t] = ~0ull;
            kpos[t] = 0u;
        }
        wave_sync();
        wave_bitonic<CAP>(keys, kpos, size, lane);
        if (cnt > keep) cnt = keep;
        if (cnt == k) {
            full = true;
            thr_d2 = __uint_as_float((uint32_t)(keys[k - 1] >> 32));
            set_limits();
        }
    };
    // keeps the k best, unsorted (cnt > k)
    auto radix_keep_slots = [&](auto slots) {
        constexpr int R = decltype(slots)::value;
        uint64_t key[R];
        uint32_t pos[R];
#pragma unroll
        for (int t = 0; t < R; ++t) {
            const uint32_t i = lane + 64u * t;
            key[t] = i < cnt ? keys[i] : ~0ull;
            pos[t] = i < cnt ? kpos[i] : 0u;
        }
        const uint64_t kth = wave_kth_key<R>(key, k);
        wave_sync();
        uint32_t base = 0;
#pragma unroll
        for (int t = 0; t < R; ++t) {
            const bool keepit = key[t] <= kth;
            const uint64_t m = ballot(keepit);
            if (keepit) {
                const uint32_t slot = base + mask_prefix(m);
                keys[slot] = key[t];
                kpos[slot] = pos[t];
            }
            base += (uint32_t)__popcll(m);
        }
        wave_sync();
        cnt = k;
        full = true;
        thr_d2 = __uint_as_float((uint32_t)(kth >> 32));
        set_limits();
    };
    // keeps the k best, unsorted (cnt > k); searches only the ceil(cnt / 64) key slots per
    // lane that hold candidates
    auto radix_keep = [&]() {
        constexpr int R = CAP / 64;
        if (R > 2 && cnt <= 128u) radix_keep_slots(std::integral_constant<int, (R > 2 ? 2 : R)>{});
        else if (R > 3 && cnt <= 192u) radix_keep_slots(std::integral_constant<int, (R > 3 ? 3 : R)>{});
        else if (R > 4 && cnt <= 256u) radix_keep_slots(std::integral_constant<int, (R > 4 ? 4 : R)>{});
        else radix_keep_slots(std::integral_constant<int, R>{});
    };
    // keeps at least k and at most CAP - 64 candidates under a high-word threshold
    auto window_keep = [&]() {
        constexpr int R = CAP / 64;
        uint64_t key[R];
        uint32_t pos[R];
#pragma unroll
        for (int t = 0; t < R; ++t) {
            const uint32_t i = lane + 64u * t;
            key[t] = i < cnt ? keys[i] : ~0ull;
            pos[t] = i < cnt ? kpos[i] : 0u;
        }
        uint32_t upper = 0;
        const uint32_t T = wave_window_hi<R>(key, k, (uint32_t)CAP - 64u, cnt, &upper);
        if (upper > (uint32_t)CAP - 64u) {  // equal distances: only the exact selection fits
            radix_keep();
            return;
        }
        wave_sync();
        uint32_t base = 0;
#pragma unroll
        for (int t = 0; t < R; ++t) {
            const uint64_t m = ballot((uint32_t)(key[t] >> 32) <= T) & ballot(key[t] != ~0ull);
            if (__builtin_amdgcn_inverse_ballot_w64(m)) {
                const uint32_t slot = base + mask_prefix(m);
                keys[slot] = key[t];
                kpos[slot] = pos[t];
            }
            base += (uint32_t)__popcll(m);
        }
        wave_sync();
        cnt = upper;
        full = true;
        thr_d2 = __uint_as_float(T);
        set_limits();
    };
    auto select = [&](uint32_t keep) {
        if (kSel == kSelWindow) window_keep();
        else if (kSel == kSelRadix && cnt > keep) radix_keep();
        else sort_keep(keep);
    };

    for (;;) {  // a second pass with r_max only if the warm-start bound held fewer than k
        for (;;) {
            node = uniform(node);
            uint2 nd = ld_uniform8(Tr.nodes, node);
            nd.x = uniform(nd.x);
            nd.y = uniform(nd.y);
            while (!is_leaf(nd.y)) {
                const uint32_t a = nd.y >> 30;
                const float s = __uint_as_float(nd.x);
                const float qa = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(qlane), (int)a));
                const float rest = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rlane), (int)a));
                const float diff = qa - s;  // > 0 exactly when qa > s
                const float d2a = diff * diff;
                // far child's box distance: the split plane replaces this axis' offset
                const float bd2 = d2a + rest;
                const uint32_t left = node + 1, right = nd.y & MNRT_INDEX_MASK;
                const bool gr = diff > 0.0f;
                // entries are far children of distinct levels of the current path, so a
                // tree of max_level <= MNRT_MAX_LEVEL < kKnnStack never fills the stack; the
                // slot clamp only keeps a malformed tree's writes inside it
                if (bd2 <= plim) {
                    if (lane == 0) {
                        const int e = min(sp, kKnnStack - 1);
                        snode[e] = gr ? left : right;
                        soff[e] = make_float4(a == 0 ? d2a : ox2, a == 1 ? d2a : oy2, a == 2 ? d2a : oz2, bd2);
                    }
                    ++sp;
                }
                node = uniform(gr ? right : left);
                nd = ld_uniform8(Tr.nodes, node);
                nd.x = uniform(nd.x);
                nd.y = uniform(nd.y);
            }
            wave_sync();
            const uint32_t count = nd.x, begin = nd.y & MNRT_INDEX_MASK;
            // the leaf's records (32 B each, position and photon index first) through a buffer
            // descriptor that ends with the leaf: lanes past its end read zeros and are masked out
            const __amdgpu_buffer_rsrc_t lr = __builtin_amdgcn_make_buffer_rsrc(
                (void*)(Tr.recs + 2ull * begin), (short)0, (int)(count * 32u), kBufferWord3);
            for (uint32_t base = 0; base < count; base += 64) {
                const auto w = __builtin_amdgcn_raw_buffer_load_b128(lr, (int)(lane32 + base * 32u), 0, 0);
                const float dx = __uint_as_float(w[0]) - qx, dy = __uint_as_float(w[1]) - qy,
                            dz = __uint_as_float(w[2]) - qz;
                const float d2 = dx * dx + dy * dy + dz * dz;
                uint64_t m = ballot(d2 < tlim) & ballot(lane < count - base);
                uint32_t nm = (uint32_t)__popcll(m);
                if (nm == 0) continue;
                if (cnt + nm > (uint32_t)CAP) {
                    select(k);
                    m &= ballot(d2 < tlim);
                    nm = (uint32_t)__popcll(m);
                }
                if (__builtin_amdgcn_inverse_ballot_w64(m)) {
                    const uint32_t slot = mask_prefix_from(m, cnt);
                    keys[slot] = ((uint64_t)__float_as_uint(d2) << 32) | (uint64_t)w[3];
                    kpos[slot] = begin + base + lane;
                }
                cnt += nm;
                wave_sync();
            }
            if (!full && cnt >= k) select(k);
            bool found = false;
            while (sp > 0) {
                --sp;
                const int e = min(sp, kKnnStack - 1);
                const float4 f = soff[e];
                if (uniformf(f.w) > plim) continue;
                node = uniform(snode[e]);
                ox2 = uniformf(f.x);
                oy2 = uniformf(f.y);
                oz2 = uniformf(f.z);
                set_rest();
                found = true;
                break;
            }
            wave_sync();
            if (!found) break;
        }
        if (cnt >= k || r2lim == r2max) break;
        // a guess that held fewer than k: the bound next; the bound holds k (the r_max pass
        // after it keeps the search exact whatever the roundings)
        r2lim = r2lim < r2next ? r2next : r2max;
        cnt = 0;
        full = false;
        thr_d2 = kInf;
        set_limits();
        sp = 0;
        node = 0;
        ox2 = oy2 = oz2 = 0.0f;
        rlane = 0.0f;
    }
    if (kSel != kSelSort && cnt > k) radix_keep();
    // the buffer now holds exactly the set; it is sorted only for callers of the lists
    if (kSel == kSelSort || A.out_nbr) sort_keep(k);
    const uint32_t found = cnt;  // min(#photons closer than r_max, k)

    float r2 = r2max;
    if (found >= k && k > 0) {
        // the k-th distance: the largest d2 of the set
        uint32_t hi = 0;
        for (uint32_t i = lane; i < found; i += 64) hi = max(hi, (uint32_t)(keys[i] >> 32));
        r2 = __uint_as_float(wave_max_u32(hi));
    }

    if (hint) {
        hint->x = qx;
        hint->y = qy;
        hint->z = qz;
        hint->r2 = (found >= k && k > 0) ? r2 : -1.0f;
    }
    if (A.out_count && lane == 0) A.out_count[q] = found;
    if (A.out_nbr) {
        for (uint32_t i = lane; i < k; i += 64)
            A.out_nbr[(size_t)q * k + i] = (i < found) ? (uint32_t)(keys[i] & 0xFFFFFFFFull) : 0xFFFFFFFFu;
    }
    if (kMode == kModeRadii) {
        if (A.out_radius && lane == 0) A.out_radius[q] = (found >= k && k > 0) ? sqrtf(r2) : rq;
        if (A.out_r2 && lane == 0) A.out_r2[q] = r2;
        return;
    }

    // Normal flipped toward the source ray (photon_knn.cu:500-505).
    f3 nq = xyz(A.nrm[q]);
    f3 rd = xyz(A.dir[q]);
    f3 nrd = mk3(-rd.x, -rd.y, -rd.z);
    if (dot3(nrd, nq) < 0.0f) nq = mk3(nq.x * -1.0f, nq.y * -1.0f, nq.z * -1.0f);

    // Front-facing powers: each lane fetches the records of its slots of the set, the
    // channel sums are exact fixed-point sums (independent of the set's order).  The set
    // has at most k <= CAP - 64 entries (the launch picks CAP >= k + 64), so ceil(k / 64)
    // slots per lane hold it; slots beyond would only add zeros.
    constexpr int RK = std::min<int>((CAP - 64) / 64, (int)((MNRT_KNN_MAX_K + 63u) / 64u));
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    auto front_sums = [&](auto slots) {
        constexpr int S = decltype(slots)::value;
        float vx[S], vy[S], vz[S];
#pragma unroll
        for (int t = 0; t < S; ++t) {
            const uint32_t i = lane + 64u * (uint32_t)t;
            vx[t] = vy[t] = vz[t] = 0.0f;
            if (i < found) {
                const float4 pw = Tr.recs[2ull * kpos[i] + 1];
                if (dot3(nq, oct_decode(__float_as_uint(pw.w))) < 0.0f) {
                    vx[t] = pw.x;
                    vy[t] = pw.y;
                    vz[t] = pw.z;
                }
            }
        }
        sx = wave_fixed_sum<S>(vx);
        sy = wave_fixed_sum<S>(vy);
        sz = wave_fixed_sum<S>(vz);
    };
    if (RK >= 2 && k <= 64u) front_sums(std::integral_constant<int, 1>{});
    else if (RK >= 3 && k <= 128u) front_sums(std::integral_constant<int, (RK >= 3 ? 2 : RK)>{});
    else if (RK >= 4 && k <= 192u) front_sums(std::integral_constant<int, (RK >= 4 ? 3 : RK)>{});
    else front_sums(std::integral_constant<int, RK>{});
    float ix = 0.0f, iy = 0.0f, iz = 0.0f;
    if (r2 > 0.0f) {
        float den = r2 * kPi;
        ix = sx / den;
        iy = sy / den;
        iz = sz / den;
    }
    float4 c4 = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    if (A.clr) {
        c4 = A.clr[q];
        if (found < 8) ix = iy = iz = 0.0f;
    }
    float4 in4 = A.infl[q];
    if (kMode == kModeBackward) {
        // Adjoint with the neighbour set fixed: L_c = ((alpha rho_c / pi) S_c / (pi r2)) infl_c,
        // so dLoss/dS_c = G_c (alpha rho_c / pi) infl_c / (pi r2) for every front-facing
        // neighbour's power, and dLoss/drho_c = G_c (alpha / pi) E_c infl_c.
        float4 g = A.grad[q];
        bool gated = (A.clr && found < 8) || !(r2 > 0.0f);
        float den = r2 * kPi;
        float wx = gated ? 0.0f : (g.x * (((c4.w * c4.x) * kInvPi) * in4.x)) / den;
        float wy = gated ? 0.0f : (g.y * (((c4.w * c4.y) * kInvPi) * in4.y)) / den;
        float wz = gated ? 0.0f : (g.z * (((c4.w * c4.z) * kInvPi) * in4.z)) / den;
        if (A.out_grad_clr && lane == 0)
            A.out_grad_clr[q] = make_float4(g.x * (((c4.w * kInvPi) * ix) * in4.x),
                                            g.y * (((c4.w * kInvPi) * iy) * in4.y),
                                            g.z * (((c4.w * kInvPi) * iz) * in4.z), 0.0f);
        // fixed-point adds are associative, so the per-photon sums do not depend on the
        // order in which queries arrive
        const double fs = kPairs ? sink->fix_scale : A.fix_scale;
        long long fx = __double2ll_rn((double)wx * fs);
        long long fy = __double2ll_rn((double)wy * fs);
        long long fz = __double2ll_rn((double)wz * fs);
        if (kPairs) {
            if (fx == 0 && fy == 0 && fz == 0) return;
            if (lane == 0) {
                sink->wtab[3 * sink->slot + 0] = fx;
                sink->wtab[3 * sink->slot + 1] = fy;
                sink->wtab[3 * sink->slot + 2] = fz;
            }
            uint32_t np = uniform(*sink->npairs);
            for (uint32_t base = 0; base < found; base += 64) {
                const uint32_t i = base + lane;
                bool front = false;
                uint64_t key = 0;
                if (i < found) {
                    const float4 pw = Tr.recs[2ull * kpos[i] + 1];
                    front = dot3(nq, oct_decode(__float_as_uint(pw.w))) < 0.0f;
                    key = (((keys[i] & 0xFFFFFFFFull) + sink->pbase) << 8) | sink->slot;
                }
                const uint64_t m = ballot(front);
                if (front) sink->pairs[np + mask_prefix(m)] = key;
                np += (uint32_t)__popcll(m);
            }
            wave_sync();
            if (lane == 0) *sink->npairs = np;
            wave_sync();
            return;
        }
        if (fx != 0 || fy != 0 || fz != 0) {
            for (uint32_t i = lane; i < found; i += 64) {
                float4 pw = Tr.recs[2ull * kpos[i] + 1];
                f3 pd = oct_decode(__float_as_uint(pw.w));
                if (dot3(nq, pd) < 0.0f) {
                    size_t p = (size_t)(keys[i] & 0xFFFFFFFFull);
                    atomicAdd(&A.acc[3 * p + 0], (unsigned long long)fx);
                    atomicAdd(&A.acc[3 * p + 1], (unsigned long long)fy);
                    atomicAdd(&A.acc[3 * p + 2], (unsigned long long)fz);
                }
            }
        }
        return;
    }
    float rx = (((c4.w * c4.x) * kInvPi) * ix) * in4.x;
    float ry = (((c4.w * c4.y) * kInvPi) * iy) * in4.y;
    float rz = (((c4.w * c4.z) * kInvPi) * iz) * in4.z;
    if (lane == 0) {
        if (A.inv_dist) {
            // reciprocal distance from the source ray's origin (photon_knn.cu:529-535)
            f3 od = sub3(qp, xyz(A.orig[q]));
            float v = 1.0f / sqrtf(dot3(od, od));
            A.inv_dist[q] = A.inv_accumulate ? A.inv_dist[q] + v : v;
        }
        if (A.out_r2) A.out_r2[q] = r2;
        if (A.out_irr) A.out_irr[q] = make_float4(ix, iy, iz, 0.0f);
        if (A.accumulate) {
            float4 o = A.rad_out[q];
            A.rad_out[q] = make_float4(o.x + rx, o.y + ry, o.z + rz, o.w + 0.0f);
        } else {
            A.rad_out[q] = make_float4(rx, ry, rz, 0.0f);
        }
    }
}

// Consecutive queries per wave: each after the first starts from the previous one's
// k-th distance (WarmHint); neighbouring pixels' hits share most of their neighbours.
constexpr uint32_t kKnnChain = 8;

template <int CAP, int kMode, int kSel>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_knn(KnnArgs A) {
    __shared__ uint64_t s_key[kWavesPerBlock][CAP];
    __shared__ uint32_t s_pos[kWavesPerBlock][CAP];
    __shared__ uint32_t s_snode[kWavesPerBlock][kKnnStack];
    __shared__ float4 s_soff[kWavesPerBlock][kKnnStack];
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t j0 = ((uint64_t)blockIdx.x * kWavesPerBlock + wave) * kKnnChain;
    if (j0 >= A.n) return;  // uniform per wave
    const uint32_t j1 = (uint32_t)min<uint64_t>(j0 + kKnnChain, A.n);
    const KnnTree Tr{A.nodes, A.recs, A.r2max, A.r_max};
    WarmHint hint{0.0f, 0.0f, 0.0f, -1.0f, A.order_far && !order_coherent(A.n, A.order_far)};
    for (uint32_t j = (uint32_t)j0; j < j1; ++j) {
        const uint32_t q = A.order ? ld_uniform(A.order, j) : j;
        knn_query<CAP, kMode, kSel>(A, Tr, q, lane, s_key[wave], s_pos[wave], s_snode[wave], s_soff[wave], nullptr,
                                    &hint);
        wave_sync();
    }
}

// Batched radiance: waves take chunks of `chunk` consecutive query slots from a work
// counter (one counter per XCD share, as k_trace_batch), slot g belonging to the scene
// whose slot range holds it.  Slots at or past a scene's count (count[s], device side:
// the number of shading points the batched compaction wrote) are left untouched.
template <int CAP, int kWaves, int kSel>
__global__ __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_knn_batch(KnnArgs G, const KnnScene* __restrict__ scenes,
                                                           const uint32_t* __restrict__ slot_begin, uint32_t ns,
                                                           const uint32_t* __restrict__ count, uint32_t chunk,
                                                           uint32_t num_parts, const uint32_t* __restrict__ part_begin,
                                                           const uint32_t* __restrict__ part_end,
                                                           uint32_t* __restrict__ work) {
    __shared__ uint64_t s_key[kWaves][CAP];
    __shared__ uint32_t s_pos[kWaves][CAP];
    __shared__ uint32_t s_snode[kWaves][kKnnStack];
    __shared__ float4 s_soff[kWaves][kKnnStack];
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t xcd = blockIdx.x % num_parts;
    for (uint32_t part = 0; part < num_parts; ++part) {
        const uint32_t p = (xcd + part) % num_parts;
        const uint32_t pb = part_begin[p], pe = part_end[p];
        for (;;) {
            uint32_t c = 0;
            if (lane == 0) c = atomicAdd(&work[p], 1u);
            c = uniform(__shfl(c, 0, 64));
            const uint64_t g0 = (uint64_t)pb + (uint64_t)c * chunk;
            if (g0 >= pe) break;
            const uint32_t g1 = (uint32_t)min<uint64_t>(g0 + chunk, pe);
            // the scene of the chunk's slots, refreshed only when a slot crosses into the next
            uint32_t s = uniform(seg_of(slot_begin, ns, (uint32_t)g0));
            uint32_t b = 0, e = 0, m = 0;
            KnnTree Tr{};
            WarmHint hint{0.0f, 0.0f, 0.0f, -1.0f, false};
            auto enter = [&]() {
                b = ld_uniform(slot_begin, s);
                e = ld_uniform(slot_begin, s + 1);
                m = count ? ld_uniform(count, s) : e - b;
                const KnnScene sc = ld_uniform(scenes, s);
                Tr = KnnTree{sc.nodes, sc.recs, sc.r2max, sc.r_max};
                hint.r2 = -1.0f;  // another scene's photons
            };
            enter();
            for (uint32_t g = (uint32_t)g0; g < g1; ++g) {
                if (g >= e) {
                    do {
                        ++s;
                    } while (g >= ld_uniform(slot_begin, s + 1));
                    enter();
                }
                if (g - b >= m) continue;
                knn_query<CAP, kModeRadiance, kSel>(G, Tr, g, lane, s_key[wave], s_pos[wave], s_snode[wave],
                                                    s_soff[wave], nullptr, &hint);
                wave_sync();
            }
        }
    }
}

// Backward pass over chunks of `chunk` consecutive queries per wave (chunk x k <= kPairs):
// the chunk's (photon, query) pairs are sorted by photon in LDS and each photon's
// fixed-point contributions summed before one atomic per channel.  Consecutive queries
// (neighbouring pixels) share most of their neighbours, so this issues a fraction of the
// per-pair atomics; the integer sums are the same in any order.
constexpr int kPairCap = 256;

// Sorts a wave's (photon << 8 | slot) pairs by photon in LDS and adds each photon's summed
// fixed-point weights with one atomic per channel.
__device__ __forceinline__ void merge_pairs(uint64_t* pairs, uint32_t np, const long long* w,
                                            unsigned long long* acc, uint32_t lane) {
    if (np == 0) return;
    // sort the pairs by (photon, slot); pad to a power of two with keys above every pair
    uint32_t size = 64;
    while (size < np) size <<= 1;
    for (uint32_t t = np + lane; t < size; t += 64) pairs[t] = ~0ull;
    wave_sync();
    for (uint32_t k2 = 2; k2 <= size; k2 <<= 1) {
        for (uint32_t j = k2 >> 1; j > 0; j >>= 1) {
            for (uint32_t t = lane; t < (size >> 1); t += 64) {
                const uint32_t i = 2 * j * (t / j) + (t % j), l = i + j;
                const uint64_t a = pairs[i], b = pairs[l];
                if ((a > b) == ((i & k2) == 0)) {
                    pairs[i] = b;
                    pairs[l] = a;
                }
            }
            wave_sync();
        }
    }
    // one lane per run of equal photons: sum its queries' weights, one atomic per channel
    for (uint32_t i = lane; i < np; i += 64) {
        const uint64_t p = pairs[i] >> 8;
        if (i > 0 && (pairs[i - 1] >> 8) == p) continue;
        long long sx = 0, sy = 0, sz = 0;
        for (uint32_t j = i; j < np && (pairs[j] >> 8) == p; ++j) {
            const uint32_t sl = (uint32_t)(pairs[j] & 0xFFu);
            sx += w[3 * sl + 0];
            sy += w[3 * sl + 1];
            sz += w[3 * sl + 2];
        }
        if (sx) atomicAdd(&acc[3 * p + 0], (unsigned long long)sx);
        if (sy) atomicAdd(&acc[3 * p + 1], (unsigned long long)sy);
        if (sz) atomicAdd(&acc[3 * p + 2], (unsigned long long)sz);
    }
}

template <int CAP, int kSel>
__global__ __launch_bounds__(256) void k_knn_bwd_chunks(KnnArgs A, uint32_t chunk) {
    __shared__ uint64_t s_key[kWavesPerBlock][CAP];
    __shared__ uint32_t s_pos[kWavesPerBlock][CAP];
    __shared__ uint32_t s_snode[kWavesPerBlock][kKnnStack];
    __shared__ float4 s_soff[kWavesPerBlock][kKnnStack];
    __shared__ uint64_t s_pairs[kWavesPerBlock][kPairCap];
    __shared__ long long s_w[kWavesPerBlock][3 * (kPairCap / 8)];
    __shared__ uint32_t s_np[kWavesPerBlock];
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t c = blockIdx.x * kWavesPerBlock + wave;
    const uint64_t j0 = (uint64_t)c * chunk;
    if (j0 >= A.n) return;  // uniform per wave
    const uint32_t j1 = (uint32_t)min<uint64_t>(j0 + chunk, A.n);
    uint64_t* pairs = s_pairs[wave];
    if (lane == 0) s_np[wave] = 0;
    wave_sync();
    const KnnTree Tr{A.nodes, A.recs, A.r2max, A.r_max};
    WarmHint hint{0.0f, 0.0f, 0.0f, -1.0f, A.order_far && !order_coherent(A.n, A.order_far)};
    for (uint32_t j = (uint32_t)j0; j < j1; ++j) {
        const uint32_t q = A.order ? ld_uniform(A.order, j) : j;
        const PairSink sink{pairs, &s_np[wave], s_w[wave], j - (uint32_t)j0, 0u, A.fix_scale};
        knn_query<CAP, kModeBackward, kSel, true>(A, Tr, q, lane, s_key[wave], s_pos[wave], s_snode[wave],
                                                  s_soff[wave], &sink, &hint);
        wave_sync();
    }
    merge_pairs(pairs, uniform(s_np[wave]), s_w[wave], A.acc, lane);
}

// Adds per-query radiance into slots given by a slot index per query (final gathering,
// PhotonMap::GatherFor, PhotonMap.cpp:234-270).  A run of equal consecutive slots is
// summed in query order by the run's first query, so sorted slot indices (as final
// gathering produces them) give a deterministic result.
__global__ __launch_bounds__(256) void k_scatter_runs(const float4* __restrict__ per_query,
                                                      const float* __restrict__ per_query_inv,
                                                      const uint32_t* __restrict__ idx_sp, uint32_t n,
                                                      float4* __restrict__ io_rad, float* __restrict__ io_inv) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t slot = idx_sp[i];
    if (i > 0 && idx_sp[i - 1] == slot) return;
    float x = 0.0f, y = 0.0f, z = 0.0f, w = 0.0f;
    for (uint32_t j = i; j < n && idx_sp[j] == slot; ++j) {
        float4 r = per_query[j];
        x = x + r.x;
        y = y + r.y;
        z = z + r.z;
        if (per_query_inv) w = w + per_query_inv[j];
    }
    atomicAdd(&io_rad[slot].x, x);
    atomicAdd(&io_rad[slot].y, y);
    atomicAdd(&io_rad[slot].z, z);
    if (per_query_inv) atomicAdd(&io_inv[slot], w);
}

// Bound on |weight| per query for the fixed-point scale of the backward pass, from the
// forward's r2 and neighbour counts (same expressions as the backward kernel).
__global__ __launch_bounds__(256) void k_bwd_wmax(uint32_t n, const float4* __restrict__ grad,
                                                  const float4* __restrict__ clr, const float4* __restrict__ infl,
                                                  const float* __restrict__ r2, const uint32_t* __restrict__ count,
                                                  uint32_t* __restrict__ wmax_bits) {
    uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    float m = 0.0f;
    if (q < n) {
        float4 g = grad[q], in4 = infl[q];
        float4 c4 = clr ? clr[q] : make_float4(1.0f, 1.0f, 1.0f, 1.0f);
        float rr = r2[q];
        bool gated = (clr && count[q] < 8) || !(rr > 0.0f);
        if (!gated) {
            float den = rr * kPi;
            float wx = (g.x * (((c4.w * c4.x) * kInvPi) * in4.x)) / den;
            float wy = (g.y * (((c4.w * c4.y) * kInvPi) * in4.y)) / den;
            float wz = (g.z * (((c4.w * c4.z) * kInvPi) * in4.z)) / den;
            m = fmaxf(fabsf(wx), fmaxf(fabsf(wy), fabsf(wz)));
            if (!(m <= kInf)) m = kInf;  // NaN / inf gradients saturate the bound
        }
    }
    for (uint32_t o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63u) == 0 && m > 0.0f) atomicMax(wmax_bits, __float_as_uint(m));
}

__global__ __launch_bounds__(256) void k_fix_to_float(uint32_t np, const unsigned long long* __restrict__ acc,
                                                      double inv_scale, float4* __restrict__ out) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    out[p] = make_float4((float)((double)(long long)acc[3ull * p] * inv_scale),
                         (float)((double)(long long)acc[3ull * p + 1] * inv_scale),
                         (float)((double)(long long)acc[3ull * p + 2] * inv_scale), 0.0f);
}

// --- batched backward (one launch sequence for every scene of a batch) ---------------

// Per-scene weight bounds: slot g of scene s (g - slot_begin[s] < count[s]) adds its
// bound to wmax[s]; a wave whose slots lie in one scene reduces first.
__global__ __launch_bounds__(256) void k_bwd_wmax_batch(uint32_t n, const uint32_t* __restrict__ slot_begin,
                                                        uint32_t ns, const uint32_t* __restrict__ count,
                                                        const float4* __restrict__ grad,
                                                        const float4* __restrict__ clr,
                                                        const float4* __restrict__ infl,
                                                        const float* __restrict__ r2,
                                                        const uint32_t* __restrict__ nbr_count,
                                                        uint32_t* __restrict__ wmax_bits) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    float m = 0.0f;
    uint32_t s = 0;
    if (g < n) {
        s = seg_of(slot_begin, ns, g);
        const uint32_t cnt = count ? count[s] : slot_begin[s + 1] - slot_begin[s];
        if (g - slot_begin[s] < cnt) {
            float4 gr = grad[g], in4 = infl[g];
            float4 c4 = clr ? clr[g] : make_float4(1.0f, 1.0f, 1.0f, 1.0f);
            float rr = r2[g];
            bool gated = (clr && nbr_count[g] < 8) || !(rr > 0.0f);
            if (!gated) {
                float den = rr * kPi;
                float wx = (gr.x * (((c4.w * c4.x) * kInvPi) * in4.x)) / den;
                float wy = (gr.y * (((c4.w * c4.y) * kInvPi) * in4.y)) / den;
                float wz = (gr.z * (((c4.w * c4.z) * kInvPi) * in4.z)) / den;
                m = fmaxf(fabsf(wx), fmaxf(fabsf(wy), fabsf(wz)));
                if (!(m <= kInf)) m = kInf;
            }
        }
    }
    const uint32_t s0 = uniform(s);
    if (ballot(s != s0) == 0ull) {
        for (uint32_t o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
        if ((threadIdx.x & 63u) == 0 && m > 0.0f) atomicMax(&wmax_bits[s0], __float_as_uint(m));
    } else if (m > 0.0f) {
        atomicMax(&wmax_bits[s], __float_as_uint(m));
    }
}

// Fixed-point unit per scene from its bound and query count: the largest power of two
// with count * wmax * unit < 2^62 (the same expression as the single-scene entry point).
__global__ __launch_bounds__(64) void k_bwd_scale_batch(uint32_t ns, const uint32_t* __restrict__ slot_begin,
                                                        const uint32_t* __restrict__ count,
                                                        const uint32_t* __restrict__ wmax_bits,
                                                        double* __restrict__ scale) {
    for (uint32_t s = threadIdx.x; s < ns; s += blockDim.x) {
        const uint32_t n = count ? count[s] : slot_begin[s + 1] - slot_begin[s];
        const float wm = __uint_as_float(wmax_bits[s]);
        double sc = 0.0;
        if (n > 0 && wm > 0.0f && wm <= kInf) {
            int e = 0;
            (void)frexp(4.611686018427387904e18 / ((double)n * (double)wm), &e);
            sc = ldexp(1.0, e - 1);
        }
        scale[s] = sc;
    }
}

// Batched backward: waves take work items of `chunk` consecutive query slots (chunk x k <=
// kPairCap) from the work counters, as k_knn_batch does.  A pair carries the photon's index
// in the concatenation of the scenes' photon arrays (scene offset + original index) and its
// weight in the scene's own fixed-point unit, so one work item may span scenes.
template <int CAP, int kSel>
__global__ __launch_bounds__(256) void k_knn_bwd_batch(KnnArgs G, const KnnScene* __restrict__ scenes,
                                                       const uint32_t* __restrict__ slot_begin, uint32_t ns,
                                                       const uint32_t* __restrict__ count,
                                                       const uint32_t* __restrict__ photon_begin,
                                                       const double* __restrict__ scale, uint32_t chunk,
                                                       uint32_t num_parts, const uint32_t* __restrict__ part_begin,
                                                       const uint32_t* __restrict__ part_end,
                                                       uint32_t* __restrict__ work) {
    __shared__ uint64_t s_key[kWavesPerBlock][CAP];
    __shared__ uint32_t s_pos[kWavesPerBlock][CAP];
    __shared__ uint32_t s_snode[kWavesPerBlock][kKnnStack];
    __shared__ float4 s_soff[kWavesPerBlock][kKnnStack];
    __shared__ uint64_t s_pairs[kWavesPerBlock][kPairCap];
    __shared__ long long s_w[kWavesPerBlock][3 * (kPairCap / 8)];
    __shared__ uint32_t s_np[kWavesPerBlock];
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t xcd = blockIdx.x % num_parts;
    uint64_t* pairs = s_pairs[wave];
    for (uint32_t part = 0; part < num_parts; ++part) {
        const uint32_t p = (xcd + part) % num_parts;
        const uint32_t pb = part_begin[p], pe = part_end[p];
        for (;;) {
            uint32_t c = 0;
            if (lane == 0) c = atomicAdd(&work[p], 1u);
            c = uniform(__shfl(c, 0, 64));
            const uint64_t g0 = (uint64_t)pb + (uint64_t)c * chunk;
            if (g0 >= pe) break;
            const uint32_t g1 = (uint32_t)min<uint64_t>(g0 + chunk, pe);
            if (lane == 0) s_np[wave] = 0;
            wave_sync();
            uint32_t s = uniform(seg_of(slot_begin, ns, (uint32_t)g0));
            uint32_t b = 0, e = 0, m = 0, poff = 0;
            double fs = 0.0;
            KnnTree Tr{};
            WarmHint hint{0.0f, 0.0f, 0.0f, -1.0f, false};
            auto enter = [&]() {
                b = ld_uniform(slot_begin, s);
                e = ld_uniform(slot_begin, s + 1);
                m = count ? ld_uniform(count, s) : e - b;
                poff = ld_uniform(photon_begin, s);
                fs = ld_uniform(scale, s);
                const KnnScene sc = ld_uniform(scenes, s);
                Tr = KnnTree{sc.nodes, sc.recs, sc.r2max, sc.r_max};
                hint.r2 = -1.0f;
            };
            enter();
            for (uint32_t g = (uint32_t)g0; g < g1; ++g) {
                if (g >= e) {
                    do {
                        ++s;
                    } while (g >= ld_uniform(slot_begin, s + 1));
                    enter();
                }
                if (g - b >= m) continue;
                const PairSink sink{pairs, &s_np[wave], s_w[wave], g - (uint32_t)g0, poff, fs};
                knn_query<CAP, kModeBackward, kSel, true>(G, Tr, g, lane, s_key[wave], s_pos[wave], s_snode[wave],
                                                          s_soff[wave], &sink, &hint);
                wave_sync();
            }
            merge_pairs(pairs, uniform(s_np[wave]), s_w[wave], G.acc, lane);
            wave_sync();
        }
    }
}

// Photon gradients of every scene from the fixed-point sums (photon p of the
// concatenation belongs to the scene whose photon range holds it).
__global__ __launch_bounds__(256) void k_fix_to_float_batch(uint32_t np, const uint32_t* __restrict__ photon_begin,
                                                            uint32_t ns, const unsigned long long* __restrict__ acc,
                                                            const double* __restrict__ scale,
                                                            float4* __restrict__ out) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const double sc = scale[seg_of(photon_begin, ns, p)];
    const double inv = sc > 0.0 ? 1.0 / sc : 0.0;
    out[p] = make_float4((float)((double)(long long)acc[3ull * p] * inv),
                         (float)((double)(long long)acc[3ull * p + 1] * inv),
                         (float)((double)(long long)acc[3ull * p + 2] * inv), 0.0f);
}

__global__ __launch_bounds__(256) void k_zero_u64(unsigned long long* __restrict__ p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// Selection strategy of the single-tree launches.
constexpr int kKnnSelDefault = kSelWindow;

// Backward launches: queries in chunks of floor(kPairCap / k) per wave (at most 8).
int launch_knn_bwd(const KnnArgs& A, hipStream_t st) {
    const uint32_t chunk = std::max<uint32_t>(1u, std::min<uint32_t>(8u, (uint32_t)kPairCap / A.k));
    const uint32_t chunks = (A.n + chunk - 1) / chunk;
    dim3 block(64 * kWavesPerBlock), grid((chunks + kWavesPerBlock - 1) / kWavesPerBlock);
    const uint32_t need = A.k + 64;
    if (need <= 128)
        hipLaunchKernelGGL((k_knn_bwd_chunks<128, kKnnSelDefault>), grid, block, 0, st, A, chunk);
    else if (need <= 256)
        hipLaunchKernelGGL((k_knn_bwd_chunks<256, kKnnSelDefault>), grid, block, 0, st, A, chunk);
    else
        hipLaunchKernelGGL((k_knn_bwd_chunks<512, kKnnSelDefault>), grid, block, 0, st, A, chunk);
    MNRT_HIP_CHECK(hipGetLastError());
    return MNRT_OK;
}

template <int kMode>
int launch_knn(const KnnArgs& A, hipStream_t st) {
    const uint64_t per_block = (uint64_t)kWavesPerBlock * kKnnChain;
    dim3 block(64 * kWavesPerBlock), grid((uint32_t)((A.n + per_block - 1) / per_block));
    uint32_t need = A.k + 64;
    if (need <= 128)
        hipLaunchKernelGGL((k_knn<128, kMode, kKnnSelDefault>), grid, block, 0, st, A);
    else if (need <= 256)
        hipLaunchKernelGGL((k_knn<256, kMode, kKnnSelDefault>), grid, block, 0, st, A);
    else
        hipLaunchKernelGGL((k_knn<512, kMode, kKnnSelDefault>), grid, block, 0, st, A);
    MNRT_HIP_CHECK(hipGetLastError());
    return MNRT_OK;
}

int launch_knn_batch(const KnnArgs& G, const KnnScene* scenes, const uint32_t* slot_begin, uint32_t ns,
                     const uint32_t* count, uint32_t chunk, uint32_t num_parts, const uint32_t* part_begin,
                     const uint32_t* part_end, uint32_t* work, uint32_t cap, uint32_t waves, uint32_t sel,
                     uint32_t grid, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, work, num_parts, 0u);
#define MNRT_KB(C, W, S)                                                                                    \
    hipLaunchKernelGGL((k_knn_batch<C, W, S>), dim3(grid), dim3(64 * W), 0, st, G, scenes, slot_begin, ns, \
                       count, chunk, num_parts, part_begin, part_end, work)
#define MNRT_KB_W(C, S)                   \
    do {                                  \
        if (waves == 1) MNRT_KB(C, 1, S); \
        else if (waves == 2) MNRT_KB(C, 2, S); \
        else MNRT_KB(C, 4, S);            \
    } while (0)
    if (waves != 1 && waves != 2 && waves != 4) return MNRT_ERR_ARG;
    if (cap == 128) {
        if (sel == 2) MNRT_KB_W(128, kSelWindow);
        else if (sel == 1) MNRT_KB_W(128, kSelRadix);
        else MNRT_KB_W(128, kSelSort);
    } else if (cap == 256) {
        if (sel == 2) MNRT_KB_W(256, kSelWindow);
        else if (sel == 1) MNRT_KB_W(256, kSelRadix);
        else MNRT_KB_W(256, kSelSort);
    } else if (cap == 512) {
        if (sel == 2) MNRT_KB_W(512, kSelWindow);
        else if (sel == 1) MNRT_KB_W(512, kSelRadix);
        else MNRT_KB_W(512, kSelSort);
    } else {
        return MNRT_ERR_ARG;
    }
#undef MNRT_KB_W
#undef MNRT_KB
    MNRT_HIP_CHECK(hipGetLastError());
    return MNRT_OK;
}

// Backward of a batch: clears the accumulators, bounds and scales the weights per scene
// (from the forward's r2 and neighbour counts), runs the pair-merging walk and converts
// the sums.  G.out_r2 / G.out_count / G.out_nbr must be null (the walk would rewrite them).
int launch_knn_bwd_batch(const KnnArgs& G, const KnnScene* scenes, const uint32_t* slot_begin, uint32_t ns,
                         const uint32_t* count, const uint32_t* photon_begin, uint32_t total_photons,
                         const float* fwd_r2, const uint32_t* fwd_count, uint32_t* wmax, double* scale, uint32_t num_parts, const uint32_t* part_begin,
                         const uint32_t* part_end, uint32_t* work, uint32_t grid, float* grad_pow4, hipStream_t st) {
    const uint32_t n = G.n;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, work, num_parts, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, wmax, ns, 0u);
    const uint64_t words = 3ull * total_photons;
    if (words) {
        const uint32_t zb = (uint32_t)std::min<uint64_t>((words + 255) / 256, 8192);
        hipLaunchKernelGGL(k_zero_u64, dim3(zb), dim3(256), 0, st, G.acc, words);
    }
    if (n)
        hipLaunchKernelGGL(k_bwd_wmax_batch, dim3((n + 255) / 256), dim3(256), 0, st, n, slot_begin, ns, count,
                           G.grad, G.clr, G.infl, fwd_r2, fwd_count, wmax);
    hipLaunchKernelGGL(k_bwd_scale_batch, dim3(1), dim3(64), 0, st, ns, slot_begin, count, wmax, scale);
    const uint32_t chunk = std::max<uint32_t>(1u, std::min<uint32_t>(8u, (uint32_t)kPairCap / G.k));
    const uint32_t need = G.k + 64;
#define MNRT_KBB(C) \
    hipLaunchKernelGGL((k_knn_bwd_batch<C, kKnnSelDefault>), dim3(grid), dim3(64 * kWavesPerBlock), 0, st, G, scenes, \
                       slot_begin, ns, count, photon_begin, scale, chunk, num_parts, part_begin, part_end, work)
    if (need <= 128) MNRT_KBB(128);
    else if (need <= 256) MNRT_KBB(256);
    else MNRT_KBB(512);
#undef MNRT_KBB
    if (total_photons)
        hipLaunchKernelGGL(k_fix_to_float_batch, dim3((total_photons + 255) / 256), dim3(256), 0, st, total_photons,
                           photon_begin, ns, G.acc, scale, (float4*)grad_pow4);
    MNRT_HIP_CHECK(hipGetLastError());
    return MNRT_OK;
}
}  // namespace mnrt

namespace {

constexpr size_t kScratchWords = sizeof(scratch) / sizeof(uint32_t);
constexpr int kScratchMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_addr[kScratchMaxDevices] = {};

// Words of scratch mnrt_knn_gather_noalloc needs for n queries: the processing order, and
// 5 words per query (radiance + inverse distance) when slot indices are given.
size_t noalloc_words(uint32_t n, bool slots) {
    if (n >= mnrt::kOrderMaxItems) return SIZE_MAX;  // far beyond the array either way
    return mnrt::spatial_order_words(n) + (slots ? 5ull * n : 0ull);
}

}  // namespace

namespace mnrt {

// Looks up this device's address of `scratch` (the device must be current); kept per
// device, so later calls on it reuse the address.
int knn_scratch_init(int device) {
    if (device < 0 || device >= kScratchMaxDevices) return MNRT_ERR_NODEV;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_addr[device]) return MNRT_OK;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MNRT_ERR_HIP;
    g_scratch_addr[device] = (uint32_t*)p;
    return MNRT_OK;
}

}  // namespace mnrt

static int knn_check(const mnrt_tree* pmap, uint32_t k) {
    if (!pmap) return MNRT_ERR_ARG;
    if (k == 0 || k > MNRT_KNN_MAX_K) return MNRT_ERR_ARG;
    if (pmap->kind != MNRT_KIND_POINT || !pmap->d_nodes || !pmap->d_recs || pmap->num_nodes == 0)
        return MNRT_ERR_STATE;
    if (pmap->max_level > MNRT_MAX_LEVEL || pmap->num_nodes >= mnrt::kUniform8MaxIndex) return MNRT_ERR_CAPACITY;
    return MNRT_OK;
}

extern "C" int mnrt_knn_gather(const mnrt_tree* pmap, uint32_t k, float r_max, uint32_t n,
                               const float* d_pos4, const float* d_nrm4, const float* d_dir4,
                               const float* d_infl4, const float* d_clr4, const uint32_t* d_idx_sp,
                               const float* d_query_radius, const float* d_orig4, float* d_io_rad4,
                               float* d_io_inv_dist, float* d_out_r2, float* d_out_irr4, uint32_t* d_out_nbr,
                               uint32_t* d_out_count, void* stream) {
    int rc = knn_check(pmap, k);
    if (rc != MNRT_OK) return rc;
    if (n == 0) return MNRT_OK;
    if (!d_pos4 || !d_nrm4 || !d_dir4 || !d_infl4 || !d_io_rad4 || !(r_max > 0.0f)) return MNRT_ERR_ARG;
    if (d_io_inv_dist && !d_orig4) return MNRT_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    mnrt::KnnArgs A{};
    A.nodes = (const uint2*)(pmap->d_qnodes ? pmap->d_qnodes : pmap->d_nodes);
    A.recs = (const float4*)pmap->d_recs;
    A.k = k;
    A.r2max = r_max * r_max;
    A.r_max = r_max;
    A.n = n;
    A.pos = (const float4*)d_pos4;
    A.nrm = (const float4*)d_nrm4;
    A.dir = (const float4*)d_dir4;
    A.infl = (const float4*)d_infl4;
    A.clr = (const float4*)d_clr4;
    A.r_query = d_query_radius;
    A.orig = (const float4*)d_orig4;
    A.out_r2 = d_out_r2;
    A.out_irr = (float4*)d_out_irr4;
    A.out_nbr = d_out_nbr;
    A.out_count = d_out_count;
    void* obuf = nullptr;
    rc = mnrt::spatial_order(pmap->root_min, pmap->root_max, n, d_pos4, st, &A.order, &obuf, &A.order_far);
    if (rc != MNRT_OK) return rc;
    float4* tmp = nullptr;
    if (d_idx_sp) {
        // per-query values first, then summed into their slots run by run
        MNRT_HIP_CHECK(mnrt::pool_alloc((void**)&tmp, (size_t)n * (sizeof(float4) + sizeof(float)), st));
        A.rad_out = tmp;
        A.accumulate = 0;
        A.inv_dist = d_io_inv_dist ? (float*)(tmp + n) : nullptr;
        A.inv_accumulate = 0;
    } else {
        A.rad_out = (float4*)d_io_rad4;
        A.accumulate = 1;
        A.inv_dist = d_io_inv_dist;
        A.inv_accumulate = 1;
    }
    rc = mnrt::launch_knn<mnrt::kModeRadiance>(A, st);
    if (rc == MNRT_OK && d_idx_sp) {
        hipLaunchKernelGGL(mnrt::k_scatter_runs, dim3((n + 255) / 256), dim3(256), 0, st, tmp,
                           d_io_inv_dist ? (const float*)(tmp + n) : (const float*)nullptr, d_idx_sp, n,
                           (float4*)d_io_rad4, d_io_inv_dist);
        if (hipGetLastError() != hipSuccess) rc = MNRT_ERR_HIP;
    }
    if (tmp && mnrt::pool_free(tmp, st) != hipSuccess && rc == MNRT_OK) rc = MNRT_ERR_HIP;
    if (obuf && mnrt::pool_free(obuf, st) != hipSuccess && rc == MNRT_OK) rc = MNRT_ERR_HIP;
    return rc;
}

extern "C" int mnrt_knn_gather_noalloc(const mnrt_tree* pmap, uint32_t k, float r_max, uint32_t n,
                                       const float* d_pos4, const float* d_nrm4, const float* d_dir4,
                                       const float* d_infl4, const float* d_clr4, const uint32_t* d_idx_sp,
                                       const float* d_query_radius, const float* d_orig4, float* d_io_rad4,
                                       float* d_io_inv_dist, float* d_out_r2, float* d_out_irr4,
                                       uint32_t* d_out_nbr, uint32_t* d_out_count, void* stream) {
    int rc = knn_check(pmap, k);
    if (rc != MNRT_OK) return rc;
    if (n == 0) return MNRT_OK;
    if (!d_pos4 || !d_nrm4 || !d_dir4 || !d_infl4 || !d_io_rad4 || !(r_max > 0.0f)) return MNRT_ERR_ARG;
    if (d_io_inv_dist && !d_orig4) return MNRT_ERR_ARG;
    if (noalloc_words(n, d_idx_sp != nullptr) > kScratchWords) return MNRT_ERR_CAPACITY;
    int dev = 0;
    MNRT_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchMaxDevices) return MNRT_ERR_NODEV;
    uint32_t* scr = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        scr = g_scratch_addr[dev];
    }
    if (!scr) {  // mnrt_device_init was not called for this device
        rc = mnrt::knn_scratch_init(dev);
        if (rc != MNRT_OK) return rc;
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        scr = g_scratch_addr[dev];
    }
    hipStream_t st = (hipStream_t)stream;
    mnrt::KnnArgs A{};
    A.nodes = (const uint2*)(pmap->d_qnodes ? pmap->d_qnodes : pmap->d_nodes);
    A.recs = (const float4*)pmap->d_recs;
    A.k = k;
    A.r2max = r_max * r_max;
    A.r_max = r_max;
    A.n = n;
    A.pos = (const float4*)d_pos4;
    A.nrm = (const float4*)d_nrm4;
    A.dir = (const float4*)d_dir4;
    A.infl = (const float4*)d_infl4;
    A.clr = (const float4*)d_clr4;
    A.r_query = d_query_radius;
    A.orig = (const float4*)d_orig4;
    A.out_r2 = d_out_r2;
    A.out_irr = (float4*)d_out_irr4;
    A.out_nbr = d_out_nbr;
    A.out_count = d_out_count;
    // scratch layout: [processing order][per-query radiance (4 n)][per-query inverse distance (n)]
    rc = mnrt::spatial_order_into(pmap->root_min, pmap->root_max, n, d_pos4, st, scr, &A.order, &A.order_far);
    if (rc != MNRT_OK) return rc;
    float4* tmp = nullptr;
    if (d_idx_sp) {
        // every query's value is written before the run sums read it: nothing to clear
        tmp = (float4*)(scr + mnrt::spatial_order_words(n));
        A.rad_out = tmp;
        A.accumulate = 0;
        A.inv_dist = d_io_inv_dist ? (float*)(tmp + n) : nullptr;
        A.inv_accumulate = 0;
    } else {
        A.rad_out = (float4*)d_io_rad4;
        A.accumulate = 1;
        A.inv_dist = d_io_inv_dist;
        A.inv_accumulate = 1;
    }
    rc = mnrt::launch_knn<mnrt::kModeRadiance>(A, st);
    if (rc == MNRT_OK && d_idx_sp) {
        hipLaunchKernelGGL(mnrt::k_scatter_runs, dim3((n + 255) / 256), dim3(256), 0, st, tmp,
                           d_io_inv_dist ? (const float*)(tmp + n) : (const float*)nullptr, d_idx_sp, n,
                           (float4*)d_io_rad4, d_io_inv_dist);
        if (hipGetLastError() != hipSuccess) rc = MNRT_ERR_HIP;
    }
    return rc;
}

extern "C" int mnrt_knn_radii(const mnrt_tree* pmap, uint32_t k, float r_max, uint32_t n,
                              const float* d_pos4, float* d_out_radius, void* stream) {
    int rc = knn_check(pmap, k);
    if (rc != MNRT_OK) return rc;
    if (n == 0) return MNRT_OK;
    if (!d_pos4 || !d_out_radius || !(r_max > 0.0f)) return MNRT_ERR_ARG;
    mnrt::KnnArgs A{};
    A.nodes = (const uint2*)(pmap->d_qnodes ? pmap->d_qnodes : pmap->d_nodes);
    A.recs = (const float4*)pmap->d_recs;
    A.k = k;
    A.r2max = r_max * r_max;
    A.r_max = r_max;
    A.n = n;
    A.pos = (const float4*)d_pos4;
    A.out_radius = d_out_radius;
    hipStream_t st = (hipStream_t)stream;
    void* obuf = nullptr;
    rc = mnrt::spatial_order(pmap->root_min, pmap->root_max, n, d_pos4, st, &A.order, &obuf, &A.order_far);
    if (rc != MNRT_OK) return rc;
    rc = mnrt::launch_knn<mnrt::kModeRadii>(A, st);
    if (obuf && mnrt::pool_free(obuf, st) != hipSuccess && rc == MNRT_OK) rc = MNRT_ERR_HIP;
    return rc;
}

extern "C" int mnrt_knn_gather_bwd(const mnrt_tree* pmap, uint32_t k, float r_max, uint32_t n,
                                   const float* d_pos4, const float* d_nrm4, const float* d_dir4,
                                   const float* d_infl4, const float* d_clr4, const float* d_query_radius,
                                   const float* d_grad_rad4,
                                   const float* d_r2, const uint32_t* d_count, float* d_grad_pow4,
                                   float* d_grad_clr4, void* stream) {
    int rc = knn_check(pmap, k);
    if (rc != MNRT_OK) return rc;
    if (!d_grad_pow4) return MNRT_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    const uint32_t np = pmap->num_elems;
    if (n == 0 || np == 0) {
        if (np) MNRT_HIP_CHECK(hipMemsetAsync(d_grad_pow4, 0, (size_t)np * 16, st));
        return MNRT_OK;
    }
    if (!d_pos4 || !d_nrm4 || !d_dir4 || !d_infl4 || !d_grad_rad4 || !d_r2 || !d_count || !(r_max > 0.0f))
        return MNRT_ERR_ARG;
    void* buf = nullptr;
    MNRT_HIP_CHECK(mnrt::pool_alloc(&buf, (size_t)np * 24 + 256, st));
    unsigned long long* acc = (unsigned long long*)buf;
    uint32_t* wmax = (uint32_t*)((char*)buf + (size_t)np * 24);
    rc = MNRT_OK;
    uint32_t h_wmax = 0;
    if (hipMemsetAsync(buf, 0, (size_t)np * 24 + 4, st) != hipSuccess) rc = MNRT_ERR_HIP;
    if (rc == MNRT_OK) {
        hipLaunchKernelGGL(mnrt::k_bwd_wmax, dim3((n + 255) / 256), dim3(256), 0, st, n, (const float4*)d_grad_rad4,
                           (const float4*)d_clr4, (const float4*)d_infl4, d_r2, d_count, wmax);
        if (hipMemcpyAsync(&h_wmax, wmax, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = MNRT_ERR_HIP;
    }
    float wm = 0.0f;
    std::memcpy(&wm, &h_wmax, 4);
    double scale = 0.0;
    if (rc == MNRT_OK && wm > 0.0f && wm <= mnrt::kInf) {
        // |sum| <= n * wmax * scale < 2^62
        int e = 0;
        std::frexp(4.611686018427387904e18 / ((double)n * (double)wm), &e);
        scale = std::ldexp(1.0, e - 1);
    } else if (rc == MNRT_OK && wm > mnrt::kInf) {
        rc = MNRT_ERR_ARG;  // non-finite upstream gradient
    }
    const uint32_t* order = nullptr;
    const uint32_t* order_far = nullptr;
    void* obuf = nullptr;
    if (rc == MNRT_OK)
        rc = mnrt::spatial_order(pmap->root_min, pmap->root_max, n, d_pos4, st, &order, &obuf, &order_far);
    if (rc == MNRT_OK && scale > 0.0) {
        mnrt::KnnArgs A{};
        A.order = order;
        A.order_far = order_far;
        A.nodes = (const uint2*)(pmap->d_qnodes ? pmap->d_qnodes : pmap->d_nodes);
        A.recs = (const float4*)pmap->d_recs;
        A.k = k;
        A.r2max = r_max * r_max;
        A.r_max = r_max;
        A.n = n;
        A.pos = (const float4*)d_pos4;
        A.nrm = (const float4*)d_nrm4;
        A.dir = (const float4*)d_dir4;
        A.infl = (const float4*)d_infl4;
        A.clr = (const float4*)d_clr4;
        A.r_query = d_query_radius;
        A.grad = (const float4*)d_grad_rad4;
        A.acc = acc;
        A.fix_scale = scale;
        A.out_grad_clr = (float4*)d_grad_clr4;
        rc = mnrt::launch_knn_bwd(A, st);
    } else if (rc == MNRT_OK && d_grad_clr4) {
        // all weights zero: dLoss/drho still needs the recomputed irradiance
        mnrt::KnnArgs A{};
        A.order = order;
        A.nodes = (const uint2*)(pmap->d_qnodes ? pmap->d_qnodes : pmap->d_nodes);
        A.recs = (const float4*)pmap->d_recs;
        A.k = k;
        A.r2max = r_max * r_max;
        A.r_max = r_max;
        A.n = n;
        A.pos = (const float4*)d_pos4;
        A.nrm = (const float4*)d_nrm4;
        A.dir = (const float4*)d_dir4;
        A.infl = (const float4*)d_infl4;
        A.clr = (const float4*)d_clr4;
        A.r_query = d_query_radius;
        A.grad = (const float4*)d_grad_rad4;
        A.acc = acc;
        A.fix_scale = 0.0;
        A.out_grad_clr = (float4*)d_grad_clr4;
        rc = mnrt::launch_knn_bwd(A, st);
    }
    if (rc == MNRT_OK) {
        hipLaunchKernelGGL(mnrt::k_fix_to_float, dim3((np + 255) / 256), dim3(256), 0, st, np, acc,
                           scale > 0.0 ? 1.0 / scale : 0.0, (float4*)d_grad_pow4);
        if (hipGetLastError() != hipSuccess) rc = MNRT_ERR_HIP;
    }
    if (obuf) (void)mnrt::pool_free(obuf, st);
    (void)mnrt::pool_free(buf, st);
    return rc;
}
