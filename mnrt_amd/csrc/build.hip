// build.hip — kd-tree construction on gfx950 for triangles (SAH) and photons (VVH).
//
// Follows the two-stage build of KDTreeGPU::BuildTree (KDTreeGPU.cpp:175-209):
//
//   large-node stage (LargeNodeStage / ProcessLargeNodes, KDTreeGPU.cpp:247-782):
//     per level, for all active nodes at once:
//       chunk list + tight bounds   k_chunk_*, k_tight     (kdtree_chunklist.cu:73-184, kdtree.cu:310-445)
//       empty-space cuts + median   k_plan                 (kdtree.cu:467-682, 1986-2033)
//       element classification      k_classify             (kdtree.cu:752-824)
//       stream compaction           scan of L/R flags      (CompactElementData, KDTreeGPU.cpp:664-782)
//       small/large partition       k_child_plan/_write    (UpdateSmallList, KDTreeGPU.cpp:784-941)
//       split clipping + scatter    k_scatter              (kdtree_tri.cu:190-489)
//     The level loop keeps all per-element work on the device; the host reads five
//     counters once per level (the reference synchronises about ten times per level).
//
//   small-node stage (PreProcessSmallNodes / ProcessSmallNodes, KDTreeGPU.cpp:945-1154):
//     one wave64 per small root builds the whole subtree in LDS: split candidates and
//     their 64-bit element masks (kdtree.cu:1035-1251), exact SAH / VVH per node with a
//     wave argmin over (cost, candidate index) (kdtree.cu:1288-1471), preorder emission.
//
//   layout: subtree sizes bottom-up, addresses top-down (KDTreeGPU::PreorderTraversal,
//   KDTreeGPU.cpp:1156-1209), in one single-workgroup pass over the large-stage nodes.
//
// Deterministic rules replace the reference's random and racy choices (elements in a
// split plane go left for even index, right for odd; equal SAH costs resolve to the
// lowest candidate index), so the tree is identical, word for word, to the one the host
// oracle builds.
#include <algorithm>
#include <vector>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "batch_dev.h"
#include "scan.h"

namespace mnrt {

constexpr uint32_t kChunk = 256;
constexpr uint32_t kNone = 0xFFFFFFFFu;

enum : uint32_t { BN_LEAF = 3, BN_SMALL = 4, BN_PENDING = 5, BN_DEAD = 6 };
// counters[] slots written by device kernels
enum : uint32_t {
    C_NCHUNKS = 0,
    C_NEWNODES,
    C_NSMALL_NEW,
    C_NNEXT,
    C_ESMALL_NEW,
    C_ENEXT,
    C_MAXLEVEL,
    C_LLCOUNT,
    C_TOTAL_NODES,
    C_TOTAL_REFS,
    C_LEAVES,
    C_NLEAFCOPY,
    C_ERR,
    C_SMALL_OVF,  // small roots whose subtree outgrew its staging area
    C_COUNT
};

// Build-node table: one entry per node of the large-stage tree (inner nodes, empty
// and capped leaves, small roots), 8 uint32 columns of `cap` entries.  Build nodes
// 0..S-1 are the roots of the S scenes built together.
constexpr int kBnCols = 8;
struct BnView {
    uint32_t* type;   // 0..2 inner split axis, BN_LEAF, BN_SMALL, BN_PENDING, BN_DEAD
    uint32_t* split;  // split position bits (inner)
    uint32_t* left;
    uint32_t* right;
    uint32_t* level;
    uint32_t* aux;  // small-root index (BN_SMALL) or first copied element (BN_LEAF)
    uint32_t* cnt;  // element count (BN_LEAF)
    uint32_t* scene;
};

struct SmallView {  // small roots
    uint32_t* first;
    uint32_t* cnt;
    uint32_t* level;
    uint32_t* bn;
    float4* lo;
    float4* hi;
    uint32_t* scene;
};

struct ElemView {  // element arrays: original index, bounds, owning active node
    uint32_t* idx;
    float4* lo;
    float4* hi;
    uint32_t* node;
};

struct NodeView {  // active large nodes
    float4* lo;
    float4* hi;
    uint32_t* first;
    uint32_t* cnt;
    uint32_t* level;
    uint32_t* bn;
    uint32_t* scene;
};

// Concatenated input of S scenes: element e of scene s is entry base[s] + e; element
// indices inside a build are scene-local, as a single-scene build would number them.
struct Geometry {
    int kind;
    const float4* v0;
    const float4* v1;
    const float4* v2;
    const uint32_t* base;  // S + 1 offsets
    uint32_t num_scenes;
};

__device__ __forceinline__ uint32_t scene_of(const uint32_t* base, uint32_t nscenes, uint32_t g) {
    uint32_t lo = 0, hi = nscenes;  // last s with base[s] <= g
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (base[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t fenc(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fdec(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}
__device__ __forceinline__ float f4c(const float4& v, uint32_t a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
__device__ __forceinline__ void f4set(float4& v, uint32_t a, float x) {
    if (a == 0) v.x = x;
    else if (a == 1) v.y = x;
    else v.z = x;
}
// -0 -> +0 so min/max and stored split words never depend on the sign of zero
__device__ __forceinline__ float canon(float x) { return x + 0.0f; }

__device__ __forceinline__ bool plane_goes_left(uint32_t idx) { return (idx & 1u) == 0u; }

// Large-node classification (kernel_MarkLeftRightElements, kdtree.cu:752-824) with the
// even/odd rule for elements in the split plane.
__device__ __forceinline__ void classify(int kind, uint32_t idx, const float4& lo, const float4& hi,
                                         uint32_t axis, float s, bool& L, bool& R) {
    if (kind == MNRT_KIND_TRI) {
        float mn = f4c(lo, axis), mx = f4c(hi, axis);
        bool planar_on = (mn == s && mn == mx);
        L = mn < s || (planar_on && plane_goes_left(idx));
        R = s < mx || (planar_on && !plane_goes_left(idx));
    } else {
        float p = f4c(lo, axis);
        if (plane_goes_left(idx)) {
            L = p <= s;
            R = s < p;
        } else {
            L = p < s;
            R = s <= p;
        }
    }
}

// ---- split clipping (dev_ClipSegment / dev_ClipTriToBounds, kdtree_tri.cu:190-316) ----

// Vertex lists of the clipping in LDS (a per-lane array indexed by the running vertex
// count would live in scratch memory): per lane two lists of up to kClipVerts vertices,
// coordinate-major with the lane innermost, so the lanes of a wave touch consecutive banks.
// k_clip runs kClipBlock lanes per workgroup.
constexpr int kClipBlock = 64;
constexpr int kClipVerts = 12;  // a triangle clipped by six planes has at most 9

struct ClipLists {
    __attribute__((address_space(3))) float* col;  // this lane's column
    int n[2];
    __device__ __forceinline__ float get(int l, int v, int c) const {
        return col[((l * kClipVerts + v) * 3 + c) * kClipBlock];
    }
    __device__ __forceinline__ f3 vert(int l, int v) const { return mk3(get(l, v, 0), get(l, v, 1), get(l, v, 2)); }
    __device__ __forceinline__ void push(int l, f3 p) {
        const int v = n[l]++;
        if (v >= kClipVerts) return;  // (cannot happen: at most 9 vertices)
        col[((l * kClipVerts + v) * 3 + 0) * kClipBlock] = p.x;
        col[((l * kClipVerts + v) * 3 + 1) * kClipBlock] = p.y;
        col[((l * kClipVerts + v) * 3 + 2) * kClipBlock] = p.z;
    }
};

__device__ __forceinline__ float f3c(const f3& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
__device__ __forceinline__ void f3set(f3& v, int a, float x) {
    if (a == 0) v.x = x;
    else if (a == 1) v.y = x;
    else v.z = x;
}

// Sutherland-Hodgman step: list `in` clipped by the plane x[axis] = val, keeping the side
// where (x[axis] - val) * sense >= 0, into list `out`.
__device__ void clip_segments(ClipLists& L, int in, int axis, float val, float sense, int out) {
    L.n[out] = 0;
    const int n = min(L.n[in], kClipVerts);
    for (int i = 0; i < n; ++i) {
        f3 a = L.vert(in, i);
        f3 b = L.vert(in, (i + 1 == n) ? 0 : i + 1);
        bool ins = (f3c(a, axis) - val) * sense >= 0.0f;
        bool ine = (f3c(b, axis) - val) * sense >= 0.0f;
        if (ins != ine) {
            float t = (val - f3c(a, axis)) / (f3c(b, axis) - f3c(a, axis));
            f3 h = mk3(a.x + t * (b.x - a.x), a.y + t * (b.y - a.y), a.z + t * (b.z - a.z));
            f3set(h, axis, val);
            L.push(out, h);
        }
        if (ine) L.push(out, b);
    }
}

// Clipped bounds of a straddling triangle inside a child box, padded outward by 2^-20 of
// the largest magnitude involved, clamped to the child box and to the previous bounds.
// col: this lane's LDS column (2 * kClipVerts * 3 words at stride kClipBlock).
__device__ void clip_bounds(const Geometry& G, uint32_t idx, float4& lo, float4& hi, const float4& bmin,
                            const float4& bmax, __attribute__((address_space(3))) float* col) {
    ClipLists L;
    L.col = col;
    L.n[0] = L.n[1] = 0;
    L.push(0, xyz(G.v0[idx]));
    L.push(0, xyz(G.v1[idx]));
    L.push(0, xyz(G.v2[idx]));
    for (int ax = 0; ax < 3; ++ax) {
        if (f4c(lo, ax) != f4c(hi, ax)) {
            clip_segments(L, 0, ax, f4c(bmin, ax), 1.0f, 1);
            clip_segments(L, 1, ax, f4c(bmax, ax), -1.0f, 0);
        }
    }
    const int na = min(L.n[0], kClipVerts);
    for (int ax = 0; ax < 3; ++ax) {
        float l, h;
        float bl = f4c(bmin, ax), bh = f4c(bmax, ax);
        if (na == 0) {
            l = f4c(lo, ax);
            h = f4c(hi, ax);
        } else {
            l = kInf;
            h = -kInf;
            for (int i = 0; i < na; ++i) {
                l = fminf(L.get(0, i, ax), l);
                h = fmaxf(L.get(0, i, ax), h);
            }
            float m = fmaxf(fmaxf(fabsf(l), fabsf(h)), fmaxf(fabsf(bl), fabsf(bh)));
            float pad = m * (1.0f / 1048576.0f);
            l = l - pad;
            h = h + pad;
        }
        l = fmaxf(l, bl);
        h = fminf(h, bh);
        l = fmaxf(l, f4c(lo, ax));
        h = fminf(h, f4c(hi, ax));
        f4set(lo, ax, canon(l));
        f4set(hi, ax, canon(h));
    }
}

// ---------------------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------------------
// Counts come as DN (host value or device word).  Per-item kernels stride over the grid,
// so any grid is correct: the host-driven build launches one item per thread, the
// recorded device-driven build launches grids sized from the first build's counts.
#define MNRT_GRID_STRIDE(i, n) \
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += gridDim.x * blockDim.x)

__global__ __launch_bounds__(256) void k_init_elems(Geometry G, uint32_t n, ElemView E) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = scene_of(G.base, G.num_scenes, i);
    float4 a = G.v0[i];
    float4 lo, hi;
    if (G.kind == MNRT_KIND_TRI) {
        float4 b = G.v1[i], c = G.v2[i];
        lo = make_float4(canon(fminf(a.x, fminf(b.x, c.x))), canon(fminf(a.y, fminf(b.y, c.y))),
                         canon(fminf(a.z, fminf(b.z, c.z))), 0.0f);
        hi = make_float4(canon(fmaxf(a.x, fmaxf(b.x, c.x))), canon(fmaxf(a.y, fmaxf(b.y, c.y))),
                         canon(fmaxf(a.z, fmaxf(b.z, c.z))), 0.0f);
    } else {
        lo = hi = make_float4(canon(a.x), canon(a.y), canon(a.z), 0.0f);
    }
    E.idx[i] = i - G.base[s];
    E.lo[i] = lo;
    E.hi[i] = hi;
    E.node[i] = s;
}

// Chunks per node, and the node's tight-bound accumulators (3 per node, ordered-int
// encoding) cleared for k_tight.
__global__ __launch_bounds__(256) void k_chunk_counts(DN dn_, const uint32_t* __restrict__ cnt,
                                                      uint32_t* __restrict__ cc, uint32_t* __restrict__ tlo,
                                                      uint32_t* __restrict__ thi) {
    const uint32_t n = dn_.get();
    MNRT_GRID_STRIDE(p, n) {
        cc[p] = (cnt[p] + kChunk - 1) / kChunk;
#pragma unroll
        for (uint32_t a = 0; a < 3; ++a) {
            tlo[3 * p + a] = 0xFFFFFFFFu;
            thi[3 * p + a] = 0u;
        }
    }
}

// Chunk -> node by binary search over the chunk offsets (kernel_GenerateChunks,
// kdtree_chunklist.cu:110-134, runs one serial thread per node instead).
__global__ __launch_bounds__(256) void k_chunk_fill(DN dn_, uint32_t max_chunks, const uint32_t* __restrict__ coff,
                                                    const uint32_t* __restrict__ counters,
                                                    const uint32_t* __restrict__ first,
                                                    const uint32_t* __restrict__ cnt, uint32_t* __restrict__ ch_node,
                                                    uint32_t* __restrict__ ch_first, uint32_t* __restrict__ ch_cnt) {
    const uint32_t n = dn_.get();
    const uint32_t nch = min(max_chunks, counters[C_NCHUNKS]);
    MNRT_GRID_STRIDE(c, nch) {
        uint32_t lo = 0, hi = n;  // last p with coff[p] <= c
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (coff[mid] <= c) lo = mid;
            else hi = mid;
        }
        uint32_t p = lo;
        uint32_t local = c - coff[p];
        ch_node[c] = p;
        ch_first[c] = first[p] + local * kChunk;
        ch_cnt[c] = min(kChunk, cnt[p] - local * kChunk);
    }
}

// Tight node bounds: per-chunk wave/LDS min/max (kernel_GenChunkAABB + two segmented
// reductions, kdtree.cu:310-445, KDTreeGPU.cpp:526-553).  A workgroup reduces
// kTightChunks consecutive chunks and issues one atomic per component each time the
// owning node changes (chunks are ordered by node, so a large node's many chunks cost
// few atomics).  min/max are exact, so the order of the atomics does not matter.
constexpr uint32_t kTightChunks = 8;

__global__ __launch_bounds__(kChunk) void k_tight(uint32_t max_chunks, const uint32_t* __restrict__ counters,
                                                  const uint32_t* __restrict__ ch_node,
                                                  const uint32_t* __restrict__ ch_first,
                                                  const uint32_t* __restrict__ ch_cnt, ElemView E, int kind,
                                                  uint32_t* __restrict__ tlo, uint32_t* __restrict__ thi) {
    __shared__ uint32_t s_red[6][kChunk / 64];
    const uint32_t nch = min(max_chunks, counters[C_NCHUNKS]);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (uint32_t c0 = blockIdx.x * kTightChunks; c0 < nch; c0 += gridDim.x * kTightChunks) {  // uniform per block
        const uint32_t c1 = min(c0 + kTightChunks, nch);
        uint32_t cur = kNone;  // node of the running accumulation (thread 0)
        uint32_t acc[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
        for (uint32_t c = c0; c < c1; ++c) {
            const uint32_t node = ch_node[c], first = ch_first[c], count = ch_cnt[c];
            uint32_t v[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
            if (threadIdx.x < count) {
                float4 lo = E.lo[first + threadIdx.x];
                float4 hi = (kind == MNRT_KIND_TRI) ? E.hi[first + threadIdx.x] : lo;
                v[0] = fenc(lo.x);
                v[1] = fenc(lo.y);
                v[2] = fenc(lo.z);
                v[3] = fenc(hi.x);
                v[4] = fenc(hi.y);
                v[5] = fenc(hi.z);
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) v[k] = wave_min_u32(v[k]);
#pragma unroll
            for (int k = 3; k < 6; ++k) v[k] = wave_max_u32(v[k]);
            if (lane == 0)
                for (int k = 0; k < 6; ++k) s_red[k][wave] = v[k];
            __syncthreads();
            if (threadIdx.x == 0) {
                for (uint32_t w = 1; w < kChunk / 64; ++w) {
                    for (int k = 0; k < 3; ++k) v[k] = min(v[k], s_red[k][w]);
                    for (int k = 3; k < 6; ++k) v[k] = max(v[k], s_red[k][w]);
                }
                if (node != cur && cur != kNone) {
                    for (int k = 0; k < 3; ++k) atomicMin(&tlo[3 * cur + k], acc[k]);
                    for (int k = 0; k < 3; ++k) atomicMax(&thi[3 * cur + k], acc[3 + k]);
                    for (int k = 0; k < 3; ++k) acc[k] = 0xFFFFFFFFu;
                    for (int k = 3; k < 6; ++k) acc[k] = 0u;
                }
                cur = node;
                for (int k = 0; k < 3; ++k) acc[k] = min(acc[k], v[k]);
                for (int k = 3; k < 6; ++k) acc[k] = max(acc[k], v[k]);
            }
            __syncthreads();  // s_red is rewritten by the next chunk
        }
        if (threadIdx.x == 0 && cur != kNone) {
            for (int k = 0; k < 3; ++k) atomicMin(&tlo[3 * cur + k], acc[k]);
            for (int k = 0; k < 3; ++k) atomicMax(&thi[3 * cur + k], acc[3 + k]);
        }
    }
}

// Per-scene placement decided on the host: an active (large) root or a small root, its
// index in that list and its first element there.
struct RootPlan {
    const uint32_t* kind;   // 0 active, 1 small
    const uint32_t* index;
    const uint32_t* off;
};

// Root boxes (given, or the tight bounds of the scene's elements; (+max, -max) for an
// empty scene) and the root entries of the active / small lists.
__global__ __launch_bounds__(256) void k_roots_setup(uint32_t nscenes, const uint32_t* __restrict__ base,
                                                     const uint32_t* __restrict__ tlo,
                                                     const uint32_t* __restrict__ thi,
                                                     const float* __restrict__ given, RootPlan R, NodeView A,
                                                     SmallView S, BnView B, float* __restrict__ out_box) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nscenes) return;
    const uint32_t cnt = base[s + 1] - base[s];
    float4 lo, hi;
    if (given) {
        lo = make_float4(canon(given[6 * s + 0]), canon(given[6 * s + 1]), canon(given[6 * s + 2]), 0.0f);
        hi = make_float4(canon(given[6 * s + 3]), canon(given[6 * s + 4]), canon(given[6 * s + 5]), 0.0f);
    } else if (cnt == 0) {
        lo = make_float4(kInf, kInf, kInf, 0.0f);
        hi = make_float4(-kInf, -kInf, -kInf, 0.0f);
    } else {
        lo = make_float4(fdec(tlo[3 * s]), fdec(tlo[3 * s + 1]), fdec(tlo[3 * s + 2]), 0.0f);
        hi = make_float4(fdec(thi[3 * s]), fdec(thi[3 * s + 1]), fdec(thi[3 * s + 2]), 0.0f);
    }
    B.level[s] = 0;
    B.scene[s] = s;
    const uint32_t i = R.index[s];
    if (R.kind[s] == 0) {
        A.lo[i] = lo;
        A.hi[i] = hi;
        A.first[i] = R.off[s];
        A.cnt[i] = cnt;
        A.level[i] = 0;
        A.bn[i] = s;
        A.scene[i] = s;
        B.type[s] = BN_PENDING;
    } else {
        S.first[i] = R.off[s];
        S.cnt[i] = cnt;
        S.level[i] = 0;
        S.bn[i] = s;
        S.lo[i] = lo;
        S.hi[i] = hi;
        S.scene[i] = s;
        B.type[s] = BN_SMALL;
        B.aux[s] = i;
    }
    out_box[6 * s + 0] = lo.x;
    out_box[6 * s + 1] = lo.y;
    out_box[6 * s + 2] = lo.z;
    out_box[6 * s + 3] = hi.x;
    out_box[6 * s + 4] = hi.y;
    out_box[6 * s + 5] = hi.z;
}

// Elements of large scenes go to the first level's arrays, those of small scenes to the
// small-root buffer; each keeps its scene order.
__global__ __launch_bounds__(256) void k_partition_elems(uint32_t n, ElemView In, int kind, RootPlan R, ElemView E,
                                                         ElemView Sm) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const uint32_t s = In.node[g], local = In.idx[g];
    const uint32_t d = R.off[s] + local;
    float4 lo = In.lo[g];
    if (R.kind[s] == 0) {
        E.idx[d] = local;
        E.lo[d] = lo;
        if (kind == MNRT_KIND_TRI) E.hi[d] = In.hi[g];
        E.node[d] = R.index[s];
    } else {
        Sm.idx[d] = local;
        Sm.lo[d] = lo;
        if (kind == MNRT_KIND_TRI) Sm.hi[d] = In.hi[g];
    }
}

struct PlanView {
    uint32_t* newcnt;  // build nodes created per active node (count phase) / base (write phase)
    uint32_t* axis;
    float* split;
    float4* lo;        // box after the empty-space cuts
    float4* hi;
    uint32_t* level;   // level of the node that is split (after cuts)
    uint32_t* childL;  // build-node id of the left child (right = +1)
    uint32_t* cur;     // build-node id of the node that is split or becomes a leaf
    uint32_t* leaf;    // 1: level cap reached, the node is a leaf with its elements
    uint32_t* noprog;  // 1: every element went to both children, the node is a leaf
};

// Empty-space cuts then the spatial median split (kernel_CanCutOffEmptySpace,
// kernel_EmptySpaceCutting, kernel_SplitLargeNodes; kdtree.cu:467-682).  Sides are cut in
// the reference's order (min x, y, z, then max x, y, z); each cut turns the node into an
// inner node with an empty leaf and a tighter copy.  The median splits the longest axis
// of the cut box, ties to the lower axis.
template <bool kWrite>
__device__ __forceinline__ void plan_node(uint32_t p, const NodeView& A, const uint32_t* __restrict__ tlo,
                                          const uint32_t* __restrict__ thi, float ratio, const PlanView& P,
                                          const uint32_t* __restrict__ base, uint32_t bn_base, const BnView& B,
                                          uint32_t* __restrict__ counters) {
    float lo[3], hi[3], tl[3], th[3];
    float4 l4 = A.lo[p], h4 = A.hi[p];
    lo[0] = l4.x; lo[1] = l4.y; lo[2] = l4.z;
    hi[0] = h4.x; hi[1] = h4.y; hi[2] = h4.z;
    for (int k = 0; k < 3; ++k) {
        tl[k] = fdec(tlo[3 * p + k]);
        th[k] = fdec(thi[3 * p + k]);
    }
    uint32_t level = A.level[p];
    uint32_t cur = A.bn[p];
    uint32_t next_id = kWrite ? bn_base + base[p] : 0u;
    uint32_t ncuts = 0;
    for (int isMax = 0; isMax < 2; ++isMax) {
        for (int ax = 0; ax < 3; ++ax) {
            if (level >= MNRT_MAX_LEVEL) continue;
            float total = hi[ax] - lo[ax];
            float empty = isMax ? (hi[ax] - th[ax]) : (tl[ax] - lo[ax]);
            if (!(empty > total * ratio)) continue;
            float s = isMax ? th[ax] : tl[ax];
            ++ncuts;
            if (kWrite) {
                uint32_t eleaf = next_id++, cont = next_id++;
                B.type[cur] = (uint32_t)ax;
                B.split[cur] = __float_as_uint(s);
                B.left[cur] = isMax ? cont : eleaf;
                B.right[cur] = isMax ? eleaf : cont;
                B.type[eleaf] = BN_LEAF;
                B.cnt[eleaf] = 0;
                B.aux[eleaf] = 0;
                B.level[eleaf] = level + 1;
                B.scene[eleaf] = A.scene[p];
                B.type[cont] = BN_PENDING;
                B.level[cont] = level + 1;
                B.scene[cont] = A.scene[p];
                cur = cont;
            }
            if (isMax) hi[ax] = s;
            else lo[ax] = s;
            level += 1;
        }
    }
    if (!kWrite) {
        P.newcnt[p] = 2 * ncuts + (level >= MNRT_MAX_LEVEL ? 0u : 2u);
        return;
    }
    atomicMax(&counters[C_MAXLEVEL], level);
    P.cur[p] = cur;
    P.level[p] = level;
    P.lo[p] = make_float4(lo[0], lo[1], lo[2], 0.0f);
    P.hi[p] = make_float4(hi[0], hi[1], hi[2], 0.0f);
    P.noprog[p] = 0;
    if (level >= MNRT_MAX_LEVEL) {
        P.leaf[p] = 1;
        P.axis[p] = 0;
        P.split[p] = 0.0f;
        P.childL[p] = kNone;
        B.type[cur] = BN_LEAF;
        return;
    }
    float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    uint32_t axis = 0;
    float emax = ex;
    if (ey > emax) { axis = 1; emax = ey; }
    if (ez > emax) axis = 2;
    float s = lo[axis] + 0.5f * (hi[axis] - lo[axis]);
    uint32_t L = next_id++, R = next_id++;
    P.leaf[p] = 0;
    P.axis[p] = axis;
    P.split[p] = s;
    P.childL[p] = L;
    B.type[cur] = axis;
    B.split[cur] = __float_as_uint(s);
    B.left[cur] = L;
    B.right[cur] = R;
    B.level[L] = level + 1;
    B.level[R] = level + 1;
    B.type[L] = BN_PENDING;
    B.type[R] = BN_PENDING;
    B.scene[L] = A.scene[p];
    B.scene[R] = A.scene[p];
}

template <bool kWrite>
__global__ __launch_bounds__(256) void k_plan(DN dn_, NodeView A, const uint32_t* __restrict__ tlo,
                                              const uint32_t* __restrict__ thi, float ratio, PlanView P,
                                              const uint32_t* __restrict__ base, DN bn_base, BnView B,
                                              uint32_t* __restrict__ counters) {
    const uint32_t n = dn_.get(), b0 = bn_base.get();
    MNRT_GRID_STRIDE(p, n) plan_node<kWrite>(p, A, tlo, thi, ratio, P, base, b0, B, counters);
}

// Flags per element: left child, right child and, for triangles, straddling the split on
// the left / right side (fl[2 ne + e], fl[3 ne + e]); one scan over them gives each element
// side its child rank and each straddling side its slot in the clip queue.
__global__ __launch_bounds__(256) void k_classify(DN dne, ElemView E, int kind, PlanView P,
                                                  uint32_t* __restrict__ fl) {
    const uint32_t ne = dne.get();
    MNRT_GRID_STRIDE(e, ne) {
        uint32_t p = E.node[e];
        bool L = false, R = false, straddle = false;
        if (!P.leaf[p]) {
            float4 lo = E.lo[e];
            float4 hi = (kind == MNRT_KIND_TRI) ? E.hi[e] : lo;
            const uint32_t axis = P.axis[p];
            const float s = P.split[p];
            classify(kind, E.idx[e], lo, hi, axis, s, L, R);
            straddle = (kind == MNRT_KIND_TRI) && (f4c(lo, axis) < s && s < f4c(hi, axis));
        }
        fl[e] = L ? 1u : 0u;
        fl[ne + e] = R ? 1u : 0u;
        if (kind == MNRT_KIND_TRI) {
            fl[2 * ne + e] = (straddle && L) ? 1u : 0u;
            fl[3 * ne + e] = (straddle && R) ? 1u : 0u;
        }
    }
}

struct ChildView {
    uint32_t* type;  // BN_SMALL, BN_PENDING (large), BN_DEAD
    uint32_t* cnt;
    uint32_t* fpos;  // flag-scan position of the child's first element
    uint32_t* sflag;
    uint32_t* lflag;
    uint32_t* scnt;
    uint32_t* lcnt;
    uint32_t* sidx;  // scans of the four columns above
    uint32_t* lidx;
    uint32_t* soff;
    uint32_t* loff;
};

// Child counts from the flag scan, the no-progress test, and the small/large split
// (kernel_MarkSmallNodes, kdtree.cu:849-875: small iff count <= smallNodeMax).
__global__ __launch_bounds__(256) void k_child_plan(DN dn_, DN dne, NodeView A, PlanView P,
                                                    const uint32_t* __restrict__ pos, uint32_t small_max,
                                                    ChildView Cv) {
    const uint32_t n = dn_.get(), ne = dne.get();
    MNRT_GRID_STRIDE(c, 2 * n) {
        uint32_t side = c >= n ? 1u : 0u;
        uint32_t p = c - side * n;
        uint32_t type = BN_DEAD, cnt = 0, fpos = 0;
        if (!P.leaf[p]) {
            uint32_t f = A.first[p], m = A.cnt[p];
            uint32_t nl = pos[f + m] - pos[f];
            uint32_t nr = pos[ne + f + m] - pos[ne + f];
            bool noprog = (nl == m && nr == m);
            if (noprog) {
                if (side == 0) P.noprog[p] = 1;
            } else {
                fpos = pos[side * ne + f];
                cnt = side ? nr : nl;
                type = (cnt <= small_max) ? BN_SMALL : BN_PENDING;
            }
        }
        Cv.type[c] = type;
        Cv.cnt[c] = cnt;
        Cv.fpos[c] = fpos;
        Cv.sflag[c] = type == BN_SMALL ? 1u : 0u;
        Cv.lflag[c] = type == BN_PENDING ? 1u : 0u;
        Cv.scnt[c] = type == BN_SMALL ? cnt : 0u;
        Cv.lcnt[c] = type == BN_PENDING ? cnt : 0u;
    }
}

__device__ __forceinline__ void child_write_one(uint32_t c, uint32_t n, const NodeView& A, const PlanView& P,
                                                const ChildView& Cv, uint32_t small_base, uint32_t small_elem_base,
                                                const SmallView& S, const NodeView& Nx, const BnView& B,
                                                uint32_t* __restrict__ counters) {
    uint32_t side = c >= n ? 1u : 0u;
    uint32_t p = c - side * n;
    if (P.leaf[p]) return;
    if (P.noprog[p]) {
        if (side == 0) {
            B.type[P.cur[p]] = BN_LEAF;
            B.type[P.childL[p]] = BN_DEAD;
            B.type[P.childL[p] + 1] = BN_DEAD;
        }
        return;
    }
    uint32_t bn = P.childL[p] + side;
    uint32_t level = P.level[p] + 1;
    atomicMax(&counters[C_MAXLEVEL], level);
    float4 lo = P.lo[p], hi = P.hi[p];
    if (side == 0) f4set(hi, P.axis[p], P.split[p]);
    else f4set(lo, P.axis[p], P.split[p]);
    uint32_t type = Cv.type[c];
    if (type == BN_SMALL) {
        uint32_t s = small_base + Cv.sidx[c];
        S.first[s] = small_elem_base + Cv.soff[c];
        S.cnt[s] = Cv.cnt[c];
        S.level[s] = level;
        S.bn[s] = bn;
        S.lo[s] = lo;
        S.hi[s] = hi;
        S.scene[s] = A.scene[p];
        B.type[bn] = BN_SMALL;
        B.aux[bn] = s;
    } else {
        uint32_t l = Cv.lidx[c];
        Nx.lo[l] = lo;
        Nx.hi[l] = hi;
        Nx.first[l] = Cv.loff[c];
        Nx.cnt[l] = Cv.cnt[c];
        Nx.level[l] = level;
        Nx.bn[l] = bn;
        Nx.scene[l] = A.scene[p];
        B.type[bn] = BN_PENDING;
    }
}

// Elements of large-stage leaves (level cap, no progress) are copied out before the
// level's element arrays are released.
__device__ __forceinline__ void leaf_copy_one(uint32_t p, const NodeView& A, const PlanView& P, const ElemView& E,
                                              uint32_t* __restrict__ ll_idx, uint32_t ll_cap, const BnView& B,
                                              uint32_t* __restrict__ counters) {
    if (!(P.leaf[p] || P.noprog[p])) return;
    uint32_t m = A.cnt[p], f = A.first[p];
    uint32_t start = atomicAdd(&counters[C_LLCOUNT], m);
    if ((uint64_t)start + m > ll_cap) {
        atomicOr(&counters[C_ERR], 32u);
        return;
    }
    for (uint32_t i = 0; i < m; ++i) ll_idx[start + i] = E.idx[f + i];
    uint32_t cur = P.cur[p];
    B.aux[cur] = start;
    B.cnt[cur] = m;
}

// The level's children (2 per node), and with the first n threads the copy of large-stage
// leaves' elements.
__global__ __launch_bounds__(256) void k_child_write(DN dn_, NodeView A, PlanView P, ChildView Cv, DN small_base,
                                                     DN small_elem_base, SmallView S, NodeView Nx, BnView B,
                                                     uint32_t* __restrict__ counters, ElemView E,
                                                     uint32_t* __restrict__ ll_idx, uint32_t ll_cap) {
    const uint32_t n = dn_.get(), sb = small_base.get(), seb = small_elem_base.get();
    MNRT_GRID_STRIDE(c, 2 * n) {
        child_write_one(c, n, A, P, Cv, sb, seb, S, Nx, B, counters);
        if (c < n) leaf_copy_one(c, A, P, E, ll_idx, ll_cap, B, counters);
    }
}

// Scatter each element into its children: small roots append to the small-element
// buffer, large children to the next level's arrays.  A straddling triangle's bounds in a
// child are its clip to the child box (kernel_PerformSplitClipping, kdtree_tri.cu:389-489),
// which k_clip computes: each straddling element side has a slot in the clip queue from
// the classification scan (pos[2 ne + side ne + e] - pos[2 ne]), where this writes the
// side's job (e << 1 | side), or kNone when the side is not scattered, so that the clips
// run 64 to a wave rather than one pass of a whole wave per straddler.
__device__ __forceinline__ void scatter_one(uint32_t e, uint32_t n, uint32_t ne, const ElemView& E,
                                            const Geometry& G, const NodeView& A, const PlanView& P,
                                            const ChildView& Cv, const uint32_t* __restrict__ pos,
                                            uint32_t small_elem_base, const ElemView& Sm, uint32_t sm_cap,
                                            const ElemView& Nx, uint32_t nx_cap, uint32_t* __restrict__ counters,
                                            uint32_t* __restrict__ clipq) {
    const uint32_t p = E.node[e];
    if (P.leaf[p]) return;  // no flags: k_classify skips leaves
    const uint32_t idx = E.idx[e];
    const float4 lo = E.lo[e];
    const float4 hi = (G.kind == MNRT_KIND_TRI) ? E.hi[e] : lo;
    const uint32_t axis = P.axis[p];
    const float s = P.split[p];
    bool L, R;
    classify(G.kind, idx, lo, hi, axis, s, L, R);
    const bool straddle = (G.kind == MNRT_KIND_TRI) && (f4c(lo, axis) < s && s < f4c(hi, axis));
    const bool noprog = P.noprog[p] != 0u;  // its elements were copied out by k_child_write
    for (uint32_t side = 0; side < 2; ++side) {
        if (!(side ? R : L)) continue;
        bool queued = false;
        const uint32_t c = side * n + p;
        const uint32_t type = noprog ? (uint32_t)BN_DEAD : Cv.type[c];
        if (type != BN_DEAD) {
            const uint32_t rank = pos[side * ne + e] - Cv.fpos[c];
            if (type == BN_SMALL) {
                const uint32_t d = small_elem_base + Cv.soff[c] + rank;
                if (d >= sm_cap) {
                    atomicOr(&counters[C_ERR], 1u);
                } else {
                    Sm.idx[d] = idx;
                    if (!straddle) {
                        Sm.lo[d] = lo;
                        if (G.kind == MNRT_KIND_TRI) Sm.hi[d] = hi;
                    }
                    queued = straddle;
                }
            } else {
                const uint32_t d = Cv.loff[c] + rank;
                if (d >= nx_cap) {
                    atomicOr(&counters[C_ERR], 2u);
                } else {
                    Nx.idx[d] = idx;
                    Nx.node[d] = Cv.lidx[c];
                    if (!straddle) {
                        Nx.lo[d] = lo;
                        if (G.kind == MNRT_KIND_TRI) Nx.hi[d] = hi;
                    }
                    queued = straddle;
                }
            }
        }
        if (straddle) clipq[pos[2 * ne + side * ne + e] - pos[2 * ne]] = queued ? (e << 1 | side) : kNone;
    }
}

__global__ __launch_bounds__(256) void k_scatter(DN dn_, DN dne, ElemView E, Geometry G, NodeView A, PlanView P,
                                                 ChildView Cv, const uint32_t* __restrict__ pos, DN small_elem_base,
                                                 ElemView Sm, uint32_t sm_cap, ElemView Nx, uint32_t nx_cap,
                                                 uint32_t* __restrict__ counters, uint32_t* __restrict__ clipq) {
    const uint32_t n = dn_.get(), ne = dne.get(), seb = small_elem_base.get();
    MNRT_GRID_STRIDE(e, ne)
    scatter_one(e, n, ne, E, G, A, P, Cv, pos, seb, Sm, sm_cap, Nx, nx_cap, counters, clipq);
}

// The clip queue of a triangle level (k_scatter): each job clips its triangle to the
// child box and writes the bounds next to the index k_scatter placed.  A grid-stride
// launch: the queue holds at most 2 ne jobs, usually a small part of that.
__global__ __launch_bounds__(kClipBlock) void k_clip(DN dn_, DN dne, ElemView E, Geometry G, NodeView A, PlanView P,
                                                     ChildView Cv, const uint32_t* __restrict__ pos,
                                                     DN small_elem_base, ElemView Sm, ElemView Nx,
                                                     const uint32_t* __restrict__ clipq) {
    __shared__ float s_lists[2 * kClipVerts * 3 * kClipBlock];
    __attribute__((address_space(3))) float* col = (__attribute__((address_space(3))) float*)(s_lists + threadIdx.x);
    const uint32_t n = dn_.get(), ne = dne.get(), seb = small_elem_base.get();
    const uint32_t nq = pos[4 * ne] - pos[2 * ne];
    MNRT_GRID_STRIDE(j, nq) {
        const uint32_t w = clipq[j];
        if (w == kNone) continue;
        const uint32_t e = w >> 1, side = w & 1u, p = E.node[e], axis = P.axis[p];
        const float s = P.split[p];
        float4 lo = E.lo[e], hi = E.hi[e];
        float4 bmin = P.lo[p], bmax = P.hi[p];
        if (side == 0) f4set(bmax, axis, s);
        else f4set(bmin, axis, s);
        clip_bounds(G, G.base[A.scene[p]] + E.idx[e], lo, hi, bmin, bmax, col);
        const uint32_t c = side * n + p, rank = pos[side * ne + e] - Cv.fpos[c];
        if (Cv.type[c] == BN_SMALL) {
            const uint32_t d = seb + Cv.soff[c] + rank;
            Sm.lo[d] = lo;
            Sm.hi[d] = hi;
        } else {
            const uint32_t d = Cv.loff[c] + rank;
            Nx.lo[d] = lo;
            Nx.hi[d] = hi;
        }
    }
}

// Workgroups of k_clip (kClipBlock lanes) for a level of up to ne elements.
inline dim3 clip_blocks(uint64_t ne) {
    return dim3((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((2 * ne + kClipBlock - 1) / kClipBlock, 4096)));
}

// ---- small-node stage: one wave64 per small root ----------------------------------------

// one small root per wave and one wave per workgroup: a workgroup's LDS (the candidate
// masks, about 10 KB for triangles) is released when its own root is done, not when the
// slowest of several roots is (measured: the recorded 1.03M-triangle build 4.97 -> 4.78 ms)
constexpr int kSmallWaves = 1;
constexpr int kSmallStack = MNRT_MAX_LEVEL + 4;

struct SmallStackEnt {
    uint64_t mask;
    float lo[3], hi[3];
    uint32_t lvl_ax;  // level | axis << 30 (axis: of the split whose right child this is)
    uint32_t patch;   // node whose right-child word gets this entry's address, or kNone
    __device__ uint32_t level() const { return lvl_ax & 0x3FFFFFFFu; }
    __device__ uint32_t axis() const { return lvl_ax >> 30; }
};

// Staging area of a small root with n elements: stage_mul * n + stage_add node words and
// as many references (4n + 64 by default; the environment variable MNRT_SMALL_STAGE_CAP
// = "mul,add" overrides it, so tests can force the fallback).  The subtree is written
// there with subtree-local addresses while its size is counted, then moved into place
// once the layout is known.  A subtree that does not fit (rare) is counted on without
// being written and emitted by a second pass at its final address.
struct StageCap {
    uint32_t mul, add;
    __host__ __device__ uint32_t of(uint32_t n) const { return mul * n + add; }
};

enum : int { kSmallStage = 0, kSmallEmit = 1 };

struct SmallOut {
    uint32_t* nsize;  // stage: nodes per small subtree
    uint32_t* rsize;  // stage: references per small subtree
    uint32_t* mlvl;   // stage: deepest level in the small subtree
    const uint32_t* st_off;  // stage: first slot of each small root's staging area
    uint2* st_nodes;         // stage: node words, subtree-local addresses
    uint32_t* st_refs;       // stage: references, subtree-local
    uint32_t* st_ovf;        // stage: 1 where the subtree did not fit; emit: only these
    StageCap st_cap;
    const uint32_t* bn_addr;  // emit: subtree start address / reference base
    const uint32_t* bn_rbase;
    uint32_t* nodes;
    uint32_t* refs;
    uint32_t* counters;
    uint32_t max_nodes;
    uint32_t max_refs;
    const uint32_t* node_base;  // per scene: first node / reference of its tree
    const uint32_t* ref_base;
};

// Returns `old` with lane `lane` replaced by the wave-uniform `v` (v_writelane_b32, which
// ignores exec; with an SGPR source the lane select has to be m0).
__device__ __forceinline__ uint32_t write_lane(uint32_t v, int lane, uint32_t old) {
    uint32_t r;
    asm("v_writelane_b32 %0, %1, m0"
        : "=v"(r)
        : "s"(__builtin_amdgcn_readfirstlane(v)), "{m0}"(__builtin_amdgcn_readfirstlane(lane)), "0"(old));
    return r;
}

// kMode kSmallStage: subtree sizes, and the subtree written to its staging area while it
// fits; kSmallEmit: written at its final address (only roots with st_ovf set
// when st_ovf is given).
// The element kind is a template argument so that LDS holds what the kind needs: point
// elements have no upper bounds and 3n candidates (triangles: 6n), which halves the mask
// arrays and lets more waves share a CU.
template <int kMode, int kKind>
__global__ __launch_bounds__(64 * kSmallWaves) void k_small(DN dnsmall, SmallView S, ElemView Sm,
                                                            const float* __restrict__ scene_qr, SmallOut O) {
    constexpr int kind = kKind;
    // candidate slots of an element: (isMax, axis), slot = isMax * 3 + axis
    constexpr int kSlots = (kKind == MNRT_KIND_TRI) ? 6 : 3;
    __shared__ SmallStackEnt s_stack[kSmallWaves][kSmallStack];

    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63u;
    const uint32_t r = blockIdx.x * kSmallWaves + wave;
    if (r >= dnsmall.get()) return;  // uniform per wave
    if (kMode == kSmallEmit && O.st_ovf && O.st_ovf[r] == 0u) return;
    if (kMode == kSmallEmit && O.counters[C_ERR] != 0u) return;  // failed build: nothing is emitted
    SmallStackEnt* stack = s_stack[wave];

    const uint32_t first = S.first[r], n = S.cnt[r];
    // lane j holds element j (n <= 64)
    const bool valid = lane < n;
    float4 mlo = make_float4(0.0f, 0.0f, 0.0f, 0.0f), mhi = mlo;
    uint32_t midx = 0;
    if (valid) {
        mlo = Sm.lo[first + lane];
        mhi = (kKind == MNRT_KIND_TRI) ? Sm.hi[first + lane] : mlo;
        midx = Sm.idx[first + lane];
    }
    const bool mleft = plane_goes_left(midx);
    // Candidates axis-major, minimums then maximums (kernel_CreateSplitCandidates,
    // kdtree.cu:1035-1082), c = isMax * 3n + axis * n + i; masks as kernel_InitSplitMasks
    // (kdtree.cu:1110-1251).  One candidate per step: every lane compares its element with
    // the plane, the comparisons' lane masks combine into the candidate's masks (element
    // i itself, and elements lying in the plane, by the rules below), and lane i keeps
    // them for its own candidate (writelane); each lane writes its candidates at the end
    // of the axis.  Lane i keeps the masks of its element's candidates in registers (slot
    // isMax * 3 + axis: candidate c = slot * n + i).
    //   triangles: L = min < s, R = s < max; the candidate's own element goes right of
    //              its minimum and left of its maximum; a flat element in the plane goes
    //              left if plane_goes_left, else right;
    //   points:    L = p < s, R = s < p; a point in the plane goes left if
    //              plane_goes_left, else right.
    uint64_t cmL[kSlots], cmR[kSlots];
    const uint64_t vmask = (n >= 64) ? ~0ull : ((1ull << n) - 1ull);
    const uint64_t leftm = __builtin_amdgcn_ballot_w64(mleft) & vmask;
    for (uint32_t ax = 0; ax < 3; ++ax) {
        const float mn = ax == 0 ? mlo.x : (ax == 1 ? mlo.y : mlo.z);
        const float mx = ax == 0 ? mhi.x : (ax == 1 ? mhi.y : mhi.z);
        const uint64_t flat = (kind == MNRT_KIND_TRI) ? __builtin_amdgcn_ballot_w64(mn == mx) & vmask : vmask;
        // this lane's minimum (0) and maximum (1) candidate: masks as 32-bit halves
        uint32_t l0lo = 0, l0hi = 0, r0lo = 0, r0hi = 0, l1lo = 0, l1hi = 0, r1lo = 0, r1hi = 0;
        for (uint32_t i = 0; i < n; ++i) {
            const uint64_t self = 1ull << i;
#pragma unroll
            for (int isMax = 0; isMax < (kind == MNRT_KIND_TRI ? 2 : 1); ++isMax) {
                const float sp = __uint_as_float(
                    __builtin_amdgcn_readlane(__float_as_uint(isMax ? mx : mn), (int)i));
                const uint64_t lt = __builtin_amdgcn_ballot_w64(mn < sp);
                const uint64_t gt = __builtin_amdgcn_ballot_w64(sp < mx);
                const uint64_t eq = __builtin_amdgcn_ballot_w64(mn == sp) & flat;
                uint64_t ml, mr;
                if (kind == MNRT_KIND_TRI) {
                    const uint64_t planar = eq & ~self;
                    ml = (lt | (planar & leftm) | (isMax ? self : 0ull)) & vmask;
                    mr = (gt | (planar & ~leftm) | (isMax ? 0ull : self)) & vmask;
                } else {
                    ml = (lt | (eq & leftm)) & vmask;
                    mr = (gt | (eq & ~leftm)) & vmask;
                }
                if (isMax) {
                    l1lo = write_lane((uint32_t)ml, (int)i, l1lo);
                    l1hi = write_lane((uint32_t)(ml >> 32), (int)i, l1hi);
                    r1lo = write_lane((uint32_t)mr, (int)i, r1lo);
                    r1hi = write_lane((uint32_t)(mr >> 32), (int)i, r1hi);
                } else {
                    l0lo = write_lane((uint32_t)ml, (int)i, l0lo);
                    l0hi = write_lane((uint32_t)(ml >> 32), (int)i, l0hi);
                    r0lo = write_lane((uint32_t)mr, (int)i, r0lo);
                    r0hi = write_lane((uint32_t)(mr >> 32), (int)i, r0hi);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {  // registers by constant index: the axis as selects
            if (ax == (uint32_t)k) {
                cmL[k] = ((uint64_t)l0hi << 32) | l0lo;
                cmR[k] = ((uint64_t)r0hi << 32) | r0lo;
                if (kind == MNRT_KIND_TRI) {
                    cmL[3 + k] = ((uint64_t)l1hi << 32) | l1lo;
                    cmR[3 + k] = ((uint64_t)r1hi << 32) | r1lo;
                }
            }
        }
    }

    uint32_t addr = 0, rbase = 0, maxlvl = 0;
    const uint32_t scene = S.scene[r];
    // emit: scene-local addresses into the scene's tree; stage: subtree-local addresses
    // into the staging area
    uint32_t* nodes = nullptr;
    uint32_t* refs = nullptr;
    uint32_t lim_nodes = 0u, lim_refs = 0u;
    bool writing = true;
    if (kMode == kSmallEmit) {
        nodes = O.nodes + 2ull * O.node_base[scene];
        refs = O.refs + O.ref_base[scene];
        lim_nodes = O.max_nodes - O.node_base[scene];
        lim_refs = O.max_refs - O.ref_base[scene];
        uint32_t bn = S.bn[r];
        addr = O.bn_addr[bn];
        rbase = O.bn_rbase[bn];
    } else if (kMode == kSmallStage) {
        const uint32_t off = O.st_off[r];
        nodes = (uint32_t*)(O.st_nodes + off);
        refs = O.st_refs + off;
        lim_nodes = lim_refs = O.st_cap.of(n);
    }
    int sp = 0;
    if (lane == 0) {
        SmallStackEnt e;
        e.mask = (n == 64) ? ~0ull : ((1ull << n) - 1ull);
        float4 lo = S.lo[r], hi = S.hi[r];
        e.lo[0] = lo.x; e.lo[1] = lo.y; e.lo[2] = lo.z;
        e.hi[0] = hi.x; e.hi[1] = hi.y; e.hi[2] = hi.z;
        e.lvl_ax = S.level[r];
        e.patch = kNone;
        stack[0] = e;
    }
    sp = 1;
    wave_sync();
    const float qr2 = 2.0f * scene_qr[scene];
    while (sp > 0) {
        --sp;
        SmallStackEnt e = stack[sp];
        wave_sync();
        if (writing && (addr >= lim_nodes || rbase > lim_refs)) {
            if (kMode == kSmallEmit) {
                if (lane == 0) atomicOr(&O.counters[C_ERR], 4u);
                break;
            }
            writing = false;  // staging area full: counted on, emitted later
        }
        if (writing && e.patch != kNone && lane == 0) nodes[2 * e.patch + 1] = (e.axis() << 30) | addr;
        maxlvl = max(maxlvl, e.level());
        const uint32_t cnt = (uint32_t)__popcll(e.mask);
        float best_cost = kInf;
        uint32_t best = kNone;
        if (cnt > 0 && e.level() < MNRT_MAX_LEVEL) {
            float ext[3] = {e.hi[0] - e.lo[0], e.hi[1] - e.lo[1], e.hi[2] - e.lo[2]};
            float denom;
            if (kind == MNRT_KIND_POINT) {
                float vol = 1.0f;
                for (int i = 0; i < 3; ++i) vol *= ext[i] + qr2;
                denom = 1.0f / vol;
            } else {
                float area = 2.0f * (ext[0] * ext[1] + ext[0] * ext[2] + ext[1] * ext[2]);
                denom = 1.0f / area;
            }
            // lane i evaluates its element's candidates, slot by slot (the axis a constant:
            // the node's bounds and extents on it need no selects); key = slot * 64 + i
            // orders the candidates as c = slot * n + i does
#pragma unroll
            for (int slot = 0; slot < kSlots; ++slot) {
                if (!valid) break;
                const int a = slot % 3;
                const float s = slot < 3 ? f4c(mlo, a) : f4c(mhi, a);
                const float lo_a = e.lo[a], hi_a = e.hi[a];
                if (!(lo_a < s && s < hi_a)) continue;
                const uint64_t bl = cmL[slot] & e.mask, br = cmR[slot] & e.mask;
                const uint32_t nl = (uint32_t)__popc((uint32_t)bl) + (uint32_t)__popc((uint32_t)(bl >> 32));
                const uint32_t nr = (uint32_t)__popc((uint32_t)br) + (uint32_t)__popc((uint32_t)(br >> 32));
                if (nl == 0 || nr == 0) continue;
                // the other two axes' extents, (a + 1) % 3 and (a + 2) % 3
                const float x1 = ext[(a + 1) % 3];
                const float x2 = ext[(a + 2) % 3];
                float cost;
                if (kind == MNRT_KIND_POINT) {
                    float vl = (x1 + qr2) * (s - lo_a + qr2) * (x2 + qr2);
                    float vr = (x1 + qr2) * (hi_a - s + qr2) * (x2 + qr2);
                    cost = ((float)nl * vl + (float)nr * vr) * denom + kCostTraverse;
                } else {
                    float al = 2.0f * (x1 * x2 + (s - lo_a) * (x1 + x2));
                    float ar = 2.0f * (x1 * x2 + (hi_a - s) * (x1 + x2));
                    cost = ((float)nl * al + (float)nr * ar) * denom + kCostTraverse;
                }
                if (cost < best_cost) {
                    best_cost = cost;
                    best = (uint32_t)slot * 64u + lane;
                }
            }
            // lowest cost, ties to the lower candidate (an order-independent minimum)
            auto take = [&](float oc, uint32_t ob) {
                if (oc < best_cost || (oc == best_cost && ob < best)) {
                    best_cost = oc;
                    best = ob;
                }
            };
#define MNRT_ARGMIN_DPP(ctrl) take(__uint_as_float(dpp_mov<ctrl>(__float_as_uint(best_cost))), dpp_mov<ctrl>(best))
            MNRT_ARGMIN_DPP(kDppPairSwap);
            MNRT_ARGMIN_DPP(kDppQuadSwap);
            MNRT_ARGMIN_DPP(kDppHalfMirror);
            MNRT_ARGMIN_DPP(kDppRowMirror);
#undef MNRT_ARGMIN_DPP
            {
                const uint32_t cb = __float_as_uint(best_cost), bb = best;
                best_cost = __uint_as_float(row_lane(cb, 0));
                best = row_lane(bb, 0);
                for (int row = 1; row < 4; ++row) take(__uint_as_float(row_lane(cb, row)), row_lane(bb, row));
            }
        }
        if (best == kNone || !(best_cost < (float)cnt)) {
            if (writing && rbase + cnt > lim_refs) {
                if (kMode == kSmallEmit) {
                    if (lane == 0) atomicOr(&O.counters[C_ERR], 8u);
                    break;
                }
                writing = false;
            }
            if (writing) {
                if (lane == 0) {
                    nodes[2 * addr] = cnt;
                    nodes[2 * addr + 1] = (MNRT_LEAF_TAG << 30) | rbase;
                }
                if (lane < n && ((e.mask >> lane) & 1ull)) {
                    uint64_t below = (lane == 0) ? 0ull : (e.mask & ((1ull << lane) - 1ull));
                    refs[rbase + (uint32_t)__popcll(below)] = midx;
                }
            }
            ++addr;
            rbase += cnt;
            continue;
        }
        // the chosen candidate: slot and element lane from the key, its plane and masks
        // read from that lane
        const uint32_t bslot = __builtin_amdgcn_readfirstlane(best) >> 6, blane = __builtin_amdgcn_readfirstlane(best) & 63u;
        const uint32_t a = bslot % 3u;
        float sv = 0.0f;
        uint64_t bmL = 0, bmR = 0;
#pragma unroll
        for (int k = 0; k < kSlots; ++k) {
            if (bslot == (uint32_t)k) {
                sv = k < 3 ? f4c(mlo, k % 3) : f4c(mhi, k % 3);
                bmL = cmL[k];
                bmR = cmR[k];
            }
        }
        // (readlane returns int: each half goes through uint32_t, or the low half would be
        // sign-extended into the high one)
        auto lane_u32 = [&](uint32_t v) { return (uint32_t)__builtin_amdgcn_readlane(v, (int)blane); };
        const float s = __uint_as_float(lane_u32(__float_as_uint(sv)));
        const uint64_t selL = ((uint64_t)lane_u32((uint32_t)(bmL >> 32)) << 32) | (uint64_t)lane_u32((uint32_t)bmL);
        const uint64_t selR = ((uint64_t)lane_u32((uint32_t)(bmR >> 32)) << 32) | (uint64_t)lane_u32((uint32_t)bmR);
        if (writing && lane == 0) {
            nodes[2 * addr] = __float_as_uint(s);
            nodes[2 * addr + 1] = a << 30;
        }
        if (lane == 0) {
            SmallStackEnt R = e, L = e;
            R.mask = selR & e.mask;
            for (uint32_t k = 0; k < 3; ++k) {  // the plane bounds the split axis (selects, not indexing)
                R.lo[k] = (a == k) ? s : e.lo[k];
                L.hi[k] = (a == k) ? s : e.hi[k];
            }
            R.lvl_ax = (e.level() + 1) | (a << 30);
            R.patch = addr;
            L.mask = selL & e.mask;
            L.lvl_ax = e.level() + 1;
            L.patch = kNone;
            stack[sp] = R;
            stack[sp + 1] = L;
        }
        sp += 2;
        ++addr;
        wave_sync();
    }
    if (kMode == kSmallStage && lane == 0) {
        O.nsize[r] = addr;
        O.rsize[r] = rbase;
        O.mlvl[r] = maxlvl;
        O.st_ovf[r] = writing ? 0u : 1u;
        if (!writing) atomicAdd(&O.counters[C_SMALL_OVF], 1u);
    }
}

__global__ __launch_bounds__(256) void k_small_caps(DN dnsmall, const uint32_t* __restrict__ cnt, StageCap cap,
                                                    uint32_t* __restrict__ caps) {
    const uint32_t nsmall = dnsmall.get();
    MNRT_GRID_STRIDE(r, nsmall) caps[r] = cap.of(cnt[r]);
}

StageCap stage_cap_setting() {
    StageCap c{4u, 64u};
    if (const char* e = std::getenv("MNRT_SMALL_STAGE_CAP")) {
        unsigned m = 0, a = 0;
        if (std::sscanf(e, "%u,%u", &m, &a) == 2 && m <= 64 && a <= 4096) c = StageCap{m, a};
    }
    return c;
}

// Moves staged small subtrees to their final addresses (one wave per small root):
// inner words get the subtree's first address added to the right-child index, leaf words
// its first reference added to the reference offset.
__global__ __launch_bounds__(64 * kSmallWaves) void k_small_place(DN dnsmall, SmallView S, SmallOut O) {
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t r = blockIdx.x * kSmallWaves + wave;
    if (r >= dnsmall.get() || O.st_ovf[r] != 0u || O.counters[C_ERR] != 0u) return;
    const uint32_t scene = S.scene[r];
    const uint32_t bn = S.bn[r];
    const uint32_t a0 = O.bn_addr[bn], r0 = O.bn_rbase[bn];
    const uint32_t nn = O.nsize[r], nr = O.rsize[r];
    const uint32_t lim_nodes = O.max_nodes - O.node_base[scene], lim_refs = O.max_refs - O.ref_base[scene];
    if (a0 + nn > lim_nodes || r0 + nr > lim_refs) {
        if (lane == 0) atomicOr(&O.counters[C_ERR], 4u);
        return;
    }
    uint2* nodes = (uint2*)(O.nodes + 2ull * O.node_base[scene]) + a0;
    uint32_t* refs = O.refs + O.ref_base[scene] + r0;
    const uint32_t off = O.st_off[r];
    for (uint32_t j = lane; j < nn; j += 64) {
        uint2 w = O.st_nodes[off + j];
        const uint32_t tag = w.y >> 30;
        w.y = (tag << 30) | ((w.y & MNRT_INDEX_MASK) + (tag == MNRT_LEAF_TAG ? r0 : a0));
        nodes[j] = w;
    }
    for (uint32_t j = lane; j < nr; j += 64) refs[j] = O.st_refs[off + j];
}

// ---- layout: subtree sizes bottom-up, addresses top-down, one workgroup ------------------

struct LayoutView {
    uint32_t* size;   // nodes in the subtree
    uint32_t* rsize;  // references in the subtree
    uint32_t* depth;  // deepest level in the subtree
    uint32_t* addr;
    uint32_t* rbase;
};

// Build nodes bucketed by level (the order inside a level does not matter).  Counts go
// through a per-workgroup LDS histogram, so each workgroup adds to a level's global
// counter once.
// top: the deepest level the layout launches cover; a node below it sets err_bit in *err.
__global__ __launch_bounds__(256) void k_level_hist(DN dnbn, BnView B, uint32_t* __restrict__ hist, uint32_t top,
                                                    uint32_t* __restrict__ err, uint32_t err_bit) {
    __shared__ uint32_t s_hist[64];
    if (threadIdx.x < 64) s_hist[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t nbn = dnbn.get();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nbn && B.type[i] < BN_PENDING) {
        const uint32_t lvl = B.level[i] & 63u;
        if (lvl > top) atomicOr(err, err_bit);
        atomicAdd(&s_hist[lvl], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 64 && s_hist[threadIdx.x] != 0u) atomicAdd(&hist[threadIdx.x], s_hist[threadIdx.x]);
}

__global__ __launch_bounds__(256) void k_level_scatter(DN dnbn, BnView B, uint32_t* __restrict__ cursor,
                                                       uint32_t* __restrict__ list) {
    __shared__ uint32_t s_cnt[64];
    __shared__ uint32_t s_base[64];
    if (threadIdx.x < 64) s_cnt[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t nbn = dnbn.get();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < nbn && B.type[i] < BN_PENDING;
    const uint32_t lvl = live ? (B.level[i] & 63u) : 0u;
    const uint32_t local = live ? atomicAdd(&s_cnt[lvl], 1u) : 0u;
    __syncthreads();
    if (threadIdx.x < 64 && s_cnt[threadIdx.x] != 0u)
        s_base[threadIdx.x] = atomicAdd(&cursor[threadIdx.x], s_cnt[threadIdx.x]);
    __syncthreads();
    if (live) list[s_base[lvl] + local] = i;
}

// Subtree sizes bottom-up, then preorder addresses top-down, one launch per level over the
// level's bucket (KDTreeGPU::PreorderTraversal launches two kernels per level over all
// nodes, KDTreeGPU.cpp:1156-1209; here only the level's nodes are touched).
struct LayoutSrc {
    const uint32_t* off;   // level bucket offsets (65)
    const uint32_t* list;  // build nodes by level
    const uint32_t* s_nsize;
    const uint32_t* s_rsize;
    const uint32_t* s_mlvl;
};

__global__ __launch_bounds__(256) void k_layout_up(uint32_t lvl, LayoutSrc Ls, BnView B, LayoutView Lv) {
    const uint32_t end = Ls.off[lvl + 1];
    for (uint32_t j = Ls.off[lvl] + blockIdx.x * blockDim.x + threadIdx.x; j < end; j += gridDim.x * blockDim.x) {
        const uint32_t i = Ls.list[j];
        const uint32_t t = B.type[i];
        if (t <= 2) {
            const uint32_t l = B.left[i], r = B.right[i];
            Lv.size[i] = 1 + Lv.size[l] + Lv.size[r];
            Lv.rsize[i] = Lv.rsize[l] + Lv.rsize[r];
            Lv.depth[i] = max(Lv.depth[l], Lv.depth[r]);
        } else if (t == BN_LEAF) {
            Lv.size[i] = 1;
            Lv.rsize[i] = B.cnt[i];
            Lv.depth[i] = B.level[i];
        } else if (t == BN_SMALL) {
            const uint32_t x = B.aux[i];
            Lv.size[i] = Ls.s_nsize[x];
            Lv.rsize[i] = Ls.s_rsize[x];
            Lv.depth[i] = Ls.s_mlvl[x];
        }
    }
}

// Level lvl's nodes place their children; every scene's root (level 0) starts its own
// tree at address 0.
__global__ __launch_bounds__(256) void k_layout_down(uint32_t lvl, LayoutSrc Ls, BnView B, LayoutView Lv) {
    const uint32_t end = Ls.off[lvl + 1];
    for (uint32_t j = Ls.off[lvl] + blockIdx.x * blockDim.x + threadIdx.x; j < end; j += gridDim.x * blockDim.x) {
        const uint32_t i = Ls.list[j];
        if (lvl == 0) {
            Lv.addr[i] = 0;
            Lv.rbase[i] = 0;
        }
        if (B.type[i] > 2) continue;
        const uint32_t l = B.left[i], r = B.right[i];
        const uint32_t a = Lv.addr[i], rb = Lv.rbase[i];
        Lv.addr[l] = a + 1;
        Lv.rbase[l] = rb;
        Lv.addr[r] = a + 1 + Lv.size[l];
        Lv.rbase[r] = rb + Lv.rsize[l];
    }
}

__global__ __launch_bounds__(256) void k_emit_large(DN dnbn, BnView B, LayoutView Lv, const uint32_t* __restrict__ ll_idx,
                                                    uint32_t* __restrict__ nodes_all, uint32_t* __restrict__ refs_all,
                                                    uint32_t* __restrict__ counters, uint32_t max_nodes_all,
                                                    uint32_t max_refs_all, const uint32_t* __restrict__ node_base,
                                                    const uint32_t* __restrict__ ref_base) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= dnbn.get() || counters[C_ERR] != 0u) return;  // failed build: nothing is emitted
    uint32_t t = B.type[i];
    if (t > BN_LEAF) return;
    const uint32_t sc = B.scene[i];
    uint32_t* nodes = nodes_all + 2ull * node_base[sc];
    uint32_t* refs = refs_all + ref_base[sc];
    const uint32_t max_nodes = max_nodes_all - node_base[sc], max_refs = max_refs_all - ref_base[sc];
    if (t <= BN_LEAF && (Lv.addr[i] >= max_nodes || (t == BN_LEAF && Lv.rbase[i] + B.cnt[i] > max_refs) ||
                         (t <= 2 && Lv.addr[B.right[i]] >= max_nodes))) {
        atomicOr(&counters[C_ERR], 16u);
        return;
    }
    if (t <= 2) {
        uint32_t a = Lv.addr[i];
        nodes[2 * a] = B.split[i];
        nodes[2 * a + 1] = (t << 30) | Lv.addr[B.right[i]];
    } else if (t == BN_LEAF) {
        uint32_t a = Lv.addr[i], rb = Lv.rbase[i], m = B.cnt[i];
        nodes[2 * a] = m;
        nodes[2 * a + 1] = (MNRT_LEAF_TAG << 30) | rb;
        for (uint32_t k = 0; k < m; ++k) refs[rb + k] = ll_idx[B.aux[i] + k];
    }
}

__global__ __launch_bounds__(256) void k_scene_counts(uint32_t nscenes, const uint32_t* __restrict__ base,
                                                      uint32_t* __restrict__ cnt) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nscenes) cnt[s] = base[s + 1] - base[s];
}

// Per scene (root = build node s): nodes, references, leaves and max level.  Every inner
// node has two children, so leaves = (nodes + 1) / 2.
__global__ __launch_bounds__(256) void k_scene_sizes(uint32_t nscenes, LayoutView Lv, uint32_t* __restrict__ nodes,
                                                     uint32_t* __restrict__ refs, uint32_t* __restrict__ leaves,
                                                     uint32_t* __restrict__ maxlvl) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nscenes) return;
    nodes[s] = Lv.size[s];
    refs[s] = Lv.rsize[s];
    leaves[s] = (Lv.size[s] + 1) / 2;
    maxlvl[s] = Lv.depth[s];
}

}  // namespace mnrt

// =========================================================================================
// host driver
// =========================================================================================
namespace {

using namespace mnrt;

// Build workspace: one device block per host thread and device, kept across builds and
// grown to the largest build seen, so a build in steady state allocates nothing but its
// output arrays (allocation calls that must grow device memory cost milliseconds each).
// Buffers are carved bump-wise and not reused within a build; a build that outgrows the
// block takes the rest from the stream-ordered pool, and the block grows to fit when that
// build ends.  The blocks live until the process ends.
constexpr int kArenaDevices = 64;
struct Arena {
    char* base = nullptr;
    size_t cap = 0;
};
thread_local Arena t_arena[kArenaDevices];

class Pool {
   public:
    explicit Pool(hipStream_t st) : st_(st) {
        int dev = -1;
        if (hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < kArenaDevices) arena_ = &t_arena[dev];
    }
    ~Pool() {
        // nothing may still use the workspace when the next build on this thread reuses it
        (void)hipStreamSynchronize(st_);
        for (void* p : ptrs_) (void)pool_free(p, st_);
        if (arena_ && total_ > arena_->cap) {
            (void)hipStreamSynchronize(st_);
            void* nb = nullptr;
            const size_t ncap = total_ + total_ / 4 + (size_t(64) << 20);
            if (hipMalloc(&nb, ncap) == hipSuccess) {
                if (arena_->base) (void)hipFree(arena_->base);
                arena_->base = (char*)nb;
                arena_->cap = ncap;
            }
        }
    }
    template <typename T>
    T* get(size_t count) {
        const size_t bytes = (std::max<size_t>(count * sizeof(T), 16) + 255) & ~size_t(255);
        total_ += bytes;
        if (arena_ && top_ + bytes <= arena_->cap) {
            void* p = arena_->base + top_;
            top_ += bytes;
            return (T*)p;
        }
        void* p = nullptr;
        if (pool_alloc(&p, bytes, st_) != hipSuccess) {
            ok = false;
            return nullptr;
        }
        ptrs_.push_back(p);
        return (T*)p;
    }
    // Returns a buffer early: pool buffers go back to the pool, workspace space stays
    // carved until the build ends.
    void release(void* p) {
        if (!p) return;
        auto it = std::find(ptrs_.begin(), ptrs_.end(), p);
        if (it != ptrs_.end()) {
            (void)pool_free(p, st_);
            ptrs_.erase(it);
        }
    }
    bool ok = true;

   private:
    hipStream_t st_;
    Arena* arena_ = nullptr;
    size_t top_ = 0;
    size_t total_ = 0;
    std::vector<void*> ptrs_;
};

inline dim3 blocks_for(uint64_t n, uint32_t bs = 256) { return dim3((uint32_t)std::max<uint64_t>(1, (n + bs - 1) / bs)); }

// Sizes a host-driven build met, from which a recorded device-driven build sizes its
// fixed buffers: per level the active nodes / elements, and the largest value of each
// quantity the level loop grows a buffer for (the same bounds the loop checks).
struct BuildStats {
    uint32_t na0 = 0, ne0 = 0, nsmall0 = 0, nse0 = 0;
    std::vector<uint32_t> level_na, level_ne;
    uint64_t max_2na = 0, max_ne = 0, max_bn = 0, max_small = 0, max_se = 0, max_ll = 0;
    uint32_t nbn = 0, nsmall = 0, nse = 0, maxlvl = 0, total_nodes = 0, total_refs = 0;
};

// Builds the trees of S scenes together: one level loop over the active nodes of all
// scenes, one small-node launch for all small roots, one layout / emission pass.  Each
// scene's tree equals the tree a single-scene build of it gives (indices are
// scene-local, all decisions are per node).
//   h_base: S + 1 element offsets into the concatenated input.
//   h_roots: 6 floats per scene (or null: tight bounds).  h_qr: VVH radius per scene
//   (or null: params->max_query_radius).
//   out[s] views into shared arrays; out[0] owns them (mnrt_tree_release(&out[0])
//   frees every tree of the batch).
int build_batch_impl(int kind, const float* d_v0, const float* d_v1, const float* d_v2, const uint32_t* h_base,
                     uint32_t nscenes, const mnrt_build_params* prm, const float* h_roots, const float* h_qr,
                     mnrt_tree* out, hipStream_t st, BuildStats* stats = nullptr) {
    if (!out || !prm || !h_base || nscenes == 0) return MNRT_ERR_ARG;
    const uint32_t ntot = h_base[nscenes];
    for (uint32_t s = 0; s < nscenes; ++s)
        if (h_base[s + 1] < h_base[s]) return MNRT_ERR_ARG;
    if (ntot > 0 && !d_v0) return MNRT_ERR_ARG;
    if (kind == MNRT_KIND_TRI && ntot > 0 && (!d_v1 || !d_v2)) return MNRT_ERR_ARG;
    if (prm->small_node_max < 1 || prm->small_node_max > 64) return MNRT_ERR_ARG;
    if (ntot > (MNRT_INDEX_MASK >> 2)) return MNRT_ERR_CAPACITY;
    const uint32_t S_ = nscenes;
    Pool pool(st);
    uint32_t* counters = pool.get<uint32_t>(C_COUNT);
    // per-scene device tables: base (S+1), plan kind/index/off (3S), qr (S), roots given (6S),
    // root boxes out (6S), node/ref counts + bases (4S + 2), leaves / max level (2S)
    uint32_t* d_tab = pool.get<uint32_t>(32ull * S_ + 64);
    if (!pool.ok) return MNRT_ERR_NOMEM;
    uint32_t* d_base = d_tab;
    uint32_t* d_pkind = d_base + S_ + 1;
    uint32_t* d_pindex = d_pkind + S_;
    uint32_t* d_poff = d_pindex + S_;
    float* d_qr = (float*)(d_poff + S_);
    float* d_given = d_qr + S_;
    float* d_rootbox = d_given + 6 * S_;
    uint32_t* d_snodes = (uint32_t*)(d_rootbox + 6 * S_);
    uint32_t* d_srefs = d_snodes + S_;
    uint32_t* d_nbase = d_srefs + S_;      // S + 1 (total at [S])
    uint32_t* d_rbase = d_nbase + S_ + 1;  // S + 1
    uint32_t* d_sleaves = d_rbase + S_ + 1;
    uint32_t* d_smaxlvl = d_sleaves + S_;

    // host staging in pinned memory: counters, the per-scene tables
    const size_t hwords = 64 + 32ull * S_;
    uint32_t* h_pin = (uint32_t*)host_staging(0, hwords * 4);
    if (!h_pin) return MNRT_ERR_NOMEM;
    uint32_t* h_counters = h_pin;
    uint32_t* h_tab = h_pin + 64;
    // RootPlan on the host: large scenes become active roots, small ones small roots
    uint32_t na = 0, nsmall = 0, ne = 0, nse = 0;
    {
        uint32_t* hb = h_tab;
        uint32_t* hk = hb + S_ + 1;
        uint32_t* hi = hk + S_;
        uint32_t* ho = hi + S_;
        float* hq = (float*)(ho + S_);
        float* hg = hq + S_;
        for (uint32_t s = 0; s <= S_; ++s) hb[s] = h_base[s];
        for (uint32_t s = 0; s < S_; ++s) {
            const uint32_t cnt = h_base[s + 1] - h_base[s];
            if (cnt > prm->small_node_max) {
                hk[s] = 0;
                hi[s] = na++;
                ho[s] = ne;
                ne += cnt;
            } else {
                hk[s] = 1;
                hi[s] = nsmall++;
                ho[s] = nse;
                nse += cnt;
            }
            hq[s] = h_qr ? h_qr[s] : prm->max_query_radius;
            for (int k = 0; k < 6; ++k) hg[6 * s + k] = h_roots ? h_roots[6 * s + k] : 0.0f;
        }
        MNRT_HIP_CHECK(hipMemcpyAsync(d_tab, h_tab, (size_t)(S_ + 1 + 4 * S_ + 6 * S_) * 4, hipMemcpyHostToDevice, st));
    }
    MNRT_HIP_CHECK(hipMemsetAsync(counters, 0, C_COUNT * sizeof(uint32_t), st));
    MNRT_HIP_CHECK(hipMemsetAsync(d_sleaves, 0, 2ull * S_ * 4, st));
    Geometry G{kind, (const float4*)d_v0, d_v1 ? (const float4*)d_v1 : (const float4*)d_v0,
               d_v2 ? (const float4*)d_v2 : (const float4*)d_v0, d_base, S_};
    RootPlan RP{d_pkind, d_pindex, d_poff};

    auto sync_counters = [&]() -> int {
        MNRT_HIP_CHECK(hipMemcpyAsync(h_counters, counters, C_COUNT * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        MNRT_HIP_CHECK(hipStreamSynchronize(st));
        return MNRT_OK;
    };
    auto alloc_elems = [&](uint32_t cap, bool with_node) {
        ElemView V{};
        V.idx = pool.get<uint32_t>(cap);
        V.lo = pool.get<float4>(cap);
        V.hi = kind == MNRT_KIND_TRI ? pool.get<float4>(cap) : V.lo;
        V.node = with_node ? pool.get<uint32_t>(cap) : nullptr;
        return V;
    };
    auto release_elems = [&](ElemView& V) {
        pool.release(V.idx);
        pool.release(V.lo);
        if (kind == MNRT_KIND_TRI) pool.release(V.hi);
        pool.release(V.node);
        V = ElemView{};
    };
    auto alloc_nodes = [&](uint32_t cap) {
        NodeView V{};
        V.lo = pool.get<float4>(cap);
        V.hi = pool.get<float4>(cap);
        V.first = pool.get<uint32_t>(cap);
        V.cnt = pool.get<uint32_t>(cap);
        V.level = pool.get<uint32_t>(cap);
        V.bn = pool.get<uint32_t>(cap);
        V.scene = pool.get<uint32_t>(cap);
        return V;
    };
    auto release_nodes = [&](NodeView& V) {
        pool.release(V.lo);
        pool.release(V.hi);
        pool.release(V.first);
        pool.release(V.cnt);
        pool.release(V.level);
        pool.release(V.bn);
        pool.release(V.scene);
        V = NodeView{};
    };
    auto alloc_small = [&](uint32_t cap) {
        SmallView V{};
        V.first = pool.get<uint32_t>(cap);
        V.cnt = pool.get<uint32_t>(cap);
        V.level = pool.get<uint32_t>(cap);
        V.bn = pool.get<uint32_t>(cap);
        V.lo = pool.get<float4>(cap);
        V.hi = pool.get<float4>(cap);
        V.scene = pool.get<uint32_t>(cap);
        return V;
    };

    // ---- input elements, per-scene tight bounds, root placement ----
    ElemView Ein = alloc_elems(std::max<uint32_t>(ntot, 1), true);
    if (!pool.ok) return MNRT_ERR_NOMEM;
    if (ntot > 0) hipLaunchKernelGGL(k_init_elems, blocks_for(ntot), dim3(256), 0, st, G, ntot, Ein);
    uint32_t* tlo = nullptr;
    uint32_t* thi = nullptr;
    if (!h_roots && ntot > 0) {
        // one node per scene over the concatenated elements
        uint32_t max_chunks = ntot / kChunk + S_;
        uint32_t* tb = pool.get<uint32_t>(3ull * max_chunks + 8ull * S_ + scan_scratch_words(S_) + 16);
        if (!pool.ok) return MNRT_ERR_NOMEM;
        uint32_t *chn = tb, *chf = chn + max_chunks, *chc = chf + max_chunks, *cc = chc + max_chunks;
        uint32_t* cnts = cc + S_;
        tlo = cnts + S_;
        thi = tlo + 3 * S_;
        uint32_t* sc = thi + 3 * S_;
        hipLaunchKernelGGL(k_scene_counts, blocks_for(S_), dim3(256), 0, st, S_, d_base, cnts);
        hipLaunchKernelGGL(k_chunk_counts, blocks_for(S_), dim3(256), 0, st, hn(S_), cnts, cc, tlo, thi);
        MNRT_HIP_CHECK(scan_exclusive(cc, cc, S_, counters + C_NCHUNKS, sc, st));
        hipLaunchKernelGGL(k_chunk_fill, blocks_for(max_chunks), dim3(256), 0, st, hn(S_), max_chunks, cc, counters,
                           d_base, cnts, chn, chf, chc);
        hipLaunchKernelGGL(k_tight, dim3((max_chunks + kTightChunks - 1) / kTightChunks), dim3(kChunk), 0, st,
                           max_chunks, counters, chn, chf, chc, Ein,
                           kind, tlo, thi);
    }
    // growable tables
    uint32_t bn_cap = std::max<uint32_t>(1024, S_ + 8 * (ntot / std::max<uint32_t>(1, prm->small_node_max)) + 64);
    uint32_t* bn_buf = pool.get<uint32_t>((size_t)kBnCols * bn_cap);
    auto bn_view = [&](uint32_t* buf, uint32_t cap) {
        return BnView{buf,           buf + cap,           buf + 2ull * cap, buf + 3ull * cap, buf + 4ull * cap,
                      buf + 5ull * cap, buf + 6ull * cap, buf + 7ull * cap};
    };
    BnView B = bn_view(bn_buf, bn_cap);
    uint32_t nbn = S_;
    uint32_t s_cap = std::max<uint32_t>(256, nsmall + 4 * (ntot / std::max<uint32_t>(1, prm->small_node_max)) + 16);
    SmallView SV = alloc_small(s_cap);
    uint32_t se_cap = std::max<uint32_t>(1024, nse + ne + ne / 2);
    ElemView Sm = alloc_elems(se_cap, false);
    uint32_t a_cap = std::max<uint32_t>(na, 1);
    NodeView A = alloc_nodes(a_cap);
    uint32_t e_cap = std::max<uint32_t>(ne, 1);
    ElemView E = alloc_elems(e_cap, true);
    if (!pool.ok) return MNRT_ERR_NOMEM;
    hipLaunchKernelGGL(k_roots_setup, blocks_for(S_), dim3(256), 0, st, S_, d_base, tlo, thi,
                       h_roots ? (const float*)d_given : (const float*)nullptr, RP, A, SV, B, d_rootbox);
    if (ntot > 0) hipLaunchKernelGGL(k_partition_elems, blocks_for(ntot), dim3(256), 0, st, ntot, Ein, kind, RP, E, Sm);
    MNRT_HIP_CHECK(hipGetLastError());
    release_elems(Ein);
    if (ne == 0) {
        release_elems(E);
        e_cap = 0;
    }

    uint32_t ll_cap = 0;
    uint32_t* ll_idx = nullptr;
    uint32_t ll_count = 0;
    uint32_t maxlvl = 0;

    // ---- large-node stage ----
    // Workspaces grow geometrically and are reused across levels; next-level arrays are
    // sized by the upper bound (every element / node goes to at most two children), so
    // the host waits on the device once per level, to read the level's counters.
    uint32_t lv_na_cap = 0, lv_ne_cap = 0;
    void* lv_buf = nullptr;
    uint32_t *cc = nullptr, *chn = nullptr, *chf = nullptr, *chc = nullptr, *fl = nullptr, *pos = nullptr,
             *cbuf = nullptr, *scr = nullptr;
    PlanView P{};
    ElemView Nx{};
    uint32_t nx_cap = 0;
    NodeView An{};
    uint32_t an_cap = 0;
    if (stats) {
        *stats = BuildStats{};
        stats->na0 = na;
        stats->ne0 = ne;
        stats->nsmall0 = nsmall;
        stats->nse0 = nse;
    }
    while (na > 0) {
        if (stats) {
            BuildStats& b = *stats;
            b.level_na.push_back(na);
            b.level_ne.push_back(ne);
            b.max_2na = std::max<uint64_t>(b.max_2na, 2ull * na);
            b.max_ne = std::max<uint64_t>(b.max_ne, ne);
            b.max_bn = std::max<uint64_t>(b.max_bn, nbn + 14ull * na);
            b.max_small = std::max<uint64_t>(b.max_small, nsmall + 2ull * na);
            b.max_se = std::max<uint64_t>(b.max_se, nse + 2ull * ne);
            b.max_ll = std::max<uint64_t>(b.max_ll, (uint64_t)ll_count + ne);
        }
        if (nbn + 14ull * na > bn_cap) {
            uint32_t ncap = std::max<uint32_t>(2 * bn_cap, nbn + 14 * na + 1024);
            uint32_t* nb = pool.get<uint32_t>((size_t)kBnCols * ncap);
            if (!pool.ok) return MNRT_ERR_NOMEM;
            for (int col = 0; col < kBnCols; ++col)
                MNRT_HIP_CHECK(hipMemcpyAsync(nb + (size_t)col * ncap, bn_buf + (size_t)col * bn_cap, (size_t)nbn * 4,
                                              hipMemcpyDeviceToDevice, st));
            pool.release(bn_buf);
            bn_buf = nb;
            bn_cap = ncap;
            B = bn_view(bn_buf, bn_cap);
        }
        if (nsmall + 2ull * na > s_cap) {
            uint32_t ncap = std::max<uint32_t>(2 * s_cap, nsmall + 2 * na + 256);
            SmallView NS = alloc_small(ncap);
            if (!pool.ok) return MNRT_ERR_NOMEM;
            MNRT_HIP_CHECK(hipMemcpyAsync(NS.first, SV.first, (size_t)nsmall * 4, hipMemcpyDeviceToDevice, st));
            MNRT_HIP_CHECK(hipMemcpyAsync(NS.cnt, SV.cnt, (size_t)nsmall * 4, hipMemcpyDeviceToDevice, st));
            MNRT_HIP_CHECK(hipMemcpyAsync(NS.level, SV.level, (size_t)nsmall * 4, hipMemcpyDeviceToDevice, st));
            MNRT_HIP_CHECK(hipMemcpyAsync(NS.bn, SV.bn, (size_t)nsmall * 4, hipMemcpyDeviceToDevice, st));
            MNRT_HIP_CHECK(hipMemcpyAsync(NS.lo, SV.lo, (size_t)nsmall * 16, hipMemcpyDeviceToDevice, st));
            MNRT_HIP_CHECK(hipMemcpyAsync(NS.hi, SV.hi, (size_t)nsmall * 16, hipMemcpyDeviceToDevice, st));
            MNRT_HIP_CHECK(hipMemcpyAsync(NS.scene, SV.scene, (size_t)nsmall * 4, hipMemcpyDeviceToDevice, st));
            pool.release(SV.first);
            pool.release(SV.cnt);
            pool.release(SV.level);
            pool.release(SV.bn);
            pool.release(SV.lo);
            pool.release(SV.hi);
            pool.release(SV.scene);
            SV = NS;
            s_cap = ncap;
        }
        if (ll_count + (uint64_t)ne > ll_cap) {
            uint32_t ncap = std::max<uint32_t>(2 * ll_cap, ll_count + ne + 1024);
            uint32_t* nl = pool.get<uint32_t>(ncap);
            if (!pool.ok) return MNRT_ERR_NOMEM;
            if (ll_count) MNRT_HIP_CHECK(hipMemcpyAsync(nl, ll_idx, (size_t)ll_count * 4, hipMemcpyDeviceToDevice, st));
            pool.release(ll_idx);
            ll_idx = nl;
            ll_cap = ncap;
        }
        if ((uint64_t)nse + 2ull * ne > se_cap) {
            uint32_t ncap = (uint32_t)std::max<uint64_t>(2ull * se_cap, (uint64_t)nse + 2ull * ne + 1024);
            ElemView NS = alloc_elems(ncap, false);
            if (!pool.ok) return MNRT_ERR_NOMEM;
            MNRT_HIP_CHECK(hipMemcpyAsync(NS.idx, Sm.idx, (size_t)nse * 4, hipMemcpyDeviceToDevice, st));
            MNRT_HIP_CHECK(hipMemcpyAsync(NS.lo, Sm.lo, (size_t)nse * 16, hipMemcpyDeviceToDevice, st));
            if (kind == MNRT_KIND_TRI)
                MNRT_HIP_CHECK(hipMemcpyAsync(NS.hi, Sm.hi, (size_t)nse * 16, hipMemcpyDeviceToDevice, st));
            release_elems(Sm);
            Sm = NS;
            se_cap = ncap;
        }
        if (2ull * ne > nx_cap) {
            release_elems(Nx);
            nx_cap = (uint32_t)std::max<uint64_t>(2ull * ne, 1024);
            Nx = alloc_elems(nx_cap, true);
        }
        if (2ull * na > an_cap) {
            release_nodes(An);
            an_cap = std::max<uint32_t>(2 * na, 256);
            An = alloc_nodes(an_cap);
        }
        const uint32_t max_chunks = ne / kChunk + na;
        if (na > lv_na_cap || ne > lv_ne_cap) {
            pool.release(lv_buf);
            lv_na_cap = std::max<uint32_t>(na, lv_na_cap + lv_na_cap / 2);
            lv_ne_cap = std::max<uint32_t>(ne, lv_ne_cap + lv_ne_cap / 2);
            const uint64_t A_ = lv_na_cap, E_ = lv_ne_cap, C_ = E_ / kChunk + A_;
            auto r4 = [](uint64_t k) { return (k + 3) & ~3ull; };
            // the columns handed out below, each rounded to 16 bytes
            // element flags (left, right; triangles also straddle left / right) and their scan
            const uint64_t flw = (kind == MNRT_KIND_TRI ? 4 : 2) * E_ + 1;
            const uint64_t words = r4(A_) + 3 * r4(C_) + 2 * r4(3 * A_) + 9 * r4(A_) + 2 * r4(4 * A_) +
                                   2 * r4(flw) + r4(22 * A_) +
                                   std::max<uint64_t>(scan_scratch_words((uint32_t)flw),
                                                      kScanSetMax * scan_scratch_words((uint32_t)(2 * A_))) +
                                   16;
            lv_buf = pool.get<uint32_t>(words);
            if (!pool.ok) return MNRT_ERR_NOMEM;
        }
        {
            uint32_t* w = (uint32_t*)lv_buf;
            const uint64_t A_ = lv_na_cap, E_ = lv_ne_cap, C_ = E_ / kChunk + A_;
            auto take = [&](uint64_t k) {
                uint32_t* r = w;
                w += (k + 3) & ~3ull;  // 16-byte alignment for float4 columns
                return r;
            };
            cc = take(A_);
            chn = take(C_);
            chf = take(C_);
            chc = take(C_);
            tlo = take(3 * A_);
            thi = take(3 * A_);
            P.newcnt = take(A_);
            P.axis = take(A_);
            P.split = (float*)take(A_);
            P.lo = (float4*)take(4 * A_);
            P.hi = (float4*)take(4 * A_);
            P.level = take(A_);
            P.childL = take(A_);
            P.cur = take(A_);
            P.leaf = take(A_);
            P.noprog = take(A_);
            fl = take((kind == MNRT_KIND_TRI ? 4 : 2) * E_ + 1);  // after the scan: the clip queue
            pos = take((kind == MNRT_KIND_TRI ? 4 : 2) * E_ + 1);
            cbuf = take(22 * A_);
            scr = w;
        }
        ChildView Cv{};
        {
            uint32_t m = 2 * na;
            Cv.type = cbuf;
            Cv.cnt = cbuf + m;
            Cv.fpos = cbuf + 2 * m;
            Cv.sflag = cbuf + 3 * m;
            Cv.lflag = cbuf + 4 * m;
            Cv.scnt = cbuf + 5 * m;
            Cv.lcnt = cbuf + 6 * m;
            Cv.sidx = cbuf + 7 * m;
            Cv.lidx = cbuf + 8 * m;
            Cv.soff = cbuf + 9 * m;
            Cv.loff = cbuf + 10 * m;
        }
        hipLaunchKernelGGL(k_chunk_counts, blocks_for(na), dim3(256), 0, st, hn(na), A.cnt, cc, tlo, thi);
        MNRT_HIP_CHECK(scan_exclusive(cc, cc, na, counters + C_NCHUNKS, scr, st));
        hipLaunchKernelGGL(k_chunk_fill, blocks_for(max_chunks), dim3(256), 0, st, hn(na), max_chunks, cc, counters,
                           A.first, A.cnt, chn, chf, chc);
        hipLaunchKernelGGL(k_tight, dim3((max_chunks + kTightChunks - 1) / kTightChunks), dim3(kChunk), 0, st,
                           max_chunks, counters, chn, chf, chc, E,
                           kind, tlo, thi);
        hipLaunchKernelGGL((k_plan<false>), blocks_for(na), dim3(256), 0, st, hn(na), A, tlo, thi,
                           prm->empty_space_ratio, P, (const uint32_t*)nullptr, hn(0u), B, counters);
        MNRT_HIP_CHECK(scan_exclusive(P.newcnt, P.newcnt, na, counters + C_NEWNODES, scr, st));
        hipLaunchKernelGGL((k_plan<true>), blocks_for(na), dim3(256), 0, st, hn(na), A, tlo, thi,
                           prm->empty_space_ratio, P, (const uint32_t*)P.newcnt, hn(nbn), B, counters);
        hipLaunchKernelGGL(k_classify, blocks_for(ne), dim3(256), 0, st, hn(ne), E, kind, P, fl);
        const uint32_t nfl = (kind == MNRT_KIND_TRI ? 4u : 2u) * ne;
        MNRT_HIP_CHECK(scan_exclusive(fl, pos, nfl, pos + nfl, scr, st));
        hipLaunchKernelGGL(k_child_plan, blocks_for(2ull * na), dim3(256), 0, st, hn(na), hn(ne), A, P, pos,
                           prm->small_node_max, Cv);
        {
            ScanSet cs{{Cv.sflag, Cv.lflag, Cv.scnt, Cv.lcnt},
                       {Cv.sidx, Cv.lidx, Cv.soff, Cv.loff},
                       {},
                       {counters + C_NSMALL_NEW, counters + C_NNEXT, counters + C_ESMALL_NEW, counters + C_ENEXT}};
            MNRT_HIP_CHECK(scan_exclusive_set(cs, 4, 2 * na, scr, st));
        }
        hipLaunchKernelGGL(k_child_write, blocks_for(2ull * na), dim3(256), 0, st, hn(na), A, P, Cv, hn(nsmall),
                           hn(nse), SV, An, B, counters, E, ll_idx, ll_cap);
        hipLaunchKernelGGL(k_scatter, blocks_for(ne), dim3(256), 0, st, hn(na), hn(ne), E, G, A, P, Cv, pos, hn(nse),
                           Sm, se_cap, Nx, nx_cap, counters, fl);
        if (kind == MNRT_KIND_TRI)
            hipLaunchKernelGGL(k_clip, clip_blocks(ne), dim3(kClipBlock), 0, st, hn(na), hn(ne), E, G, A, P, Cv,
                               (const uint32_t*)pos, hn(nse), Sm, Nx, (const uint32_t*)fl);
        MNRT_HIP_CHECK(hipGetLastError());
        int rc = sync_counters();
        if (rc != MNRT_OK) return rc;
        if (h_counters[C_ERR] != 0) return MNRT_ERR_STATE;
        const uint32_t n_next = h_counters[C_NNEXT], e_next = h_counters[C_ENEXT];
        nbn += h_counters[C_NEWNODES];
        nsmall += h_counters[C_NSMALL_NEW];
        nse += h_counters[C_ESMALL_NEW];
        ll_count = h_counters[C_LLCOUNT];
        maxlvl = h_counters[C_MAXLEVEL];
        std::swap(E, Nx);
        std::swap(e_cap, nx_cap);
        std::swap(A, An);
        std::swap(a_cap, an_cap);
        ne = e_next;
        na = n_next;
    }
    release_elems(E);
    release_elems(Nx);
    pool.release(lv_buf);

    // ---- small-node stage: subtree sizes, each subtree written to its staging area ----
    uint32_t* s_nsize = pool.get<uint32_t>(std::max<uint32_t>(nsmall, 1));
    uint32_t* s_rsize = pool.get<uint32_t>(std::max<uint32_t>(nsmall, 1));
    uint32_t* s_mlvl = pool.get<uint32_t>(std::max<uint32_t>(nsmall, 1));
    uint32_t* st_off = pool.get<uint32_t>(std::max<uint32_t>(nsmall, 1));
    uint32_t* st_ovf = pool.get<uint32_t>(std::max<uint32_t>(nsmall, 1));
    uint32_t* st_scr = pool.get<uint32_t>(scan_scratch_words(std::max<uint32_t>(nsmall, 1)));
    const StageCap st_cap = stage_cap_setting();
    const uint64_t st_total = (uint64_t)st_cap.mul * nse + (uint64_t)st_cap.add * nsmall;  // sum over the roots
    if (st_total > 0xFFFFFFFFull) return MNRT_ERR_CAPACITY;
    uint2* st_nodes = pool.get<uint2>(std::max<uint64_t>(st_total, 1));
    uint32_t* st_refs = pool.get<uint32_t>(std::max<uint64_t>(st_total, 1));
    if (!pool.ok) return MNRT_ERR_NOMEM;
    SmallOut O{};
    O.nsize = s_nsize;
    O.rsize = s_rsize;
    O.mlvl = s_mlvl;
    O.st_off = st_off;
    O.st_nodes = st_nodes;
    O.st_refs = st_refs;
    O.st_ovf = st_ovf;
    O.st_cap = st_cap;
    O.counters = counters;
    const dim3 small_grid((nsmall + kSmallWaves - 1) / kSmallWaves), small_block(64 * kSmallWaves);
    if (nsmall > 0) {
        hipLaunchKernelGGL(k_small_caps, blocks_for(nsmall), dim3(256), 0, st, hn(nsmall), (const uint32_t*)SV.cnt,
                           st_cap, st_off);
        MNRT_HIP_CHECK(scan_exclusive(st_off, st_off, nsmall, nullptr, st_scr, st));
        if (kind == MNRT_KIND_TRI)
            hipLaunchKernelGGL((k_small<kSmallStage, MNRT_KIND_TRI>), small_grid, small_block, 0, st, hn(nsmall), SV,
                               Sm, (const float*)d_qr, O);
        else
            hipLaunchKernelGGL((k_small<kSmallStage, MNRT_KIND_POINT>), small_grid, small_block, 0, st, hn(nsmall), SV,
                               Sm, (const float*)d_qr, O);
    }
    // ---- layout ----
    LayoutView Lv{};
    Lv.size = pool.get<uint32_t>(nbn);
    Lv.rsize = pool.get<uint32_t>(nbn);
    Lv.depth = pool.get<uint32_t>(nbn);
    Lv.addr = pool.get<uint32_t>(nbn);
    Lv.rbase = pool.get<uint32_t>(nbn);
    uint32_t* lvb = pool.get<uint32_t>(64 + 65 + 64 + nbn + scan_scratch_words(std::max<uint32_t>(64, S_ + 1)));
    if (!pool.ok) return MNRT_ERR_NOMEM;
    {
        uint32_t* hist = lvb;
        uint32_t* off = lvb + 64;
        uint32_t* cursor = off + 65;
        uint32_t* list = cursor + 64;
        uint32_t* sc = list + nbn;
        MNRT_HIP_CHECK(hipMemsetAsync(hist, 0, 64 * 4, st));
        hipLaunchKernelGGL(k_level_hist, blocks_for(nbn), dim3(256), 0, st, hn(nbn), B, hist, 63u,
                           (uint32_t*)nullptr, 0u);
        MNRT_HIP_CHECK(scan_exclusive(hist, off, 64, off + 64, sc, st));
        MNRT_HIP_CHECK(hipMemcpyAsync(cursor, off, 64 * 4, hipMemcpyDeviceToDevice, st));
        hipLaunchKernelGGL(k_level_scatter, blocks_for(nbn), dim3(256), 0, st, hn(nbn), B, cursor, list);
        LayoutSrc Ls{off, list, s_nsize, s_rsize, s_mlvl};
        const uint32_t top = std::min<uint32_t>(maxlvl, 63);
        const dim3 lgrid(std::min<uint32_t>(blocks_for(nbn).x, 512u));
        for (int lvl = (int)top; lvl >= 0; --lvl)
            hipLaunchKernelGGL(k_layout_up, lgrid, dim3(256), 0, st, (uint32_t)lvl, Ls, B, Lv);
        for (uint32_t lvl = 0; lvl <= top; ++lvl)
            hipLaunchKernelGGL(k_layout_down, lgrid, dim3(256), 0, st, lvl, Ls, B, Lv);
        // per-scene sizes (root = build node s) and their offsets in the shared arrays
        hipLaunchKernelGGL(k_scene_sizes, blocks_for(S_), dim3(256), 0, st, S_, Lv, d_snodes, d_srefs, d_sleaves,
                           d_smaxlvl);
        MNRT_HIP_CHECK(scan_exclusive(d_snodes, d_nbase, S_, d_nbase + S_, sc, st));
        MNRT_HIP_CHECK(scan_exclusive(d_srefs, d_rbase, S_, d_rbase + S_, sc, st));
        MNRT_HIP_CHECK(hipMemcpyAsync(h_counters + 40, d_nbase + S_, 4, hipMemcpyDeviceToHost, st));
        MNRT_HIP_CHECK(hipMemcpyAsync(h_counters + 41, d_rbase + S_, 4, hipMemcpyDeviceToHost, st));
        MNRT_HIP_CHECK(hipMemcpyAsync(h_counters + 42, counters + C_SMALL_OVF, 4, hipMemcpyDeviceToHost, st));
        MNRT_HIP_CHECK(hipStreamSynchronize(st));
    }
    const uint32_t total_nodes = h_counters[40], total_refs = h_counters[41], small_ovf = h_counters[42];
    if (total_nodes == 0 || total_nodes > MNRT_INDEX_MASK || total_refs > MNRT_INDEX_MASK) return MNRT_ERR_CAPACITY;
    if (stats) {
        stats->nbn = nbn;
        stats->nsmall = nsmall;
        stats->nse = nse;
        stats->maxlvl = maxlvl;
        stats->total_nodes = total_nodes;
        stats->total_refs = total_refs;
        stats->max_ll = std::max<uint64_t>(stats->max_ll, ll_count);
    }

    // ---- emission into the caller-visible trees ----
    uint32_t* d_nodes = nullptr;
    uint32_t* d_refs = nullptr;
    if (pool_alloc((void**)&d_nodes, (size_t)total_nodes * 8, st) != hipSuccess) return MNRT_ERR_NOMEM;
    if (pool_alloc((void**)&d_refs, (size_t)total_refs * 4, st) != hipSuccess) {
        (void)pool_free(d_nodes, st);
        return MNRT_ERR_NOMEM;
    }
    hipLaunchKernelGGL(k_emit_large, blocks_for(nbn), dim3(256), 0, st, hn(nbn), B, Lv, ll_idx, d_nodes, d_refs,
                       counters, total_nodes, total_refs, d_nbase, d_rbase);
    O.bn_addr = Lv.addr;
    O.bn_rbase = Lv.rbase;
    O.nodes = d_nodes;
    O.refs = d_refs;
    O.max_nodes = total_nodes;
    O.max_refs = total_refs;
    O.node_base = d_nbase;
    O.ref_base = d_rbase;
    if (nsmall > 0) {
        hipLaunchKernelGGL(k_small_place, small_grid, small_block, 0, st, hn(nsmall), SV, O);
        // subtrees that outgrew their staging area are built again at their final address
        if (small_ovf > 0 && kind == MNRT_KIND_TRI)
            hipLaunchKernelGGL((k_small<kSmallEmit, MNRT_KIND_TRI>), small_grid, small_block, 0, st, hn(nsmall), SV,
                               Sm, (const float*)d_qr, O);
        else if (small_ovf > 0)
            hipLaunchKernelGGL((k_small<kSmallEmit, MNRT_KIND_POINT>), small_grid, small_block, 0, st, hn(nsmall), SV,
                               Sm, (const float*)d_qr, O);
    }
    // per-scene results to the host: counts, bases, leaves, max level, root boxes
    uint32_t* h_res = h_tab;
    int rc = MNRT_OK;
    if (hipMemcpyAsync(h_res, d_snodes, (size_t)(2 * S_ + 2 * (S_ + 1) + 2 * S_) * 4, hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipMemcpyAsync(h_res + 8 * S_ + 8, d_rootbox, (size_t)6 * S_ * 4, hipMemcpyDeviceToHost, st) != hipSuccess)
        rc = MNRT_ERR_HIP;
    if (rc == MNRT_OK) rc = sync_counters();
    if (rc != MNRT_OK || hipGetLastError() != hipSuccess || h_counters[C_ERR] != 0) {
        (void)pool_free(d_nodes, st);
        (void)pool_free(d_refs, st);
        return rc != MNRT_OK ? rc : (h_counters[C_ERR] != 0 ? MNRT_ERR_STATE : MNRT_ERR_HIP);
    }
    const uint32_t* hn = h_res;
    const uint32_t* hr = hn + S_;
    const uint32_t* hnb = hr + S_;
    const uint32_t* hrb = hnb + S_ + 1;
    const uint32_t* hl = hrb + S_ + 1;
    const uint32_t* hm = hl + S_;
    const float* hbox = (const float*)(h_res + 8 * S_ + 8);
    for (uint32_t s = 0; s < S_; ++s) {
        mnrt_tree& t = out[s];
        std::memset(&t, 0, sizeof(t));
        t.kind = (uint32_t)kind;
        t.num_nodes = hn[s];
        t.num_refs = hr[s];
        t.num_elems = h_base[s + 1] - h_base[s];
        t.num_leaves = hl[s];
        t.max_level = hm[s];
        for (int k = 0; k < 3; ++k) {
            t.root_min[k] = hbox[6 * s + k];
            t.root_max[k] = hbox[6 * s + 3 + k];
        }
        t.d_nodes = d_nodes + 2ull * hnb[s];
        t.d_refs = d_refs + hrb[s];
        t.owns_nodes = (s == 0) ? 1u : 0u;
    }
    return MNRT_OK;
}


// =========================================================================================
// device-driven build, recorded once into a hipGraph
// =========================================================================================
// The same kernels as build_batch_impl, with every count a device word (DN) and every
// buffer fixed: the level loop is unrolled to a recorded number of levels whose kernels
// find the level's active nodes / elements in counters[G_NA], [G_NE]; k_graph_level_end
// moves the loop on.  Capacities come from a host-driven build of the same scenes (its
// BuildStats) with a margin; whatever would outgrow them sets an error bit instead.

// State words of the recorded loop, after the counters.
enum : uint32_t { G_NA = C_COUNT, G_NE, G_NBN, G_NSMALL, G_NSE, G_WORDS };

// counters[C_ERR] bits of a recorded build (1, 2, 4, 8, 16, 32 as in the host-driven build):
constexpr uint32_t kErrLevelRoom = 64u;    // a level's nodes / elements outgrew the fixed buffers
constexpr uint32_t kErrLevels = 128u;      // active nodes left after the last recorded level
constexpr uint32_t kErrOutputRoom = 256u;  // the trees outgrew the output arrays
constexpr uint32_t kErrLayoutDepth = 512u; // a large-stage node deeper than the layout launches
// Levels the recorded layout covers beyond the first build's deepest large-stage node.
constexpr uint32_t kLayoutSpareLevels = 6u;

struct GraphCaps {
    uint32_t na;           // node arrays of a level (hold twice the active nodes: the next level)
    uint32_t ne;           // element arrays of a level
    uint32_t bn;           // build-node table
    uint32_t small;        // small roots
    uint32_t small_elems;  // small-root elements
    uint32_t ll;           // elements of large-stage leaves
    uint32_t nodes, refs;  // output arrays (all trees)
};

__global__ __launch_bounds__(64) void k_graph_init(uint32_t* __restrict__ c, uint32_t na, uint32_t ne,
                                                   uint32_t nbn, uint32_t nsmall, uint32_t nse) {
    for (uint32_t i = threadIdx.x; i < G_WORDS; i += blockDim.x)
        c[i] = i == G_NA ? na : i == G_NE ? ne : i == G_NBN ? nbn : i == G_NSMALL ? nsmall : i == G_NSE ? nse : 0u;
}

// End of a recorded level: the running totals, and the next level's counts after the
// checks the host-driven loop makes before growing its buffers.  On any error the loop
// stops and the small stage takes no roots; emission, records and query-time nodes are
// then skipped (they test counters[C_ERR]).
__global__ __launch_bounds__(64) void k_graph_level_end(uint32_t* __restrict__ c, GraphCaps cap, uint32_t last) {
    if (threadIdx.x != 0 || blockIdx.x != 0 || c[G_NA] == 0u) return;  // G_NA == 0: no level ran
    const uint32_t nbn = c[G_NBN] + c[C_NEWNODES];
    const uint32_t nsmall = c[G_NSMALL] + c[C_NSMALL_NEW];
    const uint32_t nse = c[G_NSE] + c[C_ESMALL_NEW];
    const uint32_t na = c[C_NNEXT], ne = c[C_ENEXT];
    c[G_NBN] = nbn;
    c[G_NSMALL] = nsmall;
    c[G_NSE] = nse;
    uint32_t err = c[C_ERR];
    if (na > 0u) {
        const bool fits = 2ull * na <= cap.na && ne <= cap.ne && nbn + 14ull * na <= cap.bn &&
                          nsmall + 2ull * na <= cap.small && nse + 2ull * ne <= cap.small_elems &&
                          (uint64_t)c[C_LLCOUNT] + ne <= cap.ll;
        if (!fits) err |= kErrLevelRoom;
        else if (last) err |= kErrLevels;
    }
    if (err != 0u) {
        c[C_ERR] = err;
        c[G_NA] = 0u;
        c[G_NE] = 0u;
        c[G_NSMALL] = 0u;
        return;
    }
    c[G_NA] = na;
    c[G_NE] = ne;
}

__global__ __launch_bounds__(64) void k_graph_copy(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src,
                                                   uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

__global__ __launch_bounds__(64) void k_graph_check_totals(uint32_t* __restrict__ c, const uint32_t* __restrict__ nodes,
                                                           const uint32_t* __restrict__ refs, GraphCaps cap) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && (*nodes == 0u || *nodes > cap.nodes || *refs > cap.refs))
        c[C_ERR] |= kErrOutputRoom;
}

// Bump carving of one device block: a first pass (base null) measures, the second carves.
class Carver {
   public:
    explicit Carver(char* base) : base_(base) {}
    template <typename T>
    T* take(uint64_t count) {
        const uint64_t bytes = (std::max<uint64_t>(count * sizeof(T), 16) + 255) & ~uint64_t(255);
        T* p = base_ ? (T*)(base_ + top_) : nullptr;
        top_ += bytes;
        return p;
    }
    uint64_t size() const { return top_; }

   private:
    char* base_;
    uint64_t top_ = 0;
};

}  // namespace

struct mnrt_build_graph {
    int kind = 0;
    uint32_t ns = 0, ntot = 0;
    std::vector<uint32_t> h_base;      // ns + 1
    std::vector<uint32_t> h_plan;      // kind / index / offset per scene
    std::vector<float> h_qr, h_roots;  // per scene (roots: empty when not given)
    mnrt_build_params prm{};
    const float *v0 = nullptr, *v1 = nullptr, *v2 = nullptr, *pow4 = nullptr, *dir4 = nullptr;
    BuildStats stats;
    GraphCaps cap{};
    uint32_t levels = 0;
    float margin = 0.25f;
    char* dbuf = nullptr;
    uint64_t dbytes = 0;
    uint32_t* counters = nullptr;
    uint32_t* d_tab = nullptr;
    uint32_t* nodes = nullptr;
    uint32_t* refs = nullptr;
    float* recs = nullptr;
    uint32_t* qnodes = nullptr;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

// Per-scene device table of a recorded build, as build_batch_impl lays it out.
struct GraphTab {
    uint32_t *base, *pkind, *pindex, *poff;
    float *qr, *given, *rootbox;
    uint32_t *snodes, *srefs, *nbase, *rbase, *sleaves, *smaxlvl;
};

GraphTab graph_tab(uint32_t* t, uint32_t S_) {
    GraphTab T;
    T.base = t;
    T.pkind = T.base + S_ + 1;
    T.pindex = T.pkind + S_;
    T.poff = T.pindex + S_;
    T.qr = (float*)(T.poff + S_);
    T.given = T.qr + S_;
    T.rootbox = T.given + 6 * S_;
    T.snodes = (uint32_t*)(T.rootbox + 6 * S_);
    T.srefs = T.snodes + S_;
    T.nbase = T.srefs + S_;
    T.rbase = T.nbase + S_ + 1;
    T.sleaves = T.rbase + S_ + 1;
    T.smaxlvl = T.sleaves + S_;
    return T;
}

// Carves g's device block (base null: measures it) and, when st is given, enqueues the
// whole build on st (which is capturing).  Every launch is a kernel: the sequence has no
// memset, copy or allocation nodes.
int graph_record(mnrt_build_graph& g, char* base, hipStream_t st) {
    const uint32_t S_ = g.ns, ntot = g.ntot;
    const int kind = g.kind;
    const GraphCaps C = g.cap;
    const uint32_t NA = C.na, NE = C.ne, CH = NE / kChunk + NA;
    Carver cv(base);
    uint32_t* counters = cv.take<uint32_t>(G_WORDS);
    uint32_t* tab = cv.take<uint32_t>(32ull * S_ + 64);
    const GraphTab T = graph_tab(tab, S_);
    auto elems = [&](uint64_t cap, bool with_node) {
        ElemView V{};
        V.idx = cv.take<uint32_t>(cap);
        V.lo = cv.take<float4>(cap);
        V.hi = kind == MNRT_KIND_TRI ? cv.take<float4>(cap) : V.lo;
        V.node = with_node ? cv.take<uint32_t>(cap) : nullptr;
        return V;
    };
    auto nodesv = [&](uint64_t cap) {
        NodeView V{};
        V.lo = cv.take<float4>(cap);
        V.hi = cv.take<float4>(cap);
        V.first = cv.take<uint32_t>(cap);
        V.cnt = cv.take<uint32_t>(cap);
        V.level = cv.take<uint32_t>(cap);
        V.bn = cv.take<uint32_t>(cap);
        V.scene = cv.take<uint32_t>(cap);
        return V;
    };
    ElemView Ein = elems(std::max<uint32_t>(ntot, 1), true);
    const uint32_t ch0 = ntot / kChunk + S_;
    uint32_t* tb = cv.take<uint32_t>(3ull * ch0 + 8ull * S_ + scan_scratch_words(S_) + 16);
    uint32_t* bnb = cv.take<uint32_t>((uint64_t)kBnCols * C.bn);
    const BnView B{bnb,
                   bnb + C.bn,
                   bnb + 2ull * C.bn,
                   bnb + 3ull * C.bn,
                   bnb + 4ull * C.bn,
                   bnb + 5ull * C.bn,
                   bnb + 6ull * C.bn,
                   bnb + 7ull * C.bn};
    SmallView SV{};
    SV.first = cv.take<uint32_t>(C.small);
    SV.cnt = cv.take<uint32_t>(C.small);
    SV.level = cv.take<uint32_t>(C.small);
    SV.bn = cv.take<uint32_t>(C.small);
    SV.lo = cv.take<float4>(C.small);
    SV.hi = cv.take<float4>(C.small);
    SV.scene = cv.take<uint32_t>(C.small);
    const ElemView Sm = elems(C.small_elems, false);
    const NodeView Nv[2] = {nodesv(NA), nodesv(NA)};
    const ElemView Xv[2] = {elems(NE, true), elems(NE, true)};
    uint32_t* ll_idx = cv.take<uint32_t>(C.ll);
    // level buffers (fixed strides: NA nodes, NE elements)
    uint32_t* cc = cv.take<uint32_t>(NA);
    uint32_t* chn = cv.take<uint32_t>(CH);
    uint32_t* chf = cv.take<uint32_t>(CH);
    uint32_t* chc = cv.take<uint32_t>(CH);
    uint32_t* tlo = cv.take<uint32_t>(3ull * NA);
    uint32_t* thi = cv.take<uint32_t>(3ull * NA);
    PlanView P{};
    P.newcnt = cv.take<uint32_t>(NA);
    P.axis = cv.take<uint32_t>(NA);
    P.split = cv.take<float>(NA);
    P.lo = cv.take<float4>(NA);
    P.hi = cv.take<float4>(NA);
    P.level = cv.take<uint32_t>(NA);
    P.childL = cv.take<uint32_t>(NA);
    P.cur = cv.take<uint32_t>(NA);
    P.leaf = cv.take<uint32_t>(NA);
    P.noprog = cv.take<uint32_t>(NA);
    const uint64_t FLW = (kind == MNRT_KIND_TRI ? 4ull : 2ull) * NE;  // element flags (+ straddle flags)
    uint32_t* fl = cv.take<uint32_t>(FLW + 1);  // after the scan: the clip queue
    uint32_t* pos = cv.take<uint32_t>(FLW + 1);
    uint32_t* cbuf = cv.take<uint32_t>(22ull * NA);
    ChildView Cv{};
    {
        const uint64_t m = 2ull * NA;
        Cv.type = cbuf;
        Cv.cnt = cbuf + m;
        Cv.fpos = cbuf + 2 * m;
        Cv.sflag = cbuf + 3 * m;
        Cv.lflag = cbuf + 4 * m;
        Cv.scnt = cbuf + 5 * m;
        Cv.lcnt = cbuf + 6 * m;
        Cv.sidx = cbuf + 7 * m;
        Cv.lidx = cbuf + 8 * m;
        Cv.soff = cbuf + 9 * m;
        Cv.loff = cbuf + 10 * m;
    }
    uint32_t* scr = cv.take<uint32_t>(std::max<uint64_t>({(uint64_t)dscan_tiles(NA), (uint64_t)dscan_tiles((uint32_t)FLW),
                                                           4ull * dscan_tiles(2 * NA), (uint64_t)dscan_tiles(C.small)}) +
                                      16);
    // small stage
    const StageCap st_cap = stage_cap_setting();
    const uint64_t st_total = (uint64_t)st_cap.mul * C.small_elems + (uint64_t)st_cap.add * C.small;
    if (st_total > 0xFFFFFFFFull) return MNRT_ERR_CAPACITY;
    SmallOut O{};
    O.nsize = cv.take<uint32_t>(C.small);
    O.rsize = cv.take<uint32_t>(C.small);
    O.mlvl = cv.take<uint32_t>(C.small);
    uint32_t* st_off = cv.take<uint32_t>(C.small);
    O.st_off = st_off;
    O.st_ovf = cv.take<uint32_t>(C.small);
    O.st_nodes = cv.take<uint2>(st_total);
    O.st_refs = cv.take<uint32_t>(st_total);
    O.st_cap = st_cap;
    O.counters = counters;
    // layout
    LayoutView Lv{};
    Lv.size = cv.take<uint32_t>(C.bn);
    Lv.rsize = cv.take<uint32_t>(C.bn);
    Lv.depth = cv.take<uint32_t>(C.bn);
    Lv.addr = cv.take<uint32_t>(C.bn);
    Lv.rbase = cv.take<uint32_t>(C.bn);
    uint32_t* lvb = cv.take<uint32_t>(64 + 65 + 64 + (uint64_t)C.bn + scan_scratch_words(std::max<uint32_t>(64, S_ + 1)));
    // outputs
    uint32_t* nodes = cv.take<uint32_t>(2ull * C.nodes);
    uint32_t* refs = cv.take<uint32_t>(C.refs);
    float* recs = cv.take<float>((uint64_t)C.refs * (kind == MNRT_KIND_TRI ? 12u : 8u));
    uint32_t* qnodes = kind == MNRT_KIND_POINT ? cv.take<uint32_t>(2ull * C.nodes) : nullptr;
    if (!base) {
        g.dbytes = cv.size();
        return MNRT_OK;
    }
    g.counters = counters;
    g.d_tab = tab;
    g.nodes = nodes;
    g.refs = refs;
    g.recs = recs;
    g.qnodes = qnodes;
    if (!st) return MNRT_OK;

    // ---- the recorded sequence ----
    const BuildStats& bs = g.stats;
    const double grow = 1.0 + (double)g.margin;
    auto planned = [&](uint64_t x, uint64_t slack, uint64_t cap) {
        return std::min<uint64_t>(cap, (uint64_t)((double)x * grow) + slack);
    };
    hipLaunchKernelGGL(k_graph_init, dim3(1), dim3(64), 0, st, counters, bs.na0, bs.ne0, S_, bs.nsmall0, bs.nse0);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, T.sleaves, 2 * S_, 0u);
    Geometry G{kind, (const float4*)g.v0, g.v1 ? (const float4*)g.v1 : (const float4*)g.v0,
               g.v2 ? (const float4*)g.v2 : (const float4*)g.v0, T.base, S_};
    RootPlan RP{T.pkind, T.pindex, T.poff};
    if (ntot > 0) hipLaunchKernelGGL(k_init_elems, blocks_for(ntot), dim3(256), 0, st, G, ntot, Ein);
    uint32_t* rtlo = nullptr;
    uint32_t* rthi = nullptr;
    if (g.h_roots.empty() && ntot > 0) {
        uint32_t *rchn = tb, *rchf = rchn + ch0, *rchc = rchf + ch0, *rcc = rchc + ch0;
        uint32_t* cnts = rcc + S_;
        rtlo = cnts + S_;
        rthi = rtlo + 3 * S_;
        uint32_t* sc = rthi + 3 * S_;
        hipLaunchKernelGGL(k_scene_counts, blocks_for(S_), dim3(256), 0, st, S_, T.base, cnts);
        hipLaunchKernelGGL(k_chunk_counts, blocks_for(S_), dim3(256), 0, st, hn(S_), cnts, rcc, rtlo, rthi);
        MNRT_HIP_CHECK(scan_exclusive(rcc, rcc, S_, counters + C_NCHUNKS, sc, st));
        hipLaunchKernelGGL(k_chunk_fill, blocks_for(ch0), dim3(256), 0, st, hn(S_), ch0, rcc, counters, T.base, cnts,
                           rchn, rchf, rchc);
        hipLaunchKernelGGL(k_tight, dim3((ch0 + kTightChunks - 1) / kTightChunks), dim3(kChunk), 0, st, ch0, counters,
                           rchn, rchf, rchc, Ein, kind, rtlo, rthi);
    }
    hipLaunchKernelGGL(k_roots_setup, blocks_for(S_), dim3(256), 0, st, S_, T.base, rtlo, rthi,
                       g.h_roots.empty() ? (const float*)nullptr : (const float*)T.given, RP, Nv[0], SV, B, T.rootbox);
    if (ntot > 0)
        hipLaunchKernelGGL(k_partition_elems, blocks_for(ntot), dim3(256), 0, st, ntot, Ein, kind, RP, Xv[0], Sm);
    MNRT_HIP_CHECK(hipGetLastError());

    // large-node stage: g.levels recorded levels
    const DN na = dn(counters + G_NA, NA), ne = dn(counters + G_NE, NE);
    const DN na2 = dn(counters + G_NA, 2 * NA, 1), ne2 = dn(counters + G_NE, 2 * NE, 1);
    const DN nbn = dn(counters + G_NBN, C.bn), nsm = dn(counters + G_NSMALL, C.small);
    const DN nse = dn(counters + G_NSE, C.small_elems);
    for (uint32_t l = 0; l < g.levels; ++l) {
        const NodeView& A = Nv[l & 1u];
        const NodeView& An = Nv[(l + 1) & 1u];
        const ElemView& E = Xv[l & 1u];
        const ElemView& Nx = Xv[(l + 1) & 1u];
        const size_t pl = std::min<size_t>(l, bs.level_na.size() - 1);
        const uint64_t pna = planned(bs.level_na[pl], 256, NA), pne = planned(bs.level_ne[pl], 4096, NE);
        const dim3 gna = blocks_for(pna), gna2 = blocks_for(2 * pna), gne = blocks_for(pne);
        const dim3 gch((uint32_t)std::max<uint64_t>(1, (pne / kChunk + pna + kTightChunks - 1) / kTightChunks));
        hipLaunchKernelGGL(k_chunk_counts, gna, dim3(256), 0, st, na, A.cnt, cc, tlo, thi);
        MNRT_HIP_CHECK(dscan1(cc, cc, na, counters + C_NCHUNKS, false, scr, st));
        hipLaunchKernelGGL(k_chunk_fill, blocks_for(pne / kChunk + pna), dim3(256), 0, st, na, CH, cc, counters,
                           A.first, A.cnt, chn, chf, chc);
        hipLaunchKernelGGL(k_tight, gch, dim3(kChunk), 0, st, CH, counters, chn, chf, chc, E, kind, tlo, thi);
        hipLaunchKernelGGL((k_plan<false>), gna, dim3(256), 0, st, na, A, tlo, thi, g.prm.empty_space_ratio, P,
                           (const uint32_t*)nullptr, hn(0u), B, counters);
        MNRT_HIP_CHECK(dscan1(P.newcnt, P.newcnt, na, counters + C_NEWNODES, false, scr, st));
        hipLaunchKernelGGL((k_plan<true>), gna, dim3(256), 0, st, na, A, tlo, thi, g.prm.empty_space_ratio, P,
                           (const uint32_t*)P.newcnt, nbn, B, counters);
        hipLaunchKernelGGL(k_classify, gne, dim3(256), 0, st, ne, E, kind, P, fl);
        MNRT_HIP_CHECK(dscan1(fl, pos, kind == MNRT_KIND_TRI ? dn(counters + G_NE, 4 * NE, 2) : ne2, nullptr, true,
                              scr, st));
        hipLaunchKernelGGL(k_child_plan, gna2, dim3(256), 0, st, na, ne, A, P, pos, g.prm.small_node_max, Cv);
        {
            DScanSet cs{};
            const uint32_t* ins[4] = {Cv.sflag, Cv.lflag, Cv.scnt, Cv.lcnt};
            uint32_t* outs[4] = {Cv.sidx, Cv.lidx, Cv.soff, Cv.loff};
            uint32_t* tots[4] = {counters + C_NSMALL_NEW, counters + C_NNEXT, counters + C_ESMALL_NEW,
                                 counters + C_ENEXT};
            for (int y = 0; y < 4; ++y) {
                cs.in[y] = ins[y];
                cs.out[y] = outs[y];
                cs.total[y] = tots[y];
            }
            MNRT_HIP_CHECK(dscan(cs, 4, na2, scr, st));
        }
        hipLaunchKernelGGL(k_child_write, gna2, dim3(256), 0, st, na, A, P, Cv, nsm, nse, SV, An, B, counters, E,
                           ll_idx, C.ll);
        hipLaunchKernelGGL(k_scatter, gne, dim3(256), 0, st, na, ne, E, G, A, P, Cv, pos, nse, Sm, C.small_elems, Nx,
                           NE, counters, fl);
        if (kind == MNRT_KIND_TRI)
            hipLaunchKernelGGL(k_clip, clip_blocks(pne), dim3(kClipBlock), 0, st, na, ne, E, G, A, P, Cv, (const uint32_t*)pos,
                               nse, Sm, Nx, (const uint32_t*)fl);
        hipLaunchKernelGGL(k_graph_level_end, dim3(1), dim3(64), 0, st, counters, C, l + 1 == g.levels ? 1u : 0u);
        MNRT_HIP_CHECK(hipGetLastError());
    }

    // small-node stage
    const dim3 small_grid((C.small + kSmallWaves - 1) / kSmallWaves), small_block(64 * kSmallWaves);
    hipLaunchKernelGGL(k_small_caps, blocks_for(C.small), dim3(256), 0, st, nsm, (const uint32_t*)SV.cnt, st_cap,
                       st_off);
    MNRT_HIP_CHECK(dscan1(st_off, st_off, nsm, nullptr, false, scr, st));
    if (kind == MNRT_KIND_TRI)
        hipLaunchKernelGGL((k_small<kSmallStage, MNRT_KIND_TRI>), small_grid, small_block, 0, st, nsm, SV, Sm,
                           (const float*)T.qr, O);
    else
        hipLaunchKernelGGL((k_small<kSmallStage, MNRT_KIND_POINT>), small_grid, small_block, 0, st, nsm, SV, Sm,
                           (const float*)T.qr, O);
    // layout
    {
        uint32_t* hist = lvb;
        uint32_t* off = lvb + 64;
        uint32_t* cursor = off + 65;
        uint32_t* list = cursor + 64;
        uint32_t* sc = list + C.bn;
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, hist, 64u, 0u);
        // the layout launches cover the first build's depth and a few levels more; a deeper
        // replay sets kErrLayoutDepth here, before the layout, so nothing is emitted
        const uint32_t top = std::min<uint32_t>({bs.maxlvl + kLayoutSpareLevels, (uint32_t)MNRT_MAX_LEVEL, 63u});
        hipLaunchKernelGGL(k_level_hist, blocks_for(C.bn), dim3(256), 0, st, nbn, B, hist, top, counters + C_ERR,
                           kErrLayoutDepth);
        MNRT_HIP_CHECK(scan_exclusive(hist, off, 64, off + 64, sc, st));
        hipLaunchKernelGGL(k_graph_copy, dim3(1), dim3(64), 0, st, cursor, (const uint32_t*)off, 64u);
        hipLaunchKernelGGL(k_level_scatter, blocks_for(C.bn), dim3(256), 0, st, nbn, B, cursor, list);
        LayoutSrc Ls{off, list, O.nsize, O.rsize, O.mlvl};
        const dim3 lgrid(std::min<uint32_t>(blocks_for(C.bn).x, 512u));
        for (int lvl = (int)top; lvl >= 0; --lvl)
            hipLaunchKernelGGL(k_layout_up, lgrid, dim3(256), 0, st, (uint32_t)lvl, Ls, B, Lv);
        for (uint32_t lvl = 0; lvl <= top; ++lvl)
            hipLaunchKernelGGL(k_layout_down, lgrid, dim3(256), 0, st, lvl, Ls, B, Lv);
        hipLaunchKernelGGL(k_scene_sizes, blocks_for(S_), dim3(256), 0, st, S_, Lv, T.snodes, T.srefs, T.sleaves,
                           T.smaxlvl);
        MNRT_HIP_CHECK(scan_exclusive(T.snodes, T.nbase, S_, T.nbase + S_, sc, st));
        MNRT_HIP_CHECK(scan_exclusive(T.srefs, T.rbase, S_, T.rbase + S_, sc, st));
        hipLaunchKernelGGL(k_graph_check_totals, dim3(1), dim3(64), 0, st, counters, (const uint32_t*)(T.nbase + S_),
                           (const uint32_t*)(T.rbase + S_), C);
    }
    // emission (each kernel returns at once when counters[C_ERR] is set)
    hipLaunchKernelGGL(k_emit_large, blocks_for(C.bn), dim3(256), 0, st, nbn, B, Lv, (const uint32_t*)ll_idx, nodes,
                       refs, counters, C.nodes, C.refs, (const uint32_t*)T.nbase, (const uint32_t*)T.rbase);
    O.bn_addr = Lv.addr;
    O.bn_rbase = Lv.rbase;
    O.nodes = nodes;
    O.refs = refs;
    O.max_nodes = C.nodes;
    O.max_refs = C.refs;
    O.node_base = T.nbase;
    O.ref_base = T.rbase;
    hipLaunchKernelGGL(k_small_place, small_grid, small_block, 0, st, nsm, SV, O);
    // subtrees that outgrew their staging area, built again at their final address
    if (kind == MNRT_KIND_TRI)
        hipLaunchKernelGGL((k_small<kSmallEmit, MNRT_KIND_TRI>), small_grid, small_block, 0, st, nsm, SV, Sm,
                           (const float*)T.qr, O);
    else
        hipLaunchKernelGGL((k_small<kSmallEmit, MNRT_KIND_POINT>), small_grid, small_block, 0, st, nsm, SV, Sm,
                           (const float*)T.qr, O);
    // records and query-time node words
    const DN nrefs = dn(T.rbase + S_, C.refs), nnodes = dn(T.nbase + S_, C.nodes);
    if (kind == MNRT_KIND_TRI) {
        int rc = launch_records_batch(kind, refs, nrefs, T.rbase, T.base, S_, (const float4*)g.v0, (const float4*)g.v1,
                                      (const float4*)g.v2, (float4*)recs, counters + C_ERR, st);
        if (rc != MNRT_OK) return rc;
    } else if (g.pow4 && g.dir4) {
        int rc = launch_records_batch(kind, refs, nrefs, T.rbase, T.base, S_, (const float4*)g.v0,
                                      (const float4*)g.pow4, (const float4*)g.dir4, (float4*)recs, counters + C_ERR,
                                      st);
        if (rc == MNRT_OK)
            rc = launch_collapse_batch((const uint2*)nodes, nnodes, T.nbase, S_, kQueryLeafRefs, (uint2*)qnodes,
                                       counters + C_ERR, st);
        if (rc != MNRT_OK) return rc;
    }
    MNRT_HIP_CHECK(hipGetLastError());
    return MNRT_OK;
}

}  // namespace

extern "C" int mnrt_build_tri(const float* d_v0, const float* d_v1, const float* d_v2, uint32_t n,
                              const mnrt_build_params* params, mnrt_tree* out, void* stream) {
    if (!params || !out) return MNRT_ERR_ARG;
    uint32_t base[2] = {0, n};
    const float* root = params->has_root_box ? params->root_min : nullptr;
    float r6[6];
    if (root) {
        for (int k = 0; k < 3; ++k) {
            r6[k] = params->root_min[k];
            r6[3 + k] = params->root_max[k];
        }
    }
    int rc = build_batch_impl(MNRT_KIND_TRI, d_v0, d_v1, d_v2, base, 1, params, root ? r6 : nullptr, nullptr, out,
                              (hipStream_t)stream);
    if (rc != MNRT_OK) return rc;
    rc = mnrt_tree_attach_tri_records(out, d_v0, d_v1, d_v2, stream);
    if (rc != MNRT_OK) mnrt_tree_release(out);
    return rc;
}

extern "C" int mnrt_build_points(const float* d_pos4, uint32_t n, const mnrt_build_params* params, mnrt_tree* out,
                                 void* stream) {
    if (!params || !out) return MNRT_ERR_ARG;
    uint32_t base[2] = {0, n};
    float r6[6];
    for (int k = 0; k < 3; ++k) {
        r6[k] = params->root_min[k];
        r6[3 + k] = params->root_max[k];
    }
    return build_batch_impl(MNRT_KIND_POINT, d_pos4, nullptr, nullptr, base, 1, params,
                            params->has_root_box ? r6 : nullptr, nullptr, out, (hipStream_t)stream);
}

extern "C" int mnrt_build_tri_batch(const float* d_v0, const float* d_v1, const float* d_v2,
                                    const uint32_t* h_offsets, uint32_t num_scenes, const mnrt_build_params* params,
                                    const float* h_root_boxes, mnrt_tree* out, void* stream) {
    int rc = build_batch_impl(MNRT_KIND_TRI, d_v0, d_v1, d_v2, h_offsets, num_scenes, params, h_root_boxes, nullptr,
                              out, (hipStream_t)stream);
    if (rc != MNRT_OK) return rc;
    rc = mnrt_tree_attach_tri_records_batch(out, num_scenes, d_v0, d_v1, d_v2, h_offsets, stream);
    if (rc != MNRT_OK) mnrt_tree_release(out);
    return rc;
}

extern "C" int mnrt_build_points_batch(const float* d_pos4, const uint32_t* h_offsets, uint32_t num_scenes,
                                       const mnrt_build_params* params, const float* h_root_boxes,
                                       const float* h_max_query_radius, mnrt_tree* out, void* stream) {
    return build_batch_impl(MNRT_KIND_POINT, d_pos4, nullptr, nullptr, h_offsets, num_scenes, params, h_root_boxes,
                            h_max_query_radius, out, (hipStream_t)stream);
}

// ---- device-driven batched build (mnrt_build_graph_*) ------------------------------------

extern "C" int mnrt_build_graph_create(const mnrt_build_graph_desc* d, mnrt_tree* out, mnrt_build_graph** g_out,
                                       void* stream) {
    using namespace mnrt;
    if (!d || !out || !g_out || !d->params || !d->h_offsets || d->num_scenes == 0) return MNRT_ERR_ARG;
    if (d->kind != MNRT_KIND_TRI && d->kind != MNRT_KIND_POINT) return MNRT_ERR_ARG;
    if (!(d->margin >= 0.0f && d->margin <= 16.0f)) return MNRT_ERR_ARG;
    *g_out = nullptr;
    const uint32_t S_ = d->num_scenes;
    const uint32_t ntot = d->h_offsets[S_];
    if (ntot > 0 && (!d->d_v0 || (d->kind == MNRT_KIND_TRI && (!d->d_v1 || !d->d_v2)))) return MNRT_ERR_ARG;
    if (d->kind == MNRT_KIND_POINT && (!d->d_power4) != (!d->d_dir4)) return MNRT_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    auto* g = new (std::nothrow) mnrt_build_graph;
    if (!g) return MNRT_ERR_NOMEM;
    auto fail = [&](int rc) {
        (void)mnrt_build_graph_destroy(g);
        return rc;
    };
    g->kind = d->kind;
    g->ns = S_;
    g->ntot = ntot;
    g->prm = *d->params;
    g->v0 = d->d_v0;
    g->v1 = d->d_v1;
    g->v2 = d->d_v2;
    g->pow4 = d->d_power4;
    g->dir4 = d->d_dir4;
    g->margin = d->margin;
    g->h_base.assign(d->h_offsets, d->h_offsets + S_ + 1);
    if (d->h_root_boxes) g->h_roots.assign(d->h_root_boxes, d->h_root_boxes + 6ull * S_);
    g->h_qr.assign(S_, d->params->max_query_radius);
    if (d->h_max_query_radius)
        for (uint32_t s = 0; s < S_; ++s) g->h_qr[s] = d->h_max_query_radius[s];

    // 1. a host-driven build of the same scenes: sizes for the fixed buffers
    {
        std::vector<mnrt_tree> tmp(S_);
        int rc = build_batch_impl(d->kind, d->d_v0, d->d_v1, d->d_v2, d->h_offsets, S_, d->params, d->h_root_boxes,
                                  d->h_max_query_radius, tmp.data(), st, &g->stats);
        if (rc != MNRT_OK) return fail(rc);
        (void)hipStreamSynchronize(st);
        (void)mnrt_tree_release(&tmp[0]);
    }
    const BuildStats& bs = g->stats;
    const double grow = 1.0 + (double)d->margin;
    auto cap = [&](uint64_t x, uint64_t slack) -> uint32_t {
        return (uint32_t)std::min<uint64_t>(MNRT_INDEX_MASK, (uint64_t)((double)x * grow) + slack);
    };
    GraphCaps& C = g->cap;
    C.na = cap(std::max<uint64_t>(bs.max_2na, 2ull * bs.na0), 64);
    C.ne = cap(std::max<uint64_t>(bs.max_ne, bs.ne0), 1024);
    C.bn = cap(std::max<uint64_t>({bs.max_bn, (uint64_t)bs.nbn, (uint64_t)S_}), 1024);
    C.small = cap(std::max<uint64_t>({bs.max_small, (uint64_t)bs.nsmall, (uint64_t)bs.nsmall0}), 256);
    C.small_elems = cap(std::max<uint64_t>({bs.max_se, (uint64_t)bs.nse, (uint64_t)bs.nse0}), 1024);
    C.ll = cap(bs.max_ll, 1024);
    C.nodes = cap(bs.total_nodes, 1024);
    C.refs = cap(bs.total_refs, 1024);
    g->levels = bs.level_na.empty() ? 0u
                                    : std::min<uint32_t>((uint32_t)bs.level_na.size() + 2u, MNRT_MAX_LEVEL + 1u);
    // the root plan, as build_batch_impl makes it
    g->h_plan.assign(3ull * S_, 0u);
    {
        uint32_t na = 0, nsmall = 0, ne = 0, nse = 0;
        for (uint32_t s = 0; s < S_; ++s) {
            const uint32_t cnt = g->h_base[s + 1] - g->h_base[s];
            if (cnt > g->prm.small_node_max) {
                g->h_plan[s] = 0;
                g->h_plan[S_ + s] = na++;
                g->h_plan[2 * S_ + s] = ne;
                ne += cnt;
            } else {
                g->h_plan[s] = 1;
                g->h_plan[S_ + s] = nsmall++;
                g->h_plan[2 * S_ + s] = nse;
                nse += cnt;
            }
        }
        if (na != bs.na0 || ne != bs.ne0 || nsmall != bs.nsmall0 || nse != bs.nse0) return fail(MNRT_ERR_STATE);
    }

    // 2. the fixed device block, and the per-scene tables in it
    int rc = graph_record(*g, nullptr, nullptr);
    if (rc != MNRT_OK) return fail(rc);
    if (hipMalloc((void**)&g->dbuf, g->dbytes) != hipSuccess) {
        g->dbuf = nullptr;
        return fail(MNRT_ERR_NOMEM);
    }
    rc = graph_record(*g, g->dbuf, nullptr);
    if (rc != MNRT_OK) return fail(rc);
    {
        const GraphTab T = graph_tab(g->d_tab, S_);
        std::vector<float> fq(S_ + 6ull * S_, 0.0f);
        for (uint32_t s = 0; s < S_; ++s) fq[s] = g->h_qr[s];
        for (size_t i = 0; i < g->h_roots.size(); ++i) fq[S_ + i] = g->h_roots[i];
        if (hipMemcpy(T.base, g->h_base.data(), 4ull * (S_ + 1), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(T.pkind, g->h_plan.data(), 12ull * S_, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(T.qr, fq.data(), 4ull * fq.size(), hipMemcpyHostToDevice) != hipSuccess)
            return fail(MNRT_ERR_HIP);
    }

    // 3. record the device-driven build
    if (hipStreamCreateWithFlags(&g->cap_stream, hipStreamNonBlocking) != hipSuccess) {
        g->cap_stream = nullptr;
        return fail(MNRT_ERR_HIP);
    }
    if (hipStreamBeginCapture(g->cap_stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return fail(MNRT_ERR_HIP);
    rc = graph_record(*g, g->dbuf, g->cap_stream);
    hipGraph_t graph = nullptr;
    const hipError_t ce = hipStreamEndCapture(g->cap_stream, &graph);
    g->graph = graph;
    if (rc != MNRT_OK) return fail(rc);
    if (ce != hipSuccess || !graph) return fail(MNRT_ERR_HIP);
    if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess) {
        g->exec = nullptr;
        return fail(MNRT_ERR_HIP);
    }

    // 4. one replay: the trees in out[], which must equal the host-driven build's
    rc = mnrt_build_graph_replay(g, stream);
    uint32_t flags = 0;
    if (rc == MNRT_OK) rc = mnrt_build_graph_status(g, out, &flags, stream);
    if (rc == MNRT_OK && flags != 0u) rc = MNRT_ERR_STATE;
    if (rc != MNRT_OK) return fail(rc);
    *g_out = g;
    return MNRT_OK;
}

extern "C" int mnrt_build_graph_replay(mnrt_build_graph* g, void* stream) {
    if (!g || !g->exec) return MNRT_ERR_ARG;
    MNRT_HIP_CHECK(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return MNRT_OK;
}

extern "C" int mnrt_build_graph_status(mnrt_build_graph* g, mnrt_tree* out, uint32_t* h_flags, void* stream) {
    using namespace mnrt;
    if (!g || !g->dbuf || !h_flags) return MNRT_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    const uint32_t S_ = g->ns;
    const GraphTab T = graph_tab(g->d_tab, S_);
    std::vector<uint32_t> c(G_WORDS), tab(8ull * S_ + 2);
    std::vector<float> box(6ull * S_);
    MNRT_HIP_CHECK(hipStreamSynchronize(st));
    MNRT_HIP_CHECK(hipMemcpy(c.data(), g->counters, 4ull * G_WORDS, hipMemcpyDeviceToHost));
    // snodes, srefs, nbase (S+1), rbase (S+1), sleaves, smaxlvl are consecutive
    MNRT_HIP_CHECK(hipMemcpy(tab.data(), T.snodes, 4ull * (6ull * S_ + 2), hipMemcpyDeviceToHost));
    MNRT_HIP_CHECK(hipMemcpy(box.data(), T.rootbox, 4ull * 6 * S_, hipMemcpyDeviceToHost));
    *h_flags = c[C_ERR];
    if (c[C_ERR] != 0u || !out) return MNRT_OK;
    const uint32_t* hn_ = tab.data();
    const uint32_t* hr = hn_ + S_;
    const uint32_t* hnb = hr + S_;
    const uint32_t* hrb = hnb + S_ + 1;
    const uint32_t* hl = hrb + S_ + 1;
    const uint32_t* hm = hl + S_;
    const uint32_t per = g->kind == MNRT_KIND_TRI ? 12u : 8u;
    const bool recs = g->kind == MNRT_KIND_TRI || (g->pow4 && g->dir4);
    for (uint32_t s = 0; s < S_; ++s) {
        mnrt_tree& t = out[s];
        std::memset(&t, 0, sizeof(t));
        t.kind = (uint32_t)g->kind;
        t.num_nodes = hn_[s];
        t.num_refs = hr[s];
        t.num_elems = g->h_base[s + 1] - g->h_base[s];
        t.num_leaves = hl[s];
        t.max_level = hm[s];
        for (int k = 0; k < 3; ++k) {
            t.root_min[k] = box[6 * s + k];
            t.root_max[k] = box[6 * s + 3 + k];
        }
        t.d_nodes = g->nodes + 2ull * hnb[s];
        t.d_refs = g->refs + hrb[s];
        t.d_recs = recs ? g->recs + (uint64_t)per * hrb[s] : nullptr;
        if (g->kind == MNRT_KIND_POINT && recs) {
            t.d_qnodes = g->qnodes + 2ull * hnb[s];
            t.qleaf_max = kQueryLeafRefs;
        }
        // the arrays belong to the graph (mnrt_build_graph_destroy frees them)
        t.owns_nodes = t.owns_recs = t.owns_qnodes = 0u;
    }
    return MNRT_OK;
}

extern "C" int mnrt_build_graph_destroy(mnrt_build_graph* g) {
    if (!g) return MNRT_OK;
    (void)hipDeviceSynchronize();
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->cap_stream) (void)hipStreamDestroy(g->cap_stream);
    if (g->dbuf) (void)hipFree(g->dbuf);
    delete g;
    return MNRT_OK;
}
