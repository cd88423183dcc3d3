// order.h — processing order for incoherent kNN queries: a counting sort by spatial cell,
// kernels only plus one scan.
//
// Points are grouped by an 18-bit Morton cell (6 bits per axis) of the tree's root box,
// cells visited in code order, so a wave's consecutive queries are close (shared leaves
// in cache, the warm start).  The order inside a cell is whatever the atomics give, which
// changes no result: every query's set is exact and written to the query's own slot.
// Queries that are coherent already — fewer than half of the consecutive pairs farther
// apart than a cell diagonal, e.g. hits of camera rays in tile order — keep their order:
// inside a cell the counting sort would scatter them, and their many queries per cell
// would make the histogram atomics collide.  (Ordering diffuse bounce rays the same way by
// direction octant and origin cell was measured slower, 3.0 -> 5.4 ms per 2.07M rays: it
// loses the rays' pixel-order origin coherence.)
#pragma once

#include "scan.h"

namespace mnrt {
namespace {  // one private copy per translation unit

constexpr uint32_t kOrderBits = 6;  // per axis
constexpr uint32_t kOrderCells = 1u << (3 * kOrderBits);
constexpr uint32_t kOrderMinItems = 1u << 15;  // smaller launches keep the caller's order
constexpr uint32_t kOrderMaxItems = 1u << 29;  // ranks share a word with the histogram copy
// Histogram copies: one per XCD (accelerator complex die), picked by the die a wave runs on,
// so a cell's counter is updated from one die's L2 only instead of bouncing between all eight.
constexpr uint32_t kOrderCopies = 8;
// 16 items per lane: with 4 (4x the workgroups) the workgroups' atomics on the one counter
// word serialise, 17 -> 27 us per 2M items
constexpr uint32_t kFarPerThread = 16;

struct OrderBox {
    float lo[3], scale[3];  // cell coordinate = (p - lo) * scale, clamped to [0, 63]
    float diag2;            // squared cell diagonal
};

__device__ __forceinline__ uint32_t spread3(uint32_t v) {  // bit i of v -> bit 3i
    uint32_t r = 0;
#pragma unroll
    for (uint32_t b = 0; b < kOrderBits; ++b) r |= ((v >> b) & 1u) << (3 * b);
    return r;
}

// Consecutive pairs that are far apart; one atomic per workgroup of kFarPerThread * 256.
__global__ __launch_bounds__(256) void k_order_far(uint32_t n, const float4* __restrict__ pos, float diag2,
                                                   uint32_t* __restrict__ far_pairs) {
    int far = 0;
    const uint32_t i0 = blockIdx.x * (256u * kFarPerThread) + threadIdx.x;
    for (uint32_t t = 0; t < kFarPerThread; ++t) {
        const uint32_t i = i0 + t * 256u;
        if (i + 1 < n) {
            const float4 a = pos[i], b = pos[i + 1];
            const float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
            far += !(dx * dx + dy * dy + dz * dz <= diag2) ? 1 : 0;
        }
    }
    __shared__ int s_far[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) far += __shfl_xor(far, o, 64);
    if ((threadIdx.x & 63u) == 0) s_far[threadIdx.x >> 6] = far;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int tot = s_far[0] + s_far[1] + s_far[2] + s_far[3];
        if (tot > 0) atomicAdd(far_pairs, (uint32_t)tot);
    }
}

__device__ __forceinline__ bool order_coherent(uint32_t n, const uint32_t* far_pairs) {
    return 2ull * *far_pairs < (uint64_t)n;
}

// Morton cell of a point (NaN coordinates land in cell 0: fmaxf / fminf return the other
// operand for NaN).
__device__ __forceinline__ uint32_t order_cell(const float4 p, const OrderBox& B) {
    const float c[3] = {p.x, p.y, p.z};
    uint32_t code = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float t = fminf(fmaxf((c[a] - B.lo[a]) * B.scale[a], 0.0f), (float)((1u << kOrderBits) - 1u));
        code |= spread3((uint32_t)t) << a;
    }
    return code;
}

// Die (XCD) the calling wave runs on, 0..kOrderCopies-1.
__device__ __forceinline__ uint32_t order_copy() {
    uint32_t x;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(x));
    return x & (kOrderCopies - 1u);
}

// Histograms of the cells, one per die, for incoherent items only; each item keeps its rank
// inside its cell's copy (the value its atomic returned) and the copy in bits 29..31, so
// placing needs no second round of atomics.
__global__ __launch_bounds__(256) void k_order_cells(uint32_t n, const float4* __restrict__ pos, OrderBox B,
                                                     const uint32_t* __restrict__ far_pairs,
                                                     uint32_t* __restrict__ rank, uint32_t* __restrict__ hist) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || order_coherent(n, far_pairs)) return;
    const uint32_t x = order_copy();
    rank[i] = atomicAdd(&hist[x * kOrderCells + order_cell(pos[i], B)], 1u) | x << 29;
}

// Per cell: copy x's count becomes the sum of copies 0..x-1 (copy 0's the cell total, the
// scan's input; copy 0 itself starts at 0).
__global__ __launch_bounds__(256) void k_order_sum(uint32_t n, const uint32_t* __restrict__ far_pairs,
                                                   uint32_t* __restrict__ hist) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= kOrderCells || order_coherent(n, far_pairs)) return;
    uint32_t s = hist[c];
#pragma unroll
    for (uint32_t x = 1; x < kOrderCopies; ++x) {
        const uint32_t h = hist[x * kOrderCells + c];
        hist[x * kOrderCells + c] = s;
        s += h;
    }
    hist[c] = s;
}

// Slot of item i: its cell's start (exclusive scan of the cell totals), the counts of the
// copies before its own, and its rank in its copy.
__global__ __launch_bounds__(256) void k_order_place(uint32_t n, const float4* __restrict__ pos, OrderBox B,
                                                     const uint32_t* __restrict__ rank,
                                                     const uint32_t* __restrict__ far_pairs,
                                                     const uint32_t* __restrict__ hist, uint32_t* __restrict__ order) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (order_coherent(n, far_pairs)) {
        order[i] = i;
        return;
    }
    const uint32_t c = order_cell(pos[i], B), rk = rank[i], x = rk >> 29;
    order[hist[c] + (x ? hist[x * kOrderCells + c] : 0u) + (rk & ((1u << 29) - 1u))] = i;
}

// Zeroes the cell histograms and the far-pair counter (a kernel rather than memset nodes, so
// a caller may capture the sequence in a graph).
__global__ __launch_bounds__(256) void k_order_clear(uint32_t* __restrict__ hist, uint32_t* __restrict__ far_pairs) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < kOrderCopies * kOrderCells;
         i += gridDim.x * blockDim.x)
        hist[i] = 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) *far_pairs = 0u;
}

// Device words spatial_order_into needs for n points (0 outside kOrderMinItems ..
// kOrderMaxItems - 1: no order).
inline size_t spatial_order_words(uint32_t n) {
    if (n < kOrderMinItems || n >= kOrderMaxItems) return 0;
    return 2ull * n + (size_t)kOrderCopies * kOrderCells + 4 + scan_scratch_words(kOrderCells);
}

// Processing order of n points at d_pos4 in the box lo3 .. hi3, built in buf
// (spatial_order_words(n) device words): *order = null when spatial_order_words(n) is 0, else a pointer
// into buf valid until buf is reused; *far (when given) the device word order_coherent reads
// (null with no order).
inline int spatial_order_into(const float* lo3, const float* hi3, uint32_t n, const float* d_pos4, hipStream_t st,
                              uint32_t* buf, const uint32_t** order, const uint32_t** far = nullptr) {
    *order = nullptr;
    if (far) *far = nullptr;
    if (spatial_order_words(n) == 0) return MNRT_OK;
    uint32_t* rank = buf;
    uint32_t* ord = rank + n;
    uint32_t* hist = ord + n;  // kOrderCopies histograms; the first becomes the cell starts
    uint32_t* far_pairs = hist + kOrderCopies * kOrderCells;
    uint32_t* scr = far_pairs + 4;
    OrderBox B;
    B.diag2 = 0.0f;
    for (int a = 0; a < 3; ++a) {
        const float ext = hi3[a] - lo3[a];
        B.lo[a] = lo3[a];
        B.scale[a] = ext > 0.0f ? (float)(1u << kOrderBits) / ext : 0.0f;
        const float c = ext / (float)(1u << kOrderBits);
        B.diag2 += c * c;
    }
    const float4* pos = (const float4*)d_pos4;
    hipLaunchKernelGGL(k_order_clear, dim3(kOrderCopies * kOrderCells / 256 / 16), dim3(256), 0, st, hist, far_pairs);
    hipLaunchKernelGGL(k_order_far, dim3((n + 256 * kFarPerThread - 1) / (256 * kFarPerThread)), dim3(256), 0, st, n,
                       pos, B.diag2, far_pairs);
    hipLaunchKernelGGL(k_order_cells, dim3((n + 255) / 256), dim3(256), 0, st, n, pos, B,
                       (const uint32_t*)far_pairs, rank, hist);
    hipLaunchKernelGGL(k_order_sum, dim3(kOrderCells / 256), dim3(256), 0, st, n, (const uint32_t*)far_pairs, hist);
    MNRT_HIP_CHECK(scan_exclusive(hist, hist, kOrderCells, nullptr, scr, st));
    hipLaunchKernelGGL(k_order_place, dim3((n + 255) / 256), dim3(256), 0, st, n, pos, B, (const uint32_t*)rank,
                       (const uint32_t*)far_pairs, (const uint32_t*)hist, ord);
    MNRT_HIP_CHECK(hipGetLastError());
    *order = ord;
    if (far) *far = far_pairs;
    return MNRT_OK;
}

// As spatial_order_into with the buffer from the device pool: *buf = the allocation to
// return with pool_free after the launch that reads the order (null with no order).
inline int spatial_order(const float* lo3, const float* hi3, uint32_t n, const float* d_pos4, hipStream_t st,
                         const uint32_t** order, void** buf, const uint32_t** far = nullptr) {
    *order = nullptr;
    *buf = nullptr;
    if (far) *far = nullptr;
    const size_t words = spatial_order_words(n);
    if (words == 0) return MNRT_OK;
    if (pool_alloc(buf, words * 4, st) != hipSuccess) return MNRT_ERR_NOMEM;
    return spatial_order_into(lo3, hi3, n, d_pos4, st, (uint32_t*)*buf, order, far);
}

}  // namespace
}  // namespace mnrt
