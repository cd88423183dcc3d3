// scan.h — device-wide exclusive prefix sum of uint32 arrays (wave64 + LDS), used by
// the kd-tree build for stream compaction.  Replaces the CUDPP scans the reference calls
// through MNCudaPrimitives::Scan (MNCudaPrimitives.cpp:160).
#pragma once

#include <algorithm>

#include "mnrt_common.h"

namespace mnrt {
namespace {  // one private copy per translation unit

constexpr uint32_t kScanBlock = 256;
constexpr uint32_t kScanPerThread = 8;
constexpr uint32_t kScanTile = kScanBlock * kScanPerThread;  // 2048 elements per block

// Inclusive scan of one value per lane across a wave64.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t lane) {
#pragma unroll
    for (uint32_t o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive scan of one 2048-element tile per block; writes the tile total to sums, and
// to *total (may be null) when the scan is one tile.
__global__ __launch_bounds__(kScanBlock) void k_scan_tiles(const uint32_t* in, uint32_t n,
                                                           uint32_t* out,
                                                           uint32_t* __restrict__ sums,
                                                           uint32_t* __restrict__ total) {
    __shared__ uint32_t s_wave[kScanBlock / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)tid * kScanPerThread;
    uint32_t v[kScanPerThread];
    uint32_t local = 0;
#pragma unroll
    for (uint32_t j = 0; j < kScanPerThread; ++j) {
        size_t i = base + j;
        v[j] = (i < n) ? in[i] : 0u;
        local += v[j];
    }
    uint32_t incl = wave_incl_scan(local, lane);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t wave_off = 0;
    for (uint32_t w = 0; w < wave; ++w) wave_off += s_wave[w];
    uint32_t run = wave_off + incl - local;
#pragma unroll
    for (uint32_t j = 0; j < kScanPerThread; ++j) {
        size_t i = base + j;
        if (i < n) out[i] = run;
        run += v[j];
    }
    if (tid == kScanBlock - 1) {
        sums[blockIdx.x] = wave_off + incl;
        if (total && gridDim.x == 1) *total = wave_off + incl;
    }
}

// Adds each tile's offset; block 0 also writes the grand total (may be null).
__global__ __launch_bounds__(kScanBlock) void k_scan_add(uint32_t* __restrict__ out, uint32_t n,
                                                         const uint32_t* __restrict__ offs,
                                                         const uint32_t* __restrict__ sums, uint32_t* __restrict__ total) {
    const size_t base = (size_t)blockIdx.x * kScanTile;
    const uint32_t add = offs[blockIdx.x];
    if (total && blockIdx.x == 0 && threadIdx.x == 0) *total = offs[gridDim.x - 1] + sums[gridDim.x - 1];
    for (uint32_t j = threadIdx.x; j < kScanTile; j += kScanBlock) {
        size_t i = base + j;
        if (i < n) out[i] += add;
    }
}

__global__ void k_scan_word(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = src ? *src : 0u;
}

// Up to kScanSetMax independent scans of n elements each, run by the same launches
// (blockIdx.y picks the scan).
constexpr int kScanSetMax = 4;
struct ScanSet {
    const uint32_t* in[kScanSetMax];
    uint32_t* out[kScanSetMax];
    uint32_t* sums[kScanSetMax];
    uint32_t* total[kScanSetMax];  // may be null
};

__global__ __launch_bounds__(kScanBlock) void k_scan_tiles_set(ScanSet S, uint32_t n) {
    const uint32_t y = blockIdx.y;
    __shared__ uint32_t s_wave[kScanBlock / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)tid * kScanPerThread;
    const uint32_t* in = S.in[y];
    uint32_t* out = S.out[y];
    uint32_t v[kScanPerThread];
    uint32_t local = 0;
#pragma unroll
    for (uint32_t j = 0; j < kScanPerThread; ++j) {
        size_t i = base + j;
        v[j] = (i < n) ? in[i] : 0u;
        local += v[j];
    }
    uint32_t incl = wave_incl_scan(local, lane);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t wave_off = 0;
    for (uint32_t w = 0; w < wave; ++w) wave_off += s_wave[w];
    uint32_t run = wave_off + incl - local;
#pragma unroll
    for (uint32_t j = 0; j < kScanPerThread; ++j) {
        size_t i = base + j;
        if (i < n) out[i] = run;
        run += v[j];
    }
    if (tid == kScanBlock - 1) {
        S.sums[y][blockIdx.x] = wave_off + incl;
        if (S.total[y] && gridDim.x == 1) *S.total[y] = wave_off + incl;
    }
}

__global__ __launch_bounds__(kScanBlock) void k_scan_add_set(ScanSet S, uint32_t n, ScanSet offs) {
    const uint32_t y = blockIdx.y;
    const size_t base = (size_t)blockIdx.x * kScanTile;
    const uint32_t add = offs.out[y][blockIdx.x];
    if (S.total[y] && blockIdx.x == 0 && threadIdx.x == 0)
        *S.total[y] = offs.out[y][gridDim.x - 1] + S.sums[y][gridDim.x - 1];
    uint32_t* out = S.out[y];
    for (uint32_t j = threadIdx.x; j < kScanTile; j += kScanBlock) {
        size_t i = base + j;
        if (i < n) out[i] += add;
    }
}

// Scans of more than one tile and at most kScanBig elements (the build's per-node scans)
// run in one workgroup: each thread takes a contiguous run of at most kScanBigPer elements,
// the run totals are scanned per wave and the wave totals through LDS, so such a scan is one
// launch instead of three.  in and out may alias (a thread reads each element of its run
// before writing it, and the runs are disjoint).
constexpr uint32_t kScanBigBlock = 1024;
constexpr uint32_t kScanBigPer = 64;
constexpr uint32_t kScanBig = kScanBigBlock * kScanBigPer;

// Returns the total in thread kScanBigBlock - 1 (and writes it to *total when given).
__device__ __forceinline__ uint32_t block_scan_runs(const uint32_t* in, uint32_t n, uint32_t* out,
                                                    uint32_t* __restrict__ total) {
    __shared__ uint32_t s_wave[kScanBigBlock / 64];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t per = (n + kScanBigBlock - 1) / kScanBigBlock;
    const uint32_t b = min(tid * per, n), e = min(b + per, n);
    uint32_t local = 0;
    for (uint32_t i = b; i < e; ++i) local += in[i];
    const uint32_t incl = wave_incl_scan(local, lane);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t wave_off = 0;
    for (uint32_t w = 0; w < wave; ++w) wave_off += s_wave[w];
    uint32_t run = wave_off + incl - local;
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t v = in[i];
        out[i] = run;
        run += v;
    }
    if (tid == kScanBigBlock - 1 && total) *total = wave_off + incl;
    return wave_off + incl;
}

__global__ __launch_bounds__(kScanBigBlock) void k_scan_big(const uint32_t* in, uint32_t n, uint32_t* out,
                                                            uint32_t* __restrict__ total) {
    block_scan_runs(in, n, out, total);
}

// blockIdx.x picks the scan of the set
__global__ __launch_bounds__(kScanBigBlock) void k_scan_big_set(ScanSet S, uint32_t n) {
    const uint32_t y = blockIdx.x;
    block_scan_runs(S.in[y], n, S.out[y], S.total[y]);
}

// Scratch words needed by scan_exclusive for n elements.
inline size_t scan_scratch_words(uint32_t n) {
    size_t words = 0;
    uint64_t m = n;
    while (m > 1) {
        m = (m + kScanTile - 1) / kScanTile;
        words += 2 * m + 2;
    }
    return words + 4;
}

// out[i] = sum(in[0..i)); *total (device word, may be null) = sum(in[0..n)).
// in and out may alias.  scratch: scan_scratch_words(n) device words.
inline hipError_t scan_exclusive(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total,
                                 uint32_t* scratch, hipStream_t st) {
    // kernels only (no memset / memcpy nodes), so a captured graph replays it exactly
    if (n == 0) {
        if (total) hipLaunchKernelGGL(k_scan_word, dim3(1), dim3(1), 0, st, total, (const uint32_t*)nullptr);
        return hipGetLastError();
    }
    if (n > kScanTile && n <= kScanBig) {
        hipLaunchKernelGGL(k_scan_big, dim3(1), dim3(kScanBigBlock), 0, st, in, n, out, total);
        return hipGetLastError();
    }
    uint32_t blocks = (n + kScanTile - 1) / kScanTile;
    uint32_t* sums = scratch;
    uint32_t* offs = scratch + blocks + 1;
    hipLaunchKernelGGL(k_scan_tiles, dim3(blocks), dim3(kScanBlock), 0, st, in, n, out, sums, total);
    if (blocks == 1) return hipGetLastError();
    hipError_t e = scan_exclusive(sums, offs, blocks, nullptr, scratch + 2 * blocks + 2, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan_add, dim3(blocks), dim3(kScanBlock), 0, st, out, n, offs, (const uint32_t*)sums,
                       total);
    return hipGetLastError();
}

// m (<= kScanSetMax) independent exclusive scans of n elements: out[y][i] = sum(in[y][0..i)),
// *total[y] (may be null) = the sum.  scratch: m * scan_scratch_words(n) device words.
inline hipError_t scan_exclusive_set(const ScanSet& S0, int m, uint32_t n, uint32_t* scratch, hipStream_t st) {
    if (m <= 0 || m > kScanSetMax) return hipErrorInvalidValue;
    if (n == 0) {
        for (int y = 0; y < m; ++y)
            if (S0.total[y]) hipLaunchKernelGGL(k_scan_word, dim3(1), dim3(1), 0, st, S0.total[y], (const uint32_t*)nullptr);
        return hipGetLastError();
    }
    if (n > kScanTile && n <= kScanBig) {
        hipLaunchKernelGGL(k_scan_big_set, dim3(m), dim3(kScanBigBlock), 0, st, S0, n);
        return hipGetLastError();
    }
    const uint32_t blocks = (n + kScanTile - 1) / kScanTile;
    const size_t per = scan_scratch_words(n);
    ScanSet S = S0, O{};
    for (int y = 0; y < m; ++y) {
        uint32_t* sc = scratch + (size_t)y * per;
        S.sums[y] = sc;
        O.in[y] = sc;
        O.out[y] = sc + blocks + 1;
        O.total[y] = nullptr;
    }
    hipLaunchKernelGGL(k_scan_tiles_set, dim3(blocks, m), dim3(kScanBlock), 0, st, S, n);
    if (blocks == 1) return hipGetLastError();
    // tile sums: one scan each, through the single-scan path (it has its own scratch)
    for (int y = 0; y < m; ++y) {
        hipError_t e = scan_exclusive(O.in[y], O.out[y], blocks, nullptr, scratch + (size_t)y * per + 2 * blocks + 2, st);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_scan_add_set, dim3(blocks, m), dim3(kScanBlock), 0, st, S, n, O);
    return hipGetLastError();
}

// ---- scans of a device count: n = DN::get(), read by every kernel of the scan, so the
// launches below can be recorded once and replayed on other sizes (up to the capacity n.v).
// A capacity of at most kScanBig is scanned by one workgroup per scan (one launch);
// larger ones reduce each 2048-element tile to its total (tiles_cap words per scan), then
// scan each tile from the sum of the totals before it (two launches; blocks stride over the
// tiles and carry that sum from one of their tiles to the next).
struct DScanSet {
    const uint32_t* in[kScanSetMax];
    uint32_t* out[kScanSetMax];
    uint32_t* total[kScanSetMax];  // may be null
    uint32_t* sums;                // tiles_cap words per scan
    uint32_t tiles_cap;
    uint32_t end_total;            // 1: the total is also written to out[n]
};

// blockIdx.x picks the scan of the set; n <= kScanBig.
__global__ __launch_bounds__(kScanBigBlock) void k_dscan_one(DScanSet S, DN dn_) {
    const uint32_t n = dn_.get();
    const uint32_t y = blockIdx.x;
    const uint32_t t = block_scan_runs(S.in[y], n, S.out[y], S.total[y]);
    if (S.end_total && threadIdx.x == kScanBigBlock - 1) S.out[y][n] = t;
}

// Sum of one value per thread over the block, returned to every thread.  s_red: at least
// kScanBlock / 64 words; the block is synchronised on return.
__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t* s_red) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
#pragma unroll
    for (uint32_t o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) s_red[wave] = v;
    __syncthreads();
    uint32_t r = 0;
#pragma unroll
    for (uint32_t w = 0; w < kScanBlock / 64; ++w) r += s_red[w];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kScanBlock) void k_dscan_reduce(DScanSet S, DN dn_) {
    __shared__ uint32_t s_red[kScanBlock / 64];
    const uint32_t n = dn_.get();
    const uint32_t tiles = (n + kScanTile - 1) / kScanTile;
    const uint32_t y = blockIdx.y;
    const uint32_t* in = S.in[y];
    uint32_t* sums = S.sums + (size_t)y * S.tiles_cap;
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {  // uniform per block
        const size_t base = (size_t)t * kScanTile + (size_t)threadIdx.x * kScanPerThread;
        uint32_t local = 0;
#pragma unroll
        for (uint32_t j = 0; j < kScanPerThread; ++j) {
            const size_t i = base + j;
            local += (i < n) ? in[i] : 0u;
        }
        const uint32_t sum = block_sum(local, s_red);
        if (threadIdx.x == 0) sums[t] = sum;
    }
}

// in and out may alias: each thread reads its elements before writing them.
__global__ __launch_bounds__(kScanBlock) void k_dscan_scan(DScanSet S, DN dn_) {
    __shared__ uint32_t s_wave[kScanBlock / 64];
    __shared__ uint32_t s_red[kScanBlock / 64];
    const uint32_t n = dn_.get();
    const uint32_t tiles = (n + kScanTile - 1) / kScanTile;
    const uint32_t y = blockIdx.y;
    const uint32_t* in = S.in[y];
    uint32_t* out = S.out[y];
    const uint32_t* sums = S.sums + (size_t)y * S.tiles_cap;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    uint32_t prefix = 0, done = 0;  // prefix = sum of sums[0 .. done)
    for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {  // uniform per block
        uint32_t part = 0;
        for (uint32_t k = done + tid; k < t; k += kScanBlock) part += sums[k];
        prefix += block_sum(part, s_red);
        done = t;
        const size_t base = (size_t)t * kScanTile + (size_t)tid * kScanPerThread;
        uint32_t v[kScanPerThread];
        uint32_t local = 0;
#pragma unroll
        for (uint32_t j = 0; j < kScanPerThread; ++j) {
            const size_t i = base + j;
            v[j] = (i < n) ? in[i] : 0u;
            local += v[j];
        }
        const uint32_t incl = wave_incl_scan(local, lane);
        if (lane == 63) s_wave[wave] = incl;
        __syncthreads();
        uint32_t wave_off = 0;
        for (uint32_t w = 0; w < wave; ++w) wave_off += s_wave[w];
        uint32_t run = prefix + wave_off + incl - local;
#pragma unroll
        for (uint32_t j = 0; j < kScanPerThread; ++j) {
            const size_t i = base + j;
            if (i < n) out[i] = run;
            run += v[j];
        }
        if (t == tiles - 1 && tid == kScanBlock - 1) {
            if (S.total[y]) *S.total[y] = run;
            if (S.end_total) out[n] = run;
        }
        __syncthreads();  // s_wave is rewritten by the next tile
    }
    if (tiles == 0 && blockIdx.x == 0 && tid == 0) {
        if (S.total[y]) *S.total[y] = 0u;
        if (S.end_total) out[0] = 0u;
    }
}

// Tiles (scratch words per scan) for a device-count scan of capacity cap.
inline uint32_t dscan_tiles(uint32_t cap) { return std::max<uint32_t>(1u, (cap + kScanTile - 1) / kScanTile); }

// m (<= kScanSetMax) exclusive scans of the device count n (capacity n.v): out[y][i] =
// sum(in[y][0..i)), *total[y] (may be null) = the sum, and out[y][n] too when end_total.
// scratch: m * dscan_tiles(n.v) words.  Kernels only: the sequence can be captured.
inline hipError_t dscan(const DScanSet& S0, int m, DN n, uint32_t* scratch, hipStream_t st) {
    if (m <= 0 || m > kScanSetMax) return hipErrorInvalidValue;
    DScanSet S = S0;
    S.sums = scratch;
    S.tiles_cap = dscan_tiles(n.v);
    if (n.v <= kScanBig) {
        hipLaunchKernelGGL(k_dscan_one, dim3(m), dim3(kScanBigBlock), 0, st, S, n);
        return hipGetLastError();
    }
    const uint32_t grid = std::min<uint32_t>(S.tiles_cap, 1024u);
    hipLaunchKernelGGL(k_dscan_reduce, dim3(grid, m), dim3(kScanBlock), 0, st, S, n);
    hipLaunchKernelGGL(k_dscan_scan, dim3(grid, m), dim3(kScanBlock), 0, st, S, n);
    return hipGetLastError();
}

// One device-count scan (see dscan).
inline hipError_t dscan1(const uint32_t* in, uint32_t* out, DN n, uint32_t* total, bool end_total,
                         uint32_t* scratch, hipStream_t st) {
    DScanSet S{};
    S.in[0] = in;
    S.out[0] = out;
    S.total[0] = total;
    S.end_total = end_total ? 1u : 0u;
    return dscan(S, 1, n, scratch, st);
}

}  // namespace
}  // namespace mnrt
