// prims.hip — the reference's parallel primitives for 32-bit data, MNCudaPrimitives::Scan
// and MNCudaPrimitives::Compact (MNCudaPrimitives.h:167-211, MNCudaPrimitives.cpp:160-190,
// CUDPP calls there), on the library's wave64 scan (scan.h) and a wave-ballot compaction:
// each workgroup takes a tile of 2048 elements, every wave ballots its 64 valid flags and
// writes its valid elements at (tile offset + earlier waves' counts + mask prefix), so the
// output keeps the input order (CUDPP's compaction is stable too).
#include "scan.h"

namespace mnrt {
namespace {

constexpr uint32_t kCompactBlock = 256;
constexpr uint32_t kCompactSteps = 8;
constexpr uint32_t kCompactTile = kCompactBlock * kCompactSteps;

__device__ __forceinline__ size_t compact_index(uint32_t step) {
    return (size_t)blockIdx.x * kCompactTile + (size_t)step * kCompactBlock + threadIdx.x;
}

// Valid elements of each tile.
__global__ __launch_bounds__(kCompactBlock) void k_compact_count(const uint32_t* __restrict__ valid, uint32_t n,
                                                                 uint32_t* __restrict__ tile_count) {
    __shared__ uint32_t s_cnt[kCompactBlock / 64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t c = 0;
#pragma unroll
    for (uint32_t j = 0; j < kCompactSteps; ++j) {
        const size_t i = compact_index(j);
        c += (uint32_t)__popcll(__ballot(i < n && valid[i] != 0u));
    }
    if (lane == 0) s_cnt[wave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < kCompactBlock / 64; ++w) t += s_cnt[w];
        tile_count[blockIdx.x] = t;
    }
}

// Valid elements to out[tile offset + rank], in input order (step, then wave, then lane).
__global__ __launch_bounds__(kCompactBlock) void k_compact_scatter(const uint32_t* __restrict__ in,
                                                                   const uint32_t* __restrict__ valid, uint32_t n,
                                                                   const uint32_t* __restrict__ tile_off,
                                                                   uint32_t* __restrict__ out) {
    __shared__ uint32_t s_cnt[kCompactBlock / 64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t base = tile_off[blockIdx.x];
    for (uint32_t j = 0; j < kCompactSteps; ++j) {
        const size_t i = compact_index(j);
        const bool v = i < n && valid[i] != 0u;
        const uint64_t m = __ballot(v);
        if (lane == 0) s_cnt[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = 0, all = 0;
        for (uint32_t w = 0; w < kCompactBlock / 64; ++w) {
            before += w < wave ? s_cnt[w] : 0u;
            all += s_cnt[w];
        }
        if (v) out[base + before + mask_prefix(m)] = in[i];
        base += all;
        __syncthreads();  // s_cnt is rewritten by the next step
    }
}

// out[i] = excl[i] + in[i] (in and out may alias: each element is read and written by one lane).
__global__ __launch_bounds__(256) void k_to_inclusive(const uint32_t* in, const uint32_t* __restrict__ excl,
                                                      uint32_t n, uint32_t* out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = excl[i] + in[i];
}

}  // namespace
}  // namespace mnrt

extern "C" int mnrt_scan_u32(const uint32_t* d_in, uint32_t n, int inclusive, uint32_t* d_out, uint32_t* d_total,
                             void* stream) {
    using namespace mnrt;
    if (n > 0 && (!d_in || !d_out)) return MNRT_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    const size_t extra = inclusive ? (size_t)n : 0;
    void* buf = nullptr;
    MNRT_HIP_CHECK(pool_alloc(&buf, (scan_scratch_words(n) + extra + 4) * 4, st));
    uint32_t* scr = (uint32_t*)buf;
    uint32_t* excl = scr + scan_scratch_words(n) + 4;
    int rc = MNRT_OK;
    if (scan_exclusive(d_in, inclusive ? excl : d_out, n, d_total, scr, st) != hipSuccess) rc = MNRT_ERR_HIP;
    if (rc == MNRT_OK && inclusive && n > 0) {
        hipLaunchKernelGGL(k_to_inclusive, dim3((n + 255) / 256), dim3(256), 0, st, d_in, (const uint32_t*)excl, n,
                           d_out);
        if (hipGetLastError() != hipSuccess) rc = MNRT_ERR_HIP;
    }
    if (pool_free(buf, st) != hipSuccess && rc == MNRT_OK) rc = MNRT_ERR_HIP;
    return rc;
}

extern "C" int mnrt_scan_u32_counted(const uint32_t* d_in, const uint32_t* d_count, uint32_t capacity,
                                     uint32_t* d_out, uint32_t* d_total, int end_total, void* stream) {
    using namespace mnrt;
    if (!d_count || (!d_out && (end_total || capacity > 0)) || (capacity > 0 && !d_in)) return MNRT_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    void* buf = nullptr;
    MNRT_HIP_CHECK(pool_alloc(&buf, (size_t)dscan_tiles(capacity) * 4, st));
    DScanSet S{};
    S.in[0] = d_in;
    S.out[0] = d_out;
    S.total[0] = d_total;
    S.end_total = end_total ? 1u : 0u;
    int rc = dscan(S, 1, dn(d_count, capacity), (uint32_t*)buf, st) == hipSuccess ? MNRT_OK : MNRT_ERR_HIP;
    if (pool_free(buf, st) != hipSuccess && rc == MNRT_OK) rc = MNRT_ERR_HIP;
    return rc;
}

extern "C" int mnrt_compact_u32(const uint32_t* d_in, const uint32_t* d_is_valid, uint32_t n, uint32_t* d_out,
                                uint32_t* d_out_count, void* stream) {
    using namespace mnrt;
    if (!d_out_count) return MNRT_ERR_ARG;
    if (n > 0 && (!d_in || !d_is_valid || !d_out || d_in == d_out)) return MNRT_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    if (n == 0) {
        hipLaunchKernelGGL(k_scan_word, dim3(1), dim3(1), 0, st, d_out_count, (const uint32_t*)nullptr);
        MNRT_HIP_CHECK(hipGetLastError());
        return MNRT_OK;
    }
    const uint32_t tiles = (uint32_t)(((size_t)n + kCompactTile - 1) / kCompactTile);
    void* buf = nullptr;
    MNRT_HIP_CHECK(pool_alloc(&buf, (2ull * tiles + scan_scratch_words(tiles) + 4) * 4, st));
    uint32_t* cnt = (uint32_t*)buf;
    uint32_t* off = cnt + tiles;
    uint32_t* scr = off + tiles;
    hipLaunchKernelGGL(k_compact_count, dim3(tiles), dim3(kCompactBlock), 0, st, d_is_valid, n, cnt);
    int rc = MNRT_OK;
    if (scan_exclusive(cnt, off, tiles, d_out_count, scr, st) != hipSuccess) rc = MNRT_ERR_HIP;
    if (rc == MNRT_OK) {
        hipLaunchKernelGGL(k_compact_scatter, dim3(tiles), dim3(kCompactBlock), 0, st, d_in, d_is_valid, n,
                           (const uint32_t*)off, d_out);
        if (hipGetLastError() != hipSuccess) rc = MNRT_ERR_HIP;
    }
    if (pool_free(buf, st) != hipSuccess && rc == MNRT_OK) rc = MNRT_ERR_HIP;
    return rc;
}
