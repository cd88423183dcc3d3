// reweight.hip — photon powers and hit colours as functions of per-light emission and
// per-triangle albedo, with the photon paths and hit triangles held fixed, and the adjoint
// of that map as deterministic fixed-point sums.
//
// The reference renders and never fits parameters; what it does to a photon's power along
// its path is what the forward restates: the power leaves the light with the light's
// emission (photon_build.cu:79-130, RTCore.cpp:454-520) and is multiplied by the diffuse
// albedo of every triangle it scatters from (photon_build.cu:278-331, RTCore.cpp:556-581).
// With the path fixed the stored power is therefore
//     power[p] = base[p] * emission[light[p]] * prod over the path's triangles of albedo,
// base[p] holding everything else (the light's share of the paths, Russian-roulette
// doublings, 1 / paths shot).  A hit's diffuse colour (KernelPMPhotonGather's
// d_clrDiffuse, photon_knn.cu:833-844) is the albedo of the triangle it hit.
//
// Adjoint: dLoss/demission[l] = sum over photons of light l of g_p * base_p * prod_p and
// dLoss/dalbedo[t] = sum over (photon, path position) pairs on t of
// g_p * base_p * emission * (product of the path's other albedos) + sum over hits on t of
// the colour gradient.  The sums are 64-bit fixed-point integers with one power-of-two unit
// per output array (from the largest term and the number of terms, so no sum overflows),
// added by atomics after a wave-level reduction of lanes that share an output; integer
// addition is associative, so the result is the same for every schedule.
#include <algorithm>
#include <cmath>

#include "mnrt_common.h"

namespace mnrt {
namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kPathTris = 4;  // triangles recorded per photon path (mnrt_photons_generate)
constexpr int kShareRounds = 4;

struct RwArgs {
    uint32_t np;
    const float4* base;
    const uint32_t* light;
    const uint32_t* tris;  // kPathTris per photon
    uint32_t num_lights;
    const float* emission;  // 3 per light
    uint32_t num_tris;
    const float* albedo;  // 3 per triangle
    uint32_t nq;
    const uint32_t* hit_tri;
    const float4* g_pow;
    const float4* g_clr;
};

__device__ __forceinline__ float3 ld3(const float* a, uint32_t i) {
    return make_float3(a[3ull * i], a[3ull * i + 1], a[3ull * i + 2]);
}

// A photon's path with its indices validated: light < num_lights, triangles < num_tris.
struct Path {
    bool ok;
    uint32_t l;
    uint32_t t[kPathTris];
};

__device__ __forceinline__ Path load_path(const RwArgs& A, uint32_t p) {
    Path P;
    P.l = A.light[p];
    P.ok = P.l < A.num_lights;
#pragma unroll
    for (int k = 0; k < kPathTris; ++k) {
        P.t[k] = A.tris[(size_t)kPathTris * p + k];
        P.ok = P.ok && (P.t[k] == kNone || P.t[k] < A.num_tris);
    }
    return P;
}

__global__ __launch_bounds__(256) void k_rw_forward(RwArgs A, float4* __restrict__ out_pow,
                                                    float4* __restrict__ io_clr) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < A.np) {
        const Path P = load_path(A, i);
        const float4 b = A.base[i];
        float r = __uint_as_float(0x7FC00000u), g = r, bl = r;  // NaN: an index out of range
        if (P.ok) {
            const float3 e = ld3(A.emission, P.l);
            r = b.x * e.x;
            g = b.y * e.y;
            bl = b.z * e.z;
#pragma unroll
            for (int k = 0; k < kPathTris; ++k) {
                if (P.t[k] == kNone) continue;
                const float3 a = ld3(A.albedo, P.t[k]);
                r = r * a.x;
                g = g * a.y;
                bl = bl * a.z;
            }
        }
        out_pow[i] = make_float4(r, g, bl, 0.0f);
    }
    if (i < A.nq && io_clr) {
        const uint32_t t = A.hit_tri[i];
        if (t != kNone) {
            float4 c = io_clr[i];
            if (t < A.num_tris) {
                const float3 a = ld3(A.albedo, t);
                c.x = a.x;
                c.y = a.y;
                c.z = a.z;
            } else {
                c.x = c.y = c.z = __uint_as_float(0x7FC00000u);
            }
            io_clr[i] = c;
        }
    }
}

// The adjoint's terms of one photon: the emission term (one output, light l) and one
// albedo term per recorded triangle, per channel.  Invalid paths contribute nothing.
struct PhotonTerms {
    float e[3];
    float a[kPathTris][3];
};

__device__ __forceinline__ PhotonTerms photon_terms(const RwArgs& A, uint32_t p, const Path& P) {
    PhotonTerms T;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        T.e[c] = 0.0f;
#pragma unroll
        for (int k = 0; k < kPathTris; ++k) T.a[k][c] = 0.0f;
    }
    if (!P.ok) return T;
    const float4 b4 = A.base[p], g4 = A.g_pow[p];
    const float b[3] = {b4.x, b4.y, b4.z}, g[3] = {g4.x, g4.y, g4.z};
    const float3 e3 = ld3(A.emission, P.l);
    const float e[3] = {e3.x, e3.y, e3.z};
    float alb[kPathTris][3];
#pragma unroll
    for (int k = 0; k < kPathTris; ++k) {
        const float3 a = P.t[k] != kNone ? ld3(A.albedo, P.t[k]) : make_float3(1.0f, 1.0f, 1.0f);
        alb[k][0] = a.x;
        alb[k][1] = a.y;
        alb[k][2] = a.z;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float prod = b[c];
#pragma unroll
        for (int k = 0; k < kPathTris; ++k) prod = prod * alb[k][c];
        T.e[c] = g[c] * prod;
#pragma unroll
        for (int k = 0; k < kPathTris; ++k) {
            if (P.t[k] == kNone) continue;
            float o = b[c] * e[c];
#pragma unroll
            for (int j = 0; j < kPathTris; ++j)
                if (j != k) o = o * alb[j][c];
            T.a[k][c] = g[c] * o;
        }
    }
    return T;
}

// Largest |v| of three, +infinity if one of them is not finite (fmaxf alone would drop a NaN).
__device__ __forceinline__ float abs_max3(const float* v) {
    float m = 0.0f;
    bool finite = true;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        finite = finite && fabsf(v[c]) <= kInf;
        m = fmaxf(m, fabsf(v[c]));
    }
    return finite ? m : __uint_as_float(0x7F800000u);
}

// Bits of a bound m >= 0 (+infinity marks a non-finite term; bit order = value order).
__device__ __forceinline__ uint32_t bound_bits(float m) {
    return (m <= kInf) ? __float_as_uint(m) : 0x7F800000u;
}

// Pass 1: the largest |term| of each output array (emission, albedo), one atomic per wave.
__global__ __launch_bounds__(256) void k_rw_bound(RwArgs A, uint32_t* __restrict__ bound) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float me = 0.0f, ma = 0.0f;
    if (i < A.np) {
        const Path P = load_path(A, i);
        const PhotonTerms T = photon_terms(A, i, P);
        me = abs_max3(T.e);
#pragma unroll
        for (int k = 0; k < kPathTris; ++k) ma = fmaxf(ma, abs_max3(T.a[k]));
    }
    if (i < A.nq && A.g_clr) {
        const uint32_t t = A.hit_tri[i];
        if (t != kNone && t < A.num_tris) {
            const float4 gc = A.g_clr[i];
            const float v[3] = {gc.x, gc.y, gc.z};
            ma = fmaxf(ma, abs_max3(v));
        }
    }
    const uint32_t be = wave_max_u32(bound_bits(me)), ba = wave_max_u32(bound_bits(ma));
    if ((threadIdx.x & 63u) == 0) {
        if (be) atomicMax(&bound[0], be);
        if (ba) atomicMax(&bound[1], ba);
    }
}

// Fixed-point unit of an output array: the largest power of two with
// terms * bound * unit < 2^62, so any sum of at most `terms` terms fits; 0 when every
// term is 0; negative when a term is not finite.
__device__ __forceinline__ double fixed_unit(uint32_t bound_word, double terms) {
    const float m = __uint_as_float(bound_word);
    if (!(m <= kInf)) return -1.0;
    if (!(m > 0.0f) || !(terms > 0.0)) return 0.0;
    int ex = 0;
    (void)frexp(4.611686018427387904e18 / (terms * (double)m), &ex);
    return ldexp(1.0, ex - 1);
}

__device__ __forceinline__ long long to_fixed(float v, double unit) {
    return unit > 0.0 ? __double2ll_rn((double)v * unit) : 0ll;
}

// Adds (x, y, z) into acc[3 key ..] for every lane with key != kNone.  Lanes that share a
// key are summed by a wave reduction first (up to kShareRounds distinct keys, taken in
// lane order), the rest add on their own.  Called by every lane of the wave.
__device__ __forceinline__ void add_shared(uint32_t key, long long x, long long y, long long z,
                                           unsigned long long* __restrict__ acc, uint32_t lane) {
    bool pending = key != kNone && (x != 0 || y != 0 || z != 0);
    for (int round = 0; round < kShareRounds; ++round) {
        const uint64_t act = __ballot(pending);
        if (act == 0ull) return;
        const uint32_t leader = (uint32_t)__ffsll((long long)act) - 1u;
        const uint32_t k0 = (uint32_t)__builtin_amdgcn_readlane((int)key, (int)leader);
        const bool mine = pending && key == k0;
        if (__popcll(__ballot(mine)) == 1) break;  // the first pending key is not shared
        const long long sx = wave_sum_i64(mine ? x : 0ll), sy = wave_sum_i64(mine ? y : 0ll),
                        sz = wave_sum_i64(mine ? z : 0ll);
        if (lane == leader) {
            if (sx) atomicAdd(&acc[3ull * k0 + 0], (unsigned long long)sx);
            if (sy) atomicAdd(&acc[3ull * k0 + 1], (unsigned long long)sy);
            if (sz) atomicAdd(&acc[3ull * k0 + 2], (unsigned long long)sz);
        }
        pending = pending && !mine;
    }
    if (pending) {
        if (x) atomicAdd(&acc[3ull * key + 0], (unsigned long long)x);
        if (y) atomicAdd(&acc[3ull * key + 1], (unsigned long long)y);
        if (z) atomicAdd(&acc[3ull * key + 2], (unsigned long long)z);
    }
}

// Pass 2: every term in its array's unit, summed per output.  No lane leaves early (the
// wave reductions need them all).
__global__ __launch_bounds__(256) void k_rw_accumulate(RwArgs A, const uint32_t* __restrict__ bound,
                                                       double terms_e, double terms_a,
                                                       unsigned long long* __restrict__ acc_e,
                                                       unsigned long long* __restrict__ acc_a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    const double ue = fixed_unit(bound[0], terms_e), ua = fixed_unit(bound[1], terms_a);
    if (ue < 0.0 || ua < 0.0) return;  // non-finite gradients: the outputs become NaN (uniform)
    Path P{};
    PhotonTerms T{};
    const bool ph = i < A.np;
    if (ph) {
        P = load_path(A, i);
        T = photon_terms(A, i, P);
    }
    const bool pe = ph && P.ok;
    add_shared(pe ? P.l : kNone, to_fixed(T.e[0], ue), to_fixed(T.e[1], ue), to_fixed(T.e[2], ue), acc_e, lane);
#pragma unroll
    for (int k = 0; k < kPathTris; ++k)
        add_shared(pe ? P.t[k] : kNone, to_fixed(T.a[k][0], ua), to_fixed(T.a[k][1], ua), to_fixed(T.a[k][2], ua),
                   acc_a, lane);
    uint32_t t = kNone;
    float4 gc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (i < A.nq && A.g_clr) {
        t = A.hit_tri[i];
        if (t != kNone && t < A.num_tris) gc = A.g_clr[i];
        else t = kNone;
    }
    add_shared(t, to_fixed(gc.x, ua), to_fixed(gc.y, ua), to_fixed(gc.z, ua), acc_a, lane);
}

// Pass 3: the sums as floats (one rounding each).
__global__ __launch_bounds__(256) void k_rw_output(uint32_t n_e, uint32_t n_a, const uint32_t* __restrict__ bound,
                                                   double terms_e, double terms_a,
                                                   const unsigned long long* __restrict__ acc_e,
                                                   const unsigned long long* __restrict__ acc_a,
                                                   float* __restrict__ out_e, float* __restrict__ out_a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const double ue = fixed_unit(bound[0], terms_e), ua = fixed_unit(bound[1], terms_a);
    const bool bad = ue < 0.0 || ua < 0.0;
    const float nan = __uint_as_float(0x7FC00000u);
    if (i < 3u * n_e) out_e[i] = bad ? nan : (ue > 0.0 ? (float)((double)(long long)acc_e[i] / ue) : 0.0f);
    if (i < 3u * n_a) out_a[i] = bad ? nan : (ua > 0.0 ? (float)((double)(long long)acc_a[i] / ua) : 0.0f);
}

__global__ __launch_bounds__(256) void k_rw_zero(unsigned long long* __restrict__ p, size_t n, uint32_t* bound) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x < 2) bound[threadIdx.x] = 0u;
}

int rw_check(uint32_t np, const float* base, const uint32_t* light, const uint32_t* tris, uint32_t num_lights,
             const float* emission, uint32_t num_tris, const float* albedo, uint32_t nq, const uint32_t* hit_tri) {
    if (np > 0 && (!base || !light || !tris || !emission || num_lights == 0)) return MNRT_ERR_ARG;
    if ((np > 0 || nq > 0) && num_tris > 0 && !albedo) return MNRT_ERR_ARG;
    if (nq > 0 && !hit_tri) return MNRT_ERR_ARG;
    if ((uint64_t)num_lights * 3 > 0xFFFFFFFFull || (uint64_t)num_tris * 3 > 0xFFFFFFFFull) return MNRT_ERR_CAPACITY;
    return MNRT_OK;
}

}  // namespace
}  // namespace mnrt

extern "C" int mnrt_photons_reweight(uint32_t np, const float* d_base_pow4, const uint32_t* d_light,
                                     const uint32_t* d_path_tris, uint32_t num_lights, const float* d_emission3,
                                     uint32_t num_tris, const float* d_albedo3, uint32_t nq,
                                     const uint32_t* d_hit_tri, float* d_out_pow4, float* d_io_clr4, void* stream) {
    using namespace mnrt;
    int rc = rw_check(np, d_base_pow4, d_light, d_path_tris, num_lights, d_emission3, num_tris, d_albedo3, nq,
                      d_hit_tri);
    if (rc != MNRT_OK) return rc;
    if ((np > 0 && !d_out_pow4) || (nq > 0 && !d_io_clr4)) return MNRT_ERR_ARG;
    const uint32_t n = std::max(np, nq);
    if (n == 0) return MNRT_OK;
    RwArgs A{np, (const float4*)d_base_pow4, d_light, d_path_tris, num_lights, d_emission3, num_tris, d_albedo3,
             nq, d_hit_tri, nullptr, nullptr};
    hipLaunchKernelGGL(k_rw_forward, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, A,
                       (float4*)d_out_pow4, nq > 0 ? (float4*)d_io_clr4 : (float4*)nullptr);
    MNRT_HIP_CHECK(hipGetLastError());
    return MNRT_OK;
}

extern "C" int mnrt_photons_reweight_bwd(uint32_t np, const float* d_base_pow4, const uint32_t* d_light,
                                         const uint32_t* d_path_tris, uint32_t num_lights, const float* d_emission3,
                                         uint32_t num_tris, const float* d_albedo3, uint32_t nq,
                                         const uint32_t* d_hit_tri, const float* d_grad_pow4,
                                         const float* d_grad_clr4, float* d_grad_emission3, float* d_grad_albedo3,
                                         void* stream) {
    using namespace mnrt;
    int rc = rw_check(np, d_base_pow4, d_light, d_path_tris, num_lights, d_emission3, num_tris, d_albedo3, nq,
                      d_hit_tri);
    if (rc != MNRT_OK) return rc;
    if ((np > 0 && !d_grad_pow4) || (nq > 0 && !d_grad_clr4)) return MNRT_ERR_ARG;
    if ((num_lights > 0 && !d_grad_emission3) || (num_tris > 0 && !d_grad_albedo3)) return MNRT_ERR_ARG;
    if (num_lights == 0 && num_tris == 0) return MNRT_OK;
    hipStream_t st = (hipStream_t)stream;
    const size_t words = 3ull * num_lights + 3ull * num_tris;
    void* buf = nullptr;
    MNRT_HIP_CHECK(pool_alloc(&buf, words * 8 + 64, st));
    unsigned long long* acc_e = (unsigned long long*)buf;
    unsigned long long* acc_a = acc_e + 3ull * num_lights;
    uint32_t* bound = (uint32_t*)(acc_a + 3ull * num_tris);
    const uint32_t zb = (uint32_t)std::max<size_t>(1, std::min<size_t>((words + 255) / 256, 4096));
    hipLaunchKernelGGL(k_rw_zero, dim3(zb), dim3(256), 0, st, acc_e, words, bound);
    RwArgs A{np, (const float4*)d_base_pow4, d_light, d_path_tris, num_lights, d_emission3, num_tris, d_albedo3,
             nq, d_hit_tri, (const float4*)d_grad_pow4, nq > 0 ? (const float4*)d_grad_clr4 : nullptr};
    // terms per output: at most one per photon (emission); kPathTris per photon plus one
    // per hit (albedo)
    const double terms_e = (double)np, terms_a = (double)kPathTris * np + (double)nq;
    const uint32_t n = std::max(np, nq);
    if (n > 0) {
        hipLaunchKernelGGL(k_rw_bound, dim3((n + 255) / 256), dim3(256), 0, st, A, bound);
        hipLaunchKernelGGL(k_rw_accumulate, dim3((n + 255) / 256), dim3(256), 0, st, A, (const uint32_t*)bound,
                           terms_e, terms_a, acc_e, acc_a);
    }
    const size_t outs = 3ull * std::max(num_lights, num_tris);
    hipLaunchKernelGGL(k_rw_output, dim3((uint32_t)((outs + 255) / 256)), dim3(256), 0, st, num_lights, num_tris,
                       (const uint32_t*)bound, terms_e, terms_a, (const unsigned long long*)acc_e,
                       (const unsigned long long*)acc_a, d_grad_emission3, d_grad_albedo3);
    rc = hipGetLastError() == hipSuccess ? MNRT_OK : MNRT_ERR_HIP;
    if (pool_free(buf, st) != hipSuccess && rc == MNRT_OK) rc = MNRT_ERR_HIP;
    return rc;
}
