// photon.hip — photon tracing on gfx950 (the step that fills the photon map the kNN
// estimate reads).
//
// Restates RTCore::RebuildPhotonMaps (RTCore.cpp:454-638) with kernel_SpawnLightPhotons
// (photon_build.cu:79), kernel_TracePhotons (raytracing.cu:1146), kernel_RussianRoulette
// (photon_build.cu:220) and kernel_SpawnScatteredPhotons (photon_build.cu:278), one lane
// per photon path instead of one launch per bounce with compaction in between:
//   * a path starts on a quad light (light = path index mod light count, power
//     scaled by the light count), direction cosine-distributed about the light normal;
//   * every diffuse hit stores a photon (position, power, incoming direction) before
//     scattering, the first hit included (RTCore.cpp:556-581);
//   * at bounce index > 3 Russian roulette keeps the path with probability 0.5 and
//     doubles its power (RTCore.cpp:545-553, photon_build.cu:220-241);
//   * a scattered path continues from the hit with a cosine-distributed direction about
//     the geometric normal facing the incoming ray, power multiplied by the albedo.
// Random numbers come from a counter-based integer hash of (seed, path, bounce, draw),
// and cosine directions use rejection sampling of the unit disk (Malley's method) with
// the Duff et al. 2017 orthonormal basis, so the sampling needs only +, *, / and sqrt
// and an implementation on any IEEE fp32 machine draws identical paths.  (The reference
// draws Mersenne-Twister uniforms and maps them with sin/cos, sample_dev.h:100-279.)
#include <algorithm>
#include <vector>

#include "sample_dev.h"
#include "scan.h"

namespace mnrt {

struct PhotonArgs {
    const uint2* nodes;
    const float4* recs;
    RootBox root;
    const float4* v0;
    const float4* v1;
    const float4* v2;
    const float4* albedo;
    const float4* lights;  // 4 float4 per light: corner, edge u, edge v, power
    uint32_t num_lights;
    uint32_t path_begin;
    uint32_t num_paths;
    uint32_t max_bounces;
    uint32_t seed;
    float eps;
    float4* pos;
    float4* pw;
    float4* dir;
    uint32_t* valid;
    uint32_t* meta;       // light | bounce << 8
    uint32_t* path_tris;  // 4 per slot: triangles scattered from before this photon
};

__global__ __launch_bounds__(256) void k_photon_paths(PhotonArgs A) {
    extern __shared__ uint32_t s_stack[];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.num_paths) return;
    const uint32_t path = A.path_begin + i;
    const uint32_t slots = A.max_bounces + 1;
    const size_t base = (size_t)i * slots;
    for (uint32_t b = 0; b < slots; ++b) A.valid[base + b] = 0u;

    const uint32_t light = path % A.num_lights;
    float4 c = A.lights[4 * light + 0], eu = A.lights[4 * light + 1], ev = A.lights[4 * light + 2];
    float4 lp = A.lights[4 * light + 3];
    float u = rnd01(A.seed, path, 0xFFFFu, 0), v = rnd01(A.seed, path, 0xFFFFu, 1);
    f3 o = mk3(c.x + u * eu.x + v * ev.x, c.y + u * eu.y + v * ev.y, c.z + u * eu.z + v * ev.z);
    f3 ln = unit_normal(xyz(eu), xyz(ev));
    f3 d = cosine_dir(ln, A.seed, path, 0xFFFEu);
    float nl = (float)A.num_lights;
    float pr = lp.x * nl, pg = lp.y * nl, pb = lp.z * nl;
    uint32_t tris[4] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};

    for (uint32_t b = 0; b < slots; ++b) {
        Hit h = trace_ray<kTraceStack, kTraceLds>(A.nodes, A.recs, A.root, o, d, A.eps, s_stack + threadIdx.x,
                                                  blockDim.x);
        if (h.tri < 0) break;
        float4 alb = A.albedo[h.tri];
        if (!(alb.x > 0.0f || alb.y > 0.0f || alb.z > 0.0f)) break;
        f3 x = mk3(o.x + d.x * h.t, o.y + d.y * h.t, o.z + d.z * h.t);
        if (b > 3) {
            if (!(rnd01(A.seed, path, b, 100) <= 0.5f)) break;
            pr = pr * 2.0f;
            pg = pg * 2.0f;
            pb = pb * 2.0f;
        }
        size_t s = base + b;
        A.valid[s] = 1u;
        A.pos[s] = make_float4(x.x, x.y, x.z, 0.0f);
        A.pw[s] = make_float4(pr, pg, pb, 0.0f);
        A.dir[s] = make_float4(d.x, d.y, d.z, 0.0f);
        if (A.meta) A.meta[s] = light | (b << 8);
        if (A.path_tris)
            for (int k = 0; k < 4; ++k) A.path_tris[4 * s + k] = tris[k];
        if (b + 1 >= slots) break;
        float4 a0 = A.v0[h.tri], a1 = A.v1[h.tri], a2 = A.v2[h.tri];
        f3 n = unit_normal(mk3(a1.x - a0.x, a1.y - a0.y, a1.z - a0.z), mk3(a2.x - a0.x, a2.y - a0.y, a2.z - a0.z));
        if (dot3(n, d) > 0.0f) n = mk3(n.x * -1.0f, n.y * -1.0f, n.z * -1.0f);
        d = cosine_dir(n, A.seed, path, b);
        o = x;
        pr = pr * alb.x;
        pg = pg * alb.y;
        pb = pb * alb.z;
        if (b < 4) tris[b] = (uint32_t)h.tri;
    }
}

// Stable compaction of the valid slots into the output, up to `limit` photons.
__global__ __launch_bounds__(256) void k_photon_compact(uint32_t nslots, const uint32_t* __restrict__ valid,
                                                        const uint32_t* __restrict__ pos, uint32_t out_base,
                                                        uint32_t limit, PhotonArgs A, float4* __restrict__ opos,
                                                        float4* __restrict__ opw, float4* __restrict__ odir,
                                                        uint32_t* __restrict__ ometa, uint32_t* __restrict__ opath,
                                                        uint32_t* __restrict__ otris, uint32_t path_begin,
                                                        uint32_t slots_per_path) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots || !valid[s]) return;
    uint32_t d = out_base + pos[s];
    if (d >= limit) return;
    opos[d] = A.pos[s];
    opw[d] = A.pw[s];
    odir[d] = A.dir[s];
    if (ometa) ometa[d] = A.meta[s];
    if (opath) opath[d] = path_begin + s / slots_per_path;
    if (otris)
        for (int k = 0; k < 4; ++k) otris[4ull * d + k] = A.path_tris[4ull * s + k];
}

__global__ __launch_bounds__(256) void k_scale_power(uint32_t n, float4* __restrict__ pw, float scale) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = pw[i];
    pw[i] = make_float4(p.x * scale, p.y * scale, p.z * scale, p.w);
}

}  // namespace mnrt

// RTCore::RebuildPhotonMaps (RTCore.cpp:454-638) for one (global) map: paths are traced
// in batches until `target` photons are stored (or 100 * target paths were shot, the
// reference's bound); the first `target` photons in (path, bounce) order are kept and
// their power scaled by 1 / (paths shot up to the last kept photon) (ScaleFlux,
// RTCore.cpp:626-629).
extern "C" int mnrt_photons_generate(const mnrt_tree* tree, const float* d_v0, const float* d_v1,
                                     const float* d_v2, const float* d_albedo4, const float* d_lights,
                                     uint32_t num_lights, uint32_t target, uint32_t max_bounces, uint32_t seed,
                                     float eps, float* d_out_pos4, float* d_out_pow4, float* d_out_dir4,
                                     uint32_t* d_out_meta, uint32_t* d_out_path, uint32_t* d_out_path_tris,
                                     uint32_t* out_count, uint32_t* out_paths, void* stream) {
    using namespace mnrt;
    if (!tree || !d_v0 || !d_v1 || !d_v2 || !d_albedo4 || !d_lights || num_lights == 0 || !d_out_pos4 ||
        !d_out_pow4 || !d_out_dir4 || !out_count || !out_paths || max_bounces > 8)
        return MNRT_ERR_ARG;
    // the record holds 4 triangles: a photon stored at bounce b scattered from b of them
    if (d_out_path_tris && max_bounces > 4) return MNRT_ERR_ARG;
    if (tree->kind != MNRT_KIND_TRI || !tree->d_nodes || !tree->d_recs) return MNRT_ERR_STATE;
    hipStream_t st = (hipStream_t)stream;
    *out_count = 0;
    *out_paths = 0;
    if (target == 0) return MNRT_OK;
    const uint32_t slots = max_bounces + 1;
    uint32_t batch = std::max<uint32_t>(target, 4096);
    const uint64_t max_paths = 100ull * target;
    PhotonArgs A{};
    A.nodes = (const uint2*)tree->d_nodes;
    A.recs = (const float4*)tree->d_recs;
    for (int a = 0; a < 3; ++a) {
        A.root.lo[a] = tree->root_min[a];
        A.root.hi[a] = tree->root_max[a];
    }
    A.v0 = (const float4*)d_v0;
    A.v1 = (const float4*)d_v1;
    A.v2 = (const float4*)d_v2;
    A.albedo = (const float4*)d_albedo4;
    A.lights = (const float4*)d_lights;
    A.num_lights = num_lights;
    A.max_bounces = max_bounces;
    A.seed = seed;
    A.eps = eps;
    uint32_t stored = 0, shot = 0, last_path = 0;
    uint32_t* h_tot = (uint32_t*)mnrt::host_staging(1, 16);
    if (!h_tot) return MNRT_ERR_NOMEM;
    int rc = MNRT_OK;
    void* buf = nullptr;
    size_t cap_slots = 0;
    while (stored < target && shot < max_paths) {
        uint32_t np = (uint32_t)std::min<uint64_t>(batch, max_paths - shot);
        size_t nslots = (size_t)np * slots;
        if (nslots > 0x7FFFFFFFull) {
            rc = MNRT_ERR_CAPACITY;
            break;
        }
        if (nslots > cap_slots) {
            if (buf) (void)mnrt::pool_free(buf, st);
            size_t words = nslots * (4 * 3 + 1 + 1 + 4 + 1) + scan_scratch_words((uint32_t)nslots) + 8;
            if (mnrt::pool_alloc(&buf, words * 4, st) != hipSuccess) {
                buf = nullptr;
                rc = MNRT_ERR_NOMEM;
                break;
            }
            cap_slots = nslots;
        }
        float4* f = (float4*)buf;
        A.pos = f;
        A.pw = f + cap_slots;
        A.dir = f + 2 * cap_slots;
        uint32_t* u = (uint32_t*)(f + 3 * cap_slots);
        A.valid = u;
        A.meta = u + cap_slots;
        A.path_tris = u + 2 * cap_slots;
        uint32_t* pos = u + 6 * cap_slots;
        uint32_t* tot = pos + cap_slots;
        uint32_t* scr = tot + 4;
        A.path_begin = shot;
        A.num_paths = np;
        hipLaunchKernelGGL(k_photon_paths, dim3((np + 255) / 256), dim3(256), trace_lds_bytes(kTraceLds, 256), st, A);
        if (scan_exclusive(A.valid, pos, (uint32_t)nslots, tot, scr, st) != hipSuccess) {
            rc = MNRT_ERR_HIP;
            break;
        }
        hipLaunchKernelGGL(k_photon_compact, dim3((uint32_t)((nslots + 255) / 256)), dim3(256), 0, st,
                           (uint32_t)nslots, A.valid, pos, stored, target, A, (float4*)d_out_pos4,
                           (float4*)d_out_pow4, (float4*)d_out_dir4, d_out_meta, d_out_path, d_out_path_tris, shot,
                           slots);
        if (hipMemcpyAsync(h_tot, tot, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            rc = MNRT_ERR_HIP;
            break;
        }
        uint32_t got = *h_tot;
        if (stored + got >= target) {
            // paths shot up to the path holding the last kept photon
            uint32_t need = target - stored;
            std::vector<uint32_t> v(nslots);
            if (hipMemcpy(v.data(), A.valid, nslots * 4, hipMemcpyDeviceToHost) != hipSuccess) {
                rc = MNRT_ERR_HIP;
                break;
            }
            uint32_t seen = 0;
            for (size_t s = 0; s < nslots; ++s) {
                if (v[s] && ++seen == need) {
                    last_path = shot + (uint32_t)(s / slots);
                    break;
                }
            }
            stored = target;
            shot = last_path + 1;
            break;
        }
        stored += got;
        shot += np;
        batch = std::max<uint32_t>(4096, (uint32_t)std::min<uint64_t>(2ull * batch, 1ull << 22));
    }
    if (buf) (void)mnrt::pool_free(buf, st);
    if (rc != MNRT_OK) return rc;
    if (stored > 0 && shot > 0) {
        float scale = 1.0f / (float)shot;
        hipLaunchKernelGGL(k_scale_power, dim3((stored + 255) / 256), dim3(256), 0, st, stored, (float4*)d_out_pow4,
                           scale);
        MNRT_HIP_CHECK(hipGetLastError());
    }
    *out_count = stored;
    *out_paths = shot;
    return MNRT_OK;
}

// ---- PhotonData plumbing of the reference layout (KernelDefs.h:537-630) ---------------

namespace mnrt {

// dev_Spherical2Direction (mncudautil_dev.h:345-353): phi in position.w, theta in power.w.
__global__ __launch_bounds__(256) void k_sph_to_dir(uint32_t n, const float4* __restrict__ pos,
                                                    const float4* __restrict__ pw, float4* __restrict__ dir) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float phi = pos[i].w, theta = pw[i].w;
    float st = sinf(theta);
    dir[i] = make_float4(st * cosf(phi), st * sinf(phi), cosf(theta), 0.0f);
}

// dev_Direction2Spherical (mncudautil_dev.h:367-375): phi = atan2(y, x) in [0, 2 pi), theta = acos(z).
__global__ __launch_bounds__(256) void k_dir_to_sph(uint32_t n, const float4* __restrict__ dir,
                                                    float4* __restrict__ pos, float4* __restrict__ pw) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 d = dir[i];
    float phi = atan2f(d.y, d.x);
    if (phi < 0.0f) phi += 2.0f * kPi;
    pos[i].w = phi;
    pw[i].w = acosf(d.z);
}

__global__ __launch_bounds__(256) void k_flags_of(uint32_t n, const uint32_t* __restrict__ valid, uint32_t* __restrict__ f) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[i] = valid[i] ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_append_valid(uint32_t n, const uint32_t* __restrict__ valid,
                                                      const uint32_t* __restrict__ pos, const float4* __restrict__ a,
                                                      const float4* __restrict__ b, float4* __restrict__ da,
                                                      float4* __restrict__ db, uint32_t cap) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !valid[i]) return;
    uint32_t d = pos[i];
    if (d >= cap) return;
    da[d] = a[i];
    db[d] = b[i];
}

}  // namespace mnrt

extern "C" int mnrt_photons_spherical_to_dir(uint32_t n, const float* d_pos4, const float* d_pow4, float* d_out_dir4,
                                             void* stream) {
    if (n == 0) return MNRT_OK;
    if (!d_pos4 || !d_pow4 || !d_out_dir4) return MNRT_ERR_ARG;
    hipLaunchKernelGGL(mnrt::k_sph_to_dir, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n,
                       (const float4*)d_pos4, (const float4*)d_pow4, (float4*)d_out_dir4);
    MNRT_HIP_CHECK(hipGetLastError());
    return MNRT_OK;
}

extern "C" int mnrt_photons_dir_to_spherical(uint32_t n, const float* d_dir4, float* d_io_pos4, float* d_io_pow4,
                                             void* stream) {
    if (n == 0) return MNRT_OK;
    if (!d_dir4 || !d_io_pos4 || !d_io_pow4) return MNRT_ERR_ARG;
    hipLaunchKernelGGL(mnrt::k_dir_to_sph, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n,
                       (const float4*)d_dir4, (float4*)d_io_pos4, (float4*)d_io_pow4);
    MNRT_HIP_CHECK(hipGetLastError());
    return MNRT_OK;
}

extern "C" int mnrt_photons_append_valid(uint32_t n, const uint32_t* d_valid, const float* d_src_pos4,
                                         const float* d_src_pow4, float* d_dst_pos4, float* d_dst_pow4,
                                         uint32_t dst_capacity, uint32_t* out_added, void* stream) {
    using namespace mnrt;
    if (!out_added) return MNRT_ERR_ARG;
    *out_added = 0;
    if (n == 0) return MNRT_OK;
    if (!d_valid || !d_src_pos4 || !d_src_pow4 || (dst_capacity > 0 && (!d_dst_pos4 || !d_dst_pow4)))
        return MNRT_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    void* buf = nullptr;
    const size_t words = 2ull * n + 4 + scan_scratch_words(n);
    MNRT_HIP_CHECK(pool_alloc(&buf, words * 4, st));
    uint32_t* f = (uint32_t*)buf;
    uint32_t* pos = f + n;
    uint32_t* tot = pos + n;
    uint32_t* scr = tot + 4;
    uint32_t* h = (uint32_t*)host_staging(1, 4);
    int rc = h ? MNRT_OK : MNRT_ERR_NOMEM;
    if (rc == MNRT_OK) {
        hipLaunchKernelGGL(k_flags_of, dim3((n + 255) / 256), dim3(256), 0, st, n, d_valid, f);
        if (scan_exclusive(f, pos, n, tot, scr, st) != hipSuccess ||
            hipMemcpyAsync(h, tot, 4, hipMemcpyDeviceToHost, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
            rc = MNRT_ERR_HIP;
    }
    if (rc == MNRT_OK && *h > dst_capacity) rc = MNRT_ERR_CAPACITY;  // nothing written
    if (rc == MNRT_OK && *h > 0) {
        hipLaunchKernelGGL(k_append_valid, dim3((n + 255) / 256), dim3(256), 0, st, n, d_valid, pos,
                           (const float4*)d_src_pos4, (const float4*)d_src_pow4, (float4*)d_dst_pos4,
                           (float4*)d_dst_pow4, dst_capacity);
        if (hipGetLastError() != hipSuccess) rc = MNRT_ERR_HIP;
    }
    if (rc == MNRT_OK) *out_added = *h;
    (void)pool_free(buf, st);
    return rc;
}

extern "C" int mnrt_photons_scale_flux(uint32_t n, float* d_io_pow4, float scale, void* stream) {
    if (n == 0) return MNRT_OK;
    if (!d_io_pow4) return MNRT_ERR_ARG;
    hipLaunchKernelGGL(mnrt::k_scale_power, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n,
                       (float4*)d_io_pow4, scale);
    MNRT_HIP_CHECK(hipGetLastError());
    return MNRT_OK;
}
