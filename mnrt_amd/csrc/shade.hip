// shade.hip — the steps either side of traversal that turn hits into kNN queries.
//
//   mnrt_shading_points: RTCore::TraceRays compaction of valid hits
//     (RTCore.cpp:416-429) + KernelBSDFGetNormalsAtHit (bsdf.cu:317-367, geometric
//     normal only: the synthetic scenes carry no per-vertex normals) + the diffuse colour
//     lookup of kernel_GetDiffuseColors (raytracing.cu:1298).
//   mnrt_bounce_rays: one cosine-distributed gather ray per shading point
//     (kernel_genFinalGatherRays, raygen.cu:406-487), about the normal turned toward the
//     arriving ray; the ray's influence is the surface albedo (cosine sampling cancels
//     the cosine / pi of the diffuse BRDF).
#include <algorithm>

#include "sample_dev.h"
#include "scan.h"

namespace mnrt {

__global__ __launch_bounds__(256) void k_hit_flags(uint32_t n, const int32_t* __restrict__ tri, uint32_t* __restrict__ f) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[i] = tri[i] >= 0 ? 1u : 0u;
}

struct ShadeArgs {
    uint32_t n;
    const float4* orig;
    const float4* dir;
    const int32_t* tri;
    const float* t;
    const float4* v0;
    const float4* v1;
    const float4* v2;
    const float4* albedo;
    const float4* infl;
    const uint32_t* pos;  // compaction positions
    float4* out_pos;
    float4* out_nrm;
    float4* out_dir;
    float4* out_clr;
    float4* out_infl;
    uint32_t* out_src;
    int32_t* out_tri;
};

__global__ __launch_bounds__(256) void k_shading(ShadeArgs A) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    int32_t tr = A.tri[i];
    if (tr < 0) return;
    uint32_t d = A.pos[i];
    float4 o = A.orig[i], r = A.dir[i];
    float t = A.t[i];
    float4 a0 = A.v0[tr], a1 = A.v1[tr], a2 = A.v2[tr];
    f3 nrm = unit_normal(mk3(a1.x - a0.x, a1.y - a0.y, a1.z - a0.z), mk3(a2.x - a0.x, a2.y - a0.y, a2.z - a0.z));
    A.out_pos[d] = make_float4(o.x + r.x * t, o.y + r.y * t, o.z + r.z * t, 0.0f);
    A.out_nrm[d] = make_float4(nrm.x, nrm.y, nrm.z, 0.0f);
    A.out_dir[d] = make_float4(r.x, r.y, r.z, 0.0f);
    A.out_clr[d] = A.albedo[tr];
    A.out_infl[d] = A.infl ? A.infl[i] : make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    if (A.out_src) A.out_src[d] = i;
    if (A.out_tri) A.out_tri[d] = tr;
}

__global__ __launch_bounds__(256) void k_bounce(uint32_t n, const float4* __restrict__ pos, const float4* __restrict__ nrm,
                                                const float4* __restrict__ dir, const float4* __restrict__ clr,
                                                uint32_t seed, float4* __restrict__ oo, float4* __restrict__ od,
                                                float4* __restrict__ oinfl) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f3 nq = xyz(nrm[i]);
    f3 rd = xyz(dir[i]);
    f3 nrd = mk3(-rd.x, -rd.y, -rd.z);
    if (dot3(nrd, nq) < 0.0f) nq = mk3(nq.x * -1.0f, nq.y * -1.0f, nq.z * -1.0f);
    f3 d = cosine_dir(nq, seed, i, 0xFFFDu);
    float4 p = pos[i];
    oo[i] = make_float4(p.x, p.y, p.z, 0.0f);
    od[i] = make_float4(d.x, d.y, d.z, 0.0f);
    if (oinfl) {
        float4 c = clr[i];
        oinfl[i] = make_float4(c.x, c.y, c.z, 0.0f);
    }
}

}  // namespace mnrt

extern "C" int mnrt_shading_points(uint32_t n, const float* d_orig4, const float* d_dir4, const int32_t* d_hit_tri,
                                   const float* d_hit_t, const float* d_v0, const float* d_v1, const float* d_v2,
                                   const float* d_albedo4, const float* d_infl4, float* d_out_pos4, float* d_out_nrm4,
                                   float* d_out_dir4, float* d_out_clr4, float* d_out_infl4, uint32_t* d_out_src,
                                   int32_t* d_out_tri, uint32_t* out_count, void* stream) {
    using namespace mnrt;
    if (!out_count) return MNRT_ERR_ARG;
    *out_count = 0;
    if (n == 0) return MNRT_OK;
    if (!d_orig4 || !d_dir4 || !d_hit_tri || !d_hit_t || !d_v0 || !d_v1 || !d_v2 || !d_albedo4 || !d_out_pos4 ||
        !d_out_nrm4 || !d_out_dir4 || !d_out_clr4 || !d_out_infl4)
        return MNRT_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    size_t words = 2ull * n + scan_scratch_words(n) + 8;
    uint32_t* buf = nullptr;
    MNRT_HIP_CHECK(hipMallocAsync((void**)&buf, words * 4, st));
    uint32_t* flags = buf;
    uint32_t* pos = buf + n;
    uint32_t* tot = buf + 2ull * n;
    uint32_t* scr = tot + 4;
    int rc = MNRT_OK;
    hipLaunchKernelGGL(k_hit_flags, dim3((n + 255) / 256), dim3(256), 0, st, n, d_hit_tri, flags);
    if (scan_exclusive(flags, pos, n, tot, scr, st) != hipSuccess) rc = MNRT_ERR_HIP;
    if (rc == MNRT_OK) {
        ShadeArgs A{n, (const float4*)d_orig4, (const float4*)d_dir4, d_hit_tri, d_hit_t, (const float4*)d_v0,
                    (const float4*)d_v1, (const float4*)d_v2, (const float4*)d_albedo4, (const float4*)d_infl4, pos,
                    (float4*)d_out_pos4, (float4*)d_out_nrm4, (float4*)d_out_dir4, (float4*)d_out_clr4,
                    (float4*)d_out_infl4, d_out_src, d_out_tri};
        hipLaunchKernelGGL(k_shading, dim3((n + 255) / 256), dim3(256), 0, st, A);
        uint32_t h = 0;
        if (hipMemcpyAsync(&h, tot, 4, hipMemcpyDeviceToHost, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
            rc = MNRT_ERR_HIP;
        *out_count = h;
    }
    (void)hipFreeAsync(buf, st);
    return rc;
}

extern "C" int mnrt_bounce_rays(uint32_t n, const float* d_pos4, const float* d_nrm4, const float* d_dir4,
                                const float* d_clr4, uint32_t seed, float* d_out_orig4, float* d_out_dir4,
                                float* d_out_infl4, void* stream) {
    using namespace mnrt;
    if (n == 0) return MNRT_OK;
    if (!d_pos4 || !d_nrm4 || !d_dir4 || !d_out_orig4 || !d_out_dir4 || (d_out_infl4 && !d_clr4)) return MNRT_ERR_ARG;
    hipLaunchKernelGGL(k_bounce, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, (const float4*)d_pos4,
                       (const float4*)d_nrm4, (const float4*)d_dir4, (const float4*)d_clr4, seed, (float4*)d_out_orig4,
                       (float4*)d_out_dir4, (float4*)d_out_infl4);
    MNRT_HIP_CHECK(hipGetLastError());
    return MNRT_OK;
}
