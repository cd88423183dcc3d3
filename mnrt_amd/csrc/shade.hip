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

#include "batch_dev.h"
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

// One shading point from ray i's hit on triangle tr (vertex arrays indexed by g).
__device__ __forceinline__ void shade_one(const float4& o, const float4& r, float t, const float4* __restrict__ v0,
                                          const float4* __restrict__ v1, const float4* __restrict__ v2,
                                          const float4* __restrict__ albedo, uint32_t g, float4 infl,
                                          uint32_t d, float4* __restrict__ out_pos, float4* __restrict__ out_nrm,
                                          float4* __restrict__ out_dir, float4* __restrict__ out_clr,
                                          float4* __restrict__ out_infl) {
    float4 a0 = v0[g], a1 = v1[g], a2 = v2[g];
    f3 nrm = unit_normal(mk3(a1.x - a0.x, a1.y - a0.y, a1.z - a0.z), mk3(a2.x - a0.x, a2.y - a0.y, a2.z - a0.z));
    out_pos[d] = make_float4(o.x + r.x * t, o.y + r.y * t, o.z + r.z * t, 0.0f);
    out_nrm[d] = make_float4(nrm.x, nrm.y, nrm.z, 0.0f);
    out_dir[d] = make_float4(r.x, r.y, r.z, 0.0f);
    out_clr[d] = albedo[g];
    out_infl[d] = infl;
}

__global__ __launch_bounds__(256) void k_shading(ShadeArgs A) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    int32_t tr = A.tri[i];
    if (tr < 0) return;
    uint32_t d = A.pos[i];
    shade_one(A.orig[i], A.dir[i], A.t[i], A.v0, A.v1, A.v2, A.albedo, (uint32_t)tr,
              A.infl ? A.infl[i] : make_float4(1.0f, 1.0f, 1.0f, 0.0f), d, A.out_pos, A.out_nrm, A.out_dir,
              A.out_clr, A.out_infl);
    if (A.out_src) A.out_src[d] = i;
    if (A.out_tri) A.out_tri[d] = tr;
}

// Batched compaction: scene s's hits fill its own slots from slot_begin[s] on, in ray
// order, exactly as mnrt_shading_points would compact that scene alone.
__global__ __launch_bounds__(256) void k_shading_batch(ShadeBatchArgs A) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    int32_t tr = A.tri[i];
    if (tr < 0) return;
    const uint32_t s = seg_of(A.slot_begin, A.num_scenes, i);
    const uint32_t b = A.slot_begin[s];
    const uint32_t d = b + (A.pos[i] - A.pos[b]);
    shade_one(A.orig[i], A.dir[i], A.t[i], A.v0, A.v1, A.v2, A.albedo, A.tri_begin[s] + (uint32_t)tr,
              A.infl ? A.infl[i] : make_float4(1.0f, 1.0f, 1.0f, 0.0f), d, A.out_pos, A.out_nrm, A.out_dir,
              A.out_clr, A.out_infl);
    if (A.out_src) A.out_src[d] = i - b;
}

__global__ __launch_bounds__(256) void k_scene_hits(ShadeBatchArgs A) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < A.num_scenes) A.out_count[s] = A.pos[A.slot_begin[s + 1]] - A.pos[A.slot_begin[s]];
}

int launch_shade_batch(const ShadeBatchArgs& A, uint32_t* flags, uint32_t* scan_tmp, hipStream_t st) {
    // flags: n + 1 words; A.pos: n + 1 words (total at [n]); scan_tmp: scan_scratch_words(n + 1)
    const uint32_t n = A.n;
    hipLaunchKernelGGL(k_hit_flags, dim3((n + 255) / 256), dim3(256), 0, st, n, A.tri, flags);
    MNRT_HIP_CHECK(scan_exclusive(flags, (uint32_t*)A.pos, n, (uint32_t*)A.pos + n, scan_tmp, st));
    hipLaunchKernelGGL(k_shading_batch, dim3((n + 255) / 256), dim3(256), 0, st, A);
    hipLaunchKernelGGL(k_scene_hits, dim3((A.num_scenes + 255) / 256), dim3(256), 0, st, A);
    MNRT_HIP_CHECK(hipGetLastError());
    return MNRT_OK;
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

// Pinhole primary rays through pixel centres (kernel_genPrimaryRays, raygen.cu:171-240),
// in tile order: tiles of tile x tile pixels row by row, pixels row by row inside a tile,
// edge tiles clipped to the image.  The direction is computed in fp64 and rounded once:
// sx = (2 (px + 0.5) / W - 1) tan_half aspect, sy = (1 - 2 (py + 0.5) / H) tan_half,
// d = normalize(fwd + sx right + sy up).
struct CamArgs {
    uint32_t w, h, tile;
    double eye[3], fwd[3], right[3], up[3];
    double th, aspect;
};

__global__ __launch_bounds__(256) void k_primary_rays(CamArgs C, float4* __restrict__ oo, float4* __restrict__ od,
                                                      uint32_t* __restrict__ opix) {
    const uint32_t tiles_x = (C.w + C.tile - 1) / C.tile, tiles_y = (C.h + C.tile - 1) / C.tile;
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t tt = (uint64_t)C.tile * C.tile;
    if (g >= (uint64_t)tiles_x * tiles_y * tt) return;
    const uint32_t t = (uint32_t)(g / tt), j = (uint32_t)(g % tt);
    const uint32_t ty = t / tiles_x, tx = t % tiles_x, iy = j / C.tile, ix = j % C.tile;
    const uint32_t tw = min(C.tile, C.w - tx * C.tile), th = min(C.tile, C.h - ty * C.tile);
    if (ix >= tw || iy >= th) return;
    const uint32_t px = tx * C.tile + ix, py = ty * C.tile + iy;
    // rays before this tile: full tile rows above, full tiles to the left in this row
    const uint64_t slot = (uint64_t)C.w * ty * C.tile + (uint64_t)tx * C.tile * th + (uint64_t)iy * tw + ix;
    const double sx = (2.0 * ((double)px + 0.5) / (double)C.w - 1.0) * C.th * C.aspect;
    const double sy = (1.0 - 2.0 * ((double)py + 0.5) / (double)C.h) * C.th;
    double d[3];
    for (int a = 0; a < 3; ++a) d[a] = (C.fwd[a] + sx * C.right[a]) + sy * C.up[a];
    const double nrm = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    oo[slot] = make_float4((float)C.eye[0], (float)C.eye[1], (float)C.eye[2], 0.0f);
    od[slot] = make_float4((float)(d[0] / nrm), (float)(d[1] / nrm), (float)(d[2] / nrm), 0.0f);
    if (opix) opix[slot] = py * C.w + px;
}

}  // namespace mnrt

extern "C" int mnrt_primary_rays(uint32_t width, uint32_t height, uint32_t tile, const double* h_cam,
                                 float* d_out_orig4, float* d_out_dir4, uint32_t* d_out_pixel, void* stream) {
    using namespace mnrt;
    if (width == 0 || height == 0) return MNRT_OK;
    if (!h_cam || !d_out_orig4 || !d_out_dir4 || tile == 0 || tile > 64) return MNRT_ERR_ARG;
    CamArgs C{};
    C.w = width;
    C.h = height;
    C.tile = tile;
    for (int a = 0; a < 3; ++a) {
        C.eye[a] = h_cam[a];
        C.fwd[a] = h_cam[3 + a];
        C.right[a] = h_cam[6 + a];
        C.up[a] = h_cam[9 + a];
    }
    C.th = h_cam[12];
    C.aspect = h_cam[13];
    const uint64_t total = (uint64_t)((width + tile - 1) / tile) * ((height + tile - 1) / tile) * tile * tile;
    if (total > (1ull << 31) * 64) return MNRT_ERR_CAPACITY;
    hipLaunchKernelGGL(k_primary_rays, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream, C,
                       (float4*)d_out_orig4, (float4*)d_out_dir4, d_out_pixel);
    MNRT_HIP_CHECK(hipGetLastError());
    return MNRT_OK;
}

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
    MNRT_HIP_CHECK(mnrt::pool_alloc((void**)&buf, words * 4, st));
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
    (void)mnrt::pool_free(buf, st);
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
