// shadow.hip — any-hit ray segments and shadow rays on gfx950.
//
// Replaces kernel_SampleAreaLight, kernel_TraceShadowRaysDelta and
// kernel_TraceShadowRaysArea (raytracing.cu:963-1114), which call the kd-tree traversal
// with needClosest = false.  One lane per segment, 256-lane workgroups; the segment is
// built in registers (no intermediate ray arrays), the top of the traversal stack lives
// in LDS and the search stops at the first occluder (trace_any in trace_dev.h).
// Differences from the reference, all deliberate:
//   * the light is a kernel argument instead of __constant__ c_LightData, the tree an
//     argument instead of the bound KDTreeData;
//   * any occluding triangle decides, not the first one met in a leaf, so the result is
//     the brute-force visibility predicate whatever the tree;
//   * the disc basis uses IEEE 1/sqrt instead of rsqrtf.
#include "trace_dev.h"

namespace mnrt {

struct LightArgs {
    uint32_t type;
    f3 pos, dir, v1, v2;
    float radius;
};

static LightArgs light_args(const mnrt_light& L) {
    LightArgs A;
    A.type = L.type;
    A.pos = f3{L.position[0], L.position[1], L.position[2]};
    A.dir = f3{L.direction[0], L.direction[1], L.direction[2]};
    A.v1 = f3{L.area_v1[0], L.area_v1[1], L.area_v1[2]};
    A.v2 = f3{L.area_v2[0], L.area_v2[1], L.area_v2[2]};
    A.radius = L.area_radius;
    return A;
}

// Concentric map of the square [-1,1]^2 onto the unit disk (Shirley and Chiu 1997, "A
// Low Distortion Map Between Disk and Square"), as dev_ConcentricSampleDisk
// (sample_dev.h:100-155) states it: the square's four wedges map to the disk's four
// quarter sectors, the angle measured in units of pi/4.
__device__ __forceinline__ void concentric_disk(float r1, float r2, float& x, float& y) {
    const float a = 2.0f * r1 - 1.0f, b = 2.0f * r2 - 1.0f;
    if (a == 0.0f && b == 0.0f) {
        x = 0.0f;
        y = 0.0f;
        return;
    }
    float r, q;  // radius, angle / (pi/4)
    if (a >= -b) {
        if (a > b) {
            r = a;
            q = (b > 0.0f) ? b / r : 8.0f + b / r;
        } else {
            r = b;
            q = 2.0f - a / r;
        }
    } else if (a <= b) {
        r = -a;
        q = 4.0f - b / r;
    } else {
        r = -b;
        q = 6.0f + a / r;
    }
    const float ang = q * (kPi / 4.0f);
    x = r * cosf(ang);
    y = r * sinf(ang);
}

// Two unit vectors orthogonal to unit n (dev_BuildCoordSystem, bsdf_dev.h:158-177): the
// first zeroes the smaller of |n.x|, |n.y|, the second is n x first.
__device__ __forceinline__ void disc_basis(f3 n, f3& s, f3& t) {
    if (fabsf(n.x) > fabsf(n.y)) {
        const float il = 1.0f / sqrtf(n.x * n.x + n.z * n.z);
        s = mk3(-n.z * il, 0.0f, n.x * il);
    } else {
        const float il = 1.0f / sqrtf(n.y * n.y + n.z * n.z);
        s = mk3(0.0f, n.z * il, -n.y * il);
    }
    t = cross3(n, s);
}

__global__ __launch_bounds__(256) void k_sample_area_light(LightArgs L, uint32_t n, float ix, float inv_x,
                                                           float iy, float inv_y, const float* __restrict__ rnd1,
                                                           const float* __restrict__ rnd2,
                                                           float4* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float u = (ix + rnd1[i]) * inv_x;
    const float v = (iy + rnd2[i]) * inv_y;
    f3 p;
    if (L.type == MNRT_LIGHT_AREA_DISC) {
        f3 s, t;
        disc_basis(L.dir, s, t);
        float dx, dy;
        concentric_disk(u, v, dx, dy);
        p = mk3(L.pos.x + L.radius * (dx * s.x + dy * t.x), L.pos.y + L.radius * (dx * s.y + dy * t.y),
                L.pos.z + L.radius * (dx * s.z + dy * t.z));
    } else {
        p = mk3(L.pos.x + u * L.v1.x + v * L.v2.x, L.pos.y + u * L.v1.y + v * L.v2.y,
                L.pos.z + u * L.v1.z + v * L.v2.z);
    }
    out[i] = make_float4(p.x, p.y, p.z, 0.0f);
}

enum : int { kSegPoint = 0, kSegDirectional = 1, kSegArea = 2, kSegExplicit = 3 };

// kMode kSegPoint / kSegDirectional / kSegArea: shadow ray from pts[i] to the light,
// out = 1 when unoccluded.  kSegExplicit: ray (pts[i], dirs[i]) limited by tmax[i],
// out = 1 when something is hit.
template <int kMode>
__global__ __launch_bounds__(256) void k_any_hit(const uint2* __restrict__ nodes, const float4* __restrict__ recs,
                                                 RootBox root, LightArgs L, uint32_t n,
                                                 const float4* __restrict__ pts, const float4* __restrict__ dirs,
                                                 const float* __restrict__ tmax, float eps,
                                                 uint32_t* __restrict__ out) {
    extern __shared__ uint32_t s_stack[];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t* lane = s_stack + threadIdx.x;
    const f3 p = xyz(pts[i]);
    if (kMode == kSegExplicit) {
        const float lim = tmax ? tmax[i] : kInf;
        out[i] = trace_any<kTraceStack, kTraceLds>(nodes, recs, root, p, xyz(dirs[i]), eps, lim, 0.0f, lane,
                                                   blockDim.x)
                     ? 1u : 0u;
        return;
    }
    f3 d;
    float dist = kInf, slack = 0.0f;
    if (kMode == kSegDirectional) {
        d = mk3(-L.dir.x, -L.dir.y, -L.dir.z);
    } else {
        const f3 target = (kMode == kSegArea) ? xyz(dirs[i]) : L.pos;
        const f3 v = sub3(target, p);
        dist = sqrtf(dot3(v, v));
        if (!(dist > 0.0f)) {
            out[i] = (kMode == kSegPoint) ? 1u : 0u;
            return;
        }
        d = mk3(v.x / dist, v.y / dist, v.z / dist);
        if (kMode == kSegArea) {
            slack = eps;
            if (!(dot3(mk3(-d.x, -d.y, -d.z), L.dir) > 0.0f)) {
                out[i] = 0u;
                return;
            }
        }
    }
    out[i] = trace_any<kTraceStack, kTraceLds>(nodes, recs, root, p, d, eps, dist, slack, lane, blockDim.x) ? 0u
                                                                                                             : 1u;
}

static int launch_any_hit(int mode, const mnrt_tree* tree, const LightArgs& L, uint32_t n, const float* d_pts4,
                          const float* d_dirs4, const float* d_tmax, float eps, uint32_t* d_out, hipStream_t st) {
    if (tree->kind != MNRT_KIND_TRI || !tree->d_nodes || !tree->d_recs || tree->num_nodes == 0)
        return MNRT_ERR_STATE;
    if (tree->max_level > MNRT_MAX_LEVEL) return MNRT_ERR_CAPACITY;
    if (n == 0) return MNRT_OK;
    RootBox rb;
    for (int a = 0; a < 3; ++a) {
        rb.lo[a] = tree->root_min[a];
        rb.hi[a] = tree->root_max[a];
    }
    const dim3 grid((n + 255) / 256), block(256);
    const uint32_t lds = trace_lds_bytes(kTraceLds, 256);
    const uint2* nodes = (const uint2*)tree->d_nodes;
    const float4* recs = (const float4*)tree->d_recs;
    const float4* pts = (const float4*)d_pts4;
    const float4* dirs = (const float4*)d_dirs4;
    switch (mode) {
        case kSegPoint:
            hipLaunchKernelGGL(k_any_hit<kSegPoint>, grid, block, lds, st, nodes, recs, rb, L, n, pts, dirs, d_tmax,
                               eps, d_out);
            break;
        case kSegDirectional:
            hipLaunchKernelGGL(k_any_hit<kSegDirectional>, grid, block, lds, st, nodes, recs, rb, L, n, pts, dirs,
                               d_tmax, eps, d_out);
            break;
        case kSegArea:
            hipLaunchKernelGGL(k_any_hit<kSegArea>, grid, block, lds, st, nodes, recs, rb, L, n, pts, dirs, d_tmax,
                               eps, d_out);
            break;
        default:
            hipLaunchKernelGGL(k_any_hit<kSegExplicit>, grid, block, lds, st, nodes, recs, rb, L, n, pts, dirs,
                               d_tmax, eps, d_out);
            break;
    }
    MNRT_HIP_CHECK(hipGetLastError());
    return MNRT_OK;
}

}  // namespace mnrt

extern "C" int mnrt_trace_any(const mnrt_tree* tree, uint32_t n, const float* d_orig4, const float* d_dir4,
                              const float* d_tmax, float eps, uint32_t* d_out_hit, void* stream) {
    if (!tree || (n > 0 && (!d_orig4 || !d_dir4 || !d_out_hit))) return MNRT_ERR_ARG;
    mnrt::LightArgs L{};
    return mnrt::launch_any_hit(mnrt::kSegExplicit, tree, L, n, d_orig4, d_dir4, d_tmax, eps, d_out_hit,
                                (hipStream_t)stream);
}

extern "C" int mnrt_sample_area_light(const mnrt_light* light, uint32_t n, float idx_x, float inv_x, float idx_y,
                                      float inv_y, const float* d_rnd1, const float* d_rnd2, float* d_out_pts4,
                                      void* stream) {
    if (!light || (n > 0 && (!d_rnd1 || !d_rnd2 || !d_out_pts4))) return MNRT_ERR_ARG;
    if (light->type != MNRT_LIGHT_AREA_DISC && light->type != MNRT_LIGHT_AREA_RECT) return MNRT_ERR_ARG;
    if (n == 0) return MNRT_OK;
    hipLaunchKernelGGL(mnrt::k_sample_area_light, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       mnrt::light_args(*light), n, idx_x, inv_x, idx_y, inv_y, d_rnd1, d_rnd2,
                       (float4*)d_out_pts4);
    MNRT_HIP_CHECK(hipGetLastError());
    return MNRT_OK;
}

extern "C" int mnrt_trace_shadow(const mnrt_tree* tree, const mnrt_light* light, uint32_t n, const float* d_pts4,
                                 const float* d_samples4, float eps, uint32_t* d_out_unoccluded, void* stream) {
    if (!tree || !light || (n > 0 && (!d_pts4 || !d_out_unoccluded))) return MNRT_ERR_ARG;
    int mode;
    switch (light->type) {
        case MNRT_LIGHT_POINT: mode = mnrt::kSegPoint; break;
        case MNRT_LIGHT_DIRECTIONAL: mode = mnrt::kSegDirectional; break;
        case MNRT_LIGHT_AREA_DISC:
        case MNRT_LIGHT_AREA_RECT:
            if (n > 0 && !d_samples4) return MNRT_ERR_ARG;
            mode = mnrt::kSegArea;
            break;
        default: return MNRT_ERR_ARG;
    }
    return mnrt::launch_any_hit(mode, tree, mnrt::light_args(*light), n, d_pts4, d_samples4, nullptr, eps,
                                d_out_unoccluded, (hipStream_t)stream);
}
