// tree.hip — tree records, release, cost and status plumbing of the C ABI.
//
// Leaf-ordered records replace the reference's texture-bound element arrays
// (RTUpdateKernelData raytracing.cu:1516-1562, PMSetPhotonMap photon_knn.cu:746-771):
// each leaf reference owns one record, so a leaf's elements are contiguous in HBM and a
// wave reads them with coalesced loads.
#include <cstring>
#include <vector>

#include "mnrt_common.h"

namespace mnrt {

__global__ __launch_bounds__(256) void k_tri_records(const uint32_t* __restrict__ refs, uint32_t num_refs,
                                                     const float4* __restrict__ v0,
                                                     const float4* __restrict__ v1,
                                                     const float4* __restrict__ v2,
                                                     float4* __restrict__ recs) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= num_refs) return;
    uint32_t t = refs[r];
    float4 a = v0[t], b = v1[t], c = v2[t];
    recs[3ull * r + 0] = make_float4(a.x, a.y, a.z, __uint_as_float(t));
    recs[3ull * r + 1] = make_float4(b.x - a.x, b.y - a.y, b.z - a.z, 0.0f);
    recs[3ull * r + 2] = make_float4(c.x - a.x, c.y - a.y, c.z - a.z, 0.0f);
}

__global__ __launch_bounds__(256) void k_photon_records(const uint32_t* __restrict__ refs, uint32_t num_refs,
                                                        const float4* __restrict__ pos,
                                                        const float4* __restrict__ pw,
                                                        const float4* __restrict__ dir,
                                                        float4* __restrict__ recs) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= num_refs) return;
    uint32_t p = refs[r];
    float4 x = pos[p], w = pw[p], d = dir[p];
    recs[2ull * r + 0] = make_float4(x.x, x.y, x.z, __uint_as_float(p));
    recs[2ull * r + 1] = make_float4(w.x, w.y, w.z, __uint_as_float(oct_encode(d.x, d.y, d.z)));
}

// One lane per node: the subtree's reference range runs from its leftmost leaf's first
// reference to the end of its rightmost leaf.
__global__ __launch_bounds__(256) void k_collapse(const uint2* __restrict__ nodes, uint32_t num_nodes,
                                                  uint32_t max_refs, uint2* __restrict__ qnodes) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num_nodes) return;
    uint2 nd = nodes[i];
    if (is_leaf(nd.y)) {
        qnodes[i] = nd;
        return;
    }
    uint32_t l = i;
    uint2 a = nd;
    for (uint32_t s = 0; s <= MNRT_MAX_LEVEL + 1 && !is_leaf(a.y); ++s) {
        l = l + 1;
        a = nodes[l];
    }
    uint32_t r = i;
    uint2 b = nd;
    for (uint32_t s = 0; s <= MNRT_MAX_LEVEL + 1 && !is_leaf(b.y); ++s) {
        r = b.y & MNRT_INDEX_MASK;
        b = nodes[r];
    }
    uint32_t first = a.y & MNRT_INDEX_MASK;
    uint32_t end = (b.y & MNRT_INDEX_MASK) + b.x;
    if (is_leaf(a.y) && is_leaf(b.y) && end >= first && end - first <= max_refs)
        qnodes[i] = make_uint2(end - first, (MNRT_LEAF_TAG << 30) | first);
    else
        qnodes[i] = nd;
}

}  // namespace mnrt

extern "C" int mnrt_tree_collapse(mnrt_tree* tree, uint32_t max_refs, void* stream) {
    if (!tree || !tree->d_nodes || tree->num_nodes == 0) return MNRT_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    if (!tree->d_qnodes) {
        void* p = nullptr;
        if (hipMalloc(&p, (size_t)tree->num_nodes * 8) != hipSuccess) return MNRT_ERR_NOMEM;
        tree->d_qnodes = (uint32_t*)p;
        tree->owns_qnodes = 1;
    }
    hipLaunchKernelGGL(mnrt::k_collapse, dim3((tree->num_nodes + 255) / 256), dim3(256), 0, st,
                       (const uint2*)tree->d_nodes, tree->num_nodes, max_refs, (uint2*)tree->d_qnodes);
    MNRT_HIP_CHECK(hipGetLastError());
    tree->qleaf_max = max_refs;
    return MNRT_OK;
}

static int ensure_recs(mnrt_tree* t, size_t floats_per_ref, hipStream_t st) {
    if (t->d_recs) return MNRT_OK;
    size_t bytes = (size_t)t->num_refs * floats_per_ref * sizeof(float);
    if (bytes == 0) bytes = 16;
    void* p = nullptr;
    (void)st;
    if (hipMalloc(&p, bytes) != hipSuccess) return MNRT_ERR_NOMEM;
    t->d_recs = (float*)p;
    t->owns_recs = 1;
    return MNRT_OK;
}

extern "C" int mnrt_tree_attach_tri_records(mnrt_tree* tree, const float* d_v0, const float* d_v1,
                                            const float* d_v2, void* stream) {
    if (!tree || !tree->d_nodes || (tree->num_refs > 0 && (!tree->d_refs || !d_v0 || !d_v1 || !d_v2)))
        return MNRT_ERR_ARG;
    if (tree->kind != MNRT_KIND_TRI) return MNRT_ERR_STATE;
    hipStream_t st = (hipStream_t)stream;
    int rc = ensure_recs(tree, 12, st);
    if (rc != MNRT_OK) return rc;
    if (tree->num_refs == 0) return MNRT_OK;
    hipLaunchKernelGGL(mnrt::k_tri_records, dim3((tree->num_refs + 255) / 256), dim3(256), 0, st,
                       tree->d_refs, tree->num_refs, (const float4*)d_v0, (const float4*)d_v1,
                       (const float4*)d_v2, (float4*)tree->d_recs);
    MNRT_HIP_CHECK(hipGetLastError());
    return MNRT_OK;
}

extern "C" int mnrt_tree_attach_photon_records(mnrt_tree* tree, const float* d_pos4, const float* d_pow4,
                                               const float* d_dir4, void* stream) {
    if (!tree || !tree->d_nodes || (tree->num_refs > 0 && (!tree->d_refs || !d_pos4 || !d_pow4 || !d_dir4)))
        return MNRT_ERR_ARG;
    if (tree->kind != MNRT_KIND_POINT) return MNRT_ERR_STATE;
    hipStream_t st = (hipStream_t)stream;
    int rc = ensure_recs(tree, 8, st);
    if (rc != MNRT_OK) return rc;
    if (tree->num_refs == 0) return MNRT_OK;
    hipLaunchKernelGGL(mnrt::k_photon_records, dim3((tree->num_refs + 255) / 256), dim3(256), 0, st,
                       tree->d_refs, tree->num_refs, (const float4*)d_pos4, (const float4*)d_pow4,
                       (const float4*)d_dir4, (float4*)tree->d_recs);
    MNRT_HIP_CHECK(hipGetLastError());
    return mnrt_tree_collapse(tree, 64, stream);
}

extern "C" int mnrt_tree_release(mnrt_tree* tree) {
    if (!tree) return MNRT_ERR_ARG;
    int rc = MNRT_OK;
    if (tree->owns_recs && tree->d_recs && hipFree(tree->d_recs) != hipSuccess) rc = MNRT_ERR_HIP;
    if (tree->owns_nodes) {
        if (tree->d_nodes && hipFree(tree->d_nodes) != hipSuccess) rc = MNRT_ERR_HIP;
        if (tree->d_refs && hipFree(tree->d_refs) != hipSuccess) rc = MNRT_ERR_HIP;
    }
    if (tree->owns_qnodes && tree->d_qnodes && hipFree(tree->d_qnodes) != hipSuccess) rc = MNRT_ERR_HIP;
    if (tree->owns_qnodes) tree->d_qnodes = nullptr;
    tree->owns_qnodes = 0;
    if (tree->owns_recs) tree->d_recs = nullptr;
    if (tree->owns_nodes) {
        tree->d_nodes = nullptr;
        tree->d_refs = nullptr;
    }
    tree->owns_recs = tree->owns_nodes = 0;
    return rc;
}

// Cost over the spatial node boxes; a reporting function, evaluated after one copy of
// the node words to the host.
extern "C" int mnrt_tree_cost(const mnrt_tree* tree, float max_query_radius, double* out_cost,
                              void* stream) {
    if (!tree || !out_cost || !tree->d_nodes || tree->num_nodes == 0) return MNRT_ERR_ARG;
    std::vector<uint32_t> w(2ull * tree->num_nodes);
    hipStream_t st = (hipStream_t)stream;
    MNRT_HIP_CHECK(hipMemcpyAsync(w.data(), tree->d_nodes, w.size() * 4, hipMemcpyDeviceToHost, st));
    MNRT_HIP_CHECK(hipStreamSynchronize(st));
    const int kind = (int)tree->kind;
    auto measure = [&](const double* lo, const double* hi) {
        double e[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
        if (kind == (int)MNRT_KIND_POINT) {
            double q = 2.0 * (double)max_query_radius;
            return (e[0] + q) * (e[1] + q) * (e[2] + q);
        }
        return 2.0 * (e[0] * e[1] + e[0] * e[2] + e[1] * e[2]);
    };
    struct Ent {
        uint32_t node;
        double lo[3], hi[3];
    };
    std::vector<Ent> stack;
    Ent r{};
    for (int i = 0; i < 3; ++i) {
        r.lo[i] = tree->root_min[i];
        r.hi[i] = tree->root_max[i];
    }
    double mroot = measure(r.lo, r.hi);
    stack.push_back(r);
    double cost = 0.0;
    size_t guard = 0;
    while (!stack.empty()) {
        if (++guard > tree->num_nodes) return MNRT_ERR_STATE;
        Ent e = stack.back();
        stack.pop_back();
        if (e.node >= tree->num_nodes) return MNRT_ERR_STATE;
        double m = measure(e.lo, e.hi) / mroot;
        uint32_t w1 = w[2 * e.node + 1];
        if ((w1 >> 30) == MNRT_LEAF_TAG) {
            cost += (double)w[2 * e.node] * m;
            continue;
        }
        cost += 3.0 * m;
        int a = (int)(w1 >> 30);
        float sf;
        uint32_t sb = w[2 * e.node];
        std::memcpy(&sf, &sb, 4);
        Ent L = e, R = e;
        L.node = e.node + 1;
        L.hi[a] = sf;
        R.node = w1 & MNRT_INDEX_MASK;
        R.lo[a] = sf;
        stack.push_back(R);
        stack.push_back(L);
    }
    *out_cost = cost;
    return MNRT_OK;
}

extern "C" const char* mnrt_status_string(int status) {
    switch (status) {
        case MNRT_OK: return "ok";
        case MNRT_ERR_ARG: return "invalid argument";
        case MNRT_ERR_HIP: return "HIP runtime error";
        case MNRT_ERR_NOMEM: return "device allocation failed";
        case MNRT_ERR_NODEV: return "no gfx950 device";
        case MNRT_ERR_CAPACITY: return "size exceeds the tree layout's limits";
        case MNRT_ERR_STATE: return "tree lacks what the call needs";
        default: return "unknown status";
    }
}

extern "C" int mnrt_abi_version(void) { return MNRT_ABI_VERSION; }

extern "C" int mnrt_device_init(int device) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0) return MNRT_ERR_NODEV;
    MNRT_HIP_CHECK(hipSetDevice(device));
    hipDeviceProp_t prop;
    MNRT_HIP_CHECK(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return MNRT_ERR_NODEV;
    return MNRT_OK;
}
