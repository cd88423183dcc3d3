// tree.hip — tree records, release, cost and status plumbing of the C ABI.
//
// Leaf-ordered records replace the reference's texture-bound element arrays
// (RTUpdateKernelData raytracing.cu:1516-1562, PMSetPhotonMap photon_knn.cu:746-771):
// each leaf reference owns one record, so a leaf's elements are contiguous in HBM and a
// wave reads them with coalesced loads.
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <vector>

#include "batch_dev.h"

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

// Batched records: reference r of the shared array belongs to tree s (rbase), its element
// to input entry ebase[s] + refs[r].  err (may be null): a nonzero word there means the
// build failed and nothing is written.
__global__ __launch_bounds__(256) void k_tri_records_batch(const uint32_t* __restrict__ refs, DN dnum_refs,
                                                           const uint32_t* __restrict__ rbase,
                                                           const uint32_t* __restrict__ ebase, uint32_t ns,
                                                           const float4* __restrict__ v0,
                                                           const float4* __restrict__ v1,
                                                           const float4* __restrict__ v2,
                                                           float4* __restrict__ recs, const uint32_t* __restrict__ err) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= dnum_refs.get() || (err && *err != 0u)) return;
    uint32_t t = refs[r];
    uint32_t g = ebase[seg_of(rbase, ns, r)] + t;
    float4 a = v0[g], b = v1[g], c = v2[g];
    recs[3ull * r + 0] = make_float4(a.x, a.y, a.z, __uint_as_float(t));
    recs[3ull * r + 1] = make_float4(b.x - a.x, b.y - a.y, b.z - a.z, 0.0f);
    recs[3ull * r + 2] = make_float4(c.x - a.x, c.y - a.y, c.z - a.z, 0.0f);
}

__global__ __launch_bounds__(256) void k_photon_records_batch(const uint32_t* __restrict__ refs, DN dnum_refs,
                                                              const uint32_t* __restrict__ rbase,
                                                              const uint32_t* __restrict__ ebase, uint32_t ns,
                                                              const float4* __restrict__ pos,
                                                              const float4* __restrict__ pw,
                                                              const float4* __restrict__ dir,
                                                              float4* __restrict__ recs,
                                                              const uint32_t* __restrict__ err) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= dnum_refs.get() || (err && *err != 0u)) return;
    uint32_t p = refs[r];
    uint32_t g = ebase[seg_of(rbase, ns, r)] + p;
    float4 x = pos[g], w = pw[g], d = dir[g];
    recs[2ull * r + 0] = make_float4(x.x, x.y, x.z, __uint_as_float(p));
    recs[2ull * r + 1] = make_float4(w.x, w.y, w.z, __uint_as_float(oct_encode(d.x, d.y, d.z)));
}

// One lane per node: the subtree's reference range runs from its leftmost leaf's first
// reference to the end of its rightmost leaf.  nbase (ns + 1 entries, or null for one
// tree) places the trees of a batch in the shared node array; err as for the records.
__global__ __launch_bounds__(256) void k_collapse(const uint2* __restrict__ nodes_all, DN dnum_nodes,
                                                  const uint32_t* __restrict__ nbase, uint32_t ns,
                                                  uint32_t max_refs, uint2* __restrict__ qnodes_all,
                                                  const uint32_t* __restrict__ err) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= dnum_nodes.get() || (err && *err != 0u)) return;
    const uint32_t b = nbase ? nbase[seg_of(nbase, ns, g)] : 0u;
    const uint2* nodes = nodes_all + b;
    uint2* qnodes = qnodes_all + b;
    const uint32_t i = g - b;
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
    uint2 c = nd;
    for (uint32_t s = 0; s <= MNRT_MAX_LEVEL + 1 && !is_leaf(c.y); ++s) {
        r = c.y & MNRT_INDEX_MASK;
        c = nodes[r];
    }
    uint32_t first = a.y & MNRT_INDEX_MASK;
    uint32_t end = (c.y & MNRT_INDEX_MASK) + c.x;
    if (is_leaf(a.y) && is_leaf(c.y) && end >= first && end - first <= max_refs)
        qnodes[i] = make_uint2(end - first, (MNRT_LEAF_TAG << 30) | first);
    else
        qnodes[i] = nd;
}

// Records (and, for photon trees, query-time node words) of the trees of a batch whose
// sizes are device words (a recorded build): grids cover the capacities nrefs.v / nnodes.v.
int launch_records_batch(int kind, const uint32_t* refs, DN nrefs, const uint32_t* rbase, const uint32_t* ebase,
                         uint32_t ns, const float4* a, const float4* b, const float4* c, float4* recs,
                         const uint32_t* err, hipStream_t st) {
    const dim3 grid((std::max<uint32_t>(nrefs.v, 1u) + 255) / 256);
    if (kind == MNRT_KIND_TRI)
        hipLaunchKernelGGL(k_tri_records_batch, grid, dim3(256), 0, st, refs, nrefs, rbase, ebase, ns, a, b, c, recs,
                           err);
    else
        hipLaunchKernelGGL(k_photon_records_batch, grid, dim3(256), 0, st, refs, nrefs, rbase, ebase, ns, a, b, c,
                           recs, err);
    return hipGetLastError() == hipSuccess ? MNRT_OK : MNRT_ERR_HIP;
}

int launch_collapse_batch(const uint2* nodes, DN nnodes, const uint32_t* nbase, uint32_t ns, uint32_t max_refs,
                          uint2* qnodes, const uint32_t* err, hipStream_t st) {
    const dim3 grid((std::max<uint32_t>(nnodes.v, 1u) + 255) / 256);
    hipLaunchKernelGGL(k_collapse, grid, dim3(256), 0, st, nodes, nnodes, nbase, ns, max_refs, qnodes, err);
    return hipGetLastError() == hipSuccess ? MNRT_OK : MNRT_ERR_HIP;
}

}  // namespace mnrt

namespace mnrt {

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_pool_mutex;
hipMemPool_t g_pools[kMaxDevices] = {};

hipMemPool_t device_pool() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_pool_mutex);
    if (!g_pools[dev]) {
        hipMemPoolProps props{};
        props.allocType = hipMemAllocationTypePinned;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = dev;
        hipMemPool_t mp = nullptr;
        if (hipMemPoolCreate(&mp, &props) != hipSuccess) return nullptr;
        uint64_t keep = UINT64_MAX;
        (void)hipMemPoolSetAttribute(mp, hipMemPoolAttrReleaseThreshold, &keep);
        g_pools[dev] = mp;
    }
    return g_pools[dev];
}
}  // namespace

hipError_t pool_alloc(void** p, size_t bytes, hipStream_t st) {
    hipMemPool_t mp = device_pool();
    if (!mp) return hipErrorOutOfMemory;
    return hipMallocFromPoolAsync(p, std::max<size_t>(bytes, 16), mp, st);
}

hipError_t pool_free(void* p, hipStream_t st) { return p ? hipFreeAsync(p, st) : hipSuccess; }

void* host_staging(int slot, size_t bytes) {
    // never freed: the buffers live until the process ends (a thread-exit free could run
    // after the HIP runtime has shut down)
    thread_local void* buf[2] = {nullptr, nullptr};
    thread_local size_t cap[2] = {0, 0};
    if (slot < 0 || slot > 1) return nullptr;
    if (cap[slot] < bytes) {
        size_t want = std::max<size_t>(bytes, 4096);
        void* p = nullptr;
        if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) return nullptr;
        if (buf[slot]) (void)hipHostFree(buf[slot]);
        buf[slot] = p;
        cap[slot] = want;
    }
    return buf[slot];
}

}  // namespace mnrt

extern "C" int mnrt_tree_collapse(mnrt_tree* tree, uint32_t max_refs, void* stream) {
    if (!tree || !tree->d_nodes || tree->num_nodes == 0) return MNRT_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    if (!tree->d_qnodes) {
        void* p = nullptr;
        if (mnrt::pool_alloc(&p, (size_t)tree->num_nodes * 8, st) != hipSuccess) return MNRT_ERR_NOMEM;
        tree->d_qnodes = (uint32_t*)p;
        tree->owns_qnodes = 1;
    }
    hipLaunchKernelGGL(mnrt::k_collapse, dim3((tree->num_nodes + 255) / 256), dim3(256), 0, st,
                       (const uint2*)tree->d_nodes, mnrt::hn(tree->num_nodes), (const uint32_t*)nullptr, 1u, max_refs,
                       (uint2*)tree->d_qnodes, (const uint32_t*)nullptr);
    MNRT_HIP_CHECK(hipGetLastError());
    tree->qleaf_max = max_refs;
    return MNRT_OK;
}

static int ensure_recs(mnrt_tree* t, size_t floats_per_ref, hipStream_t st) {
    if (t->d_recs) return MNRT_OK;
    size_t bytes = (size_t)t->num_refs * floats_per_ref * sizeof(float);
    if (bytes == 0) bytes = 16;
    void* p = nullptr;
    if (mnrt::pool_alloc(&p, bytes, st) != hipSuccess) return MNRT_ERR_NOMEM;
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
    return mnrt_tree_collapse(tree, mnrt::kQueryLeafRefs, stream);
}

// ---- batches: trees[s] view shared arrays owned by trees[0] ----

namespace {

struct BatchTables {
    std::vector<uint32_t> h;  // nbase (S+1), rbase (S+1), ebase (S+1)
    uint32_t total_nodes = 0, total_refs = 0;
};

int batch_tables(const mnrt_tree* t, uint32_t ns, const uint32_t* h_offsets, uint32_t kind, BatchTables& B) {
    if (!t || ns == 0 || !h_offsets || !t[0].d_nodes) return MNRT_ERR_ARG;
    B.h.assign(3ull * (ns + 1), 0u);
    uint32_t* nb = B.h.data();
    uint32_t* rb = nb + ns + 1;
    uint32_t* eb = rb + ns + 1;
    for (uint32_t s = 0; s < ns; ++s) {
        if (t[s].kind != kind) return MNRT_ERR_STATE;
        const ptrdiff_t dn = t[s].d_nodes - t[0].d_nodes, dr = t[s].d_refs - t[0].d_refs;
        if (dn != 2 * (ptrdiff_t)B.total_nodes || dr != (ptrdiff_t)B.total_refs) return MNRT_ERR_ARG;
        if (h_offsets[s + 1] < h_offsets[s] || h_offsets[s + 1] - h_offsets[s] != t[s].num_elems) return MNRT_ERR_ARG;
        nb[s] = B.total_nodes;
        rb[s] = B.total_refs;
        eb[s] = h_offsets[s];
        B.total_nodes += t[s].num_nodes;
        B.total_refs += t[s].num_refs;
    }
    nb[ns] = B.total_nodes;
    rb[ns] = B.total_refs;
    eb[ns] = h_offsets[ns];
    return MNRT_OK;
}

int upload(const BatchTables& B, uint32_t** d, hipStream_t st) {
    if (mnrt::pool_alloc((void**)d, B.h.size() * 4, st) != hipSuccess) return MNRT_ERR_NOMEM;
    MNRT_HIP_CHECK(hipMemcpyAsync(*d, B.h.data(), B.h.size() * 4, hipMemcpyHostToDevice, st));
    return MNRT_OK;
}

}  // namespace

extern "C" int mnrt_tree_attach_tri_records_batch(mnrt_tree* trees, uint32_t num_scenes, const float* d_v0,
                                                  const float* d_v1, const float* d_v2,
                                                  const uint32_t* h_offsets, void* stream) {
    BatchTables B;
    int rc = batch_tables(trees, num_scenes, h_offsets, MNRT_KIND_TRI, B);
    if (rc != MNRT_OK) return rc;
    if (B.total_refs > 0 && (!d_v0 || !d_v1 || !d_v2)) return MNRT_ERR_ARG;
    for (uint32_t s = 0; s < num_scenes; ++s)
        if (trees[s].d_recs) return MNRT_ERR_STATE;
    hipStream_t st = (hipStream_t)stream;
    float* recs = nullptr;
    if (mnrt::pool_alloc((void**)&recs, (size_t)B.total_refs * 48, st) != hipSuccess) return MNRT_ERR_NOMEM;
    uint32_t* d_tab = nullptr;
    rc = upload(B, &d_tab, st);
    if (rc == MNRT_OK && B.total_refs > 0) {
        const uint32_t ns = num_scenes;
        hipLaunchKernelGGL(mnrt::k_tri_records_batch, dim3((B.total_refs + 255) / 256), dim3(256), 0, st,
                           trees[0].d_refs, mnrt::hn(B.total_refs), d_tab + ns + 1, d_tab + 2 * (ns + 1), ns,
                           (const float4*)d_v0, (const float4*)d_v1, (const float4*)d_v2, (float4*)recs,
                           (const uint32_t*)nullptr);
        if (hipGetLastError() != hipSuccess) rc = MNRT_ERR_HIP;
    }
    if (d_tab) (void)mnrt::pool_free(d_tab, st);
    if (hipStreamSynchronize(st) != hipSuccess && rc == MNRT_OK) rc = MNRT_ERR_HIP;
    if (rc != MNRT_OK) {
        (void)mnrt::pool_free(recs, st);
        return rc;
    }
    for (uint32_t s = 0; s < num_scenes; ++s) {
        trees[s].d_recs = recs + 12ull * B.h[num_scenes + 1 + s];
        trees[s].owns_recs = s == 0 ? 1u : 0u;
    }
    return MNRT_OK;
}

extern "C" int mnrt_tree_attach_photon_records_batch(mnrt_tree* trees, uint32_t num_scenes, const float* d_pos4,
                                                     const float* d_pow4, const float* d_dir4,
                                                     const uint32_t* h_offsets, void* stream) {
    BatchTables B;
    int rc = batch_tables(trees, num_scenes, h_offsets, MNRT_KIND_POINT, B);
    if (rc != MNRT_OK) return rc;
    if (B.total_refs > 0 && (!d_pos4 || !d_pow4 || !d_dir4)) return MNRT_ERR_ARG;
    const uint32_t ns = num_scenes;
    hipStream_t st = (hipStream_t)stream;
    bool all = true, none = true;
    for (uint32_t s = 0; s < ns; ++s) {
        all = all && trees[s].d_recs && trees[s].d_qnodes;
        none = none && !trees[s].d_recs && !trees[s].d_qnodes;
    }
    if (!all && !none) return MNRT_ERR_STATE;
    if (all) {
        // records attached by an earlier call: rewritten in place (new photon powers, say);
        // the query-time nodes depend on the tree only and stay
        for (uint32_t s = 0; s < ns; ++s)
            if (trees[s].d_recs != trees[0].d_recs + 8ull * B.h[ns + 1 + s]) return MNRT_ERR_STATE;
        uint32_t* d_tab = nullptr;
        rc = upload(B, &d_tab, st);
        if (rc == MNRT_OK && B.total_refs > 0) {
            hipLaunchKernelGGL(mnrt::k_photon_records_batch, dim3((B.total_refs + 255) / 256), dim3(256), 0, st,
                               trees[0].d_refs, mnrt::hn(B.total_refs), d_tab + ns + 1, d_tab + 2 * (ns + 1), ns,
                               (const float4*)d_pos4, (const float4*)d_pow4, (const float4*)d_dir4,
                               (float4*)trees[0].d_recs, (const uint32_t*)nullptr);
            if (hipGetLastError() != hipSuccess) rc = MNRT_ERR_HIP;
        }
        if (d_tab) (void)mnrt::pool_free(d_tab, st);
        if (hipStreamSynchronize(st) != hipSuccess && rc == MNRT_OK) rc = MNRT_ERR_HIP;
        return rc;
    }
    float* recs = nullptr;
    uint32_t* qn = nullptr;
    if (mnrt::pool_alloc((void**)&recs, (size_t)B.total_refs * 32, st) != hipSuccess) return MNRT_ERR_NOMEM;
    if (mnrt::pool_alloc((void**)&qn, (size_t)B.total_nodes * 8, st) != hipSuccess) {
        (void)mnrt::pool_free(recs, st);
        return MNRT_ERR_NOMEM;
    }
    uint32_t* d_tab = nullptr;
    rc = upload(B, &d_tab, st);
    if (rc == MNRT_OK && B.total_refs > 0)
        hipLaunchKernelGGL(mnrt::k_photon_records_batch, dim3((B.total_refs + 255) / 256), dim3(256), 0, st,
                           trees[0].d_refs, mnrt::hn(B.total_refs), d_tab + ns + 1, d_tab + 2 * (ns + 1), ns,
                           (const float4*)d_pos4, (const float4*)d_pow4, (const float4*)d_dir4, (float4*)recs,
                           (const uint32_t*)nullptr);
    if (rc == MNRT_OK)
        hipLaunchKernelGGL(mnrt::k_collapse, dim3((B.total_nodes + 255) / 256), dim3(256), 0, st,
                           (const uint2*)trees[0].d_nodes, mnrt::hn(B.total_nodes), (const uint32_t*)d_tab, ns,
                           mnrt::kQueryLeafRefs, (uint2*)qn, (const uint32_t*)nullptr);
    if (rc == MNRT_OK && hipGetLastError() != hipSuccess) rc = MNRT_ERR_HIP;
    if (d_tab) (void)mnrt::pool_free(d_tab, st);
    if (hipStreamSynchronize(st) != hipSuccess && rc == MNRT_OK) rc = MNRT_ERR_HIP;
    if (rc != MNRT_OK) {
        (void)mnrt::pool_free(recs, st);
        (void)mnrt::pool_free(qn, st);
        return rc;
    }
    for (uint32_t s = 0; s < num_scenes; ++s) {
        trees[s].d_recs = recs + 8ull * B.h[ns + 1 + s];
        trees[s].d_qnodes = qn + 2ull * B.h[s];
        trees[s].qleaf_max = mnrt::kQueryLeafRefs;
        trees[s].owns_recs = trees[s].owns_qnodes = s == 0 ? 1u : 0u;
    }
    return MNRT_OK;
}

extern "C" int mnrt_tree_release(mnrt_tree* tree) {
    if (!tree) return MNRT_ERR_ARG;
    // library arrays come from the stream-ordered pool; the null stream orders the frees
    // after all work queued on blocking streams
    int rc = MNRT_OK;
    if (tree->owns_recs && tree->d_recs && mnrt::pool_free(tree->d_recs, nullptr) != hipSuccess) rc = MNRT_ERR_HIP;
    if (tree->owns_nodes) {
        if (tree->d_nodes && mnrt::pool_free(tree->d_nodes, nullptr) != hipSuccess) rc = MNRT_ERR_HIP;
        if (tree->d_refs && mnrt::pool_free(tree->d_refs, nullptr) != hipSuccess) rc = MNRT_ERR_HIP;
    }
    if (tree->owns_qnodes && tree->d_qnodes && mnrt::pool_free(tree->d_qnodes, nullptr) != hipSuccess)
        rc = MNRT_ERR_HIP;
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
    return mnrt::knn_scratch_init(device);
}
