// export.hip — the reference's KDTreeData layout from a packed tree, and its custom bits.
//
// KDTreeData (KDKernelDefs.h:126-194) as kernel_TraversalDownPath writes it
// (kdtree.cu:1707-1770): per node its element count, its word address in the preorder
// stream, its extent (centre, half diagonal), child indices; the preorder stream holds
//   inner: [node | 0] [axis << 30 | custom bits << 28 | address of the right child] [split bits]
//   leaf : [node | 0x80000000] [count] [element indices ...]
// with the left child right after its parent's three words.  Node indices are preorder
// indices here.  The right-child address keeps all 28 bits of its field (the reference
// masks 24, kdtree.cu:1742); streams longer than 2^28 words are refused.
#include <algorithm>

#include "mnrt_common.h"
#include "scan.h"

namespace mnrt {

__global__ __launch_bounds__(256) void k_export_sizes(const uint2* __restrict__ nodes, uint32_t n,
                                                      uint32_t* __restrict__ words) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint2 nd = nodes[i];
    words[i] = is_leaf(nd.y) ? 2u + nd.x : 3u;
}

// parent[c] = p (left child) or p | 2^31 (right child); the root keeps kNoParent
__global__ __launch_bounds__(256) void k_export_parents(const uint2* __restrict__ nodes, uint32_t n,
                                                        uint32_t* __restrict__ parent) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i == 0) parent[0] = 0xFFFFFFFFu;
    uint2 nd = nodes[i];
    if (is_leaf(nd.y)) return;
    uint32_t r = nd.y & MNRT_INDEX_MASK;
    if (i + 1 < n) parent[i + 1] = i;
    if (r < n) parent[r] = i | 0x80000000u;
}

struct ExportArgs {
    const uint2* nodes;
    const uint32_t* refs;
    uint32_t n;
    float lo[3], hi[3];
    const uint32_t* parent;
    const uint32_t* addr;  // exclusive scan of the word counts
    uint32_t* tree;
    uint32_t* num_elems;
    uint32_t* node_addr;
    float4* extent;
    uint32_t* child_left;
    uint32_t* child_right;
    uint32_t* err;
};

__global__ __launch_bounds__(256) void k_export(ExportArgs A) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const uint2 nd = A.nodes[i];
    const uint32_t a = A.addr[i];
    // node box: the root box cut by the splits on the path (the nearest ancestor's cut
    // per side and axis is the tightest, min / max keeps it whatever the order)
    float lo[3] = {A.lo[0], A.lo[1], A.lo[2]}, hi[3] = {A.hi[0], A.hi[1], A.hi[2]};
    uint32_t c = i;
    for (uint32_t step = 0; step <= MNRT_MAX_LEVEL + 1; ++step) {
        const uint32_t pw = A.parent[c];
        if (pw == 0xFFFFFFFFu) break;
        const uint32_t p = pw & 0x7FFFFFFFu;
        const uint2 pn = A.nodes[p];
        const uint32_t ax = pn.y >> 30;
        const float s = __uint_as_float(pn.x);
        if (pw & 0x80000000u) lo[ax] = fmaxf(lo[ax], s);
        else hi[ax] = fminf(hi[ax], s);
        c = p;
    }
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    const float radius = 0.5f * sqrtf((dx * dx + dy * dy) + dz * dz);
    A.extent[i] = make_float4(lo[0] + 0.5f * dx, lo[1] + 0.5f * dy, lo[2] + 0.5f * dz, radius);
    A.node_addr[i] = a;
    if (is_leaf(nd.y)) {
        const uint32_t first = nd.y & MNRT_INDEX_MASK, cnt = nd.x;
        A.tree[a] = i | 0x80000000u;
        A.tree[a + 1] = cnt;
        for (uint32_t k = 0; k < cnt; ++k) A.tree[a + 2 + k] = A.refs[first + k];
        A.num_elems[i] = cnt;
        A.child_left[i] = 0;
        A.child_right[i] = 0;
        return;
    }
    const uint32_t r = nd.y & MNRT_INDEX_MASK;
    const uint32_t ar = A.addr[r];
    if (ar > 0x0FFFFFFFu) atomicOr(A.err, 1u);
    A.tree[a] = i;
    A.tree[a + 1] = ((nd.y >> 30) << 30) | (ar & 0x0FFFFFFFu);
    A.tree[a + 2] = nd.x;
    A.child_left[i] = i + 1;
    A.child_right[i] = r;
    // subtree reference count: leftmost leaf's first reference to the rightmost leaf's end
    uint32_t l = i + 1;
    uint2 ln = A.nodes[l];
    for (uint32_t step = 0; step <= MNRT_MAX_LEVEL + 1 && !is_leaf(ln.y); ++step) ln = A.nodes[++l];
    uint32_t rr = r;
    uint2 rn = A.nodes[rr];
    for (uint32_t step = 0; step <= MNRT_MAX_LEVEL + 1 && !is_leaf(rn.y); ++step) {
        rr = rn.y & MNRT_INDEX_MASK;
        rn = A.nodes[rr];
    }
    A.num_elems[i] = ((rn.y & MNRT_INDEX_MASK) + rn.x) - (ln.y & MNRT_INDEX_MASK);
}

__global__ __launch_bounds__(256) void k_custom_bits(uint32_t* __restrict__ tree, const uint32_t* __restrict__ node_addr,
                                                     uint32_t n, uint32_t bit, const uint32_t* __restrict__ values) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t a = node_addr[i];
    if (tree[a] & 0x80000000u) return;  // leaves carry no custom bits
    const uint32_t pos = 28 + bit;
    tree[a + 1] = (tree[a + 1] & ~(1u << pos)) | ((values[i] & 1u) << pos);
}

}  // namespace mnrt

extern "C" int mnrt_tree_preorder_size(const mnrt_tree* tree, uint64_t* out_words) {
    if (!tree || !out_words || tree->num_nodes == 0) return MNRT_ERR_ARG;
    const uint64_t inner = (uint64_t)tree->num_nodes - tree->num_leaves;
    *out_words = 3ull * inner + 2ull * tree->num_leaves + tree->num_refs;
    return MNRT_OK;
}

extern "C" int mnrt_tree_export_preorder(const mnrt_tree* tree, uint32_t* d_tree, uint64_t capacity_words,
                                         uint32_t* d_num_elems, uint32_t* d_node_addr, float* d_node_extent4,
                                         uint32_t* d_child_left, uint32_t* d_child_right, void* stream) {
    using namespace mnrt;
    uint64_t need = 0;
    int rc = mnrt_tree_preorder_size(tree, &need);
    if (rc != MNRT_OK) return rc;
    if (!tree->d_nodes || (tree->num_refs && !tree->d_refs)) return MNRT_ERR_STATE;
    if (!d_tree || !d_num_elems || !d_node_addr || !d_node_extent4 || !d_child_left || !d_child_right)
        return MNRT_ERR_ARG;
    if (capacity_words < need) return MNRT_ERR_ARG;
    if (need > 0x0FFFFFFFull) return MNRT_ERR_CAPACITY;
    hipStream_t st = (hipStream_t)stream;
    const uint32_t n = tree->num_nodes;
    void* buf = nullptr;
    const size_t words = 2ull * n + 2 + scan_scratch_words(n) + 4;
    MNRT_HIP_CHECK(pool_alloc(&buf, words * 4, st));
    uint32_t* sizes = (uint32_t*)buf;
    uint32_t* parent = sizes + n;
    uint32_t* err = parent + n;
    uint32_t* scr = err + 2;
    const dim3 grid((n + 255) / 256), block(256);
    hipLaunchKernelGGL(k_export_sizes, grid, block, 0, st, (const uint2*)tree->d_nodes, n, sizes);
    hipLaunchKernelGGL(k_export_parents, grid, block, 0, st, (const uint2*)tree->d_nodes, n, parent);
    rc = scan_exclusive(sizes, sizes, n, nullptr, scr, st) == hipSuccess ? MNRT_OK : MNRT_ERR_HIP;
    uint32_t* h_err = (uint32_t*)host_staging(1, 4);
    if (rc == MNRT_OK && !h_err) rc = MNRT_ERR_NOMEM;
    if (rc == MNRT_OK) {
        ExportArgs A{};
        A.nodes = (const uint2*)tree->d_nodes;
        A.refs = tree->d_refs;
        A.n = n;
        for (int a = 0; a < 3; ++a) {
            A.lo[a] = tree->root_min[a];
            A.hi[a] = tree->root_max[a];
        }
        A.parent = parent;
        A.addr = sizes;
        A.tree = d_tree;
        A.num_elems = d_num_elems;
        A.node_addr = d_node_addr;
        A.extent = (float4*)d_node_extent4;
        A.child_left = d_child_left;
        A.child_right = d_child_right;
        A.err = err;
        if (hipMemsetAsync(err, 0, 4, st) != hipSuccess) rc = MNRT_ERR_HIP;
        if (rc == MNRT_OK) hipLaunchKernelGGL(k_export, grid, block, 0, st, A);
        if (rc == MNRT_OK && (hipGetLastError() != hipSuccess ||
                              hipMemcpyAsync(h_err, err, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                              hipStreamSynchronize(st) != hipSuccess))
            rc = MNRT_ERR_HIP;
        if (rc == MNRT_OK && *h_err) rc = MNRT_ERR_CAPACITY;
    }
    (void)pool_free(buf, st);
    return rc;
}

extern "C" int mnrt_preorder_set_custom_bits(uint32_t* d_tree, const uint32_t* d_node_addr, uint32_t num_nodes,
                                             uint32_t bit, const uint32_t* d_values, void* stream) {
    if (bit > 1 || (num_nodes > 0 && (!d_tree || !d_node_addr || !d_values))) return MNRT_ERR_ARG;
    if (num_nodes == 0) return MNRT_OK;
    hipLaunchKernelGGL(mnrt::k_custom_bits, dim3((num_nodes + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       d_tree, d_node_addr, num_nodes, bit, d_values);
    MNRT_HIP_CHECK(hipGetLastError());
    return MNRT_OK;
}
