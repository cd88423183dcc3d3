/*
 * mnrt_kd.h — C ABI of the MI355X (gfx950) kd-tree hot path: SAH kd-tree build for
 * triangles and photons, closest-hit ray traversal, and the exact-kNN photon-map
 * radiance estimate.
 *
 * Every entry point is plain C: device pointers, sizes, and an opaque HIP stream
 * (hipStream_t passed as void*, NULL = the null stream).  Every entry point returns an
 * int status (MNRT_OK = 0, negative = error; mnrt_status_string() names it).  The caller
 * owns inputs and outputs; a tree owns the arrays the library allocated for it until
 * mnrt_tree_release().  Nothing here holds global per-scene state: two trees can be
 * queried concurrently on two streams (the reference binds one tree into __constant__
 * memory and texture references, raytracing.cu:125-161, photon_knn.cu:65-81).
 *
 * Vectors are float4 (16 B) arrays exactly as the reference's TriangleData.d_verts[3]
 * (KernelDefs.h:506), RayChunk.d_origins/d_dirs (RayPool.h:162-171) and
 * PhotonData.d_positions/d_powers (KernelDefs.h:537-630); the w component of inputs is
 * ignored unless stated.
 *
 * ---------------------------------------------------------------------------------
 * Tree layout in HBM (replaces KDTreeData::d_preorderTree, KDKernelDefs.h:76-194):
 *   nodes  : uint32 pairs in preorder; the left child of inner node i is node i+1.
 *     inner: w0 = IEEE bits of the split position, w1 = axis<<30 | right child index
 *     leaf : w0 = element count,                   w1 = 3<<30   | first leaf reference
 *     (30-bit indices: up to 2^30 nodes / references; the reference masks 24 bits when
 *      writing, kdtree.cu:1742, and 28 when reading, raytracing.cu:564.)
 *   refs   : uint32 original element index per leaf reference, leaves in preorder,
 *            each leaf's references ascending.
 *   recs   : leaf-ordered element records, one per reference, so a leaf's elements are
 *            contiguous in memory:
 *     triangle (3 x float4): [v0.xyz, index bits] [e1 = v1-v0, 0] [e2 = v2-v0, 0]
 *     photon   (2 x float4): [pos.xyz, index bits] [power.rgb, oct-encoded direction bits]
 * ---------------------------------------------------------------------------------
 */
#ifndef MNRT_KD_H
#define MNRT_KD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Raised whenever a declaration or struct layout here changes (2: mnrt_batch_io gained the
 * adjoint buffers). */
#define MNRT_ABI_VERSION 5

enum mnrt_status_code {
    MNRT_OK = 0,
    MNRT_ERR_ARG = -1,      /* bad argument (null pointer, size, k out of range) */
    MNRT_ERR_HIP = -2,      /* a HIP runtime call failed */
    MNRT_ERR_NOMEM = -3,    /* device allocation failed */
    MNRT_ERR_NODEV = -4,    /* no usable gfx950 device */
    MNRT_ERR_CAPACITY = -5, /* a size exceeds the layout's index width */
    MNRT_ERR_STATE = -6     /* tree lacks what the call needs (records, kind) */
};

/* Tree kinds. */
#define MNRT_KIND_TRI 0u
#define MNRT_KIND_POINT 1u

/* Packed node word helpers. */
#define MNRT_LEAF_TAG 3u
#define MNRT_INDEX_MASK 0x3FFFFFFFu

/* Deepest node level that may still be split (levels start at 0 at the root).  The
 * reference caps traversal stacks at KD_MAX_HEIGHT 50 (KDKernelDefs.h:61) but never
 * bounds the build (KDTreeGPU.cpp:254-281); here the build stops splitting at this
 * level so no traversal can exceed it. */
#define MNRT_MAX_LEVEL 48u

/* Largest kNN k supported by mnrt_knn_gather (the reference default k is 100 for the
 * global map and 200 for caustics, MNRTSettings.h:431,437). */
#define MNRT_KNN_MAX_K 256u

/* Build parameters: KDTreeGPU::SetEmptySpaceRatio/SetSmallNodeMax/SetMaxQueryRadius
 * (KDTreeGPU.h:248-280).  Defaults: triangles 0.25 / 64 (KDTreeTri.cpp:46-54),
 * points 0.1 / 32 (KDTreePoint.cpp:120-134). */
typedef struct mnrt_build_params {
    float empty_space_ratio;
    uint32_t small_node_max; /* 1..64 (ElementMask is 64 bits, KDKernelDefs.h:55) */
    float max_query_radius;  /* VVH cost extension for point trees (kdtree.cu:1370-1389) */
    uint32_t has_root_box;   /* 1: use root_min/root_max; 0: tight bounds of the input */
    float root_min[3];
    float root_max[3];
} mnrt_build_params;

/* A device-resident tree.  Filled by the builders; a caller may also fill nodes/refs
 * itself (owns = 0) and attach records with mnrt_tree_attach_*_records(). */
typedef struct mnrt_tree {
    uint32_t kind;       /* MNRT_KIND_TRI or MNRT_KIND_POINT */
    uint32_t num_nodes;
    uint32_t num_refs;
    uint32_t num_elems;  /* input element count */
    uint32_t num_leaves;
    uint32_t max_level;  /* deepest node level present */
    float root_min[3];
    float root_max[3];
    uint32_t* d_nodes;   /* 2 * num_nodes words */
    uint32_t* d_refs;    /* num_refs */
    float* d_recs;       /* num_refs * (12 | 8) floats, see header comment */
    uint32_t owns_nodes; /* library allocated d_nodes/d_refs */
    uint32_t owns_recs;  /* library allocated d_recs */
    uint32_t* d_qnodes;  /* query-time node words (NULL: d_nodes), see mnrt_tree_collapse */
    uint32_t qleaf_max;  /* subtrees with at most this many references are leaves in d_qnodes */
    uint32_t owns_qnodes;
} mnrt_tree;

const char* mnrt_status_string(int status);
int mnrt_abi_version(void);

/* Selects the device and checks that it is gfx950. */
int mnrt_device_init(int device);

/* --- build ------------------------------------------------------------------------ */

/* KDTreeTri(const TriangleData&) + KDTreeGPU::BuildTree() (KDTreeTri.cpp:46-91,
 * KDTreeGPU.cpp:175-209): triangle vertices as three float4 arrays of n entries. */
int mnrt_build_tri(const float* d_v0, const float* d_v1, const float* d_v2, uint32_t n,
                   const mnrt_build_params* params, mnrt_tree* out, void* stream);

/* KDTreePoint(d_points, n, min, max, maxQueryRadius) + BuildTree()
 * (KDTreePoint.cpp:120-134): photon positions as a float4 array.  Photon records are
 * attached separately because the reference builds the tree on positions only. */
int mnrt_build_points(const float* d_pos4, uint32_t n, const mnrt_build_params* params,
                      mnrt_tree* out, void* stream);

/* Batched builds (no reference counterpart; the reference builds one KDTreeGPU per
 * scene, KDTreeGPU.cpp:175-209): num_scenes scenes whose elements are concatenated in
 * one input, scene s holding entries h_offsets[s] .. h_offsets[s+1]-1 (host array of
 * num_scenes + 1).  All scenes go through one level loop, one small-node launch and one
 * layout pass; out[s] (num_scenes trees) equals what mnrt_build_tri / mnrt_build_points
 * gives for scene s alone, word for word, with scene-local element indices.  The trees
 * are views into shared arrays owned by out[0]: mnrt_tree_release(&out[0]) frees the
 * batch, release of out[s > 0] frees nothing.  h_root_boxes: NULL (tight bounds) or
 * 6 floats (min xyz, max xyz) per scene; h_max_query_radius: NULL (params value) or one
 * VVH radius per scene.  params->has_root_box is ignored. */
int mnrt_build_tri_batch(const float* d_v0, const float* d_v1, const float* d_v2,
                         const uint32_t* h_offsets, uint32_t num_scenes,
                         const mnrt_build_params* params, const float* h_root_boxes,
                         mnrt_tree* out, void* stream);
int mnrt_build_points_batch(const float* d_pos4, const uint32_t* h_offsets, uint32_t num_scenes,
                            const mnrt_build_params* params, const float* h_root_boxes,
                            const float* h_max_query_radius, mnrt_tree* out, void* stream);
/* Records of a batch (trees from one batched build, inputs concatenated as for the
 * build): one shared record array owned by trees[0]; photon batches are collapsed with
 * max_refs = 512 like mnrt_tree_attach_photon_records.  Called again on photon trees whose
 * records it attached, it rewrites the records in place (new photon powers, say) and
 * keeps the arrays, so batches made from these trees stay valid. */
int mnrt_tree_attach_tri_records_batch(mnrt_tree* trees, uint32_t num_scenes, const float* d_v0,
                                       const float* d_v1, const float* d_v2,
                                       const uint32_t* h_offsets, void* stream);
int mnrt_tree_attach_photon_records_batch(mnrt_tree* trees, uint32_t num_scenes,
                                          const float* d_pos4, const float* d_pow4,
                                          const float* d_dir4, const uint32_t* h_offsets,
                                          void* stream);

/* Device-driven batched build recorded into a hipGraph (no reference counterpart: the
 * reference's level loop reads its counters on the host every level, KDTreeGPU.cpp:247-281).
 * mnrt_build_graph_create builds the batch once as mnrt_build_*_batch does, sizes fixed
 * buffers from that build (each quantity the level loop grows a buffer for, times
 * 1 + margin), and records the whole build into a graph: every level (num_levels of the
 * first build + 2), the small-node stage, layout, emission, the records (photon trees
 * when d_power4 / d_dir4 are given) and, for photon trees, the query-time node words.
 * Counts stay on the device; there is no host synchronisation and no allocation in it.
 * It then replays the graph once and fills out[num_scenes] with the trees, which equal
 * mnrt_build_*_batch's word for word.  The trees are views into arrays the graph owns
 * (owns_* = 0; mnrt_build_graph_destroy frees them).
 * mnrt_build_graph_replay rebuilds the trees in place from the input addresses given at
 * creation (their contents may change; the element counts are fixed).  The trees move
 * only when their sizes change: mnrt_build_graph_status waits for the stream, returns in
 * *h_flags 0 when the last replay fit its buffers (and then refreshes out[], if given,
 * with its sizes, root boxes and addresses), nonzero when it did not (the trees are then
 * invalid until a replay that fits; create a new graph for the new sizes).  Bits: 64 a
 * level outgrew the level buffers, 128 more levels than recorded, 256 the trees outgrew
 * the output arrays, 512 the large-node stage went more than 6 levels deeper than the
 * first build (the recorded layout's depth); 1..32 as the host-driven build's errors. */
typedef struct mnrt_build_graph mnrt_build_graph;
typedef struct mnrt_build_graph_desc {
    int kind;                         /* MNRT_KIND_TRI or MNRT_KIND_POINT */
    uint32_t num_scenes;
    const uint32_t* h_offsets;        /* num_scenes + 1 element offsets (host) */
    const float* d_v0;                /* triangles: float4 vertices; points: float4 positions */
    const float* d_v1;
    const float* d_v2;
    const float* d_power4;            /* points: photon powers and directions for the */
    const float* d_dir4;              /* records (both NULL: no records) */
    const mnrt_build_params* params;
    const float* h_root_boxes;        /* NULL or 6 floats per scene, as mnrt_build_*_batch */
    const float* h_max_query_radius;  /* NULL or one per scene */
    float margin;                     /* buffer room over the first build's sizes (0..16) */
} mnrt_build_graph_desc;
int mnrt_build_graph_create(const mnrt_build_graph_desc* desc, mnrt_tree* out, mnrt_build_graph** graph,
                            void* stream);
int mnrt_build_graph_replay(mnrt_build_graph* graph, void* stream);
int mnrt_build_graph_status(mnrt_build_graph* graph, mnrt_tree* out, uint32_t* h_flags, void* stream);
int mnrt_build_graph_destroy(mnrt_build_graph* graph);

/* Leaf-ordered records for a tree whose nodes/refs exist. */
int mnrt_tree_attach_tri_records(mnrt_tree* tree, const float* d_v0, const float* d_v1,
                                 const float* d_v2, void* stream);
/* d_dir4: incoming (travel) direction of each photon; the reference stores it as
 * spherical angles in the w components (photon_dev.h:71-116). */
int mnrt_tree_attach_photon_records(mnrt_tree* tree, const float* d_pos4, const float* d_pow4,
                                    const float* d_dir4, void* stream);

/* Query-time node words: every inner node whose subtree holds at most max_refs leaf
 * references becomes a leaf over that contiguous reference range (a subtree's leaves are
 * consecutive in preorder).  A wave then tests such a subtree's elements in one step.
 * The build's node words (d_nodes) are unchanged; queries read d_qnodes.  Query results
 * do not depend on it.  mnrt_tree_attach_photon_records applies max_refs = 512. */
int mnrt_tree_collapse(mnrt_tree* tree, uint32_t max_refs, void* stream);

/* Reference KDTreeData layout of a tree (KDTreeData, KDKernelDefs.h:76-194, as
 * kernel_TraversalDownPath writes it, kdtree.cu:1707-1770), for consumers that read the
 * reference's preorder stream.  Node indices are preorder indices.  Per node (num_nodes
 * entries): element count (a leaf's count; an inner node's leaf references below it),
 * word address in the stream, extent (centre xyz, half diagonal) of the node's spatial
 * box, left / right child (0 / 0 for leaves).  The stream:
 *   inner: [node] [axis << 30 | custom bits << 28 | address of the right child] [split bits]
 *   leaf : [node | 0x80000000] [count] [element indices]
 * The right-child address keeps 28 bits (the reference masks 24, kdtree.cu:1742); a stream
 * of 2^28 words or more is refused (MNRT_ERR_CAPACITY).  mnrt_tree_preorder_size gives
 * the stream length (3 per inner node + 2 per leaf + references). */
int mnrt_tree_preorder_size(const mnrt_tree* tree, uint64_t* out_words);
int mnrt_tree_export_preorder(const mnrt_tree* tree, uint32_t* d_tree, uint64_t capacity_words,
                              uint32_t* d_num_elems, uint32_t* d_node_addr, float* d_node_extent4,
                              uint32_t* d_child_left, uint32_t* d_child_right, void* stream);
/* KDTreeGPU::SetCustomBits / kernel_SetCustomBit (KDTreeGPU.cpp:1373-1380,
 * kdtree.cu:1879-1908): bit 0 or 1 of every inner node's parent word in an exported
 * stream = d_values[node] & 1; leaves are left alone. */
int mnrt_preorder_set_custom_bits(uint32_t* d_tree, const uint32_t* d_node_addr, uint32_t num_nodes,
                                  uint32_t bit, const uint32_t* d_values, void* stream);

/* KDTreeGPU::Destroy() (KDTreeGPU.cpp:237-245): frees what the library allocated. */
int mnrt_tree_release(mnrt_tree* tree);

/* Tree SAH cost (triangles) or VVH cost (points) of a device tree, computed on the
 * device: sum_inner 3*A(n)/A(root) + sum_leaf count*A(n)/A(root) (kdtree.cu:51,1418-1436). */
int mnrt_tree_cost(const mnrt_tree* tree, float max_query_radius, double* out_cost,
                   void* stream);

/* --- traversal ----------------------------------------------------------------------- */

/* KernelRTTraceRays(const RayChunk&, ShadingPoints&, uint* d_outIsValid)
 * (raytracing.cu:854-879, 1617-1635): closest hit for n rays.  eps is the ray epsilon
 * (SceneConfig.cpp:111); hits need lambda > eps.  Outputs: triangle index or -1,
 * lambda, barycentrics (float2), optional hit point o + d*lambda (float4), optional
 * valid flag.  Equal lambdas resolve to the lower triangle index. */
int mnrt_trace_closest(const mnrt_tree* tree, uint32_t n, const float* d_orig4,
                       const float* d_dir4, float eps, int32_t* d_out_tri, float* d_out_t,
                       float* d_out_bary2, float* d_out_pt4, uint32_t* d_out_valid,
                       void* stream);

/* dev_FindNextIntersectionKDWhileWhile<false> (raytracing.cu:503-645): any-hit test of
 * n ray segments.  d_out_hit[i] = 1 iff some triangle is hit at lambda with
 * eps < lambda < min(t_exit, tmax_i) + eps and lambda < tmax_i, where t_exit is where the
 * ray leaves the root box and tmax_i = d_tmax[i] (NULL: no limit).  The result does not
 * depend on which triangle is met first. */
int mnrt_trace_any(const mnrt_tree* tree, uint32_t n, const float* d_orig4, const float* d_dir4,
                   const float* d_tmax, float eps, uint32_t* d_out_hit, void* stream);

/* LightData / LightType (KernelDefs.h:89-101, 307-323). */
enum {
    MNRT_LIGHT_POINT = 0,
    MNRT_LIGHT_DIRECTIONAL = 1,
    MNRT_LIGHT_AREA_DISC = 2,
    MNRT_LIGHT_AREA_RECT = 3
};
typedef struct mnrt_light {
    uint32_t type;         /* MNRT_LIGHT_* */
    float position[3];     /* point / disc centre / rectangle corner */
    float direction[3];    /* directional: travel direction; area: surface normal */
    float L_emit[3];       /* emitted radiance (intensity for point lights) */
    float area_v1[3];      /* rectangle edges */
    float area_v2[3];
    float area_radius;     /* disc radius */
} mnrt_light;

/* kernel_SampleAreaLight (raytracing.cu:963-1000): one point on the area light per
 * entry, with the stratified uniforms (idx_x + rnd1[i]) * inv_x, (idx_y + rnd2[i]) * inv_y:
 * rectangle: position + u*v1 + v*v2; disc: concentric disc sample (Shirley-Chiu) in the
 * plane through position with normal direction (basis of dev_BuildCoordSystem,
 * bsdf_dev.h:158-177).  Output float4 (w = 0). */
int mnrt_sample_area_light(const mnrt_light* light, uint32_t n, float idx_x, float inv_x, float idx_y,
                           float inv_y, const float* d_rnd1, const float* d_rnd2, float* d_out_pts4,
                           void* stream);

/* kernel_TraceShadowRaysDelta / kernel_TraceShadowRaysArea (raytracing.cu:1019-1114):
 * d_out_unoccluded[i] = 1 iff the light is visible from shading point d_pts4[i].
 *   point light: segment to position, occluded by a hit with lambda < distance;
 *   directional: direction -light.direction, unlimited;
 *   area lights: segment to d_samples4[i] (mnrt_sample_area_light); 0 when the sample
 *     faces away (dot(-dir, light.direction) <= 0), else occluded by a hit with
 *     lambda + eps < distance.
 * A point at zero distance from its light sample is unoccluded. */
int mnrt_trace_shadow(const mnrt_tree* tree, const mnrt_light* light, uint32_t n, const float* d_pts4,
                      const float* d_samples4, float eps, uint32_t* d_out_unoccluded, void* stream);

/* --- photon-map kNN radiance --------------------------------------------------------- */

/* KernelPMPhotonGather (photon_knn.cu:487-552, 833-844) with the range search of
 * dev_RangeSearch (photon_knn.cu:113-235) replaced by the exact k nearest photons
 * (ties by photon index) among those closer than the query's radius cap:
 *   E = sum_{front-facing of the set} power / (pi * r2),  r2 = d_k^2 when k were found,
 *   else cap^2;  E = 0 when fewer than 8 were found and d_clr4 is given;
 *   radiance[slot] += alpha * rho * (1/pi) * E * influence.
 * Queries: position, shading normal, source ray direction (normal flipped toward it),
 * influence (float4 each, n entries).  d_clr4: diffuse colour with alpha in w, or NULL
 * (photon visualisation: colour 1, no 8-photon cut).  d_idx_sp: output slot per query
 * or NULL (slot = query); several queries may share a slot (final gathering), and
 * queries of one slot that are consecutive are summed in query order.
 * d_query_radius: per-query cap (the reference's d_queryRadius) or NULL (r_max for all).
 * d_io_inv_dist (with d_orig4, the source ray origins): slot += 1 / |position - origin|
 * (the reference's d_ioInvDistAccu), or NULL.  Optional outputs per query: d_out_r2
 * (the r2 used), d_out_irr4 (E), d_out_nbr (n*k photon indices sorted by (d2, index),
 * 0xFFFFFFFF past the count), d_out_count (neighbours found).  Photon trees of 2^29 nodes
 * or more are refused (MNRT_ERR_CAPACITY). */
int mnrt_knn_gather(const mnrt_tree* pmap, uint32_t k, float r_max, uint32_t n,
                    const float* d_pos4, const float* d_nrm4, const float* d_dir4,
                    const float* d_infl4, const float* d_clr4, const uint32_t* d_idx_sp,
                    const float* d_query_radius, const float* d_orig4, float* d_io_rad4,
                    float* d_io_inv_dist, float* d_out_r2, float* d_out_irr4, uint32_t* d_out_nbr,
                    uint32_t* d_out_count, void* stream);

/* mnrt_knn_gather for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation); no counterpart in the
 * reference.  Same arguments, results bit for bit equal to mnrt_knn_gather's, but no
 * hipMalloc / pool allocation: the query processing order and, with d_idx_sp, the
 * per-query values live in one fixed 64 MiB device array of the library (one per device,
 * kept until the process ends), so n is limited to 7,339,897 queries, or 2,097,113 with
 * d_idx_sp (MNRT_ERR_CAPACITY beyond, nothing launched).  Every call on one device shares
 * that array: calls that overlap on a device (concurrent streams or host threads) race and
 * must be serialised by the caller.  mnrt_device_init looks the array's address up for the
 * device; a call on a device it was not run for looks it up first. */
int mnrt_knn_gather_noalloc(const mnrt_tree* pmap, uint32_t k, float r_max, uint32_t n,
                            const float* d_pos4, const float* d_nrm4, const float* d_dir4,
                            const float* d_infl4, const float* d_clr4, const uint32_t* d_idx_sp,
                            const float* d_query_radius, const float* d_orig4, float* d_io_rad4,
                            float* d_io_inv_dist, float* d_out_r2, float* d_out_irr4,
                            uint32_t* d_out_nbr, uint32_t* d_out_count, void* stream);

/* Adjoint of mnrt_knn_gather (no reference counterpart): given dLoss/dradiance per query
 * (float4, w ignored) and the forward's r2 and neighbour counts (d_out_r2, d_out_count),
 * recomputes every query's neighbour set (identical to the forward's: same r_max and
 * d_query_radius) and writes
 *   d_grad_pow4[p] = dLoss/dpower of photon p (original photon index, num_elems entries),
 *   d_grad_clr4[q] = dLoss/d(diffuse colour rgb) of query q (optional).
 * Per-photon sums use 64-bit fixed-point atomic adds (scale a power of two fixed from the
 * largest weight), so results do not depend on scheduling.  d_idx_sp slots are not
 * supported here (radiance per query). */
int mnrt_knn_gather_bwd(const mnrt_tree* pmap, uint32_t k, float r_max, uint32_t n,
                        const float* d_pos4, const float* d_nrm4, const float* d_dir4,
                        const float* d_infl4, const float* d_clr4, const float* d_query_radius,
                        const float* d_grad_rad4,
                        const float* d_r2, const uint32_t* d_count, float* d_grad_pow4,
                        float* d_grad_clr4, void* stream);

/* KDTreePoint::ComputeQueryRadii (KDTreePoint.cpp:229-248): radius to the k-th nearest
 * photon, capped at r_max (r_max when fewer than k are closer). */
int mnrt_knn_radii(const mnrt_tree* pmap, uint32_t k, float r_max, uint32_t n,
                   const float* d_pos4, float* d_out_radius, void* stream);

/* --- photon tracing ------------------------------------------------------------------ */

/* RTCore::RebuildPhotonMaps for the global map (RTCore.cpp:454-638; emission
 * photon_build.cu:79, tracing raytracing.cu:1146, Russian roulette photon_build.cu:220,
 * scattering photon_build.cu:278): traces photon paths from quad lights through the
 * triangle tree until `target` photons are stored at diffuse hits (at most
 * max_bounces + 1 per path, Russian roulette with p = 0.5 past bounce index 3), then
 * scales their power by 1 / paths shot.
 *   d_albedo4: per-triangle diffuse RGB (0 = absorbing, e.g. the light itself).
 *   d_lights: num_lights records of 4 float4: corner, edge u, edge v, emitted power RGB;
 *             path p starts on light p mod num_lights with power * num_lights.
 * Outputs (target entries): position, power, incoming direction (float4 each); optional
 * meta (light | bounce << 8), source path index, and the 4 triangles the path scattered
 * from before the photon (0xFFFFFFFF padded) for reweighting by emission and albedo
 * (d_out_path_tris needs max_bounces <= 4, else MNRT_ERR_ARG: a photon stored at bounce
 * b has scattered from b triangles).
 * *out_count = photons stored, *out_paths = paths shot. */
int mnrt_photons_generate(const mnrt_tree* tree, const float* d_v0, const float* d_v1,
                          const float* d_v2, const float* d_albedo4, const float* d_lights,
                          uint32_t num_lights, uint32_t target, uint32_t max_bounces, uint32_t seed,
                          float eps, float* d_out_pos4, float* d_out_pow4, float* d_out_dir4,
                          uint32_t* d_out_meta, uint32_t* d_out_path, uint32_t* d_out_path_tris,
                          uint32_t* out_count, uint32_t* out_paths, void* stream);

/* PhotonData in the reference layout (KernelDefs.h:537-630, photon_dev.h:71-116) keeps the
 * incoming direction as spherical angles: azimuth phi in position.w, polar theta in
 * power.w.  These convert between that and direction vectors (dev_Spherical2Direction /
 * dev_Direction2Spherical, mncudautil_dev.h:345-375, with HIP's sinf/cosf/atan2f/acosf). */
int mnrt_photons_spherical_to_dir(uint32_t n, const float* d_pos4, const float* d_pow4, float* d_out_dir4,
                                  void* stream);
int mnrt_photons_dir_to_spherical(uint32_t n, const float* d_dir4, float* d_io_pos4, float* d_io_pow4,
                                  void* stream);
/* PhotonData::Merge (KernelDefs.cpp:363-384): appends the entries i with d_valid[i] != 0,
 * in order, to d_dst_* (pointers at the first free entry); *out_added = count.  If more
 * than dst_capacity are valid nothing is written and MNRT_ERR_CAPACITY is returned (the
 * reference stops with MNFatal).  Synchronous. */
int mnrt_photons_append_valid(uint32_t n, const uint32_t* d_valid, const float* d_src_pos4,
                              const float* d_src_pow4, float* d_dst_pos4, float* d_dst_pow4,
                              uint32_t dst_capacity, uint32_t* out_added, void* stream);
/* PhotonMap::ScaleFlux / KernelPMScaleFlux (PhotonMap.cpp:186-193): power.rgb *= scale,
 * w (the polar angle) untouched. */
int mnrt_photons_scale_flux(uint32_t n, float* d_io_pow4, float scale, void* stream);

/* Photon powers and hit colours from per-light emission and per-triangle albedo, with the
 * photon paths and hit triangles held fixed (no counterpart in the reference, which never
 * fits parameters; the forward restates how RebuildPhotonMaps forms a stored photon's
 * power along its path: emission of its light, photon_build.cu:79-130, times the albedo of
 * every triangle it scattered from, photon_build.cu:278-331, RTCore.cpp:556-581):
 *   power[p].rgb = ((base[p] * emission[light[p]]) * albedo[t0]) * ... (per channel, over
 *                  the path's triangles t0..t3 of d_path_tris that are not 0xFFFFFFFF);
 *   clr[q].rgb   = albedo[hit_tri[q]] for hit_tri[q] != 0xFFFFFFFF (clr[q].w kept).
 * d_emission3 / d_albedo3: 3 floats per light / triangle.  An index out of range gives a
 * NaN power or colour (never an out-of-range read). */
int mnrt_photons_reweight(uint32_t np, const float* d_base_pow4, const uint32_t* d_light,
                          const uint32_t* d_path_tris, uint32_t num_lights, const float* d_emission3,
                          uint32_t num_tris, const float* d_albedo3, uint32_t nq,
                          const uint32_t* d_hit_tri, float* d_out_pow4, float* d_io_clr4, void* stream);

/* Adjoint of mnrt_photons_reweight: dLoss/demission (3 per light) and dLoss/dalbedo (3 per
 * triangle) from dLoss/dpower (d_grad_pow4) and dLoss/dclr (d_grad_clr4, nq entries).
 * Each output is a sum of 64-bit fixed-point terms in a power-of-two unit fixed from the
 * largest term and the term count, so the result is the same bits on every run; terms
 * below about 2^-38 of the largest term are lost.  A non-finite term makes every output
 * NaN.  Photons with an index out of range contribute nothing. */
int mnrt_photons_reweight_bwd(uint32_t np, const float* d_base_pow4, const uint32_t* d_light,
                              const uint32_t* d_path_tris, uint32_t num_lights, const float* d_emission3,
                              uint32_t num_tris, const float* d_albedo3, uint32_t nq,
                              const uint32_t* d_hit_tri, const float* d_grad_pow4,
                              const float* d_grad_clr4, float* d_grad_emission3, float* d_grad_albedo3,
                              void* stream);

/* --- parallel primitives (32-bit data) -------------------------------------------------- */

/* MNCudaPrimitives::Scan (MNCudaPrimitives.h:167-184, MNCudaPrimitives.cpp:160-177): add-scan
 * of n uint32 words, exclusive (inclusive = 0) or inclusive; d_out may equal d_in;
 * *d_total (device word, may be null) = the sum of all n (mod 2^32). */
int mnrt_scan_u32(const uint32_t* d_in, uint32_t n, int inclusive, uint32_t* d_out, uint32_t* d_total,
                  void* stream);

/* Exclusive add-scan of a length that lives on the device (no reference counterpart: the
 * reference's scans take a host count, MNCudaPrimitives.cpp:160-177): n = min(*d_count,
 * capacity), read by the kernels, so the call can be recorded into a graph and replayed
 * after the count changes; d_out[0..n) = the scan of d_in[0..n) (d_out may equal d_in),
 * *d_total (device word, may be null) = the sum, and d_out[n] = the sum too when end_total
 * is nonzero (d_out then holds capacity + 1 words).  The recorded builds scan their
 * per-level counts with it. */
int mnrt_scan_u32_counted(const uint32_t* d_in, const uint32_t* d_count, uint32_t capacity,
                          uint32_t* d_out, uint32_t* d_total, int end_total, void* stream);

/* MNCudaPrimitives::Compact (MNCudaPrimitives.h:187-211, MNCudaPrimitives.cpp:180-190):
 * the words d_in[i] with d_is_valid[i] != 0, in input order, to d_out; *d_out_count (device
 * word) = their number.  Not in place (d_out != d_in). */
int mnrt_compact_u32(const uint32_t* d_in, const uint32_t* d_is_valid, uint32_t n, uint32_t* d_out,
                     uint32_t* d_out_count, void* stream);

/* --- hits to queries -------------------------------------------------------------------- */

/* Compaction of valid hits (RTCore::TraceRays, RTCore.cpp:416-429) into shading points:
 * position o + d*t, unit geometric normal (KernelBSDFGetNormalsAtHit, bsdf.cu:317-367),
 * source ray direction, per-triangle diffuse colour (alpha in w) and ray influence
 * (d_infl4 per ray, or NULL for 1).  Outputs hold up to n entries; *out_count = hits.
 * Optional: source ray index and hit triangle per shading point. */
int mnrt_shading_points(uint32_t n, const float* d_orig4, const float* d_dir4, const int32_t* d_hit_tri,
                        const float* d_hit_t, const float* d_v0, const float* d_v1, const float* d_v2,
                        const float* d_albedo4, const float* d_infl4, float* d_out_pos4, float* d_out_nrm4,
                        float* d_out_dir4, float* d_out_clr4, float* d_out_infl4, uint32_t* d_out_src,
                        int32_t* d_out_tri, uint32_t* out_count, void* stream);

/* Pinhole primary rays through pixel centres (kernel_genPrimaryRays, raygen.cu:171-240)
 * in tile order (tile x tile pixel tiles row by row, pixels row by row inside a tile,
 * edge tiles clipped): width * height rays.  h_cam (host, 14 doubles): eye xyz, unit
 * forward xyz, unit right xyz, up xyz, tan(vfov / 2), aspect (width / height).  The
 * direction is computed in double and rounded once to float.  Optional pixel index
 * py * width + px per ray. */
int mnrt_primary_rays(uint32_t width, uint32_t height, uint32_t tile, const double* h_cam,
                      float* d_out_orig4, float* d_out_dir4, uint32_t* d_out_pixel, void* stream);

/* One cosine-distributed diffuse bounce ray per shading point (kernel_genFinalGatherRays,
 * raygen.cu:406-487) about the normal turned toward the arriving ray; optional influence
 * output = the point's diffuse colour. */
int mnrt_bounce_rays(uint32_t n, const float* d_pos4, const float* d_nrm4, const float* d_dir4,
                     const float* d_clr4, uint32_t seed, float* d_out_orig4, float* d_out_dir4,
                     float* d_out_infl4, void* stream);

/* --- batches of scenes (no reference counterpart) ----------------------------------- */

/* A batch is many independent scenes traced and queried together: the reference runs
 * KernelRTTraceRays / KernelPMPhotonGather once per scene with the scene bound into
 * global device state (RTCore.cpp:404-439, PhotonMap.cpp:195-232).  Scene s owns ray and
 * query slots ray_offsets[s] .. ray_offsets[s+1]-1 of every per-ray / per-query array
 * and triangles tri_offsets[s] .. of the concatenated vertex and albedo arrays.  Each
 * phase is one persistent launch (workgroups take tiles from device work counters) and
 * enqueues kernels only, with no allocation and no host synchronisation, so a phase can
 * be captured in a hipGraph (mnrt_batch_capture) and replayed.  Per scene, results are
 * identical to mnrt_trace_closest / mnrt_shading_points / mnrt_knn_gather on that scene. */
typedef struct mnrt_batch mnrt_batch;

typedef struct mnrt_batch_desc {
    uint32_t num_scenes;
    const mnrt_tree* tri_trees;    /* num_scenes triangle trees with records */
    const mnrt_tree* photon_maps;  /* num_scenes point trees with photon records, or NULL */
    const uint32_t* tri_offsets;   /* host, num_scenes + 1 */
    const uint32_t* ray_offsets;   /* host, num_scenes + 1 */
    const float* eps;              /* host, ray epsilon per scene */
    const float* r_max;            /* host, kNN radius cap per scene (photon_maps only) */
    uint32_t k;                    /* neighbours per query, 1..MNRT_KNN_MAX_K */
} mnrt_batch_desc;

/* Launch shape of the phases; mnrt_batch_autotune picks it by timing candidates. */
typedef struct mnrt_batch_tuning {
    uint32_t trace_block; /* lanes per workgroup: 64, 128, 256 */
    uint32_t trace_tile;  /* rays per work item, a multiple of trace_block */
    uint32_t trace_grid;  /* persistent workgroups */
    uint32_t trace_stack; /* traversal stack entries: 16, 24, 32, 50; >= every tree's max_level */
    uint32_t knn_waves;   /* waves (one query each) per workgroup: 1, 2, 4 */
    uint32_t knn_chunk;   /* query slots per work item */
    uint32_t knn_grid;    /* persistent workgroups */
    uint32_t knn_cap;     /* LDS candidate entries per wave: 128, 256, 512; >= k + 64 */
    uint32_t parts;       /* work-counter shares: 1, or 8 (one per XCD) */
    uint32_t knn_select;  /* candidate selection: 0 bitonic sort, 1 radix select of the k-th key,
                             2 threshold window, exact selection at the end */
} mnrt_batch_tuning;

/* Buffers of the phases (slot-indexed unless stated). */
typedef struct mnrt_batch_io {
    const float* d_orig4;
    const float* d_dir4;
    const float* d_v0;
    const float* d_v1;
    const float* d_v2;
    const float* d_albedo4; /* per triangle */
    const float* d_infl4;   /* per ray, or NULL (1) */
    int32_t* d_hit_tri;     /* trace out: scene-local triangle or -1 */
    float* d_hit_t;
    float* d_hit_bary2;     /* optional */
    float* d_sp_pos4;       /* shade out: scene s's shading points from slot ray_offsets[s] */
    float* d_sp_nrm4;
    float* d_sp_dir4;
    float* d_sp_clr4;
    float* d_sp_infl4;
    uint32_t* d_sp_src;     /* optional: scene-local source ray */
    uint32_t* d_sp_count;   /* num_scenes: shading points per scene */
    float* d_io_rad4;       /* kNN: radiance accumulated per shading point */
    float* d_out_r2;        /* optional */
    uint32_t* d_out_nbr;    /* optional, k per slot */
    uint32_t* d_out_count;  /* optional for MNRT_PHASE_KNN; MNRT_PHASE_KNN_BWD reads it */
    /* kNN adjoint (MNRT_PHASE_KNN_BWD), the batched mnrt_knn_gather_bwd: */
    const float* d_grad_rad4; /* dLoss/dradiance per slot (float4, w ignored) */
    float* d_grad_pow4;       /* out: dLoss/dpower per photon of the concatenation of the
                                 scenes' photon arrays (scene s from the sum of the photon
                                 counts of scenes 0..s-1, original photon order) */
    float* d_grad_clr4;       /* optional out: dLoss/d(diffuse colour) per slot */
} mnrt_batch_io;

#define MNRT_PHASE_TRACE 1u
#define MNRT_PHASE_SHADE 2u
#define MNRT_PHASE_KNN 4u
/* Adjoint of MNRT_PHASE_KNN for every scene (reads the forward's d_out_r2 / d_out_count and
 * the shading points; recomputes the neighbour sets).  Per scene, d_grad_pow4 and
 * d_grad_clr4 equal mnrt_knn_gather_bwd on that scene's d_sp_count[s] shading points bit for
 * bit: each scene keeps its own fixed-point unit, fixed on the device from its own largest
 * weight and query count, so the phase needs no host synchronisation and can be captured.
 * The first run (or capture) allocates the batch's accumulators (24 bytes per photon). */
#define MNRT_PHASE_KNN_BWD 8u

/* Copies the per-scene tables to the device (the trees must outlive the batch). */
int mnrt_batch_create(const mnrt_batch_desc* desc, mnrt_batch** out);
int mnrt_batch_destroy(mnrt_batch* batch);
int mnrt_batch_get_tuning(const mnrt_batch* batch, mnrt_batch_tuning* out);
/* Synchronous (re-uploads the work tables); drops captured graphs. */
int mnrt_batch_set_tuning(mnrt_batch* batch, const mnrt_batch_tuning* tuning);
/* Times candidate launch shapes of the trace and kNN phases on io (a prefix of at most
 * max_slots slots; 0 = all) and keeps the fastest.  Writes io's outputs. */
int mnrt_batch_autotune(mnrt_batch* batch, const mnrt_batch_io* io, uint32_t max_slots, void* stream);
/* Points the batch at new trees of the same scenes (rebuilt geometry or photons; same
 * ray offsets, epsilons and radii).  Synchronous.  Captured graphs stay valid unless a
 * tree is deeper than the tuned traversal stack: then the stack grows and the graphs are
 * dropped (mnrt_batch_replay returns MNRT_ERR_STATE until recaptured). */
int mnrt_batch_update_trees(mnrt_batch* batch, const mnrt_tree* tri_trees, const mnrt_tree* photon_maps);
/* Enqueue one phase (kernels only). */
int mnrt_batch_run(mnrt_batch* batch, const mnrt_batch_io* io, uint32_t phases, void* stream);
/* Captures the given phases with io into a graph kept under that phase mask. */
int mnrt_batch_capture(mnrt_batch* batch, const mnrt_batch_io* io, uint32_t phases, void* stream);
int mnrt_batch_replay(mnrt_batch* batch, uint32_t phases, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MNRT_KD_H */
