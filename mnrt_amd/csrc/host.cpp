// host.cpp — the reference's C++ call surface (mnrt_host.h) over the C ABI (mnrt_kd.h).
//
// Method-for-method counterparts of KDTreeGPU.cpp, KDTreeTri.cpp, KDTreePoint.cpp,
// PhotonMap.cpp and the kernel-glue functions of RTCore.cpp / PhotonMap.cpp; every
// computation goes through the C ABI.  Host code only.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "mnrt_host.h"

namespace {

void check(int rc, const char* what) {
    if (rc != MNRT_OK) throw mnrt_error(std::string(what) + ": " + mnrt_status_string(rc), rc);
}

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw mnrt_error(std::string(what) + ": " + hipGetErrorString(e), MNRT_ERR_HIP);
}

template <typename T>
T* dev_alloc(size_t n) {
    void* p = nullptr;
    hip_check(hipMalloc(&p, std::max<size_t>(n * sizeof(T), 16)), "hipMalloc");
    return (T*)p;
}

template <typename T>
void dev_free(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

// Exported KDTreeData -> the packed tree it came from (and the tree's VVH radius), for the
// kernel-glue functions that receive only a KDTreeData.
struct Bound {
    const mnrt_tree* tree;
    float max_query_radius;
};
std::mutex g_registry_mutex;
std::map<const KDTreeData*, Bound> g_registry;

void register_data(const KDTreeData* d, const mnrt_tree* t, float qr) {
    std::lock_guard<std::mutex> lock(g_registry_mutex);
    g_registry[d] = Bound{t, qr};
}

void unregister_data(const KDTreeData* d) {
    std::lock_guard<std::mutex> lock(g_registry_mutex);
    g_registry.erase(d);
}

Bound lookup(const KDTreeData* d, const char* what) {
    std::lock_guard<std::mutex> lock(g_registry_mutex);
    auto it = g_registry.find(d);
    if (it == g_registry.end())
        throw mnrt_error(std::string(what) + ": KDTreeData does not come from a live KDTreeGPU::GetData()",
                         MNRT_ERR_STATE);
    return it->second;
}

// State bound by RTUpdateKernelData / PMSetPhotonMap (the reference keeps it in
// __constant__ memory and texture references; here per host thread).
thread_local const mnrt_tree* t_tri_tree = nullptr;
thread_local float t_ray_eps = 0.0f;
thread_local mnrt_light t_light = {};
thread_local const mnrt_tree* t_photon_tree = nullptr;
thread_local float t_photon_qr = 0.0f;
thread_local uint t_knn_k = 0;

}  // namespace

// ---- data layouts ------------------------------------------------------------------------

void KDTreeData::Free() {
    dev_free(d_numElems);
    dev_free(d_nodeAddresses);
    dev_free(d_nodeExtent);
    dev_free(d_childLeft);
    dev_free(d_childRight);
    dev_free(d_preorderTree);
    numNodes = 0;
    sizeTree = 0;
}

void KDFinalNodeList::Free() {
    dev_free(d_idxFirstElem);
    dev_free(d_numElems);
    dev_free(d_aabbMin);
    dev_free(d_aabbMax);
    dev_free(d_nodeLevel);
    dev_free(d_splitAxis);
    dev_free(d_splitPos);
    dev_free(d_childLeft);
    dev_free(d_childRight);
    dev_free(d_elemNodeAssoc);
    numNodes = maxNodes = nextFreePos = maxElems = 0;
}

bool PhotonData::Initialize(uint _maxPhotons) {
    Destroy();
    maxPhotons = _maxPhotons;
    numPhotons = 0;
    d_positions = dev_alloc<float4>(_maxPhotons);
    d_powers = dev_alloc<float4>(_maxPhotons);
    return true;
}

void PhotonData::Destroy() {
    dev_free(d_positions);
    dev_free(d_powers);
    numPhotons = maxPhotons = 0;
}

uint PhotonData::Merge(const PhotonData& other, uint* d_isValid) {
    if (other.numPhotons == 0) return 0;
    uint32_t added = 0;
    check(mnrt_photons_append_valid(other.numPhotons, d_isValid, (const float*)other.d_positions,
                                    (const float*)other.d_powers, (float*)(d_positions + numPhotons),
                                    (float*)(d_powers + numPhotons), maxPhotons - numPhotons, &added, nullptr),
          "PhotonData::Merge");
    numPhotons += added;
    return added;
}

bool ShadingPoints::Initialize(uint _maxPoints) {
    Destroy();
    maxPoints = _maxPoints;
    numPoints = 0;
    d_pixels = dev_alloc<uint>(_maxPoints);
    d_idxTris = dev_alloc<int>(_maxPoints);
    d_ptInter = dev_alloc<float4>(_maxPoints);
    d_normalsG = dev_alloc<float4>(_maxPoints);
    d_normalsS = dev_alloc<float4>(_maxPoints);
    d_baryHit = dev_alloc<float2>(_maxPoints);
    return true;
}

void ShadingPoints::Destroy() {
    dev_free(d_pixels);
    dev_free(d_idxTris);
    dev_free(d_ptInter);
    dev_free(d_normalsG);
    dev_free(d_normalsS);
    dev_free(d_baryHit);
    numPoints = maxPoints = 0;
}

void RayChunk::Initialize(uint _maxRays) {
    Destroy();
    maxRays = _maxRays;
    numRays = 0;
    rekDepth = 0;
    d_origins = dev_alloc<float4>(_maxRays);
    d_dirs = dev_alloc<float4>(_maxRays);
    d_pixels = dev_alloc<uint>(_maxRays);
    d_influences = dev_alloc<float4>(_maxRays);
}

void RayChunk::Destroy() {
    dev_free(d_origins);
    dev_free(d_dirs);
    dev_free(d_pixels);
    dev_free(d_influences);
    numRays = maxRays = 0;
}

// ---- KDTreeGPU ---------------------------------------------------------------------------

KDTreeGPU::KDTreeGPU(size_t numInputElems, uint numElementPoints, float3 rootAABBMin, float3 rootAABBMax)
    : m_numInputElements(numInputElems),
      m_numElementPoints(numElementPoints),
      m_rootAABBMin(rootAABBMin),
      m_rootAABBMax(rootAABBMax),
      m_fEmptySpaceRatio(0.25f),
      m_nSmallNodeMax(64),
      m_fMaxQueryRadius(1.0f) {
    std::memset(&m_tree, 0, sizeof(m_tree));
}

KDTreeGPU::~KDTreeGPU() { KDTreeGPU::Destroy(); }

void KDTreeGPU::Destroy() {
    if (m_pKDData) {
        unregister_data(m_pKDData);
        m_pKDData->Free();
        delete m_pKDData;
        m_pKDData = nullptr;
    }
    if (m_tree.d_nodes) {
        (void)hipDeviceSynchronize();
        (void)mnrt_tree_release(&m_tree);
    }
    const uint32_t kind = m_tree.kind;
    std::memset(&m_tree, 0, sizeof(m_tree));
    m_tree.kind = kind;
}

void KDTreeGPU::SetSmallNodeMax(uint maximum) {
    if (maximum > 64 || maximum == 0)
        throw mnrt_error("Illegal small node maximum chosen. Maximum is: 64.", MNRT_ERR_ARG);
    m_nSmallNodeMax = maximum;
}

void KDTreeGPU::AddListener(KDTreeListener* pListener) {
    if (pListener && std::find(m_vecListeners.begin(), m_vecListeners.end(), pListener) == m_vecListeners.end())
        m_vecListeners.push_back(pListener);
}

void KDTreeGPU::RemoveListener(KDTreeListener* pListener) {
    m_vecListeners.erase(std::remove(m_vecListeners.begin(), m_vecListeners.end(), pListener), m_vecListeners.end());
}

bool KDTreeGPU::BuildTree() {
    Destroy();
    PreBuild();
    mnrt_build_params p{};
    p.empty_space_ratio = m_fEmptySpaceRatio;
    p.small_node_max = m_nSmallNodeMax;
    p.max_query_radius = m_fMaxQueryRadius;
    p.has_root_box = 1;
    p.root_min[0] = m_rootAABBMin.x;
    p.root_min[1] = m_rootAABBMin.y;
    p.root_min[2] = m_rootAABBMin.z;
    p.root_max[0] = m_rootAABBMax.x;
    p.root_max[1] = m_rootAABBMax.y;
    p.root_max[2] = m_rootAABBMax.z;
    if (BuildPacked(p) != MNRT_OK) return false;
    PostBuild();
    if (!m_vecListeners.empty()) NotifyListeners();
    return true;
}

KDTreeData* KDTreeGPU::GetData() {
    if (!m_pKDData && m_tree.d_nodes) ExportData();
    return m_pKDData;
}

void KDTreeGPU::ExportData() {
    uint64_t words = 0;
    check(mnrt_tree_preorder_size(&m_tree, &words), "KDTreeGPU::GetData");
    auto* d = new KDTreeData;
    try {
        const uint32_t n = m_tree.num_nodes;
        d->numNodes = n;
        d->aabbRootMin = make_float3(m_tree.root_min[0], m_tree.root_min[1], m_tree.root_min[2]);
        d->aabbRootMax = make_float3(m_tree.root_max[0], m_tree.root_max[1], m_tree.root_max[2]);
        d->d_numElems = dev_alloc<uint>(n);
        d->d_nodeAddresses = dev_alloc<uint>(n);
        d->d_nodeExtent = dev_alloc<float4>(n);
        d->d_childLeft = dev_alloc<uint>(n);
        d->d_childRight = dev_alloc<uint>(n);
        d->d_preorderTree = dev_alloc<uint>(words);
        d->sizeTree = (uint)words;
        check(mnrt_tree_export_preorder(&m_tree, d->d_preorderTree, words, d->d_numElems, d->d_nodeAddresses,
                                        (float*)d->d_nodeExtent, d->d_childLeft, d->d_childRight, m_stream),
              "KDTreeGPU::GetData");
    } catch (...) {
        d->Free();
        delete d;
        throw;
    }
    m_pKDData = d;
    register_data(d, &m_tree, m_fMaxQueryRadius);
}

void KDTreeGPU::SetCustomBits(uint bitNo, uint* d_values) {
    if (bitNo > 1 || !d_values) throw mnrt_error("SetCustomBits: bit 0 or 1 and values required", MNRT_ERR_ARG);
    KDTreeData* d = GetData();
    if (!d) throw mnrt_error("Setting custom bits requires final kd-tree data.", MNRT_ERR_STATE);
    check(mnrt_preorder_set_custom_bits(d->d_preorderTree, d->d_nodeAddresses, d->numNodes, bitNo, d_values,
                                        m_stream),
          "KDTreeGPU::SetCustomBits");
}

// The final node list in preorder, built on the host from the node words and the export.
void KDTreeGPU::NotifyListeners() {
    KDTreeData* d = GetData();
    const uint32_t n = m_tree.num_nodes;
    std::vector<uint32_t> w(2ull * n), ne(n), cl(n), cr(n);
    std::vector<float4> ext(n);
    hip_check(hipMemcpy(w.data(), m_tree.d_nodes, w.size() * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    hip_check(hipMemcpy(ne.data(), d->d_numElems, n * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    hip_check(hipMemcpy(cl.data(), d->d_childLeft, n * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    hip_check(hipMemcpy(cr.data(), d->d_childRight, n * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    std::vector<uint32_t> first(n), level(n, 0), axis(n, 0);
    std::vector<float> split(n, 0.0f);
    std::vector<float4> lo(n), hi(n);
    // preorder: a parent precedes its children, so level and box propagate in index order
    lo[0] = make_float4(m_tree.root_min[0], m_tree.root_min[1], m_tree.root_min[2], 0.0f);
    hi[0] = make_float4(m_tree.root_max[0], m_tree.root_max[1], m_tree.root_max[2], 0.0f);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t w1 = w[2 * i + 1];
        if ((w1 >> 30) == MNRT_LEAF_TAG) {
            first[i] = w1 & MNRT_INDEX_MASK;
            continue;
        }
        const uint32_t a = w1 >> 30, r = w1 & MNRT_INDEX_MASK;
        float s;
        std::memcpy(&s, &w[2 * i], 4);
        axis[i] = a;
        split[i] = s;
        level[i + 1] = level[r] = level[i] + 1;
        lo[i + 1] = lo[r] = lo[i];
        hi[i + 1] = hi[r] = hi[i];
        (&hi[i + 1].x)[a] = s;
        (&lo[r].x)[a] = s;
    }
    for (uint32_t i = n; i-- > 0;) {
        if ((w[2 * i + 1] >> 30) != MNRT_LEAF_TAG) first[i] = first[i + 1];
    }
    KDFinalNodeList L;
    L.numNodes = L.maxNodes = L.nextFreePos = n;
    L.maxElems = m_tree.num_refs;
    auto up = [&](auto* dst, const auto& src) {
        hip_check(hipMemcpy(dst, src.data(), src.size() * sizeof(src[0]), hipMemcpyHostToDevice), "hipMemcpy");
    };
    try {
        L.d_idxFirstElem = dev_alloc<uint>(n);
        L.d_numElems = dev_alloc<uint>(n);
        L.d_aabbMin = dev_alloc<float4>(n);
        L.d_aabbMax = dev_alloc<float4>(n);
        L.d_nodeLevel = dev_alloc<uint>(n);
        L.d_splitAxis = dev_alloc<uint>(n);
        L.d_splitPos = dev_alloc<float>(n);
        L.d_childLeft = dev_alloc<uint>(n);
        L.d_childRight = dev_alloc<uint>(n);
        L.d_elemNodeAssoc = dev_alloc<uint>(m_tree.num_refs);
        up(L.d_idxFirstElem, first);
        up(L.d_numElems, ne);
        up(L.d_aabbMin, lo);
        up(L.d_aabbMax, hi);
        up(L.d_nodeLevel, level);
        up(L.d_splitAxis, axis);
        up(L.d_splitPos, split);
        up(L.d_childLeft, cl);
        up(L.d_childRight, cr);
        if (m_tree.num_refs)
            hip_check(hipMemcpy(L.d_elemNodeAssoc, m_tree.d_refs, (size_t)m_tree.num_refs * 4, hipMemcpyDeviceToDevice),
                      "hipMemcpy");
        for (KDTreeListener* l : m_vecListeners) l->OnFinalNodeList(&L);
    } catch (...) {
        L.Free();
        throw;
    }
    L.Free();
}

// ---- KDTreeTri / KDTreePoint -------------------------------------------------------------

KDTreeTri::KDTreeTri(const TriangleData& td) : KDTreeGPU(td.numTris, 2, td.aabbMin, td.aabbMax), m_pTD(&td) {
    m_tree.kind = MNRT_KIND_TRI;
    SetEmptySpaceRatio(0.25f);
    SetSmallNodeMax(64);
}

KDTreeTri::~KDTreeTri() {}

int KDTreeTri::BuildPacked(const mnrt_build_params& params) {
    return mnrt_build_tri((const float*)m_pTD->d_verts[0], (const float*)m_pTD->d_verts[1],
                          (const float*)m_pTD->d_verts[2], m_pTD->numTris, &params, &m_tree, m_stream);
}

KDTreePoint::KDTreePoint(float4* d_points, uint numPoints, float3 sceneAABBMin, float3 sceneAABBMax,
                         float maxQueryRadius)
    : KDTreeGPU(numPoints, 1, sceneAABBMin, sceneAABBMax), m_pPoints(d_points) {
    m_tree.kind = MNRT_KIND_POINT;
    SetEmptySpaceRatio(0.1f);
    SetSmallNodeMax(32);
    SetMaxQueryRadius(maxQueryRadius);
}

KDTreePoint::~KDTreePoint() {}

int KDTreePoint::BuildPacked(const mnrt_build_params& params) {
    int rc = mnrt_build_points((const float*)m_pPoints, (uint32_t)m_numInputElements, &params, &m_tree, m_stream);
    if (rc != MNRT_OK) return rc;
    // radius queries read positions only: records of the positions (power and direction
    // fields unused until AttachPhotonRecords)
    return mnrt_tree_attach_photon_records(&m_tree, (const float*)m_pPoints, (const float*)m_pPoints,
                                           (const float*)m_pPoints, m_stream);
}

void KDTreePoint::AttachPhotonRecords(const float4* d_powers, const float4* d_dirs) {
    check(mnrt_tree_attach_photon_records(&m_tree, (const float*)m_pPoints, (const float*)d_powers,
                                          (const float*)d_dirs, m_stream),
          "KDTreePoint::AttachPhotonRecords");
}

void KDTreePoint::ComputeQueryRadii(float4* d_queryPoints, uint numQueryPoints, float* d_outRadii) {
    check(mnrt_knn_radii(&m_tree, m_knnTargetCount, m_fMaxQueryRadius, numQueryPoints, (const float*)d_queryPoints,
                         d_outRadii, m_stream),
          "KDTreePoint::ComputeQueryRadii");
}

// ---- PhotonMap ---------------------------------------------------------------------------

PhotonMap::PhotonMap(uint maxPhotons, float3 sceneAABBMin, float3 sceneAABBMax, float globalQRadiusMax,
                     MNRTSettings* pSettings)
    : m_sceneAABBMin(sceneAABBMin), m_sceneAABBMax(sceneAABBMax), m_pSettings(pSettings),
      m_fGlobalQRadiusMax(globalQRadiusMax) {
    if (maxPhotons == 0 || !pSettings) throw mnrt_error("PhotonMap: maxPhotons > 0 and settings required", MNRT_ERR_ARG);
    m_Photons.Initialize(maxPhotons);
}

PhotonMap::~PhotonMap() { Destroy(); }

void PhotonMap::Destroy() {
    if (t_photon_tree && m_pPhotonMap && t_photon_tree == m_pPhotonMap->GetTree()) PMUnsetPhotonMap();
    m_Photons.Destroy();
    DestroyTree();
}

void PhotonMap::DestroyTree() {
    delete m_pPhotonMap;
    m_pPhotonMap = nullptr;
    dev_free(d_photonDirs);
}

bool PhotonMap::BuildMap() {
    DestroyTree();
    if (m_Photons.numPhotons == 0) return true;
    m_pPhotonMap = new KDTreePoint(m_Photons.d_positions, m_Photons.numPhotons, m_sceneAABBMin, m_sceneAABBMax,
                                   m_fGlobalQRadiusMax);
    m_pPhotonMap->AddListener(this);
    m_pPhotonMap->SetKNNRefineIters(m_pSettings->GetKNNRefineIters());
    m_pPhotonMap->SetKNNTargetCount(m_knnTargetCount);
    if (!m_pPhotonMap->BuildTree()) return false;
    d_photonDirs = dev_alloc<float4>(m_Photons.numPhotons);
    check(mnrt_photons_spherical_to_dir(m_Photons.numPhotons, (const float*)m_Photons.d_positions,
                                        (const float*)m_Photons.d_powers, (float*)d_photonDirs, nullptr),
          "PhotonMap::BuildMap");
    m_pPhotonMap->AttachPhotonRecords(m_Photons.d_powers, d_photonDirs);
    return true;
}

uint PhotonMap::Merge(const PhotonData& other, uint* d_isValid) { return m_Photons.Merge(other, d_isValid); }

void PhotonMap::ScaleFlux(float scale) {
    if (m_Photons.numPhotons == 0) return;
    check(mnrt_photons_scale_flux(m_Photons.numPhotons, (float*)m_Photons.d_powers, scale, nullptr),
          "PhotonMap::ScaleFlux");
    // a map built before the scaling reads the new powers from refreshed records
    if (m_pPhotonMap && d_photonDirs) m_pPhotonMap->AttachPhotonRecords(m_Photons.d_powers, d_photonDirs);
}

void PhotonMap::GatherImpl(const ShadingPoints& sp, const RayChunk& rays, float4* d_clrDiffHit,
                           uint* d_idxShadingPts, float4* d_ioRadiance, float* d_ioReciDistAccu) {
    if (m_Photons.numPhotons == 0 || sp.numPoints == 0) return;
    check(mnrt_knn_gather(m_pPhotonMap->GetTree(), m_knnTargetCount, m_fGlobalQRadiusMax, sp.numPoints,
                          (const float*)sp.d_ptInter, (const float*)sp.d_normalsS, (const float*)rays.d_dirs,
                          (const float*)rays.d_influences, (const float*)d_clrDiffHit, d_idxShadingPts, nullptr,
                          (const float*)rays.d_origins, (float*)d_ioRadiance, d_ioReciDistAccu, nullptr, nullptr,
                          nullptr, nullptr, nullptr),
          "PhotonMap::Gather");
}

void PhotonMap::Gather(const ShadingPoints& sp, const RayChunk& rayChunk, float4* d_clrDiffHit,
                       float4* d_ioRadiance) {
    GatherImpl(sp, rayChunk, d_clrDiffHit, nullptr, d_ioRadiance, nullptr);
}

void PhotonMap::GatherFor(const ShadingPoints& sp, const RayChunk& rayChunk, float4* d_clrDiffHit,
                          uint* d_idxShadingPts, float4* d_ioRadiance, float* d_ioReciDistAccu) {
    GatherImpl(sp, rayChunk, d_clrDiffHit, d_idxShadingPts, d_ioRadiance, d_ioReciDistAccu);
}

void PhotonMap::Visualize(const ShadingPoints& sp, const RayChunk& rayChunk, float4* d_ioRadiance) {
    GatherImpl(sp, rayChunk, nullptr, nullptr, d_ioRadiance, nullptr);
}

void PhotonMap::OnFinalNodeList(KDFinalNodeList*) {
    // illumination cuts (the reference's only use of the final node list here) are not
    // part of this path; MNRTSettings::GetUseIllumCuts() is false
}

// ---- kernel glue -------------------------------------------------------------------------

void RTUpdateKernelData(const LightData& lights, const TriangleData&, const MaterialData&, const KDTreeData& kdData,
                        float fRayEpsilon) {
    Bound b = lookup(&kdData, "RTUpdateKernelData");
    if (b.tree->kind != MNRT_KIND_TRI) throw mnrt_error("RTUpdateKernelData: not a triangle tree", MNRT_ERR_STATE);
    t_tri_tree = b.tree;
    t_ray_eps = fRayEpsilon;
    mnrt_light L{};
    L.type = (uint32_t)lights.type;
    const float3* src[5] = {&lights.position, &lights.direction, &lights.L_emit, &lights.areaV1, &lights.areaV2};
    float* dst[5] = {L.position, L.direction, L.L_emit, L.area_v1, L.area_v2};
    for (int k = 0; k < 5; ++k) {
        dst[k][0] = src[k]->x;
        dst[k][1] = src[k]->y;
        dst[k][2] = src[k]->z;
    }
    L.area_radius = lights.areaRadius;
    t_light = L;
}

void KernelRTSampleAreaLight(uint numShadingPts, float idxSampleX, float invNumSamplesX, float idxSampleY,
                             float invNumSamplesY, const float* d_randoms1, const float* d_randoms2,
                             float4* d_outSamplePts) {
    check(mnrt_sample_area_light(&t_light, numShadingPts, idxSampleX, invNumSamplesX, idxSampleY, invNumSamplesY,
                                 d_randoms1, d_randoms2, (float*)d_outSamplePts, nullptr),
          "KernelRTSampleAreaLight");
}

void KernelRTTraceShadowRaysDelta(const ShadingPoints& shadingPts, uint* d_outShadowRayResult) {
    if (!t_tri_tree) throw mnrt_error("KernelRTTraceShadowRaysDelta: no tree bound (RTUpdateKernelData)", MNRT_ERR_STATE);
    if (t_light.type != MNRT_LIGHT_POINT && t_light.type != MNRT_LIGHT_DIRECTIONAL)
        throw mnrt_error("KernelRTTraceShadowRaysDelta: the bound light is not a delta light", MNRT_ERR_ARG);
    check(mnrt_trace_shadow(t_tri_tree, &t_light, shadingPts.numPoints, (const float*)shadingPts.d_ptInter, nullptr,
                            t_ray_eps, d_outShadowRayResult, nullptr),
          "KernelRTTraceShadowRaysDelta");
}

void KernelRTTraceShadowRaysArea(const ShadingPoints& shadingPts, const float4* d_samplePts,
                                 uint* d_outShadowRayResult) {
    if (!t_tri_tree) throw mnrt_error("KernelRTTraceShadowRaysArea: no tree bound (RTUpdateKernelData)", MNRT_ERR_STATE);
    if (t_light.type != MNRT_LIGHT_AREA_DISC && t_light.type != MNRT_LIGHT_AREA_RECT)
        throw mnrt_error("KernelRTTraceShadowRaysArea: the bound light is not an area light", MNRT_ERR_ARG);
    check(mnrt_trace_shadow(t_tri_tree, &t_light, shadingPts.numPoints, (const float*)shadingPts.d_ptInter,
                            (const float*)d_samplePts, t_ray_eps, d_outShadowRayResult, nullptr),
          "KernelRTTraceShadowRaysArea");
}

void KernelRTTraceRays(const RayChunk& rayChunk, ShadingPoints& outInters, uint* d_outIsValid) {
    if (!t_tri_tree) throw mnrt_error("KernelRTTraceRays: no tree bound (RTUpdateKernelData)", MNRT_ERR_STATE);
    if (rayChunk.numRays > outInters.maxPoints)
        throw mnrt_error("KernelRTTraceRays: shading points too small", MNRT_ERR_ARG);
    check(mnrt_trace_closest(t_tri_tree, rayChunk.numRays, (const float*)rayChunk.d_origins,
                             (const float*)rayChunk.d_dirs, t_ray_eps, outInters.d_idxTris, nullptr,
                             (float*)outInters.d_baryHit, (float*)outInters.d_ptInter, d_outIsValid, nullptr),
          "KernelRTTraceRays");
    if (rayChunk.numRays && rayChunk.d_pixels && outInters.d_pixels)
        hip_check(hipMemcpyAsync(outInters.d_pixels, rayChunk.d_pixels, (size_t)rayChunk.numRays * 4,
                                 hipMemcpyDeviceToDevice, nullptr),
                  "hipMemcpyAsync");
    outInters.numPoints = rayChunk.numRays;
}

void PMSetPhotonMap(const KDTreeData& pmap, const PhotonData&, uint knnTargetCount) {
    Bound b = lookup(&pmap, "PMSetPhotonMap");
    if (b.tree->kind != MNRT_KIND_POINT) throw mnrt_error("PMSetPhotonMap: not a point tree", MNRT_ERR_STATE);
    t_photon_tree = b.tree;
    t_photon_qr = b.max_query_radius;
    t_knn_k = knnTargetCount;
}

void PMUnsetPhotonMap() {
    t_photon_tree = nullptr;
    t_knn_k = 0;
}

void KernelPMPhotonGather(const ShadingPoints& sp, const RayChunk& rays, float* d_queryRadius, float4* d_clrDiffuse,
                          uint* d_idxShadingPoint, float4* d_ioRadiance, float* d_ioInvDistAccu,
                          bool bVisualizePhotons) {
    if (!t_photon_tree) throw mnrt_error("KernelPMPhotonGather: no photon map bound (PMSetPhotonMap)", MNRT_ERR_STATE);
    if (sp.numPoints == 0) return;
    check(mnrt_knn_gather(t_photon_tree, t_knn_k, t_photon_qr, sp.numPoints, (const float*)sp.d_ptInter,
                          (const float*)sp.d_normalsS, (const float*)rays.d_dirs, (const float*)rays.d_influences,
                          bVisualizePhotons ? nullptr : (const float*)d_clrDiffuse, d_idxShadingPoint, d_queryRadius,
                          (const float*)rays.d_origins, (float*)d_ioRadiance, d_ioInvDistAccu, nullptr, nullptr,
                          nullptr, nullptr, nullptr),
          "KernelPMPhotonGather");
}
