/*
 * mnrt_host.h — the reference's C++ call surface for the kd-tree hot path, over the C ABI
 * of mnrt_kd.h.  Class, struct and method names, argument order and meaning follow the
 * reference headers so a caller of the reference compiles against this one:
 *   data layouts   TriangleData, PhotonData, ShadingPoints (KernelDefs.h:160-630),
 *                  RayChunk (RayPool.h:72-124), KDTreeData / KDFinalNodeList
 *                  (KDKernelDefs.h:126-332)
 *   kd-trees       KDTreeGPU (KDTreeGPU.h:120-307), KDTreeTri (KDTreeTri.h:66-123),
 *                  KDTreePoint (KDTreePoint.h:65-181), KDTreeListener (KDTreeListener.h:61-86)
 *   photon map     PhotonMap (PhotonMap.h:89-257)
 *   kernel glue    RTUpdateKernelData / KernelRTTraceRays (RTCore.cpp:49-76),
 *                  KernelRTSampleAreaLight / KernelRTTraceShadowRaysDelta / _Area
 *                  (raytracing.cu:963-1114),
 *                  PMSetPhotonMap / PMUnsetPhotonMap / KernelPMPhotonGather
 *                  (PhotonMap.cpp:40-59)
 * Differences, all deliberate:
 *   * errors: the reference logs and aborts through MNFatal; here a failed C-ABI call
 *     throws mnrt_error (BuildTree() returns false instead);
 *   * the trees are the packed layout of mnrt_kd.h; GetData() exports the reference's
 *     KDTreeData on first use (mnrt_tree_export_preorder);
 *   * the kNN estimate is exact (k nearest within the radius cap), so
 *     KDTreePoint::ComputeQueryRadii returns the exact k-th neighbour distance and the
 *     KNN refine-iteration setting has no effect;
 *   * the kernel-glue functions bind per host thread instead of into __constant__ memory.
 */
#ifndef MNRT_HOST_H
#define MNRT_HOST_H

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "mnrt_kd.h"

typedef unsigned int uint;

class mnrt_error : public std::runtime_error {
   public:
    mnrt_error(const std::string& what, int status) : std::runtime_error(what), status(status) {}
    int status;
};

/* ---- data layouts ------------------------------------------------------------------ */

/* KDTreeData (KDKernelDefs.h:126-194): the reference's final tree layout. */
struct KDTreeData {
    void Free();
    bool IsEmpty() { return numNodes == 0; }

    uint numNodes = 0;
    float3 aabbRootMin{};
    float3 aabbRootMax{};
    uint* d_numElems = nullptr;
    uint* d_nodeAddresses = nullptr;
    float4* d_nodeExtent = nullptr;
    uint* d_childLeft = nullptr;
    uint* d_childRight = nullptr;
    uint sizeTree = 0;
    uint* d_preorderTree = nullptr;
};

/* KDFinalNodeList (KDKernelDefs.h:209-332), as handed to KDTreeListener after a build:
 * the final nodes in preorder.  d_idxFirstElem / d_elemNodeAssoc describe the leaves'
 * element lists; inner nodes' element fields cover the leaf references below them. */
struct KDFinalNodeList {
    void Free();
    bool IsEmpty() { return numNodes == 0; }

    uint numNodes = 0;
    uint maxNodes = 0;
    uint nextFreePos = 0;
    uint maxElems = 0;
    uint* d_idxFirstElem = nullptr;
    uint* d_numElems = nullptr;
    float4* d_aabbMin = nullptr;
    float4* d_aabbMax = nullptr;
    uint* d_nodeLevel = nullptr;
    uint* d_splitAxis = nullptr;
    float* d_splitPos = nullptr;
    uint* d_childLeft = nullptr;
    uint* d_childRight = nullptr;
    uint* d_elemNodeAssoc = nullptr;
};

/* TriangleData (KernelDefs.h:461-530): the members the hot path reads. */
struct TriangleData {
    uint numTris = 0;
    float3 aabbMin{};
    float3 aabbMax{};
    float4* d_verts[3] = {nullptr, nullptr, nullptr};
    float4* d_normals[3] = {nullptr, nullptr, nullptr};
    uint* d_idxMaterial = nullptr;
    float2* d_texCoords[3] = {nullptr, nullptr, nullptr};
};

/* PhotonData (KernelDefs.h:537-630): position with the azimuthal angle of the incoming
 * direction in w, power with the polar angle in w (photon_dev.h:71-116). */
struct PhotonData {
    bool Initialize(uint _maxPhotons);
    void Clear() { numPhotons = 0; }
    void Destroy();
    /* Appends the photons of `other` whose d_isValid flag is set; returns the count added. */
    uint Merge(const PhotonData& other, uint* d_isValid);

    uint numPhotons = 0;
    uint maxPhotons = 0;
    float4* d_positions = nullptr;
    float4* d_powers = nullptr;
};

/* ShadingPoints (KernelDefs.h:160-305). */
struct ShadingPoints {
    bool Initialize(uint _maxPoints);
    void Clear() { numPoints = 0; }
    void Destroy();

    uint numPoints = 0;
    uint maxPoints = 0;
    uint* d_pixels = nullptr;
    int* d_idxTris = nullptr;
    float4* d_ptInter = nullptr;
    float4* d_normalsG = nullptr;
    float4* d_normalsS = nullptr;
    float2* d_baryHit = nullptr;
};

/* RayChunk (RayPool.h:72-124). */
class RayChunk {
   public:
    void Initialize(uint _maxRays);
    void Destroy();

    float4* d_origins = nullptr;
    float4* d_dirs = nullptr;
    uint* d_pixels = nullptr;
    float4* d_influences = nullptr;
    uint rekDepth = 0;
    uint numRays = 0;
    uint maxRays = 0;
};

/* ---- kd-trees ---------------------------------------------------------------------- */

class KDTreeListener {
   public:
    KDTreeListener() {}
    virtual ~KDTreeListener() {}
    virtual void OnFinalNodeList(KDFinalNodeList* pNodesFinal) = 0;
};

class KDTreeGPU {
   public:
    KDTreeGPU(size_t numInputElems, uint numElementPoints, float3 rootAABBMin, float3 rootAABBMax);
    virtual ~KDTreeGPU();

    bool BuildTree();
    virtual void Destroy();

    KDTreeData* GetData();
    void SetCustomBits(uint bitNo, uint* d_values);
    void SetEmptySpaceRatio(float ratio) { m_fEmptySpaceRatio = ratio; }
    float GetEmptySpaceRatio() const { return m_fEmptySpaceRatio; }
    void SetSmallNodeMax(uint maximum);
    uint GetSmallNodeMax() const { return m_nSmallNodeMax; }
    void SetMaxQueryRadius(float radius) { m_fMaxQueryRadius = radius; }
    float GetMaxQueryRadius() const { return m_fMaxQueryRadius; }
    void AddListener(KDTreeListener* pListener);
    void RemoveListener(KDTreeListener* pListener);

    /* The packed tree of mnrt_kd.h (valid after BuildTree). */
    const mnrt_tree* GetTree() const { return &m_tree; }
    /* Stream used by the tree's launches (default: the null stream). */
    void SetStream(hipStream_t stream) { m_stream = stream; }

   protected:
    /* Runs the C-ABI build for this tree's kind into m_tree. */
    virtual int BuildPacked(const mnrt_build_params& params) = 0;
    virtual void PreBuild() {}
    virtual void PostBuild() {}

    size_t m_numInputElements;
    uint m_numElementPoints;
    float3 m_rootAABBMin, m_rootAABBMax;
    float m_fEmptySpaceRatio;
    uint m_nSmallNodeMax;
    float m_fMaxQueryRadius;
    mnrt_tree m_tree{};
    hipStream_t m_stream = nullptr;

   private:
    void ExportData();
    void NotifyListeners();
    KDTreeData* m_pKDData = nullptr;
    std::vector<KDTreeListener*> m_vecListeners;
};

/* KDTreeTri(const TriangleData&) (KDTreeTri.cpp:46-91): empty-space ratio 0.25, small
 * node maximum 64, root box = the scene bounds. */
class KDTreeTri : public KDTreeGPU {
   public:
    explicit KDTreeTri(const TriangleData& td);
    virtual ~KDTreeTri();

   protected:
    int BuildPacked(const mnrt_build_params& params) override;
    const TriangleData* m_pTD;
};

/* KDTreePoint (KDTreePoint.cpp:120-248): empty-space ratio 0.1, small node maximum 32,
 * VVH with the maximum query radius. */
class KDTreePoint : public KDTreeGPU {
   public:
    KDTreePoint(float4* d_points, uint numPoints, float3 sceneAABBMin, float3 sceneAABBMax, float maxQueryRadius);
    virtual ~KDTreePoint();

    /* Radius to the k-th nearest point (k = KNN target count), capped at the maximum query
     * radius. */
    void ComputeQueryRadii(float4* d_queryPoints, uint numQueryPoints, float* d_outRadii);
    void SetKNNRefineIters(uint numIters) { m_numRefineIters = numIters; }
    uint GetKNNRefineIters() const { return m_numRefineIters; }
    void SetKNNTargetCount(uint targetCount) { m_knnTargetCount = targetCount; }
    uint GetKNNTargetCount() const { return m_knnTargetCount; }
    /* Leaf-ordered photon records for kNN radiance queries (powers and incoming
     * directions per point); BuildTree attaches records of the positions alone, which
     * radius queries need. */
    void AttachPhotonRecords(const float4* d_powers, const float4* d_dirs);

   protected:
    int BuildPacked(const mnrt_build_params& params) override;
    float4* m_pPoints;
    uint m_numRefineIters = 2;
    uint m_knnTargetCount = 5;
};

/* ---- photon map -------------------------------------------------------------------- */

/* The settings PhotonMap reads (MNRTSettings.h): KNN refine iterations and the
 * illumination-cut switch (illumination cuts are not part of this path; must be false). */
class MNRTSettings {
   public:
    uint GetKNNRefineIters() const { return m_knnRefineIters; }
    void SetKNNRefineIters(uint n) { m_knnRefineIters = n; }
    bool GetUseIllumCuts() const { return false; }

   private:
    uint m_knnRefineIters = 2;
};

class PhotonMap : public KDTreeListener {
   public:
    PhotonMap(uint maxPhotons, float3 sceneAABBMin, float3 sceneAABBMax, float globalQRadiusMax,
              MNRTSettings* pSettings);
    ~PhotonMap();

    bool BuildMap();
    uint Merge(const PhotonData& other, uint* d_isValid);
    void ScaleFlux(float scale);
    /* Radiance at the shading points (sp[i] pairs with ray i of rayChunk) added into
     * d_ioRadiance[i]. */
    void Gather(const ShadingPoints& sp, const RayChunk& rayChunk, float4* d_clrDiffHit, float4* d_ioRadiance);
    /* As Gather, into d_ioRadiance[d_idxShadingPts[i]] (final gathering), with the
     * reciprocal distance from each ray's origin added into d_ioReciDistAccu if given. */
    void GatherFor(const ShadingPoints& sp, const RayChunk& rayChunk, float4* d_clrDiffHit, uint* d_idxShadingPts,
                   float4* d_ioRadiance, float* d_ioReciDistAccu);
    /* Photon visualisation: colour 1 and no 8-photon cut. */
    void Visualize(const ShadingPoints& sp, const RayChunk& rayChunk, float4* d_ioRadiance);

    uint GetPhotonCount() const { return m_Photons.numPhotons; }
    void SetKNNTargetCount(uint targetCount) { m_knnTargetCount = targetCount; }

    void OnFinalNodeList(KDFinalNodeList* pNodesFinal) override;

    PhotonData& GetPhotons() { return m_Photons; }
    const KDTreePoint* GetTree() const { return m_pPhotonMap; }

   private:
    void Destroy();
    void DestroyTree();
    void GatherImpl(const ShadingPoints& sp, const RayChunk& rayChunk, float4* d_clrDiffHit, uint* d_idxShadingPts,
                    float4* d_ioRadiance, float* d_ioReciDistAccu);

    PhotonData m_Photons;
    KDTreePoint* m_pPhotonMap = nullptr;
    float4* d_photonDirs = nullptr;  // incoming directions reconstructed from the angles
    float3 m_sceneAABBMin, m_sceneAABBMax;
    MNRTSettings* m_pSettings;
    uint m_knnTargetCount = 100;
    float m_fGlobalQRadiusMax;
};

/* ---- kernel glue (RTCore.cpp:49-76, PhotonMap.cpp:40-59) --------------------------- */

/* LightType / LightData (KernelDefs.h:89-101, 307-323). */
enum LightType { Light_Point = 0, Light_Directional = 1, Light_AreaDisc = 2, Light_AreaRect = 3 };
struct LightData {
    LightType type;
    float3 position;   /* point / disc centre / rectangle corner */
    float3 direction;  /* directional: travel direction; area: surface normal */
    float3 L_emit;
    float3 areaV1, areaV2;
    float areaRadius;
};
class MaterialData;

/* Binds the triangle tree that owns kdData (a KDTreeTri's GetData()), the light and the
 * ray epsilon for KernelRTTraceRays and the shadow-ray functions on this host thread.
 * Materials are not read (the reference binds them for shading kernels). */
void RTUpdateKernelData(const LightData& lights, const TriangleData& tris, const MaterialData& mats,
                        const KDTreeData& kdData, float fRayEpsilon);
/* Closest hits of rayChunk's rays: outInters.d_idxTris (-1: miss), d_ptInter, d_baryHit
 * and d_pixels per ray (numPoints = numRays), d_outIsValid[i] = 1 on a hit.  Normals are
 * not written (the reference computes them later, KernelBSDFGetNormalsAtHit). */
void KernelRTTraceRays(const RayChunk& rayChunk, ShadingPoints& outInters, uint* d_outIsValid);
/* kernel_SampleAreaLight (raytracing.cu:963-1000) for the bound area light: one sample
 * point per shading point from the stratified uniforms (idx + rnd) * inv. */
void KernelRTSampleAreaLight(uint numShadingPts, float idxSampleX, float invNumSamplesX, float idxSampleY,
                             float invNumSamplesY, const float* d_randoms1, const float* d_randoms2,
                             float4* d_outSamplePts);
/* kernel_TraceShadowRaysDelta / kernel_TraceShadowRaysArea (raytracing.cu:1019-1114) for
 * the bound tree and light: d_outShadowRayResult[i] = 1 iff the light (point /
 * directional) or d_samplePts[i] (area lights) is visible from shadingPts.d_ptInter[i]. */
void KernelRTTraceShadowRaysDelta(const ShadingPoints& shadingPts, uint* d_outShadowRayResult);
void KernelRTTraceShadowRaysArea(const ShadingPoints& shadingPts, const float4* d_samplePts,
                                 uint* d_outShadowRayResult);
/* Binds the photon map whose tree owns kdData, and k, for KernelPMPhotonGather. */
void PMSetPhotonMap(const KDTreeData& pmap, const PhotonData& photons, uint knnTargetCount);
void PMUnsetPhotonMap();
void KernelPMPhotonGather(const ShadingPoints& shadingPts, const RayChunk& srcRays, float* d_queryRadius,
                          float4* d_clrDiffuse, uint* d_idxShadingPoint, float4* d_ioRadiance,
                          float* d_ioInvDistAccu, bool bVisualizePhotons);

#endif /* MNRT_HOST_H */
