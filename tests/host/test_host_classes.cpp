// GPU test of the C++ call surface (mnrt_host.h): each class method gives what the C ABI
// gives for the same inputs, and the reference's contracts hold (SetSmallNodeMax limit,
// KDTreeData size, merge compaction, flux scaling).  Prints one line per check; exit 0 on
// success.  Run by tests/test_gpu_host_classes.py.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "mnrt_host.h"

class MaterialData {
   public:
    int unused;
};

#define HIPOK(x)                                                            \
    do {                                                                    \
        hipError_t e_ = (x);                                                \
        if (e_ != hipSuccess) {                                             \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));   \
            std::exit(2);                                                   \
        }                                                                   \
    } while (0)
#define ABIOK(x)                                                                 \
    do {                                                                         \
        int r_ = (x);                                                            \
        if (r_ != MNRT_OK) {                                                     \
            std::fprintf(stderr, "%s: %s\n", #x, mnrt_status_string(r_));       \
            std::exit(2);                                                        \
        }                                                                        \
    } while (0)

static int g_fail = 0;
static void expect(bool ok, const char* what) {
    std::printf("%s: %s\n", ok ? "ok  " : "FAIL", what);
    std::fflush(stdout);
    if (!ok) g_fail = 1;
}

static uint32_t g_rng = 12345u;
static float frand() {
    g_rng = g_rng * 1664525u + 1013904223u;
    return (float)(g_rng >> 8) / 16777216.0f;
}

template <typename T>
static T* upload(const std::vector<T>& h) {
    void* p = nullptr;
    HIPOK(hipMalloc(&p, std::max<size_t>(h.size() * sizeof(T), 16)));
    if (!h.empty()) HIPOK(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return (T*)p;
}

template <typename T>
static std::vector<T> download(const T* d, size_t n) {
    std::vector<T> h(n);
    if (n) HIPOK(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost));
    return h;
}

static bool same_bytes(const void* a, const void* b, size_t n) { return std::memcmp(a, b, n) == 0; }

int main() {
    // the structs this program shares with the library (mnrt_tree inside the classes) must
    // have the layout the library was built with
    if (mnrt_abi_version() != MNRT_ABI_VERSION) {
        std::fprintf(stderr, "library ABI %d, this program was built for %d: rebuild it\n", mnrt_abi_version(),
                     MNRT_ABI_VERSION);
        return 2;
    }
    ABIOK(mnrt_device_init(0));
    // scene: the 12 inward faces of [-1,1]^3, a light quad at the ceiling, 3000 small triangles
    std::vector<float4> v0, v1, v2, alb;
    auto tri = [&](float4 a, float4 b, float4 c, float4 al) {
        v0.push_back(a);
        v1.push_back(b);
        v2.push_back(c);
        alb.push_back(al);
    };
    const float4 grey = make_float4(0.7f, 0.7f, 0.7f, 1.0f);
    for (int ax = 0; ax < 3; ++ax) {
        for (int side = -1; side <= 1; side += 2) {
            float4 q[4];
            for (int k = 0; k < 4; ++k) {
                float u = (k == 1 || k == 2) ? 1.0f : -1.0f, w = (k >= 2) ? 1.0f : -1.0f;
                float c[3];
                c[ax] = (float)side;
                c[(ax + 1) % 3] = u;
                c[(ax + 2) % 3] = w;
                q[k] = make_float4(c[0], c[1], c[2], 0.0f);
            }
            tri(q[0], q[1], q[2], grey);
            tri(q[0], q[2], q[3], grey);
        }
    }
    const float4 black = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    float4 L0 = make_float4(-0.25f, 0.999f, -0.25f, 0), L1 = make_float4(0.25f, 0.999f, -0.25f, 0),
           L2 = make_float4(0.25f, 0.999f, 0.25f, 0), L3 = make_float4(-0.25f, 0.999f, 0.25f, 0);
    tri(L0, L1, L2, black);
    tri(L0, L2, L3, black);
    for (int i = 0; i < 3000; ++i) {
        float4 a = make_float4(frand() * 1.6f - 0.8f, frand() * 1.6f - 0.8f, frand() * 1.6f - 0.8f, 0.0f);
        auto jit = [&]() { return make_float4(a.x + frand() * 0.1f, a.y + frand() * 0.1f, a.z + frand() * 0.1f, 0.0f); };
        tri(a, jit(), jit(), grey);
    }
    const uint32_t nt = (uint32_t)v0.size();
    TriangleData td;
    td.numTris = nt;
    td.aabbMin = make_float3(-1, -1, -1);
    td.aabbMax = make_float3(1, 1, 1);
    td.d_verts[0] = upload(v0);
    td.d_verts[1] = upload(v1);
    td.d_verts[2] = upload(v2);
    float4* d_alb = upload(alb);

    // ---- KDTreeTri ----
    KDTreeTri kd(td);
    bool threw = false;
    try {
        kd.SetSmallNodeMax(65);
    } catch (const mnrt_error&) {
        threw = true;
    }
    expect(threw && kd.GetSmallNodeMax() == 64, "SetSmallNodeMax above 64 is refused");
    expect(kd.BuildTree(), "KDTreeTri::BuildTree");
    const mnrt_tree* T = kd.GetTree();
    mnrt_build_params bp{};
    bp.empty_space_ratio = 0.25f;
    bp.small_node_max = 64;
    bp.max_query_radius = 1.0f;
    bp.has_root_box = 1;
    for (int a = 0; a < 3; ++a) {
        bp.root_min[a] = -1.0f;
        bp.root_max[a] = 1.0f;
    }
    mnrt_tree direct{};
    ABIOK(mnrt_build_tri((const float*)td.d_verts[0], (const float*)td.d_verts[1], (const float*)td.d_verts[2], nt,
                         &bp, &direct, nullptr));
    {
        auto a = download(T->d_nodes, 2ull * T->num_nodes);
        auto b = download(direct.d_nodes, 2ull * direct.num_nodes);
        expect(T->num_nodes == direct.num_nodes && a == b, "KDTreeTri tree == mnrt_build_tri tree");
    }
    KDTreeData* kdd = kd.GetData();
    uint64_t words = 0;
    ABIOK(mnrt_tree_preorder_size(T, &words));
    expect(kdd && kdd->numNodes == T->num_nodes && kdd->sizeTree == words &&
               kdd->sizeTree == 3 * (T->num_nodes - T->num_leaves) + 2 * T->num_leaves + T->num_refs,
           "GetData: KDTreeData sizes");
    {
        auto first = download(kdd->d_preorderTree, 3);
        expect(first[0] == 0u && (first[1] >> 30) == (download(T->d_nodes, 2)[1] >> 30), "GetData: root words");
    }

    // ---- RTUpdateKernelData + KernelRTTraceRays ----
    const uint32_t nr = 4096;
    std::vector<float4> ro(nr), rd(nr), rinfl(nr, make_float4(1, 1, 1, 0));
    std::vector<uint> pix(nr);
    for (uint32_t i = 0; i < nr; ++i) {
        ro[i] = make_float4(0.0f, 0.0f, 0.9f, 0.0f);
        float x = frand() * 2 - 1, y = frand() * 2 - 1, z = -1.0f;
        float l = std::sqrt(x * x + y * y + z * z);
        rd[i] = make_float4(x / l, y / l, z / l, 0.0f);
        pix[i] = i * 7u;
    }
    RayChunk rc;
    rc.Initialize(nr);
    HIPOK(hipMemcpy(rc.d_origins, ro.data(), nr * 16, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(rc.d_dirs, rd.data(), nr * 16, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(rc.d_pixels, pix.data(), nr * 4, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(rc.d_influences, rinfl.data(), nr * 16, hipMemcpyHostToDevice));
    rc.numRays = nr;
    ShadingPoints sp;
    sp.Initialize(nr);
    uint* d_valid = nullptr;
    HIPOK(hipMalloc((void**)&d_valid, nr * 4));
    LightData lights{};
    MaterialData mats{};
    const float eps = 1e-4f;
    RTUpdateKernelData(lights, td, mats, *kdd, eps);
    KernelRTTraceRays(rc, sp, d_valid);
    HIPOK(hipDeviceSynchronize());
    int32_t* d_tri = nullptr;
    float* d_t = nullptr;
    HIPOK(hipMalloc((void**)&d_tri, nr * 4));
    HIPOK(hipMalloc((void**)&d_t, nr * 4));
    ABIOK(mnrt_trace_closest(T, nr, (const float*)rc.d_origins, (const float*)rc.d_dirs, eps, d_tri, d_t, nullptr,
                             nullptr, nullptr, nullptr));
    {
        auto a = download(sp.d_idxTris, nr);
        auto b = download(d_tri, nr);
        auto p = download(sp.d_pixels, nr);
        auto v = download(d_valid, nr);
        bool ok = sp.numPoints == nr && same_bytes(a.data(), b.data(), nr * 4) && p == pix;
        for (uint32_t i = 0; i < nr; ++i) ok = ok && (v[i] == (a[i] >= 0 ? 1u : 0u));
        expect(ok, "KernelRTTraceRays == mnrt_trace_closest (hits, pixels, valid flags)");
    }
    try {
        KDTreeData fake;
        RTUpdateKernelData(lights, td, mats, fake, eps);
        expect(false, "RTUpdateKernelData refuses a KDTreeData of no tree");
    } catch (const mnrt_error&) {
        expect(true, "RTUpdateKernelData refuses a KDTreeData of no tree");
    }

    // ---- shadow rays: KernelRTTraceShadowRaysDelta / _Area, KernelRTSampleAreaLight ----
    {
        uint *d_res = nullptr, *d_ref = nullptr;
        HIPOK(hipMalloc((void**)&d_res, nr * 4));
        HIPOK(hipMalloc((void**)&d_ref, nr * 4));
        LightData pl{};
        pl.type = Light_Point;
        pl.position = make_float3(0.1f, 0.9f, 0.0f);
        RTUpdateKernelData(pl, td, mats, *kdd, eps);
        KernelRTTraceShadowRaysDelta(sp, d_res);
        mnrt_light ml{};
        ml.type = MNRT_LIGHT_POINT;
        ml.position[0] = 0.1f;
        ml.position[1] = 0.9f;
        ABIOK(mnrt_trace_shadow(T, &ml, nr, (const float*)sp.d_ptInter, nullptr, eps, d_ref, nullptr));
        HIPOK(hipDeviceSynchronize());
        auto a = download(d_res, nr), b = download(d_ref, nr);
        uint32_t lit = 0;
        for (uint32_t i = 0; i < nr; ++i) lit += a[i];
        expect(a == b && lit > 0 && lit < nr, "KernelRTTraceShadowRaysDelta == mnrt_trace_shadow (point light)");

        LightData al{};
        al.type = Light_AreaRect;
        al.position = make_float3(-0.25f, 0.998f, -0.25f);
        al.direction = make_float3(0.0f, -1.0f, 0.0f);
        al.areaV1 = make_float3(0.5f, 0.0f, 0.0f);
        al.areaV2 = make_float3(0.0f, 0.0f, 0.5f);
        RTUpdateKernelData(al, td, mats, *kdd, eps);
        std::vector<float> r1(nr), r2(nr);
        for (uint32_t i = 0; i < nr; ++i) {
            r1[i] = frand();
            r2[i] = frand();
        }
        float *d_r1 = upload(r1), *d_r2 = upload(r2);
        float4 *d_smp = nullptr, *d_smp2 = nullptr;
        HIPOK(hipMalloc((void**)&d_smp, nr * 16));
        HIPOK(hipMalloc((void**)&d_smp2, nr * 16));
        KernelRTSampleAreaLight(nr, 1.0f, 0.5f, 0.0f, 0.5f, d_r1, d_r2, d_smp);
        mnrt_light ma{};
        ma.type = MNRT_LIGHT_AREA_RECT;
        ma.position[0] = -0.25f;
        ma.position[1] = 0.998f;
        ma.position[2] = -0.25f;
        ma.direction[1] = -1.0f;
        ma.area_v1[0] = 0.5f;
        ma.area_v2[2] = 0.5f;
        ABIOK(mnrt_sample_area_light(&ma, nr, 1.0f, 0.5f, 0.0f, 0.5f, d_r1, d_r2, (float*)d_smp2, nullptr));
        KernelRTTraceShadowRaysArea(sp, d_smp, d_res);
        ABIOK(mnrt_trace_shadow(T, &ma, nr, (const float*)sp.d_ptInter, (const float*)d_smp2, eps, d_ref, nullptr));
        HIPOK(hipDeviceSynchronize());
        auto s1 = download(d_smp, nr), s2 = download(d_smp2, nr);
        a = download(d_res, nr);
        b = download(d_ref, nr);
        lit = 0;
        for (uint32_t i = 0; i < nr; ++i) lit += a[i];
        expect(same_bytes(s1.data(), s2.data(), nr * 16), "KernelRTSampleAreaLight == mnrt_sample_area_light");
        expect(a == b && lit > 0 && lit < nr, "KernelRTTraceShadowRaysArea == mnrt_trace_shadow (area light)");
        bool threw_kind = false;
        try {
            KernelRTTraceShadowRaysDelta(sp, d_res);
        } catch (const mnrt_error&) {
            threw_kind = true;
        }
        expect(threw_kind, "KernelRTTraceShadowRaysDelta refuses a bound area light");
        RTUpdateKernelData(lights, td, mats, *kdd, eps);
        HIPOK(hipFree(d_res));
        HIPOK(hipFree(d_ref));
        HIPOK(hipFree(d_r1));
        HIPOK(hipFree(d_r2));
        HIPOK(hipFree(d_smp));
        HIPOK(hipFree(d_smp2));
    }

    // ---- photons into the reference PhotonData layout ----
    std::vector<float> light = {-0.25f, 0.999f, -0.25f, 0, 0.5f, 0, 0, 0, 0, 0, 0.5f, 0, 10, 10, 10, 0};
    float* d_light = upload(light);
    const uint32_t target = 20000;
    float4 *d_pp = nullptr, *d_pw = nullptr, *d_pd = nullptr;
    HIPOK(hipMalloc((void**)&d_pp, target * 16));
    HIPOK(hipMalloc((void**)&d_pw, target * 16));
    HIPOK(hipMalloc((void**)&d_pd, target * 16));
    uint32_t got = 0, paths = 0;
    ABIOK(mnrt_photons_generate(T, (const float*)td.d_verts[0], (const float*)td.d_verts[1],
                                (const float*)td.d_verts[2], (const float*)d_alb, d_light, 1, target, 4, 1, eps,
                                (float*)d_pp, (float*)d_pw, (float*)d_pd, nullptr, nullptr, nullptr, &got, &paths,
                                nullptr));
    ABIOK(mnrt_photons_dir_to_spherical(got, (const float*)d_pd, (float*)d_pp, (float*)d_pw, nullptr));
    std::vector<uint> valid(got);
    uint32_t nvalid = 0;
    for (uint32_t i = 0; i < got; ++i) nvalid += (valid[i] = (i % 3 != 1) ? 1u : 0u);
    uint* d_pvalid = upload(valid);
    PhotonData other;
    other.numPhotons = other.maxPhotons = got;
    other.d_positions = d_pp;
    other.d_powers = d_pw;

    MNRTSettings settings;
    const float rmax = 0.35f;
    PhotonMap pm(2 * got, td.aabbMin, td.aabbMax, rmax, &settings);
    uint added = pm.Merge(other, d_pvalid);
    expect(added == nvalid && pm.GetPhotonCount() == nvalid, "PhotonMap::Merge appends the valid photons");
    {
        auto src = download(d_pp, got);
        auto dst = download(pm.GetPhotons().d_positions, nvalid);
        bool ok = true;
        uint32_t j = 0;
        for (uint32_t i = 0; i < got; ++i)
            if (valid[i]) ok = ok && same_bytes(&src[i], &dst[j++], 16);
        expect(ok, "PhotonMap::Merge keeps order");
    }
    std::vector<float4> pw_before = download(pm.GetPhotons().d_powers, nvalid);
    pm.ScaleFlux(0.5f);
    {
        auto after = download(pm.GetPhotons().d_powers, nvalid);
        bool ok = true;
        for (uint32_t i = 0; i < nvalid; ++i)
            ok = ok && after[i].x == pw_before[i].x * 0.5f && after[i].z == pw_before[i].z * 0.5f &&
                 same_bytes(&after[i].w, &pw_before[i].w, 4);
        expect(ok, "PhotonMap::ScaleFlux scales rgb, keeps the polar angle");
    }
    pm.SetKNNTargetCount(40);
    expect(pm.BuildMap(), "PhotonMap::BuildMap");

    // shading points at the primary hits (compacted), gathered by the class and directly
    float4 *qp, *qn, *qd, *qc, *qi;
    uint32_t* qsrc;
    for (float4** p : {&qp, &qn, &qd, &qc, &qi}) HIPOK(hipMalloc((void**)p, nr * 16));
    HIPOK(hipMalloc((void**)&qsrc, nr * 4));
    uint32_t nq = 0;
    ABIOK(mnrt_shading_points(nr, (const float*)rc.d_origins, (const float*)rc.d_dirs, d_tri, d_t,
                              (const float*)td.d_verts[0], (const float*)td.d_verts[1], (const float*)td.d_verts[2],
                              (const float*)d_alb, nullptr, (float*)qp, (float*)qn, (float*)qd, (float*)qc,
                              (float*)qi, qsrc, nullptr, &nq, nullptr));
    ShadingPoints g;
    g.numPoints = nq;
    g.d_ptInter = qp;
    g.d_normalsS = qn;
    RayChunk gr;
    gr.d_dirs = qd;
    gr.d_influences = qi;
    gr.d_origins = rc.d_origins;  // every ray starts at the eye
    gr.numRays = nq;
    float4* d_rad = nullptr;
    float4* d_rad2 = nullptr;
    HIPOK(hipMalloc((void**)&d_rad, nq * 16));
    HIPOK(hipMalloc((void**)&d_rad2, nq * 16));
    HIPOK(hipMemset(d_rad, 0, nq * 16));
    HIPOK(hipMemset(d_rad2, 0, nq * 16));
    pm.Gather(g, gr, qc, d_rad);
    // direct: tree over the merged positions, records of the scaled powers and the
    // directions the map reconstructs from the angles
    float4* d_dirs = nullptr;
    HIPOK(hipMalloc((void**)&d_dirs, nvalid * 16));
    ABIOK(mnrt_photons_spherical_to_dir(nvalid, (const float*)pm.GetPhotons().d_positions,
                                        (const float*)pm.GetPhotons().d_powers, (float*)d_dirs, nullptr));
    mnrt_build_params pp{};
    pp.empty_space_ratio = 0.1f;
    pp.small_node_max = 32;
    pp.max_query_radius = rmax;
    pp.has_root_box = 1;
    for (int a = 0; a < 3; ++a) {
        pp.root_min[a] = -1.0f;
        pp.root_max[a] = 1.0f;
    }
    mnrt_tree pt{};
    ABIOK(mnrt_build_points((const float*)pm.GetPhotons().d_positions, nvalid, &pp, &pt, nullptr));
    ABIOK(mnrt_tree_attach_photon_records(&pt, (const float*)pm.GetPhotons().d_positions,
                                          (const float*)pm.GetPhotons().d_powers, (const float*)d_dirs, nullptr));
    ABIOK(mnrt_knn_gather(&pt, 40, rmax, nq, (const float*)qp, (const float*)qn, (const float*)qd, (const float*)qi,
                          (const float*)qc, nullptr, nullptr, nullptr, (float*)d_rad2, nullptr, nullptr, nullptr,
                          nullptr, nullptr, nullptr));
    HIPOK(hipDeviceSynchronize());
    {
        auto a = download(d_rad, nq), b = download(d_rad2, nq);
        bool nonzero = false;
        for (auto& r : a) nonzero = nonzero || r.x != 0.0f;
        expect(nq > 0 && nonzero && same_bytes(a.data(), b.data(), nq * 16), "PhotonMap::Gather == mnrt_knn_gather");
    }
    // PMSetPhotonMap + KernelPMPhotonGather with per-query radii from ComputeQueryRadii
    float* d_qr = nullptr;
    HIPOK(hipMalloc((void**)&d_qr, nq * 4));
    KDTreePoint* tp = const_cast<KDTreePoint*>(pm.GetTree());
    tp->SetKNNTargetCount(40);
    tp->ComputeQueryRadii(qp, nq, d_qr);
    HIPOK(hipMemset(d_rad2, 0, nq * 16));
    PMSetPhotonMap(*tp->GetData(), pm.GetPhotons(), 40);
    KernelPMPhotonGather(g, gr, d_qr, qc, nullptr, d_rad2, nullptr, false);
    PMUnsetPhotonMap();
    HIPOK(hipDeviceSynchronize());
    float4* d_rad3 = nullptr;
    HIPOK(hipMalloc((void**)&d_rad3, nq * 16));
    HIPOK(hipMemset(d_rad3, 0, nq * 16));
    ABIOK(mnrt_knn_gather(&pt, 40, rmax, nq, (const float*)qp, (const float*)qn, (const float*)qd, (const float*)qi,
                          (const float*)qc, nullptr, d_qr, nullptr, (float*)d_rad3, nullptr, nullptr, nullptr,
                          nullptr, nullptr, nullptr));
    HIPOK(hipDeviceSynchronize());
    {
        auto a = download(d_rad2, nq), b = download(d_rad3, nq);
        auto r = download(d_qr, nq);
        bool capped = true;
        for (float x : r) capped = capped && x > 0.0f && x <= rmax;
        expect(capped && same_bytes(a.data(), b.data(), nq * 16),
               "KernelPMPhotonGather (per-query caps from ComputeQueryRadii) == mnrt_knn_gather");
    }
    (void)mnrt_tree_release(&pt);
    (void)mnrt_tree_release(&direct);
    std::printf("%s\n", g_fail ? "host classes: FAILED" : "host classes: all checks passed");
    return g_fail;
}
