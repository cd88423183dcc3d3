// oracle.cpp — CPU restatement of the MNRT kd-tree hot path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load this library, and only as the checker / CPU baseline.  The
// product (mnrt_amd/) never links, loads or calls it.
//
// It restates, with its own code, the algorithms of the reference (jonike/mnrt):
//   build      KDTreeGPU.cpp:175-1209, kdtree.cu:467-1631, kdtree_tri.cu:190-489,
//              KDTreeTri.cpp:46-91, KDTreePoint.cpp:120-134
//   traversal  raytracing.cu:503-645, intersect_dev.h:77-164
//   kNN        photon_knn.cu:113-235 (range search, replaced by exact kNN),
//              photon_knn.cu:487-552 (gather / radiance)
// with the deterministic rules that replace the reference's random and racy choices:
//   * elements lying in a split plane go left for an even element index, right for an
//     odd one (the reference draws Mersenne-Twister uniforms, kdtree.cu:788-815,1226-1232);
//   * the best small-node split is the lowest candidate index among equal costs
//     (the reference races on s_idxMin, kdtree.cu:1458-1463);
//   * closest hits tie-break equal lambda to the lower triangle index (the reference
//     keeps the first found, raytracing.cu:627);
//   * kNN sets are ordered by (d2, photon index);
//   * the fp32 photon power sums are exact fixed-point sums (fixed_sum), so they do not
//     depend on the order of the set.
// Every deviation is listed in the repository's design notes.
//
// Floating point: compiled with -ffp-contract=off and without fast-math so that every
// fp32 operation rounds exactly as the HIP kernels' operations (same order, IEEE
// division and square root).  An fp64 path evaluates the radiance estimate for the
// tolerance checks.
//
// Parity pins (tests/test_oracle_*.py): closed-form intersections, brute-force closest
// hits over all triangles, brute-force sorted kNN, closed-form radiance cases, build
// invariants (every triangle referenced by every leaf its geometry overlaps, in fp64),
// SAH/VVH cost on hand-computed trees, oct encoding round trips.

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

namespace {

constexpr uint32_t kLeafTag = 3u;
constexpr uint32_t kIndexMask = 0x3FFFFFFFu;
constexpr uint32_t kMaxLevel = 48u;          // mirrors MNRT_MAX_LEVEL
constexpr float kCostTraverse = 3.0f;        // KD_COST_TRAVERSE, kdtree.cu:51
constexpr float kPi = 3.14159265358979323846f;     // MN_PI, MNMath.h:53
constexpr float kInvPi = 0.31830988618379067154f;  // MN_INV_PI, MNMath.h:55
constexpr float kInf = FLT_MAX;              // MN_INFINITY, MNMath.h:60
// Relative tolerance of the traversal's interval decisions (see trace_one).
constexpr float kTravTolLo = 1.0f - 1.0f / 65536.0f;
constexpr float kTravTolHi = 1.0f + 1.0f / 65536.0f;

inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

struct V3 { float v[3]; };
inline V3 mk(float x, float y, float z) { V3 r; r.v[0] = x; r.v[1] = y; r.v[2] = z; return r; }
inline V3 sub(const V3& a, const V3& b) { return mk(a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]); }
inline float dot(const V3& a, const V3& b) { return a.v[0] * b.v[0] + a.v[1] * b.v[1] + a.v[2] * b.v[2]; }
inline V3 cross(const V3& a, const V3& b) {
    return mk(a.v[1] * b.v[2] - a.v[2] * b.v[1], a.v[2] * b.v[0] - a.v[0] * b.v[2],
              a.v[0] * b.v[1] - a.v[1] * b.v[0]);
}
inline V3 ld4(const float* p, size_t i) { return mk(p[4 * i], p[4 * i + 1], p[4 * i + 2]); }

void parallel_for(uint32_t n, int nthreads, const std::function<void(uint32_t, uint32_t)>& fn) {
    if (nthreads <= 1 || n < 2) { fn(0, n); return; }
    std::vector<std::thread> th;
    uint32_t chunk = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        uint32_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back(fn, b, e);
    }
    for (auto& t : th) t.join();
}

// ---------------------------------------------------------------------------------
// Intersection tests (intersect_dev.h:77-99 and 124-164)
// ---------------------------------------------------------------------------------

// Moller-Trumbore as in dev_RayTriIntersect: det == 0 misses, inclusive barycentric
// edges, lambda may be negative.  e1 = b - a and e2 = c - a.
inline bool ray_tri(const V3& a, const V3& e1, const V3& e2, const V3& o, const V3& d,
                    float& lam, float& u, float& v) {
    V3 pvec = cross(d, e2);
    float det = dot(e1, pvec);
    if (det == 0.0f) return false;
    float inv_det = 1.0f / det;
    V3 tvec = sub(o, a);
    u = dot(tvec, pvec) * inv_det;
    V3 qvec = cross(tvec, e1);
    v = dot(d, qvec) * inv_det;
    lam = dot(e2, qvec) * inv_det;
    return u >= 0.0f && v >= 0.0f && (u + v) <= 1.0f;
}

// Slab test as in dev_RayBoxIntersect: no early out, interval narrowed per axis.
inline bool ray_box(const float* bmin, const float* bmax, const V3& o, const V3& inv, float t0,
                    float t1, float& tn, float& tf) {
    bool hit = true;
    for (int i = 0; i < 3; ++i) {
        float v1 = (bmin[i] - o.v[i]) * inv.v[i];
        float v2 = (bmax[i] - o.v[i]) * inv.v[i];
        float tnear = v1, tfar = v2;
        if (v1 > v2) tnear = v2;
        if (v1 > v2) tfar = v1;
        t0 = (tnear > t0) ? tnear : t0;
        t1 = (tfar < t1) ? tfar : t1;
        if (t0 > t1) hit = false;
    }
    tn = t0;
    tf = t1;
    return hit;
}

// ---------------------------------------------------------------------------------
// Octahedral direction code (photon records).  The reference stores the photon
// direction as spherical angles and rebuilds it with sinf/cosf (photon_dev.h:71-116,
// mncudautil_dev.h:345-375); records here keep two 16-bit octahedral coordinates so the
// kNN kernel needs no trigonometry.  Only the sign of dot(normal, direction) is used,
// so decoding does not normalise.
// ---------------------------------------------------------------------------------
inline float sgn1(float x) { return x >= 0.0f ? 1.0f : -1.0f; }

uint32_t oct_encode(float x, float y, float z) {
    float s = std::fabs(x) + std::fabs(y) + std::fabs(z);
    float u = x / s, v = y / s;
    if (z < 0.0f) {
        float ou = u, ov = v;
        u = (1.0f - std::fabs(ov)) * sgn1(ou);
        v = (1.0f - std::fabs(ou)) * sgn1(ov);
    }
    float qu = std::rint(u * 32767.0f), qv = std::rint(v * 32767.0f);
    qu = std::min(32767.0f, std::max(-32767.0f, qu));
    qv = std::min(32767.0f, std::max(-32767.0f, qv));
    int32_t iu = (int32_t)qu, iv = (int32_t)qv;
    return ((uint32_t)(uint16_t)(int16_t)iu) | (((uint32_t)(uint16_t)(int16_t)iv) << 16);
}

V3 oct_decode(uint32_t c) {
    int16_t iu = (int16_t)(uint16_t)(c & 0xFFFFu), iv = (int16_t)(uint16_t)(c >> 16);
    float u = (float)iu / 32767.0f, v = (float)iv / 32767.0f;
    float z = (1.0f - std::fabs(u)) - std::fabs(v);
    if (z < 0.0f) {
        float ou = u;
        u = (1.0f - std::fabs(v)) * sgn1(ou);
        v = (1.0f - std::fabs(ou)) * sgn1(v);
    }
    return mk(u, v, z);
}

// ---------------------------------------------------------------------------------
// kd-tree build
// ---------------------------------------------------------------------------------

struct Elem {
    uint32_t idx;
    float lo[3], hi[3];
};

struct BNode {
    bool leaf = true;
    int axis = 0;
    float split = 0.0f;
    int left = -1, right = -1;
    std::vector<uint32_t> elems;  // leaf only: original element indices, ascending
};

struct BuildCtx {
    int kind;  // 0 tri, 1 point
    const float* v0;
    const float* v1;
    const float* v2;  // tri vertices (float4) or points (v0)
    float ratio;
    uint32_t small_max;
    float max_qr;
    std::vector<BNode> T;
    uint32_t max_level = 0;
    // element-stream counts for the build's algorithmic bytes: elements of active large
    // nodes per level, child element records written, triangle clips (one per straddling
    // triangle and side), elements entering the small stage
    uint64_t large_elems = 0, child_elems = 0, clips = 0, small_elems = 0;
};

// Deterministic side of an element lying in the split plane (replaces the MT uniforms
// of kdtree.cu:788-815 and 1226-1232).
inline bool plane_goes_left(uint32_t idx) { return (idx & 1u) == 0u; }

// -0 -> +0, so min/max results and split words never depend on the sign of zero.
inline float canon(float x) { return x + 0.0f; }

// Classification of a large-node element (kernel_MarkLeftRightElements, kdtree.cu:752-824).
inline void classify(int kind, const Elem& e, int axis, float s, bool& L, bool& R) {
    if (kind == 0) {
        float mn = e.lo[axis], mx = e.hi[axis];
        bool planar_on = (mn == s && mn == mx);
        L = mn < s || (planar_on && plane_goes_left(e.idx));
        R = s < mx || (planar_on && !plane_goes_left(e.idx));
    } else {
        float p = e.lo[axis];
        if (plane_goes_left(e.idx)) { L = p <= s; R = s < p; }
        else { L = p < s; R = s <= p; }
    }
}

// Sutherland-Hodgman clipping (dev_ClipSegment / dev_ClipSegments, kdtree_tri.cu:190-257).
struct PList { int n = 0; V3 p[12]; };

void clip_segments(const PList& in, int axis, float val, float sense, PList& out) {
    out.n = 0;
    for (int i = 0; i < in.n; ++i) {
        const V3& a = in.p[i];
        const V3& b = in.p[(i + 1 == in.n) ? 0 : i + 1];
        bool ins = (a.v[axis] - val) * sense >= 0.0f;
        bool ine = (b.v[axis] - val) * sense >= 0.0f;
        if (ins != ine) {
            float t = (val - a.v[axis]) / (b.v[axis] - a.v[axis]);
            V3 h = mk(a.v[0] + t * (b.v[0] - a.v[0]), a.v[1] + t * (b.v[1] - a.v[1]),
                      a.v[2] + t * (b.v[2] - a.v[2]));
            h.v[axis] = val;
            out.p[out.n++] = h;
        }
        if (ine) out.p[out.n++] = b;
    }
}

// Clipped bounds of a straddling triangle in a child box (dev_ClipTriToBounds and
// kernel_PerformSplitClipping, kdtree_tri.cu:283-489), rounded outward: the clipped
// polygon's bounds are padded by 2^-20 of the largest magnitude on each axis, then
// clamped to the child box and to the element's previous bounds.  The reference clamps
// only the split axis and can lose a triangle by rounding; this keeps every triangle in
// every child its geometry reaches.
void clip_bounds(const BuildCtx& C, Elem& e, const float* bmin, const float* bmax) {
    PList a, b;
    a.n = 3;
    a.p[0] = ld4(C.v0, e.idx);
    a.p[1] = ld4(C.v1, e.idx);
    a.p[2] = ld4(C.v2, e.idx);
    for (int ax = 0; ax < 3; ++ax) {
        if (e.lo[ax] != e.hi[ax]) {
            clip_segments(a, ax, bmin[ax], 1.0f, b);
            clip_segments(b, ax, bmax[ax], -1.0f, a);
        }
    }
    for (int ax = 0; ax < 3; ++ax) {
        float lo, hi;
        if (a.n == 0) {
            lo = e.lo[ax];
            hi = e.hi[ax];
        } else {
            lo = kInf;
            hi = -kInf;
            for (int i = 0; i < a.n; ++i) {
                lo = std::fmin(a.p[i].v[ax], lo);
                hi = std::fmax(a.p[i].v[ax], hi);
            }
            float m = std::fmax(std::fmax(std::fabs(lo), std::fabs(hi)),
                                std::fmax(std::fabs(bmin[ax]), std::fabs(bmax[ax])));
            float pad = m * (1.0f / 1048576.0f);
            lo = lo - pad;
            hi = hi + pad;
        }
        lo = std::fmax(lo, bmin[ax]);
        hi = std::fmin(hi, bmax[ax]);
        lo = std::fmax(lo, e.lo[ax]);
        hi = std::fmin(hi, e.hi[ax]);
        e.lo[ax] = canon(lo);
        e.hi[ax] = canon(hi);
    }
}

int new_node(BuildCtx& C) {
    C.T.emplace_back();
    return (int)C.T.size() - 1;
}

void make_leaf(BuildCtx& C, int id, const std::vector<Elem>& el) {
    BNode& n = C.T[id];
    n.leaf = true;
    n.elems.clear();
    for (const Elem& e : el) n.elems.push_back(e.idx);
}

// Small-node stage for one small root (PreProcessSmallNodes/ProcessSmallNodes,
// KDTreeGPU.cpp:945-1154; kernel_CreateSplitCandidates, kernel_InitSplitMasks,
// kernel_FindBestSplits, kernel_SplitSmallNodes, kernel_GenerateENAFromMasks,
// kdtree.cu:1035-1631).  Candidates are the element points, axis-major, minimums before
// maximums; 64-bit left/right masks are fixed per small root; each node takes the split
// with the lowest SAH (triangles) or VVH (points) cost, lowest candidate index on ties,
// among candidates strictly inside its box, and splits iff that cost < its element count.
struct SmallRoot {
    std::vector<Elem> el;
    float bmin[3], bmax[3];
    uint32_t level;
    int id;
};

struct Cand {
    float pos;
    int axis;
    uint64_t mL, mR;
};

void small_recurse(BuildCtx& C, const SmallRoot& R, const std::vector<Cand>& cand, int id,
                   uint64_t mask, const float* bmin, const float* bmax, uint32_t level) {
    C.max_level = std::max(C.max_level, level);
    uint32_t cnt = (uint32_t)__builtin_popcountll(mask);
    int best = -1;
    float best_cost = kInf;
    if (cnt > 0 && level < kMaxLevel) {
        float ext[3] = {bmax[0] - bmin[0], bmax[1] - bmin[1], bmax[2] - bmin[2]};
        float denom;
        float qr2 = 2.0f * C.max_qr;
        if (C.kind == 1) {
            float vol = 1.0f;
            for (int i = 0; i < 3; ++i) vol *= ext[i] + qr2;
            denom = 1.0f / vol;
        } else {
            float area = 2.0f * (ext[0] * ext[1] + ext[0] * ext[2] + ext[1] * ext[2]);
            denom = 1.0f / area;
        }
        for (size_t c = 0; c < cand.size(); ++c) {
            const Cand& k = cand[c];
            int a = k.axis;
            float s = k.pos;
            if (!(bmin[a] < s && s < bmax[a])) continue;
            uint32_t nl = (uint32_t)__builtin_popcountll(k.mL & mask);
            uint32_t nr = (uint32_t)__builtin_popcountll(k.mR & mask);
            if (nl == 0 || nr == 0) continue;
            int o1 = (a + 1) % 3, o2 = (a + 2) % 3;
            float cost;
            if (C.kind == 1) {
                float vl = (ext[o1] + qr2) * (s - bmin[a] + qr2) * (ext[o2] + qr2);
                float vr = (ext[o1] + qr2) * (bmax[a] - s + qr2) * (ext[o2] + qr2);
                cost = ((float)nl * vl + (float)nr * vr) * denom + kCostTraverse;
            } else {
                float al = 2.0f * (ext[o1] * ext[o2] + (s - bmin[a]) * (ext[o1] + ext[o2]));
                float ar = 2.0f * (ext[o1] * ext[o2] + (bmax[a] - s) * (ext[o1] + ext[o2]));
                cost = ((float)nl * al + (float)nr * ar) * denom + kCostTraverse;
            }
            if (cost < best_cost) {  // strict: lowest index wins ties; NaN never wins
                best_cost = cost;
                best = (int)c;
            }
        }
    }
    if (best < 0 || !(best_cost < (float)cnt)) {
        BNode& n = C.T[id];
        n.leaf = true;
        n.elems.clear();
        for (size_t i = 0; i < R.el.size(); ++i)
            if (mask >> i & 1ull) n.elems.push_back(R.el[i].idx);
        return;
    }
    const Cand& k = cand[best];
    int l = new_node(C), r = new_node(C);
    BNode& n = C.T[id];
    n.leaf = false;
    n.axis = k.axis;
    n.split = k.pos;
    n.left = l;
    n.right = r;
    float lmax[3] = {bmax[0], bmax[1], bmax[2]};
    lmax[k.axis] = k.pos;
    float rmin[3] = {bmin[0], bmin[1], bmin[2]};
    rmin[k.axis] = k.pos;
    uint64_t ml = k.mL & mask, mr = k.mR & mask;
    small_recurse(C, R, cand, l, ml, bmin, lmax, level + 1);
    small_recurse(C, R, cand, r, mr, rmin, bmax, level + 1);
}

void small_stage(BuildCtx& C, const SmallRoot& R) {
    uint32_t n = (uint32_t)R.el.size();
    C.small_elems += n;
    std::vector<Cand> cand;
    int passes = (C.kind == 0) ? 2 : 1;
    for (int isMax = 0; isMax < passes; ++isMax) {
        for (int axis = 0; axis < 3; ++axis) {
            for (uint32_t i = 0; i < n; ++i) {
                Cand c;
                c.axis = axis;
                c.pos = isMax ? R.el[i].hi[axis] : R.el[i].lo[axis];
                c.mL = c.mR = 0;
                for (uint32_t j = 0; j < n; ++j) {
                    const Elem& e = R.el[j];
                    bool L, Rr;
                    if (C.kind == 0) {
                        float mn = e.lo[axis], mx = e.hi[axis];
                        L = mn < c.pos;
                        Rr = c.pos < mx;
                        if (j == i) {
                            if (isMax) L = true;
                            else Rr = true;
                        } else if (mn == mx && mn == c.pos) {
                            if (plane_goes_left(e.idx)) L = true;
                            else Rr = true;
                        }
                    } else {
                        float p = e.lo[axis];
                        if (plane_goes_left(e.idx)) { L = p <= c.pos; Rr = c.pos < p; }
                        else { L = p < c.pos; Rr = c.pos <= p; }
                    }
                    if (L) c.mL |= 1ull << j;
                    if (Rr) c.mR |= 1ull << j;
                }
                cand.push_back(c);
            }
        }
    }
    uint64_t all = (n == 64) ? ~0ull : ((1ull << n) - 1ull);
    small_recurse(C, R, cand, R.id, all, R.bmin, R.bmax, R.level);
}

struct Active {
    float inh_lo[3], inh_hi[3];
    std::vector<Elem> el;
    uint32_t level;
    int id;
};

// Large-node stage (LargeNodeStage/ProcessLargeNodes, KDTreeGPU.cpp:247-782; empty space
// cutting kdtree.cu:467-612 and 1986-2033; median split kdtree.cu:629-682).  A node whose
// median split would put every element into both children becomes a leaf (the
// reference's level loop never ends there: the node comes back unchanged every level).
//
// Level-parallel (the reference's kernels are level-parallel too): the active nodes'
// elements are cut into chunks; the tight boxes (min / max: exact in any order) and the
// classification with clipping run in parallel over chunks, the chunk outputs are joined
// in element order, and the per-node decisions and node allocation run in node order.
// The tree is the same for every thread count.
constexpr uint32_t kLargeChunk = 16384;

void large_stage(BuildCtx& C, std::vector<Active> active, std::vector<SmallRoot>& smalls, int nthreads) {
    struct Chunk {
        uint32_t node, b, e;
    };
    while (!active.empty()) {
        std::vector<Chunk> chunks;
        for (uint32_t i = 0; i < active.size(); ++i)
            for (uint32_t b = 0; b < active[i].el.size(); b += kLargeChunk)
                chunks.push_back({i, b, (uint32_t)std::min<size_t>(active[i].el.size(), (size_t)b + kLargeChunk)});
        // tight boxes, per chunk then per node
        std::vector<float> part(6 * chunks.size());
        parallel_for((uint32_t)chunks.size(), nthreads, [&](uint32_t cb, uint32_t ce) {
            for (uint32_t c = cb; c < ce; ++c) {
                float* t = &part[6 * c];
                for (int a = 0; a < 3; ++a) {
                    t[a] = kInf;
                    t[3 + a] = -kInf;
                }
                const Active& A = active[chunks[c].node];
                for (uint32_t k = chunks[c].b; k < chunks[c].e; ++k)
                    for (int a = 0; a < 3; ++a) {
                        t[a] = std::fmin(t[a], A.el[k].lo[a]);
                        t[3 + a] = std::fmax(t[3 + a], A.el[k].hi[a]);
                    }
            }
        });
        std::vector<float> tight(6 * active.size());
        for (size_t i = 0; i < active.size(); ++i)
            for (int a = 0; a < 3; ++a) {
                tight[6 * i + a] = kInf;
                tight[6 * i + 3 + a] = -kInf;
            }
        for (size_t c = 0; c < chunks.size(); ++c)
            for (int a = 0; a < 3; ++a) {
                float* t = &tight[6 * chunks[c].node];
                t[a] = std::fmin(t[a], part[6 * c + a]);
                t[3 + a] = std::fmax(t[3 + a], part[6 * c + 3 + a]);
            }
        // per node: empty-space cuts, level cap, the median split and the child boxes
        std::vector<int> cur(active.size()), axis(active.size(), -1);
        std::vector<float> split(active.size());
        std::vector<Active> Lc(active.size()), Rc(active.size());
        for (size_t i = 0; i < active.size(); ++i) {
            Active& A = active[i];
            C.max_level = std::max(C.max_level, A.level);
            const float* tlo = &tight[6 * i];
            const float* thi = &tight[6 * i + 3];
            int c = A.id;
            for (int isMax = 0; isMax < 2; ++isMax) {
                for (int ax = 0; ax < 3; ++ax) {
                    if (A.level >= kMaxLevel) continue;
                    float total = A.inh_hi[ax] - A.inh_lo[ax];
                    float empty = isMax ? (A.inh_hi[ax] - thi[ax]) : (tlo[ax] - A.inh_lo[ax]);
                    if (!(empty > total * C.ratio)) continue;
                    float sp = isMax ? thi[ax] : tlo[ax];
                    int eleaf = new_node(C), cont = new_node(C);
                    BNode& n = C.T[c];
                    n.leaf = false;
                    n.axis = ax;
                    n.split = sp;
                    n.left = isMax ? cont : eleaf;
                    n.right = isMax ? eleaf : cont;
                    if (isMax) A.inh_hi[ax] = sp;
                    else A.inh_lo[ax] = sp;
                    A.level += 1;
                    C.max_level = std::max(C.max_level, A.level);
                    c = cont;
                }
            }
            cur[i] = c;
            if (A.level >= kMaxLevel) {
                make_leaf(C, c, A.el);
                continue;
            }
            C.large_elems += A.el.size();
            float ex = A.inh_hi[0] - A.inh_lo[0], ey = A.inh_hi[1] - A.inh_lo[1],
                  ez = A.inh_hi[2] - A.inh_lo[2];
            // Longest axis, ties to the lower axis.  The reference's test
            // (kdtree.cu:641-647) falls back to x whenever y and z tie for longest even
            // if x is shorter, which slices cube-like scenes into ever thinner x slabs.
            int ax = 0;
            float emax = ex;
            if (ey > emax) { ax = 1; emax = ey; }
            if (ez > emax) ax = 2;
            axis[i] = ax;
            split[i] = A.inh_lo[ax] + 0.5f * (A.inh_hi[ax] - A.inh_lo[ax]);
            for (int a = 0; a < 3; ++a) {
                Lc[i].inh_lo[a] = Rc[i].inh_lo[a] = A.inh_lo[a];
                Lc[i].inh_hi[a] = Rc[i].inh_hi[a] = A.inh_hi[a];
            }
            Lc[i].inh_hi[ax] = split[i];
            Rc[i].inh_lo[ax] = split[i];
        }
        // classification and clipping, per chunk, joined in element order
        std::vector<std::vector<Elem>> lo_part(chunks.size()), hi_part(chunks.size());
        std::vector<uint64_t> clips(chunks.size(), 0);
        parallel_for((uint32_t)chunks.size(), nthreads, [&](uint32_t cb, uint32_t ce) {
            for (uint32_t c = cb; c < ce; ++c) {
                const uint32_t i = chunks[c].node;
                if (axis[i] < 0) continue;
                const Active& A = active[i];
                const int ax = axis[i];
                const float sp = split[i];
                for (uint32_t k = chunks[c].b; k < chunks[c].e; ++k) {
                    const Elem& e = A.el[k];
                    bool L, R;
                    classify(C.kind, e, ax, sp, L, R);
                    if (C.kind == 0 && (e.lo[ax] < sp && sp < e.hi[ax])) {
                        clips[c] += 2;
                        Elem el = e, er = e;
                        clip_bounds(C, el, Lc[i].inh_lo, Lc[i].inh_hi);
                        clip_bounds(C, er, Rc[i].inh_lo, Rc[i].inh_hi);
                        lo_part[c].push_back(el);
                        hi_part[c].push_back(er);
                    } else {
                        if (L) lo_part[c].push_back(e);
                        if (R) hi_part[c].push_back(e);
                    }
                }
            }
        });
        for (size_t c = 0; c < chunks.size(); ++c) {
            const uint32_t i = chunks[c].node;
            if (axis[i] < 0) continue;
            C.clips += clips[c];
            Lc[i].el.insert(Lc[i].el.end(), lo_part[c].begin(), lo_part[c].end());
            Rc[i].el.insert(Rc[i].el.end(), hi_part[c].begin(), hi_part[c].end());
            std::vector<Elem>().swap(lo_part[c]);
            std::vector<Elem>().swap(hi_part[c]);
        }
        // per node: no-progress leaf, or two children (small roots or next level)
        std::vector<Active> next;
        for (size_t i = 0; i < active.size(); ++i) {
            if (axis[i] < 0) continue;
            Active& A = active[i];
            if (Lc[i].el.size() == A.el.size() && Rc[i].el.size() == A.el.size()) {
                make_leaf(C, cur[i], A.el);
                continue;
            }
            C.child_elems += Lc[i].el.size() + Rc[i].el.size();
            int l = new_node(C), r = new_node(C);
            BNode& n = C.T[cur[i]];
            n.leaf = false;
            n.axis = axis[i];
            n.split = split[i];
            n.left = l;
            n.right = r;
            Lc[i].level = Rc[i].level = A.level + 1;
            C.max_level = std::max(C.max_level, Lc[i].level);
            Lc[i].id = l;
            Rc[i].id = r;
            for (Active* ch : {&Lc[i], &Rc[i]}) {
                if (ch->el.size() <= C.small_max) {
                    SmallRoot sr;
                    sr.el = std::move(ch->el);
                    for (int a = 0; a < 3; ++a) {
                        sr.bmin[a] = ch->inh_lo[a];
                        sr.bmax[a] = ch->inh_hi[a];
                    }
                    sr.level = ch->level;
                    sr.id = ch->id;
                    smalls.push_back(std::move(sr));
                } else {
                    next.push_back(std::move(*ch));
                }
            }
            std::vector<Elem>().swap(A.el);
        }
        active.swap(next);
    }
}

// Small stage of every small root, in parallel: each subtree is built into a node array
// of its own (local node 0 = the small root) and spliced into the tree in small-root
// order.
void small_stages(BuildCtx& C, const std::vector<SmallRoot>& smalls, int nthreads) {
    std::vector<BuildCtx> sub(smalls.size());
    parallel_for((uint32_t)smalls.size(), nthreads, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; ++i) {
            BuildCtx& L = sub[i];
            L.kind = C.kind;
            L.v0 = C.v0;
            L.v1 = C.v1;
            L.v2 = C.v2;
            L.ratio = C.ratio;
            L.small_max = C.small_max;
            L.max_qr = C.max_qr;
            L.T.emplace_back();
            SmallRoot R = smalls[i];
            R.id = 0;
            small_stage(L, R);
        }
    });
    for (size_t i = 0; i < smalls.size(); ++i) {
        BuildCtx& L = sub[i];
        const int off = (int)C.T.size() - 1;  // local node j >= 1 -> off + j
        auto map = [&](int j) { return j == 0 ? smalls[i].id : off + j; };
        for (size_t j = 1; j < L.T.size(); ++j) {
            BNode n = std::move(L.T[j]);
            if (!n.leaf) {
                n.left = map(n.left);
                n.right = map(n.right);
            }
            C.T.push_back(std::move(n));
        }
        BNode r = std::move(L.T[0]);
        if (!r.leaf) {
            r.left = map(r.left);
            r.right = map(r.right);
        }
        C.T[smalls[i].id] = std::move(r);
        C.max_level = std::max(C.max_level, L.max_level);
        C.small_elems += L.small_elems;
        BuildCtx().T.swap(L.T);
    }
}

struct OTree {
    int kind;
    std::vector<uint32_t> nodes;
    std::vector<uint32_t> refs;
    float root[6];
    uint32_t max_level;
    uint32_t num_leaves;
    uint32_t n_elems;
    uint64_t counts[4];  // BuildCtx large_elems, child_elems, clips, small_elems
};

uint32_t emit(const BuildCtx& C, int id, OTree& T) {
    uint32_t at = (uint32_t)(T.nodes.size() / 2);
    T.nodes.push_back(0);
    T.nodes.push_back(0);
    const BNode& n = C.T[id];
    if (n.leaf) {
        T.nodes[2 * at] = (uint32_t)n.elems.size();
        T.nodes[2 * at + 1] = (kLeafTag << 30) | (uint32_t)T.refs.size();
        for (uint32_t e : n.elems) T.refs.push_back(e);
        T.num_leaves++;
    } else {
        emit(C, n.left, T);
        uint32_t r = emit(C, n.right, T);
        T.nodes[2 * at] = fbits(n.split);
        T.nodes[2 * at + 1] = ((uint32_t)n.axis << 30) | r;
    }
    return at;
}

OTree* build(int kind, const float* v0, const float* v1, const float* v2, uint32_t n, float ratio,
             uint32_t small_max, float max_qr, int has_root, const float* root6, int nthreads) {
    BuildCtx C;
    C.kind = kind;
    C.v0 = v0;
    C.v1 = v1;
    C.v2 = v2;
    C.ratio = ratio;
    C.small_max = small_max;
    C.max_qr = max_qr;
    Active root;
    root.level = 0;
    root.el.resize(n);
    float lo[3] = {kInf, kInf, kInf}, hi[3] = {-kInf, -kInf, -kInf};
    for (uint32_t i = 0; i < n; ++i) {
        Elem& e = root.el[i];
        e.idx = i;
        V3 a = ld4(v0, i);
        if (kind == 0) {
            V3 b = ld4(v1, i), c = ld4(v2, i);
            for (int k = 0; k < 3; ++k) {
                e.lo[k] = canon(std::fmin(a.v[k], std::fmin(b.v[k], c.v[k])));
                e.hi[k] = canon(std::fmax(a.v[k], std::fmax(b.v[k], c.v[k])));
            }
        } else {
            for (int k = 0; k < 3; ++k) e.lo[k] = e.hi[k] = canon(a.v[k]);
        }
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::fmin(lo[k], e.lo[k]);
            hi[k] = std::fmax(hi[k], e.hi[k]);
        }
    }
    if (has_root) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = canon(root6[k]);
            hi[k] = canon(root6[3 + k]);
        }
    }
    for (int k = 0; k < 3; ++k) {
        root.inh_lo[k] = lo[k];
        root.inh_hi[k] = hi[k];
    }
    root.id = new_node(C);
    std::vector<SmallRoot> smalls;
    if (n <= small_max) {
        SmallRoot sr;
        sr.el = root.el;
        for (int k = 0; k < 3; ++k) {
            sr.bmin[k] = lo[k];
            sr.bmax[k] = hi[k];
        }
        sr.level = 0;
        sr.id = root.id;
        smalls.push_back(std::move(sr));
    } else {
        large_stage(C, {root}, smalls, nthreads);
    }
    small_stages(C, smalls, nthreads);
    OTree* T = new OTree();
    T->kind = kind;
    T->num_leaves = 0;
    T->n_elems = n;
    for (int k = 0; k < 3; ++k) {
        T->root[k] = lo[k];
        T->root[3 + k] = hi[k];
    }
    emit(C, root.id, *T);
    T->max_level = C.max_level;
    T->counts[0] = C.large_elems;
    T->counts[1] = C.child_elems;
    T->counts[2] = C.clips;
    T->counts[3] = C.small_elems;
    return T;
}

// ---------------------------------------------------------------------------------
// Traversal (dev_FindNextIntersectionKDWhileWhile<true>, raytracing.cu:503-645).
// Interval decisions are widened by a relative 2^-16 so that rounding of the split
// parameter never hides a hit that the brute force over all triangles finds:
//   NaN split parameter (origin in the plane, direction parallel): both children;
//   tsplit < 0 or tsplit > tmax*(1+tol): near child only;
//   tsplit < tmin*(1-tol): far child only;
//   otherwise both, near on [tmin, min(tmax, tsplit*(1+tol))], far on
//   [max(tmin, tsplit*(1-tol)), tmax];
// the near side is where the origin lies, an origin in the plane taking the side the
// direction leaves (the reference uses |o - split| < 1e-8 and the direction,
// raytracing.cu:584-586).  Stacked nodes starting after the best hit are skipped.
// ---------------------------------------------------------------------------------

struct TraceCounts {
    uint64_t inner = 0, leaves = 0, refs = 0, tests = 0;
};

struct TriSoup {
    const float* v0;
    const float* v1;
    const float* v2;
};

inline void tri_test(const TriSoup& S, uint32_t t, const V3& o, const V3& d, float eps,
                     float& best_t, int32_t& best_tri, float& bu, float& bv) {
    V3 a = ld4(S.v0, t), b = ld4(S.v1, t), c = ld4(S.v2, t);
    V3 e1 = sub(b, a), e2 = sub(c, a);
    float lam, u, v;
    if (ray_tri(a, e1, e2, o, d, lam, u, v) && lam > eps) {
        if (lam < best_t || (lam == best_t && best_tri >= 0 && (int32_t)t < best_tri)) {
            best_t = lam;
            best_tri = (int32_t)t;
            bu = u;
            bv = v;
        }
    }
}

void trace_one(const uint32_t* nodes, const uint32_t* refs, const float* root, const TriSoup& S,
               const V3& o, const V3& d, float eps, int32_t& out_tri, float& out_t, float& ou,
               float& ov, TraceCounts& cnt) {
    V3 inv = mk(1.0f / d.v[0], 1.0f / d.v[1], 1.0f / d.v[2]);
    float t0, t1;
    out_tri = -1;
    out_t = 0.0f;
    ou = ov = 0.0f;
    if (!ray_box(root, root + 3, o, inv, eps, kInf, t0, t1)) return;
    t1 = t1 + eps;
    float best_t = t1, bu = 0.0f, bv = 0.0f;
    int32_t best_tri = -1;
    struct Ent { uint32_t node; float tmin, tmax; };
    std::vector<Ent> stack;
    stack.reserve(64);
    uint32_t node = 0;
    float tmin = t0, tmax = t1;
    for (;;) {
        for (;;) {
            uint32_t w1 = nodes[2 * node + 1];
            if ((w1 >> 30) == kLeafTag) break;
            cnt.inner++;
            int a = (int)(w1 >> 30);
            float s = bitsf(nodes[2 * node]);
            uint32_t left = node + 1, right = w1 & kIndexMask;
            float oa = o.v[a], da = d.v[a];
            float tsplit = (s - oa) * inv.v[a];
            bool below = (oa < s) || (oa == s && da > 0.0f);
            uint32_t first = below ? left : right, second = below ? right : left;
            if (tsplit != tsplit) {
                stack.push_back({second, tmin, tmax});
                node = first;
            } else if (tsplit < 0.0f || tsplit > tmax * kTravTolHi) {
                node = first;
            } else if (tsplit < tmin * kTravTolLo) {
                node = second;
            } else {
                float fmin = std::fmax(tmin, tsplit * kTravTolLo);
                stack.push_back({second, fmin, tmax});
                tmax = std::fmin(tmax, tsplit * kTravTolHi);
                node = first;
            }
        }
        cnt.leaves++;
        uint32_t count = nodes[2 * node], begin = nodes[2 * node + 1] & kIndexMask;
        for (uint32_t i = 0; i < count; ++i) {
            cnt.refs++;
            cnt.tests++;
            tri_test(S, refs[begin + i], o, d, eps, best_t, best_tri, bu, bv);
        }
        // Stacked nodes that start after the best hit are skipped.  (The reference's
        // `outLambda < tMax` early exit, raytracing.cu:637-641, ends the whole search;
        // a sibling pushed for a ray lying in the split plane covers the same interval
        // as the subtree searched first, so it may still hold an equal or closer hit.)
        bool found = false;
        while (!stack.empty()) {
            Ent e = stack.back();
            stack.pop_back();
            if (best_tri >= 0 && best_t < e.tmin) continue;
            node = e.node;
            tmin = e.tmin;
            tmax = e.tmax;
            found = true;
            break;
        }
        if (!found) break;
    }
    out_tri = best_tri;
    if (best_tri >= 0) {
        out_t = best_t;
        ou = bu;
        ov = bv;
    }
}

// ---------------------------------------------------------------------------------
// Any-hit segments (dev_FindNextIntersectionKDWhileWhile<false>, raytracing.cu:503-645,
// as the shadow-ray kernels use it, raytracing.cu:1019-1114).  A triangle occludes when
// it is hit at lambda with
//     lambda > eps,  lambda < min(t_exit, dist) + eps,  lambda + slack < dist,
// t_exit being where the ray leaves the root box.  The reference tests only the first
// hit found against the light distance; here any occluding triangle counts, so the
// tree walk must agree with the test over all triangles (any_brute).
// ---------------------------------------------------------------------------------

inline bool occludes(const TriSoup& S, uint32_t t, const V3& o, const V3& d, float eps, float tlim,
                     float dist, float slack) {
    V3 a = ld4(S.v0, t), b = ld4(S.v1, t), c = ld4(S.v2, t);
    float lam, u, v;
    return ray_tri(a, sub(b, a), sub(c, a), o, d, lam, u, v) && lam > eps && lam < tlim &&
           lam + slack < dist;
}

bool any_one(const uint32_t* nodes, const uint32_t* refs, const float* root, const TriSoup& S,
             const V3& o, const V3& d, float eps, float dist, float slack, TraceCounts& cnt) {
    V3 inv = mk(1.0f / d.v[0], 1.0f / d.v[1], 1.0f / d.v[2]);
    float t0, t1;
    if (!ray_box(root, root + 3, o, inv, eps, kInf, t0, t1)) return false;
    const float tlim = std::fmin(t1, dist) + eps;
    struct Ent { uint32_t node; float tmin, tmax; };
    std::vector<Ent> stack;
    stack.reserve(64);
    uint32_t node = 0;
    float tmin = t0, tmax = tlim;
    for (;;) {
        for (;;) {
            uint32_t w1 = nodes[2 * node + 1];
            if ((w1 >> 30) == kLeafTag) break;
            cnt.inner++;
            int a = (int)(w1 >> 30);
            float s = bitsf(nodes[2 * node]);
            uint32_t left = node + 1, right = w1 & kIndexMask;
            float oa = o.v[a], da = d.v[a];
            float tsplit = (s - oa) * inv.v[a];
            bool below = (oa < s) || (oa == s && da > 0.0f);
            uint32_t first = below ? left : right, second = below ? right : left;
            if (tsplit != tsplit) {
                stack.push_back({second, tmin, tmax});
                node = first;
            } else if (tsplit < 0.0f || tsplit > tmax * kTravTolHi) {
                node = first;
            } else if (tsplit < tmin * kTravTolLo) {
                node = second;
            } else {
                stack.push_back({second, std::fmax(tmin, tsplit * kTravTolLo), tmax});
                tmax = std::fmin(tmax, tsplit * kTravTolHi);
                node = first;
            }
        }
        cnt.leaves++;
        uint32_t count = nodes[2 * node], begin = nodes[2 * node + 1] & kIndexMask;
        for (uint32_t i = 0; i < count; ++i) {
            cnt.refs++;
            cnt.tests++;
            if (occludes(S, refs[begin + i], o, d, eps, tlim, dist, slack)) return true;
        }
        if (stack.empty()) return false;
        node = stack.back().node;
        tmin = stack.back().tmin;
        tmax = stack.back().tmax;
        stack.pop_back();
    }
}

bool any_brute(const float* root, const TriSoup& S, uint32_t ntri, const V3& o, const V3& d, float eps,
               float dist, float slack) {
    V3 inv = mk(1.0f / d.v[0], 1.0f / d.v[1], 1.0f / d.v[2]);
    float t0, t1;
    if (!ray_box(root, root + 3, o, inv, eps, kInf, t0, t1)) return false;
    const float tlim = std::fmin(t1, dist) + eps;
    for (uint32_t t = 0; t < ntri; ++t)
        if (occludes(S, t, o, d, eps, tlim, dist, slack)) return true;
    return false;
}

// Light (LightData, KernelDefs.h:307-323) as 16 floats: type, position[3], direction[3],
// L_emit[3], v1[3], v2[3].  radius is passed separately.
struct OLight {
    int type;
    V3 pos, dir, v1, v2;
    float radius;
};

inline OLight olight(const float* L16, float radius) {
    OLight L;
    L.type = (int)L16[0];
    L.pos = mk(L16[1], L16[2], L16[3]);
    L.dir = mk(L16[4], L16[5], L16[6]);
    L.v1 = mk(L16[10], L16[11], L16[12]);
    L.v2 = mk(L16[13], L16[14], L16[15]);
    L.radius = radius;
    return L;
}

// Shadow segment of kernel_TraceShadowRaysDelta / _Area (raytracing.cu:1019-1114).
// Returns 0 / 1 when the answer is fixed without tracing (area sample facing away or
// at zero distance: 0; point light at zero distance: 1), else -1 with the segment set.
int shadow_segment(const OLight& L, const V3& p, const float* samples4, size_t i, V3& d, float& dist,
                   float& slack, float eps) {
    slack = 0.0f;
    if (L.type == 1) {
        d = mk(-L.dir.v[0], -L.dir.v[1], -L.dir.v[2]);
        dist = kInf;
        return -1;
    }
    const bool area = L.type >= 2;
    V3 target = area ? ld4(samples4, i) : L.pos;
    V3 v = sub(target, p);
    dist = std::sqrt(dot(v, v));
    if (!(dist > 0.0f)) return area ? 0 : 1;
    d = mk(v.v[0] / dist, v.v[1] / dist, v.v[2] / dist);
    if (area) {
        slack = eps;
        if (!(dot(mk(-d.v[0], -d.v[1], -d.v[2]), L.dir) > 0.0f)) return 0;
    }
    return -1;
}

// ---------------------------------------------------------------------------------
// Exact kNN (replaces the Zhou radius estimate + range search, kdtree_point.cu:229-267,
// photon_knn.cu:113-235).  Depth-first as dev_RangeSearch: near child where the query
// lies (p > split goes right), far child pushed when its plane distance^2 < r_max^2;
// a bounded max-heap of (d2, index) keeps the k nearest with d2 < r_max^2, and a popped
// far child is skipped once the heap is full and its plane distance^2 exceeds the k-th
// d2.
// ---------------------------------------------------------------------------------

struct KnnCounts {
    uint64_t inner = 0, leaves = 0, refs = 0, tests = 0, accepted = 0;
};

inline uint64_t kkey(float d2, uint32_t idx) { return ((uint64_t)fbits(d2) << 32) | idx; }

inline float dist2(const V3& p, const V3& q) {
    V3 df = sub(p, q);
    return dot(df, df);
}

uint32_t knn_one(const uint32_t* nodes, const uint32_t* refs, const float* ppos, const V3& q,
                 uint32_t k, float r2max, uint64_t* heap, KnnCounts& cnt) {
    uint32_t hn = 0;
    struct Ent { uint32_t node; float pd2; };
    Ent stack[128];
    int sp = 0;
    uint32_t node = 0;
    for (;;) {
        for (;;) {
            uint32_t w1 = nodes[2 * node + 1];
            if ((w1 >> 30) == kLeafTag) break;
            cnt.inner++;
            int a = (int)(w1 >> 30);
            float s = bitsf(nodes[2 * node]);
            float diff = q.v[a] - s;
            float pd2 = diff * diff;
            uint32_t left = node + 1, right = w1 & kIndexMask;
            uint32_t nearc = left, farc = right;
            if (q.v[a] > s) { nearc = right; farc = left; }
            if (pd2 < r2max) stack[sp++] = {farc, pd2};
            node = nearc;
        }
        cnt.leaves++;
        uint32_t count = nodes[2 * node], begin = nodes[2 * node + 1] & kIndexMask;
        for (uint32_t i = 0; i < count; ++i) {
            cnt.refs++;
            cnt.tests++;
            uint32_t pi = refs[begin + i];
            float d2 = dist2(ld4(ppos, pi), q);
            if (!(d2 < r2max)) continue;
            uint64_t key = kkey(d2, pi);
            if (hn < k) {
                heap[hn++] = key;
                std::push_heap(heap, heap + hn);
                cnt.accepted++;
            } else if (key < heap[0]) {
                std::pop_heap(heap, heap + hn);
                heap[hn - 1] = key;
                std::push_heap(heap, heap + hn);
                cnt.accepted++;
            }
        }
        bool found = false;
        while (sp > 0) {
            Ent e = stack[--sp];
            if (hn == k && e.pd2 > bitsf((uint32_t)(heap[0] >> 32))) continue;
            node = e.node;
            found = true;
            break;
        }
        if (!found) break;
    }
    std::sort_heap(heap, heap + hn);
    return hn;
}

// Radiance estimate of kernel_PhotonGather (photon_knn.cu:487-552) over an exact,
// sorted neighbour set.  fp32 path: the kernel's operation order (sequential sum in
// (d2, index) order, component-wise division by r2*pi).  fp64 path: same set and same
// front-facing decisions, continuous arithmetic in double.
struct RadIn {
    const float* ppos;
    const float* ppow;
    const uint32_t* pdir;  // oct codes
    const float* qpos;
    const float* qnrm;
    const float* qdir;
    const float* qinfl;
    const float* qclr;
};

// Exact sum of the terms as 64-bit fixed-point integers: with M the largest |v|
// (M = f 2^e, f in [0.5, 1)) every term is rounded to a multiple of 2^(e-40), the integers
// are added (any order gives the same total) and the total is rounded to float once.
float fixed_sum(const float* v, uint32_t n) {
    float m = 0.0f;
    for (uint32_t i = 0; i < n; ++i) m = std::fmax(m, std::fabs(v[i]));
    if (!(m > 0.0f)) return 0.0f;
    int ex;
    (void)std::frexp(m, &ex);
    const int sh = 40 - ex;
    long long acc = 0;
    for (uint32_t i = 0; i < n; ++i) acc += std::llrint(std::ldexp((double)v[i], sh));
    return (float)std::ldexp((double)acc, -sh);
}

void radiance_one(const RadIn& I, uint32_t q, const uint32_t* nbr, uint32_t cnt, uint32_t k,
                  float rmax, float* out32, double* out64, float* irr32, double* irr64,
                  float* r2out) {
    V3 n = ld4(I.qnrm, q), d = ld4(I.qdir, q), qp = ld4(I.qpos, q);
    V3 nd = mk(-d.v[0], -d.v[1], -d.v[2]);
    if (dot(nd, n) < 0.0f) n = mk(n.v[0] * -1.0f, n.v[1] * -1.0f, n.v[2] * -1.0f);
    float r2max = rmax * rmax;
    // front-facing powers (photon_knn.cu:506-516): fp32 as an exact fixed-point sum, fp64
    // directly
    double s64[3] = {0.0, 0.0, 0.0};
    std::vector<float> terms[3];
    for (uint32_t i = 0; i < cnt && i < k; ++i) {
        uint32_t p = nbr[i];
        V3 pd = oct_decode(I.pdir[p]);
        if (dot(n, pd) < 0.0f) {
            for (int c = 0; c < 3; ++c) {
                terms[c].push_back(I.ppow[4 * p + c]);
                s64[c] += (double)I.ppow[4 * p + c];
            }
        }
    }
    float s[3];
    for (int c = 0; c < 3; ++c) s[c] = fixed_sum(terms[c].data(), (uint32_t)terms[c].size());
    float r2 = r2max;
    double r2d = (double)rmax * (double)rmax;
    if (cnt >= k && k > 0) {
        uint32_t p = nbr[k - 1];
        r2 = dist2(ld4(I.ppos, p), qp);
        double dd = 0.0;
        for (int c = 0; c < 3; ++c) {
            double df = (double)I.ppos[4 * p + c] - (double)qp.v[c];
            dd += df * df;
        }
        r2d = dd;
    }
    if (r2out) r2out[q] = r2;
    float irr[3] = {0.0f, 0.0f, 0.0f};
    double irrd[3] = {0.0, 0.0, 0.0};
    if (r2 > 0.0f) {
        float den = r2 * kPi;
        for (int c = 0; c < 3; ++c) irr[c] = s[c] / den;
    }
    if (r2d > 0.0) {
        for (int c = 0; c < 3; ++c) irrd[c] = s64[c] / (r2d * M_PI);
    }
    if (I.qclr && cnt < 8) {
        for (int c = 0; c < 3; ++c) {
            irr[c] = 0.0f;
            irrd[c] = 0.0;
        }
    }
    float c4[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    if (I.qclr)
        for (int c = 0; c < 4; ++c) c4[c] = I.qclr[4 * q + c];
    for (int c = 0; c < 3; ++c) {
        float infl = I.qinfl[4 * q + c];
        float rad = (((c4[3] * c4[c]) * kInvPi) * irr[c]) * infl;
        if (out32) out32[4 * q + c] = rad;
        if (out64)
            out64[4 * q + c] = (double)c4[3] * (double)c4[c] / M_PI * irrd[c] * (double)infl;
        if (irr32) irr32[4 * q + c] = irr[c];
        if (irr64) irr64[4 * q + c] = irrd[c];
    }
    if (out32) out32[4 * q + 3] = 0.0f;
    if (out64) out64[4 * q + 3] = 0.0;
    if (irr32) irr32[4 * q + 3] = 0.0f;
    if (irr64) irr64[4 * q + 3] = 0.0;
}

// Tree cost over the spatial node boxes (SAH for triangles, VVH for points).
double tree_cost(const uint32_t* nodes, uint32_t num_nodes, const float* root, int kind,
                 float max_qr) {
    if (num_nodes == 0) return 0.0;
    auto measure = [&](const double* lo, const double* hi) {
        double e[3];
        for (int i = 0; i < 3; ++i) e[i] = hi[i] - lo[i];
        if (kind == 1) {
            double q = 2.0 * max_qr;
            return (e[0] + q) * (e[1] + q) * (e[2] + q);
        }
        return 2.0 * (e[0] * e[1] + e[0] * e[2] + e[1] * e[2]);
    };
    struct Ent { uint32_t node; double lo[3], hi[3]; };
    std::vector<Ent> st;
    Ent r;
    r.node = 0;
    for (int i = 0; i < 3; ++i) {
        r.lo[i] = root[i];
        r.hi[i] = root[3 + i];
    }
    double mroot = measure(r.lo, r.hi);
    st.push_back(r);
    double cost = 0.0;
    while (!st.empty()) {
        Ent e = st.back();
        st.pop_back();
        double m = measure(e.lo, e.hi) / mroot;
        uint32_t w1 = nodes[2 * e.node + 1];
        if ((w1 >> 30) == kLeafTag) {
            cost += (double)nodes[2 * e.node] * m;
            continue;
        }
        cost += (double)kCostTraverse * m;
        int a = (int)(w1 >> 30);
        double s = (double)bitsf(nodes[2 * e.node]);
        Ent L = e, R = e;
        L.node = e.node + 1;
        L.hi[a] = s;
        R.node = w1 & kIndexMask;
        R.lo[a] = s;
        st.push_back(R);
        st.push_back(L);
    }
    return cost;
}

// fp64 triangle / box overlap by separating axes (13 axes), used to check that every
// leaf whose box the triangle's geometry touches references it.
bool tri_box_overlap(const double* lo, const double* hi, const double tv[3][3]) {
    double c[3], h[3], v[3][3];
    for (int i = 0; i < 3; ++i) {
        c[i] = 0.5 * (lo[i] + hi[i]);
        h[i] = 0.5 * (hi[i] - lo[i]);
    }
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) v[j][i] = tv[j][i] - c[i];
    for (int i = 0; i < 3; ++i) {
        double mn = std::min({v[0][i], v[1][i], v[2][i]}), mx = std::max({v[0][i], v[1][i], v[2][i]});
        if (mn > h[i] || mx < -h[i]) return false;
    }
    double e[3][3];
    for (int i = 0; i < 3; ++i) {
        e[0][i] = v[1][i] - v[0][i];
        e[1][i] = v[2][i] - v[1][i];
        e[2][i] = v[0][i] - v[2][i];
    }
    double nrm[3] = {e[0][1] * e[1][2] - e[0][2] * e[1][1], e[0][2] * e[1][0] - e[0][0] * e[1][2],
                     e[0][0] * e[1][1] - e[0][1] * e[1][0]};
    {
        double d = nrm[0] * v[0][0] + nrm[1] * v[0][1] + nrm[2] * v[0][2];
        double r = h[0] * std::fabs(nrm[0]) + h[1] * std::fabs(nrm[1]) + h[2] * std::fabs(nrm[2]);
        if (std::fabs(d) > r) return false;
    }
    for (int ei = 0; ei < 3; ++ei) {
        for (int ax = 0; ax < 3; ++ax) {
            double a[3] = {0, 0, 0};
            double u[3] = {0, 0, 0};
            u[ax] = 1.0;
            a[0] = u[1] * e[ei][2] - u[2] * e[ei][1];
            a[1] = u[2] * e[ei][0] - u[0] * e[ei][2];
            a[2] = u[0] * e[ei][1] - u[1] * e[ei][0];
            double p0 = a[0] * v[0][0] + a[1] * v[0][1] + a[2] * v[0][2];
            double p1 = a[0] * v[1][0] + a[1] * v[1][1] + a[2] * v[1][2];
            double p2 = a[0] * v[2][0] + a[1] * v[2][1] + a[2] * v[2][2];
            double r = h[0] * std::fabs(a[0]) + h[1] * std::fabs(a[1]) + h[2] * std::fabs(a[2]);
            double mn = std::min({p0, p1, p2}), mx = std::max({p0, p1, p2});
            if (mn > r * (1 + 1e-12) || mx < -r * (1 + 1e-12)) return false;
        }
    }
    return true;
}

// ---------------------------------------------------------------------------------
// Photon tracing (RTCore::RebuildPhotonMaps, RTCore.cpp:454-638; emission
// photon_build.cu:79, Russian roulette photon_build.cu:220-241, scattering
// photon_build.cu:278): one path at a time, photons stored at every diffuse hit, Russian
// roulette (p = 0.5, power x2) past bounce index 3, cosine scattering about the normal
// facing the incoming ray.  Random numbers: an integer hash of (seed, path, bounce, draw)
// (lowbias32 finaliser); cosine directions: Malley's method with rejection-sampled disk
// points and the Duff et al. 2017 orthonormal basis; only +, *, / and sqrt, so any IEEE
// fp32 implementation with the same operation order draws the same paths.
// ---------------------------------------------------------------------------------
inline uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

inline float rnd01(uint32_t seed, uint32_t path, uint32_t bounce, uint32_t draw) {
    uint32_t h = hash32(seed ^ hash32(path ^ hash32(bounce * 0x9E3779B9u ^ hash32(draw))));
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}

V3 cosine_dir(const V3& n, uint32_t seed, uint32_t path, uint32_t bounce) {
    float x = 0.0f, y = 0.0f, r2 = 0.0f;
    bool ok = false;
    for (uint32_t t = 0; t < 32 && !ok; ++t) {
        x = 2.0f * rnd01(seed, path, bounce, 2 * t + 1) - 1.0f;
        y = 2.0f * rnd01(seed, path, bounce, 2 * t + 2) - 1.0f;
        r2 = x * x + y * y;
        ok = r2 < 1.0f;
    }
    if (!ok) {
        x = 0.0f;
        y = 0.0f;
        r2 = 0.0f;
    }
    float z = std::sqrt(1.0f - r2);
    const float nx = n.v[0], ny = n.v[1], nz = n.v[2];
    float sign = nz >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (sign + nz);
    float b = nx * ny * a;
    float Tx = 1.0f + sign * nx * nx * a, Ty = sign * b, Tz = -sign * nx;
    float Bx = b, By = sign + ny * ny * a, Bz = -ny;
    return mk(x * Tx + y * Bx + z * nx, x * Ty + y * By + z * ny, x * Tz + y * Bz + z * nz);
}

V3 unit_normal(const V3& e1, const V3& e2) {
    V3 n = cross(e1, e2);
    float len = std::sqrt(dot(n, n));
    return mk(n.v[0] / len, n.v[1] / len, n.v[2] / len);
}

struct PhotonSlot {
    bool valid = false;
    float pos[3], pw[3], dir[3];
    uint32_t meta;
    uint32_t tris[4];
};

void photon_path(const uint32_t* nodes, const uint32_t* refs, const float* root, const TriSoup& S,
                 const float* albedo, const float* lights, uint32_t num_lights, uint32_t path,
                 uint32_t max_bounces, uint32_t seed, float eps, PhotonSlot* out) {
    const uint32_t slots = max_bounces + 1;
    for (uint32_t b = 0; b < slots; ++b) out[b].valid = false;
    const uint32_t light = path % num_lights;
    const float* L = lights + 16 * light;
    V3 c = mk(L[0], L[1], L[2]), eu = mk(L[4], L[5], L[6]), ev = mk(L[8], L[9], L[10]);
    float u = rnd01(seed, path, 0xFFFFu, 0), v = rnd01(seed, path, 0xFFFFu, 1);
    V3 o = mk(c.v[0] + u * eu.v[0] + v * ev.v[0], c.v[1] + u * eu.v[1] + v * ev.v[1],
              c.v[2] + u * eu.v[2] + v * ev.v[2]);
    V3 ln = unit_normal(eu, ev);
    V3 d = cosine_dir(ln, seed, path, 0xFFFEu);
    float nl = (float)num_lights;
    float pw[3] = {L[12] * nl, L[13] * nl, L[14] * nl};
    uint32_t tris[4] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
    TraceCounts cnt;
    for (uint32_t b = 0; b < slots; ++b) {
        int32_t tri;
        float t, hu, hv;
        trace_one(nodes, refs, root, S, o, d, eps, tri, t, hu, hv, cnt);
        if (tri < 0) break;
        const float* alb = albedo + 4 * (size_t)tri;
        if (!(alb[0] > 0.0f || alb[1] > 0.0f || alb[2] > 0.0f)) break;
        V3 x = mk(o.v[0] + d.v[0] * t, o.v[1] + d.v[1] * t, o.v[2] + d.v[2] * t);
        if (b > 3) {
            if (!(rnd01(seed, path, b, 100) <= 0.5f)) break;
            for (int k = 0; k < 3; ++k) pw[k] = pw[k] * 2.0f;
        }
        PhotonSlot& ps = out[b];
        ps.valid = true;
        for (int k = 0; k < 3; ++k) {
            ps.pos[k] = x.v[k];
            ps.pw[k] = pw[k];
            ps.dir[k] = d.v[k];
        }
        ps.meta = light | (b << 8);
        for (int k = 0; k < 4; ++k) ps.tris[k] = tris[k];
        if (b + 1 >= slots) break;
        V3 a0 = ld4(S.v0, tri), a1 = ld4(S.v1, tri), a2 = ld4(S.v2, tri);
        V3 n = unit_normal(sub(a1, a0), sub(a2, a0));
        if (dot(n, d) > 0.0f) n = mk(n.v[0] * -1.0f, n.v[1] * -1.0f, n.v[2] * -1.0f);
        d = cosine_dir(n, seed, path, b);
        o = x;
        for (int k = 0; k < 3; ++k) pw[k] = pw[k] * alb[k];
        if (b < 4) tris[b] = (uint32_t)tri;
    }
}

}  // namespace

// =================================================================================
// C API (ctypes)
// =================================================================================
extern "C" {

void* orc_build(int kind, const float* v0, const float* v1, const float* v2, uint32_t n,
                float ratio, uint32_t small_max, float max_qr, int has_root, const float* root6) {
    if (small_max < 1 || small_max > 64) return nullptr;
    return build(kind, v0, v1, v2, n, ratio, small_max, max_qr, has_root, root6, 1);
}
void* orc_build_mt(int kind, const float* v0, const float* v1, const float* v2, uint32_t n, float ratio,
                   uint32_t small_max, float max_qr, int has_root, const float* root6, int nthreads) {
    if (small_max < 1 || small_max > 64) return nullptr;
    return build(kind, v0, v1, v2, n, ratio, small_max, max_qr, has_root, root6, nthreads);
}
uint32_t orc_tree_num_nodes(void* h) { return (uint32_t)(((OTree*)h)->nodes.size() / 2); }
uint32_t orc_tree_num_refs(void* h) { return (uint32_t)((OTree*)h)->refs.size(); }
uint32_t orc_tree_num_leaves(void* h) { return ((OTree*)h)->num_leaves; }
uint32_t orc_tree_max_level(void* h) { return ((OTree*)h)->max_level; }
void orc_tree_root(void* h, float* out6) { std::memcpy(out6, ((OTree*)h)->root, 24); }
void orc_tree_copy(void* h, uint32_t* nodes, uint32_t* refs) {
    OTree* T = (OTree*)h;
    std::memcpy(nodes, T->nodes.data(), T->nodes.size() * 4);
    if (!T->refs.empty()) std::memcpy(refs, T->refs.data(), T->refs.size() * 4);
}
void orc_tree_free(void* h) { delete (OTree*)h; }
void orc_build_counts(void* h, uint64_t* out4) { std::memcpy(out4, ((OTree*)h)->counts, 32); }

double orc_tree_cost(const uint32_t* nodes, uint32_t num_nodes, const float* root6, int kind,
                     float max_qr) {
    return tree_cost(nodes, num_nodes, root6, kind, max_qr);
}

// Structural check of a packed preorder tree: every index in range, every node reached
// exactly once, leaf reference ranges tile [0, num_refs) in preorder, each leaf's
// references ascending.  Returns 0 when valid, else a code naming the first failure.
int orc_tree_check_structure(const uint32_t* nodes, uint32_t num_nodes, const uint32_t* refs,
                             uint32_t num_refs, uint32_t num_elems) {
    if (num_nodes == 0) return 1;
    uint32_t next_ref = 0;
    std::vector<uint32_t> st{0};
    uint32_t expect = 0;
    while (!st.empty()) {
        uint32_t nd = st.back();
        st.pop_back();
        if (nd != expect) return 2;  // preorder: nodes appear in order
        expect++;
        uint32_t w1 = nodes[2 * nd + 1];
        if ((w1 >> 30) == kLeafTag) {
            uint32_t cnt = nodes[2 * nd], b = w1 & kIndexMask;
            if (b != next_ref) return 3;
            if ((uint64_t)b + cnt > num_refs) return 4;
            for (uint32_t i = 0; i < cnt; ++i) {
                if (refs[b + i] >= num_elems) return 5;
                if (i > 0 && refs[b + i] <= refs[b + i - 1]) return 6;
            }
            next_ref = b + cnt;
        } else {
            uint32_t r = w1 & kIndexMask;
            if (r <= nd + 1 || r >= num_nodes) return 7;
            st.push_back(r);
            st.push_back(nd + 1);
        }
    }
    if (expect != num_nodes) return 8;
    if (next_ref != num_refs) return 9;
    return 0;
}

// Coverage check: every leaf box the triangle's geometry touches (fp64 separating-axis
// test, boxes shrunk by rel_shrink of the root extent) must reference the triangle.
// Returns the number of violations.
uint64_t orc_tree_check_coverage(const uint32_t* nodes, const uint32_t* refs, const float* root6,
                                 const float* v0, const float* v1, const float* v2, uint32_t ntri,
                                 double rel_shrink) {
    struct Ent { uint32_t node; double lo[3], hi[3]; };
    std::vector<Ent> st;
    Ent r;
    r.node = 0;
    double ext = 0;
    for (int i = 0; i < 3; ++i) {
        r.lo[i] = root6[i];
        r.hi[i] = root6[3 + i];
        ext = std::max(ext, r.hi[i] - r.lo[i]);
    }
    st.push_back(r);
    uint64_t bad = 0;
    double sh = rel_shrink * ext;
    while (!st.empty()) {
        Ent e = st.back();
        st.pop_back();
        uint32_t w1 = nodes[2 * e.node + 1];
        if ((w1 >> 30) != kLeafTag) {
            int a = (int)(w1 >> 30);
            double s = bitsf(nodes[2 * e.node]);
            Ent L = e, R = e;
            L.node = e.node + 1;
            L.hi[a] = s;
            R.node = w1 & kIndexMask;
            R.lo[a] = s;
            st.push_back(L);
            st.push_back(R);
            continue;
        }
        uint32_t cnt = nodes[2 * e.node], b = w1 & kIndexMask;
        double lo[3], hi[3];
        bool degenerate = false;
        for (int i = 0; i < 3; ++i) {
            lo[i] = e.lo[i] + sh;
            hi[i] = e.hi[i] - sh;
            if (lo[i] > hi[i]) degenerate = true;
        }
        if (degenerate) continue;
        const uint32_t* lr = refs + b;
        for (uint32_t t = 0; t < ntri; ++t) {
            double tv[3][3];
            for (int c = 0; c < 3; ++c) {
                tv[0][c] = v0[4 * t + c];
                tv[1][c] = v1[4 * t + c];
                tv[2][c] = v2[4 * t + c];
            }
            if (!tri_box_overlap(lo, hi, tv)) continue;
            if (!std::binary_search(lr, lr + cnt, t)) bad++;
        }
    }
    return bad;
}

// Closest hits by tree traversal.  counters[4] += inner visits, leaf visits, leaf
// references read, triangle tests.
void orc_trace(const uint32_t* nodes, const uint32_t* refs, const float* root6, const float* v0,
               const float* v1, const float* v2, const float* orig4, const float* dir4, uint32_t n,
               float eps, int32_t* out_tri, float* out_t, float* out_uv, uint64_t* counters,
               int nthreads) {
    TriSoup S{v0, v1, v2};
    std::vector<TraceCounts> per(std::max(1, nthreads));
    std::atomic<int> slot{0};
    parallel_for(n, nthreads, [&](uint32_t b, uint32_t e) {
        TraceCounts& c = per[slot++];
        for (uint32_t i = b; i < e; ++i) {
            float u, v;
            trace_one(nodes, refs, root6, S, ld4(orig4, i), ld4(dir4, i), eps, out_tri[i], out_t[i],
                      u, v, c);
            if (out_uv) {
                out_uv[2 * i] = u;
                out_uv[2 * i + 1] = v;
            }
        }
    });
    if (counters) {
        for (auto& c : per) {
            counters[0] += c.inner;
            counters[1] += c.leaves;
            counters[2] += c.refs;
            counters[3] += c.tests;
        }
    }
}

// Closest hits by testing every triangle (same root clip, epsilon and tie rule).
void orc_trace_brute(const float* root6, const float* v0, const float* v1, const float* v2,
                     uint32_t ntri, const float* orig4, const float* dir4, uint32_t n, float eps,
                     int32_t* out_tri, float* out_t, float* out_uv, int nthreads) {
    TriSoup S{v0, v1, v2};
    parallel_for(n, nthreads, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; ++i) {
            V3 o = ld4(orig4, i), d = ld4(dir4, i);
            V3 inv = mk(1.0f / d.v[0], 1.0f / d.v[1], 1.0f / d.v[2]);
            float t0, t1;
            out_tri[i] = -1;
            out_t[i] = 0.0f;
            if (out_uv) out_uv[2 * i] = out_uv[2 * i + 1] = 0.0f;
            if (!ray_box(root6, root6 + 3, o, inv, eps, kInf, t0, t1)) continue;
            float best_t = t1 + eps, bu = 0, bv = 0;
            int32_t best = -1;
            for (uint32_t t = 0; t < ntri; ++t) tri_test(S, t, o, d, eps, best_t, best, bu, bv);
            out_tri[i] = best;
            if (best >= 0) {
                out_t[i] = best_t;
                if (out_uv) {
                    out_uv[2 * i] = bu;
                    out_uv[2 * i + 1] = bv;
                }
            }
        }
    });
}

// Any-hit segments: out_hit[i] = 1 iff a triangle occludes ray i within tmax[i] (NULL:
// unlimited).  ntri > 0 tests every triangle instead of walking the tree (nodes / refs
// unused).  counters[4] as orc_trace (tree walk only).
void orc_trace_any(const uint32_t* nodes, const uint32_t* refs, const float* root6, const float* v0,
                   const float* v1, const float* v2, uint32_t ntri, const float* orig4, const float* dir4,
                   const float* tmax, uint32_t n, float eps, uint32_t* out_hit, uint64_t* counters,
                   int nthreads) {
    TriSoup S{v0, v1, v2};
    std::vector<TraceCounts> per(std::max(1, nthreads));
    std::atomic<int> slot{0};
    parallel_for(n, nthreads, [&](uint32_t b, uint32_t e) {
        TraceCounts& c = per[slot++];
        for (uint32_t i = b; i < e; ++i) {
            const V3 o = ld4(orig4, i), d = ld4(dir4, i);
            const float dist = tmax ? tmax[i] : kInf;
            out_hit[i] = (ntri ? any_brute(root6, S, ntri, o, d, eps, dist, 0.0f)
                               : any_one(nodes, refs, root6, S, o, d, eps, dist, 0.0f, c))
                             ? 1u : 0u;
        }
    });
    if (counters) {
        for (auto& c : per) {
            counters[0] += c.inner;
            counters[1] += c.leaves;
            counters[2] += c.refs;
            counters[3] += c.tests;
        }
    }
}

// Shadow rays (kernel_TraceShadowRaysDelta / _Area, raytracing.cu:1019-1114):
// out_unocc[i] = 1 iff the light is visible from pts4[i].  light16 / radius as OLight;
// samples4: area-light sample per point (area lights only).  ntri > 0: brute force.
void orc_trace_shadow(const uint32_t* nodes, const uint32_t* refs, const float* root6, const float* v0,
                      const float* v1, const float* v2, uint32_t ntri, const float* light16, float radius,
                      const float* pts4, const float* samples4, uint32_t n, float eps, uint32_t* out_unocc,
                      int nthreads) {
    TriSoup S{v0, v1, v2};
    const OLight L = olight(light16, radius);
    parallel_for(n, nthreads, [&](uint32_t b, uint32_t e) {
        TraceCounts c;
        for (uint32_t i = b; i < e; ++i) {
            const V3 p = ld4(pts4, i);
            V3 d;
            float dist, slack;
            const int fixed = shadow_segment(L, p, samples4, i, d, dist, slack, eps);
            if (fixed >= 0) {
                out_unocc[i] = (uint32_t)fixed;
                continue;
            }
            const bool occ = ntri ? any_brute(root6, S, ntri, p, d, eps, dist, slack)
                                  : any_one(nodes, refs, root6, S, p, d, eps, dist, slack, c);
            out_unocc[i] = occ ? 0u : 1u;
        }
    });
}

// kernel_SampleAreaLight (raytracing.cu:963-1000): rectangle position + u v1 + v v2;
// disc: Shirley-Chiu concentric map (dev_ConcentricSampleDisk, sample_dev.h:100-155) in
// the basis of dev_BuildCoordSystem (bsdf_dev.h:158-177), IEEE 1/sqrt.  Uses the host's
// cosf / sinf, so disc samples agree with the device to a few ulps, not bit for bit.
void orc_sample_area_light(const float* light16, float radius, uint32_t n, float ix, float inv_x, float iy,
                           float inv_y, const float* rnd1, const float* rnd2, float* out4) {
    const OLight L = olight(light16, radius);
    for (uint32_t i = 0; i < n; ++i) {
        const float u = (ix + rnd1[i]) * inv_x, v = (iy + rnd2[i]) * inv_y;
        V3 p;
        if (L.type == 2) {
            const V3& nn = L.dir;
            V3 s;
            if (std::fabs(nn.v[0]) > std::fabs(nn.v[1])) {
                const float il = 1.0f / std::sqrt(nn.v[0] * nn.v[0] + nn.v[2] * nn.v[2]);
                s = mk(-nn.v[2] * il, 0.0f, nn.v[0] * il);
            } else {
                const float il = 1.0f / std::sqrt(nn.v[1] * nn.v[1] + nn.v[2] * nn.v[2]);
                s = mk(0.0f, nn.v[2] * il, -nn.v[1] * il);
            }
            const V3 t = cross(nn, s);
            // square [-1,1]^2 -> disk, angle in units of pi/4 per wedge
            const float a = 2.0f * u - 1.0f, b = 2.0f * v - 1.0f;
            float x = 0.0f, y = 0.0f;
            if (!(a == 0.0f && b == 0.0f)) {
                float r, q;
                if (a >= -b && a > b) {
                    r = a;
                    q = (b > 0.0f) ? b / r : 8.0f + b / r;
                } else if (a >= -b) {
                    r = b;
                    q = 2.0f - a / r;
                } else if (a <= b) {
                    r = -a;
                    q = 4.0f - b / r;
                } else {
                    r = -b;
                    q = 6.0f + a / r;
                }
                const float ang = q * (kPi / 4.0f);
                x = r * std::cos(ang);
                y = r * std::sin(ang);
            }
            p = mk(L.pos.v[0] + L.radius * (x * s.v[0] + y * t.v[0]),
                   L.pos.v[1] + L.radius * (x * s.v[1] + y * t.v[1]),
                   L.pos.v[2] + L.radius * (x * s.v[2] + y * t.v[2]));
        } else {
            p = mk(L.pos.v[0] + u * L.v1.v[0] + v * L.v2.v[0], L.pos.v[1] + u * L.v1.v[1] + v * L.v2.v[1],
                   L.pos.v[2] + u * L.v1.v[2] + v * L.v2.v[2]);
        }
        out4[4 * i] = p.v[0];
        out4[4 * i + 1] = p.v[1];
        out4[4 * i + 2] = p.v[2];
        out4[4 * i + 3] = 0.0f;
    }
}

// Exact kNN through the point tree.  out_nbr: nq*k indices sorted by (d2, index),
// 0xFFFFFFFF past the count.  counters[5] += inner, leaves, refs, distance tests,
// heap insertions.
void orc_knn(const uint32_t* nodes, const uint32_t* refs, const float* ppos4, const float* qpos4,
             uint32_t nq, uint32_t k, float rmax, uint32_t* out_nbr, uint32_t* out_count,
             uint64_t* counters, int nthreads) {
    float r2 = rmax * rmax;
    std::vector<KnnCounts> per(std::max(1, nthreads));
    std::atomic<int> slot{0};
    parallel_for(nq, nthreads, [&](uint32_t b, uint32_t e) {
        KnnCounts& c = per[slot++];
        std::vector<uint64_t> heap(k + 1);
        for (uint32_t q = b; q < e; ++q) {
            uint32_t hn = knn_one(nodes, refs, ppos4, ld4(qpos4, q), k, r2, heap.data(), c);
            out_count[q] = hn;
            if (out_nbr)
                for (uint32_t i = 0; i < k; ++i)
                    out_nbr[(size_t)q * k + i] = (i < hn) ? (uint32_t)(heap[i] & 0xFFFFFFFFu) : 0xFFFFFFFFu;
        }
    });
    if (counters) {
        for (auto& c : per) {
            counters[0] += c.inner;
            counters[1] += c.leaves;
            counters[2] += c.refs;
            counters[3] += c.tests;
            counters[4] += c.accepted;
        }
    }
}

// Exact kNN by testing every photon.
void orc_knn_brute(const float* ppos4, uint32_t np, const float* qpos4, uint32_t nq, uint32_t k,
                   float rmax, uint32_t* out_nbr, uint32_t* out_count, int nthreads) {
    float r2 = rmax * rmax;
    parallel_for(nq, nthreads, [&](uint32_t b, uint32_t e) {
        std::vector<uint64_t> keys;
        for (uint32_t q = b; q < e; ++q) {
            keys.clear();
            V3 qp = ld4(qpos4, q);
            for (uint32_t p = 0; p < np; ++p) {
                float d2 = dist2(ld4(ppos4, p), qp);
                if (d2 < r2) keys.push_back(kkey(d2, p));
            }
            uint32_t m = (uint32_t)std::min<size_t>(k, keys.size());
            std::partial_sort(keys.begin(), keys.begin() + m, keys.end());
            out_count[q] = m;
            for (uint32_t i = 0; i < k; ++i)
                out_nbr[(size_t)q * k + i] = (i < m) ? (uint32_t)(keys[i] & 0xFFFFFFFFu) : 0xFFFFFFFFu;
        }
    });
}

// Radiance of each query from its sorted neighbour set (fp32 kernel order and fp64).
// The accumulation target is zero-initialised here (out = estimate, not +=).
void orc_radiance(const uint32_t* nbr, const uint32_t* count, uint32_t k, float rmax,
                  const float* ppos4, const float* ppow4, const uint32_t* pdir_oct,
                  const float* qpos4, const float* qnrm4, const float* qdir4, const float* qinfl4,
                  const float* qclr4, uint32_t nq, float* out32, double* out64, float* irr32,
                  double* irr64, float* r2out, int nthreads) {
    RadIn I{ppos4, ppow4, pdir_oct, qpos4, qnrm4, qdir4, qinfl4, qclr4};
    parallel_for(nq, nthreads, [&](uint32_t b, uint32_t e) {
        for (uint32_t q = b; q < e; ++q)
            radiance_one(I, q, nbr + (size_t)q * k, count[q], k, rmax, out32, out64, irr32, irr64,
                         r2out);
    });
}

// Photon map of `target` photons (first `target` stored photons in (path, bounce) order,
// power scaled by 1 / paths shot up to the last kept photon).  Output arrays: float4
// position / power / direction, meta (light | bounce << 8), path index, 4 path triangles.
void orc_photons(const uint32_t* nodes, const uint32_t* refs, const float* root6, const float* v0,
                 const float* v1, const float* v2, const float* albedo4, const float* lights,
                 uint32_t num_lights, uint32_t target, uint32_t max_bounces, uint32_t seed, float eps,
                 float* out_pos4, float* out_pow4, float* out_dir4, uint32_t* out_meta, uint32_t* out_path,
                 uint32_t* out_tris, uint32_t* out_count, uint32_t* out_paths, int nthreads) {
    TriSoup S{v0, v1, v2};
    const uint32_t slots = max_bounces + 1;
    const uint64_t max_paths = 100ull * target;
    uint32_t stored = 0, shot = 0;
    const uint32_t chunk = 65536;
    std::vector<PhotonSlot> buf((size_t)chunk * slots);
    while (stored < target && shot < max_paths) {
        uint32_t np = (uint32_t)std::min<uint64_t>(chunk, max_paths - shot);
        parallel_for(np, nthreads, [&](uint32_t b, uint32_t e) {
            for (uint32_t i = b; i < e; ++i)
                photon_path(nodes, refs, root6, S, albedo4, lights, num_lights, shot + i, max_bounces, seed, eps,
                            &buf[(size_t)i * slots]);
        });
        bool done = false;
        for (uint32_t i = 0; i < np && !done; ++i) {
            for (uint32_t b = 0; b < slots; ++b) {
                const PhotonSlot& ps = buf[(size_t)i * slots + b];
                if (!ps.valid) continue;
                for (int k = 0; k < 3; ++k) {
                    out_pos4[4 * stored + k] = ps.pos[k];
                    out_pow4[4 * stored + k] = ps.pw[k];
                    out_dir4[4 * stored + k] = ps.dir[k];
                }
                out_pos4[4 * stored + 3] = out_pow4[4 * stored + 3] = out_dir4[4 * stored + 3] = 0.0f;
                if (out_meta) out_meta[stored] = ps.meta;
                if (out_path) out_path[stored] = shot + i;
                if (out_tris)
                    for (int k = 0; k < 4; ++k) out_tris[4 * (size_t)stored + k] = ps.tris[k];
                if (++stored == target) {
                    shot = shot + i + 1;
                    done = true;
                    break;
                }
            }
        }
        if (!done) shot += np;
    }
    if (stored > 0 && shot > 0) {
        float scale = 1.0f / (float)shot;
        for (uint32_t i = 0; i < stored; ++i)
            for (int k = 0; k < 3; ++k) out_pow4[4 * i + k] = out_pow4[4 * i + k] * scale;
    }
    *out_count = stored;
    *out_paths = shot;
}

// Adjoint of the radiance estimate with neighbour sets fixed (fp64): for upstream
// gradient G (nq float4) writes dLoss/dpower per photon (np x 4 double, accumulated) and
// dLoss/d(colour) per query (nq x 4 double).  Front-facing decisions and the 8-photon
// cut are the fp32 ones of radiance_one.
void orc_radiance_grad(const uint32_t* nbr, const uint32_t* count, uint32_t k, float rmax, const float* ppos4,
                       const float* ppow4, const uint32_t* pdir_oct, uint32_t np, const float* qpos4,
                       const float* qnrm4, const float* qdir4, const float* qinfl4, const float* qclr4, uint32_t nq,
                       const float* grad4, double* out_gpow, double* out_gclr) {
    std::memset(out_gpow, 0, sizeof(double) * 4 * (size_t)np);
    for (uint32_t q = 0; q < nq; ++q) {
        const uint32_t* nb = nbr + (size_t)q * k;
        uint32_t cnt = count[q];
        V3 n = ld4(qnrm4, q), d = ld4(qdir4, q), qp = ld4(qpos4, q);
        V3 nd = mk(-d.v[0], -d.v[1], -d.v[2]);
        if (dot(nd, n) < 0.0f) n = mk(n.v[0] * -1.0f, n.v[1] * -1.0f, n.v[2] * -1.0f);
        double r2d = (double)rmax * (double)rmax;
        if (cnt >= k && k > 0) {
            uint32_t p = nb[k - 1];
            double dd = 0.0;
            for (int c = 0; c < 3; ++c) {
                double df = (double)ppos4[4 * p + c] - (double)qp.v[c];
                dd += df * df;
            }
            r2d = dd;
        }
        double s64[3] = {0, 0, 0};
        for (uint32_t i = 0; i < cnt && i < k; ++i) {
            uint32_t p = nb[i];
            if (dot(n, oct_decode(pdir_oct[p])) < 0.0f)
                for (int c = 0; c < 3; ++c) s64[c] += (double)ppow4[4 * p + c];
        }
        double c4[4] = {1, 1, 1, 1};
        if (qclr4)
            for (int c = 0; c < 4; ++c) c4[c] = qclr4[4 * q + c];
        bool gated = (qclr4 && cnt < 8) || !(r2d > 0.0);
        for (int c = 0; c < 3; ++c) {
            double g = grad4[4 * q + c], infl = qinfl4[4 * q + c];
            double w = gated ? 0.0 : g * c4[3] * c4[c] / M_PI * infl / (M_PI * r2d);
            double e = gated ? 0.0 : s64[c] / (M_PI * r2d);
            if (out_gclr) out_gclr[4 * q + c] = g * c4[3] / M_PI * e * infl;
            if (w != 0.0)
                for (uint32_t i = 0; i < cnt && i < k; ++i) {
                    uint32_t p = nb[i];
                    if (dot(n, oct_decode(pdir_oct[p])) < 0.0f) out_gpow[4 * p + c] += w;
                }
        }
        if (out_gclr) out_gclr[4 * q + 3] = 0.0;
    }
}

float orc_fixed_sum(const float* v, uint32_t n) { return fixed_sum(v, n); }

uint32_t orc_oct_encode(float x, float y, float z) { return oct_encode(x, y, z); }
void orc_oct_encode_n(const float* dir4, uint32_t n, uint32_t* out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = oct_encode(dir4[4 * i], dir4[4 * i + 1], dir4[4 * i + 2]);
}
void orc_oct_decode(uint32_t c, float* out3) {
    V3 v = oct_decode(c);
    out3[0] = v.v[0];
    out3[1] = v.v[1];
    out3[2] = v.v[2];
}

// Single-ray / single-triangle test for closed-form checks.
int orc_ray_tri(const float* a, const float* b, const float* c, const float* o, const float* d,
                float* out3) {
    V3 A = mk(a[0], a[1], a[2]), B = mk(b[0], b[1], b[2]), Cc = mk(c[0], c[1], c[2]);
    float lam = 0.0f, u = 0.0f, v = 0.0f;
    bool hit = ray_tri(A, sub(B, A), sub(Cc, A), mk(o[0], o[1], o[2]), mk(d[0], d[1], d[2]), lam, u, v);
    out3[0] = lam;
    out3[1] = u;
    out3[2] = v;
    return hit ? 1 : 0;
}

// Reference KDTreeData layout of a packed tree (kernel_TraversalDownPath, kdtree.cu:1707-1770,
// format KDKernelDefs.h:76-117): recursive preorder walk, word addresses assigned top-down
// (left child 3 words after its parent, right child after the left subtree), node boxes
// carried down the walk.  Returns the stream length in words.
namespace {
struct ExportCtx {
    const uint32_t* nodes;
    const uint32_t* refs;
    uint32_t* tree;
    uint32_t* num_elems;
    uint32_t* addr;
    float* extent;
    uint32_t* left;
    uint32_t* right;
};

uint32_t export_words(const uint32_t* nodes, uint32_t i) {
    uint32_t w1 = nodes[2 * i + 1];
    if ((w1 >> 30) == 3u) return 2 + nodes[2 * i];
    return 3 + export_words(nodes, i + 1) + export_words(nodes, w1 & 0x3FFFFFFFu);
}

// returns the subtree's reference count
uint32_t export_node(const ExportCtx& X, uint32_t i, uint32_t a, const float* lo, const float* hi) {
    const uint32_t w0 = X.nodes[2 * i], w1 = X.nodes[2 * i + 1];
    X.addr[i] = a;
    float d[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    X.extent[4 * i + 0] = lo[0] + 0.5f * d[0];
    X.extent[4 * i + 1] = lo[1] + 0.5f * d[1];
    X.extent[4 * i + 2] = lo[2] + 0.5f * d[2];
    X.extent[4 * i + 3] = 0.5f * std::sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    if ((w1 >> 30) == 3u) {
        X.tree[a] = i | 0x80000000u;
        X.tree[a + 1] = w0;
        for (uint32_t k = 0; k < w0; ++k) X.tree[a + 2 + k] = X.refs[(w1 & 0x3FFFFFFFu) + k];
        X.num_elems[i] = w0;
        X.left[i] = X.right[i] = 0;
        return w0;
    }
    const uint32_t ax = w1 >> 30, r = w1 & 0x3FFFFFFFu;
    float s;
    std::memcpy(&s, &w0, 4);
    const uint32_t al = a + 3, ar = al + export_words(X.nodes, i + 1);
    X.tree[a] = i;
    X.tree[a + 1] = (ax << 30) | (ar & 0x0FFFFFFFu);
    X.tree[a + 2] = w0;
    X.left[i] = i + 1;
    X.right[i] = r;
    float lhi[3] = {hi[0], hi[1], hi[2]}, rlo[3] = {lo[0], lo[1], lo[2]};
    lhi[ax] = s;
    rlo[ax] = s;
    uint32_t c = export_node(X, i + 1, al, lo, lhi);
    c += export_node(X, r, ar, rlo, hi);
    X.num_elems[i] = c;
    return c;
}
}  // namespace

uint32_t orc_export_preorder(const uint32_t* nodes, const uint32_t* refs, const float* root6, uint32_t* tree,
                             uint32_t* num_elems, uint32_t* addr, float* extent4, uint32_t* left, uint32_t* right) {
    ExportCtx X{nodes, refs, tree, num_elems, addr, extent4, left, right};
    export_node(X, 0, 0, root6, root6 + 3);
    return export_words(nodes, 0);
}

}  // extern "C"
