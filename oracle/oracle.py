"""ctypes binding of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the CPU-baseline
leg of bench.py, never by the product package.  See oracle.cpp for what each function
restates.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
vp = ctypes.c_void_p


def build_lib(force: bool = False) -> str:
    if force or not os.path.exists(_LIB) or os.path.getmtime(_LIB) < os.path.getmtime(os.path.join(_HERE, "oracle.cpp")):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "liboracle.so"], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB):
        build_lib()
    L = ctypes.CDLL(_LIB)
    L.orc_build.restype = vp
    L.orc_build.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_uint32, ctypes.c_float, ctypes.c_uint32,
                            ctypes.c_float, ctypes.c_int, vp]
    L.orc_build_mt.restype = vp
    L.orc_build_mt.argtypes = L.orc_build.argtypes + [ctypes.c_int]
    for name in ("orc_tree_num_nodes", "orc_tree_num_refs", "orc_tree_num_leaves", "orc_tree_max_level"):
        getattr(L, name).restype = ctypes.c_uint32
        getattr(L, name).argtypes = [vp]
    L.orc_tree_root.argtypes = [vp, f32p]
    L.orc_tree_copy.argtypes = [vp, u32p, u32p]
    L.orc_tree_free.argtypes = [vp]
    L.orc_build_counts.argtypes = [vp, u64p]
    L.orc_tree_cost.restype = ctypes.c_double
    L.orc_tree_cost.argtypes = [u32p, ctypes.c_uint32, f32p, ctypes.c_int, ctypes.c_float]
    L.orc_tree_check_structure.restype = ctypes.c_int
    L.orc_tree_check_structure.argtypes = [u32p, ctypes.c_uint32, u32p, ctypes.c_uint32, ctypes.c_uint32]
    L.orc_tree_check_coverage.restype = ctypes.c_uint64
    L.orc_tree_check_coverage.argtypes = [u32p, u32p, f32p, f32p, f32p, f32p, ctypes.c_uint32, ctypes.c_double]
    L.orc_trace.argtypes = [u32p, u32p, f32p, f32p, f32p, f32p, f32p, f32p, ctypes.c_uint32, ctypes.c_float,
                            i32p, f32p, f32p, u64p, ctypes.c_int]
    L.orc_trace_brute.argtypes = [f32p, f32p, f32p, f32p, ctypes.c_uint32, f32p, f32p, ctypes.c_uint32,
                                  ctypes.c_float, i32p, f32p, f32p, ctypes.c_int]
    L.orc_knn.argtypes = [u32p, u32p, f32p, f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, u32p, u32p,
                          u64p, ctypes.c_int]
    L.orc_knn_brute.argtypes = [f32p, ctypes.c_uint32, f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                                u32p, u32p, ctypes.c_int]
    L.orc_radiance.argtypes = [u32p, u32p, ctypes.c_uint32, ctypes.c_float, f32p, f32p, u32p, f32p, f32p, f32p,
                               f32p, vp, ctypes.c_uint32, f32p, f64p, f32p, f64p, f32p, ctypes.c_int]
    L.orc_oct_encode.restype = ctypes.c_uint32
    L.orc_oct_encode.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float]
    L.orc_oct_encode_n.argtypes = [f32p, ctypes.c_uint32, u32p]
    L.orc_oct_decode.argtypes = [ctypes.c_uint32, f32p]
    L.orc_photons.argtypes = [u32p, u32p, f32p, f32p, f32p, f32p, f32p, f32p, ctypes.c_uint32, ctypes.c_uint32,
                              ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, f32p, f32p, f32p, u32p, u32p, u32p,
                              ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32), ctypes.c_int]
    L.orc_radiance_grad.argtypes = [u32p, u32p, ctypes.c_uint32, ctypes.c_float, f32p, f32p, u32p, ctypes.c_uint32,
                                    f32p, f32p, f32p, f32p, vp, ctypes.c_uint32, f32p, f64p, f64p]
    L.orc_export_preorder.restype = ctypes.c_uint32
    L.orc_export_preorder.argtypes = [u32p, u32p, f32p, u32p, u32p, u32p, f32p, u32p, u32p]
    L.orc_ray_tri.restype = ctypes.c_int
    L.orc_ray_tri.argtypes = [f32p, f32p, f32p, f32p, f32p, f32p]
    L.orc_fixed_sum.restype = ctypes.c_float
    L.orc_fixed_sum.argtypes = [f32p, ctypes.c_uint32]
    L.orc_trace_any.argtypes = [u32p, u32p, f32p, f32p, f32p, f32p, ctypes.c_uint32, f32p, f32p, f32p,
                                ctypes.c_uint32, ctypes.c_float, u32p, u64p, ctypes.c_int]
    L.orc_trace_shadow.argtypes = [u32p, u32p, f32p, f32p, f32p, f32p, ctypes.c_uint32, f32p, ctypes.c_float,
                                   f32p, f32p, ctypes.c_uint32, ctypes.c_float, u32p, ctypes.c_int]
    L.orc_sample_area_light.argtypes = [f32p, ctypes.c_float, ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
                                        ctypes.c_float, ctypes.c_float, f32p, f32p, f32p]
    _lib = L
    return L


def _c(a, dt=np.float32):
    return np.ascontiguousarray(a, dtype=dt)


def host_threads() -> int:
    """Host threads for the oracle: the cores this process may run on, capped by
    OMP_NUM_THREADS when set (the share of a shared host a job is given)."""
    n = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS")
    if cap and cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return max(1, n)


def cpu_model() -> str:
    """The host CPU's model string (/proc/cpuinfo), for reports."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _threads(n):
    return n if n else host_threads()


class Tree:
    """A packed preorder tree (the library's node format) held in host memory."""

    def __init__(self, kind, nodes, refs, root, max_level=0, num_leaves=0, num_elems=0):
        self.kind = kind
        self.nodes = nodes
        self.refs = refs
        self.root = root
        self.max_level = max_level
        self.num_leaves = num_leaves
        self.num_elems = num_elems

    @property
    def num_nodes(self):
        return self.nodes.size // 2


def build(kind, v0, v1=None, v2=None, ratio=None, small_max=None, max_qr=1.0, root=None, nthreads=0) -> Tree:
    """Oracle kd-tree (the tree is the same for every thread count; 0 = host_threads())."""
    L = lib()
    if ratio is None:
        ratio = 0.25 if kind == 0 else 0.1
    if small_max is None:
        small_max = 64 if kind == 0 else 32
    v0 = _c(v0)
    v1 = _c(v1) if v1 is not None else v0
    v2 = _c(v2) if v2 is not None else v0
    r6 = _c(root).ravel() if root is not None else np.zeros(6, np.float32)
    h = L.orc_build_mt(kind, v0.ctypes.data, v1.ctypes.data, v2.ctypes.data, v0.shape[0], ratio, small_max,
                       max_qr, 1 if root is not None else 0, r6.ctypes.data, _threads(nthreads))
    if not h:
        raise ValueError("oracle build rejected its parameters")
    try:
        nn, nr = L.orc_tree_num_nodes(h), L.orc_tree_num_refs(h)
        nodes = np.zeros(2 * nn, np.uint32)
        refs = np.zeros(max(nr, 1), np.uint32)
        L.orc_tree_copy(h, nodes, refs)
        rt = np.zeros(6, np.float32)
        L.orc_tree_root(h, rt)
        T = Tree(kind, nodes, refs[:nr], rt, L.orc_tree_max_level(h), L.orc_tree_num_leaves(h), v0.shape[0])
        T.counts = np.zeros(4, np.uint64)
        L.orc_build_counts(h, T.counts)
        return T
    finally:
        L.orc_tree_free(h)


def build_bytes(tree: Tree) -> float:
    """Algorithmic bytes of a build (element streams read and written by its stages), from
    the oracle build's counts: the input (36 B per triangle, 12 B per point) read and one
    element record (index + bounds: 28 B per triangle, 16 B per point: points keep no upper
    bound) written; per element of an active large node per level its bounds read for the
    tight box (24 / 12 B) and its record read for classification (28 / 16 B); every child
    record written (28 / 16 B); a clipped triangle side reads the vertices (36 B); every
    small-stage element read once (28 / 16 B); 8 B per node and 4 B per leaf reference
    written."""
    large, child, clips, small = (float(c) for c in tree.counts)
    n = float(tree.num_elems)
    if tree.kind == 0:
        inp, rec, bnd = 36.0, 28.0, 24.0
    else:
        inp, rec, bnd = 12.0, 16.0, 12.0
    return (n * (inp + rec) + large * (bnd + rec) + child * rec + clips * 36.0 + small * rec
            + 8.0 * tree.num_nodes + 4.0 * tree.refs.size)


def tree_cost(nodes, root, kind, max_qr=1.0) -> float:
    nodes = _c(nodes, np.uint32)
    return lib().orc_tree_cost(nodes, nodes.size // 2, _c(root).ravel(), kind, max_qr)


def export_preorder(nodes, refs, root, num_leaves):
    """Reference KDTreeData layout of a packed tree (kernel_TraversalDownPath)."""
    nodes = _c(nodes, np.uint32)
    n = nodes.size // 2
    refs_c = _c(refs, np.uint32) if len(refs) else np.zeros(1, np.uint32)
    size = 3 * (n - num_leaves) + 2 * num_leaves + len(refs)
    tree = np.zeros(max(size, 1), np.uint32)
    ne, addr, left, right = (np.zeros(n, np.uint32) for _ in range(4))
    ext = np.zeros((n, 4), np.float32)
    got = lib().orc_export_preorder(nodes, refs_c, _c(root).ravel(), tree, ne, addr, ext, left, right)
    assert got == size, (got, size)
    return dict(tree=tree[:size], num_elems=ne, addr=addr, extent=ext, left=left, right=right)


def check_structure(nodes, refs, num_elems) -> int:
    nodes = _c(nodes, np.uint32)
    refs = _c(refs, np.uint32) if len(refs) else np.zeros(1, np.uint32)
    return lib().orc_tree_check_structure(nodes, nodes.size // 2, refs, len(refs), num_elems)


def check_coverage(nodes, refs, root, v0, v1, v2, rel_shrink=1e-6) -> int:
    refs = _c(refs, np.uint32) if len(refs) else np.zeros(1, np.uint32)
    return int(lib().orc_tree_check_coverage(_c(nodes, np.uint32), refs, _c(root).ravel(), _c(v0), _c(v1), _c(v2),
                                             v0.shape[0], rel_shrink))


def trace(tree: Tree, v0, v1, v2, orig, dirs, eps, nthreads=0):
    n = orig.shape[0]
    tri = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    uv = np.zeros(2 * n, np.float32)
    cnt = np.zeros(4, np.uint64)
    refs = tree.refs if tree.refs.size else np.zeros(1, np.uint32)
    lib().orc_trace(tree.nodes, refs, tree.root, _c(v0), _c(v1), _c(v2), _c(orig), _c(dirs), n, eps, tri, t, uv,
                    cnt, _threads(nthreads))
    return tri, t, uv.reshape(n, 2), cnt


def trace_brute(root, v0, v1, v2, orig, dirs, eps, nthreads=0):
    n = orig.shape[0]
    tri = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    uv = np.zeros(2 * n, np.float32)
    lib().orc_trace_brute(_c(root).ravel(), _c(v0), _c(v1), _c(v2), v0.shape[0], _c(orig), _c(dirs), n, eps, tri,
                          t, uv, _threads(nthreads))
    return tri, t, uv.reshape(n, 2)


def fixed_sum(v):
    """The radiance estimate's fp32 power sum (orc_fixed_sum): exact fixed-point total of
    the terms rounded to float once."""
    v = np.ascontiguousarray(v, np.float32).ravel()
    return np.float32(lib().orc_fixed_sum(v if v.size else np.zeros(1, np.float32), v.size))


FLT_MAX = float(np.finfo(np.float32).max)
_EMPTY_TREE = (np.array([0, 3 << 30], np.uint32), np.zeros(1, np.uint32))


def _tree_arrays(tree):
    if tree is None:
        return _EMPTY_TREE
    return tree.nodes, (tree.refs if tree.refs.size else np.zeros(1, np.uint32))


def trace_any(tree, root, v0, v1, v2, orig, dirs, eps, tmax=None, brute=False, nthreads=0):
    """Any-hit segments (orc_trace_any): uint32 1 where a triangle occludes; brute=True
    tests every triangle instead of walking `tree`.  Returns (hit, counters)."""
    n = orig.shape[0]
    out = np.zeros(n, np.uint32)
    cnt = np.zeros(4, np.uint64)
    tm = _c(tmax) if tmax is not None else np.full(n, FLT_MAX, np.float32)
    nodes, refs = _tree_arrays(None if brute else tree)
    lib().orc_trace_any(nodes, refs, _c(root).ravel(), _c(v0), _c(v1), _c(v2), v0.shape[0] if brute else 0,
                        _c(orig), _c(dirs), tm, n, eps, out, cnt, _threads(nthreads))
    return out, cnt


def light16(kind, position=(0, 0, 0), direction=(0, 0, -1), L_emit=(1, 1, 1), v1=(0, 0, 0), v2=(0, 0, 0)):
    """LightData as the oracle's 16 floats: type, position, direction, L_emit, v1, v2."""
    return np.array([kind, *position, *direction, *L_emit, *v1, *v2], np.float32)


def trace_shadow(tree, root, v0, v1, v2, light, radius, pts, eps, samples=None, brute=False, nthreads=0):
    """Shadow rays (orc_trace_shadow): uint32 1 where the light is visible."""
    n = pts.shape[0]
    out = np.zeros(n, np.uint32)
    smp = _c(samples) if samples is not None else np.zeros(4, np.float32)
    nodes, refs = _tree_arrays(None if brute else tree)
    lib().orc_trace_shadow(nodes, refs, _c(root).ravel(), _c(v0), _c(v1), _c(v2), v0.shape[0] if brute else 0,
                           _c(light), radius, _c(pts), smp, n, eps, out, _threads(nthreads))
    return out


def sample_area_light(light, radius, rnd1, rnd2, ix=0.0, inv_x=1.0, iy=0.0, inv_y=1.0):
    """Area-light sample points (orc_sample_area_light): (n, 4) float32."""
    n = rnd1.shape[0]
    out = np.zeros(4 * n, np.float32)
    lib().orc_sample_area_light(_c(light), radius, n, ix, inv_x, iy, inv_y, _c(rnd1), _c(rnd2), out)
    return out.reshape(n, 4)


def knn(tree: Tree, ppos, qpos, k, rmax, nthreads=0):
    nq = qpos.shape[0]
    nbr = np.zeros(nq * k, np.uint32)
    cnt = np.zeros(nq, np.uint32)
    ctr = np.zeros(5, np.uint64)
    refs = tree.refs if tree.refs.size else np.zeros(1, np.uint32)
    lib().orc_knn(tree.nodes, refs, _c(ppos), _c(qpos), nq, k, rmax, nbr, cnt, ctr, _threads(nthreads))
    return nbr.reshape(nq, k), cnt, ctr


def knn_brute(ppos, qpos, k, rmax, nthreads=0):
    nq = qpos.shape[0]
    nbr = np.zeros(nq * k, np.uint32)
    cnt = np.zeros(nq, np.uint32)
    lib().orc_knn_brute(_c(ppos), ppos.shape[0], _c(qpos), nq, k, rmax, nbr, cnt, _threads(nthreads))
    return nbr.reshape(nq, k), cnt


def oct_encode(dirs4) -> np.ndarray:
    d = _c(dirs4)
    out = np.zeros(d.shape[0], np.uint32)
    lib().orc_oct_encode_n(d, d.shape[0], out)
    return out


def oct_decode(code: int) -> np.ndarray:
    out = np.zeros(3, np.float32)
    lib().orc_oct_decode(int(code), out)
    return out


def radiance(nbr, count, k, rmax, ppos, ppow, pdir4, qpos, qnrm, qdir, qinfl, qclr=None, nthreads=0):
    nq = qpos.shape[0]
    codes = oct_encode(pdir4)
    out32 = np.zeros(4 * nq, np.float32)
    out64 = np.zeros(4 * nq, np.float64)
    irr32 = np.zeros(4 * nq, np.float32)
    irr64 = np.zeros(4 * nq, np.float64)
    r2 = np.zeros(nq, np.float32)
    clr = _c(qclr) if qclr is not None else None
    lib().orc_radiance(_c(nbr, np.uint32).ravel(), _c(count, np.uint32), k, rmax, _c(ppos), _c(ppow), codes,
                       _c(qpos), _c(qnrm), _c(qdir), _c(qinfl), clr.ctypes.data if clr is not None else None, nq,
                       out32, out64, irr32, irr64, r2, _threads(nthreads))
    return (out32.reshape(nq, 4), out64.reshape(nq, 4), irr32.reshape(nq, 4), irr64.reshape(nq, 4), r2)


def radiance_grad(nbr, count, k, rmax, ppos, ppow, pdir4, qpos, qnrm, qdir, qinfl, qclr, grad):
    """fp64 adjoint of the estimate with the neighbour sets fixed: (dLoss/dpower, dLoss/dcolour)."""
    nq, np_ = qpos.shape[0], ppos.shape[0]
    codes = oct_encode(pdir4)
    gp = np.zeros(4 * np_, np.float64)
    gc = np.zeros(4 * nq, np.float64)
    clr = _c(qclr) if qclr is not None else None
    lib().orc_radiance_grad(_c(nbr, np.uint32).ravel(), _c(count, np.uint32), k, rmax, _c(ppos), _c(ppow), codes,
                            np_, _c(qpos), _c(qnrm), _c(qdir), _c(qinfl), clr.ctypes.data if clr is not None else None,
                            nq, _c(grad), gp, gc)
    return gp.reshape(np_, 4), gc.reshape(nq, 4)


def ray_tri(a, b, c, o, d):
    out = np.zeros(3, np.float32)
    hit = lib().orc_ray_tri(_c(a), _c(b), _c(c), _c(o), _c(d), out)
    return bool(hit), out


def photons(tree: Tree, v0, v1, v2, albedo, lights, target, max_bounces=4, seed=0, eps=1e-4, nthreads=0):
    """Traced photon map: dict of pos, power, dir (float4), meta, path, tris, count, paths."""
    lights = _c(np.asarray(lights, np.float32).reshape(-1, 4, 3))
    nl = lights.shape[0]
    lrec = np.zeros((nl, 4, 4), np.float32)
    lrec[:, :, :3] = lights
    pos = np.zeros((target, 4), np.float32)
    pw = np.zeros((target, 4), np.float32)
    dr = np.zeros((target, 4), np.float32)
    meta = np.zeros(target, np.uint32)
    path = np.zeros(target, np.uint32)
    tris = np.zeros(4 * target, np.uint32)
    cnt = ctypes.c_uint32(0)
    shot = ctypes.c_uint32(0)
    refs = tree.refs if tree.refs.size else np.zeros(1, np.uint32)
    lib().orc_photons(tree.nodes, refs, tree.root, _c(v0), _c(v1), _c(v2), _c(albedo), _c(lrec).ravel(), nl,
                      target, max_bounces, seed, eps, pos, pw, dr, meta, path, tris, ctypes.byref(cnt),
                      ctypes.byref(shot), _threads(nthreads))
    n = cnt.value
    return dict(pos=pos[:n], power=pw[:n], dir=dr[:n], meta=meta[:n], path=path[:n], tris=tris.reshape(-1, 4)[:n],
                count=n, paths=shot.value)


# --- photon reweighting by emission and albedo (fp64, numpy) ----------------------------
# Restates the forward and adjoint of mnrt_photons_reweight: with a photon's path fixed its
# stored power is the emission-only power of its light times the albedo of every triangle it
# scattered from (the tracer's bookkeeping, photon_build.cu:79-130 and :278-331,
# RTCore.cpp:545-581); a hit's colour is its triangle's albedo.  Pinned in
# tests/test_oracle_reweight.py: emission 1 and the trace-time albedo give back the powers
# the photon tracer stored (an independent code path), and the gradient equals fp64
# central differences of the forward and a loop over every (photon, path position) pair.

def _path_albedos(tris, albedo):
    """(P, 4, 3) albedo of each recorded triangle, 1 for the padding (index < 0)."""
    t = np.asarray(tris).astype(np.int64)
    t = np.where(t >= 2 ** 31, t - 2 ** 32, t)  # uint32 padding 0xFFFFFFFF = -1
    a = np.ones(t.shape + (3,), np.float64)
    ok = t >= 0
    a[ok] = np.asarray(albedo, np.float64)[t[ok]]
    return t, a


def reweight(base, light, tris, emission, albedo):
    """fp64 photon powers (P, 3) = base * emission[light] * prod of path albedos."""
    _, a = _path_albedos(tris, albedo)
    p = np.asarray(base, np.float64)[:, :3] * np.asarray(emission, np.float64)[np.asarray(light, np.int64)]
    return p * a.prod(axis=1)


def reweight_grad(base, light, tris, emission, albedo, grad_pow, hit_tri=None, grad_clr=None):
    """fp64 (dLoss/demission (L, 3), dLoss/dalbedo (T, 3)) of the reweighting given
    dLoss/dpower (P, >=3) and optionally dLoss/dclr at hits on triangles hit_tri (-1: none)."""
    t, a = _path_albedos(tris, albedo)
    b = np.asarray(base, np.float64)[:, :3]
    g = np.asarray(grad_pow, np.float64)[:, :3]
    lt = np.asarray(light, np.int64)
    e = np.asarray(emission, np.float64)
    ge = np.zeros(e.shape, np.float64)
    ga = np.zeros(np.asarray(albedo).shape[:1] + (3,), np.float64)
    np.add.at(ge, lt, g * b * a.prod(axis=1))
    for k in range(t.shape[1]):
        others = np.prod(np.delete(a, k, axis=1), axis=1)
        ok = t[:, k] >= 0
        np.add.at(ga, t[ok, k], (g * b * e[lt] * others)[ok])
    if hit_tri is not None:
        h = np.asarray(hit_tri).astype(np.int64)
        h = np.where(h >= 2 ** 31, h - 2 ** 32, h)
        ok = h >= 0
        np.add.at(ga, h[ok], np.asarray(grad_clr, np.float64)[ok, :3])
    return ge, ga


def base_powers(meta, lights, paths):
    """Emission-only powers (P, 3) of a traced map: light power * light count, doubled per
    Russian-roulette survival (bounce index > 3), / paths (fp64)."""
    meta = np.asarray(meta).astype(np.int64)
    lt, bounce = meta & 0xFF, (meta >> 8) & 0xFF
    lp = np.asarray(lights, np.float64)[:, 3, :]
    return lp[lt] * lp.shape[0] * np.exp2(np.maximum(bounce - 3, 0))[:, None] / float(paths)
