"""Python entry points over the C ABI, with torch device tensors as buffers.

Mirrors the reference call surface for the hot path:
  KDTreeTri / KDTreePoint + BuildTree   -> build_tri / build_points  (KDTreeGPU.cpp:175)
  KernelRTTraceRays                     -> trace_closest            (raytracing.cu:1617)
  KernelPMPhotonGather                  -> knn_gather               (photon_knn.cu:833)
  KDTreePoint::ComputeQueryRadii        -> knn_radii                (KDTreePoint.cpp:229)
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import native as N


def _torch():
    import torch

    return torch


class DeviceTree:
    """A device-resident tree plus the tensors that keep caller-provided arrays alive."""

    def __init__(self, kind: int, t=None, batch=None):
        self.t = t if t is not None else N.Tree()
        self.t.kind = kind
        self._keep = []
        self._batch = batch  # views of a TreeBatch never release

    @property
    def kind(self):
        return self.t.kind

    @property
    def num_nodes(self):
        return self.t.num_nodes

    @property
    def num_refs(self):
        return self.t.num_refs

    @property
    def max_level(self):
        return self.t.max_level

    @property
    def root(self):
        return np.array(list(self.t.root_min) + list(self.t.root_max), np.float32)

    def ref(self):
        return ctypes.byref(self.t)

    def nodes_host(self) -> np.ndarray:
        torch = _torch()
        n = 2 * self.t.num_nodes
        out = torch.empty(n, dtype=torch.int32, device="cuda")
        _copy_from_ptr(out, self.t.d_nodes, n * 4)
        return out.cpu().numpy().view(np.uint32).copy()

    def refs_host(self) -> np.ndarray:
        torch = _torch()
        n = self.t.num_refs
        if n == 0:
            return np.zeros(0, np.uint32)
        out = torch.empty(n, dtype=torch.int32, device="cuda")
        _copy_from_ptr(out, self.t.d_refs, n * 4)
        return out.cpu().numpy().view(np.uint32).copy()

    def recs_host(self) -> np.ndarray:
        """Leaf-ordered records as float32 rows (12 floats per triangle, 8 per photon)."""
        torch = _torch()
        per = 12 if self.kind == N.KIND_TRI else 8
        n = self.t.num_refs * per
        if n == 0 or not self.t.d_recs:
            return np.zeros((0, per), np.float32)
        out = torch.empty(n, dtype=torch.float32, device="cuda")
        _copy_from_ptr(out, self.t.d_recs, n * 4)
        return out.cpu().numpy().reshape(-1, per).copy()

    def qnodes_host(self) -> np.ndarray:
        """Query-time node words (empty when the tree has none)."""
        torch = _torch()
        n = 2 * self.t.num_nodes
        if not self.t.d_qnodes:
            return np.zeros(0, np.uint32)
        out = torch.empty(n, dtype=torch.int32, device="cuda")
        _copy_from_ptr(out, self.t.d_qnodes, n * 4)
        return out.cpu().numpy().view(np.uint32).copy()

    def cost(self, max_query_radius: float = 1.0) -> float:
        c = ctypes.c_double(0.0)
        N.check(N.load().mnrt_tree_cost(self.ref(), max_query_radius, ctypes.byref(c), N.stream_handle()),
                "mnrt_tree_cost")
        return c.value

    def _owns(self):
        return self._batch is None and (self.t.owns_nodes or self.t.owns_recs or self.t.owns_qnodes)

    def release(self):
        if self._owns():
            _torch().cuda.synchronize()
            N.check(N.load().mnrt_tree_release(self.ref()), "mnrt_tree_release")
        self._keep = []

    def __del__(self):
        try:
            if self._owns():
                self.release()
        except Exception:
            pass


class TreeBatch:
    """Trees of one batched build: views into arrays owned by the first tree."""

    def __init__(self, kind: int, offsets):
        self.kind = kind
        self.offsets = np.ascontiguousarray(offsets, np.uint32)
        self.num_scenes = len(self.offsets) - 1
        self.arr = (N.Tree * self.num_scenes)()
        self._keep = []

    def __len__(self):
        return self.num_scenes

    def __getitem__(self, s: int) -> DeviceTree:
        if not 0 <= s < self.num_scenes:
            raise IndexError(s)
        return DeviceTree(self.kind, self.arr[s], batch=self)

    def ref(self):
        return ctypes.cast(self.arr, ctypes.POINTER(N.Tree))

    def offsets_ref(self):
        return self.offsets.ctypes.data_as(ctypes.POINTER(N.u32))

    def release(self):
        if self.num_scenes and (self.arr[0].owns_nodes or self.arr[0].owns_recs or self.arr[0].owns_qnodes):
            _torch().cuda.synchronize()
            N.check(N.load().mnrt_tree_release(ctypes.byref(self.arr[0])), "mnrt_tree_release")
        self._keep = []

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def _f32_ptr(a, per: int, n: int):
    if a is None:
        return None, None
    a = np.ascontiguousarray(a, np.float32).reshape(-1)
    if a.size != per * n:
        raise N.MnrtError(f"expected {per * n} floats, got {a.size}")
    return a, a.ctypes.data_as(ctypes.POINTER(N.f32))


def build_tri_batch(v0, v1, v2, offsets, params: N.BuildParams | None = None, root_boxes=None) -> TreeBatch:
    """Trees of len(offsets)-1 scenes whose triangles are concatenated in v0/v1/v2."""
    v0, v1, v2 = dev(v0), dev(v1), dev(v2)
    B = TreeBatch(N.KIND_TRI, offsets)
    if int(B.offsets[-1]) != v0.shape[0]:
        raise N.MnrtError("offsets do not cover the input")
    params = params or default_params(N.KIND_TRI)
    rb, rbp = _f32_ptr(root_boxes, 6, B.num_scenes)
    B._keep += [v0, v1, v2]
    N.check(N.load().mnrt_build_tri_batch(N.ptr(v0), N.ptr(v1), N.ptr(v2), B.offsets_ref(), B.num_scenes,
                                          ctypes.byref(params), rbp, B.ref(), N.stream_handle()),
            "mnrt_build_tri_batch")
    return B


def build_points_batch(pos, offsets, params: N.BuildParams | None = None, root_boxes=None,
                       max_query_radius=None) -> TreeBatch:
    pos = dev(pos)
    B = TreeBatch(N.KIND_POINT, offsets)
    if int(B.offsets[-1]) != pos.shape[0]:
        raise N.MnrtError("offsets do not cover the input")
    params = params or default_params(N.KIND_POINT)
    rb, rbp = _f32_ptr(root_boxes, 6, B.num_scenes)
    qr, qrp = _f32_ptr(max_query_radius, 1, B.num_scenes)
    N.check(N.load().mnrt_build_points_batch(N.ptr(pos), B.offsets_ref(), B.num_scenes, ctypes.byref(params), rbp,
                                             qrp, B.ref(), N.stream_handle()), "mnrt_build_points_batch")
    return B


def attach_photon_records_batch(B: TreeBatch, pos, power, dirs):
    pos, power, dirs = dev(pos), dev(power), dev(dirs)
    B._keep = [pos, power, dirs]  # a second call rewrites the records in place
    N.check(N.load().mnrt_tree_attach_photon_records_batch(B.ref(), B.num_scenes, N.ptr(pos), N.ptr(power),
                                                           N.ptr(dirs), B.offsets_ref(), N.stream_handle()),
            "mnrt_tree_attach_photon_records_batch")


class BuildGraph:
    """A batched build recorded into a hipGraph (mnrt_build_graph_*): `trees` (a TreeBatch
    over the graph's arrays) is rebuilt in place by replay(), from the input tensors given
    here (their contents may change, their sizes not).  status() waits for the stream and
    returns 0 when the last replay fit the graph's buffers, refreshing the trees' sizes,
    root boxes and addresses; nonzero means the trees are invalid."""

    def __init__(self, kind, offsets, v0, v1=None, v2=None, power=None, dirs=None,
                 params: N.BuildParams | None = None, root_boxes=None, max_query_radius=None, margin=0.25):
        ins = [dev(a) if a is not None else None for a in (v0, v1, v2, power, dirs)]
        self.trees = TreeBatch(kind, offsets)
        if int(self.trees.offsets[-1]) != ins[0].shape[0]:
            raise N.MnrtError("offsets do not cover the input")
        self._keep = [a for a in ins if a is not None]
        self.params = params or default_params(kind)
        ns = self.trees.num_scenes
        self._rb, rbp = _f32_ptr(root_boxes, 6, ns)
        self._qr, qrp = _f32_ptr(max_query_radius, 1, ns)
        p = lambda a: N.ptr(a) if a is not None else None
        d = N.BuildGraphDesc(kind, ns, self.trees.offsets_ref(), p(ins[0]), p(ins[1]), p(ins[2]), p(ins[3]),
                             p(ins[4]), ctypes.pointer(self.params), rbp, qrp, float(margin))
        self.h = ctypes.c_void_p()
        N.check(N.load().mnrt_build_graph_create(ctypes.byref(d), self.trees.ref(), ctypes.byref(self.h),
                                                 N.stream_handle()), "mnrt_build_graph_create")

    def replay(self, stream=None):
        N.check(N.load().mnrt_build_graph_replay(self.h, N.stream_handle(stream)), "mnrt_build_graph_replay")

    def status(self, stream=None) -> int:
        flags = N.u32(0)
        N.check(N.load().mnrt_build_graph_status(self.h, self.trees.ref(), ctypes.byref(flags),
                                                 N.stream_handle(stream)), "mnrt_build_graph_status")
        return int(flags.value)

    def close(self):
        if self.h:
            N.check(N.load().mnrt_build_graph_destroy(self.h), "mnrt_build_graph_destroy")
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _copy_from_ptr(dst, src_ptr: int, nbytes: int):
    torch = _torch()
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    rc = hip.hipMemcpyAsync(ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(src_ptr), nbytes, 3,
                            ctypes.c_void_p(N.stream_handle()))
    if rc != 0:
        raise N.MnrtError(f"hipMemcpyAsync failed ({rc})")
    torch.cuda.current_stream().synchronize()


def dev(a, dtype=None):
    """numpy -> contiguous cuda tensor (float32 / int32 views of uint32 data)."""
    torch = _torch()
    if isinstance(a, torch.Tensor):
        return a.contiguous().cuda()
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(a).cuda()


def tree_from_host(kind, nodes: np.ndarray, refs: np.ndarray, root6, num_elems: int, max_level: int) -> DeviceTree:
    """Wraps host-built node words/references (same layout) as a device tree."""
    T = DeviceTree(kind)
    dn = dev(np.asarray(nodes, np.uint32))
    dr = dev(np.asarray(refs, np.uint32) if len(refs) else np.zeros(1, np.uint32))
    T._keep += [dn, dr]
    t = T.t
    t.num_nodes = len(nodes) // 2
    t.num_refs = len(refs)
    t.num_elems = num_elems
    t.max_level = max_level
    for i in range(3):
        t.root_min[i] = float(root6[i])
        t.root_max[i] = float(root6[3 + i])
    t.d_nodes = dn.data_ptr()
    t.d_refs = dr.data_ptr()
    t.owns_nodes = 0
    t.owns_recs = 0
    return T


def export_preorder(T: DeviceTree):
    """Reference KDTreeData arrays of a tree (mnrt_tree_export_preorder), as device tensors."""
    torch = _torch()
    size = ctypes.c_uint64(0)
    N.check(N.load().mnrt_tree_preorder_size(T.ref(), ctypes.byref(size)), "mnrt_tree_preorder_size")
    n = T.num_nodes
    i32 = dict(dtype=torch.int32, device="cuda")
    out = dict(tree=torch.empty(max(size.value, 1), **i32), num_elems=torch.empty(n, **i32),
               addr=torch.empty(n, **i32), extent=torch.empty((n, 4), dtype=torch.float32, device="cuda"),
               left=torch.empty(n, **i32), right=torch.empty(n, **i32))
    N.check(N.load().mnrt_tree_export_preorder(T.ref(), N.ptr(out["tree"]), size.value, N.ptr(out["num_elems"]),
                                               N.ptr(out["addr"]), N.ptr(out["extent"]), N.ptr(out["left"]),
                                               N.ptr(out["right"]), N.stream_handle()), "mnrt_tree_export_preorder")
    out["tree"] = out["tree"][:size.value]
    return out


def set_custom_bits(X: dict, bit: int, values):
    values = dev(np.asarray(values, np.uint32))
    N.check(N.load().mnrt_preorder_set_custom_bits(N.ptr(X["tree"]), N.ptr(X["addr"]), X["addr"].shape[0], bit,
                                                   N.ptr(values), N.stream_handle()), "mnrt_preorder_set_custom_bits")


def attach_tri_records(T: DeviceTree, v0, v1, v2):
    v0, v1, v2 = dev(v0), dev(v1), dev(v2)
    T._keep += [v0, v1, v2]
    N.check(N.load().mnrt_tree_attach_tri_records(T.ref(), N.ptr(v0), N.ptr(v1), N.ptr(v2), N.stream_handle()),
            "mnrt_tree_attach_tri_records")


def attach_photon_records(T: DeviceTree, pos, power, dirs):
    pos, power, dirs = dev(pos), dev(power), dev(dirs)
    T._keep += [pos, power, dirs]
    N.check(N.load().mnrt_tree_attach_photon_records(T.ref(), N.ptr(pos), N.ptr(power), N.ptr(dirs),
                                                     N.stream_handle()), "mnrt_tree_attach_photon_records")


def default_params(kind: int, root=None, max_query_radius: float = 1.0) -> N.BuildParams:
    p = N.BuildParams()
    p.empty_space_ratio = 0.25 if kind == N.KIND_TRI else 0.1
    p.small_node_max = 64 if kind == N.KIND_TRI else 32
    p.max_query_radius = max_query_radius
    if root is not None:
        p.has_root_box = 1
        for i in range(3):
            p.root_min[i] = float(root[i])
            p.root_max[i] = float(root[3 + i])
    return p


def build_tri(v0, v1, v2, params: N.BuildParams | None = None) -> DeviceTree:
    v0, v1, v2 = dev(v0), dev(v1), dev(v2)
    T = DeviceTree(N.KIND_TRI)
    params = params or default_params(N.KIND_TRI)
    N.check(N.load().mnrt_build_tri(N.ptr(v0), N.ptr(v1), N.ptr(v2), v0.shape[0], ctypes.byref(params), T.ref(),
                                    N.stream_handle()), "mnrt_build_tri")
    return T


def build_points(pos, params: N.BuildParams | None = None) -> DeviceTree:
    pos = dev(pos)
    T = DeviceTree(N.KIND_POINT)
    params = params or default_params(N.KIND_POINT)
    N.check(N.load().mnrt_build_points(N.ptr(pos), pos.shape[0], ctypes.byref(params), T.ref(), N.stream_handle()),
            "mnrt_build_points")
    return T


def trace_closest(T: DeviceTree, orig, dirs, eps: float, want_pt=False):
    torch = _torch()
    orig, dirs = dev(orig), dev(dirs)
    n = orig.shape[0]
    tri = torch.empty(n, dtype=torch.int32, device="cuda")
    t = torch.empty(n, dtype=torch.float32, device="cuda")
    bary = torch.empty((n, 2), dtype=torch.float32, device="cuda")
    pt = torch.empty((n, 4), dtype=torch.float32, device="cuda") if want_pt else None
    N.check(N.load().mnrt_trace_closest(T.ref(), n, N.ptr(orig), N.ptr(dirs), eps, N.ptr(tri), N.ptr(t),
                                        N.ptr(bary), N.ptr(pt), None, N.stream_handle()), "mnrt_trace_closest")
    return tri, t, bary, pt


def trace_any(T: DeviceTree, orig, dirs, eps: float, tmax=None):
    """Any-hit test of ray segments (mnrt_trace_any): uint32 1 where a triangle is hit at
    eps < lambda < tmax (tmax None: unlimited)."""
    torch = _torch()
    orig, dirs = dev(orig), dev(dirs)
    tmax = dev(tmax) if tmax is not None else None
    n = orig.shape[0]
    hit = torch.empty(n, dtype=torch.int32, device="cuda")
    N.check(N.load().mnrt_trace_any(T.ref(), n, N.ptr(orig), N.ptr(dirs), N.ptr(tmax), eps, N.ptr(hit),
                                    N.stream_handle()), "mnrt_trace_any")
    return hit


def make_light(kind: int, position=(0.0, 0.0, 0.0), direction=(0.0, 0.0, -1.0), L_emit=(1.0, 1.0, 1.0),
               v1=(0.0, 0.0, 0.0), v2=(0.0, 0.0, 0.0), radius=0.0):
    """mnrt_light of one of N.LIGHT_POINT / LIGHT_DIRECTIONAL / LIGHT_AREA_DISC / LIGHT_AREA_RECT."""
    L = N.Light()
    L.type = kind
    for name, val in (("position", position), ("direction", direction), ("L_emit", L_emit), ("area_v1", v1),
                      ("area_v2", v2)):
        arr = getattr(L, name)
        for a in range(3):
            arr[a] = float(val[a])
    L.area_radius = float(radius)
    return L


def sample_area_light(light, rnd1, rnd2, idx_x=0.0, inv_x=1.0, idx_y=0.0, inv_y=1.0):
    """Points on an area light (mnrt_sample_area_light): (n, 4) float32."""
    torch = _torch()
    rnd1, rnd2 = dev(rnd1), dev(rnd2)
    n = rnd1.shape[0]
    out = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    N.check(N.load().mnrt_sample_area_light(ctypes.byref(light), n, idx_x, inv_x, idx_y, inv_y, N.ptr(rnd1),
                                            N.ptr(rnd2), N.ptr(out), N.stream_handle()), "mnrt_sample_area_light")
    return out


def trace_shadow(T: DeviceTree, light, pts, eps: float, samples=None):
    """Shadow rays (mnrt_trace_shadow): uint32 1 where the light (or its sample) is visible."""
    torch = _torch()
    pts = dev(pts)
    samples = dev(samples) if samples is not None else None
    n = pts.shape[0]
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    N.check(N.load().mnrt_trace_shadow(T.ref(), ctypes.byref(light), n, N.ptr(pts), N.ptr(samples), eps,
                                       N.ptr(out), N.stream_handle()), "mnrt_trace_shadow")
    return out


def knn_gather(P: DeviceTree, k: int, r_max: float, pos, nrm, dirs, infl, clr=None, idx_sp=None, io_rad=None,
               want_nbr=False, want_r2=False, want_irr=False, want_count=False, query_radius=None, orig=None,
               io_inv_dist=None):
    torch = _torch()
    pos, nrm, dirs, infl = dev(pos), dev(nrm), dev(dirs), dev(infl)
    clr = dev(clr) if clr is not None else None
    idx_sp = dev(idx_sp) if idx_sp is not None else None
    query_radius = dev(query_radius) if query_radius is not None else None
    orig = dev(orig) if orig is not None else None
    n = pos.shape[0]
    if io_rad is None:
        io_rad = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    nbr = torch.empty(n * k, dtype=torch.int32, device="cuda") if want_nbr else None
    r2 = torch.empty(n, dtype=torch.float32, device="cuda") if want_r2 else None
    irr = torch.empty((n, 4), dtype=torch.float32, device="cuda") if want_irr else None
    cnt = torch.empty(n, dtype=torch.int32, device="cuda") if want_count else None
    N.check(N.load().mnrt_knn_gather(P.ref(), k, r_max, n, N.ptr(pos), N.ptr(nrm), N.ptr(dirs), N.ptr(infl),
                                     N.ptr(clr), N.ptr(idx_sp), N.ptr(query_radius), N.ptr(orig), N.ptr(io_rad),
                                     N.ptr(io_inv_dist), N.ptr(r2), N.ptr(irr), N.ptr(nbr), N.ptr(cnt),
                                     N.stream_handle()), "mnrt_knn_gather")
    return dict(rad=io_rad, nbr=nbr.view(n, k) if nbr is not None else None, r2=r2, irr=irr, count=cnt,
                inv_dist=io_inv_dist)


def knn_gather_bwd(P: DeviceTree, k: int, r_max: float, pos, nrm, dirs, infl, clr, grad_rad, r2, count,
                   want_grad_clr=True, query_radius=None):
    """Adjoint of knn_gather (neighbour sets recomputed): (dLoss/dpower per photon, dLoss/dcolour per query)."""
    torch = _torch()
    n = pos.shape[0]
    gp = torch.empty((P.t.num_elems, 4), dtype=torch.float32, device="cuda")
    gc = torch.empty((n, 4), dtype=torch.float32, device="cuda") if want_grad_clr else None
    N.check(N.load().mnrt_knn_gather_bwd(P.ref(), k, r_max, n, N.ptr(dev(pos)), N.ptr(dev(nrm)), N.ptr(dev(dirs)),
                                         N.ptr(dev(infl)), N.ptr(dev(clr) if clr is not None else None),
                                         N.ptr(dev(query_radius) if query_radius is not None else None),
                                         N.ptr(grad_rad.contiguous()), N.ptr(r2), N.ptr(count), N.ptr(gp), N.ptr(gc),
                                         N.stream_handle()), "mnrt_knn_gather_bwd")
    return gp, gc


def knn_radii(P: DeviceTree, k: int, r_max: float, pos):
    torch = _torch()
    pos = dev(pos)
    out = torch.empty(pos.shape[0], dtype=torch.float32, device="cuda")
    N.check(N.load().mnrt_knn_radii(P.ref(), k, r_max, pos.shape[0], N.ptr(pos), N.ptr(out), N.stream_handle()),
            "mnrt_knn_radii")
    return out


def light_records(lights) -> np.ndarray:
    """(m, 4, 3) corner / edge u / edge v / power -> (m, 4, 4) float4 records."""
    lights = np.asarray(lights, np.float32).reshape(-1, 4, 3)
    rec = np.zeros((lights.shape[0], 4, 4), np.float32)
    rec[:, :, :3] = lights
    return rec


def photons_generate(T: DeviceTree, v0, v1, v2, albedo, lights, target: int, max_bounces=4, seed=0, eps=1e-4,
                     want_meta=False):
    """Traced photon map of `target` photons (RTCore::RebuildPhotonMaps)."""
    torch = _torch()
    v0, v1, v2, albedo = dev(v0), dev(v1), dev(v2), dev(albedo)
    lr = dev(light_records(lights))
    pos = torch.empty((target, 4), dtype=torch.float32, device="cuda")
    pw = torch.empty((target, 4), dtype=torch.float32, device="cuda")
    dr = torch.empty((target, 4), dtype=torch.float32, device="cuda")
    meta = torch.empty(target, dtype=torch.int32, device="cuda") if want_meta else None
    path = torch.empty(target, dtype=torch.int32, device="cuda") if want_meta else None
    tris = torch.empty((target, 4), dtype=torch.int32, device="cuda") if want_meta else None
    cnt, shot = N.u32(0), N.u32(0)
    N.check(N.load().mnrt_photons_generate(T.ref(), N.ptr(v0), N.ptr(v1), N.ptr(v2), N.ptr(albedo), N.ptr(lr),
                                           lr.shape[0], target, max_bounces, seed, eps, N.ptr(pos), N.ptr(pw),
                                           N.ptr(dr), N.ptr(meta), N.ptr(path), N.ptr(tris), ctypes.byref(cnt),
                                           ctypes.byref(shot), N.stream_handle()), "mnrt_photons_generate")
    n = cnt.value
    out = dict(pos=pos[:n], power=pw[:n], dir=dr[:n], count=n, paths=shot.value)
    if want_meta:
        out.update(meta=meta[:n], path=path[:n], tris=tris[:n])
    return out


def _i32(t):
    """int32 view of an index tensor (0xFFFFFFFF = -1 for "none")."""
    torch = _torch()
    return t if t.dtype == torch.int32 and t.is_contiguous() else t.to(torch.int32).contiguous()


def photons_reweight(base, light, path_tris, emission, albedo, hit_tri=None, io_clr=None, out_pow=None):
    """Photon powers (P, 4) = base * emission[light] * albedo products along each path
    (mnrt_photons_reweight); with hit_tri / io_clr also writes io_clr[q].rgb =
    albedo[hit_tri[q]] for hit_tri[q] >= 0.  emission (L, 3), albedo (T, 3) float32."""
    torch = _torch()
    np_ = base.shape[0]
    light, path_tris = _i32(light), _i32(path_tris)
    emission, albedo = emission.detach().float().contiguous(), albedo.detach().float().contiguous()
    out = out_pow if out_pow is not None else torch.empty((max(np_, 1), 4), dtype=torch.float32, device="cuda")
    nq = 0 if hit_tri is None else hit_tri.shape[0]
    ht = _i32(hit_tri) if hit_tri is not None else None
    N.check(N.load().mnrt_photons_reweight(np_, N.ptr(base), N.ptr(light), N.ptr(path_tris), emission.shape[0],
                                           N.ptr(emission), albedo.shape[0], N.ptr(albedo), nq, N.ptr(ht),
                                           N.ptr(out), N.ptr(io_clr) if nq else None, N.stream_handle()),
            "mnrt_photons_reweight")
    return out[:np_]


def photons_reweight_bwd(base, light, path_tris, emission, albedo, grad_pow, hit_tri=None, grad_clr=None):
    """(dLoss/demission (L, 3), dLoss/dalbedo (T, 3)) from dLoss/dpower (P, 4) and, with
    hit_tri, dLoss/dclr (Q, 4) (mnrt_photons_reweight_bwd; fixed-point sums, same bits on
    every run)."""
    torch = _torch()
    np_ = base.shape[0]
    light, path_tris = _i32(light), _i32(path_tris)
    emission, albedo = emission.detach().float().contiguous(), albedo.detach().float().contiguous()
    ge = torch.empty_like(emission)
    ga = torch.empty_like(albedo)
    nq = 0 if hit_tri is None else hit_tri.shape[0]
    ht = _i32(hit_tri) if hit_tri is not None else None
    gc = grad_clr.contiguous() if nq else None
    N.check(N.load().mnrt_photons_reweight_bwd(np_, N.ptr(base), N.ptr(light), N.ptr(path_tris),
                                               emission.shape[0], N.ptr(emission), albedo.shape[0], N.ptr(albedo),
                                               nq, N.ptr(ht), N.ptr(grad_pow.contiguous()), N.ptr(gc), N.ptr(ge),
                                               N.ptr(ga), N.stream_handle()), "mnrt_photons_reweight_bwd")
    return ge, ga


def photons_spherical_to_dir(pos, power):
    """Directions (P, 4) from the reference's PhotonData angles: azimuth phi in pos.w,
    polar theta in power.w (mnrt_photons_spherical_to_dir)."""
    torch = _torch()
    pos, power = dev(pos), dev(power)
    n = pos.shape[0]
    out = torch.empty((max(n, 1), 4), dtype=torch.float32, device="cuda")
    N.check(N.load().mnrt_photons_spherical_to_dir(n, N.ptr(pos), N.ptr(power), N.ptr(out), N.stream_handle()),
            "mnrt_photons_spherical_to_dir")
    return out[:n]


def photons_dir_to_spherical(dirs, io_pos, io_power):
    """Writes the angles of unit directions into io_pos.w (phi) and io_power.w (theta)
    (mnrt_photons_dir_to_spherical)."""
    dirs = dev(dirs)
    N.check(N.load().mnrt_photons_dir_to_spherical(dirs.shape[0], N.ptr(dirs), N.ptr(io_pos), N.ptr(io_power),
                                                   N.stream_handle()), "mnrt_photons_dir_to_spherical")


def scan_u32(a, inclusive=False, out=None):
    """Add-scan of an int32/uint32 tensor (mnrt_scan_u32): (scan, total)."""
    torch = _torch()
    a = a.contiguous()
    out = out if out is not None else torch.empty_like(a)
    tot = torch.empty(1, dtype=torch.int32, device="cuda")
    N.check(N.load().mnrt_scan_u32(N.ptr(a) if a.numel() else None, a.numel(), 1 if inclusive else 0,
                                   N.ptr(out) if a.numel() else None, N.ptr(tot), N.stream_handle()), "mnrt_scan_u32")
    return out, tot


def scan_u32_counted(a, count: int, capacity: int, end_total=False, out=None):
    """Exclusive add-scan of the first min(count, capacity) words of a, the count handed over
    in a device word (mnrt_scan_u32_counted): (scan, total); out (default: capacity + 1 words
    of -1) may be a itself."""
    torch = _torch()
    a = a.contiguous()
    out = out if out is not None else torch.full((capacity + 1,), -1, dtype=torch.int32, device="cuda")
    tot = torch.empty(1, dtype=torch.int32, device="cuda")
    cnt = torch.from_numpy(np.array([count], np.uint32).view(np.int32)).cuda()
    N.check(N.load().mnrt_scan_u32_counted(N.ptr(a) if a.numel() else None, N.ptr(cnt), capacity, N.ptr(out),
                                           N.ptr(tot), 1 if end_total else 0, N.stream_handle()),
            "mnrt_scan_u32_counted")
    return out, tot


def compact_u32(a, valid):
    """The words of a where valid != 0, in order (mnrt_compact_u32): (compacted, count)."""
    torch = _torch()
    a, valid = a.contiguous(), valid.contiguous()
    out = torch.empty(max(a.numel(), 1), dtype=a.dtype, device="cuda")
    cnt = torch.empty(1, dtype=torch.int32, device="cuda")
    n = a.numel()
    N.check(N.load().mnrt_compact_u32(N.ptr(a) if n else None, N.ptr(valid) if n else None, n,
                                      N.ptr(out) if n else None, N.ptr(cnt), N.stream_handle()), "mnrt_compact_u32")
    return out, cnt


def _no_photons():
    torch = _torch()
    z4 = torch.empty((0, 4), dtype=torch.float32, device="cuda")
    zi = torch.empty((0, 4), dtype=torch.int32, device="cuda")
    return z4, zi[:, 0].contiguous(), zi, torch.empty((0, 3), dtype=torch.float32, device="cuda")


def hit_colours(albedo, hit_tri, io_clr):
    """io_clr[q].rgb = albedo[hit_tri[q]] for hit_tri[q] >= 0 (mnrt_photons_reweight
    without photons)."""
    b, l, t, e = _no_photons()
    photons_reweight(b, l, t, e, albedo, hit_tri=hit_tri, io_clr=io_clr)
    return io_clr


def hit_colours_bwd(albedo, hit_tri, grad_clr):
    """dLoss/dalbedo (T, 3) from dLoss/dclr (Q, 4) at hits on hit_tri."""
    b, l, t, e = _no_photons()
    return photons_reweight_bwd(b, l, t, e, albedo, b, hit_tri=hit_tri, grad_clr=grad_clr)[1]


def shading_points(orig, dirs, tri, t, v0, v1, v2, albedo, infl=None, want_src=False):
    """Compacted shading points at the valid hits."""
    torch = _torch()
    n = orig.shape[0]
    outs = [torch.empty((max(n, 1), 4), dtype=torch.float32, device="cuda") for _ in range(5)]
    src = torch.empty(max(n, 1), dtype=torch.int32, device="cuda") if want_src else None
    otri = torch.empty(max(n, 1), dtype=torch.int32, device="cuda") if want_src else None
    cnt = N.u32(0)
    N.check(N.load().mnrt_shading_points(n, N.ptr(orig), N.ptr(dirs), N.ptr(tri), N.ptr(t), N.ptr(v0), N.ptr(v1),
                                         N.ptr(v2), N.ptr(albedo), N.ptr(infl), *[N.ptr(o) for o in outs],
                                         N.ptr(src), N.ptr(otri), ctypes.byref(cnt), N.stream_handle()),
            "mnrt_shading_points")
    m = cnt.value
    q = dict(pos=outs[0][:m], nrm=outs[1][:m], dir=outs[2][:m], clr=outs[3][:m], infl=outs[4][:m], count=m)
    if want_src:
        q.update(src=src[:m], tri=otri[:m])
    return q


def primary_rays(width: int, height: int, eye=(0.0, 0.0, 0.9), target=(0.0, 0.0, 0.0), vfov_deg=90.0, tile=8,
                 out_orig=None, out_dir=None, want_pixel=False):
    """Camera rays on the device, equal to scene.camera_rays (mnrt_primary_rays)."""
    from . import scene as S

    torch = _torch()
    n = width * height
    cam = np.ascontiguousarray(S.camera_basis(width, height, eye, target, vfov_deg))
    o = out_orig if out_orig is not None else torch.empty((max(n, 1), 4), dtype=torch.float32, device="cuda")
    d = out_dir if out_dir is not None else torch.empty((max(n, 1), 4), dtype=torch.float32, device="cuda")
    pix = torch.empty(max(n, 1), dtype=torch.int32, device="cuda") if want_pixel else None
    N.check(N.load().mnrt_primary_rays(width, height, tile, cam.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       N.ptr(o), N.ptr(d), N.ptr(pix), N.stream_handle()), "mnrt_primary_rays")
    return o[:n], d[:n], (pix[:n] if pix is not None else None)


def bounce_rays(q, seed: int):
    torch = _torch()
    n = q["pos"].shape[0]
    o = torch.empty((max(n, 1), 4), dtype=torch.float32, device="cuda")
    d = torch.empty((max(n, 1), 4), dtype=torch.float32, device="cuda")
    infl = torch.empty((max(n, 1), 4), dtype=torch.float32, device="cuda")
    N.check(N.load().mnrt_bounce_rays(n, N.ptr(q["pos"]), N.ptr(q["nrm"]), N.ptr(q["dir"]), N.ptr(q["clr"]), seed,
                                      N.ptr(o), N.ptr(d), N.ptr(infl), N.stream_handle()), "mnrt_bounce_rays")
    return o[:n], d[:n], infl[:n]


class QueryBatch:
    """Many scenes traced, shaded and queried together (mnrt_batch_*).

    tri_trees / photon_maps: TreeBatch or list of DeviceTree (photon_maps None: trace
    only).  tri_offsets: (S+1) triangle offsets into the concatenated vertex / albedo
    arrays; ray_offsets: (S+1) ray slot offsets; eps, r_max: per scene."""

    def __init__(self, tri_trees, tri_offsets, ray_offsets, eps, photon_maps=None, r_max=None, k: int = 50):
        self.S = len(ray_offsets) - 1
        self._tri_arr = self._tree_array(tri_trees)
        self._pm_arr = self._tree_array(photon_maps) if photon_maps is not None else None
        self.tri_offsets = np.ascontiguousarray(tri_offsets, np.uint32)
        self.ray_offsets = np.ascontiguousarray(ray_offsets, np.uint32)
        self.eps = np.ascontiguousarray(eps, np.float32)
        self.r_max = np.ascontiguousarray(r_max if r_max is not None else np.ones(self.S), np.float32)
        self.k = k
        self._refs = (tri_trees, photon_maps)
        d = N.BatchDesc()
        d.num_scenes = self.S
        d.tri_trees = ctypes.cast(self._tri_arr, ctypes.POINTER(N.Tree))
        d.photon_maps = ctypes.cast(self._pm_arr, ctypes.POINTER(N.Tree)) if self._pm_arr is not None else None
        d.tri_offsets = self.tri_offsets.ctypes.data_as(ctypes.POINTER(N.u32))
        d.ray_offsets = self.ray_offsets.ctypes.data_as(ctypes.POINTER(N.u32))
        d.eps = self.eps.ctypes.data_as(ctypes.POINTER(N.f32))
        d.r_max = self.r_max.ctypes.data_as(ctypes.POINTER(N.f32))
        d.k = k
        h = ctypes.c_void_p()
        N.check(N.load().mnrt_batch_create(ctypes.byref(d), ctypes.byref(h)), "mnrt_batch_create")
        self.h = h
        self._io = None
        self._io_keep = None

    @staticmethod
    def _tree_array(trees):
        if isinstance(trees, TreeBatch):
            return trees.arr
        arr = (N.Tree * len(trees))()
        for i, t in enumerate(trees):
            arr[i] = t.t
        return arr

    def update_trees(self, tri_trees, photon_maps=None):
        """Points the batch at rebuilt trees of the same scenes (captured graphs stay valid)."""
        self._tri_arr = self._tree_array(tri_trees)
        self._pm_arr = self._tree_array(photon_maps) if photon_maps is not None else None
        self._refs = (tri_trees, photon_maps)
        N.check(N.load().mnrt_batch_update_trees(
            self.h, ctypes.cast(self._tri_arr, ctypes.POINTER(N.Tree)),
            ctypes.cast(self._pm_arr, ctypes.POINTER(N.Tree)) if self._pm_arr is not None else None),
            "mnrt_batch_update_trees")

    @property
    def photon_offsets(self) -> np.ndarray:
        """(S+1) offsets of the scenes' photons in the concatenated gradient d_grad_pow4."""
        off = np.zeros(self.S + 1, np.int64)
        if self._pm_arr is not None:
            off[1:] = np.cumsum([self._pm_arr[s].num_elems for s in range(self.S)])
        return off

    def make_io(self, **bufs):
        """BatchIO from named device tensors (names as mnrt_batch_io without d_)."""
        io = N.BatchIO()
        keep = {}
        for name in N.BATCH_IO_FIELDS:
            t = bufs.pop(name[2:], None)
            if t is not None:
                keep[name] = t
                setattr(io, name, N.ptr(t))
        if bufs:
            raise N.MnrtError(f"unknown batch buffers: {sorted(bufs)}")
        self._io, self._io_keep = io, keep
        return io

    def run(self, phases: int, stream=None):
        N.check(N.load().mnrt_batch_run(self.h, ctypes.byref(self._io), phases, N.stream_handle(stream)),
                "mnrt_batch_run")

    def capture(self, phases: int, stream=None):
        N.check(N.load().mnrt_batch_capture(self.h, ctypes.byref(self._io), phases, N.stream_handle(stream)),
                "mnrt_batch_capture")

    def replay(self, phases: int, stream=None):
        N.check(N.load().mnrt_batch_replay(self.h, phases, N.stream_handle(stream)), "mnrt_batch_replay")

    def autotune(self, max_slots: int = 0, stream=None):
        N.check(N.load().mnrt_batch_autotune(self.h, ctypes.byref(self._io), max_slots, N.stream_handle(stream)),
                "mnrt_batch_autotune")
        return self.tuning

    @property
    def tuning(self) -> dict:
        t = N.BatchTuning()
        N.check(N.load().mnrt_batch_get_tuning(self.h, ctypes.byref(t)), "mnrt_batch_get_tuning")
        return t.as_dict()

    @tuning.setter
    def tuning(self, d: dict):
        t = N.BatchTuning()
        N.check(N.load().mnrt_batch_get_tuning(self.h, ctypes.byref(t)), "mnrt_batch_get_tuning")
        for k_, v in d.items():
            setattr(t, k_, int(v))
        N.check(N.load().mnrt_batch_set_tuning(self.h, ctypes.byref(t)), "mnrt_batch_set_tuning")

    def close(self):
        if self.h:
            _torch().cuda.synchronize()
            N.check(N.load().mnrt_batch_destroy(self.h), "mnrt_batch_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
