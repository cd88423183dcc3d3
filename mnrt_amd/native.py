"""ctypes binding of libmnrt_kd.so (the C ABI declared in include/mnrt_kd.h).

Device memory and streams come from PyTorch (torch.cuda tensors and streams); the
library sees only raw device pointers.  There is no fallback: if the library is missing
or the device is not a gfx950, every call raises.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmnrt_kd.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "mnrt_kd.h")

KIND_TRI = 0
KIND_POINT = 1
MAX_LEVEL = 48
KNN_MAX_K = 256

u32 = ctypes.c_uint32
f32 = ctypes.c_float
vp = ctypes.c_void_p


class BuildParams(ctypes.Structure):
    """mnrt_build_params (KDTreeGPU::Set* parameters, KDTreeGPU.h:248-280)."""

    _fields_ = [
        ("empty_space_ratio", f32),
        ("small_node_max", u32),
        ("max_query_radius", f32),
        ("has_root_box", u32),
        ("root_min", f32 * 3),
        ("root_max", f32 * 3),
    ]


class Light(ctypes.Structure):
    """mnrt_light (LightData, KernelDefs.h:307-323)."""

    _fields_ = [
        ("type", u32),
        ("position", f32 * 3),
        ("direction", f32 * 3),
        ("L_emit", f32 * 3),
        ("area_v1", f32 * 3),
        ("area_v2", f32 * 3),
        ("area_radius", f32),
    ]


LIGHT_POINT, LIGHT_DIRECTIONAL, LIGHT_AREA_DISC, LIGHT_AREA_RECT = 0, 1, 2, 3


class BuildGraphDesc(ctypes.Structure):
    """mnrt_build_graph_desc: a batched build recorded into a hipGraph."""

    _fields_ = [
        ("kind", ctypes.c_int),
        ("num_scenes", u32),
        ("h_offsets", ctypes.POINTER(u32)),
        ("d_v0", vp),
        ("d_v1", vp),
        ("d_v2", vp),
        ("d_power4", vp),
        ("d_dir4", vp),
        ("params", ctypes.POINTER(BuildParams)),
        ("h_root_boxes", ctypes.POINTER(f32)),
        ("h_max_query_radius", ctypes.POINTER(f32)),
        ("margin", f32),
    ]


class Tree(ctypes.Structure):
    """mnrt_tree: device-resident packed preorder kd-tree."""

    _fields_ = [
        ("kind", u32),
        ("num_nodes", u32),
        ("num_refs", u32),
        ("num_elems", u32),
        ("num_leaves", u32),
        ("max_level", u32),
        ("root_min", f32 * 3),
        ("root_max", f32 * 3),
        ("d_nodes", vp),
        ("d_refs", vp),
        ("d_recs", vp),
        ("owns_nodes", u32),
        ("owns_recs", u32),
        ("d_qnodes", vp),
        ("qleaf_max", u32),
        ("owns_qnodes", u32),
    ]


class BatchDesc(ctypes.Structure):
    """mnrt_batch_desc."""

    _fields_ = [
        ("num_scenes", u32),
        ("tri_trees", ctypes.POINTER(Tree)),
        ("photon_maps", ctypes.POINTER(Tree)),
        ("tri_offsets", ctypes.POINTER(u32)),
        ("ray_offsets", ctypes.POINTER(u32)),
        ("eps", ctypes.POINTER(f32)),
        ("r_max", ctypes.POINTER(f32)),
        ("k", u32),
    ]


class BatchTuning(ctypes.Structure):
    """mnrt_batch_tuning."""

    _fields_ = [(n, u32) for n in ("trace_block", "trace_tile", "trace_grid", "trace_stack", "knn_waves",
                                   "knn_chunk", "knn_grid", "knn_cap", "parts", "knn_select")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


BATCH_IO_FIELDS = ("d_orig4", "d_dir4", "d_v0", "d_v1", "d_v2", "d_albedo4", "d_infl4", "d_hit_tri", "d_hit_t",
                   "d_hit_bary2", "d_sp_pos4", "d_sp_nrm4", "d_sp_dir4", "d_sp_clr4", "d_sp_infl4", "d_sp_src",
                   "d_sp_count", "d_io_rad4", "d_out_r2", "d_out_nbr", "d_out_count", "d_grad_rad4",
                   "d_grad_pow4", "d_grad_clr4")


class BatchIO(ctypes.Structure):
    """mnrt_batch_io (device pointers)."""

    _fields_ = [(n, vp) for n in BATCH_IO_FIELDS]


PHASE_TRACE, PHASE_SHADE, PHASE_KNN, PHASE_KNN_BWD = 1, 2, 4, 8

# MNRT_ABI_VERSION of the header these declarations mirror
ABI_VERSION = 5


# status codes (the MNRT_OK / MNRT_ERR_* values of mnrt_kd.h)
OK, ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_NODEV, ERR_CAPACITY, ERR_STATE = 0, -1, -2, -3, -4, -5, -6


class MnrtError(RuntimeError):
    pass


_lib = None


def load():
    """Loads libmnrt_kd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MnrtError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() (make -C mnrt_amd/csrc)")
    L = ctypes.CDLL(LIB_PATH)
    L.mnrt_status_string.restype = ctypes.c_char_p
    L.mnrt_status_string.argtypes = [ctypes.c_int]
    L.mnrt_abi_version.restype = ctypes.c_int
    if L.mnrt_abi_version() != ABI_VERSION:
        raise MnrtError(f"{LIB_PATH} has ABI version {L.mnrt_abi_version()}, these bindings expect {ABI_VERSION}: "
                        "rebuild it with __graft_entry__.build()")
    L.mnrt_device_init.argtypes = [ctypes.c_int]
    TP = ctypes.POINTER(Tree)
    L.mnrt_build_tri.argtypes = [vp, vp, vp, u32, ctypes.POINTER(BuildParams), TP, vp]
    L.mnrt_build_points.argtypes = [vp, u32, ctypes.POINTER(BuildParams), TP, vp]
    L.mnrt_build_tri_batch.argtypes = [vp, vp, vp, ctypes.POINTER(u32), u32, ctypes.POINTER(BuildParams),
                                       ctypes.POINTER(f32), TP, vp]
    L.mnrt_build_points_batch.argtypes = [vp, ctypes.POINTER(u32), u32, ctypes.POINTER(BuildParams),
                                          ctypes.POINTER(f32), ctypes.POINTER(f32), TP, vp]
    L.mnrt_tree_attach_tri_records_batch.argtypes = [TP, u32, vp, vp, vp, ctypes.POINTER(u32), vp]
    L.mnrt_tree_attach_photon_records_batch.argtypes = [TP, u32, vp, vp, vp, ctypes.POINTER(u32), vp]
    L.mnrt_build_graph_create.argtypes = [ctypes.POINTER(BuildGraphDesc), TP, ctypes.POINTER(vp), vp]
    L.mnrt_build_graph_replay.argtypes = [vp, vp]
    L.mnrt_build_graph_status.argtypes = [vp, TP, ctypes.POINTER(u32), vp]
    L.mnrt_build_graph_destroy.argtypes = [vp]
    BP = ctypes.POINTER(ctypes.c_void_p)
    L.mnrt_batch_create.argtypes = [ctypes.POINTER(BatchDesc), BP]
    L.mnrt_batch_destroy.argtypes = [vp]
    L.mnrt_batch_get_tuning.argtypes = [vp, ctypes.POINTER(BatchTuning)]
    L.mnrt_batch_set_tuning.argtypes = [vp, ctypes.POINTER(BatchTuning)]
    L.mnrt_batch_autotune.argtypes = [vp, ctypes.POINTER(BatchIO), u32, vp]
    L.mnrt_batch_run.argtypes = [vp, ctypes.POINTER(BatchIO), u32, vp]
    L.mnrt_batch_update_trees.argtypes = [vp, TP, TP]
    L.mnrt_batch_capture.argtypes = [vp, ctypes.POINTER(BatchIO), u32, vp]
    L.mnrt_batch_replay.argtypes = [vp, u32, vp]
    L.mnrt_tree_attach_tri_records.argtypes = [TP, vp, vp, vp, vp]
    L.mnrt_tree_attach_photon_records.argtypes = [TP, vp, vp, vp, vp]
    L.mnrt_tree_release.argtypes = [TP]
    L.mnrt_tree_collapse.argtypes = [TP, u32, vp]
    L.mnrt_tree_cost.argtypes = [TP, f32, ctypes.POINTER(ctypes.c_double), vp]
    L.mnrt_trace_closest.argtypes = [TP, u32, vp, vp, f32, vp, vp, vp, vp, vp, vp]
    LP = ctypes.POINTER(Light)
    L.mnrt_trace_any.argtypes = [TP, u32, vp, vp, vp, f32, vp, vp]
    L.mnrt_sample_area_light.argtypes = [LP, u32, f32, f32, f32, f32, vp, vp, vp, vp]
    L.mnrt_trace_shadow.argtypes = [TP, LP, u32, vp, vp, f32, vp, vp]
    L.mnrt_knn_gather.argtypes = [TP, u32, f32, u32] + [vp] * 15
    L.mnrt_knn_gather_noalloc.argtypes = [TP, u32, f32, u32] + [vp] * 15
    L.mnrt_knn_radii.argtypes = [TP, u32, f32, u32, vp, vp, vp]
    L.mnrt_knn_gather_bwd.argtypes = [TP, u32, f32, u32] + [vp] * 12
    L.mnrt_photons_generate.argtypes = [TP, vp, vp, vp, vp, vp, u32, u32, u32, u32, f32, vp, vp, vp, vp, vp, vp,
                                        ctypes.POINTER(u32), ctypes.POINTER(u32), vp]
    L.mnrt_photons_reweight.argtypes = [u32, vp, vp, vp, u32, vp, u32, vp, u32, vp, vp, vp, vp]
    L.mnrt_photons_reweight_bwd.argtypes = [u32, vp, vp, vp, u32, vp, u32, vp, u32, vp, vp, vp, vp, vp, vp]
    L.mnrt_photons_spherical_to_dir.argtypes = [u32, vp, vp, vp, vp]
    L.mnrt_photons_dir_to_spherical.argtypes = [u32, vp, vp, vp, vp]
    L.mnrt_scan_u32.argtypes = [vp, u32, ctypes.c_int, vp, vp, vp]
    L.mnrt_scan_u32_counted.argtypes = [vp, vp, u32, vp, vp, ctypes.c_int, vp]
    L.mnrt_compact_u32.argtypes = [vp, vp, u32, vp, vp, vp]
    L.mnrt_shading_points.argtypes = [u32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                      ctypes.POINTER(u32), vp]
    L.mnrt_bounce_rays.argtypes = [u32, vp, vp, vp, vp, u32, vp, vp, vp, vp]
    L.mnrt_tree_preorder_size.argtypes = [TP, ctypes.POINTER(ctypes.c_uint64)]
    L.mnrt_tree_export_preorder.argtypes = [TP, vp, ctypes.c_uint64, vp, vp, vp, vp, vp, vp]
    L.mnrt_preorder_set_custom_bits.argtypes = [vp, vp, u32, u32, vp, vp]
    L.mnrt_primary_rays.argtypes = [u32, u32, u32, ctypes.POINTER(ctypes.c_double), vp, vp, vp, vp]
    for name in ("mnrt_tree_preorder_size", "mnrt_tree_export_preorder", "mnrt_preorder_set_custom_bits",
                 "mnrt_primary_rays", "mnrt_batch_update_trees", "mnrt_batch_create", "mnrt_batch_destroy", "mnrt_batch_get_tuning", "mnrt_batch_set_tuning",
                 "mnrt_batch_autotune", "mnrt_batch_run", "mnrt_batch_capture", "mnrt_batch_replay",
                 "mnrt_build_tri_batch", "mnrt_build_points_batch", "mnrt_tree_attach_tri_records_batch",
                 "mnrt_tree_attach_photon_records_batch", "mnrt_knn_gather_bwd", "mnrt_tree_collapse", "mnrt_photons_generate", "mnrt_shading_points", "mnrt_bounce_rays", "mnrt_build_tri", "mnrt_build_points", "mnrt_tree_attach_tri_records",
                 "mnrt_tree_attach_photon_records", "mnrt_tree_release", "mnrt_tree_cost",
                 "mnrt_trace_closest", "mnrt_trace_any", "mnrt_sample_area_light", "mnrt_trace_shadow", "mnrt_knn_gather", "mnrt_knn_radii", "mnrt_device_init"):
        getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = load().mnrt_status_string(rc).decode()
        raise MnrtError(f"{what}: {msg} ({rc})")


def ptr(t) -> int | None:
    """Device pointer of a torch tensor (None passes NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise MnrtError("expected a device tensor")
    if not t.is_contiguous():
        raise MnrtError("expected a contiguous tensor")
    return t.data_ptr()


def stream_handle(stream=None) -> int | None:
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def exported_symbols_in_header() -> list[str]:
    """Names of the functions include/mnrt_kd.h declares."""
    import re

    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(mnrt_\w+)\s*\(", text, re.M)))
