"""The C ABI library loads without a GPU, exports every function include/mnrt_kd.h
declares, and rejects bad arguments before touching a device."""

import ctypes

import pytest

from mnrt_amd import native as N


def test_exports_every_declared_symbol():
    L = N.load()
    names = N.exported_symbols_in_header()
    assert len(names) >= 15
    for name in names:
        assert hasattr(L, name), name


def test_status_strings():
    L = N.load()
    assert L.mnrt_status_string(0) == b"ok"
    assert L.mnrt_status_string(-1) == b"invalid argument"
    assert L.mnrt_abi_version() == N.ABI_VERSION == 5


def test_argument_errors_without_device():
    L = N.load()
    t = N.Tree()
    assert L.mnrt_build_tri(None, None, None, 10, None, ctypes.byref(t), None) == -1
    p = N.BuildParams()
    p.small_node_max = 65
    assert L.mnrt_build_points(None, 0, ctypes.byref(p), ctypes.byref(t), None) == -1
    assert L.mnrt_trace_closest(None, 1, None, None, 0.0, None, None, None, None, None, None) == -1
    assert L.mnrt_tree_release(None) == -1
    t.kind = N.KIND_POINT
    assert L.mnrt_knn_gather(ctypes.byref(t), 0, 1.0, 1, *([None] * 15)) == -1
    assert L.mnrt_knn_gather(ctypes.byref(t), 5, 1.0, 1, *([None] * 15)) == -6  # no nodes yet
    t.kind = N.KIND_TRI
    assert L.mnrt_trace_closest(ctypes.byref(t), 1, None, None, 0.0, None, None, None, None, None, None) == -1
    # any-hit / shadow entry points: argument and state errors are found before any device call
    fake = 256  # never dereferenced: every call below returns before launching
    light = N.Light()
    light.type = N.LIGHT_POINT
    assert L.mnrt_trace_any(None, 1, fake, fake, None, 0.0, fake, None) == N.ERR_ARG
    assert L.mnrt_trace_any(ctypes.byref(t), 1, None, fake, None, 0.0, fake, None) == N.ERR_ARG
    assert L.mnrt_trace_any(ctypes.byref(t), 1, fake, fake, None, 0.0, fake, None) == N.ERR_STATE
    assert L.mnrt_trace_shadow(ctypes.byref(t), None, 1, fake, None, 0.0, fake, None) == N.ERR_ARG
    assert L.mnrt_trace_shadow(ctypes.byref(t), ctypes.byref(light), 1, fake, None, 0.0, fake, None) == N.ERR_STATE
    light.type = N.LIGHT_AREA_DISC
    assert L.mnrt_trace_shadow(ctypes.byref(t), ctypes.byref(light), 1, fake, None, 0.0, fake, None) == N.ERR_ARG
    light.type = 9
    assert L.mnrt_trace_shadow(ctypes.byref(t), ctypes.byref(light), 1, fake, fake, 0.0, fake, None) == N.ERR_ARG
    light.type = N.LIGHT_POINT
    assert L.mnrt_sample_area_light(ctypes.byref(light), 1, 0.0, 1.0, 0.0, 1.0, fake, fake, fake, None) == N.ERR_ARG
    light.type = N.LIGHT_AREA_RECT
    assert L.mnrt_sample_area_light(ctypes.byref(light), 1, 0.0, 1.0, 0.0, 1.0, None, fake, fake, None) == N.ERR_ARG
    t.kind = N.KIND_POINT
    light.type = N.LIGHT_POINT
    assert L.mnrt_trace_shadow(ctypes.byref(t), ctypes.byref(light), 1, fake, None, 0.0, fake, None) == N.ERR_STATE
    # photon trees of 2^29 nodes or more (node words addressed by 32-bit byte offsets)
    t.d_nodes, t.d_recs, t.max_level = fake, fake, 20
    t.num_nodes = (1 << 29) - 1
    assert L.mnrt_knn_gather(ctypes.byref(t), 5, 1.0, 0, *([None] * 15)) == N.OK  # n = 0: nothing to do
    t.num_nodes = 1 << 29
    assert L.mnrt_knn_gather(ctypes.byref(t), 5, 1.0, 1, *([None] * 15)) == N.ERR_CAPACITY
    assert L.mnrt_knn_radii(ctypes.byref(t), 5, 1.0, 1, fake, fake, None) == N.ERR_CAPACITY


def test_reweight_argument_errors_without_device():
    L = N.load()
    fake = 256  # never dereferenced: every call below returns before launching
    # photons but no base / light / path arrays, no emission, or no output
    assert L.mnrt_photons_reweight(4, None, fake, fake, 1, fake, 1, fake, 0, None, fake, None, None) == N.ERR_ARG
    assert L.mnrt_photons_reweight(4, fake, fake, fake, 0, fake, 1, fake, 0, None, fake, None, None) == N.ERR_ARG
    assert L.mnrt_photons_reweight(4, fake, fake, fake, 1, fake, 1, fake, 0, None, None, None, None) == N.ERR_ARG
    # hits without hit triangles or without a colour buffer
    assert L.mnrt_photons_reweight(0, None, None, None, 0, None, 1, fake, 3, None, None, fake, None) == N.ERR_ARG
    assert L.mnrt_photons_reweight(0, None, None, None, 0, None, 1, fake, 3, fake, None, None, None) == N.ERR_ARG
    assert L.mnrt_photons_reweight(0, None, None, None, 0, None, 0, None, 0, None, None, None, None) == N.OK
    # adjoint: missing upstream gradients or outputs
    assert L.mnrt_photons_reweight_bwd(4, fake, fake, fake, 1, fake, 1, fake, 0, None, None, None, fake, fake,
                                       None) == N.ERR_ARG
    assert L.mnrt_photons_reweight_bwd(4, fake, fake, fake, 1, fake, 1, fake, 0, None, fake, None, None, fake,
                                       None) == N.ERR_ARG
    assert L.mnrt_photons_reweight_bwd(4, fake, fake, fake, 1, fake, 1, fake, 2, fake, fake, None, fake, fake,
                                       None) == N.ERR_ARG


def test_primitive_argument_errors_without_device():
    L = N.load()
    fake = 256
    assert L.mnrt_scan_u32(None, 5, 0, fake, None, None) == N.ERR_ARG
    assert L.mnrt_scan_u32(fake, 5, 1, None, None, None) == N.ERR_ARG
    assert L.mnrt_scan_u32_counted(fake, None, 5, fake, None, 0, None) == N.ERR_ARG  # no count word
    assert L.mnrt_scan_u32_counted(None, fake, 5, fake, None, 0, None) == N.ERR_ARG
    assert L.mnrt_scan_u32_counted(fake, fake, 5, None, None, 0, None) == N.ERR_ARG
    assert L.mnrt_scan_u32_counted(None, fake, 0, None, None, 1, None) == N.ERR_ARG  # end total, no output
    assert L.mnrt_compact_u32(fake, fake, 5, fake, None, None) == N.ERR_ARG
    assert L.mnrt_compact_u32(fake, fake, 5, fake, fake, None) == N.ERR_ARG  # in place
    assert L.mnrt_compact_u32(fake, None, 5, 512, fake, None) == N.ERR_ARG


def test_build_graph_argument_errors_without_device():
    import ctypes

    L = N.load()
    fake = 256
    prm = N.BuildParams(0.25, 64, 1.0, 0)
    off = (N.u32 * 3)(0, 5, 9)
    trees = (N.Tree * 2)()
    g = ctypes.c_void_p()

    def desc(**kw):
        a = dict(kind=N.KIND_TRI, num_scenes=2, h_offsets=ctypes.cast(off, ctypes.POINTER(N.u32)), d_v0=fake,
                 d_v1=fake, d_v2=fake, d_power4=None, d_dir4=None, params=ctypes.pointer(prm), h_root_boxes=None,
                 h_max_query_radius=None, margin=0.25)
        a.update(kw)
        return N.BuildGraphDesc(**a)

    create = lambda d: L.mnrt_build_graph_create(ctypes.byref(d), trees, ctypes.byref(g), None)
    assert L.mnrt_build_graph_create(None, trees, ctypes.byref(g), None) == N.ERR_ARG
    assert create(desc(kind=7)) == N.ERR_ARG
    assert create(desc(num_scenes=0)) == N.ERR_ARG
    assert create(desc(margin=-0.5)) == N.ERR_ARG
    assert create(desc(margin=float("nan"))) == N.ERR_ARG
    assert create(desc(d_v1=None)) == N.ERR_ARG  # triangles need three vertex arrays
    assert create(desc(kind=N.KIND_POINT, d_power4=fake)) == N.ERR_ARG  # powers without directions
    assert create(desc(params=None)) == N.ERR_ARG
    assert L.mnrt_build_graph_replay(None, None) == N.ERR_ARG
    flags = N.u32(0)
    assert L.mnrt_build_graph_status(None, trees, ctypes.byref(flags), None) == N.ERR_ARG
    assert L.mnrt_build_graph_destroy(None) == N.OK


def test_product_fails_loudly_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(N, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(N, "_lib", None)
    with pytest.raises(N.MnrtError):
        N.load()
