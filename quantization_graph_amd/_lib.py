"""Loader for libqgraph_hip.so, the gfx950 kernels behind the C-ABI in
include/qgraph_rabitq.h.

There is no fallback: if the library is missing or fails to load, every
entry point raises. Build it with ``python -c "import __graft_entry__ as g;
g.build()"`` or ``make -C quantization_graph_amd/csrc``.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libqgraph_hip.so")

_lib = None


class QGError(RuntimeError):
    """A failure reported by libqgraph_hip (message from qg_last_error)."""


class IndexDesc(C.Structure):
    """qg_index_desc (include/qgraph_rabitq.h)."""
    _fields_ = [("codes", C.c_void_p), ("hdr", C.c_void_p), ("sfac", C.c_void_p),
                ("gate", C.c_void_p), ("rscale", C.c_void_p), ("rcodes", C.c_void_p), ("rfac", C.c_void_p),
                ("n", C.c_uint64), ("cid", C.c_void_p)]


class QueryDesc(C.Structure):
    """qg_query_desc (include/qgraph_rabitq.h)."""
    _fields_ = [("q", C.c_void_p), ("x8", C.c_void_p), ("x8p", C.c_void_p), ("qs", C.c_void_p),
                ("nq", C.c_uint64)]


def lib():
    """Return the loaded library, loading it on first use."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise QGError("libqgraph_hip.so is not built (expected at %s)" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp, sz, u64, u32, i32, f32 = C.c_void_p, C.c_size_t, C.c_uint64, C.c_uint32, C.c_int, C.c_float
    L.qg_last_error.restype = C.c_char_p
    L.qg_space_new.restype = vp
    L.qg_space_new.argtypes = [sz, u32, sz, sz, sz, C.c_bool, C.c_bool]
    L.qg_space_delete.argtypes = [vp]
    L.qg_space_set_center.argtypes = [vp, vp]
    L.qg_space_code_dim.restype = sz
    L.qg_space_code_dim.argtypes = [vp]
    L.qg_gate_row_bytes.restype = sz
    L.qg_gate_row_bytes.argtypes = [vp]
    L.qg_space_signs.restype = vp
    L.qg_space_signs.argtypes = [vp]
    L.qg_space_host_constants.argtypes = [vp, vp, vp]
    L.qg_space_rotated_center.restype = vp
    L.qg_space_rotated_center.argtypes = [vp]
    L.qg_space_set_centroids.argtypes = [vp, vp, sz]
    L.qg_space_centroid_count.restype = sz
    L.qg_space_centroid_count.argtypes = [vp]
    L.qg_space_set_rotation_qr.argtypes = [vp, vp, i32]
    L.qg_space_rotation_qr.argtypes = [vp, vp]
    L.qg_rotate_rows.argtypes = [vp, vp, u64, i32, vp, vp]
    L.qg_encode_workspace_bytes.restype = sz
    L.qg_encode_workspace_bytes.argtypes = [vp, u64]
    L.qg_encode.argtypes = [vp, vp, u64, C.POINTER(IndexDesc), vp, sz, vp]
    L.qg_pack_records.argtypes = [vp, C.POINTER(IndexDesc), vp, sz, vp, sz, vp]
    L.qg_unpack_records.argtypes = [vp, vp, sz, vp, sz, u64, C.POINTER(IndexDesc), vp]
    L.qg_prepare_workspace_bytes.restype = sz
    L.qg_prepare_workspace_bytes.argtypes = [u64]
    L.qg_prepare_queries.argtypes = [vp, vp, u64, i32, C.POINTER(QueryDesc), vp, vp]
    L.qg_scan.argtypes = [vp, i32, i32, C.POINTER(IndexDesc), C.POINTER(QueryDesc), vp, vp, u64, u64,
                          f32, vp, vp, vp, vp, vp, vp, vp, vp]
    L.qg_topk.argtypes = [vp, vp, u64, u32, u32, vp, vp, vp, vp]
    L.qg_search_workspace_bytes.restype = sz
    L.qg_search_workspace_bytes.argtypes = [u64, u32]
    L.qg_search.argtypes = [vp, C.POINTER(IndexDesc), C.POINTER(QueryDesc), i32, vp, vp, sz, u32, u32, f32,
                            u32, u32, vp, vp, vp, vp, vp, vp, sz, vp]
    L.qg_search_ex.argtypes = [vp, C.POINTER(IndexDesc), C.POINTER(QueryDesc), i32, vp, vp, sz, u32, u32, vp,
                               f32, u32, u32, u32, vp, vp, vp, vp, vp, vp, sz, vp]
    L.qg_search_noalloc_scratch_bytes.restype = sz
    L.qg_search_noalloc_scratch_bytes.argtypes = []
    L.qg_search_noalloc.argtypes = [vp, C.POINTER(IndexDesc), C.POINTER(QueryDesc), i32, vp, vp, sz, u32, u32, vp,
                                    f32, u32, u32, u32, vp, vp, vp, vp, vp, vp]
    L.qg_mlp_packed_bytes.restype = sz
    L.qg_mlp_packed_bytes.argtypes = []
    L.qg_mlp_pack.argtypes = [vp, vp, vp]
    L.qg_search_learned.argtypes = [vp, C.POINTER(IndexDesc), C.POINTER(QueryDesc), i32, vp, vp, sz, u32, u32, vp,
                                    f32, u32, u32, u32, vp, f32, u32, vp, vp, vp, vp, vp, vp, sz, vp]
    L.qg_search_lds_bytes.restype = sz
    L.qg_search_lds_bytes.argtypes = [u32, u32]
    L.qg_kmeans_workspace_bytes.restype = sz
    L.qg_kmeans_workspace_bytes.argtypes = [u64, u64, u32]
    L.qg_train_centroids.argtypes = [vp, u64, u64, u32, C.c_int, vp, vp, sz, vp]
    L.qg_build_workspace_bytes.restype = sz
    L.qg_build_workspace_bytes.argtypes = [u64, u32, u32, u32, u32]
    L.qg_build_graph.argtypes = [vp, C.POINTER(IndexDesc), u32, u32, f32, u32, u32, u32, u32, u32, vp, vp,
                                 C.POINTER(C.c_uint64), vp, sz, vp]
    L.qg_mlp_param_count.restype = sz
    L.qg_mlp_workspace_bytes.restype = sz
    L.qg_mlp_workspace_bytes.argtypes = [u64]
    L.qg_mlp_forward.argtypes = [vp, vp, u64, vp, vp]
    L.qg_mlp_grad.argtypes = [vp, vp, vp, u64, f32, vp, vp, vp, sz, vp]
    L.qg_mlp_adam_step.argtypes = [vp, vp, vp, vp, vp, u64, f32, f32, f32, f32, f32, u32, vp, vp, sz, vp]
    L.qg_residual_stats.argtypes = [vp, C.POINTER(IndexDesc), vp, vp]
    L.qg_rerank_features.argtypes = [vp, C.POINTER(IndexDesc), C.POINTER(QueryDesc), vp, vp, u32, u32, vp, vp, vp]
    L.qg_learned_rerank.argtypes = [vp, C.POINTER(IndexDesc), C.POINTER(QueryDesc), i32, vp, vp, u32, u32, vp,
                                    f32, u32, u32, vp, vp, vp, vp]
    L.qg_signscan_encode.argtypes = [vp, vp, u64, vp, vp]
    L.qg_signscan_prepare.argtypes = [vp, vp, u64, vp, vp, vp, vp]
    L.qg_signscan.argtypes = [vp, i32, vp, u64, vp, vp, vp, u64, vp, vp]
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise QGError(lib().qg_last_error().decode())
