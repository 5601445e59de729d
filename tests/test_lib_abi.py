"""The C-ABI library loads on the CPU host and exports every entry point
that include/qgraph_rabitq.h declares (no compute calls without a GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qgraph_rabitq.h")
LIB = os.path.join(ROOT, "quantization_graph_amd", "lib", "libqgraph_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b((?:qg|rabitq)_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_both_surfaces():
    names = declared_functions()
    assert "qg_encode" in names and "qg_scan" in names and "qg_prepare_queries" in names
    assert "rabitq_space_new" in names and "rabitq_query_distance_batch" in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.skip("libqgraph_hip.so not built")
    lib = C.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_errors_surface_without_a_device():
    import quantization_graph_amd as qg
    from quantization_graph_amd import _lib
    L = _lib.lib()
    assert L.qg_space_new(960, 100, 257, 4, 16, True, True) is None
    assert b"centroid_count" in L.qg_last_error()
    assert L.qg_space_new(960, 100, 1, 8, 16, True, True) is None
    assert qg.code_dim_for(960) == 1024 and qg.code_dim_for(1536) == 2048 and qg.code_dim_for(100) == 128


# The 37 extern "C" entry points of experiments/02_diskann_fair/native/rabitq_bridge.cpp.
REFERENCE_BRIDGE = """
rabitq_space_new rabitq_space_delete rabitq_space_set_center rabitq_space_set_centroids rabitq_train_centroids
rabitq_full_record_bytes rabitq_compact_record_bytes rabitq_residual_record_bytes rabitq_residual_block_size
rabitq_residual_bits rabitq_encode_full rabitq_copy_compact rabitq_copy_residual_record rabitq_build_vamana_graph
rabitq_paper_msb_code_bytes rabitq_paper_factor_bytes rabitq_export_paper_sidecar rabitq_symmetric_distance
rabitq_prepare_query rabitq_set_query_coarse_codec rabitq_release_query rabitq_query_distance
rabitq_query_distance_b1 rabitq_query_distance_b1_batch rabitq_paper_estimate rabitq_paper_estimate_sidecar
rabitq_paper_estimate_codec rabitq_paper_estimate_codec_batch rabitq_query_distance_with_paper
rabitq_query_distance_with_quantized_paper rabitq_residual_distance rabitq_residual_distance_from_record
rabitq_paper_estimate_batch rabitq_paper_estimate_batch_sidecar rabitq_query_distance_batch
rabitq_residual_distance_batch rabitq_full_residual_distance_batch
""".split()


def test_library_exports_the_whole_reference_bridge():
    assert len(REFERENCE_BRIDGE) == 37
    declared = set(declared_functions())
    assert not [n for n in REFERENCE_BRIDGE if n not in declared]
    if not os.path.exists(LIB):
        pytest.skip("libqgraph_hip.so not built")
    lib = C.CDLL(LIB)
    assert not [n for n in REFERENCE_BRIDGE if not hasattr(lib, n)]
