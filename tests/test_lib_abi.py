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
    assert L.qg_space_new(960, 100, 2, 4, 16, True, True) is None
    assert b"centroid_count" in L.qg_last_error()
    assert L.qg_space_new(960, 100, 1, 8, 16, True, True) is None
    assert qg.code_dim_for(960) == 1024 and qg.code_dim_for(1536) == 2048 and qg.code_dim_for(100) == 128
