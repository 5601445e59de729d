"""quantization_graph_amd -- the ExRaBitQ4 distance path of
Coconutstree/quantization_graph ("Ours-DiskANN") on AMD MI355X (gfx950).

Public surface:
    RabitqSpace        space lifecycle, encode, query preparation, estimators
    DeviceIndex        structure-of-arrays encoded base set in HBM
    DeviceQueries      prepared query batch in HBM
    topk               per-query top-k by (value, id)
    fixed              fixed-candidate scoring + residual rerank pipeline
    learned            learned adaptive rerank depth (MLP kernels, training, calibration)
"""
from . import fixed, learned  # noqa: F401
from .space import (CODEC_B1, CODEC_B1_MAIN, CODEC_FULL, CODEC_INT4, CODEC_INT8, SCAN_ASYM, SCAN_GATE_B1,
                    SCAN_GATE_FP32, SCAN_GATE_INT8, SCAN_INT8_FULL, SCAN_REM_FP32, SCAN_REM_INT8, SCAN_RESIDUAL,
                    SCAN_SYMMETRIC, SCAN_PAPER_RECORD, SCAN_B1, SCAN_RESIDUAL_RECORD, DeviceIndex,
                    DeviceQueries, RabitqSpace, code_dim_for, estimates_np, intervals_np, topk)

__all__ = ["RabitqSpace", "DeviceIndex", "DeviceQueries", "topk", "code_dim_for", "estimates_np",
           "intervals_np", "CODEC_FULL", "CODEC_B1", "CODEC_INT4", "CODEC_INT8", "CODEC_B1_MAIN",
           "SCAN_ASYM", "SCAN_GATE_B1", "SCAN_GATE_FP32",
           "SCAN_GATE_INT8", "SCAN_INT8_FULL", "SCAN_REM_FP32", "SCAN_REM_INT8", "SCAN_RESIDUAL",
           "SCAN_SYMMETRIC", "SCAN_PAPER_RECORD", "SCAN_B1", "SCAN_RESIDUAL_RECORD"]
