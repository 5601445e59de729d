"""The run_ours.py surface (Ours/experiments/run_ours.py) and the formats
its callers exchange: the reference's own assertions (Ours/tests/test_ours.py)
restated, plus the CSV schema of logging.rs:323, keep_only_formal_int8_rows,
fvecs / ivecs and the DiskANN graph file laid out by hand."""
import csv
import json
import struct
from pathlib import Path

import numpy as np
import pytest

from quantization_graph_amd import formats, run_ours

# logging.rs:323, the header every consumer of diskann_fair_raw.csv expects
REFERENCE_HEADER = (
    "suite dataset method status graph_build_distance graph_build_mode graph_build_distance_evaluations "
    "shared_graph_build_time_ms metric k nominal_bpd actual_bytes_per_vector codebook_bytes index_size_mb "
    "index_bytes auxiliary_bytes residual_bytes fp32_base_bytes peak_rss_mb build_time_ms graph_build_time_ms "
    "train_time_ms encode_time_ms max_degree build_beam alpha search_param_name search_param_value "
    "search_beam_width query_coarse_codec recall qps latency_mean_us latency_p95_us prepare_us traverse_us "
    "rerank_us neighbor_fetch_us visited_mark_us paper_batch_us flush_us visited_nodes distance_calls "
    "mean_relative_error p95_relative_error mean_absolute_error top10_overlap pairwise_flip_rate_top10 "
    "paper_checked paper_would_prune paper_msb_kernel_calls paper_remaining_kernel_calls ffi_calls_per_query "
    "payload_bytes_read avg_bits_read refine_calls threads repeat_id seed log_path note").split()


def _write_vecs_by_hand(path, arr, code):
    with open(path, "wb") as f:
        for row in arr:
            f.write(struct.pack("<i", len(row)))
            f.write(struct.pack("<%d%s" % (len(row), code), *row))


def test_split_files(tmp_path):
    """test_ours.py:19-54: 200 queries split at 100."""
    data = tmp_path / "data"
    (data / "smoke").mkdir(parents=True)
    queries = np.arange(200 * 4, dtype=np.float32).reshape(200, 4)
    gt = np.arange(200 * 10, dtype=np.int64).reshape(200, 10)
    _write_vecs_by_hand(data / "smoke" / "smoke_query.fvecs", queries, "f")
    _write_vecs_by_hand(data / "smoke" / "smoke_groundtruth.ivecs", gt, "i")
    val_q, val_gt, test_q, test_gt = run_ours.prepare_query_splits("smoke", data, tmp_path / "results", 100)
    for p in (val_q, val_gt, test_q, test_gt):
        assert p.exists()
    assert run_ours._fvec_count(test_q) == 100 and run_ours._fvec_count(val_q) == 100
    assert np.array_equal(formats.read_fvecs(val_q), queries[:100])
    assert np.array_equal(formats.read_fvecs(test_q), queries[100:])
    assert np.array_equal(formats.read_ivecs(test_gt), gt[100:])
    # a second call reuses the split on disk
    assert run_ours.prepare_query_splits("smoke", data, tmp_path / "results", 7)[2] == test_q
    assert run_ours._fvec_count(test_q) == 100


def test_config_valid():
    """test_ours.py:57-67 against the package's configuration."""
    cfg = json.loads((Path(run_ours.__file__).parent / "config.json").read_text())
    assert cfg["R"] == 32 and cfg["L_build"] == 400 and cfg["degrees"] == [32, 64]
    assert cfg["centroid_count"] == 1 and cfg["threads"] in (32, 64)


def test_csv_header_is_the_reference_schema():
    assert run_ours.CSV_HEADER.split(",") == REFERENCE_HEADER and len(REFERENCE_HEADER) == 61


def _row(method, degree, codec, L, recall):
    r = {k: "" for k in REFERENCE_HEADER}
    r.update(method=method, max_degree=str(degree), query_coarse_codec=codec, search_param_value=str(L),
             recall=str(recall))
    return r


def test_keep_only_formal_int8_rows(tmp_path):
    p = tmp_path / "raw.csv"
    rows = [_row("PQ", 32, "", 10, 0.1), _row("Ours", 32, '"full"', 10, 0.2), _row("Ours", 32, '"int8"', 20, 0.3),
            _row("Ours", 32, '"int8"', 10, 0.4), _row("Ours", 64, '"int8"', 10, 0.5),
            _row("Ours", 32, '"INT8"', 20, 0.6)]
    with p.open("w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=REFERENCE_HEADER)
        w.writeheader()
        w.writerows(rows)
    run_ours.keep_only_formal_int8_rows(p, 32)
    got = list(csv.DictReader(p.open()))
    # baselines and other degrees first, then the last INT8 row per L, sorted by L
    assert [(r["method"], r["max_degree"], r["recall"]) for r in got] == \
        [("PQ", "32", "0.1"), ("Ours", "64", "0.5"), ("Ours", "32", "0.4"), ("Ours", "32", "0.6")]
    with pytest.raises(RuntimeError):
        run_ours.keep_only_formal_int8_rows(p, 48)
    run_ours.keep_only_formal_int8_rows(tmp_path / "absent.csv", 32)


def test_vecs_formats(tmp_path):
    x = np.random.default_rng(0).standard_normal((37, 11)).astype(np.float32)
    _write_vecs_by_hand(tmp_path / "a.fvecs", x, "f")
    assert np.array_equal(formats.read_fvecs(tmp_path / "a.fvecs"), x)
    formats.write_fvecs(tmp_path / "b.fvecs", x)
    assert (tmp_path / "a.fvecs").read_bytes() == (tmp_path / "b.fvecs").read_bytes()
    g = np.arange(37 * 12, dtype=np.int32).reshape(37, 12)
    formats.write_ivecs(tmp_path / "g.ivecs", g)
    _write_vecs_by_hand(tmp_path / "h.ivecs", g, "i")
    assert (tmp_path / "g.ivecs").read_bytes() == (tmp_path / "h.ivecs").read_bytes()
    assert np.array_equal(formats.read_ivecs_topk(tmp_path / "g.ivecs", 10), g[:, :10])
    with pytest.raises(ValueError):
        formats.read_ivecs_topk(tmp_path / "g.ivecs", 13)
    (tmp_path / "bad.fvecs").write_bytes((tmp_path / "a.fvecs").read_bytes()[:-4])
    with pytest.raises(ValueError):
        formats.read_fvecs(tmp_path / "bad.fvecs")


def test_graph_file_layout(tmp_path):
    """storage::bin::save_graph's layout, built by hand: header {u64 size,
    u32 max degree, u32 start, u64 frozen}, then (count, ids...) per list,
    the frozen start point's list last."""
    lists = [[1, 2], [], [0, 3, 1], [2]]
    frozen = [0]
    body = b"".join(struct.pack("<I%dI" % len(l), len(l), *l) for l in lists + [frozen])
    size = 24 + len(body)
    (tmp_path / "hand.bin").write_bytes(struct.pack("<QIIQ", size, 41, 0, 1) + body)
    deg = np.array([len(l) for l in lists], np.uint32)
    edges = np.full((4, 3), 0xFFFFFFFF, np.uint32)
    for i, l in enumerate(lists):
        edges[i, :len(l)] = l
    n = formats.write_graph(tmp_path / "w.bin", deg, edges, start_point=0, frozen_lists=(frozen,), max_degree=41)
    assert n == size
    assert (tmp_path / "w.bin").read_bytes() == (tmp_path / "hand.bin").read_bytes()
    d2, e2, head = formats.read_graph(tmp_path / "hand.bin")
    assert head == dict(file_size=size, max_degree=41, start_point=0, frozen_points=1)
    assert list(d2) == [2, 0, 3, 1, 1]
    for i, l in enumerate(lists + [frozen]):
        assert list(e2[i, :len(l)]) == l
    assert formats.graph_edge_count(tmp_path / "hand.bin") == 7
    (tmp_path / "trunc.bin").write_bytes((tmp_path / "hand.bin").read_bytes()[:-4])
    with pytest.raises(ValueError):
        formats.read_graph(tmp_path / "trunc.bin")


def test_raw_csv_writer(tmp_path):
    ds = run_ours.Dataset("tiny", Path("b"), Path("q"), Path("g"))
    cfg = run_ours.RunConfig(max_degree=32)
    res = dict(search_list_size=20, recall=0.5, qps=1e5, latency_mean_us=10.0, latency_p95_us=10.0,
               query_coarse_codec="int8", status="done")
    for k in ("prepare_us", "traverse_us", "rerank_us", "neighbor_fetch_us", "visited_mark_us", "paper_batch_us",
              "flush_us", "visited_nodes", "distance_computations", "mean_relative_error", "p95_relative_error",
              "mean_absolute_error", "top10_overlap", "pairwise_flip_rate_top10", "paper_checked",
              "paper_would_prune", "paper_msb_kernel_calls", "paper_remaining_kernel_calls", "ffi_calls_per_query"):
        res[k] = float("nan")
    payload = dict(method="Ours", status="done", has_real_metrics=True, graph_build_distance="ExRaBitQ4_symmetric",
                   graph_build_mode="built_in_run", graph_build_distance_evaluations=123,
                   shared_graph_build_time_ms=None, actual_bytes_per_vector=1177.0, codebook_bytes=None,
                   index_size_mb=1.5, index_bytes=10, auxiliary_bytes=20, residual_bytes=30, fp32_base_bytes=0,
                   peak_rss_mb=100.0, build_time_ms=1.0, graph_build_time_ms=1.0, train_time_ms=1.0,
                   encode_time_ms=1.0, note='a "quoted" note', search_results=[res])
    p = tmp_path / "raw.csv"
    p.write_text("old,header\n1,2\n")
    run_ours.append_raw_csv(p, ds, cfg, payload, tmp_path / "x.log")
    run_ours.append_raw_csv(p, ds, cfg, payload, tmp_path / "x.log")
    lines = p.read_text().splitlines()
    assert lines[2] == "# schema_reset=02_diskann_fair_real_metrics" and lines[3] == run_ours.CSV_HEADER
    # the reference compares the file's first line, so a foreign first line
    # means a reset marker and the header before every batch of rows
    assert len(lines) == 8 and lines[5:7] == lines[2:4] and lines[7] == lines[4]
    fresh = tmp_path / "fresh.csv"
    run_ours.append_raw_csv(fresh, ds, cfg, payload, tmp_path / "x.log")
    run_ours.append_raw_csv(fresh, ds, cfg, payload, tmp_path / "x.log")
    assert fresh.read_text().splitlines() == [run_ours.CSV_HEADER, lines[4], lines[4]]
    row = next(csv.DictReader(lines[3:]))
    assert len(row) == 61 and row["query_coarse_codec"] == "int8" and row["note"] == 'a "quoted" note'
    assert row["alpha"] == "1.2" and row["graph_build_distance_evaluations"] == "123"
    assert row["shared_graph_build_time_ms"] == "" and row["prepare_us"] == "NaN"
    # the 03 adapter's reads (ours_adapter.py:54-71)
    for k in ("recall", "qps", "latency_mean_us", "latency_p95_us", "build_time_ms", "graph_build_time_ms",
              "index_size_mb", "peak_rss_mb"):
        float(row[k])


def test_cli_rejects_what_the_reference_rejects(tmp_path):
    with pytest.raises(SystemExit):
        run_ours.main(["--dataset", "x", "--M", "48", "--out-root", str(tmp_path)])
    with pytest.raises(SystemExit):
        run_ours.main(["--dataset", "x", "--M", "", "--out-root", str(tmp_path)])
    with pytest.raises(SystemExit):
        run_ours.main(["--dataset", "x", "--centroid-count", "0", "--out-root", str(tmp_path)])
