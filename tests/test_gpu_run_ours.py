"""run_ours.main end to end on the GPU over a small synthetic dataset:
the CSV rows, raw log, archived log and DiskANN graph file the reference's
runner leaves behind; the graph is the device build's, and the second run
reuses it from disk with identical results."""
import csv

import numpy as np
import pytest

from quantization_graph_amd import formats, run_ours

pytestmark = pytest.mark.gpu


def _dataset(root, name, n=4000, dim=96, nq=150):
    """Power-law spectrum (variance of coordinate i ~ i^-0.8), rows
    normalised, queries = perturbed base rows: neighbour ranks that a 4-bit
    code can resolve, unlike equidistant points of a wide cluster."""
    rng = np.random.default_rng(31)
    x = rng.normal(size=(n, dim)) * np.arange(1, dim + 1) ** -0.4
    x = (x / np.linalg.norm(x, axis=1, keepdims=True)).astype(np.float32)
    q = (x[rng.integers(0, n, nq)] + 0.02 * rng.normal(size=(nq, dim))).astype(np.float32)
    d = ((q[:, None, :].astype(np.float64) - x[None].astype(np.float64)) ** 2).sum(-1)
    gt = np.lexsort((np.broadcast_to(np.arange(n), d.shape), d), axis=1)[:, :100].astype(np.int32)
    (root / name).mkdir(parents=True)
    formats.write_fvecs(root / name / f"{name}_base.fvecs", x)
    formats.write_fvecs(root / name / f"{name}_query.fvecs", q)
    formats.write_ivecs(root / name / f"{name}_groundtruth.ivecs", gt)
    return x


def _rows(path):
    return [r for r in csv.DictReader(open(path)) if r["method"] == "Ours"]


def test_run_ours_end_to_end(gpu, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    x = _dataset(tmp_path / "data", "tiny")
    args = ["--dataset", "tiny", "--M", "32", "--data-root", str(tmp_path / "data"), "--out-root",
            str(tmp_path / "out"), "--val-queries", "50", "--search-list-sizes", "20,60", "--repeats", "2"]
    print("case: run_ours tiny n=4000 dim=96 M=32 L=20,60 (build)", flush=True)
    assert run_ours.main(args) == 0
    raw = tmp_path / "out" / "02_diskann_fair" / "tiny" / "csv" / "diskann_fair_raw.csv"
    first = _rows(raw)
    assert [r["search_param_value"] for r in first] == ["20", "60"]
    assert all(r["query_coarse_codec"] == "int8" and r["max_degree"] == "32" for r in first)
    assert first[0]["graph_build_mode"] == "built_in_run" and int(first[0]["graph_build_distance_evaluations"]) > 0
    assert float(first[1]["recall"]) > 0.6 and float(first[1]["recall"]) >= float(first[0]["recall"])
    assert float(first[1]["top10_overlap"]) == pytest.approx(float(first[1]["recall"]))
    gpath = run_ours.graph_path(tmp_path / "out", "tiny", 32)
    deg, edges, head = formats.read_graph(gpath)
    assert head["max_degree"] == 41 and head["frozen_points"] == 1 and deg.size == x.shape[0] + 1
    assert deg[-1] == 1 and edges[-1, 0] == 0 and deg[:-1].max() <= 32
    log = tmp_path / "out" / "02_diskann_fair" / "tiny" / "logs" / "Ours" / "tiny_Ours_R32_Lbuild400.log"
    text = log.read_text()
    assert "graph_build_mode=built_in_run" in text and "efSearch_values=20,60" in text
    assert (tmp_path / "Ours" / "logs" / "tiny" / log.name).exists()

    # the device build through the Python API gives the same graph
    import quantization_graph_amd as qg
    sp = qg.RabitqSpace(96, 100)
    sp.set_center(run_ours.train_global_center(x, 100000, 100))
    ix = sp.encode(x)
    d2, e2, _ = sp.build_graph(ix, r=32, l_build=400, alpha=1.2, beam_width=1, batch=4096, refine_passes=2)
    d2 = d2.cpu().numpy()
    e2 = e2.cpu().numpy().view(np.uint32)
    assert np.array_equal(deg[:-1], d2)
    for i in range(0, x.shape[0], 97):
        assert np.array_equal(edges[i, :deg[i]], e2[i, :d2[i]])
    # the CSV's recall is the recall of the batched INT8 search over the test split
    split = tmp_path / "out" / "03_system_fair" / "tiny" / "csv" / "_query_splits"
    tq = formats.read_fvecs(split / "test_query.fvecs")
    tgt = formats.read_ivecs_topk(split / "test_gt.ivecs", 10).astype(np.int64)
    for L, row in zip((20, 60), first):
        ids, _, _ = sp.search(ix, sp.prepare(np.asarray(tq), "int8"), d2, e2.view(np.int32), k=10, l_search=L)
        hits = (ids.cpu().numpy().astype(np.int64)[:, :, None] == tgt[:, None, :]).any(-1).sum()
        assert "%.6f" % (hits / tgt.size) == row["recall"]

    print("case: run_ours tiny (reuse graph)", flush=True)
    assert run_ours.main(args) == 0
    second = _rows(raw)
    assert len(second) == 2 and second[0]["graph_build_mode"] == "reused_graph"
    for a, b in zip(first, second):
        assert a["recall"] == b["recall"] and a["visited_nodes"] == b["visited_nodes"]
