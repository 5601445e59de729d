"""Ours-DiskANN on the GPU: the entry point of Ours/experiments/run_ours.py
together with the runner it launches (experiments/02_diskann_fair,
run_ours_exrabitq4 in diskann_runner.rs:155-552 and the writers of
logging.rs), computed in-process through libqgraph_hip.so.

Kept from the reference:
* the command line (run_ours.py:143-180): --dataset --M --centroid-count
  --centroid-train-samples --data-root --out-root --threads --val-queries
  --repeats, with M in {32, 64} and K in 1..256;
* the helpers ``_fvec_count``, ``prepare_query_splits`` and
  ``keep_only_formal_int8_rows`` with the same arguments and effects;
* the outputs: ``02_diskann_fair/<ds>/csv/diskann_fair_raw.csv`` rows under
  the 61-column header, the raw log ``.../logs/Ours/<ds>_Ours_R<M>_Lbuild400.log``
  (archived to ``Ours/logs/<ds>/`` under the working directory), the
  manifest, the payload marker and JSON, and the graph file
  ``<out>/<ds>/indexes/02_diskann_fair/Ours/<ds>_Ours_R<M>_Lbuild400.graph.bin``
  in DiskANN's format, reused on the next run when present (run_ours.py:209-222).
* the formal configuration: INT8 query codec, epsilon0 1.9, beam 1, rerank
  100, k 10, L_build 400, alpha 1.2, refine 2, batch 4096, the search-list
  sweep 10..460 (args.rs:50-68), seed 20260813.

Different by design:
* queries run as one batch per (search-list size, repeat), so QPS is the
  batch throughput and the per-query latency is the batch's wall time (the
  time at which every answer of the batch is available);
* the centre is the mean of ``centroid_train_samples`` rows drawn with
  numpy's seeded generator (seed 100) and summed in float64: the
  reference's draw uses Rust's StdRng, which is not reproduced here;
* the graph file's start-point field holds 0, the node every search starts
  from (the reference writes the id its own generator picks);
* K > 1 (multi-centroid codes) is rejected: this build implements K = 1.
"""
from __future__ import annotations

import argparse
import csv
import json
import math
import shutil
import sys
import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import formats

K = 10
SUITE = "02_diskann_fair"
METHOD = "Ours"
DEFAULT_SEARCH_LIST_SIZES = list(range(10, 31)) + list(range(40, 101, 10)) + list(range(140, 461, 40))
CSV_HEADER = (
    "suite,dataset,method,status,graph_build_distance,graph_build_mode,graph_build_distance_evaluations,"
    "shared_graph_build_time_ms,metric,k,nominal_bpd,actual_bytes_per_vector,codebook_bytes,index_size_mb,"
    "index_bytes,auxiliary_bytes,residual_bytes,fp32_base_bytes,peak_rss_mb,build_time_ms,graph_build_time_ms,"
    "train_time_ms,encode_time_ms,max_degree,build_beam,alpha,search_param_name,search_param_value,"
    "search_beam_width,query_coarse_codec,recall,qps,latency_mean_us,latency_p95_us,prepare_us,traverse_us,"
    "rerank_us,neighbor_fetch_us,visited_mark_us,paper_batch_us,flush_us,visited_nodes,distance_calls,"
    "mean_relative_error,p95_relative_error,mean_absolute_error,top10_overlap,pairwise_flip_rate_top10,"
    "paper_checked,paper_would_prune,paper_msb_kernel_calls,paper_remaining_kernel_calls,ffi_calls_per_query,"
    "payload_bytes_read,avg_bits_read,refine_calls,threads,repeat_id,seed,log_path,note")
MANIFEST_HEADER = ("suite,dataset,method,status,implementation,source_path,max_degree,build_beam,alpha,"
                   "nominal_bpd,payload_path,raw_log,note")
NOTE = ("Ours ExRaBitQ4 block16 on MI355X. Encode, the ExRaBitQ4 symmetric Vamana build, the paper-pruned "
        "query traversal with the 1-bit gate and the residual4 block16 mse/fp16 rerank run as HIP kernels of "
        "libqgraph_hip.so, queries batched per launch; the index stores no fp32 base vectors "
        "(fp32_base_bytes=0).")


# ----------------------------------------------------------------- helpers
def _fvec_count(path: Path) -> int:
    """Rows of an fvecs file (run_ours.py:28-33)."""
    return formats.vecs_count(path)


def prepare_query_splits(dataset: str, data_root: Path, out_root: Path,
                         val_queries: int) -> tuple[Path, Path, Path, Path]:
    """The shared validation / test split (run_ours.py:36-100): the first
    ``val_queries`` queries and ground-truth rows are the validation split,
    the rest the test split, written once under
    ``<out_root>/03_system_fair/<dataset>/csv/_query_splits`` and reused."""
    split_dir = Path(out_root) / "03_system_fair" / dataset / "csv" / "_query_splits"
    split_dir.mkdir(parents=True, exist_ok=True)
    paths = (split_dir / "validation_query.fvecs", split_dir / "validation_gt.ivecs",
             split_dir / "test_query.fvecs", split_dir / "test_gt.ivecs")
    if paths[3].exists():
        return paths
    src = Path(data_root) / dataset
    queries = np.asarray(formats.read_fvecs(src / f"{dataset}_query.fvecs", mmap=False))
    gt = formats.read_ivecs(src / f"{dataset}_groundtruth.ivecs")
    n_val = min(val_queries, len(queries))
    formats.write_fvecs(paths[0], queries[:n_val])
    formats.write_ivecs(paths[1], gt[:n_val])
    formats.write_fvecs(paths[2], queries[n_val:])
    formats.write_ivecs(paths[3], gt[n_val:])
    return paths


def keep_only_formal_int8_rows(csv_path: Path, degree: int) -> None:
    """Keep every other row and, for Ours at ``degree``, only the last INT8
    row of each search-list size (run_ours.py:103-140). Raises when the run
    left no such row."""
    csv_path = Path(csv_path)
    if not csv_path.exists():
        return
    with csv_path.open(newline="") as handle:
        reader = csv.DictReader(handle)
        fieldnames = list(reader.fieldnames or [])
        rows = list(reader)
    if not rows:
        return
    kept, formal = [], {}
    for row in rows:
        if row.get("method") != METHOD or int(row.get("max_degree", 0) or 0) != degree:
            kept.append(row)
            continue
        if row.get("query_coarse_codec", "").strip('"').lower() == "int8":
            formal[int(row.get("search_param_value", 0) or 0)] = row
    if not formal:
        raise RuntimeError(f"formal INT8 run produced no Ours rows with max_degree={degree}: {csv_path}")
    tmp = csv_path.with_suffix(csv_path.suffix + ".tmp")
    with tmp.open("w", newline="") as handle:
        writer = csv.DictWriter(handle, fieldnames=fieldnames)
        writer.writeheader()
        writer.writerows(kept)
        writer.writerows(row for _, row in sorted(formal.items()))
    tmp.replace(csv_path)


def _metric(v) -> str:
    """fmt_metric (logging.rs): NaN, or six decimals."""
    return "NaN" if v is None or (isinstance(v, float) and math.isnan(v)) else "%.6f" % v


def _opt_f(v) -> str:
    return "" if v is None else "%.6f" % v


def _opt_u(v) -> str:
    return "" if v is None else str(int(v))


def _quote(s: str) -> str:
    return '"' + s.replace('"', '""') + '"'


def _fmt_alpha(a: float) -> str:
    """Rust's Display of an f32: the shortest text that reads back as that
    f32, without a trailing '.0' (1.2 -> '1.2', 1.0 -> '1')."""
    return np.format_float_positional(np.float32(a), trim="-")


def peak_rss_mb():
    try:
        for line in Path("/proc/self/status").read_text().splitlines():
            if line.startswith("VmHWM:"):
                return float(line.split()[1]) / 1024.0
    except OSError:
        pass
    return None


# ------------------------------------------------------------------ config
@dataclass
class RunConfig:
    max_degree: int = 32
    build_beam: int = 400
    alpha: float = 1.2
    search_beam_width: int = 1
    search_list_sizes: list = field(default_factory=lambda: list(DEFAULT_SEARCH_LIST_SIZES))
    epsilon0: float = 1.9
    rerank_candidates: int = 100
    threads: int = 64
    build_threads: int = 64
    repeats: int = 1
    seed: int = 20260813
    build_batch_size: int = 4096
    refine_passes: int = 2
    prune_candidate_cap: int = 0
    build_early_stop_hops: int = 0
    centroid_count: int = 1
    centroid_train_samples: int = 100000
    query_coarse_codec: str = "int8"


@dataclass
class Dataset:
    name: str
    base_path: Path
    query_path: Path
    gt_path: Path

    def shape(self):
        n, dim = formats._vecs_shape(self.base_path)
        nq, qdim = formats._vecs_shape(self.query_path)
        if qdim != dim:
            raise ValueError(f"query dimension mismatch: base={dim}, query={qdim}")
        _, gt_k = formats._vecs_shape(self.gt_path)
        return n, nq, dim, gt_k


def output_paths(out_root: Path, dataset: str, degree: int, build_beam: int):
    """create_output_dirs (logging.rs:15-37)."""
    root = Path(out_root) / SUITE / dataset
    for d in (root / "logs" / METHOD, root / "csv", root / "manifests", root / "indexes/shared_graph"):
        d.mkdir(parents=True, exist_ok=True)
    return dict(raw_log=root / "logs" / METHOD / f"{dataset}_{METHOD}_R{degree}_Lbuild{build_beam}.log",
                raw_csv=root / "csv" / "diskann_fair_raw.csv",
                manifest_csv=root / "manifests" / "02_diskann_fair_manifest.csv")


def graph_path(out_root: Path, dataset: str, degree: int, build_beam: int = 400) -> Path:
    return (Path(out_root) / dataset / "indexes" / SUITE / METHOD /
            f"{dataset}_{METHOD}_R{degree}_Lbuild{build_beam}.graph.bin")


def train_global_center(x, sample_count: int, seed: int = 100):
    """Mean of min(sample_count, n) distinct rows (seeded numpy draw, ids
    sorted, float64 sum): the role of train_global_center_matrix
    (diskann_runner.rs:1677-1701) with this module's own sampler."""
    n = x.shape[0]
    m = min(sample_count, n)
    if m == 0:
        raise ValueError("empty sample for the global center")
    ids = np.sort(np.random.default_rng(seed).choice(n, size=m, replace=False))
    acc = np.zeros(x.shape[1], np.float64)
    for s in range(0, m, 65536):
        acc += np.asarray(x[ids[s:s + 65536]], np.float64).sum(0)
    return (acc / m).astype(np.float32)


def accuracy_metrics(ids, dists, queries, base, gt):
    """accumulate_accuracy + accuracy_summary (diskann_runner.rs:2022-2093):
    relative / absolute error of the reported distances against exact L2 of
    the returned ids, overlap with the ground truth, and the rate of pairs
    of the top-k whose order the reported distances flip. ids / dists are
    [nq, k]; exact distances in float64 on the device."""
    import torch
    ids = np.clip(np.asarray(ids, np.int64), 0, base.shape[0] - 1)
    ids_t = torch.as_tensor(ids, device="cuda")
    q = torch.as_tensor(np.asarray(queries, np.float32), device="cuda").double()
    rows = torch.as_tensor(np.asarray(base[np.asarray(ids, np.int64).reshape(-1)], np.float32),
                           device="cuda").double().reshape(ids_t.shape[0], ids_t.shape[1], -1)
    exact = ((rows - q[:, None, :]) ** 2).sum(-1).float().double()
    rep = torch.as_tensor(np.asarray(dists, np.float32), device="cuda").double()
    rel = torch.where(exact > 1e-12, (rep - exact).abs() / exact, torch.zeros_like(exact)).reshape(-1)
    abs_err = (rep - exact).abs().sum().item()
    gt_t = torch.as_tensor(np.asarray(gt, np.int64), device="cuda")
    overlap = (ids_t[:, :, None] == gt_t[:, None, :]).any(-1).sum().item()
    k = ids_t.shape[1]
    iu = torch.triu_indices(k, k, 1, device="cuda")
    rep_ord = rep[:, iu[0]] <= rep[:, iu[1]]
    ex_ord = exact[:, iu[0]] <= exact[:, iu[1]]
    flips = (rep_ord != ex_ord).sum().item()
    pairs = rep_ord.numel()
    count = max(1, rel.numel())
    srt = torch.sort(rel).values
    rank = min(max(1, math.ceil(srt.numel() * 0.95)), srt.numel())
    return dict(mean_relative_error=rel.mean().item() if rel.numel() else float("nan"),
                p95_relative_error=srt[rank - 1].item() if rel.numel() else float("nan"),
                mean_absolute_error=abs_err / count,
                top10_overlap=overlap / count,
                pairwise_flip_rate_top10=flips / pairs if pairs else float("nan"))


# --------------------------------------------------------------- pipeline
def run_ours_exrabitq4(ds: Dataset, cfg: RunConfig, out_root: Path, progress_log: Path, graph_file=None,
                       log=print):
    """One Ours run (run_ours_exrabitq4, diskann_runner.rs:155-552) on the
    GPU; returns the payload dict the writers take."""
    import torch
    from .space import RabitqSpace, search_counters
    if cfg.centroid_count != 1:
        raise SystemExit("ExRaBitQ centroid count K > 1 is not implemented by this build (K = 1 only)")

    def progress(stage, status):
        with open(progress_log, "a") as f:
            f.write(f"progress_stage={stage} status={status}\n")

    with open(progress_log, "w") as f:
        f.write(f"experiment_profile=02_diskann_payload_fair\ndataset={ds.name}\nmethod={METHOD}\nstatus=running\n")
    t0 = time.perf_counter()
    progress("load_data", "start")
    base = formats.read_fvecs(ds.base_path)
    queries = np.asarray(formats.read_fvecs(ds.query_path, mmap=False))
    gt = formats.read_ivecs_topk(ds.gt_path, K)
    progress("load_data", "done ms=%.3f" % ((time.perf_counter() - t0) * 1e3))
    n, dim = base.shape

    progress("train_center", "start residual_block_size=16")
    t0 = time.perf_counter()
    center = train_global_center(base, min(cfg.centroid_train_samples, n), 100)
    sp = RabitqSpace.new_with_centroids(dim, 100, 1, center)
    train_ms = (time.perf_counter() - t0) * 1e3
    progress("train_center", "done ms=%.3f residual_bits=4 residual_block_size=16 centroid_count=1" % train_ms)

    total_t0 = time.perf_counter()
    progress("payload_encode", "start target=device_index")
    t0 = time.perf_counter()
    xdev = torch.empty((n, dim), dtype=torch.float32, device="cuda")
    chunk = max(1, (1 << 28) // (4 * dim))
    for s in range(0, n, chunk):  # copy the memory-mapped base set in 256 MiB slices
        xdev[s:s + chunk].copy_(torch.from_numpy(np.ascontiguousarray(base[s:s + chunk])))
    index = sp.encode(xdev, batch=65536)
    torch.cuda.synchronize()
    del xdev
    encode_ms = (time.perf_counter() - t0) * 1e3
    progress("payload_encode", "done ms=%.3f target=device_index" % encode_ms)

    progress("graph_build", "start graph_build_distance=ExRaBitQ4_symmetric builder=native_vamana_bulk_grouped_refine "
             "refine_passes=%d residual_block_size=16" % cfg.refine_passes)
    t0 = time.perf_counter()
    evals = None
    if graph_file is not None:
        graph_file = Path(graph_file)
        if not graph_file.exists():
            raise RuntimeError(f"Ours graph file not found: {graph_file}")
        progress("graph_load", f"path={graph_file}")
        gdeg, gedges, head = formats.read_graph(graph_file)
        if gdeg.size < n:
            raise RuntimeError(f"{graph_file} holds {gdeg.size} lists for {n} points")
        if gedges.shape[1] > 64:
            raise RuntimeError(f"{graph_file}: degree {gedges.shape[1]} exceeds the search kernel's 64")
        deg = torch.as_tensor(gdeg[:n].astype(np.int32), device="cuda")
        edges = torch.as_tensor(np.where(np.arange(gedges.shape[1])[None] < gdeg[:n, None],
                                         gedges[:n], 0xFFFFFFFF).astype(np.uint32).view(np.int32), device="cuda")
        mode = "reused_graph"
        total_edges = int(gdeg.astype(np.int64).sum())
    else:
        deg, edges, evals = sp.build_graph(index, r=cfg.max_degree, l_build=cfg.build_beam, alpha=cfg.alpha,
                                           beam_width=cfg.search_beam_width, batch=cfg.build_batch_size,
                                           refine_passes=cfg.refine_passes, prune_cap=cfg.prune_candidate_cap,
                                           early_stop_hops=cfg.build_early_stop_hops)
        torch.cuda.synchronize()
        mode = "built_in_run"
        hdeg = deg.cpu().numpy().astype(np.uint32)
        total_edges = int(hdeg.astype(np.int64).sum())
        gp = graph_path(out_root, ds.name, cfg.max_degree, cfg.build_beam)
        gp.parent.mkdir(parents=True, exist_ok=True)
        nbytes = formats.write_graph(gp, hdeg, edges.cpu().numpy().view(np.uint32), start_point=0,
                                     frozen_lists=((0,),), max_degree=int(np.float32(1.3) * cfg.max_degree))
        progress("graph_save", f"path={gp} bytes={nbytes}")
    graph_ms = (time.perf_counter() - t0) * 1e3
    progress("graph_build", "mode=%s done ms=%.3f" % (mode, graph_ms))
    build_ms = (time.perf_counter() - total_t0) * 1e3

    codec = cfg.query_coarse_codec
    progress("search", f"start paper_prune=active residual_rerank=block16 query_coarse_codecs={codec}")
    sp.set_query_coarse_codec(codec)
    qdev = torch.from_numpy(queries).cuda()
    results = []
    nq = queries.shape[0]
    for L in cfg.search_list_sizes:
        lat, hits, total = [], 0, 0
        prep_s = search_s = 0.0
        stats_sum = None
        torch.cuda.synchronize()
        start = time.perf_counter()
        for _ in range(cfg.repeats):
            b0 = time.perf_counter()
            qv = sp.prepare(qdev)
            torch.cuda.synchronize()
            b1 = time.perf_counter()
            ids, dists, st = sp.search(index, qv, deg, edges, k=K, l_search=L, eps0=cfg.epsilon0,
                                       rerank=cfg.rerank_candidates)
            torch.cuda.synchronize()
            b2 = time.perf_counter()
            prep_s += b1 - b0
            search_s += b2 - b1
            lat += [(b2 - b0) * 1e6] * nq
            h_ids = ids.cpu().numpy().astype(np.int64)
            hits += int((h_ids[:, :, None] == gt[:, None, :].astype(np.int64)).any(-1).sum())
            total += K * nq
            s = {name: float(v.sum()) for name, v in search_counters(st.cpu().numpy()).items()}
            stats_sum = s if stats_sum is None else {name: stats_sum[name] + v for name, v in s.items()}
        elapsed = time.perf_counter() - start
        ctr = stats_sum
        acc = accuracy_metrics(h_ids, dists.cpu().numpy(), queries, base, gt)
        qn = float(nq * cfg.repeats)
        lat_sorted = sorted(lat)
        results.append(dict(
            search_list_size=L, recall=hits / total, qps=qn / max(elapsed, 1e-12),
            latency_mean_us=sum(lat) / len(lat),
            latency_p95_us=lat_sorted[int(round((len(lat_sorted) - 1) * 0.95))],
            prepare_us=prep_s * 1e6 / qn, traverse_us=search_s * 1e6 / qn, rerank_us=float("nan"),
            neighbor_fetch_us=float("nan"), visited_mark_us=float("nan"), paper_batch_us=float("nan"),
            flush_us=float("nan"), visited_nodes=ctr["visited_nodes"] / qn,
            distance_computations=ctr["distance_computations"] / qn,
            paper_checked=ctr["paper_checked"] / qn, paper_would_prune=ctr["paper_would_prune"] / qn,
            paper_msb_kernel_calls=ctr["paper_msb_kernel_calls"] / qn,
            paper_remaining_kernel_calls=ctr["paper_remaining_kernel_calls"] / qn,
            ffi_calls_per_query=0.0, query_coarse_codec=codec, status="done", **acc))
        progress("search", "search_list_size=%d query_coarse_codec=%s recall=%.6f qps=%.3f paper_checked=%d "
                 "traverse_us=%.3f" % (L, codec, results[-1]["recall"], results[-1]["qps"], int(ctr["paper_checked"]),
                                       results[-1]["traverse_us"]))
        log("L=%d recall@10=%.4f qps=%.1f" % (L, results[-1]["recall"], results[-1]["qps"]))
    progress("search", "done")

    method_dir = Path(out_root) / ds.name / "indexes" / SUITE / METHOD
    method_dir.mkdir(parents=True, exist_ok=True)
    payload_path = method_dir / f"{METHOD}_bpd4_payload.bin"
    payload_path.write_bytes(b"device_index_libqgraph_hip\n")
    note = f"{NOTE} query_coarse_codecs={codec}"
    actual = float(sp.compact_record_bytes + sp.residual_record_bytes)
    payload_json = payload_path.with_suffix(".json")
    payload_json.write_text(json.dumps(dict(
        suite=SUITE, dataset=ds.name, method=METHOD, status="done", has_real_metrics=True,
        nominal_bits_per_dim=4, actual_bytes_per_vector=actual,
        code_bytes_per_vector=float(sp.compact_record_bytes), codebook_bytes=None,
        payload_path=str(payload_path), note=note), indent=2) + "\n")
    primary = sp.compact_record_bytes * n
    residual = sp.residual_record_bytes * n
    sidecar = (sp.paper_msb_code_bytes + sp.paper_factor_bytes) * n
    index_bytes = n * 4 + total_edges * 4 + primary + sidecar
    auxiliary = dim * 4 * 2
    total_bytes = index_bytes + auxiliary + residual
    return dict(method=METHOD, status="done", has_real_metrics=True, payload_path=payload_path,
                payload_json_path=payload_json, nominal_bits_per_dim=4.0, actual_bytes_per_vector=actual,
                code_bytes_per_vector=float(sp.compact_record_bytes),
                metadata_bytes_per_vector=float(sp.residual_record_bytes), codebook_bytes=None,
                train_time_ms=train_ms, encode_time_ms=encode_ms, build_time_ms=build_ms,
                graph_build_time_ms=graph_ms, shared_graph_build_time_ms=None, graph_build_mode=mode,
                index_size_mb=total_bytes / 1048576.0, index_bytes=index_bytes, auxiliary_bytes=auxiliary,
                residual_bytes=residual, fp32_base_bytes=0, graph_build_distance="ExRaBitQ4_symmetric",
                graph_build_distance_evaluations=evals, peak_rss_mb=peak_rss_mb(), search_results=results,
                note=note, base_count=n, query_count=nq, dimension=dim, gt_k=gt.shape[1])


# ----------------------------------------------------------------- writers
def append_raw_csv(path: Path, ds: Dataset, cfg: RunConfig, payload: dict, raw_log: Path) -> None:
    """append_raw_csv (logging.rs:313-408): header written once; a file
    with another header gets a schema-reset marker and the header again."""
    if not payload["has_real_metrics"]:
        return
    path = Path(path)
    needs_header = not path.exists()
    matches = needs_header or path.read_text().split("\n", 1)[0] == CSV_HEADER
    with path.open("a") as f:
        if needs_header or not matches:
            if not needs_header:
                f.write("# schema_reset=02_diskann_fair_real_metrics\n")
            f.write(CSV_HEADER + "\n")
        p = payload
        for r in p["search_results"]:
            row = [SUITE, ds.name, p["method"], p["status"], p["graph_build_distance"], p["graph_build_mode"],
                   _opt_u(p["graph_build_distance_evaluations"]), _opt_f(p["shared_graph_build_time_ms"]), "L2",
                   str(K), "4", _opt_f(p["actual_bytes_per_vector"]), _opt_u(p["codebook_bytes"]),
                   _opt_f(p["index_size_mb"]), _opt_u(p["index_bytes"]), _opt_u(p["auxiliary_bytes"]),
                   _opt_u(p["residual_bytes"]), _opt_u(p["fp32_base_bytes"]), _opt_f(p["peak_rss_mb"]),
                   _opt_f(p["build_time_ms"]), _opt_f(p["graph_build_time_ms"]), _opt_f(p["train_time_ms"]),
                   _opt_f(p["encode_time_ms"]), str(cfg.max_degree), str(cfg.build_beam), _fmt_alpha(cfg.alpha),
                   "efSearch", str(r["search_list_size"]), str(cfg.search_beam_width),
                   _quote(r["query_coarse_codec"]), "%.6f" % r["recall"], "%.6f" % r["qps"],
                   "%.6f" % r["latency_mean_us"], "%.6f" % r["latency_p95_us"]]
            row += [_metric(r[k]) for k in (
                "prepare_us", "traverse_us", "rerank_us", "neighbor_fetch_us", "visited_mark_us", "paper_batch_us",
                "flush_us", "visited_nodes", "distance_computations", "mean_relative_error", "p95_relative_error",
                "mean_absolute_error", "top10_overlap", "pairwise_flip_rate_top10", "paper_checked",
                "paper_would_prune", "paper_msb_kernel_calls", "paper_remaining_kernel_calls", "ffi_calls_per_query")]
            row += ["NaN", "NaN", "0", str(cfg.threads), "aggregate", str(cfg.seed), str(raw_log), _quote(p["note"])]
            f.write(",".join(row) + "\n")


def append_manifest(path: Path, ds: Dataset, cfg: RunConfig, payload: dict, raw_log: Path) -> None:
    """append_manifest (logging.rs:410-440)."""
    path = Path(path)
    needs_header = not path.exists()
    with path.open("a") as f:
        if needs_header:
            f.write(MANIFEST_HEADER + "\n")
        f.write("%s,%s,%s,%s,adapter_slot,experiments/02_diskann_fair,%d,%d,%s,4,%s,%s,\"%s\"\n" % (
            SUITE, ds.name, payload["method"], payload["status"], cfg.max_degree, cfg.build_beam,
            _fmt_alpha(cfg.alpha), payload["payload_path"], raw_log, payload["note"].replace('"', "'")))


def write_raw_log(path: Path, ds: Dataset, cfg: RunConfig, payload: dict) -> None:
    """write_raw_log (logging.rs:39-311): key=value configuration lines, the
    build stages, storage, then one line per search point."""
    p = payload
    ms_to_us = (lambda v: 0 if v is None else int(round(v * 1000.0)))
    lines = ["experiment_profile=02_diskann_payload_fair", f"suite={SUITE}", f"dataset={ds.name}",
             f"method={p['method']}", f"status={p['status']}", "has_real_metrics=true", "metric=L2",
             f"base_count={p['base_count']}", f"query_count={p['query_count']}", f"dimension={p['dimension']}",
             f"groundtruth_k={p['gt_k']}", "graph_type=DiskANN3/Vamana",
             f"graph_build_distance={p['graph_build_distance']}", f"graph_build_mode={p['graph_build_mode']}",
             f"shared_graph_build_time_ms={_opt_f(p['shared_graph_build_time_ms'])}",
             "search_distance=Float32_to_ExRaBitQ4_plus_residual4_block16",
             "graph_build_builder=native_vamana_bulk_grouped_refine", f"graph_refine_passes={cfg.refine_passes}",
             "framework=mi355x_hip_batched_search", "provider=OursExRaBitQ4", "paper_prune=active",
             "paper_epsilon0=1.9", f"rerank_candidates={cfg.rerank_candidates}", "residual_bits=4",
             "residual_block_size=16", "residual_scale_mode=mse", "residual_scale_storage=fp16",
             "legacy_hnswlib_runtime=0", f"max_degree={cfg.max_degree}", f"build_beam={cfg.build_beam}",
             f"M={cfg.max_degree}", f"ef={cfg.build_beam}", f"efConstruction={cfg.build_beam}",
             "search_param_name=efSearch",
             "efSearch_values=" + ",".join(str(v) for v in cfg.search_list_sizes),
             f"alpha={_fmt_alpha(cfg.alpha)}", f"legacy_equivalent=M{cfg.max_degree}_ef{cfg.build_beam}",
             f"search_beam_width={cfg.search_beam_width}", f"threads={cfg.threads}",
             f"build_threads={cfg.build_threads}", f"repeats={cfg.repeats}", f"seed={cfg.seed}",
             "accuracy_metrics=mean_relative_error,p95_relative_error,mean_absolute_error,top10_overlap,"
             "pairwise_flip_rate_top10",
             "nominal_bits_per_dim=4", f"actual_bytes_per_vector={_opt_f(p['actual_bytes_per_vector'])}",
             f"code_bytes_per_vector={_opt_f(p['code_bytes_per_vector'])}",
             f"metadata_bytes_per_vector={_opt_f(p['metadata_bytes_per_vector'])}",
             f"fp32_base_bytes={_opt_u(p['fp32_base_bytes'])}", f"codebook_bytes={_opt_u(p['codebook_bytes'])}",
             f"base_path={ds.base_path}", f"query_path={ds.query_path}", f"gt_path={ds.gt_path}",
             f"payload_path={p['payload_path']}", f"payload_json_path={p['payload_json_path']}",
             f"note={p['note']}",
             f"build_stage=train_center us={ms_to_us(p['train_time_ms'])} status=done",
             f"build_stage=payload_encode us={ms_to_us(p['encode_time_ms'])} count={p['base_count']} status=done",
             "build_stage=payload_import us=0 status=in_memory_provider",
             f"build_stage=graph_build us={ms_to_us(p['graph_build_time_ms'])} status=done",
             f"build_total_us={ms_to_us(p['build_time_ms'])} status=done",
             "Graph construction time: %.6f seconds status=done" % ((p["graph_build_time_ms"] or 0) / 1000.0),
             "Build time: %.6f seconds status=done" % ((p["build_time_ms"] or 0) / 1000.0)]
    idx, aux, res = p["index_bytes"] or 0, p["auxiliary_bytes"] or 0, p["residual_bytes"] or 0
    fp32 = p["fp32_base_bytes"] or 0
    tot = idx + aux + res + fp32
    lines.append("Index storage size: %.2f MB (index=%.2f MB, auxiliary=%.2f MB, residual=%.2f MB, fp32_base=%.2f MB, "
                 "total_bytes=%d) status=computed" % (tot / 1e6, idx / 1e6, aux / 1e6, res / 1e6, fp32 / 1e6, tot))
    lines.append(f"storage_breakdown index_bytes={idx} auxiliary_bytes={aux} residual_bytes={res} "
                 f"fp32_base_bytes={fp32} total_bytes={tot}")
    lines.append(f"Peak RSS: {_opt_f(p['peak_rss_mb'])} MB status=measured")
    lines.append("search_list_size recall latency_mean_us qps latency_p95_us prepare_us traverse_us rerank_us "
                 "neighbor_fetch_us visited_mark_us paper_batch_us flush_us visited_nodes distance_computations "
                 "mean_relative_error p95_relative_error mean_absolute_error top10_overlap pairwise_flip_rate_top10 "
                 "payload_bytes_read avg_bits_read refine_calls paper_checked paper_would_prune "
                 "paper_msb_kernel_calls paper_remaining_kernel_calls ffi_calls_per_query status")
    for r in p["search_results"]:
        L = r["search_list_size"]
        lines.append(
            "%d %.6f %.6f us search_list_size=%d efSearch=%d search_beam_width=%d total_us_per_query=%.6f qps=%.6f "
            "p95_us=%.6f " % (L, r["recall"], r["latency_mean_us"], L, L, cfg.search_beam_width,
                              r["latency_mean_us"], r["qps"], r["latency_p95_us"]) +
            " ".join("%s=%s" % (k, _metric(r[k])) for k in (
                "prepare_us", "traverse_us", "rerank_us", "neighbor_fetch_us", "visited_mark_us", "paper_batch_us",
                "flush_us", "visited_nodes")) +
            " distance_computations=%s " % _metric(r["distance_computations"]) +
            " ".join("%s=%s" % (k, _metric(r[k])) for k in (
                "mean_relative_error", "p95_relative_error", "mean_absolute_error", "top10_overlap",
                "pairwise_flip_rate_top10")) +
            " payload_bytes_read=NaN avg_bits_read=NaN refine_calls=0 " +
            " ".join("%s=%s" % (k, _metric(r[k])) for k in (
                "paper_checked", "paper_would_prune", "paper_msb_kernel_calls", "paper_remaining_kernel_calls",
                "ffi_calls_per_query")) + " status=%s" % r["status"])
    Path(path).write_text("\n".join(lines) + "\n")


# -------------------------------------------------------------------- CLI
def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--dataset", required=True)
    ap.add_argument("--M", default="32,64", help="comma-separated max-degree values to build (default: 32,64)")
    ap.add_argument("--centroid-count", type=int, default=1,
                    help="ExRaBitQ centroid count K (default 1; K in 1..256, 1-byte id)")
    ap.add_argument("--centroid-train-samples", type=int, default=100000,
                    help="rows averaged into the global centre")
    ap.add_argument("--data-root", default="data")
    ap.add_argument("--out-root", default=str(Path("results") / "memory_environment"))
    ap.add_argument("--threads", type=int, default=64, help="recorded in the outputs; the GPU sets its own parallelism")
    ap.add_argument("--val-queries", type=int, default=1000)
    ap.add_argument("--repeats", type=int, default=1)
    ap.add_argument("--search-list-sizes", default=None,
                    help="comma-separated search list sizes (default: the runner's 10..460 sweep)")
    args = ap.parse_args(argv)

    degrees = [int(v.strip()) for v in args.M.split(",") if v.strip()]
    if not degrees:
        ap.error("--M must contain at least one degree")
    for degree in degrees:
        if degree not in (32, 64):
            ap.error("--M values must be 32 or 64")
    if not 1 <= args.centroid_count <= 256:
        raise SystemExit("ExRaBitQ centroid count K must be in 1..256 "
                         "(the 1-byte centroid-id layout caps K at 256).")

    out_root = Path(args.out_root)
    _, _, test_q, test_gt = prepare_query_splits(args.dataset, Path(args.data_root), out_root, args.val_queries)
    ds = Dataset(args.dataset, Path(args.data_root) / args.dataset / f"{args.dataset}_base.fvecs", test_q, test_gt)
    ds.shape()
    for degree in degrees:
        cfg = RunConfig(max_degree=degree, threads=args.threads, repeats=args.repeats,
                        centroid_count=args.centroid_count, centroid_train_samples=args.centroid_train_samples)
        if args.search_list_sizes:
            cfg.search_list_sizes = [int(v) for v in args.search_list_sizes.split(",") if v.strip()]
        name = graph_path(out_root, args.dataset, degree).name
        candidates = [out_root / SUITE / args.dataset / "indexes" / METHOD / name,
                      graph_path(out_root, args.dataset, degree)]  # where the build writes it
        graph = next((p for p in candidates if p.exists()), None)
        paths = output_paths(out_root, args.dataset, degree, cfg.build_beam)
        print(f"suite={SUITE}\ndataset={ds.name}\nbase_path={ds.base_path}\nquery_path={ds.query_path}\n"
              f"gt_path={ds.gt_path}\nmax_degree={degree}\nbuild_beam={cfg.build_beam}\nmethods={METHOD}\n"
              f"query_coarse_codecs={cfg.query_coarse_codec}" + (f"\ngraph_file={graph}" if graph else ""), flush=True)
        payload = run_ours_exrabitq4(ds, cfg, out_root, paths["raw_log"], graph_file=graph)
        write_raw_log(paths["raw_log"], ds, cfg, payload)
        append_raw_csv(paths["raw_csv"], ds, cfg, payload, paths["raw_log"])
        append_manifest(paths["manifest_csv"], ds, cfg, payload, paths["raw_log"])
        print(f"method={METHOD} status=done metrics=real raw_log={paths['raw_log']}", flush=True)
        keep_only_formal_int8_rows(paths["raw_csv"], degree)
        raw = paths["raw_log"]
        if raw.exists():
            dst = Path.cwd() / "Ours" / "logs" / args.dataset
            dst.mkdir(parents=True, exist_ok=True)
            shutil.copy2(raw, dst / raw.name)
            print(f"log archived to {dst / raw.name}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
