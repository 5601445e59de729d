#!/usr/bin/env python
"""Headline benchmark on one MI355X.

Workload (the 1M x 960 precision-form configuration): 1,000,000 x 960 fp32
vectors from a 256-component Gaussian mixture (centers N(0, 1), per-cluster
std 0.3, seed 2), ExRaBitQ4 block16 encode (seed-100 FHT rotation, MSE fp16
residual scales, mean center), 10,000 queries (base points + N(0, 0.05)
noise) quantized to INT8, 1,000 fixed candidates per query (the exact-L2
top-10 plus 990 drawn from exact ranks 11..2000, shuffled). One step =
prepare the 10k queries + INT8 x 4-bit scoring of all 10M (query,
candidate) pairs + per-query top-`depth` + FP32 restore + residual-4 rerank
+ top-10. Recall@10 is measured against brute-force fp32 ground truth; the
rerank depth is swept and the headline is the fastest depth that reaches
Recall@10 >= 0.95.

Prints one JSON line (see the keys below). ``--gpus N`` with N > 1 exits:
this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_data(torch, n, dim, nq, clusters, std, seed, qstd):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    centers = torch.randn((clusters, dim), generator=g, device="cuda")
    labels = torch.randint(0, clusters, (n,), generator=g, device="cuda")
    x = centers[labels]
    x.add_(std * torch.randn((n, dim), generator=g, device="cuda"))
    pick = torch.randint(0, n, (nq,), generator=g, device="cuda")
    q = x[pick] + qstd * torch.randn((nq, dim), generator=g, device="cuda")
    return x.contiguous(), q.contiguous(), g


def exact_topk(torch, x, q, k, chunk=512):
    """Brute-force fp32 L2 top-k ids [nq, k] (ties by index)."""
    xn = (x * x).sum(1)
    out = []
    for s in range(0, q.shape[0], chunk):
        qc = q[s:s + chunk]
        d = xn.unsqueeze(0) - 2.0 * (qc @ x.T)
        out.append(torch.topk(d, k, dim=1, largest=False, sorted=True).indices)
    return torch.cat(out).to(torch.int32)


def make_candidates(torch, top2000, per, g):
    """GT top-10 + (per - 10) drawn without replacement from ranks 11..2000,
    shuffled per query."""
    nq, r = top2000.shape
    keys = torch.rand((nq, r - 10), generator=g, device="cuda")
    pick = torch.argsort(keys, dim=1)[:, :per - 10] + 10
    cand = torch.cat([top2000[:, :10], torch.gather(top2000, 1, pick)], 1)
    perm = torch.argsort(torch.rand((nq, per), generator=g, device="cuda"), dim=1)
    return torch.gather(cand, 1, perm).contiguous()


def _ref_query(ref, comp, res, ids, q, depth):
    import numpy as np
    zeros = np.zeros(ids.size, np.uint8)
    pq = ref.prepare(q)
    d = ref.distance_batch(pq, comp, ids, zeros)
    order = np.lexsort((ids, d))[:depth]
    top = ids[order]
    iv = ref.residual_batch(pq, comp, res, top, d[order])
    return top[np.lexsort((top, iv["estimate"]))[:10]]


def cpu_reference_baseline(comp, res, cand, gt, q_host, center, dim, depth, sample, repeats, threads):
    """The reference's own fixed-candidate path (mode-0 FP32 4-bit scoring,
    top-`depth` by (distance, id), residual-4 rerank, top-10) through its C-ABI
    bridge compiled from source (oracle/_ref). Protocol of the reference
    runner: one warm-up pass, then `repeats` timed passes single-threaded,
    median reported; plus one pass over all host cores (one bridge space per
    thread; ctypes releases the GIL)."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import Ref
    if not Ref.available():
        return None, None
    ref = Ref(dim, 100)
    ref.set_center(center)
    ref.set_codec(0)
    ns = min(sample, cand.shape[0])
    ids = cand[:ns].astype(np.uint32)
    found = np.zeros((ns, 10), np.int64)
    for i in range(min(ns, 64)):
        _ref_query(ref, comp, res, ids[i], q_host[i], depth)
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        for i in range(ns):
            found[i] = _ref_query(ref, comp, res, ids[i], q_host[i], depth)
        times.append(time.perf_counter() - t0)
    single = ns / float(np.median(times))

    spaces = []
    for _ in range(threads):
        r = Ref(dim, 100)
        r.set_center(center)
        r.set_codec(0)
        spaces.append(r)

    def work(t):
        for i in range(t, ns, threads):
            _ref_query(spaces[t], comp, res, ids[i], q_host[i], depth)

    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(work, range(threads)))
    multi = ns / (time.perf_counter() - t0)
    hits = [len(set(found[i]) & set(gt[i, :10])) for i in range(ns)]
    return {"qps": single, "qps_all_cores": multi, "threads": threads, "queries": ns,
            "seconds": float(np.sum(times)), "recall_at_10": float(np.mean(hits)) / 10.0}, found


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--dim", type=int, default=960)
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--per", type=int, default=1000)
    ap.add_argument("--depths", default="100,150,200,300,400,600",
                    help="residual-rerank depths swept; the headline is the smallest with Recall@10 >= 0.95")
    ap.add_argument("--target-recall", type=float, default=0.95)
    ap.add_argument("--cpu-queries", type=int, default=2000)
    ap.add_argument("--cpu-repeats", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run 5 device-to-device copies of BYTES (a plain copy kernel of known size, "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only; --gpus %d is not supported by this build" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import quantization_graph_amd as qg
    from quantization_graph_amd import fixed

    torch.cuda.set_device(0)
    if args.calibrate_copy:
        src = torch.ones(args.calibrate_copy // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": args.calibrate_copy}))
        return
    t_setup = time.perf_counter()
    x, q, g = make_data(torch, args.n, args.dim, args.nq, 256, 0.3, 2, 0.05)
    center = x.double().mean(0).float().cpu().numpy()
    log("data ready", tuple(x.shape), tuple(q.shape))
    top2000 = exact_topk(torch, x, q, 2000)
    gt = top2000[:, :10].contiguous()
    cand = make_candidates(torch, top2000, args.per, g)
    del top2000
    log("ground truth + candidates ready")

    sp = qg.RabitqSpace(args.dim, 100)
    sp.set_center(center)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    index = sp.encode(x, batch=262144)
    torch.cuda.synchronize()
    encode_s = time.perf_counter() - t0
    log("encoded %d vectors in %.2f s" % (args.n, encode_s))

    nq, per = args.nq, args.per
    stream = torch.cuda.current_stream()
    depths = [int(d) for d in args.depths.split(",") if d]

    def step(depth, events=None):
        qv = sp.prepare(q, "int8")
        if events is not None:
            events[0].record(stream)
        first = fixed.score(sp, index, qv, cand, codec="int8")
        if events is not None:
            events[1].record(stream)
        ids, dists, _, _ = fixed.rerank(sp, index, qv, cand, first, k=10, depth=depth)
        return ids, dists

    def timed(depth, steps, warmup):
        for _ in range(warmup):
            step(depth)
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        t0 = time.perf_counter()
        for i in range(steps):
            ids, _ = step(depth, ev[i])
        torch.cuda.synchronize()
        ms = 1000.0 * (time.perf_counter() - t0) / steps
        kms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
        return ms, kms, fixed.recall_at_k(ids, gt, 10)

    sweep = []
    for d in depths:
        ms, kms, rec = timed(d, max(3, args.steps // 4), 1)
        sweep.append({"depth": d, "ms_per_step": ms, "qps": nq / (ms / 1000.0), "recall_at_10": rec})
        log("depth %4d: %.3f ms/step  %.0f q/s  recall@10 %.4f" % (d, ms, nq / (ms / 1000.0), rec))
    ok = [r for r in sweep if r["recall_at_10"] >= args.target_recall]
    best = max(r["recall_at_10"] for r in sweep)
    if not ok:
        # the method's recall plateaus below the target on this data: take the
        # fastest depth within 0.002 of the plateau
        ok = [r for r in sweep if r["recall_at_10"] >= best - 0.002]
    depth = min(ok, key=lambda r: r["ms_per_step"])["depth"]
    ms_per_step, kernel_ms, recall = timed(depth, args.steps, args.warmup)
    qps = nq / (ms_per_step / 1000.0)

    cd = sp.code_dim
    unit_bytes = cd // 2 + 8 + 4 + 4  # codes + {norm_sqr, long_scale} + id + estimate
    units = nq * per
    achieved = units * unit_bytes / (kernel_ms / 1000.0) / 1e9
    traffic = None
    pmc = os.path.join(ROOT, "profiles", "int8_full_fixed_1Mx960_pmc.json")
    if os.path.exists(pmc) and args.n == 1_000_000 and args.dim == 960:
        with open(pmc) as f:
            traffic = json.load(f).get("traffic_bytes_per_launch")

    cpu = None
    parity = None
    if not args.no_cpu:
        comp, res = index.to_records()
        comp_h, res_h = comp.cpu().numpy(), res.cpu().numpy()
        del comp, res
        ns = min(args.cpu_queries, nq)
        cand_h = cand[:ns].cpu().numpy()
        q_h = q[:ns].cpu().numpy()
        gt_h = gt[:ns].cpu().numpy()
        threads = min(16, os.cpu_count() or 1)
        base, ref_found = cpu_reference_baseline(comp_h, res_h, cand_h, gt_h, q_h, center, args.dim, depth,
                                                 ns, args.cpu_repeats, threads)
        if base is not None:
            # the same FP32-first-stage pipeline on the GPU, compared id-for-id
            qf = sp.prepare(q[:ns], "full")
            gids, _ = fixed.search(sp, index, qf, cand[:ns], k=10, depth=depth, codec="full")
            parity = float((gids.cpu().numpy() == ref_found).all(1).mean())
            cpu = {"value": base["qps"], "unit": "queries/s", "cores": 1, "kind": "reference",
                   "sample": "%d queries x %d candidates of this workload, warm-up + %d repeats (median), "
                             "mode-0 FP32 scoring + residual-4 top-%d rerank through the reference bridge "
                             "compiled from source" % (ns, per, args.cpu_repeats, depth),
                   "value_all_cores": base["qps_all_cores"], "all_cores_threads": base["threads"],
                   "recall_at_10": base["recall_at_10"]}
            log("cpu reference: %.1f q/s (1 thread), %.1f q/s (%d threads); GPU vs reference top-10 id "
                "parity %.4f" % (base["qps"], base["qps_all_cores"], base["threads"], parity))

    out = {
        "metric": "QPS at Recall@10>=0.95 (fixed candidates, residual4 rerank); Gdist/s vs HBM roofline",
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic",
        "config": {"workload": "fixed-candidate INT8 x 4-bit scoring + residual4 rerank, %d x %d GMM(256, std 0.3, "
                               "seed 2), %d queries x %d candidates" % (args.n, args.dim, nq, per),
                   "base": args.n, "dim": args.dim, "queries": nq, "candidates_per_query": per,
                   "rerank_depth": depth, "codec": "int8"},
        "recall_at_10": recall,
        "recall_target_met": recall >= args.target_recall,
        "recall_plateau": best,
        "depth_sweep": sweep,
        "gdist_per_s": units / (kernel_ms / 1000.0) / 1e9,
        "encode_seconds": encode_s,
        "encode_mvec_per_s": args.n / encode_s / 1e6,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_int8_full<true> (INT8 query x 4-bit code, fixed lists)",
                     "kernel_ms": kernel_ms, "bytes_per_unit": unit_bytes, "units_per_launch": units},
        "cpu_baseline": cpu,
        "gpu_vs_reference_id_parity": parity,
        "setup_seconds": time.perf_counter() - t_setup,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
