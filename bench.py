#!/usr/bin/env python
"""Headline benchmark on one MI355X.

Workload (the 1M x 960 precision-form configuration): 1,000,000 x 960 fp32
vectors from a 256-component Gaussian mixture (centers N(0, 1), per-cluster
std 0.3, seed 2), ExRaBitQ4 block16 encode (seed-100 FHT rotation, MSE fp16
residual scales, mean center), 10,000 queries (base points + N(0, 0.05)
noise) quantized to INT8, 1,000 fixed candidates per query (the exact-L2
top-10 plus 990 drawn from exact ranks 11..2000, shuffled). One step =
prepare the 10k queries + INT8 x 4-bit scoring of all 10M (query,
candidate) pairs + per-query top-100 + FP32 restore + residual-4 rerank +
top-10. Recall@10 is measured against brute-force fp32 ground truth.

Prints one JSON line (see the keys below). ``--gpus N`` with N > 1 exits:
this build runs on one GPU.
"""
import argparse
import ctypes as C
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


def cpu_reference_baseline(comp, res, cand, gt, q_host, center, dim, budget_s, max_queries):
    """The reference's own fixed-candidate path (mode-0 FP32 4-bit scoring,
    top-100 by (distance, id), residual-4 rerank, top-10) through its C-ABI
    bridge compiled from source (oracle/_ref), one thread."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import Ref
    if not Ref.available():
        return None, None
    ref = Ref(dim, 100)
    ref.set_center(center)
    ref.set_codec(0)
    nq = min(max_queries, cand.shape[0])
    found = np.zeros((nq, 10), np.int64)
    zeros = np.zeros(cand.shape[1], np.uint8)
    done = 0
    t0 = time.perf_counter()
    for i in range(nq):
        ids = cand[i].astype(np.uint32)
        pq = ref.prepare(q_host[i])
        d = ref.distance_batch(pq, comp, ids, zeros)
        order = np.lexsort((ids, d))[:100]
        top = ids[order]
        iv = ref.residual_batch(pq, comp, res, top, d[order])
        fin = np.lexsort((top, iv["estimate"]))[:10]
        found[i] = top[fin]
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    elapsed = time.perf_counter() - t0
    hits = [len(set(found[i]) & set(gt[i, :10])) for i in range(done)]
    return {"qps": done / elapsed, "queries": done, "seconds": elapsed,
            "recall_at_10": float(np.mean(hits)) / 10.0}, found[:done]


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--dim", type=int, default=960)
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--per", type=int, default=1000)
    ap.add_argument("--depth", type=int, default=100)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-max-queries", type=int, default=4000)
    ap.add_argument("--no-cpu", action="store_true")
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

    def step(events=None):
        qv = sp.prepare(q, "int8")
        if events is not None:
            events[0].record(stream)
        first = fixed.score(sp, index, qv, cand, codec="int8")
        if events is not None:
            events[1].record(stream)
        ids, dists, _, _ = fixed.rerank(sp, index, qv, cand, first, k=10, depth=args.depth)
        return ids, dists

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    kev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        ids, dists = step(kev[i])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1000.0 * wall / args.steps
    kernel_ms = float(np.mean([a.elapsed_time(b) for a, b in kev]))
    recall = fixed.recall_at_k(ids, gt, 10)
    qps = nq / (ms_per_step / 1000.0)

    cd = sp.code_dim
    unit_bytes = cd // 2 + 8 + 4 + 4  # codes + {norm_sqr, long_scale} + id + estimate
    units = nq * per
    achieved = units * unit_bytes / (kernel_ms / 1000.0) / 1e9
    traffic = None
    pmc = os.path.join(ROOT, "profiles", "int8_full_fixed_1Mx960_pmc.json")
    if os.path.exists(pmc):
        with open(pmc) as f:
            traffic = json.load(f).get("traffic_bytes_per_launch")

    cpu = None
    parity = None
    if not args.no_cpu:
        comp, res = index.to_records()
        comp_h, res_h = comp.cpu().numpy(), res.cpu().numpy()
        del comp, res
        ns = min(args.cpu_max_queries, nq)
        cand_h = cand[:ns].cpu().numpy()
        q_h = q[:ns].cpu().numpy()
        gt_h = gt[:ns].cpu().numpy()
        base, ref_found = cpu_reference_baseline(comp_h, res_h, cand_h, gt_h, q_h, center, args.dim,
                                                 args.cpu_seconds, ns)
        if base is not None:
            done = base["queries"]
            # the same FP32-first-stage pipeline on the GPU, compared id-for-id
            qf = sp.prepare(q[:done], "full")
            gids, _ = fixed.search(sp, index, qf, cand[:done], k=10, depth=args.depth, codec="full")
            parity = float((gids.cpu().numpy() == ref_found).all(1).mean())
            cpu = {"value": base["qps"], "unit": "queries/s", "cores": 1, "kind": "reference",
                   "sample": "%d queries x %d candidates of this workload (%.1f s), mode-0 FP32 scoring + "
                             "residual-4 top-%d rerank via the reference bridge compiled from source"
                             % (done, per, base["seconds"], args.depth),
                   "recall_at_10": base["recall_at_10"]}
            log("cpu reference: %.1f q/s on %d queries; GPU vs reference top-10 id parity %.4f"
                % (base["qps"], done, parity))

    out = {
        "metric": "QPS at Recall@10>=0.95 (fixed candidates, residual4 top-100 rerank); Gdist/s vs HBM roofline",
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
        "config": {"workload": "fixed-candidate INT8 x 4-bit scoring + residual4 top-%d rerank, %d x %d GMM(256, "
                               "std 0.3, seed 2), %d queries x %d candidates" % (args.depth, args.n, args.dim, nq, per),
                   "base": args.n, "dim": args.dim, "queries": nq, "candidates_per_query": per,
                   "rerank_depth": args.depth, "codec": "int8"},
        "recall_at_10": recall,
        "gdist_per_s": units / (kernel_ms / 1000.0) / 1e9,
        "encode_seconds": encode_s,
        "encode_gvec_per_s": args.n / encode_s / 1e9,
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
