"""Fixed-candidate scoring with the residual-4 rerank, for a batch of queries.

The reference's fixed-candidate protocol scores a per-query candidate list
with the 4-bit estimator and ranks by (estimate, index)
(experiments/01_quantizer_fair/faiss_quantizer_smoke.cpp:1399-1417,
431-478). The rerank stage follows the graph-search rerank
(experiments/02_diskann_fair/src/ours_diskann.rs:1813-1902): the first
``rerank`` candidates in (distance, id) order get their full FP32 4-bit
distance (INT8 codec: recomputed, :1850-1866) and the residual-4 correction
(rabitq_residual_distance_batch), then the final top-k by (estimate, id).

Everything runs on the device for the whole batch: score -> top-rerank ->
(FP32 restore -> residual -> top-k). With k <= rerank depth <= 128 the
bracketed part is one launch of the workgroup rerank kernel (one 256-thread
workgroup per query: block-parallel residual sums, the ordered block
recurrence per candidate, top-k by (estimate, id)); otherwise three launches
that also return every reranked candidate's estimate.
"""
import torch

from . import space as S


def score(sp, index, queries, cand_ids, codec="int8", eps0=1.9):
    """First-stage estimates [nq, P] for candidate ids [nq, P].

    codec "int8": INT8 query x full 4-bit code, exact integer dot
    (finish_int_paper_prune_estimate). codec "full": FP32 query x 4-bit code
    (queryDistanceLong)."""
    kernel = S.SCAN_INT8_FULL if S._CODECS[codec] == S.CODEC_INT8 else S.SCAN_ASYM
    return sp.scan_fixed(kernel, index, queries, cand_ids, eps0=eps0)[0]


def rerank(sp, index, queries, cand_ids, first, k=10, depth=100, restore_full=True):
    """Residual-4 rerank of the first ``depth`` candidates by (first, id).

    Returns (ids [nq, k], estimates [nq, k], rerank ids [nq, depth],
    rerank estimates [nq, depth])."""
    depth = min(depth, cand_ids.shape[1])
    top_ids, top_first, _ = S.topk(first, cand_ids, depth)
    long_d = sp.scan_fixed(S.SCAN_ASYM, index, queries, top_ids)[0] if restore_full else top_first
    est = sp.scan_fixed(S.SCAN_RESIDUAL, index, queries, top_ids, long_in=long_d, want=("f", "iv"))[0]
    ids, dists, _ = S.topk(est, top_ids, k)
    return ids, dists, top_ids, est


def rerank_topk(sp, index, queries, cand_ids, first, k=10, depth=100):
    """The final top-k of rerank() (same ids and estimates, bit for bit),
    without the per-candidate estimates: one workgroup-per-query launch after
    the top-`depth` selection when k <= depth <= 128. The FP32 long distance
    is restored for INT8 queries, as rerank(restore_full=True) does."""
    depth = min(depth, cand_ids.shape[1])
    if not (k <= depth <= 128):
        ids, dists, _, _ = rerank(sp, index, queries, cand_ids, first, k=k, depth=depth,
                                  restore_full=queries.codec == S.CODEC_INT8)
        return ids, dists
    from . import learned
    top_ids, top_first, _ = S.topk(first, cand_ids, depth)
    ids, dists, _ = learned.rerank(sp, index, queries, top_ids, top_first, None, depth=depth, k=k)
    return ids, dists


def search(sp, index, queries, cand_ids, k=10, depth=100, codec="int8", eps0=1.9):
    """Score + rerank for a prepared query batch; returns (ids, estimates)."""
    first = score(sp, index, queries, cand_ids, codec=codec, eps0=eps0)
    return rerank_topk(sp, index, queries, cand_ids, first, k=k, depth=depth)


def recall_at_k(found, truth, k=10):
    """Mean |found[:k] & truth[:k]| / k over queries (diskann_runner.rs:1922-1924)."""
    f = found[:, :k].long().cpu()
    t = truth[:, :k].long().cpu()
    hits = (f.unsqueeze(2) == t.unsqueeze(1)).any(2).sum(1)
    return float(hits.float().mean()) / k
