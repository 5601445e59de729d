"""Learned adaptive rerank depth for the batched graph search.

No counterpart in the reference: the reference reranks a fixed
min(|pool|, max(100, k)) pool entries with the residual-4 correction
(ours_diskann.rs:1813-1902). Here a small MLP (24 -> 64 -> 64 -> 1, BF16
matrix-core forward with FP32 accumulation, kernels in csrc/qg_mlp.hip)
predicts for each of a query's first 100 pool candidates a high quantile
of the 4-bit estimator's error |exact - estimate| / ||q||^2. Candidate i is
reranked only while its lower bound, estimate - k * error * ||q||^2, can
still beat the largest upper bound of the first ten; the depth is clamped
to [10, 100]. The plain top-100 rerank (preds=None) is the fallback and the
judge: the multiplier k is calibrated on held-out queries so that recall
against the plain rerank drops by no more than a set budget.

The network is trained on the GPU by the library's own forward / backward
kernels (pinball loss at quantile tau, Adam on FP32 master weights) on
(estimate, exact) pairs that the search itself produces.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check
from . import space as S
from .space import _dev, _ptr, _stream

N_FEATURES = 24
FEATURE_STRIDE = 32
N_PARAMS = 5825


def init_params(seed):
    """Uniform fan-in initialisation: W1 ~ U(+-1/sqrt(24)), W2, W3 ~
    U(+-1/8), biases 0 (same values as oracle/mlp_oracle.init_params)."""
    rng = np.random.default_rng(seed)
    P = np.zeros(N_PARAMS, np.float32)
    P[0:1536] = rng.uniform(-1, 1, 1536) / np.sqrt(24)
    P[1600:5696] = rng.uniform(-1, 1, 4096) / 8.0
    P[5760:5824] = rng.uniform(-1, 1, 64) / 8.0
    return P


class MLP:
    """Parameters (FP32, device) and the library's train / inference kernels."""

    def __init__(self, params=None, seed=0):
        assert _lib.lib().qg_mlp_param_count() == N_PARAMS
        P = init_params(seed) if params is None else np.asarray(params, np.float32)
        self.params = torch.from_numpy(P.copy()).cuda()

    def forward(self, feats):
        """feats [n, 32] int16 (bf16 bits) on the device -> predictions [n]."""
        feats = feats.contiguous()
        n = feats.shape[0]
        y = torch.empty(n, dtype=torch.float32, device="cuda")
        check(_lib.lib().qg_mlp_forward(_ptr(self.params), _ptr(feats), n, _ptr(y), _stream()))
        return y

    def _ws(self, n):
        return torch.empty(_lib.lib().qg_mlp_workspace_bytes(n), dtype=torch.uint8, device="cuda")

    def grad(self, feats, targets, tau):
        """Mean pinball loss and its gradient [N_PARAMS] over the batch."""
        feats, targets = feats.contiguous(), targets.contiguous().float()
        n = feats.shape[0]
        g = torch.empty(N_PARAMS, dtype=torch.float32, device="cuda")
        loss = torch.empty(1, dtype=torch.float32, device="cuda")
        ws = self._ws(n)
        check(_lib.lib().qg_mlp_grad(_ptr(self.params), _ptr(feats), _ptr(targets), n, C.c_float(tau), _ptr(g),
                                     _ptr(loss), _ptr(ws), ws.numel(), _stream()))
        return loss, g

    def fit(self, feats, targets, tau=0.95, epochs=3, batch=8192, lr=2e-3, seed=0, log=None):
        """Adam over shuffled mini-batches; returns the per-epoch mean loss."""
        L = _lib.lib()
        n = feats.shape[0]
        m = torch.zeros(N_PARAMS, dtype=torch.float32, device="cuda")
        v = torch.zeros(N_PARAMS, dtype=torch.float32, device="cuda")
        ws = self._ws(batch)
        loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        step = 0
        history = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g, device="cuda")
            total = torch.zeros(1, dtype=torch.float64, device="cuda")
            nb = 0
            for s in range(0, n - batch + 1 if n >= batch else 1, batch):
                idx = perm[s:s + batch]
                xb = feats[idx].contiguous()
                tb = targets[idx].contiguous().float()
                step += 1
                check(L.qg_mlp_adam_step(_ptr(self.params), _ptr(m), _ptr(v), _ptr(xb), _ptr(tb), xb.shape[0],
                                         C.c_float(tau), C.c_float(lr), C.c_float(0.9), C.c_float(0.999),
                                         C.c_float(1e-8), step, _ptr(loss), _ptr(ws), ws.numel(), _stream()))
                total += loss.double()
                nb += 1
            history.append(float(total.item() / max(nb, 1)))
            if log:
                log("epoch %d: mean pinball loss %.6f" % (ep, history[-1]))
        return history


def residual_stats(sp, index):
    out = torch.empty((index.n, 4), dtype=torch.float32, device="cuda")
    d = index.desc()
    check(_lib.lib().qg_residual_stats(sp.handle, C.byref(d), _ptr(out), _stream()))
    return out


def search_pool(sp, index, queries, degrees, edges, l_search, eps0=1.9, slots=None, workspace=None,
                check_graph=True, events=None):
    """Traversal only (qg_search without outputs): pools [nq, L] and stats.
    With check_graph=False and a preallocated `workspace` (sized for `slots`)
    nothing synchronises; read the error flag later with
    space.search_error(workspace). `events` = (start, end) CUDA events
    recorded around the launch alone."""
    degrees = _dev(degrees, torch.int32).reshape(-1)
    edges = _dev(edges, torch.int32)
    nq = queries.nq
    L = l_search
    slots = slots or max(1, min(nq, 4096))
    pool_ids = torch.empty((nq, L), dtype=torch.int32, device="cuda")
    pool_d = torch.empty((nq, L), dtype=torch.float32, device="cuda")
    stats = torch.zeros((nq, 6), dtype=torch.int64, device="cuda")
    Lb = _lib.lib()
    ws_bytes = Lb.qg_search_workspace_bytes(index.n, slots)
    ws = workspace if workspace is not None else torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    if ws.numel() < ws_bytes:
        raise ValueError("search workspace too small: need %d bytes" % ws_bytes)
    di, dq = index.desc(), queries.desc()
    if events is not None:
        events[0].record(torch.cuda.current_stream())
    check(Lb.qg_search(sp.handle, C.byref(di), C.byref(dq), queries.codec, _ptr(degrees), _ptr(edges),
                       edges.shape[1], 10, L, C.c_float(eps0), 100, slots, None, None, _ptr(stats), _ptr(pool_ids),
                       _ptr(pool_d), _ptr(ws), ws.numel(), _stream()))
    if events is not None:
        events[1].record(torch.cuda.current_stream())
    if check_graph:
        S.search_error(ws)
    return pool_ids, pool_d, stats


def features(sp, index, queries, pool_ids, pool_d, rstats, depth=100):
    """Features [nq * depth, 32] (bf16 bits as int16) of each query's first
    `depth` pool entries (zero rows past the pool)."""
    nq, L = pool_ids.shape
    out = torch.empty((nq * depth, FEATURE_STRIDE), dtype=torch.int16, device="cuda")
    di, dq = index.desc(), queries.desc()
    check(_lib.lib().qg_rerank_features(sp.handle, C.byref(di), C.byref(dq), _ptr(pool_ids), _ptr(pool_d), L, depth,
                                        _ptr(rstats), _ptr(out), _stream()))
    return out


def rerank(sp, index, queries, pool_ids, pool_d, preds=None, kmult=1.0, depth=100, min_depth=10, k=10):
    """Residual-4 rerank of each query's pool: the first min(|pool|, depth)
    entries (preds None: the plain rerank), or the learned depth. Returns
    (ids [nq, k], estimates [nq, k], depths [nq])."""
    nq, L = pool_ids.shape
    ids = torch.empty((nq, k), dtype=torch.int32, device="cuda")
    dists = torch.empty((nq, k), dtype=torch.float32, device="cuda")
    depths = torch.empty(nq, dtype=torch.int32, device="cuda")
    di, dq = index.desc(), queries.desc()
    check(_lib.lib().qg_learned_rerank(sp.handle, C.byref(di), C.byref(dq), queries.codec, _ptr(pool_ids),
                                       _ptr(pool_d), L, depth, _ptr(preds) if preds is not None else None,
                                       C.c_float(kmult), min_depth, k, _ptr(ids), _ptr(dists), _ptr(depths),
                                       _stream()))
    return ids, dists, depths


def query_norm_sqr(queries):
    return queries.qs[:, 4:8].contiguous().view(torch.float32).reshape(-1)


def pairs_from_pool(sp, index, x, q, qv, pid, pd, rstats, depth=100):
    """(features, targets) of the first `depth` entries of prepared queries'
    pools (ids pid / estimates pd, sorted): target = |exact squared L2 -
    estimate| / ||q - c||^2, exact from the raw fp32 vectors x [n, dim]."""
    f = features(sp, index, qv, pid, pd, rstats, depth)
    d = min(depth, pid.shape[1])
    ids = pid[:, :d].long()
    valid = ids >= 0
    safe = torch.where(valid, ids, torch.zeros_like(ids))
    exact = ((x[safe] - q[:, None, :]) ** 2).sum(-1)
    err = (exact - pd[:, :d]).abs() / query_norm_sqr(qv).clamp_min(1e-30)[:, None]
    t = torch.zeros((q.shape[0], depth), dtype=torch.float32, device="cuda")
    t[:, :d] = torch.where(valid, err, torch.zeros_like(err))
    keep = torch.zeros((q.shape[0], depth), dtype=torch.bool, device="cuda")
    keep[:, :d] = valid
    keep = keep.reshape(-1)
    return f[keep], t.reshape(-1)[keep]


def training_pairs(sp, index, degrees, edges, x, raw_queries, l_search, codec, rstats, depth=100, chunk=2048):
    """Training pairs from graph searches: the first `depth` entries of each
    query's traversal pool."""
    feats, targets = [], []
    for s0 in range(0, raw_queries.shape[0], chunk):
        q = raw_queries[s0:s0 + chunk]
        qv = sp.prepare(q, codec)
        pid, pd, _ = search_pool(sp, index, qv, degrees, edges, l_search)
        f, t = pairs_from_pool(sp, index, x, q, qv, pid, pd, rstats, depth)
        feats.append(f)
        targets.append(t)
    return torch.cat(feats), torch.cat(targets)


def fixed_pool(sp, index, qv, cand_ids, depth=100, codec="int8"):
    """The first-stage pool of fixed-candidate scoring: each query's `depth`
    candidates with the smallest first-stage estimates, in (estimate, id)
    order."""
    from . import fixed
    first = fixed.score(sp, index, qv, cand_ids, codec=codec)
    pid, pd, _ = S.topk(first, cand_ids, min(depth, cand_ids.shape[1]))
    return pid, pd


def training_pairs_fixed(sp, index, x, raw_queries, cand_ids, codec, rstats, depth=100, chunk=2048):
    """Training pairs from fixed-candidate scoring: the first `depth`
    entries of each query's first-stage pool."""
    feats, targets = [], []
    for s0 in range(0, raw_queries.shape[0], chunk):
        q = raw_queries[s0:s0 + chunk]
        qv = sp.prepare(q, codec)
        pid, pd = fixed_pool(sp, index, qv, cand_ids[s0:s0 + chunk], depth, codec)
        f, t = pairs_from_pool(sp, index, x, q, qv, pid, pd, rstats, depth)
        feats.append(f)
        targets.append(t)
    return torch.cat(feats), torch.cat(targets)


def recall_at_k(ids, gt, k=10):
    return float((ids[:, :k].unsqueeze(2) == gt[:, :k].unsqueeze(1)).any(2).float().sum(1).mean().item()) / k


class LearnedRerank:
    """A trained depth model plus its calibrated error multiplier."""

    def __init__(self, mlp, kmult, depth=100, min_depth=10):
        self.mlp, self.kmult, self.depth, self.min_depth = mlp, kmult, depth, min_depth

    def predict(self, sp, index, qv, pool_ids, pool_d, rstats):
        f = features(sp, index, qv, pool_ids, pool_d, rstats, self.depth)
        return self.mlp.forward(f)

    def search(self, sp, index, qv, degrees, edges, l_search, rstats, k=10, workspace=None, check_graph=True):
        """Traversal, pool write-out, features, MLP and learned rerank as
        four launches (the reference form the fused search is checked against)."""
        pid, pd, st = search_pool(sp, index, qv, degrees, edges, l_search, workspace=workspace,
                                  check_graph=check_graph)
        preds = self.predict(sp, index, qv, pid, pd, rstats)
        ids, dists, depths = rerank(sp, index, qv, pid, pd, preds, self.kmult, self.depth, self.min_depth, k)
        return ids, dists, depths, st

    def packed(self):
        """The weights rounded to bf16 and laid out per lane (qg_mlp_pack),
        packed again whenever the parameters change."""
        key = (self.mlp.params.data_ptr(), self.mlp.params._version)
        if getattr(self, "_packed_key", None) != key:
            Lb = _lib.lib()
            self._packed = torch.empty(Lb.qg_mlp_packed_bytes(), dtype=torch.uint8, device="cuda")
            check(Lb.qg_mlp_pack(_ptr(self.mlp.params), _ptr(self._packed), _stream()))
            self._packed_key = key
        return self._packed

    def search_fused(self, sp, index, qv, degrees, edges, l_search, rstats, k=10, eps0=1.9, slots=None,
                     workspace=None, check_graph=True, events=None):
        """One launch (qg_search_learned): each query's workgroup computes
        its pool's features, the MLP's errors and the depth right after its
        traversal and reranks that many; equal to search() bit for bit.
        Returns (ids [nq, k], estimates, depths [nq], stats [nq, 6])."""
        degrees = _dev(degrees, torch.int32).reshape(-1)
        edges = _dev(edges, torch.int32)
        nq = qv.nq
        slots = slots or max(1, min(nq, 4096))
        ids = torch.empty((nq, k), dtype=torch.int32, device="cuda")
        dists = torch.empty((nq, k), dtype=torch.float32, device="cuda")
        depths = torch.empty((nq,), dtype=torch.int32, device="cuda")
        stats = torch.zeros((nq, 6), dtype=torch.int64, device="cuda")
        Lb = _lib.lib()
        ws_bytes = Lb.qg_search_workspace_bytes(index.n, slots)
        ws = workspace if workspace is not None else torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
        if ws.numel() < ws_bytes:
            raise ValueError("search workspace too small: need %d bytes" % ws_bytes)
        di, dq = index.desc(), qv.desc()
        if events is not None:
            events[0].record(torch.cuda.current_stream())
        check(Lb.qg_search_learned(sp.handle, C.byref(di), C.byref(dq), qv.codec, _ptr(degrees), _ptr(edges),
                                   edges.shape[1], k, l_search, None, C.c_float(eps0), self.depth, slots, 4,
                                   _ptr(self.packed()), C.c_float(self.kmult), self.min_depth, _ptr(rstats),
                                   _ptr(ids), _ptr(dists), _ptr(stats), _ptr(depths), _ptr(ws), ws.numel(),
                                   _stream()))
        if events is not None:
            events[1].record(torch.cuda.current_stream())
        if check_graph:
            S.search_error(ws)
        return ids, dists, depths, stats


CALIBRATION_GRID = (0.25, 0.35, 0.5, 0.6, 0.75, 1.0, 1.25, 1.5, 2.0, 3.0, 4.0, 6.0)


def calibrate(sp, index, qv, pool_ids, pool_d, preds, gt, budget=0.001, grid=CALIBRATION_GRID, depth=100,
              min_depth=10):
    """Smallest multiplier whose recall is within `budget` of the plain
    rerank's on these (held-out) queries; returns (kmult, table)."""
    plain, _, _ = rerank(sp, index, qv, pool_ids, pool_d, None, 1.0, depth, min_depth)
    r_plain = recall_at_k(plain, gt)
    table = []
    best = None
    for km in grid:
        ids, _, depths = rerank(sp, index, qv, pool_ids, pool_d, preds, km, depth, min_depth)
        r = recall_at_k(ids, gt)
        table.append({"kmult": km, "recall": r, "mean_depth": float(depths.float().mean().item())})
        if best is None and r >= r_plain - budget:
            best = km
    return (best if best is not None else grid[-1]), {"plain_recall": r_plain, "grid": table}
