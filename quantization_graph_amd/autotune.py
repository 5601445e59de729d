"""Launch-shape autotuner for the batched graph search.

The search kernel's launch shape has three free choices that do not change
any result:
* waves per query workgroup (2 or 4): how many 16-lane groups share a
  query's gate and distance passes;
* slots: how many query workgroups run at once. The resident limit follows
  from LDS (static part + qg_search_lds_bytes(code_dim, L) per workgroup, 160
  KB per CU) and from registers (4 waves per SIMD);
* the order in which the work queue hands out queries: as given, or
  longest list first, which shortens the tail of a batch with uneven lists.

`SearchTuner.tune` times every candidate on the caller's own batch, with HIP
events around the kernel, and keeps the fastest per workload key (code
dimension, index size, batch size, list-length profile, codec). The table
can be saved to and loaded from JSON. Results are identical for every
candidate; the order candidate permutes the queries and the caller gets its
answers back in its own order.
"""
from __future__ import annotations

import json
import math
import os

import torch

from . import _lib

STATIC_LDS_BYTES = 19 * 1024 + 1024   # per-workgroup static LDS of the search kernel (rounded up)
LDS_PER_CU = 160 * 1024
WAVES_PER_SIMD = 4                    # register-limited (<= 128 VGPRs)


def resident_slots(cd, L, waves, cus=None):
    """Query workgroups that fit the GPU at once for code dimension cd and
    pool capacity L."""
    if cus is None:
        cus = torch.cuda.get_device_properties(0).multi_processor_count
    lds = STATIC_LDS_BYTES + _lib.lib().qg_search_lds_bytes(cd, L)
    per_cu = min(LDS_PER_CU // lds, (4 * WAVES_PER_SIMD) // waves)
    return max(1, per_cu) * cus


def workload_key(sp, n, nq, l_search, l_per_query, codec):
    if l_per_query is None:
        prof = (int(l_search),)
    else:
        v = l_per_query.float()
        prof = (int(l_search), int(v.median().item()), int(v.max().item()))
    return "cd%d_n%d_nq%d_L%s_codec%d" % (sp.code_dim, n, nq, "-".join(map(str, prof)), codec)


class SearchTuner:
    def __init__(self, path=None):
        self.path = path
        self.table = {}
        if path and os.path.exists(path):
            with open(path) as f:
                self.table = json.load(f)

    def save(self):
        if self.path:
            with open(self.path, "w") as f:
                json.dump(self.table, f, indent=1, sort_keys=True)

    def candidates(self, sp, l_search, nq):
        out = []
        for waves in (4, 2):
            cap = resident_slots(sp.code_dim, l_search, waves)
            for slots in sorted({max(1, cap // 2), cap, 2 * cap}):
                for order in ("given", "longest_first"):
                    if slots > 2 * nq:
                        continue
                    out.append({"waves": waves, "slots": min(slots, nq), "order": order})
        return out

    @staticmethod
    def order_of(l_per_query, order):
        """Permutation the work queue sees (None: as given)."""
        if order != "longest_first" or l_per_query is None:
            return None
        return torch.argsort(l_per_query.long(), descending=True, stable=True)

    def tune(self, sp, index, queries_raw, codec, degrees, edges, k=10, l_search=400, l_per_query=None,
             rerank=100, reps=2, force=False):
        """Time every candidate on this batch and return the fastest:
        {"waves", "slots", "order", "kernel_ms"} (+ "table": all timings)."""
        nq = queries_raw.shape[0]
        key = workload_key(sp, index.n, nq, l_search, l_per_query, 3 if codec == "int8" else 0)
        if key in self.table and not force:
            return self.table[key]
        timings = []
        for cand in self.candidates(sp, l_search, nq):
            perm = self.order_of(l_per_query, cand["order"])
            qx = queries_raw if perm is None else queries_raw[perm]
            lq = l_per_query if perm is None or l_per_query is None else l_per_query[perm]
            qv = sp.prepare(qx, codec)
            ws = sp.search_workspace(index.n, cand["slots"])
            ms = []
            for r in range(reps + 1):
                ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                sp.search(index, qv, degrees, edges, k=k, l_search=l_search, l_per_query=lq, rerank=rerank,
                          slots=cand["slots"], waves=cand["waves"], workspace=ws, events=ev)
                torch.cuda.synchronize()
                if r:  # the first run is a warm-up
                    ms.append(ev[0].elapsed_time(ev[1]))
            timings.append(dict(cand, kernel_ms=sorted(ms)[len(ms) // 2]))
            del ws
        best = min(timings, key=lambda c: c["kernel_ms"])
        res = dict(best, table=timings)
        self.table[key] = res
        self.save()
        return res


def lognormal_lists(nq, median, sigma, lo, hi, seed):
    """Per-query search list sizes: lognormal with the given median and
    log-space sigma, rounded and clipped to [lo, hi] (int32, on the device)."""
    g = torch.Generator(device="cpu")
    g.manual_seed(seed)
    z = torch.randn(nq, generator=g, dtype=torch.float64)
    v = torch.exp(math.log(median) + sigma * z).round().clamp(lo, hi)
    return v.to(torch.int32).cuda()
