"""The oracle's query-time search loop (ro_search, oracle/rabitq_oracle.c)
against the reference's own compiled graph search, on reference-built graphs
where the pool fills and both prune sites fire.

* vamana::search_vamana_paper_active (vamana_search.h:216-364), driven by
  oracle/search_ref.cpp: final pool ids and distances bit for bit, and the
  reference's search counters.
* VamanaIndex::searchKnnPlainThenResidualRerank (vamana_index.h:530-636)
  with setPaperPrune(true, 1.9) over a VMNARBQ1 file and the external
  residual file: top-10 ids bit for bit, reranked estimates within 1e-5
  relative (its per-record residual sum runs in another order than the
  runner's batch call), and the same counters.
* A trace of every prune decision: a candidate is pruned exactly when the
  pool is full, its gate is valid and lower_bound > the pool back at its
  slot, at the first pass and at the re-check.

The reference's C++ loop scores slot by slot; the Rust loop the oracle
restates (ours_diskann.rs:1539-1986) gates the hop's fresh list against the
pool as it stood at the start of the hop and re-checks survivors at
insertion. Both give the same pools and counters because the pool back only
falls within a hop. Only the Full (FP32) codec is compared: with an INT8
query the C++ twin takes the FP32 remaining-bit distance after the INT8
gate, where the Rust loop takes the INT8 one (mode 2), and its rerank skips
the FP32 long-distance restore."""
import numpy as np
import pytest

from conftest import gaussian_mixture
from oracle_bind import Oracle, Ref, RefGraphSearch

pytestmark = pytest.mark.skipif(not (Ref.available() and RefGraphSearch.available()),
                                reason="reference sources not built (oracle/_ref)")

CASES = {960: dict(n=2000, r=32, l_build=64, nq=8), 1536: dict(n=1500, r=64, l_build=96, nq=6)}


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request, tmp_path_factory):
    from quantization_graph_amd import formats
    dim = request.param
    cfg = CASES[dim]
    x, q = gaussian_mixture(cfg["n"], dim, 12, 0.3, 31 + dim, nq=cfg["nq"])
    c = x.mean(0).astype(np.float32)
    orc = Oracle(dim, 100)
    orc.set_center(c)
    comp, res = orc.split(orc.encode_full(x))
    msb, fac = orc.sidecar(comp)
    ref = Ref(dim, 100)
    ref.set_center(c)
    deg, edges, _, _, _ = ref.build_graph(comp, cfg["r"], cfg["l_build"], 1.2, beam=1, batch=256, refine=1, cap=128)
    d = tmp_path_factory.mktemp("search_ref_%d" % dim)
    formats.write_vamana_index(d / "index.bin", comp, deg, edges, cfg["r"], cfg["l_build"], 1.2)
    res.tofile(d / "residual.bin")
    gs = RefGraphSearch(dim, 100, c)
    gs.set_codec(0)
    gs.open_index(d / "index.bin", d / "residual.bin", comp.shape[0])
    return dict(dim=dim, x=x, q=q, orc=orc, comp=comp, res=res, msb=msb, fac=fac, deg=deg, edges=edges, gs=gs)


@pytest.mark.parametrize("L", [16, 64, 200])
def test_oracle_search_equals_reference_search(case, L):
    g = case
    orc, gs = g["orc"], g["gs"]
    edges = np.ascontiguousarray(g["edges"], np.uint32)
    deg = np.ascontiguousarray(g["deg"], np.uint32)
    fired = np.zeros(2, np.int64)
    for i in range(g["q"].shape[0]):
        oq = orc.prepare(g["q"][i], 0)
        o_ids, o_d, o_pool, o_pd, st, trace = orc.search_traced(oq, 0, g["comp"], g["res"], g["msb"], g["fac"],
                                                                deg, edges, k=10, L=L, eps0=1.9, rerank=100)
        orc.free_query(oq)
        # the traversal: pool ids and distances, and the counters
        r_pool, r_pd, r_st = gs.pool(g["q"][i], g["comp"], g["msb"], g["fac"], deg, edges, L)
        assert len(o_pool) == min(L, g["comp"].shape[0]) and len(r_pool) == len(o_pool)
        assert np.array_equal(o_pool, r_pool), "pool ids (query %d, L %d)" % (i, L)
        assert np.array_equal(o_pd.view(np.uint32), r_pd.view(np.uint32)), "pool distances"
        assert st.reference_counters() == r_st.as_tuple(), (st.as_tuple(), r_st.as_tuple())
        # searchKnnPlainThenResidualRerank: top-10 after the residual rerank
        r_ids, r_d, r_m = gs.knn(g["q"][i], 10, L, 100)
        assert np.array_equal(o_ids, r_ids), "top-10 ids (query %d, L %d)" % (i, L)
        # the index path sums each residual record with residualInt4IpAvx2
        # (space_rabitq.h:1723-1735, 2278), the runner's batch call with the
        # batched AVX2 loop (:5432-5548, which the oracle follows bit for bit
        # against the compiled bridge): the same sums in another order
        np.testing.assert_allclose(o_d, r_d, rtol=1e-5, atol=0)
        assert st.reference_counters() == r_m.as_tuple()
        # the trace: pruned exactly when full, valid and lower_bound > back
        want = (trace["full"] == 1) & (trace["valid"] == 1) & (trace["lower_bound"] > trace["back"])
        assert np.array_equal(trace["pruned"] == 1, want)
        assert int((trace["pruned"] == 1).sum()) == st.pruned
        assert int(((trace["stage"] == 1) & (trace["full"] == 1) & (trace["valid"] == 1)).sum()) == st.recheck_checked
        for stage in (0, 1):
            fired[stage] += int(((trace["stage"] == stage) & (trace["pruned"] == 1)).sum())
        # every gated candidate has a first-pass decision; survivors a re-check
        assert int((trace["stage"] == 0).sum()) == st.gate_evals
        assert int((trace["stage"] == 1).sum()) == st.full_dists - 1
    if L < 200:  # the pool fills well before the traversal ends
        assert fired[0] > 0, "no first-pass prune fired"
    if L == 16:
        assert fired[1] > 0, "no re-check prune fired"
    assert fired.sum() > 0

