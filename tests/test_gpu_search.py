"""Batched GPU graph search against the oracle's restatement of the
reference search loop (ours_diskann.rs:1539-1986): identical final ids and
estimates, identical final pools and identical hop / gate / distance /
prune counts, for both query codecs and several list sizes, on a Vamana
graph built by the reference bridge compiled from source."""
import numpy as np
import pytest

from conftest import gaussian_mixture
from oracle_bind import Oracle, Ref


def _ref_counters(st):
    """The oracle's counters in the order search_counters returns them."""
    v, dc, hops, checked, would, not_pruned, msb, rem = st.reference_counters()
    return (hops, v, dc, checked, would, would, not_pruned, msb, rem)

pytestmark = pytest.mark.gpu


CASES = {960: dict(n=3000, r=32, l_build=64), 1536: dict(n=3000, r=64, l_build=128)}


@pytest.fixture(scope="module", params=sorted(CASES), ids=lambda d: "D%d" % d)
def graph_case(request):
    """D=960 (code dimension 1024) on a degree-32 graph, and D=1536 (code
    dimension 2048, the graph workload's) on a degree-64 graph, where a hop
    can bring 64 fresh neighbours and 64 survivors and code rows are staged
    16 per batch instead of 32."""
    dim = request.param
    cfg = CASES[dim]
    n = cfg["n"]
    x, q = gaussian_mixture(n, dim, 16, 0.3, 21, nq=12)
    c = x.mean(0).astype(np.float32)
    orc = Oracle(dim, 100)
    orc.set_center(c)
    comp, res = orc.split(orc.encode_full(x))
    msb, fac = orc.sidecar(comp)
    if Ref.available():
        ref = Ref(dim, 100)
        ref.set_center(c)
        deg, edges, _, _, _ = ref.build_graph(comp, cfg["r"], cfg["l_build"], 1.2, beam=1, batch=512, refine=1,
                                              cap=128)
    else:  # a connected random graph when the reference build is absent
        rng = np.random.default_rng(0)
        deg = np.full(n, cfg["r"], np.uint32)
        edges = rng.integers(0, n, size=(n, cfg["r"])).astype(np.uint32)
        edges[:, 0] = (np.arange(n) + 1) % n
    print("case: graph of %d x %d, max degree %d" % (n, dim, int(deg.max())), flush=True)
    return dict(dim=dim, x=x, q=q, c=c, orc=orc, comp=comp, res=res, msb=msb, fac=fac, deg=deg, edges=edges)


def _compare_with_oracle(gpu, g, deg, edges, codec, L):
    import quantization_graph_amd as qg
    from quantization_graph_amd.space import search_counters
    print("case: D %d codec %d L %d" % (g["dim"], codec, L), flush=True)
    sp = qg.RabitqSpace(g["dim"], 100)
    sp.set_center(g["c"])
    ix = qg.DeviceIndex.from_records(sp, g["comp"], g["res"])
    qv = sp.prepare(gpu.from_numpy(g["q"]).cuda(), codec)
    ids, dists, stats, pool, pool_d = sp.search(ix, qv, deg.astype(np.int32), edges.astype(np.int32),
                                                k=10, l_search=L, eps0=1.9, rerank=100, want_pool=True)
    ids, dists, stats = ids.cpu().numpy(), dists.cpu().numpy(), stats.cpu().numpy()
    pool, pool_d = pool.cpu().numpy(), pool_d.cpu().numpy()
    orc = g["orc"]
    for i in range(g["q"].shape[0]):
        oq = orc.prepare(g["q"][i], codec)
        o_ids, o_d, o_pool, o_pd, st = orc.search(oq, codec, g["comp"], g["res"], g["msb"], g["fac"], deg,
                                                  edges, k=10, L=L, eps0=1.9, rerank=100)
        m = len(o_pool)
        assert np.array_equal(pool[i, :m].astype(np.uint32), o_pool), "pool ids (query %d)" % i
        assert np.array_equal(pool_d[i, :m].view(np.uint32), o_pd.view(np.uint32)), "pool distances"
        assert np.array_equal(ids[i, :len(o_ids)].astype(np.uint32), o_ids)
        assert np.array_equal(dists[i, :len(o_d)].view(np.uint32), o_d.view(np.uint32))
        assert tuple(stats[i]) == st.as_tuple()
        assert tuple(int(v) for v in search_counters(stats[i]).values()) == _ref_counters(st)
        orc.free_query(oq)


@pytest.mark.parametrize("codec", [3, 0])
@pytest.mark.parametrize("L", [16, 64, 200])
def test_search_matches_oracle(gpu, graph_case, codec, L):
    _compare_with_oracle(gpu, graph_case, graph_case["deg"], graph_case["edges"], codec, L)


@pytest.mark.parametrize("codec", [3, 0])
def test_search_duplicate_neighbours_match_oracle(gpu, graph_case, codec):
    """Adjacency lists with repeated entries: only a neighbour's first
    occurrence is a candidate (the reference marks it visited as it walks the
    list), whether the kernel prepared the list ahead during the previous
    hop or loads it on the spot."""
    g = graph_case
    deg, edges = g["deg"].copy(), g["edges"].copy()
    rng = np.random.default_rng(5)
    for v in range(edges.shape[0]):
        d = int(deg[v])
        if d < 4:
            continue
        for _ in range(3):  # copy earlier entries over later slots
            a, b = sorted(rng.choice(d, size=2, replace=False))
            edges[v, b] = edges[v, a]
    _compare_with_oracle(gpu, g, deg, edges, codec, 64)


@pytest.mark.parametrize("codec", [3, 0])
def test_search_odd_row_stride_matches_oracle(gpu, graph_case, codec):
    """Adjacency tables whose row stride is not a multiple of four words (33
    for the degree-32 graph, 63 for the degree-64 one, its lists cut to 63):
    the kernel stages the candidate next frontiers' rows word by word
    instead of in 16-byte pieces; results still equal the oracle's."""
    g = graph_case
    deg, edges = g["deg"], g["edges"]
    stride = 33 if edges.shape[1] <= 32 else 63
    keep = min(edges.shape[1], stride)
    e2 = np.full((edges.shape[0], stride), 0xFFFFFFFF, np.uint32)
    e2[:, :keep] = edges[:, :keep]
    d2 = np.minimum(deg, stride).astype(np.uint32)
    _compare_with_oracle(gpu, g, d2, e2, codec, 64)


def test_search_rejects_bad_graph(gpu, graph_case):
    import quantization_graph_amd as qg
    g = graph_case
    sp = qg.RabitqSpace(g["dim"], 100)
    sp.set_center(g["c"])
    ix = qg.DeviceIndex.from_records(sp, g["comp"], g["res"])
    qv = sp.prepare(gpu.from_numpy(g["q"][:2]).cuda(), 3)
    edges = g["edges"].astype(np.int64).copy()
    edges[0, 0] = g["x"].shape[0] + 5  # out of range from the start node
    with pytest.raises(qg._lib.QGError):
        sp.search(ix, qv, g["deg"].astype(np.int32), edges.astype(np.int32), k=10, l_search=32)
    with pytest.raises(Exception):
        sp.search(ix, qv, g["deg"].astype(np.int32), np.zeros((g["x"].shape[0], 65), np.int32), k=10)


@pytest.mark.parametrize("codec", [3, 0])
def test_search_rejects_bad_neighbour_past_the_start(gpu, graph_case, codec):
    """An out-of-range neighbour in lists reached after the first hop (which
    the kernel prepares ahead of time) is reported as well."""
    import quantization_graph_amd as qg
    g = graph_case
    sp = qg.RabitqSpace(g["dim"], 100)
    sp.set_center(g["c"])
    ix = qg.DeviceIndex.from_records(sp, g["comp"], g["res"])
    qv = sp.prepare(gpu.from_numpy(g["q"][:4]).cuda(), codec)
    edges = g["edges"].astype(np.int64).copy()
    edges[1:, 3] = g["x"].shape[0] + 7  # every list but the start node's
    with pytest.raises(qg._lib.QGError):
        sp.search(ix, qv, g["deg"].astype(np.int32), edges.astype(np.int32), k=10, l_search=32)
    # the same graph without the bad entries searches cleanly afterwards
    sp.search(ix, qv, g["deg"].astype(np.int32), g["edges"].astype(np.int32), k=10, l_search=32)


PER_QUERY_L = [10, 16, 37, 64, 100, 200, 600, 1200, 2500, 4096, 23, 77]


@pytest.mark.parametrize("waves", [4, 2])
@pytest.mark.parametrize("codec", [3, 0])
def test_search_per_query_lists(gpu, graph_case, codec, waves):
    """Uneven search lists in one launch (pool capacity 4096, above the 64 KB
    LDS default): every query equals the oracle's search at its own L, for
    two- and four-wave workgroups."""
    import quantization_graph_amd as qg
    from quantization_graph_amd.space import search_counters
    g = graph_case
    sp = qg.RabitqSpace(g["dim"], 100)
    sp.set_center(g["c"])
    ix = qg.DeviceIndex.from_records(sp, g["comp"], g["res"])
    print("case: D", g["dim"], "per-query L", PER_QUERY_L, "codec", codec, "waves", waves, flush=True)
    qv = sp.prepare(gpu.from_numpy(g["q"]).cuda(), codec)
    lq = gpu.tensor(PER_QUERY_L, dtype=gpu.int32, device="cuda")
    ids, dists, stats, pool, pool_d = sp.search(ix, qv, g["deg"].astype(np.int32), g["edges"].astype(np.int32),
                                                k=10, l_search=4096, l_per_query=lq, waves=waves, rerank=100,
                                                want_pool=True, slots=5)
    ids, dists, stats = ids.cpu().numpy(), dists.cpu().numpy(), stats.cpu().numpy()
    pool, pool_d = pool.cpu().numpy(), pool_d.cpu().numpy()
    orc = g["orc"]
    for i, L in enumerate(PER_QUERY_L):
        oq = orc.prepare(g["q"][i], codec)
        o_ids, o_d, o_pool, o_pd, st = orc.search(oq, codec, g["comp"], g["res"], g["msb"], g["fac"], g["deg"],
                                                  g["edges"], k=10, L=L, eps0=1.9, rerank=100)
        m = len(o_pool)
        assert np.array_equal(pool[i, :m].astype(np.uint32), o_pool), "pool ids (query %d, L %d)" % (i, L)
        assert np.all(pool[i, m:] == -1)
        assert np.array_equal(pool_d[i, :m].view(np.uint32), o_pd.view(np.uint32))
        assert np.array_equal(ids[i, :len(o_ids)].astype(np.uint32), o_ids)
        assert np.array_equal(dists[i, :len(o_d)].view(np.uint32), o_d.view(np.uint32))
        assert tuple(stats[i]) == st.as_tuple()
        assert tuple(int(v) for v in search_counters(stats[i]).values()) == _ref_counters(st)
        orc.free_query(oq)


def test_search_graph_capture_replays(gpu, graph_case):
    """Query preparation + search captured in one HIP graph: every replay,
    with eager searches and synchronisations in between, equals the eager
    result."""
    import quantization_graph_amd as qg
    g = graph_case
    sp = qg.RabitqSpace(g["dim"], 100)
    sp.set_center(g["c"])
    ix = qg.DeviceIndex.from_records(sp, g["comp"], g["res"])
    qx = gpu.from_numpy(g["q"]).cuda()
    deg = gpu.from_numpy(g["deg"].astype(np.int32)).cuda()
    edges = gpu.from_numpy(g["edges"].astype(np.int32)).cuda()
    lq = gpu.tensor(PER_QUERY_L, dtype=gpu.int32, device="cuda")
    ws = sp.search_workspace(ix.n, 8)
    print("case: graph capture of prepare + search, 12 queries", flush=True)
    e_ids, e_d, e_st = sp.search(ix, sp.prepare(qx, "int8"), deg, edges, k=10, l_search=4096, l_per_query=lq,
                                 slots=8, workspace=ws)
    side = gpu.cuda.Stream()
    side.wait_stream(gpu.cuda.current_stream())
    with gpu.cuda.stream(side):  # warm the allocator on the capture stream
        sp.search(ix, sp.prepare(qx, "int8"), deg, edges, k=10, l_search=4096, l_per_query=lq, slots=8,
                  workspace=ws)
    gpu.cuda.current_stream().wait_stream(side)
    gpu.cuda.synchronize()
    graph = gpu.cuda.CUDAGraph()
    with gpu.cuda.graph(graph):
        g_ids, g_d, g_st = sp.search(ix, sp.prepare(qx, "int8"), deg, edges, k=10, l_search=4096, l_per_query=lq,
                                     slots=8, workspace=ws, check_graph=False)
    for rep in range(3):
        graph.replay()
        gpu.cuda.synchronize()
        qg.space.search_error(ws)
        assert gpu.equal(g_ids, e_ids) and gpu.equal(g_d.view(gpu.int32), e_d.view(gpu.int32)), rep
        assert gpu.equal(g_st, e_st)
        # eager work between replays
        sp.search(ix, sp.prepare(qx, "full"), deg, edges, k=10, l_search=64, slots=8, workspace=ws)
        gpu.cuda.synchronize()
