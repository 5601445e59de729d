"""Properties at the workloads' full sizes, where the oracle cannot replay
everything: the batched search over 1,000,000 x 1536 vectors (code
dimension 2048, degree-64 graph built on the GPU) against the oracle's
search loop on a spread sample of queries, and the same search's results
independent of how the queries are batched (slots, waves, order, per-query
lists); the encode of 1,000,000 x 960 vectors independent of the batch a
vector is encoded in, and a spread sample of its records against the
oracle's encode."""
import numpy as np
import pytest

from oracle_bind import Oracle

pytestmark = pytest.mark.gpu


def _powerlaw(torch, n, dim, nq, seed, exponent=0.8, qnoise=0.1, chunk=131072):
    """Unit vectors with a power-law spectrum (z ~ N(0, diag(i^-exponent))
    in a seeded orthogonal basis); queries are perturbed base points."""
    basis, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(dim, dim)))
    basis = torch.from_numpy(basis.astype(np.float32)).cuda()
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    scale = torch.arange(1, dim + 1, device="cuda", dtype=torch.float32).pow(-exponent / 2.0)
    x = torch.empty((n, dim), device="cuda")
    for s in range(0, n, chunk):
        z = torch.randn((min(chunk, n - s), dim), generator=g, device="cuda") * scale
        x[s:s + z.shape[0]] = torch.nn.functional.normalize(z @ basis.T, dim=1)
    pick = torch.randint(0, n, (nq,), generator=g, device="cuda")
    noise = (qnoise / dim ** 0.5) * torch.randn((nq, dim), generator=g, device="cuda")
    return x, torch.nn.functional.normalize(x[pick] + noise, dim=1).contiguous()


@pytest.fixture(scope="module")
def full_graph(gpu):
    import quantization_graph_amd as qg
    torch = gpu
    n, dim, nq = 1_000_000, 1536, 200
    print("case: 1M x 1536 power-law, degree-64 graph", flush=True)
    x, q = _powerlaw(torch, n, dim, nq, 11)
    center = x.double().mean(0).float().cpu().numpy()
    sp = qg.RabitqSpace(dim, 100)
    sp.set_center(center)
    index = sp.encode(x, batch=131072)
    del x
    deg, edges, _ = sp.build_graph(index, r=64, l_build=128, alpha=1.2, beam_width=1, batch=4096,
                                   refine_passes=0, prune_cap=256)
    torch.cuda.synchronize()
    comp, res = index.to_records()
    return dict(sp=sp, index=index, deg=deg, edges=edges, q=q, center=center, comp=comp.cpu().numpy(),
                res=res.cpu().numpy())


@pytest.mark.parametrize("codec", [0, 3])
def test_search_full_size_matches_oracle(gpu, full_graph, codec):
    """Pools, answers, estimates and counters of 12 queries spread over the
    batch equal the oracle's search loop on the same records and graph."""
    g = full_graph
    sp = g["sp"]
    print("case: full-size search codec %d" % codec, flush=True)
    qv = sp.prepare(g["q"], codec)
    ids, dists, stats, pool, pool_d = sp.search(g["index"], qv, g["deg"], g["edges"], k=10, l_search=100,
                                                eps0=1.9, rerank=100, want_pool=True)
    ids, dists, stats = ids.cpu().numpy(), dists.cpu().numpy(), stats.cpu().numpy()
    pool, pool_d = pool.cpu().numpy(), pool_d.cpu().numpy()
    orc = Oracle(1536, 100)
    orc.set_center(g["center"])
    msb, fac = orc.sidecar(g["comp"])
    deg = g["deg"].cpu().numpy().astype(np.uint32)
    edges = g["edges"].cpu().numpy().astype(np.uint32)
    qh = g["q"].cpu().numpy()
    for i in np.linspace(0, qh.shape[0] - 1, 12).astype(int):
        oq = orc.prepare(qh[i], codec)
        o_ids, o_d, o_pool, o_pd, st = orc.search(oq, codec, g["comp"], g["res"], msb, fac, deg, edges, k=10,
                                                  L=100, eps0=1.9, rerank=100)
        orc.free_query(oq)
        m = len(o_pool)
        assert np.array_equal(pool[i, :m].astype(np.uint32), o_pool), "pool ids (query %d)" % i
        assert np.array_equal(pool_d[i, :m].view(np.uint32), o_pd.view(np.uint32)), "pool distances (query %d)" % i
        assert np.array_equal(ids[i, :len(o_ids)].astype(np.uint32), o_ids), "ids (query %d)" % i
        assert np.array_equal(dists[i, :len(o_d)].view(np.uint32), o_d.view(np.uint32)), "estimates (query %d)" % i
        assert tuple(stats[i]) == st.as_tuple(), "counters (query %d)" % i


def test_search_full_size_batch_invariance(gpu, full_graph):
    """A query's answer, estimates and counters do not depend on the launch
    shape: all queries at once on 4 waves, 13 slots on 2 waves, the batch in
    reverse order with per-query lists equal to the launch's list."""
    torch = gpu
    g = full_graph
    sp = g["sp"]
    qv = sp.prepare(g["q"], 0)
    a_ids, a_d, a_st = sp.search(g["index"], qv, g["deg"], g["edges"], k=10, l_search=100, rerank=100)
    b_ids, b_d, b_st = sp.search(g["index"], qv, g["deg"], g["edges"], k=10, l_search=100, rerank=100, slots=13,
                                 waves=2)
    rev = torch.arange(g["q"].shape[0] - 1, -1, -1, device="cuda")
    qr = sp.prepare(g["q"][rev].contiguous(), 0)
    lists = torch.full((g["q"].shape[0],), 100, dtype=torch.int32, device="cuda")
    c_ids, c_d, c_st = sp.search(g["index"], qr, g["deg"], g["edges"], k=10, l_search=400, l_per_query=lists,
                                 rerank=100)
    assert torch.equal(a_ids, b_ids) and torch.equal(a_d.view(torch.int32), b_d.view(torch.int32))
    assert torch.equal(a_st, b_st)
    assert torch.equal(a_ids, c_ids[rev]) and torch.equal(a_d.view(torch.int32), c_d[rev].view(torch.int32))
    assert torch.equal(a_st, c_st[rev])


def test_encode_full_size_batch_independent(gpu):
    """1,000,000 x 960 (the fixed-candidate workload's size): 4,096 spread
    vectors encoded on their own give the same record bytes as inside the
    full batch, and 64 of them equal the oracle's encode."""
    import quantization_graph_amd as qg
    torch = gpu
    n, dim = 1_000_000, 960
    print("case: encode 1M x 960", flush=True)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(2)
    centers = torch.randn((256, dim), generator=gen, device="cuda")
    x = centers[torch.randint(0, 256, (n,), generator=gen, device="cuda")]
    x.add_(0.3 * torch.randn((n, dim), generator=gen, device="cuda"))
    center = x.double().mean(0).float().cpu().numpy()
    sp = qg.RabitqSpace(dim, 100)
    sp.set_center(center)
    comp, res = sp.encode(x, batch=262144).to_records()
    pick = torch.from_numpy(np.linspace(0, n - 1, 4096).astype(np.int64)).cuda()
    comp2, res2 = sp.encode(x[pick].contiguous()).to_records()
    assert torch.equal(comp[pick], comp2) and torch.equal(res[pick], res2)
    orc = Oracle(dim, 100)
    orc.set_center(center)
    sub = pick[::64]
    o_comp, o_res = orc.split(orc.encode_full(x[sub].cpu().numpy()))
    assert np.array_equal(comp[sub].cpu().numpy(), o_comp)
    assert np.array_equal(res[sub].cpu().numpy(), o_res)
