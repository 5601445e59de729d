"""Pins the CPU oracle (oracle/rabitq_oracle.c) before it is trusted:

* against the reference's own outputs (tests/golden/ref_vectors.npz, written
  by tests/golden/make_golden.py from the reference compiled from source);
* against what the mathematics fixes: the standard's mt19937 check values,
  orthogonality of the rotation, exact integer dots by brute force, optimality
  of fastQuantizeAbs over all codes on tiny inputs, the FP64 form of every
  float estimator, IEEE half values, and brute-force RobustPrune / search on
  tiny inputs;
* the Vamana build against the reference bridge's graphs (fixture and, where
  oracle/_ref is built, live).
"""
import ctypes as C
import itertools
import os

import numpy as np
import pytest

from conftest import gaussian_mixture
from oracle_bind import ESTIMATE_DTYPE, FACTOR_DTYPE, INTERVAL_DTYPE, Oracle, Ref, RefSearch

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_vectors.npz")


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


def test_mt19937_check_values():
    o = Oracle(64, 1)
    assert o.lib.ro_mt32_nth(5489, 10000) == 4123659995
    assert o.lib.ro_mt64_nth(5489, 10000) == 9981545732273789042


def test_rotation_is_scaled_hadamard():
    from scipy.linalg import hadamard
    rng = np.random.default_rng(0)
    for dim in (100, 960):
        o = Oracle(dim, 100)
        x = rng.normal(size=dim).astype(np.float32)
        r = o.rotate(x)
        cd = o.cd
        xp = np.zeros(cd)
        xp[:dim] = x
        dense = hadamard(cd).astype(np.float64) @ (o.signs().astype(np.float64) * xp)
        assert np.max(np.abs(r - dense)) < 1e-4 * np.abs(dense).max()
        assert abs(np.linalg.norm(r) - np.linalg.norm(x)) < 1e-5 * np.linalg.norm(x)
        assert set(np.unique(np.abs(o.signs()))) == {np.float32(1.0 / np.sqrt(np.float32(cd)))}


def test_half_rounding_on_representable_values():
    o = Oracle(64, 1)
    h = np.arange(0, 0x7C00, dtype=np.uint16)  # every finite non-negative half
    f = h.view(np.float16).astype(np.float32)
    back = np.array([o.lib.ro_float_to_half(float(v)) for v in f[::97]], np.uint16)
    assert np.array_equal(back, h[::97])
    dec = np.array([o.lib.ro_half_to_float(int(v)) for v in h[::97]], np.float32)
    assert np.array_equal(dec, f[::97])
    # halfway cases round up (add 0x1000, truncate), unlike round-to-nearest-even
    one_half_ulp = np.float32(1.0 + 2.0 ** -11)
    assert o.lib.ro_float_to_half(float(one_half_ulp)) == 0x3C01


def _unpack(codes):
    lo = codes & 0x0F
    hi = codes >> 4
    return np.stack([lo, hi], -1).reshape(codes.shape[0], -1).astype(np.int64)


def test_integer_dots_by_brute_force():
    x, q = gaussian_mixture(64, 960, 4, 0.3, 0, nq=3)
    o = Oracle(960, 100)
    o.set_center(x.mean(0))
    comp, _ = o.split(o.encode_full(x))
    msb, _ = o.sidecar(comp)
    codes = _unpack(comp[:, 16:16 + o.cd // 2])
    bits = np.unpackbits(msb, axis=1, bitorder="little").astype(np.int64)
    assert np.array_equal(bits, codes >> 3)
    for i in range(q.shape[0]):
        oq = o.prepare(q[i], 3)
        xq = o.qx8(oq).astype(np.int64)
        assert oq.int8_code_sum == xq.sum()
        assert np.all(np.abs(xq) <= 128)
        _, _, dot = o.int8_full(oq, comp)
        assert np.array_equal(dot, codes @ xq)
        _, rdot = o.rem_int8(oq, comp, np.zeros(comp.shape[0], np.float32))
        assert np.array_equal(rdot, (codes & 7) @ xq)
        assert np.array_equal(o.masked_sum(oq, msb), bits @ xq)
        o.free_query(oq)
    for a, b in [(0, 1), (5, 9), (3, 3)]:
        want = ((2 * codes[a] - 15) * (2 * codes[b] - 15)).sum()
        assert o.symmetric_ip4(comp[a], comp[b]) == want


def test_fast_quantize_abs_is_optimal_on_tiny_inputs():
    """The chosen 3-bit magnitudes maximise <a, c+0.5>/||c+0.5|| over all
    8^d codes (the ExRaBitQ objective) on 4-dimensional inputs."""
    o = Oracle(64, 1)
    rng = np.random.default_rng(1)
    grid = np.array(list(itertools.product(range(8), repeat=4)), np.float64) + 0.5
    for _ in range(200):
        a = np.abs(rng.normal(size=4)).astype(np.float32)
        a /= np.float32(np.linalg.norm(a))
        code, ipn = o.fast_quantize_abs(a)
        score = (grid @ a.astype(np.float64)) / np.linalg.norm(grid, axis=1)
        c = code.astype(np.float64) + 0.5
        got = c @ a / np.linalg.norm(c)
        assert got >= score.max() - 1e-12
        assert abs(ipn - np.float32(1.0 / (c @ a.astype(np.float64)))) <= 1e-6 * ipn


def test_estimators_against_fp64_forms():
    x, q = gaussian_mixture(128, 960, 8, 0.3, 2, nq=4)
    o = Oracle(960, 100)
    o.set_center(x.mean(0))
    comp, res = o.split(o.encode_full(x))
    codes = _unpack(comp[:, 16:16 + o.cd // 2]).astype(np.float64)
    hdr = comp[:, :16].copy().view(np.float32).reshape(-1, 4)
    true = ((x[None] - q[:, None]) ** 2).sum(-1)
    errs_before, errs_after = [], []
    for i in range(q.shape[0]):
        oq = o.prepare(q[i], 0)
        qv = o.qvec(oq).astype(np.float64)
        x4ip = (codes - 7.5) @ qv * hdr[:, 1].astype(np.float64)
        f64 = hdr[:, 0].astype(np.float64) + float(oq.query_norm_sqr) - x4ip
        a = o.asym(oq, comp)
        assert np.allclose(o.asym_f64(oq, comp), f64, rtol=1e-12, atol=1e-9)
        scale = np.abs(f64) + float(oq.query_norm_sqr)
        assert np.max(np.abs(a - f64) / scale) < 1e-5
        r = o.residual(oq, res, a)
        r64 = o.residual_f64(oq, res, a.astype(np.float64))
        assert np.max(np.abs(r["estimate"] - r64) / scale) < 1e-5
        assert np.all(r["lower_bound"] <= r["estimate"]) and np.all(r["estimate"] <= r["upper_bound"])
        errs_before.append(np.abs(a - true[i]) / true[i])
        errs_after.append(np.abs(r["estimate"] - true[i]) / true[i])
        o.free_query(oq)
    # the estimators estimate ||x - q||^2; the residual stage tightens them
    assert np.mean(errs_before) < 0.05
    assert np.mean(errs_after) < 0.5 * np.mean(errs_before)


@pytest.mark.parametrize("dim", [960, 1536])
def test_oracle_matches_reference_fixture(golden, dim):
    p = "d%d_" % dim
    o = Oracle(dim, 100)
    o.set_center(golden[p + "center"])
    full = o.encode_full(golden[p + "x"])
    assert np.array_equal(full, golden[p + "full"])
    comp, res = o.split(full)
    assert np.array_equal(comp, golden[p + "compact"])
    assert np.array_equal(res, golden[p + "residual"])
    msb, fac = o.sidecar(comp)
    assert np.array_equal(msb, golden[p + "msb"])
    gf = golden[p + "factors"].reshape(-1).view(FACTOR_DTYPE)
    for f in ("norm_sqr", "data_norm", "cross_scale", "error_cross_scale", "valid"):
        assert np.array_equal(fac[f], gf[f]), f
    q = golden[p + "q"]
    for codec, name in ((0, "full"), (3, "int8")):
        for i in range(q.shape[0]):
            oq = o.prepare(q[i], codec)
            a = o.asym(oq, comp)
            assert np.array_equal(a.view(np.uint32), golden[p + name + "_dist"][i].view(np.uint32))
            g = golden[p + name + "_gate"][i].reshape(-1).view(ESTIMATE_DTYPE)
            e = o.gate(oq, msb, fac, codec, 1.9)
            for f in ("lower_bound", "short_ip", "ip_hat", "error_bound", "valid"):
                assert e[f].tobytes() == g[f].tobytes(), (name, f)
            rem = o.rem_int8(oq, comp, g["short_ip"])[0] if codec == 3 else o.rem_fp32(oq, comp, g["short_ip"])
            assert np.array_equal(rem.view(np.uint32), golden[p + name + "_rem"][i].view(np.uint32))
            iv = golden[p + name + "_residual"][i].reshape(-1).view(INTERVAL_DTYPE)
            r = o.residual(oq, res, a)
            for f in ("estimate", "lower_bound", "upper_bound"):
                assert np.array_equal(r[f].view(np.uint32), iv[f].view(np.uint32)), f
            o.free_query(oq)
    sym = np.array([o.symmetric(comp[a], comp[b]) for a, b in golden[p + "sym_pairs"]], np.float32)
    assert np.array_equal(sym.view(np.uint32), golden[p + "sym"].view(np.uint32))


def test_robust_prune_against_brute_force():
    x, _ = gaussian_mixture(40, 128, 4, 0.3, 5)
    o = Oracle(128, 100)
    o.set_center(x.mean(0))
    comp, _ = o.split(o.encode_full(x))
    rng = np.random.default_rng(3)
    for trial in range(10):
        src = int(rng.integers(0, 40))
        pool = rng.permutation(40)[:25].astype(np.uint32)
        pool = np.concatenate([pool, pool[:3]])  # duplicates are dropped
        d = np.array([o.symmetric(comp[src], comp[j]) for j in pool], np.float32)
        R, alpha = 6, 1.2
        got, evals = o.robust_prune(comp, src, pool, d, R, alpha, with_evals=True)
        # brute force: vamana_prune.h:41-156 re-derived, with the lazily
        # filled (candidate, selected-slot) memo whose fills are the count
        memo = {}

        def pair(j, k):
            if (j, k) not in memo or memo[(j, k)] < 0:
                memo[(j, k)] = np.float32(o.symmetric(comp[j], comp[sel[k]]))
                count[0] += 1
            return memo[(j, k)]
        count = [0]
        cand = sorted({int(j): float(dd) for j, dd in zip(pool, d)}.items(), key=lambda t: (np.float32(t[1]), t[0]))
        cand = [(j, dd) for j, dd in cand if j != src]
        sel, dis, cur = [], set(), np.float32(1.0)
        inc = np.float32(min(alpha, 1.2))
        while len(sel) < R:
            promoted = False
            for j, dd in cand:
                if len(sel) >= R:
                    break
                if j in dis:
                    continue
                occ = any(np.float32(cur * pair(j, k)) <= np.float32(dd) for k in range(len(sel)))
                if not occ:
                    sel.append(j)
                    dis.add(j)
                    promoted = True
            if cur == np.float32(alpha):
                break
            nxt = np.float32(min(np.float32(alpha), np.float32(cur * inc)))
            if nxt == cur:
                break
            cur = nxt
            if not promoted and len(sel) >= R:
                break
        assert list(got) == sel
        assert evals == count[0]
        assert src not in got and len(set(got.tolist())) == len(got) <= R


def test_build_matches_reference_fixture(golden):
    """Oracle build vs the graph the reference bridge built (300 x 128, R 16,
    L 64, batch 64, one refine pass, prune cap 128): same adjacency lists in
    order and the same distance_evaluations."""
    o = Oracle(128, 100)
    o.set_center(golden["g_center"])
    deg, edges, evals = o.build_graph(golden["g_compact"], 16, 64, 1.2, beam=1, batch=64, refine=1, cap=128)
    assert np.array_equal(deg, golden["g_degrees"])
    for i in range(deg.size):
        assert np.array_equal(edges[i, :deg[i]], golden["g_edges"][i, :deg[i]])
    assert evals == int(golden["g_evals"][0])


@pytest.mark.skipif(not Ref.available(), reason="reference bridge not built (oracle/_ref)")
@pytest.mark.parametrize("case", [(900, 64, 32, 100, 1.2, 1, 256, 0, 0, 0), (900, 64, 16, 50, 1.5, 2, 100, 2, 40, 0),
                                  (700, 96, 24, 80, 2.0, 3, 300, 1, 0, 3)])
def test_build_matches_reference_bridge(case):
    """Oracle build vs the reference bridge compiled from source, across
    beam widths, early stop, prune caps, refine passes and alpha."""
    n, dim, R, L, alpha, beam, batch, refine, cap, es = case
    x, _ = gaussian_mixture(n, dim, 8, 0.3, 5)
    c = x.mean(0).astype(np.float32)
    o = Oracle(dim, 100)
    o.set_center(c)
    comp, _ = o.split(o.encode_full(x))
    r = Ref(dim, 100)
    r.set_center(c)
    d1, e1, v1 = o.build_graph(comp, R, L, alpha, beam, batch, refine, cap, es)
    d2, e2, _, _, v2 = r.build_graph(comp, R, L, alpha, beam, batch, refine, cap, es)
    assert np.array_equal(d1, d2)
    for i in range(n):
        assert np.array_equal(e1[i, :d1[i]], e2[i, :d2[i]])
    assert v1 == v2


def test_search_visits_everything_when_pool_never_fills():
    """With L >= n and a connected graph the gate never prunes, so the pool
    is every node ranked by its traversal distance and the answer is the
    residual rerank of its first 100 entries."""
    n, dim = 200, 128
    x, q = gaussian_mixture(n, dim, 4, 0.3, 9, nq=3)
    o = Oracle(dim, 100)
    o.set_center(x.mean(0))
    comp, res = o.split(o.encode_full(x))
    msb, fac = o.sidecar(comp)
    rng = np.random.default_rng(4)
    deg = np.full(n, 8, np.uint32)
    edges = np.stack([np.roll(np.arange(n), -k - 1)[:] for k in range(8)], 1).astype(np.uint32)
    edges[:, 4:] = rng.integers(0, n, size=(n, 4))
    for codec in (0, 3):
        for i in range(q.shape[0]):
            oq = o.prepare(q[i], codec)
            ids, ds, pool, pd, st = o.search(oq, codec, comp, res, msb, fac, deg, edges, k=10, L=n + 10)
            assert len(pool) == n and st.pruned == 0
            g = o.gate(oq, msb, fac, codec, 1.9)
            if codec == 3:
                trav = o.rem_int8(oq, comp, g["short_ip"])[0]
            else:
                trav = o.rem_fp32(oq, comp, g["short_ip"])
            trav[0] = o.asym(oq, comp[:1])[0]  # start node: full distance
            order = sorted(range(n), key=lambda j: (trav[j], j))
            assert list(pool) == order
            top = order[:100]
            long_d = o.asym(oq, comp[top]) if codec == 3 else trav[top]
            est = o.residual(oq, res[top], long_d)["estimate"]
            final = sorted(range(100), key=lambda j: (est[j], top[j]))[:10]
            assert list(ids) == [top[j] for j in final]
            o.free_query(oq)


@pytest.mark.skipif(not RefSearch.available(), reason="reference bridge not built (oracle/_ref)")
def test_search_matches_reference_kernels():
    """The oracle's search (restated estimators) vs the same loop driving the
    reference bridge's compiled kernels: identical ids and estimates, for the
    FP32 and INT8 query codecs, on a graph the reference built."""
    n, dim = 1500, 128
    x, q = gaussian_mixture(n, dim, 8, 0.3, 21, nq=8)
    c = x.mean(0).astype(np.float32)
    o = Oracle(dim, 100)
    o.set_center(c)
    comp, res = o.split(o.encode_full(x))
    msb, fac = o.sidecar(comp)
    r = Ref(dim, 100)
    r.set_center(c)
    deg, edges, _, _, _ = r.build_graph(comp, 16, 64, 1.2, beam=1, batch=256, refine=1, cap=128)
    rs = RefSearch(r)
    for codec in (0, 3):
        r.set_codec(codec)
        ids, ds = rs.search(q, codec, comp, res, msb, fac, deg, edges, k=10, L=48, eps0=1.9, rerank=100)
        for i in range(q.shape[0]):
            oq = o.prepare(q[i], codec)
            o_ids, o_d, _, _, _ = o.search(oq, codec, comp, res, msb, fac, deg, edges, k=10, L=48, eps0=1.9,
                                           rerank=100)
            assert np.array_equal(ids[i, :len(o_ids)], o_ids), (codec, i)
            assert np.array_equal(ds[i, :len(o_d)].view(np.uint32), o_d.view(np.uint32))
            o.free_query(oq)
    # per-query list sizes in one call (the visited tags carry over between
    # queries) equal one call per query at that query's list size
    r.set_codec(0)
    lists = np.array([16, 48, 200, 10, 64, 33, 48, 128], np.uint32)
    ids_l, ds_l = rs.search_lists(q, 0, comp, res, msb, fac, deg, edges, lists, k=10, rerank=100)
    for i in range(q.shape[0]):
        one = rs.search(q, 0, comp, res, msb, fac, deg, edges, k=10, L=int(lists[i]), rerank=100, q0=i, q1=i + 1)
        assert np.array_equal(ids_l[i], one[0][i]) and np.array_equal(ds_l[i].view(np.uint32),
                                                                      one[1][i].view(np.uint32)), i


@pytest.mark.skipif(not RefSearch.available(), reason="reference bridge not built (oracle/_ref)")
def test_fixed_candidate_driver_matches_bridge_calls():
    """The C driver of the plain fixed-candidate pipeline (the CPU baseline's
    loop) against the same steps spelled out over the bridge's batch calls:
    mode-0 distances, the first 100 by (distance, id), residual intervals of
    those, top-10 by (estimate, id)."""
    n, dim, per, depth = 800, 128, 300, 100
    x, q = gaussian_mixture(n, dim, 8, 0.3, 23, nq=6)
    c = x.mean(0).astype(np.float32)
    o = Oracle(dim, 100)
    o.set_center(c)
    comp, res = o.split(o.encode_full(x))
    r = Ref(dim, 100)
    r.set_center(c)
    r.set_codec(0)
    rng = np.random.default_rng(5)
    cand = np.stack([rng.choice(n, per, replace=False) for _ in range(q.shape[0])]).astype(np.uint32)
    got = RefSearch(r).fixed(q, comp, res, cand, depth=depth)
    for i in range(q.shape[0]):
        pq = r.prepare(q[i])
        d = r.distance_batch(pq, comp, cand[i], np.zeros(per, np.uint8))
        order = np.lexsort((cand[i], d))[:depth]
        top = cand[i][order]
        est = r.residual_batch(pq, comp, res, top, d[order])["estimate"]
        want = top[np.lexsort((top, est))[:10]]
        assert np.array_equal(got[i], want), i


@pytest.mark.parametrize("dim", [960, 1536])
def test_record_gate_and_b1_match_reference_fixture(golden, dim):
    """rabitq_paper_estimate (gate from the compact record) and
    rabitq_query_distance_b1, bit for bit against the reference's outputs;
    the B1 dot against brute force and the record gate against its FP64
    form (short_ip, alpha)."""
    p = "d%d_" % dim
    o = Oracle(dim, 100)
    o.set_center(golden[p + "center"])
    comp = golden[p + "compact"]
    hdr = comp[:, :16].copy().view(np.float32).reshape(-1, 4)
    codes = _unpack(comp[:, 16:16 + o.cd // 2])
    q = golden[p + "q"]
    for codec, name in ((0, "full"), (3, "int8")):
        for i in range(q.shape[0]):
            oq = o.prepare(q[i], codec)
            e = o.paper_record(oq, comp, 1.9)
            g = golden[p + name + "_record_gate"][i].reshape(-1).view(ESTIMATE_DTYPE)
            for f in ("lower_bound", "short_ip", "alpha", "ip_hat", "error_bound", "valid"):
                assert e[f].tobytes() == g[f].tobytes(), (name, f)
            qv = o.qvec(oq).astype(np.float64)
            sip = ((codes >= 8) * qv).sum(1) - 0.5 * qv.sum()
            assert np.allclose(e["short_ip"], sip, rtol=1e-4, atol=1e-4 * np.abs(qv).sum())
            alpha = np.minimum(1.0, 4.0 * np.sqrt(hdr[:, 0].astype(np.float64)) /
                               (hdr[:, 2].astype(np.float64) * np.sqrt(o.cd)))
            assert np.allclose(e["alpha"], alpha, rtol=1e-6)
            o.free_query(oq)
    for i in range(q.shape[0]):
        oq = o.prepare(q[i], 0)
        d, dots, ssum = o.b1(oq, comp)
        sign = np.where(o.qvec(oq) > 0, 1, -1).astype(np.int64)
        assert np.array_equal(dots, codes @ sign) and ssum == int(sign.sum())
        assert np.array_equal(d.view(np.uint32), golden[p + "b1_dist"][i].view(np.uint32))
        o.free_query(oq)


def test_train_centroids_matches_reference_fixture(golden):
    """rabitq_train_centroids: bit for bit against the reference (random rows
    and the near tie that fixes the contraction of the distance), and against
    what the algorithm fixes: no iterations return the sampled rows (mt19937
    draws mod count), one iteration returns the means of the assignment."""
    o = Oracle(64, 100)
    x = golden["km_x"]
    for k, sc, it in ((1, 2000, 3), (16, 1000, 5), (256, 1500, 2)):
        assert np.array_equal(o.train_centroids(x, k, sc, 11, it), golden["km_k%d" % k])
    assert np.array_equal(o.train_centroids(golden["km_tie_x"], 2, 3, 33, 1), golden["km_tie"])
    n = x.shape[0]
    sample = np.array([o.lib.ro_mt32_nth(11, i + 1) % n for i in range(300)])
    assert np.array_equal(o.train_centroids(x, 7, 300, 11, 0), x[sample[np.arange(7) % 300]])
    c0 = x[sample[:7]].astype(np.float64)
    xs = x[sample].astype(np.float64)
    assign = np.argmin(((xs[:, None] - c0[None]) ** 2).sum(-1), 1)
    expect = np.stack([xs[assign == c].mean(0) if (assign == c).any() else c0[c] for c in range(7)])
    assert np.allclose(o.train_centroids(x, 7, 300, 11, 1), expect, rtol=1e-5, atol=1e-5)
    for bad in (dict(k=0), dict(k=257), dict(sc=0)):
        kw = dict(k=4, sc=10)
        kw.update(bad)
        assert o.train_centroids(x, kw["k"], kw["sc"], 1, 1) is None
