"""The host-memory bridge (rabitq_* in libqgraph_hip.so) against the
reference bridge compiled from source, called the way the reference's Rust
runner calls it (ours_diskann.rs:45-242): same arguments, same outputs."""
import ctypes as C

import numpy as np
import pytest

from conftest import gaussian_mixture
from oracle_bind import ESTIMATE_DTYPE, FACTOR_DTYPE, INTERVAL_DTYPE, Oracle, Ref

pytestmark = pytest.mark.gpu


class Ours(Ref):
    """The same ctypes surface, bound to libqgraph_hip.so."""

    def __init__(self, dim, seed):
        import oracle_bind
        from quantization_graph_amd import _lib
        saved = oracle_bind.REF_SO
        oracle_bind.REF_SO = _lib.LIB_PATH
        try:
            Ref.__init__(self, dim, seed)
        finally:
            oracle_bind.REF_SO = saved


def _pair(dim, n=96):
    x, q = gaussian_mixture(n, dim, 8, 0.3, 11, nq=3)
    c = x.mean(0).astype(np.float32)
    return x, q, c


@pytest.mark.parametrize("dim", [960, 1536])
def test_bridge_matches_reference(gpu, dim):
    if not Ref.available():
        pytest.skip("oracle/_ref not built")
    x, q, c = _pair(dim)
    ref, ours = Ref(dim, 100), Ours(dim, 100)
    for s in (ref, ours):
        s.set_center(c)
    assert (ours.full_bytes, ours.compact_bytes, ours.residual_bytes, ours.msb_bytes) == \
        (ref.full_bytes, ref.compact_bytes, ref.residual_bytes, ref.msb_bytes)
    full_r, full_o = ref.encode_full(x[:24]), ours.encode_full(x[:24])
    assert np.array_equal(full_r, full_o)
    comp, res = ref.split(ref.encode_full(x))
    comp_o, res_o = ours.split(ref.encode_full(x))
    assert np.array_equal(comp, comp_o) and np.array_equal(res, res_o)
    msb, fac = ref.sidecar(comp)
    msb_o, fac_o = ours.sidecar(comp)
    assert np.array_equal(msb, msb_o)
    for f in ("norm_sqr", "data_norm", "cross_scale", "error_cross_scale", "valid"):
        assert np.array_equal(fac[f], fac_o[f])
    rng = np.random.default_rng(0)
    ids = rng.permutation(x.shape[0]).astype(np.uint32)[:70]
    for codec in (0, 3):
        ref.set_codec(codec)
        ours.set_codec(codec)
        for i in range(q.shape[0]):
            pr = ref.prepare(q[i])
            er = ref.gate_batch(pr, msb, fac, ids, 1.9)
            d0r = ref.distance_batch(pr, comp, ids, np.zeros(ids.size, np.uint8))
            modes = rng.integers(0, 3, ids.size).astype(np.uint8)
            if codec == 0:
                modes[modes == 2] = 1
            dmr = ref.distance_batch(pr, comp, ids, modes, er["short_ip"])
            ivr = ref.residual_batch(pr, comp, res, ids[:40], d0r[:40])
            po = ours.prepare(q[i])
            eo = ours.gate_batch(po, msb, fac, ids, 1.9)
            d0o = ours.distance_batch(po, comp, ids, np.zeros(ids.size, np.uint8))
            dmo = ours.distance_batch(po, comp, ids, modes, er["short_ip"])
            ivo = ours.residual_batch(po, comp, res, ids[:40], d0r[:40])
            for f in ("lower_bound", "short_ip", "ip_hat", "error_bound", "valid"):
                assert er[f].tobytes() == eo[f].tobytes(), f
            assert d0r.tobytes() == d0o.tobytes()
            assert dmr.tobytes() == dmo.tobytes()
            for f in ("estimate", "lower_bound", "upper_bound"):
                assert ivr[f].tobytes() == ivo[f].tobytes()
    for a, b in [(0, 1), (7, 3), (5, 5)]:
        assert np.float32(ref.symmetric(comp[a], comp[b])).tobytes() == \
            np.float32(ours.symmetric(comp[a], comp[b])).tobytes()


@pytest.mark.parametrize("dim", [960, 1536])
def test_bridge_threads_share_one_space(gpu, dim):
    """Eight host threads call the bridge on one shared space at once, as
    the reference's encode workers and query threads do (ours_diskann.rs:
    2150-2199; the space is Send + Sync, :250-251): encode_full, and per
    prepared query (thread-local) the gate batch, the distance batch and the
    residual batch, with batch sizes that differ per thread so each thread's
    staging memory grows on its own. Every output equals the reference
    bridge's, byte for byte."""
    from concurrent.futures import ThreadPoolExecutor
    if not Ref.available():
        pytest.skip("oracle/_ref not built")
    x, q = gaussian_mixture(320, dim, 8, 0.3, 17, nq=16)
    c = x.mean(0).astype(np.float32)
    ref, ours = Ref(dim, 100), Ours(dim, 100)
    for s in (ref, ours):
        s.set_center(c)
    full = ref.encode_full(x)
    comp, res = ref.split(full)
    msb, fac = ref.sidecar(comp)
    rng = np.random.default_rng(3)
    perms = [rng.permutation(x.shape[0]).astype(np.uint32)[:40 + 30 * w] for w in range(8)]
    want = {}
    for codec in (0, 3):
        ref.set_codec(codec)
        for qi in range(q.shape[0]):
            ids = perms[qi % 8]
            pr = ref.prepare(q[qi])
            e = ref.gate_batch(pr, msb, fac, ids, 1.9)
            d = ref.distance_batch(pr, comp, ids, np.zeros(ids.size, np.uint8))
            iv = ref.residual_batch(pr, comp, res, ids, d)
            want[codec, qi] = (e.tobytes(), d.tobytes(), iv.tobytes())

    def worker(w):
        got = {}
        rows = np.arange(w, x.shape[0], 8)
        enc = ours.encode_full(x[rows])
        assert np.array_equal(enc, full[rows]), "encode_full, thread %d" % w
        for qi in range(w, q.shape[0], 8):
            ids = perms[qi % 8]
            po = ours.prepare(q[qi])
            e = ours.gate_batch(po, msb, fac, ids, 1.9)
            d = ours.distance_batch(po, comp, ids, np.zeros(ids.size, np.uint8))
            iv = ours.residual_batch(po, comp, res, ids, d)
            got[qi] = (e.tobytes(), d.tobytes(), iv.tobytes())
        return got

    for codec in (0, 3):
        ours.set_codec(codec)  # between sweeps, as the runner does (diskann_runner.rs:361)
        print("case: 8 threads on one space, D %d codec %d" % (dim, codec), flush=True)
        for rep in range(2):
            with ThreadPoolExecutor(8) as pool:
                results = list(pool.map(worker, range(8)))
            for got in results:
                for qi, outs in got.items():
                    assert outs == want[codec, qi], (codec, qi, rep)


def test_bridge_error_sentinels(gpu):
    ours = Ours(128, 100)
    L = ours.lib
    assert L.rabitq_space_new(128, 100, 0, 4, 16, True, True) is None      # K = 0
    assert L.rabitq_space_new(128, 100, 257, 4, 16, True, True) is None    # K > 256
    assert L.rabitq_space_new(128, 100, 1, 3, 16, True, True) is None      # bits not in {4, 8}
    assert not L.rabitq_encode_full(ours.h, None, None)
    L.rabitq_symmetric_distance.restype = C.c_float
    assert np.isinf(L.rabitq_symmetric_distance(ours.h, None, None))


@pytest.mark.parametrize("refine", [0, 1])
def test_bridge_build_matches_reference(gpu, refine):
    """rabitq_build_vamana_graph through both libraries on the same records:
    same degrees, adjacency lists, sidecar and distance_evaluations."""
    if not Ref.available():
        pytest.skip("oracle/_ref not built")
    n, dim = 700, 256
    x, _ = gaussian_mixture(n, dim, 8, 0.3, 13)
    c = x.mean(0).astype(np.float32)
    ref, ours = Ref(dim, 100), Ours(dim, 100)
    for s in (ref, ours):
        s.set_center(c)
    comp, _ = ref.split(ref.encode_full(x))
    d1, e1, m1, f1, v1 = ref.build_graph(comp, 24, 48, 1.2, beam=1, batch=128, refine=refine, cap=0)
    d2, e2, m2, f2, v2 = ours.build_graph(comp, 24, 48, 1.2, beam=1, batch=128, refine=refine, cap=0)
    assert np.array_equal(d1, d2)
    for i in range(n):
        assert np.array_equal(e1[i, :d1[i]], e2[i, :d2[i]])
    assert np.array_equal(m1, m2)
    for f in ("norm_sqr", "data_norm", "cross_scale", "error_cross_scale", "valid"):  # padding bytes excluded
        assert f1[f].tobytes() == f2[f].tobytes(), f
    assert v1 == v2


def test_bridge_build_rejects_what_the_reference_rejects(gpu):
    ours = Ours(128, 100)
    x, _ = gaussian_mixture(32, 128, 2, 0.3, 1)
    comp, _ = ours.split(ours.encode_full(x))
    with pytest.raises(AssertionError):
        ours.build_graph(comp, 0, 48, 1.2)        # R = 0
    with pytest.raises(AssertionError):
        ours.build_graph(comp, 16, 48, 0.5)       # alpha < 1
    with pytest.raises(AssertionError):
        ours.build_graph(comp, 16, 0, 1.2)        # L_build = 0


EST_FIELDS = ("lower_bound", "short_ip", "alpha", "ip_hat", "error_bound", "valid")


def _est_bytes(rows):
    a = np.array(rows, ESTIMATE_DTYPE) if isinstance(rows, list) else rows
    return {f: a[f].tobytes() for f in EST_FIELDS}


@pytest.mark.parametrize("dim", [960, 1536])
def test_bridge_single_record_entry_points(gpu, golden, dim):
    """The single-record and remaining batch symbols (rabitq_bridge.cpp:380-700,
    1044-1084) against the reference's outputs: the golden fixture for the
    record gate and B1 distances, and the live reference bridge for the rest."""
    p = "d%d_" % dim
    ours = Ours(dim, 100)
    ours.set_center(golden[p + "center"])
    comp, res = golden[p + "compact"], golden[p + "residual"]
    msb = golden[p + "msb"]
    fac = golden[p + "factors"].reshape(-1).view(FACTOR_DTYPE).copy()
    q = golden[p + "q"]
    ids = np.arange(comp.shape[0], dtype=np.uint32)[::-1].copy()
    print("case: dim", dim, "records", comp.shape[0], flush=True)
    for codec, name in ((0, "full"), (3, "int8")):
        ours.set_codec(codec)
        for i in range(q.shape[0]):
            pq = ours.prepare(q[i])
            g = golden[p + name + "_record_gate"][i].reshape(-1).view(ESTIMATE_DTYPE)
            eb = ours.paper_estimate_batch(pq, comp, ids, 1.9)
            assert _est_bytes(eb) == _est_bytes(g[ids])
            es = [ours.paper_estimate(pq, comp[j], 1.9) for j in range(0, comp.shape[0], 7)]
            assert _est_bytes(es) == _est_bytes(g[::7])
            gate = golden[p + name + "_gate"][i].reshape(-1).view(ESTIMATE_DTYPE)
            cb = ours.paper_estimate_codec_batch(pq, comp, msb, fac, ids, 1.9)
            assert _est_bytes(cb) == _est_bytes(gate[ids])
            one = [ours.paper_estimate_sidecar(pq, msb[j], fac, j, 1.9) for j in (0, 5, 47)]
            two = [ours.paper_estimate_codec(pq, comp[j], msb[j], fac, j, 1.9) for j in (0, 5, 47)]
            assert _est_bytes(one) == _est_bytes(gate[[0, 5, 47]]) == _est_bytes(two)
            rem = golden[p + name + "_rem"][i]
            d = np.array([ours.with_paper(pq, comp[j], gate["short_ip"][j], quantized=codec == 3)
                          for j in (1, 9, 30)], np.float32)
            assert d.tobytes() == rem[[1, 9, 30]].tobytes()
            dist = golden[p + name + "_dist"][i]
            iv = golden[p + name + "_residual"][i].reshape(-1).view(INTERVAL_DTYPE)
            fr = ours.full_residual_batch(pq, comp, res, ids)
            for f in ("estimate", "lower_bound", "upper_bound"):
                assert fr[f].tobytes() == iv[f][ids].tobytes(), f
            for j in (2, 40):
                got = ours.residual_from_record(pq, comp[j], res[j], dist[j])
                assert np.array(got, np.float32).tobytes() == \
                    np.array([iv["estimate"][j], iv["lower_bound"][j], iv["upper_bound"][j]], np.float32).tobytes()
                assert ours.residual_single(pq, comp[j], dist[j]) == (dist[j], dist[j], dist[j])
    ours.set_codec(1)
    for i in range(q.shape[0]):
        pq = ours.prepare(q[i])
        b = ours.b1_batch(pq, comp, ids)
        assert b.tobytes() == golden[p + "b1_dist"][i][ids].tobytes()
        assert np.float32(ours.b1(pq, comp[3])).tobytes() == golden[p + "b1_dist"][i][3].tobytes()
    for codec in (4, 5, 6):  # B1Main, FullStaged, FullRerankScalar gate with the FP32 query
        ours.set_codec(codec)
        pq = ours.prepare(q[1])
        gate = golden[p + "full_gate"][1].reshape(-1).view(ESTIMATE_DTYPE)
        assert _est_bytes(ours.gate_batch(pq, msb, fac, ids, 1.9)) == _est_bytes(gate[ids])
    ours.set_codec(0)
    pq = ours.prepare(q[0])
    assert np.isinf(ours.b1_batch(pq, comp, ids[:4])).all()  # no sign code under the Full codec


def test_bridge_train_centroids(gpu, golden):
    """rabitq_train_centroids on the device: bit for bit against the
    reference's outputs (random rows, the contraction-sensitive near tie) and
    against the live reference bridge on a larger draw; the reference's
    argument checks."""
    ours = Ours(64, 100)
    x = golden["km_x"]
    for k, sc, it in ((1, 2000, 3), (16, 1000, 5), (256, 1500, 2)):
        print("case: kmeans k", k, "samples", sc, "iters", it, flush=True)
        assert np.array_equal(ours.train_centroids(x, k, sc, 11, it), golden["km_k%d" % k])
    assert np.array_equal(ours.train_centroids(golden["km_tie_x"], 2, 3, 33, 1), golden["km_tie"])
    assert ours.train_centroids(x, 0, 10, 1, 1) is None
    assert ours.train_centroids(x, 257, 10, 1, 1) is None
    assert ours.train_centroids(x, 4, 0, 1, 1) is None
    assert np.array_equal(ours.train_centroids(x[:3], 5, 3, 7, 0), ours.train_centroids(x[:3], 5, 3, 7, 0))
    if Ref.available():
        ref = Ref(64, 100)
        big = np.random.default_rng(9).standard_normal((20000, 300)).astype(np.float32)
        print("case: kmeans 20000 x 300 k 64 iters 4 vs reference", flush=True)
        assert np.array_equal(ours.train_centroids(big, 64, 12000, 5, 4), ref.train_centroids(big, 64, 12000, 5, 4))


def test_bridge_set_centroids(gpu):
    ours = Ours(64, 100)
    c = np.random.default_rng(0).standard_normal((2, 64)).astype(np.float32)
    assert ours.set_centroids(c[0], 1)
    assert not ours.set_centroids(c, 2)  # the space was made with one centroid
    assert not ours.set_centroids(c, 0)
