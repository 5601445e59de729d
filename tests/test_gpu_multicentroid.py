"""Spaces of K > 1 centroids (space_rabitq.h:2914-2960, 2564-2683,
3404-3416; rabitq_bridge.cpp:42-99) through the bridge and the batch API,
against the reference bridge compiled from source:

* records (with each vector's nearest centroid, assignCentroid) and the
  sidecar, byte for byte;
* per prepared query (one context per centroid, the INT8 dither seeded with
  the centroid id): the distance batch in every mode and its mixes, the
  record gate, the residual intervals, the fused full + residual batch and
  the 1-bit sidecar gate (which takes centroid 0's context, as the
  reference's does), bit for bit;
* symmetric distances of pairs within one centroid and across centroids
  (the cross-centroid terms), bit for bit;
* the batch API's encode and per-record-context scans equal the bridge;
* the device build and search refuse K > 1 with an error."""
import ctypes as C

import numpy as np
import pytest

from conftest import gaussian_mixture
from oracle_bind import Ref

pytestmark = pytest.mark.gpu


class Ours(Ref):
    """Ref's ctypes surface bound to libqgraph_hip.so."""

    def __init__(self, dim, seed, centroid_count=1):
        import oracle_bind
        from quantization_graph_amd import _lib
        saved = oracle_bind.REF_SO
        oracle_bind.REF_SO = _lib.LIB_PATH
        try:
            Ref.__init__(self, dim, seed, centroid_count)
        finally:
            oracle_bind.REF_SO = saved


def _case(dim, K, seed):
    x, q = gaussian_mixture(400, dim, 2 * K, 0.3, seed, nq=4)
    rng = np.random.default_rng(seed)
    cents = x[rng.choice(x.shape[0], K, replace=False)] + 0.05 * rng.normal(size=(K, dim)).astype(np.float32)
    return x, q, np.ascontiguousarray(cents, np.float32)


@pytest.mark.parametrize("dim,K", [(960, 4), (960, 16), (1536, 4)])
def test_multicentroid_bridge_matches_reference(gpu, dim, K):
    if not Ref.available():
        pytest.skip("oracle/_ref not built")
    print("case: K %d, D %d" % (K, dim), flush=True)
    x, q, cents = _case(dim, K, 40 + K)
    ref, ours = Ref(dim, 100, K), Ours(dim, 100, K)
    for s in (ref, ours):
        assert s.set_centroids(cents, K)
        assert not s.lib.rabitq_space_set_center(s.h, C.c_void_p(cents.ctypes.data))  # K > 1: setGlobalCenter fails
    full = ref.encode_full(x)
    assert np.array_equal(ours.encode_full(x), full)
    comp, res = ref.split(full)
    cid = comp[:, -1]
    assert len(np.unique(cid)) > 1 and cid.max() < K
    msb, fac = ref.sidecar(comp)
    msb_o, fac_o = ours.sidecar(comp)
    assert np.array_equal(msb, msb_o)
    for f in ("norm_sqr", "data_norm", "cross_scale", "error_cross_scale", "valid"):  # the pad bytes are unset
        assert np.array_equal(fac[f], fac_o[f])
    rng = np.random.default_rng(1)
    ids = rng.permutation(x.shape[0]).astype(np.uint32)[:120]
    for codec in (0, 3):
        ref.set_codec(codec)
        ours.set_codec(codec)
        for i in range(q.shape[0]):
            pr, po = ref.prepare(q[i]), ours.prepare(q[i])
            for modes in (np.zeros(ids.size, np.uint8), rng.integers(0, 4, ids.size).astype(np.uint8),
                          np.full(ids.size, 1, np.uint8), np.full(ids.size, 2, np.uint8)):
                if codec == 0:
                    modes = np.where(modes == 2, 1, modes).astype(np.uint8)
                sip = rng.normal(size=ids.size).astype(np.float32)
                assert ref.distance_batch(pr, comp, ids, modes, sip).tobytes() == \
                    ours.distance_batch(po, comp, ids, modes, sip).tobytes(), (codec, i, modes[:4])
            d0 = ref.distance_batch(pr, comp, ids, np.zeros(ids.size, np.uint8))
            assert ref.residual_batch(pr, comp, res, ids, d0).tobytes() == \
                ours.residual_batch(po, comp, res, ids, d0).tobytes()
            assert ref.full_residual_batch(pr, comp, res, ids).tobytes() == \
                ours.full_residual_batch(po, comp, res, ids).tobytes()
            er, eo = ref.paper_estimate_batch(pr, comp, ids, 1.9), ours.paper_estimate_batch(po, comp, ids, 1.9)
            for f in ("lower_bound", "short_ip", "alpha", "ip_hat", "error_bound", "valid"):
                assert er[f].tobytes() == eo[f].tobytes(), f
            gr, go = ref.gate_batch(pr, msb, fac, ids, 1.9), ours.gate_batch(po, msb, fac, ids, 1.9)
            for f in ("lower_bound", "short_ip", "ip_hat", "error_bound", "valid"):
                assert gr[f].tobytes() == go[f].tobytes(), f
    # symmetric distances within and across centroids
    pairs = [(a, b) for a in range(0, 60, 7) for b in range(1, 60, 5) if a != b]
    cross = 0
    for a, b in pairs:
        cross += int(cid[a] != cid[b])
        assert np.float32(ref.symmetric(comp[a], comp[b])).tobytes() == \
            np.float32(ours.symmetric(comp[a], comp[b])).tobytes(), (a, b, cid[a], cid[b])
    assert cross > 0


@pytest.mark.parametrize("K", [4, 16])
def test_multicentroid_batch_api(gpu, K):
    """sp.encode, per-record-context scans and symmetric pairs on the device
    index equal the bridge; the device build and search refuse K > 1."""
    import torch
    import quantization_graph_amd as qg
    if not Ref.available():
        pytest.skip("oracle/_ref not built")
    dim = 960
    x, q, cents = _case(dim, K, 70 + K)
    ref = Ref(dim, 100, K)
    assert ref.set_centroids(cents, K)
    full = ref.encode_full(x)
    comp_r, res_r = ref.split(full)
    sp = qg.RabitqSpace(dim, 100, centroid_count=K)
    sp.set_centroids(cents)
    ix = sp.encode(torch.from_numpy(x).cuda())
    comp, res = ix.to_records()
    assert np.array_equal(comp.cpu().numpy(), comp_r) and np.array_equal(res.cpu().numpy(), res_r)
    rng = np.random.default_rng(3)
    ids = np.stack([rng.permutation(x.shape[0])[:150] for _ in range(q.shape[0])]).astype(np.int32)
    for codec, name in ((0, "full"), (3, "int8")):
        ref.set_codec(codec)
        qv = sp.prepare(torch.from_numpy(q).cuda(), name)
        assert qv.rows == K * q.shape[0]
        asym = sp.scan_fixed(qg.SCAN_ASYM, ix, qv, ids)[0].cpu().numpy()
        for i in range(q.shape[0]):
            pr = ref.prepare(q[i])
            want = ref.distance_batch(pr, comp_r, ids[i].astype(np.uint32), np.zeros(ids.shape[1], np.uint8))
            assert asym[i].tobytes() == want.tobytes(), (codec, i)
    l = torch.arange(0, 40, dtype=torch.int32, device="cuda")
    r = torch.arange(40, 80, dtype=torch.int32, device="cuda")
    sym = sp.scan_pairs(qg.SCAN_SYMMETRIC, ix, None, l, r)[0].cpu().numpy()
    for j in range(40):
        assert np.float32(ref.symmetric(comp_r[j], comp_r[40 + j])).tobytes() == sym[j].tobytes(), j
    deg = torch.zeros(x.shape[0], dtype=torch.int32, device="cuda")
    with pytest.raises(qg._lib.QGError, match="one centroid"):
        sp.build_graph(ix, r=8, l_build=16, alpha=1.2, batch=64)
    with pytest.raises(qg._lib.QGError, match="one query context"):
        sp.search(ix, sp.prepare(torch.from_numpy(q).cuda(), "full"), deg, deg[:, None].expand(-1, 4).contiguous(),
                  k=5, l_search=16)


@pytest.mark.parametrize("K", [1, 4])
def test_single_record_residual_matches_reference(gpu, K):
    """rabitq_residual_distance_from_record scores one record in the
    per-record form (computeResidualDistanceIntervalFromRecord,
    space_rabitq.h:1723-1735), whose lane products the reference compiler
    fuses the other way from the batch loop: every record's interval equals
    the reference's, bit for bit, in spaces of one and of four centroids."""
    if not Ref.available():
        pytest.skip("oracle/_ref not built")
    dim = 960
    x, q, cents = _case(dim, max(K, 2), 90 + K)
    ref, ours = Ref(dim, 100, K), Ours(dim, 100, K)
    for s in (ref, ours):
        assert s.set_centroids(cents[:K], K)
    comp, res = ref.split(ref.encode_full(x))
    ref.set_codec(0)
    ours.set_codec(0)
    for i in range(2):
        pr, po = ref.prepare(q[i]), ours.prepare(q[i])
        d0 = ref.distance_batch(pr, comp, np.arange(x.shape[0], dtype=np.uint32), np.zeros(x.shape[0], np.uint8))
        for j in range(x.shape[0]):
            want = ref.residual_from_record(pr, comp[j], res[j], d0[j])
            got = ours.residual_from_record(po, comp[j], res[j], d0[j])
            assert np.array(got, np.float32).tobytes() == np.array(want, np.float32).tobytes(), (i, j)
