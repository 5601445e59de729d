"""Code dimensions at both ends: 4096 (dimension 2049..4096) for encode,
query preparation, every estimator (pairs and fixed lists) and the bridge,
bit-exact against the oracle and the reference bridge compiled from source;
the graph build and search, the learned rerank and the sign scan refuse it
with an error. Code dimension 64 (dimension <= 64) runs the batched search,
against the oracle's restatement of the reference search loop."""
import numpy as np
import pytest

from conftest import gaussian_mixture
from oracle_bind import Oracle, Ref
from test_gpu_bridge import Ours

pytestmark = pytest.mark.gpu

DIM = 3072  # code dimension 4096


@pytest.fixture(scope="module")
def big(gpu):
    import quantization_graph_amd as qg
    x, q = gaussian_mixture(192, DIM, 8, 0.3, 31, nq=4)
    c = x.mean(0).astype(np.float32)
    sp = qg.RabitqSpace(DIM, 100)
    assert sp.code_dim == 4096
    sp.set_center(c)
    orc = Oracle(DIM, 100)
    orc.set_center(c)
    ix = sp.encode(gpu.from_numpy(x).cuda())
    comp, res = ix.to_records()
    return dict(qg=qg, sp=sp, orc=orc, ix=ix, comp=comp.cpu().numpy(), res=res.cpu().numpy(), x=x, q=q, c=c)


def test_code_dim_4096_encode(gpu, big):
    comp_o, res_o = big["orc"].split(big["orc"].encode_full(big["x"]))
    assert np.array_equal(big["comp"], comp_o)
    assert np.array_equal(big["res"], res_o)


@pytest.mark.parametrize("codec", [0, 3])
def test_code_dim_4096_estimators(gpu, big, codec):
    qg, sp, orc, ix, comp, res, q = (big[k] for k in ("qg", "sp", "orc", "ix", "comp", "res", "q"))
    n, nq = comp.shape[0], q.shape[0]
    qv = sp.prepare(gpu.from_numpy(q).cuda(), codec)
    ids = np.stack([np.random.default_rng(i).permutation(n)[:96] for i in range(nq)]).astype(np.int32)
    asym, _, _, _, _ = sp.scan_fixed(qg.SCAN_ASYM, ix, qv, ids)
    _, _, _, iv, _ = sp.scan_fixed(qg.SCAN_RESIDUAL, ix, qv, ids, long_in=asym, want=("iv",))
    asym, iv = asym.cpu().numpy(), qg.intervals_np(iv).reshape(nq, -1)
    if codec == 3:
        est, low, _, _, dot = sp.scan_fixed(qg.SCAN_INT8_FULL, ix, qv, ids, want=("f", "f2", "dot"))
        est, dot = est.cpu().numpy(), dot.cpu().numpy()
    msb_o, fac_o = orc.sidecar(comp)
    qidx = np.repeat(np.arange(nq), n).astype(np.int32)
    all_ids = np.tile(np.arange(n), nq).astype(np.int32)
    gate_k = qg.SCAN_GATE_INT8 if codec == 3 else qg.SCAN_GATE_FP32
    _, _, g, _, _ = sp.scan_pairs(gate_k, ix, qv, qidx, all_ids, want=("est",))
    g = qg.estimates_np(g).reshape(nq, n)
    for i in range(nq):
        oq = orc.prepare(q[i], codec)
        sel = comp[ids[i]]
        a_o = orc.asym(oq, sel)
        assert asym[i].tobytes() == a_o.tobytes()
        r_o = orc.residual(oq, res[ids[i]], a_o)
        assert iv[i]["estimate"].tobytes() == r_o["estimate"].tobytes()
        if codec == 3:
            e_o, _, d_o = orc.int8_full(oq, sel)
            assert np.array_equal(dot[i], d_o) and est[i].tobytes() == e_o.tobytes()
        want = orc.gate(oq, msb_o, fac_o, codec, 1.9)
        for f in ("lower_bound", "short_ip", "ip_hat", "error_bound", "valid"):
            assert g[i][f].tobytes() == want[f].tobytes(), f
        rem_k = qg.SCAN_REM_INT8 if codec == 3 else qg.SCAN_REM_FP32
        d, _, _, _, _ = sp.scan_pairs(rem_k, ix, qv, np.full(n, i, np.int32), np.arange(n, dtype=np.int32),
                                      short_ips=want["short_ip"].copy())
        d_o = orc.rem_int8(oq, comp, want["short_ip"])[0] if codec == 3 else orc.rem_fp32(oq, comp, want["short_ip"])
        assert d.cpu().numpy().tobytes() == d_o.tobytes()
        orc.free_query(oq)


def test_code_dim_4096_bridge(gpu, big):
    if not Ref.available():
        pytest.skip("oracle/_ref not built")
    x, q, c = big["x"][:64], big["q"], big["c"]
    ref, ours = Ref(DIM, 100), Ours(DIM, 100)
    for s in (ref, ours):
        s.set_center(c)
    full_r = ref.encode_full(x)
    assert np.array_equal(full_r, ours.encode_full(x))
    comp, res = ref.split(full_r)
    msb, fac = ref.sidecar(comp)
    ids = np.arange(64, dtype=np.uint32)
    for codec in (0, 3):
        ref.set_codec(codec)
        ours.set_codec(codec)
        pr, po = ref.prepare(q[0]), ours.prepare(q[0])
        modes = np.zeros(64, np.uint8)
        d_r, d_o = ref.distance_batch(pr, comp, ids, modes), ours.distance_batch(po, comp, ids, modes)
        assert d_r.tobytes() == d_o.tobytes()
        g_r, g_o = ref.gate_batch(pr, msb, fac, ids, 1.9), ours.gate_batch(po, msb, fac, ids, 1.9)
        assert g_r.tobytes() == g_o.tobytes()
        assert ref.residual_batch(pr, comp, res, ids, d_r).tobytes() == ours.residual_batch(po, comp, res, ids, d_r).tobytes()
        assert ref.symmetric(comp[0], comp[1]) == ours.symmetric(comp[0], comp[1])


def test_code_dim_4096_graph_paths_refuse(gpu, big):
    qg, sp, ix, q = big["qg"], big["sp"], big["ix"], big["q"]
    n = big["comp"].shape[0]
    qv = sp.prepare(gpu.from_numpy(q).cuda(), "full")
    deg = np.full(n, 4, np.int32)
    edges = np.stack([(np.arange(n) + i) % n for i in range(1, 5)], 1).astype(np.int32)
    with pytest.raises(qg._lib.QGError, match="code dimensions up to 2048"):
        sp.search(ix, qv, deg, edges, k=3, l_search=16)
    with pytest.raises(qg._lib.QGError, match="code dimensions up to 2048"):
        sp.build_graph(ix, r=8, l_build=16)


def test_code_dim_64_search_matches_oracle(gpu):
    """Dimension 48: the gate's last 64 dims take half a 16-byte MSB step,
    the remaining-bit dot its last 8 dwords."""
    import test_gpu_search as ts
    if not Ref.available():
        pytest.skip("oracle/_ref not built")
    dim, n = 48, 1500
    x, q = gaussian_mixture(n, dim, 8, 0.3, 41, nq=8)
    c = x.mean(0).astype(np.float32)
    orc = Oracle(dim, 100)
    orc.set_center(c)
    comp, res = orc.split(orc.encode_full(x))
    msb, fac = orc.sidecar(comp)
    ref = Ref(dim, 100)
    ref.set_center(c)
    deg, edges, _, _, _ = ref.build_graph(comp, 16, 32, 1.2, beam=1, batch=256, refine=1, cap=64)
    g = dict(dim=dim, x=x, q=q, c=c, orc=orc, comp=comp, res=res, msb=msb, fac=fac, deg=deg, edges=edges)
    for codec in (0, 3):
        for L in (16, 64):
            ts._compare_with_oracle(gpu, g, deg, edges, codec, L)
