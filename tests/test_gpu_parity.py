"""HIP kernels vs the CPU oracle (and the reference compiled from source,
where it travelled with the tree): bit-exact for every record byte, integer
dot, float estimate and id.

Sizes are small enough for the oracle to finish in seconds; the full-size
properties run in tests/test_gpu_scale.py.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import gaussian_mixture
from oracle_bind import Oracle, Ref

pytestmark = pytest.mark.gpu

DIMS = [960, 1536, 100, 48]


def _setup(gpu, dim, n=384, nq=6, seed=0):
    import quantization_graph_amd as qg
    x, q = gaussian_mixture(n, dim, 16, 0.3, seed, nq=nq)
    c = x.mean(0).astype(np.float32)
    sp = qg.RabitqSpace(dim, 100)
    sp.set_center(c)
    orc = Oracle(dim, 100)
    orc.set_center(c)
    return qg, sp, orc, x, q, c


@pytest.mark.parametrize("dim", DIMS)
def test_encode_records_bit_exact(gpu, dim):
    qg, sp, orc, x, q, c = _setup(gpu, dim)
    ix = sp.encode(gpu.from_numpy(x).cuda())
    comp, res = ix.to_records()
    comp_o, res_o = orc.split(orc.encode_full(x))
    comp, res = comp.cpu().numpy(), res.cpu().numpy()
    assert np.array_equal(comp, comp_o)
    assert np.array_equal(res, res_o)
    msb_o, fac_o = orc.sidecar(comp_o)
    assert np.array_equal(ix.msb.cpu().numpy(), msb_o)
    f = ix.factors.cpu().numpy()
    for k, name in enumerate(("norm_sqr", "data_norm", "cross_scale", "error_cross_scale")):
        assert np.array_equal(f[:, k].view(np.uint32), fac_o[name].view(np.uint32)), name
    assert np.array_equal(ix.fvalid.cpu().numpy(), fac_o["valid"])


def test_encode_matches_reference_build(gpu):
    if not Ref.available():
        pytest.skip("oracle/_ref not built")
    qg, sp, orc, x, q, c = _setup(gpu, 960, n=256)
    ref = Ref(960, 100)
    ref.set_center(c)
    comp_r, res_r = ref.split(ref.encode_full(x))
    comp, res = sp.encode(gpu.from_numpy(x).cuda()).to_records()
    assert np.array_equal(comp.cpu().numpy(), comp_r)
    assert np.array_equal(res.cpu().numpy(), res_r)


@pytest.mark.parametrize("dim", DIMS)
@pytest.mark.parametrize("codec", [0, 3])
def test_query_prepare(gpu, dim, codec):
    qg, sp, orc, x, q, c = _setup(gpu, dim)
    qv = sp.prepare(gpu.from_numpy(q).cuda(), codec)
    s = qv.scalars()
    qq = qv.q.cpu().numpy()
    x8 = qv.x8.cpu().numpy()
    for i in range(q.shape[0]):
        oq = orc.prepare(q[i], codec)
        assert np.array_equal(qq[i].view(np.uint32), orc.qvec(oq).view(np.uint32))
        for f in ("query_norm", "query_norm_sqr", "half_sum", "sum", "positive_sum", "int8_scale"):
            assert np.float32(s[f][i]).tobytes() == np.float32(getattr(oq, f)).tobytes(), f
        if codec == 3:
            assert np.array_equal(x8[i], orc.qx8(oq))
            assert s["int8_code_sum"][i] == oq.int8_code_sum
        orc.free_query(oq)


@pytest.mark.parametrize("dim", DIMS)
def test_estimators_fixed_lists(gpu, dim):
    qg, sp, orc, x, q, c = _setup(gpu, dim)
    n, nq, per = x.shape[0], q.shape[0], 200
    rng = np.random.default_rng(1)
    ids = np.stack([rng.permutation(n)[:per] for _ in range(nq)]).astype(np.int32)
    ix = sp.encode(gpu.from_numpy(x).cuda())
    comp_o, res_o = orc.split(orc.encode_full(x))
    qv = sp.prepare(gpu.from_numpy(q).cuda(), "int8")
    est, low, _, _, dot = sp.scan_fixed(qg.SCAN_INT8_FULL, ix, qv, ids, eps0=1.9, want=("f", "f2", "dot"))
    asym = sp.scan_fixed(qg.SCAN_ASYM, ix, qv, ids)[0]
    _, _, _, iv, _ = sp.scan_fixed(qg.SCAN_RESIDUAL, ix, qv, ids, long_in=asym, want=("iv",))
    est, low, dot, asym = (t.cpu().numpy() for t in (est, low, dot, asym))
    iv = qg.intervals_np(iv).reshape(nq, per)
    for i in range(nq):
        oq = orc.prepare(q[i], 3)
        e_o, l_o, d_o = orc.int8_full(oq, comp_o[ids[i]])
        assert np.array_equal(dot[i], d_o)
        assert np.array_equal(est[i].view(np.uint32), e_o.view(np.uint32))
        assert np.array_equal(low[i].view(np.uint32), l_o.view(np.uint32))
        a_o = orc.asym(oq, comp_o[ids[i]])
        assert np.array_equal(asym[i].view(np.uint32), a_o.view(np.uint32))
        iv_o = orc.residual(oq, res_o[ids[i]], a_o)
        for f in ("estimate", "lower_bound", "upper_bound"):
            assert np.array_equal(iv[i][f].view(np.uint32), iv_o[f].view(np.uint32)), f
        # FP64 oracle: the float estimate is within 1e-5 of the exact formula
        a64 = orc.asym_f64(oq, comp_o[ids[i]])
        scale = np.abs(a64) + float(oq.query_norm_sqr) + 1e-6
        assert np.max(np.abs(asym[i] - a64) / scale) < 1e-5
        orc.free_query(oq)


@pytest.mark.parametrize("dim", [960, 1536, 100, 48])
@pytest.mark.parametrize("codec", [0, 3])
def test_gates_and_remaining_pairs(gpu, dim, codec):
    qg, sp, orc, x, q, c = _setup(gpu, dim)
    n, nq = x.shape[0], q.shape[0]
    rng = np.random.default_rng(2)
    per = 64
    qidx = np.repeat(np.arange(nq), per).astype(np.int32)
    ids = np.concatenate([rng.permutation(n)[:per] for _ in range(nq)]).astype(np.int32)
    ix = sp.encode(gpu.from_numpy(x).cuda())
    comp_o, _ = orc.split(orc.encode_full(x))
    msb_o, fac_o = orc.sidecar(comp_o)
    qv = sp.prepare(gpu.from_numpy(q).cuda(), codec)
    gk = qg.SCAN_GATE_INT8 if codec == 3 else qg.SCAN_GATE_FP32
    est = qg.estimates_np(sp.scan_pairs(gk, ix, qv, qidx, ids, eps0=1.9, want=("est",))[2])
    rk = qg.SCAN_REM_INT8 if codec == 3 else qg.SCAN_REM_FP32
    rem = sp.scan_pairs(rk, ix, qv, qidx, ids, short_ips=est["short_ip"].copy())[0].cpu().numpy()
    for i in range(nq):
        sl = slice(i * per, (i + 1) * per)
        oq = orc.prepare(q[i], codec)
        e_o = orc.gate(oq, msb_o[ids[sl]], fac_o[ids[sl]], codec, 1.9)
        for f in ("lower_bound", "short_ip", "ip_hat", "error_bound", "valid"):
            assert np.array_equal(est[sl][f], e_o[f]), f
        if codec == 3:
            r_o, _ = orc.rem_int8(oq, comp_o[ids[sl]], e_o["short_ip"])
        else:
            r_o = orc.rem_fp32(oq, comp_o[ids[sl]], e_o["short_ip"])
        assert np.array_equal(rem[sl].view(np.uint32), r_o.view(np.uint32))
        orc.free_query(oq)


@pytest.mark.parametrize("per,k", [(1000, 100), (1024, 10), (37, 100), (64, 64), (200, 1), (3000, 100)])
@pytest.mark.parametrize("want_pos", [False, True])
def test_topk_order_and_ties(gpu, per, k, want_pos):
    """Both top-k kernels (the register sort of one wavefront for lists up to
    1024 without positions, the workgroup sort otherwise) return the first k
    entries by (value, id), values bit for bit (signed zeros included)."""
    import quantization_graph_amd as qg
    rng = np.random.default_rng(3 + per + k)
    nq = 9
    vals = rng.integers(-20, 30, size=(nq, per)).astype(np.float32)  # many exact ties
    vals[0, :10] = -0.0
    vals[0, 10:20] = 0.0
    ids = np.stack([rng.permutation(10 * per)[:per] for _ in range(nq)]).astype(np.int32)
    oid, ov, opos = qg.topk(gpu.from_numpy(vals).cuda(), gpu.from_numpy(ids).cuda(), k, want_pos=want_pos)
    oid, ov = oid.cpu().numpy(), ov.cpu().numpy()
    kk = min(k, per)
    for i in range(nq):
        order = sorted(range(per), key=lambda j: (float(vals[i, j]), int(ids[i, j])))[:kk]
        assert np.array_equal(oid[i, :kk], ids[i, order])
        assert np.array_equal(ov[i, :kk].view(np.uint32), vals[i, order].view(np.uint32))
        assert (oid[i, kk:] == -1).all() and np.isinf(ov[i, kk:]).all()
        if want_pos:
            assert np.array_equal(opos.cpu().numpy()[i, :kk], order)


def test_records_roundtrip(gpu):
    qg, sp, orc, x, q, c = _setup(gpu, 960, n=200)
    comp_o, res_o = orc.split(orc.encode_full(x))
    ix = qg.DeviceIndex.from_records(sp, comp_o, res_o)
    comp, res = ix.to_records()
    assert np.array_equal(comp.cpu().numpy(), comp_o)
    assert np.array_equal(res.cpu().numpy(), res_o)
    msb_o, fac_o = orc.sidecar(comp_o)
    assert np.array_equal(ix.msb.cpu().numpy(), msb_o)
    assert np.array_equal(ix.fvalid.cpu().numpy(), fac_o["valid"])
