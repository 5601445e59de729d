"""The precision mode (QR rotation) on the GPU.

* The FP64 QR of the Gaussian on the GPU against LAPACK's (numpy.linalg.qr
  with the columns' signs fixed so that diag(R) > 0): the same Q to 1e-10,
  orthogonal to FP64 rounding.
* The BF16-MFMA rotation GEMM against the FP64 product Q (x - c), for both
  part counts, dimensions 100 .. 1536 (code dimension 128 .. 2048), row
  counts that leave partial tiles, and a dimension that is not a multiple of
  4 (the scalar load path): the per-row relative error is at the BF16 parts'
  resolution.
* Encode: every record byte equals the oracle's precision-mode codec
  (ro_encode_full_rotated) applied to the GEMM's own output, and |x - c|
  equals the FP64 norm to FP32 rounding.
* Query preparation: q, the INT8 codes and every scalar equal the oracle's
  (ro_prepare_query_rotated) applied to the GEMM's rotated query and
  center.
* The estimators on precision-mode records: INT8 dots and estimates bit
  exact against the oracle, and the FP32 4-bit estimate within 1e-5
  relative (plus an absolute term at FP32 resolution of the terms it
  cancels) of the same estimator evaluated in FP64 with the FP64 rotation.
"""
import numpy as np
import pytest

from conftest import gaussian_mixture
from oracle_bind import Oracle

pytestmark = pytest.mark.gpu


def _qr_ref(cd, seed):
    g = np.random.default_rng(seed).standard_normal((cd, cd))
    q, r = np.linalg.qr(g)
    return g, q * np.sign(np.diag(r))[None, :]


@pytest.mark.parametrize("dim", [960, 1536])
def test_qr_rotation_matches_lapack(gpu, dim):
    import quantization_graph_amd as qg
    sp = qg.RabitqSpace(dim, 100)
    print("case: FP64 QR on the GPU, code dimension %d" % sp.code_dim, flush=True)
    g, q_ref = _qr_ref(sp.code_dim, 7)
    sp.set_qr_rotation(gaussian=g)
    q = sp.rotation_matrix()
    assert np.abs(q - q_ref).max() < 1e-10
    assert np.abs(q.T @ q - np.eye(sp.code_dim)).max() < 1e-13
    # R = Q^T G is upper triangular with a positive diagonal
    r = q.T @ g
    assert np.abs(np.tril(r, -1)).max() < 1e-9 * np.abs(r).max()
    assert (np.diag(r) > 0).all()


@pytest.mark.parametrize("parts,tol", [(2, 4e-5), (3, 1e-6)])
@pytest.mark.parametrize("dim,n", [(960, 3001), (1536, 700), (100, 129), (130, 257), (200, 64)])
def test_rotation_gemm_vs_fp64(gpu, dim, n, parts, tol):
    import quantization_graph_amd as qg
    x, _ = gaussian_mixture(n, dim, 6, 0.3, 5 + dim)
    c = x.mean(0).astype(np.float32)
    sp = qg.RabitqSpace(dim, 100)
    sp.set_center(c)
    sp.set_qr_rotation(seed=11, bf16_parts=parts)
    print("case: rotation GEMM, %d x %d, %d BF16 parts" % (n, dim, parts), flush=True)
    q = sp.rotation_matrix()
    y = sp.rotate(gpu.from_numpy(x).cuda(), subtract_center=True).cpu().numpy().astype(np.float64)
    want = (x.astype(np.float64) - c.astype(np.float64)) @ q[:, :dim].T
    rel = np.linalg.norm(y - want, axis=1) / np.linalg.norm(want, axis=1)
    print("max per-row relative error %.3g" % rel.max())
    assert rel.max() < tol
    # without the center, and one row at a time (the rotated center's path)
    y0 = sp.rotate(gpu.from_numpy(x[:3]).cuda()).cpu().numpy()
    y1 = sp.rotate(gpu.from_numpy(x[1:2]).cuda()).cpu().numpy()
    assert np.array_equal(y0[1].view(np.uint32), y1[0].view(np.uint32))
    w0 = x[:3].astype(np.float64) @ q[:, :dim].T
    assert (np.linalg.norm(y0 - w0, axis=1) / np.linalg.norm(w0, axis=1)).max() < tol


def _qr_space(qg, dim, c, parts=3):
    sp = qg.RabitqSpace(dim, 100)
    sp.set_center(c)
    sp.set_qr_rotation(seed=3, bf16_parts=parts)
    return sp


@pytest.mark.parametrize("dim", [960, 1536, 200])
@pytest.mark.parametrize("parts", [2, 3])
def test_precision_encode_matches_oracle(gpu, dim, parts):
    import quantization_graph_amd as qg
    x, _ = gaussian_mixture(300, dim, 6, 0.3, 9 + dim)
    x[7] = 0.0
    c = x.mean(0).astype(np.float32)
    x[11] = c  # a vector at the center: |x - c| = 0
    sp = _qr_space(qg, dim, c, parts)
    print("case: precision encode, D %d, %d parts" % (dim, parts), flush=True)
    xd = gpu.from_numpy(x).cuda()
    comp, res = sp.encode(xd).to_records()
    y = sp.rotate(xd, subtract_center=True).cpu().numpy()
    orc = Oracle(dim, 100)
    orc.set_center(c)
    comp_o, res_o = orc.split(orc.encode_full_rotated(x, y))
    comp, res = comp.cpu().numpy(), res.cpu().numpy()
    assert np.array_equal(comp, comp_o)
    assert np.array_equal(res, res_o)
    # |x - c| (compensated FP32) against FP64: within FP32 rounding
    rn = np.sqrt(comp[:, :4].copy().view(np.float32)[:, 0].astype(np.float64))
    rn64 = np.linalg.norm(x.astype(np.float64) - c.astype(np.float32).astype(np.float64), axis=1)
    # norm_sqr = rn * rn rounds once more: compare rn itself to 2 ulp
    assert np.all(np.abs(rn - rn64) <= 2.5 * np.spacing(rn64.astype(np.float32)) + 1e-30)


@pytest.mark.parametrize("dim", [960, 1536])
@pytest.mark.parametrize("codec", ["full", "int8"])
def test_precision_queries_match_oracle(gpu, dim, codec):
    import quantization_graph_amd as qg
    x, q = gaussian_mixture(400, dim, 6, 0.3, 13 + dim, nq=9)
    c = x.mean(0).astype(np.float32)
    sp = _qr_space(qg, dim, c)
    print("case: precision query preparation, D %d, %s" % (dim, codec), flush=True)
    qv = sp.prepare(gpu.from_numpy(q).cuda(), codec)
    rq = sp.rotate(gpu.from_numpy(q).cuda()).cpu().numpy()
    rc = sp.rotate(gpu.from_numpy(c[None, :]).cuda()).cpu().numpy()[0]
    orc = Oracle(dim, 100)
    orc.set_center(c)
    s = qv.scalars()
    qq, x8 = qv.q.cpu().numpy(), qv.x8.cpu().numpy()
    cid = 3 if codec == "int8" else 0
    for i in range(q.shape[0]):
        oq = orc.prepare_rotated(q[i], rq[i], rc, cid)
        assert np.array_equal(qq[i].view(np.uint32), orc.qvec(oq).view(np.uint32))
        for f in ("query_norm", "query_norm_sqr", "half_sum", "sum", "positive_sum", "int8_scale"):
            assert np.float32(s[f][i]).tobytes() == np.float32(getattr(oq, f)).tobytes(), f
        if cid == 3:
            assert np.array_equal(x8[i], orc.qx8(oq))
            assert s["int8_code_sum"][i] == oq.int8_code_sum
        orc.free_query(oq)


def _decode_long(comp, cd):
    """x4 (FP64) of compact records: 0.5 * long_scale * (c - 7.5) per coordinate."""
    ls = comp[:, 4:8].copy().view(np.float32)[:, 0].astype(np.float64)
    codes = comp[:, 16:16 + cd // 2]
    nib = np.stack([codes & 15, codes >> 4], axis=2).reshape(comp.shape[0], cd).astype(np.float64)
    return 0.5 * ls[:, None] * (nib - 7.5)


@pytest.mark.parametrize("dim", [960, 1536])
def test_precision_estimators(gpu, dim):
    import quantization_graph_amd as qg
    x, q = gaussian_mixture(2000, dim, 8, 0.3, 17 + dim, nq=6)
    c = x.mean(0).astype(np.float32)
    sp = _qr_space(qg, dim, c)
    cd = sp.code_dim
    print("case: precision-mode estimators, D %d" % dim, flush=True)
    xd = gpu.from_numpy(x).cuda()
    ix = sp.encode(xd)
    comp = ix.to_records()[0].cpu().numpy()
    rng = np.random.default_rng(2)
    per = 300
    ids = np.stack([rng.permutation(x.shape[0])[:per] for _ in range(q.shape[0])]).astype(np.int32)
    qv8 = sp.prepare(gpu.from_numpy(q).cuda(), "int8")
    est, dot = (t.cpu().numpy() for t in sp.scan_fixed(qg.SCAN_INT8_FULL, ix, qv8, ids, want=("f", "dot"))[::4])
    qvf = sp.prepare(gpu.from_numpy(q).cuda(), "full")
    asym = sp.scan_fixed(qg.SCAN_ASYM, ix, qvf, ids)[0].cpu().numpy()
    rq = sp.rotate(gpu.from_numpy(q).cuda()).cpu().numpy()
    rc = sp.rotate(gpu.from_numpy(c[None, :]).cuda()).cpu().numpy()[0]
    orc = Oracle(dim, 100)
    orc.set_center(c)
    Q = sp.rotation_matrix()
    x4 = _decode_long(comp, cd)
    norm_sqr = comp[:, :4].copy().view(np.float32)[:, 0].astype(np.float64)
    worst = 0.0
    for i in range(q.shape[0]):
        oq = orc.prepare_rotated(q[i], rq[i], rc, 3)
        e_o, _, d_o = orc.int8_full(oq, comp[ids[i]])
        assert np.array_equal(dot[i], d_o)
        assert np.array_equal(est[i].view(np.uint32), e_o.view(np.uint32))
        orc.free_query(oq)
        # the FP32 estimate against the same estimator in FP64 with the FP64
        # rotation: norm_sqr + |q - c|^2 - 2 <x4, Q (q - c)>
        q64 = Q[:, :dim] @ (q[i].astype(np.float64) - c.astype(np.float64))
        qn2 = float(np.dot(q[i].astype(np.float64) - c, q[i].astype(np.float64) - c))
        ip = x4[ids[i]] @ q64
        d64 = norm_sqr[ids[i]] + qn2 - 2.0 * ip
        scale = norm_sqr[ids[i]] + qn2  # the magnitude the subtraction cancels
        err = np.abs(asym[i].astype(np.float64) - d64)
        assert np.all(err <= 1e-5 * np.abs(d64) + 4e-6 * scale), (err / np.abs(d64)).max()
        worst = max(worst, float((err / np.abs(d64)).max()))
    print("max relative error of the FP32 estimate vs FP64: %.3g" % worst)
