"""The oracle's precision-mode pieces (oracle/rabitq_oracle.c), pinned on
the CPU by what the mathematics fixes:

* ro_encode_from_unit with the reference's factors (precision 0) is the
  reference codec: fed the reference rotation's unit vector and norm it
  writes the records ro_encode_full writes (which the reference bridge's
  fixtures pin).
* With precision 1 only the short and error factors change: codes, scales,
  residual codes, norm and long scale are the same bytes, and both factors
  equal their FP64 formulas (2|o| / (sum|o| / 2) and the ExRaBitQ bound
  sqrt((1 - a^2) / a^2) * 5.75 / 16 / sqrt(cd) * 2|o| with
  a = sum|o| / sqrt(cd)) to FP32 rounding.
* ro_encode_full_rotated: |x - c| from its compensated sum equals the FP64
  norm to FP32 rounding; the unit vector it quantises is y / |y| (same
  codes as ro_encode_from_unit of y / |y| in FP64 rounded once).
* ro_prepare_query_rotated of the reference rotation's outputs is
  ro_prepare_query.
* numpy's QR with the column signs fixed (the convention the GPU's
  Gram-Schmidt produces) is orthogonal with an upper-triangular R of
  positive diagonal.
"""
import numpy as np
import pytest

from conftest import gaussian_mixture
from oracle_bind import Oracle


def _fields(full, cd):
    h = full[:16].copy().view(np.float32)
    return dict(norm_sqr=h[0], long_scale=h[1], short_scale=h[2], error_scale=h[3])


@pytest.mark.parametrize("dim", [960, 200, 1536])
def test_codec_tail_is_the_reference_codec(dim):
    x, _ = gaussian_mixture(40, dim, 4, 0.3, 3 + dim)
    c = x.mean(0).astype(np.float32)
    orc = Oracle(dim, 100)
    orc.set_center(c)
    cd = orc.cd
    ref = orc.encode_full(x)
    for i in range(x.shape[0]):
        o = x[i] - c
        # the reference's rn (recovered from its norm_sqr = rn * rn; a row
        # where that is ambiguous is skipped) and its rotation of o / rn
        rn = np.sqrt(ref[i, :4].copy().view(np.float32)[0])
        full = ref[i]
        # rebuild rot exactly as the reference: unit vector then rotation
        u = (o * (np.float32(1.0) / np.float32(rn))).astype(np.float32)
        rot = orc.rotate(u)
        got0 = orc.encode_from_unit(np.float32(rn), rot, 0)
        if np.float32(rn) * np.float32(rn) != ref[i, :4].copy().view(np.float32)[0]:
            continue  # rn recovered by sqrt differs from the reference's: skip this row
        assert np.array_equal(got0[16:], full[16:]) and np.array_equal(got0[4:16], full[4:16])
        got1 = orc.encode_from_unit(np.float32(rn), rot, 1)
        # everything but the short / error factors is the same bytes
        assert np.array_equal(got1[:8], got0[:8]) and np.array_equal(got1[16:], got0[16:])
        f1 = _fields(got1, cd)
        l1 = np.sum(np.abs(rot.astype(np.float64)))
        short64 = 2.0 * float(rn) / (0.5 * l1)
        a = min(max(l1 / np.sqrt(cd), 1e-6), 0.999999)
        err64 = np.sqrt((1 - a * a) / (a * a)) * 5.75 / 16 / np.sqrt(cd) * 2 * float(rn)
        assert abs(f1["short_scale"] - short64) <= 4 * np.spacing(np.float32(short64))
        assert abs(f1["error_scale"] - err64) <= 1e-5 * err64


@pytest.mark.parametrize("dim", [960, 1536])
def test_encode_full_rotated_norms(dim):
    x, _ = gaussian_mixture(30, dim, 4, 0.3, 8 + dim)
    c = x.mean(0).astype(np.float32)
    orc = Oracle(dim, 100)
    orc.set_center(c)
    cd = orc.cd
    rng = np.random.default_rng(0)
    y = rng.normal(size=(x.shape[0], cd)).astype(np.float32)
    full = orc.encode_full_rotated(x, y)
    for i in range(x.shape[0]):
        rn64 = np.linalg.norm(x[i].astype(np.float64) - c.astype(np.float64))
        ns = full[i, :4].copy().view(np.float32)[0]
        assert abs(np.sqrt(np.float64(ns)) - rn64) <= 2.5 * np.spacing(np.float32(rn64))
        # the unit vector quantised is y / |y|: same records as the codec of
        # y / |y| computed in FP64 and rounded once, up to a last-bit flip of
        # the normalisation; compare the codes, which do not depend on it here
        unit = (y[i].astype(np.float64) / np.linalg.norm(y[i].astype(np.float64))).astype(np.float32)
        alt = orc.encode_from_unit(np.sqrt(np.float32(ns)), unit, 1)
        codes = slice(24 + 2 * (cd // 16), 24 + 2 * (cd // 16) + cd // 2)
        assert (full[i, codes] != alt[codes]).mean() < 0.01


def test_prepare_rotated_is_prepare():
    dim = 960
    x, q = gaussian_mixture(50, dim, 4, 0.3, 1, nq=4)
    c = x.mean(0).astype(np.float32)
    orc = Oracle(dim, 100)
    orc.set_center(c)
    rc = orc.rotate(c)
    for codec in (0, 3):
        for i in range(q.shape[0]):
            a = orc.prepare(q[i], codec)
            b = orc.prepare_rotated(q[i], orc.rotate(q[i]), rc, codec)
            assert np.array_equal(orc.qvec(a).view(np.uint32), orc.qvec(b).view(np.uint32))
            assert np.array_equal(orc.qx8(a), orc.qx8(b))
            for f in ("query_norm", "query_norm_sqr", "half_sum", "sum", "positive_sum", "int8_scale",
                      "int8_code_sum"):
                assert getattr(a, f) == getattr(b, f)
            orc.free_query(a)
            orc.free_query(b)


@pytest.mark.parametrize("cd", [128, 1024])
def test_qr_sign_convention(cd):
    g = np.random.default_rng(5).standard_normal((cd, cd))
    q, r = np.linalg.qr(g)
    s = np.sign(np.diag(r))
    q, r = q * s[None, :], r * s[:, None]
    assert (np.diag(r) > 0).all()
    assert np.abs(q @ r - g).max() < 1e-11
    assert np.abs(q.T @ q - np.eye(cd)).max() < 1e-13
