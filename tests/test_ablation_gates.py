"""The ablation codecs' 1-bit gates in the oracle against the reference
bridge compiled from source: B1 (the query's sign code against the MSB
plane, compute_b1_paper_prune_estimate_batch_by_id, space_rabitq.h:4607-4753)
and Int4 (the dithered 4-bit query code through the integer gate,
compute_int_paper_prune_estimate_batch_by_id, :4991-5093), with the INT4
code's remaining-bit distances (mode 2 of query_distance_batch). Every
estimate field must be equal bit for bit; the reference compiler's one
contraction in the B1 bound is pinned by dropping it.
"""
import numpy as np
import pytest

from conftest import gaussian_mixture
from oracle_bind import FMA, Oracle, Ref

FIELDS = ("lower_bound", "short_ip", "alpha", "ip_hat", "error_bound", "valid")


def _case(dim, n=300, nq=8, seed=3):
    if not Ref.available():
        pytest.skip("oracle/_ref not built")
    x, q = gaussian_mixture(n, dim, 8, 0.3, seed, nq=nq)
    c = x.mean(0).astype(np.float32)
    ref = Ref(dim, 100)
    ref.set_center(c)
    comp, res = ref.split(ref.encode_full(x))
    msb, fac = ref.sidecar(comp)
    orc = Oracle(dim, 100)
    orc.set_center(c)
    return ref, orc, comp, msb, fac, q


def _equal(a, b):
    return all(np.asarray(a[f]).tobytes() == np.asarray(b[f]).tobytes() for f in FIELDS)


@pytest.mark.parametrize("dim", [960, 1536, 100])
@pytest.mark.parametrize("codec", [1, 2])
def test_ablation_gate_matches_reference(dim, codec):
    ref, orc, comp, msb, fac, q = _case(dim)
    ref.set_codec(codec)
    ids = np.arange(comp.shape[0], dtype=np.uint32)
    valid = 0
    for i in range(q.shape[0]):
        pq = ref.prepare(q[i])
        want = ref.gate_batch(pq, msb, fac, ids, 1.9)
        oq = orc.prepare(q[i], codec)
        got = orc.gate(oq, msb, fac, codec, 1.9)
        assert _equal(got, want), "codec %d query %d" % (codec, i)
        valid += int(got["valid"].sum())
        if codec == 2:
            # the INT4 code behind the gate: its remaining-bit distances
            modes = np.full(ids.size, 2, np.uint8)
            d_ref = ref.distance_batch(pq, comp, ids, modes, got["short_ip"])
            d_orc, _ = orc.rem_int8(oq, comp, got["short_ip"])
            assert d_ref.tobytes() == d_orc.tobytes()
        orc.free_query(oq)
    assert valid > 0


def test_b1_gate_contraction_is_pinned():
    """Without the reference compiler's fused db_err + query_err the B1
    estimates differ from the compiled reference."""
    ref, orc, comp, msb, fac, q = _case(960)
    ref.set_codec(1)
    ids = np.arange(comp.shape[0], dtype=np.uint32)
    mask = orc.get_fma()
    assert mask & FMA["B1_SUM_DB"]
    differs = False
    try:
        orc.set_fma(mask & ~FMA["B1_SUM_DB"])
        for i in range(q.shape[0]):
            pq = ref.prepare(q[i])
            want = ref.gate_batch(pq, msb, fac, ids, 1.9)
            oq = orc.prepare(q[i], 1)
            got = orc.gate(oq, msb, fac, 1, 1.9)
            orc.free_query(oq)
            differs = differs or not _equal(got, want)
    finally:
        orc.set_fma(mask)
    assert differs


def test_b1_scalars_closed_form():
    """The sign code's scalars from their definitions in FP64: a_q = |q|_1 /
    |q|, alpha = min(1, a_q / sqrt(cd)), short scale 4 |q| / a_q, error
    scale 2 |q| sqrt((1 - alpha^2) / alpha^2) / sqrt(cd - 1)."""
    ref, orc, comp, msb, fac, q = _case(960)
    for i in range(q.shape[0]):
        oq = orc.prepare(q[i], 1)
        qq = orc.qvec(oq).astype(np.float64)
        qn = float(oq.query_norm)
        a_q = np.abs(qq).sum() / qn
        al = min(1.0, a_q / np.sqrt(orc.cd))
        assert abs(oq.b1_alpha - al) <= 1e-6 * al
        assert abs(oq.b1_short_scale - 4 * qn / a_q) <= 1e-6 * (4 * qn / a_q)
        es = 2 * qn * np.sqrt(max(0.0, (1 - al * al) / (al * al))) / np.sqrt(orc.cd - 1)
        assert abs(oq.b1_error_scale - es) <= 1e-5 * es
        x8 = orc.qx8(oq)
        assert np.array_equal(x8, np.where(qq > 0, 1, -1)) and oq.int8_code_sum == int(x8.sum())
        orc.free_query(oq)
