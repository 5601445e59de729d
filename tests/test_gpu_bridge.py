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
