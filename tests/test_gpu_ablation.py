"""The ablation codecs on the GPU: INT4 and B1 query preparation, the B1
sign-code gate and the INT4 integer gate with its remaining-bit distances,
against the oracle (which the CPU suite pins to the reference bridge
compiled from source), and the bridge's rabitq_* entry points for both
codecs against that compiled bridge, bit for bit."""
import numpy as np
import pytest

from conftest import gaussian_mixture
from oracle_bind import Oracle, Ref
from test_gpu_bridge import Ours

pytestmark = pytest.mark.gpu

FIELDS = ("lower_bound", "short_ip", "alpha", "ip_hat", "error_bound", "valid")


def _setup(gpu, dim, n=320, nq=6, seed=5):
    import quantization_graph_amd as qg
    x, q = gaussian_mixture(n, dim, 8, 0.3, seed, nq=nq)
    c = x.mean(0).astype(np.float32)
    sp = qg.RabitqSpace(dim, 100)
    sp.set_center(c)
    orc = Oracle(dim, 100)
    orc.set_center(c)
    ix = sp.encode(gpu.from_numpy(x).cuda())
    comp, _ = ix.to_records()
    return qg, sp, orc, ix, comp.cpu().numpy(), q


@pytest.mark.parametrize("dim", [960, 1536, 100])
@pytest.mark.parametrize("codec", [1, 2, 4])
def test_ablation_query_prepare(gpu, dim, codec):
    qg, sp, orc, ix, comp, q = _setup(gpu, dim)
    qv = sp.prepare(gpu.from_numpy(q).cuda(), codec)
    s = qv.scalars()
    x8 = qv.x8.cpu().numpy()
    for i in range(q.shape[0]):
        oq = orc.prepare(q[i], codec)
        assert np.array_equal(x8[i], orc.qx8(oq))
        assert s["int8_code_sum"][i] == oq.int8_code_sum
        if codec == 2:
            assert np.float32(s["int8_scale"][i]).tobytes() == np.float32(oq.int8_scale).tobytes()
            assert np.float32(s["int8_error_norm"][i]).tobytes() == np.float32(oq.int8_error_norm).tobytes()
        else:
            for mine, theirs in (("int8_scale", "b1_short_scale"), ("int8_error_norm", "b1_error_scale"),
                                 ("b1_alpha", "b1_alpha")):
                assert np.float32(s[mine][i]).tobytes() == np.float32(getattr(oq, theirs)).tobytes(), mine
        orc.free_query(oq)


@pytest.mark.parametrize("dim", [960, 1536, 100])
@pytest.mark.parametrize("codec", [1, 2])
def test_ablation_gates_match_oracle(gpu, dim, codec):
    qg, sp, orc, ix, comp, q = _setup(gpu, dim)
    qv = sp.prepare(gpu.from_numpy(q).cuda(), codec)
    n = comp.shape[0]
    nq = q.shape[0]
    qidx = np.repeat(np.arange(nq), n).astype(np.int32)
    ids = np.tile(np.arange(n), nq).astype(np.int32)
    kernel = qg.SCAN_GATE_B1 if codec == 1 else qg.SCAN_GATE_INT8
    _, _, est, _, _ = sp.scan_pairs(kernel, ix, qv, qidx, ids, eps0=1.9, want=("est",))
    est = qg.estimates_np(est).reshape(nq, n)
    msb_o, fac_o = orc.sidecar(comp)
    for i in range(nq):
        oq = orc.prepare(q[i], codec)
        want = orc.gate(oq, msb_o, fac_o, codec, 1.9)
        for f in FIELDS:
            assert est[i][f].tobytes() == want[f].tobytes(), "%s of query %d" % (f, i)
        if codec == 2:
            d, _, _, _, _ = sp.scan_pairs(qg.SCAN_REM_INT8, ix, qv, np.full(n, i, np.int32), np.arange(n, dtype=np.int32),
                                          short_ips=want["short_ip"].copy())
            d_o, _ = orc.rem_int8(oq, comp, want["short_ip"])
            assert d.cpu().numpy().tobytes() == d_o.tobytes()
        orc.free_query(oq)


@pytest.mark.parametrize("dim", [960, 1536])
def test_bridge_ablation_gates_match_reference(gpu, dim):
    if not Ref.available():
        pytest.skip("oracle/_ref not built")
    x, q = gaussian_mixture(200, dim, 8, 0.3, 9, nq=4)
    c = x.mean(0).astype(np.float32)
    ref, ours = Ref(dim, 100), Ours(dim, 100)
    for s in (ref, ours):
        s.set_center(c)
    comp, _ = ref.split(ref.encode_full(x))
    msb, fac = ref.sidecar(comp)
    ids = np.random.default_rng(1).permutation(200).astype(np.uint32)[:150]
    for codec in (1, 2):
        ref.set_codec(codec)
        ours.set_codec(codec)
        for i in range(q.shape[0]):
            pr = ref.prepare(q[i])
            want = ref.gate_batch(pr, msb, fac, ids, 1.9)
            po = ours.prepare(q[i])
            got = ours.gate_batch(po, msb, fac, ids, 1.9)
            for f in FIELDS:
                assert got[f].tobytes() == want[f].tobytes(), "codec %d %s" % (codec, f)
            if codec == 2:
                modes = np.full(ids.size, 2, np.uint8)
                assert ours.distance_batch(po, comp, ids, modes, want["short_ip"]).tobytes() == \
                    ref.distance_batch(pr, comp, ids, modes, want["short_ip"]).tobytes()
