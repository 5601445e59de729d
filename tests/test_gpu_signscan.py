"""HIP 1-bit query-codec scan (qg_signscan_*) against the reference's own
outputs (tests/golden/signscan_ref.npz) and the CPU oracle: database words,
query rows, scalars and the four codecs' distances bit-exact, in code
dimensions 64 to 2048; at a million vectors, through properties that do not
depend on size.
"""
import numpy as np
import pytest

from oracle_bind import Oracle

pytestmark = pytest.mark.gpu

NAMES = ("fp32_sum", "fp32_norm2", "b1_scale", "b1_constant", "int4_scale", "int4_constant", "int8_scale",
         "int8_constant")


def _f32_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def _setup(gpu, dim, seed=100):
    import quantization_graph_amd as qg
    from quantization_graph_amd.signscan import SignScan
    sp = qg.RabitqSpace(dim, seed)
    return sp, SignScan(sp)


def _u64(t):
    return t.cpu().numpy().view(np.uint64)


def test_signscan_matches_reference_fixture(gpu, golden_ss):
    g = golden_ss
    print("signscan fixture: 96 x 960, 5 queries", flush=True)
    sp, ss = _setup(gpu, 960, int(g["seed"]))
    x = gpu.from_numpy(g["x"]).cuda()
    db = ss.encode(x)
    assert np.array_equal(_u64(db), g["db"])
    pq = ss.prepare(gpu.from_numpy(g["queries"]).cuda())
    assert _f32_equal(pq.q.cpu().numpy(), g["q_rot"])
    assert np.array_equal(_u64(pq.bits).reshape(5, -1), g["bits"])
    f = pq.scalar_fields()
    for k, name in enumerate(NAMES):
        assert _f32_equal(f[name], g["fscalars"][:, k]), name
    assert np.array_equal(f["int4_sum"], g["isums"][:, 0])
    assert np.array_equal(f["int8_sum"], g["isums"][:, 1])
    for codec in range(4):
        out = ss.scan(codec, db, 96, pq).cpu().numpy()
        assert _f32_equal(out, g["dists"][:, codec]), codec


@pytest.mark.parametrize("dim", [48, 100, 960, 1536, 2000])
def test_signscan_matches_oracle(gpu, dim):
    print("signscan vs oracle: dim %d" % dim, flush=True)
    sp, ss = _setup(gpu, dim)
    orc = Oracle(dim, 100)
    W = orc.cd // 64
    rng = np.random.default_rng(dim)
    n = 32 * 37  # an odd number of blocks: the last wavefront is half full
    x = (rng.standard_normal((n, dim)) * rng.uniform(0.05, 5.0, (n, 1))).astype(np.float32)
    x[3] = 0.0
    q = (rng.standard_normal((5, dim)) * rng.uniform(0.01, 50.0, (5, 1))).astype(np.float32)
    q[1] = 0.0
    q[2] = -np.abs(q[2])
    db = ss.encode(gpu.from_numpy(x).cuda())
    db_o = orc.ss_encode(x)
    assert np.array_equal(_u64(db), db_o)
    pq = ss.prepare(gpu.from_numpy(q).cuda())
    qg_q, qg_bits, f = pq.q.cpu().numpy(), _u64(pq.bits).reshape(5, -1), pq.scalar_fields()
    outs = [ss.scan(c, db, n, pq).cpu().numpy() for c in range(4)]
    for i in range(5):
        qo, bo, so = orc.ss_prepare(q[i])
        assert _f32_equal(qg_q[i], qo)
        assert np.array_equal(qg_bits[i], bo)
        assert _f32_equal([f[k][i] for k in NAMES], [getattr(so, k) for k in NAMES])
        assert (f["int4_sum"][i], f["int8_sum"][i]) == (so.int4_sum, so.int8_sum)
        for codec in range(4):
            assert _f32_equal(outs[codec][i], orc.ss_scan(codec, db_o, n, qo, bo, so)), (i, codec)
    assert W == qg_bits.shape[1] // 13


@pytest.mark.parametrize("dim,nq", [(100, 70), (960, 9), (960, 70), (1536, 37), (2000, 70)])
def test_signscan_batched_matches_oracle(gpu, dim, nq):
    """Batches of 8 or more queries run the integer codecs on the matrix
    cores (int8 GEMM of query codes x database bits); every distance must
    still equal the oracle's bit for bit, across query chunks and partial
    16-query tiles."""
    print("signscan batched vs oracle: dim %d, %d queries" % (dim, nq), flush=True)
    sp, ss = _setup(gpu, dim)
    orc = Oracle(dim, 100)
    rng = np.random.default_rng(dim + nq)
    n = 32 * 21
    x = (rng.standard_normal((n, dim)) * rng.uniform(0.05, 5.0, (n, 1))).astype(np.float32)
    q = (rng.standard_normal((nq, dim)) * rng.uniform(0.01, 50.0, (nq, 1))).astype(np.float32)
    q[0] = 0.0
    db = ss.encode(gpu.from_numpy(x).cuda())
    db_o = orc.ss_encode(x)
    pq = ss.prepare(gpu.from_numpy(q).cuda())
    outs = {c: ss.scan(c, db, n, pq).cpu().numpy() for c in (1, 2, 3)}
    for i in range(nq):
        qo, bo, so = orc.ss_prepare(q[i])
        for codec in (1, 2, 3):
            assert _f32_equal(outs[codec][i], orc.ss_scan(codec, db_o, n, qo, bo, so)), (i, codec)


def test_signscan_rejects_partial_blocks(gpu):
    from quantization_graph_amd._lib import QGError
    sp, ss = _setup(gpu, 100)
    x = gpu.randn(40, 100, device="cuda")
    with pytest.raises(QGError, match="blocks of 32"):
        ss.encode(x)
    db = ss.encode(x[:32])
    pq = ss.prepare(x[:2])
    with pytest.raises(QGError, match="blocks of 32"):
        ss.scan(1, db, 31, pq)
    with pytest.raises(QGError, match="codec"):
        ss.scan(7, db, 32, pq)


def test_signscan_full_size_self_match(gpu):
    """1M x 960: a query equal to a database row has Hamming distance 0 to
    that row's sign code, so the B1 distance there is b1_constant - 2 *
    b1_scale * cd and no row is closer; the FP32 codec's distance there is
    cd + |q|^2 - 2 sum|q_i| (its own signs), the minimum over all sign
    vectors; the INT codecs' integer dot there is sum |code_i|."""
    torch = gpu
    print("signscan full size: 1M x 960", flush=True)
    sp, ss = _setup(gpu, 960)
    g = torch.Generator(device="cuda")
    g.manual_seed(9)
    n = 1 << 20
    x = torch.randn((n, 960), generator=g, device="cuda")
    db = ss.encode(x)
    rows = torch.tensor([0, 1, 777, 123456, n - 33, n - 1], device="cuda")
    pq = ss.prepare(x[rows])
    f = pq.scalar_fields()
    cd = 1024
    out_b1 = ss.scan(1, db, n, pq)
    out_fp = ss.scan(0, db, n, pq)
    out_i8 = ss.scan(3, db, n, pq)
    torch.cuda.synchronize()
    own_b1 = out_b1.gather(1, rows[:, None]).squeeze(1).cpu().numpy()
    want_b1 = (f["b1_constant"] - np.float32(2.0) * f["b1_scale"] * np.float32(cd)).astype(np.float32)
    assert np.abs(own_b1 - want_b1).max() <= 1e-6 * np.abs(want_b1).max()
    assert np.array_equal(out_b1.min(1).values.cpu().numpy(), own_b1)
    qrot = pq.q.double()
    want_fp = (cd + (qrot * qrot).sum(1) - 2.0 * qrot.abs().sum(1)).cpu().numpy()
    own_fp = out_fp.gather(1, rows[:, None]).squeeze(1).double().cpu().numpy()
    assert np.abs(own_fp - want_fp).max() <= 1e-5 * np.abs(want_fp).max()
    assert np.array_equal(out_fp.argmin(1).cpu().numpy(), rows.cpu().numpy())
    # the batched (matrix-core) integer scans equal the single-query scans
    # on all 1M rows: 50 queries whose first 6 are the rows above
    more = torch.cat([rows, torch.randint(0, n, (44,), generator=g, device="cuda")])
    pq50 = ss.prepare(x[more])
    for codec, single in ((1, out_b1), (3, out_i8), (2, ss.scan(2, db, n, pq))):
        batched = ss.scan(codec, db, n, pq50)[:6]
        assert torch.equal(batched.view(torch.int32), single.view(torch.int32)), codec
    # INT8: own-row dot = sum |code| (codes of the query's positive / non-positive
    # coordinates meet signs +1 / -1; zero codes add nothing)
    bits = pq.bits.cpu().numpy().view(np.uint64)[:, 5:13, :]
    own_i8 = out_i8.gather(1, rows[:, None]).squeeze(1).double().cpu().numpy()
    for i in range(rows.numel()):
        v = np.zeros(cd, np.int64)
        for p in range(8):
            b = ((bits[i, p][:, None] >> np.arange(64, dtype=np.uint64)[None, :]) & np.uint64(1)).reshape(-1)
            v += b.astype(np.int64) << p
        codes = np.where(v >= 128, v - 256, v)
        qv = pq.q[i].cpu().numpy()
        dot = int(np.where(qv > 0, codes, -codes).sum())
        assert dot == int(np.abs(codes).sum())
        back = (np.float64(f["int8_constant"][i]) - own_i8[i]) / (2.0 * np.float64(f["int8_scale"][i]))
        assert abs(back - dot) < 0.5
