"""The oracle's restatement of the 1-bit query-codec scan
(experiments/04_query_codec_1bit_scan/query_codec_1bit_scan.cpp) pinned
three ways, on the CPU:

* against the reference's own outputs in tests/golden/signscan_ref.npz and,
  where oracle/_ref was built, against the compiled reference live on fresh
  inputs (bit-exact: database words, query rows, scalars, all four codecs);
* against closed forms in every code dimension: the integer dots by brute
  force over decoded +-1 signs and query codes, the FP32 distance against
  the FP64 value of cd + |q|^2 - 2 <q, s>, the B1 distance through the
  Hamming distance, the query codes against their definition;
* the cache-file format against the reference's own writer and loader.
"""
import os
import tempfile

import numpy as np
import pytest

from oracle_bind import FMA, Oracle, RefSignScan, SSQuery
from quantization_graph_amd import formats

NAMES = ("fp32_sum", "fp32_norm2", "b1_scale", "b1_constant", "int4_scale", "int4_constant", "int8_scale",
         "int8_constant")


def _f32_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def _scalars(qs):
    return [getattr(qs, k) for k in NAMES], [qs.int4_sum, qs.int8_sum]


def test_oracle_matches_reference_fixture(golden_ss):
    g = golden_ss
    o = Oracle(960, int(g["seed"]))
    x, db_r = g["x"], g["db"]
    assert np.array_equal(o.ss_encode(x), db_r)
    for i in range(g["queries"].shape[0]):
        q, bits, qs = o.ss_prepare(g["queries"][i])
        assert _f32_equal(q, g["q_rot"][i])
        assert np.array_equal(bits, g["bits"][i])
        f, s = _scalars(qs)
        assert _f32_equal(f, g["fscalars"][i])
        assert s == list(g["isums"][i])
        for codec in range(4):
            assert _f32_equal(o.ss_scan(codec, db_r, x.shape[0], q, bits, qs), g["dists"][i, codec]), codec


def test_fixture_detects_a_dropped_contraction(golden_ss):
    """Planted mistake: without the reference compiler's fused multiply-add
    in the B1 and integer distances the oracle no longer matches."""
    g = golden_ss
    o = Oracle(960, int(g["seed"]))
    mask = o.lib.ro_get_fma_mask()  # the mask is global to the oracle library: restore it
    o.set_fma(mask & ~(FMA["SS_B1DIST"] | FMA["SS_INTDIST"]))
    try:
        miss = 0
        for i in range(g["queries"].shape[0]):
            q, bits, qs = o.ss_prepare(g["queries"][i])
            for codec in (1, 2, 3):
                miss += not _f32_equal(o.ss_scan(codec, g["db"], 96, q, bits, qs), g["dists"][i, codec])
    finally:
        o.set_fma(mask)
    assert miss > 0


@pytest.mark.skipif(not RefSignScan.available(), reason="oracle/_ref/libsignscan_ref.so not built")
def test_oracle_matches_compiled_reference():
    ref = RefSignScan(100)
    o = Oracle(960, 100)
    rng = np.random.default_rng(11)
    n = 320
    x = (rng.standard_normal((n, 960)) * rng.uniform(0.01, 10, (n, 1))).astype(np.float32)
    x[5] = 0.0
    with tempfile.TemporaryDirectory() as td:
        db_r = ref.build_database(x, td)
    assert np.array_equal(ref.signs(), o.signs())
    assert np.array_equal(o.ss_encode(x), db_r)
    for i in range(40):
        raw = (rng.standard_normal(960) * rng.uniform(1e-3, 100)).astype(np.float32)
        h, q_r, bits_r, sc = ref.prepare(raw)
        assert ref.validate(db_r, n, h)
        q, bits, qs = o.ss_prepare(raw)
        assert _f32_equal(q, q_r)
        assert np.array_equal(bits, bits_r)
        f, s = _scalars(qs)
        assert _f32_equal(f, [sc[k] for k in NAMES])
        assert s == [sc["int4_sum"], sc["int8_sum"]]
        codecs = range(4) if i < 8 else (1, 2, 3)
        for codec in codecs:
            assert _f32_equal(o.ss_scan(codec, db_r, n, q, bits, qs), ref.scan(codec, db_r, n, h)), (i, codec)
        ref.free(h)


def _signs_of(db, n, W):
    """+-1 matrix [n, 64 W] decoded from the blocked database words."""
    words = db.reshape(n // 32, W, 32).transpose(0, 2, 1).reshape(n, W)
    bits = (words[:, :, None] >> np.arange(64, dtype=np.uint64)[None, None, :]) & np.uint64(1)
    return np.where(bits.reshape(n, 64 * W) == 1, 1, -1).astype(np.int64)


def _codes_of(bits_rows, W, B):
    """Signed query codes from B two's-complement planes of W words."""
    planes = bits_rows.reshape(B, W)
    v = np.zeros(64 * W, np.int64)
    for p in range(B):
        b = ((planes[p][:, None] >> np.arange(64, dtype=np.uint64)[None, :]) & np.uint64(1)).reshape(-1)
        v += b.astype(np.int64) << p
    return np.where(v >= (1 << (B - 1)), v - (1 << B), v)


@pytest.mark.parametrize("dim", [48, 100, 960, 1536])
def test_scan_closed_forms(dim):
    o = Oracle(dim, 100)
    cd, W = o.cd, o.cd // 64
    rng = np.random.default_rng(dim)
    n = 64
    x = rng.standard_normal((n, dim)).astype(np.float32)
    db = o.ss_encode(x)
    S = _signs_of(db, n, W)
    rot = np.array([o.rotate(x[i]) for i in range(n)])
    assert np.array_equal(S == 1, rot > 0)
    for t in range(3):
        raw = (rng.standard_normal(dim) * (10.0 ** (t - 1))).astype(np.float32)
        q, bits, qs = o.ss_prepare(raw)
        assert _f32_equal(q, o.rotate(raw))
        q64 = q.astype(np.float64)
        # FP32 codec: the FP64 value of cd + |q|^2 - 2 <q, s>
        exact = cd + (q64 * q64).sum() - 2.0 * (S * q64[None, :]).sum(1)
        d = o.ss_scan(0, db, n, q, bits, qs)
        assert np.abs(d - exact).max() <= 1e-5 * max(1.0, np.abs(exact).max())
        # B1: sign code and Hamming distance
        b1 = _codes_of(bits[:W], W, 1)
        assert np.array_equal(b1 == -1, q > 0)
        ham = (S != np.where(q > 0, 1, -1)[None, :]).sum(1)
        sd = np.float32(cd) - np.float32(2.0) * ham.astype(np.float32)
        expect = np.float32(qs.b1_constant) - np.float32(2.0 * qs.b1_scale) * sd
        assert np.abs(o.ss_scan(1, db, n, q, bits, qs) - expect).max() <= 1e-6 * np.abs(expect).max()
        # scalars: sequential float32 sums in index order (np.add.accumulate
        # adds left to right), products rounded before the add
        seq = lambda v: np.add.accumulate(v.astype(np.float32), dtype=np.float32)[-1]
        assert np.float32(qs.fp32_sum) == seq(q)
        assert np.float32(qs.fp32_norm2) == seq(q * q)
        assert np.float32(qs.b1_scale) == seq(np.abs(q)) / np.float32(cd)
        assert abs(qs.b1_scale - np.abs(q64).mean()) <= 1e-4 * qs.b1_scale
        # INT4 / INT8: codes are round-to-nearest-even of q / scale, clamped;
        # the scan's integer dot is exactly sum(code * sign)
        for codec, B, row0, scale, csum, const in ((2, 4, 1, qs.int4_scale, qs.int4_sum, qs.int4_constant),
                                                   (3, 8, 5, qs.int8_scale, qs.int8_sum, qs.int8_constant)):
            qmax = (1 << (B - 1)) - 1
            codes = _codes_of(bits[row0 * W:(row0 + B) * W], W, B)
            want = np.clip(np.rint(q / np.float32(scale)), -qmax, qmax).astype(np.int64)
            assert np.array_equal(codes, want)
            assert codes.sum() == csum
            m = np.abs(q).max()
            assert np.float32(scale) == (np.float32(m / np.float32(qmax)) if m > 0 else np.float32(1.0))
            dot = (S * codes[None, :]).sum(1)
            got = o.ss_scan(codec, db, n, q, bits, qs)
            expect = np.float32(const) - np.float32(2.0 * scale) * dot.astype(np.float32)
            assert np.abs(got - expect).max() <= 1e-6 * max(1.0, np.abs(expect).max())
            # the exact integer dot is recoverable from the distance
            back = (np.float64(const) - got.astype(np.float64)) / (2.0 * np.float64(scale))
            assert np.abs(back - dot).max() < 0.5 + 1e-3 * np.abs(dot).max()


def test_zero_query_uses_unit_scale():
    o = Oracle(100, 7)
    q, bits, qs = o.ss_prepare(np.zeros(100, np.float32))
    assert qs.int4_scale == 1.0 and qs.int8_scale == 1.0 and qs.int4_sum == 0 and qs.int8_sum == 0
    assert not bits.any()
    assert qs.b1_constant == np.float32(o.cd)


@pytest.mark.skipif(not RefSignScan.available(), reason="oracle/_ref/libsignscan_ref.so not built")
def test_cache_file_round_trips_with_the_reference():
    ref = RefSignScan(100)
    o = Oracle(960, 100)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((64, 960)).astype(np.float32)
    with tempfile.TemporaryDirectory() as td:
        cache = os.path.join(td, "db.bin")
        db_r = ref.build_database(x, td, cache=cache)      # the reference writes its cache
        head, words = formats.read_sign_cache(cache, mmap=False)
        src = os.path.getsize(os.path.join(td, "qcs_base.fvecs"))
        assert head == dict(input_dim=960, code_dim=1024, batch=32, data_count=64, source_size=src, seed=100,
                            payload_bytes=64 * 16 * 8)
        assert np.array_equal(words, db_r)
        assert np.array_equal(words, o.ss_encode(x))
        # a cache written here is loaded by the reference: its payload (marked
        # by flipping every word) comes back instead of a rebuild
        mine = os.path.join(td, "mine.bin")
        formats.write_sign_cache(mine, ~db_r, 960, 1024, 64, src, 100)
        with open(mine, "rb") as f_mine, open(cache, "rb") as f_ref:
            assert f_mine.read(64) == f_ref.read(64)
        assert np.array_equal(ref.build_database(x, td, cache=mine), ~db_r)


def test_cache_reader_rejects_bad_headers():
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "c.bin")
        formats.write_sign_cache(p, np.zeros(32 * 16, np.uint64), 960, 1024, 32, 0, 100)
        formats.read_sign_cache(p)
        raw = bytearray(open(p, "rb").read())
        raw[0] ^= 1
        open(p, "wb").write(bytes(raw))
        with pytest.raises(ValueError):
            formats.read_sign_cache(p)
        with pytest.raises(ValueError):
            formats.write_sign_cache(p, np.zeros(10, np.uint64), 960, 1024, 32, 0, 100)


def test_ss_query_record_is_48_bytes():
    import ctypes
    assert ctypes.sizeof(SSQuery) == 48
