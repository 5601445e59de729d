"""Writes tests/golden/signscan_ref.npz from the reference's 1-bit
query-codec scan (experiments/04_query_codec_1bit_scan/
query_codec_1bit_scan.cpp) compiled from source behind
oracle/sign_scan_ref.cpp (oracle/_ref/libsignscan_ref.so, built by
oracle/Makefile in the container that holds the reference checkout).

The fixture holds inputs and the reference's outputs only, at the
reference's fixed sizes (input dimension 960, code dimension 1024, seed 100):
  * 96 base vectors (three blocks of 32) of mixed scales -> the sign-plane
    database words in the reference's blocked layout;
  * 5 queries (scaled Gaussian rows, one all-zero row, one all-negative
    row) -> the rotated FP32 query, the B1 code, the INT4 / INT8 bit planes,
    the eight float scalars and two integer sums of prepare_query;
  * the four codecs' distances of every query to every database vector
    (run_codec).

Run: python tests/golden/make_signscan_golden.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_bind import RefSignScan  # noqa: E402

NAMES = ("fp32_sum", "fp32_norm2", "b1_scale", "b1_constant", "int4_scale", "int4_constant", "int8_scale",
         "int8_constant")


def inputs():
    rng = np.random.default_rng(2026)
    x = (rng.standard_normal((96, 960)) * rng.uniform(0.05, 4.0, (96, 1))).astype(np.float32)
    q = (rng.standard_normal((5, 960)) * rng.uniform(0.1, 3.0, (5, 1))).astype(np.float32)
    q[3] = 0.0
    q[4] = -np.abs(q[4])
    return x, q


def main():
    ref = RefSignScan(100)
    x, q = inputs()
    with tempfile.TemporaryDirectory() as td:
        db = ref.build_database(x, td)
    out = dict(x=x, queries=q, db=db, seed=np.uint32(100))
    qrot, bits, fsc, isum, dist = [], [], [], [], []
    for i in range(q.shape[0]):
        h, qr, b, sc = ref.prepare(q[i])
        qrot.append(qr)
        bits.append(b)
        fsc.append([sc[k] for k in NAMES])
        isum.append([sc["int4_sum"], sc["int8_sum"]])
        dist.append([ref.scan(c, db, x.shape[0], h) for c in range(4)])
        ref.free(h)
    out.update(q_rot=np.array(qrot, np.float32), bits=np.array(bits, np.uint64),
               fscalars=np.array(fsc, np.float32), isums=np.array(isum, np.int64),
               dists=np.array(dist, np.float32))
    np.savez_compressed(os.path.join(HERE, "signscan_ref.npz"), **out)
    print("wrote", os.path.join(HERE, "signscan_ref.npz"), {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
