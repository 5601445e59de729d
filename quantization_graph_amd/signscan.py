"""The fixed 1-bit database scanned by four query codecs on MI355X.

Counterpart of experiments/04_query_codec_1bit_scan/query_codec_1bit_scan.cpp:
a database that keeps only the sign plane of each rotated vector (cd bits,
128 bytes at cd = 1024) is scanned in full by a query in one of four forms:

* ``CODEC_FP32``: the rotated FP32 query, summed over the set bits;
* ``CODEC_B1``: the query's own sign code and mean |q|, by XOR popcount;
* ``CODEC_INT4`` / ``CODEC_INT8``: max-abs scalar-quantised queries as
  two's-complement bit planes, by AND popcounts.

Every codec returns ``cd + |q|^2 - 2 <q, s>`` in its own approximation, with
s the +-1 sign vector of the database row (the reference's distance forms,
:529-648). The reference scans one query at a time; ``SignScan.scan`` scores
a whole batch of prepared queries in one launch, reading the database once.

The database layout is the reference's (blocks of 32 vectors, word-major
inside a block), so ``formats.read_sign_cache`` / ``write_sign_cache``
exchange it with the reference's cache files.
"""
import ctypes as C

import torch

from . import _lib
from ._lib import check

CODEC_FP32 = 0
CODEC_B1 = 1
CODEC_INT4 = 2
CODEC_INT8 = 3
CODECS = {"fp32": CODEC_FP32, "b1": CODEC_B1, "int4": CODEC_INT4, "int8": CODEC_INT8}
BATCH = 32   # vectors per database block
ROWS = 13    # bit rows per query: B1 code, 4 INT4 planes, 8 INT8 planes
SCALAR_BYTES = 48


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class PreparedQueries:
    """A batch of prepared queries in HBM: ``q`` [nq, cd] f32, ``bits``
    [nq, 13, W] (u64 words held as int64), ``scalars`` [nq, 48] bytes
    (qg_signscan_query)."""

    def __init__(self, nq, cd, device="cuda"):
        self.nq, self.cd = nq, cd
        self.q = torch.empty((nq, cd), dtype=torch.float32, device=device)
        self.bits = torch.empty((nq, ROWS, cd // 64), dtype=torch.int64, device=device)
        self.scalars = torch.empty((nq, SCALAR_BYTES), dtype=torch.uint8, device=device)

    def scalar_fields(self):
        """The scalars as a dict of numpy arrays (host copy)."""
        import numpy as np
        dt = np.dtype([("fp32_sum", "<f4"), ("fp32_norm2", "<f4"), ("b1_scale", "<f4"),
                       ("b1_constant", "<f4"), ("int4_scale", "<f4"), ("int4_constant", "<f4"),
                       ("int8_scale", "<f4"), ("int8_constant", "<f4"), ("int4_sum", "<i8"),
                       ("int8_sum", "<i8")])
        rec = self.scalars.cpu().numpy().view(dt).reshape(-1)
        return {k: rec[k] for k in dt.names}


class SignScan:
    """The 1-bit scan over one space's rotation (its dim, code_dim and signs)."""

    def __init__(self, space):
        self.space = space
        self.dim, self.cd = space.dim, space.code_dim
        if self.cd % 64 != 0:
            raise ValueError("code_dim must be a multiple of 64")
        self.words = self.cd // 64

    def encode(self, x, out=None):
        """Database of the rows of x (device [n, dim] f32, n a multiple of 32):
        int64 tensor of n * W words in the reference's blocked layout."""
        x = x.to(device="cuda", dtype=torch.float32).contiguous()
        n = x.shape[0]
        if x.dim() != 2 or x.shape[1] != self.dim:
            raise ValueError("expected [n, %d] rows" % self.dim)
        if out is None:
            out = torch.empty(n * self.words, dtype=torch.int64, device="cuda")
        elif (out.dtype != torch.int64 or not out.is_cuda or out.numel() < n * self.words
              or not out.is_contiguous()):
            raise ValueError("out must be a contiguous int64 CUDA tensor of n * code_dim / 64 words")
        check(_lib.lib().qg_signscan_encode(self.space.handle, _ptr(x), n, _ptr(out), _stream()))
        return out

    def prepare(self, queries):
        queries = queries.to(device="cuda", dtype=torch.float32).contiguous()
        if queries.dim() != 2 or queries.shape[1] != self.dim:
            raise ValueError("expected [nq, %d] queries" % self.dim)
        pq = PreparedQueries(queries.shape[0], self.cd)
        check(_lib.lib().qg_signscan_prepare(self.space.handle, _ptr(queries), pq.nq, _ptr(pq.q), _ptr(pq.bits),
                                             _ptr(pq.scalars), _stream()))
        return pq

    def scan(self, codec, db, n, pq, out=None):
        """Distances [nq, n] of every prepared query to every database vector."""
        codec = CODECS.get(codec, codec)
        if db.numel() < n * self.words:
            raise ValueError("database holds fewer than %d vectors" % n)
        if db.dtype != torch.int64 or not db.is_cuda or not db.is_contiguous():
            raise ValueError("db must be a contiguous int64 CUDA tensor of database words")
        if out is None:
            out = torch.empty((pq.nq, n), dtype=torch.float32, device="cuda")
        elif (out.dtype != torch.float32 or not out.is_cuda or out.numel() < pq.nq * n
              or not out.is_contiguous()):
            raise ValueError("out must be a contiguous [nq, n] float32 CUDA tensor")
        check(_lib.lib().qg_signscan(self.space.handle, int(codec), _ptr(db), n, _ptr(pq.q), _ptr(pq.bits),
                                     _ptr(pq.scalars), pq.nq, _ptr(out), _stream()))
        return out
