"""Host-side mirror of the reference's ExRaBitQ4 space on MI355X.

``RabitqSpace`` plays the role of the Rust wrapper ``RabitqSpace``
(experiments/02_diskann_fair/src/ours_diskann.rs:244-330) and of
``hnswlib::RaBitQSpace`` behind the bridge: same constructor arguments
(``rabitq_space_new(dim, seed, K, residual_bits, block, mse, fp16)``,
rabitq_bridge.cpp:42-74), same record sizes, same codecs. Instead of one
query per call it works on device-resident batches: ``encode`` turns N base
vectors into a structure-of-arrays ``DeviceIndex``; ``prepare`` turns N
queries into ``DeviceQueries``; the estimator entry points score many
(query, candidate) pairs per launch.

PyTorch supplies device memory and the current HIP stream; all arithmetic
runs in libqgraph_hip.so.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import IndexDesc, QueryDesc, check

# QueryCoarseCodec (space_rabitq.h:36-44): the formal codecs Full and Int8, and
# the ablation codecs B1 (sign code), Int4 and B1Main
CODEC_FULL = 0
CODEC_B1 = 1
CODEC_INT4 = 2
CODEC_INT8 = 3
CODEC_B1_MAIN = 4
_CODECS = {"full": CODEC_FULL, "b1": CODEC_B1, "int4": CODEC_INT4, "int8": CODEC_INT8, "b1main": CODEC_B1_MAIN}
_CODECS.update({v: v for v in list(_CODECS.values())})

SCAN_INT8_FULL = 0
SCAN_ASYM = 1
SCAN_REM_FP32 = 2
SCAN_REM_INT8 = 3
SCAN_GATE_FP32 = 4
SCAN_GATE_INT8 = 5
SCAN_RESIDUAL = 6
SCAN_SYMMETRIC = 7
SCAN_PAPER_RECORD = 8
SCAN_B1 = 9
SCAN_RESIDUAL_RECORD = 10  # E1 in the per-record form (single records, spaces of several centroids)
SCAN_GATE_B1 = 11  # the B1 codec's sign-code gate (pairs; INT4 queries take SCAN_GATE_INT8)

# RabitqPaperEstimate (rabitq_bridge.cpp:21-28) and RabitqDistanceInterval
# (:15-19) as numpy record types.
ESTIMATE_DTYPE = np.dtype([("lower_bound", "<f4"), ("short_ip", "<f4"), ("alpha", "<f4"),
                           ("ip_hat", "<f4"), ("error_bound", "<f4"), ("valid", "u1"),
                           ("pad", "u1", 3)])
INTERVAL_DTYPE = np.dtype([("estimate", "<f4"), ("lower_bound", "<f4"), ("upper_bound", "<f4")])
QSCALAR_DTYPE = np.dtype([("query_norm", "<f4"), ("query_norm_sqr", "<f4"), ("half_sum", "<f4"),
                          ("sum", "<f4"), ("positive_sum", "<f4"), ("int8_scale", "<f4"),
                          ("int8_error_norm", "<f4"), ("max_abs_bits", "<i4"),
                          ("int8_code_sum", "<i8"), ("b1_alpha", "<f4"), ("pad", "<u4")])


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _dev(t, dtype):
    if not isinstance(t, torch.Tensor):
        a = np.asarray(t)
        if not a.flags.writeable:  # read-only views (memory-mapped files): torch needs its own copy
            a = a.copy()
        t = torch.as_tensor(a)
    return t.to(device="cuda", dtype=dtype).contiguous()


def code_dim_for(dim):
    """roundUp64 (space_rabitq.h:217-224): next power of two of dim rounded up to 64."""
    rounded = ((dim + 63) // 64) * 64
    p = 1
    while p < rounded:
        p <<= 1
    return p


class DeviceIndex:
    """Structure-of-arrays encoded base set resident in HBM.

    codes   uint8 [n, cd/2]   primary 4-bit codes (reference nibble layout)
    hdr     f32   [n, 2]      norm_sqr, long_scale
    sfac    f32   [n, 2]      short_scale, error_scale
    gate    uint8 [n, G]      gate rows (G = cd/8 rounded up to 16, plus 32):
                              the MSB plane of the codes, the paper-prune
                              factors, a copy of hdr and the sidecar validity,
                              read by the gate as one row; the views msb,
                              factors and fvalid below
    rscale  int16 [n, cd/16]  residual block scales (fp16 bit patterns)
    rcodes  uint8 [n, cd/2]   residual 4-bit two's-complement codes
    rfac    f32   [n, 2]      residual_norm_sqr, long_residual_ip
    cid     uint8 [n]         centroid id (0 in a space of one centroid)
    """

    def __init__(self, space, n):
        cd = space.code_dim
        dev = "cuda"
        self.space = space
        self.n = n
        self.codes = torch.empty((n, cd // 2), dtype=torch.uint8, device=dev)
        self.hdr = torch.empty((n, 2), dtype=torch.float32, device=dev)
        self.sfac = torch.empty((n, 2), dtype=torch.float32, device=dev)
        self.gate = torch.zeros((n, _lib.lib().qg_gate_row_bytes(space.handle)), dtype=torch.uint8, device=dev)
        self.rscale = torch.empty((n, cd // 16), dtype=torch.int16, device=dev)
        self.rcodes = torch.empty((n, cd // 2), dtype=torch.uint8, device=dev)
        self.rfac = torch.empty((n, 2), dtype=torch.float32, device=dev)
        self.cid = torch.zeros((n,), dtype=torch.uint8, device=dev)

    @property
    def _span(self):
        return self.gate.shape[1] - 32

    @property
    def msb(self):
        """MSB plane [n, cd/8] (a view of the gate rows)."""
        return self.gate[:, :self.space.code_dim // 8]

    @property
    def factors(self):
        """Paper-prune factors [n, 4] (a copy out of the gate rows)."""
        return self.gate[:, self._span:self._span + 16].contiguous().view(torch.float32)

    @property
    def fvalid(self):
        """Sidecar validity [n] (a view of the gate rows)."""
        return self.gate[:, self._span + 24]

    def desc(self):
        return IndexDesc(self.codes.data_ptr(), self.hdr.data_ptr(), self.sfac.data_ptr(),
                         self.gate.data_ptr(), self.rscale.data_ptr(), self.rcodes.data_ptr(),
                         self.rfac.data_ptr(), self.n, self.cid.data_ptr())

    def nbytes(self):
        return sum(t.numel() * t.element_size() for t in (
            self.codes, self.hdr, self.sfac, self.gate, self.rscale, self.rcodes, self.rfac))

    def to_records(self):
        """Reference AoS records: compact [n, 16+cd/2+1] and residual
        [n, 8+2*cd/16+cd/2] (copyCompactPayloadFromFull /
        copyResidualRecordFromFull, space_rabitq.h:6005-6029)."""
        sp = self.space
        comp = torch.empty((self.n, sp.compact_record_bytes), dtype=torch.uint8, device="cuda")
        res = torch.empty((self.n, sp.residual_record_bytes), dtype=torch.uint8, device="cuda")
        d = self.desc()
        check(_lib.lib().qg_pack_records(sp.handle, C.byref(d), _ptr(comp), comp.shape[1], _ptr(res),
                                         res.shape[1], _stream()))
        return comp, res

    @classmethod
    def from_records(cls, space, compact, residual=None):
        """Build the SoA index from reference AoS records (device or host)."""
        compact = _dev(compact, torch.uint8)
        n = compact.shape[0]
        ix = cls(space, n)
        if residual is not None:
            residual = _dev(residual, torch.uint8)
        else:
            ix.rscale.zero_()
            ix.rcodes.zero_()
            ix.rfac.zero_()
        d = ix.desc()
        check(_lib.lib().qg_unpack_records(space.handle, _ptr(compact), compact.shape[1],
                                           _ptr(residual), residual.shape[1] if residual is not None else 0,
                                           n, C.byref(d), _stream()))
        return ix


class DeviceQueries:
    """Prepared queries resident in HBM (QueryContext, space_rabitq.h:101-132)."""

    def __init__(self, space, nq, codec):
        cd = space.code_dim
        self.space = space
        self.nq = nq
        self.codec = codec
        # one context per centroid: row k * nq + i is query i's context for centroid k
        rows = nq * space.centroid_count
        self.rows = rows
        self.q = torch.empty((rows, cd), dtype=torch.float32, device="cuda")
        self.x8 = torch.empty((rows, cd), dtype=torch.int8, device="cuda")
        self.x8p = torch.empty((rows, cd), dtype=torch.int8, device="cuda")
        self.qs = torch.empty((rows, 48), dtype=torch.uint8, device="cuda")

    def desc(self):
        return QueryDesc(self.q.data_ptr(), self.x8.data_ptr(), self.x8p.data_ptr(),
                         self.qs.data_ptr(), self.rows)

    def scalars(self):
        """Per-query scalars as a numpy record array (QSCALAR_DTYPE)."""
        return self.qs.cpu().numpy().view(QSCALAR_DTYPE).reshape(self.rows)


class RabitqSpace:
    """One ExRaBitQ4 space: rabitq_space_new(dim, seed, K, bits, block, mse,
    fp16) (rabitq_bridge.cpp:42-74). The formal configuration of the
    reference runner is (dim, 100, 1, 4, 16, True, True)
    (ours_diskann.rs:277)."""

    def __init__(self, dim, seed=100, centroid_count=1, residual_bits=4, residual_block_size=16,
                 residual_mse=True, residual_fp16=True):
        L = _lib.lib()
        h = L.qg_space_new(dim, seed, centroid_count, residual_bits, residual_block_size,
                           bool(residual_mse), bool(residual_fp16))
        if not h:
            raise _lib.QGError(L.qg_last_error().decode())
        self.handle = C.c_void_p(h)
        self.dim = dim
        self.seed = seed
        self.code_dim = L.qg_space_code_dim(self.handle)
        self.block_count = self.code_dim // residual_block_size
        self.compact_record_bytes = 16 + self.code_dim // 2 + 1
        self.residual_record_bytes = 8 + 2 * self.block_count + self.code_dim // 2
        self.full_record_bytes = 24 + 2 * self.block_count + self.code_dim + 1
        self.paper_msb_code_bytes = (self.code_dim + 7) // 8
        self.paper_factor_bytes = 20
        self.residual_block_size = residual_block_size
        self.residual_mse = bool(residual_mse)
        self.residual_fp16 = bool(residual_fp16)
        self.codec = CODEC_INT8
        self.rotation = "fht"
        self.centroid_count = L.qg_space_centroid_count(self.handle)

    @classmethod
    def new_with_centroids(cls, dim, seed, centroid_count, centroids):
        """RabitqSpace::new_with_centroids (ours_diskann.rs:264-297)."""
        centroids = np.asarray(centroids, np.float32).reshape(-1)
        if centroids.size != dim * centroid_count:
            raise ValueError("ExRaBitQ centroids dimension mismatch: centroids=%d dim*K=%d"
                             % (centroids.size, dim * centroid_count))
        sp = cls(dim, seed, centroid_count, 4, 16, True, True)
        sp.set_center(centroids)
        return sp

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                _lib.lib().qg_space_delete(self.handle)
                self.handle = None
        except Exception:
            pass

    def set_center(self, center):
        """rabitq_space_set_center (rabitq_bridge.cpp:80-87)."""
        c = np.ascontiguousarray(np.asarray(center, np.float32).reshape(-1))
        if c.size != self.dim:
            raise ValueError("center has %d values, expected %d" % (c.size, self.dim))
        check(_lib.lib().qg_space_set_center(self.handle, c.ctypes.data_as(C.c_void_p)))
        self._centroids = c.reshape(1, self.dim).copy()

    def set_centroids(self, centroids):
        """rabitq_space_set_centroids (rabitq_bridge.cpp:89-99): K x dim
        centroids; records then carry their nearest centroid's id and queries
        one context per centroid."""
        c = np.ascontiguousarray(np.asarray(centroids, np.float32).reshape(-1))
        if c.size != self.dim * self.centroid_count:
            raise ValueError("centroids hold %d values, expected %d x %d" % (c.size, self.centroid_count, self.dim))
        check(_lib.lib().qg_space_set_centroids(self.handle, c.ctypes.data_as(C.c_void_p), self.centroid_count))
        self._centroids = c.reshape(self.centroid_count, self.dim).copy()

    def set_qr_rotation(self, seed=None, gaussian=None, bf16_parts=3):
        """Precision mode: rotate with Q of the FP64 QR factorisation (on the
        GPU) of a Gaussian matrix instead of the sign-flipped Walsh-Hadamard
        transform. The matrix is `gaussian` (code_dim x code_dim) or else
        numpy's default_rng(seed).standard_normal((code_dim, code_dim)).
        Encode and query preparation then run the rotation on the BF16
        matrix cores with each operand split in bf16_parts parts (2: about 16
        significant bits, 3: FP32 class) (qg_space_set_rotation_qr)."""
        if gaussian is None:
            if seed is None:
                raise ValueError("give a seed or a Gaussian matrix")
            gaussian = np.random.default_rng(seed).standard_normal((self.code_dim, self.code_dim))
        g = np.ascontiguousarray(np.asarray(gaussian, np.float64))
        if g.shape != (self.code_dim, self.code_dim):
            raise ValueError("the Gaussian matrix must be %d x %d" % (self.code_dim, self.code_dim))
        check(_lib.lib().qg_space_set_rotation_qr(self.handle, g.ctypes.data_as(C.c_void_p), int(bf16_parts)))
        self.rotation = "qr"

    def rotation_matrix(self):
        """Q of a QR-rotated space (code_dim x code_dim FP64, rotated = Q x)."""
        q = np.zeros((self.code_dim, self.code_dim), np.float64)
        check(_lib.lib().qg_space_rotation_qr(self.handle, q.ctypes.data_as(C.c_void_p)))
        return q

    def rotate(self, rows, subtract_center=False):
        """The space's rotation of device rows [n, dim] -> [n, code_dim]
        (subtract_center: Q (x - c), the product encode takes; QR spaces)."""
        import torch
        rows = _dev(rows, torch.float32)
        out = torch.empty((rows.shape[0], self.code_dim), dtype=torch.float32, device="cuda")
        check(_lib.lib().qg_rotate_rows(self.handle, _ptr(rows), rows.shape[0], 1 if subtract_center else 0,
                                        _ptr(out), _stream()))
        return out

    def host_constants(self):
        """(rotation signs [code_dim], center [dim]) as float32 host arrays."""
        signs = np.zeros(self.code_dim, np.float32)
        center = np.zeros(self.dim, np.float32)
        check(_lib.lib().qg_space_host_constants(self.handle, signs.ctypes.data_as(C.c_void_p),
                                                 center.ctypes.data_as(C.c_void_p)))
        return signs, center

    def save_state(self, path):
        """RaBitQSpace::saveState (space_rabitq.h:3139-3197): the EXRBTQ31
        file the reference's loadState reads."""
        from . import formats
        signs, center = self.host_constants()
        centroids = center.reshape(1, self.dim)
        if self.centroid_count > 1:
            centroids = getattr(self, "_centroids", None)
            if centroids is None or centroids.shape[0] != self.centroid_count:
                raise ValueError("save_state: the %d centroids are not set" % self.centroid_count)
        formats.write_space_state(path, self.dim, self.seed, centroids, signs, self.residual_block_size,
                                  self.residual_mse, self.residual_fp16)

    @classmethod
    def load_state(cls, path):
        """RaBitQSpace::loadState (space_rabitq.h:3200-3345) of an EXRBTQ31
        file: a space of the stored dimension, seed, centroid count, residual
        settings and centroids (the center when K = 1). Raises ValueError on
        rotation signs other than its seed's."""
        from . import formats
        head, seed, centroids, signs = formats.read_space_state(path)
        flags = head["flags"]
        k = head["centroid_count"]
        sp = cls(head["dim"], seed, k, head["residual_bits"], head["residual_block_size"], bool(flags & 4),
                 bool(flags & 8))
        want = formats.space_state_fields(head["dim"], k, head["residual_block_size"], bool(flags & 4),
                                          bool(flags & 8))
        if head != want:
            raise ValueError("%s: record layout differs from this space's" % path)
        if not np.array_equal(sp.host_constants()[0], signs):
            raise ValueError("%s: rotation signs are not those of seed %d" % (path, seed))
        if k == 1:
            sp.set_center(centroids[0])
        else:
            sp.set_centroids(centroids)
        return sp

    def set_query_coarse_codec(self, codec):
        """RabitqSpace::set_query_coarse_codec (ours_diskann.rs:299-312):
        'full' (0) or 'int8' (3)."""
        if codec not in _CODECS:
            raise ValueError("unsupported query coarse codec %r" % (codec,))
        self.codec = _CODECS[codec]

    # ----------------------------------------------------------- encode
    def encode(self, x, batch=None):
        """Encode base vectors x [n, dim] (float32) into a DeviceIndex."""
        x = _dev(x, torch.float32)
        n = x.shape[0]
        if x.shape[1] != self.dim:
            raise ValueError("x has dimension %d, expected %d" % (x.shape[1], self.dim))
        ix = DeviceIndex(self, n)
        if n == 0:
            return ix
        L = _lib.lib()
        rows = n if batch is None else max(1, min(n, batch))
        ws_bytes = L.qg_encode_workspace_bytes(self.handle, rows)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
        d = ix.desc()
        check(L.qg_encode(self.handle, _ptr(x), n, C.byref(d), _ptr(ws), ws_bytes, _stream()))
        return ix

    # ------------------------------------------------------------ query
    def prepare(self, queries, codec=None):
        """Prepare queries [nq, dim] (rabitq_prepare_query for a batch)."""
        codec = self.codec if codec is None else _CODECS[codec]
        qx = _dev(queries, torch.float32)
        if qx.dim() == 1:
            qx = qx.reshape(1, -1)
        nq = qx.shape[0]
        out = DeviceQueries(self, nq, codec)
        L = _lib.lib()
        ws = torch.empty(L.qg_prepare_workspace_bytes(nq), dtype=torch.uint8, device="cuda")
        d = out.desc()
        check(L.qg_prepare_queries(self.handle, _ptr(qx), nq, codec, C.byref(d), _ptr(ws), _stream()))
        return out

    # -------------------------------------------------------- estimators
    def _scan(self, kernel, fixed, index, queries, ids, qidx=None, per_query=0, n_pairs=0, eps0=1.9,
              short_ips=None, long_in=None, want=("f",)):
        n = ids.numel()
        out_f = torch.empty(ids.shape, dtype=torch.float32, device="cuda") if "f" in want else None
        out_f2 = torch.empty(ids.shape, dtype=torch.float32, device="cuda") if "f2" in want else None
        out_est = torch.empty((n, 24), dtype=torch.uint8, device="cuda") if "est" in want else None
        out_iv = torch.empty((n, 12), dtype=torch.uint8, device="cuda") if "iv" in want else None
        out_dot = torch.empty(ids.shape, dtype=torch.int64, device="cuda") if "dot" in want else None
        di = index.desc()
        dq = C.byref(queries.desc()) if queries is not None else None  # symmetric pairs take no query
        check(_lib.lib().qg_scan(self.handle, kernel, 1 if fixed else 0, C.byref(di), dq,
                                 _ptr(ids), _ptr(qidx), per_query, n_pairs, C.c_float(eps0),
                                 _ptr(short_ips), _ptr(long_in), _ptr(out_f), _ptr(out_f2),
                                 _ptr(out_est), _ptr(out_iv), _ptr(out_dot), _stream()))
        return out_f, out_f2, out_est, out_iv, out_dot

    def scan_fixed(self, kernel, index, queries, ids, eps0=1.9, long_in=None, want=("f",)):
        """Score fixed candidate lists ids [nq, P] for every prepared query."""
        ids = _dev(ids, torch.int32)
        if ids.dim() != 2 or ids.shape[0] != queries.nq:
            raise ValueError("ids must be [nq, per_query]")
        if long_in is not None:
            long_in = _dev(long_in, torch.float32)
        if self.centroid_count > 1:
            # each record is scored with its centroid's context (queryForEncoded,
            # space_rabitq.h:3404-3416); the 1-bit sidecar gates use centroid 0's
            nq = queries.nq
            if kernel == SCAN_RESIDUAL:  # the reference scores each record alone here
                kernel = SCAN_RESIDUAL_RECORD
            if kernel in (SCAN_GATE_FP32, SCAN_GATE_INT8, SCAN_GATE_B1):
                qidx = torch.arange(nq, dtype=torch.int32, device="cuda")[:, None].expand_as(ids)
            else:
                qidx = index.cid[ids.long()].int() * nq + torch.arange(nq, dtype=torch.int32, device="cuda")[:, None]
            out = self._scan(kernel, False, index, queries, ids.reshape(-1), qidx=qidx.contiguous().reshape(-1),
                             n_pairs=ids.numel(), eps0=eps0,
                             long_in=None if long_in is None else long_in.reshape(-1), want=want)
            return tuple(None if t is None else (t.reshape(ids.shape) if t.dim() == 1 else t) for t in out)
        return self._scan(kernel, True, index, queries, ids, per_query=ids.shape[1], eps0=eps0,
                          long_in=long_in, want=want)

    def scan_pairs(self, kernel, index, queries, qidx, ids, eps0=1.9, short_ips=None, long_in=None,
                   want=("f",)):
        """Score (qidx[i], ids[i]) pairs."""
        ids = _dev(ids, torch.int32).reshape(-1)
        qidx = _dev(qidx, torch.int32).reshape(-1)
        if qidx.numel() != ids.numel():
            raise ValueError("qidx and ids must have the same length")
        if short_ips is not None:
            short_ips = _dev(short_ips, torch.float32).reshape(-1)
        if long_in is not None:
            long_in = _dev(long_in, torch.float32).reshape(-1)
        return self._scan(kernel, False, index, queries, ids, qidx=qidx, n_pairs=ids.numel(),
                          eps0=eps0, short_ips=short_ips, long_in=long_in, want=want)

    # ------------------------------------------------------------ search
    def search(self, index, queries, degrees, edges, k=10, l_search=100, eps0=1.9, rerank=100,
               slots=None, want_pool=False, events=None, l_per_query=None, waves=4, check_graph=True,
               workspace=None):
        """Batched graph search (search_ours_paper_active_inner, ours_diskann.rs:1539-1911)
        over a graph given as degrees [n] and edges [n, R] (R <= 64), start
        node 0, with the query codec the queries were prepared with.

        l_per_query (int32 [nq] on the device) gives each query its own
        search list size, clamped to [k, l_search]; l_search is then the pool
        capacity. waves (2 or 4) sets the wavefronts per query workgroup.
        With check_graph=False and a preallocated `workspace` the call
        enqueues kernels only (no synchronisation), so it can be captured in a
        HIP graph; read the error flag later with `search_error(workspace)`.

        Returns (ids [nq, k] int32, estimates [nq, k], stats [nq, 6] int64 =
        hops, gate evaluations, distances computed, prunes, re-checks with
        the pool full, re-check prunes[, pool ids, pool distances [nq,
        max(l_search, k)]]); `search_counters` turns stats into the
        reference's counters. `events` = (start, end) CUDA events recorded on
        the launch stream around the search kernel."""
        degrees = _dev(degrees, torch.int32).reshape(-1)
        edges = _dev(edges, torch.int32)
        if degrees.numel() != index.n or edges.shape[0] != index.n:
            raise ValueError("graph does not match the index size")
        nq = queries.nq
        L = max(l_search, k)
        if slots is None:
            slots = max(1, min(nq, 4096))
        lq = None
        if l_per_query is not None:
            lq = _dev(l_per_query, torch.int32).reshape(-1)
            if lq.numel() != nq:
                raise ValueError("l_per_query needs one entry per query")
        ids = torch.empty((nq, k), dtype=torch.int32, device="cuda")
        dists = torch.empty((nq, k), dtype=torch.float32, device="cuda")
        stats = torch.zeros((nq, 6), dtype=torch.int64, device="cuda")
        pool_ids = torch.empty((nq, L), dtype=torch.int32, device="cuda") if want_pool else None
        pool_d = torch.empty((nq, L), dtype=torch.float32, device="cuda") if want_pool else None
        L_ = _lib.lib()
        ws_bytes = L_.qg_search_workspace_bytes(index.n, slots)
        ws = workspace if workspace is not None else torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
        if ws.numel() < ws_bytes:
            raise ValueError("search workspace too small: need %d bytes" % ws_bytes)
        di, dq = index.desc(), queries.desc()
        if events is not None:
            events[0].record(torch.cuda.current_stream())
        check(L_.qg_search_ex(self.handle, C.byref(di), C.byref(dq), queries.codec, _ptr(degrees), _ptr(edges),
                              edges.shape[1], k, l_search, _ptr(lq), C.c_float(eps0), rerank, slots, waves,
                              _ptr(ids), _ptr(dists), _ptr(stats), _ptr(pool_ids), _ptr(pool_d), _ptr(ws),
                              ws.numel(), _stream()))
        if events is not None:
            events[1].record(torch.cuda.current_stream())
        if check_graph:
            search_error(ws)
        if want_pool:
            return ids, dists, stats, pool_ids, pool_d
        return ids, dists, stats

    @staticmethod
    def search_workspace(n, slots):
        """Device workspace of a search over n points with `slots` concurrent queries."""
        return torch.empty(_lib.lib().qg_search_workspace_bytes(n, slots), dtype=torch.uint8, device="cuda")

    def build_graph(self, index, r=32, l_build=400, alpha=1.2, beam_width=1, batch=4096, refine_passes=0,
                    prune_cap=0, early_stop_hops=0):
        """Vamana graph over the index's codes with the symmetric distance
        (rabitq_build_vamana_graph, rabitq_bridge.cpp:224-303; defaults are
        the reference's: R 32, L_build 400, alpha 1.2, beam 1, batch 4096).

        Returns (degrees [n] int32, edges [n, r] int32 with -1 past each
        degree, distance_evaluations)."""
        n = index.n
        degrees = torch.empty(n, dtype=torch.int32, device="cuda")
        edges = torch.empty((n, r), dtype=torch.int32, device="cuda")
        L_ = _lib.lib()
        ws_bytes = L_.qg_build_workspace_bytes(n, r, l_build, batch, refine_passes)
        if ws_bytes == 0:
            raise _lib.QGError("build workspace sizing failed")
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
        evals = C.c_uint64()
        di = index.desc()
        check(L_.qg_build_graph(self.handle, C.byref(di), r, l_build, C.c_float(alpha), beam_width, batch,
                                refine_passes, prune_cap, early_stop_hops, _ptr(degrees), _ptr(edges),
                                C.byref(evals), _ptr(ws), ws_bytes, _stream()))
        del ws
        return degrees, edges, evals.value


def search_counters(stats):
    """The reference runner's per-query search counters from qg_search's
    stats rows ([..., 6] integers: hops, gate evaluations, distances
    computed, prunes, re-checks with the pool full and the gate valid,
    re-check prunes), following where ours_diskann.rs:1539-1986 counts them:
    visited_nodes is the start node plus every fresh neighbour (:1604, 1659);
    distance_computations and paper_remaining_kernel_calls count the
    distances that pass the re-check at insertion (:1960-1961);
    paper_checked counts first-pass prunes plus re-checks made with the pool
    full and the gate valid (:1720, 1947); paper_would_prune and
    paper_full_saved count prunes at both sites; paper_msb_kernel_calls the
    gate evaluations. Returns a dict of arrays (or numbers) of those names."""
    import numpy as np
    st = np.asarray(stats, dtype=np.int64)
    hops, gates, dists, pruned, rchecked, rpruned = (st[..., i] for i in range(6))
    computed = dists - rpruned
    return dict(hops=hops, visited_nodes=1 + gates, distance_computations=computed,
                paper_checked=pruned - rpruned + rchecked, paper_would_prune=pruned, paper_full_saved=pruned,
                paper_not_pruned=rchecked - rpruned, paper_msb_kernel_calls=gates,
                paper_remaining_kernel_calls=computed)


def search_error(workspace):
    """Raise if the last search that used `workspace` met an out-of-range
    neighbour (the flag the kernel sets; synchronises)."""
    if workspace[4:8].view(torch.int32).item():
        raise _lib.QGError("graph contains an out-of-range neighbour")


def topk(values, ids, k, want_pos=False):
    """Per-row top-k of values [nq, P] by (value, id); returns (ids, values,
    positions), each [nq, k] (positions in the row only with want_pos, else
    None). Rows shorter than k are padded with id 0xFFFFFFFF and +inf."""
    values = values.contiguous()
    ids = _dev(ids, torch.int32)
    nq, p = values.shape
    out_ids = torch.empty((nq, k), dtype=torch.int32, device="cuda")
    out_v = torch.empty((nq, k), dtype=torch.float32, device="cuda")
    out_pos = torch.empty((nq, k), dtype=torch.int32, device="cuda") if want_pos else None
    check(_lib.lib().qg_topk(_ptr(values), _ptr(ids), nq, p, k, _ptr(out_ids), _ptr(out_v),
                             _ptr(out_pos), _stream()))
    return out_ids, out_v, out_pos


def estimates_np(t):
    """Device [n, 24] estimate bytes -> numpy ESTIMATE_DTYPE array."""
    return t.cpu().numpy().reshape(-1).view(ESTIMATE_DTYPE)


def intervals_np(t):
    """Device [n, 12] interval bytes -> numpy INTERVAL_DTYPE array."""
    return t.cpu().numpy().reshape(-1).view(INTERVAL_DTYPE)
