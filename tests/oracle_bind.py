"""ctypes bindings for the CPU checkers under oracle/ (test infrastructure).

* ``Oracle`` wraps oracle/lib/liboracle.so, the C restatement of the
  reference algorithm (oracle/rabitq_oracle.c).
* ``Ref`` wraps oracle/_ref/librabitq_ref.so, the reference's own C-ABI bridge
  (experiments/02_diskann_fair/native/rabitq_bridge.cpp) compiled from its
  sources by oracle/Makefile. It exists only where it was built.

Nothing in the product package imports this module.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "lib", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "librabitq_ref.so")
SIGNSCAN_REF_SO = os.path.join(ROOT, "oracle", "_ref", "libsignscan_ref.so")

# Contraction sites of the reference compiler, pinned against oracle/_ref by
# tests/test_oracle_vs_ref.py (bit names follow oracle/rabitq_oracle.h).
FMA = dict(ENC_NORM=1 << 0, FQA_T=1 << 1, ENC_OOBAR=1 << 2, ENC_DECNORM=1 << 3,
           ENC_X4IP=1 << 4, C1_LANE=1 << 5, C1_SUB7=1 << 6, C1_EST=1 << 7,
           GATE_UPPER=1 << 8, RES_HI=1 << 9, RES_LO=1 << 10, RES_ACC=1 << 11,
           PAPER_LOWER=1 << 12, KMEANS_DIST=1 << 13, SS_NORM2=1 << 14, SS_B1CONST=1 << 15,
           SS_INTCONST=1 << 16, SS_B1DIST=1 << 17, SS_INTDIST=1 << 18, B1_SUM_DB=1 << 19,
           B1_SUM_Q=1 << 20, B1_UPPER=1 << 21)


class ROSpace(C.Structure):
    _fields_ = [("dim", C.c_size_t), ("code_dim", C.c_size_t), ("block_size", C.c_size_t),
                ("block_count", C.c_size_t), ("scale_bytes", C.c_size_t),
                ("compact_code_bytes", C.c_size_t), ("residual_code_bytes", C.c_size_t),
                ("compact_bytes", C.c_size_t), ("residual_bytes", C.c_size_t),
                ("full_bytes", C.c_size_t), ("seed", C.c_uint32), ("mse", C.c_int),
                ("fp16", C.c_int), ("inv_sqrt_code_dim", C.c_float),
                ("signs", C.POINTER(C.c_float)), ("center", C.POINTER(C.c_float)),
                ("rotated_center", C.POINTER(C.c_float)), ("center_norm_sqr", C.c_float)]


class ROQuery(C.Structure):
    _fields_ = [("q", C.POINTER(C.c_float)), ("x8", C.POINTER(C.c_int8)),
                ("query_norm", C.c_float), ("query_norm_sqr", C.c_float),
                ("half_sum", C.c_float), ("sum", C.c_float), ("positive_sum", C.c_float),
                ("int8_scale", C.c_float), ("int8_error_norm", C.c_float),
                ("int8_code_sum", C.c_int64), ("b1_short_scale", C.c_float),
                ("b1_error_scale", C.c_float), ("b1_alpha", C.c_float)]


class SSQuery(C.Structure):
    """Scalars of one sign-scan query (ro_ss_query / qg_signscan_query), 48 bytes."""
    _fields_ = [("fp32_sum", C.c_float), ("fp32_norm2", C.c_float), ("b1_scale", C.c_float),
                ("b1_constant", C.c_float), ("int4_scale", C.c_float), ("int4_constant", C.c_float),
                ("int8_scale", C.c_float), ("int8_constant", C.c_float), ("int4_sum", C.c_int64),
                ("int8_sum", C.c_int64)]


SS_ROWS = 13  # bit rows per query: B1 code, 4 INT4 planes, 8 INT8 planes


class Estimate(C.Structure):
    """RabitqPaperEstimate (rabitq_bridge.cpp:21-28), 24 bytes."""
    _fields_ = [("lower_bound", C.c_float), ("short_ip", C.c_float), ("alpha", C.c_float),
                ("ip_hat", C.c_float), ("error_bound", C.c_float), ("valid", C.c_uint8)]


class Interval(C.Structure):
    """RabitqDistanceInterval (rabitq_bridge.cpp:15-19), 12 bytes."""
    _fields_ = [("estimate", C.c_float), ("lower_bound", C.c_float), ("upper_bound", C.c_float)]


class SearchStats(C.Structure):
    _fields_ = [("hops", C.c_uint64), ("gate_evals", C.c_uint64),
                ("full_dists", C.c_uint64), ("pruned", C.c_uint64),
                ("recheck_checked", C.c_uint64), ("recheck_pruned", C.c_uint64)]

    def as_tuple(self):
        return (self.hops, self.gate_evals, self.full_dists, self.pruned, self.recheck_checked,
                self.recheck_pruned)

    def reference_counters(self):
        """The reference's counters (ours_diskann.rs:1604-1968, the same
        values as vamana::SearchStats of search_vamana_paper_active) from
        these: (visited_nodes, distance_computations, hops, paper_checked,
        paper_would_prune, paper_not_pruned, paper_msb_kernel_calls,
        paper_remaining_kernel_calls)."""
        dc = self.full_dists - self.recheck_pruned
        first_pruned = self.pruned - self.recheck_pruned
        return (1 + self.gate_evals, dc, self.hops, first_pruned + self.recheck_checked, self.pruned,
                self.recheck_checked - self.recheck_pruned, self.gate_evals, dc)


class SearchEvent(C.Structure):
    """ro_search_event: one prune decision of the oracle's search."""
    _fields_ = [("hop", C.c_uint32), ("id", C.c_uint32), ("lower_bound", C.c_float), ("back", C.c_float),
                ("stage", C.c_uint8), ("full", C.c_uint8), ("valid", C.c_uint8), ("pruned", C.c_uint8)]


SEARCH_EVENT_DTYPE = np.dtype([("hop", "<u4"), ("id", "<u4"), ("lower_bound", "<f4"), ("back", "<f4"),
                               ("stage", "u1"), ("full", "u1"), ("valid", "u1"), ("pruned", "u1")])


class RefStats(C.Structure):
    """SrStats of oracle/search_ref.cpp: the reference's search counters."""
    _fields_ = [(n, C.c_uint64) for n in ("visited_nodes", "distance_computations", "hops", "paper_checked",
                                          "paper_would_prune", "paper_not_pruned", "paper_msb_kernel_calls",
                                          "paper_remaining_kernel_calls")]

    def as_tuple(self):
        return tuple(getattr(self, n) for n, _ in self._fields_)


ESTIMATE_DTYPE = np.dtype([("lower_bound", "<f4"), ("short_ip", "<f4"), ("alpha", "<f4"),
                           ("ip_hat", "<f4"), ("error_bound", "<f4"), ("valid", "u1"),
                           ("pad", "u1", 3)])
INTERVAL_DTYPE = np.dtype([("estimate", "<f4"), ("lower_bound", "<f4"), ("upper_bound", "<f4")])
FACTOR_DTYPE = np.dtype([("norm_sqr", "<f4"), ("data_norm", "<f4"), ("cross_scale", "<f4"),
                         ("error_cross_scale", "<f4"), ("valid", "u1"), ("pad", "u1", 3)])


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """The C restatement, one space per instance (K = 1, residual 4-bit)."""

    def __init__(self, dim, seed, block=16, mse=True, fp16=True, fma_mask=None):
        self.lib = C.CDLL(ORACLE_SO)
        L = self.lib
        L.ro_space_init.argtypes = [C.POINTER(ROSpace), C.c_size_t, C.c_uint32, C.c_size_t, C.c_int, C.c_int]
        L.ro_set_fma_mask.argtypes = [C.c_uint]
        L.ro_asym_distance.restype = C.c_float
        L.ro_asym_distance_f64.restype = C.c_double
        L.ro_remaining_distance_fp32.restype = C.c_float
        L.ro_remaining_distance_fp32.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float]
        L.ro_remaining_distance_int8.restype = C.c_float
        L.ro_remaining_distance_int8.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float]
        L.ro_gate_fp32.restype = Estimate
        L.ro_gate_fp32.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float]
        L.ro_gate_int8.restype = Estimate
        L.ro_gate_int8.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float]
        L.ro_gate_b1.restype = Estimate
        L.ro_gate_b1.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float]
        L.ro_int8_full_estimate.restype = C.c_float
        L.ro_int8_full_estimate.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float,
                                            C.POINTER(C.c_float)]
        L.ro_mt32_nth.restype = C.c_uint32
        L.ro_mt32_nth.argtypes = [C.c_uint32, C.c_size_t]
        L.ro_mt64_nth.restype = C.c_uint64
        L.ro_mt64_nth.argtypes = [C.c_uint64, C.c_size_t]
        L.ro_int_codec_estimate.restype = Estimate
        L.ro_int_codec_estimate.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float]
        L.ro_masked_sum.restype = C.c_int64
        L.ro_masked_sum.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.ro_remaining_dot_int8.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.ro_full_dot_int8.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.ro_symmetric_ip4.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.ro_symmetric_distance.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.ro_asym_distance.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.ro_asym_distance_f64.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.ro_remaining_dot_int8.restype = C.c_int64
        L.ro_full_dot_int8.restype = C.c_int64
        L.ro_symmetric_ip4.restype = C.c_int64
        L.ro_symmetric_distance.restype = C.c_float
        L.ro_residual_interval.restype = Interval
        L.ro_residual_interval.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float]
        L.ro_residual_estimate_f64.restype = C.c_double
        L.ro_residual_estimate_f64.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double]
        L.ro_paper_estimate_record.restype = Estimate
        L.ro_paper_estimate_record.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_float]
        L.ro_b1_distance.restype = C.c_float
        L.ro_b1_distance.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.ro_b1_dot.restype = C.c_int64
        L.ro_b1_dot.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int64)]
        L.ro_train_centroids.argtypes = [C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t,
                                         C.c_uint32, C.c_int, C.c_void_p]
        L.ro_search.restype = C.c_size_t
        L.ro_search.argtypes = [C.c_void_p, C.c_void_p, C.c_int,
                                C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t,
                                C.c_void_p, C.c_size_t, C.c_void_p,
                                C.c_void_p, C.c_void_p, C.c_size_t,
                                C.c_size_t, C.c_size_t, C.c_size_t, C.c_float, C.c_size_t,
                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_void_p]
        L.ro_search_traced.restype = C.c_size_t
        L.ro_search_traced.argtypes = list(L.ro_search.argtypes) + [C.c_void_p, C.c_size_t, C.c_void_p]
        L.ro_robust_prune.restype = C.c_size_t
        L.ro_robust_prune.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_uint32,
                                      C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t,
                                      C.c_float, C.c_void_p, C.c_void_p]
        L.ro_build_vamana.restype = C.c_int
        L.ro_build_vamana.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t,
                                      C.c_float, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t,
                                      C.c_void_p, C.c_void_p, C.c_void_p]
        L.ro_fast_quantize_abs.argtypes = [C.c_size_t, C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]
        L.ro_float_to_half.restype = C.c_uint16
        L.ro_float_to_half.argtypes = [C.c_float]
        L.ro_half_to_float.restype = C.c_float
        L.ro_half_to_float.argtypes = [C.c_uint16]
        self.s = ROSpace()
        if L.ro_space_init(C.byref(self.s), dim, seed, block, int(mse), int(fp16)) != 0:
            raise ValueError("bad space")
        if fma_mask is not None:
            self.set_fma(fma_mask)
        self.dim, self.cd = dim, self.s.code_dim
        self.full_bytes = self.s.full_bytes
        self.compact_bytes = self.s.compact_bytes
        self.residual_bytes = self.s.residual_bytes
        self.msb_bytes = (self.cd + 7) // 8
        self.sp = C.byref(self.s)

    def __del__(self):
        try:
            self.lib.ro_space_free(C.byref(self.s))
        except Exception:
            pass

    def set_fma(self, mask):
        self.lib.ro_set_fma_mask(mask)

    def get_fma(self):
        """The library-wide contraction mask (shared by every Oracle)."""
        self.lib.ro_get_fma_mask.restype = C.c_uint
        return int(self.lib.ro_get_fma_mask())

    def signs(self):
        return np.ctypeslib.as_array(self.s.signs, (self.cd,)).copy()

    def set_center(self, c):
        c = np.ascontiguousarray(c, np.float32)
        self.lib.ro_space_set_center(self.sp, _p(c))

    def rotate(self, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros(self.cd, np.float32)
        self.lib.ro_rotate(self.sp, _p(x), _p(out))
        return out

    def encode_full(self, X):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, self.dim)
        out = np.zeros((X.shape[0], self.full_bytes), np.uint8)
        for i in range(X.shape[0]):
            self.lib.ro_encode_full(self.sp, _p(X[i]), _p(out[i]))
        return out

    def encode_full_rotated(self, X, Y):
        """Precision-mode records of X given Y = Q (X - c), the rotation
        GEMM's output (ro_encode_full_rotated)."""
        X = np.ascontiguousarray(X, np.float32).reshape(-1, self.dim)
        Y = np.ascontiguousarray(Y, np.float32).reshape(-1, self.cd)
        out = np.zeros((X.shape[0], self.full_bytes), np.uint8)
        for i in range(X.shape[0]):
            self.lib.ro_encode_full_rotated(self.sp, _p(X[i]), _p(Y[i]), _p(out[i]))
        return out

    def encode_from_unit(self, rn, rot, precision):
        """Records of the rotated unit vector rot of a residual of norm rn
        (ro_encode_from_unit)."""
        rot = np.ascontiguousarray(rot, np.float32)
        out = np.zeros(self.full_bytes, np.uint8)
        self.lib.ro_encode_from_unit(self.sp, C.c_float(rn), _p(rot), int(precision), _p(out))
        return out

    def prepare_rotated(self, raw, rq, rc, codec=3):
        """Query preparation from the rotated raw query rq and rotated
        center rc (ro_prepare_query_rotated)."""
        q = ROQuery()
        self.lib.ro_query_alloc(self.sp, C.byref(q))
        raw = np.ascontiguousarray(raw, np.float32)
        rq = np.ascontiguousarray(rq, np.float32)
        rc = np.ascontiguousarray(rc, np.float32)
        self.lib.ro_prepare_query_rotated(self.sp, _p(raw), _p(rq), _p(rc), codec, C.byref(q))
        q._raw = raw
        return q

    def split(self, full):
        n = full.shape[0]
        comp = np.zeros((n, self.compact_bytes), np.uint8)
        res = np.zeros((n, self.residual_bytes), np.uint8)
        for i in range(n):
            self.lib.ro_copy_compact(self.sp, _p(full[i]), _p(comp[i]))
            self.lib.ro_copy_residual(self.sp, _p(full[i]), _p(res[i]))
        return comp, res

    def sidecar(self, comp):
        n = comp.shape[0]
        msb = np.zeros((n, self.msb_bytes), np.uint8)
        fac = np.zeros(n, FACTOR_DTYPE)
        for i in range(n):
            self.lib.ro_extract_sidecar(self.sp, _p(comp[i]), _p(msb[i]),
                                        C.c_void_p(fac.ctypes.data + 20 * i))
        return msb, fac

    def fast_quantize_abs(self, a):
        a = np.ascontiguousarray(a, np.float32)
        code = np.zeros(a.size, np.uint8)
        ipn = C.c_float()
        self.lib.ro_fast_quantize_abs(a.size, _p(a), _p(code), C.byref(ipn))
        return code, ipn.value

    def prepare(self, raw, codec=3):
        q = ROQuery()
        self.lib.ro_query_alloc(self.sp, C.byref(q))
        raw = np.ascontiguousarray(raw, np.float32)
        self.lib.ro_prepare_query(self.sp, _p(raw), codec, C.byref(q))
        q._raw = raw
        return q

    def free_query(self, q):
        self.lib.ro_query_free(C.byref(q))

    def qvec(self, q):
        return np.ctypeslib.as_array(q.q, (self.cd,)).copy()

    def qx8(self, q):
        return np.ctypeslib.as_array(q.x8, (self.cd,)).copy()

    # ---- 1-bit sign-plane scan (query_codec_1bit_scan.cpp restated)
    def ss_encode(self, X):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, self.dim)
        db = np.zeros(X.shape[0] * (self.cd // 64), np.uint64)
        if self.lib.ro_ss_encode(self.sp, _p(X), C.c_uint64(X.shape[0]), _p(db)) != 0:
            raise ValueError("ro_ss_encode rejected n=%d" % X.shape[0])
        return db

    def ss_prepare(self, raw):
        raw = np.ascontiguousarray(raw, np.float32)
        q = np.zeros(self.cd, np.float32)
        bits = np.zeros(SS_ROWS * (self.cd // 64), np.uint64)
        qs = SSQuery()
        self.lib.ro_ss_prepare(self.sp, _p(raw), _p(q), _p(bits), C.byref(qs))
        return q, bits, qs

    def ss_scan(self, codec, db, n, q, bits, qs):
        out = np.zeros(n, np.float32)
        self.lib.ro_ss_scan(self.sp, C.c_int(codec), _p(db), C.c_uint64(n), _p(q), _p(bits), C.byref(qs),
                            _p(out))
        return out

    def asym(self, q, comp):
        return np.array([self.lib.ro_asym_distance(self.sp, C.byref(q), _p(comp[i]))
                         for i in range(comp.shape[0])], np.float32)

    def asym_f64(self, q, comp):
        return np.array([self.lib.ro_asym_distance_f64(self.sp, C.byref(q), _p(comp[i]))
                         for i in range(comp.shape[0])], np.float64)

    # ---- batched helpers over compact records / sidecars (rows of arrays)
    def _each(self, fn, rows):
        return [fn(rows[i]) for i in range(rows.shape[0])]

    def int8_full(self, q, comp, eps0=1.9):
        est = np.zeros(comp.shape[0], np.float32)
        low = np.zeros(comp.shape[0], np.float32)
        dot = np.zeros(comp.shape[0], np.int64)
        lo = C.c_float()
        for i in range(comp.shape[0]):
            est[i] = self.lib.ro_int8_full_estimate(self.sp, C.byref(q), _p(comp[i]), C.c_float(eps0), C.byref(lo))
            low[i] = lo.value
            dot[i] = self.lib.ro_full_dot_int8(self.sp, C.byref(q), _p(comp[i]))
        return est, low, dot

    def rem_int8(self, q, comp, short_ips):
        d = np.array([self.lib.ro_remaining_distance_int8(self.sp, C.byref(q), _p(comp[i]), C.c_float(short_ips[i]))
                      for i in range(comp.shape[0])], np.float32)
        dots = np.array([self.lib.ro_remaining_dot_int8(self.sp, C.byref(q), _p(comp[i]))
                         for i in range(comp.shape[0])], np.int64)
        return d, dots

    def rem_fp32(self, q, comp, short_ips):
        return np.array([self.lib.ro_remaining_distance_fp32(self.sp, C.byref(q), _p(comp[i]), C.c_float(short_ips[i]))
                         for i in range(comp.shape[0])], np.float32)

    def gate(self, q, msb, fac, codec, eps0=1.9):
        fn = self.lib.ro_gate_int8 if codec in (2, 3) else (self.lib.ro_gate_b1 if codec == 1 else self.lib.ro_gate_fp32)
        out = np.zeros(msb.shape[0], ESTIMATE_DTYPE)
        for i in range(msb.shape[0]):
            e = fn(self.sp, C.byref(q), _p(msb[i]), C.c_void_p(fac.ctypes.data + 20 * i), C.c_float(eps0))
            out[i] = (e.lower_bound, e.short_ip, e.alpha, e.ip_hat, e.error_bound, e.valid, (0, 0, 0))
        return out

    def paper_record(self, q, comp, eps0=1.9):
        out = np.zeros(comp.shape[0], ESTIMATE_DTYPE)
        for i in range(comp.shape[0]):
            e = self.lib.ro_paper_estimate_record(self.sp, C.byref(q), _p(comp[i]), C.c_float(eps0))
            out[i] = (e.lower_bound, e.short_ip, e.alpha, e.ip_hat, e.error_bound, e.valid, (0, 0, 0))
        return out

    def b1(self, q, comp):
        d = np.array([self.lib.ro_b1_distance(self.sp, C.byref(q), _p(comp[i])) for i in range(comp.shape[0])],
                     np.float32)
        ssum = C.c_int64()
        dots = np.array([self.lib.ro_b1_dot(self.sp, C.byref(q), _p(comp[i]), C.byref(ssum))
                         for i in range(comp.shape[0])], np.int64)
        return d, dots, ssum.value

    def train_centroids(self, data, k, sample_count, seed, iters):
        data = np.ascontiguousarray(data, np.float32)
        out = np.zeros((k, data.shape[1]), np.float32)
        rc = self.lib.ro_train_centroids(_p(data), data.shape[0], data.shape[1], k, sample_count, seed, iters,
                                         _p(out))
        return out if rc == 0 else None

    def masked_sum(self, q, msb):
        return np.array([self.lib.ro_masked_sum(self.sp, C.byref(q), _p(msb[i])) for i in range(msb.shape[0])],
                        np.int64)

    def residual(self, q, res, long_d):
        out = np.zeros(res.shape[0], INTERVAL_DTYPE)
        for i in range(res.shape[0]):
            iv = self.lib.ro_residual_interval(self.sp, C.byref(q), _p(res[i]), C.c_float(long_d[i]))
            out[i] = (iv.estimate, iv.lower_bound, iv.upper_bound)
        return out

    def residual_f64(self, q, res, long_d):
        return np.array([self.lib.ro_residual_estimate_f64(self.sp, C.byref(q), _p(res[i]), float(long_d[i]))
                         for i in range(res.shape[0])], np.float64)

    def symmetric(self, a, b):
        return self.lib.ro_symmetric_distance(self.sp, _p(a), _p(b))

    def symmetric_ip4(self, a, b):
        return self.lib.ro_symmetric_ip4(self.sp, _p(a[16:]), _p(b[16:]))

    def search(self, q, codec, comp, res, msb, fac, degrees, edges, k=10, L=100, eps0=1.9, rerank=100):
        n = comp.shape[0]
        ids = np.zeros(k, np.uint32)
        ds = np.zeros(k, np.float32)
        pool_ids = np.zeros(max(L, k) + 1, np.uint32)
        pool_d = np.zeros(max(L, k) + 1, np.float32)
        plen = C.c_size_t()
        st = SearchStats()
        got = self.lib.ro_search(self.sp, C.byref(q), codec, _p(comp), comp.shape[1], _p(res), res.shape[1],
                                 _p(msb), msb.shape[1], _p(fac), _p(degrees), _p(edges), edges.shape[1],
                                 n, k, L, C.c_float(eps0), rerank, _p(ids), _p(ds), _p(pool_ids), _p(pool_d),
                                 C.byref(plen), C.byref(st))
        return ids[:got], ds[:got], pool_ids[:plen.value], pool_d[:plen.value], st

    def search_traced(self, q, codec, comp, res, msb, fac, degrees, edges, k=10, L=100, eps0=1.9, rerank=100,
                      trace_cap=1 << 20):
        """search() plus the trace of every prune decision (SEARCH_EVENT_DTYPE)."""
        n = comp.shape[0]
        ids = np.zeros(k, np.uint32)
        ds = np.zeros(k, np.float32)
        pool_ids = np.zeros(max(L, k) + 1, np.uint32)
        pool_d = np.zeros(max(L, k) + 1, np.float32)
        plen = C.c_size_t()
        st = SearchStats()
        trace = np.zeros(trace_cap, SEARCH_EVENT_DTYPE)
        tlen = C.c_size_t()
        got = self.lib.ro_search_traced(self.sp, C.byref(q), codec, _p(comp), comp.shape[1], _p(res), res.shape[1],
                                        _p(msb), msb.shape[1], _p(fac), _p(degrees), _p(edges), edges.shape[1],
                                        n, k, L, C.c_float(eps0), rerank, _p(ids), _p(ds), _p(pool_ids),
                                        _p(pool_d), C.byref(plen), C.byref(st), _p(trace), trace_cap,
                                        C.byref(tlen))
        assert tlen.value <= trace_cap, "trace buffer too small"
        return (ids[:got], ds[:got], pool_ids[:plen.value], pool_d[:plen.value], st,
                trace[:tlen.value].copy())

    def robust_prune(self, comp, source, pool_ids, pool_d, r, alpha, with_evals=False):
        pool_ids = np.ascontiguousarray(pool_ids, np.uint32)
        pool_d = np.ascontiguousarray(pool_d, np.float32)
        out = np.zeros(r, np.uint32)
        ev = C.c_uint64()
        got = self.lib.ro_robust_prune(self.sp, _p(comp), comp.shape[1], source, _p(pool_ids), _p(pool_d),
                                       pool_ids.size, r, C.c_float(alpha), _p(out), C.byref(ev))
        return (out[:got], ev.value) if with_evals else out[:got]

    def build_graph(self, comp, r, l_build, alpha, beam=1, batch=4096, refine=0, cap=0, early_stop=0):
        n = comp.shape[0]
        deg = np.zeros(n, np.uint32)
        edges = np.zeros((n, r), np.uint32)
        ev = C.c_uint64()
        rc = self.lib.ro_build_vamana(self.sp, _p(comp), comp.shape[1], n, r, l_build, C.c_float(alpha), beam,
                                      batch, refine, cap, early_stop, _p(deg), _p(edges), C.byref(ev))
        if rc != 0:
            raise ValueError("ro_build_vamana failed: %d" % rc)
        return deg, edges, ev.value


class Ref:
    """The reference bridge compiled from source (oracle/_ref)."""

    @staticmethod
    def available():
        return os.path.exists(REF_SO)

    def __init__(self, dim, seed, centroid_count=1, residual_bits=4, block=16, mse=True, fp16=True):
        self.lib = C.CDLL(REF_SO)
        L = self.lib
        L.rabitq_space_new.restype = C.c_void_p
        L.rabitq_space_new.argtypes = [C.c_size_t, C.c_uint32, C.c_size_t, C.c_size_t,
                                       C.c_size_t, C.c_bool, C.c_bool]
        L.rabitq_space_delete.argtypes = [C.c_void_p]
        L.rabitq_space_set_center.restype = C.c_bool
        L.rabitq_space_set_center.argtypes = [C.c_void_p, C.c_void_p]
        for f in ("full", "compact", "residual"):
            getattr(L, "rabitq_%s_record_bytes" % f).restype = C.c_size_t
            getattr(L, "rabitq_%s_record_bytes" % f).argtypes = [C.c_void_p]
        L.rabitq_paper_msb_code_bytes.restype = C.c_size_t
        L.rabitq_paper_msb_code_bytes.argtypes = [C.c_void_p]
        L.rabitq_encode_full.restype = C.c_bool
        L.rabitq_encode_full.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.rabitq_copy_compact.restype = C.c_bool
        L.rabitq_copy_compact.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.rabitq_copy_residual_record.restype = C.c_bool
        L.rabitq_copy_residual_record.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.rabitq_export_paper_sidecar.restype = C.c_bool
        L.rabitq_export_paper_sidecar.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t,
                                                  C.c_void_p, C.c_void_p]
        L.rabitq_prepare_query.restype = C.c_void_p
        L.rabitq_prepare_query.argtypes = [C.c_void_p, C.c_void_p]
        L.rabitq_set_query_coarse_codec.argtypes = [C.c_void_p, C.c_int]
        L.rabitq_release_query.argtypes = [C.c_void_p, C.c_void_p]
        L.rabitq_query_distance.restype = C.c_float
        L.rabitq_query_distance.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.rabitq_symmetric_distance.restype = C.c_float
        L.rabitq_symmetric_distance.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.rabitq_query_distance_batch.restype = C.c_bool
        L.rabitq_query_distance_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                  C.c_size_t, C.c_void_p, C.c_size_t,
                                                  C.c_void_p, C.c_void_p]
        L.rabitq_paper_estimate_batch_sidecar.restype = C.c_bool
        L.rabitq_paper_estimate_batch_sidecar.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p,
                                                          C.c_size_t, C.c_void_p, C.c_size_t,
                                                          C.c_void_p, C.c_size_t, C.c_float,
                                                          C.c_void_p]
        L.rabitq_residual_distance_batch.restype = C.c_bool
        L.rabitq_residual_distance_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                                     C.c_void_p, C.c_size_t, C.c_void_p,
                                                     C.c_size_t, C.c_void_p, C.c_void_p]
        if hasattr(L, "rabitq_train_centroids"):
            self._bind_singles()
        if not hasattr(L, "rabitq_build_vamana_graph"):
            self._init_space(dim, seed, centroid_count, residual_bits, block, mse, fp16)
            return
        L.rabitq_build_vamana_graph.restype = C.c_bool
        L.rabitq_build_vamana_graph.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t,
                                                C.c_size_t, C.c_size_t, C.c_float, C.c_size_t,
                                                C.c_size_t, C.c_size_t, C.c_size_t, C.c_size_t,
                                                C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        self._init_space(dim, seed, centroid_count, residual_bits, block, mse, fp16)

    def _bind_singles(self):
        """The bridge's single-record and remaining batch entry points
        (rabitq_bridge.cpp:89-172, 380-700, 1044-1084)."""
        L, V, Z, F = self.lib, C.c_void_p, C.c_size_t, C.c_float
        sigs = {
            "rabitq_space_set_centroids": (C.c_bool, [V, V, Z]),
            "rabitq_train_centroids": (C.c_bool, [V, Z, Z, Z, Z, C.c_uint32, C.c_int, V]),
            "rabitq_query_distance_b1": (F, [V, V, V]),
            "rabitq_query_distance_b1_batch": (C.c_bool, [V, V, V, Z, V, Z, V]),
            "rabitq_paper_estimate": (Estimate, [V, V, V, F]),
            "rabitq_paper_estimate_sidecar": (Estimate, [V, V, V, V, F]),
            "rabitq_paper_estimate_codec": (Estimate, [V, V, V, V, V, F]),
            "rabitq_paper_estimate_batch": (C.c_bool, [V, V, V, Z, V, Z, F, V]),
            "rabitq_paper_estimate_codec_batch": (C.c_bool, [V, V, V, Z, V, Z, V, Z, V, Z, F, V]),
            "rabitq_query_distance_with_paper": (F, [V, V, V, F]),
            "rabitq_query_distance_with_quantized_paper": (F, [V, V, V, F]),
            "rabitq_residual_distance": (Interval, [V, V, V, F]),
            "rabitq_residual_distance_from_record": (Interval, [V, V, V, V, F]),
            "rabitq_full_residual_distance_batch": (C.c_bool, [V, V, V, Z, V, Z, V, Z, V]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args

    @staticmethod
    def _est_row(e):
        return (e.lower_bound, e.short_ip, e.alpha, e.ip_hat, e.error_bound, e.valid, (0, 0, 0))

    def set_centroids(self, c, count):
        c = np.ascontiguousarray(c, np.float32)
        return bool(self.lib.rabitq_space_set_centroids(self.h, _p(c), count))

    def train_centroids(self, data, k, sample_count, seed, iters):
        data = np.ascontiguousarray(data, np.float32)
        out = np.zeros((k, data.shape[1]), np.float32)
        ok = self.lib.rabitq_train_centroids(_p(data), data.shape[0], data.shape[1], k, sample_count, seed, iters,
                                             _p(out))
        return out if ok else None

    def b1(self, pq, comp_row):
        return self.lib.rabitq_query_distance_b1(self.h, pq, _p(comp_row))

    def b1_batch(self, pq, comp, ids):
        ids = np.ascontiguousarray(ids, np.uint32)
        out = np.zeros(ids.size, np.float32)
        assert self.lib.rabitq_query_distance_b1_batch(self.h, pq, _p(ids), ids.size, _p(comp), comp.shape[1],
                                                       _p(out))
        return out

    def paper_estimate(self, pq, comp_row, eps0):
        return self._est_row(self.lib.rabitq_paper_estimate(self.h, pq, _p(comp_row), C.c_float(eps0)))

    def paper_estimate_sidecar(self, pq, msb_row, fac, i, eps0):
        return self._est_row(self.lib.rabitq_paper_estimate_sidecar(
            self.h, pq, _p(msb_row), C.c_void_p(fac.ctypes.data + 20 * i), C.c_float(eps0)))

    def paper_estimate_codec(self, pq, comp_row, msb_row, fac, i, eps0):
        return self._est_row(self.lib.rabitq_paper_estimate_codec(
            self.h, pq, _p(comp_row), _p(msb_row), C.c_void_p(fac.ctypes.data + 20 * i), C.c_float(eps0)))

    def paper_estimate_batch(self, pq, comp, ids, eps0):
        ids = np.ascontiguousarray(ids, np.uint32)
        out = np.zeros(ids.size, ESTIMATE_DTYPE)
        assert self.lib.rabitq_paper_estimate_batch(self.h, pq, _p(ids), ids.size, _p(comp), comp.shape[1],
                                                    C.c_float(eps0), _p(out))
        return out

    def paper_estimate_codec_batch(self, pq, comp, msb, fac, ids, eps0):
        ids = np.ascontiguousarray(ids, np.uint32)
        out = np.zeros(ids.size, ESTIMATE_DTYPE)
        assert self.lib.rabitq_paper_estimate_codec_batch(self.h, pq, _p(ids), ids.size, _p(comp), comp.shape[1],
                                                          _p(msb), msb.shape[1], _p(fac), 20, C.c_float(eps0),
                                                          _p(out))
        return out

    def with_paper(self, pq, comp_row, short_ip, quantized=False):
        fn = self.lib.rabitq_query_distance_with_quantized_paper if quantized else \
            self.lib.rabitq_query_distance_with_paper
        return fn(self.h, pq, _p(comp_row), C.c_float(short_ip))

    def residual_single(self, pq, row, long_d):
        iv = self.lib.rabitq_residual_distance(self.h, pq, _p(row), C.c_float(long_d))
        return (iv.estimate, iv.lower_bound, iv.upper_bound)

    def residual_from_record(self, pq, comp_row, res_row, long_d):
        iv = self.lib.rabitq_residual_distance_from_record(self.h, pq, _p(comp_row), _p(res_row), C.c_float(long_d))
        return (iv.estimate, iv.lower_bound, iv.upper_bound)

    def full_residual_batch(self, pq, comp, res, ids):
        ids = np.ascontiguousarray(ids, np.uint32)
        out = np.zeros(ids.size, INTERVAL_DTYPE)
        assert self.lib.rabitq_full_residual_distance_batch(self.h, pq, _p(ids), ids.size, _p(comp), comp.shape[1],
                                                            _p(res), res.shape[1], _p(out))
        return out

    def _init_space(self, dim, seed, centroid_count, residual_bits, block, mse, fp16):
        L = self.lib
        self.h = L.rabitq_space_new(dim, seed, centroid_count, residual_bits, block, mse, fp16)
        if not self.h:
            raise ValueError("rabitq_space_new failed")
        self.dim = dim
        self.full_bytes = L.rabitq_full_record_bytes(self.h)
        self.compact_bytes = L.rabitq_compact_record_bytes(self.h)
        self.residual_bytes = L.rabitq_residual_record_bytes(self.h)
        self.msb_bytes = L.rabitq_paper_msb_code_bytes(self.h)

    def __del__(self):
        try:
            self.lib.rabitq_space_delete(self.h)
        except Exception:
            pass

    def set_center(self, c):
        c = np.ascontiguousarray(c, np.float32)
        assert self.lib.rabitq_space_set_center(self.h, _p(c))

    def set_codec(self, codec):
        self.lib.rabitq_set_query_coarse_codec(self.h, codec)

    def encode_full(self, X):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, self.dim)
        out = np.zeros((X.shape[0], self.full_bytes), np.uint8)
        for i in range(X.shape[0]):
            assert self.lib.rabitq_encode_full(self.h, _p(X[i]), _p(out[i]))
        return out

    def split(self, full):
        n = full.shape[0]
        comp = np.zeros((n, self.compact_bytes), np.uint8)
        res = np.zeros((n, self.residual_bytes), np.uint8)
        for i in range(n):
            assert self.lib.rabitq_copy_compact(self.h, _p(full[i]), _p(comp[i]))
            assert self.lib.rabitq_copy_residual_record(self.h, _p(full[i]), _p(res[i]))
        return comp, res

    def sidecar(self, comp):
        n = comp.shape[0]
        msb = np.zeros((n, self.msb_bytes), np.uint8)
        fac = np.zeros(n, FACTOR_DTYPE)
        assert self.lib.rabitq_export_paper_sidecar(self.h, _p(comp), n, comp.shape[1],
                                                    _p(msb), _p(fac))
        return msb, fac

    def prepare(self, raw):
        """The returned handle is thread-local storage: valid until the next
        prepare on this thread (space_rabitq.h:3700-3706)."""
        self._raw = np.ascontiguousarray(raw, np.float32)
        return self.lib.rabitq_prepare_query(self.h, _p(self._raw))

    def distance_batch(self, pq, comp, ids, modes, short_ips=None):
        ids = np.ascontiguousarray(ids, np.uint32)
        modes = np.ascontiguousarray(modes, np.uint8)
        if short_ips is None:
            short_ips = np.zeros(ids.size, np.float32)
        short_ips = np.ascontiguousarray(short_ips, np.float32)
        out = np.zeros(ids.size, np.float32)
        assert self.lib.rabitq_query_distance_batch(self.h, pq, _p(ids), _p(modes), ids.size,
                                                    _p(comp), comp.shape[1], _p(short_ips), _p(out))
        return out

    def gate_batch(self, pq, msb, fac, ids, eps0):
        ids = np.ascontiguousarray(ids, np.uint32)
        out = np.zeros(ids.size, ESTIMATE_DTYPE)
        assert self.lib.rabitq_paper_estimate_batch_sidecar(self.h, pq, _p(ids), ids.size,
                                                            _p(msb), msb.shape[1], _p(fac), 20,
                                                            C.c_float(eps0), _p(out))
        return out

    def residual_batch(self, pq, comp, res, ids, long_d):
        ids = np.ascontiguousarray(ids, np.uint32)
        long_d = np.ascontiguousarray(long_d, np.float32)
        out = np.zeros(ids.size, INTERVAL_DTYPE)
        assert self.lib.rabitq_residual_distance_batch(self.h, pq, _p(ids), ids.size, _p(comp),
                                                       comp.shape[1], _p(res), res.shape[1],
                                                       _p(long_d), _p(out))
        return out

    def symmetric(self, a, b):
        return self.lib.rabitq_symmetric_distance(self.h, _p(a), _p(b))

    def build_graph(self, comp, r, l_build, alpha, beam=1, batch=4096, refine=0, cap=0,
                    early_stop=0):
        n = comp.shape[0]
        deg = np.zeros(n, np.uint32)
        edges = np.zeros((n, r), np.uint32)
        msb = np.zeros((n, self.msb_bytes), np.uint8)
        fac = np.zeros(n, FACTOR_DTYPE)
        ms, fb, de = C.c_size_t(), C.c_size_t(), C.c_uint64()
        ok = self.lib.rabitq_build_vamana_graph(self.h, _p(comp), n, comp.shape[1], r, l_build,
                                                C.c_float(alpha), beam, batch, refine, cap,
                                                early_stop, _p(deg), _p(edges), r, _p(msb),
                                                _p(fac), C.byref(ms), C.byref(fb), C.byref(de))
        assert ok
        return deg, edges, msb, fac, de.value


REF_SEARCH_SO = os.path.join(ROOT, "oracle", "lib", "libref_search.so")


class RefSearch:
    """The Rust search loop restated in C (oracle/ref_search.c) driving the
    reference bridge's compiled kernels; one instance per Ref space."""

    _lib = None

    @staticmethod
    def available():
        return Ref.available() and os.path.exists(REF_SEARCH_SO)

    def __init__(self, ref):
        if RefSearch._lib is None:
            L = C.CDLL(REF_SEARCH_SO)
            L.rs_open.restype = C.c_int
            L.rs_open.argtypes = [C.c_char_p]
            if L.rs_open(REF_SO.encode()) != 0:
                raise OSError("cannot bind the reference bridge")
            L.rs_search_range.restype = C.c_int
            L.rs_search_range.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_int,
                                          C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t,
                                          C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t,
                                          C.c_size_t, C.c_size_t, C.c_float, C.c_size_t, C.c_void_p, C.c_void_p]
            L.rs_search_lists.restype = C.c_int
            L.rs_search_lists.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_int,
                                          C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t,
                                          C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t,
                                          C.c_size_t, C.c_void_p, C.c_float, C.c_size_t, C.c_void_p, C.c_void_p]
            L.rs_fixed_range.restype = C.c_int
            L.rs_fixed_range.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_void_p,
                                         C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_size_t,
                                         C.c_size_t, C.c_void_p]
            RefSearch._lib = L
        self.ref = ref

    def search(self, queries, codec, comp, res, msb, fac, degrees, edges, k=10, L=100, eps0=1.9, rerank=100,
               q0=0, q1=None, out=None):
        """Searches queries[q0:q1] into out = (ids [nq, k], dists [nq, k]);
        the space's codec must already be set to `codec`."""
        queries = np.ascontiguousarray(queries, np.float32)
        nq, dim = queries.shape
        q1 = nq if q1 is None else q1
        if out is None:
            out = (np.zeros((nq, k), np.uint32), np.zeros((nq, k), np.float32))
        rc = RefSearch._lib.rs_search_range(self.ref.h, _p(queries), dim, q0, q1, codec, _p(comp), comp.shape[1],
                                            _p(res), res.shape[1], _p(msb), msb.shape[1], _p(fac), fac.itemsize,
                                            _p(degrees), _p(edges), edges.shape[1], comp.shape[0], k, L,
                                            C.c_float(eps0), rerank, _p(out[0]), _p(out[1]))
        if rc != 0:
            raise RuntimeError("reference search failed")
        return out

    def fixed(self, queries, comp, res, cand, depth=100, k=10, q0=0, q1=None, out=None):
        """The plain fixed-candidate pipeline (mode-0 distances of each
        query's candidates, top-`depth` by (distance, id), residual rerank,
        top-k by (estimate, id)) for queries[q0:q1] into out [nq, k]; the
        space's codec must be Full."""
        queries = np.ascontiguousarray(queries, np.float32)
        cand = np.ascontiguousarray(cand, np.uint32)
        nq, dim = queries.shape
        q1 = nq if q1 is None else q1
        if out is None:
            out = np.zeros((nq, k), np.uint32)
        rc = RefSearch._lib.rs_fixed_range(self.ref.h, _p(queries), dim, q0, q1, _p(comp), comp.shape[1], _p(res),
                                           res.shape[1], _p(cand), cand.shape[1], depth, k, _p(out))
        if rc != 0:
            raise RuntimeError("reference fixed-candidate pipeline failed")
        return out

    def search_lists(self, queries, codec, comp, res, msb, fac, degrees, edges, lists, k=10, eps0=1.9, rerank=100,
                     q0=0, q1=None, out=None):
        """search() with query i's own list size lists[i] (uint32 [nq])."""
        queries = np.ascontiguousarray(queries, np.float32)
        lists = np.ascontiguousarray(lists, np.uint32)
        nq, dim = queries.shape
        q1 = nq if q1 is None else q1
        if out is None:
            out = (np.zeros((nq, k), np.uint32), np.zeros((nq, k), np.float32))
        rc = RefSearch._lib.rs_search_lists(self.ref.h, _p(queries), dim, q0, q1, codec, _p(comp), comp.shape[1],
                                            _p(res), res.shape[1], _p(msb), msb.shape[1], _p(fac), fac.itemsize,
                                            _p(degrees), _p(edges), edges.shape[1], comp.shape[0], k, _p(lists),
                                            C.c_float(eps0), rerank, _p(out[0]), _p(out[1]))
        if rc != 0:
            raise RuntimeError("reference search failed")
        return out


class RefSignScan:
    """The reference's 1-bit query-codec scan
    (experiments/04_query_codec_1bit_scan/query_codec_1bit_scan.cpp) compiled
    from source behind oracle/sign_scan_ref.cpp. Fixed at 960 -> 1024."""

    @staticmethod
    def available():
        return os.path.exists(SIGNSCAN_REF_SO)

    def __init__(self, seed=100):
        L = C.CDLL(SIGNSCAN_REF_SO)
        L.qcs_input_dim.restype = C.c_size_t
        L.qcs_code_dim.restype = C.c_size_t
        L.qcs_signs.argtypes = [C.c_uint32, C.c_void_p]
        L.qcs_build_database.argtypes = [C.c_char_p, C.c_uint64, C.c_uint32, C.c_void_p]
        L.qcs_build_database_cached.argtypes = [C.c_char_p, C.c_uint64, C.c_uint32, C.c_char_p, C.c_void_p]
        L.qcs_prepare.restype = C.c_void_p
        L.qcs_prepare.argtypes = [C.c_void_p, C.c_uint32]
        L.qcs_query_fields.argtypes = [C.c_void_p] + [C.c_void_p] * 6
        L.qcs_query_free.argtypes = [C.c_void_p]
        L.qcs_scan.argtypes = [C.c_int, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]
        L.qcs_validate.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p]
        self.lib, self.seed = L, seed
        self.dim, self.cd = L.qcs_input_dim(), L.qcs_code_dim()
        self.W = self.cd // 64

    def signs(self):
        out = np.zeros(self.cd, np.float32)
        self.lib.qcs_signs(self.seed, _p(out))
        return out

    def write_fvecs(self, X, tmpdir):
        X = np.ascontiguousarray(X, np.float32)
        path = os.path.join(str(tmpdir), "qcs_base.fvecs")
        rows = np.empty((X.shape[0], 1 + self.dim), np.float32)
        rows[:, 0] = np.array([self.dim], np.int32).view(np.float32)[0]
        rows[:, 1:] = X
        rows.tofile(path)
        return path

    def build_database(self, X, tmpdir, cache=None):
        """The reference's database build over X written as an fvecs file;
        with cache, through the reference's cache file at that path (loaded
        when its header matches, written otherwise)."""
        n = X.shape[0]
        path = self.write_fvecs(X, tmpdir)
        db = np.zeros(n * self.W, np.uint64)
        if cache is None:
            rc = self.lib.qcs_build_database(path.encode(), n, self.seed, _p(db))
        else:
            rc = self.lib.qcs_build_database_cached(path.encode(), n, self.seed, str(cache).encode(), _p(db))
        if rc != 0:
            raise RuntimeError("reference database build failed")
        return db

    def prepare(self, raw):
        """Returns (handle, q, bits [13*W], scalars dict) for one raw query."""
        raw = np.ascontiguousarray(raw, np.float32)
        h = self.lib.qcs_prepare(_p(raw), self.seed)
        q = np.zeros(self.cd, np.float32)
        bits = np.zeros(SS_ROWS * self.W, np.uint64)
        fs = np.zeros(8, np.float32)
        isum = np.zeros(2, np.int64)
        W = self.W
        self.lib.qcs_query_fields(h, _p(q), _p(bits[:W]), C.c_void_p(bits.ctypes.data + 8 * W),
                                  C.c_void_p(bits.ctypes.data + 8 * 5 * W), _p(fs), _p(isum))
        names = ("fp32_sum", "fp32_norm2", "b1_scale", "b1_constant", "int4_scale", "int4_constant",
                 "int8_scale", "int8_constant")
        sc = {k: np.float32(v) for k, v in zip(names, fs)}
        sc["int4_sum"], sc["int8_sum"] = int(isum[0]), int(isum[1])
        return h, q, bits, sc

    def free(self, h):
        self.lib.qcs_query_free(h)

    def scan(self, codec, db, n, h):
        out = np.full(n, np.nan, np.float32)
        self.lib.qcs_scan(codec, _p(db), n, h, _p(out))
        return out

    def validate(self, db, n, h):
        return self.lib.qcs_validate(_p(db), n, h) == 0


INDEXIO_REF_SO = os.path.join(ROOT, "oracle", "_ref", "libindexio_ref.so")


class RefIndexIO:
    """The reference's space-state and Vamana-index serialisers behind
    oracle/index_io_ref.cpp (oracle/_ref/libindexio_ref.so); spaces and
    records come from the bridge compiled into the same library."""

    @staticmethod
    def available():
        return os.path.exists(INDEXIO_REF_SO)

    def __init__(self):
        L = C.CDLL(INDEXIO_REF_SO)
        L.iio_save_state.argtypes = [C.c_size_t, C.c_uint32, C.c_bool, C.c_bool, C.c_void_p, C.c_char_p]
        L.iio_load_save_state.argtypes = [C.c_size_t, C.c_bool, C.c_bool, C.c_char_p, C.c_char_p]
        L.iio_load_save_state_k.argtypes = [C.c_size_t, C.c_size_t, C.c_bool, C.c_bool, C.c_char_p, C.c_char_p]
        L.iio_build_save_index.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_float,
                                           C.c_size_t, C.c_size_t, C.c_size_t, C.c_char_p]
        L.iio_load_save_index.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p]
        L.rabitq_space_new.restype = C.c_void_p
        L.rabitq_space_new.argtypes = [C.c_size_t, C.c_uint32, C.c_size_t, C.c_size_t, C.c_size_t, C.c_bool,
                                       C.c_bool]
        L.rabitq_space_delete.argtypes = [C.c_void_p]
        L.rabitq_space_set_center.argtypes = [C.c_void_p, C.c_void_p]
        L.rabitq_encode_full.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.rabitq_copy_compact.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        for f in ("rabitq_full_record_bytes", "rabitq_compact_record_bytes"):
            getattr(L, f).restype = C.c_size_t
            getattr(L, f).argtypes = [C.c_void_p]
        self.lib = L

    def save_state(self, dim, seed, path, center=None, mse=True, fp16=True):
        c = None if center is None else np.ascontiguousarray(center, np.float32)
        return self.lib.iio_save_state(dim, seed, mse, fp16, _p(c) if c is not None else None, str(path).encode())

    def load_save_state(self, dim, src, dst, mse=True, fp16=True, centroid_count=1):
        return self.lib.iio_load_save_state_k(dim, centroid_count, mse, fp16, str(src).encode(), str(dst).encode())

    def space(self, dim, seed, center):
        h = self.lib.rabitq_space_new(dim, seed, 1, 4, 16, True, True)
        c = np.ascontiguousarray(center, np.float32)
        self.lib.rabitq_space_set_center(h, _p(c))
        return h

    def compact_records(self, h, X):
        L = self.lib
        X = np.ascontiguousarray(X, np.float32)
        full = np.zeros(L.rabitq_full_record_bytes(h), np.uint8)
        comp = np.zeros((X.shape[0], L.rabitq_compact_record_bytes(h)), np.uint8)
        for i in range(X.shape[0]):
            L.rabitq_encode_full(h, _p(X[i]), _p(full))
            L.rabitq_copy_compact(h, _p(full), _p(comp[i]))
        return comp

    def build_save_index(self, h, comp, r, l_build, alpha, beam, batch, refine, path):
        comp = np.ascontiguousarray(comp, np.uint8)
        return self.lib.iio_build_save_index(h, _p(comp), comp.shape[0], r, l_build, C.c_float(alpha), beam, batch,
                                             refine, str(path).encode())

    def load_save_index(self, h, src, dst):
        return self.lib.iio_load_save_index(h, str(src).encode(), str(dst).encode())

    def free(self, h):
        self.lib.rabitq_space_delete(h)


SEARCH_REF_SO = os.path.join(ROOT, "oracle", "_ref", "libsearch_ref.so")


class RefGraphSearch:
    """The reference's own graph search behind oracle/search_ref.cpp
    (oracle/_ref/libsearch_ref.so), over a space made by the bridge compiled
    into the same library: VamanaIndex::searchKnnPlainThenResidualRerank
    (vamana_index.h:530-636) and vamana::search_vamana_paper_active
    (vamana_search.h:216-364)."""

    @staticmethod
    def available():
        return os.path.exists(SEARCH_REF_SO)

    def __init__(self, dim, seed, center):
        L = C.CDLL(SEARCH_REF_SO)
        L.rabitq_space_new.restype = C.c_void_p
        L.rabitq_space_new.argtypes = [C.c_size_t, C.c_uint32, C.c_size_t, C.c_size_t, C.c_size_t, C.c_bool,
                                       C.c_bool]
        L.rabitq_space_delete.argtypes = [C.c_void_p]
        L.rabitq_space_set_center.argtypes = [C.c_void_p, C.c_void_p]
        L.rabitq_set_query_coarse_codec.argtypes = [C.c_void_p, C.c_int]
        L.sr_index_open.restype = C.c_void_p
        L.sr_index_open.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p, C.c_size_t, C.c_float]
        L.sr_index_close.argtypes = [C.c_void_p]
        L.sr_index_knn.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_void_p]
        L.sr_pool.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p,
                              C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_float, C.c_void_p,
                              C.c_void_p, C.c_void_p, C.c_void_p]
        self.lib = L
        self.h = L.rabitq_space_new(dim, seed, 1, 4, 16, True, True)
        self.center = np.ascontiguousarray(center, np.float32)
        L.rabitq_space_set_center(self.h, _p(self.center))
        self.ix = None

    def __del__(self):
        if getattr(self, "ix", None):
            self.lib.sr_index_close(self.ix)
        if getattr(self, "h", None):
            self.lib.rabitq_space_delete(self.h)

    def set_codec(self, codec):
        self.lib.rabitq_set_query_coarse_codec(self.h, codec)

    def open_index(self, index_path, residual_path, n, eps0=1.9):
        self.ix = self.lib.sr_index_open(self.h, str(index_path).encode(), str(residual_path).encode(), n,
                                         C.c_float(eps0))
        if not self.ix:
            raise RuntimeError("the reference rejected the index or residual file")

    def knn(self, query, k, L, rerank=100):
        query = np.ascontiguousarray(query, np.float32)
        labels = np.zeros(k, np.uint64)
        d = np.zeros(k, np.float32)
        cnt = C.c_size_t()
        st = RefStats()
        if self.lib.sr_index_knn(self.ix, _p(query), k, L, rerank, _p(labels), _p(d), C.byref(cnt),
                                 C.byref(st)) != 0:
            raise RuntimeError("reference search failed")
        return labels[:cnt.value].astype(np.uint32), d[:cnt.value], st

    def pool(self, query, comp, msb, fac, degrees, edges, L, eps0=1.9):
        query = np.ascontiguousarray(query, np.float32)
        n = comp.shape[0]
        ids = np.zeros(max(L, 1) + 1, np.uint32)
        d = np.zeros(max(L, 1) + 1, np.float32)
        plen = C.c_size_t()
        st = RefStats()
        degrees = np.ascontiguousarray(degrees, np.uint32)
        edges = np.ascontiguousarray(edges, np.uint32)
        if self.lib.sr_pool(self.h, _p(query), _p(comp), comp.shape[1], _p(msb), msb.shape[1], _p(fac),
                            _p(degrees), _p(edges), edges.shape[1], n, L, C.c_float(eps0), _p(ids), _p(d),
                            C.byref(plen), C.byref(st)) != 0:
            raise RuntimeError("reference traversal failed")
        return ids[:plen.value], d[:plen.value], st
