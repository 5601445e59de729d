/*
 * rabitq_oracle.h -- CPU restatement of the ExRaBitQ4 distance path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP
 * implementation is compared against. Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it; the product library
 * (quantization_graph_amd/csrc) never links, calls or falls back to it.
 *
 * Every function restates the algorithm of the reference header
 * Ours/core/hnswlib/space_rabitq.h (and the Rust search loop in
 * experiments/02_diskann_fair/src/ours_diskann.rs); the file:line each one
 * follows is cited at its definition in rabitq_oracle.c.
 *
 * Floating-point order: sums are accumulated in the lane order of the
 * reference's AVX-512 / AVX2 kernels (16-lane or 8-lane accumulators and
 * the 8/4/2/1 tree of _mm512_reduce_add_ps). Where the reference compiler
 * may contract a multiply and an add into one FMA, the oracle exposes the
 * choice as a bit of ro_fma_mask; the tests pin the mask against the
 * reference compiled from source (oracle/_ref).
 */
#ifndef RABITQ_ORACLE_H
#define RABITQ_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* FMA-contraction sites (bits of ro_fma_mask). */
#define RO_FMA_ENC_NORM    (1u << 0)  /* encode: residual_norm_sqr += r*r (float)         */
#define RO_FMA_FQA_T       (1u << 1)  /* fastQuantizeAbs: int(t*a + eps) (double)         */
#define RO_FMA_ENC_OOBAR   (1u << 2)  /* encode: o_obar += (rot*sign)*inv_sqrt_d          */
#define RO_FMA_ENC_DECNORM (1u << 3)  /* encode: decoded_norm += d*d (double)             */
#define RO_FMA_ENC_X4IP    (1u << 4)  /* encode: x4_ip += x4*d (double)                   */
#define RO_FMA_C1_LANE     (1u << 5)  /* asym: lane += code*q (float)                     */
#define RO_FMA_C1_SUB7     (1u << 6)  /* asym: (8s + r) - 7*half_sum                      */
#define RO_FMA_C1_EST      (1u << 7)  /* asym: (n + q) - ls*signed                        */
#define RO_FMA_GATE_UPPER  (1u << 8)  /* gate: cross + (eps*ecs)*qn                       */
#define RO_FMA_RES_HI      (1u << 9)  /* residual: products = fma(hi, odd, lo*even)       */
#define RO_FMA_RES_LO      (1u << 10) /* residual: products = fma(lo, even, hi*odd)       */
#define RO_FMA_RES_ACC     (1u << 11) /* residual: ips += scale*block_ip                  */
#define RO_FMA_PAPER_LOWER (1u << 12) /* record gate: (n + q) - two_norms*ip_upper         */
#define RO_FMA_KMEANS_DIST (1u << 13) /* k-means: d += diff*diff                          */
#define RO_FMA_SS_NORM2    (1u << 14) /* sign scan query: norm2 += v*v                    */
#define RO_FMA_SS_B1CONST  (1u << 15) /* sign scan query: cd*(1 + s*s)                    */
#define RO_FMA_SS_INTCONST (1u << 16) /* sign scan query: cd + (scale*scale)*code_norm2   */
#define RO_FMA_SS_B1DIST   (1u << 17) /* sign scan B1: constant - (2*scale)*signed_dot     */
#define RO_FMA_SS_INTDIST  (1u << 18) /* sign scan INT4/INT8: constant - (2*scale)*dot     */
#define RO_FMA_B1_SUM_DB   (1u << 19) /* B1 gate: fma(ecs, qn, b1_error_scale*dn)          */
#define RO_FMA_B1_SUM_Q    (1u << 20) /* B1 gate: fma(b1_error_scale, dn, ecs*qn)          */
#define RO_FMA_B1_UPPER    (1u << 21) /* B1 gate: cross + eps*(db_err + query_err)          */
/* The pattern the reference compiler produces (g++ 11, -O3, AVX-512). */
#define RO_FMA_REFERENCE (RO_FMA_FQA_T | RO_FMA_ENC_OOBAR | RO_FMA_ENC_DECNORM | RO_FMA_ENC_X4IP | \
                          RO_FMA_C1_LANE | RO_FMA_C1_SUB7 | RO_FMA_C1_EST | RO_FMA_RES_HI | \
                          RO_FMA_RES_ACC | RO_FMA_PAPER_LOWER | RO_FMA_KMEANS_DIST | \
                          RO_FMA_SS_B1CONST | RO_FMA_SS_INTCONST | RO_FMA_SS_B1DIST | RO_FMA_SS_INTDIST | \
                          RO_FMA_B1_SUM_DB)

extern unsigned ro_fma_mask;
void ro_set_fma_mask(unsigned mask);
unsigned ro_get_fma_mask(void);

/* Space configuration (reference: RaBitQSpace ctor, space_rabitq.h:2836-2889;
 * bridge rabitq_space_new, rabitq_bridge.cpp:42-74). K = 1, external
 * residual storage, residual bits 4, block 16. */
typedef struct ro_space {
    size_t dim;
    size_t code_dim;
    size_t block_size;
    size_t block_count;
    size_t scale_bytes;        /* block_count * (fp16 ? 2 : 4) */
    size_t compact_code_bytes; /* code_dim / 2 */
    size_t residual_code_bytes;
    size_t compact_bytes;      /* 16 + code_dim/2 + 1 */
    size_t residual_bytes;     /* 8 + scale_bytes + code_dim/2 */
    size_t full_bytes;
    uint32_t seed;
    int mse;
    int fp16;
    float inv_sqrt_code_dim;
    float *signs;              /* code_dim */
    float *center;             /* dim */
    float *rotated_center;     /* code_dim */
    float center_norm_sqr;
} ro_space;

uint32_t ro_mt32_nth(uint32_t seed, size_t n);
uint64_t ro_mt64_nth(uint64_t seed, size_t n);
int ro_space_init(ro_space *s, size_t dim, uint32_t seed, size_t block, int mse, int fp16);
void ro_space_free(ro_space *s);
void ro_space_set_center(ro_space *s, const float *center);
size_t ro_code_dim(size_t dim);

void ro_rotate(const ro_space *s, const float *x, float *out);
void ro_fast_quantize_abs(size_t code_dim, const float *abs_unit, uint8_t *code, float *ip_norm);
uint16_t ro_float_to_half(float v);
float ro_half_to_float(uint16_t h);
/* Precision mode (QR rotation): encode of x given y = Q (x - c) from the
 * rotation GEMM, and query preparation given the rotated raw query rq and
 * the rotated center rc. */
void ro_encode_full_rotated(const ro_space *s, const float *x, const float *y, uint8_t *out);
/* The codec after the rotation, shared by both encodes: records of the
 * rotated unit vector rot (cd floats) of a residual of norm rn;
 * precision = 1 takes the short / error factors from compensated FP32 sums. */
void ro_encode_from_unit(const ro_space *s, float rn, const float *rot, int precision, uint8_t *out);
void ro_encode_full(const ro_space *s, const float *x, uint8_t *full_out);
void ro_copy_compact(const ro_space *s, const uint8_t *full, uint8_t *compact_out);
void ro_copy_residual(const ro_space *s, const uint8_t *full, uint8_t *residual_out);

/* Paper-prune sidecar (space_rabitq.h:4364-4392). */
typedef struct ro_factors {
    float norm_sqr;
    float data_norm;
    float cross_scale;
    float error_cross_scale;
    uint8_t valid;
    uint8_t pad[3];
} ro_factors;
void ro_extract_sidecar(const ro_space *s, const uint8_t *compact, uint8_t *msb_out, ro_factors *f);

/* Prepared query (QueryContext, space_rabitq.h:101-132). */
typedef struct ro_query {
    float *q;          /* rotated residual, code_dim */
    int8_t *x8;        /* INT8 code, code_dim */
    float query_norm;
    float query_norm_sqr;
    float half_sum;
    float sum;
    float positive_sum;
    float int8_scale;      /* INT8 or INT4 scale of the prepared integer code */
    float int8_error_norm; /* its error norm */
    int64_t int8_code_sum; /* its code sum (B1: the signed sum of the sign code) */
    float b1_short_scale;  /* B1 / B1Main: the sign code's scale factors */
    float b1_error_scale;
    float b1_alpha;
} ro_query;
void ro_query_alloc(const ro_space *s, ro_query *q);
void ro_query_free(ro_query *q);
/* codec: 0 = Full, 1 = B1, 2 = Int4, 3 = Int8, 4 = B1Main (QueryCoarseCodec,
 * space_rabitq.h:36-44). Int4 and Int8 fill x8 with the dithered code, B1
 * and B1Main with the +-1 sign code. */
void ro_prepare_query(const ro_space *s, const float *raw, int codec, ro_query *q);
void ro_prepare_query_rotated(const ro_space *s, const float *raw, const float *rq, const float *rc, int codec,
                              ro_query *q);

typedef struct ro_estimate {
    float lower_bound;
    float short_ip;
    float alpha;
    float ip_hat;
    float error_bound;
    uint8_t valid;
    uint8_t pad[3];
} ro_estimate;

typedef struct ro_interval {
    float estimate;
    float lower;
    float upper;
} ro_interval;

float ro_asym_distance(const ro_space *s, const ro_query *q, const uint8_t *compact);
double ro_asym_distance_f64(const ro_space *s, const ro_query *q, const uint8_t *compact);
float ro_remaining_distance_fp32(const ro_space *s, const ro_query *q, const uint8_t *compact, float short_ip);
ro_estimate ro_gate_fp32(const ro_space *s, const ro_query *q, const uint8_t *msb, const ro_factors *f, float eps0);
ro_estimate ro_gate_int8(const ro_space *s, const ro_query *q, const uint8_t *msb, const ro_factors *f, float eps0);
/* compute_b1_paper_prune_estimate_batch_by_id (space_rabitq.h:4607-4753):
 * the symmetric 1-bit gate of a B1 query. */
ro_estimate ro_gate_b1(const ro_space *s, const ro_query *q, const uint8_t *msb, const ro_factors *f, float eps0);
int64_t ro_masked_sum(const ro_space *s, const ro_query *q, const uint8_t *msb);
int64_t ro_remaining_dot_int8(const ro_space *s, const ro_query *q, const uint8_t *compact);
int64_t ro_full_dot_int8(const ro_space *s, const ro_query *q, const uint8_t *compact);
float ro_remaining_distance_int8(const ro_space *s, const ro_query *q, const uint8_t *compact, float short_ip);
float ro_int8_full_estimate(const ro_space *s, const ro_query *q, const uint8_t *compact, float eps0,
                            float *lower_out);
ro_estimate ro_int_codec_estimate(const ro_space *s, const ro_query *q, const uint8_t *compact, float eps0);
int64_t ro_symmetric_ip4(const ro_space *s, const uint8_t *lhs, const uint8_t *rhs);
float ro_symmetric_distance(const ro_space *s, const uint8_t *lhs, const uint8_t *rhs);
ro_interval ro_residual_interval(const ro_space *s, const ro_query *q, const uint8_t *residual, float long_distance);
/* compute_paper_prune_estimate (space_rabitq.h:3926-3959): the gate from a
 * compact record, short_ip over the nibble MSBs in pair-lane order. */
ro_estimate ro_paper_estimate_record(const ro_space *s, const ro_query *q, const uint8_t *compact, float eps0);
/* query_distance_b1 (space_rabitq.h:3742-3755): 4-bit code against the
 * query's sign code (+1 where q_i > 0, else -1). */
float ro_b1_distance(const ro_space *s, const ro_query *q, const uint8_t *compact);
int64_t ro_b1_dot(const ro_space *s, const ro_query *q, const uint8_t *compact, int64_t *signed_sum);
/* rabitq_train_centroids (rabitq_bridge.cpp:101-172): mt19937 sample with
 * replacement, first k samples as seeds, `iters` Lloyd rounds. Returns 0,
 * or -1 for the arguments the reference rejects. */
int ro_train_centroids(const float *data, size_t count, size_t dim, size_t k, size_t sample_count, uint32_t seed,
                       int iters, float *out);
double ro_residual_estimate_f64(const ro_space *s, const ro_query *q, const uint8_t *residual, double long_distance);

/* Beam search with the 1-bit gate + residual rerank
 * (ours_diskann.rs:1539-1911). Returns the number of results written. */
typedef struct ro_search_stats {
    uint64_t hops;
    uint64_t gate_evals;
    uint64_t full_dists;      /* start node + every distance computed */
    uint64_t pruned;          /* first-pass and re-check prunes */
    uint64_t recheck_checked; /* re-checks with the pool full and the gate valid */
    uint64_t recheck_pruned;  /* of those, pruned */
} ro_search_stats;
/* One prune decision of the search (trace of ro_search_traced): the
 * candidate, its gate bound, the pool back it was tested against, and
 * whether the pool was full, the gate valid and the candidate pruned.
 * stage 0 = first pass against the pool as it stood at the start of the
 * hop, 1 = re-check at insertion. */
typedef struct ro_search_event {
    uint32_t hop;
    uint32_t id;
    float lower_bound;
    float back;
    uint8_t stage, full, valid, pruned;
} ro_search_event;
size_t ro_search_traced(const ro_space *s, const ro_query *q, int codec,
                        const uint8_t *compact_base, size_t compact_stride,
                        const uint8_t *residual_base, size_t residual_stride,
                        const uint8_t *msb_base, size_t msb_stride,
                        const ro_factors *factors,
                        const uint32_t *degrees, const uint32_t *edges, size_t edge_stride,
                        size_t n_points, size_t k, size_t l_value, float eps0, size_t rerank,
                        uint32_t *out_ids, float *out_dists,
                        uint32_t *pool_ids_out, float *pool_dists_out, size_t *pool_len_out,
                        ro_search_stats *stats, ro_search_event *trace, size_t trace_cap, size_t *trace_len);
size_t ro_search(const ro_space *s, const ro_query *q, int codec,
                 const uint8_t *compact_base, size_t compact_stride,
                 const uint8_t *residual_base, size_t residual_stride,
                 const uint8_t *msb_base, size_t msb_stride,
                 const ro_factors *factors,
                 const uint32_t *degrees, const uint32_t *edges, size_t edge_stride,
                 size_t n_points, size_t k, size_t l_value, float eps0, size_t rerank,
                 uint32_t *out_ids, float *out_dists,
                 uint32_t *pool_ids_out, float *pool_dists_out, size_t *pool_len_out,
                 ro_search_stats *stats);

/* RobustPrune over a candidate pool (vamana_prune.h:41-156); *evals (may
 * be NULL) receives the distance_between calls it makes. */
size_t ro_robust_prune(const ro_space *s, const uint8_t *compact_base, size_t compact_stride,
                       uint32_t source, const uint32_t *pool_ids, const float *pool_dists,
                       size_t pool_len, size_t r, float alpha, uint32_t *out, uint64_t *evals);

/* Vamana build with the symmetric distance (rabitq_build_vamana_graph,
 * rabitq_bridge.cpp:224-303): batched inserts, grouped back-edges, refine
 * passes. edges has stride r; distance_evaluations as the reference counts
 * them. Returns 0, -1 on bad parameters, -2 on an out-of-range neighbour. */
int ro_build_vamana(const ro_space *s, const uint8_t *compact_base, size_t compact_stride, size_t n, size_t r,
                    size_t l_build, float alpha, size_t beam, size_t batch, size_t refine, size_t cap,
                    size_t early_stop, uint32_t *degrees, uint32_t *edges, uint64_t *distance_evaluations);

/* ------------------------------------------------ 1-bit sign-plane scan
 * experiments/04_query_codec_1bit_scan/query_codec_1bit_scan.cpp, general
 * in the code dimension (the reference fixes 960 -> 1024). The database is
 * the sign plane of each vector's rotation (ro_rotate of the space: signs,
 * padding, Walsh-Hadamard; no center, no scaling) in blocks of 32 vectors,
 * word-major inside a block: db[(block * W + word) * 32 + lane], W = cd/64.
 * A query carries its rotation q[cd], 13 bit rows of W words (row 0 the B1
 * sign code, rows 1-4 the INT4 planes, rows 5-12 the INT8 planes) and the
 * scalars below. Codecs: 0 FP32, 1 B1, 2 INT4, 3 INT8. */
typedef struct ro_ss_query {
    float fp32_sum, fp32_norm2, b1_scale, b1_constant;
    float int4_scale, int4_constant, int8_scale, int8_constant;
    int64_t int4_sum, int8_sum;
} ro_ss_query;
#define RO_SS_BATCH 32
#define RO_SS_ROWS 13
/* n must be a multiple of 32. */
int ro_ss_encode(const ro_space *s, const float *x, uint64_t n, uint64_t *db);
void ro_ss_prepare(const ro_space *s, const float *raw, float *q, uint64_t *bits, ro_ss_query *out);
/* distances of the first (n / 32) * 32 database vectors */
void ro_ss_scan(const ro_space *s, int codec, const uint64_t *db, uint64_t n, const float *q,
                const uint64_t *bits, const ro_ss_query *qs, float *out);

#ifdef __cplusplus
}
#endif
#endif
