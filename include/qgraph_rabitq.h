/*
 * qgraph_rabitq.h -- C-ABI of libqgraph_hip.so, the MI355X (gfx950)
 * implementation of the ExRaBitQ4 distance path.
 *
 * Two groups of entry points:
 *
 *  1. Device batch API (qg_*): many queries per call, device pointers,
 *     an explicit hipStream_t (passed as void*), structure-of-arrays index
 *     views. Status is an int (0 = success); qg_last_error() describes the
 *     last failure on the calling thread.
 *
 *  2. Bridge API (rabitq_*): the exact symbols, argument meanings, struct
 *     layouts and error sentinels of the reference C-ABI
 *     experiments/02_diskann_fair/native/rabitq_bridge.cpp, operating on
 *     host memory. Each call stages its inputs to the GPU, runs the same
 *     kernels as the batch API, and copies the results back. A caller
 *     linked against the reference bridge links against this library
 *     unchanged.
 *
 * Numerics: distances are bit-identical to the reference CPU path (same
 * accumulation order and fused multiply-adds); integer dots are exact.
 */
#ifndef QGRAPH_RABITQ_H
#define QGRAPH_RABITQ_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ POD */

/* rabitq_bridge.cpp:15-19 */
typedef struct RabitqDistanceInterval {
    float estimate;
    float lower_bound;
    float upper_bound;
} RabitqDistanceInterval;

/* rabitq_bridge.cpp:21-28 */
typedef struct RabitqPaperEstimate {
    float lower_bound;
    float short_ip;
    float alpha;
    float ip_hat;
    float error_bound;
    uint8_t valid;
} RabitqPaperEstimate;

/* rabitq_bridge.cpp:30-36 (PaperPruneFactors<float>, hnswlib.h:298-305) */
typedef struct RabitqPaperFactors {
    float norm_sqr;
    float data_norm;
    float cross_scale;
    float error_cross_scale;
    uint8_t valid;
} RabitqPaperFactors;

/* Structure-of-arrays device index (row i = base vector i).
 *
 * Gate rows hold what the 1-bit gate reads about a record in one 16-byte
 * aligned row of qg_gate_row_bytes(cd) bytes, so that a gated neighbour
 * costs the cache lines of its MSB plane and nothing more. With m = cd/8
 * rounded up to 16: bytes [0, cd/8) the MSB plane (zero up to m), [m, m+16)
 * the paper-prune factors as four floats {norm_sqr, data_norm, cross_scale,
 * error_cross_scale}, [m+16, m+24) a copy of the header {norm_sqr,
 * long_scale}, byte m+24 the sidecar validity, [m+25, m+32) zero. */
typedef struct qg_index_desc {
    uint8_t *codes;     /* [n][cd/2] primary 4-bit codes (reference nibble layout)  */
    float *hdr;         /* [n][2]    norm_sqr, long_scale                           */
    float *sfac;        /* [n][2]    short_scale, error_scale                       */
    uint8_t *gate;      /* [n][qg_gate_row_bytes(cd)] gate rows, below              */
    uint16_t *rscale;   /* [n][cd/16] residual block scales, fp16 bits              */
    uint8_t *rcodes;    /* [n][cd/2] residual 4-bit codes                           */
    float *rfac;        /* [n][2]    residual_norm_sqr, long_residual_ip            */
    uint64_t n;
    uint8_t *cid;       /* [n]       centroid id (the compact record's last byte);
                         * may be NULL for a space of one centroid             */
} qg_index_desc;

/* Prepared queries on the device (QueryContext, space_rabitq.h:101-132). */
typedef struct qg_query_desc {
    float *q;           /* [nq][cd] rotated residual                              */
    int8_t *x8;         /* [nq][cd] integer code: INT8 or INT4 (dithered), or the
                         * +-1 sign code of B1 / B1Main                           */
    int8_t *x8p;        /* [nq][cd] the same code regrouped: per 8 dims, 4 even, 4 odd */
    void *qs;           /* [nq] 48-byte scalar records (qg_query_scalars)          */
    uint64_t nq;
} qg_query_desc;

/* int8_scale / int8_error_norm / int8_code_sum describe the prepared
 * integer code: INT8 or INT4 scale, error norm and code sum; for B1 and
 * B1Main the sign code's short scale, error scale and signed sum, with
 * b1_alpha its alpha (space_rabitq.h:3493-3524). */
typedef struct qg_query_scalars {
    float query_norm;
    float query_norm_sqr;
    float half_sum;
    float sum;
    float positive_sum;
    float int8_scale;
    float int8_error_norm;
    int32_t max_abs_bits;
    int64_t int8_code_sum;
    float b1_alpha;
    uint32_t pad;
} qg_query_scalars;

/* Query coarse codecs (QueryCoarseCodec, space_rabitq.h:36-44). */
enum { QG_CODEC_FULL = 0, QG_CODEC_B1 = 1, QG_CODEC_INT4 = 2, QG_CODEC_INT8 = 3, QG_CODEC_B1_MAIN = 4 };

/* Estimator kernels for qg_scan. */
enum {
    QG_SCAN_INT8_FULL = 0, /* C7: INT8 query x full code; out_f = estimate, out_f2 = lower   */
    QG_SCAN_ASYM = 1,      /* C1/C2: FP32 query x 4-bit code (mode 0)                         */
    QG_SCAN_REM_FP32 = 2,  /* C3: FP32 remaining bits given short_ips (mode 1)                */
    QG_SCAN_REM_INT8 = 3,  /* C6: INT8 remaining bits given short_ips (mode 2)                */
    QG_SCAN_GATE_FP32 = 4, /* C4: FP32 1-bit gate -> RabitqPaperEstimate                      */
    QG_SCAN_GATE_INT8 = 5, /* C5: INT8 1-bit gate -> RabitqPaperEstimate                      */
    QG_SCAN_RESIDUAL = 6,  /* E1: residual-4 interval from long_in -> RabitqDistanceInterval  */
    QG_SCAN_SYMMETRIC = 7, /* C8: code<->code distance, pairs (qidx = lhs row, ids = rhs row)  */
    QG_SCAN_PAPER_RECORD = 8, /* gate from the compact record (compute_paper_prune_estimate,
                                 space_rabitq.h:3926-3959) -> RabitqPaperEstimate              */
    QG_SCAN_B1 = 9,        /* 4-bit code x B1 sign code (query_distance_b1, space_rabitq.h:
                              3742-3755); queries prepared with QG_CODEC_B1                    */
    QG_SCAN_RESIDUAL_RECORD = 10, /* E1 in the per-record form (computeResidualDistanceInterval
                              FromRecord, :1723-1735): what single-record calls and spaces of
                              several centroids take; pairs only                              */
    QG_SCAN_GATE_B1 = 11   /* the B1 codec's symmetric 1-bit gate (compute_b1_paper_prune_
                              estimate_batch_by_id, space_rabitq.h:4607-4753); pairs only;
                              INT4 queries take QG_SCAN_GATE_INT8                             */
};

/* -------------------------------------------------------- device batch API */

const char *qg_last_error(void);

/* Space lifecycle. Replaces rabitq_space_new / _delete / _set_center
 * (rabitq_bridge.cpp:42-87) with explicit device constants. */
void *qg_space_new(size_t dim, uint32_t seed, size_t centroid_count, size_t residual_bits,
                   size_t residual_block_size, bool residual_mse, bool residual_fp16);
void qg_space_delete(void *space);
int qg_space_set_center(void *space, const float *center_host);
size_t qg_space_dim(const void *space);
size_t qg_space_code_dim(const void *space);
/* Bytes per gate row of a space: cd/8 rounded up to 16, plus 32. */
size_t qg_gate_row_bytes(const void *space);
/* Device pointers to the rotation signs [cd] and the rotated center [cd]. */
const float *qg_space_signs(const void *space);
const float *qg_space_rotated_center(const void *space);
/* Host copies of the rotation signs [code_dim] and the center [dim] (either
 * output may be NULL): the random_seed_ / centroids_ / fht_signs_ payload
 * RaBitQSpace::saveState writes (space_rabitq.h:3190-3192). */
int qg_space_host_constants(const void *space, float *signs_out, float *center_out);

/* Encode n base vectors (device, row-major [n][dim]) into the SoA index.
 * Replaces rabitq_encode_full + copy_compact + copy_residual_record +
 * export_paper_sidecar (rabitq_bridge.cpp:195-222, 317-343). */
size_t qg_encode_workspace_bytes(const void *space, uint64_t n);
int qg_encode(const void *space, const float *x, uint64_t n, qg_index_desc *out,
              void *workspace, size_t workspace_bytes, void *stream);

/* SoA <-> reference AoS records (compact 16+cd/2+1 bytes; residual
 * 8+scale_bytes+cd/2 bytes; full record), all device pointers. */
int qg_pack_records(const void *space, const qg_index_desc *ix, uint8_t *compact, size_t compact_stride,
                    uint8_t *residual, size_t residual_stride, void *stream);
int qg_unpack_records(const void *space, const uint8_t *compact, size_t compact_stride,
                      const uint8_t *residual, size_t residual_stride, uint64_t n,
                      qg_index_desc *out, void *stream);

/* Prepare nq queries (device [nq][dim]). Replaces rabitq_prepare_query +
 * rabitq_set_query_coarse_codec (rabitq_bridge.cpp:355-370). Workspace:
 * qg_prepare_workspace_bytes(nq) bytes of device memory. A space of K
 * centroids writes K * nq query rows (out->nq = K * nq), row k * nq + i the
 * context of query i for centroid k (buildCentroidQuery, space_rabitq.h:
 * 3429-3663, the INT8 dither seeded with the centroid id). */
size_t qg_prepare_workspace_bytes(uint64_t nq);
int qg_prepare_queries(const void *space, const float *queries, uint64_t nq, int codec,
                       qg_query_desc *out, void *workspace, void *stream);

/* Estimators over candidate lists. fixed != 0: ids is [nq][per_query];
 * fixed == 0: (qidx[i], ids[i]) pairs, i < n_pairs (for QG_SCAN_SYMMETRIC
 * both are rows of ix and qv is unused). In a space of K > 1 centroids the
 * query rows are qg_prepare_queries' K contexts per query (row k * nq + i);
 * a record is scored with its centroid's context (qidx = cid * nq + i, the
 * reference's queryForEncoded), the 1-bit sidecar gates with centroid 0's,
 * and QG_SCAN_SYMMETRIC adds the cross-centroid terms (it reads ix->cid). Replaces
 * rabitq_query_distance_batch (:885), rabitq_paper_estimate_batch_sidecar
 * (:702), rabitq_residual_distance_batch (:972) for many queries at once. */
int qg_scan(const void *space, int kernel, int fixed, const qg_index_desc *ix, const qg_query_desc *qv,
            const uint32_t *ids, const uint32_t *qidx, uint64_t per_query, uint64_t n_pairs, float eps0,
            const float *short_ips, const float *long_in, float *out_f, float *out_f2,
            RabitqPaperEstimate *out_est, RabitqDistanceInterval *out_iv, int64_t *out_dot, void *stream);

/* Per-query top-k by (value, id) over [nq][per_query] lists (per_query <= 4096);
 * out_values and out_pos may be NULL (lists up to 1024 without positions take
 * a faster register sort with the same results). */
int qg_topk(const float *values, const uint32_t *ids, uint64_t nq, uint32_t per_query, uint32_t k,
            uint32_t *out_ids, float *out_values, uint32_t *out_pos, void *stream);

/* ------------------------------------------------ precision mode
 * A QR rotation in place of the space's sign-flipped Walsh-Hadamard one
 * (the reference's formal rotation, space_rabitq.h:2900-2912; that remains
 * the default and the parity mode). No counterpart in the reference's
 * formal codec; the QR-orthogonalised Gaussian it follows is
 * experiments/05_disk_system_fair/quantizers.py:383-390.
 * qg_space_set_rotation_qr takes a Gaussian matrix (host, FP64,
 * code_dim x code_dim row-major), factors it G = QR in FP64 on the GPU
 * (classical Gram-Schmidt with reorthogonalisation, diag(R) > 0), and from
 * then on the space rotates with rotated = Q * pad(x): encode and query
 * preparation run the product on the BF16 matrix cores with FP32
 * accumulation, both operands split into bf16_parts BF16 parts (2: hi + lo,
 * 3 products, about 16 significant bits; 3: hi + mid + lo, 6 products, FP32
 * class), encode normalises the rotated residual
 * by its compensated FP32 norm and takes |x - c|, the short and error
 * factors from compensated FP32 sums; everything else is the codec as
 * before. Needs code_dim >= 128. The 1-bit sign-plane scan refuses such a
 * space. qg_space_rotation_qr copies Q (code_dim x code_dim row-major) to
 * the host; qg_rotate_rows applies the space's rotation to n device rows
 * (out [n][code_dim]; subtract_center only for QR spaces: Q * pad(x - c),
 * the product encode uses). */
int qg_space_set_rotation_qr(void *space, const double *gaussian, int bf16_parts);
int qg_space_rotation_qr(const void *space, double *q_out);
int qg_rotate_rows(const void *space, const float *rows, uint64_t n, int subtract_center, float *out,
                   void *stream);

/* Batched graph search: greedy beam search (beam width 1, start node 0)
 * with the 1-bit gate, remaining-bit distances and the residual-4 rerank of
 * the first max(rerank, k) pool entries, one 4-wavefront workgroup per
 * query.
 * Replaces the per-query loop search_ours_paper_active_inner +
 * flush_distance_batch + rerank (ours_diskann.rs:1539-1986) and its bridge
 * calls. Graph: degrees [n], edges [n][edge_stride], edge_stride <= 64.
 * Outputs [nq][k] ids / estimates (missing entries: 0xFFFFFFFF / +inf);
 * optional stats [nq][6] = hops, gate evaluations, distances computed
 * (start node included), prunes (first pass and re-check), re-checks made
 * with the pool full and the gate valid, and re-check prunes; and the final
 * pool [nq][max(l_search, k)]. The reference's counters follow from these
 * (ours_diskann.rs:1604-1968): visited_nodes = 1 + gate evaluations,
 * distance_computations = distances - re-check prunes, paper_checked =
 * prunes - re-check prunes + re-checks, paper_would_prune = prunes.
 * Workspace:
 * qg_search_workspace_bytes(n, slots) of device memory, 16-byte aligned
 * (per-slot visited bitsets and the work-queue head); slots = concurrent
 * queries. max(l_search, k) <= 4096, max(rerank, k) <= 1024 and
 * code_dim <= 2048. */
size_t qg_search_workspace_bytes(uint64_t n, uint32_t slots);
int qg_search(const void *space, const qg_index_desc *ix, const qg_query_desc *qv, int codec,
              const uint32_t *degrees, const uint32_t *edges, size_t edge_stride, uint32_t k, uint32_t l_search,
              float eps0, uint32_t rerank, uint32_t slots, uint32_t *out_ids, float *out_dists,
              uint64_t *out_stats, uint32_t *pool_ids, float *pool_dists, void *workspace, size_t workspace_bytes,
              void *stream);
/* With out_ids == out_dists == NULL, qg_search stops after the traversal
 * and only writes the pools (and stats): the input of the learned rerank. */

/* qg_search with uneven lists: l_per_query (device, [nq], may be NULL) gives
 * each query its own search list size, clamped to [k, max(l_search, k)]
 * (l_search is then the pool capacity and the pool outputs' row stride);
 * waves (2 or 4) sets the wavefronts per query workgroup. One launch serves
 * every query through the work queue; it enqueues no allocation, copy or
 * memset, so it can be captured in a HIP graph. This batched form has no
 * single counterpart in the reference, whose runner loops over queries
 * (diskann_runner.rs:360-439). Dynamic LDS per workgroup:
 * qg_search_lds_bytes(code_dim, max(l_search, k)) plus about 19 KB static. */
int qg_search_ex(const void *space, const qg_index_desc *ix, const qg_query_desc *qv, int codec,
                 const uint32_t *degrees, const uint32_t *edges, size_t edge_stride, uint32_t k, uint32_t l_search,
                 const uint32_t *l_per_query, float eps0, uint32_t rerank, uint32_t slots, uint32_t waves,
                 uint32_t *out_ids, float *out_dists, uint64_t *out_stats, uint32_t *pool_ids, float *pool_dists,
                 void *workspace, size_t workspace_bytes, void *stream);
size_t qg_search_lds_bytes(uint32_t cd, uint32_t L);

/* qg_search_ex with the learned rerank depth fused after each query's
 * traversal (no counterpart in the reference; the plain top-100 rerank of
 * qg_search_ex is its fallback and judge): the feature rows of the first
 * min(|pool|, max(rerank, k)) pool entries, the MLP's predicted errors
 * (mlp_packed: qg_mlp_pack of the qg_mlp_* parameter vector, the weights
 * rounded to bf16 in each lane's order, qg_mlp_packed_bytes() of device
 * memory) and the depth rule (kmult, min_depth) pick how many of them the
 * residual-4 rerank scores; out_depth [nq] (optional) receives it. The
 * results equal qg_rerank_features + qg_mlp_forward + qg_learned_rerank on
 * the traversal's pools, bit for bit, without the pool write-out or the
 * extra launches. rstats: qg_residual_stats of the index. max(rerank, k)
 * <= 112. */
size_t qg_mlp_packed_bytes(void);
int qg_mlp_pack(const float *params, void *packed, void *stream);
int qg_search_learned(const void *space, const qg_index_desc *ix, const qg_query_desc *qv, int codec,
                      const uint32_t *degrees, const uint32_t *edges, size_t edge_stride, uint32_t k,
                      uint32_t l_search, const uint32_t *l_per_query, float eps0, uint32_t rerank, uint32_t slots,
                      uint32_t waves, const void *mlp_packed, float kmult, uint32_t min_depth, const float *rstats,
                      uint32_t *out_ids, float *out_dists, uint64_t *out_stats, uint32_t *out_depth,
                      void *workspace, size_t workspace_bytes, void *stream);

/* qg_search_ex without the workspace argument; it allocates nothing. No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The work-queue head, error flag and
 * per-slot visited bitsets live in a file-scope device array of
 * qg_search_noalloc_scratch_bytes() (32 MiB) that each GPU holds until the
 * process ends; calls whose qg_search_workspace_bytes(n, slots) exceeds it
 * fail with an error. Every call on one GPU shares that array, so calls that
 * overlap on one GPU race. Outputs and errors as qg_search_ex. */
size_t qg_search_noalloc_scratch_bytes(void);
int qg_search_noalloc(const void *space, const qg_index_desc *ix, const qg_query_desc *qv, int codec,
                      const uint32_t *degrees, const uint32_t *edges, size_t edge_stride, uint32_t k,
                      uint32_t l_search, const uint32_t *l_per_query, float eps0, uint32_t rerank, uint32_t slots,
                      uint32_t waves, uint32_t *out_ids, float *out_dists, uint64_t *out_stats, uint32_t *pool_ids,
                      float *pool_dists, void *stream);

/* ------------------------------------- learned adaptive rerank depth
 * No counterpart in the reference. A 24-64-64-1 MLP (BF16 MFMA, FP32
 * accumulation; parameters are qg_mlp_param_count() FP32 values laid out
 * W1[64][24], b1[64], W2[64][64], b2[64], W3[64], b3) predicts a quantile
 * of the 4-bit estimator's error |exact - estimate| / ||q||^2 for each of a
 * query's first `depth` pool candidates; the rerank depth per query is the
 * last candidate whose lower bound can beat the largest upper bound of the
 * first min_depth, and the residual-4 rerank then runs on that many. With
 * preds == NULL, qg_learned_rerank is the plain rerank of the search
 * kernel (ours_diskann.rs:1813-1902). */
size_t qg_mlp_param_count(void);
size_t qg_mlp_workspace_bytes(uint64_t n);
int qg_mlp_forward(const float *params, const uint16_t *features, uint64_t n, float *out, void *stream);
int qg_mlp_grad(const float *params, const uint16_t *features, const float *targets, uint64_t n, float tau,
                float *grad_out, float *loss_out, void *workspace, size_t workspace_bytes, void *stream);
int qg_mlp_adam_step(float *params, float *m, float *v, const uint16_t *features, const float *targets, uint64_t n,
                     float tau, float lr, float beta1, float beta2, float eps, uint32_t step, float *loss_out,
                     void *workspace, size_t workspace_bytes, void *stream);
/* per-vector residual block stats [n][4] (rms scale, max scale, zero-scale fraction, 0) */
int qg_residual_stats(const void *space, const qg_index_desc *ix, float *out, void *stream);
/* features [nq][depth][32] bf16 of the pools [nq][L] (first min(|pool|, depth) rows) */
int qg_rerank_features(const void *space, const qg_index_desc *ix, const qg_query_desc *qv, const uint32_t *pool_ids,
                       const float *pool_dists, uint32_t L, uint32_t depth, const float *rstats, uint16_t *features,
                       void *stream);
int qg_learned_rerank(const void *space, const qg_index_desc *ix, const qg_query_desc *qv, int codec,
                      const uint32_t *pool_ids, const float *pool_dists, uint32_t L, uint32_t depth,
                      const float *preds, float kmult, uint32_t min_depth, uint32_t k, uint32_t *out_ids,
                      float *out_dists, uint32_t *out_depth, void *stream);

/* Vamana graph build with the symmetric 4-bit distance over the index's
 * primary codes: VamanaIndex::buildEncodedSymmetricBulk + refineGraphSymmetric
 * as driven by rabitq_build_vamana_graph (rabitq_bridge.cpp:224-303,
 * vamana_index.h:308-528). Start node 0, batches of inserts from id 1,
 * grouped back-edges, refine_passes full passes. degrees [n] and edges
 * [n][r] are device arrays (unused edge slots 0xFFFFFFFF);
 * *distance_evaluations receives the reference's counter. r <= 64,
 * l_build <= 1024, beam_width <= 64. Synchronises the stream.
 * Workspace: qg_build_workspace_bytes(n, r, l_build, batch, refine_passes). */
size_t qg_build_workspace_bytes(uint64_t n, uint32_t r, uint32_t l_build, uint32_t batch, uint32_t refine_passes);
int qg_build_graph(const void *space, const qg_index_desc *ix, uint32_t r, uint32_t l_build, float alpha,
                   uint32_t beam_width, uint32_t batch, uint32_t refine_passes, uint32_t prune_cap,
                   uint32_t early_stop_hops, uint32_t *degrees, uint32_t *edges, uint64_t *distance_evaluations,
                   void *workspace, size_t workspace_bytes, void *stream);

/* Lloyd k-means of rabitq_train_centroids (rabitq_bridge.cpp:101-172) over
 * device samples [sample_count][dim] in draw order: centroid c starts at
 * sample c % sample_count, then `iters` rounds of first-minimum assignment
 * (d = fma(diff, diff, d) in dimension order) and in-order member sums.
 * centroids: device [k][dim] output; k <= 256.
 * Workspace: qg_kmeans_workspace_bytes(sample_count, dim, k). */
size_t qg_kmeans_workspace_bytes(uint64_t sample_count, uint64_t dim, uint32_t k);
int qg_train_centroids(const float *samples, uint64_t sample_count, uint64_t dim, uint32_t k, int iters,
                       float *centroids, void *workspace, size_t workspace_bytes, void *stream);

/* ------------------------------------------------ 1-bit sign-plane scan
 * The fixed 1-bit database scanned with four query codecs of
 * experiments/04_query_codec_1bit_scan/query_codec_1bit_scan.cpp, general in
 * the code dimension (the reference fixes 960 -> 1024) and batched over
 * queries (the reference scans one query at a time, :868-925). Bit i of a
 * database vector is set when coordinate i of its rotation is > 0; the
 * rotation is the space's (signs, zero padding, unnormalised Walsh-Hadamard;
 * no centering, no scaling; seed 0 is the identity). The database is stored
 * in blocks of 32 vectors, word-major inside a block:
 * db[(v / 32 * W + w) * 32 + v % 32], W = code_dim / 64 u64 words
 * (build_or_load_database, :341-384), so n must be a multiple of 32. */
enum { QG_SIGNSCAN_FP32 = 0, QG_SIGNSCAN_B1 = 1, QG_SIGNSCAN_INT4 = 2, QG_SIGNSCAN_INT8 = 3 }; /* Codec, :735 */
#define QG_SIGNSCAN_BATCH 32
#define QG_SIGNSCAN_ROWS 13
/* Scalars of QueryRepresentations (:405-425). */
typedef struct qg_signscan_query {
    float fp32_sum, fp32_norm2, b1_scale, b1_constant;
    float int4_scale, int4_constant, int8_scale, int8_constant;
    int64_t int4_sum, int8_sum;
} qg_signscan_query;
/* Database build of n vectors (device [n][dim]) into db (device,
 * n * code_dim / 64 u64); replaces build_or_load_database (:341-384). */
int qg_signscan_encode(const void *space, const float *x, uint64_t n, uint64_t *db, void *stream);
/* prepare_query (:458-492) of nq queries (device [nq][dim]): q [nq][code_dim]
 * the rotated FP32 query, bits [nq][13][W] (row 0 the B1 sign code, rows 1-4
 * the INT4 planes, rows 5-12 the INT8 planes, plane 0 first, two's
 * complement), scalars [nq]. */
int qg_signscan_prepare(const void *space, const float *queries, uint64_t nq, float *q, uint64_t *bits,
                        qg_signscan_query *scalars, void *stream);
/* scan_fp32 / scan_b1 / scan_integer<4|8> (:529-648; codec QG_SIGNSCAN_*)
 * of nq prepared queries against all n database vectors in one launch:
 * out [nq][n] distances. q is read by the FP32 codec, bits by the others. */
int qg_signscan(const void *space, int codec, const uint64_t *db, uint64_t n, const float *q, const uint64_t *bits,
                const qg_signscan_query *scalars, uint64_t nq, float *out, void *stream);

/* ------------------------------------------------------------ bridge API */
/* Host-memory drop-ins for rabitq_bridge.cpp; see that file for semantics. */

void *rabitq_space_new(size_t dim, uint32_t seed, size_t centroid_count, size_t residual_bits,
                       size_t residual_block_size, bool residual_mse, bool residual_fp16); /* :42  */
void rabitq_space_delete(void *space);                                                 /* :76  */
bool rabitq_space_set_center(void *space, const float *center);                       /* :80  */
/* centroid_count must equal the space's (1): the K > 1 codec is not built. */
bool rabitq_space_set_centroids(void *space, const float *centroids, size_t centroid_count); /* :89 */
bool rabitq_train_centroids(const float *data, size_t count, size_t dim, size_t k, size_t sample_count,
                            uint32_t seed, int iters, float *out);                    /* :101 */
size_t rabitq_full_record_bytes(const void *space);                                    /* :174 */
size_t rabitq_compact_record_bytes(const void *space);                                 /* :178 */
size_t rabitq_residual_record_bytes(const void *space);                                /* :183 */
size_t rabitq_residual_block_size(const void *space);                                  /* :187 */
size_t rabitq_residual_bits(const void *space);                                        /* :191 */
bool rabitq_encode_full(const void *space, const float *raw, void *encoded_out);       /* :195 */
bool rabitq_copy_compact(const void *space, const void *full, void *compact_out);      /* :204 */
bool rabitq_copy_residual_record(const void *space, const void *full, void *out);      /* :214 */
bool rabitq_build_vamana_graph(void *space, const void *compact_records, size_t record_count,
                               size_t compact_stride, size_t r, size_t l_build, float alpha, size_t beam_width,
                               size_t batch_size, size_t refine_passes, size_t prune_candidate_cap,
                               size_t build_early_stop_hops, uint32_t *degrees_out, uint32_t *edges_out,
                               size_t edge_stride, uint8_t *paper_msb_out, void *paper_factors_out,
                               size_t *paper_msb_stride_out, size_t *paper_factor_bytes_out,
                               uint64_t *distance_evaluations_out);                   /* :224 */
size_t rabitq_paper_msb_code_bytes(const void *space);                                 /* :305 */
size_t rabitq_paper_factor_bytes(void);                                                /* :313 */
bool rabitq_export_paper_sidecar(const void *space, const void *compact_records, size_t record_count,
                                 size_t compact_stride, uint8_t *msb_out, void *factors_out); /* :317 */
float rabitq_symmetric_distance(const void *space, const void *lhs, const void *rhs); /* :345 */
const void *rabitq_prepare_query(const void *space, const float *query);              /* :355 */
void rabitq_set_query_coarse_codec(void *space, int codec);                           /* :364 */
void rabitq_release_query(const void *space, const void *prepared);                   /* :372 */
float rabitq_query_distance(const void *space, const void *prepared, const void *encoded); /* :380 */
/* B1 sign-code distance; needs a query prepared under codec 1 or 4, else +inf. */
float rabitq_query_distance_b1(const void *space, const void *prepared, const void *encoded); /* :389 */
bool rabitq_query_distance_b1_batch(const void *space, const void *prepared, const uint32_t *ids, size_t count,
                                    const void *encoded_base, size_t encoded_stride,
                                    float *out_distances);                            /* :398 */
RabitqPaperEstimate rabitq_paper_estimate(const void *space, const void *prepared, const void *encoded,
                                          float epsilon0);                            /* :425 */
RabitqPaperEstimate rabitq_paper_estimate_sidecar(const void *space, const void *prepared,
                                                  const uint8_t *msb_code, const void *factors,
                                                  float epsilon0);                    /* :448 */
RabitqPaperEstimate rabitq_paper_estimate_codec(const void *space, const void *prepared, const void *encoded,
                                                const uint8_t *msb_code, const void *factors,
                                                float epsilon0);                      /* :475 */
bool rabitq_paper_estimate_codec_batch(const void *space, const void *prepared, const uint32_t *ids,
                                       size_t count, const void *encoded_base, size_t encoded_stride,
                                       const uint8_t *msb_base, size_t msb_stride, const void *factors_base,
                                       size_t factors_stride, float epsilon0,
                                       RabitqPaperEstimate *out);                     /* :506 */
float rabitq_query_distance_with_paper(const void *space, const void *prepared, const void *encoded,
                                       float short_ip);                               /* :595 */
float rabitq_query_distance_with_quantized_paper(const void *space, const void *prepared, const void *encoded,
                                                 float short_ip);                     /* :609 */
/* The bridge's spaces keep residuals in separate records, so this returns
 * {long, long, long}, the reference's failure value for that case. */
RabitqDistanceInterval rabitq_residual_distance(const void *space, const void *prepared, const void *encoded,
                                                float long_distance);                 /* :624 */
RabitqDistanceInterval rabitq_residual_distance_from_record(const void *space, const void *prepared,
                                                            const void *encoded, const void *residual_record,
                                                            float long_distance);     /* :644 */
bool rabitq_paper_estimate_batch(const void *space, const void *prepared, const uint32_t *ids, size_t count,
                                 const void *encoded_base, size_t encoded_stride, float epsilon0,
                                 RabitqPaperEstimate *out);                           /* :665 */
bool rabitq_query_distance_batch(const void *space, const void *prepared, const uint32_t *ids,
                                 const uint8_t *modes, size_t count, const void *encoded_base,
                                 size_t encoded_stride, const float *short_ips, float *out); /* :885 */
bool rabitq_paper_estimate_batch_sidecar(const void *space, const void *prepared, const uint32_t *ids,
                                         size_t count, const uint8_t *msb_base, size_t msb_stride,
                                         const void *factors_base, size_t factors_stride,
                                         float epsilon0, RabitqPaperEstimate *out);   /* :702 */
bool rabitq_residual_distance_batch(const void *space, const void *prepared, const uint32_t *ids,
                                    size_t count, const void *encoded_base, size_t encoded_stride,
                                    const void *residual_base, size_t residual_stride,
                                    const float *long_distances, RabitqDistanceInterval *out); /* :972 */
/* Mode-0 distance then residual interval per id (bit-identical to the two
 * calls in sequence, space_rabitq.h:5557-5778). */
bool rabitq_full_residual_distance_batch(const void *space, const void *prepared, const uint32_t *ids,
                                         size_t count, const void *encoded_base, size_t encoded_stride,
                                         const void *residual_base, size_t residual_stride,
                                         RabitqDistanceInterval *out);                /* :1044 */

#ifdef __cplusplus
}
#endif
#endif
