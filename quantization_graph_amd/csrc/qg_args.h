// qg_args.h -- kernel argument records shared by the launchers (.hip) and
// the C-ABI layer (qg_api.cpp).
#pragma once

#include "qg_common.h"

namespace qg {

struct PrepParams {
    const float *raw;          // [nq][dim]
    uint64_t nq;
    uint32_t dim, cd;
    int codec;                 // 0 Full, 1 / 4 B1 sign code, 3 Int8
    int identity;              // seed == 0: no rotation
    const float *signs;        // [cd]
    const float *rot_center;   // [cd]
    float center_norm_sqr;
    const uint64_t *seed;      // [nq] FNV-1a of the raw bytes
    uint32_t centroid_id;      // mixed into the INT8 dither seed (space_rabitq.h:3549-3558)
    const double *raw_norm;    // [nq]
    const double *raw_dot;     // [nq]
    float *q_out;              // [nq][cd]; holds the rotated raw rows on entry when pre_rotated
    int pre_rotated;           // QR rotation: q_out already holds Q * pad(raw)
    int8_t *x8_out;            // [nq][cd]
    int8_t *x8p_out;           // [nq][cd]
    QScalars *qs_out;          // [nq]
};

// Scalars of one prepared sign-scan query (qg_signscan_query; the scalar
// fields of QueryRepresentations, query_codec_1bit_scan.cpp:405-425).
struct SsQuery {
    float fp32_sum, fp32_norm2, b1_scale, b1_constant;
    float int4_scale, int4_constant, int8_scale, int8_constant;
    int64_t int4_sum, int8_sum;
};
static_assert(sizeof(SsQuery) == 48, "SsQuery layout");

struct ScanArgs {
    IndexView ix;
    QueryView qv;
    uint32_t cd;
    const uint32_t *ids;        // fixed: [nq][per_query]; pairs: [n_pairs]
    const uint32_t *qidx;       // pairs only
    uint64_t per_query;         // fixed
    uint64_t n_pairs;           // pairs
    float eps0;
    const float *short_ips;     // remaining-bit distances
    const float *long_in;       // residual rerank input
    float *out_f;               // distances / estimates
    float *out_f2;              // lower bounds (INT8 full estimate)
    PaperEstimate *out_est;     // gates
    Interval *out_iv;           // residual intervals
    int64_t *out_dot;           // raw integer dots (optional)
    // symmetric distances of spaces with K > 1 centroids: the cross-centroid
    // terms (space_rabitq.h:2634-2683)
    uint32_t centroid_count;
    const float *rot_centroids; // [K][cd] rotated centroids
    const float2 *pair_terms;   // [K][K] {|c_a - c_b|^2, half sum of r_a - r_b}
    int record_form;            // residual pairs: the per-record lane form (see k_residual)
};

struct SelectArgs {
    const float *values;     // [nq][per_query]
    const uint32_t *ids;     // [nq][per_query]
    uint64_t nq;
    uint32_t per_query;
    uint32_t k;
    uint32_t *out_ids;       // [nq][k]
    float *out_values;       // [nq][k]
    uint32_t *out_pos;       // [nq][k] position in the input list (optional)
};

struct EncArgs {
    const float *x;            // [n][dim]
    uint64_t n;                // vectors in this batch
    uint32_t dim, cd, block, block_count;
    int identity;              // seed == 0
    int mse;
    int precision;             // QR rotation: compensated FP32 norms and factors, rot filled by the GEMM
    const float *signs;        // [cd]
    const float *center;       // [K][dim] centroids; vector v uses row cid[v]
    uint32_t centroid_count;   // K
    uint8_t *cid;              // [n] assigned centroid (K > 1: written by k_enc_assign; may be null for K = 1)
    const float *rot;          // precision mode: ws [n][cd], the rotation GEMM's Q (x - c)
    // outputs (SoA rows for this batch)
    uint8_t *codes;
    float2 *hdr;
    float2 *sfac;
    uint8_t *gate;             // [n][cd/8 + 32] gate rows
    uint16_t *rscale;
    uint8_t *rcodes;
    float2 *rfac;
};

// Rotation GEMM of the QR rotation: out[r][i] = sum_k Q[i][k] * u_r[k] with
// u_r = (x_r - center) * row_scale[r] (center / row_scale optional), Q as
// BF16 parts [parts][cd][kpad] (k contiguous, zero past dim).
struct RotGemmArgs {
    const float *x;            // [n][dim]
    uint64_t n;
    uint32_t dim, cd, kpad;
    const float *center;       // [dim] or null
    const float *row_scale;    // [n] or null
    const uint16_t *q;         // [parts][cd][kpad] bf16 bits: Q rounded, then the rounded remainders
    uint32_t parts;            // 2 or 3 BF16 parts per operand
    float *out;                // [n][cd]
};

struct PackArgs {
    IndexView ix;
    const float2 *rfac;
    uint32_t cd, block_count;
    uint8_t *compact; uint64_t compact_stride;
    uint8_t *residual; uint64_t residual_stride;
    // unpack direction
    const uint8_t *in_compact; const uint8_t *in_residual;
    uint8_t *codes; float2 *hdr; float2 *sfac; uint16_t *rscale; uint8_t *rcodes; float2 *rfac_out;
    uint8_t *gate;             // unpack: gate rows
    uint8_t *cid_out;          // unpack: centroid ids (optional)
};

struct SearchArgs {
    IndexView ix;
    QueryView qv;
    uint32_t cd;
    int codec;                 // 0 Full, 1 / 4 B1 sign code, 3 Int8
    uint32_t L, k, rerank;
    float eps0;
    const uint32_t *degrees;   // [n]
    const uint32_t *edges;     // [n][edge_stride]
    uint64_t edge_stride;
    uint64_t n;
    uint32_t *visited;         // [slots][visited_words] scratch bitsets
    uint64_t visited_words;
    uint32_t *counter;         // work-queue head
    uint32_t *err;             // set to non-zero on an out-of-range neighbour
    uint32_t *out_ids;         // [nq][k]
    float *out_d;              // [nq][k]
    uint64_t *stats;           // [nq][6] hops, gate evaluations, distances, pruned, re-checks, re-check prunes (optional)
    uint32_t *pool_ids;        // [nq][L] final pool (optional)
    int pool_only;             // skip the rerank (out_ids / out_d unused)
    float *pool_d;             // [nq][L]
    const uint32_t *lq;        // [nq] per-query search list sizes (optional; L is then the pool capacity)
    // learned rerank depth fused after the traversal (optional: mlp null)
    const float *mlp;          // MLP weights packed per lane (qg_mlp_pack)
    float kmult;               // error multiplier of the depth rule
    uint32_t min_depth;
    const float4 *rstats;      // [n] residual block stats
    const float2 *rfac;        // [n] residual factors
    uint32_t *out_depth;       // [nq] rerank depth taken (optional)
};

// Vamana build (qg_build.hip)
struct BuildSearchArgs {
    const uint8_t *codes;      // [n][cd/2]
    const float2 *hdr;         // [n]
    uint32_t cd;
    const uint32_t *degrees;   // [n]
    const uint32_t *edges;     // [n][R]
    uint32_t R;
    uint64_t node_count;       // neighbours must be < node_count
    uint32_t L, beam, early_stop;
    uint32_t jobs;
    uint64_t qid_base;         // job j searches for record qid_base + j
    uint32_t *tags;            // [slots][tag_stride] visited epochs
    uint64_t tag_stride;
    uint32_t tag_base;         // job j marks with tag_base + j + 1
    uint32_t *counter;         // work-queue head
    uint32_t *err;             // bit 0: out-of-range neighbour
    unsigned long long *evals; // distance evaluations
    float *pool_d;             // [jobs][L]
    uint32_t *pool_id;         // [jobs][L]
    uint32_t *pool_len;        // [jobs]
};

struct PruneArgs {
    const uint8_t *codes;
    const float2 *hdr;
    uint32_t cd;
    uint32_t jobs;
    const uint32_t *job_ids;   // optional: block b runs job job_ids[b]
    uint64_t src_base;         // source of job j: sources[j] or src_base + j
    const uint32_t *sources;
    // part A: a sorted search pool, first min(len, cap) entries
    const float *pool_d;
    const uint32_t *pool_id;
    const uint32_t *pool_len;
    uint64_t pool_stride;
    uint32_t cap;
    // part B: ids whose distance from the source is computed here, either a
    // graph row of the source (graph_deg != null) or a list (extra)
    const uint32_t *graph_deg;
    const uint32_t *graph_edges;
    const uint32_t *extra;
    const uint64_t *extra_off;
    const uint32_t *extra_len;
    uint32_t R;
    float alpha;
    // per-candidate state for pools above the LDS capacity (part B lists only),
    // indexed like extra
    unsigned long long *big_keys;
    float *big_min;
    float *big_pre;
    uint32_t *big_state;
    // output rows: out_edges[src][R], out_deg[src]
    uint32_t *out_edges;
    uint32_t *out_deg;
    uint32_t *err;             // bit 1: pool above LDS capacity without state storage
    unsigned long long *evals;
};

struct MergeArgs {
    const uint32_t *group_src;   // [ngroups] source of each back-edge group
    const uint32_t *group_begin; // [ngroups] offset in the sorted targets
    const uint32_t *group_len;   // [ngroups]
    const uint32_t *targets;     // sorted (source, target) pairs, target half
    uint32_t ngroups;
    uint32_t R;
    uint32_t *degrees;           // graph, updated in place for groups that fit
    uint32_t *edges;
    const uint64_t *merged_off;  // [ngroups]
    uint32_t *merged;            // merged neighbour lists of overflowing groups
    uint32_t *job_src;           // prune jobs (overflowing groups), one region of
    uint64_t *job_off;           // job_cap entries per size class
    uint32_t *job_len;
    uint32_t *job_count;         // [4] jobs per size class
    uint64_t job_cap;
};

// learned rerank (qg_mlp.hip)
struct RerankArgs {
    IndexView ix;
    const float2 *rfac;        // [n] {residual_norm_sqr, long_residual_ip}
    const float4 *rstats;      // [n] {rms scale, max scale, zero-scale fraction, 0}
    QueryView qv;
    uint32_t cd;
    int codec;                 // 0 Full, 1 / 4 B1 sign code, 3 Int8
    uint64_t nq;
    uint32_t L;                // pool row stride
    uint32_t depth;            // largest rerank depth (feature rows per query)
    uint32_t k;
    uint32_t min_depth;
    float kmult;
    const uint32_t *pool_ids;  // [nq][L], 0xFFFFFFFF past the pool
    const float *pool_d;       // [nq][L]
    const float *preds;        // [nq][depth] predicted errors, or null (plain rerank)
    uint16_t *feats;           // [nq][depth][32] bf16
    uint32_t *out_ids;         // [nq][k]
    float *out_d;              // [nq][k]
    uint32_t *out_depth;       // [nq] (optional)
};

}  // namespace qg
