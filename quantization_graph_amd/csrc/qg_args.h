// qg_args.h -- kernel argument records shared by the launchers (.hip) and
// the C-ABI layer (qg_api.cpp).
#pragma once

#include "qg_common.h"

namespace qg {

struct PrepParams {
    const float *raw;          // [nq][dim]
    uint64_t nq;
    uint32_t dim, cd;
    int codec;                 // 0 Full, 3 Int8
    int identity;              // seed == 0: no rotation
    const float *signs;        // [cd]
    const float *rot_center;   // [cd]
    float center_norm_sqr;
    const uint64_t *seed;      // [nq] FNV-1a of the raw bytes
    const double *raw_norm;    // [nq]
    const double *raw_dot;     // [nq]
    float *q_out;              // [nq][cd]
    int8_t *x8_out;            // [nq][cd]
    int8_t *x8p_out;           // [nq][cd]
    QScalars *qs_out;          // [nq]
};

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

struct EncMeta {
    double max_o;
    double best_t;
    float rn;
    float ip_norm;
    uint16_t lv_start[7];
    uint16_t lv_end[7];
    uint32_t pad;
};

struct EncArgs {
    const float *x;            // [n][dim]
    uint64_t n;                // vectors in this batch
    uint32_t dim, cd, block, block_count;
    int identity;              // seed == 0
    int mse;
    const float *signs;        // [cd]
    const float *center;       // [dim]
    float *rot;                // ws [n][cd]
    float *sa;                 // ws [n][cd]  |o| sorted descending
    uint16_t *si;              // ws [n][cd]  matching dimension index
    EncMeta *meta;             // ws [n]
    // outputs (SoA rows for this batch)
    uint8_t *codes;
    float2 *hdr;
    float2 *sfac;
    uint8_t *msb;
    float4 *factors;
    uint8_t *fvalid;
    uint16_t *rscale;
    uint8_t *rcodes;
    float2 *rfac;
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
    uint8_t *msb; float4 *factors; uint8_t *fvalid;
};

struct SearchArgs {
    IndexView ix;
    QueryView qv;
    uint32_t cd;
    int codec;                 // 0 Full, 3 Int8
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
    uint64_t *stats;           // [nq][4] hops, gate evaluations, distances, pruned (optional)
    uint32_t *pool_ids;        // [nq][L] final pool (optional)
    float *pool_d;             // [nq][L]
};

}  // namespace qg
