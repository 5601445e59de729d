// qg_scan.hip -- the estimator kernels over candidate lists.
//
// Two list shapes:
//  * fixed:  every query q owns `per_query` candidate ids at ids[q*per_query..];
//            grid = (ceil(per_query / 256), nq), so a workgroup serves one
//            query and stages that query's vectors in LDS once.
//  * pairs:  a flat list of (query index, id) pairs; query vectors are read
//            through the cache (used for ragged batches).
// One lane evaluates one (query, candidate) pair and reads the candidate's
// code row with 16-byte loads. Integer dots are exact (v_dot4 on unpacked
// nibbles); float estimators keep the reference's 16-lane accumulation order.
#include "qg_args.h"
#include "qg_device.h"

namespace qg {

constexpr int kScanThreads = 256;



__device__ __forceinline__ bool fixed_item(const ScanArgs &a, uint64_t &qi, uint64_t &pos) {
    qi = blockIdx.y;
    pos = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    return pos < a.per_query;
}

// ---------------------------------------------------------------- C7 -----
// INT8 query x full 4-bit code: dot = sum c_i x_i (exact int32), then
// finish_int_paper_prune_estimate (space_rabitq.h:5230-5260).
template <bool kFixed>
__global__ void __launch_bounds__(kScanThreads) k_int8_full(ScanArgs a) {
    __shared__ uint32_t xs[2048 / 4];
    uint64_t qi, pos, item;
    bool live;
    if (kFixed) {
        live = fixed_item(a, qi, pos);
        item = qi * a.per_query + pos;
        if (a.cd <= 2048) {
            const uint32_t *src = (const uint32_t *)(a.qv.x8p + qi * a.cd);
            for (uint32_t i = threadIdx.x; i < a.cd / 4; i += blockDim.x) xs[i] = src[i];
            __syncthreads();
        }
    } else {
        item = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
        live = item < a.n_pairs;
        qi = live ? a.qidx[item] : 0;
    }
    if (!live) return;
    const uint32_t id = a.ids[item];
    const uint4 *row = (const uint4 *)(a.ix.codes + (uint64_t)id * (a.cd / 2));
    const uint32_t *xg = (const uint32_t *)(a.qv.x8p + qi * a.cd);
    const bool staged = kFixed && a.cd <= 2048;  // above, the query is read from global memory
    int acc = 0;
    const uint32_t chunks = a.cd / 32;  // 16 code bytes = 32 dims per chunk
    for (uint32_t c = 0; c < chunks; ++c) {
        const uint4 w = row[c];
        const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t d = 4 * c + k;  // code dword: dims 8d..8d+7
            const uint32_t xl = staged ? xs[2 * d] : xg[2 * d];
            const uint32_t xh = staged ? xs[2 * d + 1] : xg[2 * d + 1];
            acc = __builtin_amdgcn_sdot4((int)(ws[k] & 0x0F0F0F0Fu), (int)xl, acc, false);
            acc = __builtin_amdgcn_sdot4((int)((ws[k] >> 4) & 0x0F0F0F0Fu), (int)xh, acc, false);
        }
    }
    const QScalars s = a.qv.qs[qi];
    const float2 h = a.ix.hdr[id];
    const float scale = s.int8_scale;
    float est, lower;
    if (!long_scale_ok(h.y) || !(scale > 0.0f) || !__builtin_isfinite(scale) || s.query_norm <= 0.0f) {
        est = h.x + s.query_norm_sqr;
        lower = est;
    } else {
        const double centered = (double)acc - 7.5 * (double)s.int8_code_sum;
        const double sip = (double)scale * centered;
        const double est_ip = (double)h.y * sip;
        est = h.x + s.query_norm_sqr - (float)est_ip;
        lower = est;
        if (a.out_f2) {
            const double db_err = (double)a.ix.sfac[id].y * (double)s.query_norm;
            lower = est - (float)((double)a.eps0 * db_err);
        }
    }
    a.out_f[item] = est;
    if (a.out_f2) a.out_f2[item] = lower;
    if (a.out_dot) a.out_dot[item] = acc;
}

// C7 for fixed lists, coalesced form: 8 lanes per candidate. Lane `sub` of
// a group owns the 16-byte pieces sub, sub+8, sub+16, ... of the code row, so
// every load instruction reads 128 contiguous bytes per candidate, and keeps
// the matching slice of the INT8 query in registers. The integer dot is
// exact in any order; the 8 partial sums meet through lane shuffles.
template <int CD>
__global__ void __launch_bounds__(kScanThreads) k_int8_full_grouped(ScanArgs a) {
    constexpr int kLanes = 8;
    constexpr int kPieces = CD / 2 / 16 / kLanes;  // 16-byte pieces per lane
    const uint64_t qi = blockIdx.y;
    const int sub = threadIdx.x & (kLanes - 1);
    const int slot = threadIdx.x / kLanes;
    constexpr int kSlots = kScanThreads / kLanes;
    // query slice: piece p of this lane covers code bytes 16*(sub + 8p) ..,
    // i.e. dims 32*(sub + 8p) .. +32 -> 8 dwords of x8p
    uint32_t xv[kPieces][8];
    const uint32_t *xq = (const uint32_t *)(a.qv.x8p + qi * CD);
#pragma unroll
    for (int p = 0; p < kPieces; ++p)
#pragma unroll
        for (int w = 0; w < 8; ++w) xv[p][w] = xq[8 * (sub + kLanes * p) + w];
    const QScalars s = a.qv.qs[qi];
    const uint64_t first = (uint64_t)blockIdx.x * kSlots * 4;
    for (int r = 0; r < 4; ++r) {
        const uint64_t pos = first + (uint64_t)r * kSlots + slot;
        const bool live = pos < a.per_query;
        const uint64_t item = qi * a.per_query + (live ? pos : 0);
        const uint32_t id = a.ids[item];
        const uint4 *row = (const uint4 *)(a.ix.codes + (uint64_t)id * (CD / 2));
        uint4 w[kPieces];
#pragma unroll
        for (int p = 0; p < kPieces; ++p) w[p] = row[sub + kLanes * p];
        int acc = 0;
#pragma unroll
        for (int p = 0; p < kPieces; ++p) {
            const uint32_t ws[4] = {w[p].x, w[p].y, w[p].z, w[p].w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                acc = __builtin_amdgcn_sdot4((int)(ws[k] & 0x0F0F0F0Fu), (int)xv[p][2 * k], acc, false);
                acc = __builtin_amdgcn_sdot4((int)((ws[k] >> 4) & 0x0F0F0F0Fu), (int)xv[p][2 * k + 1], acc, false);
            }
        }
        acc += __shfl_xor(acc, 1);
        acc += __shfl_xor(acc, 2);
        acc += __shfl_xor(acc, 4);
        if (sub != 0 || !live) continue;
        const float2 h = a.ix.hdr[id];
        const float scale = s.int8_scale;
        float est, lower;
        if (!long_scale_ok(h.y) || !(scale > 0.0f) || !__builtin_isfinite(scale) || s.query_norm <= 0.0f) {
            est = h.x + s.query_norm_sqr;
            lower = est;
        } else {
            const double centered = (double)acc - 7.5 * (double)s.int8_code_sum;
            const double sip = (double)scale * centered;
            const double est_ip = (double)h.y * sip;
            est = h.x + s.query_norm_sqr - (float)est_ip;
            lower = est;
            if (a.out_f2) {
                const double db_err = (double)a.ix.sfac[id].y * (double)s.query_norm;
                lower = est - (float)((double)a.eps0 * db_err);
            }
        }
        a.out_f[item] = est;
        if (a.out_f2) a.out_f2[item] = lower;
        if (a.out_dot) a.out_dot[item] = acc;
    }
}

// ---------------------------------------------------------------- C1 -----
// FP32 query x 4-bit code (queryDistanceLong + longCodeIpsAvx512,
// space_rabitq.h:1551-1607): 16 lane accumulators per sum, lane j takes
// pairs j, j+16, ...; remaining bits fused (FMA), MSB half by masked add.
template <bool kFixed>
__global__ void __launch_bounds__(kScanThreads) k_asym(ScanArgs a) {
    __shared__ float qsh[2048];
    uint64_t qi, pos, item;
    bool live;
    if (kFixed) {
        live = fixed_item(a, qi, pos);
        item = qi * a.per_query + pos;
        if (a.cd <= 2048) {
            const float4 *src = (const float4 *)(a.qv.q + qi * a.cd);
            for (uint32_t i = threadIdx.x; i < a.cd / 4; i += blockDim.x) ((float4 *)qsh)[i] = src[i];
            __syncthreads();
        }
    } else {
        item = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
        live = item < a.n_pairs;
        qi = live ? a.qidx[item] : 0;
    }
    if (!live) return;
    const uint32_t id = a.ids[item];
    const QScalars s = a.qv.qs[qi];
    const float2 h = a.ix.hdr[id];
    if (!long_scale_ok(h.y)) { a.out_f[item] = h.x + s.query_norm_sqr; return; }
    const float *qg = (kFixed && a.cd <= 2048) ? qsh : a.qv.q + qi * a.cd;  // above 2048 from global memory
    const uint4 *row = (const uint4 *)(a.ix.codes + (uint64_t)id * (a.cd / 2));
    float sh[16], rm[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) { sh[j] = 0.0f; rm[j] = 0.0f; }
    for (uint32_t c = 0; c < a.cd / 32; ++c) {
        const uint4 w = row[c];
        const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
        const float4 *qp = (const float4 *)(qg + 32 * c);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float4 q0 = qp[2 * k], q1 = qp[2 * k + 1];  // dims of code bytes 4k..4k+3
            const float qe[4] = {q0.x, q0.z, q1.x, q1.z};
            const float qo[4] = {q0.y, q0.w, q1.y, q1.w};
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int j = 4 * k + b;
                const uint32_t byte = (ws[k] >> (8 * b)) & 0xFFu;
                const uint32_t lo = byte & 0x0Fu, hi = byte >> 4;
                rm[j] = fmaf((float)(lo & 7u), qe[b], rm[j]);
                rm[j] = fmaf((float)(hi & 7u), qo[b], rm[j]);
                if (lo > 7u) sh[j] = sh[j] + qe[b];
                if (hi > 7u) sh[j] = sh[j] + qo[b];
            }
        }
    }
    const float short_ip = reduce16(sh) - s.half_sum;
    const float rem_ip = reduce16(rm);
    a.out_f[item] = finish_long(h.x, h.y, s.query_norm_sqr, s.half_sum, short_ip, rem_ip);
}

// ---------------------------------------------------------------- C3 -----
// FP32 remaining three bits after the gate (query_distance_with_paper_msb_
// batch_by_id, space_rabitq.h:3972-4073).
__global__ void __launch_bounds__(kScanThreads) k_rem_fp32(ScanArgs a) {
    const uint64_t item = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= a.n_pairs) return;
    const uint64_t qi = a.qidx[item];
    const uint32_t id = a.ids[item];
    const QScalars s = a.qv.qs[qi];
    const float2 h = a.ix.hdr[id];
    if (!long_scale_ok(h.y)) { a.out_f[item] = h.x + s.query_norm_sqr; return; }
    const float *qg = a.qv.q + qi * a.cd;
    const uint4 *row = (const uint4 *)(a.ix.codes + (uint64_t)id * (a.cd / 2));
    float rm[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) rm[j] = 0.0f;
    for (uint32_t c = 0; c < a.cd / 32; ++c) {
        const uint4 w = row[c];
        const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
        const float4 *qp = (const float4 *)(qg + 32 * c);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float4 q0 = qp[2 * k], q1 = qp[2 * k + 1];
            const float qe[4] = {q0.x, q0.z, q1.x, q1.z};
            const float qo[4] = {q0.y, q0.w, q1.y, q1.w};
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int j = 4 * k + b;
                const uint32_t byte = (ws[k] >> (8 * b)) & 0xFFu;
                rm[j] = fmaf((float)(byte & 7u), qe[b], rm[j]);
                rm[j] = fmaf((float)((byte >> 4) & 7u), qo[b], rm[j]);
            }
        }
    }
    a.out_f[item] = finish_long(h.x, h.y, s.query_norm_sqr, s.half_sum, a.short_ips[item], reduce16(rm));
}

// ---------------------------------------------------------------- C6 -----
// INT8 remaining three bits (query_distance_with_quantized_paper_msb_batch_
// by_id + finishQuantizedTraversalDistance, space_rabitq.h:4234-4358).
__global__ void __launch_bounds__(kScanThreads) k_rem_int8(ScanArgs a) {
    const uint64_t item = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= a.n_pairs) return;
    const uint64_t qi = a.qidx[item];
    const uint32_t id = a.ids[item];
    const QScalars s = a.qv.qs[qi];
    const float2 h = a.ix.hdr[id];
    if (!long_scale_ok(h.y)) { a.out_f[item] = h.x + s.query_norm_sqr; return; }
    const uint4 *row = (const uint4 *)(a.ix.codes + (uint64_t)id * (a.cd / 2));
    const uint32_t *xg = (const uint32_t *)(a.qv.x8p + qi * a.cd);
    int acc = 0;
    for (uint32_t c = 0; c < a.cd / 32; ++c) {
        const uint4 w = row[c];
        const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t d = 4 * c + k;
            acc = __builtin_amdgcn_sdot4((int)(ws[k] & 0x07070707u), (int)xg[2 * d], acc, false);
            acc = __builtin_amdgcn_sdot4((int)((ws[k] >> 4) & 0x07070707u), (int)xg[2 * d + 1], acc, false);
        }
    }
    const double centered = (double)s.int8_scale * ((double)acc - 3.5 * (double)s.int8_code_sum);
    const double signed_ip = 8.0 * (double)a.short_ips[item] + centered;
    a.out_f[item] = h.x + s.query_norm_sqr - (float)((double)h.y * signed_ip);
    if (a.out_dot) a.out_dot[item] = acc;
}

// ---------------------------------------------------------------- C4 -----
// FP32 1-bit gate (compute_full_paper_prune_estimate_batch_by_id,
// space_rabitq.h:4427-4508): 16 lane accumulators over the MSB plane.
__global__ void __launch_bounds__(kScanThreads) k_gate_fp32(ScanArgs a) {
    const uint64_t item = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= a.n_pairs) return;
    const uint64_t qi = a.qidx[item];
    const uint32_t id = a.ids[item];
    const QScalars s = a.qv.qs[qi];
    const float *qg = a.qv.q + qi * a.cd;
    const uint8_t *grow = gate_row(a.ix.gate, a.cd, id);
    const uint2 *m = (const uint2 *)grow;
    float acc[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = 0.0f;
    for (uint32_t c = 0; c < a.cd / 64; ++c) {  // 8 bytes of MSB = 64 dims (code_dim >= 64)
        const uint2 w = m[c];
        const uint32_t ws[2] = {w.x, w.y};
#pragma unroll
        for (int k = 0; k < 2; ++k) {
#pragma unroll
            for (int half = 0; half < 2; ++half) {  // one 16-dim group per 16 bits
                const uint32_t bits = (ws[k] >> (16 * half)) & 0xFFFFu;
                const float4 *qp = (const float4 *)(qg + 64 * c + 32 * k + 16 * half);
                const float4 q4[4] = {qp[0], qp[1], qp[2], qp[3]};
                const float qv[16] = {q4[0].x, q4[0].y, q4[0].z, q4[0].w, q4[1].x, q4[1].y, q4[1].z, q4[1].w,
                                      q4[2].x, q4[2].y, q4[2].z, q4[2].w, q4[3].x, q4[3].y, q4[3].z, q4[3].w};
#pragma unroll
                for (int j = 0; j < 16; ++j) acc[j] = acc[j] + masked_q(bits, j, qv[j]);
            }
        }
    }
    const float short_ip = reduce16(acc) - s.half_sum;
    a.out_est[item] = finish_gate(short_ip, gate_factors(grow, a.cd), gate_valid(grow, a.cd), s.query_norm,
                                  s.query_norm_sqr, a.eps0);
}

// ---------------------------------------------------------------- C5 -----
// INT8 1-bit gate (compute_int_paper_prune_estimate_batch_by_id,
// space_rabitq.h:4991-5093): selected = sum of x_i over set MSB bits (exact).
__global__ void __launch_bounds__(kScanThreads) k_gate_int8(ScanArgs a) {
    const uint64_t item = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= a.n_pairs) return;
    const uint64_t qi = a.qidx[item];
    const uint32_t id = a.ids[item];
    const QScalars s = a.qv.qs[qi];
    const uint8_t *grow = gate_row(a.ix.gate, a.cd, id);
    const float4 f = gate_factors(grow, a.cd);
    const uint8_t fv = gate_valid(grow, a.cd);
    PaperEstimate r;
    r.lower_bound = r.short_ip = r.alpha = r.ip_hat = r.error_bound = 0.0f;
    r.valid = 0; r.pad[0] = r.pad[1] = r.pad[2] = 0;
    if (!fv || s.query_norm <= 0.0f || f.z <= 0.0f || !__builtin_isfinite(f.z)) { a.out_est[item] = r; return; }
    const uint32_t *xg = (const uint32_t *)(a.qv.x8 + qi * a.cd);
    const uint2 *m = (const uint2 *)grow;
    int sel = 0;
    for (uint32_t c = 0; c < a.cd / 64; ++c) {  // 8 bytes of MSB = 64 dims (code_dim >= 64)
        const uint2 w = m[c];
        const uint32_t ws[2] = {w.x, w.y};
#pragma unroll
        for (int k = 0; k < 2; ++k) {
#pragma unroll
            for (int nib = 0; nib < 8; ++nib) {  // 4 dims per MSB nibble -> one dot4
                const uint32_t bits = (ws[k] >> (4 * nib)) & 0xFu;
                // expand 4 bits to a byte mask 0x01 per selected dim
                const uint32_t mask = nibble_bytes(bits);
                const uint32_t d4 = 64 * c + 32 * k + 4 * nib;  // first dim of this group
                sel = __builtin_amdgcn_sdot4((int)mask, (int)xg[d4 / 4], sel, false);
            }
        }
    }
    const float short_ip = 0.5f * s.int8_scale * (float)((int64_t)2 * (int64_t)sel - s.int8_code_sum);
    a.out_est[item] = finish_gate(short_ip, f, fv, s.query_norm, s.query_norm_sqr, a.eps0);
}

// The B1 codec's symmetric 1-bit gate (compute_b1_paper_prune_estimate_
// batch_by_id, space_rabitq.h:4607-4753). With the +-1 sign code x, the
// masked sum sel = sum of x_i over set MSB bits gives the reference's
// cd - 2 mismatches as 2 sel - sum(x), exactly.
__global__ void __launch_bounds__(kScanThreads) k_gate_b1(ScanArgs a) {
    const uint64_t item = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= a.n_pairs) return;
    const uint64_t qi = a.qidx[item];
    const uint32_t id = a.ids[item];
    const QScalars s = a.qv.qs[qi];
    const uint8_t *grow = gate_row(a.ix.gate, a.cd, id);
    const float4 f = gate_factors(grow, a.cd);
    const uint8_t fv = gate_valid(grow, a.cd);
    PaperEstimate r;
    r.lower_bound = r.short_ip = r.alpha = r.ip_hat = r.error_bound = 0.0f;
    r.valid = 0; r.pad[0] = r.pad[1] = r.pad[2] = 0;
    const bool query_valid = s.int8_scale > 0.0f && __builtin_isfinite(s.int8_scale) && s.query_norm > 0.0f;
    if (!query_valid || !fv || f.z <= 0.0f || !__builtin_isfinite(f.z) || f.y <= 0.0f) { a.out_est[item] = r; return; }
    const uint32_t *xg = (const uint32_t *)(a.qv.x8 + qi * a.cd);
    const uint2 *m = (const uint2 *)grow;
    int sel = 0;
    for (uint32_t c = 0; c < a.cd / 64; ++c) {
        const uint2 w = m[c];
        const uint32_t ws[2] = {w.x, w.y};
#pragma unroll
        for (int k2 = 0; k2 < 2; ++k2) {
#pragma unroll
            for (int nib = 0; nib < 8; ++nib) {
                const uint32_t mask = nibble_bytes((ws[k2] >> (4 * nib)) & 0xFu);
                const uint32_t d4 = 64 * c + 32 * k2 + 4 * nib;
                sel = __builtin_amdgcn_sdot4((int)mask, (int)xg[d4 / 4], sel, false);
            }
        }
    }
    const float signed_dot = (float)((int64_t)2 * (int64_t)sel - s.int8_code_sum);
    const float short_ip = signed_dot * s.int8_scale / 8.0f;
    a.out_est[item] = finish_gate_b1(short_ip, f, s.query_norm, s.query_norm_sqr, s.int8_error_norm, s.b1_alpha, a.eps0);
}

// ---------------------------------------------------------------- E1 -----
// Residual-4 rerank (compute_residual_distance_intervals_batch_by_id,
// space_rabitq.h:5432-5548): per 16-dim block, 8 lane products
// fma(hi, odd, lo*even), lane sum in order, ips = fma(scale, block, ips);
// blocks whose fp16 scale decodes to 0 or non-finite are skipped. With
// record_form (the per-record computeResidualDistanceIntervalFromRecord,
// :1723-1735, which single-record calls and spaces of several centroids
// take) the lane products are fma(lo, even, hi*odd).
template <bool kFixed>
__global__ void __launch_bounds__(kScanThreads) k_residual(ScanArgs a) {
    __shared__ float qsh[2048];
    uint64_t qi, pos, item;
    bool live;
    if (kFixed) {
        live = fixed_item(a, qi, pos);
        item = qi * a.per_query + pos;
        if (a.cd <= 2048) {
            const float4 *src = (const float4 *)(a.qv.q + qi * a.cd);
            for (uint32_t i = threadIdx.x; i < a.cd / 4; i += blockDim.x) ((float4 *)qsh)[i] = src[i];
            __syncthreads();
        }
    } else {
        item = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
        live = item < a.n_pairs;
        qi = live ? a.qidx[item] : 0;
    }
    if (!live) return;
    const uint32_t id = a.ids[item];
    const QScalars s = a.qv.qs[qi];
    const float *qg = (kFixed && a.cd <= 2048) ? qsh : a.qv.q + qi * a.cd;  // above 2048 from global memory
    const uint32_t blocks = a.cd / 16;
    const uint16_t *sc = a.ix.rscale + (uint64_t)id * blocks;
    const uint2 *rc = (const uint2 *)(a.ix.rcodes + (uint64_t)id * (a.cd / 2));
    float ips = 0.0f;
    double errsq = 0.0;
    for (uint32_t b = 0; b < blocks; ++b) {
        const float scale = half_to_float_ref(sc[b]);
        if (scale == 0.0f || !__builtin_isfinite(scale)) continue;
        const uint2 w = rc[b];
        const float4 *qp = (const float4 *)(qg + 16 * b);
        const float4 q4[4] = {qp[0], qp[1], qp[2], qp[3]};
        const float qv[16] = {q4[0].x, q4[0].y, q4[0].z, q4[0].w, q4[1].x, q4[1].y, q4[1].z, q4[1].w,
                              q4[2].x, q4[2].y, q4[2].z, q4[2].w, q4[3].x, q4[3].y, q4[3].z, q4[3].w};
        float bip = 0.0f;
#pragma unroll
        for (int l = 0; l < 8; ++l) {
            const uint32_t byte = ((l < 4 ? w.x : w.y) >> (8 * (l & 3))) & 0xFFu;
            int lo = (int)(byte & 0x0Fu), hi = (int)(byte >> 4);
            lo -= (lo > 7) ? 16 : 0;
            hi -= (hi > 7) ? 16 : 0;
            // the batch loop fuses the odd product into the even one; the
            // per-record form (residualInt4IpAvx2, :2278-2325) the other way
            float prod;
            if (a.record_form) prod = fmaf((float)lo, qv[2 * l], (float)hi * qv[2 * l + 1]);
            else prod = fmaf((float)hi, qv[2 * l + 1], (float)lo * qv[2 * l]);
            bip = bip + prod;
        }
        ips = fmaf(scale, bip, ips);
        const double pde = 0.5 * fabs((double)scale);
        errsq += 16.0 * pde * pde;
    }
    const float dist = a.long_in[item] - 2.0f * ips;
    const float err = (float)(2.0 * (double)s.query_norm * sqrt(errsq));
    Interval iv;
    iv.estimate = dist;
    iv.lower_bound = dist - err;
    iv.upper_bound = dist + err;
    a.out_iv[item] = iv;
    if (a.out_f) a.out_f[item] = dist;
}

// ---------------------------------------------------------------- C8 -----
// Symmetric code<->code distance (distanceBetweenEncoded / symmetric
// BuildDistanceFromPrepared, space_rabitq.h:2634-2747, K = 1): ip4 =
// sum (2a-15)(2b-15) exactly, then the double-precision finish with the lhs
// (prepared) side first. Pairs (qidx[i] = lhs row, ids[i] = rhs row).
__global__ void __launch_bounds__(kScanThreads) k_symmetric(ScanArgs a) {
    const uint64_t item = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= a.n_pairs) return;
    const uint32_t l = a.qidx[item], r = a.ids[item];
    const int acc = symmetric_ip4_rows(a.ix.codes + (uint64_t)l * (a.cd / 2), a.ix.codes + (uint64_t)r * (a.cd / 2),
                                       a.cd);
    a.out_f[item] = symmetric_finish(a.ix.hdr[l], a.ix.hdr[r], acc);
    if (a.out_dot) a.out_dot[item] = acc;
}

// E1 for fixed lists, two-phase form. Phase 1: 8 lanes per candidate load
// the residual row with 128-byte coalesced pieces (lane `sub` owns 16-byte
// pieces sub, sub+8, ... = blocks 2(sub+8p), 2(sub+8p)+1) and compute each
// block's 8-lane product sum in the reference order; block sums and decoded
// scales go to LDS. Phase 2: one lane per candidate runs the order-dependent
// recurrences ips = fma(scale, block_ip, ips) and the error norm over the
// blocks in index order.
template <int CD>
__global__ void __launch_bounds__(kScanThreads) k_residual_grouped(ScanArgs a) {
    constexpr int kLanes = 8;
    constexpr int kSlots = kScanThreads / kLanes;      // candidates per pass
    constexpr int kBlocks = CD / 16;
    constexpr int kPieces = CD / 2 / 16 / kLanes;       // 16-byte pieces per lane
    __shared__ float qsh[CD];
    __shared__ float bip_s[kSlots][kBlocks + 1];
    __shared__ float scl_s[kSlots][kBlocks + 1];
    const uint64_t qi = blockIdx.y;
    const float4 *src = (const float4 *)(a.qv.q + qi * CD);
    for (uint32_t i = threadIdx.x; i < CD / 4; i += blockDim.x) ((float4 *)qsh)[i] = src[i];
    const QScalars s = a.qv.qs[qi];
    const int sub = threadIdx.x & (kLanes - 1);
    const int slot = threadIdx.x / kLanes;
    const uint64_t first = (uint64_t)blockIdx.x * kSlots * 2;
    __syncthreads();
    for (int r = 0; r < 2; ++r) {
        const uint64_t pos = first + (uint64_t)r * kSlots + slot;
        const bool live = pos < a.per_query;
        const uint64_t item = qi * a.per_query + (live ? pos : 0);
        const uint32_t id = a.ids[item];
        const uint4 *row = (const uint4 *)(a.ix.rcodes + (uint64_t)id * (CD / 2));
        const uint32_t *srow = (const uint32_t *)(a.ix.rscale + (uint64_t)id * kBlocks);
        uint4 w[kPieces];
        uint32_t sc2[kPieces];
#pragma unroll
        for (int p = 0; p < kPieces; ++p) {
            w[p] = row[sub + kLanes * p];
            sc2[p] = srow[sub + kLanes * p];  // fp16 scales of the piece's two blocks
        }
#pragma unroll
        for (int p = 0; p < kPieces; ++p) {
            const uint32_t ws[4] = {w[p].x, w[p].y, w[p].z, w[p].w};
#pragma unroll
            for (int hb = 0; hb < 2; ++hb) {
                const int blk = 2 * (sub + kLanes * p) + hb;
                const float4 *qp = (const float4 *)(qsh + 16 * blk);
                const float4 q4[4] = {qp[0], qp[1], qp[2], qp[3]};
                const float qv[16] = {q4[0].x, q4[0].y, q4[0].z, q4[0].w, q4[1].x, q4[1].y, q4[1].z, q4[1].w,
                                      q4[2].x, q4[2].y, q4[2].z, q4[2].w, q4[3].x, q4[3].y, q4[3].z, q4[3].w};
                float bip = 0.0f;
#pragma unroll
                for (int l = 0; l < 8; ++l) {
                    const uint32_t byte = (ws[2 * hb + (l >> 2)] >> (8 * (l & 3))) & 0xFFu;
                    int lo = (int)(byte & 0x0Fu), hi = (int)(byte >> 4);
                    lo -= (lo > 7) ? 16 : 0;
                    hi -= (hi > 7) ? 16 : 0;
                    const float pl = (float)lo * qv[2 * l];
                    bip = bip + fmaf((float)hi, qv[2 * l + 1], pl);
                }
                bip_s[slot][blk] = bip;
                scl_s[slot][blk] = half_to_float_ref((uint16_t)(sc2[p] >> (16 * hb)));
            }
        }
        __syncthreads();
        if (sub == 0 && live) {
            float ips = 0.0f;
            double errsq = 0.0;
            for (int b = 0; b < kBlocks; ++b) {
                const float scale = scl_s[slot][b];
                if (scale == 0.0f || !__builtin_isfinite(scale)) continue;
                ips = fmaf(scale, bip_s[slot][b], ips);
                const double pde = 0.5 * fabs((double)scale);
                errsq += 16.0 * pde * pde;
            }
            const float dist = a.long_in[item] - 2.0f * ips;
            const float err = (float)(2.0 * (double)s.query_norm * sqrt(errsq));
            Interval iv;
            iv.estimate = dist;
            iv.lower_bound = dist - err;
            iv.upper_bound = dist + err;
            a.out_iv[item] = iv;
            if (a.out_f) a.out_f[item] = dist;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------- record gate -----
// compute_paper_prune_estimate (space_rabitq.h:3926-3959): the gate read
// from the compact record. short_ip over the nibble MSBs in the pair-lane
// order of shortCodeIpFromFullCodeAvx512 (:1307-1329); the lower bound is
// (norm_sqr + |q|^2) - two_norms * ip_upper with the product contracted,
// as the reference compiler emits it.
__global__ void __launch_bounds__(kScanThreads) k_paper_record(ScanArgs a) {
    const uint64_t item = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= a.n_pairs) return;
    const uint64_t qi = a.qidx[item];
    const uint32_t id = a.ids[item];
    const QScalars s = a.qv.qs[qi];
    const float2 h = a.ix.hdr[id];
    const float short_scale = a.ix.sfac[id].x;
    PaperEstimate r;
    r.lower_bound = r.short_ip = r.alpha = r.ip_hat = r.error_bound = 0.0f;
    r.valid = 0; r.pad[0] = r.pad[1] = r.pad[2] = 0;
    if (a.cd <= 1 || a.eps0 < 0.0f || !__builtin_isfinite(a.eps0) || h.x <= 0.0f || s.query_norm <= 0.0f ||
        short_scale <= 0.0f || !__builtin_isfinite(short_scale)) { a.out_est[item] = r; return; }
    const float data_norm = sqrtf(h.x);
    const float alpha_raw = 4.0f * data_norm / (short_scale * sqrtf((float)a.cd));
    if (!(alpha_raw > 0.0f) || !__builtin_isfinite(alpha_raw)) { a.out_est[item] = r; return; }
    const float alpha = (alpha_raw < 1.0f) ? alpha_raw : 1.0f;
    const float *qg = a.qv.q + qi * a.cd;
    const uint4 *row = (const uint4 *)(a.ix.codes + (uint64_t)id * (a.cd / 2));
    float sh[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) sh[j] = 0.0f;
    for (uint32_t c = 0; c < a.cd / 32; ++c) {
        const uint4 w = row[c];
        const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
        const float4 *qp = (const float4 *)(qg + 32 * c);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float4 q0 = qp[2 * k], q1 = qp[2 * k + 1];
            const float qe[4] = {q0.x, q0.z, q1.x, q1.z};
            const float qo[4] = {q0.y, q0.w, q1.y, q1.w};
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int j = 4 * k + b;
                const uint32_t bits = ws[k] >> (8 * b);
                sh[j] = sh[j] + masked_q(bits, 3, qe[b]);
                sh[j] = sh[j] + masked_q(bits, 7, qo[b]);
            }
        }
    }
    const float short_ip = reduce16(sh) - s.half_sum;
    const float cross = short_ip * short_scale;
    const float two_norms = 2.0f * data_norm * s.query_norm;
    const float ip_hat = cross / two_norms;
    const float alpha2 = alpha * alpha;
    const float t = (1.0f - alpha2) / alpha2;
    const float eb = sqrtf(t > 0.0f ? t : 0.0f) * a.eps0 / sqrtf((float)(a.cd - 1));
    float up = ip_hat + eb;
    up = (up < 1.0f) ? up : 1.0f;
    up = (-1.0f < up) ? up : -1.0f;
    r.lower_bound = fmaf(-two_norms, up, h.x + s.query_norm_sqr);
    r.short_ip = short_ip;
    r.alpha = alpha;
    r.ip_hat = ip_hat;
    r.error_bound = eb;
    r.valid = __builtin_isfinite(r.lower_bound) ? 1 : 0;
    a.out_est[item] = r;
}

// ------------------------------------------------------------------ B1 -----
// query_distance_b1 (space_rabitq.h:3742-3755): the full 4-bit code against
// the query's +/-1 sign code (prepared in the INT8 slots), exact integer dot,
// est_ip = long_scale * (dot - 7.5 * sum) in double.
__global__ void __launch_bounds__(kScanThreads) k_b1(ScanArgs a) {
    const uint64_t item = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= a.n_pairs) return;
    const uint64_t qi = a.qidx[item];
    const uint32_t id = a.ids[item];
    const QScalars s = a.qv.qs[qi];
    const float2 h = a.ix.hdr[id];
    if (!long_scale_ok(h.y)) { a.out_f[item] = h.x + s.query_norm_sqr; return; }
    const uint4 *row = (const uint4 *)(a.ix.codes + (uint64_t)id * (a.cd / 2));
    const uint32_t *xg = (const uint32_t *)(a.qv.x8p + qi * a.cd);
    int acc = 0;
    for (uint32_t c = 0; c < a.cd / 32; ++c) {
        const uint4 w = row[c];
        const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t d = 4 * c + k;
            acc = __builtin_amdgcn_sdot4((int)(ws[k] & 0x0F0F0F0Fu), (int)xg[2 * d], acc, false);
            acc = __builtin_amdgcn_sdot4((int)((ws[k] >> 4) & 0x0F0F0F0Fu), (int)xg[2 * d + 1], acc, false);
        }
    }
    const double signed_long_ip = (double)acc - 7.5 * (double)s.int8_code_sum;
    const double est_ip = (double)h.y * signed_long_ip;
    a.out_f[item] = h.x + s.query_norm_sqr - (float)est_ip;
    if (a.out_dot) a.out_dot[item] = acc;
}

// estimateTwiceResidualIp (space_rabitq.h:2612-2626) with the centroid
// delta context r_a - r_b (buildCentroidDeltaContexts, :2564-2610) as the
// query: long_scale * (8 short_ip + remaining_ip - 7 half_sum), the inner
// products in asym_one's lane order, the delta taken on the fly.
__device__ __forceinline__ float twice_residual_ip(const float *ra, const float *rb, float half_sum,
                                                   const uint8_t *code_row, uint32_t cd, float2 h) {
    if (!long_scale_ok(h.y)) return 0.0f;
    const uint4 *row = (const uint4 *)code_row;
    float sh[16], rm[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) { sh[j] = 0.0f; rm[j] = 0.0f; }
    for (uint32_t c = 0; c < cd / 32; ++c) {
        const uint4 w = row[c];
        const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int j = 4 * k + b;
                const uint32_t i = 32 * c + 8 * k + 2 * b;  // dims i (low nibble), i + 1 (high)
                const float qe = ra[i] - rb[i], qo = ra[i + 1] - rb[i + 1];
                const uint32_t byte = (ws[k] >> (8 * b)) & 0xFFu;
                const uint32_t lo = byte & 0x0Fu, hi = byte >> 4;
                rm[j] = fmaf((float)(lo & 7u), qe, rm[j]);
                rm[j] = fmaf((float)(hi & 7u), qo, rm[j]);
                if (lo > 7u) sh[j] = sh[j] + qe;
                if (hi > 7u) sh[j] = sh[j] + qo;
            }
        }
    }
    const float short_ip = reduce16(sh) - half_sum;
    const float t = 8.0f * short_ip + reduce16(rm);
    return h.y * fmaf(-7.0f, half_sum, t);
}

// The cross-centroid terms of distanceBetweenEncoded (:2634-2683) on top of
// k_symmetric's distances, for pairs whose records hold different
// centroids: |c_a - c_b|^2 enters the invalid-record sum, else the distance
// gains |c_a - c_b|^2 + 2<d, r_l> - 2<d, r_r> estimated with d = r_a - r_b.
__global__ void __launch_bounds__(kScanThreads) k_sym_cross(ScanArgs a) {
    const uint64_t item = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= a.n_pairs) return;
    const uint32_t l = a.qidx[item], r = a.ids[item];
    const uint32_t K = a.centroid_count;
    const uint32_t ca = a.ix.cid[l], cb = a.ix.cid[r];
    if (!(ca < K && cb < K && ca != cb)) return;
    const float2 pt = a.pair_terms[ca * K + cb];
    const float2 lh = a.ix.hdr[l], rh = a.ix.hdr[r];
    const double ln = sqrt((0.0 < (double)lh.x) ? (double)lh.x : 0.0);
    const double rn = sqrt((0.0 < (double)rh.x) ? (double)rh.x : 0.0);
    if (!(ln > 0.0) || !(rn > 0.0) || !(lh.y > 0.0f) || !(rh.y > 0.0f) || !__builtin_isfinite(lh.y) ||
        !__builtin_isfinite(rh.y)) {
        a.out_f[item] = (float)((double)lh.x + (double)rh.x + (double)pt.x);
        return;
    }
    const float *ra = a.rot_centroids + (uint64_t)ca * a.cd, *rb = a.rot_centroids + (uint64_t)cb * a.cd;
    const float tl = twice_residual_ip(ra, rb, pt.y, a.ix.codes + (uint64_t)l * (a.cd / 2), a.cd, lh);
    const float tr = twice_residual_ip(ra, rb, pt.y, a.ix.codes + (uint64_t)r * (a.cd / 2), a.cd, rh);
    a.out_f[item] = a.out_f[item] + ((pt.x + tl) - tr);
}

// Pair terms of buildCentroidDeltaContexts (:2564-2610), one lane per
// (a, b): |c_a - c_b|^2 summed in FP64 in index order (stored as float),
// and the half sum of r_a - r_b over coordinate pairs in float.
__global__ void k_centroid_pairs(const float *c, const float *rc, uint32_t K, uint32_t dim, uint32_t cd,
                                 float2 *pairs) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= K * K) return;
    const uint32_t ia = e / K, ib = e % K;
    const float *ca = c + (uint64_t)ia * dim, *cb = c + (uint64_t)ib * dim;
    double pn = 0.0;
    for (uint32_t i = 0; i < dim; ++i) {
        const double d = (double)ca[i] - (double)cb[i];
        pn = fma(d, d, pn);
    }
    const float *ra = rc + (uint64_t)ia * cd, *rb = rc + (uint64_t)ib * cd;
    float half = 0.0f;
    for (uint32_t p = 0; p < cd / 2; ++p) {
        const float ev = ra[2 * p] - rb[2 * p], od = ra[2 * p + 1] - rb[2 * p + 1];
        half = half + (ev + od);
    }
    pairs[e] = make_float2((float)pn, half * 0.5f);
}

}  // namespace qg

extern "C" {

int qg_launch_centroid_pairs(const float *centroids, const float *rot_centroids, uint32_t K, uint32_t dim,
                             uint32_t cd, float2 *pairs, hipStream_t st) {
    const uint32_t total = K * K;
    hipLaunchKernelGGL(qg::k_centroid_pairs, dim3((total + 63) / 64), dim3(64), 0, st, centroids, rot_centroids, K,
                       dim, cd, pairs);
    return (int)hipGetLastError();
}

enum { QG_K_INT8_FULL = 0, QG_K_ASYM = 1, QG_K_REM_FP32 = 2, QG_K_REM_INT8 = 3, QG_K_GATE_FP32 = 4,
       QG_K_GATE_INT8 = 5, QG_K_RESIDUAL = 6, QG_K_SYMMETRIC = 7, QG_K_PAPER_RECORD = 8, QG_K_B1 = 9,
       QG_K_GATE_B1 = 11 };

int qg_launch_scan(int kernel, int fixed, const qg::ScanArgs *a, hipStream_t stream) {
    const dim3 block(qg::kScanThreads);
    if (fixed) {
        if (a->per_query == 0 || a->qv.nq == 0) return 0;
        const dim3 grid((unsigned)((a->per_query + qg::kScanThreads - 1) / qg::kScanThreads), (unsigned)a->qv.nq);
        switch (kernel) {
            case QG_K_INT8_FULL: {
                const unsigned per_block = qg::kScanThreads / 8 * 4;
                const dim3 g2((unsigned)((a->per_query + per_block - 1) / per_block), (unsigned)a->qv.nq);
                if (a->cd == 1024) hipLaunchKernelGGL(qg::k_int8_full_grouped<1024>, g2, block, 0, stream, *a);
                else if (a->cd == 2048) hipLaunchKernelGGL(qg::k_int8_full_grouped<2048>, g2, block, 0, stream, *a);
                else hipLaunchKernelGGL(qg::k_int8_full<true>, grid, block, 0, stream, *a);
                break;
            }
            case QG_K_ASYM: hipLaunchKernelGGL(qg::k_asym<true>, grid, block, 0, stream, *a); break;
            case QG_K_RESIDUAL: {
                const unsigned per_block = qg::kScanThreads / 8 * 2;
                const dim3 g2((unsigned)((a->per_query + per_block - 1) / per_block), (unsigned)a->qv.nq);
                if (a->cd == 1024) hipLaunchKernelGGL(qg::k_residual_grouped<1024>, g2, block, 0, stream, *a);
                else if (a->cd == 2048) hipLaunchKernelGGL(qg::k_residual_grouped<2048>, g2, block, 0, stream, *a);
                else hipLaunchKernelGGL(qg::k_residual<true>, grid, block, 0, stream, *a);
                break;
            }
            default: return (int)hipErrorInvalidValue;
        }
    } else {
        if (a->n_pairs == 0) return 0;
        const dim3 grid((unsigned)((a->n_pairs + qg::kScanThreads - 1) / qg::kScanThreads));
        switch (kernel) {
            case QG_K_INT8_FULL: hipLaunchKernelGGL(qg::k_int8_full<false>, grid, block, 0, stream, *a); break;
            case QG_K_ASYM: hipLaunchKernelGGL(qg::k_asym<false>, grid, block, 0, stream, *a); break;
            case QG_K_REM_FP32: hipLaunchKernelGGL(qg::k_rem_fp32, grid, block, 0, stream, *a); break;
            case QG_K_REM_INT8: hipLaunchKernelGGL(qg::k_rem_int8, grid, block, 0, stream, *a); break;
            case QG_K_GATE_FP32: hipLaunchKernelGGL(qg::k_gate_fp32, grid, block, 0, stream, *a); break;
            case QG_K_GATE_INT8: hipLaunchKernelGGL(qg::k_gate_int8, grid, block, 0, stream, *a); break;
            case QG_K_RESIDUAL: hipLaunchKernelGGL(qg::k_residual<false>, grid, block, 0, stream, *a); break;
            case QG_K_SYMMETRIC:
                hipLaunchKernelGGL(qg::k_symmetric, grid, block, 0, stream, *a);
                if (a->centroid_count > 1) hipLaunchKernelGGL(qg::k_sym_cross, grid, block, 0, stream, *a);
                break;
            case QG_K_PAPER_RECORD: hipLaunchKernelGGL(qg::k_paper_record, grid, block, 0, stream, *a); break;
            case QG_K_B1: hipLaunchKernelGGL(qg::k_b1, grid, block, 0, stream, *a); break;
            case QG_K_GATE_B1: hipLaunchKernelGGL(qg::k_gate_b1, grid, block, 0, stream, *a); break;
            default: return (int)hipErrorInvalidValue;
        }
    }
    return (int)hipGetLastError();
}

}  // extern "C"
