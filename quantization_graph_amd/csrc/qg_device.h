// qg_device.h -- one-candidate forms of the estimators, for kernels in which
// a single lane owns a (query, candidate) pair (graph search, rerank).
// The arithmetic is the one of the batch kernels in qg_scan.hip, so results
// are bit-identical between the two.
#pragma once

#include "qg_common.h"

namespace qg {

// bit j of `bits` ? q : +0.0, as the AND of q with the sign-extended bit
// (the same value; v_bfe_i32 + v_and instead of compare, hazard wait, select)
__device__ __forceinline__ float masked_q(uint32_t bits, int j, float q) {
    int m = __builtin_amdgcn_sbfe((int)bits, j, 1);
    asm("" : "+v"(m));  // keep the mask an integer: no fold back into compare + select
    return __uint_as_float(__float_as_uint(q) & (uint32_t)m);
}

// four bits b0..b3 -> bytes 0x0b0 .. 0x0b3 (b * 0x204081 places bit i at
// position 7i + i without carries)
__device__ __forceinline__ uint32_t nibble_bytes(uint32_t b4) { return (b4 * 0x204081u) & 0x01010101u; }

// C1: FP32 query x 4-bit code (queryDistanceLong + longCodeIpsAvx512,
// space_rabitq.h:1551-1607).
__device__ __forceinline__ float asym_one(const float *q, const uint8_t *code_row, uint32_t cd, float2 h,
                                          const QScalars &s) {
    if (!long_scale_ok(h.y)) return h.x + s.query_norm_sqr;
    const uint4 *row = (const uint4 *)code_row;
    float sh[16], rm[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) { sh[j] = 0.0f; rm[j] = 0.0f; }
    for (uint32_t c = 0; c < cd / 32; ++c) {
        const uint4 w = row[c];
        const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
        const float4 *qp = (const float4 *)(q + 32 * c);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float4 q0 = qp[2 * k], q1 = qp[2 * k + 1];
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
    return finish_long(h.x, h.y, s.query_norm_sqr, s.half_sum, short_ip, reduce16(rm));
}

// C3: FP32 remaining three bits (query_distance_with_paper_msb_batch_by_id).
__device__ __forceinline__ float rem_fp32_one(const float *q, const uint8_t *code_row, uint32_t cd, float2 h,
                                              const QScalars &s, float short_ip) {
    if (!long_scale_ok(h.y)) return h.x + s.query_norm_sqr;
    const uint4 *row = (const uint4 *)code_row;
    float rm[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) rm[j] = 0.0f;
    for (uint32_t c = 0; c < cd / 32; ++c) {
        const uint4 w = row[c];
        const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
        const float4 *qp = (const float4 *)(q + 32 * c);
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
    return finish_long(h.x, h.y, s.query_norm_sqr, s.half_sum, short_ip, reduce16(rm));
}

// C6: INT8 remaining three bits (finishQuantizedTraversalDistance).
__device__ __forceinline__ float rem_int8_one(const int8_t *x8p, const uint8_t *code_row, uint32_t cd, float2 h,
                                              const QScalars &s, float short_ip) {
    if (!long_scale_ok(h.y)) return h.x + s.query_norm_sqr;
    const uint4 *row = (const uint4 *)code_row;
    const uint32_t *xg = (const uint32_t *)x8p;
    int acc = 0;
    for (uint32_t c = 0; c < cd / 32; ++c) {
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
    const double signed_ip = 8.0 * (double)short_ip + centered;
    return h.x + s.query_norm_sqr - (float)((double)h.y * signed_ip);
}

// E1: residual-4 interval for one candidate.
__device__ __forceinline__ Interval residual_one(const float *q, const uint16_t *scale_row, const uint8_t *rcode_row,
                                                 uint32_t cd, const QScalars &s, float long_d) {
    const uint2 *rc = (const uint2 *)rcode_row;
    float ips = 0.0f;
    double errsq = 0.0;
    for (uint32_t b = 0; b < cd / 16; ++b) {
        const float scale = half_to_float_ref(scale_row[b]);
        if (scale == 0.0f || !__builtin_isfinite(scale)) continue;
        const uint2 w = rc[b];
        const float4 *qp = (const float4 *)(q + 16 * b);
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
            const float pl = (float)lo * qv[2 * l];
            bip = bip + fmaf((float)hi, qv[2 * l + 1], pl);
        }
        ips = fmaf(scale, bip, ips);
        const double pde = 0.5 * fabs((double)scale);
        errsq += 16.0 * pde * pde;
    }
    const float dist = long_d - 2.0f * ips;
    const float err = (float)(2.0 * (double)s.query_norm * sqrt(errsq));
    Interval iv;
    iv.estimate = dist;
    iv.lower_bound = dist - err;
    iv.upper_bound = dist + err;
    return iv;
}

// E1 split in two for block-parallel reranks: the 8-lane product sum of one
// 16-dimension block (q16: the block's query values, w: its 8 code bytes),
// and the order-dependent recurrence ips = fma(scale, block_ip, ips) over
// the blocks in index order (blocks with a zero or non-finite scale
// skipped). residual_chain(scales, {residual_block_ip(b)}, ...) equals
// residual_one(...).estimate bit for bit.
__device__ __forceinline__ float residual_block_ip(const float *q16, uint2 w) {
    const float4 *qp = (const float4 *)q16;
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
        const float pl = (float)lo * qv[2 * l];
        bip = bip + fmaf((float)hi, qv[2 * l + 1], pl);
    }
    return bip;
}

// Block sums of candidates 0..cnt) into work[c * stride + b] by the T
// threads of a workgroup (t = thread index); id_of(c) names candidate c's
// residual code row. Each thread issues U blocks' 8-byte loads before it
// computes any, so U memory round trips overlap.
template <uint32_t T, uint32_t U, typename IdOf>
__device__ __forceinline__ void residual_block_sums(float *work, uint32_t stride, const float *q,
                                                    const uint8_t *__restrict__ rcodes, uint32_t cd, uint32_t cnt,
                                                    IdOf id_of, uint32_t t) {
    const uint32_t nb = cd / 16, lg = (uint32_t)__builtin_ctz(nb), total = cnt * nb;
    for (uint32_t it0 = t; it0 < total; it0 += U * T) {
        uint2 w[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const uint32_t it = it0 + u * T;
            w[u] = it < total ? ((const uint2 *)(rcodes + (uint64_t)id_of(it >> lg) * (cd / 2)))[it & (nb - 1u)]
                              : make_uint2(0u, 0u);
        }
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const uint32_t it = it0 + u * T;
            if (it < total) {
                const uint32_t b = it & (nb - 1u);
                work[(it >> lg) * stride + b] = residual_block_ip(q + 16 * b, w[u]);
            }
        }
    }
}

// nb (= cd / 16) is a multiple of 4; scale_row is 8-byte aligned. The fp16
// scales are widened by the hardware conversion, which equals
// half_to_float_ref on every finite value (subnormals included: f16
// denormals are preserved); the non-finite ones, where the two may differ
// in a NaN payload, are skipped either way.
__device__ __forceinline__ float residual_chain(const uint16_t *scale_row, const float *bip, uint32_t nb,
                                               float long_d) {
    const uint2 *srow = (const uint2 *)scale_row;
    float ips = 0.0f;
#pragma unroll 4
    for (uint32_t b4 = 0; b4 < nb / 4; ++b4) {
        const uint2 sv = srow[b4];
        const uint32_t sw[2] = {sv.x, sv.y};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float scale =
                (float)__builtin_bit_cast(_Float16, (uint16_t)(sw[e >> 1] >> (16 * (e & 1))));
            const bool use = scale != 0.0f && __builtin_isfinite(scale);
            const float next = fmaf(scale, bip[4 * b4 + e], ips);
            ips = use ? next : ips;
        }
    }
    return long_d - 2.0f * ips;
}

// residual_chain for up to T / 8 candidates at once, eight lanes per
// candidate (lane group c = t / 8; nb <= 128, a power of two >= 4). Lane j
// of a group owns the blocks [j * per, (j + 1) * per) of its candidate's
// chain. residual_chain_scales issues the loads of its fp16 scales (at most
// 16 blocks: four 8-byte loads), so that they fly together with the block
// sums' loads that follow; residual_chains_x8 then runs the eight spans in
// block order, each lane continuing from the running sum of the lane before
// it, so every step is residual_chain's step. out[c] = long_d[c] - 2 * ips.
__device__ __forceinline__ uint32_t residual_chain_per(uint32_t nb) { return (nb + 31u) / 32u * 4u; }

template <typename IdOf>
__device__ __forceinline__ void residual_chain_scales(uint2 (&sv)[4], const uint16_t *__restrict__ rscale,
                                                      uint32_t nb, uint32_t cnt, IdOf id_of, uint32_t t) {
    const uint32_t c = t >> 3, per = residual_chain_per(nb), b0 = (t & 7u) * per;
    const bool live = c < cnt && b0 < nb;
    const uint2 *srow = (const uint2 *)(rscale + (uint64_t)(live ? id_of(c) : 0u) * nb) + b0 / 4u;
#pragma unroll
    for (uint32_t u = 0; u < 4; ++u) sv[u] = (live && 4u * u < per) ? srow[u] : make_uint2(0u, 0u);
}

__device__ __forceinline__ void residual_chains_x8(const uint2 (&sv)[4], const float *work, uint32_t stride,
                                                   uint32_t nb, uint32_t cnt, const float *long_d, float *out,
                                                   uint32_t t) {
    const uint32_t c = t >> 3, j = t & 7u, per = residual_chain_per(nb), b0 = j * per;
    const bool live = c < cnt;
    const uint32_t own = (live && b0 < nb) ? per : 0u;
    const float *bip = work + (live ? c : 0u) * stride + (b0 < nb ? b0 : 0u);
    const uint32_t base = (t & 63u) & ~7u;
    float ips = 0.0f;
    for (uint32_t r = 0; r < 8; ++r) {
        if (j == r) {
#pragma unroll
            for (uint32_t u = 0; u < 4; ++u) {
                const uint32_t sw[2] = {sv[u].x, sv[u].y};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float scale =
                        (float)__builtin_bit_cast(_Float16, (uint16_t)(sw[e >> 1] >> (16 * (e & 1))));
                    const bool use = 4u * u + e < own && scale != 0.0f && __builtin_isfinite(scale);
                    const float next = fmaf(scale, 4u * u + e < own ? bip[4 * u + e] : 0.0f, ips);
                    ips = use ? next : ips;
                }
            }
        }
        ips = __shfl(ips, (int)(base + r));  // lane r's running sum, to the whole group
    }
    if (live && j == 0) out[c] = long_d[c] - 2.0f * ips;
}

// ---------------------------------------------------------------------------
// 16-lane group forms: lane j (= lane % 16) of an aligned group of 16 owns
// accumulator j of the reference's 16-lane vector and runs its chain in the
// reference order; the 8/4/2/1 tree of reduce16 is the xor-8/4/2/1
// butterfly over the group. Every lane of the group returns the same value,
// bit-identical to the one-lane forms above. Integer sums are exact and are
// split freely. Rows are first copied to the group's LDS staging area with
// coalesced 16-byte loads (stage_row_group); LDS operations of one
// wavefront complete in order, so a wave-level scheduling barrier orders
// the copy before the strided reads.
__device__ __forceinline__ float group_reduce16(float v) {
    v = v + __shfl_xor(v, 8);
    v = v + __shfl_xor(v, 4);
    v = v + __shfl_xor(v, 2);
    return v + __shfl_xor(v, 1);
}

__device__ __forceinline__ int group_sum_int(int v) {
    v += __shfl_xor(v, 8);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 2);
    return v + __shfl_xor(v, 1);
}

// copy `bytes` (a multiple of 8) from global memory to the group's LDS area
__device__ __forceinline__ void stage_row_group(uint4 *dst, const uint8_t *src, uint32_t bytes, int j) {
    if ((bytes & 15u) == 0) {
        const uint4 *s4 = (const uint4 *)src;
        for (uint32_t i = j; i < bytes / 16; i += 16) dst[i] = s4[i];
    } else {
        const uint2 *s2 = (const uint2 *)src;
        uint2 *d2 = (uint2 *)dst;
        for (uint32_t i = j; i < bytes / 8; i += 16) d2[i] = s2[i];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// C1 on a staged code row: accumulator j sums dims 2(16c + j), 2(16c + j) + 1
// (uniform trip count; whole chunks of 8 steps, unrolled so that their LDS
// reads are in flight together: the mask's asm barrier counts as a convergent
// instruction, which rules out a remainder loop)
__device__ __forceinline__ float asym_group(const float *q, const uint8_t *row_lds, uint32_t cd, float2 h,
                                           const QScalars &s, int j) {
    if (!long_scale_ok(h.y)) return h.x + s.query_norm_sqr;
    float sh = 0.0f, rm = 0.0f;
    auto step = [&](uint32_t c) {
        const uint32_t byte = row_lds[16 * c + j];
        const float2 qq = *(const float2 *)(q + 32 * c + 2 * j);
        const uint32_t lo = byte & 0x0Fu, hi = byte >> 4;
        rm = fmaf((float)(lo & 7u), qq.x, rm);
        rm = fmaf((float)(hi & 7u), qq.y, rm);
        sh = sh + masked_q(byte, 3, qq.x);
        sh = sh + masked_q(byte, 7, qq.y);
    };
    const uint32_t n = cd / 32;
    if ((n & 7u) == 0) {
        for (uint32_t c0 = 0; c0 < n; c0 += 8) {
#pragma unroll
            for (uint32_t u = 0; u < 8; ++u) step(c0 + u);
        }
    } else {
        for (uint32_t c = 0; c < n; ++c) step(c);
    }
    const float short_ip = group_reduce16(sh) - s.half_sum;
    return finish_long(h.x, h.y, s.query_norm_sqr, s.half_sum, short_ip, group_reduce16(rm));
}

// C3 on a staged code row
__device__ __forceinline__ float rem_fp32_group(const float *q, const uint8_t *row_lds, uint32_t cd, float2 h,
                                               const QScalars &s, float short_ip, int j) {
    if (!long_scale_ok(h.y)) return h.x + s.query_norm_sqr;
    float rm = 0.0f;
#pragma unroll 8
    for (uint32_t c = 0; c < cd / 32; ++c) {
        const uint32_t byte = row_lds[16 * c + j];
        const float2 qq = *(const float2 *)(q + 32 * c + 2 * j);
        rm = fmaf((float)(byte & 7u), qq.x, rm);
        rm = fmaf((float)((byte >> 4) & 7u), qq.y, rm);
    }
    return finish_long(h.x, h.y, s.query_norm_sqr, s.half_sum, short_ip, group_reduce16(rm));
}

// C6 on a staged code row: code dwords d = j, j + 16, ...
__device__ __forceinline__ float rem_int8_group(const int8_t *x8p, const uint8_t *row_lds, uint32_t cd, float2 h,
                                               const QScalars &s, float short_ip, int j) {
    if (!long_scale_ok(h.y)) return h.x + s.query_norm_sqr;
    const uint32_t *row = (const uint32_t *)row_lds;
    const uint2 *xg = (const uint2 *)x8p;
    int acc = 0;
#pragma unroll 8
    for (uint32_t i = 0; i < cd / 128; ++i) {
        const uint32_t d = (uint32_t)j + 16u * i;
        const uint32_t w = row[d];
        const uint2 x = xg[d];
        acc = __builtin_amdgcn_sdot4((int)(w & 0x07070707u), (int)x.x, acc, false);
        acc = __builtin_amdgcn_sdot4((int)((w >> 4) & 0x07070707u), (int)x.y, acc, false);
    }
    if ((cd & 64u) && j < 8) {  // the last 64 dims (code dimension 64): 8 dwords
        const uint32_t d = 16u * (cd / 128) + (uint32_t)j;
        const uint32_t w = row[d];
        const uint2 x = xg[d];
        acc = __builtin_amdgcn_sdot4((int)(w & 0x07070707u), (int)x.x, acc, false);
        acc = __builtin_amdgcn_sdot4((int)((w >> 4) & 0x07070707u), (int)x.y, acc, false);
    }
    acc = group_sum_int(acc);
    const double centered = (double)s.int8_scale * ((double)acc - 3.5 * (double)s.int8_code_sum);
    const double signed_ip = 8.0 * (double)short_ip + centered;
    return h.x + s.query_norm_sqr - (float)((double)h.y * signed_ip);
}

// C4 on staged MSB rows, four candidates per 16-lane group in one pass (rows
// r0..r3 in LDS): lane j's accumulator for each candidate sums dims 16m + j
// in order, as the reference's 16-lane loop does; MSB word w holds dims
// 32w .. 32w + 31, of which lane j owns 32w + j (bit j) and 32w + 16 + j
// (bit 16 + j). The four chains are independent and share the query loads.
// Returns the four lanes' partial sums; the caller reduces (reduce16 order)
// and finishes each.
__device__ __forceinline__ void gate_fp32_group4(const float *q, const uint4 *r0, const uint4 *r1,
                                                 const uint4 *r2, const uint4 *r3, uint32_t cd, int j,
                                                 float (&acc)[4]) {
    const int j2 = 16 + j;
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
    const float *qj = q + j;
#pragma unroll 1
    for (uint32_t u = 0; u < cd / 128; ++u) {
        const uint4 m0 = r0[u], m1 = r1[u], m2 = r2[u], m3 = r3[u];
        const float *qu = qj + 128 * u;
        // the four chains advance in turn (masked_q's register barrier keeps
        // source order, so the interleaving has to be written out)
#define QG_GATE_STEP(W, JJ, QV)                                                    \
        a0 = a0 + masked_q(m0.W, JJ, QV); a1 = a1 + masked_q(m1.W, JJ, QV);        \
        a2 = a2 + masked_q(m2.W, JJ, QV); a3 = a3 + masked_q(m3.W, JJ, QV);
        QG_GATE_STEP(x, j, qu[0])
        QG_GATE_STEP(x, j2, qu[16])
        QG_GATE_STEP(y, j, qu[32])
        QG_GATE_STEP(y, j2, qu[48])
        QG_GATE_STEP(z, j, qu[64])
        QG_GATE_STEP(z, j2, qu[80])
        QG_GATE_STEP(w, j, qu[96])
        QG_GATE_STEP(w, j2, qu[112])
    }
    if (cd & 64u) {
        // code dimension 64 (or an odd multiple): the last 64 dims, the low
        // half of a 16-byte MSB step (the gate row's MSB span is zero-padded
        // to 16 bytes)
        const uint32_t u = cd / 128;
        const uint4 m0 = r0[u], m1 = r1[u], m2 = r2[u], m3 = r3[u];
        const float *qu = qj + 128 * u;
        QG_GATE_STEP(x, j, qu[0])
        QG_GATE_STEP(x, j2, qu[16])
        QG_GATE_STEP(y, j, qu[32])
        QG_GATE_STEP(y, j2, qu[48])
    }
#undef QG_GATE_STEP
    acc[0] = a0; acc[1] = a1; acc[2] = a2; acc[3] = a3;
}

// gate_fp32_group4 over the pair-interleaved query, qp2[16w + j] =
// (q[32w + j], q[32w + 16 + j]): one 8-byte LDS read per MSB word and lane
__device__ __forceinline__ void gate_fp32_group4_pairs(const float2 *qp2, const uint4 *r0, const uint4 *r1,
                                                       const uint4 *r2, const uint4 *r3, uint32_t cd, int j,
                                                       float (&acc)[4]) {
    const int j2 = 16 + j;
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
#pragma unroll 1
    for (uint32_t u = 0; u < cd / 128; ++u) {
        const uint4 m0 = r0[u], m1 = r1[u], m2 = r2[u], m3 = r3[u];
        const float2 q0 = qp2[64 * u + j], q1 = qp2[64 * u + 16 + j];
        const float2 q2 = qp2[64 * u + 32 + j], q3 = qp2[64 * u + 48 + j];
#define QG_GATE_STEP(W, JJ, QV)                                                    \
        a0 = a0 + masked_q(m0.W, JJ, QV); a1 = a1 + masked_q(m1.W, JJ, QV);        \
        a2 = a2 + masked_q(m2.W, JJ, QV); a3 = a3 + masked_q(m3.W, JJ, QV);
        QG_GATE_STEP(x, j, q0.x)
        QG_GATE_STEP(x, j2, q0.y)
        QG_GATE_STEP(y, j, q1.x)
        QG_GATE_STEP(y, j2, q1.y)
        QG_GATE_STEP(z, j, q2.x)
        QG_GATE_STEP(z, j2, q2.y)
        QG_GATE_STEP(w, j, q3.x)
        QG_GATE_STEP(w, j2, q3.y)
    }
    if (cd & 64u) {
        const uint32_t u = cd / 128;
        const uint4 m0 = r0[u], m1 = r1[u], m2 = r2[u], m3 = r3[u];
        const float2 q0 = qp2[64 * u + j], q1 = qp2[64 * u + 16 + j];
        QG_GATE_STEP(x, j, q0.x)
        QG_GATE_STEP(x, j2, q0.y)
        QG_GATE_STEP(y, j, q1.x)
        QG_GATE_STEP(y, j2, q1.y)
    }
#undef QG_GATE_STEP
    acc[0] = a0; acc[1] = a1; acc[2] = a2; acc[3] = a3;
}

// gate_fp32_group4 for one candidate (the same chain, alone)
__device__ __forceinline__ float gate_fp32_group1(const float *q, const uint4 *r0, uint32_t cd, int j) {
    const int j2 = 16 + j;
    float a0 = 0.0f;
    const float *qj = q + j;
    auto step = [&](uint32_t u) {
        const uint4 m0 = r0[u];
        const float *qu = qj + 128 * u;
        a0 = a0 + masked_q(m0.x, j, qu[0]);
        a0 = a0 + masked_q(m0.x, j2, qu[16]);
        a0 = a0 + masked_q(m0.y, j, qu[32]);
        a0 = a0 + masked_q(m0.y, j2, qu[48]);
        a0 = a0 + masked_q(m0.z, j, qu[64]);
        a0 = a0 + masked_q(m0.z, j2, qu[80]);
        a0 = a0 + masked_q(m0.w, j, qu[96]);
        a0 = a0 + masked_q(m0.w, j2, qu[112]);
    };
    // two MSB steps per trip (their loads issue together), an odd last one alone
    const uint32_t n = cd / 128;
#pragma unroll 1
    for (uint32_t u = 0; u + 1 < n; u += 2) {
        step(u);
        step(u + 1);
    }
    if (n & 1u) step(n - 1);
    if (cd & 64u) {
        const uint32_t u = cd / 128;
        const uint4 m0 = r0[u];
        const float *qu = qj + 128 * u;
        a0 = a0 + masked_q(m0.x, j, qu[0]);
        a0 = a0 + masked_q(m0.x, j2, qu[16]);
        a0 = a0 + masked_q(m0.y, j, qu[32]);
        a0 = a0 + masked_q(m0.y, j2, qu[48]);
    }
    return a0;
}

// C5 on a staged MSB row: 4-dim groups g4 = j + 16i (uniform trip count)
__device__ __forceinline__ PaperEstimate gate_int8_group(const int8_t *x8, const uint8_t *msb_lds, uint32_t cd,
                                                        float4 f, uint8_t fv, const QScalars &s, float eps0, int j) {
    if (!fv || s.query_norm <= 0.0f || f.z <= 0.0f || !__builtin_isfinite(f.z)) {
        PaperEstimate r;
        r.lower_bound = r.short_ip = r.alpha = r.ip_hat = r.error_bound = 0.0f;
        r.valid = 0; r.pad[0] = r.pad[1] = r.pad[2] = 0;
        return r;
    }
    const uint32_t *xg = (const uint32_t *)x8;
    const uint32_t *mw = (const uint32_t *)msb_lds;
    const uint32_t sh = 4u * ((uint32_t)j & 7u);
    const uint32_t half = (uint32_t)j >> 3;
    int sel = 0;
#pragma unroll 8
    for (uint32_t i = 0; i < cd / 64; ++i) {
        const uint32_t g4 = (uint32_t)j + 16u * i;  // dims 4 g4 .. 4 g4 + 3: word 2i + j/8, nibble j % 8
        sel = __builtin_amdgcn_sdot4((int)nibble_bytes((mw[2u * i + half] >> sh) & 0xFu), (int)xg[g4], sel, false);
    }
    sel = group_sum_int(sel);
    const float short_ip = 0.5f * s.int8_scale * (float)((int64_t)2 * (int64_t)sel - s.int8_code_sum);
    return finish_gate(short_ip, f, fv, s.query_norm, s.query_norm_sqr, eps0);
}

// C8: symmetric code<->code distance (distanceBetweenEncoded /
// symmetricBuildDistanceFromPrepared, space_rabitq.h:2634-2747, K = 1).
// ip4 = sum (2a-15)(2b-15) is exact in int32 (|term| <= 225, cd <= 2048);
// the finish runs in double with the lhs (prepared) side first.
__device__ __forceinline__ uint32_t centered_bytes(uint32_t nib4) {
    // four nibbles c (one per byte) -> four signed bytes 2c - 15, carry-free
    const uint32_t x = nib4 ^ 0x08080808u;                           // c - 8, 4-bit two's complement
    const uint32_t sx = x | (((x & 0x08080808u) >> 3) * 0xF0u);      // sign-extend each byte
    return ((sx << 1) & 0xFEFEFEFEu) | 0x01010101u;                   // 2(c - 8) + 1
}

__device__ __forceinline__ float symmetric_finish(float2 lh, float2 rh, int64_t ip4) {
    const double ln = sqrt((0.0 < (double)lh.x) ? (double)lh.x : 0.0);
    const double rn = sqrt((0.0 < (double)rh.x) ? (double)rh.x : 0.0);
    const bool lvalid = ln > 0.0 && lh.y > 0.0f && __builtin_isfinite(lh.y);
    const double lipn = lvalid ? (double)lh.y / (2.0 * ln) : 0.0;
    if (!lvalid || !(rn > 0.0) || !(rh.y > 0.0f) || !__builtin_isfinite(rh.y))
        return (float)((double)lh.x + (double)rh.x + 0.0);
    const double ripn = (double)rh.y / (2.0 * rn);
    const double code_ip = 0.25 * (double)ip4;
    double u = code_ip * lipn * ripn;
    if (!__builtin_isfinite(u)) u = 0.0;
    u = (u < 1.0) ? u : 1.0;
    u = (-1.0 < u) ? u : -1.0;
    const double rip = ln * rn * u;
    return (float)((double)lh.x + (double)rh.x - 2.0 * rip);
}

// ip4 of two raw code rows (cd/2 bytes each, 16-byte aligned)
__device__ __forceinline__ int symmetric_ip4_rows(const uint8_t *lhs, const uint8_t *rhs, uint32_t cd) {
    const uint4 *lr = (const uint4 *)lhs;
    const uint4 *rr = (const uint4 *)rhs;
    int acc = 0;
    for (uint32_t c = 0; c < cd / 32; ++c) {
        const uint4 wl = lr[c], wr = rr[c];
        const uint32_t ls[4] = {wl.x, wl.y, wl.z, wl.w}, rs[4] = {wr.x, wr.y, wr.z, wr.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            acc = __builtin_amdgcn_sdot4((int)centered_bytes(ls[k] & 0x0F0F0F0Fu),
                                         (int)centered_bytes(rs[k] & 0x0F0F0F0Fu), acc, false);
            acc = __builtin_amdgcn_sdot4((int)centered_bytes((ls[k] >> 4) & 0x0F0F0F0Fu),
                                         (int)centered_bytes((rs[k] >> 4) & 0x0F0F0F0Fu), acc, false);
        }
    }
    return acc;
}

// ip4 with the lhs already centred: lc[2w] / lc[2w+1] are the low / high
// nibbles of code dword w as signed bytes 2c - 15 (cd/4 dwords)
__device__ __forceinline__ int symmetric_ip4_prepared(const uint32_t *lc, const uint8_t *rhs, uint32_t cd) {
    const uint4 *rr = (const uint4 *)rhs;
    int acc = 0;
#pragma unroll 8
    for (uint32_t c = 0; c < cd / 32; ++c) {
        const uint4 w = rr[c];
        const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t d = 4 * c + k;
            acc = __builtin_amdgcn_sdot4((int)lc[2 * d], (int)centered_bytes(ws[k] & 0x0F0F0F0Fu), acc, false);
            acc = __builtin_amdgcn_sdot4((int)lc[2 * d + 1], (int)centered_bytes((ws[k] >> 4) & 0x0F0F0F0Fu), acc,
                                         false);
        }
    }
    return acc;
}

}  // namespace qg
