// qg_learned.h -- the learned rerank depth's per-candidate pieces, shared by
// the learned-rerank kernels (qg_mlp.hip) and the search kernel's fused
// learned rerank (qg_search.hip) so both give the same bits: the 24 feature
// values of a pool entry, the MLP forward of a 16-row tile on one wavefront
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation), and the per-query depth
// rule. No counterpart in the reference.
#pragma once

#include "qg_common.h"

namespace qg {

constexpr int kF = 24, kFP = 32, kH = 64;
constexpr int kOffW1 = 0, kOffB1 = kH * kF, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH, kOffW3 = kOffB2 + kH,
              kOffB3 = kOffW3 + kH, kParams = kOffB3 + 1;  // 5825

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef short short8_t __attribute__((ext_vector_type(8)));
typedef float f4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even, finite inputs
    uint32_t b = __float_as_uint(f);
    b += 0x7FFFu + ((b >> 16) & 1u);
    return (uint16_t)(b >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x8_t pack8(const uint16_t *v) {
    short8_t s;
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] = (short)v[j];
    return __builtin_bit_cast(bf16x8_t, s);
}

__device__ __forceinline__ float feat_clamp(float v) {
    if (!__builtin_isfinite(v)) return 0.0f;
    return v < -1e4f ? -1e4f : (v > 1e4f ? 1e4f : v);
}

// The feature row of pool entry i (of len, pool distances pd) with id's
// header h, short factors sf, residual factors rf and residual block stats
// rs, as 16 words of bf16 pairs (24 features, then zeros).
__device__ __forceinline__ void learned_feature_row(uint32_t i, uint32_t len, const float *pd, const QScalars &s,
                                                    float2 h, float2 sf, float2 rf, float4 rs, uint32_t cd,
                                                    uint32_t *w) {
    const float S = s.query_norm_sqr > 0.0f ? s.query_norm_sqr : 1.0f;
    const float qn = s.query_norm;
    const float e = pd[i];
    const float e0 = pd[0], e9 = pd[len > 9 ? 9 : len - 1];
    const float enext = i + 1 < len ? pd[i + 1] : e, eprev = i > 0 ? pd[i - 1] : e;
    const float elo = pd[i >= 5 ? i - 5 : 0], ehi = pd[i + 5 < len ? i + 5 : len - 1];
    const float sqcd = sqrtf((float)cd);
    const float nbk = (float)(cd / 16);
    const float bound = sf.y * qn;
    const float f[kF] = {
        e / S, (e - e0) / S, (e - e9) / S, (float)i / 100.0f, logf(1.0f + (float)i) / logf(101.0f),
        (enext - e) / S, (e - eprev) / S, bound / S, h.x / S, sqrtf(h.x > 0.0f ? h.x : 0.0f) * qn / S,
        h.y * qn / S, sf.x * qn / S, e0 / S, e9 / S, rf.x / S, rf.y / S, rs.x * qn * sqcd / S,
        rs.y * qn * sqcd / S, rs.z, 4.0f * qn * rs.x * sqrtf(nbk) / S, (float)i / (float)len, (ehi - elo) / S,
        e / (e9 > 0.0f ? e9 : 1.0f), (e - e9) / (bound + 1e-12f)};
#pragma unroll
    for (int k = 0; k < kFP / 2; ++k)
        w[k] = 2 * k < kF ? (uint32_t)f2bf(feat_clamp(f[2 * k])) | ((uint32_t)f2bf(feat_clamp(f[2 * k + 1])) << 16)
                          : 0u;
}


// The MLP weights a lane needs for its part of every 16-row tile, rounded
// to bf16: lane l = (g = l / 16, c = l % 16) uses W1 rows 16 nt + c, inputs
// 8g .. 8g + 7 (B1[nt]) and W2 rows 16 nt + c, inputs 32 ks + 8g .. + 7
// (B2[nt][ks]); biases b1, b2 and W3 of rows 16 nt + c. The packed form
// holds them per lane as 12 uint4 (B1[0..3], B2[0..3][0..1]) and 12 floats,
// so a wavefront fetches its weights with a few wide loads.
struct MlpLane {
    bf16x8_t B1[4], B2[4][2];
    float b1[4], b2[4], w3[4], b3;
};

constexpr uint32_t kMlpPackedU4 = 12, kMlpPackedF = 13;  // per lane
constexpr size_t kMlpPackedBytes = 64 * (kMlpPackedU4 * 16 + 16 * 4);  // 16 floats per lane reserved

__device__ __forceinline__ void mlp_lane_from_params(const float *P, int lane, MlpLane &m) {
    const int g = lane >> 4, c = lane & 15;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const int oc = 16 * nt + c;
        uint16_t wb[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = 8 * g + j;
            wb[j] = k < kF ? f2bf(P[kOffW1 + oc * kF + k]) : (uint16_t)0;
        }
        m.B1[nt] = pack8(wb);
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
#pragma unroll
            for (int j = 0; j < 8; ++j) wb[j] = f2bf(P[kOffW2 + oc * kH + 32 * ks + 8 * g + j]);
            m.B2[nt][ks] = pack8(wb);
        }
        m.b1[nt] = P[kOffB1 + oc];
        m.b2[nt] = P[kOffB2 + oc];
        m.w3[nt] = P[kOffW3 + oc];
    }
    m.b3 = P[kOffB3];
}

// packed layout: uint4 [12][64] (lane-minor), then float [16][64]
__device__ __forceinline__ void mlp_lane_store_packed(const MlpLane &m, int lane, void *packed) {
    uint4 *u = (uint4 *)packed;
    float *f = (float *)(u + kMlpPackedU4 * 64);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        u[nt * 64 + lane] = __builtin_bit_cast(uint4, m.B1[nt]);
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) u[(4 + 2 * nt + ks) * 64 + lane] = __builtin_bit_cast(uint4, m.B2[nt][ks]);
        f[nt * 64 + lane] = m.b1[nt];
        f[(4 + nt) * 64 + lane] = m.b2[nt];
        f[(8 + nt) * 64 + lane] = m.w3[nt];
    }
    f[12 * 64 + lane] = m.b3;
}

__device__ __forceinline__ void wave_lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// MLP forward of one 16-row tile on the calling wavefront: rows x16 (16 rows
// of kFP bf16, LDS or global, 16-byte aligned; valid_rows of them real, the
// rest read as zero features), weights m, sh1 the wave's own 16 x kH bf16
// LDS scratch. Lane l = (g = l / 16, c = l % 16) returns y of rows 4g + r in
// y[r] (the same in every lane of group g).
__device__ __forceinline__ void mlp_forward_tile(const uint16_t *x16, uint32_t valid_rows, const MlpLane &m,
                                                 uint16_t *sh1, float *y) {
    const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
    uint16_t xa[8];
    {
        const uint4 v = (uint32_t)c < valid_rows ? *(const uint4 *)(x16 + c * kFP + 8 * g) : make_uint4(0, 0, 0, 0);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) { xa[2 * j] = (uint16_t)(w[j] & 0xFFFF); xa[2 * j + 1] = (uint16_t)(w[j] >> 16); }
    }
    const bf16x8_t A1 = pack8(xa);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, m.B1[nt], acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float z = acc[r] + m.b1[nt];
            sh1[(4 * g + r) * kH + 16 * nt + c] = f2bf(z > 0.0f ? z : 0.0f);
        }
    }
    // sh1 belongs to this wave alone: a wave-level barrier orders the writes
    // above before the reads below (and those reads before the next tile's writes)
    wave_lds_order();
    bf16x8_t A2[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) A2[ks] = pack8(sh1 + c * kH + 32 * ks + 8 * g);
    float part[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A2[ks], m.B2[nt][ks], acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float z = acc[r] + m.b2[nt];
            part[r] = fmaf(z > 0.0f ? z : 0.0f, m.w3[nt], part[r]);
        }
    }
    wave_lds_order();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        float v = part[r];
        for (int off = 1; off < 16; off <<= 1) v += __shfl_xor(v, off);
        y[r] = v + m.b3;
    }
}

// mlp_forward_tile with the weights read from the packed per-lane layout
// as each layer needs them (wide loads, few registers held): the same
// operands in the same order, so the same y.
__device__ __forceinline__ void mlp_forward_tile_packed(const uint16_t *x16, uint32_t valid_rows, const void *packed,
                                                        uint16_t *sh1, float *y, int lane) {
    const int g = lane >> 4, c = lane & 15;
    const uint4 *u = (const uint4 *)packed;
    const float *f = (const float *)(u + kMlpPackedU4 * 64);
    uint16_t xa[8];
    {
        const uint4 v = (uint32_t)c < valid_rows ? *(const uint4 *)(x16 + c * kFP + 8 * g) : make_uint4(0, 0, 0, 0);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) { xa[2 * j] = (uint16_t)(w[j] & 0xFFFF); xa[2 * j + 1] = (uint16_t)(w[j] >> 16); }
    }
    const bf16x8_t A1 = pack8(xa);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, __builtin_bit_cast(bf16x8_t, u[nt * 64 + lane]), acc, 0, 0,
                                                      0);
        const float b1 = f[nt * 64 + lane];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float z = acc[r] + b1;
            sh1[(4 * g + r) * kH + 16 * nt + c] = f2bf(z > 0.0f ? z : 0.0f);
        }
    }
    wave_lds_order();
    bf16x8_t A2[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) A2[ks] = pack8(sh1 + c * kH + 32 * ks + 8 * g);
    float part[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                A2[ks], __builtin_bit_cast(bf16x8_t, u[(4 + 2 * nt + ks) * 64 + lane]), acc, 0, 0, 0);
        const float b2 = f[(4 + nt) * 64 + lane], w3 = f[(8 + nt) * 64 + lane];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float z = acc[r] + b2;
            part[r] = fmaf(z > 0.0f ? z : 0.0f, w3, part[r]);
        }
    }
    wave_lds_order();
    const float b3 = f[12 * 64 + lane];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        float v = part[r];
        for (int off = 1; off < 16; off <<= 1) v += __shfl_xor(v, off);
        y[r] = v + b3;
    }
}

// Per-query rerank depth on one wavefront (every lane returns it): rc
// candidates, pool distances pd, predicted errors pr (as fractions of
// |q|^2): the last position whose lower bound pd - k * err * |q|^2 can
// still beat the largest upper bound among the first min_depth, at least
// min(rc, min_depth).
template <typename PredAt>
__device__ __forceinline__ uint32_t learned_depth(uint32_t rc, const float *pd, PredAt pr_at, float qn2, float kmult,
                                                  uint32_t min_depth, uint32_t lane) {
    if (rc == 0) return 0;
    const float S = qn2 > 0.0f ? qn2 : 1.0f;
    const uint32_t top = rc < min_depth ? rc : min_depth;
    float ub = -__builtin_huge_valf();
    for (uint32_t j = lane; j < top; j += 64) {
        const float pr = pr_at(j);
        ub = fmaxf(ub, pd[j] + kmult * (pr > 0.0f ? pr : 0.0f) * S);
    }
    for (int off = 32; off > 0; off >>= 1) ub = fmaxf(ub, __shfl_xor(ub, off));
    uint32_t last = 0;
    for (uint32_t j = lane; j < rc; j += 64) {
        const float pr = pr_at(j);
        const float lb = pd[j] - kmult * (pr > 0.0f ? pr : 0.0f) * S;
        if (lb <= ub) last = j + 1 > last ? j + 1 : last;
    }
    for (int off = 32; off > 0; off >>= 1) last = max(last, (uint32_t)__shfl_xor((int)last, off));
    return last < top ? top : last;
}

}  // namespace qg
