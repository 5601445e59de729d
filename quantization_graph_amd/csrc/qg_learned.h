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
    float f[kF];
    f[0] = e / S;
    f[1] = (e - e0) / S;
    f[2] = (e - e9) / S;
    f[3] = (float)i / 100.0f;
    f[4] = logf(1.0f + (float)i) / logf(101.0f);
    f[5] = (enext - e) / S;
    f[6] = (e - eprev) / S;
    f[7] = bound / S;
    f[8] = h.x / S;
    f[9] = sqrtf(h.x > 0.0f ? h.x : 0.0f) * qn / S;
    f[10] = h.y * qn / S;
    f[11] = sf.x * qn / S;
    f[12] = e0 / S;
    f[13] = e9 / S;
    f[14] = rf.x / S;
    f[15] = rf.y / S;
    f[16] = rs.x * qn * sqcd / S;
    f[17] = rs.y * qn * sqcd / S;
    f[18] = rs.z;
    f[19] = 4.0f * qn * rs.x * sqrtf(nbk) / S;
    f[20] = (float)i / (float)len;
    f[21] = (ehi - elo) / S;
    f[22] = e / (e9 > 0.0f ? e9 : 1.0f);
    f[23] = (e - e9) / (bound + 1e-12f);
#pragma unroll
    for (int k = 0; k < kFP / 2; ++k)
        w[k] = 2 * k < kF ? (uint32_t)f2bf(feat_clamp(f[2 * k])) | ((uint32_t)f2bf(feat_clamp(f[2 * k + 1])) << 16)
                          : 0u;
}

__device__ __forceinline__ void wave_lds_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// MLP forward of one 16-row tile on the calling wavefront: rows x16 (16 rows
// of kFP bf16, LDS or global, 16-byte aligned), weights P (FP32 master
// copy, rounded to bf16 as loaded), sh1 the wave's own 16 x kH bf16 LDS
// scratch. Lane l = (g = l / 16, c = l % 16) returns y of rows 4g + r in
// y[r] (the same in every lane of group g). valid_rows rows are real; the
// rest read as zero features.
__device__ __forceinline__ void mlp_forward_tile(const uint16_t *x16, uint32_t valid_rows, const float *P,
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
        const int oc = 16 * nt + c;
        uint16_t wb[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = 8 * g + j;
            wb[j] = k < kF ? f2bf(P[kOffW1 + oc * kF + k]) : (uint16_t)0;
        }
        f4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, pack8(wb), acc, 0, 0, 0);
        const float b1 = P[kOffB1 + oc];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float z = acc[r] + b1;
            sh1[(4 * g + r) * kH + oc] = f2bf(z > 0.0f ? z : 0.0f);
        }
    }
    wave_lds_order();
    bf16x8_t A2[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) A2[ks] = pack8(sh1 + c * kH + 32 * ks + 8 * g);
    float part[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        const int oc = 16 * nt + c;
        f4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            uint16_t wb[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) wb[j] = f2bf(P[kOffW2 + oc * kH + 32 * ks + 8 * g + j]);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A2[ks], pack8(wb), acc, 0, 0, 0);
        }
        const float b2 = P[kOffB2 + oc], w3 = P[kOffW3 + oc];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float z = acc[r] + b2;
            part[r] = fmaf(z > 0.0f ? z : 0.0f, w3, part[r]);
        }
    }
    wave_lds_order();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        float v = part[r];
        for (int off = 1; off < 16; off <<= 1) v += __shfl_xor(v, off);
        y[r] = v + P[kOffB3];
    }
}

// Per-query rerank depth on one wavefront (every lane returns it): rc
// candidates, pool distances pd, predicted errors pr (as fractions of
// |q|^2): the last position whose lower bound pd - k * err * |q|^2 can
// still beat the largest upper bound among the first min_depth, at least
// min(rc, min_depth).
template <typename PredAt>
__device__ __forceinline__ uint32_t learned_depth(uint32_t rc, const float *pd, PredAt pr_at, float qn2, float kmult,
                                                  uint32_t min_depth) {
    const int lane = threadIdx.x & 63;
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
