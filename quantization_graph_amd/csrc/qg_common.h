// qg_common.h -- shared device-side definitions for the ExRaBitQ4 HIP kernels.
//
// Floating-point contract: every kernel is compiled with -ffp-contract=off
// and spells out each fused multiply-add explicitly (fmaf / fma) at exactly
// the points where the reference's compiled AVX-512 code fuses them, so that
// distances come out bit-identical to the reference CPU path
// (Ours/core/hnswlib/space_rabitq.h). Sums are accumulated in the lane order
// of the reference kernels (16 lane accumulators + the 8/4/2/1 reduction
// tree of _mm512_reduce_add_ps).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qg {

// Device view of an encoded index, structure-of-arrays. Row i of every array
// belongs to base vector i. Byte contents of `codes`, `rscale` and `rcodes`
// are the reference's own encodings (sequential nibble codes, fp16 block
// scales with the reference's rounding, two's-complement residual nibbles).
struct IndexView {
    const uint8_t *codes;     // [n][cd/2]   primary 4-bit codes
    const float2 *hdr;        // [n]         {norm_sqr, long_scale}
    const float2 *sfac;       // [n]         {short_scale, error_scale}
    const uint8_t *gate;      // [n][cd/8 + 32] gate rows (below)
    const uint16_t *rscale;   // [n][cd/16]  residual block scales (fp16 bits)
    const uint8_t *rcodes;    // [n][cd/2]   residual 4-bit codes
    uint64_t n;
    const uint8_t *cid;       // [n]         centroid id (null: all 0)
};

// Gate rows: everything the 1-bit gate reads about one record, in one
// 16-byte-aligned row, so that a gated neighbour costs the lines of its MSB
// plane and nothing more. With m = cd/8 rounded up to 16 bytes:
//   [0, cd/8)         MSB plane (the paper-prune sidecar's bit plane),
//                     zero up to m
//   [m, m + 16)       sidecar factors {norm_sqr, data_norm, cross_scale,
//                     error_cross_scale}
//   [m + 16, m + 24)  a copy of the header {norm_sqr, long_scale}
//   m + 24            sidecar validity byte
//   [m + 25, m + 32)  zero
__host__ __device__ constexpr uint64_t gate_msb_span(uint64_t cd) { return (cd / 8 + 15) / 16 * 16; }
__host__ __device__ constexpr uint64_t gate_row_bytes(uint64_t cd) { return gate_msb_span(cd) + 32; }
__device__ __forceinline__ const uint8_t *gate_row(const uint8_t *gate, uint32_t cd, uint64_t id) {
    return gate + id * gate_row_bytes(cd);
}
__device__ __forceinline__ float4 gate_factors(const uint8_t *row, uint32_t cd) {
    return *(const float4 *)(row + gate_msb_span(cd));
}
__device__ __forceinline__ float2 gate_hdr(const uint8_t *row, uint32_t cd) {
    return *(const float2 *)(row + gate_msb_span(cd) + 16);
}
__device__ __forceinline__ uint8_t gate_valid(const uint8_t *row, uint32_t cd) { return row[gate_msb_span(cd) + 24]; }
// the 32-byte tail of a row as two 16-byte stores
__device__ __forceinline__ void gate_store_tail(uint8_t *row, uint32_t cd, float4 f, float2 h, uint8_t valid) {
    if (cd / 8 < gate_msb_span(cd)) *(uint2 *)(row + cd / 8) = make_uint2(0u, 0u);  // cd = 64
    uint4 *t = (uint4 *)(row + gate_msb_span(cd));
    t[0] = make_uint4(__float_as_uint(f.x), __float_as_uint(f.y), __float_as_uint(f.z), __float_as_uint(f.w));
    t[1] = make_uint4(__float_as_uint(h.x), __float_as_uint(h.y), (uint32_t)valid, 0u);
}

// Per-query scalars of the prepared query context (QueryContext,
// space_rabitq.h:101-132). The int8_* fields describe the prepared integer
// code: INT8 or INT4 scale, error norm and code sum; for the B1 codecs the
// sign code's short scale, error scale and signed sum (with b1_alpha).
struct QScalars {
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
};
static_assert(sizeof(QScalars) == 48, "QScalars layout");

struct QueryView {
    const float *q;       // [nq][cd] rotated residual
    const int8_t *x8;     // [nq][cd] INT8 query code, natural order
    const int8_t *x8p;    // [nq][cd] INT8 code regrouped per 8 dims: 4 even dims then 4 odd dims
    const QScalars *qs;   // [nq]
    uint64_t nq;
};

// RabitqPaperEstimate (rabitq_bridge.cpp:21-28).
struct PaperEstimate {
    float lower_bound;
    float short_ip;
    float alpha;
    float ip_hat;
    float error_bound;
    uint8_t valid;
    uint8_t pad[3];
};
static_assert(sizeof(PaperEstimate) == 24, "PaperEstimate layout");

// RabitqDistanceInterval (rabitq_bridge.cpp:15-19).
struct Interval {
    float estimate;
    float lower_bound;
    float upper_bound;
};

// In-LDS Walsh-Hadamard butterflies of the whole workgroup, stage step 1,
// 2, 4, ... (hadamard, space_rabitq.h:944-988; the 1-bit scan's copy,
// query_codec_1bit_scan.cpp:248-270, has the same per-element steps). Only
// add/sub: bit-exact in any schedule. Ends with a barrier.
__device__ inline void fht_lds(float *v, uint32_t cd) {
    for (uint32_t step = 1; step < cd; step <<= 1) {
        for (uint32_t b = threadIdx.x; b < cd / 2; b += blockDim.x) {
            const uint32_t blk = (b / step) * (step << 1);
            const uint32_t i = blk + (b % step);
            const float a = v[i], c = v[i + step];
            v[i] = a + c;
            v[i + step] = a - c;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ float reduce16(const float *v) {
    float t8[8], t4[4];
#pragma unroll
    for (int i = 0; i < 8; ++i) t8[i] = v[8 + i] + v[i];
#pragma unroll
    for (int i = 0; i < 4; ++i) t4[i] = t8[4 + i] + t8[i];
    const float a = t4[0] + t4[2];
    const float b = t4[1] + t4[3];
    return a + b;
}

__device__ __forceinline__ bool long_scale_ok(float ls) {
    return ls > 0.0f && __builtin_isfinite(ls);
}

// queryDistanceLongWithIps (space_rabitq.h:1621-1632) as the reference
// compiler evaluates it: 8*s + r (exact scaling), then two fused steps.
__device__ __forceinline__ float finish_long(float norm_sqr, float ls, float qn2, float half_sum,
                                             float short_ip, float rem_ip) {
    const float t = 8.0f * short_ip + rem_ip;
    const float signed_ip = fmaf(-7.0f, half_sum, t);
    const float base = norm_sqr + qn2;
    return fmaf(-ls, signed_ip, base);
}

// finish_paper_prune_estimate_sidecar (space_rabitq.h:4701-4721).
__device__ __forceinline__ PaperEstimate finish_gate(float short_ip, float4 f, uint8_t fvalid,
                                                     float qn, float qn2, float eps0) {
    PaperEstimate r;
    r.lower_bound = 0.0f; r.short_ip = 0.0f; r.alpha = 0.0f; r.ip_hat = 0.0f; r.error_bound = 0.0f;
    r.valid = 0; r.pad[0] = r.pad[1] = r.pad[2] = 0;
    if (!fvalid) return r;
    const float cross = short_ip * f.z;
    const float max_cross = 2.0f * f.y * qn;
    const float ee = eps0 * f.w;
    const float eq = ee * qn;
    // std::max(-m, std::min(m, v)) with std::min(a, b) = b < a ? b : a.
    float up = cross + eq;
    up = (up < max_cross) ? up : max_cross;
    up = (-max_cross < up) ? up : -max_cross;
    r.lower_bound = f.x + qn2 - up;
    r.short_ip = short_ip;
    r.ip_hat = max_cross > 0.0f ? cross / max_cross : 0.0f;
    r.error_bound = max_cross > 0.0f ? eq / max_cross : 0.0f;
    r.valid = __builtin_isfinite(r.lower_bound) ? 1 : 0;
    return r;
}

// finish_b1_paper_prune_estimate (space_rabitq.h:4724-4753): the bound adds
// the query side's error to the database side's; the reference compiler
// fuses db_err into the sum (fma(ecs, |q|, b1_error_scale * data_norm)).
__device__ __forceinline__ PaperEstimate finish_gate_b1(float short_ip, float4 f, float qn, float qn2, float b1e,
                                                        float b1_alpha, float eps0) {
    PaperEstimate r;
    const float cross = short_ip * f.z;
    const float max_cross = 2.0f * f.y * qn;
    const float sum = fmaf(f.w, qn, b1e * f.y);
    const float es = eps0 * sum;
    float up = cross + es;
    up = (up < max_cross) ? up : max_cross;
    up = (-max_cross < up) ? up : -max_cross;
    r.lower_bound = f.x + qn2 - up;
    r.short_ip = short_ip;
    r.alpha = b1_alpha;
    r.ip_hat = max_cross > 0.0f ? cross / max_cross : 0.0f;
    r.error_bound = max_cross > 0.0f ? es / max_cross : 0.0f;
    r.valid = __builtin_isfinite(r.lower_bound) ? 1 : 0;
    r.pad[0] = r.pad[1] = r.pad[2] = 0;
    return r;
}

// Reference fp16 rounding (floatToHalfScalar / halfToFloatScalar,
// space_rabitq.h:392-438): truncating add-0x1000, not round-to-nearest-even.
__device__ __forceinline__ uint16_t float_to_half_ref(float value) {
    const uint32_t bits = __float_as_uint(value);
    const uint32_t sign = (bits >> 16) & 0x8000u;
    uint32_t mant = bits & 0x007FFFFFu;
    const int32_t e = (int32_t)((bits >> 23) & 0xFFu) - 127 + 15;
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        mant = (mant | 0x00800000u) >> (uint32_t)(1 - e);
        return (uint16_t)(sign | ((mant + 0x00001000u) >> 13));
    }
    if (e >= 31) return (uint16_t)(sign | 0x7C00u);
    return (uint16_t)(sign | ((uint32_t)e << 10) | ((mant + 0x00001000u) >> 13));
}

__device__ __forceinline__ float half_to_float_ref(uint16_t v) {
    const uint32_t sign = (uint32_t)(v & 0x8000u) << 16;
    uint32_t e = (v >> 10) & 0x1Fu;
    uint32_t m = v & 0x03FFu;
    uint32_t bits;
    if (e == 0) {
        if (m == 0) {
            bits = sign;
        } else {
            e = 1;
            while ((m & 0x0400u) == 0) { m <<= 1; --e; }
            m &= 0x03FFu;
            bits = sign | ((e + 127 - 15) << 23) | (m << 13);
        }
    } else if (e == 31) {
        bits = sign | 0x7F800000u | (m << 13);
    } else {
        bits = sign | ((e + 127 - 15) << 23) | (m << 13);
    }
    return __uint_as_float(bits);
}

// Total order used by every ranking in the reference: (distance, id).
// Maps a float to an unsigned key whose integer order matches float order.
// -0.0 and +0.0 compare equal in the reference, so both map to one key.
__device__ __forceinline__ uint32_t float_order_key(float f) {
    const uint32_t b = (f == 0.0f) ? 0u : __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float float_from_order_key(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

}  // namespace qg
