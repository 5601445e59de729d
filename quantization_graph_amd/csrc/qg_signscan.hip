// qg_signscan.hip -- the fixed 1-bit database scanned by four query codecs
// (experiments/04_query_codec_1bit_scan/query_codec_1bit_scan.cpp), general
// in the code dimension and batched over queries: one launch scores every
// query of a batch against every database vector, reading the database once.
//
//   k_ss_encode       WG/vector   signs * x, Walsh-Hadamard in LDS, sign bits
//                                 by ballot into 32-vector blocks
//   k_ss_prepare      WG/query    rotation; B1 code and INT4 / INT8 planes by
//                                 ballot; the FP32 sums in index order
//   k_ss_scan_fp32    lane/vector 16 masked-add chains in the reference's lane
//                                 order, query values in scalar registers
//   k_ss_scan_bits<C,W> lane/vector  the vector's W words held in registers
//                                 for every query; XOR / AND popcounts
//
// Database layout (build_or_load_database, :341-384): blocks of 32 vectors,
// word-major inside a block, db[(v / 32 * W + w) * 32 + v % 32], W = cd / 64.
// A wavefront's 64 lanes own 64 consecutive vectors, so every word load of a
// wavefront reads two contiguous 256-byte runs.
#include "qg_args.h"
#include "qg_device.h"

namespace qg {

constexpr int kSsThreads = 256;
constexpr uint32_t kSsBatch = 32;

// signs * x (zero padded to cd), then the butterflies: rotate, :272-277.
__device__ void ss_rotate(float *v, const float *x, uint32_t dim, uint32_t cd, int identity, const float *signs) {
    for (uint32_t i = threadIdx.x; i < cd; i += blockDim.x)
        v[i] = (i < dim) ? (identity ? x[i] : x[i] * signs[i]) : 0.0f;
    __syncthreads();
    if (!identity) fht_lds(v, cd);
}

__global__ void __launch_bounds__(kSsThreads) k_ss_encode(const float *x, uint64_t n, uint32_t dim, uint32_t cd,
                                                          int identity, const float *signs, uint64_t *db) {
    extern __shared__ __align__(16) unsigned char smem[];
    float *v = (float *)smem;
    const uint64_t vi = blockIdx.x;
    ss_rotate(v, x + vi * dim, dim, cd, identity, signs);
    const uint32_t W = cd / 64, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t *dst = db + (vi / kSsBatch) * W * kSsBatch + vi % kSsBatch;
    for (uint32_t w = wave; w < W; w += kSsThreads / 64) {
        const uint64_t bits = __ballot(v[w * 64 + lane] > 0.0f);
        if (lane == 0) dst[(uint64_t)w * kSsBatch] = bits;
    }
}

// prepare_query + make_integer_query<4/8>, :427-492. Bit rows per query:
// 0 = B1 sign code, 1..4 = INT4 planes, 5..12 = INT8 planes.
__global__ void __launch_bounds__(kSsThreads) k_ss_prepare(const float *raw, uint64_t nq, uint32_t dim, uint32_t cd,
                                                           int identity, const float *signs, float *q_out,
                                                           uint64_t *bits_out, SsQuery *qs_out) {
    extern __shared__ __align__(16) unsigned char smem[];
    float *v = (float *)smem;
    __shared__ float s_max[kSsThreads / 64];
    __shared__ int s_sum4[kSsThreads / 64], s_sum8[kSsThreads / 64];
    __shared__ int s_nrm4[kSsThreads / 64], s_nrm8[kSsThreads / 64];
    const uint64_t qi = blockIdx.x;
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6, W = cd / 64;
    ss_rotate(v, raw + qi * dim, dim, cd, identity, signs);
    float m = 0.0f;
    for (uint32_t i = t; i < cd; i += blockDim.x) {
        const float r = v[i];
        q_out[qi * cd + i] = r;
        const float a = fabsf(r);
        m = (m < a) ? a : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const float o = __shfl_xor(m, off);
        m = (m < o) ? o : m;
    }
    if (lane == 0) s_max[wave] = m;
    __syncthreads();
    float max_abs = s_max[0];
    for (int w = 1; w < kSsThreads / 64; ++w) max_abs = (max_abs < s_max[w]) ? s_max[w] : max_abs;
    const float scale4 = max_abs > 0.0f ? max_abs / 7.0f : 1.0f;
    const float scale8 = max_abs > 0.0f ? max_abs / 127.0f : 1.0f;

    int sum4 = 0, sum8 = 0, nrm4 = 0, nrm8 = 0;
    uint64_t *rows = bits_out + qi * 13 * W;
    for (uint32_t w = wave; w < W; w += kSsThreads / 64) {
        const float r = v[w * 64 + lane];
        const int c4 = (int)fminf(fmaxf(rintf(r / scale4), -7.0f), 7.0f);
        const int c8 = (int)fminf(fmaxf(rintf(r / scale8), -127.0f), 127.0f);
        sum4 += c4;
        sum8 += c8;
        nrm4 += c4 * c4;
        nrm8 += c8 * c8;
        const uint32_t e4 = (uint32_t)c4 & 0xFu, e8 = (uint32_t)c8 & 0xFFu;
        uint64_t row[13];
        row[0] = __ballot(r > 0.0f);
#pragma unroll
        for (int p = 0; p < 4; ++p) row[1 + p] = __ballot((e4 >> p) & 1u);
#pragma unroll
        for (int p = 0; p < 8; ++p) row[5 + p] = __ballot((e8 >> p) & 1u);
        if (lane < 13) {
            uint64_t mine = row[0];
#pragma unroll
            for (int k = 1; k < 13; ++k) mine = (lane == (uint32_t)k) ? row[k] : mine;
            rows[(uint64_t)lane * W + w] = mine;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        sum4 += __shfl_xor(sum4, off);
        sum8 += __shfl_xor(sum8, off);
        nrm4 += __shfl_xor(nrm4, off);
        nrm8 += __shfl_xor(nrm8, off);
    }
    if (lane == 0) { s_sum4[wave] = sum4; s_sum8[wave] = sum8; s_nrm4[wave] = nrm4; s_nrm8[wave] = nrm8; }
    __syncthreads();
    if (t == 0) {
        // FP32 sum, squared norm and |q| sum in index order (:463-475); the
        // reference compiler keeps norm2 += v*v as a separate multiply and add
        float sum = 0.0f, norm2 = 0.0f, abs_sum = 0.0f;
        for (uint32_t i = 0; i < cd; i += 4) {
            const float4 r = *(const float4 *)(v + i);
            const float e[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                sum = sum + e[k];
                const float p = e[k] * e[k];
                norm2 = norm2 + p;
                abs_sum = abs_sum + fabsf(e[k]);
            }
        }
        int64_t s4 = 0, s8 = 0, n4 = 0, n8 = 0;
        for (int w = 0; w < kSsThreads / 64; ++w) { s4 += s_sum4[w]; s8 += s_sum8[w]; n4 += s_nrm4[w]; n8 += s_nrm8[w]; }
        SsQuery s;
        s.fp32_sum = sum;
        s.fp32_norm2 = norm2;
        s.b1_scale = abs_sum / (float)cd;
        s.b1_constant = (float)cd * fmaf(s.b1_scale, s.b1_scale, 1.0f);
        s.int4_scale = scale4;
        s.int4_constant = fmaf(scale4 * scale4, (float)n4, (float)cd);
        s.int8_scale = scale8;
        s.int8_constant = fmaf(scale8 * scale8, (float)n8, (float)cd);
        s.int4_sum = s4;
        s.int8_sum = s8;
        qs_out[qi] = s;
    }
}

// scan_fp32, :529-563: lane j of the 16 accumulators takes dimensions j,
// j+16, j+32, ... in order (each word's four 16-bit quarters), then the
// _mm512_reduce_add_ps tree. Masked-out dimensions add +0.0, which leaves a
// sum unchanged, so they are added as q AND mask. The query's 64 values of
// a word are wave-uniform and live in scalar registers.
__global__ void __launch_bounds__(kSsThreads) k_ss_scan_fp32(const uint64_t *__restrict__ db, uint64_t n,
                                                             uint32_t W, const float *__restrict__ q,
                                                             const SsQuery *__restrict__ qs, uint64_t nq,
                                                             float *__restrict__ out) {
    const uint64_t v = (uint64_t)blockIdx.x * kSsThreads + threadIdx.x;
    const bool valid = v < n;
    const uint64_t *row = db + (v / kSsBatch) * W * kSsBatch + v % kSsBatch;
    const uint32_t cd = W * 64;
    for (uint64_t qi = 0; qi < nq; ++qi) {
        const float *qr = q + qi * cd;
        float acc[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) acc[j] = 0.0f;
        for (uint32_t w = 0; w < W; ++w) {
            const uint64_t word = valid ? row[(uint64_t)w * kSsBatch] : 0;
            const uint32_t lo = (uint32_t)word, hi = (uint32_t)(word >> 32);
            const float *qw = qr + w * 64;
#pragma unroll
            for (int b = 0; b < 32; ++b) acc[b & 15] += masked_q(lo, b, qw[b]);
#pragma unroll
            for (int b = 0; b < 32; ++b) acc[b & 15] += masked_q(hi, b, qw[32 + b]);
        }
        const SsQuery s = qs[qi];
        const float dot = 2.0f * reduce16(acc) - s.fp32_sum;
        if (valid) out[qi * n + v] = ((float)cd + s.fp32_norm2) - 2.0f * dot;
    }
}

// scan_b1 (:565-598) and scan_integer<Bits> (:600-648). The vector's W words
// are loaded once into registers and reused for every query; the query's bit
// rows are wave-uniform scalar loads. CODEC 1: B1 mismatches by XOR popcount;
// CODEC 2 / 3: INT4 / INT8 selected sums as plane popcounts weighted 1, 2,
// ..., -2^(Bits-1) (two's complement), all exact integers.
template <int CODEC, int W>
__global__ void __launch_bounds__(kSsThreads) k_ss_scan_bits(const uint64_t *__restrict__ db, uint64_t n,
                                                             const uint64_t *__restrict__ bits,
                                                             const SsQuery *__restrict__ qs, uint64_t nq,
                                                             float *__restrict__ out) {
    constexpr int kBits = CODEC == 2 ? 4 : 8;
    constexpr int kRow0 = CODEC == 1 ? 0 : (CODEC == 2 ? 1 : 5);
    constexpr uint32_t kCd = W * 64;
    const uint64_t v = (uint64_t)blockIdx.x * kSsThreads + threadIdx.x;
    const bool valid = v < n;
    const uint64_t *row = db + (v / kSsBatch) * W * kSsBatch + v % kSsBatch;
    uint64_t word[W];
#pragma unroll
    for (int w = 0; w < W; ++w) word[w] = valid ? row[(uint64_t)w * kSsBatch] : 0;
    for (uint64_t qi = 0; qi < nq; ++qi) {
        const uint64_t *qb = bits + (qi * 13 + kRow0) * W;
        const SsQuery s = qs[qi];
        float d;
        if (CODEC == 1) {
            uint32_t mism = 0;
#pragma unroll
            for (int w = 0; w < W; ++w) mism += (uint32_t)__popcll(word[w] ^ qb[w]);
            const float signed_dot = (float)kCd - 2.0f * (float)mism;
            d = fmaf(-(2.0f * s.b1_scale), signed_dot, s.b1_constant);
        } else {
            int32_t selected = 0;
#pragma unroll
            for (int p = 0; p < kBits; ++p) {
                uint32_t cnt = 0;
#pragma unroll
                for (int w = 0; w < W; ++w) cnt += (uint32_t)__popcll(word[w] & qb[p * W + w]);
                const int32_t weight = p + 1 == kBits ? -(1 << (kBits - 1)) : (1 << p);
                selected += (int32_t)cnt * weight;
            }
            const int64_t dot = 2 * (int64_t)selected - (CODEC == 2 ? s.int4_sum : s.int8_sum);
            const float scale = CODEC == 2 ? s.int4_scale : s.int8_scale;
            const float constant = CODEC == 2 ? s.int4_constant : s.int8_constant;
            d = fmaf(-(2.0f * scale), (float)dot, constant);
        }
        if (valid) out[qi * n + v] = d;
    }
}

template <int CODEC>
int launch_bits(uint32_t W, dim3 grid, hipStream_t st, const uint64_t *db, uint64_t n, const uint64_t *bits,
                const SsQuery *qs, uint64_t nq, float *out) {
    switch (W) {
#define QG_SS_CASE(w) \
    case w: hipLaunchKernelGGL((k_ss_scan_bits<CODEC, w>), grid, dim3(kSsThreads), 0, st, db, n, bits, qs, nq, out); break;
        QG_SS_CASE(1) QG_SS_CASE(2) QG_SS_CASE(4) QG_SS_CASE(8) QG_SS_CASE(16) QG_SS_CASE(32)
#undef QG_SS_CASE
        default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

}  // namespace qg

extern "C" {

int qg_launch_ss_encode(const float *x, uint64_t n, uint32_t dim, uint32_t cd, int identity, const float *signs,
                        uint64_t *db, hipStream_t st) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(qg::k_ss_encode, dim3((unsigned)n), dim3(qg::kSsThreads), (size_t)cd * 4, st, x, n, dim, cd,
                       identity, signs, db);
    return (int)hipGetLastError();
}

int qg_launch_ss_prepare(const float *raw, uint64_t nq, uint32_t dim, uint32_t cd, int identity, const float *signs,
                         float *q_out, uint64_t *bits_out, qg::SsQuery *qs_out, hipStream_t st) {
    if (nq == 0) return 0;
    hipLaunchKernelGGL(qg::k_ss_prepare, dim3((unsigned)nq), dim3(qg::kSsThreads), (size_t)cd * 4, st, raw, nq, dim,
                       cd, identity, signs, q_out, bits_out, qs_out);
    return (int)hipGetLastError();
}

int qg_launch_ss_scan(int codec, const uint64_t *db, uint64_t n, uint32_t W, const float *q, const uint64_t *bits,
                      const qg::SsQuery *qs, uint64_t nq, float *out, hipStream_t st) {
    if (n == 0 || nq == 0) return 0;
    const dim3 grid((unsigned)((n + qg::kSsThreads - 1) / qg::kSsThreads));
    switch (codec) {
        case 0:
            hipLaunchKernelGGL(qg::k_ss_scan_fp32, grid, dim3(qg::kSsThreads), 0, st, db, n, W, q, qs, nq, out);
            return (int)hipGetLastError();
        case 1: return qg::launch_bits<1>(W, grid, st, db, n, bits, qs, nq, out);
        case 2: return qg::launch_bits<2>(W, grid, st, db, n, bits, qs, nq, out);
        case 3: return qg::launch_bits<3>(W, grid, st, db, n, bits, qs, nq, out);
        default: return (int)hipErrorInvalidValue;
    }
}

}  // extern "C"
