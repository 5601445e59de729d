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
// a word are wave-uniform and live in scalar registers. QB queries share one
// pass: each dimension's mask is extracted once for all of them (QB = 4:
// 2.25 VALU operations per query and dimension instead of 3).
template <int QB>
__device__ __forceinline__ void ss_fp32_pass(const uint64_t *row, bool valid, uint64_t v, uint64_t n, uint32_t W,
                                             const float *__restrict__ q, const SsQuery *__restrict__ qs,
                                             uint64_t qi0, float *__restrict__ out) {
    const uint32_t cd = W * 64;
    float acc[QB][16];
#pragma unroll
    for (int qq = 0; qq < QB; ++qq)
#pragma unroll
        for (int j = 0; j < 16; ++j) acc[qq][j] = 0.0f;
    for (uint32_t w = 0; w < W; ++w) {
        const uint64_t word = valid ? row[(uint64_t)w * kSsBatch] : 0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t bits = (uint32_t)(word >> (32 * h));
#pragma unroll
            for (int b = 0; b < 32; ++b) {
                int m = __builtin_amdgcn_sbfe((int)bits, b, 1);
                asm("" : "+v"(m));  // keep the mask an integer: no fold back into compare + select
#pragma unroll
                for (int qq = 0; qq < QB; ++qq) {
                    const float qv = q[(qi0 + qq) * cd + w * 64 + 32 * h + b];
                    acc[qq][b & 15] += __uint_as_float(__float_as_uint(qv) & (uint32_t)m);
                }
            }
        }
    }
#pragma unroll
    for (int qq = 0; qq < QB; ++qq) {
        const SsQuery s = qs[qi0 + qq];
        const float dot = 2.0f * reduce16(acc[qq]) - s.fp32_sum;
        if (valid) out[(qi0 + qq) * n + v] = ((float)cd + s.fp32_norm2) - 2.0f * dot;
    }
}

__global__ void __launch_bounds__(kSsThreads) k_ss_scan_fp32(const uint64_t *__restrict__ db, uint64_t n,
                                                             uint32_t W, const float *__restrict__ q,
                                                             const SsQuery *__restrict__ qs, uint64_t nq,
                                                             float *__restrict__ out) {
    const uint64_t v = (uint64_t)blockIdx.x * kSsThreads + threadIdx.x;
    const bool valid = v < n;
    const uint64_t *row = db + (v / kSsBatch) * W * kSsBatch + v % kSsBatch;
    uint64_t qi = 0;
    for (; qi + 4 <= nq; qi += 4) ss_fp32_pass<4>(row, valid, v, n, W, q, qs, qi, out);
    for (; qi < nq; ++qi) ss_fp32_pass<1>(row, valid, v, n, W, q, qs, qi, out);
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

// ------------------------------------------- batched integer codecs on MFMA
// For a batch of queries the integer codecs are one int8 GEMM: with the
// database row as 0/1 bytes s_i and the query as int8 codes x_i,
// G = sum_i s_i x_i is exact in int32, and
//   INT4 / INT8: integer_dot = 2 G - code_sum        (scan_integer<Bits>)
//   B1 (x = +-1): mismatches = popcount(q) - G, signed_dot = cd - 2 mism
// are the reference's integers, so every distance is bit-identical to the
// popcount kernels (v_mfma_i32_16x16x64_i8, gfx950's 64-deep int8 step). A workgroup stages QT x 16 queries' codes in LDS (the
// 8 x 8 bit transpose of their planes) and its waves walk 16 VT-vector
// tiles: each lane holds its vector's words in registers and expands a
// step's 16 bits to 0/1 bytes once for all query tiles (A operand); the
// MFMAs read the query codes from LDS (B operand).
// Lane (g = lane / 16, r = lane % 16) supplies 16 consecutive dimensions
// 64 k + 16 g .. +15 of step k on both sides (vector r as the A row, query
// r as the B column); accumulator j holds vector 4 g + j against query r, so
// a lane stores 4 consecutive distances of one query row.
typedef int v4i_ss __attribute__((ext_vector_type(4)));
typedef float v4f_ss __attribute__((ext_vector_type(4)));

// byte j of the result holds bit j of every input byte i at bit i
__device__ __forceinline__ uint64_t transpose8x8(uint64_t x) {
    uint64_t t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
    x = x ^ t ^ (t << 7);
    t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
    x = x ^ t ^ (t << 14);
    t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
    return x ^ t ^ (t << 28);
}

// two waves per SIMD (<= 256 VGPRs): the vectors' words stay in registers
// (2 VT W VGPRs) and each step's bytes are expanded just before its MFMAs
template <int CODEC, int W, int VT, int QT>
__global__ void __launch_bounds__(kSsThreads) __attribute__((amdgpu_waves_per_eu(2))) k_ss_scan_mfma(const uint64_t *__restrict__ db, uint64_t n,
                                                             const uint64_t *__restrict__ bits,
                                                             const SsQuery *__restrict__ qs, uint64_t nq,
                                                             float *__restrict__ out) {
    constexpr int KS = W, CD = 64 * W, QC = 16 * QT;  // one 64-dimension step per database word
    constexpr int RS = CD + 16;  // LDS row stride: 8 lanes reading 16 bytes at one offset of 8 rows hit 32 banks
    constexpr int kBits = CODEC == 1 ? 1 : (CODEC == 2 ? 4 : 8);
    constexpr int kRow0 = CODEC == 1 ? 0 : (CODEC == 2 ? 1 : 5);
    extern __shared__ __align__(16) unsigned char smem[];
    int8_t *sq = (int8_t *)smem;  // [QC][RS] query codes
    __shared__ float s_scale2[QC], s_const[QC];
    __shared__ int s_off[QC];
    const uint32_t t = threadIdx.x, lane = t & 63u, g = lane >> 4, rl = lane & 15u;
    const uint64_t tiles = (n + 16 * VT - 1) / (16 * VT);
    const uint64_t wave0 = (uint64_t)blockIdx.x * (kSsThreads / 64) + (t >> 6);
    const uint64_t waves = (uint64_t)gridDim.x * (kSsThreads / 64);
    for (uint64_t q0 = 0; q0 < nq; q0 += QC) {
        __syncthreads();
        for (uint32_t item = t; item < (uint32_t)QC * (CD / 8); item += kSsThreads) {
            const uint32_t ql = item / (CD / 8), b = item % (CD / 8);
            const uint64_t q = q0 + ql;
            uint64_t code = 0;
            if (q < nq) {
                const uint64_t *rows = bits + (q * 13 + kRow0) * W + b / 8;
                const uint32_t sh = 8 * (b % 8);
                uint64_t x = 0;
#pragma unroll
                for (int p = 0; p < kBits; ++p) x |= ((rows[p * W] >> sh) & 0xFFull) << (8 * p);
                const uint64_t y = transpose8x8(x);
                if (CODEC == 1) code = ~(y * 0xFEull);                               // 1 -> +1, 0 -> -1
                else if (CODEC == 2) code = y | (((y >> 3) & 0x0101010101010101ull) * 0xF0ull);  // sign-extend
                else code = y;
            }
            *(uint64_t *)(sq + (uint64_t)ql * RS + 8 * b) = code;
        }
        if (t < (uint32_t)QC) {
            const uint64_t q = q0 + t;
            if (q < nq) {
                const SsQuery s = qs[q];
                if (CODEC == 1) {
                    int pc = 0;
                    for (int w = 0; w < W; ++w) pc += __popcll(bits[q * 13 * W + w]);
                    s_scale2[t] = -(2.0f * s.b1_scale);
                    s_const[t] = s.b1_constant;
                    s_off[t] = pc;
                } else {
                    s_scale2[t] = -(2.0f * (CODEC == 2 ? s.int4_scale : s.int8_scale));
                    s_const[t] = CODEC == 2 ? s.int4_constant : s.int8_constant;
                    s_off[t] = -(int)(CODEC == 2 ? s.int4_sum : s.int8_sum);
                }
            }
        }
        __syncthreads();
        // this lane's queries (column rl of each query tile) and their scalars
        float sc2[QT], cst[QT];
        int off[QT];
        float *orow[QT];
#pragma unroll
        for (int qt = 0; qt < QT; ++qt) {
            const uint32_t ql = 16 * qt + rl;
            sc2[qt] = s_scale2[ql];
            cst[qt] = s_const[ql];
            off[qt] = s_off[ql];
            orow[qt] = q0 + ql < nq ? out + (q0 + ql) * n : nullptr;
        }
        // a tile's words are loaded while the previous tile computes (two
        // register sets in turn); loads are unconditional (past the end they
        // re-read the last tile, whose results are not stored), so the
        // compiler waits on exactly the loads a step uses
        auto load_words = [&](uint64_t tile, uint64_t (&dst)[VT][W]) {
            const uint64_t tc = tile < tiles ? tile : tiles - 1;
#pragma unroll
            for (int vt = 0; vt < VT; ++vt) {
                uint64_t v = tc * 16 * VT + 16 * vt + rl;
                v = v < n ? v : n - 1;
                const uint64_t *row = db + (v / kSsBatch) * W * kSsBatch + v % kSsBatch;
#pragma unroll
                for (int w = 0; w < W; ++w) dst[vt][w] = row[(uint64_t)w * kSsBatch];
            }
        };
        auto compute_tile = [&](uint64_t tile, const uint64_t (&word)[VT][W]) {
            const uint64_t v0 = tile * 16 * VT;
            v4i_ss acc[VT][QT];
#pragma unroll
            for (int vt = 0; vt < VT; ++vt)
#pragma unroll
                for (int qt = 0; qt < QT; ++qt) acc[vt][qt] = (v4i_ss){0, 0, 0, 0};
            // the query codes of step k + kAhead are read while step k runs
            // (a ring of kAhead + 1 register sets); the scheduling barrier
            // keeps the compiler from hoisting every step's reads, and their
            // registers, to the top
            constexpr int kAhead = QT >= 4 ? 1 : 2;
            const int8_t *arow = sq + rl * RS + 16 * g;
            const uint32_t sh16 = 16 * (g & 1);  // this lane's 16 bits inside its half of the word
            v4i_ss ring[kAhead + 1][QT];
#pragma unroll
            for (int k = 0; k < kAhead && k < KS; ++k)
#pragma unroll
                for (int qt = 0; qt < QT; ++qt) ring[k][qt] = *(const v4i_ss *)(arow + 16 * qt * RS + 64 * k);
#pragma unroll
            for (int k = 0; k < KS; ++k) {
                if (k + kAhead < KS) {
#pragma unroll
                    for (int qt = 0; qt < QT; ++qt)
                        ring[(k + kAhead) % (kAhead + 1)][qt] =
                            *(const v4i_ss *)(arow + 16 * qt * RS + 64 * (k + kAhead));
                }
                v4i_ss bx[VT];
#pragma unroll
                for (int vt = 0; vt < VT; ++vt) {
                    const uint32_t half = g >= 2 ? (uint32_t)(word[vt][k] >> 32) : (uint32_t)word[vt][k];
#pragma unroll
                    for (int i = 0; i < 4; ++i) bx[vt][i] = (int)nibble_bytes(__builtin_amdgcn_ubfe(half, sh16 + 4 * i, 4));
                }
#pragma unroll
                for (int qt = 0; qt < QT; ++qt)
#pragma unroll
                    for (int vt = 0; vt < VT; ++vt)
                        acc[vt][qt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(bx[vt], ring[k % (kAhead + 1)][qt],
                                                                            acc[vt][qt], 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
            }
            // acc[vt][qt][j]: vector v0 + 16 vt + 4 g + j against query 16 qt + rl;
            // n is a multiple of 32, so a 16-vector tile is all in or all out
#pragma unroll
            for (int vt = 0; vt < VT; ++vt) {
                const uint64_t vb = v0 + 16 * vt;
                if (vb >= n) continue;
#pragma unroll
                for (int qt = 0; qt < QT; ++qt) {
                    if (!orow[qt]) continue;
                    v4f_ss d;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int gsum = acc[vt][qt][j];
                        const int dot = CODEC == 1 ? CD - 2 * (off[qt] - gsum) : 2 * gsum + off[qt];
                        d[j] = fmaf(sc2[qt], (float)dot, cst[qt]);
                    }
                    *(v4f_ss *)(orow[qt] + vb + 4 * g) = d;
                }
            }
        };
        uint64_t wa[VT][W], wb[VT][W];
        load_words(wave0, wa);
        for (uint64_t tile = wave0; tile < tiles; tile += 2 * waves) {
            load_words(tile + waves, wb);
            compute_tile(tile, wa);
            if (tile + waves >= tiles) break;
            load_words(tile + 2 * waves, wa);
            compute_tile(tile + waves, wb);
        }
    }
}

template <int CODEC, int W, int VT, int QT>
int launch_mfma_wq(uint64_t n, hipStream_t st, const uint64_t *db, const uint64_t *bits, const SsQuery *qs,
                   uint64_t nq, float *out) {
    const uint64_t tiles = (n + 16 * VT - 1) / (16 * VT);
    uint64_t blocks = (tiles + 3) / 4;
    if (blocks > 512) blocks = 512;  // two resident workgroups per CU; staging amortised over the tiles
    const size_t smem = (size_t)16 * QT * (64 * W + 16);
    static bool raised = false;  // dynamic LDS above the 64 KB default needs the attribute once
    if (!raised && smem > 65536) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_ss_scan_mfma<CODEC, W, VT, QT>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
        if (e != hipSuccess) return (int)e;
        raised = true;
    }
    hipLaunchKernelGGL((k_ss_scan_mfma<CODEC, W, VT, QT>), dim3((unsigned)blocks), dim3(kSsThreads), smem, st, db, n,
                       bits, qs, nq, out);
    return (int)hipGetLastError();
}

// query tiles per staged chunk: as many as the batch needs, up to 66 KB of
// codes (4 tiles up to code dimension 1024, 2 at 2048): two workgroups per CU
template <int CODEC, int W, int VT, int QTMAX>
int launch_mfma_w(uint64_t n, hipStream_t st, const uint64_t *db, const uint64_t *bits, const SsQuery *qs,
                  uint64_t nq, float *out) {
    if (nq <= 16) return launch_mfma_wq<CODEC, W, VT, 1>(n, st, db, bits, qs, nq, out);
    if (nq <= 32 || QTMAX == 2) return launch_mfma_wq<CODEC, W, VT, 2>(n, st, db, bits, qs, nq, out);
    return launch_mfma_wq<CODEC, W, VT, (QTMAX >= 4 ? 4 : 2)>(n, st, db, bits, qs, nq, out);
}

template <int CODEC>
int launch_mfma(uint32_t W, hipStream_t st, const uint64_t *db, uint64_t n, const uint64_t *bits, const SsQuery *qs,
                uint64_t nq, float *out) {
    switch (W) {
        case 2: return launch_mfma_w<CODEC, 2, 2, 4>(n, st, db, bits, qs, nq, out);
        case 4: return launch_mfma_w<CODEC, 4, 2, 4>(n, st, db, bits, qs, nq, out);
        case 8: return launch_mfma_w<CODEC, 8, 2, 4>(n, st, db, bits, qs, nq, out);
        case 16: return launch_mfma_w<CODEC, 16, 1, 4>(n, st, db, bits, qs, nq, out);
        case 32: return launch_mfma_w<CODEC, 32, 1, 2>(n, st, db, bits, qs, nq, out);
        default: return (int)hipErrorInvalidValue;
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
        case 1:
        case 2:
        case 3:
            // batches on the matrix cores (code dimension >= 128, 16-byte
            // aligned output rows): from 8 queries for INT4 / INT8; B1's XOR
            // popcount costs 70 operations per pair and stays ahead up to
            // about 48 queries on MI355X
            if (nq >= (codec == 1 ? 48u : 8u) && W >= 2 && ((uintptr_t)out & 15u) == 0) {
                return codec == 1 ? qg::launch_mfma<1>(W, st, db, n, bits, qs, nq, out)
                                  : (codec == 2 ? qg::launch_mfma<2>(W, st, db, n, bits, qs, nq, out)
                                                : qg::launch_mfma<3>(W, st, db, n, bits, qs, nq, out));
            }
            return codec == 1 ? qg::launch_bits<1>(W, grid, st, db, n, bits, qs, nq, out)
                              : (codec == 2 ? qg::launch_bits<2>(W, grid, st, db, n, bits, qs, nq, out)
                                            : qg::launch_bits<3>(W, grid, st, db, n, bits, qs, nq, out));
        default: return (int)hipErrorInvalidValue;
    }
}

}  // extern "C"
