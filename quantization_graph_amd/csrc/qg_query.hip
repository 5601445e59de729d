// qg_query.hip -- batched query preparation (FP32 rotated residual + INT8
// dithered code) for N queries in one launch.
//
// Restates initializePreparedQuery + buildCentroidQuery
// (Ours/core/hnswlib/space_rabitq.h:3429-3698) for K = 1:
//   rotated   = FHT(diag(signs) * pad(query))            (:990-1000, 944-988)
//   q         = rotated - rotated_center
//   sum       = sequential float sum of (even + odd) per pair; half = 0.5*sum
//   pos       = sequential float sum of max(0,even) + max(0,odd)
//   ||q-c||^2 = float(||raw||^2 + ||c||^2 - 2<raw,c>)   (doubles, :3447-3454)
//   INT8: scale = max|q|/128, x = clamp(floor(q/scale + u), -128, 127) with u
//   from mt19937_64 seeded by FNV-1a over the raw bytes (:3540-3659), two
//   draws per dimension, the first used for INT8 (the other codecs draw
//   nothing, so their rows are not hashed).
//   B1 / B1Main: sign code x = +1 where q_i > 0 else -1, and its sum.
// Serial recurrences (the float sums, FNV-1a, the mt19937_64 seeding) run on
// single lanes in the reference order; everything else is data-parallel.
#include "qg_args.h"

namespace qg {

constexpr int kPrepThreads = 256;

// FNV-1a over a row's bytes (the INT8 dither is its one consumer), folded
// in the reference's byte order, one value's four bytes at a time.
__device__ __forceinline__ void fnv_fold(float v, uint64_t &h) {
    const uint32_t b = __float_as_uint(v);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        h ^= (uint64_t)((b >> (8 * k)) & 0xFFu);
        h *= 1099511628211ull;
    }
}

// A row's serial recurrences are independent chains, and each goes to a
// wavefront of its own, one row per lane, in index order: wave 0 sums
// ||row||^2 and wave 1 <row, c> (FP64 add chains), and with HASH wave 2 runs
// the row's FNV-1a (only the INT8 dither needs it). The center is widened to
// double in LDS once per block. Splitting the chains divides each lane's
// issue between waves on different SIMDs, and 64-value chunks keep the next
// loads in flight meanwhile.
constexpr uint32_t kNormChunk = 64;

template <bool HASH>
__global__ void __launch_bounds__(HASH ? 192 : 128) k_row_norms(const float *__restrict__ rows, uint64_t n,
                                                                uint32_t dim, const float *__restrict__ center,
                                                                uint64_t *__restrict__ seed_out,
                                                                double *__restrict__ norm_out,
                                                                double *__restrict__ dot_out) {
    extern __shared__ __align__(16) double cs[];  // [dim] when center is given
    const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
    if (center)
        for (uint32_t i = t; i < dim; i += blockDim.x) cs[i] = (double)center[i];
    __syncthreads();
    const uint64_t r = (uint64_t)blockIdx.x * 64 + lane;
    const bool live = r < n;
    if (HASH && wv == 2) {
        const float *x = rows + (live ? r : 0) * dim;  // idle lanes walk row 0
        uint64_t h = 1469598103934665603ull;
        const uint32_t full = dim / kNormChunk * kNormChunk;
        float cur[kNormChunk], nxt[kNormChunk];
        if (full) {
#pragma unroll
            for (uint32_t u = 0; u < kNormChunk; ++u) cur[u] = x[u];
        }
        for (uint32_t base = 0; base < full; base += kNormChunk) {
            if (base + kNormChunk < full) {
#pragma unroll
                for (uint32_t u = 0; u < kNormChunk; ++u) nxt[u] = x[base + kNormChunk + u];
            }
#pragma unroll
            for (uint32_t u = 0; u < kNormChunk; ++u) fnv_fold(cur[u], h);
#pragma unroll
            for (uint32_t u = 0; u < kNormChunk; ++u) cur[u] = nxt[u];
        }
        for (uint32_t i = full; i < dim; ++i) fnv_fold(x[i], h);
        if (live) seed_out[r] = h;
        return;
    }
    if (wv == 1 && !center) {
        if (live && dot_out) dot_out[r] = 0.0;
        return;
    }
    const float *x = rows + (live ? r : 0) * dim;  // idle lanes walk row 0
    double acc = 0.0;
    const uint32_t full = dim / kNormChunk * kNormChunk;
    float cur[kNormChunk], nxt[kNormChunk];
    if (full) {
#pragma unroll
        for (uint32_t u = 0; u < kNormChunk; ++u) cur[u] = x[u];
    }
    for (uint32_t base = 0; base < full; base += kNormChunk) {
        if (base + kNormChunk < full) {
#pragma unroll
            for (uint32_t u = 0; u < kNormChunk; ++u) nxt[u] = x[base + kNormChunk + u];
        }
        if (wv == 0) {
#pragma unroll
            for (uint32_t u = 0; u < kNormChunk; ++u) {
                const double dv = (double)cur[u];
                acc += dv * dv;  // exact product: ordering of the add is what matters
            }
        } else {
#pragma unroll
            for (uint32_t u = 0; u < kNormChunk; ++u) acc += (double)cur[u] * cs[base + u];
        }
#pragma unroll
        for (uint32_t u = 0; u < kNormChunk; ++u) cur[u] = nxt[u];
    }
    for (uint32_t i = full; i < dim; ++i) {
        const double dv = (double)x[i];
        acc += wv == 0 ? dv * dv : dv * cs[i];
    }
    if (!live) return;
    if (wv == 0) {
        if (norm_out) norm_out[r] = acc;
    } else if (dot_out) {
        dot_out[r] = acc;
    }
}

__device__ __forceinline__ uint64_t mt64_temper(uint64_t x) {
    x ^= (x >> 29) & 0x5555555555555555ull;
    x ^= (x << 17) & 0x71D67FFFEDA60000ull;
    x ^= (x << 37) & 0xFFF7EEE000000000ull;
    x ^= x >> 43;
    return x;
}

// One std::mt19937_64 state refill (n = 312, m = 156), parallel form: the
// first 156 words depend only on old words, the rest on old words and the
// first half's new words.
__device__ void mt64_twist(uint64_t *mt) {
    const uint64_t UM = 0xFFFFFFFF80000000ull, LM = 0x7FFFFFFFull, A = 0xB5026F5AA96619E9ull;
    uint64_t v[2] = {0, 0};
    int c = 0;
    for (uint32_t i = threadIdx.x; i < 156; i += blockDim.x, ++c) {
        const uint64_t y = (mt[i] & UM) | (mt[i + 1] & LM);
        v[c] = mt[i + 156] ^ (y >> 1) ^ ((y & 1ull) ? A : 0ull);
    }
    __syncthreads();
    c = 0;
    for (uint32_t i = threadIdx.x; i < 156; i += blockDim.x, ++c) mt[i] = v[c];
    __syncthreads();
    c = 0;
    for (uint32_t i = 156 + threadIdx.x; i < 312; i += blockDim.x, ++c) {
        const uint64_t y = (mt[i] & UM) | (mt[(i + 1) % 312] & LM);
        v[c] = mt[i - 156] ^ (y >> 1) ^ ((y & 1ull) ? A : 0ull);
    }
    __syncthreads();
    c = 0;
    for (uint32_t i = 156 + threadIdx.x; i < 312; i += blockDim.x, ++c) mt[i] = v[c];
    __syncthreads();
}

__global__ void __launch_bounds__(kPrepThreads) k_query_prep(PrepParams p) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t cd = p.cd;
    float *v = (float *)smem;                          // cd floats
    uint64_t *mt = (uint64_t *)(smem + 4 * cd);        // 312 words
    __shared__ float s_sum, s_pos, s_max, s_abs;
    __shared__ long long s_csum;
    const uint64_t qi = blockIdx.x;
    const float *raw = p.raw + qi * p.dim;
    const uint32_t t = threadIdx.x;

    if (p.pre_rotated) {  // QR rotation: q_out holds Q * pad(raw) from the rotation GEMM
        for (uint32_t i = t; i < cd; i += blockDim.x) v[i] = p.q_out[qi * cd + i];
    } else {
        for (uint32_t i = t; i < cd; i += blockDim.x)
            v[i] = (i < p.dim) ? (p.identity ? raw[i] : raw[i] * p.signs[i]) : 0.0f;
    }
    if (t == 0) { s_max = 0.0f; s_csum = 0; }
    __syncthreads();
    if (!p.identity && !p.pre_rotated) fht_lds(v, cd);
    for (uint32_t i = t; i < cd; i += blockDim.x) {
        const float d = v[i] - p.rot_center[i];
        v[i] = d;
        p.q_out[qi * cd + i] = d;
    }
    __syncthreads();
    // Serial float recurrences of buildCentroidQuery (:3471-3486) on two
    // lanes of different waves; max|q| by reduction (order-free).
    if (t == 0) {
        float s = 0.0f;
        for (uint32_t pr = 0; pr < cd / 2; ++pr) { const float e = v[2 * pr], o = v[2 * pr + 1]; s += e + o; }
        s_sum = s;
    } else if (t == 64) {
        float s = 0.0f;
        for (uint32_t pr = 0; pr < cd / 2; ++pr) {
            const float e = v[2 * pr], o = v[2 * pr + 1];
            const float pe = (0.0f < e) ? e : 0.0f, po = (0.0f < o) ? o : 0.0f;  // std::max(0.0f, x)
            s += pe + po;
        }
        s_pos = s;
    } else if (t == 128 && (p.codec == 1 || p.codec == 4)) {
        // the sign code's |q| sum in index order (:3497-3499)
        float s = 0.0f;
        for (uint32_t i = 0; i < cd; ++i) s += fabsf(v[i]);
        s_abs = s;
    }
    float m = 0.0f;
    for (uint32_t i = t; i < cd; i += blockDim.x) { const float a = fabsf(v[i]); m = (m < a) ? a : m; }
    for (int off = 32; off > 0; off >>= 1) { const float o = __shfl_xor(m, off); m = (m < o) ? o : m; }
    if ((t & 63) == 0) atomicMax((int *)&s_max, __float_as_int(m));  // non-negative floats order as ints
    __syncthreads();

    double qn2d = p.raw_norm[qi] + (double)p.center_norm_sqr - 2.0 * p.raw_dot[qi];
    float qn2 = (float)qn2d;
    qn2 = (0.0f < qn2) ? qn2 : 0.0f;  // std::max(0.0f, x)
    const float qn = sqrtf(qn2);
    const float max_abs = s_max;
    float scale = 1.0f;
    // INT8 or INT4 (:3536-3652): scale max|q| / 128 or / 8, the first or the
    // second of each dimension's two dither draws, clamp to the code's range
    const bool int8 = (p.codec == 3 || p.codec == 2) && (max_abs > 0.0f);
    const bool is4 = p.codec == 2;
    if (int8) {
        scale = is4 ? max_abs / 8.0f : max_abs / 128.0f;
        if (t == 0) {
            uint64_t sd = p.seed[qi];
            sd ^= (uint64_t)p.centroid_id + 0x9E3779B97F4A7C15ull + (sd << 6) + (sd >> 2);
            mt[0] = sd;
            for (int i = 1; i < 312; ++i) mt[i] = 6364136223846793005ull * (mt[i - 1] ^ (mt[i - 1] >> 62)) + (uint64_t)i;
        }
        __syncthreads();
        const uint32_t draws = 2 * cd;
        long long csum = 0;
        for (uint32_t base = 0; base < draws; base += 312) {
            mt64_twist(mt);
            for (uint32_t k = t; k < 312; k += blockDim.x) {
                const uint32_t idx = base + k;
                if (idx < draws && (idx & 1u) == (is4 ? 1u : 0u)) {
                    const uint32_t i = idx >> 1;
                    double u = (double)mt64_temper(mt[k]) * 0x1p-64;
                    if (u >= 1.0) u = 0x1.fffffffffffffp-1;  // nextafter(1.0, 0.0)
                    int x = (int)floor((double)v[i] / (double)scale + u);
                    const int hi = is4 ? 7 : 127, lo = is4 ? -8 : -128;
                    x = x > hi ? hi : (x < lo ? lo : x);
                    p.x8_out[qi * cd + i] = (int8_t)x;
                    // regroup: per 8 dims, 4 even dims then 4 odd dims
                    const uint32_t g = i >> 3, w = i & 7u;
                    p.x8p_out[qi * cd + 8 * g + ((w & 1u) ? 4 : 0) + (w >> 1)] = (int8_t)x;
                    csum += x;
                }
            }
            __syncthreads();
        }
        for (int off = 32; off > 0; off >>= 1) csum += __shfl_xor(csum, off);
        if ((t & 63) == 0) atomicAdd((unsigned long long *)&s_csum, (unsigned long long)csum);
        // the error norm's ordered FP64 chain runs in k_query_err, one lane per query
    } else if (p.codec == 1 || p.codec == 4) {
        // B1 / B1Main sign code (:3493-3511): +1 where q_i > 0, else -1; the
        // signed sum goes where the INT8 code sum goes.
        long long csum = 0;
        for (uint32_t i = t; i < cd; i += blockDim.x) {
            const int x = (v[i] > 0.0f) ? 1 : -1;
            p.x8_out[qi * cd + i] = (int8_t)x;
            const uint32_t g = i >> 3, w = i & 7u;
            p.x8p_out[qi * cd + 8 * g + ((w & 1u) ? 4 : 0) + (w >> 1)] = (int8_t)x;
            csum += x;
        }
        for (int off = 32; off > 0; off >>= 1) csum += __shfl_xor(csum, off);
        if ((t & 63) == 0) atomicAdd((unsigned long long *)&s_csum, (unsigned long long)csum);
    } else {
        for (uint32_t i = t; i < cd; i += blockDim.x) { p.x8_out[qi * cd + i] = 0; p.x8p_out[qi * cd + i] = 0; }
    }
    __syncthreads();
    if (t == 0) {
        QScalars s;
        s.query_norm = qn;
        s.query_norm_sqr = qn2;
        s.half_sum = s_sum * 0.5f;
        s.sum = s_sum;
        s.positive_sum = s_pos;
        s.int8_scale = scale;
        s.int8_error_norm = 0.0f;  // filled by k_query_err for INT8 / INT4 queries
        s.max_abs_bits = __float_as_int(max_abs);
        s.int8_code_sum = (int8 || p.codec == 1 || p.codec == 4) ? (int64_t)s_csum : 0;
        s.b1_alpha = 0.0f;
        s.pad = 0;
        if (p.codec == 1 || p.codec == 4) {
            // the sign code's scale factors (:3512-3524)
            const float abs_sum = s_abs;
            s.int8_scale = 0.0f;
            if (qn > 0.0f && abs_sum > 0.0f && cd > 1) {
                const float a_q = abs_sum / qn;
                float alpha_q = a_q / sqrtf((float)cd);
                alpha_q = (1.0f < alpha_q) ? 1.0f : alpha_q;  // std::min(1.0f, .)
                const float alpha2 = alpha_q * alpha_q;
                s.b1_alpha = alpha_q;
                s.int8_scale = 4.0f * qn / a_q;
                float r = (1.0f - alpha2) / alpha2;
                r = (0.0f < r) ? r : 0.0f;  // std::max(0.0f, .)
                s.int8_error_norm = 2.0f * qn * sqrtf(r) / sqrtf((float)(cd - 1));
            }
        }
        p.qs_out[qi] = s;
    }
}

// The INT8 error norm (:3592-3600): sqrt of sum_i e_i^2 in index order as
// one FP64 fma chain, e_i = q_i - scale x_i as the dither computed it. One
// 64-thread workgroup per query: its lanes form the e_i into LDS, then lane 0
// runs the chain, one LDS value and one fused multiply-add per dimension (the
// e_i are formed off the chain; the product scale * x_i is exact in double).
__global__ void __launch_bounds__(64) k_query_err(const float *q, const int8_t *x8, uint64_t nq, uint32_t cd,
                                                  QScalars *qs) {
    extern __shared__ __align__(16) double es[];  // [cd]
    const uint64_t qi = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    QScalars &s = qs[qi];
    const float max_abs = __int_as_float(s.max_abs_bits);
    if (!(max_abs > 0.0f)) return;  // no INT8 code: the norm stays 0 (the whole block returns)
    const double scale = (double)s.int8_scale;
    const float *qr = q + qi * cd;
    const int8_t *xr = x8 + qi * cd;
    for (uint32_t i = lane; i < cd; i += 64) es[i] = (double)qr[i] - scale * (double)xr[i];
    __syncthreads();
    if (lane != 0) return;
    double a = 0.0;
    for (uint32_t i = 0; i < cd; i += 2) {  // cd is a multiple of 64
        const double2 e = *(const double2 *)(es + i);
        a = fma(e.x, e.x, a);
        a = fma(e.y, e.y, a);
    }
    s.int8_error_norm = (float)sqrt(a);
}

// Rotation only (setCentroids' rotate, :2918-2936): rows -> FHT(signs * x).
__global__ void __launch_bounds__(kPrepThreads) k_rotate_rows(const float *rows, uint64_t n, uint32_t dim,
                                                              uint32_t cd, int identity,
                                                              const float *signs, float *out) {
    extern __shared__ __align__(16) unsigned char smem[];
    float *v = (float *)smem;
    const float *x = rows + (uint64_t)blockIdx.x * dim;
    for (uint32_t i = threadIdx.x; i < cd; i += blockDim.x)
        v[i] = (i < dim) ? (identity ? x[i] : x[i] * signs[i]) : 0.0f;
    __syncthreads();
    if (!identity) fht_lds(v, cd);
    for (uint32_t i = threadIdx.x; i < cd; i += blockDim.x) out[(uint64_t)blockIdx.x * cd + i] = v[i];
}

}  // namespace qg

extern "C" {

int qg_launch_row_seed(const float *rows, uint64_t n, uint32_t dim, const float *center,
                       uint64_t *seed, double *norm, double *dot, hipStream_t stream) {
    if (n == 0) return 0;
    const dim3 grid((unsigned)((n + 63) / 64));
    const size_t smem = center ? (size_t)dim * 8 : 0;
    if (seed)
        hipLaunchKernelGGL(qg::k_row_norms<true>, grid, dim3(192), smem, stream, rows, n, dim, center, seed, norm,
                           dot);
    else
        hipLaunchKernelGGL(qg::k_row_norms<false>, grid, dim3(128), smem, stream, rows, n, dim, center, nullptr,
                           norm, dot);
    return (int)hipGetLastError();
}

int qg_launch_query_prep(const qg::PrepParams *p, hipStream_t stream) {
    if (p->nq == 0) return 0;
    const size_t smem = (size_t)p->cd * 4 + 312 * 8;
    hipLaunchKernelGGL(qg::k_query_prep, dim3((unsigned)p->nq), dim3(qg::kPrepThreads), smem, stream, *p);
    if (p->codec == 3 || p->codec == 2)
        hipLaunchKernelGGL(qg::k_query_err, dim3((unsigned)p->nq), dim3(64), (size_t)p->cd * 8, stream, p->q_out,
                           p->x8_out, p->nq, p->cd, p->qs_out);
    return (int)hipGetLastError();
}

int qg_launch_rotate_rows(const float *rows, uint64_t n, uint32_t dim, uint32_t cd, int identity,
                          const float *signs, float *out, hipStream_t stream) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(qg::k_rotate_rows, dim3((unsigned)n), dim3(qg::kPrepThreads), (size_t)cd * 4, stream,
                       rows, n, dim, cd, identity, signs, out);
    return (int)hipGetLastError();
}

}  // extern "C"
