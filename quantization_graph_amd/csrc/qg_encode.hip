// qg_encode.hip -- ExRaBitQ4 block16 encode of N base vectors
// (encodeVectorFullWithCentroid, Ours/core/hnswlib/space_rabitq.h:5845-5999,
// with fastQuantizeAbs :867-942 and the residual scales :729-786).
//
// The reference encodes one vector at a time with several strictly
// sequential floating-point recurrences (norms, the running numerator of
// the critical-point scan, the decoded-residual sums). To keep every output
// byte identical while using the whole GPU, the work is split into
// data-parallel phases (one workgroup per vector, in LDS) and recurrence
// phases (one lane per vector, 64 vectors per wave):
//
//   k_enc_norm   lane/vector   ||x - c||^2 in the reference's order
//   k_enc_rotate WG/vector     normalise, FHT, |.|, max, sort dims by
//                              (|o| desc, index asc), per-level event ranges
//   k_enc_scan   lane/vector   critical-point scan of fastQuantizeAbs as a
//                              7-way merge of the per-level event lists,
//                              ip_norm, o_obar, short/error scales
//   k_enc_codes  WG/vector     nibble codes, MSB sidecar, quantisation
//                              errors, per-block residual scales + codes
//   k_enc_rfac   lane/vector   decoded-residual norm and <x4, residual>
//
// Event structure used by k_enc_scan: dimension i with initial code c0_i
// contributes events at t = (k+1)/|o_i| for k = c0_i (if |o_i| > eps) and
// every k in (c0_i, 6] with t < t_end. Sorting dimensions by |o| descending
// makes each level k's event list a contiguous slice of that order, already
// sorted by (t, i); merging the seven slices by (t, i) reproduces the pop
// order of the reference's priority queue exactly.
#include "qg_args.h"

namespace qg {

constexpr int kEncThreads = 256;
constexpr double kEps = 1e-5;

// Compensated FP32 arithmetic of the precision mode (QR rotation): Knuth's
// branch-free TwoSum and the fma TwoProduct carry each rounding error into a
// second accumulator, so a sum of squares or of magnitudes is accurate to
// about FP32 rounding of the exact result (Ogita, Rump and Oishi's Sum2 /
// Dot2).
struct Comp {
    float s = 0.0f, c = 0.0f;
    __device__ __forceinline__ void add(float p) {
        const float t = s + p;
        const float bp = t - s;
        c = c + ((s - (t - bp)) + (p - bp));
        s = t;
    }
    __device__ __forceinline__ void add_sq(float r) {
        const float p = r * r;
        const float e = fmaf(r, r, -p);
        add(p);
        c = c + e;
    }
    __device__ __forceinline__ void merge(Comp o) {
        add(o.s);
        c = c + o.c;
    }
    __device__ __forceinline__ float value() const { return s + c; }
};

// block-wide merge of per-thread compensated sums in a fixed tree order
__device__ __forceinline__ float comp_block(Comp v, float *red_s, float *red_c) {
    const uint32_t t = threadIdx.x;
    red_s[t] = v.s;
    red_c[t] = v.c;
    __syncthreads();
    for (uint32_t h = kEncThreads / 2; h > 0; h >>= 1) {
        if (t < h) {
            Comp a;
            a.s = red_s[t];
            a.c = red_c[t];
            Comp b;
            b.s = red_s[t + h];
            b.c = red_c[t + h];
            a.merge(b);
            red_s[t] = a.s;
            red_c[t] = a.c;
        }
        __syncthreads();
    }
    Comp r;
    r.s = red_s[0];
    r.c = red_c[0];
    __syncthreads();
    return r.value();
}





// assignCentroid (space_rabitq.h:2944-2960): the first centroid of least
// float squared distance, the sum over dimensions in index order with the
// reference compiler's fused multiply-add. One lane per vector.
__global__ void k_enc_assign(EncArgs a) {
    const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.n) return;
    const float *x = a.x + v * a.dim;
    uint32_t best = 0;
    float best_d = __builtin_huge_valf();
    for (uint32_t k = 0; k < a.centroid_count; ++k) {
        const float *c = a.center + (uint64_t)k * a.dim;
        float d = 0.0f;
        for (uint32_t i = 0; i < a.dim; ++i) {
            const float diff = x[i] - c[i];
            d = fmaf(diff, diff, d);
        }
        if (d < best_d) { best_d = d; best = k; }
    }
    a.cid[v] = (uint8_t)best;
}

__device__ __forceinline__ const float *center_of(const EncArgs &a, uint64_t v) {
    return a.centroid_count > 1 ? a.center + (uint64_t)a.cid[v] * a.dim : a.center;
}

// ||x - c||^2 (float). The reference compiler vectorises this loop: the
// products of its 16- and 8-wide blocks are rounded and added in index
// order; only the scalar tail (D mod 8 elements) is fused.
__global__ void k_enc_norm(EncArgs a) {
    const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.n) return;
    const float *x = a.x + v * a.dim;
    const float *center = center_of(a, v);
    const uint32_t D = a.dim;
    if (a.cid && a.centroid_count == 1) a.cid[v] = 0;
    if (a.precision) {  // compensated sum of squares
        Comp c;
        for (uint32_t i = 0; i < D; ++i) c.add_sq(x[i] - center[i]);
        a.meta[v].rn = sqrtf(c.value());
        return;
    }
    const uint32_t vec_end = D >= 16 ? (D / 16) * 16 + ((D % 16) >= 8 ? 8 : 0) : 0;
    float s = 0.0f;
    for (uint32_t i = 0; i < D; ++i) {
        const float r = x[i] - center[i];
        if (i < vec_end) { const float p = r * r; s = s + p; }
        else s = fmaf(r, r, s);
    }
    a.meta[v].rn = sqrtf(s);
}

__device__ __forceinline__ int init_code(double t_start, float a) {
    int c = (int)fma(t_start, (double)a, kEps);
    return c < 7 ? c : 7;
}

__device__ __forceinline__ bool level_member(int k, float a, int c0, double t_end) {
    if (!((double)a > kEps) || c0 > k) return false;
    if (k == c0) return true;
    return (double)(k + 1) / (double)a < t_end;
}

__global__ void __launch_bounds__(kEncThreads) k_enc_rotate(EncArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t cd = a.cd;
    float *v = (float *)smem;                                   // cd
    unsigned long long *key = (unsigned long long *)(smem + 4 * cd);  // cd
    __shared__ int s_max_bits;
    __shared__ int s_start[7], s_end[7];
    const uint64_t vi = blockIdx.x;
    const uint32_t t = threadIdx.x;
    const float *x = a.x + vi * a.dim;
    const float rn = a.meta[vi].rn;
    if (a.precision) {
        // rot holds Q (x - c) from the rotation GEMM: the unit vector is it
        // divided by its compensated norm, which corrects the GEMM's scale
        // error as well (Q is orthogonal: the exact norm is |x - c|)
        __shared__ float red_s[kEncThreads], red_c[kEncThreads];
        Comp c;
        for (uint32_t i = t; i < cd; i += blockDim.x) {
            const float y = a.rot[vi * cd + i];
            v[i] = y;
            c.add_sq(y);
        }
        const float n2 = comp_block(c, red_s, red_c);
        const float inv = n2 > 0.0f ? 1.0f / sqrtf(n2) : 0.0f;
        for (uint32_t i = t; i < cd; i += blockDim.x) v[i] = v[i] * inv;
    } else {
        const bool scale = rn > 0.0f;
        const float inv = scale ? 1.0f / rn : 1.0f;
        const float *center = center_of(a, vi);
        for (uint32_t i = t; i < cd; i += blockDim.x) {
            float u = 0.0f;
            if (i < a.dim) {
                u = x[i] - center[i];
                if (scale) u = u * inv;
                if (!a.identity) u = u * a.signs[i];
            }
            v[i] = u;
        }
    }
    if (t == 0) s_max_bits = 0;
    if (t < 7) { s_start[t] = 0x7FFFFFFF; s_end[t] = 0; }
    __syncthreads();
    if (!a.identity && !a.precision) fht_lds(v, cd);
    float m = 0.0f;
    for (uint32_t i = t; i < cd; i += blockDim.x) {
        const float r = v[i];
        a.rot[vi * cd + i] = r;
        const float ab = fabsf(r);
        m = (m < ab) ? ab : m;
        key[i] = ((unsigned long long)(~__float_as_uint(ab)) << 32) | i;
    }
    for (int off = 32; off > 0; off >>= 1) { const float o = __shfl_xor(m, off); m = (m < o) ? o : m; }
    if ((t & 63) == 0) atomicMax(&s_max_bits, __float_as_int(m));
    __syncthreads();
    // bitonic sort of (|o| desc, index asc)
    for (uint32_t size = 2; size <= cd; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t b = t; b < cd / 2; b += blockDim.x) {
                const uint32_t lo = 2 * stride * (b / stride) + (b % stride);
                const uint32_t hi = lo + stride;
                const bool up = ((lo & size) == 0);
                const unsigned long long k0 = key[lo], k1 = key[hi];
                if (up ? (k0 > k1) : (k0 < k1)) { key[lo] = k1; key[hi] = k0; }
            }
            __syncthreads();
        }
    }
    const double max_o = (double)__int_as_float(s_max_bits);
    const double t_start = 2.0 / max_o;
    const double t_end = 17.0 / max_o;
    for (uint32_t p = t; p < cd; p += blockDim.x) {
        const unsigned long long k = key[p];
        const uint32_t idx = (uint32_t)(k & 0xFFFFFFFFu);
        const float ab = __uint_as_float(~(uint32_t)(k >> 32));
        a.sa[vi * cd + p] = ab;
        a.si[vi * cd + p] = (uint16_t)idx;
        if (max_o > kEps) {
            const int c0 = init_code(t_start, ab);
            for (int lv = 0; lv < 7; ++lv) {
                if (level_member(lv, ab, c0, t_end)) {
                    atomicMin(&s_start[lv], (int)p);
                    atomicMax(&s_end[lv], (int)p + 1);
                }
            }
        }
    }
    __syncthreads();
    if (t == 0) {
        EncMeta &mt = a.meta[vi];
        mt.max_o = max_o;
        for (int lv = 0; lv < 7; ++lv) {
            const bool any = s_end[lv] > 0;
            mt.lv_start[lv] = any ? (uint16_t)s_start[lv] : 0;
            mt.lv_end[lv] = any ? (uint16_t)s_end[lv] : 0;
        }
    }
}

// fastQuantizeAbs critical-point scan + the o_obar / L1 recurrences, one
// lane per vector.
__global__ void k_enc_scan(EncArgs a) {
    const uint64_t vi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (vi >= a.n) return;
    const uint32_t cd = a.cd;
    EncMeta mt = a.meta[vi];
    const float *rot = a.rot + vi * cd;
    const float *sa = a.sa + vi * cd;
    const uint16_t *si = a.si + vi * cd;
    double best_t = 0.0;
    float ip_norm = 1.0f;
    if (mt.max_o > kEps) {
        const double t_start = 2.0 / mt.max_o;
        double den = (double)cd * 0.25;
        double num = 0.0;
        for (uint32_t i = 0; i < cd; ++i) {
            const float ab = fabsf(rot[i]);
            const int c = init_code(t_start, ab);
            den += (double)(c * c + c);
            num += ((double)c + 0.5) * (double)ab;
        }
        uint32_t cur[7], end[7];
        double ht[7];
        uint32_t hi[7];
        float ha[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            cur[k] = mt.lv_start[k];
            end[k] = mt.lv_end[k];
            if (cur[k] < end[k]) {
                ha[k] = sa[cur[k]];
                hi[k] = si[cur[k]];
                ht[k] = (double)(k + 1) / (double)ha[k];
            } else {
                ha[k] = 0.0f; hi[k] = 0xFFFFFFFFu; ht[k] = __builtin_inf();
            }
        }
        double max_ip = 0.0;
        for (;;) {
            int kb = -1;
            double tb = __builtin_inf();
            uint32_t ib = 0xFFFFFFFFu;
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                if (cur[k] < end[k] && (ht[k] < tb || (ht[k] == tb && hi[k] < ib))) {
                    kb = k; tb = ht[k]; ib = hi[k];
                }
            }
            if (kb < 0) break;
            float ab = 0.0f;
#pragma unroll
            for (int k = 0; k < 7; ++k) if (k == kb) ab = ha[k];
            num += (double)ab;
            den += 2.0 * (double)(kb + 1);
            const double ip = num / sqrt(den);
            if (ip > max_ip) { max_ip = ip; best_t = tb; }
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                if (k == kb) {
                    ++cur[k];
                    if (cur[k] < end[k]) {
                        ha[k] = sa[cur[k]];
                        hi[k] = si[cur[k]];
                        ht[k] = (double)(k + 1) / (double)ha[k];
                    }
                }
            }
        }
        double num2 = 0.0;
        for (uint32_t i = 0; i < cd; ++i) {
            const float ab = fabsf(rot[i]);
            const int c = init_code(best_t, ab);
            num2 += ((double)c + 0.5) * (double)ab;
        }
        ip_norm = num2 > kEps ? (float)(1.0 / num2) : 1.0f;
        if (!__builtin_isfinite(ip_norm)) ip_norm = 1.0f;
    }
    const float rn = mt.rn;
    const float long_scale = 2.0f * rn * ip_norm;
    if (a.precision) {
        // the factors from a compensated FP32 sum of |o| (no FP64)
        Comp l1c;
        for (uint32_t i = 0; i < cd; ++i) l1c.add(fabsf(rot[i]));
        const float l1 = l1c.value(), half = 0.5f * l1;
        float short_scale = 0.0f, error_scale = 0.0f;
        if (rn > 0.0f && half > 1e-12f) {
            short_scale = 2.0f * rn / half;
            const float sq = sqrtf((float)cd);
            float o = l1 / sq;
            if (!__builtin_isfinite(o)) o = 0.8f;
            o = (1e-6f < o) ? o : 1e-6f;
            o = (0.999999f < o) ? 0.999999f : o;
            const float o2 = o * o;
            float r = (1.0f - o2) / o2;
            r = (0.0f < r) ? r : 0.0f;
            error_scale = sqrtf(r) * (0.359375f / sq) * 2.0f * rn;
        }
        a.hdr[vi] = make_float2(rn * rn, long_scale);
        a.sfac[vi] = make_float2(short_scale, error_scale);
        a.meta[vi].best_t = best_t;
        a.meta[vi].ip_norm = ip_norm;
        return;
    }
    double o_obar = 0.0, half_l1 = 0.0;
    const double inv_sqrt_d = 1.0 / sqrt((double)cd);
    for (uint32_t i = 0; i < cd; ++i) {
        const float r = rot[i];
        const double sign = r > 0.0f ? 1.0 : -1.0;
        const double rs = (double)r * sign;
        o_obar = fma(rs, inv_sqrt_d, o_obar);
        half_l1 += 0.5 * fabs((double)r);
    }
    float short_scale = 0.0f, error_scale = 0.0f;
    if (rn > 0.0f && half_l1 > 1e-12) {
        short_scale = (float)(2.0 * (double)rn / half_l1);
        if (!__builtin_isfinite(o_obar)) o_obar = 0.8;
        o_obar = (1e-6 < o_obar) ? o_obar : 1e-6;                // std::max(1e-6, o)
        o_obar = (0.999999 < o_obar) ? 0.999999 : o_obar;        // std::min(0.999999, o)
        const double o2 = o_obar * o_obar;
        const double fac = ldexp(5.75, -4) / sqrt((double)cd);
        double r = (1.0 - o2) / o2;
        r = (0.0 < r) ? r : 0.0;                                   // std::max(0.0, r)
        error_scale = (float)(sqrt(r) * fac * 2.0 * (double)rn);
    }
    a.hdr[vi] = make_float2(rn * rn, long_scale);
    a.sfac[vi] = make_float2(short_scale, error_scale);
    a.meta[vi].best_t = best_t;
    a.meta[vi].ip_norm = ip_norm;
}

__device__ __forceinline__ int clamp7_round(double v) {
    int q = (int)round(v);
    q = q < 7 ? q : 7;      // std::min(qmax, .)
    return q > -7 ? q : -7; // std::max(qmin, .)
}

__global__ void __launch_bounds__(kEncThreads) k_enc_codes(EncArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t cd = a.cd;
    float *err = (float *)smem;                 // cd
    uint8_t *nib = (uint8_t *)(smem + 4 * cd);  // cd primary nibbles
    const uint64_t vi = blockIdx.x;
    const uint32_t t = threadIdx.x;
    const EncMeta mt = a.meta[vi];
    const float *rot = a.rot + vi * cd;
    const float2 h = a.hdr[vi];
    const float ls = h.y;
    const double t0 = mt.best_t;
    const bool quant = mt.max_o > kEps;
    for (uint32_t i = t; i < cd; i += blockDim.x) {
        const float r = rot[i];
        const int m = quant ? init_code(t0, fabsf(r)) : 0;
        const uint8_t c = r > 0.0f ? (uint8_t)(8 + m) : (uint8_t)(7 - m);
        nib[i] = c;
        const double tv = (double)mt.rn * (double)r;
        const double x4 = 0.5 * (double)ls * ((double)c - 7.5);
        err[i] = (float)(tv - x4);
    }
    __syncthreads();
    uint8_t *code_row = a.codes + vi * (cd / 2);
    uint8_t *msb_row = a.gate + vi * gate_row_bytes(cd);
    for (uint32_t p = t; p < cd / 2; p += blockDim.x)
        code_row[p] = (uint8_t)(nib[2 * p] | (nib[2 * p + 1] << 4));
    for (uint32_t b = t; b < cd / 8; b += blockDim.x) {
        uint8_t m = 0;
        for (int j = 0; j < 8; ++j) m |= (uint8_t)((nib[8 * b + j] >> 3) << j);
        msb_row[b] = m;
    }
    // residual block scales and codes, one thread per 16-dim block
    uint8_t *rrow = a.rcodes + vi * (cd / 2);
    for (uint32_t blk = t; blk < a.block_count; blk += blockDim.x) {
        const uint32_t b0 = blk * a.block;
        const uint32_t b1 = (b0 + a.block < cd) ? b0 + a.block : cd;
        float mx = 0.0f;
        for (uint32_t i = b0; i < b1; ++i) { const float e = fabsf(err[i]); mx = (mx < e) ? e : mx; }
        float scale = mx <= 0.0f ? 0.0f : mx / 7.0f;
        if (a.mse && scale != 0.0f && __builtin_isfinite(scale)) {
            for (int it = 0; it < 8; ++it) {
                double num = 0.0, den = 0.0;
                for (uint32_t i = b0; i < b1; ++i) {
                    const int q = clamp7_round((double)err[i] / (double)scale);
                    num += (double)err[i] * (double)q;
                    den += (double)q * (double)q;
                }
                if (den <= 0.0) break;
                const float next = (float)(num / den);
                if (next <= 0.0f || !__builtin_isfinite(next)) break;
                const float lim = 1e-6f * ((1.0f < scale) ? scale : 1.0f);
                if (fabsf(next - scale) <= lim) { scale = next; break; }
                scale = next;
            }
        }
        // fp16 scale storage (the formal configuration); codes are quantised
        // against the re-loaded scale (:5963-5978)
        const uint16_t hb = float_to_half_ref(scale);
        a.rscale[vi * a.block_count + blk] = hb;
        const float stored = half_to_float_ref(hb);
        const bool live = stored != 0.0f && __builtin_isfinite(stored);
        for (uint32_t i = b0; i < b1; i += 2) {
            int q0 = 0, q1 = 0;
            if (live) {
                q0 = clamp7_round((double)err[i] / (double)stored);
                q1 = (i + 1 < b1) ? clamp7_round((double)err[i + 1] / (double)stored) : 0;
            }
            rrow[i / 2] = (uint8_t)(((uint32_t)q0 & 0xFu) | (((uint32_t)q1 & 0xFu) << 4));
        }
    }
    // paper-prune sidecar factors (extract_paper_prune_sidecar, :4364-4392)
    if (t == 0) {
        const float2 sf = a.sfac[vi];
        float4 f = make_float4(h.x, 0.0f, 0.0f, 0.0f);
        uint8_t valid = 0;
        if (cd > 1 && h.x > 0.0f && sf.x > 0.0f) {
            const float dn = sqrtf(h.x);
            f.y = dn;
            float al = 4.0f * dn / (sf.x * sqrtf((float)cd));
            al = (al < 1.0f) ? al : 1.0f;  // std::min(1.0f, .)
            if (al > 0.0f && __builtin_isfinite(al)) {
                const float a2 = al * al;
                f.z = sf.x;
                float r = (1.0f - a2) / a2;
                r = (0.0f < r) ? r : 0.0f;
                f.w = 2.0f * dn * sqrtf(r) / sqrtf((float)(cd - 1));
                valid = __builtin_isfinite(f.w) ? 1 : 0;
            }
        }
        gate_store_tail(a.gate + vi * gate_row_bytes(cd), cd, f, h, valid);
    }
}

// Decoded-residual recurrences (:5980-5991), one lane per vector.
__global__ void k_enc_rfac(EncArgs a) {
    const uint64_t vi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (vi >= a.n) return;
    const uint32_t cd = a.cd;
    const float ls = a.hdr[vi].y;
    const uint8_t *code = a.codes + vi * (cd / 2);
    const uint8_t *rc = a.rcodes + vi * (cd / 2);
    const uint16_t *sc = a.rscale + vi * a.block_count;
    double dec = 0.0, x4ip = 0.0;
    for (uint32_t blk = 0; blk < a.block_count; ++blk) {
        const float st = half_to_float_ref(sc[blk]);
        if (st == 0.0f || !__builtin_isfinite(st)) continue;
        const uint32_t b0 = blk * a.block;
        const uint32_t b1 = (b0 + a.block < cd) ? b0 + a.block : cd;
        for (uint32_t i = b0; i < b1; ++i) {
            const uint32_t byte = rc[i >> 1];
            int q = (int)((i & 1u) ? (byte >> 4) : (byte & 0xFu));
            q -= (q > 7) ? 16 : 0;
            const uint32_t cb = code[i >> 1];
            const uint32_t c = (i & 1u) ? (cb >> 4) : (cb & 0xFu);
            const double dr = (double)st * (double)q;
            const double x4 = 0.5 * (double)ls * ((double)c - 7.5);
            dec = fma(dr, dr, dec);
            x4ip = fma(x4, dr, x4ip);
        }
    }
    a.rfac[vi] = make_float2((float)dec, (float)x4ip);
}

// ------------------------------------------------------ record transcoding


// SoA -> reference AoS (copyCompactPayloadFromFull / copyResidualRecord
// FromFull layouts, :6005-6029): compact = hdr|short factors|codes|cid,
// residual = residual factors|fp16 scales|codes.
__global__ void k_pack(PackArgs a) {
    const uint64_t v = blockIdx.x;
    const uint32_t cd = a.cd;
    uint8_t *c = a.compact + v * a.compact_stride;
    uint8_t *r = a.residual + v * a.residual_stride;
    const uint32_t nc = 16 + cd / 2 + 1;
    for (uint32_t i = threadIdx.x; i < nc; i += blockDim.x) {
        uint8_t b;
        if (i < 8) b = ((const uint8_t *)&a.ix.hdr[v])[i];
        else if (i < 16) b = ((const uint8_t *)&a.ix.sfac[v])[i - 8];
        else if (i < 16 + cd / 2) b = a.ix.codes[v * (cd / 2) + (i - 16)];
        else b = a.ix.cid ? a.ix.cid[v] : 0;  // centroid id
        c[i] = b;
    }
    const uint32_t sb = 2 * a.block_count;
    const uint32_t nr = 8 + sb + cd / 2;
    for (uint32_t i = threadIdx.x; i < nr; i += blockDim.x) {
        uint8_t b;
        if (i < 8) b = ((const uint8_t *)&a.rfac[v])[i];
        else if (i < 8 + sb) b = ((const uint8_t *)(a.ix.rscale + v * a.block_count))[i - 8];
        else b = a.ix.rcodes[v * (cd / 2) + (i - 8 - sb)];
        r[i] = b;
    }
}

// AoS -> SoA, re-deriving the MSB plane and sidecar factors from the
// compact record exactly as extract_paper_prune_sidecar does.
__global__ void k_unpack(PackArgs a) {
    const uint64_t v = blockIdx.x;
    const uint32_t cd = a.cd;
    const uint8_t *c = a.in_compact + v * a.compact_stride;
    const uint8_t *r = a.in_residual + v * a.residual_stride;
    for (uint32_t i = threadIdx.x; i < cd / 2; i += blockDim.x) a.codes[v * (cd / 2) + i] = c[16 + i];
    if (a.cid_out && threadIdx.x == 0) a.cid_out[v] = c[16 + cd / 2];
    for (uint32_t b = threadIdx.x; b < cd / 8; b += blockDim.x) {
        uint8_t m = 0;
        for (int j = 0; j < 4; ++j) {
            const uint8_t byte = c[16 + 4 * b + j];
            m |= (uint8_t)(((byte >> 3) & 1u) << (2 * j));
            m |= (uint8_t)(((byte >> 7) & 1u) << (2 * j + 1));
        }
        a.gate[v * gate_row_bytes(cd) + b] = m;
    }
    const uint32_t sb = 2 * a.block_count;
    if (a.in_residual) {
        for (uint32_t i = threadIdx.x; i < a.block_count; i += blockDim.x)
            a.rscale[v * a.block_count + i] = (uint16_t)(r[8 + 2 * i] | (r[9 + 2 * i] << 8));
        for (uint32_t i = threadIdx.x; i < cd / 2; i += blockDim.x) a.rcodes[v * (cd / 2) + i] = r[8 + sb + i];
    }
    if (threadIdx.x == 0) {
        float hv[4];
        for (int k = 0; k < 4; ++k) {
            uint32_t w = (uint32_t)c[4 * k] | ((uint32_t)c[4 * k + 1] << 8) | ((uint32_t)c[4 * k + 2] << 16) |
                         ((uint32_t)c[4 * k + 3] << 24);
            hv[k] = __uint_as_float(w);
        }
        a.hdr[v] = make_float2(hv[0], hv[1]);
        a.sfac[v] = make_float2(hv[2], hv[3]);
        if (a.in_residual && a.rfac_out) {
            float rv[2];
            for (int k = 0; k < 2; ++k) {
                uint32_t w = (uint32_t)r[4 * k] | ((uint32_t)r[4 * k + 1] << 8) | ((uint32_t)r[4 * k + 2] << 16) |
                             ((uint32_t)r[4 * k + 3] << 24);
                rv[k] = __uint_as_float(w);
            }
            a.rfac_out[v] = make_float2(rv[0], rv[1]);
        }
        float4 f = make_float4(hv[0], 0.0f, 0.0f, 0.0f);
        uint8_t valid = 0;
        if (cd > 1 && hv[0] > 0.0f && hv[2] > 0.0f) {
            const float dn = sqrtf(hv[0]);
            f.y = dn;
            float al = 4.0f * dn / (hv[2] * sqrtf((float)cd));
            al = (al < 1.0f) ? al : 1.0f;
            if (al > 0.0f && __builtin_isfinite(al)) {
                const float a2 = al * al;
                f.z = hv[2];
                float rr = (1.0f - a2) / a2;
                rr = (0.0f < rr) ? rr : 0.0f;
                f.w = 2.0f * dn * sqrtf(rr) / sqrtf((float)(cd - 1));
                valid = __builtin_isfinite(f.w) ? 1 : 0;
            }
        }
        gate_store_tail(a.gate + v * gate_row_bytes(cd), cd, f, make_float2(hv[0], hv[1]), valid);
    }
}

}  // namespace qg

extern "C" {

int qg_launch_encode(const qg::EncArgs *a, hipStream_t stream) {
    if (a->n == 0) return 0;
    const unsigned lane_blocks = (unsigned)((a->n + 63) / 64);
    if (a->cid && a->centroid_count > 1) hipLaunchKernelGGL(qg::k_enc_assign, dim3(lane_blocks), dim3(64), 0, stream, *a);
    hipLaunchKernelGGL(qg::k_enc_norm, dim3(lane_blocks), dim3(64), 0, stream, *a);
    hipLaunchKernelGGL(qg::k_enc_rotate, dim3((unsigned)a->n), dim3(qg::kEncThreads), (size_t)a->cd * 12, stream, *a);
    hipLaunchKernelGGL(qg::k_enc_scan, dim3(lane_blocks), dim3(64), 0, stream, *a);
    hipLaunchKernelGGL(qg::k_enc_codes, dim3((unsigned)a->n), dim3(qg::kEncThreads), (size_t)a->cd * 5, stream, *a);
    hipLaunchKernelGGL(qg::k_enc_rfac, dim3(lane_blocks), dim3(64), 0, stream, *a);
    return (int)hipGetLastError();
}

int qg_launch_pack(const qg::PackArgs *a, uint64_t n, int unpack, hipStream_t stream) {
    if (n == 0) return 0;
    if (unpack) hipLaunchKernelGGL(qg::k_unpack, dim3((unsigned)n), dim3(128), 0, stream, *a);
    else hipLaunchKernelGGL(qg::k_pack, dim3((unsigned)n), dim3(128), 0, stream, *a);
    return (int)hipGetLastError();
}

}  // extern "C"
