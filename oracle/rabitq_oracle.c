/*
 * rabitq_oracle.c -- CPU restatement of the ExRaBitQ4 distance path.
 *
 * TEST INFRASTRUCTURE ONLY (see rabitq_oracle.h). Compiled with
 * -ffp-contract=off so that every multiply-add that the reference compiler
 * contracts is spelled out explicitly through ro_fma_mask.
 *
 * Citations are to the reference checkout: Ours/core/hnswlib/space_rabitq.h
 * unless another file is named.
 */
#include "rabitq_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Contraction pattern of the reference compiled with g++ 11 -O3 (pinned
 * against oracle/_ref by tests/test_oracle_vs_ref.py). */
unsigned ro_fma_mask = RO_FMA_REFERENCE;
void ro_set_fma_mask(unsigned mask) { ro_fma_mask = mask; }
unsigned ro_get_fma_mask(void) { return ro_fma_mask; }

#define SITE(bit) ((ro_fma_mask & (bit)) != 0)

static float madd_f(unsigned site, float a, float b, float c) {
    if (ro_fma_mask & site) return fmaf(a, b, c);
    float p = a * b;
    return p + c;
}

static double madd_d(unsigned site, double a, double b, double c) {
    if (ro_fma_mask & site) return fma(a, b, c);
    double p = a * b;
    return p + c;
}

/* _mm512_reduce_add_ps tree: halves 8, then 4, then 2, then 1. */
static float reduce16(const float *v) {
    float t8[8], t4[4], t2[2];
    for (int i = 0; i < 8; ++i) t8[i] = v[8 + i] + v[i];
    for (int i = 0; i < 4; ++i) t4[i] = t8[4 + i] + t8[i];
    for (int i = 0; i < 2; ++i) t2[i] = t4[i] + t4[2 + i];
    return t2[0] + t2[1];
}

/* ---------------------------------------------------------------- mt19937 */
/* std::mt19937 / std::mt19937_64 as specified by [rand.eng.mers]; the
 * reference draws rotation signs from the 32-bit engine (space_rabitq.h:
 * 2900-2906) and the INT8 dither from the 64-bit one (:3559-3560). */
typedef struct { uint32_t mt[624]; int idx; } mt32;
static void mt32_seed(mt32 *g, uint32_t seed) {
    g->mt[0] = seed;
    for (int i = 1; i < 624; ++i)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->idx = 624;
}
static uint32_t mt32_next(mt32 *g) {
    if (g->idx >= 624) {
        for (int i = 0; i < 624; ++i) {
            uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
            uint32_t v = g->mt[(i + 397) % 624] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908b0dfu;
            g->mt[i] = v;
        }
        g->idx = 0;
    }
    uint32_t y = g->mt[g->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

typedef struct { uint64_t mt[312]; int idx; } mt64;
static void mt64_seed(mt64 *g, uint64_t seed) {
    g->mt[0] = seed;
    for (int i = 1; i < 312; ++i)
        g->mt[i] = 6364136223846793005ull * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
    g->idx = 312;
}
static uint64_t mt64_next(mt64 *g) {
    if (g->idx >= 312) {
        for (int i = 0; i < 312; ++i) {
            uint64_t y = (g->mt[i] & 0xFFFFFFFF80000000ull) | (g->mt[(i + 1) % 312] & 0x7FFFFFFFull);
            uint64_t v = g->mt[(i + 156) % 312] ^ (y >> 1);
            if (y & 1ull) v ^= 0xB5026F5AA96619E9ull;
            g->mt[i] = v;
        }
        g->idx = 0;
    }
    uint64_t x = g->mt[g->idx++];
    x ^= (x >> 29) & 0x5555555555555555ull;
    x ^= (x << 17) & 0x71D67FFFEDA60000ull;
    x ^= (x << 37) & 0xFFF7EEE000000000ull;
    x ^= x >> 43;
    return x;
}
/* uniform_real_distribution<double>(0,1) over a 64-bit engine: one draw,
 * converted to double and scaled by 2^-64, clamped below 1. */
static double u01_from(uint64_t x) {
    double r = (double)x * 0x1p-64;
    if (r >= 1.0) r = nextafter(1.0, 0.0);
    return r;
}

/* n-th output (1-based) of each engine from a seed: the standard's check
 * values are mt19937(5489)[10000] = 4123659995 and mt19937_64(5489)[10000]
 * = 9981545732273789042 ([rand.predef]). */
uint32_t ro_mt32_nth(uint32_t seed, size_t n) {
    mt32 g;
    uint32_t v = 0;
    mt32_seed(&g, seed);
    for (size_t i = 0; i < n; ++i) v = mt32_next(&g);
    return v;
}
uint64_t ro_mt64_nth(uint64_t seed, size_t n) {
    mt64 g;
    uint64_t v = 0;
    mt64_seed(&g, seed);
    for (size_t i = 0; i < n; ++i) v = mt64_next(&g);
    return v;
}

/* ------------------------------------------------------------------ space */
size_t ro_code_dim(size_t dim) { /* roundUp64, space_rabitq.h:217-224 */
    size_t rounded = ((dim + 63u) / 64u) * 64u;
    size_t p = 1;
    while (p < rounded) p <<= 1u;
    return p;
}

int ro_space_init(ro_space *s, size_t dim, uint32_t seed, size_t block, int mse, int fp16) {
    memset(s, 0, sizeof(*s));
    if (dim == 0 || block == 0) return -1;
    s->dim = dim;
    s->code_dim = ro_code_dim(dim);
    s->block_size = block;
    s->block_count = (s->code_dim + block - 1) / block;
    s->scale_bytes = s->block_count * (fp16 ? 2u : 4u);
    s->compact_code_bytes = (s->code_dim + 1) / 2;
    s->residual_code_bytes = (s->code_dim * 4 + 7) / 8;
    s->compact_bytes = 8 + 8 + s->compact_code_bytes + 1;
    s->residual_bytes = 8 + s->scale_bytes + s->residual_code_bytes;
    s->full_bytes = 8 + 8 + 8 + s->scale_bytes + s->compact_code_bytes + 1 + s->residual_code_bytes;
    s->seed = seed;
    s->mse = mse;
    s->fp16 = fp16;
    s->inv_sqrt_code_dim = 1.0f / sqrtf((float)s->code_dim);
    s->signs = (float *)calloc(s->code_dim, sizeof(float));
    s->center = (float *)calloc(dim, sizeof(float));
    s->rotated_center = (float *)calloc(s->code_dim, sizeof(float));
    /* setRandomRotation, space_rabitq.h:2900-2906: uniform_int_distribution
     * <int>(0,1) over the 32-bit engine keeps the top bit of each draw. */
    if (seed != 0) {
        mt32 g;
        mt32_seed(&g, seed);
        for (size_t i = 0; i < s->code_dim; ++i)
            s->signs[i] = ((mt32_next(&g) >> 31) ? 1.0f : -1.0f) * s->inv_sqrt_code_dim;
    } else {
        for (size_t i = 0; i < s->code_dim; ++i) s->signs[i] = s->inv_sqrt_code_dim;
    }
    return 0;
}

void ro_space_free(ro_space *s) {
    free(s->signs);
    free(s->center);
    free(s->rotated_center);
    memset(s, 0, sizeof(*s));
}

/* hadamard + rotate, space_rabitq.h:944-1000. */
void ro_rotate(const ro_space *s, const float *x, float *out) {
    const size_t cd = s->code_dim;
    memset(out, 0, cd * sizeof(float));
    if (s->seed == 0) {
        memcpy(out, x, s->dim * sizeof(float));
        return;
    }
    for (size_t i = 0; i < s->dim; ++i) out[i] = x[i] * s->signs[i];
    for (size_t step = 1; step < cd; step <<= 1u) {
        for (size_t blk = 0; blk < cd; blk += step << 1u) {
            for (size_t i = 0; i < step; ++i) {
                float a = out[blk + i], b = out[blk + step + i];
                out[blk + i] = a + b;
                out[blk + step + i] = a - b;
            }
        }
    }
}

/* setCentroids with K = 1, space_rabitq.h:2918-2936. */
void ro_space_set_center(ro_space *s, const float *center) {
    memcpy(s->center, center, s->dim * sizeof(float));
    double norm = 0.0;
    for (size_t i = 0; i < s->dim; ++i) norm += (double)center[i] * (double)center[i];
    s->center_norm_sqr = (float)norm;
    ro_rotate(s, s->center, s->rotated_center);
}

/* ------------------------------------------------------- fastQuantizeAbs */
typedef struct { double t; size_t i; } ro_event;
static int ev_less(ro_event a, ro_event b) { return a.t < b.t || (a.t == b.t && a.i < b.i); }
static void heap_push(ro_event *h, size_t *n, ro_event e) {
    size_t c = (*n)++;
    h[c] = e;
    while (c > 0) {
        size_t p = (c - 1) / 2;
        if (!ev_less(h[c], h[p])) break;
        ro_event t = h[c]; h[c] = h[p]; h[p] = t;
        c = p;
    }
}
static ro_event heap_pop(ro_event *h, size_t *n) {
    ro_event top = h[0];
    h[0] = h[--(*n)];
    size_t c = 0;
    for (;;) {
        size_t l = 2 * c + 1, r = l + 1, m = c;
        if (l < *n && ev_less(h[l], h[m])) m = l;
        if (r < *n && ev_less(h[r], h[m])) m = r;
        if (m == c) break;
        ro_event t = h[c]; h[c] = h[m]; h[m] = t;
        c = m;
    }
    return top;
}

/* fastQuantizeAbs, space_rabitq.h:867-942. The pop order of the reference's
 * std::priority_queue over (t, i) pairs is the (t, i) total order, which any
 * binary min-heap on that key reproduces. */
void ro_fast_quantize_abs(size_t cd, const float *a, uint8_t *code, float *ip_norm) {
    const double eps = 1e-5;
    double max_o = 0.0;
    for (size_t i = 0; i < cd; ++i) if ((double)a[i] > max_o) max_o = (double)a[i];
    if (max_o <= eps) {
        memset(code, 0, cd);
        *ip_norm = 1.0f;
        return;
    }
    const double t_start = 2.0 / max_o;          /* kRemainingMax / 3 = 2 */
    const double t_end = (7.0 + 10.0) / max_o;   /* (kRemainingMax + n_enum) */
    int *cur = (int *)calloc(cd, sizeof(int));
    ro_event *heap = (ro_event *)malloc(cd * sizeof(ro_event));
    size_t hn = 0;
    double sqr_den = (double)cd * 0.25;
    double num = 0.0;
    for (size_t i = 0; i < cd; ++i) {
        int c = (int)madd_d(RO_FMA_FQA_T, t_start, (double)a[i], eps);
        if (c > 7) c = 7;
        cur[i] = c;
        sqr_den += (double)(c * c + c);
        num += ((double)c + 0.5) * (double)a[i];
    }
    for (size_t i = 0; i < cd; ++i) {
        if ((double)a[i] > eps) {
            ro_event e = {(double)(cur[i] + 1) / (double)a[i], i};
            heap_push(heap, &hn, e);
        }
    }
    double best_t = 0.0, max_ip = 0.0;
    while (hn > 0) {
        ro_event e = heap_pop(heap, &hn);
        int uc = ++cur[e.i];
        sqr_den += 2.0 * (double)uc;
        num += (double)a[e.i];
        double ip = num / sqrt(sqr_den);
        if (ip > max_ip) { max_ip = ip; best_t = e.t; }
        if (uc < 7) {
            double ct = (double)(uc + 1) / (double)a[e.i];
            if (ct < t_end) { ro_event n = {ct, e.i}; heap_push(heap, &hn, n); }
        }
    }
    num = 0.0;
    for (size_t i = 0; i < cd; ++i) {
        int v = (int)madd_d(RO_FMA_FQA_T, best_t, (double)a[i], eps);
        if (v > 7) v = 7;
        code[i] = (uint8_t)v;
        num += ((double)v + 0.5) * (double)a[i];
    }
    float ipn = num > eps ? (float)(1.0 / num) : 1.0f;
    if (!isfinite(ipn)) ipn = 1.0f;
    *ip_norm = ipn;
    free(cur);
    free(heap);
}

/* floatToHalfScalar / halfToFloatScalar, space_rabitq.h:392-438: truncating
 * add-0x1000 rounding, not IEEE round-to-nearest-even. */
uint16_t ro_float_to_half(float value) {
    uint32_t bits;
    memcpy(&bits, &value, 4);
    uint32_t sign = (bits >> 16) & 0x8000u;
    uint32_t mant = bits & 0x007FFFFFu;
    int32_t e = (int32_t)((bits >> 23) & 0xFFu) - 127 + 15;
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        mant = (mant | 0x00800000u) >> (uint32_t)(1 - e);
        return (uint16_t)(sign | ((mant + 0x00001000u) >> 13));
    }
    if (e >= 31) return (uint16_t)(sign | 0x7C00u);
    return (uint16_t)(sign | ((uint32_t)e << 10) | ((mant + 0x00001000u) >> 13));
}

float ro_half_to_float(uint16_t v) {
    uint32_t sign = (uint32_t)(v & 0x8000u) << 16;
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
    float f;
    memcpy(&f, &bits, 4);
    return f;
}

static void store_scale(const ro_space *s, uint8_t *dst, size_t blk, float v) {
    if (s->fp16) { uint16_t h = ro_float_to_half(v); memcpy(dst + 2 * blk, &h, 2); }
    else memcpy(dst + 4 * blk, &v, 4);
}
static float load_scale(const ro_space *s, const uint8_t *src, size_t blk) {
    if (s->fp16) { uint16_t h; memcpy(&h, src + 2 * blk, 2); return ro_half_to_float(h); }
    float v; memcpy(&v, src + 4 * blk, 4); return v;
}

/* residualMaxAbsScale / residualMseOptimalScale, space_rabitq.h:729-786
 * (4-bit: qmax = 7). */
static float max_abs_scale(const float *e, size_t b, size_t n) {
    float m = 0.0f;
    for (size_t i = b; i < n; ++i) { float v = fabsf(e[i]); if (v > m) m = v; }
    if (m <= 0.0f) return 0.0f;
    return m / 7.0f;
}
static int32_t clamp7(double v) {
    int32_t q = (int32_t)round(v);
    if (q > 7) q = 7;
    if (q < -7) q = -7;
    return q;
}
static float mse_scale(const float *e, size_t b, size_t n) {
    float scale = max_abs_scale(e, b, n);
    if (scale == 0.0f || !isfinite(scale)) return scale;
    for (int it = 0; it < 8; ++it) {
        double num = 0.0, den = 0.0;
        for (size_t i = b; i < n; ++i) {
            int32_t q = clamp7((double)e[i] / (double)scale);
            num += (double)e[i] * (double)q;
            den += (double)q * (double)q;
        }
        if (den <= 0.0) break;
        float next = (float)(num / den);
        if (next <= 0.0f || !isfinite(next)) break;
        float lim = 1e-6f * (scale > 1.0f ? scale : 1.0f);
        if (fabsf(next - scale) <= lim) { scale = next; break; }
        scale = next;
    }
    return scale;
}

static void set_nibble(uint8_t *code, size_t i, uint8_t v) {
    v &= 0x0Fu;
    if (i & 1u) code[i >> 1] = (uint8_t)((code[i >> 1] & 0x0Fu) | (v << 4));
    else code[i >> 1] = (uint8_t)((code[i >> 1] & 0xF0u) | v);
}
static uint8_t get_nibble(const uint8_t *code, size_t i) {
    uint8_t b = code[i >> 1];
    return (uint8_t)((i & 1u) ? (b >> 4) : (b & 0x0Fu));
}

/* encodeVectorFullWithCentroid, space_rabitq.h:5845-5999 (K = 1, sequential
 * nibble layout). Full record: header{norm_sqr, long_scale} | short factors
 * {short_scale, error_scale} | residual factors | scales | codes | cid |
 * residual codes (:261-266, 2857-2861). */
void ro_encode_full(const ro_space *s, const float *x, uint8_t *out) {
    const size_t D = s->dim, cd = s->code_dim;
    float *res = (float *)calloc(D, sizeof(float));
    /* The reference compiler vectorises this loop: products of the 16- and
     * 8-wide blocks are rounded and then added in index order; only the
     * scalar tail (D mod 8 elements) is contracted into FMAs. */
    float rn2 = 0.0f;
    const size_t vec_end = D >= 16 ? (D / 16) * 16 + ((D % 16) >= 8 ? 8 : 0) : 0;
    for (size_t i = 0; i < D; ++i) {
        res[i] = x[i] - s->center[i];
        if (i < vec_end && !SITE(RO_FMA_ENC_NORM)) { const float p = res[i] * res[i]; rn2 = rn2 + p; }
        else rn2 = fmaf(res[i], res[i], rn2);
    }
    const float rn = sqrtf(rn2);
    if (rn > 0.0f) {
        const float inv = 1.0f / rn;
        for (size_t i = 0; i < D; ++i) res[i] *= inv;
    }
    float *rot = (float *)malloc(cd * sizeof(float));
    ro_rotate(s, res, rot);
    ro_encode_from_unit(s, rn, rot, 0, out);
    free(res); free(rot);
}

/* Compensated FP32 sums of the precision mode: Knuth's TwoSum and the fma
 * TwoProduct (Ogita, Rump and Oishi's Sum2 / Dot2), in the kernels' order. */
typedef struct { float s, c; } comp_t;
static void comp_add(comp_t *a, float p) {
    const float t = a->s + p;
    const float bp = t - a->s;
    a->c = a->c + ((a->s - (t - bp)) + (p - bp));
    a->s = t;
}
static void comp_add_sq(comp_t *a, float r) {
    const float p = r * r;
    const float e = fmaf(r, r, -p);
    comp_add(a, p);
    a->c = a->c + e;
}
static void comp_merge(comp_t *a, comp_t b) {
    comp_add(a, b.s);
    a->c = a->c + b.c;
}

/* Precision-mode encode of x given y = Q (x - c), the rotation GEMM's
 * output: |x - c| by a compensated sum in index order; the unit vector is y
 * divided by its compensated norm, summed as the encode workgroup does (256
 * strided partial sums, then a pairwise tree over them); the short and error
 * factors from a compensated sum of |o| in FP32; the codec as before. */
void ro_encode_full_rotated(const ro_space *s, const float *x, const float *y, uint8_t *out) {
    const size_t D = s->dim, cd = s->code_dim;
    comp_t rc = {0.0f, 0.0f};
    for (size_t i = 0; i < D; ++i) comp_add_sq(&rc, x[i] - s->center[i]);
    const float rn = sqrtf(rc.s + rc.c);
    comp_t part[256];
    for (size_t t = 0; t < 256; ++t) {
        part[t].s = 0.0f; part[t].c = 0.0f;
        for (size_t i = t; i < cd; i += 256) comp_add_sq(&part[t], y[i]);
    }
    for (size_t h = 128; h > 0; h >>= 1)
        for (size_t t = 0; t < h; ++t) comp_merge(&part[t], part[t + h]);
    const float n2 = part[0].s + part[0].c;
    const float inv = n2 > 0.0f ? 1.0f / sqrtf(n2) : 0.0f;
    float *rot = (float *)malloc(cd * sizeof(float));
    for (size_t i = 0; i < cd; ++i) rot[i] = y[i] * inv;
    ro_encode_from_unit(s, rn, rot, 1, out);
    free(rot);
}

/* Everything after the rotation: fastQuantizeAbs, nibble codes, scales,
 * residual codes and factors of the rotated unit vector rot of a residual
 * of norm rn; precision = 1 takes the short and error factors from a
 * compensated FP32 sum of |rot| instead of the reference's FP64 sums. */
void ro_encode_from_unit(const ro_space *s, float rn, const float *rot, int precision, uint8_t *out) {
    const size_t cd = s->code_dim;
    memset(out, 0, s->full_bytes);
    uint8_t *scales = out + 24;
    uint8_t *code = scales + s->scale_bytes;
    uint8_t *cid = code + s->compact_code_bytes;
    uint8_t *rcode = cid + 1;
    float norm_sqr = rn * rn;
    float long_scale = 0.0f, short_scale = 0.0f, error_scale = 0.0f;

    float *a = (float *)malloc(cd * sizeof(float));
    uint8_t *m = (uint8_t *)malloc(cd);
    for (size_t i = 0; i < cd; ++i) a[i] = fabsf(rot[i]);
    float ip_norm = 1.0f;
    ro_fast_quantize_abs(cd, a, m, &ip_norm);
    long_scale = 2.0f * rn * ip_norm;

    double o_obar = 0.0, half_l1 = 0.0;
    const double inv_sqrt_d = 1.0 / sqrt((double)cd);
    for (size_t i = 0; i < cd; ++i) {
        const int pos = rot[i] > 0.0f;
        set_nibble(code, i, pos ? (uint8_t)(8 + m[i]) : (uint8_t)(7 - m[i]));
        const double sign = pos ? 1.0 : -1.0;
        o_obar = madd_d(RO_FMA_ENC_OOBAR, (double)rot[i] * sign, inv_sqrt_d, o_obar);
        half_l1 += 0.5 * fabs((double)rot[i]);
    }
    if (precision) {
        comp_t l1c = {0.0f, 0.0f};
        for (size_t i = 0; i < cd; ++i) comp_add(&l1c, fabsf(rot[i]));
        const float l1 = l1c.s + l1c.c, half = 0.5f * l1;
        if (rn > 0.0f && half > 1e-12f) {
            short_scale = 2.0f * rn / half;
            const float sq = sqrtf((float)cd);
            float o = l1 / sq;
            if (!isfinite(o)) o = 0.8f;
            if (o < 1e-6f) o = 1e-6f;
            if (o > 0.999999f) o = 0.999999f;
            const float o2 = o * o;
            float r = (1.0f - o2) / o2;
            if (r < 0.0f) r = 0.0f;
            error_scale = sqrtf(r) * (0.359375f / sq) * 2.0f * rn;
        }
    } else if (rn > 0.0f && half_l1 > 1e-12) {
        short_scale = (float)(2.0 * (double)rn / half_l1);
        if (!isfinite(o_obar)) o_obar = 0.8;
        if (o_obar < 1e-6) o_obar = 1e-6;
        if (o_obar > 0.999999) o_obar = 0.999999;
        const double o2 = o_obar * o_obar;
        const double fac = ldexp(5.75, -4) / sqrt((double)cd);
        double r = (1.0 - o2) / o2;
        if (r < 0.0) r = 0.0;
        error_scale = (float)(sqrt(r) * fac * 2.0 * (double)rn);
    }

    float *qerr = (float *)malloc(cd * sizeof(float));
    for (size_t i = 0; i < cd; ++i) {
        const double tv = (double)rn * (double)rot[i];
        const double x4 = 0.5 * (double)long_scale * ((double)get_nibble(code, i) - 7.5);
        qerr[i] = (float)(tv - x4);
    }
    double dec_norm = 0.0, x4_ip = 0.0;
    int32_t *rc = (int32_t *)calloc(cd, sizeof(int32_t));
    for (size_t blk = 0; blk < s->block_count; ++blk) {
        const size_t b = blk * s->block_size;
        const size_t e = (b + s->block_size < cd) ? b + s->block_size : cd;
        const float sc = s->mse ? mse_scale(qerr, b, e) : max_abs_scale(qerr, b, e);
        store_scale(s, scales, blk, sc);
        const float st = load_scale(s, scales, blk);
        if (st == 0.0f || !isfinite(st)) continue;
        for (size_t i = b; i < e; ++i) {
            const int32_t q = clamp7((double)qerr[i] / (double)st);
            rc[i] = q;
            const double dr = (double)st * (double)q;
            const double x4 = 0.5 * (double)long_scale * ((double)get_nibble(code, i) - 7.5);
            dec_norm = madd_d(RO_FMA_ENC_DECNORM, dr, dr, dec_norm);
            x4_ip = madd_d(RO_FMA_ENC_X4IP, x4, dr, x4_ip);
        }
    }
    /* pack_residual_codes 4-bit, space_rabitq.h:580-595 */
    for (size_t i = 0; i < cd; ++i) set_nibble(rcode, i, (uint8_t)((uint32_t)rc[i] & 0x0Fu));
    float rf[2] = {(float)dec_norm, (float)x4_ip};
    float hdr[4] = {norm_sqr, long_scale, short_scale, error_scale};
    memcpy(out, hdr, 16);
    memcpy(out + 16, rf, 8);
    *cid = 0;
    free(a); free(m); free(qerr); free(rc);
}

/* copyCompactPayloadFromFull / copyResidualRecordFromFull, :6005-6029. */
void ro_copy_compact(const ro_space *s, const uint8_t *full, uint8_t *out) {
    memcpy(out, full, 16);
    memcpy(out + 16, full + 24 + s->scale_bytes, s->compact_code_bytes);
    out[16 + s->compact_code_bytes] = full[24 + s->scale_bytes + s->compact_code_bytes];
}
void ro_copy_residual(const ro_space *s, const uint8_t *full, uint8_t *out) {
    memcpy(out, full + 16, 8);
    memcpy(out + 8, full + 24, s->scale_bytes);
    memcpy(out + 8 + s->scale_bytes, full + 24 + s->scale_bytes + s->compact_code_bytes + 1,
           s->residual_code_bytes);
}

/* extract_paper_prune_sidecar, space_rabitq.h:4364-4392. */
void ro_extract_sidecar(const ro_space *s, const uint8_t *compact, uint8_t *msb, ro_factors *f) {
    const size_t cd = s->code_dim;
    const uint8_t *code = compact + 16;
    memset(msb, 0, (cd + 7) / 8);
    for (size_t i = 0; i < cd; ++i)
        if (get_nibble(code, i) >> 3) msb[i >> 3] |= (uint8_t)(1u << (i & 7u));
    float hdr[4];
    memcpy(hdr, compact, 16);
    memset(f, 0, sizeof(*f));
    f->norm_sqr = hdr[0];
    if (cd <= 1 || hdr[0] <= 0.0f || hdr[2] <= 0.0f) return;
    const float dn = sqrtf(hdr[0]);
    f->data_norm = dn;
    float alpha = 4.0f * dn / (hdr[2] * sqrtf((float)cd));
    if (alpha > 1.0f) alpha = 1.0f;
    if (!(alpha > 0.0f) || !isfinite(alpha)) return;
    const float a2 = alpha * alpha;
    f->cross_scale = hdr[2];
    float r = (1.0f - a2) / a2;
    if (r < 0.0f) r = 0.0f;
    f->error_cross_scale = 2.0f * dn * sqrtf(r) / sqrtf((float)(cd - 1));
    f->valid = isfinite(f->error_cross_scale) ? 1 : 0;
}

/* ------------------------------------------------------------------ query */
void ro_query_alloc(const ro_space *s, ro_query *q) {
    memset(q, 0, sizeof(*q));
    q->q = (float *)calloc(s->code_dim, sizeof(float));
    q->x8 = (int8_t *)calloc(s->code_dim, 1);
}
void ro_query_free(ro_query *q) { free(q->q); free(q->x8); memset(q, 0, sizeof(*q)); }

/* initializePreparedQuery + buildCentroidQuery, space_rabitq.h:3429-3698
 * (K = 1; Full or Int8 codec). */
void ro_prepare_query(const ro_space *s, const float *raw, int codec, ro_query *q) {
    float *rq = (float *)malloc(s->code_dim * sizeof(float));
    ro_rotate(s, raw, rq);
    ro_prepare_query_rotated(s, raw, rq, s->rotated_center, codec, q);
    free(rq);
}

/* Query preparation from the rotated raw query rq and rotated center rc
 * (the precision mode passes the rotation GEMM's outputs). */
void ro_prepare_query_rotated(const ro_space *s, const float *raw, const float *rq, const float *rc, int codec,
                              ro_query *q) {
    const size_t D = s->dim, cd = s->code_dim;
    double raw_n2 = 0.0;
    for (size_t i = 0; i < D; ++i) raw_n2 += (double)raw[i] * (double)raw[i];
    for (size_t i = 0; i < cd; ++i) q->q[i] = rq[i] - rc[i];
    double dot = 0.0;
    for (size_t i = 0; i < D; ++i) dot += (double)raw[i] * (double)s->center[i];
    float qn2 = (float)(raw_n2 + (double)s->center_norm_sqr - 2.0 * dot);
    if (qn2 < 0.0f) qn2 = 0.0f;
    q->query_norm_sqr = qn2;
    float half = 0.0f, sum = 0.0f, pos = 0.0f, mx = 0.0f;
    for (size_t p = 0; p < cd / 2; ++p) {
        const float ev = q->q[2 * p], od = q->q[2 * p + 1];
        half += ev + od;
        sum += ev + od;
        pos += (ev > 0.0f ? ev : 0.0f) + (od > 0.0f ? od : 0.0f);
        const float ae = fabsf(ev), ao = fabsf(od);
        const float am = ae > ao ? ae : ao;
        if (am > mx) mx = am;
    }
    q->half_sum = half * 0.5f;
    q->sum = sum;
    q->positive_sum = pos;
    q->query_norm = sqrtf(q->query_norm_sqr);
    q->int8_scale = 1.0f;
    q->int8_error_norm = 0.0f;
    q->int8_code_sum = 0;
    q->b1_short_scale = q->b1_error_scale = q->b1_alpha = 0.0f;
    memset(q->x8, 0, cd);
    if (codec == 1 || codec == 4) {
        /* the sign code and its scale factors (:3493-3524): |q| summed in
         * float in index order */
        float abs_sum = 0.0f;
        for (size_t i = 0; i < cd; ++i) {
            const float r = q->q[i];
            abs_sum += fabsf(r);
            q->x8[i] = r > 0.0f ? 1 : -1;
            q->int8_code_sum += q->x8[i];
        }
        if (q->query_norm > 0.0f && abs_sum > 0.0f && cd > 1) {
            const float a_q = abs_sum / q->query_norm;
            float alpha_q = a_q / sqrtf((float)cd);
            alpha_q = (1.0f < alpha_q) ? 1.0f : alpha_q;   /* std::min(1.0f, .) */
            const float alpha2 = alpha_q * alpha_q;
            q->b1_alpha = alpha_q;
            q->b1_short_scale = 4.0f * q->query_norm / a_q;
            float r = (1.0f - alpha2) / alpha2;
            r = (0.0f < r) ? r : 0.0f;                      /* std::max(0.0f, .) */
            q->b1_error_scale = 2.0f * q->query_norm * sqrtf(r) / sqrtf((float)(cd - 1));
        }
    }
    if ((codec == 3 || codec == 2) && mx > 0.0f) {
        const int int8 = codec == 3;
        q->int8_scale = int8 ? mx / 128.0f : mx / 8.0f;
        /* FNV-1a over the raw query bytes, mixed with the centroid id
         * (:3549-3558), seeds mt19937_64; two uniform draws per dimension,
         * the first for INT8, the second for INT4 (:3585-3590, 3628-3652). */
        uint64_t seed = 1469598103934665603ull;
        const uint8_t *bytes = (const uint8_t *)raw;
        for (size_t i = 0; i < D * 4; ++i) { seed ^= bytes[i]; seed *= 1099511628211ull; }
        seed ^= 0ull + 0x9E3779B97F4A7C15ull + (seed << 6) + (seed >> 2);
        mt64 g;
        mt64_seed(&g, seed);
        double err = 0.0;
        const int32_t lo = int8 ? -128 : -8, hi = int8 ? 127 : 7;
        for (size_t i = 0; i < cd; ++i) {
            const double u8 = u01_from(mt64_next(&g));
            const double u4 = u01_from(mt64_next(&g));
            int32_t x = (int32_t)floor((double)q->q[i] / (double)q->int8_scale + (int8 ? u8 : u4));
            if (x > hi) x = hi;
            if (x < lo) x = lo;
            const double e = (double)q->q[i] - (double)q->int8_scale * (double)x;
            err += e * e;
            q->x8[i] = (int8_t)x;
            q->int8_code_sum += x;
        }
        q->int8_error_norm = (float)sqrt(err);
    }
}

/* ------------------------------------------------------------- estimators */
static int long_scale_ok(float ls) { return ls > 0.0f && isfinite(ls); }

/* queryDistanceLongWithIps, space_rabitq.h:1621-1632. */
static float finish_long(const ro_query *q, float norm_sqr, float ls, float short_ip, float rem_ip) {
    const float t = 8.0f * short_ip + rem_ip;
    const float signed_ip = madd_f(RO_FMA_C1_SUB7, -7.0f, q->half_sum, t);
    const float base = norm_sqr + q->query_norm_sqr;
    if (SITE(RO_FMA_C1_EST)) return fmaf(-ls, signed_ip, base);
    const float est = ls * signed_ip;
    return base - est;
}

/* queryDistanceLong + longCodeIpsAvx512 (space_rabitq.h:1551-1607): lane j
 * of the 16-wide accumulators takes pairs j, j+16, j+32, ... */
float ro_asym_distance(const ro_space *s, const ro_query *q, const uint8_t *compact) {
    float hdr[2];
    memcpy(hdr, compact, 8);
    if (!long_scale_ok(hdr[1])) return hdr[0] + q->query_norm_sqr;
    const uint8_t *code = compact + 16;
    const size_t pairs = s->code_dim / 2;
    float sh[16] = {0}, rm[16] = {0};
    for (size_t p0 = 0; p0 < pairs; p0 += 16) {
        for (int j = 0; j < 16; ++j) {
            const uint8_t b = code[p0 + j];
            const uint8_t lo = b & 0x0Fu, hi = b >> 4;
            const float ev = q->q[2 * (p0 + j)], od = q->q[2 * (p0 + j) + 1];
            rm[j] = madd_f(RO_FMA_C1_LANE, (float)(lo & 7u), ev, rm[j]);
            rm[j] = madd_f(RO_FMA_C1_LANE, (float)(hi & 7u), od, rm[j]);
            if (lo > 7) sh[j] = sh[j] + ev;
            if (hi > 7) sh[j] = sh[j] + od;
        }
    }
    const float short_ip = reduce16(sh) - q->half_sum;
    const float rem_ip = reduce16(rm);
    return finish_long(q, hdr[0], hdr[1], short_ip, rem_ip);
}

/* FP64 restatement of the same estimate: ||o||^2 + ||q||^2 - 2<x4, q>,
 * with x4_i = 0.5 * long_scale * (c_i - 7.5) (space_rabitq.h:5896, 1621). */
double ro_asym_distance_f64(const ro_space *s, const ro_query *q, const uint8_t *compact) {
    float hdr[2];
    memcpy(hdr, compact, 8);
    if (!long_scale_ok(hdr[1])) return (double)hdr[0] + (double)q->query_norm_sqr;
    const uint8_t *code = compact + 16;
    double ip = 0.0;
    for (size_t i = 0; i < s->code_dim; ++i)
        ip += ((double)get_nibble(code, i) - 7.5) * (double)q->q[i];
    return (double)hdr[0] + (double)q->query_norm_sqr - (double)hdr[1] * ip;
}

/* query_distance_with_paper_msb_batch_by_id (:3972-4073): remaining 3 bits
 * in 16 lanes, MSB part supplied by the gate's short_ip. */
float ro_remaining_distance_fp32(const ro_space *s, const ro_query *q, const uint8_t *compact, float short_ip) {
    float hdr[2];
    memcpy(hdr, compact, 8);
    if (!long_scale_ok(hdr[1])) return hdr[0] + q->query_norm_sqr;
    const uint8_t *code = compact + 16;
    const size_t pairs = s->code_dim / 2;
    float rm[16] = {0};
    for (size_t p0 = 0; p0 < pairs; p0 += 16) {
        for (int j = 0; j < 16; ++j) {
            const uint8_t b = code[p0 + j];
            rm[j] = madd_f(RO_FMA_C1_LANE, (float)(b & 7u), q->q[2 * (p0 + j)], rm[j]);
            rm[j] = madd_f(RO_FMA_C1_LANE, (float)((b >> 4) & 7u), q->q[2 * (p0 + j) + 1], rm[j]);
        }
    }
    return finish_long(q, hdr[0], hdr[1], short_ip, reduce16(rm));
}

/* finish_paper_prune_estimate_sidecar, space_rabitq.h:4701-4721. */
static ro_estimate finish_gate(float short_ip, const ro_factors *f, const ro_query *q, float eps0) {
    ro_estimate r;
    memset(&r, 0, sizeof(r));
    if (!f->valid) return r;
    const float cross = short_ip * f->cross_scale;
    const float max_cross = 2.0f * f->data_norm * q->query_norm;
    const float ee = eps0 * f->error_cross_scale;
    float up = madd_f(RO_FMA_GATE_UPPER, ee, q->query_norm, cross);
    up = (up < max_cross) ? up : max_cross;     /* std::min(max_cross, up) */
    up = (-max_cross < up) ? up : -max_cross;   /* std::max(-max_cross, up) */
    r.lower_bound = f->norm_sqr + q->query_norm_sqr - up;
    r.short_ip = short_ip;
    r.ip_hat = max_cross > 0.0f ? cross / max_cross : 0.0f;
    r.error_bound = max_cross > 0.0f ? eps0 * f->error_cross_scale * q->query_norm / max_cross : 0.0f;
    r.valid = isfinite(r.lower_bound) ? 1 : 0;
    return r;
}

/* compute_full_paper_prune_estimate_batch_by_id (:4427-4508): FP32 query
 * summed over the MSB plane in 16 lanes. */
ro_estimate ro_gate_fp32(const ro_space *s, const ro_query *q, const uint8_t *msb, const ro_factors *f, float eps0) {
    float acc[16] = {0};
    for (size_t d0 = 0; d0 < s->code_dim; d0 += 16) {
        const uint16_t bits = (uint16_t)(msb[d0 >> 3] | ((uint16_t)msb[(d0 >> 3) + 1] << 8));
        for (int j = 0; j < 16; ++j) acc[j] = acc[j] + (((bits >> j) & 1u) ? q->q[d0 + j] : 0.0f);
    }
    return finish_gate(reduce16(acc) - q->half_sum, f, q, eps0);
}

/* maskedQuantizedQuerySum (:5169-5202): exact integer sum. */
int64_t ro_masked_sum(const ro_space *s, const ro_query *q, const uint8_t *msb) {
    int64_t t = 0;
    for (size_t i = 0; i < s->code_dim; ++i)
        if ((msb[i >> 3] >> (i & 7u)) & 1u) t += q->x8[i];
    return t;
}

/* compute_int_paper_prune_estimate_batch_by_id (:4991-5093). */
ro_estimate ro_gate_int8(const ro_space *s, const ro_query *q, const uint8_t *msb, const ro_factors *f, float eps0) {
    ro_estimate r;
    memset(&r, 0, sizeof(r));
    if (!f->valid) return r;
    if (q->query_norm <= 0.0f || f->cross_scale <= 0.0f || !isfinite(f->cross_scale)) return r;
    const int64_t sel = ro_masked_sum(s, q, msb);
    const float short_ip = 0.5f * q->int8_scale * (float)(2 * sel - q->int8_code_sum);
    return finish_gate(short_ip, f, q, eps0);
}

/* compute_b1_paper_prune_estimate_batch_by_id (:4607-4753): mismatches of
 * the query's sign code against the MSB plane, signed_dot = cd - 2
 * mismatches, and finish_b1_paper_prune_estimate (:4724-4753), whose bound
 * adds the query side's error to the database side's. */
ro_estimate ro_gate_b1(const ro_space *s, const ro_query *q, const uint8_t *msb, const ro_factors *f, float eps0) {
    ro_estimate r;
    memset(&r, 0, sizeof(r));
    const int query_valid = q->b1_short_scale > 0.0f && isfinite(q->b1_short_scale) && q->query_norm > 0.0f;
    if (!query_valid || !f->valid || f->cross_scale <= 0.0f || !isfinite(f->cross_scale) || f->data_norm <= 0.0f)
        return r;
    uint64_t mism = 0;
    for (size_t i = 0; i < s->code_dim; ++i) {
        const unsigned qb = q->q[i] > 0.0f ? 1u : 0u;
        mism += qb ^ ((msb[i >> 3] >> (i & 7u)) & 1u);
    }
    const float signed_dot = (float)s->code_dim - 2.0f * (float)mism;
    const float short_ip = signed_dot * q->b1_short_scale / 8.0f;
    const float cross = short_ip * f->cross_scale;
    const float max_cross = 2.0f * f->data_norm * q->query_norm;
    float sum;
    if (SITE(RO_FMA_B1_SUM_DB)) sum = fmaf(f->error_cross_scale, q->query_norm, q->b1_error_scale * f->data_norm);
    else if (SITE(RO_FMA_B1_SUM_Q)) sum = fmaf(q->b1_error_scale, f->data_norm, f->error_cross_scale * q->query_norm);
    else {
        const float db_err = f->error_cross_scale * q->query_norm;
        const float query_err = q->b1_error_scale * f->data_norm;
        sum = db_err + query_err;
    }
    float up = SITE(RO_FMA_B1_UPPER) ? fmaf(eps0, sum, cross) : cross + eps0 * sum;
    up = (up < max_cross) ? up : max_cross;     /* std::min(max_cross, up) */
    up = (-max_cross < up) ? up : -max_cross;   /* std::max(-max_cross, up) */
    r.lower_bound = f->norm_sqr + q->query_norm_sqr - up;
    r.short_ip = short_ip;
    r.alpha = q->b1_alpha;
    r.ip_hat = max_cross > 0.0f ? cross / max_cross : 0.0f;
    r.error_bound = max_cross > 0.0f ? eps0 * sum / max_cross : 0.0f;
    r.valid = isfinite(r.lower_bound) ? 1 : 0;
    return r;
}

/* quantizedCodeDotDispatch (:5099-5155): exact integer dot. */
static int64_t qdot(const ro_space *s, const ro_query *q, const uint8_t *code, uint8_t mask) {
    int64_t t = 0;
    for (size_t i = 0; i < s->code_dim; ++i) t += (int64_t)(get_nibble(code, i) & mask) * (int64_t)q->x8[i];
    return t;
}
int64_t ro_remaining_dot_int8(const ro_space *s, const ro_query *q, const uint8_t *compact) {
    return qdot(s, q, compact + 16, 7u);
}
int64_t ro_full_dot_int8(const ro_space *s, const ro_query *q, const uint8_t *compact) {
    return qdot(s, q, compact + 16, 15u);
}

/* finishQuantizedTraversalDistance (:4342-4358). */
float ro_remaining_distance_int8(const ro_space *s, const ro_query *q, const uint8_t *compact, float short_ip) {
    float hdr[2];
    memcpy(hdr, compact, 8);
    if (!long_scale_ok(hdr[1])) return hdr[0] + q->query_norm_sqr;
    const int64_t rem = ro_remaining_dot_int8(s, q, compact);
    const double centered = (double)q->int8_scale * ((double)rem - 3.5 * (double)q->int8_code_sum);
    const double signed_ip = 8.0 * (double)short_ip + centered;
    return hdr[0] + q->query_norm_sqr - (float)((double)hdr[1] * signed_ip);
}

/* compute_paper_prune_estimate_int_codec + finish_int_paper_prune_estimate
 * (:5204-5260). */
ro_estimate ro_int_codec_estimate(const ro_space *s, const ro_query *q, const uint8_t *compact, float eps0) {
    ro_estimate r;
    memset(&r, 0, sizeof(r));
    float hdr[4];
    memcpy(hdr, compact, 16);
    const float scale = q->int8_scale;
    if (!long_scale_ok(hdr[1]) || !(scale > 0.0f) || !isfinite(scale) || q->query_norm <= 0.0f) {
        r.lower_bound = hdr[0] + q->query_norm_sqr;
        r.valid = isfinite(r.lower_bound) ? 1 : 0;
        return r;
    }
    const int64_t dot = ro_full_dot_int8(s, q, compact);
    const double centered = (double)dot - 7.5 * (double)q->int8_code_sum;
    const double sip = (double)scale * centered;
    const double est_ip = (double)hdr[1] * sip;
    const float est = hdr[0] + q->query_norm_sqr - (float)est_ip;
    const double db_err = (double)hdr[3] * (double)q->query_norm;
    r.lower_bound = est - (float)((double)eps0 * db_err);
    r.short_ip = (float)sip;
    r.error_bound = est - r.lower_bound;
    r.valid = isfinite(r.lower_bound) ? 1 : 0;
    return r;
}

/* The estimate and lower bound of finish_int_paper_prune_estimate
 * (:5238-5252) for the INT8 query x full 4-bit code. */
float ro_int8_full_estimate(const ro_space *s, const ro_query *q, const uint8_t *compact, float eps0,
                            float *lower_out) {
    float hdr[4];
    memcpy(hdr, compact, 16);
    const float scale = q->int8_scale;
    if (!long_scale_ok(hdr[1]) || !(scale > 0.0f) || !isfinite(scale) || q->query_norm <= 0.0f) {
        const float v = hdr[0] + q->query_norm_sqr;
        if (lower_out) *lower_out = v;
        return v;
    }
    const int64_t dot = ro_full_dot_int8(s, q, compact);
    const double centered = (double)dot - 7.5 * (double)q->int8_code_sum;
    const double sip = (double)scale * centered;
    const double est_ip = (double)hdr[1] * sip;
    const float est = hdr[0] + q->query_norm_sqr - (float)est_ip;
    const double db_err = (double)hdr[3] * (double)q->query_norm;
    if (lower_out) *lower_out = est - (float)((double)eps0 * db_err);
    return est;
}

/* centeredNibbleDotTimesFour (:2405-2538): sum (2a-15)(2b-15), exact. */
int64_t ro_symmetric_ip4(const ro_space *s, const uint8_t *l, const uint8_t *r) {
    int64_t t = 0;
    for (size_t i = 0; i < s->code_dim; ++i)
        t += (int64_t)(2 * (int)get_nibble(l, i) - 15) * (int64_t)(2 * (int)get_nibble(r, i) - 15);
    return t;
}

/* distanceBetweenEncoded / symmetricBuildDistanceFromPrepared, K = 1
 * (:2634-2683, 2706-2747); lhs is the prepared side. */
float ro_symmetric_distance(const ro_space *s, const uint8_t *lhs, const uint8_t *rhs) {
    float lh[2], rh[2];
    memcpy(lh, lhs, 8);
    memcpy(rh, rhs, 8);
    const double ln = sqrt((double)lh[0] > 0.0 ? (double)lh[0] : 0.0);
    const double rn = sqrt((double)rh[0] > 0.0 ? (double)rh[0] : 0.0);
    const int lvalid = ln > 0.0 && lh[1] > 0.0f && isfinite(lh[1]);
    const double lipn = lvalid ? (double)lh[1] / (2.0 * ln) : 0.0;
    if (!lvalid || !(rn > 0.0) || !(rh[1] > 0.0f) || !isfinite(rh[1]))
        return (float)((double)lh[0] + (double)rh[0] + 0.0);
    const double ripn = (double)rh[1] / (2.0 * rn);
    const double code_ip = 0.25 * (double)ro_symmetric_ip4(s, lhs + 16, rhs + 16);
    double u = code_ip * lipn * ripn;
    if (!isfinite(u)) u = 0.0;
    if (u < -1.0) u = -1.0;
    if (u > 1.0) u = 1.0;
    const double rip = ln * rn * u;
    return (float)((double)lh[0] + (double)rh[0] - 2.0 * rip);
}

/* compute_residual_distance_intervals_batch_by_id (:5432-5548): per block of
 * 16 dims, 8 lane products (lo*even + hi*odd), sequential lane sum, then
 * ips += scale * block_ip; skipped blocks have scale 0 or non-finite. */
ro_interval ro_residual_interval(const ro_space *s, const ro_query *q, const uint8_t *residual, float long_distance) {
    const uint8_t *scales = residual + 8;
    const uint8_t *code = residual + 8 + s->scale_bytes;
    float ips = 0.0f;
    double errsq = 0.0;
    for (size_t blk = 0; blk < s->block_count; ++blk) {
        const float sc = load_scale(s, scales, blk);
        if (sc == 0.0f || !isfinite(sc)) continue;
        const size_t pair = blk * 8;
        float bip = 0.0f;
        for (int l = 0; l < 8; ++l) {
            const uint8_t b = code[pair + l];
            int lo = b & 0x0F, hi = b >> 4;
            if (lo > 7) lo -= 16;
            if (hi > 7) hi -= 16;
            const float ev = q->q[2 * (pair + l)], od = q->q[2 * (pair + l) + 1];
            float prod;
            if (SITE(RO_FMA_RES_HI)) prod = fmaf((float)hi, od, (float)lo * ev);
            else if (SITE(RO_FMA_RES_LO)) prod = fmaf((float)lo, ev, (float)hi * od);
            else { const float pl = (float)lo * ev, ph = (float)hi * od; prod = pl + ph; }
            bip += prod;
        }
        ips = madd_f(RO_FMA_RES_ACC, sc, bip, ips);
        const double pde = 0.5 * fabs((double)sc);
        errsq += 16.0 * pde * pde;
    }
    const float dist = long_distance - 2.0f * ips;
    const float err = (float)(2.0 * (double)q->query_norm * sqrt(errsq));
    ro_interval r = {dist, dist - err, dist + err};
    return r;
}

/* FP64 form of the residual correction: long - 2 * sum_b s_b <r_b, q_b>. */
double ro_residual_estimate_f64(const ro_space *s, const ro_query *q, const uint8_t *residual, double long_distance) {
    const uint8_t *scales = residual + 8;
    const uint8_t *code = residual + 8 + s->scale_bytes;
    double ips = 0.0;
    for (size_t blk = 0; blk < s->block_count; ++blk) {
        const float sc = load_scale(s, scales, blk);
        if (sc == 0.0f || !isfinite(sc)) continue;
        double bip = 0.0;
        for (size_t i = blk * 16; i < blk * 16 + 16; ++i) {
            int v = get_nibble(code, i);
            if (v > 7) v -= 16;
            bip += (double)v * (double)q->q[i];
        }
        ips += (double)sc * bip;
    }
    return long_distance - 2.0 * ips;
}

/* ----------------------------------------------------------------- search */
typedef struct { float d; uint32_t id; uint8_t expanded; } ro_cand;
static int cand_less(float ad, uint32_t aid, float bd, uint32_t bid) {
    return ad < bd || (ad == bd && aid < bid);
}
/* insert_candidate_sorted, ours_diskann.rs:1100-1126. */
static void pool_insert(ro_cand *pool, size_t *len, size_t L, float d, uint32_t id) {
    if (L == 0) return;
    if (*len >= L && !cand_less(d, id, pool[*len - 1].d, pool[*len - 1].id)) return;
    size_t pos = 0;
    while (pos < *len && cand_less(pool[pos].d, pool[pos].id, d, id)) ++pos;
    size_t n = *len < L ? *len : L - 1;
    memmove(pool + pos + 1, pool + pos, (n - pos) * sizeof(ro_cand));
    pool[pos].d = d;
    pool[pos].id = id;
    pool[pos].expanded = 0;
    if (*len < L) ++*len;
}

/* search_ours_paper_active_inner + flush_distance_batch
 * (ours_diskann.rs:1539-1986) with beam_width 1, start node 0, kth/early
 * stop off; codec 3 = Int8 (gate C5, mode-2 flush) or 0 = Full (gate C4,
 * mode-1 flush). Rerank: top min(|pool|, max(rerank, k)) ids, FP32 long
 * distance recomputed for Int8, residual interval, sort by (estimate, id). */
static void trace_event(ro_search_event *trace, size_t cap, size_t *len, uint32_t hop, uint32_t id, float lb,
                        float back, int stage, int full, int valid, int pruned) {
    if (!trace || *len >= cap) { if (len) ++*len; return; }
    ro_search_event *e = &trace[(*len)++];
    e->hop = hop; e->id = id; e->lower_bound = lb; e->back = back;
    e->stage = (uint8_t)stage; e->full = (uint8_t)full; e->valid = (uint8_t)valid; e->pruned = (uint8_t)pruned;
}

size_t ro_search(const ro_space *s, const ro_query *q, int codec,
                 const uint8_t *cbase, size_t cstride,
                 const uint8_t *rbase, size_t rstride,
                 const uint8_t *mbase, size_t mstride,
                 const ro_factors *factors,
                 const uint32_t *degrees, const uint32_t *edges, size_t estride,
                 size_t n, size_t k, size_t L, float eps0, size_t rerank,
                 uint32_t *out_ids, float *out_d,
                 uint32_t *pool_ids_out, float *pool_d_out, size_t *pool_len_out,
                 ro_search_stats *st) {
    return ro_search_traced(s, q, codec, cbase, cstride, rbase, rstride, mbase, mstride, factors, degrees, edges,
                            estride, n, k, L, eps0, rerank, out_ids, out_d, pool_ids_out, pool_d_out, pool_len_out,
                            st, NULL, 0, NULL);
}

/* trace (may be NULL) receives up to trace_cap prune decisions; *trace_len
 * (may be NULL) counts them all. */
size_t ro_search_traced(const ro_space *s, const ro_query *q, int codec,
                        const uint8_t *cbase, size_t cstride,
                        const uint8_t *rbase, size_t rstride,
                        const uint8_t *mbase, size_t mstride,
                        const ro_factors *factors,
                        const uint32_t *degrees, const uint32_t *edges, size_t estride,
                        size_t n, size_t k, size_t L, float eps0, size_t rerank,
                        uint32_t *out_ids, float *out_d,
                        uint32_t *pool_ids_out, float *pool_d_out, size_t *pool_len_out,
                        ro_search_stats *st, ro_search_event *trace, size_t trace_cap, size_t *trace_len) {
    ro_search_stats local;
    size_t tl_local = 0;
    if (!st) st = &local;
    if (!trace_len) trace_len = &tl_local;
    *trace_len = 0;
    memset(st, 0, sizeof(*st));
    if (k == 0 || n == 0) return 0;
    if (L < k) L = k;
    ro_cand *pool = (ro_cand *)calloc(L + 1, sizeof(ro_cand));
    uint8_t *visited = (uint8_t *)calloc(n, 1);
    size_t len = 0;
    visited[0] = 1;
    pool_insert(pool, &len, L, ro_asym_distance(s, q, cbase), 0);
    st->full_dists++;
    uint32_t fresh[4096];
    ro_estimate est[4096];
    for (;;) {
        size_t f = len;
        for (size_t i = 0; i < len; ++i) if (!pool[i].expanded) { f = i; break; }
        if (f == len) break;
        pool[f].expanded = 1;
        const uint32_t node = pool[f].id;
        st->hops++;
        size_t nf = 0;
        const uint32_t deg = degrees[node];
        for (uint32_t j = 0; j < deg; ++j) {
            const uint32_t nb = edges[(size_t)node * estride + j];
            if (!visited[nb]) { visited[nb] = 1; fresh[nf++] = nb; }
        }
        if (nf == 0) continue;
        for (size_t j = 0; j < nf; ++j) {
            const uint32_t id = fresh[j];
            est[j] = codec == 3
                ? ro_gate_int8(s, q, mbase + (size_t)id * mstride, &factors[id], eps0)
                : ro_gate_fp32(s, q, mbase + (size_t)id * mstride, &factors[id], eps0);
            st->gate_evals++;
        }
        /* gate survivors are batched (at most 64, DIST_BATCH_SIZE) and
         * re-checked against the current pool when the batch is flushed */
        uint32_t bid[64];
        size_t bslot[64], nb = 0;
        for (size_t j = 0; j <= nf; ++j) {
            if (j < nf) {
                const int full = len >= L;
                const float lb = full ? pool[len - 1].d : 3.402823466e38f;
                const int cut = full && est[j].valid && est[j].lower_bound > lb;
                trace_event(trace, trace_cap, trace_len, (uint32_t)st->hops, fresh[j], est[j].lower_bound, lb, 0,
                            full, est[j].valid != 0, cut);
                if (cut) { st->pruned++; continue; }
                bid[nb] = fresh[j];
                bslot[nb] = j;
                nb++;
                if (nb < 64) continue;
            }
            for (size_t b = 0; b < nb; ++b) {
                const uint32_t id = bid[b];
                const ro_estimate *e = &est[bslot[b]];
                const uint8_t *rec = cbase + (size_t)id * cstride;
                float d;
                if (!e->valid) d = ro_asym_distance(s, q, rec);
                else if (codec == 3) d = ro_remaining_distance_int8(s, q, rec, e->short_ip);
                else d = ro_remaining_distance_fp32(s, q, rec, e->short_ip);
                st->full_dists++;
                const int full = len >= L;
                const float lb = full ? pool[len - 1].d : 3.402823466e38f;
                const int cut = full && e->valid && e->lower_bound > lb;
                trace_event(trace, trace_cap, trace_len, (uint32_t)st->hops, id, e->lower_bound, lb, 1, full,
                            e->valid != 0, cut);
                if (full && e->valid) st->recheck_checked++;
                if (cut) { st->pruned++; st->recheck_pruned++; continue; }
                pool_insert(pool, &len, L, d, id);
            }
            nb = 0;
        }
    }
    if (pool_ids_out) for (size_t i = 0; i < len; ++i) { pool_ids_out[i] = pool[i].id; pool_d_out[i] = pool[i].d; }
    if (pool_len_out) *pool_len_out = len;
    size_t rc = len < (rerank > k ? rerank : k) ? len : (rerank > k ? rerank : k);
    float *rd = (float *)malloc((rc + 1) * sizeof(float));
    uint32_t *ri = (uint32_t *)malloc((rc + 1) * sizeof(uint32_t));
    for (size_t i = 0; i < rc; ++i) {
        const uint32_t id = pool[i].id;
        float ld = pool[i].d;
        if (codec == 3) ld = ro_asym_distance(s, q, cbase + (size_t)id * cstride);
        ro_interval iv = ro_residual_interval(s, q, rbase + (size_t)id * rstride, ld);
        rd[i] = iv.estimate;
        ri[i] = id;
    }
    /* sort by (estimate, id) */
    for (size_t i = 1; i < rc; ++i) {
        float d = rd[i]; uint32_t id = ri[i];
        size_t j = i;
        while (j > 0 && cand_less(d, id, rd[j - 1], ri[j - 1])) { rd[j] = rd[j - 1]; ri[j] = ri[j - 1]; --j; }
        rd[j] = d; ri[j] = id;
    }
    size_t out = rc < k ? rc : k;
    for (size_t i = 0; i < out; ++i) { out_ids[i] = ri[i]; out_d[i] = rd[i]; }
    free(rd); free(ri); free(pool); free(visited);
    return out;
}

/* robust_prune_4bit (vamana_prune.h:41-156) with the symmetric distance as
 * distance_between(candidate, selected) (vamana_index.h:844-849). */
static int pc_less(const void *a, const void *b) {
    const ro_cand *x = (const ro_cand *)a, *y = (const ro_cand *)b;
    if (x->d != y->d) return x->d < y->d ? -1 : 1;
    return x->id < y->id ? -1 : (x->id > y->id ? 1 : 0);
}
static size_t robust_prune_counted(const ro_space *s, const uint8_t *cbase, size_t cstride,
                                   uint32_t source, const uint32_t *ids, const float *ds,
                                   size_t n, size_t R, float alpha, uint32_t *out, uint64_t *evals) {
    if (R == 0) return 0;
    ro_cand *c = (ro_cand *)malloc((n + 1) * sizeof(ro_cand));
    for (size_t i = 0; i < n; ++i) { c[i].d = ds[i]; c[i].id = ids[i]; c[i].expanded = 0; }
    qsort(c, n, sizeof(ro_cand), pc_less);
    size_t u = 0;
    for (size_t i = 0; i < n; ++i) {
        if (c[i].id == source) continue;
        int dup = 0;
        for (size_t j = 0; j < u; ++j) if (c[j].id == c[i].id) { dup = 1; break; }
        if (!dup) c[u++] = c[i];
    }
    uint8_t *dis = (uint8_t *)calloc(u + 1, 1);
    float *memo = (float *)malloc((u * R + 1) * sizeof(float));
    for (size_t i = 0; i < u * R; ++i) memo[i] = -1.0f;
    size_t sel = 0;
    float cur = 1.0f;
    const float inc = alpha < 1.2f ? alpha : 1.2f;
    while (sel < R) {
        int promoted = 0;
        for (size_t i = 0; i < u && sel < R; ++i) {
            if (dis[i]) continue;
            int occ = 0;
            for (size_t kk = 0; kk < sel; ++kk) {
                float *m = &memo[i * R + kk];
                if (*m < 0.0f) {
                    *m = ro_symmetric_distance(s, cbase + (size_t)c[i].id * cstride,
                                               cbase + (size_t)out[kk] * cstride);
                    if (evals) ++*evals;
                }
                if (cur * *m <= c[i].d) { occ = 1; break; }
            }
            if (occ) continue;
            out[sel++] = c[i].id;
            dis[i] = 1;
            promoted = 1;
        }
        if (cur == alpha) break;
        float next = cur * inc;
        if (next > alpha) next = alpha;
        if (next == cur) break;
        cur = next;
        if (!promoted && sel >= R) break;
    }
    free(c); free(dis); free(memo);
    return sel;
}

size_t ro_robust_prune(const ro_space *s, const uint8_t *cbase, size_t cstride,
                       uint32_t source, const uint32_t *ids, const float *ds,
                       size_t n, size_t R, float alpha, uint32_t *out, uint64_t *evals) {
    if (evals) *evals = 0;
    return robust_prune_counted(s, cbase, cstride, source, ids, ds, n, R, alpha, out, evals);
}

/* ------------------------------------------------------------------ build */
/* search_vamana (vamana_search.h:97-198) with BuildEvaluator
 * (vamana_index.h:41-69): lhs = the inserted record, rhs = graph nodes.
 * Visited marks are epoch tags; returns the pool length or (size_t)-1 on an
 * out-of-range neighbour. */
static size_t build_search(const ro_space *s, const uint8_t *cbase, size_t cstride, uint32_t qid,
                           const uint32_t *deg, const uint32_t *edges, size_t R, size_t node_count,
                           size_t L, size_t beam, size_t early_stop, uint32_t *tags, uint32_t tag,
                           ro_cand *pool, uint64_t *evals) {
    if (node_count == 0 || L == 0) return 0;
    const uint8_t *qrec = cbase + (size_t)qid * cstride;
    size_t len = 0;
    tags[0] = tag;
    pool[len].d = ro_symmetric_distance(s, qrec, cbase);
    pool[len].id = 0;
    pool[len].expanded = 0;
    ++len;
    ++*evals;
    uint32_t frontier[64];
    size_t stall = 0;
    for (;;) {
        const float back_before = pool[len - 1].d;
        size_t nfr = 0;
        for (size_t i = 0; i < len && nfr < beam; ++i)
            if (!pool[i].expanded) { pool[i].expanded = 1; frontier[nfr++] = pool[i].id; }
        if (nfr == 0) break;
        for (size_t f = 0; f < nfr; ++f) {
            const uint32_t node = frontier[f];
            uint32_t fresh[256];
            size_t nf = 0;
            for (uint32_t j = 0; j < deg[node]; ++j) {
                const uint32_t nb = edges[(size_t)node * R + j];
                if (nb >= node_count) return (size_t)-1;
                if (tags[nb] != tag) { tags[nb] = tag; fresh[nf++] = nb; }
            }
            for (size_t j = 0; j < nf; ++j) {
                const float d = ro_symmetric_distance(s, qrec, cbase + (size_t)fresh[j] * cstride);
                ++*evals;
                pool_insert(pool, &len, L, d, fresh[j]);
            }
        }
        if (early_stop != 0 && len >= L) {
            stall = pool[len - 1].d >= back_before ? stall + 1 : 0;
            if (stall >= early_stop) break;
        }
    }
    return len;
}

typedef struct { uint32_t src, dst; } be_pair;
static int be_less(const void *a, const void *b) {
    const be_pair *x = (const be_pair *)a, *y = (const be_pair *)b;
    if (x->src != y->src) return x->src < y->src ? -1 : 1;
    return x->dst < y->dst ? -1 : (x->dst > y->dst ? 1 : 0);
}

/* addBackedgesGroupedAndPrune + addBackedgeGroupAndPrune
 * (vamana_index.h:995-1117). */
static void add_backedges(const ro_space *s, const uint8_t *cbase, size_t cstride, be_pair *pairs, size_t np,
                          uint32_t *deg, uint32_t *edges, size_t R, float alpha, uint64_t *evals) {
    if (np == 0) return;
    qsort(pairs, np, sizeof(be_pair), be_less);
    uint32_t *nb = (uint32_t *)malloc((R + np + 1) * sizeof(uint32_t));
    float *nd = (float *)malloc((R + np + 1) * sizeof(float));
    for (size_t b = 0; b < np;) {
        size_t e = b;
        const uint32_t src = pairs[b].src;
        while (e < np && pairs[e].src == src) ++e;
        size_t len = deg[src];
        memcpy(nb, edges + (size_t)src * R, len * sizeof(uint32_t));
        size_t added = 0;
        for (size_t i = b; i < e; ++i) {
            const uint32_t t = pairs[i].dst;
            if (t == src || (i > b && pairs[i - 1].dst == t)) continue;
            int present = 0;
            for (size_t j = 0; j < len; ++j) if (nb[j] == t) { present = 1; break; }
            if (!present) { nb[len++] = t; ++added; }
        }
        if (added != 0) {
            if (len <= R) {
                memcpy(edges + (size_t)src * R, nb, len * sizeof(uint32_t));
                deg[src] = (uint32_t)len;
            } else {
                size_t m = 0;
                for (size_t j = 0; j < len; ++j) if (nb[j] != src) nb[m++] = nb[j];
                for (size_t j = 0; j < m; ++j)
                    nd[j] = ro_symmetric_distance(s, cbase + (size_t)src * cstride, cbase + (size_t)nb[j] * cstride);
                *evals += m;
                uint32_t out[256];
                const size_t got = robust_prune_counted(s, cbase, cstride, src, nb, nd, m, R, alpha, out, evals);
                memcpy(edges + (size_t)src * R, out, got * sizeof(uint32_t));
                deg[src] = (uint32_t)got;
            }
        }
        b = e;
    }
    free(nb); free(nd);
}

static int cap_less(const void *a, const void *b) { return pc_less(a, b); }

/* VamanaIndex::buildEncodedSymmetricBulk + refineGraphSymmetric
 * (vamana_index.h:308-422, 452-528), driven as rabitq_build_vamana_graph
 * does (rabitq_bridge.cpp:224-303). Start node 0; batches of inserts from
 * id 1 search the graph as it stood when the batch began. */
int ro_build_vamana(const ro_space *s, const uint8_t *cbase, size_t cstride, size_t n, size_t R, size_t L,
                    float alpha, size_t beam, size_t batch, size_t refine, size_t cap, size_t early_stop,
                    uint32_t *deg, uint32_t *edges, uint64_t *evals_out) {
    if (R == 0 || R > 256 || beam == 0 || beam > 64 || !(alpha >= 1.0f) || !isfinite(alpha)) return -1;
    if (batch == 0) batch = 1;
    uint64_t evals = 0;
    memset(deg, 0, n * sizeof(uint32_t));
    uint32_t *tags = (uint32_t *)calloc(n + 1, sizeof(uint32_t));
    uint32_t tag = 0;
    ro_cand *pool = (ro_cand *)malloc((L + 1) * sizeof(ro_cand));
    uint32_t *pid = (uint32_t *)malloc((L + R + 1) * sizeof(uint32_t));
    float *pd = (float *)malloc((L + R + 1) * sizeof(float));
    uint32_t *nbr = (uint32_t *)malloc((batch * R + 1) * sizeof(uint32_t));
    uint32_t *cnt = (uint32_t *)malloc((batch + 1) * sizeof(uint32_t));
    be_pair *pairs = (be_pair *)malloc((n * R + 1) * sizeof(be_pair));
    int rc = 0;
    for (size_t begin = 1; begin < n && rc == 0; begin += batch) {
        const size_t end = begin + batch < n ? begin + batch : n;
        for (size_t id = begin; id < end; ++id) {
            size_t len = build_search(s, cbase, cstride, (uint32_t)id, deg, edges, R, begin, L, beam, early_stop,
                                      tags, ++tag, pool, &evals);
            if (len == (size_t)-1) { rc = -2; break; }
            if (cap != 0 && len > cap) { qsort(pool, len, sizeof(ro_cand), cap_less); len = cap; }
            for (size_t i = 0; i < len; ++i) { pid[i] = pool[i].id; pd[i] = pool[i].d; }
            cnt[id - begin] = (uint32_t)robust_prune_counted(s, cbase, cstride, (uint32_t)id, pid, pd, len, R, alpha,
                                                             nbr + (id - begin) * R, &evals);
        }
        size_t np = 0;
        for (size_t id = begin; id < end && rc == 0; ++id) {
            const uint32_t *row = nbr + (id - begin) * R;
            memcpy(edges + id * R, row, cnt[id - begin] * sizeof(uint32_t));
            deg[id] = cnt[id - begin];
            for (uint32_t j = 0; j < cnt[id - begin]; ++j) { pairs[np].src = row[j]; pairs[np].dst = (uint32_t)id; ++np; }
        }
        if (rc == 0) add_backedges(s, cbase, cstride, pairs, np, deg, edges, R, alpha, &evals);
    }
    if (refine != 0 && n > 1 && rc == 0) {
        uint32_t *rdeg = (uint32_t *)malloc(n * sizeof(uint32_t));
        uint32_t *redges = (uint32_t *)malloc((n * R + 1) * sizeof(uint32_t));
        for (size_t pass = 0; pass < refine && rc == 0; ++pass) {
            for (size_t id = 0; id < n; ++id) {
                size_t len = build_search(s, cbase, cstride, (uint32_t)id, deg, edges, R, n, L, beam, early_stop,
                                          tags, ++tag, pool, &evals);
                if (len == (size_t)-1) { rc = -2; break; }
                if (cap != 0 && len > cap) { qsort(pool, len, sizeof(ro_cand), cap_less); len = cap; }
                for (size_t i = 0; i < len; ++i) { pid[i] = pool[i].id; pd[i] = pool[i].d; }
                const uint8_t *rec = cbase + id * cstride;
                for (uint32_t j = 0; j < deg[id]; ++j) {
                    const uint32_t e = edges[id * R + j];
                    pid[len] = e;
                    pd[len] = ro_symmetric_distance(s, rec, cbase + (size_t)e * cstride);
                    ++evals;
                    ++len;
                }
                rdeg[id] = (uint32_t)robust_prune_counted(s, cbase, cstride, (uint32_t)id, pid, pd, len, R, alpha,
                                                          redges + id * R, &evals);
            }
            if (rc != 0) break;
            memcpy(deg, rdeg, n * sizeof(uint32_t));
            memcpy(edges, redges, n * R * sizeof(uint32_t));
            size_t np = 0;
            for (size_t id = 0; id < n; ++id)
                for (uint32_t j = 0; j < deg[id]; ++j) { pairs[np].src = edges[id * R + j]; pairs[np].dst = (uint32_t)id; ++np; }
            add_backedges(s, cbase, cstride, pairs, np, deg, edges, R, alpha, &evals);
        }
        free(rdeg); free(redges);
    }
    free(tags); free(pool); free(pid); free(pd); free(nbr); free(cnt); free(pairs);
    if (evals_out) *evals_out = evals;
    return rc;
}

/* ------------------------------------------------ record gate, B1, k-means */
/* compute_paper_prune_estimate (space_rabitq.h:3926-3959). short_ip is
 * shortCodeIpFromFullCodeAvx512 (:1307-1329): lane j of 16 takes pairs j,
 * j+16, ..., even dim then odd dim, then the _mm512_reduce_add_ps tree. */
ro_estimate ro_paper_estimate_record(const ro_space *s, const ro_query *q, const uint8_t *compact, float eps0) {
    ro_estimate r;
    memset(&r, 0, sizeof(r));
    float hdr[4];
    memcpy(hdr, compact, 16);  /* norm_sqr, long_scale, short_scale, error_scale */
    const float norm_sqr = hdr[0], short_scale = hdr[2];
    if (s->code_dim <= 1 || eps0 < 0.0f || !isfinite(eps0) || norm_sqr <= 0.0f || q->query_norm <= 0.0f ||
        short_scale <= 0.0f || !isfinite(short_scale))
        return r;
    const float data_norm = sqrtf(norm_sqr);
    const float alpha_raw = 4.0f * data_norm / (short_scale * sqrtf((float)s->code_dim));
    if (!(alpha_raw > 0.0f) || !isfinite(alpha_raw)) return r;
    const float alpha = (alpha_raw < 1.0f) ? alpha_raw : 1.0f; /* std::min(1.0f, a) */
    const uint8_t *code = compact + 16;
    float sh[16] = {0};
    for (size_t p0 = 0; p0 < s->code_dim / 2; p0 += 16)
        for (int j = 0; j < 16; ++j) {
            const uint8_t b = code[p0 + j];
            if ((b >> 3) & 1u) sh[j] = sh[j] + q->q[2 * (p0 + j)];
            if ((b >> 7) & 1u) sh[j] = sh[j] + q->q[2 * (p0 + j) + 1];
        }
    const float short_ip = reduce16(sh) - q->half_sum;
    const float cross = short_ip * short_scale;
    const float two_norms = 2.0f * data_norm * q->query_norm;
    const float ip_hat = cross / two_norms;
    const float alpha2 = alpha * alpha;
    const float t = (1.0f - alpha2) / alpha2;
    const float eb = sqrtf(t > 0.0f ? t : 0.0f) * eps0 / sqrtf((float)(s->code_dim - 1));
    float up = ip_hat + eb;
    up = (up < 1.0f) ? up : 1.0f;    /* std::min(1.0f, v)  */
    up = (-1.0f < up) ? up : -1.0f;  /* std::max(-1.0f, v) */
    const float base = norm_sqr + q->query_norm_sqr;
    r.lower_bound = SITE(RO_FMA_PAPER_LOWER) ? fmaf(-two_norms, up, base) : base - two_norms * up;
    r.short_ip = short_ip;
    r.alpha = alpha;
    r.ip_hat = ip_hat;
    r.error_bound = eb;
    r.valid = isfinite(r.lower_bound) ? 1 : 0;
    return r;
}

/* The B1 query code (buildCentroidQuery, space_rabitq.h:3493-3511) and
 * quantizedLongCodeDotDispatch against it: exact integers. */
int64_t ro_b1_dot(const ro_space *s, const ro_query *q, const uint8_t *compact, int64_t *signed_sum) {
    int64_t dot = 0, sum = 0;
    for (size_t i = 0; i < s->code_dim; ++i) {
        const int x = q->q[i] > 0.0f ? 1 : -1;
        sum += x;
        dot += (int64_t)get_nibble(compact + 16, i) * x;
    }
    if (signed_sum) *signed_sum = sum;
    return dot;
}

float ro_b1_distance(const ro_space *s, const ro_query *q, const uint8_t *compact) {
    float hdr[2];
    memcpy(hdr, compact, 8);
    if (!long_scale_ok(hdr[1])) return hdr[0] + q->query_norm_sqr;
    int64_t sum = 0;
    const int64_t dot = ro_b1_dot(s, q, compact, &sum);
    const double signed_long_ip = (double)dot - 7.5 * (double)sum;
    const double est_ip = (double)hdr[1] * signed_long_ip;
    return hdr[0] + q->query_norm_sqr - (float)est_ip;
}

int ro_train_centroids(const float *data, size_t count, size_t dim, size_t k, size_t sample_count, uint32_t seed,
                       int iters, float *out) {
    if (!data || !out || count == 0 || dim == 0 || k == 0 || k > 256 || sample_count == 0) return -1;
    if (sample_count > count) sample_count = count;
    mt32 g;
    mt32_seed(&g, seed);
    size_t *sample = (size_t *)malloc(sample_count * sizeof(size_t));
    for (size_t i = 0; i < sample_count; ++i) sample[i] = (size_t)mt32_next(&g) % count;
    float *cen = (float *)malloc(k * dim * sizeof(float));
    float *next = (float *)malloc(k * dim * sizeof(float));
    size_t *cnt = (size_t *)malloc(k * sizeof(size_t));
    for (size_t c = 0; c < k; ++c) memcpy(cen + c * dim, data + sample[c % sample_count] * dim, dim * sizeof(float));
    for (int it = 0; it < iters; ++it) {
        memset(next, 0, k * dim * sizeof(float));
        memset(cnt, 0, k * sizeof(size_t));
        for (size_t i = 0; i < sample_count; ++i) {
            const float *v = data + sample[i] * dim;
            size_t best = 0;
            float best_d = INFINITY;
            for (size_t c = 0; c < k; ++c) {
                const float *cc = cen + c * dim;
                float d = 0.0f;
                for (size_t j = 0; j < dim; ++j) {
                    const float diff = v[j] - cc[j];
                    d = madd_f(RO_FMA_KMEANS_DIST, diff, diff, d);
                }
                if (d < best_d) { best_d = d; best = c; }
            }
            float *dst = next + best * dim;
            for (size_t j = 0; j < dim; ++j) dst[j] += v[j];
            ++cnt[best];
        }
        for (size_t c = 0; c < k; ++c) {
            if (cnt[c] == 0) continue;
            const float inv = 1.0f / (float)cnt[c];
            for (size_t j = 0; j < dim; ++j) cen[c * dim + j] = next[c * dim + j] * inv;
        }
    }
    memcpy(out, cen, k * dim * sizeof(float));
    free(sample); free(cen); free(next); free(cnt);
    return 0;
}

/* ------------------------------------------------ 1-bit sign-plane scan */

/* build_or_load_database, query_codec_1bit_scan.cpp:341-384: bit b of word
 * w is set when rotated coordinate w*64+b is > 0. */
int ro_ss_encode(const ro_space *s, const float *x, uint64_t n, uint64_t *db) {
    const size_t cd = s->code_dim, W = cd / 64;
    if (cd % 64 != 0 || n % RO_SS_BATCH != 0) return -1;
    float *rot = (float *)malloc(cd * sizeof(float));
    for (uint64_t row = 0; row < n; ++row) {
        ro_rotate(s, x + row * s->dim, rot);
        const uint64_t block = row / RO_SS_BATCH, lane = row % RO_SS_BATCH;
        for (size_t w = 0; w < W; ++w) {
            uint64_t bits = 0;
            for (size_t b = 0; b < 64; ++b)
                if (rot[w * 64 + b] > 0.0f) bits |= (uint64_t)1 << b;
            db[(block * W + w) * RO_SS_BATCH + lane] = bits;
        }
    }
    free(rot);
    return 0;
}

/* make_integer_query<Bits>, :427-456: max-abs scale, round-to-nearest-even
 * codes clamped to +-(2^(Bits-1) - 1), two's-complement bit planes. */
static void ss_integer_query(const float *q, size_t cd, int bits, uint64_t *planes, float *scale_out,
                             float *constant_out, int64_t *sum_out) {
    const int qmax = (1 << (bits - 1)) - 1;
    const size_t W = cd / 64;
    float max_abs = 0.0f;
    for (size_t i = 0; i < cd; ++i) {
        const float a = fabsf(q[i]);
        max_abs = max_abs < a ? a : max_abs;
    }
    const float scale = max_abs > 0.0f ? max_abs / (float)qmax : 1.0f;
    memset(planes, 0, (size_t)bits * W * sizeof(uint64_t));
    int64_t sum = 0, norm2 = 0;
    for (size_t i = 0; i < cd; ++i) {
        long c = lrintf(q[i] / scale);
        if (c < -qmax) c = -qmax;
        if (c > qmax) c = qmax;
        sum += c;
        norm2 += (int64_t)c * c;
        const uint8_t enc = (uint8_t)(int8_t)c;
        for (int p = 0; p < bits; ++p)
            if ((enc >> p) & 1u) planes[(size_t)p * W + i / 64] |= (uint64_t)1 << (i % 64);
    }
    *scale_out = scale;
    *constant_out = madd_f(RO_FMA_SS_INTCONST, scale * scale, (float)norm2, (float)cd);
    *sum_out = sum;
}

/* prepare_query, :458-492: the rotation, then in index order the FP32 sum
 * and squared norm, the B1 sign code with its mean-|q| scale, and the INT4
 * and INT8 bit planes. */
void ro_ss_prepare(const ro_space *s, const float *raw, float *q, uint64_t *bits, ro_ss_query *out) {
    const size_t cd = s->code_dim, W = cd / 64;
    memset(out, 0, sizeof(*out));
    ro_rotate(s, raw, q);
    float sum = 0.0f, norm2 = 0.0f, abs_sum = 0.0f;
    for (size_t i = 0; i < cd; ++i) {
        sum += q[i];
        norm2 = madd_f(RO_FMA_SS_NORM2, q[i], q[i], norm2);
    }
    memset(bits, 0, W * sizeof(uint64_t));
    for (size_t i = 0; i < cd; ++i) {
        if (q[i] > 0.0f) bits[i / 64] |= (uint64_t)1 << (i % 64);
        abs_sum += fabsf(q[i]);
    }
    out->fp32_sum = sum;
    out->fp32_norm2 = norm2;
    out->b1_scale = abs_sum / (float)cd;
    out->b1_constant = (float)cd * madd_f(RO_FMA_SS_B1CONST, out->b1_scale, out->b1_scale, 1.0f);
    ss_integer_query(q, cd, 4, bits + W, &out->int4_scale, &out->int4_constant, &out->int4_sum);
    ss_integer_query(q, cd, 8, bits + 5 * W, &out->int8_scale, &out->int8_constant, &out->int8_sum);
}

static uint64_t ss_word(const uint64_t *db, size_t W, uint64_t cand, size_t w) {
    return db[((cand / RO_SS_BATCH) * W + w) * RO_SS_BATCH + cand % RO_SS_BATCH];
}

/* scan_fp32 / scan_b1 / scan_integer<Bits>, :529-648. FP32: 16 lane
 * accumulators, lane j summing the query values of the set bits among
 * dimensions j, j+16, j+32, ... in order, then the _mm512_reduce_add_ps
 * tree; B1: XOR popcount; INT4/INT8: per-plane AND popcounts weighted by
 * the two's-complement plane weights (exact integers). */
void ro_ss_scan(const ro_space *s, int codec, const uint64_t *db, uint64_t n, const float *q,
                const uint64_t *bits, const ro_ss_query *qs, float *out) {
    const size_t cd = s->code_dim, W = cd / 64;
    const uint64_t count = (n / RO_SS_BATCH) * RO_SS_BATCH;
    for (uint64_t c = 0; c < count; ++c) {
        if (codec == 0) {
            float acc[16] = {0};
            for (size_t w = 0; w < W; ++w) {
                const uint64_t word = ss_word(db, W, c, w);
                for (size_t b = 0; b < 64; ++b)
                    if ((word >> b) & 1u) acc[b % 16] += q[w * 64 + b];
            }
            const float selected = reduce16(acc);
            const float dot = 2.0f * selected - qs->fp32_sum;
            out[c] = ((float)cd + qs->fp32_norm2) - 2.0f * dot;
        } else if (codec == 1) {
            uint64_t mism = 0;
            for (size_t w = 0; w < W; ++w) mism += (uint64_t)__builtin_popcountll(ss_word(db, W, c, w) ^ bits[w]);
            const float signed_dot = (float)cd - 2.0f * (float)mism;
            out[c] = madd_f(RO_FMA_SS_B1DIST, -(2.0f * qs->b1_scale), signed_dot, qs->b1_constant);
        } else {
            const int B = codec == 2 ? 4 : 8;
            const uint64_t *planes = bits + (codec == 2 ? 1 : 5) * W;
            int32_t selected = 0;
            for (int p = 0; p < B; ++p) {
                const int32_t weight = p + 1 == B ? -(1 << (B - 1)) : (1 << p);
                int32_t cnt = 0;
                for (size_t w = 0; w < W; ++w)
                    cnt += __builtin_popcountll(ss_word(db, W, c, w) & planes[(size_t)p * W + w]);
                selected += cnt * weight;
            }
            const int64_t dot = 2LL * selected - (codec == 2 ? qs->int4_sum : qs->int8_sum);
            const float scale = codec == 2 ? qs->int4_scale : qs->int8_scale;
            const float constant = codec == 2 ? qs->int4_constant : qs->int8_constant;
            out[c] = madd_f(RO_FMA_SS_INTDIST, -(2.0f * scale), (float)dot, constant);
        }
    }
}
