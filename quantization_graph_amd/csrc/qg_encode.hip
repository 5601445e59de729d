// qg_encode.hip -- ExRaBitQ4 block16 encode of N base vectors
// (encodeVectorFullWithCentroid, Ours/core/hnswlib/space_rabitq.h:5845-5999,
// with fastQuantizeAbs :867-942 and the residual scales :729-786).
//
// One workgroup per vector does the whole encode in LDS (k_encode): the
// input row is read once and only the records are written. The reference's
// strictly sequential floating-point recurrences that decide output bits
// (the norm, the decoded-residual sums, the scale factors) run on single
// lanes in its order; the critical-point scan of fastQuantizeAbs, which is
// sequential in the reference, is recast as a data-parallel computation
// that provably picks the same event (see k_encode), with the sequential
// merge kept for exact ties; the sum behind ip_norm is parallel with an
// error bound, falling back to the reference's order when the bound could
// change its float. Spaces of several centroids first assign each vector
// its centroid (k_enc_assign).
#include <mutex>

#include "qg_args.h"

namespace qg {

constexpr uint32_t kEncThreads = 256;
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

__device__ __forceinline__ int init_code(double t, float a) {
    int c = (int)fma(t, (double)a, kEps);
    return c < 7 ? c : 7;
}

__device__ __forceinline__ int clamp7_round(double v) {
    int q = (int)round(v);
    q = q < 7 ? q : 7;      // std::min(qmax, .)
    return q > -7 ? q : -7; // std::max(qmin, .)
}

// lane k of each group of 4 lanes (DPP quad broadcast), for doubles
__device__ __forceinline__ double quad_bcast(double v, int k) {
    const int ctrl = k | (k << 2) | (k << 4) | (k << 6);
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    int lo = (int)(uint32_t)b, hi = (int)(uint32_t)(b >> 32);
    switch (ctrl) {
    case 0x00: lo = __builtin_amdgcn_mov_dpp(lo, 0x00, 0xF, 0xF, false); hi = __builtin_amdgcn_mov_dpp(hi, 0x00, 0xF, 0xF, false); break;
    case 0x55: lo = __builtin_amdgcn_mov_dpp(lo, 0x55, 0xF, 0xF, false); hi = __builtin_amdgcn_mov_dpp(hi, 0x55, 0xF, 0xF, false); break;
    case 0xAA: lo = __builtin_amdgcn_mov_dpp(lo, 0xAA, 0xF, 0xF, false); hi = __builtin_amdgcn_mov_dpp(hi, 0xAA, 0xF, 0xF, false); break;
    default: lo = __builtin_amdgcn_mov_dpp(lo, 0xFF, 0xF, 0xF, false); hi = __builtin_amdgcn_mov_dpp(hi, 0xFF, 0xF, 0xF, false); break;
    }
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// residualMseOptimalScale / residualMaxAbsScale (space_rabitq.h:729-786)
// for a block of 16 errors spread over a group of 4 lanes (4 each, lane s
// holding elements 4s..4s+3): the max-abs scale, then up to 8 least-squares
// steps when MSE is on. Each lane computes its own 4 codes; the 16 exact
// products e q are gathered so that every lane adds them in index order,
// as the reference's double sum does, and all 4 lanes agree on each step.
__device__ __forceinline__ float residual_scale_quad(const float (&e)[4], float mx, int mse) {
    float scale = mx <= 0.0f ? 0.0f : mx / 7.0f;
    if (mse && scale != 0.0f && __builtin_isfinite(scale)) {
        for (int it = 0; it < 8; ++it) {
            double p[4];
            double den = 0.0;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int q = clamp7_round((double)e[u] / (double)scale);
                p[u] = (double)e[u] * (double)q;
                den += (double)q * (double)q;
            }
            // den: small integers, exact in any order
            const double dsum = quad_bcast(den, 0) + quad_bcast(den, 1) + quad_bcast(den, 2) + quad_bcast(den, 3);
            double num = 0.0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
#pragma unroll
                for (int u = 0; u < 4; ++u) num += quad_bcast(p[u], k);
            }
            if (dsum <= 0.0) break;
            const float next = (float)(num / dsum);
            if (next <= 0.0f || !__builtin_isfinite(next)) break;
            const float lim = 1e-6f * ((1.0f < scale) ? scale : 1.0f);
            if (fabsf(next - scale) <= lim) { scale = next; break; }
            scale = next;
        }
    }
    return scale;
}

// ------------------------------------------------- workgroup reductions
// 256 threads = 4 wavefronts; `red` is 4 slots of LDS scratch. Each ends
// with every thread holding the result and the scratch free again.
__device__ __forceinline__ double block_max_d(double v, double *red) {
    for (int off = 32; off > 0; off >>= 1) { const double o = __shfl_xor(v, off); v = v < o ? o : v; }
    __syncthreads();
    if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = red[0];
    for (int w = 1; w < 4; ++w) r = r < red[w] ? red[w] : r;
    return r;
}
__device__ __forceinline__ double block_sum_d(double v, double *red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    __syncthreads();
    if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

// Bitonic sort of CD 64-bit keys, ascending, E = CD / 256 consecutive keys
// per thread (threads past CD / E hold none). Partners closer than E keys
// are in the same thread, closer than 64 E in the same wavefront (xor
// shuffles), farther ones go through LDS.
template <int CD>
__device__ __forceinline__ void enc_sort(uint64_t (&k)[CD >= 256 ? CD / 256 : 1], uint64_t *lds, uint32_t t) {
    constexpr int E = CD >= 256 ? CD / 256 : 1;
    constexpr uint32_t NT = CD / E;
    const bool act = t < NT;
#pragma unroll
    for (uint32_t size = 2; size <= (uint32_t)CD; size <<= 1) {
#pragma unroll
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            if (stride < (uint32_t)E) {
#pragma unroll
                for (int j = 0; j < E; ++j) {
                    const int pj = j ^ (int)stride;
                    if (pj > j) {
                        const bool up = (((t * E + j) & size) == 0);
                        const uint64_t lo = k[j], hi = k[pj];
                        const bool sw = up ? (lo > hi) : (lo < hi);
                        k[j] = sw ? hi : lo;
                        k[pj] = sw ? lo : hi;
                    }
                }
            } else {
                const uint32_t m = stride / E;  // partner thread = t ^ m
                uint64_t other[E];
                if (m < 64) {
#pragma unroll
                    for (int j = 0; j < E; ++j) {
                        const uint32_t lo = __shfl_xor((uint32_t)k[j], (int)m);
                        const uint32_t hi = __shfl_xor((uint32_t)(k[j] >> 32), (int)m);
                        other[j] = ((uint64_t)hi << 32) | lo;
                    }
                } else {
                    __syncthreads();
                    if (act) {
#pragma unroll
                        for (int j = 0; j < E; ++j) lds[t * E + j] = k[j];
                    }
                    __syncthreads();
                    if (act) {
#pragma unroll
                        for (int j = 0; j < E; ++j) other[j] = lds[(t ^ m) * E + j];
                    }
                }
                const bool lower = (t & m) == 0;
#pragma unroll
                for (int j = 0; j < E; ++j) {
                    const bool up = (((t * E + j) & size) == 0);
                    const uint64_t mn = k[j] < other[j] ? k[j] : other[j];
                    const uint64_t mx = k[j] < other[j] ? other[j] : k[j];
                    k[j] = (up == lower) ? mn : mx;
                }
            }
        }
    }
}

// The reference's sequential critical-point scan (fastQuantizeAbs,
// space_rabitq.h:886-925) as a 7-way merge of the per-level event lists over
// the sorted |o| in LDS, one lane; the fallback of the parallel scan below.
__device__ double enc_scan_sequential(const float *V, const float *SA, const uint16_t *SI, const uint32_t *lvs,
                                      const uint32_t *lve, uint32_t cd, double max_o) {
    const double t_start = 2.0 / max_o;
    double den = (double)cd * 0.25, num = 0.0;
    for (uint32_t i = 0; i < cd; ++i) {
        const float ab = fabsf(V[i]);
        const int c = init_code(t_start, ab);
        den += (double)(c * c + c);
        num += ((double)c + 0.5) * (double)ab;
    }
    uint32_t cur[7];
    double ht[7];
    for (int k = 0; k < 7; ++k) {
        cur[k] = lvs[k];
        ht[k] = cur[k] < lve[k] ? (double)(k + 1) / (double)SA[cur[k]] : __builtin_inf();
    }
    double best_t = 0.0, max_ip = 0.0;
    for (;;) {
        int kb = -1;
        double tb = __builtin_inf();
        uint32_t ib = 0xFFFFFFFFu;
        for (int k = 0; k < 7; ++k) {
            if (cur[k] < lve[k]) {
                const uint32_t ik = SI[cur[k]];
                if (ht[k] < tb || (ht[k] == tb && ik < ib)) { kb = k; tb = ht[k]; ib = ik; }
            }
        }
        if (kb < 0) break;
        num += (double)SA[cur[kb]];
        den += 2.0 * (double)(kb + 1);
        const double ip = num / sqrt(den);
        if (ip > max_ip) { max_ip = ip; best_t = tb; }
        ++cur[kb];
        ht[kb] = cur[kb] < lve[kb] ? (double)(kb + 1) / (double)SA[cur[kb]] : __builtin_inf();
    }
    return best_t;
}

// How many of the 8 sorted positions q0, q0 + step, ..., q0 + 7 step (those
// below `end`) hold an event before the one of |o| = a_p (dimension sip) in
// the order of the critical-point scan: |o_q| > th, or |o_q| == th when the
// threshold is exact and the dimension is lower. The positions of a level
// slice that precede form a prefix, so this is also the length of the
// preceding prefix among them. Reads past `end` stay inside the workgroup's
// LDS and are masked.
__device__ __forceinline__ uint32_t prec_count8(const float *SA, const uint16_t *SI, uint32_t q0, uint32_t step,
                                                uint32_t end, float th, bool exact, uint32_t sip) {
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = SA[q0 + (uint32_t)i * step];
    uint32_t c = 0, eq = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const bool in = q0 + (uint32_t)i * step < end;
        c += (in && v[i] > th) ? 1u : 0u;
        eq |= (in && exact && v[i] == th) ? (1u << i) : 0u;
    }
    if (eq) {
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if ((eq >> i) & 1u) c += ((uint32_t)SI[q0 + (uint32_t)i * step] < sip) ? 1u : 0u;
    }
    return c;
}

// Float-bit buckets of [2^-20, 8): 128 per binade. U[b] = #{q : bucket(SA[q])
// >= b} turns "how many sorted |o| exceed a threshold" into two table reads
// and a scan of the threshold's own bucket.
constexpr uint32_t kBucketBase = (127u - 20u) << 7;
constexpr uint32_t kBuckets = 23u << 7;
__device__ __forceinline__ uint32_t enc_bucket(float x) {
    const int b = (int)(__float_as_uint(x) >> 16) - (int)kBucketBase;
    return b < 0 ? 0u : (b >= (int)kBuckets ? kBuckets - 1u : (uint32_t)b);
}

// short_scale and error_scale of the formal encode (:5912-5938) from
// half_l1 = sum 0.5 |o_i| and o_obar = sum |o_i| / sqrt(cd), in the
// reference's double operations; both fall as their input grows
struct EncScales {
    float ss, es;
    bool live;
};
__device__ __forceinline__ EncScales enc_scales(float rn, double half_l1, double o_obar, uint32_t cd) {
    EncScales r{0.0f, 0.0f, false};
    if (rn > 0.0f && half_l1 > 1e-12) {
        r.live = true;
        r.ss = (float)(2.0 * (double)rn / half_l1);
        if (!__builtin_isfinite(o_obar)) o_obar = 0.8;
        o_obar = (1e-6 < o_obar) ? o_obar : 1e-6;                // std::max(1e-6, o)
        o_obar = (0.999999 < o_obar) ? 0.999999 : o_obar;        // std::min(0.999999, o)
        const double o2 = o_obar * o_obar;
        const double fac = ldexp(5.75, -4) / sqrt((double)cd);
        double q = (1.0 - o2) / o2;
        q = (0.0 < q) ? q : 0.0;                                   // std::max(0.0, r)
        r.es = (float)(sqrt(q) * fac * 2.0 * (double)rn);
    }
    return r;
}

// LDS of k_encode past V, SA and SI: the reused region (PA needs 8 CD + 8;
// the precision tree 2 KB), then the bucket table
__host__ __device__ constexpr uint32_t enc_x_bytes(uint32_t cd) { return cd * 8 + 16 > 2048 ? cd * 8 + 16 : 2048; }
__host__ __device__ constexpr uint32_t enc_u_bytes() { return ((kBuckets + 1) * 2 + 15) / 16 * 16; }

struct EncEval {
    const float *SA;
    const uint16_t *SI;
    const double *PA;
    const uint16_t *U;
    double num0, den0;
    uint32_t ls[7], le[7], offs[8];
};

// The value of the critical-point scan at event j (level k, the (j -
// offs[k])-th event of that level): x = num0 + the exact sum of |o| of all
// events up to it in the scan's order, d = den0 + its denominator; returns
// x / sqrt(d). The events of level k2 before it are the positions of
// level k2's slice with (k2+1) a_p < (k+1) a_q in exact double products, or
// equal products and a lower dimension: |o_q| > th, or |o_q| == th when the
// threshold (k2+1) a_p / (k+1) is a float. th is that quotient rounded down
// to float: no such quotient lies within a double ulp of a float it does not
// equal, so the product by the reciprocal rounds the same way.
__device__ __forceinline__ double enc_eval(const EncEval &e, int k, uint32_t j, double &x, double &d) {
    uint32_t p = 0, lk = 0;
#pragma unroll
    for (int q = 0; q < 7; ++q)
        if (q == k) { p = e.ls[q] + (j - e.offs[q]); lk = e.ls[q]; }
    const float ap = e.SA[p];
    const uint32_t sip = e.SI[p];
    const double kf = (double)(k + 1), rk = 1.0 / kf, apd = (double)ap;
    double S = e.PA[p + 1] - e.PA[lk];
    uint32_t den = 2u * (uint32_t)(k + 1) * (p - lk + 1);
#pragma unroll
    for (int k2 = 0; k2 < 7; ++k2) {
        if (k2 == k) continue;
        const double m = (double)(k2 + 1) * apd;
        float th = (float)(m * rk);
        if ((double)th * kf > m) th = __uint_as_float(__float_as_uint(th) - 1u);
        const bool exact = (double)th * kf == m;
        const uint32_t b = enc_bucket(th);
        uint32_t lo = e.U[b + 1];
        const uint32_t hi = e.U[b];
#pragma clang loop unroll(disable)
        for (;;) {
            const uint32_t c8 = prec_count8(e.SA, e.SI, lo, 1, hi, th, exact, sip);
            lo += c8;
            if (c8 < 8) break;
        }
        const uint32_t l2 = e.ls[k2], h2 = e.le[k2];
        const uint32_t c = (lo < l2 ? l2 : (lo > h2 ? h2 : lo)) - l2;
        den += 2u * (uint32_t)(k2 + 1) * c;
        S += e.PA[l2 + c] - e.PA[l2];
    }
    x = e.num0 + S;
    d = e.den0 + (double)den;
    return x / sqrt(d);
}

// Bound for the events after a valued event f (numerator x, denominator
// d) and before a later event whose denominator is d_end, all of one level
// k: each event g in between comes later in the scan's order, so a_g / (2
// (k_g+1)) = 1 / (2 t_g) is at most rho = a_f / (2 (k+1)): the numerator
// grows by at most rho per unit of denominator, and (x + rho u) / sqrt(d +
// u) takes its largest value over u in [0, d_end - d] at an end. Inflated by
// 2^-40 and rounded up to float, it is still a bound.
__device__ __forceinline__ float enc_chunk_bound(double x, double d, double rho, double d_end) {
    const double u = d_end - d;
    const double ip0 = x / sqrt(d);
    double ub = (x + rho * u) / sqrt(d + u);
    ub = (ub > ip0 ? ub : ip0) * (1.0 + 0x1p-40);
    float f = (float)ub;
    if ((double)f < ub) f = __uint_as_float(__float_as_uint(f) + 1u);
    return f;
}

// 8-event chunks per thread (chunks never cross a level): enough for every
// possible event up to code dimension 1024 (7 cd events at most); above,
// enough for 4 cd events, which covers the typical 2.6 cd. A vector with
// more events takes the sequential merge.
template <int CD>
struct EncChunks {
    static constexpr int ALL = (7 * CD / 8 + 7 + 255) / 256;
    static constexpr int PER_THREAD = CD <= 1024 ? ALL : (CD / 2 + 7 + 255) / 256;
};

constexpr uint32_t kRefine = 256;

// bucket table and refine list placed over V (code dimension 2048 and up)
template <int CD>
struct EncOverlay { static constexpr bool ON = CD >= 2048; };
static_assert(((kBuckets + 1) * 2 + 15) / 16 * 16 + 4 * kRefine <= 4 * 2048, "the overlay fits V");

// waves per SIMD the encode is compiled for (its register budget)
template <int CD>
struct EncOcc { static constexpr int WAVES = CD <= 1024 ? 6 : 4; };

// One base vector per workgroup of 256 threads, every intermediate in LDS:
//   load x - c; ||x - c|| (one lane, the reference's order); normalise,
//   sign-flip, FHT (or take the QR rotation's output); |o| and its maximum;
//   sort the dimensions by (|o| desc, index asc); the seven per-level event
//   lists as slices of that order; the critical-point scan (below); codes,
//   MSB plane, quantisation errors, residual block scales and codes; the
//   scale factors and the decoded-residual sums.
//
// The critical-point scan. Dimension i with |o_i| = a_i, initial code c0_i,
// has events at t = (k+1)/a_i for k = c0_i (always, if a_i > eps) and each
// k in (c0_i, 6] with t < t_end. Level k's events, in (t, i) order, are a
// contiguous slice of the sorted order. The reference pops all events in
// (t, i) order, adding a_i to a running double numerator and 2(k+1) to the
// denominator, and keeps the first event of largest num / sqrt(den). Here an
// event's value is computed on its own (enc_eval): num0 plus the exact sum
// of the a's of all events up to it (each a > eps is a multiple of 2^-40 and
// the sums stay below 2^13, so the double sums are exact), over the square
// root of the exact denominator. That value differs from the reference's
// sequentially rounded one by at most (cd + events + 18) units of 2^-53
// relative, so the reference's winner lies within that window of the
// largest value. The events are cut into chunks of 8 within a level; the
// chunk ends are valued first, and since between them the numerator and the
// denominator only grow, num(last) / sqrt(den(first)) bounds the events in
// between: only chunks whose bound reaches the window are valued in full.
// When exactly one event lies in twice the window, it is the reference's
// winner; otherwise (exact ties, as in a vector whose |o| are all equal) one
// lane runs the reference's sequential merge.
template <int CD>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(EncOcc<CD>::WAVES))) k_encode(EncArgs a) {
    constexpr uint32_t T = 256;
    constexpr int E = CD >= 256 ? CD / 256 : 1;
    constexpr uint32_t NT = CD / E;
    constexpr int CPT = EncChunks<CD>::PER_THREAD;
    extern __shared__ __align__(16) unsigned char smem[];
    float *V = (float *)smem;                          // CD unit vector
    float *SA = (float *)(smem + 4 * CD);              // CD |o| sorted descending
    uint16_t *SI = (uint16_t *)(smem + 8 * CD);        // CD their dimension
    unsigned char *X = smem + 10 * CD;                 // 8 CD + 16 (at least 2 KB), reused:
    uint64_t *KEY = (uint64_t *)X;                     //   sort exchange
    double *PA = (double *)X;                          //   prefix sums of SA (CD + 1)
    float *ERR = (float *)X;                           //   quantisation errors
    uint8_t *NIB = X + 4 * CD;                         //   primary nibbles
    int8_t *QV = (int8_t *)(X + 5 * CD);               //   residual codes
    uint16_t *RS = (uint16_t *)(X + 6 * CD);           //   fp16 block scales
    float *XC = (float *)X;                            //   precision mode: x - c
    // the bucket table and the refine list: past the reused region, or, from
    // code dimension 2048, over V, which is rebuilt after the scan from SA, SI
    // and the sign bits (4 workgroups per CU at code dimension 2048)
    unsigned char *UR = EncOverlay<CD>::ON ? smem : smem + 10 * CD + enc_x_bytes(CD);
    uint16_t *U = (uint16_t *)UR;                                    // kBuckets + 1
    uint32_t *REF = (uint32_t *)(UR + enc_u_bytes());                // kRefine
    __shared__ uint32_t s_sign[EncOverlay<CD>::ON ? CD / 32 : 1];
    __shared__ double red[4];
    __shared__ float s_f[4];
    __shared__ uint32_t s_pos[16], s_lvs[7], s_lve[7], s_off[8], s_choff[8], s_cnt, s_best, s_nref;
    __shared__ double s_best_t;
    const uint32_t t = threadIdx.x;
    const uint64_t vi = blockIdx.x;
    const uint32_t D = a.dim;
    const float *x = a.x + vi * D;
    const float *center = center_of(a, vi);
    if (t == 0 && a.cid && a.centroid_count == 1) a.cid[vi] = 0;

    // x - c (and the QR rotation's Q (x - c) in precision mode)
    if (a.precision) {
        for (uint32_t i = t; i < CD; i += T) {
            V[i] = a.rot[vi * CD + i];
            if (i < D) XC[i] = x[i] - center[i];
        }
    } else {
        for (uint32_t i = t; i < CD; i += T) V[i] = i < D ? x[i] - center[i] : 0.0f;
    }
    __syncthreads();
    // ||x - c||: the reference compiler vectorises this loop, so the products
    // of its 16- and 8-wide blocks are rounded and added in index order and
    // only the scalar tail (D mod 8) is fused; precision mode takes a
    // compensated sum
    if (t == 0) {
        float rn;
        if (a.precision) {
            Comp c;
            for (uint32_t i = 0; i < D; ++i) c.add_sq(XC[i]);
            rn = sqrtf(c.value());
        } else {
            const uint32_t vec_end = D >= 16 ? (D / 16) * 16 + ((D % 16) >= 8 ? 8 : 0) : 0;
            float s = 0.0f;
            uint32_t i = 0;
            for (; i < vec_end; i += 8) {
                const float4 u0 = *(const float4 *)(V + i), u1 = *(const float4 *)(V + i + 4);
                float p;
                p = u0.x * u0.x; s = s + p; p = u0.y * u0.y; s = s + p;
                p = u0.z * u0.z; s = s + p; p = u0.w * u0.w; s = s + p;
                p = u1.x * u1.x; s = s + p; p = u1.y * u1.y; s = s + p;
                p = u1.z * u1.z; s = s + p; p = u1.w * u1.w; s = s + p;
            }
            for (; i < D; ++i) s = fmaf(V[i], V[i], s);
            rn = sqrtf(s);
        }
        s_f[0] = rn;
    }
    __syncthreads();
    const float rn = s_f[0];
    if (a.precision) {
        // unit vector = Q (x - c) over its compensated norm, which corrects the
        // GEMM's scale error too (Q is orthogonal); fixed-order tree of the
        // 256 per-thread compensated sums
        Comp c;
        for (uint32_t i = t; i < CD; i += T) c.add_sq(V[i]);
        float *rs_ = (float *)X, *rc_ = rs_ + T;
        __syncthreads();
        rs_[t] = c.s;
        rc_[t] = c.c;
        __syncthreads();
        for (uint32_t h = T / 2; h > 0; h >>= 1) {
            if (t < h) {
                Comp l, r;
                l.s = rs_[t]; l.c = rc_[t];
                r.s = rs_[t + h]; r.c = rc_[t + h];
                l.merge(r);
                rs_[t] = l.s; rc_[t] = l.c;
            }
            __syncthreads();
        }
        Comp tot;
        tot.s = rs_[0];
        tot.c = rc_[0];
        const float n2 = tot.value();
        const float inv = n2 > 0.0f ? 1.0f / sqrtf(n2) : 0.0f;
        __syncthreads();
        for (uint32_t i = t; i < CD; i += T) V[i] = V[i] * inv;
        __syncthreads();
    } else {
        const bool scale = rn > 0.0f;
        const float inv = scale ? 1.0f / rn : 1.0f;
        for (uint32_t i = t; i < D; i += T) {
            float u = V[i];
            if (scale) u = u * inv;
            if (!a.identity) u = u * a.signs[i];
            V[i] = u;
        }
        __syncthreads();
        if (!a.identity) fht_lds(V, CD);
    }

    // |o|, its maximum, and the sort keys (|o| descending, index ascending)
    uint64_t key[E];
    float amax = 0.0f;
    uint32_t sbits = 0;
#pragma unroll
    for (int j = 0; j < E; ++j) {
        const uint32_t i = t * E + j;
        const float v = i < CD ? V[i] : 0.0f;
        const float ab = fabsf(v);
        sbits |= (__float_as_uint(v) >> 31) << j;
        amax = amax < ab ? ab : amax;
        key[j] = ((uint64_t)(~__float_as_uint(ab)) << 32) | i;
    }
    if (EncOverlay<CD>::ON) {
        // E = CD / 256 >= 8 bits per thread: whole bytes of the sign mask
        if (t < NT) ((uint8_t *)s_sign)[t * (E / 8)] = (uint8_t)sbits;
        if (E == 16 && t < NT) ((uint8_t *)s_sign)[t * 2 + 1] = (uint8_t)(sbits >> 8);
    }
    const double max_o = block_max_d((double)amax, red);
    const bool quant = max_o > kEps;
    double best_t = 0.0;
    if (quant) {
        const double t_start = 2.0 / max_o, t_end = 17.0 / max_o;
        // num0 (any order: it enters only through the window) and den0
        double nsum = 0.0, dsum = 0.0;
#pragma unroll
        for (int j = 0; j < E; ++j) {
            const uint32_t i = t * E + j;
            if (i < CD) {
                const float ab = __uint_as_float(~(uint32_t)(key[j] >> 32));
                const int c = init_code(t_start, ab);
                nsum += ((double)c + 0.5) * (double)ab;
                dsum += (double)(c * c + c);
            }
        }
        const double num0 = block_sum_d(nsum, red);
        const double den0 = (double)CD * 0.25 + block_sum_d(dsum, red);
        enc_sort<CD>(key, KEY, t);
        if (t < NT) {
#pragma unroll
            for (int j = 0; j < E; ++j) {
                SA[t * E + j] = __uint_as_float(~(uint32_t)(key[j] >> 32));
                SI[t * E + j] = (uint16_t)(key[j] & 0xFFFFu);
            }
        }
        __syncthreads();
        // level slices: P_k = #{c0 > k}, Q_k = #{(k+1)/a < t_end} and
        // N = #{a > eps}, each a prefix of the sorted order; level k is
        // [P_k, min(N, max(P_{k-1}, Q_k))) with P_{-1} = CD
        if (t < 15) {
            uint32_t lo = 0, hi = CD;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                const float am = SA[mid];
                bool in;
                if (t < 7) in = init_code(t_start, am) > (int)t;
                else if (t < 14) in = (double)(t - 6) / (double)am < t_end;
                else in = (double)am > kEps;
                if (in) lo = mid + 1; else hi = mid;
            }
            s_pos[t] = lo;
        }
        // bucket table; each thread a run of buckets: one search, then steps
        {
            constexpr uint32_t per = (kBuckets + 1 + T - 1) / T;
            const uint32_t b0 = t * per;
            uint32_t q = 0;
            if (b0 <= kBuckets) {
                uint32_t lo = 0, hi = CD;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (enc_bucket(SA[mid]) >= b0) lo = mid + 1; else hi = mid;
                }
                q = lo;
            }
            for (uint32_t b = b0; b < b0 + per && b <= kBuckets; ++b) {
                while (q > 0 && enc_bucket(SA[q - 1]) < b) --q;
                U[b] = (uint16_t)q;
            }
        }
        __syncthreads();
        if (t == 0) {
            const uint32_t n_eps = s_pos[14];
            uint32_t off = 0, choff = 0;
            for (int k = 0; k < 7; ++k) {
                const uint32_t s0 = s_pos[k];
                const uint32_t prev = k == 0 ? (uint32_t)CD : s_pos[k - 1];
                const uint32_t q = s_pos[7 + k];
                uint32_t e0 = prev > q ? prev : q;
                e0 = e0 < n_eps ? e0 : n_eps;
                e0 = e0 > s0 ? e0 : s0;
                s_lvs[k] = s0;
                s_lve[k] = e0;
                s_off[k] = off;
                s_choff[k] = choff;
                off += e0 - s0;
                choff += (e0 - s0 + 7) / 8;
            }
            s_off[7] = off;
            s_choff[7] = choff;
            s_cnt = 0;
            s_nref = 0;
            s_best = 0xFFFFFFFFu;
        }
        // exact prefix sums of the sorted |o| (a > eps: multiples of 2^-40, sums below 2^13)
        {
            double loc[E];
            double run = 0.0;
#pragma unroll
            for (int j = 0; j < E; ++j) {
                const uint32_t p = t * E + j;
                const float am = (t < NT) ? SA[p] : 0.0f;
                run += ((double)am > kEps) ? (double)am : 0.0;
                loc[j] = run;
            }
            double incl = run;
            const uint32_t lane = t & 63u;
            for (int off = 1; off < 64; off <<= 1) {
                const double o = __shfl_up(incl, off);
                if ((int)lane >= off) incl += o;
            }
            __syncthreads();
            if (lane == 63) red[t >> 6] = incl;
            __syncthreads();
            double base = incl - run;
            for (uint32_t w = 0; w < (t >> 6); ++w) base += red[w];
            if (t < NT) {
#pragma unroll
                for (int j = 0; j < E; ++j) PA[t * E + j + 1] = base + loc[j];
            }
            if (t == 0) PA[0] = 0.0;
        }
        __syncthreads();
        EncEval ev;
        ev.SA = SA; ev.SI = SI; ev.PA = PA; ev.U = U; ev.num0 = num0; ev.den0 = den0;
#pragma unroll
        for (int q = 0; q < 7; ++q) { ev.ls[q] = s_lvs[q]; ev.le[q] = s_lve[q]; }
#pragma unroll
        for (int q = 0; q < 8; ++q) ev.offs[q] = s_off[q];
        uint32_t choff[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) choff[q] = s_choff[q];
        const uint32_t etot = ev.offs[7], nchunk = choff[7];
        // pass 1: the first event of every chunk (and the last event of each
        // level), and a bound for the events after it up to the next chunk;
        // each thread takes a run of consecutive chunks
        double m1 = -1.0, m2 = -1.0;
        uint32_t b1 = 0xFFFFFFFFu;
        float cub[CPT];
        const uint32_t cpt = (nchunk + T - 1) / T;
        const bool too_many = cpt > (uint32_t)CPT;  // then the sequential merge decides
        const uint32_t c0 = t * cpt, c1 = too_many ? c0 : (c0 + cpt < nchunk ? c0 + cpt : nchunk);
        double px = 0.0, pd = 0.0, prho = 0.0, dfirst = -1.0;
        int pr = -1;
#pragma unroll
        for (int r = 0; r < CPT; ++r) {
            cub[r] = -1.0f;
            const uint32_t c = c0 + (uint32_t)r;
            if (c < c1) {
                int k = 0;
#pragma unroll
                for (int q = 1; q < 7; ++q) k += (c >= choff[q]) ? 1 : 0;
                uint32_t jf = 0, lend = 0, pf = 0;
#pragma unroll
                for (int q = 0; q < 7; ++q)
                    if (q == k) {
                        jf = ev.offs[q] + 8u * (c - choff[q]);
                        lend = ev.offs[q + 1];
                        pf = ev.ls[q] + 8u * (c - choff[q]);
                    }
                double x1, d1;
                const double ip1 = enc_eval(ev, k, jf, x1, d1);
                if (ip1 > m1) { m2 = m1; m1 = ip1; b1 = jf; }
                else if (ip1 > m2) m2 = ip1;
                if (r == 0) dfirst = d1;
                // the previous chunk (same level) ends where this one starts
#pragma unroll
                for (int q = 0; q < CPT; ++q)
                    if (q + 1 == r && pr == q) cub[q] = enc_chunk_bound(px, pd, prho, d1);
                const double rho = (double)SA[pf] / (2.0 * (double)(k + 1));
                pr = -1;
                if (jf + 8u < lend) {
                    px = x1; pd = d1; prho = rho; pr = r;  // bounded by the next chunk's start
                } else if (jf + 1u < lend) {
                    const uint32_t jl = lend - 1u;       // the level's last event
                    double x2, d2;
                    const double ip2 = enc_eval(ev, k, jl, x2, d2);
                    if (ip2 > m1) { m2 = m1; m1 = ip2; b1 = jl; }
                    else if (ip2 > m2) m2 = ip2;
                    if (jl > jf + 1u) cub[r] = enc_chunk_bound(x1, d1, rho, d2);
                }
            }
        }
        // the thread's last chunk ends where the next thread's first starts
        {
            double dn = __shfl_down(dfirst, 1);
            __syncthreads();
            if ((t & 63u) == 0) red[t >> 6] = dfirst;
            __syncthreads();
            if ((t & 63u) == 63u) dn = (t >> 6) < 3 ? red[(t >> 6) + 1] : -1.0;
#pragma unroll
            for (int q = 0; q < CPT; ++q)
                if (pr == q && dn >= 0.0) cub[q] = enc_chunk_bound(px, pd, prho, dn);
        }
        const double M = block_max_d(m1, red);
        const double cwin = ((double)(CD + etot) + 32.0) * 0x1p-52;
        // pass 2: the insides of the chunks whose bound reaches the window
        const double lim1 = M - 2.0 * cwin * M;
#pragma unroll
        for (int r = 0; r < CPT; ++r) {
            if ((double)cub[r] >= lim1) {
                const uint32_t slot = atomicAdd(&s_nref, 1u);
                if (slot < kRefine) REF[slot] = c0 + (uint32_t)r;
            }
        }
        __syncthreads();
        const uint32_t nref = s_nref;
        const bool overflow = nref > kRefine || too_many;
        if (!overflow) {
            for (uint32_t r = t; r < nref * 7u; r += T) {
                const uint32_t c = REF[r / 7u], o = r % 7u;
                int k = 0;
#pragma unroll
                for (int q = 1; q < 7; ++q) k += (c >= choff[q]) ? 1 : 0;
                uint32_t j = 0, jend = 0;
#pragma unroll
                for (int q = 0; q < 7; ++q)
                    if (q == k) {
                        const uint32_t jf = ev.offs[q] + 8u * (c - choff[q]);
                        j = jf + 1u + o;
                        jend = jf + 8u < ev.offs[q + 1] ? jf + 8u : ev.offs[q + 1] - 1u;
                    }
                if (j < jend) {
                    double xx, dd;
                    const double ip = enc_eval(ev, k, j, xx, dd);
                    if (ip > m1) { m2 = m1; m1 = ip; b1 = j; }
                    else if (ip > m2) m2 = ip;
                }
            }
        }
        const double G = block_max_d(m1, red);
        const double lim = G - 2.0 * cwin * G;
        const uint32_t in = (m1 >= lim ? 1u : 0u) + (m2 >= lim ? 1u : 0u);
        if (in) atomicAdd(&s_cnt, in);
        if (m1 >= lim) s_best = b1;
        __syncthreads();
        if (EncOverlay<CD>::ON) {
            // V back from |o|, its dimension and its sign
            for (uint32_t p = t; p < CD; p += T) {
                const uint32_t i = SI[p];
                const uint32_t sg = (s_sign[i >> 5] >> (i & 31u)) & 1u;
                V[i] = __uint_as_float(__float_as_uint(SA[p]) | (sg << 31));
            }
            __syncthreads();
        }
        if (s_cnt == 1 && !overflow) {
            if (t == 0) {
                const uint32_t jb = s_best;
                int kb = 0;
                for (int q = 1; q < 7; ++q) kb += (jb >= s_off[q]) ? 1 : 0;
                const uint32_t pb = s_lvs[kb] + (jb - s_off[kb]);
                s_best_t = (double)(kb + 1) / (double)SA[pb];
            }
        } else if (t == 0) {
            s_best_t = enc_scan_sequential(V, SA, SI, s_lvs, s_lve, CD, max_o);
        }
        __syncthreads();
        best_t = s_best_t;
    }

    // codes and ip_norm = 1 / sum (c_i + 0.5) |o_i|; the sum is taken in any
    // order and the float of its reciprocal accepted when the sum's error
    // bound cannot move it; otherwise one lane takes the reference's order
    float ip_norm = 1.0f;
    {
        double nsum = 0.0;
        for (uint32_t i = t; i < CD; i += T) {
            const float r = V[i];
            const float ab = fabsf(r);
            const int m = quant ? init_code(best_t, ab) : 0;
            NIB[i] = r > 0.0f ? (uint8_t)(8 + m) : (uint8_t)(7 - m);
            nsum += ((double)m + 0.5) * (double)ab;
        }
        const double num = block_sum_d(nsum, red);
        if (quant) {
            if (t == 0) {
                const double b = ((double)CD + 16.0) * 0x1p-52 * num;
                float f = 1.0f;
                bool sure = false;
                if (num - b > kEps) {
                    const float f_lo = (float)(1.0 / (num + b)), f_hi = (float)(1.0 / (num - b));
                    sure = f_lo == f_hi;
                    f = f_lo;
                }
                if (!sure) {
                    double s = 0.0;
                    for (uint32_t i = 0; i < CD; ++i) {
                        const float ab = fabsf(V[i]);
                        s += ((double)init_code(best_t, ab) + 0.5) * (double)ab;
                    }
                    f = s > kEps ? (float)(1.0 / s) : 1.0f;
                }
                if (!__builtin_isfinite(f)) f = 1.0f;
                s_f[1] = f;
            }
            __syncthreads();
            ip_norm = s_f[1];
        }
    }
    const float long_scale = 2.0f * rn * ip_norm;
    for (uint32_t i = t; i < CD; i += T) {
        const float r = V[i];
        const double tv = (double)rn * (double)r;
        const double x4 = 0.5 * (double)long_scale * ((double)NIB[i] - 7.5);
        ERR[i] = (float)(tv - x4);
    }
    __syncthreads();
    uint8_t *code_row = a.codes + vi * (CD / 2);
    uint8_t *msb_row = a.gate + vi * gate_row_bytes(CD);
    for (uint32_t p = t; p < CD / 2; p += T) code_row[p] = (uint8_t)(NIB[2 * p] | (NIB[2 * p + 1] << 4));
    for (uint32_t b = t; b < CD / 8; b += T) {
        uint8_t m = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) m |= (uint8_t)((NIB[8 * b + j] >> 3) << j);
        msb_row[b] = m;
    }
    // residual block scales (max-abs start, 8 least-squares steps when MSE),
    // fp16 storage, codes against the re-loaded scale (:5940-5978)
    uint8_t *rrow = a.rcodes + vi * (CD / 2);
    if (a.block == 16) {
        // the formal block: 4 lanes per block, 4 errors each in registers
        for (uint32_t g = t >> 2; g < CD / 16; g += T / 4) {
            const uint32_t sub = t & 3u;
            const float4 v4 = ((const float4 *)(ERR + 16 * g))[sub];
            const float e[4] = {v4.x, v4.y, v4.z, v4.w};
            float mx = 0.0f;
#pragma unroll
            for (int u = 0; u < 4; ++u) { const float ab = fabsf(e[u]); mx = (mx < ab) ? ab : mx; }
            for (int off = 1; off < 4; off <<= 1) { const float o = __shfl_xor(mx, off, 4); mx = (mx < o) ? o : mx; }
            const uint16_t hb = float_to_half_ref(residual_scale_quad(e, mx, a.mse));
            if (sub == 0) {
                a.rscale[vi * a.block_count + g] = hb;
                RS[g] = hb;
            }
            const float stored = half_to_float_ref(hb);
            const bool live = stored != 0.0f && __builtin_isfinite(stored);
            uint32_t qw = 0u, rw = 0u;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int q = live ? clamp7_round((double)e[u] / (double)stored) : 0;
                qw |= ((uint32_t)q & 0xFFu) << (8 * u);
                rw |= ((uint32_t)q & 0xFu) << (4 * u);
            }
            ((uint32_t *)QV)[4 * g + sub] = qw;
            ((uint16_t *)rrow)[4 * g + sub] = (uint16_t)rw;
        }
    }
    const uint32_t bsp = a.block_count >= T ? 1u : T / a.block_count;  // threads per block: spread over the waves
    for (uint32_t blk = (t % bsp == 0 && a.block != 16) ? t / bsp : a.block_count; blk < a.block_count; blk += T / bsp) {
        const uint32_t b0 = blk * a.block;
        const uint32_t b1e = (b0 + a.block < CD) ? b0 + a.block : CD;
        float mx = 0.0f;
        for (uint32_t i = b0; i < b1e; ++i) { const float e = fabsf(ERR[i]); mx = (mx < e) ? e : mx; }
        float scale = mx <= 0.0f ? 0.0f : mx / 7.0f;
        if (a.mse && scale != 0.0f && __builtin_isfinite(scale)) {
            for (int it = 0; it < 8; ++it) {
                double num = 0.0, den = 0.0;
                for (uint32_t i = b0; i < b1e; ++i) {
                    const int q = clamp7_round((double)ERR[i] / (double)scale);
                    num += (double)ERR[i] * (double)q;
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
        const uint16_t hb = float_to_half_ref(scale);
        a.rscale[vi * a.block_count + blk] = hb;
        RS[blk] = hb;
        const float stored = half_to_float_ref(hb);
        const bool live = stored != 0.0f && __builtin_isfinite(stored);
        for (uint32_t i = b0; i < b1e; i += 2) {
            int q0 = 0, q1 = 0;
            if (live) {
                q0 = clamp7_round((double)ERR[i] / (double)stored);
                q1 = (i + 1 < b1e) ? clamp7_round((double)ERR[i + 1] / (double)stored) : 0;
            }
            QV[i] = (int8_t)q0;
            if (i + 1 < b1e) QV[i + 1] = (int8_t)q1;
            rrow[i / 2] = (uint8_t)(((uint32_t)q0 & 0xFu) | (((uint32_t)q1 & 0xFu) << 4));
        }
    }
    __syncthreads();
    const float2 h = make_float2(rn * rn, long_scale);
    // The scale factors (:5912-5938) and the decoded-residual sums
    // (:5980-5991) are sequential double recurrences in the reference whose
    // results are rounded to float. Their terms are summed here in parallel
    // (every product is exact in double); a result is taken when its error
    // bound against the sequential order, (cd + 32) units of 2^-52 of the sum
    // of |terms|, cannot move the float it rounds to. Otherwise one lane runs
    // the reference's order. Precision mode keeps its compensated FP32 sum
    // on one lane.
    const double hls = 0.5 * (double)long_scale;
    double sh = 0.0, so = 0.0, sd = 0.0, sx = 0.0, sxa = 0.0;
    for (uint32_t i = t; i < CD; i += T) {
        const float r = V[i];
        sh += 0.5 * fabs((double)r);
        so += (double)r * (r > 0.0f ? 1.0 : -1.0);
        const float st = half_to_float_ref(RS[i / a.block]);
        if (st != 0.0f && __builtin_isfinite(st)) {
            const double dr = (double)st * (double)QV[i];
            const double tx = hls * ((double)NIB[i] - 7.5) * dr;
            sd += dr * dr;
            sx += tx;
            sxa += fabs(tx);
        }
    }
    __syncthreads();
    {
        double v5[5] = {sh, so, sd, sx, sxa};
#pragma unroll
        for (int u = 0; u < 5; ++u)
            for (int off = 32; off > 0; off >>= 1) v5[u] += __shfl_xor(v5[u], off);
        double *r5 = (double *)ERR;  // the errors are consumed: 4 waves x 5 sums
        if ((t & 63u) == 0) {
#pragma unroll
            for (int u = 0; u < 5; ++u) r5[(t >> 6) * 5 + u] = v5[u];
        }
    }
    __syncthreads();
    if (t == 0) {
        const double *r5 = (const double *)ERR;
        double tot[5];
        for (int u = 0; u < 5; ++u) tot[u] = (r5[u] + r5[5 + u]) + (r5[10 + u] + r5[15 + u]);
        const double ub = ((double)CD + 32.0) * 0x1p-52;
        const double inv_sqrt_d = 1.0 / sqrt((double)CD);
        float short_scale = 0.0f, error_scale = 0.0f, fdec = 0.0f, fx4 = 0.0f;
        bool sure = false;
        if (!a.precision) {
            // short_scale falls as half_l1 grows, error_scale as o_obar grows
            const double H = tot[0], O = inv_sqrt_d * tot[1];
            const double bh = ub * H, bo = ub * O;
            const EncScales lo = enc_scales(rn, H - bh, O - bo, CD), hi = enc_scales(rn, H + bh, O + bo, CD);
            const float dlo = (float)(tot[2] - ub * tot[2]), dhi = (float)(tot[2] + ub * tot[2]);
            const float xlo = (float)(tot[3] - ub * tot[4]), xhi = (float)(tot[3] + ub * tot[4]);
            sure = lo.live == hi.live && lo.ss == hi.ss && lo.es == hi.es && dlo == dhi && xlo == xhi;
            short_scale = lo.ss;
            error_scale = lo.es;
            fdec = dlo;
            fx4 = xlo;
        }
        if (!sure) {
            double dec = 0.0, x4ip = 0.0;
            for (uint32_t blk = 0; blk < a.block_count; ++blk) {
                const float st = half_to_float_ref(RS[blk]);
                if (st == 0.0f || !__builtin_isfinite(st)) continue;
                const uint32_t b0 = blk * a.block;
                const uint32_t b1e = (b0 + a.block < CD) ? b0 + a.block : CD;
                for (uint32_t i = b0; i < b1e; ++i) {
                    const double dr = (double)st * (double)QV[i];
                    const double x4 = hls * ((double)NIB[i] - 7.5);
                    dec = fma(dr, dr, dec);
                    x4ip = fma(x4, dr, x4ip);
                }
            }
            fdec = (float)dec;
            fx4 = (float)x4ip;
            if (a.precision) {
                Comp l1c;
                for (uint32_t i = 0; i < CD; ++i) l1c.add(fabsf(V[i]));
                const float l1 = l1c.value(), half = 0.5f * l1;
                short_scale = error_scale = 0.0f;
                if (rn > 0.0f && half > 1e-12f) {
                    short_scale = 2.0f * rn / half;
                    const float sq = sqrtf((float)CD);
                    float o = l1 / sq;
                    if (!__builtin_isfinite(o)) o = 0.8f;
                    o = (1e-6f < o) ? o : 1e-6f;
                    o = (0.999999f < o) ? 0.999999f : o;
                    const float o2 = o * o;
                    float r = (1.0f - o2) / o2;
                    r = (0.0f < r) ? r : 0.0f;
                    error_scale = sqrtf(r) * (0.359375f / sq) * 2.0f * rn;
                }
            } else {
                double o_obar = 0.0, half_l1 = 0.0;
                for (uint32_t i = 0; i < CD; ++i) {
                    const float r = V[i];
                    const double sign = r > 0.0f ? 1.0 : -1.0;
                    const double rs = (double)r * sign;
                    o_obar = fma(rs, inv_sqrt_d, o_obar);
                    half_l1 += 0.5 * fabs((double)r);
                }
                const EncScales sc = enc_scales(rn, half_l1, o_obar, CD);
                short_scale = sc.ss;
                error_scale = sc.es;
            }
        }
        a.hdr[vi] = h;
        a.sfac[vi] = make_float2(short_scale, error_scale);
        a.rfac[vi] = make_float2(fdec, fx4);
        // paper-prune sidecar factors (extract_paper_prune_sidecar, :4364-4392)
        float4 f = make_float4(h.x, 0.0f, 0.0f, 0.0f);
        uint8_t valid = 0;
        if (CD > 1 && h.x > 0.0f && short_scale > 0.0f) {
            const float dn = sqrtf(h.x);
            f.y = dn;
            float al = 4.0f * dn / (short_scale * sqrtf((float)CD));
            al = (al < 1.0f) ? al : 1.0f;  // std::min(1.0f, .)
            if (al > 0.0f && __builtin_isfinite(al)) {
                const float a2 = al * al;
                f.z = short_scale;
                float r = (1.0f - a2) / a2;
                r = (0.0f < r) ? r : 0.0f;
                f.w = 2.0f * dn * sqrtf(r) / sqrtf((float)(CD - 1));
                valid = __builtin_isfinite(f.w) ? 1 : 0;
            }
        }
        gate_store_tail(a.gate + vi * gate_row_bytes(CD), CD, f, h, valid);
    }
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

namespace {

// dynamic LDS of k_encode<CD>: V, SA, SI, then the reused region (at least
// 2 KB for the compensated tree of precision mode)
static size_t encode_lds_bytes(uint32_t cd) {
    const size_t ur = qg::enc_u_bytes() + 4 * qg::kRefine;
    return (size_t)cd * 10 + qg::enc_x_bytes(cd) + (cd >= 2048 ? 0 : ur);
}

template <int CD>
static int launch_encode_cd(const qg::EncArgs *a, hipStream_t stream) {
    const size_t lds = encode_lds_bytes(CD);
    if (lds > 65536) {
        static std::once_flag once;
        static hipError_t rc = hipSuccess;
        std::call_once(once, [&] {
            rc = hipFuncSetAttribute(reinterpret_cast<const void *>(&qg::k_encode<CD>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        });
        if (rc != hipSuccess) return (int)rc;
    }
    hipLaunchKernelGGL(qg::k_encode<CD>, dim3((unsigned)a->n), dim3(qg::kEncThreads), lds, stream, *a);
    return 0;
}

}  // namespace

extern "C" {

int qg_launch_encode(const qg::EncArgs *a, hipStream_t stream) {
    if (a->n == 0) return 0;
    const unsigned lane_blocks = (unsigned)((a->n + 63) / 64);
    if (a->cid && a->centroid_count > 1) hipLaunchKernelGGL(qg::k_enc_assign, dim3(lane_blocks), dim3(64), 0, stream, *a);
    int rc;
    switch (a->cd) {
    case 64: rc = launch_encode_cd<64>(a, stream); break;
    case 128: rc = launch_encode_cd<128>(a, stream); break;
    case 256: rc = launch_encode_cd<256>(a, stream); break;
    case 512: rc = launch_encode_cd<512>(a, stream); break;
    case 1024: rc = launch_encode_cd<1024>(a, stream); break;
    case 2048: rc = launch_encode_cd<2048>(a, stream); break;
    case 4096: rc = launch_encode_cd<4096>(a, stream); break;
    default: return (int)hipErrorInvalidValue;
    }
    if (rc) return rc;
    return (int)hipGetLastError();
}

int qg_launch_pack(const qg::PackArgs *a, uint64_t n, int unpack, hipStream_t stream) {
    if (n == 0) return 0;
    if (unpack) hipLaunchKernelGGL(qg::k_unpack, dim3((unsigned)n), dim3(128), 0, stream, *a);
    else hipLaunchKernelGGL(qg::k_pack, dim3((unsigned)n), dim3(128), 0, stream, *a);
    return (int)hipGetLastError();
}

}  // extern "C"
