// qg_search.hip -- batched greedy beam search over a Vamana graph with the
// 1-bit gate, INT8 / FP32 remaining-bit distances and the residual-4 rerank,
// one workgroup of W wavefronts (2 or 4, chosen at launch) per query in a
// persistent work-queue launch; each query may carry its own search list
// size.
//
// Semantics follow search_ours_paper_active_inner + flush_distance_batch
// (experiments/02_diskann_fair/src/ours_diskann.rs:1539-1986) with beam
// width 1, start node 0 and the k-th / early stops off:
//   hop: expand the first unexpanded pool entry; its unvisited neighbours
//        (adjacency order, first occurrence) are gated against the pool as
//        it stood at the start of the hop (pruned if the pool is full, the
//        gate is valid and lower_bound > pool.back); survivors get their
//        distance (mode 2 INT8 / mode 1 FP32 reuse of the gate's short_ip,
//        mode 0 when the gate is invalid) and are inserted in order, each
//        re-checked against the current pool back (insert_candidate_sorted,
//        :1100-1126, (distance, id) order, truncated to L).
//   end:  the first min(|pool|, max(rerank, k)) entries get the FP32 4-bit
//        distance (recomputed for the INT8 codec, :1850-1866) and the
//        residual-4 interval; the answer is the top-k by (estimate, id).
// Degrees are at most 64, so a hop's fresh neighbours form one distance
// batch (DIST_BATCH_SIZE = 64), as in the reference.
//
// Work split: the serial steps of a hop (neighbour list, visited marks, the
// prune decision, the ordered insertion) run on wave 0; the gate and the
// survivors' distances run on 16-lane groups (lane j owns the reference's
// accumulator j), 4 * W candidates at a time, from MSB and code
// rows staged in LDS by gfx950's global-to-LDS loads, one memory round trip
// per batch.
#include <mutex>

#include "qg_args.h"
#include "qg_device.h"
#include "qg_learned.h"
#include "qg_space.h"

namespace qg {

constexpr uint32_t kMaxL = 4096;       // pool capacity (search list size)

__device__ __forceinline__ bool cand_less(float ad, uint32_t aid, float bd, uint32_t bid) {
    return ad < bd || (ad == bd && aid < bid);
}

// float_order_key of a number (not NaN): keys of -inf .. +inf fill
// [0x007FFFFF, 0xFF800000]; NaNs lie outside
__device__ __forceinline__ bool key_is_num(uint32_t k) { return k - 0x007FFFFFu <= 0xFF800000u - 0x007FFFFFu; }

// cand_less of two entries given as order keys (float_order_key): false
// when either distance is NaN, as the float comparisons are
__device__ __forceinline__ bool key_less(uint32_t ak, uint32_t aid, uint32_t bk, uint32_t bid) {
    return key_is_num(ak) && key_is_num(bk) && (ak < bk || (ak == bk && aid < bid));
}

// the same order without short-circuit evaluation (no branches on uniform values)
__device__ __forceinline__ bool cand_less_nb(float ad, uint32_t aid, float bd, uint32_t bid) {
    return (ad < bd) | ((ad == bd) & (aid < bid));
}

// a value every lane holds alike (read from LDS, say), moved to a scalar
// register so the compiler keeps it and what derives from it out of VGPRs
__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ float uni(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Copy rows ids[0..count) (ids in LDS) of a row-major device table (row_u4
// uint4 per row) to LDS rows 0..count) with gfx950's direct global-to-LDS
// loads, the W waves of the group taking interleaved 1 KB pieces; every
// piece is in flight at once and no registers hold the data. A piece's row
// is e / row_u4 by a multiply-high with magic = 2^32 / row_u4 rounded up
// (exact for e < 2^32 / row_u4). Lanes past the end re-read the last piece
// into stage slots below the 64-piece round-up of count * row_u4 that nobody
// reads; that round-up is at most kStageU4. The caller's barrier publishes
// the rows to the other waves.
constexpr uint32_t kStageU4 = 1152;  // 64 gate rows of cd 2048 (18 uint4 each)

__host__ __device__ constexpr uint32_t div_magic(uint32_t d) { return 0xFFFFFFFFu / d + 1u; }

template <int W>
__device__ __forceinline__ void stage_rows_issue(uint4 *dst, const uint32_t *ids, uint32_t count, uint32_t row_u4,
                                                 uint32_t magic, const uint4 *__restrict__ src, uint32_t wv,
                                                 uint32_t lane) {
    const uint32_t total = count * row_u4;
    for (uint32_t k0 = 64u * wv; k0 < total; k0 += 64u * W) {
        const uint32_t e = k0 + lane < total ? k0 + lane : total - 1u;
        const uint32_t r = __umulhi(e, magic);
        const uint4 *g = src + (uint64_t)ids[r] * row_u4 + (e - r * row_u4);
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)g,
                                         (__attribute__((address_space(3))) void *)(dst + k0), 16, 0, 0);
    }
}

template <int W>
__device__ __forceinline__ void stage_rows(uint4 *dst, const uint32_t *ids, uint32_t count, uint32_t row_u4,
                                           uint32_t magic, const uint4 *__restrict__ src, uint32_t wv,
                                           uint32_t lane) {
    stage_rows_issue<W>(dst, ids, count, row_u4, magic, src, wv, lane);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's pieces are in LDS
}

// The next frontier is the best survivor of this hop or the first
// unexpanded pool entry. Their adjacency rows and degrees go to LDS together
// with the survivors' code rows, so the prediction of the next frontier
// starts from its neighbour list instead of loading it. They live in the
// stage's last 2 KB (uint4 kAdjAt onwards), which a batch of code rows never
// reaches (at most 1024 / cch rows of cch uint4, rounded up to 64 pieces):
// the degrees in the first 32 bytes, then row r at 8 + r * stride words.
// Row r belongs to the id in lane r of idl, r < cnt (cnt <= adj_rows_fit).
// 16-byte global-to-LDS pieces when a row is a whole number of them, else
// 4-byte ones; lanes past the end stay idle. The issuing wave's vmcnt(0)
// and the caller's barrier publish them.
constexpr uint32_t kAdjAt = 1024, kAdjMax = 7;  // survivors' rows: kAdjMax - 1, then the pool entry's

__device__ __forceinline__ uint32_t adj_rows_fit(uint64_t stride) {
    const uint64_t fit = ((kStageU4 - kAdjAt) * 16u - 32u) / (4u * (stride ? stride : 1u));
    return fit < kAdjMax ? (uint32_t)fit : kAdjMax;
}

__device__ __forceinline__ void stage_adjacency(const SearchArgs &a, uint32_t idl, uint32_t cnt, uint4 *stage,
                                                uint32_t lane) {
    const uint32_t stride = (uint32_t)a.edge_stride;
    uint32_t *adj = (uint32_t *)(stage + kAdjAt);
    if ((stride & 3u) == 0u) {
        const uint32_t row_u4 = stride / 4u, total = cnt * row_u4;
        for (uint32_t k0 = 0; k0 < total; k0 += 64u) {
            const uint32_t e = k0 + lane < total ? k0 + lane : total - 1u;
            const uint32_t r = e / row_u4;
            const uint32_t id = (uint32_t)__shfl((int)idl, (int)r);
            const uint4 *g = (const uint4 *)(a.edges + (uint64_t)id * stride) + (e - r * row_u4);
            if (k0 + lane < total)
                __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)g,
                                                 (__attribute__((address_space(3))) void *)(adj + 8u + 4u * k0), 16,
                                                 0, 0);
        }
    } else {
        const uint32_t total = cnt * stride;
        for (uint32_t k0 = 0; k0 < total; k0 += 64u) {
            const uint32_t e = k0 + lane < total ? k0 + lane : total - 1u;
            const uint32_t r = e / stride;
            const uint32_t id = (uint32_t)__shfl((int)idl, (int)r);
            const uint32_t *g = a.edges + (uint64_t)id * stride + (e - r * stride);
            if (k0 + lane < total)
                __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)g,
                                                 (__attribute__((address_space(3))) void *)(adj + 8u + k0), 4, 0, 0);
        }
    }
    if (lane < cnt)
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(a.degrees + idl),
                                         (__attribute__((address_space(3))) void *)adj, 4, 0, 0);
}

// ---------------------------------------------------------------- gate filter
// The FP32 gate's float sum S_f (the query summed over a candidate's MSB
// bits: 16 lane chains of cd/16 adds, then the reduce16 tree) lies within w
// of sc * X, where X = sum m_i x_i is an exact integer sum over the
// fixed-point query x_i = rint(q_i / sc), sc = max|q| / 16256:
//   |S_f - sc X| <= sum |q_i - sc x_i| + 134 u sum |q_i| = w
// (u = 2^-24; each term passes at most cd/16 + 4 <= 132 roundings). A wave
// forms X for 16 candidates at once on the int8 matrix cores (x = 128 xh +
// xl, the halves as B columns 0 and 1, the MSB bits as 0/1 bytes in A).
// finish_gate's lower bound is monotone in short_ip = S_f - half_sum, so its
// values at the interval's ends bracket the exact one: when both are finite
// and above the pool's back (the pool full, the factors valid), the
// candidate is pruned whatever S_f is, and its float sum is not needed.
// Every other candidate gets the exact sum, so decisions, estimates and
// counters are those of the plain gate.
constexpr double kFilterX = 16256.0;  // |x| <= 127 * 128
typedef int v4i_gf __attribute__((ext_vector_type(4)));

// index of set bit number p (from 0) of m, for p < popcount(m)
__device__ __forceinline__ uint32_t nth_set_bit(uint64_t m, uint32_t p) {
    uint32_t base = 0;
#pragma unroll
    for (uint32_t w = 32; w >= 1; w >>= 1) {
        const uint64_t low = m & ((1ull << w) - 1ull);
        const uint32_t c = (uint32_t)__popcll(low);
        const bool up = p >= c;
        p = up ? p - c : p;
        m = up ? (m >> w) : low;
        base += up ? w : 0u;
    }
    return base;
}

// The query's fixed-point form (xh, xl in LDS) and per-wave partial sums of
// |q_i - sc x_i| and |q_i|; s_qmx holds each wave's largest |q| bits. Every
// thread of the group calls it after the barrier that publishes q and s_qmx.
template <int W>
__device__ void gate_filter_query(const float *q, int8_t *xh, int8_t *xl, uint32_t cd, const uint32_t *s_qmx,
                                  double *s_qe, double *s_qa, uint32_t t) {
    constexpr uint32_t T = 64u * W;
    uint32_t mb = 0u;
    for (int k = 0; k < W; ++k) mb = max(mb, s_qmx[k]);
    const bool ok = mb > 0u && mb < 0x7F800000u;
    const double sc = ok ? (double)__uint_as_float(mb) / kFilterX : 1.0;
    double e = 0.0, qa = 0.0;
    for (uint32_t i = t; i < cd; i += T) {
        const double v = (double)q[i];
        const double xr = ok ? rint(v / sc) : 0.0;
        const int x = (int)xr;
        const int h = (x + 64) >> 7;  // floor((x + 64) / 128): xh in [-127, 127], xl in [-64, 63]
        xh[i] = (int8_t)h;
        xl[i] = (int8_t)(x - 128 * h);
        e += fabs(v - sc * xr);
        qa += fabs(v);
    }
    for (int off = 32; off > 0; off >>= 1) {
        e += __shfl_xor(e, off);
        qa += __shfl_xor(qa, off);
    }
    if ((t & 63u) == 0u) {
        s_qe[t >> 6] = e;
        s_qa[t >> 6] = qa;
    }
}

// The scale and the half-width w from the partials (one thread, after the
// barrier that publishes them); the inflations cover the double sums' own
// rounding.
template <int W>
__device__ void gate_filter_constants(const uint32_t *s_qmx, const double *s_qe, const double *s_qa, double *fsc,
                                      double *fw, uint32_t *fok) {
    uint32_t mb = 0u;
    double e = 0.0, qa = 0.0;
    for (int k = 0; k < W; ++k) {
        mb = max(mb, s_qmx[k]);
        e += s_qe[k];
        qa += s_qa[k];
    }
    const double mx = (double)__uint_as_float(mb);
    const bool ok = mb > 0u && mb < 0x7F800000u && __builtin_isfinite(e) && __builtin_isfinite(qa);
    *fsc = ok ? mx / kFilterX : 1.0;
    *fw = e * (1.0 + 0x1p-30) + qa * (1.0 + 0x1p-30) * (134.0 * 0x1p-24) + mx * 0x1p-38;
    *fok = ok ? 1u : 0u;
}

// The filter over the hop's nf fresh neighbours (rows staged, frow[c] their
// stage rows): wave wv takes candidates 16 wv .. 16 wv + 15, then every
// 16 W-th tile. Lane (g, r) supplies dims 64 k + 16 g .. + 15 of candidate
// c0 + r (A) and of xh (r = 0) or xl (r = 1) (B); accumulator j holds
// candidate c0 + 4 g + j against column r. Writes f_sure[c] for every c <
// nf, and for a sure prune f_lb[c] = +inf, f_valid[c] = 1 (the prune test
// then removes it).
template <int W>
__device__ __forceinline__ void gate_filter_tiles(const uint4 *stage, const uint8_t *frow, uint32_t gch,
                                                  uint32_t cd, uint32_t nf, const int8_t *xh, const int8_t *xl,
                                                  double sc, double w, float back, const QScalars &s, float eps0,
                                                  uint8_t *f_sure, float *f_lb, uint8_t *f_valid, uint32_t wv,
                                                  uint32_t lane) {
    const uint32_t g = lane >> 4, r = lane & 15u;
    const uint32_t nk = cd / 64;
    for (uint32_t c0 = 16u * wv; c0 < nf; c0 += 16u * W) {
        const uint32_t ca = c0 + r < nf ? c0 + r : c0;
        const uint16_t *abits = (const uint16_t *)((const uint8_t *)(stage + (uint32_t)frow[ca] * gch) + 2u * g);
        const int8_t *bsrc = (r == 0u ? xh : xl) + 16u * g;
        v4i_gf acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0}, bq = {0, 0, 0, 0};
        auto step = [&](uint32_t kb, v4i_gf &acc) {
            const uint32_t bits = abits[4u * kb];
            v4i_gf av;
#pragma unroll
            for (int i = 0; i < 4; ++i) av[i] = (int)nibble_bytes(__builtin_amdgcn_ubfe(bits, 4u * i, 4u));
            if (r < 2u) bq = *(const v4i_gf *)(bsrc + 64u * kb);
            acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, bq, acc, 0, 0, 0);
        };
#pragma unroll 1
        for (uint32_t kb = 0; kb + 1 < nk; kb += 2) {
            step(kb, acc0);
            step(kb + 1, acc1);
        }
        if (nk & 1u) step(nk - 1, acc0);
        const v4i_gf acc = acc0 + acc1;
        int hx = 0, lx = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int hj = __shfl(acc[j], (int)(lane & ~15u));
            const int lj = __shfl(acc[j], (int)((lane & ~15u) | 1u));
            if (r == (uint32_t)j) {
                hx = hj;
                lx = lj;
            }
        }
        const uint32_t c = c0 + 4u * g + r;
        if (r < 4u && c < nf) {
            const uint8_t *row = (const uint8_t *)(stage + (uint32_t)frow[c] * gch);
            const float4 f = gate_factors(row, cd);
            const uint8_t fv = gate_valid(row, cd);
            const double A = sc * (double)(128 * hx + lx);
            const double slack = w + fabs(A) * 0x1p-50;
            const float lo = __double2float_rd(A - slack), hi = __double2float_ru(A + slack);
            const PaperEstimate e0 = finish_gate(lo - s.half_sum, f, fv, s.query_norm, s.query_norm_sqr, eps0);
            const PaperEstimate e1 = finish_gate(hi - s.half_sum, f, fv, s.query_norm, s.query_norm_sqr, eps0);
            const bool sure = fv != 0 && __builtin_isfinite(f.z) && e0.valid && e1.valid &&
                              fminf(e0.lower_bound, e1.lower_bound) > back;
            f_sure[c] = sure ? 1 : 0;
            if (sure) {
                f_lb[c] = __builtin_huge_valf();
                f_valid[c] = 1;
            }
        }
    }
}

// The learned depth after a query's traversal (the LEARNED search kernel):
// the feature rows of the rc pool entries, the MLP's predicted errors
// (16-row tiles across the waves) and the depth rule, all in the stage,
// which is idle after the traversal: features 112 x 64 B, each wave's hidden
// layer 2 KB, 112 predictions. Returns the depth (at most rc).
// Kernel-argument fields the hop loop does not use (outputs, per-query
// inputs, the rerank's and the learned depth's tables), read from the
// kernarg segment where they are needed through a pointer the compiler
// cannot hoist, so they do not occupy scalar registers across the traversal.
// The segment holds the SearchArgs record at offset 0.
__device__ __forceinline__ const __attribute__((address_space(4))) unsigned char *cold_kernarg() {
    const __attribute__((address_space(4))) unsigned char *p =
        (const __attribute__((address_space(4))) unsigned char *)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(p));
    return p;
}
#define COLD(field) \
    (*(const __attribute__((address_space(4))) decltype(SearchArgs::field) *)(cold_kernarg() + \
                                                                             offsetof(SearchArgs, field)))

template <int W>
__device__ __forceinline__ uint32_t search_learned_depth(const SearchArgs &a, uint32_t qi, uint32_t rc,
                                                                   const float *pd, const uint32_t *pid,
                                                                   const QScalars &s, uint4 *stage, uint32_t *s_out) {
    constexpr uint32_t T = 64u * W;
    // the thread id passes through an opaque move, so that the per-lane
    // addresses of this section are computed here and not hoisted out of the
    // query loop, where they would hold registers through the traversal
    uint32_t t = threadIdx.x;
    asm volatile("" : "+v"(t));
    const uint32_t lane = t & 63u, wv = t >> 6;
    uint16_t *feat = (uint16_t *)stage;
    uint16_t *sh1 = (uint16_t *)((unsigned char *)stage + 7168) + wv * 16 * kH;
    float *pred = (float *)((unsigned char *)stage + 7168 + 4 * 2048);
    for (uint32_t i = t; i < rc; i += T) {
        const uint32_t id = pid[i];
        uint32_t w[kFP / 2];
        learned_feature_row(i, rc, pd, s, COLD(ix.hdr)[id], COLD(ix.sfac)[id], COLD(rfac)[id], COLD(rstats)[id], a.cd, w);
        uint4 *o = (uint4 *)(feat + i * kFP);
#pragma unroll
        for (int k = 0; k < kFP / 8; ++k) o[k] = make_uint4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]);
    }
    __syncthreads();
    const uint32_t tiles = (rc + 15) / 16;
    for (uint32_t tl = wv; tl < tiles; tl += W) {
        float y[4];
        const uint32_t rows = rc - 16 * tl < 16 ? rc - 16 * tl : 16;
        mlp_forward_tile_packed(feat + 16 * tl * kFP, rows, COLD(mlp), sh1, y, (int)lane);
        if ((lane & 15u) == 0) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t row = 16 * tl + 4 * (lane >> 4) + (uint32_t)r;
                if (row < rc) pred[row] = y[r];
            }
        }
    }
    __syncthreads();
    if (wv == 0) {
        const uint32_t d = learned_depth(rc, pd, [&](uint32_t j) { return pred[j]; }, s.query_norm_sqr, COLD(kmult),
                                         COLD(min_depth), lane);
        if (lane == 0) {
            *s_out = d;
            if (COLD(out_depth)) COLD(out_depth)[qi] = d;
        }
    }
    __syncthreads();
    return *s_out;
}

// 4 waves per SIMD (<= 128 VGPRs): 1024 workgroups fit the chip at once.
template <int W, bool LEARNED, bool FILTER>
__global__ void __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(4))) k_search(SearchArgs a) {
    constexpr uint32_t T = 64u * W, G = 4u * W;
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t cd = a.cd;
    // query: FP32 residual, then one codec-specific region: with the gate
    // filter (FILTER: FP32 codec, code dimension 2048 and up) its
    // fixed-point query in 2 cd bytes; otherwise 4 cd bytes holding the
    // pair-interleaved FP32 query of the gate, or the INT8 codes
    constexpr uint32_t kQ = FILTER ? 6u : 8u;                   // query bytes per dimension
    float *q = (float *)smem;                                   // cd
    int8_t *xh = (int8_t *)(smem + 4 * cd);                     // cd     (FILTER)
    int8_t *xl = (int8_t *)(smem + 5 * cd);                     // cd     (FILTER)
    float2 *qp2 = (float2 *)(smem + 4 * cd);                    // cd / 2 (FP32 codec, no FILTER)
    int8_t *x8 = (int8_t *)(smem + 4 * cd);                     // cd     (INT8 codec)
    int8_t *x8p = (int8_t *)(smem + 5 * cd);                    // cd     (INT8 codec)
    float *pd = (float *)(smem + kQ * cd);                      // a.L
    uint32_t *pid = (uint32_t *)(smem + kQ * cd + 4 * a.L);     // a.L
    uint8_t *pexp = (uint8_t *)(smem + kQ * cd + 8 * a.L);      // a.L (+pad)
    // gate rows / code rows of a batch; rerank block sums after the traversal
    __shared__ uint4 stage[kStageU4];
    __shared__ uint32_t fresh[64], surv[64], sid[64];
    __shared__ uint8_t frow[64];  // stage row of fresh neighbour c
    __shared__ uint8_t f_sure[64];  // the filter proved fresh neighbour c pruned
    // the gate filter's query constants: per-wave partials, then the scale
    // and the interval half-width (FP32 codec)
    __shared__ uint32_t s_qmx[W];
    __shared__ double s_qe[W], s_qa[W];
    __shared__ double s_fsc, s_fw;
    __shared__ uint32_t s_fok;
    __shared__ float f_lb[64], f_sip[64], s_d[64];
    __shared__ float2 f_h[64];  // header of fresh neighbour c (from its gate row)
    __shared__ uint8_t f_valid[64];
    __shared__ uint32_t s_qi, s_nf, s_ns, s_len;
    // next hop's neighbour list and visited words, loaded during this hop's
    // insertion for the predicted next frontier s_sp_node
    __shared__ uint32_t sp_nb[64];
    __shared__ unsigned long long s_sp_fresh;  // lanes whose neighbour is fresh (first occurrence, unvisited)
    __shared__ uint32_t s_sp_node, s_sp_bad, s_spa_id, s_hit;
    __shared__ float s_spa_d;
    const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6, grp = t >> 4;
    const int j16 = (int)(t & 15u);
    const uint32_t gch = (uint32_t)gate_row_bytes(cd) / 16;  // uint4 per gate row
    const uint32_t gmagic = div_magic(gch);
    const uint32_t cch = cd / 32;   // uint4 per code row
    const uint32_t cmagic = div_magic(cch);
    const uint32_t per = 1024u / cch;
    const uint32_t arows = adj_rows_fit(a.edge_stride);  // adjacency rows the stage's tail holds
    // the point count and the adjacency row stride fit 32 bits (the launcher
    // takes at most 2^32 - 2 points)
    const uint32_t npts = (uint32_t)a.n, estr = (uint32_t)a.edge_stride;
    uint32_t *visited = COLD(visited) + (uint64_t)blockIdx.x * COLD(visited_words);
    bool first_query = true;
    constexpr bool use_filter = FILTER;  // the FP32 gate's filter (below)

    for (;;) {
        if (t == 0) s_qi = atomicAdd(COLD(counter), 1u);
        __syncthreads();
        const uint32_t qi = uni(s_qi);
        __syncthreads();
        if (qi >= COLD(qv.nq)) break;
        // this query's search list size: its own (at least k) or the launch's
        uint32_t L = COLD(L);
        if (COLD(lq)) {
            const uint32_t l = COLD(lq)[qi];
            L = l < COLD(k) ? COLD(k) : (l > COLD(L) ? COLD(L) : l);
        }
        // stage the prepared query (and its largest |q| bits), clear this
        // slot's visited bits
        uint32_t qmx = 0u;
        for (uint32_t i = t; i < cd / 4; i += T) {
            const float4 v = ((const float4 *)(COLD(qv.q) + (uint64_t)qi * cd))[i];
            ((float4 *)q)[i] = v;
            const uint32_t m0 = __float_as_uint(v.x) & 0x7FFFFFFFu, m1 = __float_as_uint(v.y) & 0x7FFFFFFFu;
            const uint32_t m2 = __float_as_uint(v.z) & 0x7FFFFFFFu, m3 = __float_as_uint(v.w) & 0x7FFFFFFFu;
            qmx = max(qmx, max(max(m0, m1), max(m2, m3)));
        }
        if (a.codec == 3) {
            for (uint32_t i = t; i < cd / 4; i += T) {
                ((uint32_t *)x8)[i] = ((const uint32_t *)(COLD(qv.x8) + (uint64_t)qi * cd))[i];
                ((uint32_t *)x8p)[i] = ((const uint32_t *)(COLD(qv.x8p) + (uint64_t)qi * cd))[i];
            }
        } else if (use_filter) {
            for (int off = 32; off > 0; off >>= 1) qmx = max(qmx, (uint32_t)__shfl_xor((int)qmx, off));
            if (lane == 0) s_qmx[wv] = qmx;
        } else {
            for (uint32_t i = t; i < cd / 2; i += T) {
                const float *qq = COLD(qv.q) + (uint64_t)qi * cd + 32 * (i >> 4) + (i & 15u);
                qp2[i] = make_float2(qq[0], qq[16]);
            }
        }
        // the launch zeroed every slot's bitset: only a slot's later queries
        // clear the marks of the one before
        if (!first_query)
            for (uint32_t i = t; i < COLD(visited_words) / 4; i += T) ((uint4 *)visited)[i] = make_uint4(0u, 0u, 0u, 0u);
        first_query = false;
        if (t == 0) s_sp_node = 0xFFFFFFFFu;
        const QScalars s = COLD(qv.qs)[qi];
        // the bitset belongs to this workgroup alone: the barrier's
        // workgroup-scope release (vmcnt(0)) puts the clear in L2 before any
        // of the workgroup's atomics on it; a device-scope fence here would
        // write back and invalidate the XCD's L2 once per query
        __syncthreads();
        if (use_filter) gate_filter_query<W>(q, xh, xl, cd, s_qmx, s_qe, s_qa, t);

        uint32_t len = 0;
        uint32_t hops = 0, gates = 0, dists = 1, pruned = 0, rchecked = 0, rpruned = 0;
        // start node 0 with its full FP32 distance
        if (npts > 0 && L > 0 && wv == 0) {
            if (lane == 0) sid[0] = 0u;
            wave_sync_lds();
            stage_rows<1>(stage, sid, 1, cch, cmagic, (const uint4 *)a.ix.codes, 0u, lane);
            wave_sync_lds();
            if (grp == 0) {
                const float d0 = asym_group(q, (const uint8_t *)stage, cd, COLD(ix.hdr)[0], s, j16);
                if (j16 == 0) {
                    atomicOr(&visited[0], 1u);
                    pd[0] = d0;
                    pid[0] = 0;
                    pexp[0] = 0;
                }
            }
        }
        len = (npts > 0 && L > 0) ? 1 : 0;
        __syncthreads();
        if (use_filter && t == 0) gate_filter_constants<W>(s_qmx, s_qe, s_qa, &s_fsc, &s_fw, &s_fok);

        for (;;) {
            // the thread id through an opaque move: addresses derived from the
            // lane are formed inside the hop, not kept (and spilled) across it
            uint32_t t_hop = threadIdx.x;
            asm volatile("" : "+v"(t_hop));
            const uint32_t t = t_hop, lane = t_hop & 63u, grp = t_hop >> 4;
            const int j16 = (int)(t_hop & 15u);
            // first unexpanded entry (every wave reads the same pool)
            uint32_t f = 0xFFFFFFFFu;
            for (uint32_t c0 = 0; c0 < len; c0 += 64u) {
                const uint32_t i = c0 + lane;
                const uint64_t m = __ballot(i < len && !pexp[i]);
                if (m) { f = c0 + (uint32_t)__builtin_ctzll(m); break; }
            }
            if (f == 0xFFFFFFFFu) break;
            const uint32_t node = uni(pid[f]);
            __syncthreads();
            ++hops;
            if (wv == 0) {
                // wave 0 runs each hop's serial steps (neighbour step, prune,
                // insertion), which hold up the whole group: from its first
                // hop on it issues ahead of the other queries' waves on its
                // SIMD. It keeps the priority between its steps (dropping it
                // let other queries' raised waves hold it back); raising it
                // for the query setup as well measured slower.
                __builtin_amdgcn_s_setprio(3);
                if (lane == 0) pexp[f] = 1;
                // the neighbour list with its first occurrences and their visited
                // test: worked out during the last hop's insertion when this node
                // was the predicted frontier (the visited bits have not changed
                // since: only this step of a hop writes them), else here
                bool is_fresh = false;
                uint32_t nb;
                if (s_sp_node == node) {
                    nb = sp_nb[lane];
                    is_fresh = ((s_sp_fresh >> lane) & 1ull) != 0ull;
                    if (lane == 0 && s_sp_bad) atomicOr(a.err, 1u);  // out-of-range neighbour
                    if (is_fresh) (void)atomicOr(&visited[nb >> 5], 1u << (nb & 31u));  // no reply needed
                } else {
                    const uint32_t deg = a.degrees[node];
                    const uint32_t raw_nb = lane < estr ? a.edges[(uint64_t)node * estr + lane]
                                                                 : 0xFFFFFFFFu;
                    nb = lane < deg ? raw_nb : 0xFFFFFFFFu;
                    // first occurrence within the list, then test-and-set visited
                    bool first_occ = lane < deg;
                    for (int jj = 0; jj < 63; ++jj) {
                        const uint32_t other = (uint32_t)__builtin_amdgcn_readlane((int)nb, jj);
                        if ((uint32_t)jj < lane && other == nb) first_occ = false;
                    }
                    if (first_occ && nb >= npts) {  // out-of-range neighbour: the reference stops with an error
                        atomicOr(a.err, 1u);
                        first_occ = false;
                    }
                    if (first_occ) {
                        const uint32_t bit = 1u << (nb & 31u);
                        const uint32_t old = atomicOr(&visited[nb >> 5], bit);
                        is_fresh = (old & bit) == 0u;
                    }
                }
                const uint64_t fm = __ballot(is_fresh);
                const bool hit = s_sp_node == node;  // fresh list, factors and MSB rows already staged
                if (!hit && is_fresh) {
                    const uint32_t r = (uint32_t)__popcll(fm & ((1ull << lane) - 1ull));
                    fresh[r] = nb;
                    frow[r] = (uint8_t)r;
                }
                if (lane == 0) { s_nf = (uint32_t)__popcll(fm); s_hit = hit ? 1u : 0u; }
            }
            __syncthreads();
            const uint32_t nf = uni(s_nf);
            if (nf == 0) continue;
            gates += nf;
            // gate: the fresh neighbours' gate rows (MSB plane, factors, header,
            // validity) in one round trip; when the frontier was predicted,
            // every listed neighbour's row was staged during the last
            // insertion and frow[] places the fresh ones
            if (!s_hit) {
                stage_rows<W>(stage, fresh, nf, gch, gmagic, (const uint4 *)a.ix.gate, wv, lane);
                __syncthreads();
            }
            // FILTER with the pool full: the filter first (its sure prunes
            // need no float sum), then the exact sums of the candidates it
            // left open. At code dimensions below 2048 the exact sums are
            // short and the filter's extra barrier costs more than it saves
            // (1M x 1024, 20k queries with lists of median 400: 270 ms
            // without, 285 ms with), so that kernel is built without it.
            const bool full_h = len >= L;
            const bool filter = use_filter && full_h && s_fok != 0u;
            uint32_t nx = nf;        // candidates scored exactly
            uint64_t open = ~0ull;   // which ones (bit c: fresh neighbour c)
            if (filter) {
                gate_filter_tiles<W>(stage, frow, gch, cd, nf, xh, xl, s_fsc, s_fw, pd[len - 1u], s, a.eps0, f_sure,
                                     f_lb, f_valid, wv, lane);
                __syncthreads();
                open = __ballot(lane < nf && f_sure[lane] == 0);
                nx = (uint32_t)__popcll(open);
            }
            if (filter && nx <= G) {
                // one open candidate per group: a single chain
                if (grp < nx) {
                    const uint32_t c = nth_set_bit(open, grp);
                    const uint8_t *row = (const uint8_t *)(stage + (uint32_t)frow[c] * gch);
                    const float acc1 = gate_fp32_group1(q, (const uint4 *)row, cd, j16);
                    const PaperEstimate e = finish_gate(group_reduce16(acc1) - s.half_sum, gate_factors(row, cd),
                                                        gate_valid(row, cd), s.query_norm, s.query_norm_sqr, a.eps0);
                    if (j16 == 0) {
                        f_lb[c] = e.lower_bound;
                        f_sip[c] = e.short_ip;
                        f_valid[c] = e.valid;
                        f_h[c] = gate_hdr(row, cd);
                    }
                }
                nx = 0;  // done
            }
            // each group scores open candidates grp, grp + G, grp + 2G, grp + 3G
            // (FP32: in one pass, four independent chains; slots past the
            // count re-read a live row and are dropped)
            for (uint32_t p0 = grp; p0 < nx; p0 += 4u * G) {
                uint32_t cc[4];
                const uint4 *rr[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t p = p0 + (uint32_t)k * G;
                    const uint32_t pc = p < nx ? p : p0;
                    const uint32_t c = filter ? nth_set_bit(open, pc) : pc;
                    cc[k] = p < nx ? c : nf;
                    rr[k] = stage + (uint32_t)frow[c] * gch;
                }
                if (a.codec == 3) {
                    // integer sums: one candidate at a time, unrolled
                    for (int k = 0; k < 4; ++k) {
                        if (cc[k] >= nf) continue;
                        const uint8_t *row = (const uint8_t *)rr[k];
                        const PaperEstimate e = gate_int8_group(x8, row, cd, gate_factors(row, cd),
                                                                gate_valid(row, cd), s, a.eps0, j16);
                        if (j16 == 0) {
                            f_lb[cc[k]] = e.lower_bound;
                            f_sip[cc[k]] = e.short_ip;
                            f_valid[cc[k]] = e.valid;
                            f_h[cc[k]] = gate_hdr(row, cd);
                        }
                    }
                } else {
                    float acc[4];
                    if constexpr (FILTER) gate_fp32_group4(q, rr[0], rr[1], rr[2], rr[3], cd, j16, acc);
                    else gate_fp32_group4_pairs(qp2, rr[0], rr[1], rr[2], rr[3], cd, j16, acc);
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        if (cc[k] >= nf) continue;
                        const uint8_t *row = (const uint8_t *)rr[k];
                        const PaperEstimate e = finish_gate(group_reduce16(acc[k]) - s.half_sum, gate_factors(row, cd),
                                                            gate_valid(row, cd), s.query_norm, s.query_norm_sqr,
                                                            a.eps0);
                        if (j16 == 0) {
                            f_lb[cc[k]] = e.lower_bound;
                            f_sip[cc[k]] = e.short_ip;
                            f_valid[cc[k]] = e.valid;
                            f_h[cc[k]] = gate_hdr(row, cd);
                        }
                    }
                }
            }
            __syncthreads();
            // prune against the pool as it stood at the start of the hop; wave 1
            // notes the first unexpanded pool entry for the prediction of the
            // next frontier (the pool is not written until the insertion)
            if (wv == 1) {
                uint32_t fa = 0xFFFFFFFFu;
                for (uint32_t c0 = 0; c0 < len; c0 += 64u) {
                    const uint64_t m = __ballot(c0 + lane < len && !pexp[c0 + lane]);
                    if (m) { fa = c0 + (uint32_t)__builtin_ctzll(m); break; }
                }
                if (lane == 0) {
                    s_spa_id = fa == 0xFFFFFFFFu ? 0xFFFFFFFFu : pid[fa];
                    s_spa_d = fa == 0xFFFFFFFFu ? 3.402823466e38f : pd[fa];
                }
            }
            if (wv == 0) {
                // this hop's visited marks are complete before any later read
                __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
                const bool full0 = len >= L;
                const float back0 = full0 ? pd[len - 1] : 3.402823466e38f;
                const bool survive = lane < nf && !(full0 && f_valid[lane] && f_lb[lane] > back0);
                const uint64_t sm = __ballot(survive);
                if (survive) {
                    const uint32_t r = (uint32_t)__popcll(sm & ((1ull << lane) - 1ull));
                    surv[r] = lane;
                    sid[r] = fresh[lane];
                }
                if (lane == 0) s_ns = (uint32_t)__popcll(sm);
            }
            __syncthreads();
            const uint32_t ns = uni(s_ns);
            pruned += nf - ns;
            // survivors' distances: code rows in one round trip per batch of
            // `per` rows (the headers came with the gate rows), one 16-lane
            // group per candidate
            // (with the first batch, the last wave stages the adjacency rows of
            // the first survivors, rows 0 .. nadj - 1, and of the first
            // unexpanded pool entry, row nadj)
            const uint32_t nadj = arows == 0u ? 0u : (ns < arows - 1u ? ns : arows - 1u);
            const bool adj_spa = ns > 0 && arows > 0 && s_spa_id != 0xFFFFFFFFu;
            for (uint32_t b0 = 0; b0 < ns; b0 += per) {
                const uint32_t nb_rows = ns - b0 < per ? ns - b0 : per;
                stage_rows_issue<W>(stage, sid + b0, nb_rows, cch, cmagic, (const uint4 *)a.ix.codes, wv, lane);
                if (b0 == 0 && wv == W - 1 && arows > 0) {
                    const uint32_t idl = lane < nadj ? sid[lane] : s_spa_id;
                    stage_adjacency(a, idl, nadj + (adj_spa ? 1u : 0u), stage, lane);
                }
                __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
                __syncthreads();
                for (uint32_t b = grp; b < nb_rows; b += G) {
                    const uint32_t slot = surv[b0 + b];
                    const uint8_t *row8 = (const uint8_t *)(stage + b * cch);
                    const float2 h = f_h[slot];
                    float d;
                    if (!f_valid[slot]) d = asym_group(q, row8, cd, h, s, j16);
                    else if (a.codec == 3) d = rem_int8_group(x8p, row8, cd, h, s, f_sip[slot], j16);
                    else d = rem_fp32_group(q, row8, cd, h, s, f_sip[slot], j16);
                    if (j16 == 0) s_d[b0 + b] = d;
                }
                __syncthreads();
            }
            dists += ns;
            // wave 1, while wave 0 inserts: the next frontier is the better of
            // the first unexpanded entry and the best survivor (a survivor the
            // insertion prunes by its bound makes this a miss, handled by
            // loading again); load its adjacency row and visited words
            if (wv == 1) {
                // the prediction runs beside the insertion on the hop's
                // critical path: the same issue priority as wave 0's steps
                __builtin_amdgcn_s_setprio(3);
                float bd = lane < ns ? s_d[lane] : 3.402823466e38f;
                uint32_t bid = lane < ns ? sid[lane] : 0xFFFFFFFFu;
                uint32_t bslot = lane;  // the best survivor's index in sid[]
                bool bv = lane < ns;
                for (int off = 32; off > 0; off >>= 1) {
                    const float od = __shfl_xor(bd, off);
                    const uint32_t oid = (uint32_t)__shfl_xor((int)bid, off);
                    const uint32_t oslot = (uint32_t)__shfl_xor((int)bslot, off);
                    const bool ov = __shfl_xor((int)bv, off) != 0;
                    if (ov && (!bv || cand_less(od, oid, bd, bid))) { bd = od; bid = oid; bslot = oslot; bv = true; }
                }
                const uint32_t sp_aid = s_spa_id;
                const bool sp_a = sp_aid != 0xFFFFFFFFu;
                uint32_t nxt = 0xFFFFFFFFu;
                // row of the next frontier's adjacency in adj[] (staged with the
                // survivors' code rows when ns > 0), or none
                uint32_t arow = 0xFFFFFFFFu;
                if (bv && (!sp_a || cand_less(bd, bid, s_spa_d, sp_aid))) {
                    nxt = bid;
                    if (bslot < nadj) arow = bslot;
                } else if (sp_a) {
                    nxt = sp_aid;
                    if (adj_spa) arow = nadj;
                }
                if (nxt != 0xFFFFFFFFu && nxt < npts) {
                    uint32_t deg2, raw2 = 0xFFFFFFFFu;
                    if (arow != 0xFFFFFFFFu) {
                        const uint32_t *adj = (const uint32_t *)(stage + kAdjAt);
                        deg2 = adj[arow];
                        if (lane < estr) raw2 = adj[8u + arow * (uint32_t)estr + lane];
                    } else {
                        deg2 = a.degrees[nxt];
                        if (lane < estr) raw2 = a.edges[(uint64_t)nxt * estr + lane];
                    }
                    const uint32_t nb2 = lane < deg2 ? raw2 : 0xFFFFFFFFu;
                    uint32_t w2 = 0u;
                    if (nb2 < npts)
                        w2 = __hip_atomic_load(&visited[nb2 >> 5], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    // first occurrence within the list (while the loads fly): a
                    // 2^17-bit hash filter in the stage (idle during the
                    // insertion) clears lists without repeats; a possible repeat
                    // (or a collision, about 1.5% of 64-entry lists) takes the
                    // exact readlane pass
                    bool first2 = lane < deg2;
                    {
                        uint32_t *filt = (uint32_t *)stage;
                        const uint32_t h = (nb2 * 0x9E3779B1u) >> 15;
                        if (first2) filt[h >> 5] = 0u;
                        wave_sync_lds();
                        bool seen = false;
                        if (first2) seen = (atomicOr(&filt[h >> 5], 1u << (h & 31u)) & (1u << (h & 31u))) != 0u;
                        if (__ballot(seen)) {
                            for (int jj = 0; jj < 63; ++jj) {
                                const uint32_t other = (uint32_t)__builtin_amdgcn_readlane((int)nb2, jj);
                                if ((uint32_t)jj < lane && other == nb2) first2 = false;
                            }
                        }
                    }
                    // the next hop's gate rows: the row of every listed
                    // neighbour, in list order, in flight together with the
                    // visited words (the stage is idle during the insertion and
                    // the filter above is done with it; an out-of-range id
                    // loads row 0 and is flagged below)
                    sp_nb[lane] = nb2 < npts ? nb2 : 0u;
                    if constexpr (FILTER) {
                        wave_sync_lds();
                        stage_rows<1>(stage, sp_nb, deg2 < 64u ? deg2 : 64u, gch, gmagic, (const uint4 *)a.ix.gate,
                                      0u, lane);
                    }
                    const bool bad2 = first2 && nb2 >= npts;
                    const bool fresh2 = first2 && !bad2 && (w2 & (1u << (nb2 & 31u))) == 0u;
                    const uint64_t fm2 = __ballot(fresh2);
                    const uint64_t bm2 = __ballot(bad2);
                    // the fresh list and where each fresh neighbour's row sits
                    if (fresh2) {
                        const uint32_t r = (uint32_t)__popcll(fm2 & ((1ull << lane) - 1ull));
                        fresh[r] = nb2;
                        frow[r] = (uint8_t)(FILTER ? lane : r);
                    }
                    if constexpr (!FILTER) {
                        // the fresh rows only, compacted, once the visited
                        // words are in
                        wave_sync_lds();
                        stage_rows<1>(stage, fresh, (uint32_t)__popcll(fm2), gch, gmagic, (const uint4 *)a.ix.gate,
                                      0u, lane);
                    }
                    if (lane == 0) { s_sp_node = nxt; s_sp_fresh = fm2; s_sp_bad = bm2 != 0ull ? 1u : 0u; }
                } else if (lane == 0) {
                    s_sp_node = 0xFFFFFFFFu;
                }
                __builtin_amdgcn_s_setprio(0);
            }
            // ordered insertion with the re-check against the current back
            // (insert_candidate_sorted in survivor order), on wave 0. The
            // decisions only need the pool's back, which is the larger of the
            // last retained pool entry and the last accepted survivor: the
            // pool's last 64 entries and the accepted survivors (sorted) sit in
            // lane registers, and the pool is rewritten once at the end.
            if (wv == 0) {
                const uint32_t tb = len > 64u ? len - 64u : 0u;  // pool tail window [tb, len)
                const float bt_d = tb + lane < len ? pd[tb + lane] : 0.0f;
                const uint32_t bt_id = tb + lane < len ? pid[tb + lane] : 0u;
                const uint32_t my_slot = lane < ns ? surv[lane] : 0u;
                const float my_d = lane < ns ? s_d[lane] : 0.0f;
                const uint32_t my_id = lane < ns ? sid[lane] : 0u;
                const float my_lb = f_lb[my_slot];
                const uint32_t my_v = f_valid[my_slot];
                uint32_t e = len, na = 0;  // retained pool entries, accepted survivors
                float a_d = 0.0f;
                uint32_t a_id = 0xFFFFFFFFu;
                // Branch-free: every operand is read (lane indices clamped into
                // range) and the decision, the counters and the accepted-list
                // update are selects, so the loop body has no data-dependent
                // branch. When the pool is full, the back is the larger of the
                // last retained pool entry (window lane e - 1 - tb) and the
                // last accepted survivor (lane na - 1).
                uint32_t npruned = 0, nchecked = 0;
                // The decisions on order keys (unsigned integers in float
                // order, both zeros equal, NaN outside the range of the
                // others and comparing false as a float would): every
                // comparison and the loop state are scalar, and the vector
                // part (the accepted list's sorted insert) runs only for a
                // taken survivor.
                {
                    const uint32_t my_k = float_order_key(my_d), my_lbk = float_order_key(my_lb);
                    const uint32_t bt_k = float_order_key(bt_d);
                    uint32_t a_k = 0u;
                    for (uint32_t b = 0; b < ns; ++b) {
                        const uint32_t kd = (uint32_t)__builtin_amdgcn_readlane((int)my_k, (int)b);
                        const uint32_t id = (uint32_t)__builtin_amdgcn_readlane((int)my_id, (int)b);
                        const uint32_t klb = (uint32_t)__builtin_amdgcn_readlane((int)my_lbk, (int)b);
                        const uint32_t v = (uint32_t)__builtin_amdgcn_readlane((int)my_v, (int)b);
                        const int pi = e > 0 ? (int)(e - 1 - tb) : 0;
                        const int ai = na > 0 ? (int)(na - 1) : 0;
                        const uint32_t pbk = (uint32_t)__builtin_amdgcn_readlane((int)bt_k, pi);
                        const uint32_t pbi = (uint32_t)__builtin_amdgcn_readlane((int)bt_id, pi);
                        const uint32_t abk = (uint32_t)__builtin_amdgcn_readlane((int)a_k, ai);
                        const uint32_t abi = (uint32_t)__builtin_amdgcn_readlane((int)a_id, ai);
                        const bool from_a = na > 0 && (e == 0 || key_less(pbk, pbi, abk, abi));
                        const uint32_t bk = from_a ? abk : pbk;
                        const uint32_t bk_id = from_a ? abi : pbi;
                        const bool full = e + na >= L;
                        const bool prune = full && v != 0u && key_is_num(bk) && klb > bk;
                        const bool take = !full || (!prune && key_less(kd, id, bk, bk_id));
                        npruned += prune ? 1u : 0u;
                        nchecked += (full && v != 0u) ? 1u : 0u;
                        e -= (full && take && !from_a) ? 1u : 0u;
                        const uint32_t nna = na - ((full && take && from_a) ? 1u : 0u);
                        if (take) {
                            // sorted insert into the accepted list: lanes past the
                            // slot take their left neighbour (DPP wave_shr:1)
                            const float d = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(my_d), (int)b));
                            const uint32_t pos = (uint32_t)__popcll(__ballot(lane < nna && key_less(a_k, a_id, kd, id)));
                            const int sd = __builtin_amdgcn_update_dpp(__float_as_int(a_d), __float_as_int(a_d), 0x138, 0xF, 0xF, false);
                            const int sk = __builtin_amdgcn_update_dpp((int)a_k, (int)a_k, 0x138, 0xF, 0xF, false);
                            const int si = __builtin_amdgcn_update_dpp((int)a_id, (int)a_id, 0x138, 0xF, 0xF, false);
                            const bool shift = lane > pos && lane <= nna;
                            const bool put = lane == pos;
                            a_d = put ? d : (shift ? __int_as_float(sd) : a_d);
                            a_k = put ? kd : (shift ? (uint32_t)sk : a_k);
                            a_id = put ? id : (shift ? (uint32_t)si : a_id);
                        }
                        na = nna + (take ? 1u : 0u);
                    }
                }
                pruned += npruned;
                rpruned += npruned;
                rchecked += nchecked;
                // merge pool[0, e) with the accepted list into pool[0, e + na), in
                // place: retained entries only move right, so they move in
                // descending chunks of 64 (each read before any write), then
                // the accepted entries drop into the gaps
                if (na > 0) {
                    // accepted entry i lands at i + #(pool entries before it)
                    uint32_t lo = 0, hi = e;
                    if (lane < na) {
                        while (lo < hi) {
                            const uint32_t mid = (lo + hi) >> 1;
                            if (cand_less(pd[mid], pid[mid], a_d, a_id)) lo = mid + 1; else hi = mid;
                        }
                    }
                    const uint32_t first = (uint32_t)__builtin_amdgcn_readlane((int)lo, 0);
                    // retained entry k >= first moves right by #(accepted entries before it)
                    for (uint32_t top = e; top > first;) {
                        const uint32_t base = top - first > 64u ? top - 64u : first;
                        const uint32_t k = base + lane;
                        const bool mv = k < top;
                        float kd = 0.0f;
                        uint32_t kid = 0, c = 0;
                        uint8_t kx = 0;
                        if (mv) {
                            kd = pd[k]; kid = pid[k]; kx = pexp[k];
                            for (uint32_t i = 0; i < na; ++i) {
                                const float ad = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a_d), (int)i));
                                const uint32_t aid = (uint32_t)__builtin_amdgcn_readlane((int)a_id, (int)i);
                                c += cand_less(ad, aid, kd, kid) ? 1u : 0u;
                            }
                        }
                        wave_sync_lds();
                        if (mv && c) { pd[k + c] = kd; pid[k + c] = kid; pexp[k + c] = kx; }
                        wave_sync_lds();
                        top = base;
                    }
                    if (lane < na) { pd[lane + lo] = a_d; pid[lane + lo] = a_id; pexp[lane + lo] = 0; }
                    wave_sync_lds();
                }
                len = e + na;
                if (lane == 0) s_len = len;
            }
            __syncthreads();
            len = uni(s_len);
        }

        // the traversal pool goes out before the rerank reuses its arrays
        if (COLD(pool_ids)) {
            for (uint32_t i = t; i < COLD(L); i += T) {
                COLD(pool_ids)[(uint64_t)qi * COLD(L) + i] = i < len ? pid[i] : 0xFFFFFFFFu;
                COLD(pool_d)[(uint64_t)qi * COLD(L) + i] = i < len ? pd[i] : __builtin_huge_valf();
            }
            __syncthreads();
        }
        if (!COLD(pool_only)) {
            // rerank: the first rc pool entries; their estimates replace pd[0..rc)
            const uint32_t want = COLD(rerank) > COLD(k) ? COLD(rerank) : COLD(k);
            uint32_t rc = len < want ? len : want;
            if (LEARNED && rc > 0) rc = search_learned_depth<W>(a, qi, rc, pd, pid, s, (uint4 *)stage, &s_len);
            if (a.codec == 3) {
                // FP32 long distance (recomputed for the INT8 codec): one
                // candidate per 16-lane group over its code row staged in LDS
                // (cd <= 2048: a row is at most 4 uint4 per lane; the next
                // row's loads are in flight while the current one is scored)
                uint4 *row = stage + grp * cch;
                uint4 buf[4];
                float2 h = make_float2(0.0f, 0.0f);
                auto fetch = [&](uint32_t c) {
                    const uint32_t id = pid[c];
                    const uint4 *src = (const uint4 *)(a.ix.codes + (uint64_t)id * (cd / 2));
#pragma unroll
                    for (uint32_t k = 0; k < 4; ++k) {
                        const uint32_t i = (uint32_t)j16 + 16u * k;
                        buf[k] = i < cch ? src[i] : make_uint4(0u, 0u, 0u, 0u);
                    }
                    h = COLD(ix.hdr)[id];
                };
                if (grp < rc) fetch(grp);
                for (uint32_t c = grp; c < rc; c += G) {
#pragma unroll
                    for (uint32_t k = 0; k < 4; ++k) {
                        const uint32_t i = (uint32_t)j16 + 16u * k;
                        if (i < cch) row[i] = buf[k];
                    }
                    wave_sync_lds();
                    const float2 hc = h;
                    if (c + G < rc) fetch(c + G);
                    const float ld = asym_group(q, (const uint8_t *)row, cd, hc, s, j16);
                    if (j16 == 0) pd[c] = ld;
                    wave_sync_lds();
                }
                __syncthreads();
            }
            // residual-4: each (candidate, block) product sum on its own
            // thread into the stage (coalesced loads of consecutive blocks),
            // then eight lanes per candidate run the block recurrence; the
            // recurrence's scales are loaded first and fly with the block
            // loads, so a pass of up to T / 8 candidates waits on memory once
            {
                float *work = (float *)stage;
                const uint32_t nbk = cd / 16, stride = nbk + 1;
                const uint32_t pp = 4096u / stride < T / 8u ? 4096u / stride : T / 8u;
                for (uint32_t c0 = 0; c0 < rc; c0 += pp) {
                    const uint32_t cnt = rc - c0 < pp ? rc - c0 : pp;
                    uint2 sv[4];
                    residual_chain_scales(sv, COLD(ix.rscale), nbk, cnt, [&](uint32_t c) { return pid[c0 + c]; }, t);
                    residual_block_sums<T, 8>(work, stride, q, COLD(ix.rcodes), cd, cnt,
                                               [&](uint32_t c) { return pid[c0 + c]; }, t);
                    __syncthreads();
                    residual_chains_x8(sv, work, stride, nbk, cnt, pd + c0, pd + c0, t);
                    __syncthreads();
                }
            }
            // top-k by (estimate, id): k rounds of a wave-wide argmin on wave 0
            if (wv == 0) {
                const uint32_t kout = rc < COLD(k) ? rc : COLD(k);
                for (uint32_t r = 0; r < COLD(k); ++r) {
                    if (r >= kout) {
                        if (lane == 0) {
                            COLD(out_ids)[(uint64_t)qi * COLD(k) + r] = 0xFFFFFFFFu;
                            COLD(out_d)[(uint64_t)qi * COLD(k) + r] = __builtin_huge_valf();
                        }
                        continue;
                    }
                    float bd = __builtin_huge_valf();
                    uint32_t bid = 0xFFFFFFFFu, bpos = 0xFFFFFFFFu;
                    for (uint32_t i = lane; i < rc; i += 64u) {
                        if (pid[i] == 0xFFFFFFFFu) continue;
                        if (bpos == 0xFFFFFFFFu || cand_less(pd[i], pid[i], bd, bid)) { bd = pd[i]; bid = pid[i]; bpos = i; }
                    }
                    for (int off = 32; off > 0; off >>= 1) {
                        const float od = __shfl_xor(bd, off);
                        const uint32_t oid = (uint32_t)__shfl_xor((int)bid, off);
                        const uint32_t opos = (uint32_t)__shfl_xor((int)bpos, off);
                        const bool take = opos != 0xFFFFFFFFu && (bpos == 0xFFFFFFFFu || cand_less(od, oid, bd, bid));
                        if (take) { bd = od; bid = oid; bpos = opos; }
                    }
                    if (lane == 0) {
                        COLD(out_ids)[(uint64_t)qi * COLD(k) + r] = bid;
                        COLD(out_d)[(uint64_t)qi * COLD(k) + r] = bd;
                        pid[bpos] = 0xFFFFFFFFu;
                    }
                    wave_sync_lds();
                }
            }
        }
        if (t == 0 && COLD(stats)) {
            uint64_t *st = COLD(stats) + (uint64_t)qi * 6;
            st[0] = hops; st[1] = gates; st[2] = dists; st[3] = pruned; st[4] = rchecked; st[5] = rpruned;
        }
        if (wv == 0) __builtin_amdgcn_s_setprio(0);  // the next query's setup runs at the default priority
        __syncthreads();
    }
}

__global__ void k_zero_u32(uint32_t *p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// zero n4 uint4 (grid-stride): the visited bitsets of a launch's slots
__global__ void __launch_bounds__(256) k_zero_u4(uint4 *p, uint64_t n4) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (uint64_t)gridDim.x * 256)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}

}  // namespace qg

template <int W, bool LEARNED, bool FILTER>
static int launch_search_w(const qg::SearchArgs *a, uint32_t slots, size_t smem, hipStream_t stream) {
    static size_t raised = 0;  // dynamic LDS above the 64 KB default needs the attribute once per size
    if (smem > raised) {
        const hipError_t e = hipFuncSetAttribute((const void *)qg::k_search<W, LEARNED, FILTER>,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
        if (e != hipSuccess) return (int)e;
        raised = smem;
    }
    hipLaunchKernelGGL((qg::k_search<W, LEARNED, FILTER>), dim3(slots), dim3(64 * W), smem, stream, *a);
    return (int)hipGetLastError();
}

// dynamic LDS of the larger variant (the one without the gate filter)
extern "C" size_t qg_search_lds_bytes(uint32_t cd, uint32_t L) { return (size_t)cd * 8 + (size_t)L * 9 + 16; }

extern "C" int qg_launch_search(const qg::SearchArgs *a, uint32_t slots, uint32_t waves, hipStream_t stream) {
    if (a->qv.nq == 0) return 0;
    // no more workgroups than queries; every slot's bitset starts at zero
    // (a fill at full memory bandwidth, so a slot's first query skips its
    // own clear; later queries on the slot clear theirs)
    if (slots > a->qv.nq) slots = (uint32_t)a->qv.nq;
    hipLaunchKernelGGL(qg::k_zero_u32, dim3(1), dim3(64), 0, stream, a->counter, 2u);  // head + error flag
    {
        const uint64_t n4 = (uint64_t)slots * a->visited_words / 4;
        const uint64_t blocks = (n4 + 255) / 256 < 8192 ? (n4 + 255) / 256 : 8192;
        hipLaunchKernelGGL(qg::k_zero_u4, dim3((unsigned)blocks), dim3(256), 0, stream, (uint4 *)a->visited, n4);
    }
    // the FP32 gate's filter from code dimension 2048 (its query takes 2 cd
    // bytes of LDS where the pair-interleaved query takes 4 cd)
    const bool filter = a->codec != 3 && a->cd >= 2048;
    const size_t smem = (size_t)a->cd * (filter ? 6 : 8) + (size_t)a->L * 9 + 16;
    // the learned-depth variant is a kernel of its own, so the plain search
    // keeps its register allocation
    if (a->mlp)
        return filter ? (waves == 2 ? launch_search_w<2, true, true>(a, slots, smem, stream)
                                    : launch_search_w<4, true, true>(a, slots, smem, stream))
                      : (waves == 2 ? launch_search_w<2, true, false>(a, slots, smem, stream)
                                    : launch_search_w<4, true, false>(a, slots, smem, stream));
    return filter ? (waves == 2 ? launch_search_w<2, false, true>(a, slots, smem, stream)
                                : launch_search_w<4, false, true>(a, slots, smem, stream))
                  : (waves == 2 ? launch_search_w<2, false, false>(a, slots, smem, stream)
                                : launch_search_w<4, false, false>(a, slots, smem, stream));
}

// ------------------------------------------ search without a workspace
// qg_search_noalloc takes its work-queue head, error flag and per-slot
// visited bitsets from `scratch`, a file-scope device array that each GPU
// holds from the library's load to the end of the process, so the call
// allocates nothing. Host code never passes `scratch` itself (its host-side
// name is not a device address): the array's address on each device comes
// from hipGetSymbolAddress on that device's first call and is kept per
// device. Every call on one GPU shares that GPU's array; the search kernel
// clears each slot's bitset before it uses it and a zero-fill kernel clears
// the queue head and error flag at launch, so leftovers of an earlier call do
// not matter, but calls that overlap on one GPU race.
namespace {
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;  // 32 MiB per GPU
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

extern "C" size_t qg_search_noalloc_scratch_bytes(void) { return kNoAllocScratchBytes; }

extern "C" int qg_search_noalloc(const void *space, const qg_index_desc *ix, const qg_query_desc *qv, int codec,
                                 const uint32_t *degrees, const uint32_t *edges, size_t edge_stride, uint32_t k,
                                 uint32_t l_search, const uint32_t *l_per_query, float eps0, uint32_t rerank,
                                 uint32_t slots, uint32_t waves, uint32_t *out_ids, float *out_dists,
                                 uint64_t *out_stats, uint32_t *pool_ids, float *pool_dists, void *stream) {
    if (!ix) return qg::fail("null argument to qg_search_noalloc");
    if (slots == 0) slots = 1;
    const size_t need = qg_search_workspace_bytes(ix->n, slots);
    if (need > kNoAllocScratchBytes)
        return qg::fail("qg_search_noalloc: %u slots over %llu points need %zu bytes of scratch, above its "
                        "%zu-byte maximum (use fewer slots, or qg_search_ex with a workspace)",
                        slots, (unsigned long long)ix->n, need, kNoAllocScratchBytes);
    int dev = 0;
    if (qg::check_hip((int)hipGetDevice(&dev), "qg_search_noalloc: hipGetDevice")) return -1;
    if (dev < 0 || dev >= kMaxDevices) return qg::fail("qg_search_noalloc: device %d out of range", dev);
    void *ptr = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_ptr[dev]) {
            void *p = nullptr;
            if (qg::check_hip((int)hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                              "qg_search_noalloc: address of the scratch array"))
                return -1;
            g_scratch_ptr[dev] = p;
        }
        ptr = g_scratch_ptr[dev];
    }
    return qg_search_ex(space, ix, qv, codec, degrees, edges, edge_stride, k, l_search, l_per_query, eps0, rerank,
                        slots, waves, out_ids, out_dists, out_stats, pool_ids, pool_dists, ptr, kNoAllocScratchBytes,
                        stream);
}
