// qg_build.hip -- Vamana graph construction on the GPU with the symmetric
// 4-bit distance (rabitq_build_vamana_graph, rabitq_bridge.cpp:224-303 ->
// VamanaIndex::buildEncodedSymmetricBulk / refineGraphSymmetric,
// Ours/core/hnswlib/vamana_index.h:308-528, 995-1117).
//
// The reference build is deterministic by construction: each batch of
// inserts searches the graph as it stood when the batch began, prunes
// privately, commits, and then adds back-edges grouped by source so that no
// two groups touch the same adjacency row. The GPU build keeps that
// structure, one kernel per stage:
//   k_build_search   one wavefront per inserted id: search_vamana
//                    (vamana_search.h:97-198) with the BuildEvaluator
//                    distance (lhs = inserted record), node_count limit,
//                    beam width, early stop. Visited marks are per-slot
//                    epoch tags, so nothing is cleared between jobs.
//   k_prune          one workgroup per job: robust_prune_4bit
//                    (vamana_prune.h:41-156). The pool is sorted by
//                    (distance, id) in LDS, de-duplicated and swept level by
//                    level. Every round evaluates, for all open candidates
//                    at or after the scan position in parallel, the lazy
//                    memoised occlusion test of the reference (per candidate:
//                    count of memoised pair distances and their minimum;
//                    float rounding of alpha * x is monotone in x, so
//                    alpha * min <= d is the same "some selected node
//                    occludes" predicate); the first open candidate that is
//                    not occluded is selected. Candidates are committed as
//                    visited exactly when the reference's sequential sweep
//                    reaches them, which reproduces its distance_evaluations.
//   back-edges       (neighbour << 32 | id) keys radix-sorted, run-length
//                    encoded into groups per source, merged one wavefront per
//                    group (addBackedgeGroupAndPrune, vamana_index.h:1047-1117);
//                    groups whose list exceeds R become prune jobs.
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>
#include <utility>

#include "qg_args.h"
#include "qg_device.h"
#include "qg_space.h"

namespace qg {

int fail(const char *fmt, ...);
int check_hip(int rc, const char *what);

constexpr int kSearchLanes = 64;
constexpr int kPruneThreads = 256;
constexpr uint32_t kPruneLds = 2048;      // candidates whose state fits in LDS
constexpr uint32_t kMaxCodeDwords = 512;  // centred code of cd <= 2048
constexpr uint32_t kNone = 0xFFFFFFFFu;

// prune state word: byte 0 = memoised pair distances, byte 1 = the count at
// the last visit, byte 2 = index of the first negative memoised distance
// (0xFF: none), byte 3 = flags
constexpr uint32_t kSelected = 1u << 24;
constexpr uint32_t kKeep = 1u << 25;

__device__ __forceinline__ bool cand_less(float ad, uint32_t aid, float bd, uint32_t bid) {
    // candidate_less / prune_candidate_less: distance, then id
    return ad < bd || (ad == bd && aid < bid);
}

__device__ __forceinline__ unsigned long long prune_key(float d, uint32_t id) {
    return ((unsigned long long)float_order_key(d) << 32) | id;
}

__device__ __forceinline__ void stage_centered(uint32_t *dst, const uint8_t *row, uint32_t cd, int tid, int nthreads) {
    const uint32_t *w = (const uint32_t *)row;
    for (uint32_t d = tid; d < cd / 8; d += nthreads) {
        const uint32_t v = w[d];
        dst[2 * d] = centered_bytes(v & 0x0F0F0F0Fu);
        dst[2 * d + 1] = centered_bytes((v >> 4) & 0x0F0F0F0Fu);
    }
}

__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------- search --
// Without a gate, inserting the fresh neighbours one by one into a pool
// truncated at L (insert_candidate_sorted) leaves exactly the L smallest of
// pool + fresh by (distance, id): the pool's back only shrinks, so a
// candidate rejected at its turn would be rejected later too. Each
// expansion therefore sorts its fresh neighbours across the wavefront
// (bitonic network over lane shuffles) and merges them with the pool into
// the second LDS buffer at their merge ranks.
__device__ __forceinline__ void wave_sort_keys(unsigned long long &key, uint32_t &payload, int lane) {
    for (int k = 2; k <= 64; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            const unsigned long long ok = (unsigned long long)__shfl_xor((long long)key, j);
            const uint32_t op = (uint32_t)__shfl_xor((int)payload, j);
            const bool up = (lane & k) == 0;     // ascending block
            const bool lower = (lane & j) == 0;  // lower index of the pair
            const bool take = (lower == up) ? (ok < key) : (ok > key);
            if (take) { key = ok; payload = op; }
        }
    }
}

__global__ void __launch_bounds__(kSearchLanes) k_build_search(BuildSearchArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t cd = a.cd, L = a.L;
    uint32_t *lc = (uint32_t *)smem;                                // cd/4 dwords
    float *bd[2] = {(float *)(smem + cd), (float *)(smem + cd + 4 * (size_t)L)};
    uint32_t *bid[2] = {(uint32_t *)(smem + cd + 8 * (size_t)L), (uint32_t *)(smem + cd + 12 * (size_t)L)};
    uint8_t *bexp[2] = {smem + cd + 16 * (size_t)L, smem + cd + 17 * (size_t)L};
    __shared__ float fd_s[kSearchLanes];
    __shared__ uint32_t fid_s[kSearchLanes];
    __shared__ uint32_t frontier[kSearchLanes];
    const int lane = threadIdx.x;
    uint32_t *tags = a.tags + (uint64_t)blockIdx.x * a.tag_stride;
    for (;;) {
        uint32_t job = 0;
        if (lane == 0) job = atomicAdd(a.counter, 1u);
        job = (uint32_t)__shfl((int)job, 0);
        if (job >= a.jobs) break;
        const uint32_t qid = (uint32_t)(a.qid_base + job);
        const uint32_t tag = a.tag_base + job + 1u;
        stage_centered(lc, a.codes + (uint64_t)qid * (cd / 2), cd, lane, kSearchLanes);
        const float2 qh = a.hdr[qid];
        __syncthreads();
        int cur = 0;
        unsigned long long evals = 1;
        if (lane == 0) {
            tags[0] = tag;
            bd[0][0] = symmetric_finish(qh, a.hdr[0], symmetric_ip4_prepared(lc, a.codes, cd));
            bid[0][0] = 0;
            bexp[0][0] = 0;
        }
        uint32_t len = 1, stall = 0;
        bool bad = false;
        __syncthreads();
        for (;;) {
            float *pd = bd[cur];
            uint32_t *pid = bid[cur];
            uint8_t *pexp = bexp[cur];
            const float back_before = pd[len - 1];
            // frontier: the first `beam` unexpanded entries in pool order
            uint32_t nfr = 0;
            for (uint32_t base = 0; base < len && nfr < a.beam; base += kSearchLanes) {
                const uint32_t i = base + lane;
                const bool un = i < len && pexp[i] == 0;
                const uint64_t m = __ballot(un);
                const uint32_t r = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                const uint32_t room = a.beam - nfr;
                if (un && r < room) {
                    frontier[nfr + r] = pid[i];
                    pexp[i] = 1;
                }
                const uint32_t cnt = (uint32_t)__popcll(m);
                nfr += cnt < room ? cnt : room;
            }
            __syncthreads();
            if (nfr == 0) break;
            for (uint32_t f = 0; f < nfr; ++f) {
                const uint32_t node = frontier[f];
                const uint32_t deg = a.degrees[node];
                const bool in = (uint32_t)lane < deg;
                const uint32_t nb = in ? a.edges[(uint64_t)node * a.R + lane] : kNone;
                if (__ballot(in && nb >= a.node_count)) {  // vamana_search.h:157-159
                    bad = true;
                    break;
                }
                // try_mark in adjacency order: the first occurrence of an id
                bool first_occ = in;
                for (int j = 0; j < kSearchLanes - 1; ++j) {
                    const uint32_t o = (uint32_t)__shfl((int)nb, j);
                    if (j < lane && o == nb) first_occ = false;
                }
                bool is_fresh = false;
                if (first_occ) {
                    is_fresh = tags[nb] != tag;
                    if (is_fresh) tags[nb] = tag;
                }
                const uint32_t nf = (uint32_t)__popcll(__ballot(is_fresh));
                if (nf == 0) continue;
                evals += nf;
                float d = 0.0f;
                if (is_fresh)
                    d = symmetric_finish(qh, a.hdr[nb], symmetric_ip4_prepared(lc, a.codes + (uint64_t)nb * (cd / 2), cd));
                unsigned long long key = is_fresh ? prune_key(d, nb) : ~0ull;
                uint32_t dbits = __float_as_uint(d);
                wave_sort_keys(key, dbits, lane);
                if ((uint32_t)lane < nf) {
                    fd_s[lane] = __uint_as_float(dbits);
                    fid_s[lane] = (uint32_t)key;
                }
                __syncthreads();
                float *nd = bd[cur ^ 1];
                uint32_t *nid = bid[cur ^ 1];
                uint8_t *nexp = bexp[cur ^ 1];
                // pool entries move up by the number of fresh entries before them
                for (uint32_t i = lane; i < len; i += kSearchLanes) {
                    const float pdv = pd[i];
                    const uint32_t piv = pid[i];
                    uint32_t lo = 0, hi = nf;
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (cand_less(fd_s[mid], fid_s[mid], pdv, piv)) lo = mid + 1; else hi = mid;
                    }
                    const uint32_t at = i + lo;
                    if (at < L) { nd[at] = pdv; nid[at] = piv; nexp[at] = pexp[i]; }
                }
                // fresh entries land after the pool entries before them
                if ((uint32_t)lane < nf) {
                    const float fv = fd_s[lane];
                    const uint32_t fi = fid_s[lane];
                    uint32_t lo = 0, hi = len;
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (cand_less(pd[mid], pid[mid], fv, fi)) lo = mid + 1; else hi = mid;
                    }
                    const uint32_t at = lane + lo;
                    if (at < L) { nd[at] = fv; nid[at] = fi; nexp[at] = 0; }
                }
                len = (len + nf < L) ? len + nf : L;
                cur ^= 1;
                pd = nd;
                pid = nid;
                pexp = nexp;
                __syncthreads();
            }
            if (bad) break;
            if (a.early_stop != 0 && len >= L) {
                const float back_after = bd[cur][len - 1];
                stall = (back_after >= back_before) ? stall + 1 : 0;
                if (stall >= a.early_stop) break;
            }
        }
        for (uint32_t i = lane; i < len; i += kSearchLanes) {
            a.pool_d[(uint64_t)job * L + i] = bd[cur][i];
            a.pool_id[(uint64_t)job * L + i] = bid[cur][i];
        }
        if (lane == 0) {
            a.pool_len[job] = bad ? 0u : len;
            if (bad) atomicOr(a.err, 1u);
            atomicAdd(a.evals, evals);
        }
        __syncthreads();
    }
}

// -------------------------------------------------------------- prune --
__device__ __forceinline__ uint32_t block_min_u32(uint32_t v, uint32_t *red) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    uint32_t r = red[0];
    for (int w = 1; w < kPruneThreads / 64; ++w) r = min(r, red[w]);
    __syncthreads();
    return r;
}

__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v, unsigned long long *red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    unsigned long long r = 0;
    for (int w = 0; w < kPruneThreads / 64; ++w) r += red[w];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(kPruneThreads) k_prune(PruneArgs a) {
    __shared__ unsigned long long s_keys[kPruneLds];
    __shared__ float s_min[kPruneLds];
    __shared__ float s_pre[kPruneLds];
    __shared__ uint32_t s_state[kPruneLds];
    __shared__ uint32_t s_src[kMaxCodeDwords];
    __shared__ uint32_t s_sel[64];
    __shared__ uint32_t s_red[kPruneThreads / 64];
    __shared__ unsigned long long s_red64[kPruneThreads / 64];
    const int t = threadIdx.x;
    const int lane = t & 63, wv = t >> 6;
    const uint32_t job = a.job_ids ? a.job_ids[blockIdx.x] : blockIdx.x;
    const uint32_t cd = a.cd, R = a.R;
    const uint32_t src = a.sources ? a.sources[job] : (uint32_t)(a.src_base + job);
    uint32_t na = 0;
    if (a.pool_len) {
        na = a.pool_len[job];
        if (a.cap != 0 && na > a.cap) na = a.cap;  // capBuildPrunePool on a sorted pool
    }
    const uint32_t *ex = nullptr;
    uint32_t nb = 0;
    uint64_t off = 0;
    if (a.graph_deg) {
        nb = a.graph_deg[src];
        ex = a.graph_edges + (uint64_t)src * R;
    } else if (a.extra) {
        off = a.extra_off[job];
        nb = a.extra_len[job];
        ex = a.extra + off;
    }
    const uint32_t n = na + nb;
    const bool big = n > kPruneLds;
    if (big && (na != 0 || a.extra == nullptr || a.big_keys == nullptr)) {
        if (t == 0) {
            atomicOr(a.err, 2u);
            a.out_deg[src] = 0;
        }
        return;
    }
    unsigned long long *keys = big ? a.big_keys + off : s_keys;
    float *mind = big ? a.big_min + off : s_min;
    float *pre = big ? a.big_pre + off : s_pre;
    uint32_t *state = big ? a.big_state + off : s_state;

    // pool: search entries as given, listed ids with their distance from
    // the source (BuildEvaluator::distance_batch / distance_batch_from: lhs
    // = source)
    if (nb) stage_centered(s_src, a.codes + (uint64_t)src * (cd / 2), cd, t, kPruneThreads);
    __syncthreads();
    const float2 sh = a.hdr[src];
    unsigned long long ev = 0;
    for (uint32_t i = t; i < n; i += kPruneThreads) {
        if (i < na) {
            keys[i] = prune_key(a.pool_d[(uint64_t)job * a.pool_stride + i], a.pool_id[(uint64_t)job * a.pool_stride + i]);
        } else {
            const uint32_t id = ex[i - na];
            const float d = symmetric_finish(sh, a.hdr[id], symmetric_ip4_prepared(s_src, a.codes + (uint64_t)id * (cd / 2), cd));
            keys[i] = prune_key(d, id);
            ++ev;
        }
    }
    __syncthreads();
    // bitonic sort, ascending, over the next power of two; entries past n are
    // +inf and never move, so pairs with the upper index >= n are skipped
    uint32_t np2 = 1;
    while (np2 < n) np2 <<= 1;
    for (uint32_t size = 2; size <= np2; size <<= 1) {
        const uint32_t half = size >> 1;
        for (uint32_t b = t; b < np2 / 2; b += kPruneThreads) {
            const uint32_t blk = b / half, r = b % half;
            const uint32_t lo = blk * size + r, hi = blk * size + size - 1 - r;
            if (hi < n) {
                const unsigned long long k0 = keys[lo], k1 = keys[hi];
                if (k0 > k1) { keys[lo] = k1; keys[hi] = k0; }
            }
        }
        __syncthreads();
        for (uint32_t stride = half >> 1; stride > 0; stride >>= 1) {
            for (uint32_t b = t; b < np2 / 2; b += kPruneThreads) {
                const uint32_t lo = 2 * stride * (b / stride) + (b % stride), hi = lo + stride;
                if (hi < n) {
                    const unsigned long long k0 = keys[lo], k1 = keys[hi];
                    if (k0 > k1) { keys[lo] = k1; keys[hi] = k0; }
                }
            }
            __syncthreads();
        }
    }
    // drop the source and repeated ids. A repeated id carries the same
    // distance (one deterministic function of the pair), so its copies are
    // adjacent after the sort and the first is the one kept.
    for (uint32_t i = t; i < n; i += kPruneThreads) {
        const unsigned long long k = keys[i];
        const bool keep = (uint32_t)k != src && !(i > 0 && keys[i - 1] == k);
        state[i] = keep ? kKeep : 0u;
    }
    __syncthreads();
    uint32_t u = 0;
    for (uint32_t base = 0; base < n; base += kPruneThreads) {
        const uint32_t i = base + t;
        const bool keep = i < n && (state[i] & kKeep);
        const unsigned long long k = keep ? keys[i] : 0ull;
        const uint64_t m = __ballot(keep);
        if (lane == 0) s_red[wv] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (int w = 0; w < kPruneThreads / 64; ++w) {
            before += (w < wv) ? s_red[w] : 0u;
            total += s_red[w];
        }
        if (keep) keys[u + before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = k;
        u += total;
        __syncthreads();
    }
    for (uint32_t i = t; i < u; i += kPruneThreads) {
        mind[i] = __builtin_nanf("");
        pre[i] = __builtin_nanf("");
        state[i] = 0x00FF0000u;
    }
    __syncthreads();

    const float alpha = a.alpha;
    const float inc = (alpha < 1.2f) ? alpha : 1.2f;  // std::min(alpha, 1.2f)
    float cur = 1.0f;
    uint32_t nsel = 0;
    while (nsel < R) {
        bool promoted = false;
        uint32_t start = 0;
        while (nsel < R) {
            // lazy occlusion test of every open candidate at or after start
            uint32_t first = kNone;
            for (uint32_t j = start + t; j < u; j += kPruneThreads) {
                const uint32_t st = state[j];
                if (st & kSelected) continue;
                const unsigned long long k = keys[j];
                const float D = float_from_order_key((uint32_t)(k >> 32));
                float m = mind[j];
                uint32_t c = st & 0xFFu;
                bool occ = c > 0 && cur * m <= D;
                if (!occ && c < nsel) {
                    const uint32_t cj = (uint32_t)k;
                    const uint8_t *crow = a.codes + (uint64_t)cj * (cd / 2);
                    const float2 ch = a.hdr[cj];
                    uint32_t g = (st >> 16) & 0xFFu;
                    float p = pre[j];
                    do {
                        // distance_between(candidate, selected): lhs = candidate
                        const uint32_t sid = s_sel[c];
                        const float d = symmetric_finish(
                            ch, a.hdr[sid], symmetric_ip4_rows(crow, a.codes + (uint64_t)sid * (cd / 2), cd));
                        if (d < 0.0f && g == 0xFFu) { g = c; p = m; }
                        m = (m != m || d < m) ? d : m;
                        ++c;
                        occ = cur * d <= D;
                    } while (!occ && c < nsel);
                    mind[j] = m;
                    pre[j] = p;
                    state[j] = (st & 0xFF00FF00u) | (g << 16) | c;
                }
                if (!occ) { first = j; break; }
            }
            first = block_min_u32(first, s_red);
            // the sequential sweep visits every open candidate in
            // [start, first] (all of [start, u) when none is selectable)
            const uint32_t vend = (first == kNone) ? u : first + 1;
            for (uint32_t j = start + t; j < vend; j += kPruneThreads) {
                uint32_t st = state[j];
                if (st & kSelected) continue;
                const uint32_t c = st & 0xFFu, com = (st >> 8) & 0xFFu, g = (st >> 16) & 0xFFu;
                ev += c - com;
                if (g != 0xFFu && g < com) {
                    // a memoised negative distance (cached < 0) is computed
                    // again when the sweep reaches it
                    const float D = float_from_order_key((uint32_t)(keys[j] >> 32));
                    if (!(cur * pre[j] <= D)) ++ev;
                }
                st = (st & ~0x0000FF00u) | (c << 8);
                if (j == first) st |= kSelected;
                state[j] = st;
            }
            // the next sweep maps candidates to other threads
            __syncthreads();
            if (first == kNone) break;
            if (t == 0) s_sel[nsel] = (uint32_t)keys[first];
            __syncthreads();
            ++nsel;
            start = first + 1;
            promoted = true;
        }
        if (cur == alpha) break;
        const float nx = cur * inc;
        const float next = (nx < alpha) ? nx : alpha;  // std::min(alpha, current_alpha * increment)
        if (next == cur) break;
        cur = next;
        if (!promoted && nsel >= R) break;
    }
    __syncthreads();
    for (uint32_t i = t; i < R; i += kPruneThreads) a.out_edges[(uint64_t)src * R + i] = i < nsel ? s_sel[i] : kNone;
    if (t == 0) a.out_deg[src] = nsel;
    ev = block_sum_u64(ev, s_red64);
    if (t == 0 && ev) atomicAdd(a.evals, ev);
}

// --------------------------------------------------------- gram prune --
// robust_prune_4bit for a listed pool of up to MP-1 candidates whose
// distances from the source are computed here (the back-edge re-prunes of
// addBackedgeGroupAndPrune, vamana_index.h:1091-1116). One wavefront per
// job. The exact integer parts ip4(a, b) = sum (2a_i - 15)(2b_i - 15) of every
// pair among {source} + pool are one int8 GEMM G = C C^T on the matrix cores
// (v_mfma_i32_16x16x32_i8, exact int32 accumulation; the dimension order
// inside a K block is the same permutation for both operands, which leaves
// every dot product unchanged); G is kept as a triangle in LDS. The sweep
// then follows the reference's sequential loop: the k-th selected node sits
// in lane k, so "some selected node occludes the visited candidate" is a
// ballot and the first occluder in selection order is its lowest set bit,
// which gives the memoised-pair counts of the reference exactly.
typedef int v4i_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t tri_index(uint32_t i, uint32_t j) {  // i < j
    return j * (j - 1) / 2 + i;
}

__device__ __forceinline__ long pack_centered(uint32_t w) {
    return (long)(((unsigned long long)centered_bytes((w >> 4) & 0x0F0F0F0Fu) << 32) |
                  centered_bytes(w & 0x0F0F0F0Fu));
}

template <int CD, int MP>
__global__ void __launch_bounds__(64) k_prune_gram(PruneArgs a) {
    constexpr int KQ = CD / 32;  // code dwords per lane group
    constexpr int NP2 = MP <= 128 ? 128 : 256;
    __shared__ int sG[MP * (MP - 1) / 2];
    __shared__ unsigned long long sKey[NP2];
    __shared__ uint16_t sRow[NP2];
    __shared__ float2 sH[MP];
    __shared__ uint32_t sId[MP];
    __shared__ float sD[NP2];
    __shared__ uint8_t sCnt[NP2];
    __shared__ uint8_t sSel[NP2];
    const int lane = threadIdx.x;
    const uint32_t job = blockIdx.x;
    const uint32_t R = a.R;
    const uint32_t src = a.sources[job];
    const uint32_t m = a.extra_len[job];
    const uint32_t *ex = a.extra + a.extra_off[job];
    if (m + 1 > MP) {  // the host sorts jobs into size classes
        if (lane == 0) { atomicOr(a.err, 2u); a.out_deg[src] = 0; }
        return;
    }
    for (uint32_t r = lane; r < MP; r += 64) {
        const uint32_t id = r == 0 ? src : (r <= m ? ex[r - 1] : src);
        sId[r] = id;
        sH[r] = a.hdr[id];
    }
    __syncthreads();
    // G over rows 0..m (row 0 = source), 16-row blocks
    const uint32_t nb = (m + 16) / 16;
    const int g = lane >> 4, rl = lane & 15;
    for (uint32_t bi = 0; bi < nb; ++bi) {
        const uint32_t ra = 16 * bi + rl;
        uint32_t A[KQ];
        {
            const uint32_t *rp = (const uint32_t *)(a.codes + (uint64_t)sId[ra < MP ? ra : 0] * (CD / 2)) + g * KQ;
            const bool va = ra <= m;
#pragma unroll
            for (int t = 0; t < KQ; ++t) A[t] = va ? rp[t] : 0u;
        }
        for (uint32_t bj = bi; bj < nb; ++bj) {
            const uint32_t rb = 16 * bj + rl;
            const uint32_t *rp = (const uint32_t *)(a.codes + (uint64_t)sId[rb < MP ? rb : 0] * (CD / 2)) + g * KQ;
            const bool vb = rb <= m;
            v4i_t acc = {0, 0, 0, 0};
#pragma unroll
            for (int t = 0; t < KQ; ++t) {
                const uint32_t b = vb ? rp[t] : 0u;
                acc = __builtin_amdgcn_mfma_i32_16x16x32_i8(pack_centered(A[t]), pack_centered(b), acc, 0, 0, 0);
            }
            // lane holds rows 16bi + 4g + r, column 16bj + rl
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t i = 16 * bi + 4 * g + r, j = rb;
                if (i < j && j <= m) sG[tri_index(i, j)] = acc[r];
            }
        }
    }
    __syncthreads();
    // pool distances from the source (lhs = source), sorted by (distance, id)
    const float2 sh = sH[0];
    for (uint32_t r = 1 + lane; r <= m; r += 64) {
        const float d = symmetric_finish(sh, sH[r], sG[tri_index(0, r)]);
        sKey[r - 1] = prune_key(d, sId[r]);
        sRow[r - 1] = (uint16_t)r;
    }
    unsigned long long ev = m;
    __syncthreads();
    uint32_t np2 = 1;
    while (np2 < m) np2 <<= 1;
    for (uint32_t size = 2; size <= np2; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t b = lane; b < np2 / 2; b += 64) {
                uint32_t lo, hi;
                if (stride == size >> 1) {
                    const uint32_t blk = b / stride, r = b % stride;
                    lo = blk * size + r;
                    hi = blk * size + size - 1 - r;
                } else {
                    lo = 2 * stride * (b / stride) + (b % stride);
                    hi = lo + stride;
                }
                if (hi < m) {
                    const unsigned long long k0 = sKey[lo], k1 = sKey[hi];
                    if (k0 > k1) {
                        sKey[lo] = k1;
                        sKey[hi] = k0;
                        const uint16_t t0 = sRow[lo];
                        sRow[lo] = sRow[hi];
                        sRow[hi] = t0;
                    }
                }
            }
            __syncthreads();
        }
    }
    // drop the source and repeated ids (adjacent after the sort), compact
    uint32_t u = 0;
    unsigned long long prev = ~0ull;
    for (uint32_t base = 0; base < m; base += 64) {
        const uint32_t p = base + lane;
        const unsigned long long k = p < m ? sKey[p] : ~0ull;
        const uint16_t row = p < m ? sRow[p] : 0;
        const unsigned long long before = (unsigned long long)__shfl_up((long long)k, 1);
        const unsigned long long pk = lane == 0 ? prev : before;
        const bool keep = p < m && (uint32_t)k != src && k != pk;
        const uint64_t msk = __ballot(keep);
        prev = (unsigned long long)__shfl((long long)k, 63);
        __syncthreads();
        if (keep) {
            const uint32_t dst = u + (uint32_t)__popcll(msk & ((1ull << lane) - 1ull));
            sKey[dst] = k;
            sRow[dst] = row;
        }
        u += (uint32_t)__popcll(msk);
        __syncthreads();
    }
    for (uint32_t p = lane; p < u; p += 64) {
        sD[p] = float_from_order_key((uint32_t)(sKey[p] >> 32));
        sCnt[p] = 0;
        sSel[p] = 0;
    }
    __syncthreads();

    const float alpha = a.alpha;
    const float inc = (alpha < 1.2f) ? alpha : 1.2f;
    float cur = 1.0f;
    uint32_t nsel = 0;
    uint32_t my_row = 0;  // row of the selected node held by this lane
    float2 my_h = make_float2(0.0f, 0.0f);
    while (nsel < R) {
        bool promoted = false;
        for (uint32_t p = 0; p < u && nsel < R; ++p) {
            if (sSel[p]) continue;
            const uint32_t ri = sRow[p];
            const float Di = sD[p];
            const float2 hi = sH[ri];
            bool occ = false, neg = false;
            if ((uint32_t)lane < nsel) {
                const uint32_t lo = ri < my_row ? ri : my_row, hi_r = ri < my_row ? my_row : ri;
                const float d = symmetric_finish(hi, my_h, sG[tri_index(lo, hi_r)]);  // lhs = candidate
                occ = cur * d <= Di;
                neg = d < 0.0f;
            }
            const uint64_t om = __ballot(occ), nm = __ballot(neg);
            const uint32_t e = om ? (uint32_t)__ffsll((unsigned long long)om) : nsel;
            const uint32_t cp = sCnt[p];
            const uint32_t lim = e < cp ? e : cp;
            // new memo fills, plus memoised negative distances computed again
            ev += (e > cp ? e - cp : 0u) + (uint32_t)__popcll(nm & (lim >= 64 ? ~0ull : ((1ull << lim) - 1ull)));
            sCnt[p] = (uint8_t)(e > cp ? e : cp);
            if (!om) {
                sSel[p] = 1;
                if ((uint32_t)lane == nsel) { my_row = ri; my_h = hi; }
                ++nsel;
                promoted = true;
            }
        }
        if (cur == alpha) break;
        const float nx = cur * inc;
        const float next = (nx < alpha) ? nx : alpha;
        if (next == cur) break;
        cur = next;
        if (!promoted && nsel >= R) break;
    }
    for (uint32_t i = lane; i < R; i += 64)
        a.out_edges[(uint64_t)src * R + i] = i < nsel ? sId[my_row] : kNone;
    if (lane == 0) {
        a.out_deg[src] = nsel;
        atomicAdd(a.evals, ev);
    }
}

// ---------------------------------------------------- pool gram prune --
// robust_prune_4bit for search pools (the batch inserts and refine passes,
// vamana_index.h:372-396, 486-499): part A = the first min(len, cap)
// entries of the job's sorted search pool, part B = the record's current
// adjacency row (refine only; its distances from the record are counted).
// Four wavefronts: ids are de-duplicated (a repeated id carries one
// distance, so dropping repeats by id equals the reference's dedup of
// adjacent equal (distance, id) pairs) and the source is dropped; the Gram
// matrix of {source} + pool comes from int8 MFMA with the block rows spread
// over the waves; every pool distance is then re-derived from G with the
// source as lhs (the function the search used, so the values are the
// search's). G lives in dynamic LDS as 24-bit entries (|ip4| <= 225 cd <
// 2^19). Wave 0 then runs the same sequential sweep as k_prune_gram. Jobs
// whose de-duplicated pool does not fit mp rows are appended to the
// overflow list and pruned by k_prune.
__device__ __forceinline__ int g24(const uint16_t *lo, const int8_t *hi, uint32_t t) {
    return ((int)hi[t] << 16) | (int)lo[t];
}

template <int CD>
__global__ void __launch_bounds__(256) k_prune_pool_gram(PruneArgs a, uint32_t mp, uint32_t *overflow,
                                                         uint32_t *overflow_count) {
    constexpr int KQ = CD / 32;
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t t_count = mp * (mp - 1) / 2;
    // layout: [ids, later keys + rows: 8 KB][G lo16][G hi8][sH][sId][sD][sCnt][sSel]
    uint32_t *sIds = (uint32_t *)smem;                          // [<= 2048] ids, later keys
    unsigned long long *sKey = (unsigned long long *)smem;      // [<= 512]
    uint16_t *sRow = (uint16_t *)(smem + 4096);                 // [<= 512]
    uint16_t *gLo = (uint16_t *)(smem + 8192);
    int8_t *gHi = (int8_t *)(smem + 8192 + 2 * (size_t)t_count);
    size_t o = (8192 + 3 * (size_t)t_count + 15) & ~(size_t)15;
    float2 *sH = (float2 *)(smem + o);
    o += 8 * (size_t)mp;
    uint32_t *sId = (uint32_t *)(smem + o);
    o += 4 * (size_t)mp;
    float *sD = (float *)(smem + o);
    o += 4 * (size_t)mp;
    uint8_t *sCnt = smem + o;
    o += mp;
    uint8_t *sSel = smem + o;
    __shared__ uint32_t s_red[4];
    __shared__ uint32_t s_mu;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const uint32_t job = blockIdx.x;
    const uint32_t R = a.R;
    const uint32_t src = (uint32_t)(a.src_base + job);
    uint32_t na = a.pool_len[job];
    if (a.cap != 0 && na > a.cap) na = a.cap;
    uint32_t nb = 0;
    const uint32_t *ex = nullptr;
    if (a.graph_deg) {
        nb = a.graph_deg[src];
        ex = a.graph_edges + (uint64_t)src * R;
    }
    const uint32_t m = na + nb;  // <= 2048 (L <= 1024, R <= 64)
    for (uint32_t i = t; i < m; i += 256) sIds[i] = i < na ? a.pool_id[(uint64_t)job * a.pool_stride + i] : ex[i - na];
    __syncthreads();
    uint32_t np2 = 1;
    while (np2 < m) np2 <<= 1;
    for (uint32_t size = 2; size <= np2; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t b = t; b < np2 / 2; b += 256) {
                uint32_t lo, hi;
                if (stride == size >> 1) {
                    lo = (b / stride) * size + (b % stride);
                    hi = (b / stride) * size + size - 1 - (b % stride);
                } else {
                    lo = 2 * stride * (b / stride) + (b % stride);
                    hi = lo + stride;
                }
                if (hi < m) {
                    const uint32_t k0 = sIds[lo], k1 = sIds[hi];
                    if (k0 > k1) { sIds[lo] = k1; sIds[hi] = k0; }
                }
            }
            __syncthreads();
        }
    }
    // unique ids other than the source, in id order
    uint32_t mu = 0;
    for (uint32_t base = 0; base < m; base += 256) {
        const uint32_t i = base + t;
        const uint32_t id = i < m ? sIds[i] : kNone;
        const bool keep = i < m && id != src && !(i > 0 && sIds[i - 1] == id);
        const uint64_t msk = __ballot(keep);
        if (lane == 0) s_red[wv] = (uint32_t)__popcll(msk);
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (int w = 0; w < 4; ++w) { before += w < wv ? s_red[w] : 0u; total += s_red[w]; }
        if (keep && mu + before + (uint32_t)__popcll(msk & ((1ull << lane) - 1ull)) + 1 < mp)
            sId[1 + mu + before + (uint32_t)__popcll(msk & ((1ull << lane) - 1ull))] = id;
        mu += total;
        __syncthreads();
    }
    if (mu + 1 > mp) {
        if (t == 0) overflow[atomicAdd(overflow_count, 1u)] = job;
        return;
    }
    for (uint32_t r = t; r <= mu; r += 256) {
        const uint32_t id = r == 0 ? src : sId[r];
        if (r == 0) sId[0] = src;
        sH[r] = a.hdr[id];
    }
    __syncthreads();
    // Gram matrix over rows 0..mu, block rows bi = wv, wv + 4, ...
    const uint32_t nblk = (mu + 16) / 16;
    const int g = lane >> 4, rl = lane & 15;
    for (uint32_t bi = wv; bi < nblk; bi += 4) {
        const uint32_t ra = 16 * bi + rl;
        uint32_t A[KQ];
        {
            const bool va = ra <= mu;
            const uint32_t *rp = (const uint32_t *)(a.codes + (uint64_t)sId[va ? ra : 0] * (CD / 2)) + g * KQ;
#pragma unroll
            for (int q = 0; q < KQ; ++q) A[q] = va ? rp[q] : 0u;
        }
        for (uint32_t bj = bi; bj < nblk; ++bj) {
            const uint32_t rb = 16 * bj + rl;
            const bool vb = rb <= mu;
            const uint32_t *rp = (const uint32_t *)(a.codes + (uint64_t)sId[vb ? rb : 0] * (CD / 2)) + g * KQ;
            v4i_t acc = {0, 0, 0, 0};
#pragma unroll
            for (int q = 0; q < KQ; ++q) {
                const uint32_t bw = vb ? rp[q] : 0u;
                acc = __builtin_amdgcn_mfma_i32_16x16x32_i8(pack_centered(A[q]), pack_centered(bw), acc, 0, 0, 0);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t i = 16 * bi + 4 * g + r, j = rb;
                if (i < j && j <= mu) {
                    const uint32_t ti = tri_index(i, j);
                    gLo[ti] = (uint16_t)(acc[r] & 0xFFFF);
                    gHi[ti] = (int8_t)(acc[r] >> 16);
                }
            }
        }
    }
    __syncthreads();
    // pool distances with the source as lhs, sorted by (distance, id)
    const float2 sh = sH[0];
    for (uint32_t r = 1 + t; r <= mu; r += 256) {
        const float d = symmetric_finish(sh, sH[r], g24(gLo, gHi, tri_index(0, r)));
        sKey[r - 1] = prune_key(d, sId[r]);
        sRow[r - 1] = (uint16_t)r;
    }
    __syncthreads();
    np2 = 1;
    while (np2 < mu) np2 <<= 1;
    for (uint32_t size = 2; size <= np2; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t b = t; b < np2 / 2; b += 256) {
                uint32_t lo, hi;
                if (stride == size >> 1) {
                    lo = (b / stride) * size + (b % stride);
                    hi = (b / stride) * size + size - 1 - (b % stride);
                } else {
                    lo = 2 * stride * (b / stride) + (b % stride);
                    hi = lo + stride;
                }
                if (hi < mu) {
                    const unsigned long long k0 = sKey[lo], k1 = sKey[hi];
                    if (k0 > k1) {
                        sKey[lo] = k1;
                        sKey[hi] = k0;
                        const uint16_t t0 = sRow[lo];
                        sRow[lo] = sRow[hi];
                        sRow[hi] = t0;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t p = t; p < mu; p += 256) {
        sD[p] = float_from_order_key((uint32_t)(sKey[p] >> 32));
        sCnt[p] = 0;
        sSel[p] = 0;
    }
    __syncthreads();
    if (wv != 0) return;
    // sequential sweep on wave 0 (as k_prune_gram)
    const uint32_t u = mu;
    unsigned long long ev = nb;
    const float alpha = a.alpha;
    const float inc = (alpha < 1.2f) ? alpha : 1.2f;
    float cur = 1.0f;
    uint32_t nsel = 0;
    uint32_t my_row = 0;
    float2 my_h = make_float2(0.0f, 0.0f);
    while (nsel < R) {
        bool promoted = false;
        for (uint32_t p = 0; p < u && nsel < R; ++p) {
            if (sSel[p]) continue;
            const uint32_t ri = sRow[p];
            const float Di = sD[p];
            const float2 hi = sH[ri];
            bool occ = false, neg = false;
            if ((uint32_t)lane < nsel) {
                const uint32_t lo = ri < my_row ? ri : my_row, hr = ri < my_row ? my_row : ri;
                const float d = symmetric_finish(hi, my_h, g24(gLo, gHi, tri_index(lo, hr)));
                occ = cur * d <= Di;
                neg = d < 0.0f;
            }
            const uint64_t om = __ballot(occ), nm = __ballot(neg);
            const uint32_t e = om ? (uint32_t)__ffsll((unsigned long long)om) : nsel;
            const uint32_t cp = sCnt[p];
            const uint32_t lim = e < cp ? e : cp;
            ev += (e > cp ? e - cp : 0u) + (uint32_t)__popcll(nm & (lim >= 64 ? ~0ull : ((1ull << lim) - 1ull)));
            sCnt[p] = (uint8_t)(e > cp ? e : cp);
            if (!om) {
                sSel[p] = 1;
                if ((uint32_t)lane == nsel) { my_row = ri; my_h = hi; }
                ++nsel;
                promoted = true;
            }
        }
        if (cur == alpha) break;
        const float nx = cur * inc;
        const float next = (nx < alpha) ? nx : alpha;
        if (next == cur) break;
        cur = next;
        if (!promoted && nsel >= R) break;
    }
    for (uint32_t i = lane; i < R; i += 64) a.out_edges[(uint64_t)src * R + i] = i < nsel ? sId[my_row] : kNone;
    if (lane == 0) {
        a.out_deg[src] = nsel;
        atomicAdd(a.evals, ev);
    }
}

static size_t pool_gram_lds(uint32_t mp) {
    const size_t t_count = (size_t)mp * (mp - 1) / 2;
    size_t o = (8192 + 3 * t_count + 15) & ~(size_t)15;
    return o + 8 * (size_t)mp + 4 * (size_t)mp + 4 * (size_t)mp + 2 * (size_t)mp;
}

// ---------------------------------------------------------- back-edges --
// (neighbour << 32 | row) for rows [row_base, row_base + rows); empty slots
// are ~0 and sort last
__global__ void k_edge_keys(const uint32_t *deg, const uint32_t *edges, uint32_t R, uint64_t row_base, uint64_t rows,
                            unsigned long long *keys) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= rows * R) return;
    const uint64_t r = row_base + g / R;
    const uint32_t i = (uint32_t)(g % R);
    keys[g] = i < deg[r] ? (((unsigned long long)edges[r * R + i] << 32) | r) : ~0ull;
}

__global__ void k_split_keys(const unsigned long long *keys, uint64_t n, uint32_t *src, uint32_t *tgt) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long k = keys[i];
    src[i] = (uint32_t)(k >> 32);
    tgt[i] = (uint32_t)k;
}

__global__ void k_merged_sizes(const uint32_t *group_src, const uint32_t *group_len, const uint32_t *ngroups,
                               const uint32_t *deg, uint64_t *sizes, uint64_t cap) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= cap) return;
    uint64_t s = 0;
    if (g < *ngroups && group_src[g] != kNone) s = (uint64_t)deg[group_src[g]] + group_len[g];
    sizes[g] = s;
}

// prune size classes of listed pools: k_prune_gram<., 80 / 128 / 168> and,
// above 167 candidates, k_prune
constexpr int kGramMP[3] = {80, 128, 168};
constexpr uint32_t kPoolGramMaxRows = 312;  // largest pool triangle that fits 160 KB of LDS
__device__ __forceinline__ uint32_t gram_class(uint32_t m) {
    return m < 80 ? 0u : (m < 128 ? 1u : (m < 168 ? 2u : 3u));
}

// addBackedgeGroupAndPrune, merge half: the current list, then each sorted
// unique target that is neither the source nor already present
__global__ void __launch_bounds__(64) k_backedge_merge(MergeArgs a) {
    const uint32_t g = blockIdx.x;
    if (g >= a.ngroups) return;
    const uint32_t src = a.group_src[g];
    if (src == kNone) return;
    const int lane = threadIdx.x;
    const uint32_t begin = a.group_begin[g], glen = a.group_len[g];
    const uint32_t deg = a.degrees[src];
    const uint32_t *cur = a.edges + (uint64_t)src * a.R;
    uint32_t *out = a.merged + a.merged_off[g];
    const uint32_t mine = (uint32_t)lane < deg ? cur[lane] : kNone;
    if ((uint32_t)lane < deg) out[lane] = mine;
    uint32_t total = deg;
    for (uint32_t base = begin; base < begin + glen; base += 64) {
        const uint32_t i = base + lane;
        bool take = false;
        uint32_t tg = kNone;
        if (i < begin + glen) {
            tg = a.targets[i];
            take = tg != src && (i == begin || a.targets[i - 1] != tg);
        }
        for (int j = 0; j < 64; ++j) {
            const uint32_t o = (uint32_t)__shfl((int)mine, j);
            if ((uint32_t)j < deg && o == tg) take = false;
        }
        const uint64_t m = __ballot(take);
        if (take) out[total + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = tg;
        total += (uint32_t)__popcll(m);
    }
    if (total == deg) return;  // nothing added
    __syncthreads();
    if (total <= a.R) {
        for (uint32_t i = lane; i < total; i += 64) a.edges[(uint64_t)src * a.R + i] = out[i];
        if (lane == 0) a.degrees[src] = total;
    } else if (lane == 0) {
        const uint32_t c = gram_class(total);
        const uint64_t j = (uint64_t)c * a.job_cap + atomicAdd(&a.job_count[c], 1u);
        a.job_src[j] = src;
        a.job_off[j] = a.merged_off[g];
        a.job_len[j] = total;
    }
}

// ------------------------------------------------------------ host side --
struct BuildLayout {
    uint64_t n, R, L, J, P, M, Jc, slots;
    size_t cub_bytes;
    size_t off_ctl, off_tags, off_pool_d, off_pool_id, off_pool_len, off_rdeg, off_redges, off_keys_a, off_keys_b,
        off_src, off_tgt, off_overflow, off_gsrc, off_glen, off_gbeg, off_msize, off_moff, off_merged, off_bkeys, off_bmin,
        off_bpre, off_bstate, off_jsrc, off_joff, off_jlen, off_cub, total;
};

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

static int cub_temp_bytes(uint64_t P, size_t *out) {
    size_t a = 0, b = 0, c = 0, d = 0;
    int rc = (int)hipcub::DeviceRadixSort::SortKeys(nullptr, a, (const unsigned long long *)nullptr,
                                                    (unsigned long long *)nullptr, (int)P, 0, 64);
    if (rc) return rc;
    rc = (int)hipcub::DeviceRunLengthEncode::Encode(nullptr, b, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                                    (uint32_t *)nullptr, (uint32_t *)nullptr, (int)P);
    if (rc) return rc;
    rc = (int)hipcub::DeviceScan::ExclusiveSum(nullptr, c, (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)P);
    if (rc) return rc;
    rc = (int)hipcub::DeviceScan::ExclusiveSum(nullptr, d, (const uint64_t *)nullptr, (uint64_t *)nullptr, (int)P);
    if (rc) return rc;
    size_t m = a;
    if (b > m) m = b;
    if (c > m) m = c;
    if (d > m) m = d;
    *out = m;
    return 0;
}

static int build_layout(uint64_t n, uint32_t R, uint32_t L, uint32_t batch, uint32_t refine, BuildLayout *l) {
    l->n = n;
    l->R = R;
    l->L = L;
    const uint64_t jb = batch < n ? batch : n;
    l->J = refine ? (n < 65536 ? n : (jb > 65536 ? jb : 65536)) : jb;
    if (l->J == 0) l->J = 1;
    const uint64_t rows = refine ? n : jb;  // rows whose edges feed one back-edge round
    l->P = rows * R;
    if (l->P == 0) l->P = 1;
    l->M = (l->P < n ? l->P : n) * R + l->P;  // groups <= min(P, n), each grows by at most R
    // concurrent searches: enough wavefronts to hide row-fetch latency, with
    // the visit tags (slots x n u32) held to 16 GiB
    uint64_t slots = (16ull << 30) / (4 * (n ? n : 1));
    slots = slots < 256 ? 256 : (slots > 8192 ? 8192 : slots);
    l->slots = l->J < slots ? l->J : slots;
    int rc = cub_temp_bytes(l->P, &l->cub_bytes);
    if (rc) return rc;
    size_t o = 0;
    auto take = [&](size_t bytes) { const size_t at = o; o = align256(o + bytes); return at; };
    l->off_ctl = take(256);
    l->off_tags = take(l->slots * n * 4);
    l->off_pool_d = take(l->J * L * 4);
    l->off_pool_id = take(l->J * L * 4);
    l->off_pool_len = take(l->J * 4);
    l->off_overflow = take(l->J * 4);
    l->off_rdeg = take(refine ? n * 4 : 0);
    l->off_redges = take(refine ? n * R * 4 : 0);
    l->off_keys_a = take(l->P * 8);
    l->off_keys_b = take(l->P * 8);
    l->off_src = take(l->P * 4);
    l->off_tgt = take(l->P * 4);
    l->off_gsrc = take(l->P * 4);
    l->off_glen = take(l->P * 4);
    l->off_gbeg = take(l->P * 4);
    l->off_msize = take(l->P * 8);
    l->off_moff = take(l->P * 8);
    l->off_merged = take(l->M * 4);
    l->off_bkeys = take(l->M * 8);
    l->off_bmin = take(l->M * 4);
    l->off_bpre = take(l->M * 4);
    l->off_bstate = take(l->M * 4);
    l->Jc = l->P < n ? l->P : n;  // prune jobs per back-edge round <= distinct sources
    if (l->Jc == 0) l->Jc = 1;
    l->off_jsrc = take(4 * l->Jc * 4);
    l->off_joff = take(4 * l->Jc * 8);
    l->off_jlen = take(4 * l->Jc * 4);
    l->off_cub = take(l->cub_bytes);
    l->total = o;
    return 0;
}

struct Ctl {
    uint32_t counter, err, ngroups, noverflow;
    uint32_t njobs[4];
    unsigned long long evals;
};

static int launch_fill(uint32_t *p, uint64_t n, uint32_t v, hipStream_t st) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
    return (int)hipGetLastError();
}

struct Builder {
    const Space *s;
    const uint8_t *codes;
    const float2 *hdr;
    BuildLayout l;
    char *w;
    hipStream_t st;
    uint32_t R, L, beam, early_stop, cap;
    float alpha;
    uint32_t tag_next = 0;

    template <class T> T *at(size_t off) { return reinterpret_cast<T *>(w + off); }
    Ctl *ctl() { return at<Ctl>(l.off_ctl); }

    int read_ctl(Ctl *h) {
        int rc = (int)hipMemcpyAsync(h, ctl(), sizeof(Ctl), hipMemcpyDeviceToHost, st);
        if (rc) return check_hip(rc, "build control read");
        return check_hip((int)hipStreamSynchronize(st), "build synchronize");
    }

    int reset_counter() { return check_hip(launch_fill(&ctl()->counter, 1, 0u, st), "build counter reset"); }

    // searches for records qid_base .. qid_base + jobs - 1 into the pools
    int search(uint64_t qid_base, uint32_t jobs, uint64_t node_count, const uint32_t *deg, const uint32_t *edges) {
        if (reset_counter()) return -1;
        BuildSearchArgs a;
        a.codes = codes;
        a.hdr = hdr;
        a.cd = (uint32_t)s->cd;
        a.degrees = deg;
        a.edges = edges;
        a.R = R;
        a.node_count = node_count;
        a.L = L;
        a.beam = beam;
        a.early_stop = early_stop;
        a.jobs = jobs;
        a.qid_base = qid_base;
        a.tags = at<uint32_t>(l.off_tags);
        a.tag_stride = l.n;
        a.tag_base = tag_next;
        tag_next += jobs;
        a.counter = &ctl()->counter;
        a.err = &ctl()->err;
        a.evals = &ctl()->evals;
        a.pool_d = at<float>(l.off_pool_d);
        a.pool_id = at<uint32_t>(l.off_pool_id);
        a.pool_len = at<uint32_t>(l.off_pool_len);
        const uint32_t slots = (uint32_t)(jobs < l.slots ? jobs : l.slots);
        const size_t smem = (size_t)s->cd + (size_t)L * 18;
        hipLaunchKernelGGL(k_build_search, dim3(slots), dim3(kSearchLanes), smem, st, a);
        return check_hip((int)hipGetLastError(), "build search launch");
    }

    PruneArgs prune_args() {
        PruneArgs p;
        memset(&p, 0, sizeof(p));
        p.codes = codes;
        p.hdr = hdr;
        p.cd = (uint32_t)s->cd;
        p.R = R;
        p.alpha = alpha;
        p.err = &ctl()->err;
        p.evals = &ctl()->evals;
        return p;
    }

    int prune(const PruneArgs &p) {
        if (p.jobs == 0) return 0;
        hipLaunchKernelGGL(k_prune, dim3(p.jobs), dim3(kPruneThreads), 0, st, p);
        return check_hip((int)hipGetLastError(), "prune launch");
    }

    // prune the pools of records id_base .. id_base + jobs - 1 into out rows;
    // with existing lists (refine), each record's current row joins its pool
    int prune_pools(uint64_t id_base, uint32_t jobs, const uint32_t *exist_deg, const uint32_t *exist_edges,
                    uint32_t *out_deg, uint32_t *out_edges) {
        PruneArgs p = prune_args();
        p.jobs = jobs;
        p.src_base = id_base;
        p.pool_d = at<float>(l.off_pool_d);
        p.pool_id = at<uint32_t>(l.off_pool_id);
        p.pool_len = at<uint32_t>(l.off_pool_len);
        p.pool_stride = L;
        p.cap = cap;
        p.graph_deg = exist_deg;
        p.graph_edges = exist_edges;
        p.out_deg = out_deg;
        p.out_edges = out_edges;
        // Gram prune when the capped pool fits the LDS triangle; jobs whose
        // de-duplicated pool does not fit go to k_prune
        const uint32_t pool_max = (cap != 0 && cap < L) ? cap : L;
        if (jobs == 0 || pool_max + 1 > kPoolGramMaxRows) return prune(p);
        uint32_t need = pool_max + (exist_deg ? R : 0) + 1;
        need = (need + 15) & ~15u;
        const uint32_t mp = need < kPoolGramMaxRows ? need : kPoolGramMaxRows;
        uint32_t *ovf = at<uint32_t>(l.off_overflow);
        if (check_hip(launch_fill(&ctl()->noverflow, 1, 0u, st), "overflow reset")) return -1;
        const size_t lds = pool_gram_lds(mp);
        int rc = 0;
        switch (s->cd) {
#define QG_POOL_CASE(CD)                                                                                  \
    case CD:                                                                                              \
        rc = (int)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_prune_pool_gram<CD>),            \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);              \
        if (rc == 0)                                                                                      \
            hipLaunchKernelGGL(k_prune_pool_gram<CD>, dim3(jobs), dim3(256), lds, st, p, mp, ovf,         \
                               &ctl()->noverflow);                                                        \
        break;
            QG_POOL_CASE(64)
            QG_POOL_CASE(128)
            QG_POOL_CASE(256)
            QG_POOL_CASE(512)
            QG_POOL_CASE(1024)
            QG_POOL_CASE(2048)
#undef QG_POOL_CASE
            default: return prune(p);
        }
        if (check_hip(rc, "pool Gram attribute") || check_hip((int)hipGetLastError(), "pool Gram prune launch"))
            return -1;
        Ctl h;
        if (read_ctl(&h)) return -1;
        if (h.noverflow == 0) return 0;
        p.jobs = h.noverflow;
        p.job_ids = ovf;
        return prune(p);
    }

    // addBackedgesGroupedAndPrune over the edges of rows [row_base, row_base + rows)
    int backedges(uint64_t row_base, uint64_t rows, uint32_t *deg, uint32_t *edges) {
        const uint64_t P = rows * R;
        if (P == 0) return 0;
        unsigned long long *ka = at<unsigned long long>(l.off_keys_a), *kb = at<unsigned long long>(l.off_keys_b);
        hipLaunchKernelGGL(k_edge_keys, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, deg, edges, R, row_base,
                           rows, ka);
        if (check_hip((int)hipGetLastError(), "edge keys")) return -1;
        int end_bit = 32;
        while (end_bit < 64 && (1ull << (end_bit - 32)) <= l.n) ++end_bit;
        size_t tb = l.cub_bytes;
        void *temp = w + l.off_cub;
        if (check_hip((int)hipcub::DeviceRadixSort::SortKeys(temp, tb, ka, kb, (int)P, 0, end_bit, st), "edge sort"))
            return -1;
        uint32_t *src = at<uint32_t>(l.off_src), *tgt = at<uint32_t>(l.off_tgt);
        hipLaunchKernelGGL(k_split_keys, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, kb, P, src, tgt);
        if (check_hip((int)hipGetLastError(), "edge split")) return -1;
        uint32_t *gsrc = at<uint32_t>(l.off_gsrc), *glen = at<uint32_t>(l.off_glen), *gbeg = at<uint32_t>(l.off_gbeg);
        tb = l.cub_bytes;
        if (check_hip((int)hipcub::DeviceRunLengthEncode::Encode(temp, tb, src, gsrc, glen, &ctl()->ngroups, (int)P, st),
                      "edge groups"))
            return -1;
        tb = l.cub_bytes;
        if (check_hip((int)hipcub::DeviceScan::ExclusiveSum(temp, tb, glen, gbeg, (int)P, st), "group offsets"))
            return -1;
        uint64_t *msize = at<uint64_t>(l.off_msize), *moff = at<uint64_t>(l.off_moff);
        hipLaunchKernelGGL(k_merged_sizes, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, gsrc, glen,
                           &ctl()->ngroups, deg, msize, P);
        if (check_hip((int)hipGetLastError(), "merged sizes")) return -1;
        tb = l.cub_bytes;
        if (check_hip((int)hipcub::DeviceScan::ExclusiveSum(temp, tb, msize, moff, (int)P, st), "merged offsets"))
            return -1;
        if (check_hip(launch_fill(ctl()->njobs, 4, 0u, st), "job count reset")) return -1;
        Ctl h;
        if (read_ctl(&h)) return -1;
        MergeArgs m;
        m.group_src = gsrc;
        m.group_begin = gbeg;
        m.group_len = glen;
        m.targets = tgt;
        m.ngroups = h.ngroups;
        m.R = R;
        m.degrees = deg;
        m.edges = edges;
        m.merged_off = moff;
        m.merged = at<uint32_t>(l.off_merged);
        m.job_src = at<uint32_t>(l.off_jsrc);
        m.job_off = at<uint64_t>(l.off_joff);
        m.job_len = at<uint32_t>(l.off_jlen);
        m.job_count = ctl()->njobs;
        m.job_cap = l.Jc;
        if (h.ngroups) {
            hipLaunchKernelGGL(k_backedge_merge, dim3(h.ngroups), dim3(64), 0, st, m);
            if (check_hip((int)hipGetLastError(), "back-edge merge")) return -1;
        }
        if (read_ctl(&h)) return -1;
        for (int c = 0; c < 4; ++c) {
            PruneArgs p = prune_args();
            p.jobs = h.njobs[c];
            p.sources = m.job_src + c * l.Jc;
            p.extra = m.merged;
            p.extra_off = m.job_off + c * l.Jc;
            p.extra_len = m.job_len + c * l.Jc;
            p.big_keys = at<unsigned long long>(l.off_bkeys);
            p.big_min = at<float>(l.off_bmin);
            p.big_pre = at<float>(l.off_bpre);
            p.big_state = at<uint32_t>(l.off_bstate);
            p.out_deg = deg;
            p.out_edges = edges;
            if ((c < 3 ? prune_gram(p, c) : prune(p)) != 0) return -1;
        }
        return 0;
    }

    int prune_gram(const PruneArgs &p, int c) {
        if (p.jobs == 0) return 0;
        switch (s->cd * 4 + c) {
#define QG_GRAM_CASE(CD)                                                                                        \
    case CD * 4 + 0: hipLaunchKernelGGL((k_prune_gram<CD, kGramMP[0]>), dim3(p.jobs), dim3(64), 0, st, p); break; \
    case CD * 4 + 1: hipLaunchKernelGGL((k_prune_gram<CD, kGramMP[1]>), dim3(p.jobs), dim3(64), 0, st, p); break; \
    case CD * 4 + 2: hipLaunchKernelGGL((k_prune_gram<CD, kGramMP[2]>), dim3(p.jobs), dim3(64), 0, st, p); break;
            QG_GRAM_CASE(64)
            QG_GRAM_CASE(128)
            QG_GRAM_CASE(256)
            QG_GRAM_CASE(512)
            QG_GRAM_CASE(1024)
            QG_GRAM_CASE(2048)
#undef QG_GRAM_CASE
            default: return fail("no Gram prune kernel for code dimension %zu", s->cd);
        }
        return check_hip((int)hipGetLastError(), "Gram prune launch");
    }
};

}  // namespace qg

extern "C" {

size_t qg_build_workspace_bytes(uint64_t n, uint32_t r, uint32_t l_build, uint32_t batch, uint32_t refine_passes) {
    qg::BuildLayout l;
    if (batch == 0) batch = 1;
    if (qg::build_layout(n, r, l_build, batch, refine_passes, &l)) return 0;
    return l.total;
}

int qg_build_graph(const void *space, const qg_index_desc *ix, uint32_t r, uint32_t l_build, float alpha,
                   uint32_t beam_width, uint32_t batch, uint32_t refine_passes, uint32_t prune_cap,
                   uint32_t early_stop_hops, uint32_t *degrees, uint32_t *edges, uint64_t *distance_evaluations,
                   void *workspace, size_t workspace_bytes, void *stream) {
    using namespace qg;
    const Space *s = static_cast<const Space *>(space);
    if (!s || !ix || !degrees || !edges || !workspace) return fail("null argument to qg_build_graph");
    if (s->centroid_count != 1) return fail("the device build has no cross-centroid terms: spaces of one centroid only");
    if (s->cd > kMaxGraphCodeDim) return fail("the graph build takes code dimensions up to %u", kMaxGraphCodeDim);
    if (r == 0 || r > 64) return fail("graph degree R must be in [1, 64]");
    if (l_build == 0 || l_build > 1024) return fail("build list size must be in [1, 1024]");
    if (!(alpha >= 1.0f) || !std::isfinite(alpha)) return fail("Vamana alpha must be finite and >= 1.0");
    if (beam_width == 0 || beam_width > 64) return fail("beam width must be in [1, 64]");
    if (batch == 0) batch = 1;
    const uint64_t n = ix->n;
    if (n > 0xFFFFFFF0ull) return fail("index too large for 32-bit node ids");
    if (n * (uint64_t)(1 + refine_passes) >= 0xFFFFFFF0ull) return fail("too many searches for 32-bit visit tags");
    Builder b;
    b.s = s;
    b.codes = ix->codes;
    b.hdr = reinterpret_cast<const float2 *>(ix->hdr);
    if (build_layout(n, r, l_build, batch, refine_passes, &b.l)) return fail("build workspace sizing failed");
    if (workspace_bytes < b.l.total) return fail("build workspace too small: need %zu bytes", b.l.total);
    b.w = static_cast<char *>(workspace);
    b.st = (hipStream_t)stream;
    b.R = r;
    b.L = l_build;
    b.beam = beam_width;
    b.early_stop = early_stop_hops;
    b.cap = prune_cap;
    b.alpha = alpha;
    hipStream_t st = b.st;
    if (check_hip(launch_fill(reinterpret_cast<uint32_t *>(b.w + b.l.off_ctl), 64, 0u, st), "build init") ||
        check_hip(launch_fill(reinterpret_cast<uint32_t *>(b.w + b.l.off_tags), b.l.slots * n, 0u, st), "tag init") ||
        check_hip(launch_fill(degrees, n, 0u, st), "degree init") ||
        check_hip(launch_fill(edges, n * r, 0xFFFFFFFFu, st), "edge init"))
        return -1;
    // buildGraphFromLoadedPayloads (vamana_index.h:452-528)
    for (uint64_t begin = 1; begin < n; begin += batch) {
        const uint64_t end = begin + batch < n ? begin + batch : n;
        const uint32_t jobs = (uint32_t)(end - begin);
        if (b.search(begin, jobs, begin, degrees, edges)) return -1;
        if (b.prune_pools(begin, jobs, nullptr, nullptr, degrees, edges)) return -1;
        if (b.backedges(begin, jobs, degrees, edges)) return -1;
        Ctl h;
        if (b.read_ctl(&h)) return -1;
        if (h.err & 1u) return fail("Vamana graph contains out-of-range neighbor");
        if (h.err) return fail("prune pool exceeds the workspace layout");
    }
    // refineGraphSymmetric (vamana_index.h:328-422): every record searches
    // the frozen graph; results go to the second graph and are swapped in
    if (refine_passes != 0 && n > 1) {
        uint32_t *cur_deg = degrees, *cur_edges = edges;
        uint32_t *alt_deg = reinterpret_cast<uint32_t *>(b.w + b.l.off_rdeg);
        uint32_t *alt_edges = reinterpret_cast<uint32_t *>(b.w + b.l.off_redges);
        for (uint32_t pass = 0; pass < refine_passes; ++pass) {
            for (uint64_t c0 = 0; c0 < n; c0 += b.l.J) {
                const uint64_t c1 = c0 + b.l.J < n ? c0 + b.l.J : n;
                const uint32_t jobs = (uint32_t)(c1 - c0);
                if (b.search(c0, jobs, n, cur_deg, cur_edges)) return -1;
                if (b.prune_pools(c0, jobs, cur_deg, cur_edges, alt_deg, alt_edges)) return -1;
            }
            std::swap(cur_deg, alt_deg);
            std::swap(cur_edges, alt_edges);
            if (b.backedges(0, n, cur_deg, cur_edges)) return -1;
            Ctl h;
            if (b.read_ctl(&h)) return -1;
            if (h.err & 1u) return fail("Vamana graph contains out-of-range neighbor");
            if (h.err) return fail("prune pool exceeds the workspace layout");
        }
        if (cur_deg != degrees) {
            if (check_hip((int)hipMemcpyAsync(degrees, cur_deg, n * 4, hipMemcpyDeviceToDevice, st), "graph copy") ||
                check_hip((int)hipMemcpyAsync(edges, cur_edges, n * r * 4, hipMemcpyDeviceToDevice, st), "graph copy"))
                return -1;
        }
    }
    Ctl h;
    if (b.read_ctl(&h)) return -1;
    if (distance_evaluations) *distance_evaluations = h.evals;
    return 0;
}

}  // extern "C"
