/*
 * ref_search.c -- the reference's CPU search path, for the CPU baseline and
 * as a second checker of the GPU search.
 *
 * TEST / BASELINE INFRASTRUCTURE ONLY. Only tests/ and bench.py's
 * cpu_baseline leg load it; the product library never links or calls it.
 *
 * The reference search loop lives in Rust
 * (experiments/02_diskann_fair/src/ours_diskann.rs:1539-1911, flush
 * :1914-1986), which has no toolchain here. This file restates that loop in
 * C and drives the reference's own compiled kernels through its C-ABI bridge
 * (oracle/_ref/librabitq_ref.so, rabitq_bridge.cpp), loaded with dlopen:
 *   - per expansion, one rabitq_paper_estimate_batch_sidecar over the fresh
 *     neighbours (the 1-bit gate, :1683-1700);
 *   - survivors queued with mode 1 (FP32 query, MSB reuse), 2 (INT query,
 *     MSB reuse) or 0 (invalid gate: full recompute) and flushed through
 *     rabitq_query_distance_batch in batches of 64 (DIST_BATCH_SIZE, :1160),
 *     re-checked against the pool when flushed;
 *   - rerank: the first min(|pool|, max(rerank, k)) pool ids, FP32 long
 *     distance recomputed (mode 0) for INT queries, then
 *     rabitq_residual_distance_batch, sorted by (estimate, id) (:1813-1902).
 * Start node 0, beam width 1, kth / early stop off (the formal settings).
 */
#include <dlfcn.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { float estimate, lower_bound, upper_bound; } rs_interval;
typedef struct { float lower_bound, short_ip, alpha, ip_hat, error_bound; uint8_t valid, pad[3]; } rs_estimate;

typedef const void *(*prepare_fn)(const void *, const float *);
typedef int (*distance_batch_fn)(const void *, const void *, const uint32_t *, const uint8_t *, size_t,
                                 const void *, size_t, const float *, float *);
typedef int (*estimate_batch_fn)(const void *, const void *, const uint32_t *, size_t, const uint8_t *, size_t,
                                 const void *, size_t, float, rs_estimate *);
typedef int (*residual_batch_fn)(const void *, const void *, const uint32_t *, size_t, const void *, size_t,
                                 const void *, size_t, const float *, rs_interval *);

static prepare_fn p_prepare;
static distance_batch_fn p_distance;
static estimate_batch_fn p_estimate;
static residual_batch_fn p_residual;

/* Binds the bridge entry points from the library at `path`; 0 on success. */
int rs_open(const char *path) {
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) return -1;
    p_prepare = (prepare_fn)dlsym(h, "rabitq_prepare_query");
    p_distance = (distance_batch_fn)dlsym(h, "rabitq_query_distance_batch");
    p_estimate = (estimate_batch_fn)dlsym(h, "rabitq_paper_estimate_batch_sidecar");
    p_residual = (residual_batch_fn)dlsym(h, "rabitq_residual_distance_batch");
    return (p_prepare && p_distance && p_estimate && p_residual) ? 0 : -2;
}

typedef struct { float d; uint32_t id; uint8_t expanded; } cand;

static int less(float ad, uint32_t aid, float bd, uint32_t bid) { return ad < bd || (ad == bd && aid < bid); }

/* insert_candidate_sorted (ours_diskann.rs:1100-1126) */
static void insert(cand *pool, size_t *len, size_t L, float d, uint32_t id) {
    if (*len >= L && !less(d, id, pool[*len - 1].d, pool[*len - 1].id)) return;
    size_t pos = 0;
    while (pos < *len && less(pool[pos].d, pool[pos].id, d, id)) ++pos;
    const size_t n = *len < L ? *len : L - 1;
    memmove(pool + pos + 1, pool + pos, (n - pos) * sizeof(cand));
    pool[pos].d = d;
    pool[pos].id = id;
    pool[pos].expanded = 0;
    if (*len < L) ++*len;
}

typedef struct {
    uint32_t ids[64];
    uint32_t slots[64];
    uint8_t modes[64];
    float short_ips[64];
    float dists[64];
    size_t n;
} batch;

static int flush(const void *space, const void *pq, const uint8_t *comp, size_t cstride, const rs_estimate *est,
                 batch *b, cand *pool, size_t *len, size_t L) {
    if (b->n == 0) return 0;
    if (!p_distance(space, pq, b->ids, b->modes, b->n, comp, cstride, b->short_ips, b->dists)) return -1;
    for (size_t i = 0; i < b->n; ++i) {
        const rs_estimate *e = &est[b->slots[i]];
        const int full = *len >= L;
        const float lb = full ? pool[*len - 1].d : 3.402823466e38f;
        if (full && e->valid && e->lower_bound > lb) continue;
        insert(pool, len, L, b->dists[i], b->ids[i]);
    }
    b->n = 0;
    return 0;
}

/* One query; codec 0 (Full) or 3 (Int8) must match the codec set on the
 * space. Returns the number of results, or -1 on a bridge failure / bad
 * graph. */
long rs_search(const void *space, const float *query, int codec, const uint8_t *comp, size_t cstride,
               const uint8_t *res, size_t rstride, const uint8_t *msb, size_t mstride, const void *factors,
               size_t fstride, const uint32_t *degrees, const uint32_t *edges, size_t estride, size_t n, size_t k,
               size_t L, float eps0, size_t rerank, uint32_t *out_ids, float *out_d, uint32_t *visited_tags,
               uint32_t tag) {
    if (k == 0 || n == 0) return 0;
    if (L < k) L = k;
    const void *pq = p_prepare(space, query);
    if (!pq) return -1;
    cand *pool = (cand *)malloc((L + 1) * sizeof(cand));
    size_t len = 0;
    batch b;
    b.n = 0;
    uint32_t fresh[4096];
    rs_estimate est[4096];
    long rc = 0;
    uint8_t zero = 0;
    float d0 = 0.0f, zf = 0.0f;
    uint32_t start = 0;
    visited_tags[0] = tag;
    if (!p_distance(space, pq, &start, &zero, 1, comp, cstride, &zf, &d0)) { free(pool); return -1; }
    insert(pool, &len, L, d0, 0);
    for (;;) {
        size_t f = len;
        for (size_t i = 0; i < len; ++i) if (!pool[i].expanded) { f = i; break; }
        if (f == len) break;
        pool[f].expanded = 1;
        const uint32_t node = pool[f].id;
        size_t nf = 0;
        for (uint32_t j = 0; j < degrees[node]; ++j) {
            const uint32_t nb = edges[(size_t)node * estride + j];
            if (nb >= n) { rc = -1; goto done; }
            if (visited_tags[nb] != tag) { visited_tags[nb] = tag; fresh[nf++] = nb; }
        }
        if (nf == 0) continue;
        if (!p_estimate(space, pq, fresh, nf, msb, mstride, factors, fstride, eps0, est)) { rc = -1; goto done; }
        for (size_t s = 0; s < nf; ++s) {
            const int full = len >= L;
            const float lb = full ? pool[len - 1].d : 3.402823466e38f;
            if (full && est[s].valid && est[s].lower_bound > lb) continue;
            b.ids[b.n] = fresh[s];
            b.slots[b.n] = (uint32_t)s;
            b.modes[b.n] = !est[s].valid ? 0 : (codec == 3 ? 2 : 1);
            b.short_ips[b.n] = est[s].short_ip;
            ++b.n;
            if (b.n >= 64 && flush(space, pq, comp, cstride, est, &b, pool, &len, L)) { rc = -1; goto done; }
        }
        if (flush(space, pq, comp, cstride, est, &b, pool, &len, L)) { rc = -1; goto done; }
    }
    {
        size_t rcount = rerank > k ? rerank : k;
        if (rcount > len) rcount = len;
        uint32_t *rid = (uint32_t *)malloc((rcount + 1) * sizeof(uint32_t));
        float *rlong = (float *)malloc((rcount + 1) * sizeof(float));
        uint8_t *modes = (uint8_t *)calloc(rcount + 1, 1);
        float *sips = (float *)calloc(rcount + 1, sizeof(float));
        rs_interval *iv = (rs_interval *)malloc((rcount + 1) * sizeof(rs_interval));
        for (size_t i = 0; i < rcount; ++i) { rid[i] = pool[i].id; rlong[i] = pool[i].d; }
        int ok = 1;
        if (rcount && codec == 3) ok = p_distance(space, pq, rid, modes, rcount, comp, cstride, sips, rlong);
        if (ok && rcount) ok = p_residual(space, pq, rid, rcount, comp, cstride, res, rstride, rlong, iv);
        if (ok) {
            for (size_t i = 1; i < rcount; ++i) {  /* sort by (estimate, id) */
                const float d = iv[i].estimate;
                const uint32_t id = rid[i];
                size_t j = i;
                while (j > 0 && less(d, id, iv[j - 1].estimate, rid[j - 1])) { iv[j] = iv[j - 1]; rid[j] = rid[j - 1]; --j; }
                iv[j].estimate = d;
                rid[j] = id;
            }
            const size_t out = rcount < k ? rcount : k;
            for (size_t i = 0; i < out; ++i) { out_ids[i] = rid[i]; out_d[i] = iv[i].estimate; }
            rc = (long)out;
        } else {
            rc = -1;
        }
        free(rid); free(rlong); free(modes); free(sips); free(iv);
    }
done:
    free(pool);
    return rc;
}

/* Queries q0 .. q1-1 of `queries` [nq][dim] on the calling thread (the
 * prepared query is thread-local in the bridge). Results [nq][k], missing
 * entries 0xFFFFFFFF. Returns 0, or -1 on the first failure. */
int rs_search_range(const void *space, const float *queries, size_t dim, size_t q0, size_t q1, int codec,
                    const uint8_t *comp, size_t cstride, const uint8_t *res, size_t rstride, const uint8_t *msb,
                    size_t mstride, const void *factors, size_t fstride, const uint32_t *degrees,
                    const uint32_t *edges, size_t estride, size_t n, size_t k, size_t L, float eps0, size_t rerank,
                    uint32_t *out_ids, float *out_d) {
    uint32_t *tags = (uint32_t *)calloc(n + 1, sizeof(uint32_t));
    if (!tags) return -1;
    int rc = 0;
    for (size_t qi = q0; qi < q1 && rc == 0; ++qi) {
        for (size_t j = 0; j < k; ++j) { out_ids[qi * k + j] = 0xFFFFFFFFu; out_d[qi * k + j] = 0.0f; }
        const long got = rs_search(space, queries + qi * dim, codec, comp, cstride, res, rstride, msb, mstride,
                                   factors, fstride, degrees, edges, estride, n, k, L, eps0, rerank,
                                   out_ids + qi * k, out_d + qi * k, tags, (uint32_t)(qi - q0 + 1));
        if (got < 0) rc = -1;
    }
    free(tags);
    return rc;
}

/* rs_search_range with each query's own list size lists[qi] (the per-query
 * loop of experiments/02_diskann_fair/src/diskann_runner.rs:360-439, whose
 * visited-tag scratch persists across queries, ours_diskann.rs:1130-1151). */
int rs_search_lists(const void *space, const float *queries, size_t dim, size_t q0, size_t q1, int codec,
                    const uint8_t *comp, size_t cstride, const uint8_t *res, size_t rstride, const uint8_t *msb,
                    size_t mstride, const void *factors, size_t fstride, const uint32_t *degrees,
                    const uint32_t *edges, size_t estride, size_t n, size_t k, const uint32_t *lists, float eps0,
                    size_t rerank, uint32_t *out_ids, float *out_d) {
    uint32_t *tags = (uint32_t *)calloc(n + 1, sizeof(uint32_t));
    if (!tags) return -1;
    int rc = 0;
    for (size_t qi = q0; qi < q1 && rc == 0; ++qi) {
        for (size_t j = 0; j < k; ++j) { out_ids[qi * k + j] = 0xFFFFFFFFu; out_d[qi * k + j] = 0.0f; }
        const long got = rs_search(space, queries + qi * dim, codec, comp, cstride, res, rstride, msb, mstride,
                                   factors, fstride, degrees, edges, estride, n, k, lists[qi], eps0, rerank,
                                   out_ids + qi * k, out_d + qi * k, tags, (uint32_t)(qi - q0 + 1));
        if (got < 0) rc = -1;
    }
    free(tags);
    return rc;
}

/* The fixed-candidate pipeline of the plain method (the harness of
 * experiments/01_quantizer_fair/faiss_quantizer_smoke.cpp:1312-1426) through
 * the bridge, for queries q0 .. q1-1: mode-0 distances of the query's `per`
 * candidate ids (cands [nq][per]), the first `depth` by (distance, id),
 * rabitq_residual_distance_batch on them, and the top `k` by (estimate, id)
 * into out_ids [nq][k]. Returns 0, or -1 on a bridge failure. */
int rs_fixed_range(const void *space, const float *queries, size_t dim, size_t q0, size_t q1, const uint8_t *comp,
                   size_t cstride, const uint8_t *res, size_t rstride, const uint32_t *cands, size_t per, size_t depth,
                   size_t k, uint32_t *out_ids) {
    if (depth > per) depth = per;
    if (k > depth) k = depth;
    float *d = (float *)malloc((per + 1) * sizeof(float));
    float *sips = (float *)calloc(per + 1, sizeof(float));
    uint8_t *modes = (uint8_t *)calloc(per + 1, 1);
    cand *pool = (cand *)malloc((depth + 1) * sizeof(cand));
    uint32_t *rid = (uint32_t *)malloc((depth + 1) * sizeof(uint32_t));
    float *rlong = (float *)malloc((depth + 1) * sizeof(float));
    rs_interval *iv = (rs_interval *)malloc((depth + 1) * sizeof(rs_interval));
    int rc = (d && sips && modes && pool && rid && rlong && iv) ? 0 : -1;
    for (size_t qi = q0; qi < q1 && rc == 0; ++qi) {
        const uint32_t *ids = cands + qi * per;
        const void *pq = p_prepare(space, queries + qi * dim);
        if (!pq || !p_distance(space, pq, ids, modes, per, comp, cstride, sips, d)) { rc = -1; break; }
        size_t len = 0;
        for (size_t i = 0; i < per; ++i) insert(pool, &len, depth, d[i], ids[i]);
        for (size_t i = 0; i < len; ++i) { rid[i] = pool[i].id; rlong[i] = pool[i].d; }
        if (len && !p_residual(space, pq, rid, len, comp, cstride, res, rstride, rlong, iv)) { rc = -1; break; }
        for (size_t i = 1; i < len; ++i) {  /* sort by (estimate, id) */
            const float e = iv[i].estimate;
            const uint32_t id = rid[i];
            size_t j = i;
            while (j > 0 && less(e, id, iv[j - 1].estimate, rid[j - 1])) { iv[j] = iv[j - 1]; rid[j] = rid[j - 1]; --j; }
            iv[j].estimate = e;
            rid[j] = id;
        }
        for (size_t i = 0; i < k; ++i) out_ids[qi * k + i] = i < len ? rid[i] : 0xFFFFFFFFu;
    }
    free(d); free(sips); free(modes); free(pool); free(rid); free(rlong); free(iv);
    return rc;
}
