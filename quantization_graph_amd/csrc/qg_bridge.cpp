// qg_bridge.cpp -- host-memory drop-ins for the reference C-ABI
// (experiments/02_diskann_fair/native/rabitq_bridge.cpp).
//
// Every entry point keeps the reference's signature, argument meaning and
// failure sentinels (false / nullptr / +inf / zeroed estimates / intervals
// equal to the long distance). Work is staged through the calling thread's
// device buffer, computed by the same kernels as the batch API, and copied
// back synchronously, so threads may share one space as the reference's
// callers do. The prepared-query handle lives in thread-local storage,
// valid on the calling thread until its next rabitq_prepare_query, as in
// the reference (space_rabitq.h:3700-3706).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <random>
#include <vector>

#include "qg_args.h"
#include "qg_space.h"

namespace qg {
namespace {

struct Arena {
    char *base;
    size_t used, cap;
    void *take(size_t bytes) {
        const size_t off = (used + 255) & ~size_t(255);
        if (off + bytes > cap) return nullptr;
        used = off + bytes;
        return base + off;
    }
};

// Device staging memory of the calling thread. The reference's space is
// shared read-only by many threads (Send + Sync, ours_diskann.rs:250-251)
// whose workers call the bridge at once (encode workers, :2150-2199), so
// nothing a call stages lives in the space: each thread grows and frees only
// its own buffer. Calls go to the null stream, which orders every thread's
// copies and kernels; each call waits for its own results before returning.
struct Staging {
    void *base = nullptr;
    size_t cap = 0;
    ~Staging() {
        if (base) (void)hipFree(base);
    }
};

thread_local Staging t_staging;

bool arena_for(const Space *s, size_t bytes, Arena *a) {
    (void)s;
    bytes += 64 * 256;
    Staging &st = t_staging;
    if (st.cap < bytes) {
        if (st.base) (void)hipFree(st.base);
        st.base = nullptr;
        st.cap = 0;
        size_t cap = 1 << 20;
        while (cap < bytes) cap <<= 1;
        if (hipMalloc(&st.base, cap) != hipSuccess) return fail("bridge staging allocation failed") == 0;
        st.cap = cap;
    }
    a->base = static_cast<char *>(st.base);
    a->used = 0;
    a->cap = st.cap;
    return true;
}

struct StagedIndex {
    qg_index_desc d;
};

// Device SoA rows for `n` records carved from the arena.
bool stage_index(const Space *s, Arena *a, uint64_t n, qg_index_desc *d) {
    const size_t cd = s->cd;
    std::memset(d, 0, sizeof(*d));
    d->codes = static_cast<uint8_t *>(a->take(n * cd / 2));
    d->hdr = static_cast<float *>(a->take(n * 8));
    d->sfac = static_cast<float *>(a->take(n * 8));
    d->gate = static_cast<uint8_t *>(a->take(n * gate_row_bytes(cd)));
    d->rscale = static_cast<uint16_t *>(a->take(n * 2 * s->block_count));
    d->rcodes = static_cast<uint8_t *>(a->take(n * cd / 2));
    d->rfac = static_cast<float *>(a->take(n * 8));
    d->cid = static_cast<uint8_t *>(a->take(n));
    d->n = n;
    return d->codes && d->hdr && d->sfac && d->gate && d->rscale && d->rcodes && d->rfac && d->cid;
}

// The sidecar of `count` device gate rows in the reference's form: MSB rows
// of cd/8 bytes and PaperPruneFactors records.
bool download_sidecar(const Space *s, const qg_index_desc &d, size_t count, uint8_t *msb_out, void *factors_out) {
    const size_t gb = gate_row_bytes(s->cd), mb = s->cd / 8, m = gate_msb_span(s->cd);
    std::vector<uint8_t> g(count * gb);
    if (hipMemcpy(g.data(), d.gate, g.size(), hipMemcpyDeviceToHost) != hipSuccess) return false;
    RabitqPaperFactors *o = static_cast<RabitqPaperFactors *>(factors_out);
    for (size_t i = 0; i < count; ++i) {
        const uint8_t *row = g.data() + i * gb;
        std::memcpy(msb_out + i * mb, row, mb);
        float f[4];
        std::memcpy(f, row + m, 16);
        o[i].norm_sqr = f[0];
        o[i].data_norm = f[1];
        o[i].cross_scale = f[2];
        o[i].error_cross_scale = f[3];
        o[i].valid = row[m + 24];
    }
    return true;
}

// Thread-local prepared query: device buffers for one query.
struct BridgeQuery {
    const Space *owner = nullptr;
    int codec = 0;
    float *q = nullptr;
    int8_t *x8 = nullptr;
    int8_t *x8p = nullptr;
    void *qs = nullptr;
    void *ws = nullptr;
    float *raw = nullptr;
    size_t cd = 0, dim = 0;
    ~BridgeQuery() { release(); }
    void release() {
        if (q) (void)hipFree(q);
        if (x8) (void)hipFree(x8);
        if (x8p) (void)hipFree(x8p);
        if (qs) (void)hipFree(qs);
        if (ws) (void)hipFree(ws);
        if (raw) (void)hipFree(raw);
        q = nullptr; x8 = nullptr; x8p = nullptr; qs = nullptr; ws = nullptr; raw = nullptr;
    }
    size_t contexts = 0;  // one per centroid of the space
    bool ensure(const Space *s) {
        if (owner == s && cd == s->cd && dim == s->dim && contexts == s->centroid_count) return true;
        release();
        cd = s->cd;
        dim = s->dim;
        contexts = s->centroid_count;
        owner = s;
        const size_t K = contexts;
        return hipMalloc(&q, K * cd * 4) == hipSuccess && hipMalloc(&x8, K * cd) == hipSuccess &&
               hipMalloc(&x8p, K * cd) == hipSuccess && hipMalloc(&qs, K * 48) == hipSuccess &&
               hipMalloc(&ws, 256) == hipSuccess && hipMalloc(&raw, dim * 4) == hipSuccess;
    }
    qg_query_desc desc() const {
        qg_query_desc d;
        d.q = q; d.x8 = x8; d.x8p = x8p; d.qs = qs; d.nq = contexts;
        return d;
    }
};

thread_local BridgeQuery t_query;

// Gather rows ids[] of a strided host array into a dense host buffer.
void gather(const void *base, size_t stride, size_t row_bytes, const uint32_t *ids, size_t count,
            std::vector<uint8_t> *out) {
    out->resize(count * row_bytes);
    const uint8_t *b = static_cast<const uint8_t *>(base);
    for (size_t i = 0; i < count; ++i)
        std::memcpy(out->data() + i * row_bytes, b + static_cast<size_t>(ids[i]) * stride, row_bytes);
}

// Stage `count` compact records (+ optional residual records) and unpack
// them to SoA rows 0..count-1.
bool upload_records(const Space *s, Arena *a, const std::vector<uint8_t> &comp, const std::vector<uint8_t> *res,
                    size_t count, qg_index_desc *d) {
    if (!stage_index(s, a, count, d)) return false;
    void *dc = a->take(comp.size());
    void *dr = res ? a->take(res->size()) : nullptr;
    if (!dc || (res && !dr)) return false;
    if (hipMemcpy(dc, comp.data(), comp.size(), hipMemcpyHostToDevice) != hipSuccess) return false;
    if (res && hipMemcpy(dr, res->data(), res->size(), hipMemcpyHostToDevice) != hipSuccess) return false;
    if (!res) {
        (void)hipMemset(d->rscale, 0, count * 2 * s->block_count);
        (void)hipMemset(d->rcodes, 0, count * s->cd / 2);
    }
    return qg_unpack_records(s, static_cast<const uint8_t *>(dc), s->compact_bytes, static_cast<const uint8_t *>(dr),
                             s->residual_bytes, count, d, nullptr) == 0;
}

uint32_t *upload_iota(Arena *a, size_t count) {
    std::vector<uint32_t> h(count);
    for (size_t i = 0; i < count; ++i) h[i] = static_cast<uint32_t>(i);
    uint32_t *d = static_cast<uint32_t *>(a->take(count * 4 + 4));
    if (d && count) (void)hipMemcpy(d, h.data(), count * 4, hipMemcpyHostToDevice);
    return d;
}

uint32_t *upload_zeros_u32(Arena *a, size_t count) {
    uint32_t *d = static_cast<uint32_t *>(a->take(count * 4 + 4));
    if (d && count) (void)hipMemset(d, 0, count * 4);
    return d;
}

// The query context of each staged record: its centroid id (the compact
// record's last byte; queryForEncoded, space_rabitq.h:3404-3416), one
// prepared context per centroid.
uint32_t *upload_centroid_qidx(const Space *s, Arena *a, const std::vector<uint8_t> &comp, size_t count) {
    std::vector<uint32_t> h(count);
    for (size_t i = 0; i < count; ++i) h[i] = comp[i * s->compact_bytes + 16 + s->cd / 2];
    uint32_t *d = static_cast<uint32_t *>(a->take(count * 4 + 4));
    if (d && count) (void)hipMemcpy(d, h.data(), count * 4, hipMemcpyHostToDevice);
    return d;
}

bool centroid_ids_valid(const Space *s, const std::vector<uint8_t> &comp, size_t count) {
    for (size_t i = 0; i < count; ++i)
        if (comp[i * s->compact_bytes + 16 + s->cd / 2] >= s->centroid_count) return false;
    return true;
}

// Stage the compact records ids[] (strided host rows) and run one pair scan
// of `kernel` for the prepared query; out_f / out_est copied back.
bool scan_compact(const Space *s, const BridgeQuery *bq, int kernel, const uint32_t *ids, size_t count,
                  const void *base, size_t stride, float eps0, float *out_f, RabitqPaperEstimate *out_est) {
    std::vector<uint8_t> comp;
    gather(base, stride, s->compact_bytes, ids, count, &comp);
    Arena a;
    if (!arena_for(s, count * (s->compact_bytes + 2 * s->cd + 128) + 4096, &a)) return false;
    if (!centroid_ids_valid(s, comp, count)) return false;
    qg_index_desc d;
    if (!upload_records(s, &a, comp, nullptr, count, &d)) return false;
    uint32_t *rows = upload_iota(&a, count);
    uint32_t *qidx = upload_centroid_qidx(s, &a, comp, count);
    float *df = static_cast<float *>(a.take(count * 4 + 4));
    RabitqPaperEstimate *de = static_cast<RabitqPaperEstimate *>(a.take(count * 24 + 24));
    if (!rows || !qidx || !df || !de) return false;
    const qg_query_desc qd = bq->desc();
    if (qg_scan(s, kernel, 0, &d, &qd, rows, qidx, 0, count, eps0, nullptr, nullptr, out_f ? df : nullptr, nullptr,
                out_est ? de : nullptr, nullptr, nullptr, nullptr) != 0)
        return false;
    if (out_f && hipMemcpy(out_f, df, count * 4, hipMemcpyDeviceToHost) != hipSuccess) return false;
    if (out_est && hipMemcpy(out_est, de, count * 24, hipMemcpyDeviceToHost) != hipSuccess) return false;
    return true;
}

const BridgeQuery *query_of(const Space *s, const void *prepared) {
    const BridgeQuery *bq = static_cast<const BridgeQuery *>(prepared);
    return (s && bq && bq->owner == s) ? bq : nullptr;
}

}  // namespace
}  // namespace qg

using namespace qg;

extern "C" {

void *rabitq_space_new(size_t dim, uint32_t seed, size_t centroid_count, size_t residual_bits,
                       size_t residual_block_size, bool residual_mse, bool residual_fp16) {
    return space_create(dim, seed, centroid_count, residual_bits, residual_block_size, residual_mse, residual_fp16);
}

void rabitq_space_delete(void *space) { space_destroy(static_cast<Space *>(space)); }

bool rabitq_space_set_center(void *space, const float *center) {
    return space_set_center(static_cast<Space *>(space), center) == 0;
}

size_t rabitq_full_record_bytes(const void *space) { return static_cast<const Space *>(space)->full_bytes; }
size_t rabitq_compact_record_bytes(const void *space) { return static_cast<const Space *>(space)->compact_bytes; }
size_t rabitq_residual_record_bytes(const void *space) { return static_cast<const Space *>(space)->residual_bytes; }
size_t rabitq_residual_block_size(const void *space) { return static_cast<const Space *>(space)->block; }
size_t rabitq_residual_bits(const void *space) { (void)space; return 4; }
size_t rabitq_paper_msb_code_bytes(const void *space) { return (static_cast<const Space *>(space)->cd + 7) / 8; }
size_t rabitq_paper_factor_bytes(void) { return 20; }

bool rabitq_encode_full(const void *space, const float *raw, void *encoded_out) {
    const Space *s = static_cast<const Space *>(space);
    if (!s || !raw || !encoded_out) return false;
    Arena a;
    const size_t ws_bytes = qg_encode_workspace_bytes(s, 1);
    if (!arena_for(s, s->dim * 4 + ws_bytes + s->compact_bytes + s->residual_bytes + 4096 + s->cd * 8, &a)) return false;
    float *dx = static_cast<float *>(a.take(s->dim * 4));
    qg_index_desc d;
    if (!dx || !stage_index(s, &a, 1, &d)) return false;
    void *ws = a.take(ws_bytes);
    uint8_t *dc = static_cast<uint8_t *>(a.take(s->compact_bytes));
    uint8_t *dr = static_cast<uint8_t *>(a.take(s->residual_bytes));
    if (!ws || !dc || !dr) return false;
    if (hipMemcpy(dx, raw, s->dim * 4, hipMemcpyHostToDevice) != hipSuccess) return false;
    if (qg_encode(s, dx, 1, &d, ws, ws_bytes, nullptr) != 0) return false;
    if (qg_pack_records(s, &d, dc, s->compact_bytes, dr, s->residual_bytes, nullptr) != 0) return false;
    std::vector<uint8_t> c(s->compact_bytes), r(s->residual_bytes);
    if (hipMemcpy(c.data(), dc, c.size(), hipMemcpyDeviceToHost) != hipSuccess) return false;
    if (hipMemcpy(r.data(), dr, r.size(), hipMemcpyDeviceToHost) != hipSuccess) return false;
    // full record (space_rabitq.h:261-266): header | short factors |
    // residual factors | scales | codes | centroid id | residual codes
    uint8_t *o = static_cast<uint8_t *>(encoded_out);
    const size_t sb = 2 * s->block_count, cb = s->cd / 2;
    std::memcpy(o, c.data(), 16);
    std::memcpy(o + 16, r.data(), 8);
    std::memcpy(o + 24, r.data() + 8, sb);
    std::memcpy(o + 24 + sb, c.data() + 16, cb);
    o[24 + sb + cb] = c[16 + cb];  // centroid id (assignCentroid, space_rabitq.h:2944-2960)
    std::memcpy(o + 25 + sb + cb, r.data() + 8 + sb, cb);
    return true;
}

bool rabitq_copy_compact(const void *space, const void *full_encoded, void *compact_out) {
    const Space *s = static_cast<const Space *>(space);
    if (!s || !full_encoded || !compact_out) return false;
    const uint8_t *f = static_cast<const uint8_t *>(full_encoded);
    uint8_t *o = static_cast<uint8_t *>(compact_out);
    const size_t sb = 2 * s->block_count, cb = s->cd / 2;
    std::memcpy(o, f, 16);
    std::memcpy(o + 16, f + 24 + sb, cb);
    o[16 + cb] = f[24 + sb + cb];
    return true;
}

bool rabitq_copy_residual_record(const void *space, const void *full_encoded, void *residual_out) {
    const Space *s = static_cast<const Space *>(space);
    if (!s || !full_encoded || !residual_out) return false;
    const uint8_t *f = static_cast<const uint8_t *>(full_encoded);
    uint8_t *o = static_cast<uint8_t *>(residual_out);
    const size_t sb = 2 * s->block_count, cb = s->cd / 2;
    std::memcpy(o, f + 16, 8);
    std::memcpy(o + 8, f + 24, sb);
    std::memcpy(o + 8 + sb, f + 25 + sb + cb, cb);
    return true;
}

bool rabitq_export_paper_sidecar(const void *space, const void *compact_records_in, size_t record_count,
                                 size_t compact_stride, uint8_t *msb_out, void *factors_out) {
    const Space *s = static_cast<const Space *>(space);
    if (!s || (!compact_records_in && record_count) || !msb_out || !factors_out) return false;
    if (record_count == 0) return true;
    if (compact_stride < s->compact_bytes) return false;
    std::vector<uint32_t> ids(record_count);
    for (size_t i = 0; i < record_count; ++i) ids[i] = static_cast<uint32_t>(i);
    std::vector<uint8_t> comp;
    gather(compact_records_in, compact_stride, s->compact_bytes, ids.data(), record_count, &comp);
    Arena a;
    if (!arena_for(s, record_count * (s->compact_bytes + 2 * s->cd + 64), &a)) return false;
    qg_index_desc d;
    if (!upload_records(s, &a, comp, nullptr, record_count, &d)) return false;
    return download_sidecar(s, d, record_count, msb_out, factors_out);
}

// rabitq_build_vamana_graph (rabitq_bridge.cpp:224-303): the records are
// staged to the device once, the graph is built by qg_build_graph, and
// degrees, edges (degree entries per row; the rest of each row untouched),
// the paper sidecar and the evaluation counter are copied back. R <= 64 and
// L_build <= 1024 are limits of the device build (false beyond them).
bool rabitq_build_vamana_graph(void *space, const void *compact_records_in, size_t record_count,
                               size_t compact_stride, size_t r, size_t l_build, float alpha, size_t beam_width,
                               size_t batch_size, size_t refine_passes, size_t prune_candidate_cap,
                               size_t build_early_stop_hops, uint32_t *degrees_out, uint32_t *edges_out,
                               size_t edge_stride, uint8_t *paper_msb_out, void *paper_factors_out,
                               size_t *paper_msb_stride_out, size_t *paper_factor_bytes_out,
                               uint64_t *distance_evaluations_out) {
    Space *s = static_cast<Space *>(space);
    if (!s || (!compact_records_in && record_count) || !degrees_out || !edges_out || edge_stride < r ||
        !paper_msb_out || !paper_factors_out || !paper_msb_stride_out || !paper_factor_bytes_out)
        return false;
    if (compact_stride < s->compact_bytes) return false;
    if (r == 0 || l_build == 0 || beam_width == 0) return false;  // VamanaIndex constructor
    if (record_count > 1 && (!(alpha >= 1.0f) || !std::isfinite(alpha))) return false;
    if (r > 64 || l_build > 1024 || beam_width > 64 || batch_size > 0xFFFFFFFFu || refine_passes > 0xFFFFu ||
        prune_candidate_cap > 0xFFFFFFFFu || build_early_stop_hops > 0xFFFFFFFFu || record_count > 0xFFFFFFF0u)
        return false;
    *paper_msb_stride_out = (s->cd + 7) / 8;
    *paper_factor_bytes_out = sizeof(RabitqPaperFactors);
    if (record_count == 0) {
        if (distance_evaluations_out) *distance_evaluations_out = 0;
        return true;
    }
    std::vector<uint32_t> ids(record_count);
    for (size_t i = 0; i < record_count; ++i) ids[i] = static_cast<uint32_t>(i);
    std::vector<uint8_t> comp;
    gather(compact_records_in, compact_stride, s->compact_bytes, ids.data(), record_count, &comp);
    Arena a;
    if (!arena_for(s, record_count * (s->compact_bytes + 2 * s->cd + 64) + 2 * record_count * (r + 1) * 4, &a))
        return false;
    qg_index_desc d;
    if (!upload_records(s, &a, comp, nullptr, record_count, &d)) return false;
    uint32_t *deg = static_cast<uint32_t *>(a.take(record_count * 4));
    uint32_t *edges = static_cast<uint32_t *>(a.take(record_count * r * 4));
    if (!deg || !edges) return false;
    uint64_t evals = 0;
    std::vector<uint32_t> hdeg(record_count, 0), hedges(record_count * r);
    if (record_count > 1) {
        const size_t ws_bytes = qg_build_workspace_bytes(record_count, (uint32_t)r, (uint32_t)l_build,
                                                         batch_size ? (uint32_t)batch_size : 1u,
                                                         (uint32_t)refine_passes);
        void *ws = nullptr;
        if (ws_bytes == 0 || hipMalloc(&ws, ws_bytes) != hipSuccess) return false;
        const int rc = qg_build_graph(s, &d, (uint32_t)r, (uint32_t)l_build, alpha, (uint32_t)beam_width,
                                      (uint32_t)batch_size, (uint32_t)refine_passes, (uint32_t)prune_candidate_cap,
                                      (uint32_t)build_early_stop_hops, deg, edges, &evals, ws, ws_bytes, nullptr);
        (void)hipFree(ws);
        if (rc != 0) return false;
        if (hipMemcpy(hdeg.data(), deg, record_count * 4, hipMemcpyDeviceToHost) != hipSuccess) return false;
        if (hipMemcpy(hedges.data(), edges, record_count * r * 4, hipMemcpyDeviceToHost) != hipSuccess) return false;
    }
    if (!download_sidecar(s, d, record_count, paper_msb_out, paper_factors_out)) return false;
    for (size_t i = 0; i < record_count; ++i) {
        degrees_out[i] = hdeg[i];
        std::memcpy(edges_out + i * edge_stride, hedges.data() + i * r, hdeg[i] * 4);
    }
    if (distance_evaluations_out) *distance_evaluations_out = evals;
    return true;
}

float rabitq_symmetric_distance(const void *space, const void *lhs, const void *rhs) {
    const Space *s = static_cast<const Space *>(space);
    const float inf = std::numeric_limits<float>::infinity();
    if (!s || !lhs || !rhs) return inf;
    std::vector<uint8_t> comp(2 * s->compact_bytes);
    std::memcpy(comp.data(), lhs, s->compact_bytes);
    std::memcpy(comp.data() + s->compact_bytes, rhs, s->compact_bytes);
    Arena a;
    if (!arena_for(s, 2 * (s->compact_bytes + 2 * s->cd + 64) + 64, &a)) return inf;
    qg_index_desc d;
    if (!upload_records(s, &a, comp, nullptr, 2, &d)) return inf;
    uint32_t *l = upload_zeros_u32(&a, 1);
    uint32_t *r = upload_iota(&a, 2) + 1;
    float *out = static_cast<float *>(a.take(4));
    if (!l || !out) return inf;
    if (qg_scan(s, QG_SCAN_SYMMETRIC, 0, &d, nullptr, r, l, 0, 1, 0.0f, nullptr, nullptr, out, nullptr, nullptr,
                nullptr, nullptr, nullptr) != 0)
        return inf;
    float v = inf;
    if (hipMemcpy(&v, out, 4, hipMemcpyDeviceToHost) != hipSuccess) return inf;
    return v;
}

const void *rabitq_prepare_query(const void *space, const float *query) {
    const Space *s = static_cast<const Space *>(space);
    if (!s || !query) return nullptr;
    BridgeQuery &bq = t_query;
    if (!bq.ensure(s)) return nullptr;
    bq.codec = s->codec;
    if (hipMemcpy(bq.raw, query, s->dim * 4, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    qg_query_desc d = bq.desc();
    // The codecs this library prepares: Int8 / Int4 (dithered codes), B1 /
    // B1Main (sign code); every other codec reads the FP32 residual only.
    const int codec = (s->codec == QG_CODEC_INT8 || s->codec == QG_CODEC_INT4 || s->codec == QG_CODEC_B1 ||
                       s->codec == QG_CODEC_B1_MAIN)
                          ? s->codec : QG_CODEC_FULL;
    if (qg_prepare_queries(s, bq.raw, 1, codec, &d, bq.ws, nullptr) != 0) return nullptr;
    if (hipDeviceSynchronize() != hipSuccess) return nullptr;
    return &bq;
}

void rabitq_set_query_coarse_codec(void *space, int codec) {
    Space *s = static_cast<Space *>(space);
    if (s) s->codec = codec;
}

void rabitq_release_query(const void *space, const void *prepared) {
    (void)space;
    (void)prepared;
}

bool rabitq_query_distance_batch(const void *space, const void *prepared, const uint32_t *ids, const uint8_t *modes,
                                 size_t count, const void *encoded_base, size_t encoded_stride,
                                 const float *short_ips, float *out) {
    const Space *s = static_cast<const Space *>(space);
    const BridgeQuery *bq = static_cast<const BridgeQuery *>(prepared);
    if (!s || !bq || bq->owner != s || (count && (!ids || !modes || !encoded_base || !out))) return false;
    if (count == 0) return true;
    std::vector<uint8_t> comp;
    gather(encoded_base, encoded_stride, s->compact_bytes, ids, count, &comp);
    if (!centroid_ids_valid(s, comp, count)) return false;
    Arena a;
    if (!arena_for(s, count * (s->compact_bytes + 2 * s->cd + 96) + 4096, &a)) return false;
    qg_index_desc d;
    if (!upload_records(s, &a, comp, nullptr, count, &d)) return false;
    uint32_t *rows = upload_iota(&a, count);
    uint32_t *qidx = upload_centroid_qidx(s, &a, comp, count);
    float *dsi = static_cast<float *>(a.take(count * 4 + 4));
    float *dout = static_cast<float *>(a.take(count * 4 + 4));
    if (!rows || !qidx || !dsi || !dout) return false;
    std::vector<float> si(count, 0.0f);
    if (short_ips) std::memcpy(si.data(), short_ips, count * 4);
    (void)hipMemcpy(dsi, si.data(), count * 4, hipMemcpyHostToDevice);
    const qg_query_desc qd = bq->desc();
    // mode 0: full recompute; 1 / 3: FP32 reuse of the gate's short_ip;
    // 2: INT8 reuse (rabitq_bridge.cpp:885-970)
    std::vector<float> res(count), tmp(count);
    const int kernels[4] = {QG_SCAN_ASYM, QG_SCAN_REM_FP32, QG_SCAN_REM_INT8, QG_SCAN_REM_FP32};
    bool used[4] = {false, false, false, false};
    for (size_t i = 0; i < count; ++i) used[modes[i] < 4 ? modes[i] : 0] = true;
    for (int m = 0; m < 4; ++m) {
        if (!used[m]) continue;
        if (qg_scan(s, kernels[m], 0, &d, &qd, rows, qidx, 0, count, 0.0f, dsi, nullptr, dout, nullptr, nullptr,
                    nullptr, nullptr, nullptr) != 0)
            return false;
        if (hipMemcpy(tmp.data(), dout, count * 4, hipMemcpyDeviceToHost) != hipSuccess) return false;
        for (size_t i = 0; i < count; ++i)
            if ((modes[i] < 4 ? modes[i] : 0) == m) res[i] = tmp[i];
    }
    std::memcpy(out, res.data(), count * 4);
    return true;
}

float rabitq_query_distance(const void *space, const void *prepared, const void *encoded) {
    const uint32_t id = 0;
    const uint8_t mode = 0;
    float out = std::numeric_limits<float>::infinity();
    const Space *s = static_cast<const Space *>(space);
    if (!s || !encoded) return out;
    if (!rabitq_query_distance_batch(space, prepared, &id, &mode, 1, encoded, s->compact_bytes, nullptr, &out))
        return std::numeric_limits<float>::infinity();
    return out;
}

bool rabitq_paper_estimate_batch_sidecar(const void *space, const void *prepared, const uint32_t *ids, size_t count,
                                         const uint8_t *msb_base, size_t msb_stride, const void *factors_base,
                                         size_t factors_stride, float epsilon0, RabitqPaperEstimate *out) {
    const Space *s = static_cast<const Space *>(space);
    const BridgeQuery *bq = static_cast<const BridgeQuery *>(prepared);
    if (!s || !bq || bq->owner != s || (count && (!ids || !msb_base || !factors_base || !out))) return false;
    if (count == 0) return true;
    // rabitq_paper_estimate_batch_sidecar (rabitq_bridge.cpp:702-883): B1
    // takes its symmetric sign-code gate, Int4 and Int8 the integer gate on
    // their dithered codes, every other codec the FP32 gate.
    // the caller's sidecar rows as device gate rows (the header copy is not
    // read by the gate and stays zero)
    const size_t mb = s->cd / 8, gb = gate_row_bytes(s->cd), m = gate_msb_span(s->cd);
    std::vector<uint8_t> rowsg(count * gb, 0);
    const uint8_t *mbase = static_cast<const uint8_t *>(msb_base);
    const uint8_t *fb = static_cast<const uint8_t *>(factors_base);
    for (size_t i = 0; i < count; ++i) {
        uint8_t *row = rowsg.data() + i * gb;
        std::memcpy(row, mbase + static_cast<size_t>(ids[i]) * msb_stride, mb);
        RabitqPaperFactors f;
        std::memcpy(&f, fb + static_cast<size_t>(ids[i]) * factors_stride, 20);
        const float fv[4] = {f.norm_sqr, f.data_norm, f.cross_scale, f.error_cross_scale};
        std::memcpy(row + m, fv, 16);
        row[m + 24] = f.valid;
    }
    Arena a;
    if (!arena_for(s, count * (gb + 64) + 4096, &a)) return false;
    qg_index_desc d;
    std::memset(&d, 0, sizeof(d));
    d.gate = static_cast<uint8_t *>(a.take(count * gb));
    d.n = count;
    uint32_t *rows = upload_iota(&a, count);
    uint32_t *qidx = upload_zeros_u32(&a, count);
    RabitqPaperEstimate *dout = static_cast<RabitqPaperEstimate *>(a.take(count * 24));
    if (!d.gate || !rows || !qidx || !dout) return false;
    (void)hipMemcpy(d.gate, rowsg.data(), rowsg.size(), hipMemcpyHostToDevice);
    const qg_query_desc qd = bq->desc();
    const int k = (bq->codec == QG_CODEC_INT8 || bq->codec == QG_CODEC_INT4) ? QG_SCAN_GATE_INT8
                  : bq->codec == QG_CODEC_B1 ? QG_SCAN_GATE_B1 : QG_SCAN_GATE_FP32;
    if (qg_scan(s, k, 0, &d, &qd, rows, qidx, 0, count, epsilon0, nullptr, nullptr, nullptr, nullptr, dout, nullptr,
                nullptr, nullptr) != 0)
        return false;
    return hipMemcpy(out, dout, count * 24, hipMemcpyDeviceToHost) == hipSuccess;
}

}  // extern "C"

namespace qg {
namespace {
bool residual_batch(const void *space, const void *prepared, const uint32_t *ids, size_t count,
                    const void *encoded_base, size_t encoded_stride, const void *residual_base,
                    size_t residual_stride, const float *long_distances, RabitqDistanceInterval *out, bool single) {
    const Space *s = static_cast<const Space *>(space);
    const BridgeQuery *bq = static_cast<const BridgeQuery *>(prepared);
    if (!s || !bq || bq->owner != s ||
        (count && (!ids || !encoded_base || !residual_base || !long_distances || !out)))
        return false;
    if (count == 0) return true;
    std::vector<uint8_t> comp, res;
    gather(encoded_base, encoded_stride, s->compact_bytes, ids, count, &comp);
    gather(residual_base, residual_stride, s->residual_bytes, ids, count, &res);
    if (!centroid_ids_valid(s, comp, count)) return false;
    Arena a;
    if (!arena_for(s, count * (s->compact_bytes + s->residual_bytes + 2 * s->cd + 96) + 4096, &a)) return false;
    qg_index_desc d;
    if (!upload_records(s, &a, comp, &res, count, &d)) return false;
    uint32_t *rows = upload_iota(&a, count);
    uint32_t *qidx = upload_centroid_qidx(s, &a, comp, count);
    float *dl = static_cast<float *>(a.take(count * 4));
    RabitqDistanceInterval *dout = static_cast<RabitqDistanceInterval *>(a.take(count * 12));
    if (!rows || !qidx || !dl || !dout) return false;
    (void)hipMemcpy(dl, long_distances, count * 4, hipMemcpyHostToDevice);
    const qg_query_desc qd = bq->desc();
    // the batch loop holds one centroid; otherwise the reference scores each
    // record alone (space_rabitq.h:5446-5461)
    const int kernel = (s->centroid_count > 1 || single) ? QG_SCAN_RESIDUAL_RECORD : QG_SCAN_RESIDUAL;
    if (qg_scan(s, kernel, 0, &d, &qd, rows, qidx, 0, count, 0.0f, nullptr, dl, nullptr, nullptr, nullptr,
                dout, nullptr, nullptr) != 0) {
        for (size_t i = 0; i < count; ++i) out[i] = RabitqDistanceInterval{long_distances[i], long_distances[i], long_distances[i]};
        return true;
    }
    return hipMemcpy(out, dout, count * 12, hipMemcpyDeviceToHost) == hipSuccess;
}
}  // namespace
}  // namespace qg

extern "C" {

bool rabitq_residual_distance_batch(const void *space, const void *prepared, const uint32_t *ids, size_t count,
                                    const void *encoded_base, size_t encoded_stride, const void *residual_base,
                                    size_t residual_stride, const float *long_distances, RabitqDistanceInterval *out) {
    return residual_batch(space, prepared, ids, count, encoded_base, encoded_stride, residual_base, residual_stride,
                          long_distances, out, false);
}

// ------------------------------------------- single-record and remaining ---

bool rabitq_space_set_centroids(void *space, const float *centroids, size_t centroid_count) {
    // setCentroids (space_rabitq.h:2918-2936): the count must match the
    // space's centroid count.
    if (!space || !centroids) return false;
    return space_set_centroids(static_cast<Space *>(space), centroids, centroid_count) == 0;
}

bool rabitq_train_centroids(const float *data, size_t count, size_t dim, size_t k, size_t sample_count,
                            uint32_t seed, int iters, float *out) {
    if (data == nullptr || out == nullptr || count == 0 || dim == 0 || k == 0 || k > 256 || sample_count == 0)
        return false;
    sample_count = std::min(sample_count, count);
    // The sample draw (std::mt19937(seed)() % count, with replacement) is the
    // reference's index stream; the rows are staged in draw order.
    std::mt19937 rng(seed);
    std::vector<float> rows(sample_count * dim);
    for (size_t i = 0; i < sample_count; ++i) {
        const size_t id = static_cast<size_t>(rng()) % count;
        std::memcpy(rows.data() + i * dim, data + id * dim, dim * sizeof(float));
    }
    const size_t ws_bytes = qg_kmeans_workspace_bytes(sample_count, dim, (uint32_t)k);
    float *d_rows = nullptr, *d_cen = nullptr;
    void *d_ws = nullptr;
    bool ok = hipMalloc(&d_rows, rows.size() * sizeof(float)) == hipSuccess &&
              hipMalloc(&d_cen, k * dim * sizeof(float)) == hipSuccess && hipMalloc(&d_ws, ws_bytes) == hipSuccess;
    ok = ok && hipMemcpy(d_rows, rows.data(), rows.size() * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && qg_train_centroids(d_rows, sample_count, dim, (uint32_t)k, iters, d_cen, d_ws, ws_bytes, nullptr) == 0;
    ok = ok && hipMemcpy(out, d_cen, k * dim * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess;
    if (d_rows) (void)hipFree(d_rows);
    if (d_cen) (void)hipFree(d_cen);
    if (d_ws) (void)hipFree(d_ws);
    return ok;
}

bool rabitq_query_distance_b1_batch(const void *space, const void *prepared, const uint32_t *ids, size_t count,
                                    const void *encoded_base, size_t encoded_stride, float *out_distances) {
    const Space *s = static_cast<const Space *>(space);
    const BridgeQuery *bq = query_of(s, prepared);
    if (!bq || (count && (!ids || !encoded_base || !out_distances))) return false;
    const bool b1 = bq->codec == QG_CODEC_B1 || bq->codec == QG_CODEC_B1_MAIN;
    if (!b1 || !scan_compact(s, bq, QG_SCAN_B1, ids, count, encoded_base, encoded_stride, 0.0f, out_distances,
                             nullptr))
        for (size_t i = 0; i < count; ++i) out_distances[i] = std::numeric_limits<float>::infinity();
    return true;
}

float rabitq_query_distance_b1(const void *space, const void *prepared, const void *encoded) {
    const Space *s = static_cast<const Space *>(space);
    const uint32_t id = 0;
    float out = std::numeric_limits<float>::infinity();
    if (!s || !encoded) return out;
    if (!rabitq_query_distance_b1_batch(space, prepared, &id, 1, encoded, s->compact_bytes, &out))
        return std::numeric_limits<float>::infinity();
    return out;
}

bool rabitq_paper_estimate_batch(const void *space, const void *prepared, const uint32_t *ids, size_t count,
                                 const void *encoded_base, size_t encoded_stride, float epsilon0,
                                 RabitqPaperEstimate *out) {
    const Space *s = static_cast<const Space *>(space);
    const BridgeQuery *bq = query_of(s, prepared);
    if (!bq || (count && (!ids || !encoded_base || !out))) return false;
    if (count == 0) return true;
    if (!scan_compact(s, bq, QG_SCAN_PAPER_RECORD, ids, count, encoded_base, encoded_stride, epsilon0, nullptr, out))
        std::memset(out, 0, count * sizeof(RabitqPaperEstimate));
    return true;
}

RabitqPaperEstimate rabitq_paper_estimate(const void *space, const void *prepared, const void *encoded,
                                          float epsilon0) {
    RabitqPaperEstimate r;
    std::memset(&r, 0, sizeof(r));
    const Space *s = static_cast<const Space *>(space);
    const uint32_t id = 0;
    if (!s || !encoded || !rabitq_paper_estimate_batch(space, prepared, &id, 1, encoded, s->compact_bytes, epsilon0, &r))
        std::memset(&r, 0, sizeof(r));
    return r;
}

bool rabitq_paper_estimate_codec_batch(const void *space, const void *prepared, const uint32_t *ids,
                                       size_t count, const void *encoded_base, size_t encoded_stride,
                                       const uint8_t *msb_base, size_t msb_stride, const void *factors_base,
                                       size_t factors_stride, float epsilon0, RabitqPaperEstimate *out) {
    // The codec comparison path always reads the 1-bit sidecar; the compact
    // records are not used (rabitq_bridge.cpp:506-593).
    (void)encoded_base;
    (void)encoded_stride;
    return rabitq_paper_estimate_batch_sidecar(space, prepared, ids, count, msb_base, msb_stride, factors_base,
                                               factors_stride, epsilon0, out);
}

RabitqPaperEstimate rabitq_paper_estimate_sidecar(const void *space, const void *prepared, const uint8_t *msb_code,
                                                  const void *factors, float epsilon0) {
    RabitqPaperEstimate r;
    std::memset(&r, 0, sizeof(r));
    const Space *s = static_cast<const Space *>(space);
    const uint32_t id = 0;
    if (!s || !msb_code || !factors ||
        !rabitq_paper_estimate_batch_sidecar(space, prepared, &id, 1, msb_code, s->cd / 8, factors, 20, epsilon0, &r))
        std::memset(&r, 0, sizeof(r));
    return r;
}

RabitqPaperEstimate rabitq_paper_estimate_codec(const void *space, const void *prepared, const void *encoded,
                                                const uint8_t *msb_code, const void *factors, float epsilon0) {
    (void)encoded;
    return rabitq_paper_estimate_sidecar(space, prepared, msb_code, factors, epsilon0);
}

static float single_mode_distance(const void *space, const void *prepared, const void *encoded, float short_ip,
                                  uint8_t mode) {
    const Space *s = static_cast<const Space *>(space);
    const uint32_t id = 0;
    float out = std::numeric_limits<float>::infinity();
    if (!s || !encoded ||
        !rabitq_query_distance_batch(space, prepared, &id, &mode, 1, encoded, s->compact_bytes, &short_ip, &out))
        return std::numeric_limits<float>::infinity();
    return out;
}

float rabitq_query_distance_with_paper(const void *space, const void *prepared, const void *encoded,
                                       float short_ip) {
    return single_mode_distance(space, prepared, encoded, short_ip, 1);
}

float rabitq_query_distance_with_quantized_paper(const void *space, const void *prepared, const void *encoded,
                                                 float short_ip) {
    return single_mode_distance(space, prepared, encoded, short_ip, 2);
}

RabitqDistanceInterval rabitq_residual_distance(const void *space, const void *prepared, const void *encoded,
                                                float long_distance) {
    // Spaces made by rabitq_space_new store residuals externally, and the
    // inline-record form throws for them (space_rabitq.h:1706-1712); the
    // bridge turns that into {long, long, long} (rabitq_bridge.cpp:637-641).
    (void)space;
    (void)prepared;
    (void)encoded;
    return RabitqDistanceInterval{long_distance, long_distance, long_distance};
}

RabitqDistanceInterval rabitq_residual_distance_from_record(const void *space, const void *prepared,
                                                            const void *encoded, const void *residual_record,
                                                            float long_distance) {
    const Space *s = static_cast<const Space *>(space);
    const uint32_t id = 0;
    RabitqDistanceInterval iv{long_distance, long_distance, long_distance};
    if (!s || !encoded || !residual_record ||
        !residual_batch(space, prepared, &id, 1, encoded, s->compact_bytes, residual_record, s->residual_bytes,
                        &long_distance, &iv, true))
        return RabitqDistanceInterval{long_distance, long_distance, long_distance};
    return iv;
}

bool rabitq_full_residual_distance_batch(const void *space, const void *prepared, const uint32_t *ids,
                                         size_t count, const void *encoded_base, size_t encoded_stride,
                                         const void *residual_base, size_t residual_stride,
                                         RabitqDistanceInterval *out) {
    const Space *s = static_cast<const Space *>(space);
    const BridgeQuery *bq = query_of(s, prepared);
    if (!bq || (count && (!ids || !encoded_base || !residual_base || !out))) return false;
    if (count == 0) return true;
    const float inf = std::numeric_limits<float>::infinity();
    std::vector<uint8_t> comp, res;
    gather(encoded_base, encoded_stride, s->compact_bytes, ids, count, &comp);
    gather(residual_base, residual_stride, s->residual_bytes, ids, count, &res);
    Arena a;
    bool ok = arena_for(s, count * (s->compact_bytes + s->residual_bytes + 2 * s->cd + 128) + 4096, &a);
    qg_index_desc d;
    ok = ok && upload_records(s, &a, comp, &res, count, &d);
    ok = ok && centroid_ids_valid(s, comp, count);
    uint32_t *rows = ok ? upload_iota(&a, count) : nullptr;
    uint32_t *qidx = ok ? upload_centroid_qidx(s, &a, comp, count) : nullptr;
    float *dl = ok ? static_cast<float *>(a.take(count * 4)) : nullptr;
    RabitqDistanceInterval *dout = ok ? static_cast<RabitqDistanceInterval *>(a.take(count * 12)) : nullptr;
    ok = rows && qidx && dl && dout;
    const qg_query_desc qd = bq->desc();
    // mode-0 long distance, then the residual interval from it, on the device
    ok = ok && qg_scan(s, QG_SCAN_ASYM, 0, &d, &qd, rows, qidx, 0, count, 0.0f, nullptr, nullptr, dl, nullptr,
                       nullptr, nullptr, nullptr, nullptr) == 0;
    ok = ok && qg_scan(s, s->centroid_count > 1 ? QG_SCAN_RESIDUAL_RECORD : QG_SCAN_RESIDUAL, 0, &d, &qd, rows, qidx,
                       0, count, 0.0f, nullptr, dl, nullptr, nullptr,
                       nullptr, dout, nullptr, nullptr) == 0;
    ok = ok && hipMemcpy(out, dout, count * 12, hipMemcpyDeviceToHost) == hipSuccess;
    if (!ok)
        for (size_t i = 0; i < count; ++i) out[i] = RabitqDistanceInterval{inf, inf, inf};
    return true;
}

}  // extern "C"
