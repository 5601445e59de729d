// qg_api.cpp -- C-ABI of libqgraph_hip.so: space lifecycle and the device
// batch entry points declared in include/qgraph_rabitq.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "qg_args.h"
#include "qg_space.h"

// kernel launchers (qg_query.hip, qg_scan.hip, qg_select.hip, qg_encode.hip)
extern "C" {
int qg_launch_row_seed(const float *rows, uint64_t n, uint32_t dim, const float *center, uint64_t *seed,
                       double *norm, double *dot, hipStream_t stream);
int qg_launch_rotate_rows(const float *rows, uint64_t n, uint32_t dim, uint32_t cd, int identity,
                          const float *signs, float *out, hipStream_t stream);
int qg_launch_query_prep(const qg::PrepParams *p, hipStream_t stream);
int qg_launch_scan(int kernel, int fixed, const qg::ScanArgs *a, hipStream_t stream);
int qg_launch_topk(const qg::SelectArgs *a, hipStream_t stream);
int qg_launch_encode(const qg::EncArgs *a, hipStream_t stream);
int qg_launch_pack(const qg::PackArgs *a, uint64_t n, int unpack, hipStream_t stream);
int qg_launch_search(const qg::SearchArgs *a, uint32_t slots, uint32_t waves, hipStream_t stream);
int qg_launch_mlp_forward(const float *P, const uint16_t *X, uint64_t n, float *y, hipStream_t st);
int qg_launch_mlp_pack(const float *P, void *packed, hipStream_t st);
uint32_t qg_mlp_grad_waves(uint64_t n);
int qg_launch_mlp_grad(const float *P, const uint16_t *X, const float *t, uint64_t n, float tau, float *partial,
                       float *partial_loss, hipStream_t st);
int qg_launch_mlp_reduce(const float *partial, const float *partial_loss, uint32_t waves, float *grad,
                         float *loss_out, float *P, float *m, float *v, float lr, float b1, float b2, float eps,
                         float bc1, float bc2, hipStream_t st);
int qg_launch_residual_stats(const uint16_t *rscale, uint64_t n, uint32_t nb, float4 *out, hipStream_t st);
int qg_launch_rerank_features(const qg::RerankArgs *a, hipStream_t st);
int qg_launch_learned_rerank(const qg::RerankArgs *a, hipStream_t st);
size_t qg_kmeans_workspace_bytes(uint64_t sample_count, uint64_t dim, uint32_t k);
int qg_launch_kmeans(const float *samples, uint64_t sc, uint32_t dim, uint32_t k, int iters, float *centroids,
                     void *workspace, hipStream_t stream);
int qg_launch_qr(double *A, uint32_t cd, double *r_scratch, uint32_t *bad, hipStream_t st);
int qg_launch_centroid_pairs(const float *centroids, const float *rot_centroids, uint32_t K, uint32_t dim,
                             uint32_t cd, float2 *pairs, hipStream_t st);
int qg_launch_qr_split(const double *QT, uint32_t cd, uint32_t dim, uint32_t kpad, uint32_t parts, uint16_t *qp,
                       hipStream_t st);
int qg_launch_rot_gemm(const qg::RotGemmArgs *a, hipStream_t st);
int qg_launch_ss_encode(const float *x, uint64_t n, uint32_t dim, uint32_t cd, int identity, const float *signs,
                        uint64_t *db, hipStream_t st);
int qg_launch_ss_prepare(const float *raw, uint64_t nq, uint32_t dim, uint32_t cd, int identity, const float *signs,
                         float *q_out, uint64_t *bits_out, qg::SsQuery *qs_out, hipStream_t st);
int qg_launch_ss_scan(int codec, const uint64_t *db, uint64_t n, uint32_t W, const float *q, const uint64_t *bits,
                      const qg::SsQuery *qs, uint64_t nq, float *out, hipStream_t st);
}

namespace qg {

thread_local std::string g_last_error;

int fail(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return -1;
}

int check_hip(int rc, const char *what) {
    if (rc != 0) return fail("%s: %s", what, hipGetErrorString((hipError_t)rc));
    return 0;
}

size_t round_up64_pow2(size_t value) {  // roundUp64, space_rabitq.h:217-224
    const size_t rounded = ((value + 63u) / 64u) * 64u;
    size_t p = 1;
    while (p < rounded) p <<= 1u;
    return p;
}

Space *space_create(size_t dim, uint32_t seed, size_t centroid_count, size_t residual_bits, size_t block,
                    bool mse, bool fp16) {
    if (dim == 0) { fail("dimension must be positive"); return nullptr; }
    if (centroid_count == 0 || centroid_count > 256) { fail("centroid_count must be in [1, 256]"); return nullptr; }
    if (residual_bits != 4) { fail("only 4-bit residual codes are implemented"); return nullptr; }
    if (block != 16) { fail("only residual block size 16 is implemented"); return nullptr; }
    if (!fp16) { fail("only fp16 residual scale storage is implemented"); return nullptr; }
    Space *s = new Space();
    s->dim = dim;
    s->cd = round_up64_pow2(dim);
    // encode, query preparation, every estimator and the bridge take code
    // dimensions up to 4096; the graph build and search, the learned rerank
    // and the 1-bit sign scan up to 2048 (each refuses larger spaces)
    if (s->cd > kMaxCodeDim) { delete s; fail("code dimension %zu exceeds %u", round_up64_pow2(dim), kMaxCodeDim); return nullptr; }
    s->seed = seed;
    s->block = block;
    s->block_count = (s->cd + block - 1) / block;
    s->mse = mse;
    s->fp16 = fp16;
    s->codec = 0;
    s->compact_bytes = 16 + s->cd / 2 + 1;
    s->residual_bytes = 8 + 2 * s->block_count + s->cd / 2;
    s->full_bytes = 24 + 2 * s->block_count + s->cd / 2 + 1 + s->cd / 2;
    // setRandomRotation (space_rabitq.h:2900-2906): the same libstdc++
    // engine and distribution, so the sign sequence is identical.
    const float inv_sqrt = 1.0f / std::sqrt(static_cast<float>(s->cd));
    s->h_signs.assign(s->cd, inv_sqrt);
    if (seed != 0) {
        std::mt19937 rng(seed);
        std::uniform_int_distribution<int> bernoulli(0, 1);
        for (size_t i = 0; i < s->cd; ++i) s->h_signs[i] = (bernoulli(rng) ? 1.0f : -1.0f) * inv_sqrt;
    }
    const size_t K = centroid_count;
    s->centroid_count = K;
    s->h_center.assign(K * dim, 0.0f);
    s->h_cnorm.assign(K, 0.0f);
    if (check_hip(hipMalloc(&s->d_signs, s->cd * 4), "hipMalloc signs") ||
        check_hip(hipMalloc(&s->d_center, K * dim * 4), "hipMalloc center") ||
        check_hip(hipMalloc(&s->d_rot_center, K * s->cd * 4), "hipMalloc rotated center") ||
        check_hip(hipMalloc(&s->d_pair, K * K * sizeof(float2)), "hipMalloc centroid pairs")) {
        space_destroy(s);
        return nullptr;
    }
    if (check_hip(hipMemcpy(s->d_signs, s->h_signs.data(), s->cd * 4, hipMemcpyHostToDevice), "upload signs") ||
        check_hip(hipMemset(s->d_center, 0, K * dim * 4), "clear center") ||
        check_hip(hipMemset(s->d_rot_center, 0, K * s->cd * 4), "clear rotated center") ||
        check_hip(hipMemset(s->d_pair, 0, K * K * sizeof(float2)), "clear centroid pairs")) {
        space_destroy(s);
        return nullptr;
    }
    s->center_norm_sqr = 0.0f;
    return s;
}

void space_destroy(Space *s) {
    if (!s) return;
    // teardown: a failed free leaves nothing to recover, so its status is dropped
    if (s->d_signs) (void)hipFree(s->d_signs);
    if (s->d_center) (void)hipFree(s->d_center);
    if (s->d_rot_center) (void)hipFree(s->d_rot_center);
    if (s->d_pair) (void)hipFree(s->d_pair);
    if (s->d_qt) (void)hipFree(s->d_qt);
    if (s->d_qp) (void)hipFree(s->d_qp);
    delete s;
}

// Rows through the QR rotation's GEMM: out = Q * pad(row - center) (center
// subtracted when requested).
int qr_rotate(const Space *s, const float *rows, uint64_t n, bool subtract_center, float *out, hipStream_t st) {
    RotGemmArgs g;
    g.x = rows;
    g.n = n;
    g.dim = (uint32_t)s->dim;
    g.cd = (uint32_t)s->cd;
    g.kpad = s->kpad;
    g.center = subtract_center ? s->d_center : nullptr;
    g.row_scale = nullptr;
    g.q = s->d_qp;
    g.parts = s->parts;
    g.out = out;
    return qg_launch_rot_gemm(&g, st);
}

// setCentroids (space_rabitq.h:2918-2936): the centroids, each one's
// |c|^2 (FP64 sum, stored as float) and rotation, and for K > 1 the
// pairwise terms of buildCentroidDeltaContexts (:2564-2610).
int space_set_centroids(Space *s, const float *centroids_host, size_t count) {
    if (!s || !centroids_host) return fail("null space or centroids");
    const size_t K = s->centroid_count;
    if (count != K) return fail("the space holds %zu centroids, not %zu", K, count);
    s->h_center.assign(centroids_host, centroids_host + K * s->dim);
    if (check_hip(hipMemcpy(s->d_center, centroids_host, K * s->dim * 4, hipMemcpyHostToDevice), "upload centroids"))
        return -1;
    double *d_norm = nullptr;
    if (check_hip(hipMalloc(&d_norm, K * sizeof(double)), "hipMalloc")) return -1;
    int rc = qg_launch_row_seed(s->d_center, K, (uint32_t)s->dim, nullptr, nullptr, d_norm, nullptr, 0);
    if (!rc && s->rotation == 1) rc = qr_rotate(s, s->d_center, K, false, s->d_rot_center, 0);
    else if (!rc) rc = qg_launch_rotate_rows(s->d_center, K, (uint32_t)s->dim, (uint32_t)s->cd, s->seed == 0,
                                             s->d_signs, s->d_rot_center, 0);
    if (!rc && K > 1)
        rc = qg_launch_centroid_pairs(s->d_center, s->d_rot_center, (uint32_t)K, (uint32_t)s->dim, (uint32_t)s->cd,
                                      s->d_pair, 0);
    std::vector<double> norm(K, 0.0);
    if (!rc) rc = hipMemcpy(norm.data(), d_norm, K * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d_norm);
    if (rc) return check_hip(rc, "set centroids");
    for (size_t k = 0; k < K; ++k) s->h_cnorm[k] = static_cast<float>(norm[k]);
    s->center_norm_sqr = s->h_cnorm[0];
    return 0;
}

// setGlobalCenter: setCentroids with one centroid, which fails for K > 1.
int space_set_center(Space *s, const float *center_host) {
    if (!s || !center_host) return fail("null space or center");
    if (s->centroid_count != 1) return fail("the space holds %zu centroids: set them all", s->centroid_count);
    return space_set_centroids(s, center_host, 1);
}

IndexView view_of(const qg_index_desc *d) {
    IndexView v;
    v.codes = d->codes;
    v.hdr = reinterpret_cast<const float2 *>(d->hdr);
    v.sfac = reinterpret_cast<const float2 *>(d->sfac);
    v.gate = d->gate;
    v.rscale = d->rscale;
    v.rcodes = d->rcodes;
    v.n = d->n;
    v.cid = d->cid;
    return v;
}

QueryView view_of(const qg_query_desc *d) {
    QueryView v;
    v.q = d->q;
    v.x8 = d->x8;
    v.x8p = d->x8p;
    v.qs = reinterpret_cast<const QScalars *>(d->qs);
    v.nq = d->nq;
    return v;
}

// The encode keeps every intermediate in LDS; only the QR rotation's GEMM
// output passes through the workspace.
size_t encode_row_ws_bytes(const Space *s) { return s->rotation == 1 ? s->cd * 4 : 16; }

int encode(const Space *s, const float *x, uint64_t n, qg_index_desc *out, void *ws, size_t ws_bytes,
           hipStream_t stream) {
    if (!s || (!x && n) || !out) return fail("null argument to encode");
    if (n == 0) return 0;
    if (!out->codes || !out->hdr || !out->sfac || !out->gate || !out->rscale ||
        !out->rcodes || !out->rfac)
        return fail("encode output view has a null array");
    if (s->centroid_count > 1 && !out->cid) return fail("a space of several centroids encodes into a centroid-id array");
    if (s->centroid_count > 1 && s->rotation != 0) return fail("the QR rotation holds one centroid");
    const size_t per = encode_row_ws_bytes(s);
    uint64_t batch = ws_bytes / per;
    if (batch == 0 || !ws) return fail("encode workspace too small: need at least %zu bytes", per);
    if (batch > (1u << 24)) batch = 1u << 24;  // one launch's grid
    const uint32_t cd = (uint32_t)s->cd;
    for (uint64_t b0 = 0; b0 < n; b0 += batch) {
        const uint64_t m = (n - b0 < batch) ? n - b0 : batch;
        char *w = static_cast<char *>(ws);
        EncArgs a;
        a.x = x + b0 * s->dim;
        a.n = m;
        a.dim = (uint32_t)s->dim;
        a.cd = cd;
        a.block = (uint32_t)s->block;
        a.block_count = (uint32_t)s->block_count;
        a.identity = s->seed == 0;
        a.mse = s->mse;
        a.precision = s->rotation == 1 ? 1 : 0;
        a.signs = s->d_signs;
        a.center = s->d_center;
        a.centroid_count = (uint32_t)s->centroid_count;
        a.cid = out->cid ? out->cid + b0 : nullptr;
        float *rot = reinterpret_cast<float *>(w);
        a.rot = rot;
        a.codes = out->codes + b0 * (cd / 2);
        a.hdr = reinterpret_cast<float2 *>(out->hdr) + b0;
        a.sfac = reinterpret_cast<float2 *>(out->sfac) + b0;
        a.gate = out->gate + b0 * gate_row_bytes(cd);
        a.rscale = out->rscale + b0 * s->block_count;
        a.rcodes = out->rcodes + b0 * (cd / 2);
        a.rfac = reinterpret_cast<float2 *>(out->rfac) + b0;
        // the QR rotation of (x - c) on the matrix cores fills rot; the encode
        // kernels then normalise it
        if (a.precision && check_hip(qr_rotate(s, a.x, m, true, rot, stream), "rotation GEMM launch")) return -1;
        if (check_hip(qg_launch_encode(&a, stream), "encode launch")) return -1;
    }
    return 0;
}

}  // namespace qg

using namespace qg;

extern "C" {

const char *qg_last_error(void) { return g_last_error.c_str(); }

void *qg_space_new(size_t dim, uint32_t seed, size_t centroid_count, size_t residual_bits, size_t block,
                   bool mse, bool fp16) {
    return space_create(dim, seed, centroid_count, residual_bits, block, mse, fp16);
}

void qg_space_delete(void *space) { space_destroy(static_cast<Space *>(space)); }

int qg_space_set_center(void *space, const float *center_host) {
    return space_set_center(static_cast<Space *>(space), center_host);
}

int qg_space_set_centroids(void *space, const float *centroids_host, size_t count) {
    return space_set_centroids(static_cast<Space *>(space), centroids_host, count);
}

size_t qg_space_centroid_count(const void *space) {
    return space ? static_cast<const Space *>(space)->centroid_count : 0;
}

int qg_space_set_rotation_qr(void *space, const double *gaussian, int bf16_parts) {
    Space *s = static_cast<Space *>(space);
    if (!s || !gaussian) return fail("null space or Gaussian matrix");
    if (bf16_parts != 2 && bf16_parts != 3) return fail("bf16_parts must be 2 or 3");
    const size_t cd = s->cd;
    if (cd < 128) return fail("the QR rotation needs a code dimension of at least 128 (dimension above 64)");
    if (cd > kMaxGraphCodeDim) return fail("the QR rotation takes code dimensions up to %u", kMaxGraphCodeDim);
    if (s->centroid_count != 1) return fail("the QR rotation holds one centroid");
    // A = G^T: row j of A is column j of G
    std::vector<double> at(cd * cd);
    for (size_t i = 0; i < cd; ++i)
        for (size_t j = 0; j < cd; ++j) at[j * cd + i] = gaussian[i * cd + j];
    if (!s->d_qt && check_hip(hipMalloc(&s->d_qt, cd * cd * 8), "hipMalloc rotation")) return -1;
    double *r = nullptr;
    uint32_t *bad = nullptr;
    if (check_hip(hipMalloc(&r, cd * 8), "hipMalloc") || check_hip(hipMalloc(&bad, 4), "hipMalloc")) {
        if (r) (void)hipFree(r);
        return -1;
    }
    uint32_t h_bad = 0;
    int rc = hipMemcpy(s->d_qt, at.data(), cd * cd * 8, hipMemcpyHostToDevice);
    if (!rc) rc = hipMemset(bad, 0, 4);
    if (!rc) rc = qg_launch_qr(s->d_qt, (uint32_t)cd, r, bad, 0);
    if (!rc) rc = hipMemcpy(&h_bad, bad, 4, hipMemcpyDeviceToHost);
    (void)hipFree(r);
    (void)hipFree(bad);
    if (rc) return check_hip(rc, "QR factorisation");
    if (h_bad) return fail("the Gaussian matrix is numerically rank-deficient");
    const uint32_t kpad = (uint32_t)((s->dim + 31) / 32 * 32);
    if (s->d_qp) (void)hipFree(s->d_qp);
    s->d_qp = nullptr;
    if (check_hip(hipMalloc(&s->d_qp, (size_t)bf16_parts * cd * kpad * 2), "hipMalloc rotation parts")) return -1;
    s->kpad = kpad;
    s->parts = (uint32_t)bf16_parts;
    if (check_hip(qg_launch_qr_split(s->d_qt, (uint32_t)cd, (uint32_t)s->dim, kpad, s->parts, s->d_qp, 0),
                  "rotation split"))
        return -1;
    s->rotation = 1;
    if (check_hip(qr_rotate(s, s->d_center, 1, false, s->d_rot_center, 0), "rotated center")) return -1;
    return check_hip(hipDeviceSynchronize(), "QR rotation setup");
}

int qg_space_rotation_qr(const void *space, double *q_out) {
    const Space *s = static_cast<const Space *>(space);
    if (!s || !q_out) return fail("null space or output");
    if (s->rotation != 1) return fail("the space has no QR rotation");
    const size_t cd = s->cd;
    std::vector<double> qt(cd * cd);
    if (check_hip(hipMemcpy(qt.data(), s->d_qt, cd * cd * 8, hipMemcpyDeviceToHost), "download rotation")) return -1;
    for (size_t i = 0; i < cd; ++i)
        for (size_t j = 0; j < cd; ++j) q_out[i * cd + j] = qt[j * cd + i];
    return 0;
}

int qg_rotate_rows(const void *space, const float *rows, uint64_t n, int subtract_center, float *out,
                   void *stream) {
    const Space *s = static_cast<const Space *>(space);
    if (!s || (n && (!rows || !out))) return fail("null argument to qg_rotate_rows");
    if (s->rotation == 1) return check_hip(qr_rotate(s, rows, n, subtract_center != 0, out, (hipStream_t)stream),
                                           "rotation GEMM launch");
    if (subtract_center) return fail("qg_rotate_rows subtracts the center for QR-rotated spaces only");
    return check_hip(qg_launch_rotate_rows(rows, n, (uint32_t)s->dim, (uint32_t)s->cd, s->seed == 0, s->d_signs, out,
                                           (hipStream_t)stream),
                     "rotate launch");
}

size_t qg_space_dim(const void *space) { return space ? static_cast<const Space *>(space)->dim : 0; }
size_t qg_space_code_dim(const void *space) { return space ? static_cast<const Space *>(space)->cd : 0; }
size_t qg_gate_row_bytes(const void *space) {
    return space ? (size_t)gate_row_bytes(static_cast<const Space *>(space)->cd) : 0;
}
const float *qg_space_signs(const void *space) { return space ? static_cast<const Space *>(space)->d_signs : nullptr; }
int qg_space_host_constants(const void *space, float *signs_out, float *center_out) {
    const Space *s = static_cast<const Space *>(space);
    if (!s) return fail("null space");
    if (signs_out) std::memcpy(signs_out, s->h_signs.data(), s->cd * sizeof(float));
    if (center_out) std::memcpy(center_out, s->h_center.data(), s->dim * sizeof(float));
    return 0;
}
const float *qg_space_rotated_center(const void *space) {
    return space ? static_cast<const Space *>(space)->d_rot_center : nullptr;
}

size_t qg_encode_workspace_bytes(const void *space, uint64_t n) {
    if (!space) return 0;
    const Space *s = static_cast<const Space *>(space);
    return n * encode_row_ws_bytes(s) + 64;
}

int qg_encode(const void *space, const float *x, uint64_t n, qg_index_desc *out, void *ws, size_t ws_bytes,
              void *stream) {
    return encode(static_cast<const Space *>(space), x, n, out, ws, ws_bytes, (hipStream_t)stream);
}

int qg_pack_records(const void *space, const qg_index_desc *ix, uint8_t *compact, size_t compact_stride,
                    uint8_t *residual, size_t residual_stride, void *stream) {
    const Space *s = static_cast<const Space *>(space);
    if (!s || !ix || !compact || !residual) return fail("null argument to qg_pack_records");
    if (compact_stride < s->compact_bytes || residual_stride < s->residual_bytes)
        return fail("record stride smaller than the record size");
    PackArgs a;
    std::memset(&a, 0, sizeof(a));
    a.ix = view_of(ix);
    a.rfac = reinterpret_cast<const float2 *>(ix->rfac);
    a.cd = (uint32_t)s->cd;
    a.block_count = (uint32_t)s->block_count;
    a.compact = compact;
    a.compact_stride = compact_stride;
    a.residual = residual;
    a.residual_stride = residual_stride;
    return check_hip(qg_launch_pack(&a, ix->n, 0, (hipStream_t)stream), "pack launch");
}

int qg_unpack_records(const void *space, const uint8_t *compact, size_t compact_stride, const uint8_t *residual,
                      size_t residual_stride, uint64_t n, qg_index_desc *out, void *stream) {
    const Space *s = static_cast<const Space *>(space);
    if (!s || !compact || !out) return fail("null argument to qg_unpack_records");
    if (compact_stride < s->compact_bytes || (residual && residual_stride < s->residual_bytes))
        return fail("record stride smaller than the record size");
    PackArgs a;
    std::memset(&a, 0, sizeof(a));
    a.cd = (uint32_t)s->cd;
    a.block_count = (uint32_t)s->block_count;
    a.compact_stride = compact_stride;
    a.residual_stride = residual_stride;
    a.in_compact = compact;
    a.in_residual = residual;
    a.codes = out->codes;
    a.cid_out = out->cid;
    a.hdr = reinterpret_cast<float2 *>(out->hdr);
    a.sfac = reinterpret_cast<float2 *>(out->sfac);
    a.rscale = out->rscale;
    a.rcodes = out->rcodes;
    a.rfac_out = reinterpret_cast<float2 *>(out->rfac);
    a.gate = out->gate;
    out->n = n;
    return check_hip(qg_launch_pack(&a, n, 1, (hipStream_t)stream), "unpack launch");
}

size_t qg_prepare_workspace_bytes(uint64_t nq) { return nq * 24 + 64; }

int qg_prepare_queries(const void *space, const float *queries, uint64_t nq, int codec, qg_query_desc *out,
                       void *ws, void *stream) {
    const Space *s = static_cast<const Space *>(space);
    if (!s || (!queries && nq) || !out || (!ws && nq)) return fail("null argument to qg_prepare_queries");
    if (codec != QG_CODEC_FULL && codec != QG_CODEC_INT8 && codec != QG_CODEC_INT4 && codec != QG_CODEC_B1 &&
        codec != QG_CODEC_B1_MAIN)
        return fail("unsupported query codec %d", codec);
    if (!out->q || !out->x8 || !out->x8p || !out->qs) return fail("query output view has a null array");
    // one context per centroid (the eager centroid queries of
    // initializePreparedQuery, space_rabitq.h:3665-3698): rows k * nq + i
    const size_t K = s->centroid_count;
    out->nq = nq * K;
    if (nq == 0) return 0;
    char *w = static_cast<char *>(ws);
    uint64_t *seed = reinterpret_cast<uint64_t *>(w);
    double *nrm = reinterpret_cast<double *>(w + nq * 8);
    double *dot = reinterpret_cast<double *>(w + nq * 16);
    hipStream_t st = (hipStream_t)stream;
    const size_t cd = s->cd;
    for (size_t k = 0; k < K; ++k) {
        // the FNV-1a seed feeds only the INT8 / INT4 dither; |q|^2 and <q, c_k> in FP64
        const bool dither = codec == QG_CODEC_INT8 || codec == QG_CODEC_INT4;
        if (check_hip(qg_launch_row_seed(queries, nq, (uint32_t)s->dim, s->d_center + k * s->dim,
                                         dither ? seed : nullptr, nrm, dot, st),
                      "row seed"))
            return -1;
        PrepParams p;
        p.raw = queries;
        p.nq = nq;
        p.dim = (uint32_t)s->dim;
        p.cd = (uint32_t)cd;
        p.codec = codec;
        p.identity = s->seed == 0;
        p.signs = s->d_signs;
        p.rot_center = s->d_rot_center + k * cd;
        p.center_norm_sqr = s->h_cnorm[k];
        p.seed = seed;
        p.centroid_id = (uint32_t)k;
        p.raw_norm = nrm;
        p.raw_dot = dot;
        p.q_out = out->q + k * nq * cd;
        p.pre_rotated = s->rotation == 1 ? 1 : 0;
        if (p.pre_rotated && check_hip(qr_rotate(s, queries, nq, false, p.q_out, st), "rotation GEMM launch"))
            return -1;
        p.x8_out = out->x8 + k * nq * cd;
        p.x8p_out = out->x8p + k * nq * cd;
        p.qs_out = reinterpret_cast<QScalars *>(out->qs) + k * nq;
        if (check_hip(qg_launch_query_prep(&p, st), "query prep")) return -1;
    }
    return 0;
}

int qg_scan(const void *space, int kernel, int fixed, const qg_index_desc *ix, const qg_query_desc *qv,
            const uint32_t *ids, const uint32_t *qidx, uint64_t per_query, uint64_t n_pairs, float eps0,
            const float *short_ips, const float *long_in, float *out_f, float *out_f2, RabitqPaperEstimate *out_est,
            RabitqDistanceInterval *out_iv, int64_t *out_dot, void *stream) {
    const Space *s = static_cast<const Space *>(space);
    if (!s || !ix || !ids || (!qv && kernel != QG_SCAN_SYMMETRIC)) return fail("null argument to qg_scan");
    if (kernel < QG_SCAN_INT8_FULL || kernel > QG_SCAN_GATE_B1) return fail("unknown scan kernel %d", kernel);
    const bool record_form = kernel == QG_SCAN_RESIDUAL_RECORD;
    if (record_form) {
        if (fixed) return fail("the per-record residual form scores pairs");
        kernel = QG_SCAN_RESIDUAL;
    }
    if (!fixed && !qidx) return fail("pair scans need qidx");
    if (fixed && kernel != QG_SCAN_INT8_FULL && kernel != QG_SCAN_ASYM && kernel != QG_SCAN_RESIDUAL)
        return fail("kernel %d has no fixed-list form", kernel);
    const bool gate = kernel == QG_SCAN_GATE_FP32 || kernel == QG_SCAN_GATE_INT8 || kernel == QG_SCAN_PAPER_RECORD ||
                      kernel == QG_SCAN_GATE_B1;
    if (gate && !out_est) return fail("gate kernels need out_est");
    if (kernel == QG_SCAN_RESIDUAL && (!out_iv || !long_in)) return fail("residual kernel needs long_in and out_iv");
    if (!gate && kernel != QG_SCAN_RESIDUAL && !out_f) return fail("kernel needs out_f");
    if ((kernel == QG_SCAN_REM_FP32 || kernel == QG_SCAN_REM_INT8) && !short_ips) return fail("remaining-bit kernels need short_ips");
    ScanArgs a;
    a.ix = view_of(ix);
    if (qv) a.qv = view_of(qv);
    else std::memset(&a.qv, 0, sizeof(a.qv));
    a.cd = (uint32_t)s->cd;
    a.ids = ids;
    a.qidx = qidx;
    a.per_query = per_query;
    a.n_pairs = n_pairs;
    a.eps0 = eps0;
    a.short_ips = short_ips;
    a.long_in = long_in;
    a.out_f = out_f;
    a.out_f2 = out_f2;
    a.out_est = reinterpret_cast<PaperEstimate *>(out_est);
    a.out_iv = reinterpret_cast<Interval *>(out_iv);
    a.out_dot = out_dot;
    a.centroid_count = (uint32_t)s->centroid_count;
    a.rot_centroids = s->d_rot_center;
    a.pair_terms = s->d_pair;
    a.record_form = record_form ? 1 : 0;
    if (kernel == QG_SCAN_SYMMETRIC && s->centroid_count > 1 && !ix->cid)
        return fail("symmetric distances of a space of several centroids need the index's centroid ids");
    return check_hip(qg_launch_scan(kernel, fixed, &a, (hipStream_t)stream), "scan launch");
}

int qg_topk(const float *values, const uint32_t *ids, uint64_t nq, uint32_t per_query, uint32_t k,
            uint32_t *out_ids, float *out_values, uint32_t *out_pos, void *stream) {
    if ((!values || !ids || !out_ids) && nq) return fail("null argument to qg_topk");
    if (per_query > 4096) return fail("qg_topk supports at most 4096 entries per query");
    SelectArgs a;
    a.values = values;
    a.ids = ids;
    a.nq = nq;
    a.per_query = per_query;
    a.k = k;
    a.out_ids = out_ids;
    a.out_values = out_values;
    a.out_pos = out_pos;
    return check_hip(qg_launch_topk(&a, (hipStream_t)stream), "topk launch");
}

// one visited bit per point (and one spare), each slot's bitset a whole
// number of 16-byte words, which the search clears with 16-byte stores
static uint64_t visited_words_of(uint64_t n) { return ((n + 32) / 32 + 3) & ~uint64_t(3); }

size_t qg_search_workspace_bytes(uint64_t n, uint32_t slots) {
    return (size_t)slots * visited_words_of(n) * 4 + 256;
}

struct LearnedDepth {
    const float *mlp;
    float kmult;
    uint32_t min_depth;
    const float *rstats, *rfac;
    uint32_t *out_depth;
};

static int search_common(const void *space, const qg_index_desc *ix, const qg_query_desc *qv, int codec,
                         const uint32_t *degrees, const uint32_t *edges, size_t edge_stride, uint32_t k,
                         uint32_t l_search, const uint32_t *l_per_query, float eps0, uint32_t rerank, uint32_t slots,
                         uint32_t waves, uint32_t *out_ids, float *out_dists, uint64_t *out_stats, uint32_t *pool_ids,
                         float *pool_dists, void *workspace, size_t workspace_bytes, const LearnedDepth *learned,
                         void *stream) {
    const Space *s = static_cast<const Space *>(space);
    const bool pool_only = out_ids == nullptr && out_dists == nullptr;
    if (!s || !ix || !qv || !degrees || !edges || !workspace || (!pool_only && (!out_ids || !out_dists)))
        return fail("null argument to qg_search");
    if (pool_only && !pool_ids) return fail("qg_search without outputs needs the pool outputs");
    if (s->centroid_count != 1) return fail("the batched search holds one query context: spaces of one centroid only");
    if (codec != QG_CODEC_FULL && codec != QG_CODEC_INT8) return fail("unsupported query codec %d", codec);
    // the search stages MSB rows in 16-byte pieces: cd / 8 bytes must fill one
    if (s->cd > kMaxGraphCodeDim) return fail("the batched search takes code dimensions up to %u", kMaxGraphCodeDim);
    if (k == 0) return fail("k must be positive");
    if (edge_stride == 0 || edge_stride > 64) return fail("graph degree bound must be in [1, 64]");
    if (ix->n == 0 || ix->n > 0xFFFFFFFEull) return fail("index must hold 1 .. 2^32-2 points");
    const uint32_t L = l_search > k ? l_search : k;
    if (L > 4096) return fail("search list size %u exceeds 4096", L);
    if ((rerank > k ? rerank : k) > 1024) return fail("rerank count exceeds 1024");
    if (waves != 2 && waves != 4) return fail("waves per query must be 2 or 4");
    if ((pool_ids == nullptr) != (pool_dists == nullptr)) return fail("pool outputs come in pairs");
    if (slots == 0) slots = 1;
    if (workspace_bytes < qg_search_workspace_bytes(ix->n, slots))
        return fail("search workspace too small: need %zu bytes", qg_search_workspace_bytes(ix->n, slots));
    if (reinterpret_cast<uintptr_t>(workspace) % 16 != 0) return fail("search workspace must be 16-byte aligned");
    if (!qv->nq) return 0;
    SearchArgs a;
    a.ix = view_of(ix);
    a.qv = view_of(qv);
    a.cd = (uint32_t)s->cd;
    a.codec = codec;
    a.L = L;
    a.k = k;
    a.rerank = rerank;
    a.eps0 = eps0;
    a.degrees = degrees;
    a.edges = edges;
    a.edge_stride = edge_stride;
    a.n = ix->n;
    char *w = static_cast<char *>(workspace);
    a.counter = reinterpret_cast<uint32_t *>(w);
    a.err = reinterpret_cast<uint32_t *>(w + 4);
    a.visited = reinterpret_cast<uint32_t *>(w + 256);
    a.visited_words = visited_words_of(ix->n);
    a.out_ids = out_ids;
    a.out_d = out_dists;
    a.stats = out_stats;
    a.pool_ids = pool_ids;
    a.pool_d = pool_dists;
    a.pool_only = pool_only ? 1 : 0;
    a.lq = l_per_query;
    a.mlp = nullptr;
    a.kmult = 0.0f;
    a.min_depth = 0;
    a.rstats = nullptr;
    a.rfac = nullptr;
    a.out_depth = nullptr;
    if (learned) {
        if (pool_only) return fail("the learned rerank depth needs the rerank outputs");
        if ((rerank > k ? rerank : k) > 112) return fail("the learned rerank depth serves at most 112 candidates");
        if (!learned->mlp || !learned->rstats || !learned->rfac) return fail("null learned-depth argument");
        a.mlp = learned->mlp;
        a.kmult = learned->kmult;
        a.min_depth = learned->min_depth;
        a.rstats = reinterpret_cast<const float4 *>(learned->rstats);
        a.rfac = reinterpret_cast<const float2 *>(learned->rfac);
        a.out_depth = learned->out_depth;
    }
    return check_hip(qg_launch_search(&a, slots, waves, (hipStream_t)stream), "search launch");
}

int qg_search_ex(const void *space, const qg_index_desc *ix, const qg_query_desc *qv, int codec,
                 const uint32_t *degrees, const uint32_t *edges, size_t edge_stride, uint32_t k, uint32_t l_search,
                 const uint32_t *l_per_query, float eps0, uint32_t rerank, uint32_t slots, uint32_t waves,
                 uint32_t *out_ids, float *out_dists, uint64_t *out_stats, uint32_t *pool_ids, float *pool_dists,
                 void *workspace, size_t workspace_bytes, void *stream) {
    return search_common(space, ix, qv, codec, degrees, edges, edge_stride, k, l_search, l_per_query, eps0, rerank,
                         slots, waves, out_ids, out_dists, out_stats, pool_ids, pool_dists, workspace, workspace_bytes,
                         nullptr, stream);
}

int qg_search_learned(const void *space, const qg_index_desc *ix, const qg_query_desc *qv, int codec,
                      const uint32_t *degrees, const uint32_t *edges, size_t edge_stride, uint32_t k,
                      uint32_t l_search, const uint32_t *l_per_query, float eps0, uint32_t rerank, uint32_t slots,
                      uint32_t waves, const void *mlp_packed, float kmult, uint32_t min_depth, const float *rstats,
                      uint32_t *out_ids, float *out_dists, uint64_t *out_stats, uint32_t *out_depth,
                      void *workspace, size_t workspace_bytes, void *stream) {
    if (!ix) return fail("null argument to qg_search_learned");
    const LearnedDepth ld{static_cast<const float *>(mlp_packed), kmult, min_depth, rstats, ix->rfac, out_depth};
    return search_common(space, ix, qv, codec, degrees, edges, edge_stride, k, l_search, l_per_query, eps0, rerank,
                         slots, waves, out_ids, out_dists, out_stats, nullptr, nullptr, workspace, workspace_bytes, &ld,
                         stream);
}

int qg_search(const void *space, const qg_index_desc *ix, const qg_query_desc *qv, int codec,
              const uint32_t *degrees, const uint32_t *edges, size_t edge_stride, uint32_t k, uint32_t l_search,
              float eps0, uint32_t rerank, uint32_t slots, uint32_t *out_ids, float *out_dists,
              uint64_t *out_stats, uint32_t *pool_ids, float *pool_dists, void *workspace, size_t workspace_bytes,
              void *stream) {
    return qg_search_ex(space, ix, qv, codec, degrees, edges, edge_stride, k, l_search, nullptr, eps0, rerank, slots,
                        4, out_ids, out_dists, out_stats, pool_ids, pool_dists, workspace, workspace_bytes, stream);
}

// ------------------------------------------------- learned rerank depth
size_t qg_mlp_param_count(void) { return 5825; }

size_t qg_mlp_workspace_bytes(uint64_t n) {
    return ((size_t)qg_mlp_grad_waves(n) * (5825 + 1) + 64) * 4;
}

int qg_mlp_pack(const float *params, void *packed, void *stream) {
    if (!params || !packed) return fail("null argument to qg_mlp_pack");
    return check_hip(qg_launch_mlp_pack(params, packed, (hipStream_t)stream), "mlp pack launch");
}

int qg_mlp_forward(const float *params, const uint16_t *features, uint64_t n, float *out, void *stream) {
    if (!params || (!features && n) || (!out && n)) return fail("null argument to qg_mlp_forward");
    return check_hip(qg_launch_mlp_forward(params, features, n, out, (hipStream_t)stream), "mlp forward launch");
}

static int mlp_grad_common(const float *params, const uint16_t *features, const float *targets, uint64_t n,
                           float tau, void *ws, size_t ws_bytes, float **partial, float **ploss, uint32_t *waves,
                           hipStream_t st) {
    if (!params || !features || !targets || !ws || n == 0) return fail("null argument or empty batch to the MLP trainer");
    if (!(tau > 0.0f && tau < 1.0f)) return fail("quantile tau must be in (0, 1)");
    if (ws_bytes < qg_mlp_workspace_bytes(n)) return fail("MLP workspace too small: need %zu bytes", qg_mlp_workspace_bytes(n));
    *waves = qg_mlp_grad_waves(n);
    *ploss = static_cast<float *>(ws);
    *partial = static_cast<float *>(ws) + ((*waves + 63) / 64) * 64;
    return check_hip(qg_launch_mlp_grad(params, features, targets, n, tau, *partial, *ploss, st), "mlp grad launch");
}

int qg_mlp_grad(const float *params, const uint16_t *features, const float *targets, uint64_t n, float tau,
                float *grad_out, float *loss_out, void *workspace, size_t workspace_bytes, void *stream) {
    float *partial, *ploss;
    uint32_t waves;
    hipStream_t st = (hipStream_t)stream;
    if (!grad_out) return fail("null gradient output");
    if (mlp_grad_common(params, features, targets, n, tau, workspace, workspace_bytes, &partial, &ploss, &waves, st))
        return -1;
    return check_hip(qg_launch_mlp_reduce(partial, ploss, waves, grad_out, loss_out, nullptr, nullptr, nullptr, 0, 0,
                                          0, 0, 1, 1, st), "mlp reduce launch");
}

int qg_mlp_adam_step(float *params, float *m, float *v, const uint16_t *features, const float *targets, uint64_t n,
                     float tau, float lr, float beta1, float beta2, float eps, uint32_t step, float *loss_out,
                     void *workspace, size_t workspace_bytes, void *stream) {
    float *partial, *ploss;
    uint32_t waves;
    hipStream_t st = (hipStream_t)stream;
    if (!m || !v || step == 0) return fail("Adam needs moment buffers and a step count >= 1");
    if (mlp_grad_common(params, features, targets, n, tau, workspace, workspace_bytes, &partial, &ploss, &waves, st))
        return -1;
    const float bc1 = 1.0f - std::pow(beta1, (float)step), bc2 = 1.0f - std::pow(beta2, (float)step);
    return check_hip(qg_launch_mlp_reduce(partial, ploss, waves, nullptr, loss_out, params, m, v, lr, beta1, beta2,
                                          eps, bc1, bc2, st), "mlp adam launch");
}

int qg_residual_stats(const void *space, const qg_index_desc *ix, float *out, void *stream) {
    const Space *s = static_cast<const Space *>(space);
    if (!s || !ix || (!out && ix->n)) return fail("null argument to qg_residual_stats");
    return check_hip(qg_launch_residual_stats(ix->rscale, ix->n, (uint32_t)s->block_count, reinterpret_cast<float4 *>(out),
                                              (hipStream_t)stream), "residual stats launch");
}

static int rerank_args(const Space *s, const qg_index_desc *ix, const qg_query_desc *qv, const uint32_t *pool_ids,
                       const float *pool_dists, uint32_t L, uint32_t depth, qg::RerankArgs *a) {
    if (!s || !ix || !qv || !pool_ids || !pool_dists) return fail("null argument to the rerank stage");
    if (depth == 0 || depth > 128) return fail("rerank depth must be in [1, 128]");
    if (L == 0) return fail("pool stride must be positive");
    std::memset(a, 0, sizeof(*a));
    a->ix = view_of(ix);
    a->rfac = reinterpret_cast<const float2 *>(ix->rfac);
    a->qv = view_of(qv);
    a->cd = (uint32_t)s->cd;
    a->nq = qv->nq;
    a->L = L;
    a->depth = depth;
    a->pool_ids = pool_ids;
    a->pool_d = pool_dists;
    return 0;
}

int qg_rerank_features(const void *space, const qg_index_desc *ix, const qg_query_desc *qv, const uint32_t *pool_ids,
                       const float *pool_dists, uint32_t L, uint32_t depth, const float *rstats, uint16_t *features,
                       void *stream) {
    qg::RerankArgs a;
    if (rerank_args(static_cast<const Space *>(space), ix, qv, pool_ids, pool_dists, L, depth, &a)) return -1;
    if (a.cd > kMaxGraphCodeDim) return fail("the rerank features take code dimensions up to %u", kMaxGraphCodeDim);
    if (!rstats || !features) return fail("null residual stats or feature output");
    if (reinterpret_cast<uintptr_t>(features) & 15u) return fail("feature output must be 16-byte aligned");
    a.rstats = reinterpret_cast<const float4 *>(rstats);
    a.feats = features;
    return check_hip(qg_launch_rerank_features(&a, (hipStream_t)stream), "rerank features launch");
}

int qg_learned_rerank(const void *space, const qg_index_desc *ix, const qg_query_desc *qv, int codec,
                      const uint32_t *pool_ids, const float *pool_dists, uint32_t L, uint32_t depth,
                      const float *preds, float kmult, uint32_t min_depth, uint32_t k, uint32_t *out_ids,
                      float *out_dists, uint32_t *out_depth, void *stream) {
    qg::RerankArgs a;
    if (rerank_args(static_cast<const Space *>(space), ix, qv, pool_ids, pool_dists, L, depth, &a)) return -1;
    if (a.cd > kMaxGraphCodeDim) return fail("the workgroup rerank takes code dimensions up to %u", kMaxGraphCodeDim);
    if (codec != QG_CODEC_FULL && codec != QG_CODEC_INT8) return fail("unsupported query codec %d", codec);
    if (k == 0 || k > 128) return fail("k must be in [1, 128]");
    if (!out_ids || !out_dists) return fail("null rerank output");
    if (preds && !(kmult >= 0.0f)) return fail("error multiplier must be >= 0");
    if (preds && k > depth) return fail("k must not exceed the rerank depth when predictions are given");
    a.codec = codec;
    a.k = k;
    a.min_depth = min_depth;
    a.kmult = kmult;
    a.preds = preds;
    a.out_ids = out_ids;
    a.out_d = out_dists;
    a.out_depth = out_depth;
    return check_hip(qg_launch_learned_rerank(&a, (hipStream_t)stream), "learned rerank launch");
}

int qg_train_centroids(const float *samples, uint64_t sample_count, uint64_t dim, uint32_t k, int iters,
                       float *centroids, void *workspace, size_t workspace_bytes, void *stream) {
    if (!samples || !centroids || !workspace) return fail("null argument to qg_train_centroids");
    if (sample_count == 0 || dim == 0 || dim > 0xFFFFFFFFull) return fail("sample_count and dim must be positive");
    if (k == 0 || k > 256) return fail("k must be in [1, 256]");
    if (workspace_bytes < qg_kmeans_workspace_bytes(sample_count, dim, k))
        return fail("k-means workspace too small: %zu < %zu", workspace_bytes,
                    qg_kmeans_workspace_bytes(sample_count, dim, k));
    return check_hip(qg_launch_kmeans(samples, sample_count, (uint32_t)dim, k, iters, centroids, workspace,
                                      (hipStream_t)stream),
                     "k-means launch");
}

static_assert(sizeof(qg_signscan_query) == sizeof(SsQuery), "qg_signscan_query layout");

int qg_signscan_encode(const void *space, const float *x, uint64_t n, uint64_t *db, void *stream) {
    const Space *s = static_cast<const Space *>(space);
    if (!s || !x || !db) return fail("null argument to qg_signscan_encode");
    if (s->cd > kMaxGraphCodeDim) return fail("the 1-bit sign scan takes code dimensions up to %u", kMaxGraphCodeDim);
    if (s->rotation != 0) return fail("the 1-bit sign-plane scan uses the Walsh-Hadamard rotation, not a QR one");
    if (n % QG_SIGNSCAN_BATCH != 0) return fail("sign-scan databases hold whole blocks of 32 vectors (n = %llu)",
                                                (unsigned long long)n);
    if (n / QG_SIGNSCAN_BATCH > 0x7FFFFFFFull / QG_SIGNSCAN_BATCH) return fail("n too large for one launch");
    return check_hip(qg_launch_ss_encode(x, n, (uint32_t)s->dim, (uint32_t)s->cd, s->seed == 0, s->d_signs, db,
                                         (hipStream_t)stream),
                     "sign-scan encode launch");
}

int qg_signscan_prepare(const void *space, const float *queries, uint64_t nq, float *q, uint64_t *bits,
                        qg_signscan_query *scalars, void *stream) {
    const Space *s = static_cast<const Space *>(space);
    if (!s || !queries || !q || !bits || !scalars) return fail("null argument to qg_signscan_prepare");
    if (s->cd > kMaxGraphCodeDim) return fail("the 1-bit sign scan takes code dimensions up to %u", kMaxGraphCodeDim);
    if (s->rotation != 0) return fail("the 1-bit sign-plane scan uses the Walsh-Hadamard rotation, not a QR one");
    if (nq > 0x7FFFFFFFull) return fail("nq too large for one launch");
    return check_hip(qg_launch_ss_prepare(queries, nq, (uint32_t)s->dim, (uint32_t)s->cd, s->seed == 0, s->d_signs,
                                          q, bits, reinterpret_cast<SsQuery *>(scalars), (hipStream_t)stream),
                     "sign-scan prepare launch");
}

int qg_signscan(const void *space, int codec, const uint64_t *db, uint64_t n, const float *q, const uint64_t *bits,
                const qg_signscan_query *scalars, uint64_t nq, float *out, void *stream) {
    const Space *s = static_cast<const Space *>(space);
    if (!s || !db || !scalars || !out) return fail("null argument to qg_signscan");
    if (s->cd > kMaxGraphCodeDim) return fail("the 1-bit sign scan takes code dimensions up to %u", kMaxGraphCodeDim);
    if (codec < QG_SIGNSCAN_FP32 || codec > QG_SIGNSCAN_INT8) return fail("unknown sign-scan codec %d", codec);
    if (codec == QG_SIGNSCAN_FP32 ? !q : !bits) return fail("null query rows for sign-scan codec %d", codec);
    if (n % QG_SIGNSCAN_BATCH != 0) return fail("sign-scan databases hold whole blocks of 32 vectors (n = %llu)",
                                                (unsigned long long)n);
    if (n > 0x7FFFFFFFull * 64ull) return fail("n too large for one launch");
    return check_hip(qg_launch_ss_scan(codec, db, n, (uint32_t)(s->cd / 64), q, bits,
                                       reinterpret_cast<const SsQuery *>(scalars), nq, out, (hipStream_t)stream),
                     "sign-scan launch");
}

}  // extern "C"
