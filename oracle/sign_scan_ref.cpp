// sign_scan_ref.cpp -- C-ABI driver around the reference's 1-bit query-codec
// scan (experiments/04_query_codec_1bit_scan/query_codec_1bit_scan.cpp),
// compiled from that source where it lies (the Makefile passes its path as
// QCS_SOURCE and renames its main). TEST INFRASTRUCTURE ONLY: tests/ use it
// to pin the oracle's restatement of the scan, and bench.py times it as the
// CPU baseline of the scan. Nothing of the reference is copied here; this file
// only calls the reference's own functions, which sit in an unnamed namespace
// and are therefore reachable from this translation unit alone.
//
// The reference fixes the input dimension at 960 and the code dimension at
// 1024 (:31-32), so every entry point below works at those sizes.
#include QCS_SOURCE

#include <cstdio>

namespace {

struct RefQuery {
    QueryRepresentations rep;
};

}  // namespace

extern "C" {

size_t qcs_input_dim(void) { return kInputDim; }
size_t qcs_code_dim(void) { return kCodeDim; }

// make_fht_signs (:239-246).
void qcs_signs(uint32_t seed, float *out) {
    const std::vector<float> s = make_fht_signs(seed);
    std::memcpy(out, s.data(), s.size() * sizeof(float));
}

// build_or_load_database (:341-384) over an fvecs file of n rows; the cache
// file is neither read nor written. db: n * kWords u64 words in the
// reference's blocked layout. Returns 0, or -1 on a reference error.
int qcs_build_database(const char *fvecs_path, uint64_t n, uint32_t seed, uint64_t *db) {
    try {
        Options options;
        options.base = fvecs_path;
        options.cache = "/nonexistent/qcs_no_cache.bin";
        options.data_count = n;
        options.seed = seed;
        options.write_cache = false;
        double seconds = 0.0;
        const std::vector<float> signs = make_fht_signs(seed);
        AlignedBuffer<std::uint64_t> database = build_or_load_database(options, signs, seconds);
        std::memcpy(db, database.data(), database.size() * sizeof(std::uint64_t));
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "qcs_build_database: %s\n", e.what());
        return -1;
    }
}

// build_or_load_database with the reference's cache file at cache_path:
// loads it when its header matches (load_cache, :291-312), otherwise builds
// from the fvecs file and writes it (save_cache, :314-339).
int qcs_build_database_cached(const char *fvecs_path, uint64_t n, uint32_t seed, const char *cache_path,
                              uint64_t *db) {
    try {
        Options options;
        options.base = fvecs_path;
        options.cache = cache_path;
        options.data_count = n;
        options.seed = seed;
        options.write_cache = true;
        double seconds = 0.0;
        const std::vector<float> signs = make_fht_signs(seed);
        AlignedBuffer<std::uint64_t> database = build_or_load_database(options, signs, seconds);
        std::memcpy(db, database.data(), database.size() * sizeof(std::uint64_t));
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "qcs_build_database_cached: %s\n", e.what());
        return -1;
    }
}

// prepare_query (:458-492) of one raw query [kInputDim]. Returns an opaque
// handle; fields are read with qcs_query_fields and freed with qcs_query_free.
void *qcs_prepare(const float *raw, uint32_t seed) {
    const std::vector<float> signs = make_fht_signs(seed);
    RefQuery *q = new RefQuery;
    q->rep = prepare_query(raw, signs);
    return q;
}

// fp32 [kCodeDim], b1 [kWords], int4 planes [4*kWords], int8 planes
// [8*kWords], fscal {fp32_sum, fp32_norm2, b1_scale, b1_constant,
// int4_scale, int4_constant, int8_scale, int8_constant}, isum {int4, int8}.
void qcs_query_fields(const void *handle, float *fp32, uint64_t *b1, uint64_t *int4_planes,
                      uint64_t *int8_planes, float *fscal, int64_t *isum) {
    const QueryRepresentations &r = static_cast<const RefQuery *>(handle)->rep;
    std::memcpy(fp32, r.fp32.data(), kCodeDim * sizeof(float));
    std::memcpy(b1, r.b1.data(), kWords * sizeof(std::uint64_t));
    std::memcpy(int4_planes, r.int4_planes.data(), r.int4_planes.size() * sizeof(std::uint64_t));
    std::memcpy(int8_planes, r.int8_planes.data(), r.int8_planes.size() * sizeof(std::uint64_t));
    fscal[0] = r.fp32_sum;
    fscal[1] = r.fp32_norm2;
    fscal[2] = r.b1_scale;
    fscal[3] = r.b1_constant;
    fscal[4] = r.int4_scale;
    fscal[5] = r.int4_constant;
    fscal[6] = r.int8_scale;
    fscal[7] = r.int8_constant;
    isum[0] = r.int4_sum;
    isum[1] = r.int8_sum;
}

void qcs_query_free(void *handle) { delete static_cast<RefQuery *>(handle); }

// run_codec (:745-768): codec 0 FP32, 1 B1, 2 INT4, 3 INT8; distances for
// the first (count / 32) * 32 database vectors.
void qcs_scan(int codec, const uint64_t *db, uint64_t count, const void *handle, float *out) {
    const QueryRepresentations &r = static_cast<const RefQuery *>(handle)->rep;
    run_codec(static_cast<Codec>(codec), db, count, r, out, 0);
}

// validate_kernels (:670-733): the reference's own scalar check of its four
// SIMD scans on the first 64 vectors. Returns 0 when it passes.
int qcs_validate(const uint64_t *db, uint64_t count, const void *handle) {
    try {
        AlignedBuffer<float> distances(count);
        validate_kernels(db, count, static_cast<const RefQuery *>(handle)->rep, distances);
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "qcs_validate: %s\n", e.what());
        return -1;
    }
}

}  // extern "C"
