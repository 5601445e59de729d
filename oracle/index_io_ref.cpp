// index_io_ref.cpp -- C-ABI driver around the reference's own serialisers:
// RaBitQSpace::saveState / loadState (the "EXRBTQ31" space state,
// Ours/core/hnswlib/space_rabitq.h:3139-3345) and VamanaIndex::saveIndex /
// loadIndex (the "VMNARBQ1" index file, Ours/core/hnswlib/vamana_index.h:
// 727-829). It compiles the reference bridge where it lies (the Makefile
// passes its path as REF_BRIDGE) so spaces are built exactly as
// rabitq_space_new builds them. TEST INFRASTRUCTURE ONLY: tests/ use it to
// pin quantization_graph_amd/formats.py's readers and writers. Nothing of
// the reference is copied here; this file only calls its functions.
#include REF_BRIDGE

#include <cstdio>
#include <fstream>

extern "C" {

// saveState of a space made by rabitq_space_new(dim, seed, 1, 4, 16, mse,
// fp16) with its center set (center may be null: the zero center).
int iio_save_state(std::size_t dim, std::uint32_t seed, bool mse, bool fp16, const float *center,
                   const char *path) {
    void *space = rabitq_space_new(dim, seed, 1, 4, 16, mse, fp16);
    if (space == nullptr) return -1;
    int rc = 0;
    try {
        auto *typed = static_cast<hnswlib::RaBitQSpace *>(space);
        if (center != nullptr) typed->setGlobalCenter(center);
        std::ofstream out(path, std::ios::binary);
        typed->saveState(out);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "iio_save_state: %s\n", e.what());
        rc = -1;
    }
    rabitq_space_delete(space);
    return rc;
}

// loadState of the file at in_path into a space of the same configuration
// (made with another seed; centroid_count centroids), then saveState to
// out_path. -1 when the reference's loader rejects the file.
int iio_load_save_state_k(std::size_t dim, std::size_t centroid_count, bool mse, bool fp16, const char *in_path,
                          const char *out_path) {
    void *space = rabitq_space_new(dim, 12345u, centroid_count, 4, 16, mse, fp16);
    if (space == nullptr) return -1;
    int rc = 0;
    try {
        auto *typed = static_cast<hnswlib::RaBitQSpace *>(space);
        std::ifstream in(in_path, std::ios::binary);
        typed->loadState(in);
        std::ofstream out(out_path, std::ios::binary);
        typed->saveState(out);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "iio_load_save_state: %s\n", e.what());
        rc = -1;
    }
    rabitq_space_delete(space);
    return rc;
}

int iio_load_save_state(std::size_t dim, bool mse, bool fp16, const char *in_path, const char *out_path) {
    return iio_load_save_state_k(dim, 1, mse, fp16, in_path, out_path);
}

// buildEncodedSymmetricBulk (+ refineGraphSymmetric) over n compact records
// as rabitq_build_vamana_graph runs it, then saveIndex.
int iio_build_save_index(void *space, const void *compact, std::size_t n, std::size_t r, std::size_t l_build,
                         float alpha, std::size_t beam, std::size_t batch, std::size_t refine, const char *path) {
    try {
        auto *typed = static_cast<hnswlib::RaBitQSpace *>(space);
        hnswlib::VamanaIndex index(typed, n, r, l_build, alpha, beam);
        index.buildEncodedSymmetricBulk(compact, n, batch);
        if (refine != 0) index.refineGraphSymmetric(refine, batch);
        index.saveIndex(path);
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "iio_build_save_index: %s\n", e.what());
        return -1;
    }
}

// loadIndex of the file at in_path, then saveIndex to out_path. -1 when the
// reference's loader rejects the file.
int iio_load_save_index(void *space, const char *in_path, const char *out_path) {
    try {
        auto *typed = static_cast<hnswlib::RaBitQSpace *>(space);
        hnswlib::VamanaIndex index(typed, 0);
        index.loadIndex(in_path);
        index.saveIndex(out_path);
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "iio_load_save_index: %s\n", e.what());
        return -1;
    }
}

}  // extern "C"
