// qg_space.h -- host-side space object: configuration of one ExRaBitQ4
// space (RaBitQSpace ctor, space_rabitq.h:2836-2889) plus its device
// constants (rotation signs, center, rotated center).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/qgraph_rabitq.h"

namespace qg {

// Code dimensions a space may have (encode, query preparation, estimators,
// bridge), and those the graph build and search, the learned rerank, the
// QR rotation and the 1-bit sign scan take.
constexpr uint32_t kMaxCodeDim = 4096;
constexpr uint32_t kMaxGraphCodeDim = 2048;

struct Space {
    size_t dim = 0, cd = 0, block = 16, block_count = 0;
    uint32_t seed = 100;
    bool mse = true, fp16 = true;
    int codec = 0;  // query coarse codec selected through the bridge API
    size_t compact_bytes = 0, residual_bytes = 0, full_bytes = 0;
    std::vector<float> h_signs, h_center;  // h_center: [K][dim]
    size_t centroid_count = 1;             // K
    float *d_signs = nullptr;
    float *d_center = nullptr;             // [K][dim] centroids (row 0 = the global center when K = 1)
    float *d_rot_center = nullptr;         // [K][cd] rotated centroids
    float center_norm_sqr = 0.0f;          // |c_0|^2
    std::vector<float> h_cnorm;            // [K] |c_k|^2
    float2 *d_pair = nullptr;              // [K][K] {|c_a - c_b|^2, half sum of r_a - r_b}
    // rotation: 0 = sign-flipped Walsh-Hadamard (the reference's), 1 = the
    // QR rotation of the precision mode (qg_space_set_rotation_qr)
    int rotation = 0;
    double *d_qt = nullptr;      // QR: Q^T [cd][cd], FP64
    uint16_t *d_qp = nullptr;    // QR: BF16 parts of Q [parts][cd][kpad]
    uint32_t kpad = 0;           // dim rounded up to 32
    uint32_t parts = 3;          // BF16 parts per GEMM operand (2 or 3)
};

Space *space_create(size_t dim, uint32_t seed, size_t centroid_count, size_t residual_bits, size_t block,
                    bool mse, bool fp16);
void space_destroy(Space *s);
int space_set_center(Space *s, const float *center_host);
int space_set_centroids(Space *s, const float *centroids_host, size_t count);
int fail(const char *fmt, ...);
int check_hip(int rc, const char *what);

}  // namespace qg
