// qg_kmeans.hip -- Lloyd k-means of rabitq_train_centroids
// (experiments/02_diskann_fair/native/rabitq_bridge.cpp:101-172) on the GPU.
//
// Same arithmetic as the reference loop, reorganised for the device:
//  * assignment: every (sample, centroid) distance is its own sequential
//    chain d = fma(diff, diff, d) over j = 0..dim-1 (the reference compiler
//    contracts d += diff * diff); the winner is the first centroid with the
//    smallest distance (strict <, so NaN never wins and an all-inf row goes
//    to centroid 0);
//  * update: next[c][j] is the float sum of the member rows in sample order,
//    so the members of each centroid are listed stably first; then
//    centroid = next * (1.0f / count) for non-empty clusters, empty ones keep
//    their previous value.
// Centroids are kept twice: row-major [k][dim] (the output) and transposed
// [dim][k] so that the 256 threads of an assignment block, one per centroid,
// read a coalesced row of centroid values per dimension.
#include "qg_args.h"

namespace qg {

constexpr int kKmThreads = 256;
constexpr int kKmSamples = 8;    // samples per assignment block
constexpr int kKmChunk = 256;    // dimensions staged in LDS per step

__global__ void __launch_bounds__(kKmThreads) k_km_init(const float *__restrict__ samples, uint64_t sc,
                                                        uint32_t dim, uint32_t k, float *__restrict__ cen,
                                                        float *__restrict__ cen_t) {
    const uint32_t c = blockIdx.y;
    const uint64_t src = (uint64_t)(c % sc);
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < dim; j += gridDim.x * blockDim.x) {
        const float v = samples[src * dim + j];
        cen[(uint64_t)c * dim + j] = v;
        cen_t[(uint64_t)j * k + c] = v;
    }
}

__global__ void __launch_bounds__(kKmThreads) k_km_assign(const float *__restrict__ samples, uint64_t sc,
                                                          uint32_t dim, uint32_t k, const float *__restrict__ cen_t,
                                                          uint32_t *__restrict__ assign) {
    __shared__ float vs[kKmSamples][kKmChunk];
    __shared__ float red_d[kKmSamples][kKmThreads / 64];
    __shared__ uint32_t red_c[kKmSamples][kKmThreads / 64];
    const uint32_t t = threadIdx.x;
    const uint64_t s0 = (uint64_t)blockIdx.x * kKmSamples;
    const uint32_t c = t;
    float d[kKmSamples];
#pragma unroll
    for (int s = 0; s < kKmSamples; ++s) d[s] = 0.0f;
    for (uint32_t j0 = 0; j0 < dim; j0 += kKmChunk) {
        const uint32_t len = (dim - j0 < (uint32_t)kKmChunk) ? dim - j0 : (uint32_t)kKmChunk;
        __syncthreads();
        for (uint32_t e = t; e < kKmSamples * (uint32_t)kKmChunk; e += blockDim.x) {
            const uint32_t s = e / kKmChunk, j = e % kKmChunk;
            vs[s][j] = (s0 + s < sc && j < len) ? samples[(s0 + s) * dim + j0 + j] : 0.0f;
        }
        __syncthreads();
        if (c < k) {
            for (uint32_t j = 0; j < len; ++j) {
                const float cj = cen_t[(uint64_t)(j0 + j) * k + c];
#pragma unroll
                for (int s = 0; s < kKmSamples; ++s) {
                    const float diff = vs[s][j] - cj;
                    d[s] = fmaf(diff, diff, d[s]);
                }
            }
        }
    }
    // first minimum over centroids: key (d, c), NaN and padding as +inf
    const uint32_t lane = t & 63u, wave = t >> 6;
#pragma unroll
    for (int s = 0; s < kKmSamples; ++s) {
        float bd = (c < k && !__builtin_isnan(d[s])) ? d[s] : __builtin_inff();
        uint32_t bc = (c < k) ? c : 0xFFFFFFFFu;
        for (int off = 32; off > 0; off >>= 1) {
            const float od = __shfl_xor(bd, off);
            const uint32_t oc = __shfl_xor(bc, off);
            if (od < bd || (od == bd && oc < bc)) { bd = od; bc = oc; }
        }
        if (lane == 0) { red_d[s][wave] = bd; red_c[s][wave] = bc; }
    }
    __syncthreads();
    if (t < kKmSamples && s0 + t < sc) {
        float bd = red_d[t][0];
        uint32_t bc = red_c[t][0];
        for (uint32_t w = 1; w < blockDim.x / 64; ++w) {
            const float od = red_d[t][w];
            const uint32_t oc = red_c[t][w];
            if (od < bd || (od == bd && oc < bc)) { bd = od; bc = oc; }
        }
        assign[s0 + t] = (bd < __builtin_inff()) ? bc : 0u;
    }
}

// Members of centroid blockIdx.x, counted (count_only) or listed in sample
// order at offset sum(count[0..c-1]).
__global__ void __launch_bounds__(kKmThreads) k_km_members(const uint32_t *__restrict__ assign, uint64_t sc,
                                                           uint32_t k, int count_only, uint32_t *__restrict__ count,
                                                           uint32_t *__restrict__ members) {
    __shared__ uint32_t wave_n[kKmThreads / 64];
    __shared__ uint64_t s_base;
    const uint32_t c = blockIdx.x, t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    if (t == 0) {
        uint64_t off = 0;
        if (!count_only)
            for (uint32_t i = 0; i < c; ++i) off += count[i];
        s_base = off;
    }
    __syncthreads();
    uint64_t base = s_base;
    for (uint64_t i0 = 0; i0 < sc; i0 += blockDim.x) {
        const uint64_t i = i0 + t;
        const bool mine = i < sc && assign[i] == c;
        const uint64_t ball = __ballot(mine);
        if (lane == 0) wave_n[wave] = (uint32_t)__popcll(ball);
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (uint32_t w = 0; w < blockDim.x / 64; ++w) {
            if (w < wave) before += wave_n[w];
            total += wave_n[w];
        }
        if (mine && !count_only) {
            const uint32_t rank = before + (uint32_t)__popcll(ball & ((1ull << lane) - 1ull));
            members[base + rank] = (uint32_t)i;
        }
        base += total;
        __syncthreads();
    }
    if (count_only && t == 0) count[c] = (uint32_t)(base - s_base);
}

__global__ void __launch_bounds__(kKmThreads) k_km_update(const float *__restrict__ samples, uint32_t dim,
                                                          uint32_t k, const uint32_t *__restrict__ count,
                                                          const uint32_t *__restrict__ members,
                                                          float *__restrict__ cen, float *__restrict__ cen_t) {
    const uint32_t c = blockIdx.y;
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= dim) return;
    const uint32_t n = count[c];
    if (n == 0) return;
    uint64_t off = 0;
    for (uint32_t i = 0; i < c; ++i) off += count[i];
    float acc = 0.0f;
    for (uint32_t r = 0; r < n; ++r) acc = acc + samples[(uint64_t)members[off + r] * dim + j];
    const float inv = 1.0f / (float)n;
    const float v = acc * inv;
    cen[(uint64_t)c * dim + j] = v;
    cen_t[(uint64_t)j * k + c] = v;
}

}  // namespace qg

extern "C" {

size_t qg_kmeans_workspace_bytes(uint64_t sample_count, uint64_t dim, uint32_t k) {
    return (size_t)k * dim * 4 + 2 * (size_t)sample_count * 4 + (size_t)k * 4 + 4 * 256;
}

// samples: device [sample_count][dim] in draw order; centroids: device
// [k][dim] output. Returns a hipError_t value.
int qg_launch_kmeans(const float *samples, uint64_t sc, uint32_t dim, uint32_t k, int iters, float *centroids,
                     void *workspace, hipStream_t stream) {
    char *w = static_cast<char *>(workspace);
    auto take = [&w](size_t bytes) {
        char *p = w;
        w += (bytes + 255) & ~size_t(255);
        return p;
    };
    float *cen_t = reinterpret_cast<float *>(take((size_t)k * dim * 4));
    uint32_t *assign = reinterpret_cast<uint32_t *>(take(sc * 4));
    uint32_t *members = reinterpret_cast<uint32_t *>(take(sc * 4));
    uint32_t *count = reinterpret_cast<uint32_t *>(take((size_t)k * 4));
    const dim3 block(qg::kKmThreads);
    const dim3 gdim((dim + qg::kKmThreads - 1) / qg::kKmThreads, k);
    hipLaunchKernelGGL(qg::k_km_init, gdim, block, 0, stream, samples, sc, dim, k, centroids, cen_t);
    const dim3 gs((unsigned)((sc + qg::kKmSamples - 1) / qg::kKmSamples));
    for (int it = 0; it < iters; ++it) {
        hipLaunchKernelGGL(qg::k_km_assign, gs, block, 0, stream, samples, sc, dim, k, cen_t, assign);
        hipLaunchKernelGGL(qg::k_km_members, dim3(k), block, 0, stream, assign, sc, k, 1, count, members);
        hipLaunchKernelGGL(qg::k_km_members, dim3(k), block, 0, stream, assign, sc, k, 0, count, members);
        hipLaunchKernelGGL(qg::k_km_update, gdim, block, 0, stream, samples, dim, k, count, members, centroids,
                           cen_t);
    }
    return (int)hipGetLastError();
}

}  // extern "C"
