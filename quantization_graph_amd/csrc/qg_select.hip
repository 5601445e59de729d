// qg_select.hip -- per-query top-k under the reference's total order
// (value, id): insert_candidate_sorted (ours_diskann.rs:1096-1126), the
// rerank sort by (estimate, id) (ours_diskann.rs:1891-1896) and the 01
// harness top-k by (value, index) (faiss_quantizer_smoke.cpp:440-449).
//
// One workgroup per query: a bitonic sort in LDS of 64-bit keys
// (order-preserving float bits << 32 | id) over the query's list padded to
// a power of two, then the first k entries are written out. Lists of up to
// 1024 entries whose positions are not asked for sort in one wavefront's
// registers instead (k_topk_wave).
#include "qg_args.h"

namespace qg {

constexpr int kSelectThreads = 256;



__global__ void __launch_bounds__(kSelectThreads) k_topk(SelectArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t n = a.per_query;
    uint32_t np = 1;
    while (np < n) np <<= 1;
    unsigned long long *key = (unsigned long long *)smem;
    uint32_t *pos = (uint32_t *)(smem + 8ull * np);
    const uint64_t qi = blockIdx.x;
    const float *v = a.values + qi * n;
    const uint32_t *id = a.ids + qi * n;
    for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) {
        if (i < n) {
            key[i] = ((unsigned long long)float_order_key(v[i]) << 32) | id[i];
            pos[i] = i;
        } else {
            key[i] = ~0ull;
            pos[i] = 0xFFFFFFFFu;
        }
    }
    __syncthreads();
    for (uint32_t size = 2; size <= np; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t t = threadIdx.x; t < np / 2; t += blockDim.x) {
                const uint32_t lo = 2 * stride * (t / stride) + (t % stride);
                const uint32_t hi = lo + stride;
                const bool up = ((lo & size) == 0);
                const unsigned long long a0 = key[lo], a1 = key[hi];
                const bool swap = up ? (a0 > a1) : (a0 < a1);
                if (swap) {
                    key[lo] = a1; key[hi] = a0;
                    const uint32_t p0 = pos[lo]; pos[lo] = pos[hi]; pos[hi] = p0;
                }
            }
            __syncthreads();
        }
    }
    const uint32_t k = a.k < n ? a.k : n;
    for (uint32_t i = threadIdx.x; i < a.k; i += blockDim.x) {
        const uint64_t o = qi * a.k + i;
        if (i < k) {
            const uint32_t p = pos[i];
            a.out_ids[o] = id[p];
            if (a.out_values) a.out_values[o] = v[p];
            if (a.out_pos) a.out_pos[o] = p;
        } else {
            a.out_ids[o] = 0xFFFFFFFFu;
            if (a.out_values) a.out_values[o] = __builtin_huge_valf();
            if (a.out_pos) a.out_pos[o] = 0xFFFFFFFFu;
        }
    }
}

// Lists of up to 64 E entries without positions: one wavefront per query
// sorts in registers, element e of lane l being list entry 64 e + l. The
// bitonic network's exchanges across lanes are xor shuffles, the others
// swaps inside a lane; no LDS and no barriers. A zero's key does not keep
// its sign, so an output whose value is zero reads it back from the list.
template <int E>
__global__ void __launch_bounds__(kSelectThreads) k_topk_wave(SelectArgs a) {
    constexpr uint32_t NP = 64u * E;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t qi = (uint64_t)blockIdx.x * (kSelectThreads / 64) + (threadIdx.x >> 6);
    if (qi >= a.nq) return;  // whole wavefronts only
    const uint32_t n = a.per_query;
    const float *v = a.values + qi * n;
    const uint32_t *id = a.ids + qi * n;
    unsigned long long key[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const uint32_t i = 64u * e + lane;
        key[e] = i < n ? ((unsigned long long)float_order_key(v[i]) << 32) | id[i] : ~0ull;
    }
#pragma unroll
    for (uint32_t size = 2; size <= NP; size <<= 1) {
#pragma unroll
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            if (stride >= 64) {
                const int s = (int)(stride / 64);
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    if (e & s) continue;
                    const bool up = ((64u * e + lane) & size) == 0;
                    const unsigned long long a0 = key[e], a1 = key[e | s];
                    const bool sw = up ? (a0 > a1) : (a0 < a1);
                    key[e] = sw ? a1 : a0;
                    key[e | s] = sw ? a0 : a1;
                }
            } else {
                const bool lower = (lane & stride) == 0;
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const unsigned long long other = __shfl_xor(key[e], (int)stride);
                    const bool up = ((64u * e + lane) & size) == 0;
                    const unsigned long long lo = key[e] < other ? key[e] : other;
                    const unsigned long long hi = key[e] < other ? other : key[e];
                    key[e] = (lower == up) ? lo : hi;
                }
            }
        }
    }
    const uint32_t k = a.k < n ? a.k : n;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const uint32_t i = 64u * e + lane;
        if (i >= a.k) continue;
        const uint64_t o = qi * a.k + i;
        if (i < k) {
            const uint32_t kid = (uint32_t)key[e];
            float val = float_from_order_key((uint32_t)(key[e] >> 32));
            if (val == 0.0f) {  // +0.0 or -0.0: the list's own value
                for (uint32_t j = 0; j < n; ++j)
                    if (id[j] == kid && v[j] == 0.0f) { val = v[j]; break; }
            }
            a.out_ids[o] = kid;
            if (a.out_values) a.out_values[o] = val;
        } else {
            a.out_ids[o] = 0xFFFFFFFFu;
            if (a.out_values) a.out_values[o] = __builtin_huge_valf();
        }
    }
    for (uint32_t i = NP + lane; i < a.k; i += 64) {  // k beyond the list's padded length
        const uint64_t o = qi * a.k + i;
        a.out_ids[o] = 0xFFFFFFFFu;
        if (a.out_values) a.out_values[o] = __builtin_huge_valf();
    }
}

}  // namespace qg

extern "C" int qg_launch_topk(const qg::SelectArgs *a, hipStream_t stream) {
    if (a->nq == 0 || a->per_query == 0) return 0;
    uint32_t np = 1;
    while (np < a->per_query) np <<= 1;
    if (!a->out_pos && np <= 1024) {
        const dim3 grid((unsigned)((a->nq + 3) / 4));
        switch (np <= 64 ? 64u : np) {
            case 64: hipLaunchKernelGGL(qg::k_topk_wave<1>, grid, dim3(qg::kSelectThreads), 0, stream, *a); break;
            case 128: hipLaunchKernelGGL(qg::k_topk_wave<2>, grid, dim3(qg::kSelectThreads), 0, stream, *a); break;
            case 256: hipLaunchKernelGGL(qg::k_topk_wave<4>, grid, dim3(qg::kSelectThreads), 0, stream, *a); break;
            case 512: hipLaunchKernelGGL(qg::k_topk_wave<8>, grid, dim3(qg::kSelectThreads), 0, stream, *a); break;
            default: hipLaunchKernelGGL(qg::k_topk_wave<16>, grid, dim3(qg::kSelectThreads), 0, stream, *a); break;
        }
        return (int)hipGetLastError();
    }
    const size_t smem = (size_t)np * 12;
    if (smem > 64 * 1024) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(qg::k_topk, dim3((unsigned)a->nq), dim3(qg::kSelectThreads), smem, stream, *a);
    return (int)hipGetLastError();
}
