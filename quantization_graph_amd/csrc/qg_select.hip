// qg_select.hip -- per-query top-k under the reference's total order
// (value, id): insert_candidate_sorted (ours_diskann.rs:1096-1126), the
// rerank sort by (estimate, id) (ours_diskann.rs:1891-1896) and the 01
// harness top-k by (value, index) (faiss_quantizer_smoke.cpp:440-449).
//
// One workgroup per query: a bitonic sort in LDS of 64-bit keys
// (order-preserving float bits << 32 | id) over the query's list padded to
// a power of two, then the first k entries are written out.
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

}  // namespace qg

extern "C" int qg_launch_topk(const qg::SelectArgs *a, hipStream_t stream) {
    if (a->nq == 0 || a->per_query == 0) return 0;
    uint32_t np = 1;
    while (np < a->per_query) np <<= 1;
    const size_t smem = (size_t)np * 12;
    if (smem > 64 * 1024) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(qg::k_topk, dim3((unsigned)a->nq), dim3(qg::kSelectThreads), smem, stream, *a);
    return (int)hipGetLastError();
}
