// qg_rotation.hip -- the precision-mode rotation of a space: a random
// orthogonal Q from the FP64 QR factorisation of a Gaussian matrix, computed
// on the GPU, and the rotation GEMM rotated = Q * pad(u) on the BF16 matrix
// cores with FP32 accumulation.
//
// The reference's formal codec rotates with a sign-flipped Walsh-Hadamard
// transform (space_rabitq.h:2900-2912, 944-1000); that stays the parity
// mode. This mode is the other pattern the codec admits: a dense random
// orthogonal rotation (as the QR-orthogonalised Gaussian of
// experiments/05_disk_system_fair/quantizers.py:383-390), everything after
// the rotation unchanged.
//
// QR: classical Gram-Schmidt with one reorthogonalisation pass (CGS2) in
// FP64, column by column, on the columns of G stored as rows of A (A = G^T,
// cd x cd): a_j -= sum_k <q_k, a_j> q_k twice, then q_j = a_j / |a_j|. The
// diagonal of R is positive, so Q is the unique QR factor with that sign
// convention (numpy.linalg.qr's Q with columns flipped where diag(R) < 0).
// Orthogonality of CGS2 is at the level of FP64 rounding.
//
// GEMM: every operand is split into BF16 parts, v = hi + lo with
// hi = bf16(v), lo = bf16(v - hi) (two parts), or hi + mid + lo (three),
// and the product is the sum of the part products whose weight reaches
// 2^-16 relative, each on v_mfma_f32_32x32x16_bf16 with FP32 accumulation:
// about 16 significant bits with two parts (3 products), FP32 class with
// three (6 products), instead of BF16's 8. Tiles of 128 rows x 128 outputs x 32 inputs
// per 256-thread workgroup, 2 x 2 MFMA tiles of 32 x 32 per wavefront, A and
// B staged in LDS with a 16-byte row pad; the workgroups of one row block
// are placed on one XCD so its rows are fetched into one L2.
#include "qg_args.h"

namespace qg {
namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr uint32_t kQrThreads = 256;

__device__ __forceinline__ double block_sum(double v, double *red) {
    // fixed-order tree over the block (deterministic)
    const uint32_t t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (uint32_t s = kQrThreads / 2; s > 0; s >>= 1) {
        if (t < s) red[t] += red[t + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// r[k] = <A[k], A[j]> for k < j, one block per k
__global__ void __launch_bounds__(kQrThreads) k_qr_dots(const double *A, uint32_t cd, uint32_t j, double *r) {
    __shared__ double red[kQrThreads];
    const uint32_t k = blockIdx.x;
    const double *ak = A + (uint64_t)k * cd, *aj = A + (uint64_t)j * cd;
    double s = 0.0;
    for (uint32_t i = threadIdx.x; i < cd; i += kQrThreads) s = fma(ak[i], aj[i], s);
    s = block_sum(s, red);
    if (threadIdx.x == 0) r[k] = s;
}

// A[j] -= sum_{k<j} r[k] A[k], one thread per coordinate
__global__ void __launch_bounds__(kQrThreads) k_qr_update(double *A, uint32_t cd, uint32_t j, const double *r) {
    const uint32_t i = blockIdx.x * kQrThreads + threadIdx.x;
    if (i >= cd) return;
    double s = A[(uint64_t)j * cd + i];
    for (uint32_t k = 0; k < j; ++k) s = fma(-r[k], A[(uint64_t)k * cd + i], s);
    A[(uint64_t)j * cd + i] = s;
}

// A[j] /= |A[j]|; *bad = 1 if the column is (numerically) dependent
__global__ void __launch_bounds__(kQrThreads) k_qr_normalize(double *A, uint32_t cd, uint32_t j, uint32_t *bad) {
    __shared__ double red[kQrThreads];
    double *aj = A + (uint64_t)j * cd;
    double s = 0.0;
    for (uint32_t i = threadIdx.x; i < cd; i += kQrThreads) s = fma(aj[i], aj[i], s);
    s = block_sum(s, red);
    if (!(s > 1e-200)) {
        if (threadIdx.x == 0) *bad = 1u;
        return;
    }
    const double inv = 1.0 / sqrt(s);
    for (uint32_t i = threadIdx.x; i < cd; i += kQrThreads) aj[i] *= inv;
}

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even, finite inputs
    uint32_t b = __float_as_uint(f);
    b += 0x7FFFu + ((b >> 16) & 1u);
    return (uint16_t)(b >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Q[n][k] = QT[k][n] as S BF16 parts [S][cd][kpad] (hi, then the rounded
// remainders), zero for k >= dim
__global__ void k_qr_split(const double *QT, uint32_t cd, uint32_t dim, uint32_t kpad, uint32_t S, uint16_t *qp) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t plane = (uint64_t)cd * kpad;
    if (e >= plane) return;
    const uint32_t n = (uint32_t)(e / kpad), k = (uint32_t)(e % kpad);
    double q = k < dim ? QT[(uint64_t)k * cd + n] : 0.0;
    for (uint32_t p = 0; p < S; ++p) {
        const uint16_t h = f2bf((float)q);
        qp[p * plane + e] = h;
        q -= (double)bf2f(h);
    }
}

constexpr uint32_t kTM = 128, kTN = 128, kTK = 32, kLdsRow = kTK + 8;  // bf16 per LDS row (16-byte pad)

__device__ __forceinline__ bf16x8_t lds8(const uint16_t *p) {
    return __builtin_bit_cast(bf16x8_t, *(const uint4 *)p);
}

// S = 2: v = hi + lo, products hi*hi + hi*lo + lo*hi (about 16 significant
// bits); S = 3: v = hi + mid + lo, the six products whose weight is at
// least 2^-16 relative (hi*hi, hi*mid, mid*hi, hi*lo, lo*hi, mid*mid; about
// 24 bits, FP32 class). Small terms are accumulated first.
template <int S>
__global__ void __launch_bounds__(256) k_rot_gemm(RotGemmArgs a, uint32_t mblocks, uint32_t nblocks) {
    __shared__ __align__(16) uint16_t as[S][kTM * kLdsRow];
    __shared__ __align__(16) uint16_t bs[S][kTN * kLdsRow];
    // XCD-aware placement: dispatch i goes to XCD i % 8; the nblocks
    // workgroups of a row block share one XCD and run back to back
    const uint32_t xcd = blockIdx.x & 7u, local = blockIdx.x >> 3;
    const uint32_t mb = (local / nblocks) * 8u + xcd, nb = local % nblocks;
    if (mb >= mblocks) return;
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    const uint32_t wm = w >> 1, wn = w & 1u;
    const uint64_t m0 = (uint64_t)mb * kTM;
    const uint32_t n0 = nb * kTN;
    const uint64_t plane = (uint64_t)a.cd * a.kpad;
    f32x16_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;
    // staging roles: thread t loads row t / 2, inputs 16 * (t & 1) .. + 16
    const uint32_t sr = t >> 1, sk = (t & 1u) * 16u;
    const uint64_t arow = m0 + sr;
    const bool arow_ok = arow < a.n;
    const float scale = (arow_ok && a.row_scale) ? a.row_scale[arow] : 1.0f;
    const float *xrow = a.x + (arow_ok ? arow : 0) * a.dim;
    const bool vec4 = (a.dim & 3u) == 0u;
    for (uint32_t k0 = 0; k0 < a.kpad; k0 += kTK) {
        // A: (x - c) * scale, split into BF16 parts
        float u[16];
        const uint32_t kb = k0 + sk;
        if (arow_ok && vec4 && kb + 16u <= a.dim) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float4 v = *(const float4 *)(xrow + kb + 4 * q);
                u[4 * q] = v.x; u[4 * q + 1] = v.y; u[4 * q + 2] = v.z; u[4 * q + 3] = v.w;
            }
        } else {
#pragma unroll
            for (int q = 0; q < 16; ++q) u[q] = (arow_ok && kb + q < a.dim) ? xrow[kb + q] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const bool in = kb + q < a.dim;
            float v = u[q];
            if (a.center && in) v = v - a.center[kb + q];
            v = v * scale;
#pragma unroll
            for (int p = 0; p < S; ++p) {
                const uint16_t h = f2bf(v);
                as[p][sr * kLdsRow + sk + q] = h;
                v = v - bf2f(h);
            }
        }
        // B: Q rows n0 + sr, inputs kb .. kb + 16 (pre-split, 2 x 16 bytes per part)
        const uint64_t boff = (uint64_t)(n0 + sr) * a.kpad + kb;
#pragma unroll
        for (int p = 0; p < S; ++p) {
            const uint4 *g = (const uint4 *)(a.q + p * plane + boff);
            *(uint4 *)(&bs[p][sr * kLdsRow + sk]) = g[0];
            *(uint4 *)(&bs[p][sr * kLdsRow + sk + 8]) = g[1];
        }
        __syncthreads();
#pragma unroll
        for (uint32_t s = 0; s < kTK / 16; ++s) {
            const uint32_t kk = 16u * s + 8u * (lane >> 5);
            bf16x8_t av[S][2], bv[S][2];
#pragma unroll
            for (int p = 0; p < S; ++p)
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const uint32_t r = wm * 64u + 32u * i + (lane & 31u);
                    av[p][i] = lds8(&as[p][r * kLdsRow + kk]);
                    const uint32_t c = wn * 64u + 32u * i + (lane & 31u);
                    bv[p][i] = lds8(&bs[p][c * kLdsRow + kk]);
                }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    if (S == 3) {
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[1][i], bv[1][j], acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[2][i], bv[0][j], acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[0][i], bv[S - 1][j], acc[i][j], 0, 0, 0);
                    }
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[1][i], bv[0][j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[0][i], bv[1][j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[0][i], bv[0][j], acc[i][j], 0, 0, 0);
                }
        }
        __syncthreads();
    }
    // D[i][j] of a 32 x 32 tile: lane j + 32 * ((i / 4) % 2), item (i % 4) + 4 * (i / 8)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const uint64_t row = m0 + wm * 64u + 32u * i + 8u * (v >> 2) + 4u * (lane >> 5) + (v & 3);
                const uint32_t col = n0 + wn * 64u + 32u * j + (lane & 31u);
                if (row < a.n) a.out[row * a.cd + col] = acc[i][j][v];
            }
}

}  // namespace
}  // namespace qg

extern "C" {

// FP64 CGS2 QR in place: A (device, cd x cd, row j = column j of G) becomes
// Q^T (row j = column j of Q). *bad (device) is set when a column is
// numerically dependent. r_scratch: cd doubles.
int qg_launch_qr(double *A, uint32_t cd, double *r_scratch, uint32_t *bad, hipStream_t st) {
    const unsigned rows = (cd + qg::kQrThreads - 1) / qg::kQrThreads;
    for (uint32_t j = 0; j < cd; ++j) {
        for (int pass = 0; pass < 2 && j > 0; ++pass) {
            hipLaunchKernelGGL(qg::k_qr_dots, dim3(j), dim3(qg::kQrThreads), 0, st, A, cd, j, r_scratch);
            hipLaunchKernelGGL(qg::k_qr_update, dim3(rows), dim3(qg::kQrThreads), 0, st, A, cd, j, r_scratch);
        }
        hipLaunchKernelGGL(qg::k_qr_normalize, dim3(1), dim3(qg::kQrThreads), 0, st, A, cd, j, bad);
    }
    return (int)hipGetLastError();
}

int qg_launch_qr_split(const double *QT, uint32_t cd, uint32_t dim, uint32_t kpad, uint32_t parts, uint16_t *qp,
                       hipStream_t st) {
    const uint64_t total = (uint64_t)cd * kpad;
    hipLaunchKernelGGL(qg::k_qr_split, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, QT, cd, dim, kpad,
                       parts, qp);
    return (int)hipGetLastError();
}

// cd must be a multiple of 128 and kpad of 32
int qg_launch_rot_gemm(const qg::RotGemmArgs *a, hipStream_t st) {
    if (a->n == 0) return 0;
    const uint32_t mblocks = (uint32_t)((a->n + qg::kTM - 1) / qg::kTM), nblocks = a->cd / qg::kTN;
    const uint64_t grid = (uint64_t)((mblocks + 7) / 8) * 8 * nblocks;
    if (a->parts == 3) hipLaunchKernelGGL(qg::k_rot_gemm<3>, dim3((unsigned)grid), dim3(256), 0, st, *a, mblocks, nblocks);
    else hipLaunchKernelGGL(qg::k_rot_gemm<2>, dim3((unsigned)grid), dim3(256), 0, st, *a, mblocks, nblocks);
    return (int)hipGetLastError();
}

}  // extern "C"
