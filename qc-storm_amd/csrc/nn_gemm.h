// nn_gemm.h -- BF16 MFMA building blocks of the warm-start network (gfx950).
//
// C[M][N] (fp32) = A[M][K] * B[N][K]^T with A and B bf16, both K-contiguous ("NT"), K a multiple
// of 16, lda/ldb multiples of 4.  One work-group of 4 waves computes a 64 x 64 tile, each wave a
// 32 x 32 block of four v_mfma_f32_16x16x16_bf16 tiles.  MFMA operand layout (16 x 16 x 16): lane
// l supplies A[l % 16][4 (l / 16) + 0..3] and B^T[l % 16][4 (l / 16) + 0..3] (that is
// B[k][n] with n = l % 16); the result D[4 (l / 16) + i][l % 16] sits in accumulator i of lane l.
// Both operands are read with one 8-byte load per lane per tile straight from global memory:
// the matrices here are at most a few MB and stay L2-resident.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qcs {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// round-to-nearest-even float -> bf16 (NaN stays NaN)
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// accumulate: C += A B^T instead of C = A B^T
template <bool ACC>
__global__ void __launch_bounds__(256) gemm_nt_kernel(const uint16_t *__restrict__ A, const uint16_t *__restrict__ B,
                                                      float *__restrict__ C, int M, int N, int K, int lda, int ldb, int ldc)
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m0 = blockIdx.y * 64 + (w >> 1) * 32, n0 = blockIdx.x * 64 + (w & 1) * 32;
    const int lr = lane & 15, lk = (lane >> 4) * 4;
    f32x4_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
    const bf16x4_t zero = {0, 0, 0, 0};
    const bool ra0 = m0 + lr < M, ra1 = m0 + 16 + lr < M, rb0 = n0 + lr < N, rb1 = n0 + 16 + lr < N;
    const uint16_t *a0 = A + (size_t)(m0 + lr) * lda + lk, *a1 = a0 + (size_t)16 * lda;
    const uint16_t *b0 = B + (size_t)(n0 + lr) * ldb + lk, *b1 = b0 + (size_t)16 * ldb;
    for (int k0 = 0; k0 < K; k0 += 16) {
        const bf16x4_t va0 = ra0 ? *reinterpret_cast<const bf16x4_t *>(a0 + k0) : zero;
        const bf16x4_t va1 = ra1 ? *reinterpret_cast<const bf16x4_t *>(a1 + k0) : zero;
        const bf16x4_t vb0 = rb0 ? *reinterpret_cast<const bf16x4_t *>(b0 + k0) : zero;
        const bf16x4_t vb1 = rb1 ? *reinterpret_cast<const bf16x4_t *>(b1 + k0) : zero;
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(va0, vb0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(va0, vb1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(va1, vb0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(va1, vb1, acc[1][1], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int row = m0 + i * 16 + lk + t, col = n0 + j * 16 + lr;
                if (row < M && col < N) {
                    float *cp = C + (size_t)row * ldc + col;
                    *cp = ACC ? *cp + acc[i][j][t] : acc[i][j][t];
                }
            }
}

}  // namespace qcs
