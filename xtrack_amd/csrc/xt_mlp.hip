// xt_mlp.hip -- BF16 matrix products with FP32 accumulation on the CDNA4
// matrix cores, for the survival-turn surrogate network (training and
// inference).  No counterpart in the reference.
//
// C[M x N] = op(A)[M x K] . op(B)[K x N] (+ bias[n]) (ReLU), where op(A)(m, k)
// = A[m * sam + k * sak] and op(B)(k, n) = B[k * sbk + n * sbn]: strides
// select plain or transposed operands, which is all the forward pass
// (X W), the input gradient (dZ W^T) and the weight gradient (X^T dZ)
// need.  The result is written in FP32 and, optionally, rounded to BF16
// (round to nearest even) for the next product.
//
// Each 256-thread block computes a 64 x 64 tile; each wave a 32 x 32
// quarter as 2 x 2 tiles of v_mfma_f32_16x16x16_bf16.  Operand layout of
// that instruction (64 lanes, lane l): A holds row l % 16 and the four
// consecutive k = 4 (l / 16) + j; B holds column l % 16 and the same k;
// the accumulator holds column l % 16 and rows 4 (l / 16) + j, j < 4.
// Fragments are read directly from global memory with bounds checks
// (out-of-range elements read as zero); the network's matrices are small
// enough (at most 256 x 256 weights, batches of 10^4 - 10^5 rows) that the
// products are bound by these reads, not by the matrix cores.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/xtrack_amd.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu))   // NaN: keep it quiet
        return (uint16_t)((u >> 16) | 0x0040u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__global__ __launch_bounds__(256) void gemm_bf16_kernel(
    const uint16_t *__restrict__ A, int64_t sam, int64_t sak, const uint16_t *__restrict__ B,
    int64_t sbk, int64_t sbn, float *__restrict__ C, int64_t ldc, uint16_t *__restrict__ Cb,
    const float *__restrict__ bias, int32_t relu, int64_t M, int64_t N, int64_t K) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t tm = (int64_t)blockIdx.y * 64 + 32 * (w >> 1);
    const int64_t tn = (int64_t)blockIdx.x * 64 + 32 * (w & 1);
    const int r16 = lane & 15, kq = (lane >> 4) * 4;
    f32x4 acc[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int64_t k0 = 0; k0 < K; k0 += 16) {
        bf16x4 a[2], b[2];
        for (int i = 0; i < 2; i++) {
            const int64_t m = tm + 16 * i + r16;
            const int64_t n = tn + 16 * i + r16;
            for (int j = 0; j < 4; j++) {
                const int64_t k = k0 + kq + j;
                a[i][j] = (m < M && k < K) ? (short)A[m * sam + k * sak] : (short)0;
                b[i][j] = (n < N && k < K) ? (short)B[k * sbk + n * sbn] : (short)0;
            }
        }
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0,
                                                                       0, 0);
    }
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int q = 0; q < 4; q++) {
                const int64_t m = tm + 16 * i + 4 * (lane >> 4) + q;
                const int64_t n = tn + 16 * j + r16;
                if (m >= M || n >= N) continue;
                float v = acc[i][j][q];
                if (bias) v += bias[n];
                if (relu && !(v > 0.f)) v = 0.f;
                C[m * ldc + n] = v;
                if (Cb) Cb[m * ldc + n] = f32_to_bf16_rne(v);
            }
}

}  // namespace

extern "C" {

int32_t xtm_gemm_bf16(const uint16_t *A, int64_t sam, int64_t sak, const uint16_t *B,
                      int64_t sbk, int64_t sbn, float *C, int64_t ldc, uint16_t *Cb,
                      const float *bias, int32_t relu, int64_t M, int64_t N, int64_t K,
                      void *stream) {
    if (!A || !B || !C || M < 0 || N < 0 || K < 0 || ldc < N) return XTM_ERR_ARG;
    if (M == 0 || N == 0) return XTM_OK;
    const int64_t gy = (M + 63) / 64, gx = (N + 63) / 64;
    if (gy > 65535 || gx > 0x7fffffff) return XTM_ERR_ARG;
    hipLaunchKernelGGL(gemm_bf16_kernel, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0,
                       (hipStream_t)stream, A, sam, sak, B, sbk, sbn, C, ldc, Cb, bias, relu, M,
                       N, K);
    return hipGetLastError() == hipSuccess ? XTM_OK : XTM_ERR_HIP;
}

}  // extern "C"
