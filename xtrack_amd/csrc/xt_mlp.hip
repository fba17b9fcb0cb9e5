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
// Each 256-thread block computes a 64 x 64 tile of C over a range of k: the
// four waves stage 64 x 32 tiles of op(A) and op(B) in LDS (k contiguous
// per row, so that a lane's fragment is one 16-byte LDS read), then each
// wave multiplies its 32 x 32 quarter as 2 x 2 tiles of gfx950's
// v_mfma_f32_16x16x32_bf16.  Operand layout of that instruction (64 lanes,
// lane l): A holds row l % 16 and the eight consecutive k = 8 (l / 16) + j;
// B holds column l % 16 and the same k; the accumulator holds column l % 16
// and rows 4 (l / 16) + j, j < 4.
//
// Products with few output tiles and a long k (the weight gradient X^T dZ
// over a batch) split k over up to 256 / tiles blocks (grid z); each block
// writes its partial tile to a workspace and a second kernel sums the
// partials in split order (deterministic) and applies bias, ReLU and the
// BF16 rounding.  The workspace is at most GEMM_WS_FLOATS floats: the
// library allocates it once per device for xtm_gemm_bf16;
// xtm_gemm_bf16_noalloc uses a file-scope device array instead.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "../../include/xtrack_amd.h"

int32_t xt_report_error(int32_t code, const char *msg);   // xt_kernel.hip

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int TM = 64, TN = 64, TK = 32;
constexpr int LDK = TK + 8;     // LDS row: 40 bf16 = 80 bytes (16-byte aligned fragments)
constexpr int64_t GEMM_WS_FLOATS = 1 << 20;   // split partials: splits x M x N <= 256 x 64 x 64

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu))   // NaN: keep it quiet
        return (uint16_t)((u >> 16) | 0x0040u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

/* op(X) tile rows r0.., k k0.. into S[row][k] (bf16 bits, zero outside
   [0, R) x [0, Kend)); X(r, k) = X[r * sr + k * sk].  256 threads. */
__device__ __forceinline__ void stage(uint16_t *S, const uint16_t *__restrict__ X, int64_t sr,
                                      int64_t sk, int64_t r0, int64_t k0, int64_t R,
                                      int64_t Kend, int tid) {
    if (sk == 1) {
        // k contiguous: 8 consecutive k of one row per thread
        const int r = tid >> 2, c = (tid & 3) * 8;
        const int64_t row = r0 + r;
        for (int j = 0; j < 8; j++) {
            const int64_t k = k0 + c + j;
            S[r * LDK + c + j] = (row < R && k < Kend) ? X[row * sr + k] : (uint16_t)0;
        }
    } else if (sr == 1) {
        // rows contiguous: 8 consecutive rows of one k per thread
        const int kk = tid >> 3, r = (tid & 7) * 8;
        const int64_t k = k0 + kk;
        for (int j = 0; j < 8; j++) {
            const int64_t row = r0 + r + j;
            S[(r + j) * LDK + kk] = (row < R && k < Kend) ? X[row + k * sk] : (uint16_t)0;
        }
    } else {
        for (int e = tid; e < TM * TK; e += 256) {
            const int r = e / TK, c = e % TK;
            const int64_t row = r0 + r, k = k0 + c;
            S[r * LDK + c] = (row < R && k < Kend) ? X[row * sr + k * sk] : (uint16_t)0;
        }
    }
}

/* Block (x, y, z): C tile rows 64 y.., columns 64 x.., k range
   [z kc, min(K, (z + 1) kc)).  With ws == nullptr the tile gets the
   epilogue and goes to C/Cb; else the raw partial goes to ws[z][m][n]. */
__global__ __launch_bounds__(256) void gemm_bf16_kernel(
    const uint16_t *__restrict__ A, int64_t sam, int64_t sak, const uint16_t *__restrict__ B,
    int64_t sbk, int64_t sbn, float *__restrict__ C, int64_t ldc, uint16_t *__restrict__ Cb,
    const float *__restrict__ bias, int32_t relu, int64_t M, int64_t N, int64_t K, int64_t kc,
    float *__restrict__ ws) {
    __shared__ __align__(16) uint16_t As[TM * LDK];
    __shared__ __align__(16) uint16_t Bs[TN * LDK];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t m0 = (int64_t)blockIdx.y * TM, n0 = (int64_t)blockIdx.x * TN;
    const int wm = 32 * (w >> 1), wn = 32 * (w & 1);
    const int r16 = lane & 15, kq = (lane >> 4) * 8;
    const int64_t kbeg = (int64_t)blockIdx.z * kc;
    const int64_t kend = kbeg + kc < K ? kbeg + kc : K;
    f32x4 acc[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int64_t k0 = kbeg; k0 < kend; k0 += TK) {
        stage(As, A, sam, sak, m0, k0, M, kend, tid);
        // op(B)(k, n): rows of the staged tile are n, its k stride sbk
        stage(Bs, B, sbn, sbk, n0, k0, N, kend, tid);
        __syncthreads();
        bf16x8 a[2], b[2];
        for (int i = 0; i < 2; i++) {
            a[i] = *reinterpret_cast<const bf16x8 *>(&As[(wm + 16 * i + r16) * LDK + kq]);
            b[i] = *reinterpret_cast<const bf16x8 *>(&Bs[(wn + 16 * i + r16) * LDK + kq]);
        }
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int q = 0; q < 4; q++) {
                const int64_t m = m0 + wm + 16 * i + 4 * (lane >> 4) + q;
                const int64_t n = n0 + wn + 16 * j + r16;
                if (m >= M || n >= N) continue;
                float v = acc[i][j][q];
                if (ws) {
                    ws[((int64_t)blockIdx.z * M + m) * N + n] = v;
                    continue;
                }
                if (bias) v += bias[n];
                if (relu && !(v > 0.f)) v = 0.f;
                C[m * ldc + n] = v;
                if (Cb) Cb[m * ldc + n] = f32_to_bf16_rne(v);
            }
}

/* Sum of the `splits` partial tiles in split order, then the epilogue. */
__global__ __launch_bounds__(256) void gemm_reduce_kernel(
    const float *__restrict__ ws, int64_t splits, float *__restrict__ C, int64_t ldc,
    uint16_t *__restrict__ Cb, const float *__restrict__ bias, int32_t relu, int64_t M,
    int64_t N) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= M * N) return;
    const int64_t m = e / N, n = e % N;
    float v = 0.f;
    for (int64_t s = 0; s < splits; s++) v += ws[s * M * N + e];
    if (bias) v += bias[n];
    if (relu && !(v > 0.f)) v = 0.f;
    C[m * ldc + n] = v;
    if (Cb) Cb[m * ldc + n] = f32_to_bf16_rne(v);
}

/* k split: more blocks for products with few output tiles and a long k */
int64_t gemm_splits(int64_t M, int64_t N, int64_t K, int64_t *kc) {
    const int64_t tiles = ((M + TM - 1) / TM) * ((N + TN - 1) / TN);
    int64_t s = 1;
    if (tiles < 128 && K > 8 * TK) {
        s = (K + 8 * TK - 1) / (8 * TK);          // at least 8 k-steps per block
        const int64_t cap = 256 / tiles > 1 ? 256 / tiles : 1;
        if (s > cap) s = cap;
    }
    int64_t c = (K + s - 1) / s;
    c = (c + TK - 1) / TK * TK;
    *kc = c > 0 ? c : TK;
    return (K + *kc - 1) / *kc > 0 ? (K + *kc - 1) / *kc : 1;
}

int32_t gemm_launch(const uint16_t *A, int64_t sam, int64_t sak, const uint16_t *B,
                    int64_t sbk, int64_t sbn, float *C, int64_t ldc, uint16_t *Cb,
                    const float *bias, int32_t relu, int64_t M, int64_t N, int64_t K,
                    hipStream_t stream, float *ws, int64_t ws_floats, int64_t splits,
                    int64_t kc) {
    const int64_t gy = (M + TM - 1) / TM, gx = (N + TN - 1) / TN;
    if (splits > 1) {
        if (!ws || splits * M * N > ws_floats)
            return xt_report_error(XTM_ERR_ARG, "xtm_gemm_bf16: split partials exceed the "
                                                "workspace");
        hipLaunchKernelGGL(gemm_bf16_kernel, dim3((unsigned)gx, (unsigned)gy, (unsigned)splits),
                           dim3(256), 0, stream, A, sam, sak, B, sbk, sbn, C, ldc, Cb, bias,
                           relu, M, N, K, kc, ws);
        if (hipGetLastError() != hipSuccess)
            return xt_report_error(XTM_ERR_HIP, "xtm_gemm_bf16: kernel launch failed");
        const int64_t nb = (M * N + 255) / 256;
        hipLaunchKernelGGL(gemm_reduce_kernel, dim3((unsigned)nb), dim3(256), 0, stream, ws,
                           splits, C, ldc, Cb, bias, relu, M, N);
    } else {
        hipLaunchKernelGGL(gemm_bf16_kernel, dim3((unsigned)gx, (unsigned)gy, 1), dim3(256), 0,
                           stream, A, sam, sak, B, sbk, sbn, C, ldc, Cb, bias, relu, M, N, K,
                           K > 0 ? K : 1, (float *)nullptr);
    }
    return hipGetLastError() == hipSuccess
               ? XTM_OK
               : xt_report_error(XTM_ERR_HIP, "xtm_gemm_bf16: kernel launch failed");
}

int32_t gemm_check(const uint16_t *A, const uint16_t *B, const float *C, int64_t ldc,
                   int64_t M, int64_t N, int64_t K) {
    if (!A || !B || !C || M < 0 || N < 0 || K < 0 || ldc < N)
        return xt_report_error(XTM_ERR_ARG, "xtm_gemm_bf16: null operand, negative size or "
                                            "ldc < N");
    const int64_t gy = (M + TM - 1) / TM, gx = (N + TN - 1) / TN;
    if (gy > 65535 || gx > 0x7fffffff)
        return xt_report_error(XTM_ERR_ARG, "xtm_gemm_bf16: M above 65535 * 64 rows");
    return XTM_OK;
}

/* workspace of xtm_gemm_bf16: allocated once per device on first need */
constexpr int kMaxDevices = 64;
std::mutex g_ws_mutex;
float *g_ws[kMaxDevices] = {};

/* Scratch of xtm_gemm_bf16_noalloc: one copy per GPU, held until the
   process ends.  In host code `scratch` names a host-side stand-in, not a
   device address: its device address is taken with hipGetSymbolAddress
   once per device, with that device current, and only that pointer
   (g_scratch[device]) is handed to the kernels.  Every call on a GPU
   shares the array; split partials are written before they are read, so
   nothing needs clearing between calls, and calls that overlap on one GPU
   race. */
__device__ __align__(256) float scratch[GEMM_WS_FLOATS];
float *g_scratch[kMaxDevices] = {};

}  // namespace

extern "C" {

int32_t xtm_gemm_bf16(const uint16_t *A, int64_t sam, int64_t sak, const uint16_t *B,
                      int64_t sbk, int64_t sbn, float *C, int64_t ldc, uint16_t *Cb,
                      const float *bias, int32_t relu, int64_t M, int64_t N, int64_t K,
                      void *stream) {
    int32_t st = gemm_check(A, B, C, ldc, M, N, K);
    if (st != XTM_OK || M == 0 || N == 0) return st;
    int64_t kc;
    const int64_t splits = gemm_splits(M, N, K, &kc);
    float *ws = nullptr;
    if (splits > 1) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
            return xt_report_error(XTM_ERR_HIP, "xtm_gemm_bf16: no current device");
        std::lock_guard<std::mutex> lock(g_ws_mutex);
        if (!g_ws[dev] && hipMalloc(&g_ws[dev], GEMM_WS_FLOATS * sizeof(float)) != hipSuccess) {
            g_ws[dev] = nullptr;
            return xt_report_error(XTM_ERR_HIP, "xtm_gemm_bf16: workspace allocation failed");
        }
        ws = g_ws[dev];
    }
    return gemm_launch(A, sam, sak, B, sbk, sbn, C, ldc, Cb, bias, relu, M, N, K,
                       (hipStream_t)stream, ws, GEMM_WS_FLOATS, splits, kc);
}

int32_t xtm_gemm_bf16_noalloc(const uint16_t *A, int64_t sam, int64_t sak, const uint16_t *B,
                              int64_t sbk, int64_t sbn, float *C, int64_t ldc, uint16_t *Cb,
                              const float *bias, int32_t relu, int64_t M, int64_t N, int64_t K,
                              void *stream) {
    int32_t st = gemm_check(A, B, C, ldc, M, N, K);
    if (st != XTM_OK || M == 0 || N == 0) return st;
    int64_t kc;
    const int64_t splits = gemm_splits(M, N, K, &kc);
    float *ws = nullptr;
    if (splits > 1) {
        if (splits * M * N > GEMM_WS_FLOATS)
            return xt_report_error(XTM_ERR_ARG, "xtm_gemm_bf16_noalloc: split partials exceed "
                                                "the scratch array (2^20 floats)");
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
            return xt_report_error(XTM_ERR_HIP, "xtm_gemm_bf16_noalloc: no current device");
        std::lock_guard<std::mutex> lock(g_ws_mutex);
        if (!g_scratch[dev]) {
            void *p = nullptr;
            const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return xt_report_error(XTM_ERR_HIP, "xtm_gemm_bf16_noalloc: "
                                                    "hipGetSymbolAddress(scratch) failed");
            g_scratch[dev] = static_cast<float *>(p);
        }
        ws = g_scratch[dev];
    }
    return gemm_launch(A, sam, sak, B, sbk, sbn, C, ldc, Cb, bias, relu, M, N, K,
                       (hipStream_t)stream, ws, GEMM_WS_FLOATS, splits, kc);
}

int64_t xtm_gemm_bf16_scratch_floats(void) { return GEMM_WS_FLOATS; }

}  // extern "C"
