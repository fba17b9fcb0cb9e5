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
// per row, so that a lane's fragment is one 16-byte LDS read; 16-byte
// global loads where the operand allows, the next k step's tiles loaded
// into registers while this step's products run), then each wave
// multiplies its 32 x 32 quarter as 2 x 2 tiles of gfx950's
// v_mfma_f32_16x16x32_bf16.  The result tile goes back through LDS so
// that each thread writes 16 consecutive columns of one row with vector
// stores (FP32, and the BF16 copy).  Block ids map to tiles so that the
// column tiles of one row tile run on one XCD (shared op(A) rows in one L2).  Operand layout of that instruction (64 lanes,
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
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

constexpr int TM = 64, TN = 64, TK = 32;
constexpr int LDK = TK + 8;     // LDS row: 40 bf16 = 80 bytes (16-byte aligned fragments)
constexpr int CLD = TN + 4;     // LDS row of the result tile: 68 floats (16-byte aligned)
constexpr int64_t GEMM_WS_FLOATS = 1 << 20;   // split partials: splits x M x N <= 256 x 64 x 64

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu))   // NaN: keep it quiet
        return (uint16_t)((u >> 16) | 0x0040u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

/* A thread's share of the op(X) tile rows r0.., k k0.. (bf16 bits, zero
   outside [0, R) x [0, Kend)); X(r, k) = X[r * sr + k * sk].  256 threads,
   8 elements each: with k contiguous (sk == 1) 8 consecutive k of one row,
   with rows contiguous (sr == 1) 8 consecutive rows of one k, else 8
   elements 256 apart in the tile.  Contiguous elements inside the operand
   that start on a 16-byte boundary are one 16-byte load. */
__device__ __forceinline__ u16x8 load_tile(const uint16_t *__restrict__ X, int64_t sr,
                                           int64_t sk, int64_t r0, int64_t k0, int64_t R,
                                           int64_t Kend, int tid) {
    const bool base16 = ((uintptr_t)X & 15) == 0;
    u16x8 v;
    if (sk == 1) {
        const int r = tid >> 2, c = (tid & 3) * 8;
        const int64_t row = r0 + r, k = k0 + c, off = row * sr + k;
        if (base16 && row < R && k + 8 <= Kend && (off & 7) == 0)
            return *reinterpret_cast<const u16x8 *>(&X[off]);
        for (int j = 0; j < 8; j++) v[j] = (row < R && k + j < Kend) ? X[off + j] : (uint16_t)0;
    } else if (sr == 1) {
        const int kk = tid >> 3, r = (tid & 7) * 8;
        const int64_t k = k0 + kk, row = r0 + r, off = row + k * sk;
        if (base16 && k < Kend && row + 8 <= R && (off & 7) == 0)
            return *reinterpret_cast<const u16x8 *>(&X[off]);
        for (int j = 0; j < 8; j++) v[j] = (row + j < R && k < Kend) ? X[off + j] : (uint16_t)0;
    } else {
        for (int j = 0; j < 8; j++) {
            const int e = tid + 256 * j, r = e / TK, c = e % TK;
            const int64_t row = r0 + r, k = k0 + c;
            v[j] = (row < R && k < Kend) ? X[row * sr + k * sk] : (uint16_t)0;
        }
    }
    return v;
}

/* The thread's share (load_tile) into the LDS tile S[row][k]. */
__device__ __forceinline__ void store_tile(uint16_t *S, const u16x8 v, int64_t sr, int64_t sk,
                                           int tid) {
    if (sk == 1) {
        const int r = tid >> 2, c = (tid & 3) * 8;
        *reinterpret_cast<u16x8 *>(&S[r * LDK + c]) = v;
    } else if (sr == 1) {
        const int kk = tid >> 3, r = (tid & 7) * 8;
        for (int j = 0; j < 8; j++) S[(r + j) * LDK + kk] = v[j];
    } else {
        for (int j = 0; j < 8; j++) {
            const int e = tid + 256 * j;
            S[(e / TK) * LDK + e % TK] = v[j];
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
    // operand tiles during the k loop, the result tile after it
    __shared__ __align__(16) unsigned char smem[TM * CLD * 4];
    uint16_t *As = reinterpret_cast<uint16_t *>(smem);
    uint16_t *Bs = As + TM * LDK;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    // Workgroups go to the eight XCDs round-robin by linear id, and each XCD
    // has its own L2: map the ids so that the column tiles of one row tile
    // (which read the same op(A) rows) run on one XCD.  Row tile t goes to
    // XCD t % 8; a bijection when the row-tile count is a multiple of 8.
    int64_t ty = blockIdx.y, tx = blockIdx.x;
    if ((gridDim.y & 7) == 0) {
        const int64_t lin = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
        const int64_t xcd = lin & 7, local = lin >> 3;
        tx = local % gridDim.x;
        ty = xcd + 8 * (local / gridDim.x);
    }
    const int64_t m0 = ty * TM, n0 = tx * TN;
    const int wm = 32 * (w >> 1), wn = 32 * (w & 1);
    const int r16 = lane & 15, kq = (lane >> 4) * 8;
    const int64_t kbeg = (int64_t)blockIdx.z * kc;
    const int64_t kend = kbeg + kc < K ? kbeg + kc : K;
    f32x4 acc[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    // op(B)(k, n): rows of the staged tile are n, its k stride sbk.  The
    // next k step's tiles are loaded into registers while this one's
    // products run.
    u16x8 va, vb;
    if (kbeg < kend) {
        va = load_tile(A, sam, sak, m0, kbeg, M, kend, tid);
        vb = load_tile(B, sbn, sbk, n0, kbeg, N, kend, tid);
    }
    for (int64_t k0 = kbeg; k0 < kend; k0 += TK) {
        store_tile(As, va, sam, sak, tid);
        store_tile(Bs, vb, sbn, sbk, tid);
        __syncthreads();
        if (k0 + TK < kend) {
            va = load_tile(A, sam, sak, m0, k0 + TK, M, kend, tid);
            vb = load_tile(B, sbn, sbk, n0, k0 + TK, N, kend, tid);
        }
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
    // the tile through LDS, so that each thread writes 16 consecutive
    // columns of one row (vector stores where the row lies inside C)
    float *Cs = reinterpret_cast<float *>(smem);
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int q = 0; q < 4; q++)
                Cs[(wm + 16 * i + 4 * (lane >> 4) + q) * CLD + wn + 16 * j + r16] = acc[i][j][q];
    __syncthreads();
    const int row = tid >> 2, c0 = (tid & 3) * 16;
    const int64_t m = m0 + row;
    if (m >= M) return;
    float *out = ws ? ws + ((int64_t)blockIdx.z * M + m) * N : C + m * ldc;
    const bool vec = (((uintptr_t)out & 15) == 0) && (ws || !Cb || ((uintptr_t)(Cb + m * ldc) & 7) == 0);
    for (int c = c0; c < c0 + 16; c += 4) {
        const int64_t n = n0 + c;
        if (n >= N) break;
        f32x4 v = *reinterpret_cast<const f32x4 *>(&Cs[row * CLD + c]);
        if (!ws) {
            for (int e = 0; e < 4; e++) {
                if (bias && n + e < N) v[e] += bias[n + e];
                if (relu && !(v[e] > 0.f)) v[e] = 0.f;
            }
        }
        if (vec && n + 4 <= N) {
            *reinterpret_cast<f32x4 *>(&out[n]) = v;
            if (!ws && Cb) {
                uint16_t b4[4] = {f32_to_bf16_rne(v[0]), f32_to_bf16_rne(v[1]),
                                  f32_to_bf16_rne(v[2]), f32_to_bf16_rne(v[3])};
                uint64_t packed;
                __builtin_memcpy(&packed, b4, 8);
                *reinterpret_cast<uint64_t *>(&Cb[m * ldc + n]) = packed;
            }
        } else {
            for (int e = 0; e < 4 && n + e < N; e++) {
                out[n + e] = v[e];
                if (!ws && Cb) Cb[m * ldc + n + e] = f32_to_bf16_rne(v[e]);
            }
        }
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
