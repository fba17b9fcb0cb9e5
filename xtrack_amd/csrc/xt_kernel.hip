// xt_kernel.hip -- persistent multi-turn tracking kernel for gfx950 and the
// C ABI of libxtrack_amd.so (declarations in include/xtrack_amd.h).
//
// One wave64 lane per particle.  A lane loads its particle once, keeps the
// 6D state and bookkeeping in registers through every element of every
// turn of the launch, and writes it back once: the reference kernel
// (xtrack/tracker.py:546-749) instead re-reads and re-writes the particle
// arrays in global memory at every getter/setter call.  All lanes of a
// wave visit the same element at the same time, so the element's op words
// and constants are loaded with wave-uniform addresses (scalar loads
// through the scalar cache) and the op dispatch is a scalar branch.
//
// Per-particle arithmetic follows the reference expressions operation by
// operation; the file must be compiled with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/xtrack_amd.h"
#include "xt_builder.h"
#include "xt_flops.h"
#include "xt_math.h"
#include "xt_program.h"
#include "xt_ops.h"

namespace {

constexpr int BLOCK = 256;

/* Occupancy target (waves per SIMD).  Four waves cap the tracking kernel
   at 128 VGPRs; on MI355X this beat both the unconstrained allocation
   (~168 VGPRs, 3 waves) and 5 waves (heavy spilling) on the thick ring.
   0 leaves the allocation to the compiler. */
#ifndef XTM_WAVES_PER_EU
#define XTM_WAVES_PER_EU 4
#endif
#if XTM_WAVES_PER_EU > 0
#define XTM_OCCUPANCY __attribute__((amdgpu_waves_per_eu(XTM_WAVES_PER_EU)))
#else
#define XTM_OCCUPANCY
#endif

__global__ __launch_bounds__(BLOCK) XTM_OCCUPANCY void track_kernel(
    const int64_t *__restrict__ entry, const double *__restrict__ prog, double line_length,
    xtm_particles P, int32_t num_turns, int32_t ele_start, int32_t num_ele_track,
    int32_t flag_end_turn_actions, int32_t flag_reset_s, int32_t flag_monitor, xtm_monitor M,
    uint64_t flags) {
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= P.capacity) return;
    xt::track_particle(P, i, entry, prog, line_length, num_turns, ele_start, num_ele_track,
                       flag_end_turn_actions, flag_reset_s, flag_monitor, M, flags);
}

__global__ void math_kernel(int32_t fn, const double *__restrict__ x, double *__restrict__ y,
                            int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = x[i];
    double r;
    switch (fn) {
    case 0: r = xt_sin(v); break;
    case 1: r = xt_cos(v); break;
    case 2: r = xt_sinh(v); break;
    case 3: r = xt_cosh(v); break;
    case 4: r = xt_exp(v); break;
    case 5: r = xt_tan(v); break;
    case 6: r = XT_SQRT(v); break;
    default: r = v / x[i + n]; break;
    }
    y[i] = r;
}

thread_local std::string g_err;

int32_t fail(int32_t code, const std::string &msg) {
    g_err = msg;
    return code;
}

int32_t hip_fail(hipError_t e, const char *where) {
    return fail(XTM_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

}  // namespace

struct xtm_line {
    int32_t device;
    int64_t num_elements;
    int64_t num_records;
    int64_t *d_entry;
    double *d_prog;
    double line_length;
    std::vector<int64_t> entry;   // host copies (operation counting)
    std::vector<double> prog;
};

extern "C" {

int32_t xtm_abi_version(void) { return XTM_ABI_VERSION; }

const char *xtm_last_error(void) { return g_err.c_str(); }

int32_t xtm_device_count(void) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    return n;
}

int32_t xtm_line_create(const int32_t *types, const int64_t *offsets, const double *params,
                        int64_t num_elements, double line_length, int32_t device,
                        xtm_line **out) {
    if (!out || num_elements < 0 || (num_elements > 0 && (!types || !offsets || !params)))
        return fail(XTM_ERR_ARG, "xtm_line_create: invalid arguments");
    *out = nullptr;
    std::vector<int64_t> entry;
    std::vector<double> prog;
    std::string err;
    int st = xt_build_program(types, offsets, params, num_elements, entry, prog, err);
    if (st != XTM_OK) return fail(st, err);
    xtm_line *L = new xtm_line();
    L->device = device;
    L->num_elements = num_elements;
    L->num_records = xt_count_records(entry);
    L->line_length = line_length;
    L->d_entry = nullptr;
    L->d_prog = nullptr;
    if (device < 0) {   // host-only handle: inspection (records, op counts), no tracking
        L->entry = std::move(entry);
        L->prog = std::move(prog);
        *out = L;
        return XTM_OK;
    }
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        delete L;
        return hip_fail(e, "hipSetDevice");
    }
    const size_t ne = entry.empty() ? 1 : entry.size();
    if ((e = hipMalloc(&L->d_entry, ne * sizeof(int64_t))) != hipSuccess ||
        (e = hipMalloc(&L->d_prog, prog.size() * sizeof(double))) != hipSuccess) {
        xtm_line_destroy(L);
        return hip_fail(e, "hipMalloc");
    }
    if (!entry.empty() &&
        (e = hipMemcpy(L->d_entry, entry.data(), entry.size() * sizeof(int64_t),
                       hipMemcpyHostToDevice)) != hipSuccess) {
        xtm_line_destroy(L);
        return hip_fail(e, "hipMemcpy");
    }
    if ((e = hipMemcpy(L->d_prog, prog.data(), prog.size() * sizeof(double),
                       hipMemcpyHostToDevice)) != hipSuccess) {
        xtm_line_destroy(L);
        return hip_fail(e, "hipMemcpy");
    }
    L->entry = std::move(entry);
    L->prog = std::move(prog);
    *out = L;
    return XTM_OK;
}

int32_t xtm_line_destroy(xtm_line *L) {
    if (!L) return XTM_OK;
    if (L->device >= 0) (void)hipSetDevice(L->device);
    if (L->d_entry) (void)hipFree(L->d_entry);
    if (L->d_prog) (void)hipFree(L->d_prog);
    delete L;
    return XTM_OK;
}

int64_t xtm_line_num_elements(const xtm_line *L) { return L ? L->num_elements : -1; }

int64_t xtm_line_num_records(const xtm_line *L) { return L ? L->num_records : -1; }

int32_t xtm_track_line(xtm_line *L, const xtm_particles *P, int32_t num_turns,
                       int32_t ele_start, int32_t num_ele_track, int32_t flag_end_turn_actions,
                       int32_t flag_reset_s_at_end_turn, int32_t flag_monitor,
                       const xtm_monitor *monitor, uint64_t track_flags, void *stream) {
    if (!L || !P) return fail(XTM_ERR_ARG, "xtm_track_line: null line or particles");
    if (num_turns < 0 || ele_start < 0 || num_ele_track < 0 ||
        (int64_t)ele_start + num_ele_track > L->num_elements)
        return fail(XTM_ERR_ARG, "xtm_track_line: element range outside the line");
    if (flag_monitor < 0 || flag_monitor > 2 || (flag_monitor && !monitor))
        return fail(XTM_ERR_ARG, "xtm_track_line: invalid monitor arguments");
    if (flag_monitor && monitor->n_repetitions > 1 && monitor->repetition_period <= 0)
        return fail(XTM_ERR_ARG, "xtm_track_line: invalid monitor repetition period");
    if ((track_flags >> XTM_FLAG_BACKTRACK) & 1)
        return fail(XTM_ERR_UNSUPPORTED, "backtracking is not supported");
    if ((track_flags >> XTM_FLAG_SR_TAPER) & 1)
        return fail(XTM_ERR_UNSUPPORTED, "SR tapering is not supported");
    if (P->capacity < 0) return fail(XTM_ERR_ARG, "xtm_track_line: negative capacity");
    if (L->device < 0) return fail(XTM_ERR_ARG, "xtm_track_line: line was created without a device");
    if (P->capacity == 0 || num_turns == 0) return XTM_OK;
    hipError_t e = hipSetDevice(L->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    xtm_monitor M;
    if (flag_monitor) M = *monitor;
    else std::memset(&M, 0, sizeof(M));
    const int64_t nblocks = (P->capacity + BLOCK - 1) / BLOCK;
    if (nblocks > 0x7fffffff) return fail(XTM_ERR_ARG, "too many particles");
    hipLaunchKernelGGL(track_kernel, dim3((unsigned)nblocks), dim3(BLOCK), 0,
                       (hipStream_t)stream, L->d_entry, L->d_prog, L->line_length, *P,
                       num_turns, ele_start, num_ele_track, flag_end_turn_actions,
                       flag_reset_s_at_end_turn, flag_monitor, M, track_flags);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "track_kernel launch");
    return XTM_OK;
}

int32_t xtm_line_count_ops(const xtm_line *L, const double *coords, double p0c,
                           double mass0, double q0, double *out) {
    if (!L || !coords || !out) return fail(XTM_ERR_ARG, "xtm_line_count_ops: null argument");
    int st = xt_count_turn(L->entry, L->prog, L->line_length, coords, p0c, mass0, q0, out);
    if (st != XTM_OK) return fail(st, "xtm_line_count_ops: probe particle lost");
    return XTM_OK;
}

int32_t xtm_synchronize(void *stream) {
    hipError_t e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    return XTM_OK;
}

int32_t xtm_math_eval(int32_t fn, const double *x, double *y, int64_t n, void *stream) {
    if (fn < 0 || fn > 7 || n < 0 || (n > 0 && (!x || !y)))
        return fail(XTM_ERR_ARG, "xtm_math_eval: invalid arguments");
    if (n == 0) return XTM_OK;
    const int64_t nb = (n + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL(math_kernel, dim3((unsigned)nb), dim3(BLOCK), 0, (hipStream_t)stream,
                       fn, x, y, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "math_kernel launch");
    return XTM_OK;
}

}  // extern "C"
