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
#include <hip/hiprtc.h>

#include <cstdint>
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/xtrack_amd.h"
#include "xt_builder.h"
#include "xt_codegen.h"
#include "xt_flops.h"
#include "xt_math.h"
#include "xt_program.h"
#include "xt_ops.h"

namespace {

constexpr int BLOCK = 256;
/* dynamic LDS of a tracking launch of n threads per block: the per-thread
   slots of the particle state (xt_ops.h, track_particle_loop) */
constexpr unsigned lds_bytes(int n) { return (unsigned)n * xt::XT_LDS_BYTES_PER_THREAD; }
constexpr unsigned kLdsLine = lds_bytes(BLOCK);

/* Occupancy target (waves per SIMD).  Four waves cap the tracking kernel
   at 128 VGPRs; on MI355X this beat both the unconstrained allocation
   (~168 VGPRs, 3 waves) and 5 waves (heavy spilling) on the thick ring.
   0 leaves the allocation to the compiler. */
#ifndef XTM_WAVES_PER_EU
#define XTM_WAVES_PER_EU 4
#endif
/* Occupancy target of compiled line code (xtm_line_specialize with
   waves_per_eu = 0).  Its per-record straight-line code needs fewer
   registers per op than the interpreter's dispatch loop; on MI355X five
   waves per SIMD measured fastest on the thick ring (3, 4, 6 slower). */
#ifndef XTM_COMPILED_WAVES_PER_EU
#define XTM_COMPILED_WAVES_PER_EU 5
#endif
#if XTM_WAVES_PER_EU > 0
#define XTM_OCCUPANCY __attribute__((amdgpu_waves_per_eu(XTM_WAVES_PER_EU)))
#else
#define XTM_OCCUPANCY
#endif

/* The particles struct is read through Pd, a device copy, so that passing
   it by reference to the out-of-line loops copies nothing into the waves'
   scratch frames. */
__global__ __launch_bounds__(BLOCK) XTM_OCCUPANCY void track_kernel(
    const int64_t *__restrict__ entry, const double *__restrict__ prog, double line_length,
    int64_t num_ele_line, const xtm_particles *__restrict__ Pd, int32_t num_turns,
    int32_t ele_start, int32_t num_ele_track, int32_t flag_end_turn_actions,
    int32_t flag_reset_s, int32_t flag_monitor, const xtm_monitor *__restrict__ Md,
    uint64_t flags, int32_t skip_chi1) {
    const xtm_particles &P = *Pd;
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= P.capacity) return;
    // after the compiled line kernel: only the particles it left (chi != 1)
    if (skip_chi1 && P.chi[i] == 1.0) return;
    xt::track_particle(P, i, entry, prog, line_length, num_ele_line, num_turns, ele_start,
                       num_ele_track, flag_end_turn_actions, flag_reset_s, flag_monitor,
                       xt::MonRef{Md, Pd}, flags);
}

/* One block of a batch: particles [first, first + count) of line `line`. */
struct BatchBlock {
    int64_t first;
    int32_t line;
    int32_t count;
};

/* Batched lattice variants: block b tracks its particles through its line
   (element programs entries[line], progs[line]) for num_turns full turns.
   The block's line is uniform across the block, so element dispatch stays
   a scalar branch as in track_kernel. */
template <int WPE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WPE))) void batch_kernel(
    const BatchBlock *__restrict__ blocks, const int64_t *const *__restrict__ entries,
    const double *const *__restrict__ progs, const int64_t *__restrict__ nele,
    const double *__restrict__ lengths, const xtm_particles *__restrict__ Pd,
    int32_t num_turns, uint64_t flags, int32_t skip_chi1) {
    const BatchBlock b = xt::cptr(blocks)[blockIdx.x];
    if ((int32_t)threadIdx.x >= b.count) return;
    const xtm_particles &P = *Pd;
    const int64_t i = b.first + threadIdx.x;
    // after the compiled batch kernel: only the particles it left (chi != 1)
    if (skip_chi1 && P.chi[i] == 1.0) return;
    const int64_t ne = xt::cptr(nele)[b.line];
    const xt::TrackCtx c{flags, xt::cptr(lengths)[b.line], P.q0, P.t_sim, ne};
    const xt::ProgramRunner run{xt::cptr(entries)[b.line], xt::cptr(progs)[b.line]};
    xt::track_particle_with<false>(P, i, run, run, c, num_turns, 0, (int32_t)ne, 1, 1, 0,
                                   xt::MonRef{nullptr, Pd});
}

/* device copies of the particle struct and the monitor for the tracking
   kernels (stream-ordered before them; kernel arguments are captured at
   launch, so the caller's structs need not outlive the call) */
__global__ void args_kernel(xtm_particles P, xtm_monitor M, xtm_particles *Pd,
                            xtm_monitor *Md) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *Pd = P;
        *Md = M;
    }
}

/* Feature accumulators of every active particle at the start of its
   current turn (the state the last tracking launch left), one thread per
   particle slot: runs between one-turn tracking launches, so the tracking
   kernels carry no trace of it. */
__global__ __launch_bounds__(256) void features_kernel(xtm_features F, xtm_particles P) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= P.capacity || !(P.state[i] > 0)) return;
    xt::feature_record(F, i, P.x[i], P.px[i], P.y[i], P.py[i], P.at_turn[i]);
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
    case 7: r = v / x[i + n]; break;
    case 8: r = xt::rcp_ir(v); break;
    case 9: r = xt::sqrt_ir(v); break;
    case 10: r = xt::div_ir(v, x[i + n]); break;
    case 11: r = xt_atan2(v, x[i + n]); break;
    default: r = xt_asin(v); break;
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

/* the message xtm_last_error() returns, for the library's other source
   files (matrix products) */
int32_t xt_report_error(int32_t code, const char *msg) { return fail(code, msg); }

struct xtm_line {
    int32_t device;
    int64_t num_elements;
    int64_t num_records;
    int64_t *d_entry;
    int64_t *d_entry_bt;   // backtracking programs (XS_FLAG_BACKTRACK)
    double *d_prog;
    double line_length;
    std::vector<int64_t> entry;   // host copies (operation counting, code generation)
    std::vector<int64_t> entry_bt;
    std::vector<double> prog;
    // device copies of the particle and monitor structs read by monitor
    // records (written by monitor_args_kernel before a launch with a monitor)
    xtm_particles *d_mon_p = nullptr;
    xtm_monitor *d_mon_m = nullptr;
    // generated line code (xtm_line_specialize)
    int32_t *d_rec = nullptr;
    hipModule_t module = nullptr;
    hipFunction_t kernel = nullptr;
    uint32_t spec_options = 0;
};

struct xtm_batch {
    int32_t device;
    std::vector<xtm_line *> lines;
    std::vector<int64_t> nele;
    std::vector<int64_t> line_start;      // current layout (num_lines + 1)
    std::vector<BatchBlock> blocks;       // current block table
    int32_t block_size = 0, waves_per_eu = 0;
    // device tables
    const int64_t **d_entries = nullptr;
    const double **d_progs = nullptr;
    int64_t *d_nele = nullptr;
    double *d_lengths = nullptr;
    BatchBlock *d_blocks = nullptr;
    int64_t d_blocks_cap = 0;
    xtm_particles *d_p = nullptr;
    xtm_monitor *d_m = nullptr;
    // generated batch code (xtm_batch_specialize)
    std::vector<int32_t *> d_recs;
    const int32_t **d_rec_ptrs = nullptr;
    hipModule_t module = nullptr;
    hipFunction_t kernel = nullptr;
};

#ifndef XTM_RTC_DEFINES
#define XTM_RTC_DEFINES ""
#endif

extern const int xt_embedded_header_count;
extern const char *const xt_embedded_header_names[];
extern const char *const xt_embedded_header_texts[];

namespace {

// code objects of generated line code, by source text (one process may
// build many lines with identical element records: tests, optics probes)
std::mutex g_rtc_mutex;
std::map<std::string, std::vector<char>> g_rtc_cache;

int32_t rtc_compile(const std::string &src, std::vector<char> &code) {
    {
        std::lock_guard<std::mutex> lk(g_rtc_mutex);
        auto it = g_rtc_cache.find(src);
        if (it != g_rtc_cache.end()) {
            code = it->second;
            return XTM_OK;
        }
    }
    hiprtcProgram prog;
    hiprtcResult r = hiprtcCreateProgram(&prog, src.c_str(), "xtm_line.hip",
                                         xt_embedded_header_count, xt_embedded_header_texts,
                                         xt_embedded_header_names);
    if (r != HIPRTC_SUCCESS)
        return fail(XTM_ERR_HIP, std::string("hiprtcCreateProgram: ") + hiprtcGetErrorString(r));
    std::vector<std::string> opts = {"--offload-arch=gfx950", "-O3", "-std=c++17",
                                     "-ffp-contract=off", "-fno-fast-math", "-Xclang",
                                     "-target-feature", "-Xclang", "-packed-fp32-ops"};
    {
        const std::string defs = XTM_RTC_DEFINES;
        size_t a = 0;
        while (a < defs.size()) {
            size_t b = defs.find(' ', a);
            if (b == std::string::npos) b = defs.size();
            if (b > a) opts.push_back(defs.substr(a, b - a));
            a = b + 1;
        }
    }
    std::vector<const char *> copts;
    for (auto &o : opts) copts.push_back(o.c_str());
    r = hiprtcCompileProgram(prog, (int)copts.size(), copts.data());
    if (r != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(prog, &n);
        std::string log(n, '\0');
        if (n) hiprtcGetProgramLog(prog, &log[0]);
        hiprtcDestroyProgram(&prog);
        return fail(XTM_ERR_HIP, "line code compilation failed: " + log.substr(0, 4000));
    }
    size_t n = 0;
    hiprtcGetCodeSize(prog, &n);
    code.resize(n);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    std::lock_guard<std::mutex> lk(g_rtc_mutex);
    g_rtc_cache[src] = code;
    return XTM_OK;
}

}  // namespace

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
    std::vector<int64_t> entry, entry_bt;
    std::vector<double> prog;
    std::string err;
    int st = xt_build_program(types, offsets, params, num_elements, entry, prog, err, &entry_bt);
    if (st != XTM_OK) return fail(st, err);
    xtm_line *L = new xtm_line();
    L->device = device;
    L->num_elements = num_elements;
    L->num_records = xt_count_records(entry);
    L->line_length = line_length;
    L->d_entry = nullptr;
    L->d_entry_bt = nullptr;
    L->d_prog = nullptr;
    if (device < 0) {   // host-only handle: inspection (records, op counts), no tracking
        L->entry = std::move(entry);
        L->entry_bt = std::move(entry_bt);
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
        (e = hipMalloc(&L->d_entry_bt, ne * sizeof(int64_t))) != hipSuccess ||
        (e = hipMalloc(&L->d_prog, prog.size() * sizeof(double))) != hipSuccess ||
        (e = hipMalloc(&L->d_mon_p, sizeof(xtm_particles))) != hipSuccess ||
        (e = hipMalloc(&L->d_mon_m, sizeof(xtm_monitor))) != hipSuccess) {
        xtm_line_destroy(L);
        return hip_fail(e, "hipMalloc");
    }
    if (!entry.empty() &&
        ((e = hipMemcpy(L->d_entry, entry.data(), entry.size() * sizeof(int64_t),
                        hipMemcpyHostToDevice)) != hipSuccess ||
         (e = hipMemcpy(L->d_entry_bt, entry_bt.data(), entry_bt.size() * sizeof(int64_t),
                        hipMemcpyHostToDevice)) != hipSuccess)) {
        xtm_line_destroy(L);
        return hip_fail(e, "hipMemcpy");
    }
    if ((e = hipMemcpy(L->d_prog, prog.data(), prog.size() * sizeof(double),
                       hipMemcpyHostToDevice)) != hipSuccess) {
        xtm_line_destroy(L);
        return hip_fail(e, "hipMemcpy");
    }
    L->entry = std::move(entry);
    L->entry_bt = std::move(entry_bt);
    L->prog = std::move(prog);
    *out = L;
    return XTM_OK;
}

int32_t xtm_line_destroy(xtm_line *L) {
    if (!L) return XTM_OK;
    if (L->device >= 0) (void)hipSetDevice(L->device);
    if (L->d_entry) (void)hipFree(L->d_entry);
    if (L->d_entry_bt) (void)hipFree(L->d_entry_bt);
    if (L->d_prog) (void)hipFree(L->d_prog);
    if (L->d_rec) (void)hipFree(L->d_rec);
    if (L->d_mon_p) (void)hipFree(L->d_mon_p);
    if (L->d_mon_m) (void)hipFree(L->d_mon_m);
    if (L->module) (void)hipModuleUnload(L->module);
    delete L;
    return XTM_OK;
}

int32_t xtm_line_specialize(xtm_line *L, int32_t waves_per_eu) {
    return xtm_line_specialize_ex(L, waves_per_eu, 0);
}

int32_t xtm_line_specialize_ex(xtm_line *L, int32_t waves_per_eu, uint32_t options) {
    if (!L) return fail(XTM_ERR_ARG, "xtm_line_specialize: null line");
    if (options & ~XTM_SPEC_FAST_ARITH)
        return fail(XTM_ERR_ARG, "xtm_line_specialize: unknown option bits");
    if (L->device < 0)
        return fail(XTM_ERR_ARG, "xtm_line_specialize: line was created without a device");
    if (waves_per_eu < 0 || waves_per_eu > 8)
        return fail(XTM_ERR_ARG, "xtm_line_specialize: waves_per_eu outside 0..8");
    if (waves_per_eu == 0) waves_per_eu = XTM_COMPILED_WAVES_PER_EU;
    std::string src, err;
    std::vector<int32_t> rec;
    int st = xt_generate_line_source(L->entry, L->prog, waves_per_eu, options, src, rec, err);
    if (st != XTM_OK) return fail(st, "xtm_line_specialize: " + err);
    std::vector<char> code;
    st = rtc_compile(src, code);
    if (st != XTM_OK) return st;
    hipError_t e = hipSetDevice(L->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipModule_t mod = nullptr;
    if ((e = hipModuleLoadData(&mod, code.data())) != hipSuccess)
        return hip_fail(e, "hipModuleLoadData");
    hipFunction_t fn = nullptr;
    if ((e = hipModuleGetFunction(&fn, mod, "xtm_line_kernel")) != hipSuccess) {
        (void)hipModuleUnload(mod);
        return hip_fail(e, "hipModuleGetFunction");
    }
    int32_t *d_rec = nullptr;
    const size_t nr = rec.empty() ? 1 : rec.size();
    if ((e = hipMalloc(&d_rec, nr * sizeof(int32_t))) != hipSuccess) {
        (void)hipModuleUnload(mod);
        return hip_fail(e, "hipMalloc");
    }
    if (!rec.empty() &&
        (e = hipMemcpy(d_rec, rec.data(), rec.size() * sizeof(int32_t),
                       hipMemcpyHostToDevice)) != hipSuccess) {
        (void)hipFree(d_rec);
        (void)hipModuleUnload(mod);
        return hip_fail(e, "hipMemcpy");
    }
    if (L->d_rec) (void)hipFree(L->d_rec);
    if (L->module) (void)hipModuleUnload(L->module);
    L->d_rec = d_rec;
    L->module = mod;
    L->kernel = fn;
    L->spec_options = options;
    return XTM_OK;
}

int32_t xtm_line_is_specialized(const xtm_line *L) { return (L && L->kernel) ? 1 : 0; }

int32_t xtm_line_specialized_options(const xtm_line *L) {
    return (L && L->kernel) ? (int32_t)L->spec_options : -1;
}

int32_t xtm_line_generate_source(const xtm_line *L, int32_t waves_per_eu, char *buf,
                                 int64_t cap, int64_t *len, int32_t *rec_of_elem) {
    return xtm_line_generate_source_ex(L, waves_per_eu, 0, buf, cap, len, rec_of_elem);
}

int32_t xtm_line_generate_source_ex(const xtm_line *L, int32_t waves_per_eu, uint32_t options,
                                    char *buf, int64_t cap, int64_t *len,
                                    int32_t *rec_of_elem) {
    if (!L || !len) return fail(XTM_ERR_ARG, "xtm_line_generate_source: null argument");
    if (options & ~XTM_SPEC_FAST_ARITH)
        return fail(XTM_ERR_ARG, "xtm_line_generate_source: unknown option bits");
    std::string src, err;
    std::vector<int32_t> rec;
    int st = xt_generate_line_source(L->entry, L->prog, waves_per_eu, options, src, rec, err);
    if (st != XTM_OK) return fail(st, "xtm_line_generate_source: " + err);
    *len = (int64_t)src.size();
    if (buf && cap > 0) {
        const int64_t n = (int64_t)src.size() < cap - 1 ? (int64_t)src.size() : cap - 1;
        std::memcpy(buf, src.data(), (size_t)n);
        buf[n] = '\0';
    }
    if (rec_of_elem) std::memcpy(rec_of_elem, rec.data(), rec.size() * sizeof(int32_t));
    return XTM_OK;
}

int64_t xtm_line_program_size(const xtm_line *L) { return L ? (int64_t)L->prog.size() : -1; }

int32_t xtm_line_copy_entry(const xtm_line *L, int64_t *out) {
    if (!L || !out) return fail(XTM_ERR_ARG, "xtm_line_copy_entry: null argument");
    std::memcpy(out, L->entry.data(), L->entry.size() * sizeof(int64_t));
    return XTM_OK;
}

int32_t xtm_line_copy_program(const xtm_line *L, double *out) {
    if (!L || !out) return fail(XTM_ERR_ARG, "xtm_line_copy_program: null argument");
    std::memcpy(out, L->prog.data(), L->prog.size() * sizeof(double));
    return XTM_OK;
}

int64_t xtm_line_num_elements(const xtm_line *L) { return L ? L->num_elements : -1; }

int64_t xtm_line_num_records(const xtm_line *L) { return L ? L->num_records : -1; }

static bool features_ok(const xtm_features *F, int64_t capacity) {
    return !F || (F->acc && F->capacity == capacity && F->start_at_turn <= F->stop_at_turn);
}

static int32_t track_line_impl(xtm_line *L, const xtm_particles *P, int32_t num_turns,
                               int32_t ele_start, int32_t num_ele_track,
                               int32_t flag_end_turn_actions,
                               int32_t flag_reset_s_at_end_turn, int32_t flag_monitor,
                               const xtm_monitor *monitor, uint64_t track_flags,
                               void *stream) {
    if (!L || !P) return fail(XTM_ERR_ARG, "xtm_track_line: null line or particles");
    if (num_turns < 0 || ele_start < 0 || num_ele_track < 0 ||
        (int64_t)ele_start + num_ele_track > L->num_elements)
        return fail(XTM_ERR_ARG, "xtm_track_line: element range outside the line");
    if (flag_monitor < 0 || flag_monitor > 2 || (flag_monitor && !monitor))
        return fail(XTM_ERR_ARG, "xtm_track_line: invalid monitor arguments");
    if (flag_monitor && monitor->n_repetitions > 1 && monitor->repetition_period <= 0)
        return fail(XTM_ERR_ARG, "xtm_track_line: invalid monitor repetition period");
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
    hipLaunchKernelGGL(args_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, *P, M,
                       L->d_mon_p, L->d_mon_m);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "args_kernel launch");
    const xtm_monitor *Md = L->d_mon_m;
    const xtm_particles *Pd = L->d_mon_p;
    const int64_t nblocks = (P->capacity + BLOCK - 1) / BLOCK;
    if (nblocks > 0x7fffffff) return fail(XTM_ERR_ARG, "too many particles");
    const bool backtrack = (track_flags >> XTM_FLAG_BACKTRACK) & 1;
    if (L->kernel && !backtrack) {
        xtm_particles Pv = *P;
        const int32_t *rec = L->d_rec;
        const int64_t *entry = L->d_entry;
        const double *prog = L->d_prog;
        double line_length = L->line_length;
        int32_t a_turns = num_turns, a_start = ele_start, a_n = num_ele_track;
        int32_t a_end = flag_end_turn_actions, a_reset = flag_reset_s_at_end_turn;
        int32_t a_mon = flag_monitor;
        uint64_t a_flags = track_flags;
        void *args[] = {&rec, &entry, &prog, &line_length, &Pv, &a_turns, &a_start, &a_n,
                        &a_end, &a_reset, &a_mon, &Md, &Pd, &a_flags};
        e = hipModuleLaunchKernel(L->kernel, (unsigned)nblocks, 1, 1, BLOCK, 1, 1, kLdsLine,
                                  (hipStream_t)stream, args, nullptr);
        if (e != hipSuccess) return hip_fail(e, "xtm_line_kernel launch");
        /* particles with chi != 1 on the interpreter, same stream */
        hipLaunchKernelGGL(track_kernel, dim3((unsigned)nblocks), dim3(BLOCK), kLdsLine,
                           (hipStream_t)stream, L->d_entry, L->d_prog, L->line_length,
                           (int64_t)L->num_elements, Pd, num_turns, ele_start, num_ele_track,
                           flag_end_turn_actions, flag_reset_s_at_end_turn, flag_monitor, Md,
                           track_flags, (int32_t)1);
        e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "track_kernel launch");
        return XTM_OK;
    }
    /* backtracking always runs on the interpreter, with the backtracking
       programs */
    hipLaunchKernelGGL(track_kernel, dim3((unsigned)nblocks), dim3(BLOCK), kLdsLine,
                       (hipStream_t)stream, backtrack ? L->d_entry_bt : L->d_entry, L->d_prog,
                       L->line_length, (int64_t)L->num_elements, Pd,
                       num_turns, ele_start, num_ele_track, flag_end_turn_actions,
                       flag_reset_s_at_end_turn, flag_monitor, Md, track_flags, (int32_t)0);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "track_kernel launch");
    return XTM_OK;
}

int32_t xtm_track_line(xtm_line *L, const xtm_particles *P, int32_t num_turns,
                       int32_t ele_start, int32_t num_ele_track, int32_t flag_end_turn_actions,
                       int32_t flag_reset_s_at_end_turn, int32_t flag_monitor,
                       const xtm_monitor *monitor, uint64_t track_flags, void *stream) {
    return track_line_impl(L, P, num_turns, ele_start, num_ele_track, flag_end_turn_actions,
                           flag_reset_s_at_end_turn, flag_monitor, monitor, track_flags,
                           stream);
}

int32_t xtm_features_record(const xtm_particles *P, const xtm_features *F, void *stream) {
    if (!P || !F) return fail(XTM_ERR_ARG, "xtm_features_record: null argument");
    if (!features_ok(F, P->capacity))
        return fail(XTM_ERR_ARG, "xtm_features_record: accumulators missing or of the wrong "
                                 "capacity");
    if (P->capacity == 0) return XTM_OK;
    const int64_t nblocks = (P->capacity + 255) / 256;
    if (nblocks > 0x7fffffff) return fail(XTM_ERR_ARG, "too many particles");
    hipLaunchKernelGGL(features_kernel, dim3((unsigned)nblocks), dim3(256), 0,
                       (hipStream_t)stream, *F, *P);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? XTM_OK : hip_fail(e, "features_kernel launch");
}

int32_t xtm_track_line_features(xtm_line *L, const xtm_particles *P, int32_t num_turns,
                                const xtm_features *features, uint64_t track_flags,
                                void *stream) {
    if (!L || !P || !features)
        return fail(XTM_ERR_ARG, "xtm_track_line_features: null argument");
    if ((track_flags >> XTM_FLAG_BACKTRACK) & 1)
        return fail(XTM_ERR_UNSUPPORTED, "xtm_track_line_features: tracks forward only");
    for (int32_t t = 0; t < num_turns; t++) {
        int32_t st = xtm_features_record(P, features, stream);
        if (st == XTM_OK)
            st = track_line_impl(L, P, 1, 0, (int32_t)L->num_elements, 1, 1, 0, nullptr,
                                 track_flags, stream);
        if (st != XTM_OK) return st;
    }
    return XTM_OK;
}


int32_t xtm_line_count_ops(const xtm_line *L, const double *coords, double p0c,
                           double mass0, double q0, double *out) {
    if (!L || !coords || !out) return fail(XTM_ERR_ARG, "xtm_line_count_ops: null argument");
    int st = xt_count_turn(L->entry, L->prog, L->line_length, coords, p0c, mass0, q0, out);
    if (st != XTM_OK) return fail(st, "xtm_line_count_ops: probe particle lost");
    return XTM_OK;
}

int32_t xtm_batch_destroy(xtm_batch *B) {
    if (!B) return XTM_OK;
    if (B->device >= 0) {
        (void)hipSetDevice(B->device);
        if (B->d_entries) (void)hipFree(B->d_entries);
        if (B->d_progs) (void)hipFree(B->d_progs);
        if (B->d_nele) (void)hipFree(B->d_nele);
        if (B->d_lengths) (void)hipFree(B->d_lengths);
        if (B->d_blocks) (void)hipFree(B->d_blocks);
        if (B->d_p) (void)hipFree(B->d_p);
        if (B->d_m) (void)hipFree(B->d_m);
        for (int32_t *r : B->d_recs)
            if (r) (void)hipFree(r);
        if (B->d_rec_ptrs) (void)hipFree(B->d_rec_ptrs);
        if (B->module) (void)hipModuleUnload(B->module);
    }
    delete B;
    return XTM_OK;
}

int32_t xtm_batch_create(xtm_line *const *lines, int32_t num_lines, xtm_batch **out) {
    if (!out || !lines || num_lines <= 0)
        return fail(XTM_ERR_ARG, "xtm_batch_create: invalid arguments");
    *out = nullptr;
    const int32_t device = lines[0] ? lines[0]->device : -1;
    for (int32_t k = 0; k < num_lines; k++) {
        if (!lines[k]) return fail(XTM_ERR_ARG, "xtm_batch_create: null line");
        if (lines[k]->device != device)
            return fail(XTM_ERR_ARG, "xtm_batch_create: lines on different devices");
        if (lines[k]->num_elements > 0x7fffffff)
            return fail(XTM_ERR_ARG, "xtm_batch_create: line too long");
    }
    xtm_batch *B = new xtm_batch();
    B->device = device;
    B->lines.assign(lines, lines + num_lines);
    std::vector<const int64_t *> entries(num_lines);
    std::vector<const double *> progs(num_lines);
    std::vector<double> lengths(num_lines);
    for (int32_t k = 0; k < num_lines; k++) {
        B->nele.push_back(lines[k]->num_elements);
        entries[k] = lines[k]->d_entry;
        progs[k] = lines[k]->d_prog;
        lengths[k] = lines[k]->line_length;
    }
    if (device < 0) {
        *out = B;
        return XTM_OK;
    }
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        xtm_batch_destroy(B);
        return hip_fail(e, "hipSetDevice");
    }
    const size_t n = (size_t)num_lines;
    if ((e = hipMalloc(&B->d_entries, n * sizeof(int64_t *))) != hipSuccess ||
        (e = hipMalloc(&B->d_progs, n * sizeof(double *))) != hipSuccess ||
        (e = hipMalloc(&B->d_nele, n * sizeof(int64_t))) != hipSuccess ||
        (e = hipMalloc(&B->d_lengths, n * sizeof(double))) != hipSuccess ||
        (e = hipMalloc(&B->d_p, sizeof(xtm_particles))) != hipSuccess ||
        (e = hipMalloc(&B->d_m, sizeof(xtm_monitor))) != hipSuccess) {
        xtm_batch_destroy(B);
        return hip_fail(e, "hipMalloc");
    }
    if ((e = hipMemcpy(B->d_entries, entries.data(), n * sizeof(int64_t *),
                       hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(B->d_progs, progs.data(), n * sizeof(double *),
                       hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(B->d_nele, B->nele.data(), n * sizeof(int64_t),
                       hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(B->d_lengths, lengths.data(), n * sizeof(double),
                       hipMemcpyHostToDevice)) != hipSuccess) {
        xtm_batch_destroy(B);
        return hip_fail(e, "hipMemcpy");
    }
    *out = B;
    return XTM_OK;
}

int32_t xtm_batch_set_layout(xtm_batch *B, const int64_t *line_start, int32_t block_size,
                             int32_t waves_per_eu) {
    if (!B || !line_start) return fail(XTM_ERR_ARG, "xtm_batch_set_layout: null argument");
    if ((block_size != 64 && block_size != 128 && block_size != 256) || waves_per_eu < 3 ||
        waves_per_eu > 5)
        return fail(XTM_ERR_ARG, "xtm_batch_set_layout: no kernel for this block size and "
                                 "occupancy (64, 128 or 256 threads, 3 to 5 waves per SIMD)");
    const size_t nl = B->lines.size();
    if (line_start[0] != 0) return fail(XTM_ERR_ARG, "xtm_batch_set_layout: line_start[0] != 0");
    std::vector<BatchBlock> blocks;
    for (size_t k = 0; k < nl; k++) {
        if (line_start[k + 1] < line_start[k])
            return fail(XTM_ERR_ARG, "xtm_batch_set_layout: line_start decreases");
        for (int64_t f = line_start[k]; f < line_start[k + 1]; f += block_size) {
            BatchBlock bb;
            bb.first = f;
            bb.line = (int32_t)k;
            bb.count = (int32_t)std::min<int64_t>(block_size, line_start[k + 1] - f);
            blocks.push_back(bb);
        }
    }
    if (blocks.size() > 0x7fffffffu) return fail(XTM_ERR_ARG, "xtm_batch_set_layout: too many blocks");
    // longest lines first: the last blocks to start are the short ones
    std::stable_sort(blocks.begin(), blocks.end(), [&](const BatchBlock &a, const BatchBlock &b) {
        return B->nele[a.line] > B->nele[b.line];
    });
    if (B->device >= 0 && !blocks.empty()) {
        hipError_t e = hipSetDevice(B->device);
        if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
        if ((int64_t)blocks.size() > B->d_blocks_cap) {
            if (B->d_blocks) (void)hipFree(B->d_blocks);
            B->d_blocks = nullptr;
            B->d_blocks_cap = 0;
            if ((e = hipMalloc(&B->d_blocks, blocks.size() * sizeof(BatchBlock))) != hipSuccess)
                return hip_fail(e, "hipMalloc");
            B->d_blocks_cap = (int64_t)blocks.size();
        }
        if ((e = hipMemcpy(B->d_blocks, blocks.data(), blocks.size() * sizeof(BatchBlock),
                           hipMemcpyHostToDevice)) != hipSuccess)
            return hip_fail(e, "hipMemcpy");
    }
    B->line_start.assign(line_start, line_start + nl + 1);
    B->blocks = std::move(blocks);
    B->block_size = block_size;
    B->waves_per_eu = waves_per_eu;
    return XTM_OK;
}

int64_t xtm_batch_num_blocks(const xtm_batch *B) { return B ? (int64_t)B->blocks.size() : -1; }

int32_t xtm_batch_copy_blocks(const xtm_batch *B, int64_t *out) {
    if (!B || !out) return fail(XTM_ERR_ARG, "xtm_batch_copy_blocks: null argument");
    for (size_t b = 0; b < B->blocks.size(); b++) {
        out[3 * b] = B->blocks[b].line;
        out[3 * b + 1] = B->blocks[b].first;
        out[3 * b + 2] = B->blocks[b].count;
    }
    return XTM_OK;
}

static int32_t batch_track_impl(xtm_batch *B, const xtm_particles *P, int32_t num_turns,
                                uint64_t track_flags, void *stream) {
    if (!B || !P) return fail(XTM_ERR_ARG, "xtm_batch_track: null batch or particles");
    if (B->device < 0) return fail(XTM_ERR_ARG, "xtm_batch_track: batch was created without a device");
    if (B->line_start.empty()) return fail(XTM_ERR_ARG, "xtm_batch_track: no layout set");
    if (num_turns < 0) return fail(XTM_ERR_ARG, "xtm_batch_track: negative number of turns");
    if (P->capacity < B->line_start.back())
        return fail(XTM_ERR_ARG, "xtm_batch_track: particle set smaller than the layout");
    if ((track_flags >> XTM_FLAG_SR_TAPER) & 1)
        return fail(XTM_ERR_UNSUPPORTED, "SR tapering is not supported");
    if ((track_flags >> XTM_FLAG_BACKTRACK) & 1)
        return fail(XTM_ERR_UNSUPPORTED, "xtm_batch_track: backtracking is not supported");

    if (num_turns == 0 || B->blocks.empty()) return XTM_OK;
    hipError_t e = hipSetDevice(B->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    xtm_monitor M;
    std::memset(&M, 0, sizeof(M));
    const hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(args_kernel, dim3(1), dim3(64), 0, s, *P, M, B->d_p, B->d_m);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "args_kernel launch");
    const dim3 grid((unsigned)B->blocks.size());
    const int64_t *const *entries = B->d_entries;
    const double *const *progs = B->d_progs;
    int32_t skip = 0;
    if (B->kernel) {
        const void *blocks = B->d_blocks;
        const int32_t *const *recs = B->d_rec_ptrs;
        const int64_t *nele = B->d_nele;
        const double *lengths = B->d_lengths;
        const xtm_particles *Pd = B->d_p;
        int32_t a_turns = num_turns;
        uint64_t a_flags = track_flags;
        void *args[] = {&blocks, &recs, &entries, &progs, &nele, &lengths, &Pd, &a_turns,
                        &a_flags};
        e = hipModuleLaunchKernel(B->kernel, grid.x, 1, 1, (unsigned)B->block_size, 1, 1,
                                  lds_bytes(B->block_size), s, args, nullptr);
        if (e != hipSuccess) return hip_fail(e, "xtm_batch_line_kernel launch");
        skip = 1;   // the interpreter follows for the particles with chi != 1
    }
    const dim3 block((unsigned)B->block_size);
    const unsigned lds = lds_bytes(B->block_size);
    if (B->waves_per_eu == 5)
        hipLaunchKernelGGL((batch_kernel<5>), grid, block, lds, s, B->d_blocks, entries, progs,
                           B->d_nele, B->d_lengths, B->d_p, num_turns, track_flags, skip);
    else if (B->waves_per_eu == 4)
        hipLaunchKernelGGL((batch_kernel<4>), grid, block, lds, s, B->d_blocks, entries, progs,
                           B->d_nele, B->d_lengths, B->d_p, num_turns, track_flags, skip);
    else
        hipLaunchKernelGGL((batch_kernel<3>), grid, block, lds, s, B->d_blocks, entries, progs,
                           B->d_nele, B->d_lengths, B->d_p, num_turns, track_flags, skip);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "batch_kernel launch");
    return XTM_OK;
}

int32_t xtm_batch_track(xtm_batch *B, const xtm_particles *P, int32_t num_turns,
                        uint64_t track_flags, void *stream) {
    return batch_track_impl(B, P, num_turns, track_flags, stream);
}

int32_t xtm_batch_track_features(xtm_batch *B, const xtm_particles *P, int32_t num_turns,
                                 const xtm_features *features, uint64_t track_flags,
                                 void *stream) {
    if (!B || !P || !features)
        return fail(XTM_ERR_ARG, "xtm_batch_track_features: null argument");
    for (int32_t t = 0; t < num_turns; t++) {
        int32_t st = xtm_features_record(P, features, stream);
        if (st == XTM_OK) st = batch_track_impl(B, P, 1, track_flags, stream);
        if (st != XTM_OK) return st;
    }
    return XTM_OK;
}

int32_t xtm_batch_generate_source(const xtm_batch *B, int32_t waves_per_eu, char *buf,
                                  int64_t cap, int64_t *len, int32_t *recs) {
    if (!B || !len) return fail(XTM_ERR_ARG, "xtm_batch_generate_source: null argument");
    std::vector<const std::vector<int64_t> *> entries;
    std::vector<const std::vector<double> *> progs;
    for (const xtm_line *L : B->lines) {
        entries.push_back(&L->entry);
        progs.push_back(&L->prog);
    }
    std::string src, err;
    std::vector<std::vector<int32_t>> rec;
    int st = xt_generate_batch_source(entries, progs, waves_per_eu, src, rec, err);
    if (st != XTM_OK) return fail(st, "xtm_batch_generate_source: " + err);
    *len = (int64_t)src.size();
    if (buf && cap > 0) {
        const int64_t n = (int64_t)src.size() < cap - 1 ? (int64_t)src.size() : cap - 1;
        std::memcpy(buf, src.data(), (size_t)n);
        buf[n] = '\0';
    }
    if (recs) {
        int64_t o = 0;
        for (const auto &r : rec) {
            if (!r.empty()) std::memcpy(recs + o, r.data(), r.size() * sizeof(int32_t));
            o += (int64_t)r.size();
        }
    }
    return XTM_OK;
}

int32_t xtm_batch_specialize(xtm_batch *B, int32_t waves_per_eu) {
    if (!B) return fail(XTM_ERR_ARG, "xtm_batch_specialize: null batch");
    if (B->device < 0)
        return fail(XTM_ERR_ARG, "xtm_batch_specialize: batch was created without a device");
    if (waves_per_eu < 0 || waves_per_eu > 8)
        return fail(XTM_ERR_ARG, "xtm_batch_specialize: waves_per_eu outside 0..8");
    if (waves_per_eu == 0) waves_per_eu = XTM_COMPILED_WAVES_PER_EU;
    std::vector<const std::vector<int64_t> *> entries;
    std::vector<const std::vector<double> *> progs;
    for (const xtm_line *L : B->lines) {
        entries.push_back(&L->entry);
        progs.push_back(&L->prog);
    }
    std::string src, err;
    std::vector<std::vector<int32_t>> rec;
    int st = xt_generate_batch_source(entries, progs, waves_per_eu, src, rec, err);
    if (st != XTM_OK) return fail(st, "xtm_batch_specialize: " + err);
    std::vector<char> code;
    st = rtc_compile(src, code);
    if (st != XTM_OK) return st;
    hipError_t e = hipSetDevice(B->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipModule_t mod = nullptr;
    if ((e = hipModuleLoadData(&mod, code.data())) != hipSuccess)
        return hip_fail(e, "hipModuleLoadData");
    hipFunction_t fn = nullptr;
    if ((e = hipModuleGetFunction(&fn, mod, "xtm_batch_line_kernel")) != hipSuccess) {
        (void)hipModuleUnload(mod);
        return hip_fail(e, "hipModuleGetFunction");
    }
    std::vector<int32_t *> d_recs(rec.size(), nullptr);
    const int32_t **d_ptrs = nullptr;
    auto cleanup = [&]() {
        for (int32_t *r : d_recs)
            if (r) (void)hipFree(r);
        if (d_ptrs) (void)hipFree(d_ptrs);
        (void)hipModuleUnload(mod);
    };
    for (size_t l = 0; l < rec.size(); l++) {
        const size_t n = rec[l].empty() ? 1 : rec[l].size();
        if ((e = hipMalloc(&d_recs[l], n * sizeof(int32_t))) != hipSuccess) {
            cleanup();
            return hip_fail(e, "hipMalloc");
        }
        if (!rec[l].empty() &&
            (e = hipMemcpy(d_recs[l], rec[l].data(), rec[l].size() * sizeof(int32_t),
                           hipMemcpyHostToDevice)) != hipSuccess) {
            cleanup();
            return hip_fail(e, "hipMemcpy");
        }
    }
    if ((e = hipMalloc(&d_ptrs, rec.size() * sizeof(int32_t *))) != hipSuccess ||
        (e = hipMemcpy(d_ptrs, d_recs.data(), rec.size() * sizeof(int32_t *),
                       hipMemcpyHostToDevice)) != hipSuccess) {
        cleanup();
        return hip_fail(e, "hipMalloc/hipMemcpy");
    }
    for (int32_t *r : B->d_recs)
        if (r) (void)hipFree(r);
    if (B->d_rec_ptrs) (void)hipFree(B->d_rec_ptrs);
    if (B->module) (void)hipModuleUnload(B->module);
    B->d_recs = std::move(d_recs);
    B->d_rec_ptrs = d_ptrs;
    B->module = mod;
    B->kernel = fn;
    return XTM_OK;
}

int32_t xtm_batch_is_specialized(const xtm_batch *B) { return (B && B->kernel) ? 1 : 0; }

int32_t xtm_synchronize(void *stream) {
    hipError_t e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    return XTM_OK;
}

int32_t xtm_math_eval(int32_t fn, const double *x, double *y, int64_t n, void *stream) {
    if (fn < 0 || fn > 12 || n < 0 || (n > 0 && (!x || !y)))
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
