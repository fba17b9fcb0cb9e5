// Host build of the tracking kernel's per-particle code
// (xtrack_amd/csrc/xt_ops.h + the element-program builder) for the CPU test
// suite: TEST INFRASTRUCTURE ONLY.  It lets the CPU tests compare the
// kernel's arithmetic with the oracle without a GPU; the GPU tests then
// show that the device build of the same code gives the same bits.
#include <cstdint>
#include <string>
#include <vector>

#include "xt_builder.h"
#include "xt_ops.h"

extern "C" int32_t xth_track_line(const int32_t *types, const int64_t *offsets,
                                  const double *params, int64_t n, double line_length,
                                  const xtm_particles *P, int32_t num_turns, int32_t ele_start,
                                  int32_t num_ele_track, int32_t flag_end_turn_actions,
                                  int32_t flag_reset_s, int32_t flag_monitor,
                                  const xtm_monitor *M, uint64_t flags) {
    std::vector<int64_t> entry, entry_bt;
    std::vector<double> prog;
    std::string err;
    int st = xt_build_program(types, offsets, params, n, entry, prog, err, &entry_bt);
    if (st) return st;
    xtm_monitor m0 = {};
    const xt::MonRef MM{M ? M : &m0, P};
    const int64_t *ent = (flags & 1) ? entry_bt.data() : entry.data();
    for (int64_t i = 0; i < P->capacity; i++)
        xt::track_particle(*P, i, ent, prog.data(), line_length, n, num_turns, ele_start,
                           num_ele_track, flag_end_turn_actions, flag_reset_s, flag_monitor,
                           MM, flags);
    return 0;
}

// div_rcp (reciprocal division used by the kernel) vs IEEE division on
// n pairs; returns the number of mismatches (bit patterns).
extern "C" int64_t xth_check_div_rcp(const double *a, const double *b, int64_t n) {
    int64_t bad = 0;
    for (int64_t i = 0; i < n; i++) {
        const double y = 1.0 / b[i];
        const double q = xt::div_rcp(a[i], b[i], y);
        const double r = a[i] / b[i];
        if (__builtin_memcmp(&q, &r, 8) != 0) bad++;
    }
    return bad;
}
