// Host-side code of the library (element-program builder, line and batch
// code generators, operation counter, the kernel's element code run on the
// host) and the oracle, built with AddressSanitizer and
// UndefinedBehaviorSanitizer and driven over element tables written by
// tests/test_sanitize.py.  TEST INFRASTRUCTURE ONLY.
//
// Table file: int64 n, int64 n_params, int32 types[n], int64 offsets[n],
// double params[n_params], double line_length.  Particles: a handful of
// coordinates around the axis, tracked 2 turns by the host kernel code and
// by the oracle, forward and (host kernel) backward.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "xt_builder.h"
#include "xt_codegen.h"
#include "xt_flops.h"
#include "xt_ops.h"

extern "C" int32_t xto_track_line(const int32_t *types, const int64_t *offsets,
                                  const double *params, int64_t num_elements,
                                  double line_length, const xtm_particles *P,
                                  int32_t num_turns, int32_t ele_start, int32_t num_ele_track,
                                  int32_t flag_end_turn_actions,
                                  int32_t flag_reset_s_at_end_turn, int32_t flag_monitor,
                                  const xtm_monitor *monitor, uint64_t track_flags,
                                  int32_t num_threads);

namespace {

struct Table {
    std::vector<int32_t> types;
    std::vector<int64_t> offsets;
    std::vector<double> params;
    double line_length = 0.0;
};

bool read_table(const char *path, Table &t) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    int64_t n = 0, np = 0;
    bool ok = std::fread(&n, 8, 1, f) == 1 && std::fread(&np, 8, 1, f) == 1 && n >= 0 &&
              np >= 0;
    if (ok) {
        t.types.resize(n);
        t.offsets.resize(n);
        t.params.resize(np);
        ok = std::fread(t.types.data(), 4, n, f) == (size_t)n &&
             std::fread(t.offsets.data(), 8, n, f) == (size_t)n &&
             std::fread(t.params.data(), 8, np, f) == (size_t)np &&
             std::fread(&t.line_length, 8, 1, f) == 1;
    }
    std::fclose(f);
    return ok;
}

/* particle arrays for the SoA struct */
struct Beam {
    int64_t n;
    std::vector<double> f[22];
    std::vector<int64_t> l[6];
    std::vector<uint32_t> u[4];
    xtm_particles P;

    explicit Beam(int64_t n_, double p0c) : n(n_) {
        for (auto &v : f) v.assign(n, 0.0);
        for (auto &v : l) v.assign(n, 0);
        for (auto &v : u) v.assign(n, 0);
        std::memset(&P, 0, sizeof(P));
        P.capacity = n;
        P.q0 = 1.0;
        P.mass0 = 938272088.16;
        double **fp[22] = {&P.p0c, &P.gamma0, &P.beta0, &P.s, &P.zeta, &P.x, &P.y, &P.px,
                           &P.py, &P.ptau, &P.delta, &P.rpp, &P.rvv, &P.chi, &P.charge_ratio,
                           &P.weight, &P.ax, &P.ay, &P.spin_x, &P.spin_y, &P.spin_z,
                           &P.anomalous_magnetic_moment};
        for (int k = 0; k < 22; k++) *fp[k] = f[k].data();
        int64_t **lp[6] = {&P.pdg_id, &P.particle_id, &P.at_element, &P.at_turn, &P.state,
                           &P.parent_particle_id};
        for (int k = 0; k < 6; k++) *lp[k] = l[k].data();
        uint32_t **up[4] = {&P.rng_s1, &P.rng_s2, &P.rng_s3, &P.rng_s4};
        for (int k = 0; k < 4; k++) *up[k] = u[k].data();
        const double m = P.mass0;
        const double gamma0 = std::sqrt(p0c * p0c + m * m) / m;
        const double beta0 = p0c / std::sqrt(p0c * p0c + m * m);
        for (int64_t i = 0; i < n; i++) {
            P.p0c[i] = p0c;
            P.gamma0[i] = gamma0;
            P.beta0[i] = beta0;
            P.x[i] = 1e-4 * (double)(i % 7 - 3);
            P.y[i] = 1e-4 * (double)(i % 5 - 2);
            P.px[i] = 1e-6 * (double)(i % 3 - 1);
            P.delta[i] = 1e-4 * (double)(i % 4 - 2);
            // rpp, rvv, ptau from delta (particles.py:1960-1990)
            const double opd = 1.0 + P.delta[i];
            const double beta = opd * beta0 / std::sqrt(1.0 + beta0 * beta0 *
                                                                  (opd * opd - 1.0));
            P.rpp[i] = 1.0 / opd;
            P.rvv[i] = beta / beta0;
            P.ptau[i] = (std::sqrt(opd * opd * beta0 * beta0 * gamma0 * gamma0 + 1.0) -
                         gamma0) / (beta0 * gamma0);
            P.chi[i] = 1.0;
            P.charge_ratio[i] = 1.0;
            P.weight[i] = 1.0;
            P.particle_id[i] = i;
            P.state[i] = 1;
        }
    }
};

int run(const char *path) {
    Table t;
    if (!read_table(path, t)) {
        std::fprintf(stderr, "%s: unreadable table\n", path);
        return 2;
    }
    const int64_t n = (int64_t)t.types.size();
    std::vector<int64_t> entry, entry_bt;
    std::vector<double> prog;
    std::string err;
    int st = xt_build_program(t.types.data(), t.offsets.data(), t.params.data(), n, entry, prog,
                              err, &entry_bt);
    if (st) {
        std::fprintf(stderr, "%s: build: %s\n", path, err.c_str());
        return 3;
    }
    std::string src;
    std::vector<int32_t> rec;
    for (uint32_t opt : {0u, (uint32_t)XTM_SPEC_FAST_ARITH}) {
        st = xt_generate_line_source(entry, prog, 5, opt, src, rec, err);
        if (st != XTM_OK && st != XTM_ERR_UNSUPPORTED) {
            std::fprintf(stderr, "%s: line source: %s\n", path, err.c_str());
            return 4;
        }
    }
    std::vector<std::vector<int32_t>> recs;
    st = xt_generate_batch_source({&entry, &entry}, {&prog, &prog}, 5, src, recs, err);
    if (st != XTM_OK && st != XTM_ERR_UNSUPPORTED) {
        std::fprintf(stderr, "%s: batch source: %s\n", path, err.c_str());
        return 5;
    }
    double ops[6];
    const double coords[6] = {1e-5, 1e-7, 1e-5, 1e-7, 1e-3, 1e-5};
    (void)xt_count_turn(entry, prog, t.line_length, coords, 7e12, 938272088.16, 1.0, ops);
    // the kernel's element code on the host, forward and backward, and the oracle
    Beam a(64, 7e12), b(64, 7e12);
    xtm_monitor m0;
    std::memset(&m0, 0, sizeof(m0));
    const xt::MonRef mon{&m0, &a.P};
    for (int64_t i = 0; i < a.n; i++)
        xt::track_particle(a.P, i, entry.data(), prog.data(), t.line_length, n, 2, 0,
                           (int32_t)n, 1, 1, 0, mon, 0);
    for (int64_t i = 0; i < a.n; i++)
        xt::track_particle(a.P, i, entry_bt.data(), prog.data(), t.line_length, n, 1, 0,
                           (int32_t)n, 1, 1, 0, mon, 1);
    st = xto_track_line(t.types.data(), t.offsets.data(), t.params.data(), n, t.line_length,
                        &b.P, 2, 0, (int32_t)n, 1, 1, 0, nullptr, 0, 1);
    if (st) {
        std::fprintf(stderr, "%s: oracle status %d\n", path, st);
        return 6;
    }
    std::printf("%s: %lld elements, %zu program words, ok\n", path, (long long)n, prog.size());
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    for (int k = 1; k < argc; k++) {
        const int st = run(argv[k]);
        if (st) return st;
    }
    return 0;
}
