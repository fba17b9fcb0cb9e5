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
#include "xt_math.h"
#include "xt_program.h"

#define DEV __device__ __forceinline__
#define COLD __device__ __attribute__((noinline))

namespace {

constexpr double C_LIGHT = 299792458.0;
constexpr double PI = 3.1415926535897932384626433832795028841971693993751;
constexpr double DEG2RAD = 0.0174532925199432957692369076848861271344287188854;
constexpr double GLOBAL_XY_LIMIT = 1.0;   // xtrack/line.py:287
constexpr int BLOCK = 256;

struct PState {
    double x, px, y, py, zeta, delta, ptau, rpp, rvv, s, ax, ay;
    double chi, charge_ratio, beta0, p0c;
    int64_t at_element, at_turn, state;
};

DEV int64_t wbits(double d) { return __double_as_longlong(d); }

// ---- drift maps ---------------------------------------------------------
// track_drift.h:11-23
DEV void drift_expanded(PState &p, double length) {
    double const rpp = p.rpp;
    double const rv0v = 1. / p.rvv;
    double const xp = p.px * rpp;
    double const yp = p.py * rpp;
    double const dzeta = 1 - rv0v * (1. + (xp * xp + yp * yp) / 2.);
    p.x += xp * length;
    p.y += yp * length;
    p.s += length;
    p.zeta += length * dzeta;
}

// track_drift.h:26-41
DEV void drift_exact(PState &p, double length) {
    double const px = p.px, py = p.py;
    double const rv0v = 1. / p.rvv;
    double const one_plus_delta = 1. + p.delta;
    double const one_over_pz =
        1. / XT_SQRT(one_plus_delta * one_plus_delta - px * px - py * py);
    double const dzeta = 1 - rv0v * one_plus_delta * one_over_pz;
    p.x += px * one_over_pz * length;
    p.y += py * one_over_pz * length;
    p.zeta += dzeta * length;
    p.s += length;
}

// polar drift, track_magnet_drift.h:45-87, with rho = 1/h, ca = cos(h s),
// sa = sin(h s), t2 = 2 sin(h s/2)^2 supplied by the element program
DEV void polar_drift(PState &p, double length, double rho, double ca, double sa,
                     double t2) {
    const double rvv = p.rvv;
    const double x = p.x, y = p.y, px = p.px, py = p.py;
    const double one_plus_delta = p.delta + 1.0;
    const double pz = XT_SQRT(one_plus_delta * one_plus_delta - px * px - py * py);
    const double _pz = 1 / pz;
    const double pxt = px * _pz;
    const double _ptt = 1 / (ca - sa * pxt);
    const double pst = (x + rho) * sa * _pz * _ptt;
    const double new_x = (x + rho * (t2 + sa * pxt)) * _ptt;
    const double new_px = ca * px + sa * pz;
    const double new_y = y + pst * py;
    const double delta_ell =
        one_plus_delta * (x + rho) * sa / ca / pz / (1 - px * sa / ca / pz);
    p.x = new_x;
    p.px = new_px;
    p.y = new_y;
    p.zeta += length - delta_ell / rvv;
    p.s += length;
}

// track_magnet_drift.h:91-213
COLD void dipole_quad(PState &p, const double length, const double k0_, const double k1_,
                     const double h) {
    const double x = p.x, y = p.y, px = p.px, py = p.py, rvv = p.rvv;
    const double delta_plus_1 = p.delta + 1;
    const double chi = p.chi;
    const double k0 = chi * k0_ / delta_plus_1;
    const double k1 = chi * k1_ / delta_plus_1;
    const double Kx = k0 * h + k1;
    const double Ky = -k1;
    double Sx, Sy, Cx, Cy;
    if (Kx > 0.0) {
        double sqrt_Kx = XT_SQRT(Kx), sn, cs;
        xt_sincos(sqrt_Kx * length, &sn, &cs);
        Sx = sn / sqrt_Kx;
        Cx = cs;
    } else if (Kx < 0.0) {
        double sqrt_Kx = XT_SQRT(-Kx), sh, ch;
        xt_sinhcosh(sqrt_Kx * length, &sh, &ch);
        Sx = sh / sqrt_Kx;
        Cx = ch;
    } else {
        Sx = length;
        Cx = 1.0;
    }
    if (Ky > 0.0) {
        double sqrt_Ky = XT_SQRT(Ky), sn, cs;
        xt_sincos(sqrt_Ky * length, &sn, &cs);
        Sy = sn / sqrt_Ky;
        Cy = cs;
    } else if (Ky < 0.0) {
        double sqrt_Ky = XT_SQRT(-Ky), sh, ch;
        xt_sinhcosh(sqrt_Ky * length, &sh, &ch);
        Sy = sh / sqrt_Ky;
        Cy = ch;
    } else {
        Sy = length;
        Cy = 1.0;
    }
    const double xp = px / delta_plus_1;
    const double yp = py / delta_plus_1;
    const double A = -Kx * x - k0 + h;
    const double B = xp;
    const double C = -Ky * y;
    const double D = yp;
    double x_ = x * Cx + xp * Sx;
    const double y_ = y * Cy + yp * Sy;
    const double px_ = (A * Sx + B * Cx) * delta_plus_1;
    const double py_ = (C * Sy + D * Cy) * delta_plus_1;
    if (Kx != 0.0)
        x_ = x_ + (k0 - h) * (Cx - 1.0) / Kx;
    else
        x_ = x_ - (k0 - h) * 0.5 * (length * length);
    double length_ = length;
    if (Kx != 0.0) {
        length_ -= (h * ((Cx - 1.0) * xp + Sx * A + length * (k0 - h))) / Kx;
        length_ += 0.5 * (-((A * A) * Cx * Sx) / (2.0 * Kx) + ((B * B) * Cx * Sx) / 2.0 +
                          ((A * A) * length) / (2.0 * Kx) + ((B * B) * length) / 2.0 -
                          (A * B * (Cx * Cx)) / Kx + (A * B) / Kx);
    } else {
        length_ += h * length * (3.0 * length * xp + 6.0 * x - (k0 - h) * (length * length)) / 6.0;
        length_ += 0.5 * (B * B) * length;
    }
    if (Ky != 0.0) {
        length_ += 0.5 * (-((C * C) * Cy * Sy) / (2.0 * Ky) + ((D * D) * Cy * Sy) / 2.0 +
                          ((C * C) * length) / (2.0 * Ky) + ((D * D) * length) / 2.0 -
                          (C * D * (Cy * Cy)) / Ky + (C * D) / Ky);
    } else {
        length_ += 0.5 * (D * D) * length;
    }
    const double dzeta = length - length_ / rvv;
    p.x = x_;
    p.px = px_;
    p.y = y_;
    p.py = py_;
    p.zeta += dzeta;
    p.s += length;
}

// ---- multipole kick (track_magnet_kick.h:24-144, 183-228) ------------------
// f: inv-factorial chain f[order], f[order-1], ..., f[0]
DEV void kick_pass(PState &p, int64_t order, const double *__restrict__ f,
                   const double *__restrict__ knl, const double *__restrict__ ksl,
                   double factor, double kick_weight, double x, double y, double chi) {
    int64_t index = order;
    double dpx_mul = chi * knl[index] * factor * f[0];
    double dpy_mul = chi * ksl[index] * factor * f[0];
    int64_t j = 0;
    while (index > 0) {
        double const zre = dpx_mul * x - dpy_mul * y;
        double const zim = dpx_mul * y + dpy_mul * x;
        index -= 1;
        j += 1;
        double this_knl = chi * knl[index] * factor;
        double this_ksl = chi * ksl[index] * factor;
        dpx_mul = this_knl * f[j] + zre;
        dpy_mul = this_ksl * f[j] + zim;
    }
    dpx_mul = -dpx_mul;
    p.px += kick_weight * dpx_mul;
    p.py += kick_weight * dpy_mul;
}

COLD void magnet_kick(PState &p, double kick_weight, const double *__restrict__ d) {
    const double chi = p.chi, x = p.x, y = p.y;
    const int64_t order = wbits(d[KD_ORDER]);
    const int64_t order_rel = wbits(d[KD_ORDER_REL]);
    const double length = d[KD_LENGTH];
    const double *f1 = d + KD_ARR;
    const double *knl = f1 + (order + 1);
    const double *ksl = knl + (order + 1);
    const double *f2 = ksl + (order + 1);
    const double *knl_rel = f2 + (order_rel + 1);
    const double *ksl_rel = knl_rel + (order_rel + 1);
    const double *f3 = ksl_rel + (order_rel + 1);
    kick_pass(p, order, f1, knl, ksl, d[KD_FACTOR], kick_weight, x, y, chi);
    kick_pass(p, order_rel, f2, knl_rel, ksl_rel, d[KD_FACTOR_REL], kick_weight, x, y, chi);
    kick_pass(p, 3, f3, d + KD_MAIN, d + KD_MAINS, 1, kick_weight, x, y, chi);
    double dpx = 0, dpy = 0, dzeta = 0;
    if (wbits(d[KD_ROT_FRAME])) {
        double const hl = d[KD_H] * length * kick_weight + d[KD_HXL] * kick_weight;
        dpx += hl * (1. + p.delta);
        double const rv0v = 1. / p.rvv;
        dzeta += -rv0v * hl * x;
    }
    const double htot = d[KD_HTOT];
    const double k0c = d[KD_K0HC] * length + d[KD_K0L];
    const double k1c = d[KD_K1HC] * length + d[KD_K1L];
    dpx += -chi * k0c * kick_weight * htot * x;
    dpx += htot * chi * k1c * kick_weight * (-x * x + 0.5 * y * y);
    dpy += htot * chi * k1c * kick_weight * x * y;
    p.px += dpx;
    p.py += dpy;
    p.zeta += dzeta;
}

// ---- RF (track_rf.h:18-167) and energy update (custom_api.h:21-32,196-216)
DEV void update_ptau(PState &p, double new_ptau) {
    double const beta0 = p.beta0;
    double const ptau = new_ptau;
    double const irpp = XT_SQRT(ptau * ptau + 2 * ptau / beta0 + 1);
    double const new_rpp = 1. / irpp;
    double const new_rvv = irpp / (1 + beta0 * ptau);
    p.delta = irpp - 1;
    p.rvv = new_rvv;
    p.ptau = ptau;
    p.rpp = new_rpp;
}

COLD void rf_kick(PState &p, const double *__restrict__ a, double line_length, double q0,
                 double t_sim) {
    const double voltage = a[0];
    double frequency = a[1];
    const double harmonic = a[2];
    const double lag = a[3];
    const double phase = a[4];
    const int64_t absolute_time = wbits(a[5]);
    double phase0 = 0;
    double const beta0 = p.beta0;
    if (harmonic != 0) {
        double const t_rev0 = line_length / (beta0 * C_LIGHT);
        frequency += (harmonic / t_rev0);
    }
    if (absolute_time == 1) {
        phase0 += 2 * PI * (double)p.at_turn * frequency * t_sim;
    }
    double const zeta = p.zeta;
    double const q = XT_FABS(q0) * p.charge_ratio;
    double const tau = zeta / beta0;
    double const energy_kick =
        q * voltage * xt_sin(phase0 + DEG2RAD * lag + phase - (2.0 * PI) / C_LIGHT * frequency * tau);
    double const rfmultipole_energy_kick = 0;
    // LocalParticle_add_to_energy(part, dE, pz_only = 1)
    double ptau = p.ptau;
    double const mass_ratio = p.charge_ratio / p.chi;
    ptau += (energy_kick + rfmultipole_energy_kick) / p.p0c / mass_ratio;
    update_ptau(p, ptau);
}

// ---- monitor (monitors/particles_monitor.h:13-77) ---------------------------
DEV void monitor_store(const xtm_monitor &M, const xtm_particles &P, int64_t i,
                       const PState &p, int64_t store_at) {
    const xtm_particles &D = M.data;
    D.p0c[store_at] = p.p0c;
    D.gamma0[store_at] = P.gamma0[i];
    D.beta0[store_at] = p.beta0;
    D.s[store_at] = p.s;
    D.zeta[store_at] = p.zeta;
    D.x[store_at] = p.x;
    D.y[store_at] = p.y;
    D.px[store_at] = p.px;
    D.py[store_at] = p.py;
    D.ptau[store_at] = p.ptau;
    D.delta[store_at] = p.delta;
    D.rpp[store_at] = p.rpp;
    D.rvv[store_at] = p.rvv;
    D.chi[store_at] = p.chi;
    D.charge_ratio[store_at] = p.charge_ratio;
    D.weight[store_at] = P.weight[i];
    D.ax[store_at] = p.ax;
    D.ay[store_at] = p.ay;
    D.spin_x[store_at] = P.spin_x[i];
    D.spin_y[store_at] = P.spin_y[i];
    D.spin_z[store_at] = P.spin_z[i];
    D.anomalous_magnetic_moment[store_at] = P.anomalous_magnetic_moment[i];
    D.pdg_id[store_at] = P.pdg_id[i];
    D.particle_id[store_at] = P.particle_id[i];
    D.at_element[store_at] = p.at_element;
    D.at_turn[store_at] = p.at_turn;
    D.state[store_at] = p.state;
    D.parent_particle_id[store_at] = P.parent_particle_id[i];
    D.rng_s1[store_at] = P.rng_s1[i];
    D.rng_s2[store_at] = P.rng_s2[i];
    D.rng_s3[store_at] = P.rng_s3[i];
    D.rng_s4[store_at] = P.rng_s4[i];
}

COLD void monitor_record(const xtm_monitor &M, const xtm_particles &P, int64_t i,
                        const PState &p, int64_t pid) {
    const int64_t n_turns_record = M.stop_at_turn - M.start_at_turn;
    const int64_t at_turn = M.ebe_mode ? p.at_element : p.at_turn;
    if (M.n_repetitions == 1) {
        if (at_turn >= M.start_at_turn && at_turn < M.stop_at_turn &&
            pid < M.part_id_end && pid >= M.part_id_start) {
            monitor_store(M, P, i, p,
                          n_turns_record * (pid - M.part_id_start) + at_turn - M.start_at_turn);
        }
    } else if (M.n_repetitions > 1 && at_turn >= M.start_at_turn) {
        const int64_t i_frame = (at_turn - M.start_at_turn) / M.repetition_period;
        if (i_frame < M.n_repetitions &&
            at_turn >= M.start_at_turn + i_frame * M.repetition_period &&
            at_turn < M.stop_at_turn + i_frame * M.repetition_period &&
            pid < M.part_id_end && pid >= M.part_id_start) {
            monitor_store(M, P, i, p,
                          n_turns_record * (M.part_id_end - M.part_id_start) * i_frame +
                              n_turns_record * (pid - M.part_id_start) +
                              (at_turn - i_frame * M.repetition_period) - M.start_at_turn);
        }
    }
}

// ---- element program interpreter -------------------------------------------
DEV void run_element(PState &p, const double *__restrict__ prog, int64_t pc,
                     uint64_t flags, double line_length, double q0, double t_sim) {
    const bool ignore_local = (flags >> XTM_FLAG_IGNORE_LOCAL_APERTURE) & 1;
    const bool ignore_global = (flags >> XTM_FLAG_IGNORE_GLOBAL_APERTURE) & 1;
    const bool kill_cavity = (flags >> XTM_FLAG_KILL_CAVITY_KICK) & 1;
    for (;;) {
        const int op = (int)wbits(prog[pc]);
        const double *__restrict__ a = prog + pc + 1;
        switch (op) {
        case OP_END:
            return;
        case OP_DRIFT_EXP:
            drift_expanded(p, a[0]);
            pc += 2;
            break;
        case OP_DRIFT_EXACT:
            drift_exact(p, a[0]);
            pc += 2;
            break;
        case OP_POLAR:
            polar_drift(p, a[0], a[1], a[2], a[3], a[4]);
            pc += 6;
            break;
        case OP_DIPQUAD:
            dipole_quad(p, a[0], a[1], a[2], a[3]);
            pc += 5;
            break;
        case OP_YOSHIDA4_BEND: {
            // track_magnet_drift.h:500-509
            const double L = a[0], rho = a[1], c1 = a[2], c2 = a[3];
            polar_drift(p, a[4], rho, a[5], a[6], a[7]);
            p.px = p.px - c1 * p.chi * L;
            polar_drift(p, a[8], rho, a[9], a[10], a[11]);
            p.px = p.px - c2 * p.chi * L;
            polar_drift(p, a[8], rho, a[9], a[10], a[11]);
            p.px = p.px - c1 * p.chi * L;
            polar_drift(p, a[4], rho, a[5], a[6], a[7]);
            pc += 13;
            break;
        }
        case OP_YOSHIDA6_BEND: {
            // track_magnet_drift.h:510-549
            const double L = a[0], rho = a[1];
            const double *kc = a + 2;
            const double *dd = a + 6;
            const int di[8] = {0, 1, 2, 3, 3, 2, 1, 0};
            const int ki[7] = {0, 1, 2, 3, 2, 1, 0};
            for (int j = 0; j < 8; j++) {
                const double *s = dd + 4 * di[j];
                polar_drift(p, s[0], rho, s[1], s[2], s[3]);
                if (j < 7) p.px = p.px - kc[ki[j]] * p.chi * L;
            }
            pc += 23;
            break;
        }
        case OP_KICK:
            magnet_kick(p, a[0], prog + wbits(a[1]));
            pc += 3;
            break;
        case OP_EDGE_AXAY:
            p.ax = a[0] * (p.y - 0.0);
            p.ay = a[1] * (p.x - 0.0);
            pc += 3;
            break;
        case OP_EDGE_CLEAR_AXAY:
            p.ax = 0.;
            p.ay = 0.;
            pc += 1;
            break;
        case OP_EDGE_LINEAR: {
            const double x = p.x, y = p.y, chi = p.chi;
            p.px += chi * a[0] * x;
            p.py += chi * a[1] * y;
            pc += 3;
            break;
        }
        case OP_RF_KICK:
            if (!kill_cavity) rf_kick(p, a, line_length, q0, t_sim);
            pc += 7;
            break;
        case OP_LIMIT_RECT:
            if (!ignore_local) {
                const double x = p.x, y = p.y;
                const bool is_alive = (x >= a[0]) && (x <= a[1]) && (y >= a[2]) && (y <= a[3]);
                if (!is_alive) p.state = 0;
            }
            pc += 5;
            break;
        case OP_LIMIT_ELLIPSE:
            if (!ignore_local) {
                const double x = p.x, y = p.y;
                const double temp = x * x * a[1] + y * y * a[0];
                if (!(temp <= a[2])) p.state = 0;
            }
            pc += 4;
            break;
        case OP_GLOBAL_APER:
            if (!ignore_global) {
                const double x = p.x, y = p.y;
                const bool within = (x >= -GLOBAL_XY_LIMIT) && (x <= GLOBAL_XY_LIMIT) &&
                                    (y >= -GLOBAL_XY_LIMIT) && (y <= GLOBAL_XY_LIMIT);
                if (p.state > 0 && !within) p.state = -1;
            }
            pc += 1;
            break;
        default:
            p.state = -999;   // corrupt program: never produced by the builder
            return;
        }
    }
}

__global__ __launch_bounds__(BLOCK) void track_kernel(
    const int64_t *__restrict__ entry, const double *__restrict__ prog, double line_length,
    xtm_particles P, int32_t num_turns, int32_t ele_start, int32_t num_ele_track,
    int32_t flag_end_turn_actions, int32_t flag_reset_s, int32_t flag_monitor, xtm_monitor M,
    uint64_t flags) {
    const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= P.capacity) return;
    if (!(P.state[i] > 0)) return;
    PState p;
    p.x = P.x[i]; p.px = P.px[i]; p.y = P.y[i]; p.py = P.py[i];
    p.zeta = P.zeta[i]; p.delta = P.delta[i]; p.ptau = P.ptau[i];
    p.rpp = P.rpp[i]; p.rvv = P.rvv[i]; p.s = P.s[i];
    p.ax = P.ax[i]; p.ay = P.ay[i];
    p.chi = P.chi[i]; p.charge_ratio = P.charge_ratio[i];
    p.beta0 = P.beta0[i]; p.p0c = P.p0c[i];
    p.at_element = P.at_element[i]; p.at_turn = P.at_turn[i]; p.state = P.state[i];
    const int64_t pid = flag_monitor ? P.particle_id[i] : 0;
    const double q0 = P.q0, t_sim = P.t_sim;
    const int64_t ele_stop = (int64_t)ele_start + num_ele_track;

    for (int32_t iturn = 0; iturn < num_turns; iturn++) {
        if (flag_monitor == 1) monitor_record(M, P, i, p, pid);
        bool lost = false;
        for (int64_t e = ele_start; e < ele_stop; e++) {
            if (flag_monitor == 2) monitor_record(M, P, i, p, pid);
            run_element(p, prog, entry[e], flags, line_length, q0, t_sim);
            if (p.state <= 0) { lost = true; break; }
            p.at_element += 1;
        }
        if (lost) break;
        if (flag_monitor == 2) monitor_record(M, P, i, p, pid);
        if (flag_end_turn_actions > 0) {
            p.at_turn += 1;
            p.at_element = 0;
            if (flag_reset_s > 0) p.s = 0.;
        }
    }
    P.x[i] = p.x; P.px[i] = p.px; P.y[i] = p.y; P.py[i] = p.py;
    P.zeta[i] = p.zeta; P.delta[i] = p.delta; P.ptau[i] = p.ptau;
    P.rpp[i] = p.rpp; P.rvv[i] = p.rvv; P.s[i] = p.s;
    P.ax[i] = p.ax; P.ay[i] = p.ay;
    P.at_element[i] = p.at_element; P.at_turn[i] = p.at_turn; P.state[i] = p.state;
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
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    xtm_line *L = new xtm_line();
    L->device = device;
    L->num_elements = num_elements;
    L->num_records = xt_count_records(entry);
    L->line_length = line_length;
    L->d_entry = nullptr;
    L->d_prog = nullptr;
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
    *out = L;
    return XTM_OK;
}

int32_t xtm_line_destroy(xtm_line *L) {
    if (!L) return XTM_OK;
    (void)hipSetDevice(L->device);
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
