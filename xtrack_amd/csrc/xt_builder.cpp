// xt_builder.cpp -- raw element table -> device op programs (host side).
//
// For each element this file performs, once on the host, the per-element
// steps the reference executes at every element call on every particle:
// the per-class argument mapping (elements_src/{multipole,quadrupole,bend,
// sextupole,octupole,cavity}.h), model/integrator defaults and the automatic
// kick count (track_magnet.h:519-550), the model configuration
// (track_magnet_configure.h:7-195), the integrator step sequence
// (track_magnet.h:191-274, track_rf.h:198-310), the linear edge coefficients
// (track_dipole_edge_linear.h:12-27) and the element-uniform factors of the
// drift maps and kicks.  Every hoisted quantity is evaluated with the same
// operation order and the same elementary functions (xt_math.h) as the
// unhoisted expression, so the kernel reproduces the literal computation.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/xtrack_amd.h"
#include "xt_math.h"
#include "xt_program.h"
#include "xt_builder.h"

namespace {

// DEG2RAD of headers/constants.h
constexpr double kDeg2Rad = 0.0174532925199432957692369076848861271344287188854;

double ibits(int64_t v) { double d; std::memcpy(&d, &v, 8); return d; }

struct Prog {
    std::vector<double> &p;
    explicit Prog(std::vector<double> &v) : p(v) {}
    void op(int o) { p.push_back(ibits(o)); }
    void arg(double v) { p.push_back(v); }
    size_t pos() const { return p.size(); }
};

// 1/n! table of headers/factorial.h (correctly rounded, n < 20), tgamma above
double one_over_factorial(int64_t n) {
    static const double t[20] = {
        0x1p+0, 0x1p+0, 0x1p-1, 0x1.5555555555555p-3, 0x1.5555555555555p-5,
        0x1.1111111111111p-7, 0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13,
        0x1.a01a01a01a01ap-16, 0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22,
        0x1.ae64567f544e4p-26, 0x1.1eed8eff8d898p-29, 0x1.6124613a86d09p-33,
        0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-41, 0x1.ae7f3e733b81fp-45,
        0x1.952c77030ad4ap-49, 0x1.6827863b97d97p-53, 0x1.2f49b46814157p-57};
    if (n < 0) return 0.;
    if (n < 20) return t[n];
    return 1. / std::tgamma((double)(n + 1));
}

struct MagnetArgs {
    double length;
    int64_t order; double inv_factorial_order; const double *knl, *ksl;
    int64_t order_rel; double inv_factorial_order_rel; const double *knl_rel, *ksl_rel;
    double rel_ref_strength;
    int64_t num_multipole_kicks, model, default_model, integrator, default_integrator;
    double h, hxl, k0, k1, k2, k3, k0s, k1s, k2s, k3s;
    int64_t edge_entry_active, edge_exit_active, edge_entry_model, edge_exit_model;
    double edge_entry_angle, edge_exit_angle, edge_entry_fdown, edge_exit_fdown;
    double edge_entry_fint, edge_exit_fint, edge_entry_hgap, edge_exit_hgap;
};

struct ModelCfg {
    int64_t drift_model;
    double k0_drift, k1_drift, h_drift;
    double k0_kick, k1_kick, h_kick, k0_h_correction, k1_h_correction;
    int64_t kick_rot_frame;
};

// track_magnet_configure.h:7-195
bool configure(int64_t model, double k0, double k1, double h, ModelCfg &c) {
    if (model == 1) model = 3;
    const bool h_is_zero = std::fabs(h) < 1e-8;
    int64_t dm;
    if (model == 2) dm = h_is_zero ? 5 : 4;
    else if (model == 3) dm = h_is_zero ? 1 : 7;
    else if (model == 4) dm = 3;
    else if (model == 5) dm = 1;
    else if (model == 6) dm = 0;
    else if (model == -1) dm = -1;
    else if (model == 7) dm = h_is_zero ? 1 : 2;
    else if (model == 8) dm = h_is_zero ? 1 : 8;
    else return false;
    c.drift_model = dm;
    if (dm == -1 || dm == 0 || dm == 1) {
        c.k0_drift = 0.0; c.k1_drift = 0.0; c.h_drift = 0.0;
        c.k0_kick = k0; c.k1_kick = k1; c.h_kick = h;
        c.k0_h_correction = k0; c.k1_h_correction = k1; c.kick_rot_frame = 1;
    } else if (dm == 2) {
        c.k0_drift = 0.0; c.k1_drift = 0.0; c.h_drift = h;
        c.k0_kick = k0; c.k1_kick = k1; c.h_kick = h;
        c.k0_h_correction = k0; c.k1_h_correction = k1; c.kick_rot_frame = 0;
    } else if (dm == 3) {
        c.k0_drift = k0; c.k1_drift = k1; c.h_drift = h;
        c.k0_kick = 0.0; c.k1_kick = 0.0; c.h_kick = h;
        c.k0_h_correction = 0.; c.k1_h_correction = k1; c.kick_rot_frame = 0;
    } else if (dm == 4) {
        c.k0_drift = k0; c.k1_drift = 0.0; c.h_drift = h;
        c.k0_kick = 0.0; c.k1_kick = k1; c.h_kick = h;
        c.k0_h_correction = 0.; c.k1_h_correction = k1; c.kick_rot_frame = 0;
    } else if (dm == 5) {
        c.k0_drift = k0; c.k1_drift = 0.0; c.h_drift = 0.0;
        c.k0_kick = 0.0; c.k1_kick = k1; c.h_kick = 0.0;
        c.k0_h_correction = 0.; c.k1_h_correction = 0.; c.kick_rot_frame = 0;
    } else {  // 7, 8
        c.k0_drift = k0; c.k1_drift = 0.0; c.h_drift = h;
        c.k0_kick = 0.0; c.k1_kick = k1; c.h_kick = h;
        c.k0_h_correction = k0; c.k1_h_correction = k1; c.kick_rot_frame = 0;
    }
    return true;
}

// constants of one polar drift of length s (track_magnet_drift.h:62-70);
// ica = 1/ca, or 0 when |ca| is outside [2^-16, 2^16] (plain divisions then)
void polar_consts(double h, double s, double *ca, double *sa, double *t2, double *ica) {
    double c, sn, s2, c2;
    xt_sincos(h * s, &sn, &c);
    xt_sincos(0.5 * h * s, &s2, &c2);
    *ca = c;
    *sa = sn;
    *t2 = 2 * s2 * s2;
    // ica = RN(1/ca) enables the kernel's reciprocal-division form of the
    // polar drift; its range argument needs |h s| <= 0.1 (so ca >= 0.995,
    // |sa| <= 0.1), |1/h| in [2, 2^400] and |sa| >= 2^-400.  Otherwise 0:
    // the drift then runs with IEEE divisions only.
    const double hs = std::fabs(h * s), rho = std::fabs(1 / h);
    const bool in_range = hs <= 0.1 && rho >= 2 && rho <= 0x1p400 && std::fabs(sn) >= 0x1p-400;
    *ica = in_range ? 1.0 / c : 0.0;
}

// a slice of a magnet (track_magnet.h:286-560 as the slice elements call
// it): kind -1 is the whole element, else XTM_SLICE_*
struct SliceCfg {
    double weight = 1.0;
    int kind = -1;
};

struct Builder {
    std::vector<double> &prog;
    std::string &err;
    bool bt = false;   // build the backtracking program (XS_FLAG_BACKTRACK)
    Builder(std::vector<double> &p, std::string &e) : prog(p), err(e) {}

    // MAGNET_DRIFT / RF_DRIFT (track_magnet_drift.h:468-555)
    bool drift(double dl, const ModelCfg &c) {
        if (c.drift_model == -1) return true;
        if (dl == 0.0) return true;
        Prog P(prog);
        const double h = c.h_drift;
        switch (c.drift_model) {
        case 0: P.op(OP_DRIFT_EXP); P.arg(dl); return true;
        case 1: P.op(OP_DRIFT_EXACT); P.arg(dl); return true;
        case 2: {
            double ca, sa, t2, ica;
            polar_consts(h, dl, &ca, &sa, &t2, &ica);
            P.op(OP_POLAR); P.arg(dl); P.arg(1 / h); P.arg(ca); P.arg(sa); P.arg(t2);
            P.arg(ica);
            return true; }
        case 3:
            P.op(OP_DIPQUAD); P.arg(dl); P.arg(c.k0_drift); P.arg(c.k1_drift); P.arg(h);
            return true;
        case 4: {
            // track_magnet_drift.h:271-346: hs = h s; sin, cos of hs and of hs/2
            double ca, sa, t2, ica, s2, c2;
            polar_consts(h, dl, &ca, &sa, &t2, &ica);
            const double hs = h * dl;
            xt_sincos(hs / 2, &s2, &c2);
            P.op(OP_BEND_CURVED); P.arg(dl); P.arg(c.k0_drift); P.arg(1 / h); P.arg(hs);
            P.arg(sa); P.arg(ca); P.arg(s2 * s2); P.arg(t2); P.arg(ica);
            return true; }
        case 5: P.op(OP_BEND_STRAIGHT); P.arg(dl); P.arg(c.k0_drift); return true;
        case 7: {
            const double a = 0.6756035959798289 * dl;
            const double b = -0.17560359597982889 * dl;
            double ca, sa, t2, ia, cb, sb, tb, ib;
            polar_consts(h, a, &ca, &sa, &t2, &ia);
            polar_consts(h, b, &cb, &sb, &tb, &ib);
            P.op(OP_YOSHIDA4_BEND);
            P.arg(dl); P.arg(1 / h);
            P.arg(1.3512071919596578 * c.k0_drift);
            P.arg((-1.7024143839193155) * c.k0_drift);
            P.arg(a); P.arg(ca); P.arg(sa); P.arg(t2); P.arg(ia);
            P.arg(b); P.arg(cb); P.arg(sb); P.arg(tb); P.arg(ib);
            return true; }
        case 8: {
            static const double d[4] = {3.922568052387799819591407413100e-01,
                                        5.100434119184584780271052295575e-01,
                                        -4.710533854097565531482416645304e-01,
                                        6.875316825251809316199569366290e-02};
            static const double k[4] = {7.845136104775599639182814826199e-01,
                                        2.355732133593569921359289764951e-01,
                                        -1.177679984178870098432412305556e+00,
                                        1.315186320683906284756403692882e+00};
            P.op(OP_YOSHIDA6_BEND);
            P.arg(dl); P.arg(1 / h);
            for (int i = 0; i < 4; i++) P.arg(k[i] * c.k0_drift);
            for (int i = 0; i < 4; i++) {
                double s = d[i] * dl, cc, ss, tt, ii;
                polar_consts(h, s, &cc, &ss, &tt, &ii);
                P.arg(s); P.arg(cc); P.arg(ss); P.arg(tt); P.arg(ii);
            }
            return true; }
        default:
            err = "drift model " + std::to_string(c.drift_model) + " is not supported";
            return false;
        }
    }

    // kick descriptor (track_magnet_kick.h:24-144) appended at the current end
    size_t kick_desc(const MagnetArgs &m, const ModelCfg &c, double core_length,
                     double factor) {
        size_t base = prog.size();
        prog.resize(base + KD_ARR, 0.0);
        auto at = [&](int i) -> double & { return prog[base + i]; };
        at(KD_ORDER) = ibits(m.order);
        at(KD_ORDER_REL) = ibits(m.order_rel);
        at(KD_LENGTH) = core_length;
        at(KD_FACTOR) = factor;
        at(KD_FACTOR_REL) = factor * m.rel_ref_strength;
        at(KD_ROT_FRAME) = ibits(c.kick_rot_frame);
        at(KD_H) = c.h_kick;
        at(KD_HXL) = m.hxl;
        at(KD_K0HC) = c.k0_h_correction;
        at(KD_K1HC) = c.k1_h_correction;
        double htot = c.h_kick;
        if (core_length != 0) htot += m.hxl / core_length;
        at(KD_HTOT) = htot;
        double k0l = 0;
        if (m.order >= 0) k0l = m.knl[0] * factor;
        if (m.order_rel >= 0) k0l += m.knl_rel[0] * factor * m.rel_ref_strength;
        double k1l = 0;
        if (m.order >= 1) k1l = m.knl[1] * factor;
        if (m.order_rel >= 1) k1l += m.knl_rel[1] * factor * m.rel_ref_strength;
        at(KD_K0L) = k0l;
        at(KD_K1L) = k1l;
        const double km[4] = {c.k0_kick, c.k1_kick, m.k2, m.k3};
        const double ks[4] = {m.k0s, m.k1s, m.k2s, m.k3s};
        for (int i = 0; i < 4; i++) {
            at(KD_MAIN + i) = km[i] * core_length;
            at(KD_MAINS + i) = ks[i] * core_length;
        }
        auto chain = [&](int64_t order, double inv_f) {
            // f[order] = inv_f, f[k-1] = f[k]*k; stored f[order] .. f[0]
            std::vector<double> f(order + 1);
            f[order] = inv_f;
            for (int64_t k = order; k > 0; k--) f[k - 1] = f[k] * k;
            for (int64_t k = order; k >= 0; k--) prog.push_back(f[k]);
        };
        chain(m.order, m.inv_factorial_order);
        for (int64_t i = 0; i <= m.order; i++) prog.push_back(m.knl[i]);
        for (int64_t i = 0; i <= m.order; i++) prog.push_back(m.ksl[i]);
        chain(m.order_rel, m.inv_factorial_order_rel);
        for (int64_t i = 0; i <= m.order_rel; i++) prog.push_back(m.knl_rel[i]);
        for (int64_t i = 0; i <= m.order_rel; i++) prog.push_back(m.ksl_rel[i]);
        chain(3, 1. / (3 * 2));
        // fast blocks: a[k] = (knl[k]*factor)*f[k] for chi == 1
        auto fast = [&](int slot, int64_t order, double inv_f, const double *knl,
                        const double *ksl, double fac) {
            std::vector<double> f(order + 1), a(order + 1), b(order + 1);
            f[order] = inv_f;
            for (int64_t k = order; k > 0; k--) f[k - 1] = f[k] * k;
            int64_t top = -1;
            for (int64_t k = 0; k <= order; k++) {
                a[k] = (knl[k] * fac) * f[k];
                b[k] = (ksl[k] * fac) * f[k];
                if (a[k] != 0.0 || b[k] != 0.0) top = k;
            }
            prog[base + KD_FASTOFF + slot] = ibits((int64_t)prog.size());
            prog.push_back(ibits(top));
            for (int64_t k = 0; k <= top; k++) prog.push_back(a[k]);
            for (int64_t k = 0; k <= top; k++) prog.push_back(b[k]);
        };
        fast(0, m.order, m.inv_factorial_order, m.knl, m.ksl, factor);
        fast(1, m.order_rel, m.inv_factorial_order_rel, m.knl_rel, m.ksl_rel,
             factor * m.rel_ref_strength);
        const double kmL[4] = {km[0] * core_length, km[1] * core_length,
                               km[2] * core_length, km[3] * core_length};
        const double ksL[4] = {ks[0] * core_length, ks[1] * core_length,
                               ks[2] * core_length, ks[3] * core_length};
        fast(2, 3, 1. / (3 * 2), kmL, ksL, 1.0);
        return base;
    }

    // highest index with a non-zero fast coefficient of one pass (-1: none)
    static int64_t pass_top(int64_t order, double inv_f, const double *knl, const double *ksl,
                            double fac) {
        std::vector<double> f(order + 1);
        f[order] = inv_f;
        for (int64_t k = order; k > 0; k--) f[k - 1] = f[k] * k;
        int64_t top = -1;
        for (int64_t k = 0; k <= order; k++)
            if ((knl[k] * fac) * f[k] != 0.0 || (ksl[k] * fac) * f[k] != 0.0) top = k;
        return top;
    }

    // linear edge (track_magnet_edge.h:28-80 model 0; ax/ay for -1, 3)
    bool edge(const MagnetArgs &m, bool is_exit, int64_t model, double angle,
              double fdown, double fint, double hgap, double factor_edge,
              double fbe) {
        if (!(model == 0 || model == -1 || model == 3)) {
            err = "edge model " + std::to_string(model) + " is not supported";
            return false;
        }
        Prog P(prog);
        const double ksol = 0.0;
        if (is_exit) {
            P.op(OP_EDGE_CLEAR_AXAY);
        } else {
            P.op(OP_EDGE_AXAY); P.arg(-0.5 * ksol); P.arg(0.5 * ksol);
        }
        if (model == 0) {
            double k0 = 0;
            k0 += m.k0;                                   // knorm[0], k_order = 3
            if (std::fabs(m.length) > 1e-10 && m.order > -1)
                k0 += factor_edge * m.knl[0] / m.length;
            if (std::fabs(m.length) > 1e-10 && m.order_rel > -1)
                k0 += (factor_edge * m.rel_ref_strength) * m.knl_rel[0] / m.length;
            double const corr = 2.0 * k0 * hgap * fint;
            double r21 = k0 * xt_tan(angle);
            double const e1_v = angle + fdown;
            double const sin_e1_v = xt_sin(e1_v);
            double const temp = corr / xt_cos(e1_v) * (1.0 + sin_e1_v * sin_e1_v);
            double r43 = -k0 * xt_tan(e1_v - temp);
            r21 = r21 * fbe;
            r43 = r43 * fbe;
            P.op(OP_EDGE_LINEAR); P.arg(r21); P.arg(r43);
        }
        return true;
    }

    bool magnet(int type, const double *p, const SliceCfg &sl = SliceCfg{}) {
        MagnetArgs m{};
        m.order = (int64_t)p[XTM_M_ORDER];
        const int64_t nrel = (int64_t)p[XTM_M_NREL];
        if (m.order < 0 || nrel < 1) { err = "invalid multipole order"; return false; }
        m.knl = p + XTM_M_ARRAYS;
        m.ksl = m.knl + (m.order + 1);
        m.knl_rel = m.ksl + (m.order + 1);
        m.ksl_rel = m.knl_rel + nrel;
        m.length = p[XTM_M_LENGTH];
        m.inv_factorial_order = p[XTM_M_INV_FACTORIAL_ORDER];
        m.order_rel = nrel - 1;
        m.inv_factorial_order_rel = one_over_factorial(nrel - 1);
        m.num_multipole_kicks = (int64_t)p[XTM_M_NUM_MULTIPOLE_KICKS];
        m.model = (int64_t)p[XTM_M_MODEL];
        m.integrator = (int64_t)p[XTM_M_INTEGRATOR];
        const bool skew = p[XTM_M_MAIN_IS_SKEW] != 0;
        if (p[XTM_M_RADIATION_FLAG] != 0) { err = "radiation is not supported"; return false; }
        bool static_thick = true;
        switch (type) {
        case XTM_MULTIPOLE: {
            int64_t mo = (int64_t)p[XTM_M_MAIN_ORDER];
            if (mo < 0 || mo > m.order) { err = "main_order out of range"; return false; }
            m.rel_ref_strength = skew ? m.ksl[mo] : m.knl[mo];
            if (p[XTM_M_ISTHICK] <= 0) m.model = -1;
            m.default_model = 6; m.default_integrator = 3;
            m.hxl = p[XTM_M_HXL];
            static_thick = false;
            break; }
        case XTM_QUADRUPOLE:
            m.rel_ref_strength = m.length * (skew ? p[XTM_M_K1S] : p[XTM_M_K1]);
            m.default_model = 4; m.default_integrator = 3;
            m.k1 = p[XTM_M_K1]; m.k1s = p[XTM_M_K1S];
            m.edge_entry_active = (int64_t)p[XTM_M_EDGE_ENTRY_ACTIVE];
            m.edge_exit_active = (int64_t)p[XTM_M_EDGE_EXIT_ACTIVE];
            m.edge_entry_model = 1; m.edge_exit_model = 1;
            break;
        case XTM_BEND:
            m.rel_ref_strength = p[XTM_M_K0] * m.length;
            m.default_model = 3; m.default_integrator = 2;
            m.h = p[XTM_M_H]; m.k0 = p[XTM_M_K0]; m.k1 = p[XTM_M_K1]; m.k2 = p[XTM_M_K2];
            m.edge_entry_active = (int64_t)p[XTM_M_EDGE_ENTRY_ACTIVE];
            m.edge_exit_active = (int64_t)p[XTM_M_EDGE_EXIT_ACTIVE];
            m.edge_entry_model = (int64_t)p[XTM_M_EDGE_ENTRY_MODEL];
            m.edge_exit_model = (int64_t)p[XTM_M_EDGE_EXIT_MODEL];
            m.edge_entry_angle = p[XTM_M_EDGE_ENTRY_ANGLE];
            m.edge_exit_angle = p[XTM_M_EDGE_EXIT_ANGLE];
            m.edge_entry_fdown = p[XTM_M_EDGE_ENTRY_ANGLE_FDOWN];
            m.edge_exit_fdown = p[XTM_M_EDGE_EXIT_ANGLE_FDOWN];
            m.edge_entry_fint = p[XTM_M_EDGE_ENTRY_FINT];
            m.edge_exit_fint = p[XTM_M_EDGE_EXIT_FINT];
            m.edge_entry_hgap = p[XTM_M_EDGE_ENTRY_HGAP];
            m.edge_exit_hgap = p[XTM_M_EDGE_EXIT_HGAP];
            break;
        case XTM_SEXTUPOLE:
            m.rel_ref_strength = m.length * (skew ? p[XTM_M_K2S] : p[XTM_M_K2]);
            m.default_model = 6; m.default_integrator = 3;
            m.k2 = p[XTM_M_K2]; m.k2s = p[XTM_M_K2S];
            m.edge_entry_active = (int64_t)p[XTM_M_EDGE_ENTRY_ACTIVE];
            m.edge_exit_active = (int64_t)p[XTM_M_EDGE_EXIT_ACTIVE];
            m.edge_entry_model = 1; m.edge_exit_model = 1;
            break;
        case XTM_OCTUPOLE:
            m.rel_ref_strength = m.length * (skew ? p[XTM_M_K3S] : p[XTM_M_K3]);
            m.default_model = 6; m.default_integrator = 3;
            m.k3 = p[XTM_M_K3]; m.k3s = p[XTM_M_K3S];
            m.edge_entry_active = (int64_t)p[XTM_M_EDGE_ENTRY_ACTIVE];
            m.edge_exit_active = (int64_t)p[XTM_M_EDGE_EXIT_ACTIVE];
            m.edge_entry_model = 1; m.edge_exit_model = 1;
            break;
        default:
            err = "not a magnet";
            return false;
        }
        // slices (elements_src/thin_slice_*.h, thick_slice_*.h, the edge
        // slices): thin = kick only, edges off; thick = body only; entry /
        // exit = that edge only (edge slices pass weight 0, unused)
        bool body_active = true;
        if (sl.kind == XTM_SLICE_THIN) {
            m.model = -1;
            m.edge_entry_active = 0; m.edge_exit_active = 0;
        } else if (sl.kind == XTM_SLICE_THICK) {
            m.edge_entry_active = 0; m.edge_exit_active = 0;
        } else if (sl.kind == XTM_SLICE_ENTRY) {
            body_active = false; m.edge_exit_active = 0;
        } else if (sl.kind == XTM_SLICE_EXIT) {
            body_active = false; m.edge_entry_active = 0;
        }
        // forward tracking or backtracking (track_magnet.h:413-443):
        // backtracking negates the body length, the body factor and hxl,
        // swaps the entry and exit edges and flips the linear edge kicks
        const double weight = sl.weight;
        const double factor_knl_ksl = 1.0;
        const double core_length = bt ? -m.length * weight : m.length * weight;
        const double factor_body = bt ? -factor_knl_ksl * weight : factor_knl_ksl * weight;
        const double factor_edge = factor_knl_ksl;
        const double fbe = bt ? -1. : 1.;
        if (bt) {
            m.hxl = -m.hxl;
            std::swap(m.edge_entry_active, m.edge_exit_active);
            std::swap(m.edge_entry_model, m.edge_exit_model);
            std::swap(m.edge_entry_angle, m.edge_exit_angle);
            std::swap(m.edge_entry_fdown, m.edge_exit_fdown);
            std::swap(m.edge_entry_fint, m.edge_exit_fint);
            std::swap(m.edge_entry_hgap, m.edge_exit_hgap);
        }

        if (m.edge_entry_active &&
            !edge(m, false, m.edge_entry_model, m.edge_entry_angle, m.edge_entry_fdown,
                  m.edge_entry_fint, m.edge_entry_hgap, factor_edge, fbe))
            return false;

        if (!body_active) {
            if (m.edge_exit_active &&
                !edge(m, true, m.edge_exit_model, m.edge_exit_angle, m.edge_exit_fdown,
                      m.edge_exit_fint, m.edge_exit_hgap, factor_edge, fbe))
                return false;
            return true;
        }
        int64_t integrator = m.integrator, model = m.model, nmk = m.num_multipole_kicks;
        if (integrator == 0) integrator = m.default_integrator;
        if (model == 0) model = m.default_model;
        if (model == -1) { integrator = 3; nmk = 1; }
        // the number of kicks follows the weight (track_magnet.h:529-532)
        if (weight != 1.0 && nmk > 0) nmk = (int64_t)std::ceil(nmk * weight);
        if (nmk == 0) {
            bool inactive = (m.h == 0 && m.k0 == 0 && m.k1 == 0 && m.k2 == 0 && m.k3 == 0 &&
                             m.k0s == 0 && m.k1s == 0 && m.k2s == 0 && m.k3s == 0);
            for (int64_t i = m.order; inactive && i >= 0; i--)
                if (m.knl[i] != 0 || m.ksl[i] != 0) inactive = false;
            if (!inactive) {
                if (std::fabs(m.h) < 1e-8) {
                    nmk = 1;
                } else {
                    double b_circum = 2 * 3.14159 / std::fabs(m.h);
                    nmk = (int64_t)(std::fabs(core_length) / b_circum / 0.5e-3);
                    if (nmk < 1) nmk = 1;
                }
            }
        }
        ModelCfg c{};
        if (!configure(model, m.k0, m.k1, m.h, c)) {
            err = "magnet model " + std::to_string(model) + " is not supported";
            return false;
        }
        // body (track_magnet.h:177-274)
        const double length = core_length;
        if (nmk == 0 && c.k0_kick == 0 && c.k1_kick == 0 && c.h_kick == 0) {
            if (!drift(length, c)) return false;
        } else {
            // reserve the kick ops' descriptor offset: emitted after OP_END
            std::vector<std::pair<size_t, double>> kicks;   // (position of desc arg, weight)
            // element-uniform parts of the kick (track_magnet_kick.h:24-144)
            const double km[4] = {c.k0_kick * core_length, c.k1_kick * core_length,
                                  m.k2 * core_length, m.k3 * core_length};
            const double ks[4] = {m.k0s * core_length, m.k1s * core_length,
                                  m.k2s * core_length, m.k3s * core_length};
            const int64_t top1 = pass_top(m.order, m.inv_factorial_order, m.knl, m.ksl,
                                          factor_body);
            const int64_t top2 = pass_top(m.order_rel, m.inv_factorial_order_rel, m.knl_rel,
                                          m.ksl_rel, factor_body * m.rel_ref_strength);
            const int64_t top3 = pass_top(3, 1. / (3 * 2), km, ks, 1.0);
            double htot = c.h_kick;
            if (core_length != 0) htot += m.hxl / core_length;
            double k0l = 0;
            if (m.order >= 0) k0l = m.knl[0] * factor_body;
            if (m.order_rel >= 0) k0l += m.knl_rel[0] * factor_body * m.rel_ref_strength;
            double k1l = 0;
            if (m.order >= 1) k1l = m.knl[1] * factor_body;
            if (m.order_rel >= 1) k1l += m.knl_rel[1] * factor_body * m.rel_ref_strength;
            const double k0c = c.k0_h_correction * core_length + k0l;
            const double k1c = c.k1_h_correction * core_length + k1l;
            auto kick = [&](double w) {
                const double hl = c.kick_rot_frame ? c.h_kick * core_length * w + m.hxl * w : 0.0;
                const double g0 = ((-k0c) * w) * htot;
                const double h1 = (htot * k1c) * w;
                int64_t mask = 0;
                if (top1 >= 0) mask |= 1;
                if (top2 >= 0) mask |= 2;
                if (top3 >= 0) mask |= 4;
                if (c.kick_rot_frame && hl != 0.0) mask |= 8;
                if (g0 != 0.0) mask |= 16;
                if (h1 != 0.0) mask |= 32;
                Prog P(prog);
                P.op(OP_KICK); P.arg(w);
                kicks.push_back({prog.size(), w});
                P.arg(0.0);  // descriptor offset, patched below
                P.arg(ibits(mask)); P.arg(hl); P.arg(g0); P.arg(h1);
            };
            if (integrator == 1) {
                const double kick_weight = 1. / nmk;
                double edge_drift_weight = 0.5;
                double inside_drift_weight = 0;
                if (nmk > 1) {
                    edge_drift_weight = 1. / (2 * (1 + nmk));
                    inside_drift_weight = (((double)nmk) / ((double)(nmk * nmk) - 1));
                }
                if (!drift(edge_drift_weight * length, c)) return false;
                for (int64_t i = 0; i < nmk - 1; i++) {
                    kick(kick_weight);
                    if (!drift(inside_drift_weight * length, c)) return false;
                }
                kick(kick_weight);
                if (!drift(edge_drift_weight * length, c)) return false;
            } else if (integrator == 3) {
                const double kick_weight = 1. / nmk;
                const double drift_weight = kick_weight;
                for (int64_t i = 0; i < nmk; i++) {
                    if (!drift(0.5 * drift_weight * length, c)) return false;
                    kick(kick_weight);
                    if (!drift(0.5 * drift_weight * length, c)) return false;
                }
            } else if (integrator == 2) {
                const int64_t num_slices = (nmk / 7 + (nmk % 7 != 0));
                const double slice_length = length / (num_slices);
                const double kick_weight = 1. / num_slices;
                static const double dy[4] = {3.922568052387799819591407413100e-01,
                                             5.100434119184584780271052295575e-01,
                                             -4.710533854097565531482416645304e-01,
                                             6.875316825251809316199569366290e-02};
                static const double ky[4] = {7.845136104775599639182814826199e-01,
                                             2.355732133593569921359289764951e-01,
                                             -1.177679984178870098432412305556e+00,
                                             1.315186320683906284756403692882e+00};
                static const int di[8] = {0, 1, 2, 3, 3, 2, 1, 0};
                static const int ki[7] = {0, 1, 2, 3, 2, 1, 0};
                for (int64_t s = 0; s < num_slices; s++) {
                    for (int j = 0; j < 7; j++) {
                        if (!drift(slice_length * dy[di[j]], c)) return false;
                        kick(kick_weight * ky[ki[j]]);
                    }
                    if (!drift(slice_length * dy[di[7]], c)) return false;
                }
            } else {
                err = "integrator " + std::to_string(integrator) + " is not supported";
                return false;
            }
            // the descriptor goes after the element's OP_END; the caller
            // appends OP_END, so stash kicks for patching
            pending_kicks = kicks;
            pending_desc = true;
            pending_m = m;
            pending_c = c;
            pending_core_length = core_length;
            pending_factor = factor_body;
        }
        if (m.edge_exit_active &&
            !edge(m, true, m.edge_exit_model, m.edge_exit_angle, m.edge_exit_fdown,
                  m.edge_exit_fint, m.edge_exit_hgap, factor_edge, fbe))
            return false;
        if (static_thick) global_aper = true;
        return true;
    }

    bool cavity(const double *p) {
        double length = p[XTM_C_LENGTH];
        double voltage = p[XTM_C_VOLTAGE];
        const double frequency = p[XTM_C_FREQUENCY];
        const double harmonic = p[XTM_C_HARMONIC];
        const double lag = p[XTM_C_LAG];
        const double phase = p[XTM_C_PHASE];
        const int64_t absolute_time = (int64_t)p[XTM_C_ABSOLUTE_TIME];
        int64_t num_kicks = (int64_t)p[XTM_C_NUM_KICKS];
        int64_t model = (int64_t)p[XTM_C_MODEL];
        int64_t integrator = (int64_t)p[XTM_C_INTEGRATOR];
        const double weight = 1.;
        // backtracking (track_rf.h:364-373): negative body length and voltage
        const double body_length = bt ? -length : length;
        if (bt) voltage = -voltage;
        if (integrator == 0) integrator = 3;
        if (model == 0) model = 6;
        if (model == -1) { integrator = 3; num_kicks = 1; }
        if (num_kicks == 0) num_kicks = 1;
        ModelCfg c{};
        if (!configure(model, 0, 0, 0, c)) {
            err = "cavity model " + std::to_string(model) + " is not supported";
            return false;
        }
        const double blen = body_length * weight;
        const double vw = voltage * weight;
        auto rfkick = [&](double w) {
            Prog P(prog);
            P.op(OP_RF_KICK);
            P.arg(vw * (w)); P.arg(frequency); P.arg(harmonic); P.arg(lag); P.arg(phase);
            P.arg(ibits(absolute_time));
        };
        if (integrator == 1) {
            const double kick_weight = 1. / num_kicks;
            double edge_drift_weight = 0.5;
            double inside_drift_weight = 0;
            if (num_kicks > 1) {
                edge_drift_weight = 1. / (2 * (1 + num_kicks));
                inside_drift_weight = (((double)num_kicks) / ((double)(num_kicks * num_kicks) - 1));
            }
            if (!drift(edge_drift_weight * blen, c)) return false;
            for (int64_t i = 0; i < num_kicks - 1; i++) {
                rfkick(kick_weight);
                if (!drift(inside_drift_weight * blen, c)) return false;
            }
            rfkick(kick_weight);
            if (!drift(edge_drift_weight * blen, c)) return false;
        } else if (integrator == 3) {
            const double kick_weight = 1. / num_kicks;
            const double drift_weight = kick_weight;
            for (int64_t i = 0; i < num_kicks; i++) {
                if (!drift(0.5 * drift_weight * blen, c)) return false;
                rfkick(kick_weight);
                if (!drift(0.5 * drift_weight * blen, c)) return false;
            }
        } else if (integrator == 2) {
            const int64_t num_slices = (num_kicks / 7 + (num_kicks % 7 != 0));
            const double slice_length = blen / (num_slices);
            const double kick_weight = 1. / num_slices;
            static const double dy[4] = {3.922568052387799819591407413100e-01,
                                         5.100434119184584780271052295575e-01,
                                         -4.710533854097565531482416645304e-01,
                                         6.875316825251809316199569366290e-02};
            static const double ky[4] = {7.845136104775599639182814826199e-01,
                                         2.355732133593569921359289764951e-01,
                                         -1.177679984178870098432412305556e+00,
                                         1.315186320683906284756403692882e+00};
            static const int di[8] = {0, 1, 2, 3, 3, 2, 1, 0};
            static const int ki[7] = {0, 1, 2, 3, 2, 1, 0};
            for (int64_t s = 0; s < num_slices; s++) {
                for (int j = 0; j < 7; j++) {
                    if (!drift(slice_length * dy[di[j]], c)) return false;
                    rfkick(kick_weight * ky[ki[j]]);
                }
                if (!drift(slice_length * dy[di[7]], c)) return false;
            }
        } else {
            err = "integrator " + std::to_string(integrator) + " is not supported";
            return false;
        }
        global_aper = true;
        return true;
    }

    // rfmultipole.h:11-47 -> track_rf.h:317-453 with model -1: one kick of
    // weight 1 between two zero-length drifts (skipped); backtracking
    // negates the voltage and the multipole factor
    bool rf_multipole(const double *p) {
        const int64_t order = (int64_t)p[XTM_RFM_ORDER];
        if (order < 0) { err = "RF multipole order < 0"; return false; }
        const double factor = bt ? -1.0 : 1.0;
        const double voltage = bt ? -p[XTM_RFM_VOLTAGE] : p[XTM_RFM_VOLTAGE];
        const double *knl = p + XTM_RFM_ARRAYS, *ksl = knl + (order + 1);
        const double *pn = ksl + (order + 1), *ps = pn + (order + 1);
        const double *phase_n = ps + (order + 1), *phase_s = phase_n + (order + 1);
        std::vector<double> d(RFD_TERMS + RFD_STRIDE * (order + 1));
        d[RFD_ORDER] = ibits(order);
        double factorial = 1.0;
        for (int64_t kk = 0; kk <= order; kk++) {
            if (kk > 0) factorial *= kk;
            double *t = &d[RFD_TERMS + RFD_STRIDE * kk];
            t[0] = kDeg2Rad * pn[kk];
            t[1] = phase_n[kk];
            t[2] = kDeg2Rad * ps[kk];
            t[3] = phase_s[kk];
            t[4] = factor * knl[kk] / factorial;
            t[5] = factor * ksl[kk] / factorial;
        }
        Prog P(prog);
        P.op(OP_RF_MULT);
        P.arg(voltage * 1.0); P.arg(p[XTM_RFM_FREQUENCY]); P.arg(p[XTM_RFM_LAG]);
        P.arg(p[XTM_RFM_PHASE]);
        pending_rf.push_back({prog.size(), std::move(d)});
        P.arg(0.0);   // descriptor offset, patched after OP_END
        return true;
    }

    // slice elements (elements_src/{thin,thick,drift}_slice_*.h and the
    // edge slices); the global aperture follows thick and drift slices,
    // whose classes declare isthick = True
    bool slice(const double *p) {
        const double weight = p[XTM_SL_WEIGHT];
        const int kind = (int)p[XTM_SL_KIND];
        const int parent = (int)p[XTM_SL_PARENT_TYPE];
        const double *q = p + XTM_SL_PARENT;
        if (kind == XTM_SLICE_DRIFT) {
            // Drift_single_particle (expanded) over weight x parent length
            double plen;
            if (parent == XTM_DRIFT) plen = q[XTM_D_LENGTH];
            else if (parent >= XTM_MULTIPOLE && parent <= XTM_OCTUPOLE) plen = q[XTM_M_LENGTH];
            else { err = "drift slice of an unsupported parent"; return false; }
            Prog P(prog);
            P.op(OP_DRIFT_EXP);
            P.arg(bt ? -weight * plen : weight * plen);
            global_aper = true;
            return true;
        }
        if (!(parent >= XTM_MULTIPOLE && parent <= XTM_OCTUPOLE) ||
            kind < XTM_SLICE_THIN || kind > XTM_SLICE_EXIT ||
            ((kind == XTM_SLICE_ENTRY || kind == XTM_SLICE_EXIT) && parent == XTM_MULTIPOLE)) {
            err = "unsupported slice";
            return false;
        }
        SliceCfg sl;
        sl.weight = weight;
        sl.kind = kind;
        if (!magnet(parent, q, sl)) return false;
        global_aper = kind == XTM_SLICE_THICK;
        return true;
    }

    // ---- frame transformations (track_misalignments.h) ----------------------
    // XY_SHIFT(dx, dy): add -dx, -dy; skipped only where both additions
    // are of -0.0, which leave every value unchanged
    void xy_shift(double dx, double dy) {
        const double ax = -dx, ay = -dy;
        if (std::signbit(ax) && ax == 0.0 && std::signbit(ay) && ay == 0.0) return;
        Prog P(prog);
        P.op(OP_XY_SHIFT); P.arg(ax); P.arg(ay);
    }
    void s_shift(double ds) {            // IF_NONZERO(DS, ...)
        if (ds == 0.0) return;
        Prog P(prog);
        P.op(OP_S_SHIFT); P.arg(ds);
    }
    void y_rotate(double theta) {        // YRotation(sin, cos, tan)
        if (theta == 0.0) return;
        Prog P(prog);
        P.op(OP_Y_ROT); P.arg(xt_sin(theta)); P.arg(xt_cos(theta)); P.arg(xt_tan(theta));
    }
    void x_rotate(double phi) {          // XRotation(-sin, cos, -tan)
        if (phi == 0.0) return;
        Prog P(prog);
        P.op(OP_X_ROT); P.arg(-xt_sin(phi)); P.arg(xt_cos(phi)); P.arg(-xt_tan(phi));
    }
    void s_rotate(double psi) {          // SRotation(sin, cos)
        if (psi == 0.0) return;
        Prog P(prog);
        P.op(OP_S_ROT); P.arg(xt_sin(psi)); P.arg(xt_cos(psi));
    }

    struct Misalign {
        double dx, dy, ds, theta, phi, psi_no_frame, anchor, length, angle, h, psi_with_frame;
    };

    // track_misalignments.h:38-74
    void entry_straight(const Misalign &g, bool backtrack) {
        const double mis_x = g.dx - g.anchor * xt_cos(g.phi) * xt_sin(g.theta);
        const double mis_y = g.dy - g.anchor * xt_sin(g.phi);
        const double mis_s = g.ds - g.anchor * (xt_cos(g.phi) * xt_cos(g.theta) - 1);
        if (!backtrack) {
            xy_shift(mis_x, mis_y);
            s_shift(mis_s);
            y_rotate(g.theta);
            x_rotate(g.phi);
            s_rotate(g.psi_no_frame);
            s_rotate(g.psi_with_frame);
        } else {
            s_rotate(-g.psi_with_frame);
            s_rotate(-g.psi_no_frame);
            x_rotate(-g.phi);
            y_rotate(-g.theta);
            s_shift(-mis_s);
            xy_shift(-mis_x, -mis_y);
        }
    }

    // track_misalignments.h:77-112
    void exit_straight(const Misalign &g, bool backtrack) {
        const double neg_part_length = g.anchor - g.length;
        const double mis_x = neg_part_length * xt_cos(g.phi) * xt_sin(g.theta) - g.dx;
        const double mis_y = neg_part_length * xt_sin(g.phi) - g.dy;
        const double mis_s = neg_part_length * (xt_cos(g.phi) * xt_cos(g.theta) - 1) - g.ds;
        if (!backtrack) {
            s_rotate(-g.psi_with_frame);
            s_rotate(-g.psi_no_frame);
            x_rotate(-g.phi);
            y_rotate(-g.theta);
            s_shift(mis_s);
            xy_shift(mis_x, mis_y);
        } else {
            xy_shift(-mis_x, -mis_y);
            s_shift(-mis_s);
            y_rotate(g.theta);
            x_rotate(g.phi);
            s_rotate(g.psi_no_frame);
            s_rotate(g.psi_with_frame);
        }
    }

    typedef double M4[4][4];
    // track_misalignments.h:389-398
    static void mat_mul(const M4 a, const M4 b, M4 r) {
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++)
                r[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j] +
                          a[i][3] * b[3][j];
    }
    // track_misalignments.h:401-429
    static void mat_rigid_inverse(const M4 m, M4 inv) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) inv[i][j] = m[j][i];
        for (int i = 0; i < 3; i++)
            inv[i][3] = -m[0][i] * m[0][3] - m[1][i] * m[1][3] - m[2][i] * m[2][3];
        inv[3][0] = 0; inv[3][1] = 0; inv[3][2] = 0; inv[3][3] = 1;
    }
    // the misalignment matrix (track_misalignments.h:155-174)
    static void misalignment_matrix(const Misalign &g, M4 m) {
        const double s_phi = xt_sin(g.phi), c_phi = xt_cos(g.phi);
        const double s_theta = xt_sin(g.theta), c_theta = xt_cos(g.theta);
        const double s_psi = xt_sin(g.psi_no_frame), c_psi = xt_cos(g.psi_no_frame);
        const M4 t = {
            {-s_phi * s_psi * s_theta + c_psi * c_theta, -c_psi * s_phi * s_theta - c_theta * s_psi,
             c_phi * s_theta, g.dx},
            {c_phi * s_psi, c_phi * c_psi, s_phi, g.dy},
            {-c_theta * s_phi * s_psi - c_psi * s_theta, -c_psi * c_theta * s_phi + s_psi * s_theta,
             c_phi * c_theta, g.ds},
            {0, 0, 0, 1}};
        std::memcpy(m, t, sizeof(M4));
    }
    // basic transformations of a rigid frame change (:215-222, :354-361)
    static void decompose(const M4 e, double *mx, double *my, double *ms, double *theta,
                          double *phi, double *psi) {
        *mx = e[0][3];
        *my = e[1][3];
        *ms = e[2][3];
        *theta = std::atan2(e[0][2], e[2][2]);
        *phi = std::atan2(e[1][2], std::sqrt(e[1][0] * e[1][0] + e[1][1] * e[1][1]));
        *psi = std::atan2(e[1][0], e[1][1]);
    }

    static bool straight_frame(const Misalign &g) {     // :131-137, :245-251
        return g.angle == 0.0 && (g.length != 0.0 || g.h == 0.0);
    }

    // track_misalignments.h:115-241
    void entry_curved(const Misalign &g, bool backtrack) {
        if (straight_frame(g)) { entry_straight(g, backtrack); return; }
        M4 mis;
        misalignment_matrix(g, mis);
        double h = g.h;
        if (g.length != 0.0) h = g.angle / g.length;
        const double part_angle = g.anchor * h;
        const double pwf = g.psi_with_frame;
        const double delta_x_first_part = (xt_cos(part_angle) - 1) * xt_cos(pwf) / h;
        const double delta_y_first_part = (xt_cos(part_angle) - 1) * xt_sin(pwf) / h;
        const double delta_s_first_part = xt_sin(part_angle) / h;
        const M4 first = {
            {(xt_cos(part_angle) - 1) * (xt_cos(pwf) * xt_cos(pwf)) + 1,
             (xt_cos(part_angle) - 1) * xt_cos(pwf) * xt_sin(pwf), -xt_cos(pwf) * xt_sin(part_angle),
             delta_x_first_part},
            {(xt_cos(part_angle) - 1) * xt_cos(pwf) * xt_sin(pwf),
             (xt_cos(part_angle) - 1) * (xt_sin(pwf) * xt_sin(pwf)) + 1,
             -xt_sin(part_angle) * xt_sin(pwf), delta_y_first_part},
            {xt_cos(pwf) * xt_sin(part_angle), xt_sin(part_angle) * xt_sin(pwf), xt_cos(part_angle),
             delta_s_first_part},
            {0, 0, 0, 1}};
        M4 inv_first, temp, e;
        mat_rigid_inverse(first, inv_first);
        mat_mul(first, mis, temp);
        mat_mul(temp, inv_first, e);
        double mis_x, mis_y, mis_s, rot_theta, rot_phi, rot_psi;
        decompose(e, &mis_x, &mis_y, &mis_s, &rot_theta, &rot_phi, &rot_psi);
        if (!backtrack) {
            xy_shift(mis_x, mis_y);
            s_shift(mis_s);
            y_rotate(rot_theta);
            x_rotate(rot_phi);
            s_rotate(rot_psi);
            s_rotate(pwf);
        } else {
            s_rotate(-pwf);
            s_rotate(-rot_psi);
            x_rotate(-rot_phi);
            y_rotate(-rot_theta);
            s_shift(-mis_s);
            xy_shift(-mis_x, -mis_y);
        }
    }

    // track_misalignments.h:244-386
    void exit_curved(const Misalign &g, bool backtrack) {
        if (straight_frame(g)) { exit_straight(g, backtrack); return; }
        M4 mis, inv_mis;
        misalignment_matrix(g, mis);
        mat_rigid_inverse(mis, inv_mis);
        double h = g.h;
        if (g.length != 0.0) h = g.angle / g.length;
        const double part_angle = g.angle - h * g.anchor;
        const double s_part_angle = xt_sin(part_angle), c_part_angle = xt_cos(part_angle);
        const double s_tilt = xt_sin(g.psi_with_frame), c_tilt = xt_cos(g.psi_with_frame);
        const double delta_x_second_part = (c_part_angle - 1) * c_tilt / h;
        const double delta_y_second_part = (c_part_angle - 1) * s_tilt / h;
        const double delta_s_second_part = s_part_angle / h;
        const M4 second = {
            {(c_part_angle - 1) * (c_tilt * c_tilt) + 1, (c_part_angle - 1) * c_tilt * s_tilt,
             c_tilt * -s_part_angle, delta_x_second_part},
            {(c_part_angle - 1) * c_tilt * s_tilt, (c_part_angle - 1) * (s_tilt * s_tilt) + 1,
             -s_part_angle * s_tilt, delta_y_second_part},
            {c_tilt * s_part_angle, s_part_angle * s_tilt, c_part_angle, delta_s_second_part},
            {0, 0, 0, 1}};
        M4 inv_second, temp, realign;
        mat_rigid_inverse(second, inv_second);
        mat_mul(inv_second, inv_mis, temp);
        mat_mul(temp, second, realign);
        double mis_x, mis_y, mis_s, rot_theta, rot_phi, rot_psi;
        decompose(realign, &mis_x, &mis_y, &mis_s, &rot_theta, &rot_phi, &rot_psi);
        if (!backtrack) {
            s_rotate(-g.psi_with_frame);
            xy_shift(mis_x, mis_y);
            s_shift(mis_s);
            y_rotate(rot_theta);
            x_rotate(rot_phi);
            s_rotate(rot_psi);
        } else {
            s_rotate(-rot_psi);
            x_rotate(-rot_phi);
            y_rotate(-rot_theta);
            s_shift(-mis_s);
            xy_shift(-mis_x, -mis_y);
            s_rotate(g.psi_with_frame);
        }
    }

    // track_local_particle_with_transformations.h:88-168: the entry
    // transformation, the element in its own frame, the exit
    // transformation; when backtracking the exit's inverse comes first
    bool transformed(const double *p) {
        const int inner = (int)p[XTM_TR_INNER_TYPE];
        switch (inner) {
        case XTM_MULTIPOLE: case XTM_QUADRUPOLE: case XTM_BEND: case XTM_SEXTUPOLE:
        case XTM_OCTUPOLE: case XTM_CAVITY: case XTM_LIMITRECT: case XTM_LIMITELLIPSE:
        case XTM_DIPOLEEDGE: case XTM_RFMULTIPOLE: case XTM_SLICE: break;
        default:
            err = "frame transformations are not supported for element type " +
                  std::to_string(inner);
            return false;
        }
        Misalign g;
        g.dx = p[XTM_TR_SHIFT_X];
        g.dy = p[XTM_TR_SHIFT_Y];
        g.ds = p[XTM_TR_SHIFT_S];
        g.theta = p[XTM_TR_ROT_Y];
        g.phi = p[XTM_TR_ROT_X];
        g.psi_no_frame = p[XTM_TR_ROT_S_NO_FRAME];
        g.anchor = p[XTM_TR_ANCHOR];
        g.psi_with_frame = p[XTM_TR_ROT_S];
        g.length = p[XTM_TR_LENGTH];
        g.angle = p[XTM_TR_ANGLE];
        g.h = p[XTM_TR_H];
        const bool curved = p[XTM_TR_CURVED] != 0.0;
        if (!bt) {
            if (curved) entry_curved(g, false); else entry_straight(g, false);
        } else {
            if (curved) exit_curved(g, true); else exit_straight(g, true);
        }
        if (!body(inner, p + XTM_TR_INNER)) return false;
        if (!bt) {
            if (curved) exit_curved(g, false); else exit_straight(g, false);
        } else {
            if (curved) entry_curved(g, true); else entry_straight(g, true);
        }
        return true;
    }

    // ops of one element without its global aperture check and OP_END
    bool body(int type, const double *p) {
        Prog P(prog);
        bool ok = true;
        switch (type) {
        case XTM_MARKER: break;
        case XTM_DRIFT: {
            // drift.h:14-40 (backtracking: negative length)
            int64_t model = (int64_t)p[XTM_D_MODEL];
            if (model == 0) model = 1;
            const double dl = bt ? -p[XTM_D_LENGTH] : p[XTM_D_LENGTH];
            if (model == 1) { P.op(OP_DRIFT_EXP); P.arg(dl); }
            else if (model == 2) { P.op(OP_DRIFT_EXACT); P.arg(dl); }
            else { err = "drift model not supported"; ok = false; }
            global_aper = true;
            break; }
        case XTM_MULTIPOLE: case XTM_QUADRUPOLE: case XTM_BEND:
        case XTM_SEXTUPOLE: case XTM_OCTUPOLE:
            ok = magnet(type, p);
            break;
        case XTM_CAVITY: ok = cavity(p); break;
        case XTM_LIMITRECT:
            P.op(OP_LIMIT_RECT);
            for (int i = 0; i < 4; i++) P.arg(p[i]);
            break;
        case XTM_LIMITELLIPSE:
            P.op(OP_LIMIT_ELLIPSE);
            for (int i = 0; i < 3; i++) P.arg(p[i]);
            break;
        case XTM_DIPOLEEDGE: {
            // dipoleedge.h:17-50: linear model r21, r43 scaled by
            // (1 + delta_taper), negated when backtracking
            const int64_t model = (int64_t)p[XTM_DE_MODEL];
            if (model == 0) {
                double r21 = p[XTM_DE_R21] * (1 + p[XTM_DE_DELTA_TAPER]);
                double r43 = p[XTM_DE_R43] * (1 + p[XTM_DE_DELTA_TAPER]);
                if (bt) { r21 = -r21; r43 = -r43; }
                P.op(OP_EDGE_LINEAR); P.arg(r21); P.arg(r43);
            } else if (model != -1) {
                err = "dipole edge model " + std::to_string(model) + " is not supported";
                ok = false;
            }
            break; }
        case XTM_TRANSFORMED: ok = transformed(p); break;
        case XTM_RFMULTIPOLE: ok = rf_multipole(p); break;
        case XTM_SLICE: ok = slice(p); break;
        default:
            err = "unknown element type " + std::to_string(type);
            ok = false;
        }
        return ok;
    }

    // element -> ops; returns start offset or -1.  The global aperture
    // check of statically thick classes (tracker.py:681-689) follows the
    // element and its frame transformations.
    int64_t element(int type, const double *p) {
        const int64_t start = (int64_t)prog.size();
        Prog P(prog);
        pending_desc = false;
        pending_kicks.clear();
        pending_rf.clear();
        global_aper = false;
        if (!body(type, p)) return -1;
        if (global_aper) P.op(OP_GLOBAL_APER);
        P.op(OP_END);
        if (pending_desc) {
            size_t desc = kick_desc(pending_m, pending_c, pending_core_length, pending_factor);
            for (auto &k : pending_kicks) prog[k.first] = ibits((int64_t)desc);
        }
        for (auto &r : pending_rf) {
            prog[r.first] = ibits((int64_t)prog.size());
            prog.insert(prog.end(), r.second.begin(), r.second.end());
        }
        pending_rf.clear();
        return start;
    }
    std::vector<std::pair<size_t, std::vector<double>>> pending_rf;

    bool global_aper = false;
    bool pending_desc = false;
    std::vector<std::pair<size_t, double>> pending_kicks;
    MagnetArgs pending_m{};
    ModelCfg pending_c{};
    double pending_core_length = 0, pending_factor = 0;
};

size_t raw_size(int type, const double *p) {
    switch (type) {
    case XTM_MARKER: return 0;
    case XTM_DRIFT: return XTM_D_SIZE;
    case XTM_MULTIPOLE: case XTM_QUADRUPOLE: case XTM_BEND:
    case XTM_SEXTUPOLE: case XTM_OCTUPOLE: {
        int64_t order = (int64_t)p[XTM_M_ORDER], nrel = (int64_t)p[XTM_M_NREL];
        return XTM_M_ARRAYS + 2 * (order + 1) + 2 * nrel; }
    case XTM_CAVITY: return XTM_C_SIZE;
    case XTM_DIPOLEEDGE: return XTM_DE_SIZE;
    case XTM_LIMITRECT: return XTM_LR_SIZE;
    case XTM_LIMITELLIPSE: return XTM_LE_SIZE;
    case XTM_RFMULTIPOLE: return XTM_RFM_ARRAYS + 6 * ((int64_t)p[XTM_RFM_ORDER] + 1);
    case XTM_SLICE: {
        const int parent = (int)p[XTM_SL_PARENT_TYPE];
        if (parent == XTM_SLICE || parent == XTM_TRANSFORMED) return 0;
        return XTM_SL_PARENT + raw_size(parent, p + XTM_SL_PARENT); }
    case XTM_TRANSFORMED: {
        const int inner = (int)p[XTM_TR_INNER_TYPE];
        if (inner == XTM_TRANSFORMED) return 0;
        return XTM_TR_INNER + raw_size(inner, p + XTM_TR_INNER); }
    default: return 0;
    }
}

}  // namespace

int xt_build_program(const int32_t *types, const int64_t *offsets, const double *params,
                     int64_t n, std::vector<int64_t> &entry, std::vector<double> &prog,
                     std::string &err, std::vector<int64_t> *entry_bt) {
    entry.assign(n, 0);
    prog.clear();
    prog.reserve(1024);
    Builder B(prog, err);
    // (direction, raw record bytes) -> program offset
    std::map<std::string, int64_t> seen;
    for (int dir = 0; dir < (entry_bt ? 2 : 1); dir++) {
    std::vector<int64_t> &ent = dir == 0 ? entry : *entry_bt;
    if (dir == 1) ent.assign(n, 0);
    B.bt = dir == 1;
    for (int64_t i = 0; i < n; i++) {
        const int type = types[i];
        if (type < 0 || type >= XTM_NUM_TYPES) {
            err = "element " + std::to_string(i) + ": unknown type";
            return XTM_ERR_ARG;
        }
        const double *p = params + offsets[i];
        if (type >= XTM_MULTIPOLE && type <= XTM_OCTUPOLE) {
            if (p[XTM_M_ORDER] < 0 || p[XTM_M_NREL] < 1) {
                err = "element " + std::to_string(i) + ": invalid multipole arrays";
                return XTM_ERR_ARG;
            }
        }
        if (type == XTM_SLICE || (type == XTM_TRANSFORMED && p[XTM_TR_INNER_TYPE] == XTM_SLICE)) {
            const double *q = type == XTM_SLICE ? p : p + XTM_TR_INNER;
            const double pt = q[XTM_SL_PARENT_TYPE];
            const double *r = q + XTM_SL_PARENT;
            if (!(pt == XTM_DRIFT || (pt >= XTM_MULTIPOLE && pt <= XTM_OCTUPOLE)) ||
                (pt >= XTM_MULTIPOLE && (r[XTM_M_ORDER] < 0 || r[XTM_M_NREL] < 1))) {
                err = "element " + std::to_string(i) + ": invalid slice record";
                return XTM_ERR_ARG;
            }
        }
        if (type == XTM_RFMULTIPOLE && !(p[XTM_RFM_ORDER] >= 0)) {
            err = "element " + std::to_string(i) + ": invalid RF multipole order";
            return XTM_ERR_ARG;
        }
        if (type == XTM_TRANSFORMED) {
            const double it = p[XTM_TR_INNER_TYPE];
            const double *q = p + XTM_TR_INNER;
            if (!((it >= XTM_MULTIPOLE && it <= XTM_DIPOLEEDGE) || it == XTM_RFMULTIPOLE ||
                  it == XTM_SLICE) ||
                (it >= XTM_MULTIPOLE && it <= XTM_OCTUPOLE &&
                 (q[XTM_M_ORDER] < 0 || q[XTM_M_NREL] < 1)) ||
                (it == XTM_RFMULTIPOLE && !(q[XTM_RFM_ORDER] >= 0))) {
                err = "element " + std::to_string(i) + ": invalid transformed record";
                return XTM_ERR_ARG;
            }
        }
        const size_t sz = raw_size(type, p);
        std::string key(1, (char)dir);
        key.append(reinterpret_cast<const char *>(&type), sizeof(type));
        key.append(reinterpret_cast<const char *>(p), sz * sizeof(double));
        auto it = seen.find(key);
        if (it != seen.end()) { ent[i] = it->second; continue; }
        int64_t off = B.element(type, p);
        if (off < 0) {
            err = "element " + std::to_string(i) + (dir ? " (backtracking)" : "") + ": " + err;
            return XTM_ERR_UNSUPPORTED;
        }
        seen.emplace(std::move(key), off);
        ent[i] = off;
    }
    }
    if (prog.empty()) prog.push_back(ibits(OP_END));
    return XTM_OK;
}

int64_t xt_count_records(const std::vector<int64_t> &entry) {
    std::map<int64_t, int> m;
    for (auto e : entry) m[e] = 1;
    return (int64_t)m.size();
}
