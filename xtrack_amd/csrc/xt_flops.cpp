// xt_flops.cpp -- algorithmic work of the tracking kernel, counted by
// executing the kernel's own element maps (xt_ops.h) on the host with an
// operation-counting scalar for one particle and one turn.
//
// Counting rules: each executed +, -, *, / and sqrt on a particle-dependent
// value is one operation; negation is free; element-uniform work hoisted to
// the element program is not counted (it is not executed per particle);
// each sin/cos/sinh/cosh/exp evaluation is counted as one elementary-function
// call (the caller weights calls, 20 flops each by convention).
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/xtrack_amd.h"
#include "xt_flops.h"
#include "xt_math.h"

namespace {

struct Counts {
    double add = 0, mul = 0, div = 0, sqrt = 0, trans = 0;
};
thread_local Counts g;

struct FC {
    double v;
    FC() : v(0) {}
    FC(double d) : v(d) {}
};

inline FC operator+(FC a, FC b) { g.add++; return FC(a.v + b.v); }
inline FC operator-(FC a, FC b) { g.add++; return FC(a.v - b.v); }
inline FC operator*(FC a, FC b) { g.mul++; return FC(a.v * b.v); }
inline FC operator/(FC a, FC b) { g.div++; return FC(a.v / b.v); }
inline FC operator+(FC a, double b) { g.add++; return FC(a.v + b); }
inline FC operator-(FC a, double b) { g.add++; return FC(a.v - b); }
inline FC operator*(FC a, double b) { g.mul++; return FC(a.v * b); }
inline FC operator/(FC a, double b) { g.div++; return FC(a.v / b); }
inline FC operator+(double a, FC b) { g.add++; return FC(a + b.v); }
inline FC operator-(double a, FC b) { g.add++; return FC(a - b.v); }
inline FC operator*(double a, FC b) { g.mul++; return FC(a * b.v); }
inline FC operator/(double a, FC b) { g.div++; return FC(a / b.v); }
inline FC operator-(FC a) { return FC(-a.v); }
inline bool operator>(FC a, double b) { return a.v > b; }
inline bool operator<(FC a, double b) { return a.v < b; }
inline bool operator!=(FC a, double b) { return a.v != b; }

}  // namespace

// overloads picked up by the templates in xt_ops.h
namespace xt {
inline FC xsqrt(FC v) { g.sqrt++; return FC(std::sqrt(v.v)); }
inline void xsincos(FC v, FC *s, FC *c) {
    g.trans += 2;
    double a, b;
    xt_sincos(v.v, &a, &b);
    *s = FC(a);
    *c = FC(b);
}
inline void xsinhcosh(FC v, FC *s, FC *c) {
    g.trans += 2;
    double a, b;
    xt_sinhcosh(v.v, &a, &b);
    *s = FC(a);
    *c = FC(b);
}
inline FC xsin(FC v) { g.trans += 1; return FC(xt_sin(v.v)); }
inline FC xatan2(FC y, FC x) { g.trans += 1; return FC(xt_atan2(y.v, x.v)); }
inline FC xasin(FC v) { g.trans += 1; return FC(xt_asin(v.v)); }
inline double xval(FC v) { return v.v; }
}  // namespace xt

#include "xt_ops.h"

int xt_count_turn(const std::vector<int64_t> &entry, const std::vector<double> &prog,
                  double line_length, const double *coords, double p0c, double mass0,
                  double q0, double out[6]) {
    using P = xt::PState<FC>;
    P p;
    const double x = coords[0], px = coords[1], y = coords[2], py = coords[3];
    const double zeta = coords[4], delta = coords[5];
    // reference quantities (particles.py:1900-2030)
    const double energy0 = std::sqrt(p0c * p0c + mass0 * mass0);
    const double beta0 = p0c / energy0;
    const double delta_beta0 = delta * beta0;
    const double ptau_beta0 = std::sqrt(delta_beta0 * delta_beta0 + 2 * delta_beta0 * beta0 + 1) - 1;
    FC slow[xt::XS_N];
    p.slow = slow;
    p.slow_stride = 1;
    p.x = x; p.px = px; p.y = y; p.py = py; p.zeta = zeta; p.put(xt::XS_DELTA, FC(delta));
    p.put(xt::XS_RPP, FC(1 / (1 + delta)));
    p.put(xt::XS_RVV, FC((1 + delta) / (1 + ptau_beta0)));
    p.set_s(FC(0.0));
    p.chi = 1.0;
    p.put(xt::XS_INV_RVV, FC(1.0) / p.rvv());
    p.at_turn = 0; p.state = 1;
    xt::refresh_flags(p);
    // fields the maps read from and write to the particle arrays
    double m_ptau = ptau_beta0 / beta0, m_beta0 = beta0, m_charge_ratio = 1.0, m_p0c = p0c;
    double m_ax = 0.0, m_ay = 0.0;
    xtm_particles mp = {};
    mp.capacity = 1;
    mp.ptau = &m_ptau; mp.beta0 = &m_beta0; mp.charge_ratio = &m_charge_ratio;
    mp.p0c = &m_p0c; mp.ax = &m_ax; mp.ay = &m_ay;
    const xt::PRef m{&mp, 0};
    g = Counts();
    for (size_t e = 0; e < entry.size(); e++) {
        xt::run_element(p, m, prog.data(), entry[e], xt::TrackCtx{0, line_length, q0, 0.0});
        if (p.state <= 0) return XTM_ERR_ARG;   // reference particle must survive
    }
    out[0] = g.add;
    out[1] = g.mul;
    out[2] = g.div;
    out[3] = g.sqrt;
    out[4] = g.trans;
    out[5] = g.add + g.mul + g.div + g.sqrt;
    return XTM_OK;
}
