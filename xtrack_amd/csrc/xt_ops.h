/* xt_ops.h -- per-particle element maps of the tracking kernel.
 *
 * Templates on the scalar type R: the kernel instantiates them with double;
 * the host instantiates them with an operation-counting scalar to measure
 * the algorithmic work of exactly this code (xt_flops.cpp).
 *
 * Every map restates the reference expression operation by operation
 * (association and operand order preserved; build with -ffp-contract=off).
 * Element-uniform sub-expressions arrive precomputed in the element program
 * (xt_program.h) and were evaluated on the host with the same operations.
 *
 * Exact skipping of zero terms (multipole kick).  In
 * track_magnet_kick.h:24-144 a magnet of the path evaluates three Horner
 * passes and three curvature terms whose coefficients are mostly zero for a
 * given element (e.g. a sector bend's body kick reduces to the weak-focusing
 * term).  A term whose coefficient is exactly +-0 contributes
 * (+-0) * v = +-0 to a sum when v is finite, which leaves the sum unchanged
 * except possibly for the sign of a zero result.  The fast path therefore
 * skips such terms only for a lane whose operands are all finite and whose
 * chi is exactly 1 (so chi*k == k); any other lane runs the literal code.
 * Results are identical to the literal evaluation as IEEE values (+0 and -0
 * compare equal).
 */
#ifndef XT_OPS_H
#define XT_OPS_H

#include <stdint.h>

#include "xt_math.h"
#include "xt_program.h"
#include "xtrack_amd.h"

#if defined(__HIPCC__)
#define XT_HD __host__ __device__ __forceinline__
#define XT_HD_COLD __host__ __device__ __attribute__((noinline))
#else
#define XT_HD inline
#define XT_HD_COLD inline
#endif

namespace xt {

constexpr double C_LIGHT = 299792458.0;
constexpr double PI = 3.1415926535897932384626433832795028841971693993751;
constexpr double DEG2RAD = 0.0174532925199432957692369076848861271344287188854;
constexpr double GLOBAL_XY_LIMIT = 1.0;   // default line config, xtrack/line.py:287

/* flag bits of OP_KICK's mask argument (which folded terms are present) */
enum { KF_P1 = 1, KF_P2 = 2, KF_P3 = 4, KF_HL = 8, KF_G0 = 16, KF_H1 = 32 };

template <class R>
struct PState {
    R x, px, y, py, zeta, delta, ptau, rpp, rvv, s, ax, ay;
    R chi, charge_ratio, beta0, p0c;
    /* 1./rvv, kept equal to that IEEE quotient: set when the particle is
       loaded and whenever rvv changes (update_ptau) */
    R inv_rvv;
    int64_t at_element, at_turn, state;
    /* per-particle predicates of quantities that change only at load and in
       update_ptau (refresh_flags): the folded multipole kick may be used
       (chi == 1, finite delta, rvv and 1/rvv, rvv != 0); rvv is in the
       range of the reciprocal division (den_ok) */
    bool fast_ok, rvv_ok;
};

XT_HD int64_t wbits(double d) { int64_t u; __builtin_memcpy(&u, &d, 8); return u; }

/* elementary functions for double; the counting scalar overloads these */
XT_HD double xsqrt(double v) { return XT_SQRT(v); }
XT_HD void xsincos(double v, double *s, double *c) { xt_sincos(v, s, c); }
XT_HD void xsinhcosh(double v, double *s, double *c) { xt_sinhcosh(v, s, c); }
XT_HD double xsin(double v) { return xt_sin(v); }
XT_HD double xfabs(double v) { return XT_FABS(v); }
XT_HD bool xfinite(double v) { return XT_FABS(v) < XT_INF; }
XT_HD double xval(double v) { return v; }

/* a/b given y = RN(1/b): Markstein's correction applied twice,
   q0 = RN(a y), r = a - b q (exact with fma), q' = RN(q + r y).  After the
   first step q is within one ulp of a/b; the second step then yields the
   correctly rounded quotient (P. Markstein, IBM J. Res. Dev. 34 (1990);
   J.-M. Muller et al., Handbook of Floating-Point Arithmetic, 2nd ed.,
   section 4.9).  Valid when a, b, y are finite, b normal and no
   intermediate over/underflows; callers check ranges (div_ok). */
XT_HD double div_rcp(double a, double b, double y) {
    double q = a * y;
    double r = XT_FMA(-q, b, a);
    q = XT_FMA(r, y, q);
    r = XT_FMA(-q, b, a);
    return XT_FMA(r, y, q);
}


/* |v| in [2^-900, 2^900] (zero and non-finite values excluded) */
XT_HD bool num_ok(double v) {
    const double a = XT_FABS(v);
    return a >= 0x1p-900 && a <= 0x1p900;
}
/* |v| in [2^-16, 2^16] */
XT_HD bool den_ok(double v) {
    const double a = XT_FABS(v);
    return a >= 0x1p-16 && a <= 0x1p16;
}

template <class R>
XT_HD void refresh_flags(PState<R> &p) {
    p.fast_ok = (xval(p.chi) == 1.0) && xfinite(xval(p.delta)) && xfinite(xval(p.rvv)) &&
                (xval(p.rvv) != 0.0) && xfinite(xval(p.inv_rvv));
    p.rvv_ok = den_ok(xval(p.rvv));
}

/* ---- drift maps --------------------------------------------------------- */
/* track_drift.h:11-23 */
template <class R>
XT_HD void drift_expanded(PState<R> &p, const double length) {
    R const rpp = p.rpp;
    R const rv0v = p.inv_rvv;
    R const xp = p.px * rpp;
    R const yp = p.py * rpp;
    R const dzeta = 1 - rv0v * (1. + (xp * xp + yp * yp) / 2.);
    p.x = p.x + xp * length;
    p.y = p.y + yp * length;
    p.s = p.s + length;
    p.zeta = p.zeta + length * dzeta;
}

/* track_drift.h:26-41 */
template <class R>
XT_HD void drift_exact(PState<R> &p, const double length) {
    R const px = p.px, py = p.py;
    R const rv0v = p.inv_rvv;
    R const one_plus_delta = 1. + p.delta;
    R const one_over_pz = 1. / xsqrt(one_plus_delta * one_plus_delta - px * px - py * py);
    R const dzeta = 1 - rv0v * one_plus_delta * one_over_pz;
    p.x = p.x + px * one_over_pz * length;
    p.y = p.y + py * one_over_pz * length;
    p.zeta = p.zeta + dzeta * length;
    p.s = p.s + length;
}

/* Path-length term of the polar drift, the reference's literal divisions:
   length - (n1/ca/pz/(1 - n3/ca/pz))/rvv.  Out of line so that the
   reciprocal-division path below cannot be if-converted into computing
   both forms on every lane. */
XT_HD_COLD double polar_dz_plain(double n1, double n3, double ca, double pz, double rvv,
                                 double length) {
    const double delta_ell = n1 / ca / pz / (1 - n3 / ca / pz);
    return length - delta_ell / rvv;
}

/* Same value through div_rcp with the reciprocals ica = RN(1/ca) (element
   program), _pz = RN(1/pz) and inv_rvv = RN(1/rvv), which the drift has at
   hand.  Range argument for exactness: with |n1|,|n3| in [2^-900, 2^900]
   and ca, pz, den, rvv in [2^-16, 2^16] every dividend is within
   [2^-948, 2^948], far from over- and underflow of the remainder.  Lanes
   outside those ranges (including n3 == 0 for px == 0) take the literal
   form. */
XT_HD double polar_dz(double n1, double n3, double ca, double ica, double pz, double _pz,
                      double rvv, double inv_rvv, bool rvv_ok, double length) {
#ifdef XTM_PLAIN_DIVISION
    (void)ica; (void)_pz; (void)inv_rvv; (void)rvv_ok;
    const double delta_ell = n1 / ca / pz / (1 - n3 / ca / pz);
    return length - delta_ell / rvv;
#else
    const double q2 = div_rcp(div_rcp(n1, ca, ica), pz, _pz);
    const double q4 = div_rcp(div_rcp(n3, ca, ica), pz, _pz);
    const double den = 1 - q4;
    const double delta_ell = q2 / den;
    double dz = length - div_rcp(delta_ell, rvv, inv_rvv);
    const bool ok = rvv_ok && (ica != 0.0) && num_ok(n1) && num_ok(n3) && den_ok(pz) &&
                    den_ok(den);
    if (!ok) dz = polar_dz_plain(n1, n3, ca, pz, rvv, length);
    return dz;
#endif
}

/* operation-counting and other scalar types: the literal form */
template <class R>
XT_HD R polar_dz(R n1, R n3, double ca, double ica, R pz, R _pz, R rvv, R inv_rvv,
                 bool rvv_ok, double length) {
    (void)ica; (void)_pz; (void)inv_rvv; (void)rvv_ok;
    const R delta_ell = n1 / ca / pz / (1 - n3 / ca / pz);
    return length - delta_ell / rvv;
}

/* track_magnet_drift.h:45-87 with rho = 1/h, ca = cos(h s), sa = sin(h s),
   t2 = 2 sin(h s / 2)^2 and ica = 1/ca (0 when ca is outside [2^-16, 2^16]
   in magnitude) from the element program. */
template <class R>
XT_HD void polar_drift(PState<R> &p, const double length, const double rho, const double ca,
                       const double sa, const double t2, const double ica) {
    const R rvv = p.rvv;
    const R x = p.x, y = p.y, px = p.px, py = p.py;
    const R one_plus_delta = p.delta + 1.0;
    const R pz = xsqrt(one_plus_delta * one_plus_delta - px * px - py * py);
    const R _pz = 1 / pz;
    const R pxt = px * _pz;
    const R sapxt = sa * pxt;
    const R _ptt = 1 / (ca - sapxt);
    const R xr = x + rho;
    const R pst = xr * sa * _pz * _ptt;
    const R new_x = (x + rho * (t2 + sapxt)) * _ptt;
    const R new_px = ca * px + sa * pz;
    const R new_y = y + pst * py;
    /* delta_ell = opd*(x+rho)*sa/ca/pz/(1 - px*sa/ca/pz) */
    const R n1 = one_plus_delta * xr * sa;
    const R n3 = px * sa;
    const R dz = polar_dz(n1, n3, ca, ica, pz, _pz, rvv, p.inv_rvv, p.rvv_ok, length);
    p.x = new_x;
    p.px = new_px;
    p.y = new_y;
    p.zeta = p.zeta + dz;
    p.s = p.s + length;
}

template <class R>
struct DQOut { R x, px, y, py, dzeta; };

/* track_magnet_drift.h:91-213.  Out of line with scalar arguments and a
   five-value result (both passed in VGPRs), so the hot loop keeps neither
   its registers nor a memory copy of the particle. */
template <class R>
XT_HD_COLD DQOut<R> dipole_quad_map(const R x, const R px, const R y, const R py,
                                    const R delta, const R rvv, const R chi,
                                    const double length, const double k0_, const double k1_,
                                    const double h) {
    const R delta_plus_1 = delta + 1;
    const R k0 = chi * k0_ / delta_plus_1;
    const R k1 = chi * k1_ / delta_plus_1;
    const R Kx = k0 * h + k1;
    const R Ky = -k1;
    R Sx, Sy, Cx, Cy;
    if (Kx > 0.0) {
        R sqrt_Kx = xsqrt(Kx), sn, cs;
        xsincos(sqrt_Kx * length, &sn, &cs);
        Sx = sn / sqrt_Kx;
        Cx = cs;
    } else if (Kx < 0.0) {
        R sqrt_Kx = xsqrt(-Kx), sh, ch;
        xsinhcosh(sqrt_Kx * length, &sh, &ch);
        Sx = sh / sqrt_Kx;
        Cx = ch;
    } else {
        Sx = R(length);
        Cx = R(1.0);
    }
    if (Ky > 0.0) {
        R sqrt_Ky = xsqrt(Ky), sn, cs;
        xsincos(sqrt_Ky * length, &sn, &cs);
        Sy = sn / sqrt_Ky;
        Cy = cs;
    } else if (Ky < 0.0) {
        R sqrt_Ky = xsqrt(-Ky), sh, ch;
        xsinhcosh(sqrt_Ky * length, &sh, &ch);
        Sy = sh / sqrt_Ky;
        Cy = ch;
    } else {
        Sy = R(length);
        Cy = R(1.0);
    }
    const R xp = px / delta_plus_1;
    const R yp = py / delta_plus_1;
    const R A = -Kx * x - k0 + h;
    const R B = xp;
    const R C = -Ky * y;
    const R D = yp;
    R x_ = x * Cx + xp * Sx;
    const R y_ = y * Cy + yp * Sy;
    const R px_ = (A * Sx + B * Cx) * delta_plus_1;
    const R py_ = (C * Sy + D * Cy) * delta_plus_1;
    if (Kx != 0.0)
        x_ = x_ + (k0 - h) * (Cx - 1.0) / Kx;
    else
        x_ = x_ - (k0 - h) * 0.5 * (length * length);
    R length_ = R(length);
    if (Kx != 0.0) {
        const R AA = A * A, BB = B * B, K2 = 2.0 * Kx, AB = A * B;
        length_ = length_ - (h * ((Cx - 1.0) * xp + Sx * A + length * (k0 - h))) / Kx;
        length_ = length_ + 0.5 * (-(AA * Cx * Sx) / K2 + (BB * Cx * Sx) / 2.0 +
                                   (AA * length) / K2 + (BB * length) / 2.0 -
                                   (AB * (Cx * Cx)) / Kx + AB / Kx);
    } else {
        length_ = length_ + h * length * (3.0 * length * xp + 6.0 * x - (k0 - h) * (length * length)) / 6.0;
        length_ = length_ + 0.5 * (B * B) * length;
    }
    if (Ky != 0.0) {
        const R CC = C * C, DD = D * D, K2 = 2.0 * Ky, CD = C * D;
        length_ = length_ + 0.5 * (-(CC * Cy * Sy) / K2 + (DD * Cy * Sy) / 2.0 +
                                   (CC * length) / K2 + (DD * length) / 2.0 -
                                   (CD * (Cy * Cy)) / Ky + CD / Ky);
    } else {
        length_ = length_ + 0.5 * (D * D) * length;
    }
    const R dzeta = length - length_ / rvv;
    return DQOut<R>{x_, px_, y_, py_, dzeta};
}

template <class R>
XT_HD void dipole_quad(PState<R> &p, const double length, const double k0, const double k1,
                       const double h) {
    const DQOut<R> o = dipole_quad_map(p.x, p.px, p.y, p.py, p.delta, p.rvv, p.chi, length,
                                       k0, k1, h);
    p.x = o.x;
    p.px = o.px;
    p.y = o.y;
    p.py = o.py;
    p.zeta = p.zeta + o.dzeta;
    p.s = p.s + length;
}

/* ---- multipole kick (track_magnet_kick.h:24-144, 183-228) ----------------- */
/* literal Horner pass; f = inv-factorial chain f[order], ..., f[0] */
template <class R>
XT_HD void kick_pass(R &px, R &py, int64_t order, const double *__restrict__ f,
                     const double *__restrict__ knl, const double *__restrict__ ksl,
                     double factor, double kick_weight, R x, R y, R chi) {
    int64_t index = order;
    R dpx_mul = chi * knl[index] * factor * f[0];
    R dpy_mul = chi * ksl[index] * factor * f[0];
    int64_t j = 0;
    while (index > 0) {
        R const zre = dpx_mul * x - dpy_mul * y;
        R const zim = dpx_mul * y + dpy_mul * x;
        index -= 1;
        j += 1;
        R this_knl = chi * knl[index] * factor;
        R this_ksl = chi * ksl[index] * factor;
        dpx_mul = this_knl * f[j] + zre;
        dpy_mul = this_ksl * f[j] + zim;
    }
    dpx_mul = -dpx_mul;
    px = px + kick_weight * dpx_mul;
    py = py + kick_weight * dpy_mul;
}

template <class R>
struct KickOut { R px, py, zeta; };

/* the unfolded kick of track_magnet_kick.h:24-144; out of line with scalar
   arguments (see dipole_quad_map) */
template <class R>
XT_HD_COLD KickOut<R> magnet_kick_literal(const R x, const R y, R px, R py, R zeta,
                                          const R delta, const R chi, const R inv_rvv,
                                          double kick_weight, const double *__restrict__ d) {
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
    kick_pass(px, py, order, f1, knl, ksl, d[KD_FACTOR], kick_weight, x, y, chi);
    kick_pass(px, py, order_rel, f2, knl_rel, ksl_rel, d[KD_FACTOR_REL], kick_weight, x, y, chi);
    kick_pass(px, py, 3, f3, d + KD_MAIN, d + KD_MAINS, 1, kick_weight, x, y, chi);
    R dpx = R(0.0), dpy = R(0.0), dzeta = R(0.0);
    if (wbits(d[KD_ROT_FRAME])) {
        double const hl = d[KD_H] * length * kick_weight + d[KD_HXL] * kick_weight;
        dpx = dpx + hl * (1. + delta);
        R const rv0v = inv_rvv;
        dzeta = dzeta + -rv0v * hl * x;
    }
    const double htot = d[KD_HTOT];
    const double k0c = d[KD_K0HC] * length + d[KD_K0L];
    const double k1c = d[KD_K1HC] * length + d[KD_K1L];
    dpx = dpx + -chi * k0c * kick_weight * htot * x;
    dpx = dpx + htot * chi * k1c * kick_weight * (-x * x + 0.5 * y * y);
    dpy = dpy + htot * chi * k1c * kick_weight * x * y;
    return KickOut<R>{px + dpx, py + dpy, zeta + dzeta};
}

/* Horner pass on precomputed coefficients a[k] = (knl[k]*factor)*f[k],
   b[k] likewise (valid for chi == 1), starting at the highest non-zero
   index `top` (> = 0); zero coefficients skip their addition. */
template <class R>
XT_HD void kick_pass_fast(PState<R> &p, int64_t top, const double *__restrict__ a,
                          const double *__restrict__ b, double kick_weight, R x, R y) {
    R dpx_mul = R(a[top]);
    R dpy_mul = R(b[top]);
    for (int64_t k = top - 1; k >= 0; k--) {
        R const zre = dpx_mul * x - dpy_mul * y;
        R const zim = dpx_mul * y + dpy_mul * x;
        dpx_mul = (a[k] != 0.0) ? a[k] + zre : zre;
        dpy_mul = (b[k] != 0.0) ? b[k] + zim : zim;
    }
    p.px = p.px + kick_weight * (-dpx_mul);
    p.py = p.py + kick_weight * dpy_mul;
}

/* OP_KICK args: w desc mask hl g0 h1 (top1 top2 top3 as int bits in desc
   extension, see builder) */
template <class R>
XT_HD void magnet_kick(PState<R> &p, const double *__restrict__ a, const double *__restrict__ prog) {
    const double w = a[0];
    const double *__restrict__ d = prog + wbits(a[1]);
    const int64_t mask = wbits(a[2]);
    /* The folded form drops terms that the literal kick computes as an
       exact zero times a polynomial in x, y and then adds.  That is exact
       when (i) chi == 1 and delta, rvv are finite with rvv != 0 (fast_ok),
       (ii) |x|, |y| < 2^511, so x*x and y*y cannot overflow into a NaN
       through 0*inf, and (iii) px, py, zeta != 0, so adding a zero of
       either sign leaves them unchanged. */
    const double ax = xfabs(xval(p.x)), ay = xfabs(xval(p.y));
    const bool sane = p.fast_ok && ax < 0x1p511 && ay < 0x1p511 && xval(p.px) != 0.0 &&
                      xval(p.py) != 0.0 && xval(p.zeta) != 0.0;
    if (!sane) {
        const KickOut<R> o = magnet_kick_literal(p.x, p.y, p.px, p.py, p.zeta, p.delta, p.chi,
                                                 p.inv_rvv, w, d);
        p.px = o.px;
        p.py = o.py;
        p.zeta = o.zeta;
        return;
    }
    const R x = p.x, y = p.y;
    /* fast coefficients follow the literal descriptor: per pass
       [top, a[0..n], b[0..n]] starting at d[KD_FAST] offset stored in d */
    const double *fp = prog + wbits(d[KD_FASTOFF]);
    if (mask & KF_P1) {
        const int64_t top = wbits(fp[0]);
        kick_pass_fast(p, top, fp + 1, fp + 2 + top, w, x, y);
    }
    fp = prog + wbits(d[KD_FASTOFF + 1]);
    if (mask & KF_P2) {
        const int64_t top = wbits(fp[0]);
        kick_pass_fast(p, top, fp + 1, fp + 2 + top, w, x, y);
    }
    fp = prog + wbits(d[KD_FASTOFF + 2]);
    if (mask & KF_P3) {
        const int64_t top = wbits(fp[0]);
        kick_pass_fast(p, top, fp + 1, fp + 2 + top, w, x, y);
    }
    if (mask & (KF_HL | KF_G0 | KF_H1)) {
        R dpx, dpy;
        bool hx = false, hy = false;
        R dzeta;
        bool hz = false;
        if (mask & KF_HL) {
            const double hl = a[3];
            dpx = hl * (1. + p.delta);
            hx = true;
            R const rv0v = p.inv_rvv;
            dzeta = -rv0v * hl * x;
            hz = true;
        }
        if (mask & KF_G0) {
            /* -chi*k0c*w*htot*x with chi == 1: (((-k0c)*w)*htot) * x */
            const R t = a[4] * x;
            dpx = hx ? dpx + t : t;
            hx = true;
        }
        if (mask & KF_H1) {
            /* htot*chi*k1c*w*(...) with chi == 1: ((htot*k1c)*w) */
            const double h1 = a[5];
            const R t = h1 * (-x * x + 0.5 * y * y);
            dpx = hx ? dpx + t : t;
            hx = true;
            dpy = h1 * x * y;
            hy = true;
        }
        if (hx) p.px = p.px + dpx;
        if (hy) p.py = p.py + dpy;
        if (hz) p.zeta = p.zeta + dzeta;
    }
}

/* ---- RF (track_rf.h:18-167) and energy update (custom_api.h:21-32, 196-216) */
template <class R>
XT_HD void update_ptau(PState<R> &p, R new_ptau) {
    R const beta0 = p.beta0;
    R const ptau = new_ptau;
    R const irpp = xsqrt(ptau * ptau + 2 * ptau / beta0 + 1);
    R const new_rpp = 1. / irpp;
    R const new_rvv = irpp / (1 + beta0 * ptau);
    p.delta = irpp - 1;
    p.rvv = new_rvv;
    p.ptau = ptau;
    p.rpp = new_rpp;
    p.inv_rvv = 1. / new_rvv;
    refresh_flags(p);
}

/* new ptau after the cavity's energy kick (out of line, scalar arguments) */
template <class R>
XT_HD_COLD R rf_kick_ptau(const R zeta, const R ptau_in, const R beta0, const R charge_ratio,
                          const R chi, const R p0c, const int64_t at_turn,
                          const double *__restrict__ a, double line_length, double q0,
                          double t_sim) {
    const double voltage = a[0];
    const double frequency0 = a[1];
    const double harmonic = a[2];
    const double lag = a[3];
    const double phase = a[4];
    const int64_t absolute_time = wbits(a[5]);
    R phase0 = R(0.0);
    R frequency = R(frequency0);
    if (harmonic != 0) {
        R const t_rev0 = line_length / (beta0 * C_LIGHT);
        frequency = frequency + (harmonic / t_rev0);
    }
    if (absolute_time == 1) {
        phase0 = phase0 + 2 * PI * (double)at_turn * frequency * t_sim;
    }
    R const q = xfabs(q0) * charge_ratio;
    R const tau = zeta / beta0;
    R const energy_kick =
        q * voltage * xsin(phase0 + DEG2RAD * lag + phase - (2.0 * PI) / C_LIGHT * frequency * tau);
    double const rfmultipole_energy_kick = 0;
    R ptau = ptau_in;
    R const mass_ratio = charge_ratio / chi;
    ptau = ptau + (energy_kick + rfmultipole_energy_kick) / p0c / mass_ratio;
    return ptau;
}

/* ---- element program interpreter ------------------------------------------ */
template <class R>
XT_HD void run_element(PState<R> &p, const double *__restrict__ prog, int64_t pc, uint64_t flags,
                       double line_length, double q0, double t_sim) {
    const bool ignore_local = (flags >> 4) & 1;   // XS_FLAG_IGNORE_LOCAL_APERTURE
    const bool ignore_global = (flags >> 3) & 1;  // XS_FLAG_IGNORE_GLOBAL_APERTURE
    const bool kill_cavity = (flags >> 2) & 1;    // XS_FLAG_KILL_CAVITY_KICK
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
            polar_drift(p, a[0], a[1], a[2], a[3], a[4], a[5]);
            pc += 7;
            break;
        case OP_DIPQUAD:
            dipole_quad(p, a[0], a[1], a[2], a[3]);
            pc += 5;
            break;
        case OP_YOSHIDA4_BEND: {
            /* track_magnet_drift.h:500-509 */
            const double L = a[0], rho = a[1], c1 = a[2], c2 = a[3];
            polar_drift(p, a[4], rho, a[5], a[6], a[7], a[8]);
            p.px = p.px - c1 * p.chi * L;
            polar_drift(p, a[9], rho, a[10], a[11], a[12], a[13]);
            p.px = p.px - c2 * p.chi * L;
            polar_drift(p, a[9], rho, a[10], a[11], a[12], a[13]);
            p.px = p.px - c1 * p.chi * L;
            polar_drift(p, a[4], rho, a[5], a[6], a[7], a[8]);
            pc += 15;
            break;
        }
        case OP_YOSHIDA6_BEND: {
            /* track_magnet_drift.h:510-549 */
            const double L = a[0], rho = a[1];
            const double *kc = a + 2;
            const double *dd = a + 6;
            for (int j = 0; j < 8; j++) {
                const int di = j < 4 ? j : 7 - j;
                const double *s = dd + 5 * di;
                polar_drift(p, s[0], rho, s[1], s[2], s[3], s[4]);
                if (j < 7) {
                    const int ki = j < 4 ? j : 6 - j;
                    p.px = p.px - kc[ki] * p.chi * L;
                }
            }
            pc += 27;
            break;
        }
        case OP_KICK:
            magnet_kick(p, a, prog);
            pc += 1 + XT_NARGS_KICK;
            break;
        case OP_EDGE_AXAY:
            p.ax = a[0] * (p.y - 0.0);
            p.ay = a[1] * (p.x - 0.0);
            pc += 3;
            break;
        case OP_EDGE_CLEAR_AXAY:
            p.ax = R(0.);
            p.ay = R(0.);
            pc += 1;
            break;
        case OP_EDGE_LINEAR: {
            const R x = p.x, y = p.y, chi = p.chi;
            p.px = p.px + chi * a[0] * x;
            p.py = p.py + chi * a[1] * y;
            pc += 3;
            break;
        }
        case OP_RF_KICK:
            if (!kill_cavity)
                update_ptau(p, rf_kick_ptau(p.zeta, p.ptau, p.beta0, p.charge_ratio, p.chi,
                                            p.p0c, p.at_turn, a, line_length, q0, t_sim));
            pc += 7;
            break;
        case OP_LIMIT_RECT:
            if (!ignore_local) {
                const double x = xval(p.x), y = xval(p.y);
                const bool is_alive = (x >= a[0]) && (x <= a[1]) && (y >= a[2]) && (y <= a[3]);
                if (!is_alive) p.state = 0;
            }
            pc += 5;
            break;
        case OP_LIMIT_ELLIPSE:
            if (!ignore_local) {
                const R temp = p.x * p.x * a[1] + p.y * p.y * a[0];
                if (!(xval(temp) <= a[2])) p.state = 0;
            }
            pc += 4;
            break;
        case OP_GLOBAL_APER:
            if (!ignore_global) {
                const double x = xval(p.x), y = xval(p.y);
                const bool within = (x >= -GLOBAL_XY_LIMIT) && (x <= GLOBAL_XY_LIMIT) &&
                                    (y >= -GLOBAL_XY_LIMIT) && (y <= GLOBAL_XY_LIMIT);
                if (p.state > 0 && !within) p.state = -1;
            }
            pc += 1;
            break;
        default:
            p.state = -999;   /* corrupt program: never produced by the builder */
            return;
        }
    }
}

// ---- monitor (monitors/particles_monitor.h:13-77) ---------------------------
XT_HD void monitor_store(const xtm_monitor &M, const xtm_particles &P, int64_t i,
                          const PState<double> &p, int64_t store_at) {
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

XT_HD_COLD void monitor_record(const xtm_monitor &M, const xtm_particles &P, int64_t i,
                               const PState<double> p, int64_t pid) {
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


/* ---- one particle through [ele_start, ele_start+num_ele_track) for
   num_turns turns: the per-thread body of track_line (tracker.py:596-734)
   with the particle held in registers; monitor records only for active
   particles as in the reference CPU context (track.h:33-50). */
XT_HD void track_particle(const xtm_particles &P, int64_t i, const int64_t *__restrict__ entry,
                          const double *__restrict__ prog, double line_length,
                          int32_t num_turns, int32_t ele_start, int32_t num_ele_track,
                          int32_t flag_end_turn_actions, int32_t flag_reset_s,
                          int32_t flag_monitor, const xtm_monitor &M, uint64_t flags) {
    if (!(P.state[i] > 0)) return;
    PState<double> p;
    p.x = P.x[i]; p.px = P.px[i]; p.y = P.y[i]; p.py = P.py[i];
    p.zeta = P.zeta[i]; p.delta = P.delta[i]; p.ptau = P.ptau[i];
    p.rpp = P.rpp[i]; p.rvv = P.rvv[i]; p.s = P.s[i];
    p.ax = P.ax[i]; p.ay = P.ay[i];
    p.chi = P.chi[i]; p.charge_ratio = P.charge_ratio[i];
    p.beta0 = P.beta0[i]; p.p0c = P.p0c[i];
    p.at_element = P.at_element[i]; p.at_turn = P.at_turn[i]; p.state = P.state[i];
    p.inv_rvv = 1. / p.rvv;
    refresh_flags(p);
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

}  // namespace xt

#endif /* XT_OPS_H */
