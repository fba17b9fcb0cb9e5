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
 * skips such terms only for a lane where every such v is finite (chi == 1,
 * finite delta, rvv, 1/rvv and |x|, |y| < 2^511) and where px, py and zeta
 * are non-zero, so that a zero of either sign added to them changes no bit;
 * any other lane runs the literal code.  Results are bit-identical to the
 * literal evaluation, the sign of zero included.
 *
 * Division and square root.  Every quotient and square root is the
 * correctly rounded IEEE result, as in the reference.  Where operands are
 * inside ranges that rule out over- and underflow, the device computes
 * them with shorter exact sequences (reciprocal multiplication with
 * Markstein's correction, and the hardware estimates refined by Newton
 * steps); a lane outside the ranges recomputes the map with IEEE
 * operations.
 */
#ifndef XT_OPS_H
#define XT_OPS_H

#include <stdint.h>

#include <type_traits>

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

/* Particle state held in registers through a launch.  Fields that the
   element maps use rarely -- ptau (cavity only), ax/ay (dipole edges only),
   charge_ratio, beta0, p0c (cavity only) -- stay in the particle arrays
   and are read and written there (PRef); at_element is not held at all:
   every element passed adds one, so it is a per-particle base plus the
   element index (track_particle). */
/* slots of the energy fields in PState::slow */
enum { XS_DELTA = 0, XS_RPP = 1, XS_RVV = 2, XS_INV_RVV = 3, XS_S = 4, XS_ZETA0 = 5, XS_N = 6 };
/* dynamic LDS bytes per thread that a tracking kernel launch must provide */
constexpr int XT_LDS_BYTES_PER_THREAD = XS_N * 8;

template <class R>
struct PState {
    R x, px, y, py, zeta;
    R chi;
    /* delta, rpp, rvv and 1./rvv change only when the energy changes (RF
       kicks) but are read by most elements; s only advances by element
       lengths.  They live in per-thread memory, slot k at
       slow[k * slow_stride] (LDS in the kernels), and are read there by
       each element: held in registers they would be live across the whole
       element loop, and with the loop's out-of-line calls the register
       allocator keeps them in scratch memory instead and writes them back
       on every element.  inv_rvv is kept equal to the IEEE quotient
       1./rvv: set when the particle is loaded and whenever rvv changes
       (update_ptau). */
    R *slow;
    int32_t slow_stride;
    int64_t at_turn, state;
    /* per-particle predicates of quantities that change only at load and in
       update_ptau (refresh_flags): the folded multipole kick may be used
       (chi == 1, finite delta, rvv and 1/rvv, rvv != 0); delta and rvv
       are in the ranges of the polar drift's reciprocal-division form
       (delta in [-0.5, 1], rvv in [0.5, 2]) */
    bool fast_ok, polar_ok;

    XT_HD R get(int k) const {
        return slow[k * slow_stride];
    }
    XT_HD void put(int k, const R &v) const { slow[k * slow_stride] = v; }
    XT_HD R delta() const { return get(XS_DELTA); }
    XT_HD R rpp() const { return get(XS_RPP); }
    XT_HD R rvv() const { return get(XS_RVV); }
    XT_HD R inv_rvv() const { return get(XS_INV_RVV); }
    XT_HD R s() const { return get(XS_S); }
    XT_HD void set_s(const R &v) const { put(XS_S, v); }
};

/* the particle's slot in the arrays, for the fields not held in PState */
struct PRef {
    const xtm_particles *P;
    int64_t i;
};

XT_HD int64_t wbits(double d) { int64_t u; __builtin_memcpy(&u, &d, 8); return u; }

/* Element programs, entry offsets and record keys are read-only for the
   whole launch.  Reading them through the constant address space (4) lets
   the compiler load wave-uniform words with scalar loads into SGPRs
   through the scalar cache; through a generic pointer it cannot rule out
   that the kernel's own stores change them and loads every constant into
   vector registers, once per lane.  CP<T> is the pointer type for such
   reads; gptr converts back for the out-of-line functions. */
#if defined(__HIP_DEVICE_COMPILE__)
#define XT_CONST_AS __attribute__((address_space(4)))
#else
#define XT_CONST_AS
#endif
template <class T>
using CP = const XT_CONST_AS T *;
using CProg = CP<double>;
template <class T>
XT_HD CP<T> cptr(const T *p) { return (CP<T>)p; }
template <class T>
XT_HD const T *gptr(CP<T> p) { return (const T *)p; }

/* An element key (record index, program offset) is the same on every
   active lane of a wave: the lanes walk the elements together and a lost
   lane leaves the loop.  The compiler cannot always prove it -- a key
   carried out of a loop that lanes leave at different elements counts as
   divergent -- and would then address the element's constants with vector
   loads and dispatch on it lane by lane.  uniform() states it: the first
   active lane's value, in scalar registers. */
XT_HD int32_t uniform(const int32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_readfirstlane(v);
#else
    return v;
#endif
}
XT_HD int64_t uniform(const int64_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
#else
    return v;
#endif
}

/* elementary functions for double; the counting scalar overloads these */
XT_HD double xsqrt(double v) { return XT_SQRT(v); }
XT_HD void xsincos(double v, double *s, double *c) { xt_sincos(v, s, c); }
XT_HD void xsinhcosh(double v, double *s, double *c) { xt_sinhcosh(v, s, c); }
XT_HD double xsin(double v) { return xt_sin(v); }
XT_HD double xfabs(double v) { return XT_FABS(v); }
XT_HD bool xfinite(double v) { return XT_FABS(v) < XT_INF; }
XT_HD double xval(double v) { return v; }
XT_HD double xatan2(double y, double x) { return xt_atan2(y, x); }
XT_HD double xasin(double v) { return xt_asin(v); }

/* a/b given y = RN(1/b): Markstein's correction applied twice,
   q0 = RN(a y), r = a - b q (exact with fma), q' = RN(q + r y).  After the
   first step q is within one ulp of a/b; the second step then yields the
   correctly rounded quotient (P. Markstein, IBM J. Res. Dev. 34 (1990);
   J.-M. Muller et al., Handbook of Floating-Point Arithmetic, 2nd ed.,
   section 4.9).  Valid when a, b, y are finite, b normal and no
   intermediate over/underflows; callers establish the ranges. */
XT_HD double div_rcp(double a, double b, double y) {
    double q = a * y;
    double r = XT_FMA(-q, b, a);
    q = XT_FMA(r, y, q);
    r = XT_FMA(-q, b, a);
    return XT_FMA(r, y, q);
}


/* ---- in-range IEEE reciprocal, quotient and square root (device) --------
   The general IEEE sequences for 1/b, a/b and sqrt(x) on CDNA are: scale
   the operands away from over/underflow (v_div_scale / ldexp), refine the
   hardware estimate (v_rcp_f64 / v_rsq_f64) by Newton steps, apply one
   remainder correction, and patch special values (v_div_fixup / class
   test).  For operands in the ranges below the scaling and the patching are
   identities, so the functions here run only the refinement and the
   correction -- the same operations on the same values, hence the same
   correctly rounded results, in 7-10 instead of 11-17 instructions.
   Ranges (checked by the caller): rcp_ir: |b| in [2^-16, 2^16];
   div_ir: that and |a| in [2^-940, 2^940]; sqrt_ir: x in [2^-32, 2^32]. */
#if defined(__HIPCC__)
__device__ __forceinline__ double rcp_ir(double b) {
    double y = __builtin_amdgcn_rcp(b);
    double e = XT_FMA(-b, y, 1.0);
    y = XT_FMA(y, e, y);
    e = XT_FMA(-b, y, 1.0);
    y = XT_FMA(y, e, y);
    const double r = XT_FMA(-b, y, 1.0);
    return XT_FMA(r, y, y);
}
__device__ __forceinline__ double div_ir(double a, double b) {
    double y = __builtin_amdgcn_rcp(b);
    double e = XT_FMA(-b, y, 1.0);
    y = XT_FMA(y, e, y);
    e = XT_FMA(-b, y, 1.0);
    y = XT_FMA(y, e, y);
    const double q = a * y;
    const double r = XT_FMA(-b, q, a);
    return XT_FMA(r, y, q);
}
__device__ __forceinline__ double sqrt_ir(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = XT_FMA(-h, g, 0.5);
    g = XT_FMA(g, r, g);
    h = XT_FMA(h, r, h);
    double d = XT_FMA(-g, g, x);
    g = XT_FMA(d, h, g);
    d = XT_FMA(-g, g, x);
    return XT_FMA(d, h, g);
}
#endif

template <class R>
XT_HD void refresh_flags(PState<R> &p) {
    p.fast_ok = (xval(p.chi) == 1.0) && xfinite(xval(p.delta())) && xfinite(xval(p.rvv())) &&
                (xval(p.rvv()) != 0.0) && xfinite(xval(p.inv_rvv()));
    p.polar_ok = xval(p.delta()) >= -0.5 && xval(p.delta()) <= 1.0 && xval(p.rvv()) >= 0.5 &&
                 xval(p.rvv()) <= 2.0;
}

/* Lane condition of the polar drift's reciprocal-division form: with the
   particle in polar_ok, the element's ica != 0 (|h s| <= 0.1, |1/h| in
   [2, 2^400], |sa| >= 2^-400, checked by the builder) and |x| <= 1,
   |px|, |py| <= 0.25, every operand stays in a safe range:
   1 + delta in [0.5, 2], the radicand in [0.125, 4], pz in [0.35, 2],
   ca - sa px/pz and 1 - n3/ca/pz in [0.92, 1.08], |n1| in
   [2^-402, 2^400], |n3/ca/pz| <= 0.08, the dividend of the last two
   divisions within [2^-404, 2^403].  The quotients by ca and pz of a
   zero or tiny n3 (|n3| < 2^-900) may differ from the IEEE ones by the
   sign of zero or one ulp of a number below 2^-890, which
   1 - n3/ca/pz = 1 absorbs exactly. */
template <class R>
XT_HD bool polar_fast_ok(const PState<R> &p, const double ica) {
    return p.polar_ok && (ica != 0.0) && xfabs(xval(p.x)) <= 1.0 &&
           xfabs(xval(p.px)) <= 0.25 && xfabs(xval(p.py)) <= 0.25;
}

/* ---- drift maps --------------------------------------------------------- */
/* track_drift.h:11-23 */
template <class R>
XT_HD void drift_expanded(PState<R> &p, const double length) {
    R const rpp = p.rpp();
    R const rv0v = p.inv_rvv();
    R const xp = p.px * rpp;
    R const yp = p.py * rpp;
    R const dzeta = 1 - rv0v * (1. + (xp * xp + yp * yp) / 2.);
    p.x = p.x + xp * length;
    p.y = p.y + yp * length;
    p.set_s(p.s() + length);
    p.zeta = p.zeta + length * dzeta;
}

/* track_drift.h:26-41 */
template <class R>
XT_HD void drift_exact(PState<R> &p, const double length) {
    R const px = p.px, py = p.py;
    R const rv0v = p.inv_rvv();
    R const one_plus_delta = 1. + p.delta();
    R const one_over_pz = 1. / xsqrt(one_plus_delta * one_plus_delta - px * px - py * py);
    R const dzeta = 1 - rv0v * one_plus_delta * one_over_pz;
    p.x = p.x + px * one_over_pz * length;
    p.y = p.y + py * one_over_pz * length;
    p.zeta = p.zeta + dzeta * length;
    p.set_s(p.s() + length);
}

/* Fast arithmetic (line code compiled with XTM_ARITH_FAST): the polar
   drift's path-length term re-associated.  The reference evaluates

       delta_ell = (1+delta) (x+rho) sa / ca / pz / (1 - px sa / ca / pz)

   with five divisions; the same quantity is (1+delta) * pst with
   pst = (x+rho) sa (1/pz) (1/(ca - sa px/pz)), the factor the map already
   forms for y.  DzReassoc<double> selects that form: one product instead
   of four correctly rounded quotients, then the reference's quotient by
   rvv.  x, px, y and s keep the reference's operations and therefore its
   bits; zeta differs by the rounding of the two forms (a few ulp of
   delta_ell per drift, no bias).  The operation-counting scalar keeps the
   reference form, so the algorithmic count does not change. */
template <class R>
struct DzReassoc {
    static constexpr bool value = false;
};
#if defined(XTM_ARITH_FAST)
template <>
struct DzReassoc<double> {
    static constexpr bool value = true;
};
#endif

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
   hand, on lanes where polar_fast_ok holds; other lanes take the literal
   form. */
XT_HD double polar_dz(double n1, double n3, double ca, double ica, double pz, double _pz,
                      double rvv, double inv_rvv, bool fast, double length) {
#ifdef XTM_PLAIN_DIVISION
    (void)ica; (void)_pz; (void)inv_rvv; (void)fast;
    const double delta_ell = n1 / ca / pz / (1 - n3 / ca / pz);
    return length - delta_ell / rvv;
#else
    const double q2 = div_rcp(div_rcp(n1, ca, ica), pz, _pz);
    const double q4 = div_rcp(div_rcp(n3, ca, ica), pz, _pz);
    const double den = 1 - q4;
    const double delta_ell = q2 / den;
    double dz = length - div_rcp(delta_ell, rvv, inv_rvv);
    if (!fast) dz = polar_dz_plain(n1, n3, ca, pz, rvv, length);
    return dz;
#endif
}

/* operation-counting and other scalar types: the literal form */
template <class R>
XT_HD R polar_dz(R n1, R n3, double ca, double ica, R pz, R _pz, R rvv, R inv_rvv,
                 bool fast, double length) {
    (void)ica; (void)_pz; (void)inv_rvv; (void)fast;
    const R delta_ell = n1 / ca / pz / (1 - n3 / ca / pz);
    return length - delta_ell / rvv;
}

struct PolarOut { double x, px, y, dz; };

/* track_magnet_drift.h:45-87 with IEEE operations throughout: the fallback
   of polar_drift_ir, out of line by default (scalar arguments, four results
   in VGPRs). */
template <bool RE>
#ifdef XTM_POLAR_FALLBACK_INLINE
XT_HD PolarOut polar_plain(
#else
XT_HD_COLD PolarOut polar_plain(
#endif
                                const double x, const double px, const double y,
                                const double py, const double one_plus_delta,
                                const double rvv, const double length, const double rho,
                                const double ca, const double sa, const double t2) {
    const double pz = XT_SQRT(one_plus_delta * one_plus_delta - px * px - py * py);
    const double _pz = 1 / pz;
    const double pxt = px * _pz;
    const double sapxt = sa * pxt;
    const double _ptt = 1 / (ca - sapxt);
    const double xr = x + rho;
    const double pst = xr * sa * _pz * _ptt;
    PolarOut o;
    o.x = (x + rho * (t2 + sapxt)) * _ptt;
    o.px = ca * px + sa * pz;
    o.y = y + pst * py;
    if (RE) {
        o.dz = length - (one_plus_delta * pst) / rvv;
        return o;
    }
    const double delta_ell = one_plus_delta * xr * sa / ca / pz / (1 - px * sa / ca / pz);
    o.dz = length - delta_ell / rvv;
    return o;
}

#if defined(__HIP_DEVICE_COMPILE__)
/* The polar drift with in-range arithmetic: sqrt_ir for pz, rcp_ir for 1/pz
   and 1/(ca - sa px/pz), div_rcp with the reciprocals ica = RN(1/ca),
   1/pz and inv_rvv for the five quotients by ca, pz and rvv, div_ir for the
   division by 1 - n3/ca/pz, on lanes where polar_fast_ok holds (its
   ranges are inside those of sqrt_ir, rcp_ir, div_ir and div_rcp); other
   lanes take polar_plain. */
__device__ __forceinline__ void polar_drift_ir(PState<double> &p, const double length,
                                               const double rho, const double ca,
                                               const double sa, const double t2,
                                               const double ica) {
    const double rvv = p.rvv();
    const double x = p.x, y = p.y, px = p.px, py = p.py;
    const double one_plus_delta = p.delta() + 1.0;
    const double rad = one_plus_delta * one_plus_delta - px * px - py * py;
    const double pz = sqrt_ir(rad);
    const double _pz = rcp_ir(pz);
    const double pxt = px * _pz;
    const double sapxt = sa * pxt;
    const double tden = ca - sapxt;
    const double _ptt = rcp_ir(tden);
    const double xr = x + rho;
    const double pst = xr * sa * _pz * _ptt;
    double new_x = (x + rho * (t2 + sapxt)) * _ptt;
    double new_px = ca * px + sa * pz;
    double new_y = y + pst * py;
    constexpr bool RE = DzReassoc<double>::value;
    double dz;
    if constexpr (RE) {
        dz = length - div_rcp(one_plus_delta * pst, rvv, p.inv_rvv());
    } else {
        const double n1 = one_plus_delta * xr * sa;
        const double n3 = px * sa;
        const double q2 = div_rcp(div_rcp(n1, ca, ica), pz, _pz);
        const double q4 = div_rcp(div_rcp(n3, ca, ica), pz, _pz);
        const double den = 1 - q4;
        const double delta_ell = div_ir(q2, den);
        dz = length - div_rcp(delta_ell, rvv, p.inv_rvv());
    }
    if (!polar_fast_ok(p, ica)) {
        const PolarOut o =
            polar_plain<RE>(x, px, y, py, one_plus_delta, rvv, length, rho, ca, sa, t2);
        new_x = o.x;
        new_px = o.px;
        new_y = o.y;
        dz = o.dz;
    }
    p.x = new_x;
    p.px = new_px;
    p.y = new_y;
    p.zeta = p.zeta + dz;
    p.set_s(p.s() + length);
}
#endif

/* track_magnet_drift.h:45-87 with rho = 1/h, ca = cos(h s), sa = sin(h s),
   t2 = 2 sin(h s / 2)^2 and ica = 1/ca (0 when ca is outside [2^-16, 2^16]
   in magnitude) from the element program. */
template <class R>
XT_HD void polar_drift(PState<R> &p, const double length, const double rho, const double ca,
                       const double sa, const double t2, const double ica) {
#if defined(__HIP_DEVICE_COMPILE__) && !defined(XTM_PLAIN_DIVISION)
    if constexpr (std::is_same<R, double>::value) {
        polar_drift_ir(p, length, rho, ca, sa, t2, ica);
        return;
    }
#endif
    const R rvv = p.rvv();
    const R x = p.x, y = p.y, px = p.px, py = p.py;
    const R one_plus_delta = p.delta() + 1.0;
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
    R dz;
    if constexpr (DzReassoc<R>::value) {
        dz = length - (one_plus_delta * pst) / rvv;
    } else {
        const R n1 = one_plus_delta * xr * sa;
        const R n3 = px * sa;
        dz = polar_dz(n1, n3, ca, ica, pz, _pz, rvv, p.inv_rvv(), polar_fast_ok(p, ica), length);
    }
    p.x = new_x;
    p.px = new_px;
    p.y = new_y;
    p.zeta = p.zeta + dz;
    p.set_s(p.s() + length);
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
    const DQOut<R> o = dipole_quad_map(p.x, p.px, p.y, p.py, p.delta(), p.rvv(), p.chi, length,
                                       k0, k1, h);
    p.x = o.x;
    p.px = o.px;
    p.y = o.y;
    p.py = o.py;
    p.zeta = p.zeta + o.dzeta;
    p.set_s(p.s() + length);
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

/* ---- folded multipole kick, in steps shared by the element-program
   interpreter and by generated line code ----------------------------------- */

/* The folded form drops terms that the literal kick computes as an exact
   zero times a polynomial in x, y and then adds.  That is exact when
   (i) chi == 1 and delta, rvv, 1/rvv are finite with rvv != 0 (fast_ok),
   (ii) |x|, |y| < 2^511, so x*x and y*y cannot overflow into a NaN through
   0*inf, and (iii) px, py, zeta != 0, so adding a zero of either sign
   leaves them unchanged. */
template <class R>
XT_HD bool kick_sane(const PState<R> &p) {
    const double ax = xfabs(xval(p.x)), ay = xfabs(xval(p.y));
    return p.fast_ok && ax < 0x1p511 && ay < 0x1p511 && xval(p.px) != 0.0 &&
           xval(p.py) != 0.0 && xval(p.zeta) != 0.0;
}

/* the literal kick for lanes that fail kick_sane; d = kick descriptor */
template <class R>
XT_HD void kick_literal_apply(PState<R> &p, double w, const double *__restrict__ d) {
    const KickOut<R> o = magnet_kick_literal(p.x, p.y, p.px, p.py, p.zeta, p.delta(), p.chi,
                                             p.inv_rvv(), w, d);
    p.px = o.px;
    p.py = o.py;
    p.zeta = o.zeta;
}

/* One Horner step on the precomputed coefficients a = (knl[k]*factor)*f[k]
   and b likewise (chi == 1); a zero coefficient skips its addition. */
template <class R>
XT_HD void horner_step(R &dpx, R &dpy, const R x, const R y, const double a, const double b,
                       const bool has_a, const bool has_b) {
    R const zre = dpx * x - dpy * y;
    R const zim = dpx * y + dpy * x;
    dpx = has_a ? a + zre : zre;
    dpy = has_b ? b + zim : zim;
}

template <class R>
XT_HD void kick_pass_apply(PState<R> &p, const double w, const R dpx, const R dpy) {
    p.px = p.px + w * (-dpx);
    p.py = p.py + w * dpy;
}

/* curvature terms of track_magnet_kick.h:120-140 for chi == 1 with the
   element-uniform factors hl, g0 = ((-k0c)*w)*htot, h1 = (htot*k1c)*w */
template <class R>
XT_HD void kick_curvature(PState<R> &p, const R x, const R y, const bool has_hl,
                          const bool has_g0, const bool has_h1, const double hl,
                          const double g0, const double h1) {
    R dpx = R(0.0), dpy = R(0.0), dzeta = R(0.0);
    bool hx = false;
    if (has_hl) {
        dpx = hl * (1. + p.delta());
        hx = true;
        R const rv0v = p.inv_rvv();
        dzeta = -rv0v * hl * x;
    }
    if (has_g0) {
        /* -chi*k0c*w*htot*x with chi == 1: (((-k0c)*w)*htot) * x */
        const R t = g0 * x;
        dpx = hx ? dpx + t : t;
        hx = true;
    }
    if (has_h1) {
        /* htot*chi*k1c*w*(...) with chi == 1: ((htot*k1c)*w) */
        const R t = h1 * (-x * x + 0.5 * y * y);
        dpx = hx ? dpx + t : t;
        hx = true;
        dpy = h1 * x * y;
    }
    if (hx) p.px = p.px + dpx;
    if (has_h1) p.py = p.py + dpy;
    if (has_hl) p.zeta = p.zeta + dzeta;
}

/* Horner pass on a coefficient block [top, a[0..top], b[0..top]] */
template <class R, class PP>
XT_HD void kick_pass_fast(PState<R> &p, const PP fp, double kick_weight, R x, R y) {
    const int64_t top = wbits(fp[0]);
    const PP a = fp + 1, b = fp + 2 + top;
    R dpx_mul = R(a[top]);
    R dpy_mul = R(b[top]);
    for (int64_t k = top - 1; k >= 0; k--)
        horner_step(dpx_mul, dpy_mul, x, y, a[k], b[k], a[k] != 0.0, b[k] != 0.0);
    kick_pass_apply(p, kick_weight, dpx_mul, dpy_mul);
}

/* OP_KICK args: w desc mask hl g0 h1; the descriptor holds the literal
   parameters and the offsets of the three coefficient blocks */
template <class R, class PP>
XT_HD void magnet_kick(PState<R> &p, const PP a, const PP prog) {
    const double w = a[0];
    const PP d = prog + wbits(a[1]);
    const int64_t mask = wbits(a[2]);
    if (!kick_sane(p)) {
        kick_literal_apply(p, w, (const double *)d);
        return;
    }
    const R x = p.x, y = p.y;
    if (mask & KF_P1) kick_pass_fast(p, prog + wbits(d[KD_FASTOFF]), w, x, y);
    if (mask & KF_P2) kick_pass_fast(p, prog + wbits(d[KD_FASTOFF + 1]), w, x, y);
    if (mask & KF_P3) kick_pass_fast(p, prog + wbits(d[KD_FASTOFF + 2]), w, x, y);
    if (mask & (KF_HL | KF_G0 | KF_H1))
        kick_curvature(p, x, y, (mask & KF_HL) != 0, (mask & KF_G0) != 0, (mask & KF_H1) != 0,
                       a[3], a[4], a[5]);
}

/* ---- RF (track_rf.h:18-167) and energy update (custom_api.h:21-32, 196-216) */
template <class R>
XT_HD void update_ptau(PState<R> &p, R new_ptau, const PRef &m) {
    R const beta0 = R(m.P->beta0[m.i]);
    R const ptau = new_ptau;
    R const irpp = xsqrt(ptau * ptau + 2 * ptau / beta0 + 1);
    R const new_rpp = 1. / irpp;
    R const new_rvv = irpp / (1 + beta0 * ptau);
    p.put(XS_DELTA, irpp - 1);
    p.put(XS_RVV, new_rvv);
    m.P->ptau[m.i] = xval(ptau);
    p.put(XS_RPP, new_rpp);
    p.put(XS_INV_RVV, 1. / new_rvv);
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

template <class R>
struct RFMOut { R px, py, ptau; };

/* track_rf.h:18-167 for an RF multipole (harmonic 0, absolute_time 0, no
   transverse voltage): the transverse kicks of orders 0..order and the new
   ptau after the energy kick (unchanged when kill_energy_kick).  The
   element-uniform terms come from the descriptor d (RFD_* in
   xt_program.h).  Out of line, scalar arguments. */
template <class R>
XT_HD_COLD RFMOut<R> rf_multipole_kick(const R x, const R y, const R px, const R py,
                                       const R zeta, const R ptau_in, const R beta0,
                                       const R charge_ratio, const R chi, const R p0c,
                                       const double voltage, const double frequency,
                                       const double lag, const double phase,
                                       const double *__restrict__ d, const double q0,
                                       const bool kill_energy_kick) {
    const R phase0 = R(0.0);
    const R q = xfabs(q0) * charge_ratio;
    const R tau = zeta / beta0;
    const R energy_kick =
        q * voltage * xsin(phase0 + DEG2RAD * lag + phase - (2.0 * PI) / C_LIGHT * frequency * tau);
    const int64_t order = wbits(d[RFD_ORDER]);
    R dpx = R(0.0), dpy = R(0.0), dptr = R(0.0), zre = R(1.0), zim = R(0.0);
    for (int64_t kk = 0; kk <= order; kk++) {
        const double *t = d + RFD_TERMS + RFD_STRIDE * kk;
        const R pn_kk = phase0 + t[0] + t[1] - (2.0 * PI) / C_LIGHT * frequency * tau;
        const R ps_kk = phase0 + t[2] + t[3] - (2.0 * PI) / C_LIGHT * frequency * tau;
        const double bal_n_kk = t[4], bal_s_kk = t[5];
        R sn, cn, ss, cs;
        xsincos(pn_kk, &sn, &cn);
        xsincos(ps_kk, &ss, &cs);
        dpx = dpx + (cn * (bal_n_kk * zre) - cs * (bal_s_kk * zim));
        dpy = dpy + (cs * (bal_s_kk * zre) + cn * (bal_n_kk * zim));
        const R zret = zre * x - zim * y;
        zim = zim * x + zre * y;
        zre = zret;
        dptr = dptr + (sn * (bal_n_kk * zre) - ss * (bal_s_kk * zim));
    }
    const R rfmultipole_energy_kick = -q * ((frequency * (2.0 * PI / C_LIGHT) * p0c) * dptr);
    RFMOut<R> o;
    o.px = px + (-chi * dpx);
    o.py = py + chi * dpy;
    o.ptau = ptau_in;
    if (!kill_energy_kick) {
        const R mass_ratio = charge_ratio / chi;
        o.ptau = ptau_in + (energy_kick + rfmultipole_energy_kick) / p0c / mass_ratio;
    }
    return o;
}

/* track_magnet_drift.h:271-346: exact map of a sector bend with field k0
   and curvature h (bend-kick-bend, h != 0); sin(hs), cos(hs) and
   sin(hs/2)^2 hoisted from the element; a particle with |k0 chi| < 1e-8
   takes the polar drift (rho = 1/h, t2, ica as for OP_POLAR). */
template <class R>
XT_HD void op_bend_curved(PState<R> &p, const double L, const double k0, const double rho,
                          const double hs, const double sin_hs, const double cos_hs,
                          const double s2sq, const double t2, const double ica) {
    const R k0_chi = k0 * p.chi;
    if (xfabs(xval(k0_chi)) < 1e-8) {
        polar_drift(p, L, rho, cos_hs, sin_hs, t2, ica);
        return;
    }
    const R rvv = p.rvv();
    const R x0 = p.x, y0 = p.y, px0 = p.px, py = p.py;
    const R one_plus_delta = p.delta() + 1.0;
    const R pz0 = xsqrt(one_plus_delta * one_plus_delta - px0 * px0 - py * py);
    const R C = pz0 - k0_chi * (rho + x0);
    const R pxs = px0 * cos_hs + C * sin_hs;
    const R pzs = xsqrt(one_plus_delta * one_plus_delta - pxs * pxs - py * py);
    const R delta_pz = (px0 - pxs) * (px0 + pxs) / (pz0 + pzs);
    const R delta_D = -2 * C * s2sq - px0 * sin_hs;
    const R delta_x = (delta_pz - delta_D) / k0_chi;
    const R delta_px = -2 * px0 * s2sq + C * sin_hs;
    const R N_a = px0 * delta_pz - pz0 * delta_px;
    const R D_a = pz0 * pzs + px0 * pxs;
    const R delta_a = xatan2(N_a, D_a);
    const R integ = (hs + delta_a) / k0_chi;
    const R new_y = y0 + py * integ;
    const R delta_ell = one_plus_delta * integ;
    p.x = p.x + delta_x;
    p.px = pxs;
    p.y = new_y;
    p.zeta = p.zeta + (L - delta_ell / rvv);
    p.set_s(p.s() + L);
}

/* track_magnet_drift.h:348-394: exact map of a straight bend (h = 0) with
   field k0; |k0 chi| < 1e-8 takes the exact drift */
template <class R>
XT_HD void op_bend_straight(PState<R> &p, const double L, const double k0) {
    const R k0_chi = k0 * p.chi;
    if (xfabs(xval(k0_chi)) < 1e-8) {
        drift_exact(p, L);
        return;
    }
    const R rvv = p.rvv();
    const R x = p.x, y = p.y, px = p.px, py = p.py;
    const R one_plus_delta = p.delta() + 1.0;
    const R A = 1.0 / xsqrt(one_plus_delta * one_plus_delta - py * py);
    const R pz = xsqrt(one_plus_delta * one_plus_delta - px * px - py * py);
    const R new_px = px - k0_chi * L;
    const R new_x = x + (xsqrt(one_plus_delta * one_plus_delta - new_px * new_px - py * py) - pz) /
                            k0_chi;
    const R D = xasin(A * px) - xasin(A * new_px);
    const R new_y = y + (py / k0_chi) * D;
    const R delta_ell = (one_plus_delta / k0_chi) * D;
    p.x = new_x;
    p.px = new_px;
    p.y = new_y;
    p.zeta = p.zeta + (L - delta_ell / rvv);
    p.set_s(p.s() + L);
}

/* ---- element ops, shared by the interpreter and generated line code ------- */

/* track_magnet_drift.h:500-509: four polar drifts with the k0 kicks of the
   fourth-order Yoshida step; sub-steps a and b carry their own constants */
template <class R>
XT_HD void op_yoshida4(PState<R> &p, const double L, const double rho, const double c1,
                       const double c2, const double la, const double caa, const double saa,
                       const double t2a, const double ia, const double lb, const double cab,
                       const double sab, const double t2b, const double ib) {
    polar_drift(p, la, rho, caa, saa, t2a, ia);
    p.px = p.px - c1 * p.chi * L;
    polar_drift(p, lb, rho, cab, sab, t2b, ib);
    p.px = p.px - c2 * p.chi * L;
    polar_drift(p, lb, rho, cab, sab, t2b, ib);
    p.px = p.px - c1 * p.chi * L;
    polar_drift(p, la, rho, caa, saa, t2a, ia);
}

/* track_magnet_drift.h:510-549 (a = the 26 op arguments) */
template <class R, class PP>
XT_HD void op_yoshida6(PState<R> &p, const PP a) {
    const double L = a[0], rho = a[1];
    const PP kc = a + 2;
    const PP dd = a + 6;
    for (int j = 0; j < 8; j++) {
        const int di = j < 4 ? j : 7 - j;
        const PP s = dd + 5 * di;
        polar_drift(p, s[0], rho, s[1], s[2], s[3], s[4]);
        if (j < 7) {
            const int ki = j < 4 ? j : 6 - j;
            p.px = p.px - kc[ki] * p.chi * L;
        }
    }
}

template <class R>
XT_HD void op_edge_axay(const PState<R> &p, const PRef &m, const double ca, const double cb) {
    m.P->ax[m.i] = xval(ca * (p.y - 0.0));
    m.P->ay[m.i] = xval(cb * (p.x - 0.0));
}

XT_HD void op_edge_clear_axay(const PRef &m) {
    m.P->ax[m.i] = 0.;
    m.P->ay[m.i] = 0.;
}

template <class R>
XT_HD void op_edge_linear(PState<R> &p, const double r21, const double r43) {
    const R x = p.x, y = p.y, chi = p.chi;
    p.px = p.px + chi * r21 * x;
    p.py = p.py + chi * r43 * y;
}

/* a = the six RF arguments */
template <class R, class PP>
XT_HD void op_rf_kick(PState<R> &p, const PRef &m, const PP a, const double line_length,
                      const double q0, const double t_sim, const bool kill_cavity) {
    if (kill_cavity) return;
    update_ptau(p,
                rf_kick_ptau(p.zeta, R(m.P->ptau[m.i]), R(m.P->beta0[m.i]),
                             R(m.P->charge_ratio[m.i]), p.chi, R(m.P->p0c[m.i]), p.at_turn,
                             (const double *)a, line_length, q0, t_sim),
                m);
}

/* rfmultipole.h:11-47: one kick (model -1); the energy update with
   pz_only = 1 (custom_api.h:196-216) */
template <class R, class PP>
XT_HD void op_rf_mult(PState<R> &p, const PRef &m, const PP a, const PP prog, const double q0,
                      const bool kill_cavity) {
    const RFMOut<R> o = rf_multipole_kick(
        p.x, p.y, p.px, p.py, p.zeta, R(m.P->ptau[m.i]), R(m.P->beta0[m.i]),
        R(m.P->charge_ratio[m.i]), p.chi, R(m.P->p0c[m.i]), a[0], a[1], a[2], a[3],
        (const double *)(prog + wbits(a[4])), q0, kill_cavity);
    p.px = o.px;
    p.py = o.py;
    if (!kill_cavity) update_ptau(p, o.ptau, m);
}

/* limitrect.h:10-38 */
template <class R>
XT_HD void op_limit_rect(PState<R> &p, const double min_x, const double max_x,
                         const double min_y, const double max_y, const bool ignore_local) {
    if (ignore_local) return;
    const double x = xval(p.x), y = xval(p.y);
    const bool is_alive = (x >= min_x) && (x <= max_x) && (y >= min_y) && (y <= max_y);
    if (!is_alive) p.state = 0;
}

/* limitellipse.h:13-38 */
template <class R>
XT_HD void op_limit_ellipse(PState<R> &p, const double a_squ, const double b_squ,
                            const double a_b_squ, const bool ignore_local) {
    if (ignore_local) return;
    const R temp = p.x * p.x * b_squ + p.y * p.y * a_squ;
    if (!(xval(temp) <= a_b_squ)) p.state = 0;
}

/* local_particle_custom_api.h:259-287 */
template <class R>
XT_HD void op_global_aper(PState<R> &p, const bool ignore_global) {
    if (ignore_global) return;
    const double x = xval(p.x), y = xval(p.y);
    const bool within = (x >= -GLOBAL_XY_LIMIT) && (x <= GLOBAL_XY_LIMIT) &&
                        (y >= -GLOBAL_XY_LIMIT) && (y <= GLOBAL_XY_LIMIT);
    if (p.state > 0 && !within) p.state = -1;
}

/* track_xyshift.h (add_to_x(-dx), add_to_y(-dy) with the negation done
   when the program was built) */
template <class R>
XT_HD void op_xy_shift(PState<R> &p, const double ax, const double ay) {
    p.x = p.x + ax;
    p.y = p.y + ay;
}

/* S_SHIFT of track_misalignments.h:26-30: Drift_single_particle_exact over
   ds (track_drift.h:26-41), then zeta and s moved back by ds */
template <class R>
XT_HD void op_s_shift(PState<R> &p, const double ds) {
    drift_exact(p, ds);
    p.zeta = p.zeta + (-ds);
    p.set_s(p.s() + (-ds));
}

/* track_srotation.h:11-40 (spin is not tracked) */
template <class R>
XT_HD void op_s_rot(PState<R> &p, const double sin_z, const double cos_z) {
    const R x = p.x, y = p.y, px = p.px, py = p.py;
    p.x = cos_z * x + sin_z * y;
    p.y = -sin_z * x + cos_z * y;
    p.px = cos_z * px + sin_z * py;
    p.py = -sin_z * px + cos_z * py;
}

template <class R>
struct RotOut { R x, y, p, zeta; };

/* track_xrotation.h:11-45 (spin is not tracked): returns x, y, py, zeta.
   Out of line: frame rotations are rare and need ptau and beta0 from
   memory. */
template <class R>
XT_HD_COLD RotOut<R> x_rotation(const R x, const R y, const R px, const R py, const R zeta,
                                const R ptau, const R beta0, const double sin_angle,
                                const double cos_angle, const double tan_angle) {
    const R t = zeta / beta0;
    const R pt = (ptau / beta0) * beta0;   /* LocalParticle_get_pzeta * beta0 */
    const R pz = xsqrt(1.0 + 2.0 * pt / beta0 + pt * pt - px * px - py * py);
    const R ptt = 1.0 - tan_angle * py / pz;
    RotOut<R> o;
    o.y = y / (cos_angle * ptt);
    o.p = cos_angle * py + sin_angle * pz;
    o.x = x + tan_angle * y * px / (pz * ptt);
    const R t_hat = t - tan_angle * y * (1.0 / beta0 + pt) / (pz * ptt);
    o.zeta = t_hat * beta0;
    return o;
}

/* track_yrotation.h:11-45 (spin is not tracked): returns x, y, px, zeta */
template <class R>
XT_HD_COLD RotOut<R> y_rotation(const R x, const R y, const R px, const R py, const R zeta,
                                const R ptau, const R beta0, const double sin_angle,
                                const double cos_angle, const double tan_angle) {
    const R t = zeta / beta0;
    const R pt = (ptau / beta0) * beta0;
    const R pz = xsqrt(1.0 + 2.0 * pt / beta0 + pt * pt - px * px - py * py);
    const R ptt = 1.0 + tan_angle * px / pz;
    RotOut<R> o;
    o.x = x / (cos_angle * ptt);
    o.p = cos_angle * px - sin_angle * pz;
    o.y = y - tan_angle * x * py / (pz * ptt);
    const R t_hat = t + tan_angle * x * (1.0 / beta0 + pt) / (pz * ptt);
    o.zeta = t_hat * beta0;
    return o;
}

template <class R>
XT_HD void op_x_rot(PState<R> &p, const PRef &m, const double sin_angle, const double cos_angle,
                    const double tan_angle) {
    const RotOut<R> o = x_rotation(p.x, p.y, p.px, p.py, p.zeta, R(m.P->ptau[m.i]),
                                   R(m.P->beta0[m.i]), sin_angle, cos_angle, tan_angle);
    p.x = o.x;
    p.py = o.p;
    p.y = o.y;
    p.zeta = o.zeta;
}

template <class R>
XT_HD void op_y_rot(PState<R> &p, const PRef &m, const double sin_angle, const double cos_angle,
                    const double tan_angle) {
    const RotOut<R> o = y_rotation(p.x, p.y, p.px, p.py, p.zeta, R(m.P->ptau[m.i]),
                                   R(m.P->beta0[m.i]), sin_angle, cos_angle, tan_angle);
    p.x = o.x;
    p.px = o.p;
    p.y = o.y;
    p.zeta = o.zeta;
}

struct TrackCtx {
    uint64_t flags;
    double line_length, q0, t_sim;
    int64_t num_ele_line;   // elements in the line (backtracking turn start)
    XT_HD bool backtrack() const { return flags & 1; }              // XS_FLAG_BACKTRACK
    XT_HD bool ignore_local() const { return (flags >> 4) & 1; }   // XS_FLAG_IGNORE_LOCAL_APERTURE
    XT_HD bool ignore_global() const { return (flags >> 3) & 1; }  // XS_FLAG_IGNORE_GLOBAL_APERTURE
    XT_HD bool kill_cavity() const { return (flags >> 2) & 1; }    // XS_FLAG_KILL_CAVITY_KICK
    XT_HD bool freeze_long() const { return (flags >> 16) & 1; }   // XTM_FLAG_FREEZE_LONGITUDINAL
    /* energy kicks skipped: killed cavities, or frozen longitudinal motion */
    XT_HD bool kill_energy() const { return kill_cavity() || freeze_long(); }
};

/* ---- element program interpreter ------------------------------------------ */
template <class R>
XT_HD void run_element(PState<R> &p, const PRef &m, const CProg prog, int64_t pc,
                       const TrackCtx &c) {
    for (;;) {
        const int op = (int)wbits(prog[pc]);
        const CProg a = prog + pc + 1;
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
        case OP_BEND_CURVED:
            op_bend_curved(p, a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8]);
            pc += 10;
            break;
        case OP_BEND_STRAIGHT:
            op_bend_straight(p, a[0], a[1]);
            pc += 3;
            break;
        case OP_DIPQUAD:
            dipole_quad(p, a[0], a[1], a[2], a[3]);
            pc += 5;
            break;
        case OP_YOSHIDA4_BEND:
            op_yoshida4(p, a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10],
                        a[11], a[12], a[13]);
            pc += 15;
            break;
        case OP_YOSHIDA6_BEND:
            op_yoshida6(p, a);
            pc += 27;
            break;
        case OP_KICK:
            magnet_kick(p, a, prog);
            pc += 1 + XT_NARGS_KICK;
            break;
        case OP_EDGE_AXAY:
            op_edge_axay(p, m, a[0], a[1]);
            pc += 3;
            break;
        case OP_EDGE_CLEAR_AXAY:
            op_edge_clear_axay(m);
            pc += 1;
            break;
        case OP_EDGE_LINEAR:
            op_edge_linear(p, a[0], a[1]);
            pc += 3;
            break;
        case OP_RF_KICK:
            op_rf_kick(p, m, a, c.line_length, c.q0, c.t_sim, c.kill_energy());
            pc += 7;
            break;
        case OP_LIMIT_RECT:
            op_limit_rect(p, a[0], a[1], a[2], a[3], c.ignore_local());
            pc += 5;
            break;
        case OP_LIMIT_ELLIPSE:
            op_limit_ellipse(p, a[0], a[1], a[2], c.ignore_local());
            pc += 4;
            break;
        case OP_GLOBAL_APER:
            op_global_aper(p, c.ignore_global());
            pc += 1;
            break;
        case OP_XY_SHIFT:
            op_xy_shift(p, a[0], a[1]);
            pc += 3;
            break;
        case OP_S_SHIFT:
            op_s_shift(p, a[0]);
            pc += 2;
            break;
        case OP_X_ROT:
            op_x_rot(p, m, a[0], a[1], a[2]);
            pc += 4;
            break;
        case OP_Y_ROT:
            op_y_rot(p, m, a[0], a[1], a[2]);
            pc += 4;
            break;
        case OP_S_ROT:
            op_s_rot(p, a[0], a[1]);
            pc += 3;
            break;
        case OP_RF_MULT:
            op_rf_mult(p, m, a, prog, c.q0, c.kill_energy());
            pc += 6;
            break;
        default:
            p.state = -999;   /* corrupt program: never produced by the builder */
            return;
        }
    }
}

// ---- monitor (monitors/particles_monitor.h:13-77) ---------------------------
XT_HD void monitor_store(const xtm_monitor &M, const xtm_particles &P, int64_t i,
                          const PState<double> &p, int64_t at_element, int64_t store_at) {
    const xtm_particles &D = M.data;
    D.p0c[store_at] = P.p0c[i];
    D.gamma0[store_at] = P.gamma0[i];
    D.beta0[store_at] = P.beta0[i];
    D.s[store_at] = p.s();
    D.zeta[store_at] = p.zeta;
    D.x[store_at] = p.x;
    D.y[store_at] = p.y;
    D.px[store_at] = p.px;
    D.py[store_at] = p.py;
    D.ptau[store_at] = P.ptau[i];
    D.delta[store_at] = p.delta();
    D.rpp[store_at] = p.rpp();
    D.rvv[store_at] = p.rvv();
    D.chi[store_at] = p.chi;
    D.charge_ratio[store_at] = P.charge_ratio[i];
    D.weight[store_at] = P.weight[i];
    D.ax[store_at] = P.ax[i];
    D.ay[store_at] = P.ay[i];
    D.spin_x[store_at] = P.spin_x[i];
    D.spin_y[store_at] = P.spin_y[i];
    D.spin_z[store_at] = P.spin_z[i];
    D.anomalous_magnetic_moment[store_at] = P.anomalous_magnetic_moment[i];
    D.pdg_id[store_at] = P.pdg_id[i];
    D.particle_id[store_at] = P.particle_id[i];
    D.at_element[store_at] = at_element;
    D.at_turn[store_at] = p.at_turn;
    D.state[store_at] = p.state;
    D.parent_particle_id[store_at] = P.parent_particle_id[i];
    D.rng_s1[store_at] = P.rng_s1[i];
    D.rng_s2[store_at] = P.rng_s2[i];
    D.rng_s3[store_at] = P.rng_s3[i];
    D.rng_s4[store_at] = P.rng_s4[i];
}

/* Out of line and taking the state fields it stores as scalars, so that a
   launch with monitoring off carries no cost for it. */
XT_HD_COLD void monitor_record(const xtm_monitor &M, const xtm_particles &P, int64_t i,
                               double x, double px, double y, double py, double zeta,
                               double delta, double rpp, double rvv, double s, double chi,
                               int64_t at_turn, int64_t state, int64_t at_element,
                               bool backtrack) {
    /* particles_monitor.h: a turn-by-turn record is not taken while
       backtracking (element-by-element records are) */
    if (backtrack && !M.ebe_mode) return;
    PState<double> p;
    double slow[XS_N];
    p.slow = slow;
    p.slow_stride = 1;
    p.x = x; p.px = px; p.y = y; p.py = py; p.zeta = zeta; p.put(XS_DELTA, delta);
    p.put(XS_RPP, rpp); p.put(XS_RVV, rvv); p.set_s(s); p.chi = chi; p.at_turn = at_turn;
    p.state = state;
    const int64_t pid = P.particle_id[i];
    const int64_t n_turns_record = M.stop_at_turn - M.start_at_turn;
    const int64_t at_turn_m = M.ebe_mode ? at_element : at_turn;
    if (M.n_repetitions == 1) {
        if (at_turn_m >= M.start_at_turn && at_turn_m < M.stop_at_turn &&
            pid < M.part_id_end && pid >= M.part_id_start) {
            monitor_store(M, P, i, p, at_element,
                          n_turns_record * (pid - M.part_id_start) + at_turn_m - M.start_at_turn);
        }
    } else if (M.n_repetitions > 1 && at_turn_m >= M.start_at_turn) {
        const int64_t i_frame = (at_turn_m - M.start_at_turn) / M.repetition_period;
        if (i_frame < M.n_repetitions &&
            at_turn_m >= M.start_at_turn + i_frame * M.repetition_period &&
            at_turn_m < M.stop_at_turn + i_frame * M.repetition_period &&
            pid < M.part_id_end && pid >= M.part_id_start) {
            monitor_store(M, P, i, p, at_element,
                          n_turns_record * (M.part_id_end - M.part_id_start) * i_frame +
                              n_turns_record * (pid - M.part_id_start) +
                              (at_turn_m - i_frame * M.repetition_period) - M.start_at_turn);
        }
    }
}

/* What the monitor records need: the monitor and the tracked particles'
   struct, by pointer.  On the GPU both point to device copies, so that no
   kernel argument has its address taken (which would copy it into every
   wave's scratch frame). */
struct MonRef {
    const xtm_monitor *M;
    const xtm_particles *P;
};

/* Feature accumulators of particle slot i at the start of a turn (see
   xtm_features), from its coordinates between two tracking launches. */
XT_HD void feature_record(const xtm_features &F, const int64_t i, const double x,
                               const double px, const double y, const double py,
                               const int64_t at_turn) {
    if (at_turn < F.start_at_turn || at_turn >= F.stop_at_turn) return;
    const double d[4] = {x - F.co[0], px - F.co[1], y - F.co[2], py - F.co[3]};
    double z[4];
    for (int r = 0; r < 4; r++)
        z[r] = F.inv_w[4 * r] * d[0] + F.inv_w[4 * r + 1] * d[1] + F.inv_w[4 * r + 2] * d[2] +
               F.inv_w[4 * r + 3] * d[3];
    const double zx = z[0] * F.inv_sigma[0], zpx = z[1] * F.inv_sigma[0];
    const double zy = z[2] * F.inv_sigma[1], zpy = z[3] * F.inv_sigma[1];
    const double jx = 0.5 * (zx * zx + zpx * zpx), jy = 0.5 * (zy * zy + zpy * zpy);
    double *a = F.acc;
    const int64_t c = F.capacity;
    double &n = a[XTM_FEAT_COUNT * c + i];
    if (n == 0.0) {
        a[XTM_FEAT_JX0 * c + i] = jx;
        a[XTM_FEAT_JY0 * c + i] = jy;
        a[XTM_FEAT_JX_MAX * c + i] = jx;
        a[XTM_FEAT_JY_MAX * c + i] = jy;
    } else {
        /* rotation from the previous (z, zp) to this one: the phase
           advance for z = sqrt(2J) cos(phi), zp = -sqrt(2J) sin(phi) */
        const double qx = a[XTM_FEAT_ZX * c + i], qpx = a[XTM_FEAT_ZPX * c + i];
        const double qy = a[XTM_FEAT_ZY * c + i], qpy = a[XTM_FEAT_ZPY * c + i];
        const double twopi = 2.0 * PI;
        double px_ = xt_atan2(zx * qpx - zpx * qx, qx * zx + qpx * zpx);
        double py_ = xt_atan2(zy * qpy - zpy * qy, qy * zy + qpy * zpy);
        if (px_ < 0.0) px_ += twopi;
        if (py_ < 0.0) py_ += twopi;
        a[XTM_FEAT_PHX_SUM * c + i] += px_;
        a[XTM_FEAT_PHY_SUM * c + i] += py_;
        if (2 * at_turn >= F.start_at_turn + F.stop_at_turn) {
            a[XTM_FEAT_PHX_H2 * c + i] += px_;
            a[XTM_FEAT_PHY_H2 * c + i] += py_;
        }
        double &mx = a[XTM_FEAT_JX_MAX * c + i];
        double &my = a[XTM_FEAT_JY_MAX * c + i];
        if (jx > mx) mx = jx;
        if (jy > my) my = jy;
    }
    a[XTM_FEAT_JX_SUM * c + i] += jx;
    a[XTM_FEAT_JX_SQ * c + i] += jx * jx;
    a[XTM_FEAT_JY_SUM * c + i] += jy;
    a[XTM_FEAT_JY_SQ * c + i] += jy * jy;
    a[XTM_FEAT_ZX * c + i] = zx;
    a[XTM_FEAT_ZPX * c + i] = zpx;
    a[XTM_FEAT_ZY * c + i] = zy;
    a[XTM_FEAT_ZPY * c + i] = zpy;
    n = n + 1.0;
}

#define XT_MONITOR_RECORD(AT_ELEMENT)                                                     \
    monitor_record(*mon.M, *mon.P, i, p.x, p.px, p.y, p.py, p.zeta, p.delta(), p.rpp(),       \
                   p.rvv(), p.s(), p.chi, p.at_turn, p.state, (AT_ELEMENT), BT)

/* ---- one particle through [ele_start, ele_start+num_ele_track) for
   num_turns turns: the per-thread body of track_line (tracker.py:596-734)
   with the particle held in registers; monitor records only for active
   particles as in the reference CPU context (track.h:33-50).
   at_element: the reference adds 1 after every element the particle
   survives and sets 0 at the end of a turn; here at_element = base + e at
   element e, with base updated at the end of each turn.
   run(p, m, run.key(e), c) tracks element e: the program interpreter, or
   the line's generated code.
   CHI1: the particle's chi is exactly 1.0 and the element code sees the
   constant, so products with chi fold away (chi * k == k exactly) instead
   of being evaluated, or hoisted out of the loops as per-particle values
   that occupy registers for the whole launch. */
template <bool CHI1, bool BT, class Runner>
XT_HD void track_particle_loop(const xtm_particles &P, int64_t i, const Runner &run,
                               const TrackCtx &c, int32_t num_turns, int32_t ele_start,
                               int32_t num_ele_track, int32_t flag_end_turn_actions,
                               int32_t flag_reset_s, int32_t flag_monitor, const MonRef &mon) {
    PState<double> p;
#if defined(__HIP_DEVICE_COMPILE__)
    /* dynamic LDS of the launch, XT_LDS_BYTES_PER_THREAD bytes per thread
       of the block (a multiple of 64 threads): wave-major, each wave's
       XS_N slots of 64 consecutive doubles, so a slot access is one
       conflict-free 64-lane read at a constant offset, whatever the block
       size */
    extern __shared__ double slow_lds[];
    p.slow = slow_lds + (threadIdx.x >> 6) * (XS_N * 64) + (threadIdx.x & 63);
    p.slow_stride = 64;
#else
    double slow_buf[XS_N];
    p.slow = slow_buf;
    p.slow_stride = 1;
#endif
    p.x = P.x[i]; p.px = P.px[i]; p.y = P.y[i]; p.py = P.py[i];
    p.zeta = P.zeta[i];
    const double rvv0 = P.rvv[i];
    p.put(XS_DELTA, P.delta[i]);
    p.put(XS_RPP, P.rpp[i]);
    p.put(XS_RVV, rvv0);
    p.set_s(P.s[i]);
    p.chi = CHI1 ? 1.0 : P.chi[i];
    p.at_turn = P.at_turn[i]; p.state = P.state[i];
    p.put(XS_INV_RVV, 1. / rvv0);
    p.put(XS_ZETA0, p.zeta);
    refresh_flags(p);
    const PRef m{&P, i};
    const int64_t ele_stop = (int64_t)ele_start + num_ele_track;
    int64_t e_lost = -1;
    int64_t at_element;
    if (!BT) {
        int64_t base = P.at_element[i] - ele_start;
        for (int32_t iturn = 0; iturn < num_turns; iturn++) {
            if (flag_monitor == 1) XT_MONITOR_RECORD(base + ele_start);
            /* the next element's key (record index / program offset) is
               loaded one element ahead, so its scalar load overlaps the
               current element's arithmetic */
            typename Runner::Key key{};
            if (ele_start < ele_stop) key = run.key(ele_start);
            for (int64_t e = ele_start; e < ele_stop;) {
                /* a runner with segments tracks a run of n consecutive
                   elements in one call (not with element-by-element
                   records) and returns the position j in the run of the
                   element where the particle was lost */
                int64_t n = 1, j = 0;
                if constexpr (Runner::kSegments) {
                    const int32_t span = Runner::span(key);
                    if (span > 1 && flag_monitor != 2 && e + span <= ele_stop) n = span;
                }
                const typename Runner::Key next = run.key(e + n < ele_stop ? e + n : e);
                if (flag_monitor == 2) XT_MONITOR_RECORD(base + e);
                if constexpr (Runner::kSegments) {
                    if (n > 1) j = run.segment(p, m, key, c, e);
                    else run(p, m, key, c);
                } else {
                    run(p, m, key, c);
                }
                if (c.freeze_long()) p.zeta = p.get(XS_ZETA0);
                if (p.state <= 0) { e_lost = e + j; break; }
                e += n;
                key = next;
            }
            if (e_lost >= 0) break;
            if (flag_monitor == 2) XT_MONITOR_RECORD(base + ele_stop);
            if (flag_end_turn_actions > 0) {
                p.at_turn += 1;
                base = -(int64_t)ele_start;   /* at_element = 0 at the next turn's start */
                if (flag_reset_s > 0) p.set_s(0.);
            } else {
                base += num_ele_track;
            }
        }
        at_element = base + (e_lost >= 0 ? e_lost : (int64_t)ele_start);
    } else {
        /* backtracking (tracker.py:629-642, 699-731): elements in reverse
           order, at_element decreasing, the previous turn entered before
           the element loop (at_turn - 1, at_element = number of elements,
           s = line length); at element e, at_element = base + e */
        int64_t at_el = P.at_element[i];
        int64_t base = 0;
        for (int32_t iturn = 0; iturn < num_turns; iturn++) {
            if (flag_end_turn_actions > 0) {
                p.at_turn -= 1;
                at_el = c.num_ele_line;
                if (flag_reset_s > 0) p.set_s(c.line_length);
            }
            base = at_el - (ele_stop - 1);
            for (int64_t e = ele_stop - 1; e >= ele_start; e--) {
                if (flag_monitor == 2) XT_MONITOR_RECORD(base + e);
                run(p, m, run.key(e), c);
                if (c.freeze_long()) p.zeta = p.get(XS_ZETA0);
                if (p.state <= 0) { e_lost = e; break; }
            }
            if (e_lost >= 0) break;
            at_el = base + ele_start - 1;
            if (flag_monitor == 2) XT_MONITOR_RECORD(at_el);
            if (flag_monitor == 1) XT_MONITOR_RECORD(at_el);
        }
        at_element = e_lost >= 0 ? base + e_lost : at_el;
    }
    P.x[i] = p.x; P.px[i] = p.px; P.y[i] = p.y; P.py[i] = p.py;
    P.zeta[i] = p.zeta; P.delta[i] = p.delta();
    P.rpp[i] = p.rpp(); P.rvv[i] = p.rvv(); P.s[i] = p.s();
    P.at_element[i] = at_element;
    P.at_turn[i] = p.at_turn; P.state[i] = p.state;
}

/* the program interpreter as element runner */
struct ProgramRunner {
    const int64_t *__restrict__ entry;
    const double *__restrict__ prog;
    using Key = int64_t;
    static constexpr bool kSegments = false;
    XT_HD Key key(int64_t e) const { return cptr(entry)[e]; }
    XT_HD void operator()(PState<double> &p, const PRef &m, Key pc, const TrackCtx &c) const {
        run_element(p, m, cptr(prog), uniform(pc), c);
    }
};

/* run_chi1 serves particles with chi == 1.0 (track_particle_loop<true>),
   run_other all others.  With ALLOW_BT, c.backtrack() selects the
   backtracking loop (run_other, which must then execute the backtracking
   programs); without, the caller guarantees forward tracking. */
/* The loops for particles with chi != 1 and for backtracking, out of
   line: their code (the general interpreter) then takes no part in the
   register allocation of the kernel's chi == 1 loop. */
template <bool BT, class Runner>
XT_HD_COLD void track_particle_loop_cold(const xtm_particles &P, int64_t i, const Runner &run,
                                         const TrackCtx &c, int32_t num_turns,
                                         int32_t ele_start, int32_t num_ele_track,
                                         int32_t flag_end_turn_actions, int32_t flag_reset_s,
                                         int32_t flag_monitor, const MonRef &mon) {
    track_particle_loop<false, BT>(P, i, run, c, num_turns, ele_start, num_ele_track,
                                   flag_end_turn_actions, flag_reset_s, flag_monitor, mon);
}

template <bool ALLOW_BT, class Runner1, class RunnerOther>
XT_HD void track_particle_with(const xtm_particles &P, int64_t i, const Runner1 &run_chi1,
                               const RunnerOther &run_other, const TrackCtx &c,
                               int32_t num_turns, int32_t ele_start, int32_t num_ele_track,
                               int32_t flag_end_turn_actions, int32_t flag_reset_s,
                               int32_t flag_monitor, const MonRef &mon) {
    if (!(P.state[i] > 0)) return;
    if (ALLOW_BT && c.backtrack())
        track_particle_loop_cold<true>(P, i, run_other, c, num_turns, ele_start, num_ele_track,
                                       flag_end_turn_actions, flag_reset_s, flag_monitor, mon);
    else if (P.chi[i] == 1.0)
        track_particle_loop<true, false>(P, i, run_chi1, c, num_turns, ele_start,
                                         num_ele_track, flag_end_turn_actions, flag_reset_s,
                                         flag_monitor, mon);
    else
        track_particle_loop_cold<false>(P, i, run_other, c, num_turns, ele_start, num_ele_track,
                                        flag_end_turn_actions, flag_reset_s, flag_monitor, mon);
}

/* The compiled line kernel's body: forward tracking of the particles with
   chi == 1 only; the interpreter kernel, launched after it on the same
   stream with skip_chi1 set, tracks the others.  Keeping the interpreter
   out of this kernel keeps its register allocation and its scratch frame
   to those of the generated code. */
template <class Runner1>
XT_HD void track_particle_chi1(const xtm_particles &P, int64_t i, const Runner1 &run_chi1,
                               const TrackCtx &c, int32_t num_turns, int32_t ele_start,
                               int32_t num_ele_track, int32_t flag_end_turn_actions,
                               int32_t flag_reset_s, int32_t flag_monitor,
                               const MonRef &mon) {
    if (!(P.state[i] > 0) || !(P.chi[i] == 1.0)) return;
    track_particle_loop<true, false>(P, i, run_chi1, c, num_turns, ele_start, num_ele_track,
                                     flag_end_turn_actions, flag_reset_s, flag_monitor, mon);
}

/* entry: the forward or, when flags hold XS_FLAG_BACKTRACK, the
   backtracking programs; num_ele_line: elements in the line */
XT_HD void track_particle(const xtm_particles &P, int64_t i, const int64_t *__restrict__ entry,
                          const double *__restrict__ prog, double line_length,
                          int64_t num_ele_line, int32_t num_turns, int32_t ele_start,
                          int32_t num_ele_track, int32_t flag_end_turn_actions,
                          int32_t flag_reset_s, int32_t flag_monitor, const MonRef &mon,
                          uint64_t flags) {
    const TrackCtx c{flags, line_length, P.q0, P.t_sim, num_ele_line};
    const ProgramRunner run{entry, prog};
    track_particle_with<true>(P, i, run, run, c, num_turns, ele_start, num_ele_track,
                              flag_end_turn_actions, flag_reset_s, flag_monitor, mon);
}

#undef XT_MONITOR_RECORD

}  // namespace xt

#endif /* XT_OPS_H */
