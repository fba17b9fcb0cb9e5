/* xt_math.h -- FP64 elementary functions shared by the HIP tracking kernel,
 * the host-side element-table builder and the CPU restatement used in tests.
 *
 * Why a project library: the reference calls the platform libm (glibc on
 * CPU, NVRTC/OpenCL device libraries on GPU), e.g. sin/cos/sinh/cosh in
 * xtrack/beam_elements/elements_src/track_magnet_drift.h:119-147 and sin in
 * track_rf.h:60-62.  Chaotic orbits amplify 1-ulp differences, so bit-exact
 * loss bookkeeping between the GPU and a CPU run needs the *same* function
 * evaluated with the *same* IEEE operations on both sides.  Everything below
 * uses only +, -, *, /, sqrt, fma and rint (all correctly rounded / exact on
 * x86-64 and on gfx950), so a host build and a device build return identical
 * bits for identical inputs.  Compile with -ffp-contract=off.
 *
 * Methods (all published, none taken from a vendor library):
 *  - sin/cos: Cody & Waite argument reduction (Cody, Waite, "Software Manual
 *    for the Elementary Functions", 1980) with a four-part pi/2 and
 *    error-free transformations (Dekker TwoProduct via fma, Knuth TwoSum);
 *    Payne & Hanek reduction ("Radian reduction for trigonometric
 *    functions", SIGNUM 1983) for |x| >= 2^27 using a table of bits of 2/pi;
 *    Taylor polynomials on |r| <= pi/4 (coefficients (-1)^k/n! rounded to
 *    nearest), evaluated by Horner's rule with fma and a first-order
 *    correction for the low part of the reduced argument.
 *  - exp: Cody-Waite reduction by ln2 (two-part constant) and a degree-13
 *    Taylor polynomial on |r| <= ln2/2, then exact scaling by 2^k.
 *  - sinh/cosh: Taylor series for |x| < 1, (e^x -/+ e^-x)/2 otherwise.
 *  - tan: sin/cos from one shared reduction.
 * Constants (pi/2 split, 2/pi bits, ln2 split) were generated with mpmath at
 * 2000-bit precision.
 */
#ifndef XT_MATH_H
#define XT_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define XT_MATH_FN static inline __host__ __device__
#define XT_MATH_COLD static __host__ __device__ __attribute__((noinline))
#else
#define XT_MATH_FN static inline
#define XT_MATH_COLD static
#endif

#define XT_FMA(a, b, c) __builtin_fma((a), (b), (c))
#define XT_RINT(a) __builtin_rint(a)
#define XT_FABS(a) __builtin_fabs(a)
#define XT_SQRT(a) __builtin_sqrt(a)
#define XT_INF __builtin_inf()

/* pi/2 = XT_PIO2_1 + XT_PIO2_2 + XT_PIO2_3 (+ ~2^-164) */
#define XT_PIO2_1 0x1.921fb54442d18p+0
#define XT_PIO2_2 0x1.1a62633145c07p-54
#define XT_PIO2_3 (-0x1.f1976b7ed8fbcp-110)
#define XT_INV_PIO2 0x1.45f306dc9c883p-1
#define XT_PIO4 0x1.921fb54442d18p-1

#define XT_LN2_HI 0x1.62e42fee00000p-1   /* 32 significant bits */
#define XT_LN2_LO 0x1.a39ef35793c76p-33
#define XT_INV_LN2 0x1.71547652b82fep+0

/* ---- bit helpers ------------------------------------------------------ */
XT_MATH_FN uint64_t xt_dbits(double d) { uint64_t u; __builtin_memcpy(&u, &d, 8); return u; }
XT_MATH_FN double xt_bitsd(uint64_t u) { double d; __builtin_memcpy(&d, &u, 8); return d; }

/* 2^k for integer k in [-1022, 1023] */
XT_MATH_FN double xt_pow2i(int k) { return xt_bitsd((uint64_t)(int64_t)(k + 1023) << 52); }

/* 64x64 -> 128 unsigned product from 32-bit halves (portable) */
XT_MATH_FN void xt_mul64(uint64_t a, uint64_t b, uint64_t *hi, uint64_t *lo) {
    uint64_t a0 = a & 0xffffffffu, a1 = a >> 32, b0 = b & 0xffffffffu, b1 = b >> 32;
    uint64_t p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0, p11 = a1 * b1;
    uint64_t mid = (p00 >> 32) + (p01 & 0xffffffffu) + (p10 & 0xffffffffu);
    *lo = (mid << 32) | (p00 & 0xffffffffu);
    *hi = p11 + (p01 >> 32) + (p10 >> 32) + (mid >> 32);
}

/* bits of 2/pi after the binary point, 64 per word, most significant first */
#define XT_TWO_OVER_PI_NWORDS 24
XT_MATH_FN uint64_t xt_two_over_pi_word(int i) {
    /* table returned through a switch so the same code serves host and
       device without a device-side global array */
    switch (i) {
    case 0: return 0xa2f9836e4e441529ULL;  case 1: return 0xfc2757d1f534ddc0ULL;
    case 2: return 0xdb6295993c439041ULL;  case 3: return 0xfe5163abdebbc561ULL;
    case 4: return 0xb7246e3a424dd2e0ULL;  case 5: return 0x06492eea09d1921cULL;
    case 6: return 0xfe1deb1cb129a73eULL;  case 7: return 0xe88235f52ebb4484ULL;
    case 8: return 0xe99c7026b45f7e41ULL;  case 9: return 0x3991d639835339f4ULL;
    case 10: return 0x9c845f8bbdf9283bULL; case 11: return 0x1ff897ffde05980fULL;
    case 12: return 0xef2f118b5a0a6d1fULL; case 13: return 0x6d367ecf27cb09b7ULL;
    case 14: return 0x4f463f669e5fea2dULL; case 15: return 0x7527bac7ebe5f17bULL;
    case 16: return 0x3d0739f78a5292eaULL; case 17: return 0x6bfb5fb11f8d5d08ULL;
    case 18: return 0x56033046fc7b6babULL; case 19: return 0xf0cfbc209af4361dULL;
    case 20: return 0xa9e391615ee61b08ULL; case 21: return 0x6599855f14a06840ULL;
    case 22: return 0x8dffd8804d732731ULL; case 23: return 0x06061556ca73a8c9ULL;
    default: return 0;
    }
}

/* bits [pos, pos+64) of 2/pi (bit 0 = weight 2^-1) */
XT_MATH_FN uint64_t xt_two_over_pi_bits(int pos) {
    int w = pos >> 6, b = pos & 63;
    uint64_t hi = xt_two_over_pi_word(w);
    if (b == 0) return hi;
    return (hi << b) | (xt_two_over_pi_word(w + 1) >> (64 - b));
}

/* Payne-Hanek: x finite, |x| >= 2^27.  Returns quadrant, r = rh + rl with
   x = (4n + q) pi/2 + r, |r| <= pi/4. */
XT_MATH_COLD int xt_rem_pio2_large(double x, double *rh, double *rl) {
    uint64_t ux = xt_dbits(x);
    int neg = (int)(ux >> 63);
    int ex = (int)((ux >> 52) & 0x7ff);
    uint64_t m = (ux & 0xfffffffffffffULL) | 0x10000000000000ULL;
    int E = ex - 1075;                 /* x = m * 2^E */
    int i0 = E - 1;                    /* first needed bit of 2/pi (1-based) */
    if (i0 < 1) i0 = 1;
    /* window W = bits i0 .. i0+191 as three 64-bit words */
    uint64_t w0 = xt_two_over_pi_bits(i0 - 1);
    uint64_t w1 = xt_two_over_pi_bits(i0 - 1 + 64);
    uint64_t w2 = xt_two_over_pi_bits(i0 - 1 + 128);
    /* P = m * W  (256 bits: p3 p2 p1 p0) */
    uint64_t h0, l0, h1, l1, h2, l2;
    xt_mul64(m, w2, &h2, &l2);
    xt_mul64(m, w1, &h1, &l1);
    xt_mul64(m, w0, &h0, &l0);
    uint64_t p0 = l2;
    uint64_t p1 = h2 + l1; uint64_t c1 = (p1 < h2);
    uint64_t t2 = h1 + l0; uint64_t c2 = (t2 < h1);
    uint64_t p2 = t2 + c1; c2 += (p2 < t2);
    uint64_t p3 = h0 + c2;
    /* value y = P * 2^-(s) with s = i0 + 191 - E; y mod 4 and its fraction */
    int s = i0 + 191 - E;              /* 190 <= s <= 217 */
    /* shift P right by (s - 128) so the binary point sits between words 1|0 of Q:
       Q = P >> (s-128): Q.hi64 = integer+top fraction bits */
    int sh = s - 128;                  /* 62 .. 89 */
    uint64_t q0, q1, q2;               /* Q = q2:q1:q0, binary point after q1's top? */
    if (sh < 64) {
        q0 = (p0 >> sh) | (p1 << (64 - sh));
        q1 = (p1 >> sh) | (p2 << (64 - sh));
        q2 = (p2 >> sh) | (p3 << (64 - sh));
    } else if (sh == 64) {
        q0 = p1; q1 = p2; q2 = p3;
    } else {
        int t = sh - 64;
        q0 = (p1 >> t) | (p2 << (64 - t));
        q1 = (p2 >> t) | (p3 << (64 - t));
        q2 = (p3 >> t);
    }
    /* now y = Q * 2^-128: integer part = q2, fraction = q1:q0 */
    int quad = (int)(q2 & 3);
    uint64_t f1 = q1, f0 = q0;
    int fneg = 0;
    if (f1 >> 63) {                    /* fraction >= 1/2: use f - 1 */
        quad = (quad + 1) & 3;
        f0 = ~f0 + 1; f1 = ~f1 + (f0 == 0);
        fneg = 1;
    }
    /* fraction (<= 1/2) as three exact chunks: 52 + 52 + 24 bits */
    double c1d = (double)(f1 >> 12) * 0x1p-52;
    double c2d = (double)(((f1 & 0xfffULL) << 40) | (f0 >> 24)) * 0x1p-104;
    double c3d = (double)(f0 & 0xffffffULL) * 0x1p-128;
    double fh = c1d + c2d;
    double fl = (c1d - fh) + c2d + c3d;
    /* r = f * pi/2 in double-double */
    double rhh = fh * XT_PIO2_1;
    double rll = XT_FMA(fh, XT_PIO2_1, -rhh) + (fh * XT_PIO2_2 + fl * XT_PIO2_1);
    double h = rhh + rll;
    double l = rll - (h - rhh);
    if (fneg) { h = -h; l = -l; }
    if (neg) { h = -h; l = -l; quad = (4 - quad) & 3; }
    *rh = h; *rl = l;
    return quad;
}

/* x = (4n + q) pi/2 + (rh + rl), |rh + rl| <= pi/4 (1 + 2^-50); q returned */
XT_MATH_FN int xt_rem_pio2(double x, double *rh, double *rl) {
    double ax = XT_FABS(x);
    if (ax <= XT_PIO4) { *rh = x; *rl = 0.0; return 0; }
    if (ax < 0x1p27) {
        double k = XT_RINT(x * XT_INV_PIO2);
        double r1 = XT_FMA(-k, XT_PIO2_1, x);      /* exact */
        double p2 = k * XT_PIO2_2;
        double e2 = XT_FMA(k, XT_PIO2_2, -p2);     /* exact error of p2 */
        double h = r1 - p2;                         /* TwoSum(r1, -p2) */
        double bb = h - r1;
        double l = (r1 - (h - bb)) + (-p2 - bb);
        l = l - e2 - k * XT_PIO2_3;
        double rh_ = h + l;
        *rl = l - (rh_ - h);
        *rh = rh_;
        int64_t ki = (int64_t)k;
        return (int)(ki & 3);
    }
    return xt_rem_pio2_large(x, rh, rl);
}

/* 1/n! rounded to nearest double (generated with fractions.Fraction) */
#define XT_IF2 0x1.0000000000000p-1
#define XT_IF3 0x1.5555555555555p-3
#define XT_IF4 0x1.5555555555555p-5
#define XT_IF5 0x1.1111111111111p-7
#define XT_IF6 0x1.6c16c16c16c17p-10
#define XT_IF7 0x1.a01a01a01a01ap-13
#define XT_IF8 0x1.a01a01a01a01ap-16
#define XT_IF9 0x1.71de3a556c734p-19
#define XT_IF10 0x1.27e4fb7789f5cp-22
#define XT_IF11 0x1.ae64567f544e4p-26
#define XT_IF12 0x1.1eed8eff8d898p-29
#define XT_IF13 0x1.6124613a86d09p-33
#define XT_IF14 0x1.93974a8c07c9dp-37
#define XT_IF15 0x1.ae7f3e733b81fp-41
#define XT_IF16 0x1.ae7f3e733b81fp-45
#define XT_IF17 0x1.952c77030ad4ap-49
#define XT_IF18 0x1.6827863b97d97p-53
#define XT_IF19 0x1.2f49b46814157p-57
#define XT_IF20 0x1.e542ba4020225p-62
#define XT_IF21 0x1.71b8ef6dcf572p-66
#define XT_IF22 0x1.0ce396db7f853p-70
#define XT_IF23 0x1.761b41316381ap-75

/* sin(rh + rl), |rh| <= pi/4 (1+eps), |rl| <= ulp(rh).
   sin(rh+rl) = rh + rh^3 S(z) + rl (1 - z/2) + O(rl^2), z = rh^2,
   S(z) = -1/3! + z/5! - ... + z^8/19!  (truncation < 2^-70 relative) */
XT_MATH_FN double xt_kernel_sin(double rh, double rl) {
    double z = rh * rh;
    double p = -XT_IF19;
    p = XT_FMA(p, z, XT_IF17);
    p = XT_FMA(p, z, -XT_IF15);
    p = XT_FMA(p, z, XT_IF13);
    p = XT_FMA(p, z, -XT_IF11);
    p = XT_FMA(p, z, XT_IF9);
    p = XT_FMA(p, z, -XT_IF7);
    p = XT_FMA(p, z, XT_IF5);
    p = XT_FMA(p, z, -XT_IF3);
    double corr = XT_FMA(rh * z, p, rl - 0.5 * z * rl);
    return rh + corr;
}

/* cos(rh + rl) = 1 - z/2 + z^2 C(z) - rh rl + O(rl^2),
   C(z) = 1/4! - z/6! + ... + z^9/22! */
XT_MATH_FN double xt_kernel_cos(double rh, double rl) {
    double z = rh * rh;
    double p = XT_IF22;
    p = XT_FMA(p, z, -XT_IF20);
    p = XT_FMA(p, z, XT_IF18);
    p = XT_FMA(p, z, -XT_IF16);
    p = XT_FMA(p, z, XT_IF14);
    p = XT_FMA(p, z, -XT_IF12);
    p = XT_FMA(p, z, XT_IF10);
    p = XT_FMA(p, z, -XT_IF8);
    p = XT_FMA(p, z, XT_IF6);
    p = XT_FMA(p, z, -XT_IF4);
    /* w = 1 - z/2 with its rounding error recovered (z/2 <= 0.31) */
    double hz = 0.5 * z;
    double w = 1.0 - hz;
    double werr = (1.0 - w) - hz;
    double tail = XT_FMA(z * z, -p, werr - rh * rl);
    /* note: C(z) above has the opposite sign convention; p = -C(z) */
    return w + tail;
}

/* sin and cos of x from one reduction */
XT_MATH_FN void xt_sincos(double x, double *s, double *c) {
    double rh, rl;
    if (!(XT_FABS(x) < XT_INF)) { *s = x - x; *c = x - x; return; }
    int q = xt_rem_pio2(x, &rh, &rl);
    double sr = xt_kernel_sin(rh, rl);
    double cr = xt_kernel_cos(rh, rl);
    switch (q) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
    }
}

XT_MATH_FN double xt_sin(double x) { double s, c; xt_sincos(x, &s, &c); return s; }
XT_MATH_FN double xt_cos(double x) { double s, c; xt_sincos(x, &s, &c); return c; }
XT_MATH_FN double xt_tan(double x) { double s, c; xt_sincos(x, &s, &c); return s / c; }

/* exp(x): x = k ln2 + r, |r| <= ln2/2; e^r by Taylor to r^13 */
XT_MATH_FN double xt_exp(double x) {
    if (x != x) return x;
    if (x > 0x1.62e42fefa39efp+9) return XT_INF;          /* > ln(DBL_MAX) */
    if (x < -0x1.74910d52d3051p+9) return 0.0;            /* < ln(2^-1075) */
    double k = XT_RINT(x * XT_INV_LN2);
    double r = (x - k * XT_LN2_HI) - k * XT_LN2_LO;       /* k*LN2_HI exact */
    double p = XT_IF13;
    p = XT_FMA(p, r, XT_IF12);
    p = XT_FMA(p, r, XT_IF11);
    p = XT_FMA(p, r, XT_IF10);
    p = XT_FMA(p, r, XT_IF9);
    p = XT_FMA(p, r, XT_IF8);
    p = XT_FMA(p, r, XT_IF7);
    p = XT_FMA(p, r, XT_IF6);
    p = XT_FMA(p, r, XT_IF5);
    p = XT_FMA(p, r, XT_IF4);
    p = XT_FMA(p, r, XT_IF3);
    p = XT_FMA(p, r, XT_IF2);
    double er = 1.0 + XT_FMA(r * r, p, r);
    int ki = (int)k;
    if (ki > 1023) return er * 2.0 * xt_pow2i(ki - 1);
    if (ki < -1021) return er * xt_pow2i(ki + 1000) * 0x1p-1000;
    return er * xt_pow2i(ki);
}

/* sinh/cosh of |x| < 1 by Taylor series (truncation < 2^-70 relative) */
XT_MATH_FN double xt_sinh_small(double x) {
    double z = x * x;
    double p = XT_IF21;
    p = XT_FMA(p, z, XT_IF19);
    p = XT_FMA(p, z, XT_IF17);
    p = XT_FMA(p, z, XT_IF15);
    p = XT_FMA(p, z, XT_IF13);
    p = XT_FMA(p, z, XT_IF11);
    p = XT_FMA(p, z, XT_IF9);
    p = XT_FMA(p, z, XT_IF7);
    p = XT_FMA(p, z, XT_IF5);
    p = XT_FMA(p, z, XT_IF3);
    return XT_FMA(x * z, p, x);
}

XT_MATH_FN double xt_cosh_small(double x) {
    double z = x * x;
    double p = XT_IF22;
    p = XT_FMA(p, z, XT_IF20);
    p = XT_FMA(p, z, XT_IF18);
    p = XT_FMA(p, z, XT_IF16);
    p = XT_FMA(p, z, XT_IF14);
    p = XT_FMA(p, z, XT_IF12);
    p = XT_FMA(p, z, XT_IF10);
    p = XT_FMA(p, z, XT_IF8);
    p = XT_FMA(p, z, XT_IF6);
    p = XT_FMA(p, z, XT_IF4);
    p = XT_FMA(p, z, XT_IF2);
    return XT_FMA(z, p, 1.0);
}

XT_MATH_FN void xt_sinhcosh(double x, double *sh, double *ch) {
    double ax = XT_FABS(x);
    if (ax < 1.0) { *sh = xt_sinh_small(x); *ch = xt_cosh_small(x); return; }
    if (!(ax == ax)) { *sh = x; *ch = x; return; }
    double e, ei;
    if (ax < 709.0) {
        e = xt_exp(ax);
        ei = 1.0 / e;
        double s = 0.5 * (e - ei);
        *sh = (x < 0) ? -s : s;
        *ch = 0.5 * (e + ei);
    } else {
        double eh = xt_exp(0.5 * ax);
        double v = (0.5 * eh) * eh;
        *sh = (x < 0) ? -v : v;
        *ch = v;
    }
}

XT_MATH_FN double xt_sinh(double x) { double s, c; xt_sinhcosh(x, &s, &c); return s; }
XT_MATH_FN double xt_cosh(double x) { double s, c; xt_sinhcosh(x, &s, &c); return c; }

/* ---- atan2 and asin ------------------------------------------------------
   atan(u) for |u| <= 1: u > tan(pi/8) maps to pi/4 + atan((u - 1)/(u + 1)),
   one argument halving u -> u / (1 + sqrt(1 + u^2)) (atan u = 2 atan u')
   leaves |u'| <= 0.2, then the Taylor series sum (-1)^k u'^(2k+1)/(2k+1)
   with 12 terms (truncation below 2^-60 relative); coefficients are the
   rationals 1/(2k+1) rounded by the compiler.  atan2 folds the quadrants
   with pi and pi/2 as two-double constants; asin(z) = atan2(z,
   sqrt((1 - z)(1 + z))).  Used per particle by the exact bend maps. */
XT_MATH_FN double xt_atan_small(double u) {      /* |u| <= 0.2 */
    const double u2 = u * u;
    double r = -1.0 / 23.0;
    r = XT_FMA(r, u2, 1.0 / 21.0);
    r = XT_FMA(r, u2, -1.0 / 19.0);
    r = XT_FMA(r, u2, 1.0 / 17.0);
    r = XT_FMA(r, u2, -1.0 / 15.0);
    r = XT_FMA(r, u2, 1.0 / 13.0);
    r = XT_FMA(r, u2, -1.0 / 11.0);
    r = XT_FMA(r, u2, 1.0 / 9.0);
    r = XT_FMA(r, u2, -1.0 / 7.0);
    r = XT_FMA(r, u2, 1.0 / 5.0);
    r = XT_FMA(r, u2, -1.0 / 3.0);
    return XT_FMA(r * u2, u, u);
}

/* atan(u) for 0 <= u <= 1, as hi + lo */
XT_MATH_FN void xt_atan_unit(double u, double *hi, double *lo) {
    const double pio4_hi = 0x1.921fb54442d18p-1, pio4_lo = 0x1.1a62633145c07p-55;
    double base_hi = 0.0, base_lo = 0.0;
    if (u > 0x1.a827999fcef32p-2) {      /* tan(pi/8) */
        u = (u - 1.0) / (u + 1.0);
        base_hi = pio4_hi;
        base_lo = pio4_lo;
    }
    const double h = u / (1.0 + XT_SQRT(XT_FMA(u, u, 1.0)));
    const double a = 2.0 * xt_atan_small(h);
    *hi = base_hi + a;
    *lo = (base_hi - *hi) + a + base_lo;
}

XT_MATH_FN double xt_atan2(double y, double x) {
    const double pi_hi = 0x1.921fb54442d18p+1, pi_lo = 0x1.1a62633145c07p-53;
    const double pio2_hi = 0x1.921fb54442d18p+0, pio2_lo = 0x1.1a62633145c07p-54;
    if (y != y || x != x) return x + y;
    const double ax = x < 0 ? -x : x, ay = y < 0 ? -y : y;
    const int neg_y = (int)(xt_dbits(y) >> 63), neg_x = (int)(xt_dbits(x) >> 63);
    double r;
    if (ay == 0.0) {
        r = neg_x ? pi_hi : 0.0;
    } else if (ax == 0.0) {
        r = pio2_hi;
    } else if (ax == __builtin_inf() && ay == __builtin_inf()) {
        r = neg_x ? 3 * (pio2_hi * 0.5) : pio2_hi * 0.5;
    } else if (ax == __builtin_inf()) {
        r = neg_x ? pi_hi : 0.0;
    } else if (ay == __builtin_inf()) {
        r = pio2_hi;
    } else {
        double hi, lo;
        if (ay <= ax) {
            xt_atan_unit(ay / ax, &hi, &lo);
            if (neg_x) { const double t = pi_hi - hi; lo = ((pi_hi - t) - hi) + pi_lo - lo; hi = t; }
        } else {
            xt_atan_unit(ax / ay, &hi, &lo);
            if (neg_x) { const double t = pio2_hi + hi; lo = ((pio2_hi - t) + hi) + pio2_lo + lo; hi = t; }
            else { const double t = pio2_hi - hi; lo = ((pio2_hi - t) - hi) + pio2_lo - lo; hi = t; }
        }
        r = hi + lo;
    }
    return neg_y ? -r : r;
}

XT_MATH_FN double xt_asin(double z) {
    return xt_atan2(z, XT_SQRT((1.0 - z) * (1.0 + z)));
}

#endif /* XT_MATH_H */
