/* xt_program.h -- device representation of a line (shared by the host
 * builder and the HIP kernel).
 *
 * A line is an array `entry[num_elements]` of offsets into one FP64 array
 * `prog`.  At each offset starts the op list of that element: an op word
 * (an integer stored in the bits of a double) followed by the op's
 * constants, repeated until OP_END.  Op words and constants are read with
 * wave-uniform addresses (every lane of a wave tracks the same element), so
 * they are scalar loads and the op switch is a scalar branch.
 *
 * The constants are the element-uniform sub-expressions of the reference's
 * element code (e.g. cos(h*s), 1/h in the polar drift of
 * track_magnet_drift.h:45-87, the 1/n! chain of the multipole kick in
 * track_magnet_kick.h:183-228), evaluated on the host with the same
 * operations and the same elementary-function library as the expression
 * would have on the device, so the per-particle arithmetic that remains in
 * the kernel produces the same bits as the unhoisted expression.
 */
#ifndef XT_PROGRAM_H
#define XT_PROGRAM_H

#include <stdint.h>

enum xt_op {
    OP_END = 0,
    OP_DRIFT_EXP = 1,      /* L                                               */
    OP_DRIFT_EXACT = 2,    /* L                                               */
    OP_POLAR = 3,          /* L rho ca sa t2 ica  (t2 = 2*sa2*sa2, ica = 1/ca) */
    OP_DIPQUAD = 4,        /* L k0 k1 h                                        */
    OP_YOSHIDA4_BEND = 5,  /* L rho c1k0 c2k0 | a ca sa t2 ica | b ...        */
    OP_YOSHIDA6_BEND = 6,  /* L rho kc0..kc3 | 4 x (s ca sa t2 ica)           */
    OP_KICK = 7,           /* w desc mask hl g0 h1                            */
    OP_EDGE_AXAY = 8,      /* ca cb   (ax = ca*(y - 0), ay = cb*(x - 0))       */
    OP_EDGE_CLEAR_AXAY = 9,/*                                                  */
    OP_EDGE_LINEAR = 10,   /* r21 r43                                          */
    OP_RF_KICK = 11,       /* V freq harmonic lag phase abs_time              */
    OP_LIMIT_RECT = 12,    /* min_x max_x min_y max_y                          */
    OP_LIMIT_ELLIPSE = 13, /* a_squ b_squ a_b_squ                              */
    OP_GLOBAL_APER = 14,   /*                                                  */
    OP_XY_SHIFT = 15,      /* ax ay   (x = x + ax, y = y + ay)                 */
    OP_S_SHIFT = 16,       /* ds      (exact drift ds, zeta and s back by ds)  */
    OP_X_ROT = 17,         /* sin cos tan as XRotation_single_particle takes   */
    OP_Y_ROT = 18,         /* sin cos tan as YRotation_single_particle takes   */
    OP_S_ROT = 19,         /* sin cos                                          */
    OP_RF_MULT = 20,       /* V freq lag phase desc  (RF multipole kick)       */
    OP_BEND_CURVED = 21,   /* L k0 1/h hs sin(hs) cos(hs) sin(hs/2)^2 t2 ica   */
    OP_BEND_STRAIGHT = 22, /* L k0                                              */
    OP_NUM = 23
};

/* number of constants following each op word */
#define XT_NARGS_KICK 6
#define XT_OP_NARGS_INIT {0, 1, 1, 6, 4, 14, 26, XT_NARGS_KICK, 2, 0, 2, 6, 4, 3, 0, 2, 1, 3, 3, 2, 5, 9, 2}

/* RF multipole descriptor, referenced by OP_RF_MULT through an absolute
 * offset into prog: the order (bits), then for k = 0..order the six
 * element-uniform terms of track_rf.h:96-118: DEG2RAD*pn[k], phase_n[k],
 * DEG2RAD*ps[k], phase_s[k], factor*knl[k]/k!, factor*ksl[k]/k! (factor =
 * +1, or -1 when backtracking). */
#define RFD_ORDER 0
#define RFD_TERMS 1
#define RFD_STRIDE 6

/* Multipole-kick descriptor (track_magnet_kick.h:24-144), referenced by
 * OP_KICK through an absolute offset into prog:
 *   KD_ORDER, KD_ORDER_REL       integers (bits)
 *   KD_LENGTH                    core length
 *   KD_FACTOR, KD_FACTOR_REL     factor_knl_ksl, factor_knl_ksl*rel_ref_strength
 *   KD_ROT_FRAME                 integer
 *   KD_H, KD_HXL, KD_K0HC, KD_K1HC
 *   KD_HTOT                      h + hxl/length (length != 0) else h
 *   KD_K0L, KD_K1L               k0l_mult, k1l_mult
 *   KD_MAIN  .. +3               knl_main = {k0,k1,k2,k3}_kick * length
 *   KD_MAINS .. +3               ksl_main
 *   KD_ARR: pass 1: inv-factorial chain f[order..0], knl[0..order], ksl[0..order]
 *           pass 2: f_rel[order_rel..0], knl_rel, ksl_rel
 *           pass 3: f_main[3..0]
 * The chain f[k] is the value inv_factorial holds when coefficient k is
 * used: f[order] = inv_factorial_order, f[k-1] = f[k] * k.
 *
 * Fast blocks (one per pass, offsets in KD_FASTOFF..+2): [top, a[0..top],
 * b[0..top]] with a[k] = (knl[k]*factor)*f[k], b[k] = (ksl[k]*factor)*f[k]
 * (the literal this_knl*inv_factorial for chi == 1) and top the highest
 * index with a non-zero coefficient.
 *
 * OP_KICK args: w, desc offset, mask of present folded terms (KF_* in
 * xt_ops.h), hl = h*length*w + hxl*w, g0 = ((-k0c)*w)*htot,
 * h1 = (htot*k1c)*w with k0c = k0_h_correction*length + k0l_mult and
 * k1c = k1_h_correction*length + k1l_mult. */
#define KD_ORDER 0
#define KD_ORDER_REL 1
#define KD_LENGTH 2
#define KD_FACTOR 3
#define KD_FACTOR_REL 4
#define KD_ROT_FRAME 5
#define KD_H 6
#define KD_HXL 7
#define KD_K0HC 8
#define KD_K1HC 9
#define KD_HTOT 10
#define KD_K0L 11
#define KD_K1L 12
#define KD_MAIN 13
#define KD_MAINS 17
#define KD_FASTOFF 21          /* three absolute offsets of the fast blocks    */
#define KD_ARR 24

#endif /* XT_PROGRAM_H */
