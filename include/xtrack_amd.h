/* xtrack_amd.h -- C ABI of the MI355X multi-turn tracker (libxtrack_amd.so).
 *
 * The library replaces the generated `track_line` kernel of the reference
 * (xtrack/tracker.py:546-564, compiled and launched by xobjects) with a
 * persistent HIP kernel for gfx950.  Element parameters arrive as a raw
 * element table whose per-type field layout mirrors the reference's
 * element xofields (files under xtrack/beam_elements/); particle data arrive as the
 * reference's SoA Particles layout (xtrack/particles/particles.py:26-83)
 * given as plain device pointers.
 *
 * Every function returns 0 on success and a negative status on error;
 * xtm_last_error() returns a message for the calling thread's last error.
 * Nothing in this header depends on torch or on HIP types (streams are
 * passed as void*; NULL means the default stream).
 */
#ifndef XTRACK_AMD_H
#define XTRACK_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XTM_ABI_VERSION 1

/* ---- status codes -------------------------------------------------------- */
#define XTM_OK 0
#define XTM_ERR_ARG (-1)        /* invalid argument                          */
#define XTM_ERR_UNSUPPORTED (-2)/* element/model/flag not supported          */
#define XTM_ERR_HIP (-3)        /* HIP runtime error                         */
#define XTM_ERR_NOMEM (-4)      /* allocation failed                         */

/* ---- element type ids (raw table) -----------------------------------------
 * Statically thick classes (those whose class declares isthick = True:
 * Drift, Quadrupole, Bend, Sextupole, Octupole, Cavity) get the global
 * |x|,|y| <= 1 m check after tracking (xtrack/tracker.py:681-689). */
#define XTM_MARKER 0
#define XTM_DRIFT 1
#define XTM_MULTIPOLE 2
#define XTM_QUADRUPOLE 3
#define XTM_BEND 4
#define XTM_SEXTUPOLE 5
#define XTM_OCTUPOLE 6
#define XTM_CAVITY 7
#define XTM_LIMITRECT 8
#define XTM_LIMITELLIPSE 9
#define XTM_DIPOLEEDGE 10
#define XTM_TRANSFORMED 11
#define XTM_RFMULTIPOLE 12
#define XTM_SLICE 13
#define XTM_NUM_TYPES 14

/* ---- raw parameter layout ---------------------------------------------------
 * Element i occupies params[offsets[i] ...].  Integer fields are stored as
 * doubles (exact for |v| < 2^53).
 *
 * DRIFT (xtrack/beam_elements/drift.py:31-34) */
#define XTM_D_LENGTH 0
#define XTM_D_MODEL 1
#define XTM_D_SIZE 2

/* Magnets: MULTIPOLE, QUADRUPOLE, BEND, SEXTUPOLE, OCTUPOLE share one layout.
 * Fields a class does not have are 0.  Field meaning follows
 * xtrack/beam_elements/{multipole,quadrupole,bend,sextupole,octupole}.py and
 * _common.py:559-600.  After XTM_M_ARRAYS come knl[order+1], ksl[order+1],
 * knl_rel[nrel], ksl_rel[nrel]. */
#define XTM_M_LENGTH 0
#define XTM_M_ORDER 1
#define XTM_M_INV_FACTORIAL_ORDER 2
#define XTM_M_NREL 3            /* len(knl_rel) == len(ksl_rel) >= 1          */
#define XTM_M_MODEL 4
#define XTM_M_INTEGRATOR 5
#define XTM_M_NUM_MULTIPOLE_KICKS 6
#define XTM_M_MAIN_IS_SKEW 7
#define XTM_M_MAIN_ORDER 8      /* Multipole only                             */
#define XTM_M_ISTHICK 9         /* Multipole only                             */
#define XTM_M_HXL 10            /* Multipole only                             */
#define XTM_M_K0 11
#define XTM_M_K1 12
#define XTM_M_K2 13
#define XTM_M_K3 14
#define XTM_M_K0S 15
#define XTM_M_K1S 16
#define XTM_M_K2S 17
#define XTM_M_K3S 18
#define XTM_M_H 19              /* Bend only (angle / length)                 */
#define XTM_M_EDGE_ENTRY_ACTIVE 20
#define XTM_M_EDGE_EXIT_ACTIVE 21
#define XTM_M_EDGE_ENTRY_MODEL 22
#define XTM_M_EDGE_EXIT_MODEL 23
#define XTM_M_EDGE_ENTRY_ANGLE 24
#define XTM_M_EDGE_EXIT_ANGLE 25
#define XTM_M_EDGE_ENTRY_ANGLE_FDOWN 26
#define XTM_M_EDGE_EXIT_ANGLE_FDOWN 27
#define XTM_M_EDGE_ENTRY_FINT 28
#define XTM_M_EDGE_EXIT_FINT 29
#define XTM_M_EDGE_ENTRY_HGAP 30
#define XTM_M_EDGE_EXIT_HGAP 31
#define XTM_M_DELTA_TAPER 32
#define XTM_M_RADIATION_FLAG 33
#define XTM_M_ARRAYS 34

/* CAVITY (xtrack/beam_elements/cavity.py:63-76) */
#define XTM_C_LENGTH 0
#define XTM_C_VOLTAGE 1
#define XTM_C_FREQUENCY 2
#define XTM_C_HARMONIC 3
#define XTM_C_LAG 4
#define XTM_C_PHASE 5
#define XTM_C_ABSOLUTE_TIME 6
#define XTM_C_NUM_KICKS 7
#define XTM_C_MODEL 8
#define XTM_C_INTEGRATOR 9
#define XTM_C_LAG_TAPER 10
#define XTM_C_PHASE_TAPER 11
#define XTM_C_SIZE 12

/* LIMITRECT (limit_rect.py), LIMITELLIPSE (limit_ellipse.py:25-29) */
#define XTM_LR_MIN_X 0
#define XTM_LR_MAX_X 1
#define XTM_LR_MIN_Y 2
#define XTM_LR_MAX_Y 3
#define XTM_LR_SIZE 4
#define XTM_LE_A_SQU 0
#define XTM_LE_B_SQU 1
#define XTM_LE_A_B_SQU 2
#define XTM_LE_SIZE 3

/* DIPOLEEDGE (dipole_edge.py:38-49): r21, r43 as the class derives them
 * from k, e1, e1_fd, hgap, fint; model 0 linear, 1 full, -1 suppressed */
#define XTM_DE_R21 0
#define XTM_DE_R43 1
#define XTM_DE_MODEL 2
#define XTM_DE_DELTA_TAPER 3
#define XTM_DE_SIZE 4

/* RFMultipole (rf_multipole.py, rfmultipole.h -> track_rf.h:18-167 with
 * model -1, one kick): voltage, frequency, lag (deg), phase (rad), order,
 * then knl, ksl, pn (deg), ps (deg), phase_n, phase_s, each order+1 long. */
#define XTM_RFM_VOLTAGE 0
#define XTM_RFM_FREQUENCY 1
#define XTM_RFM_LAG 2
#define XTM_RFM_PHASE 3
#define XTM_RFM_ORDER 4
#define XTM_RFM_ARRAYS 5

/* SLICE: a slice of a thick element made by Line.slice_thick_elements
 * (beam_elements/slice_elements_{thin,thick,drift,edge}.py and the
 * elements_src/<kind>_slice_<class>.h): the slice weight, the kind (XTM_SLICE_*), the
 * parent's type, then the parent's record (without its misalignment,
 * which wraps the slice instead).  Thin: the parent's magnet kick only
 * (model -1) with the weight; thick: the parent's body with the weight;
 * drift: an expanded drift of weight x parent length; entry/exit: the
 * parent's entry/exit edge only.  Parents: Multipole (thick), Quadrupole,
 * Bend, Sextupole, Octupole; drift slices also Drift. */
#define XTM_SL_WEIGHT 0
#define XTM_SL_KIND 1
#define XTM_SL_PARENT_TYPE 2
#define XTM_SL_PARENT 3
#define XTM_SLICE_THIN 0
#define XTM_SLICE_THICK 1
#define XTM_SLICE_DRIFT 2
#define XTM_SLICE_ENTRY 3
#define XTM_SLICE_EXIT 4

/* TRANSFORMED: an element with non-zero misalignment fields, wrapped as in
 * track_local_particle_with_transformations.h:88-168 with the frame
 * transforms of track_misalignments.h (straight :38-112, curved :115-386).
 * Record: the element's shift_x, shift_y, shift_s, rot_x_rad, rot_y_rad,
 * rot_s_rad_no_frame, rot_shift_anchor, rot_s_rad; its length (0 for
 * classes that are not thick), angle and h; 1 for classes that curve the
 * reference frame (Bend), else 0; the inner type; then the inner element's
 * record.  Inner types: Multipole, Quadrupole, Bend, Sextupole, Octupole,
 * Cavity, LimitRect, LimitEllipse, DipoleEdge, RFMultipole, and slices
 * (thin, thick and edge slices take their parent's misalignment). */
#define XTM_TR_SHIFT_X 0
#define XTM_TR_SHIFT_Y 1
#define XTM_TR_SHIFT_S 2
#define XTM_TR_ROT_X 3
#define XTM_TR_ROT_Y 4
#define XTM_TR_ROT_S_NO_FRAME 5
#define XTM_TR_ANCHOR 6
#define XTM_TR_ROT_S 7
#define XTM_TR_LENGTH 8
#define XTM_TR_ANGLE 9
#define XTM_TR_H 10
#define XTM_TR_CURVED 11
#define XTM_TR_INNER_TYPE 12
#define XTM_TR_INNER 13

/* ---- track flags (xtrack/track_flags.py:5-12) ------------------------------ */
#define XTM_FLAG_BACKTRACK 0
#define XTM_FLAG_KILL_CAVITY_KICK 2
#define XTM_FLAG_IGNORE_GLOBAL_APERTURE 3
#define XTM_FLAG_IGNORE_LOCAL_APERTURE 4
#define XTM_FLAG_SR_TAPER 5
/* Line.track(freeze_longitudinal=True) (xtrack/tracker.py:1237-1242): the
 * reference recompiles with zeta and the energy variables frozen (their
 * setters become no-ops, particles/particles.py:185-230); here a flag bit
 * beyond the reference's register: energy kicks are skipped and zeta is
 * restored after every element, which is the same for every element of
 * this build (none reads zeta after changing it). */
#define XTM_FLAG_FREEZE_LONGITUDINAL 16

/* ---- particles: SoA, one pointer per per-particle field --------------------
 * Mirrors ParticlesData (xtrack/particles/particles.py:26-83).  All pointers
 * refer to arrays of length `capacity` in the memory space the consumer
 * expects (device memory for xtm_track_line).  Unused slots carry
 * state = -999999999. */
typedef struct xtm_particles {
    int64_t capacity;
    double q0;
    double mass0;
    double t_sim;
    double *p0c, *gamma0, *beta0, *s, *zeta, *x, *y, *px, *py;
    double *ptau, *delta, *rpp, *rvv;
    double *chi, *charge_ratio, *weight, *ax, *ay;
    double *spin_x, *spin_y, *spin_z, *anomalous_magnetic_moment;
    int64_t *pdg_id, *particle_id, *at_element, *at_turn, *state;
    int64_t *parent_particle_id;
    uint32_t *rng_s1, *rng_s2, *rng_s3, *rng_s4;
} xtm_particles;

/* ---- turn-by-turn monitor (xtrack/monitors/particles_monitor.py:190-203) --
 * `data` holds n_records slots per field; record index for particle id p and
 * turn t (n_repetitions == 1) is (stop-start)*(p - part_id_start) + t - start
 * (xtrack/monitors/particles_monitor.h:36-44). */
typedef struct xtm_monitor {
    int64_t start_at_turn;
    int64_t stop_at_turn;
    int64_t part_id_start;
    int64_t part_id_end;
    int64_t ebe_mode;
    int64_t n_records;
    int64_t n_repetitions;
    int64_t repetition_period;
    xtm_particles data;
} xtm_monitor;

/* ---- turn-by-turn feature accumulators (no counterpart in the reference) --
 * Inputs of the dynamic-aperture surrogate, gathered on the device at the
 * start of every turn t in [start_at_turn, stop_at_turn) for each active
 * particle, instead of a full turn-by-turn monitor record:
 * z = inv_w (x - co_x, px - co_px, y - co_y, py - co_py) scaled by
 * inv_sigma (normalised coordinates), jx = (zx^2 + zpx^2) / 2, jy likewise,
 * and the phase advance since the previous recorded turn in each plane
 * (atan2 of the rotation between consecutive (z, zp), in [0, 2 pi)), summed
 * over all recorded turns and, separately, over the turns of the second half
 * of the window (at_turn >= (start_at_turn + stop_at_turn) / 2): the tunes of
 * the two halves, whose difference is the tune diffusion.
 * acc holds XTM_FEAT_N x capacity doubles, accumulator k of particle slot i
 * at acc[k * capacity + i]; zero it before the first recorded turn. */
#define XTM_FEAT_N 17
enum {
    XTM_FEAT_COUNT = 0,      /* turns recorded */
    XTM_FEAT_JX_SUM = 1, XTM_FEAT_JX_SQ = 2, XTM_FEAT_JX_MAX = 3,
    XTM_FEAT_JY_SUM = 4, XTM_FEAT_JY_SQ = 5, XTM_FEAT_JY_MAX = 6,
    XTM_FEAT_PHX_SUM = 7, XTM_FEAT_PHY_SUM = 8,
    XTM_FEAT_ZX = 9, XTM_FEAT_ZPX = 10, XTM_FEAT_ZY = 11, XTM_FEAT_ZPY = 12,   /* last turn */
    XTM_FEAT_JX0 = 13, XTM_FEAT_JY0 = 14,
    XTM_FEAT_PHX_H2 = 15, XTM_FEAT_PHY_H2 = 16   /* phase sums, second half */
};
typedef struct xtm_features {
    double *acc;
    int64_t capacity;
    int64_t start_at_turn;
    int64_t stop_at_turn;
    double inv_w[16];        /* row-major 4 x 4 */
    double co[4];
    double inv_sigma[2];
} xtm_features;

/* Opaque device-resident element table for one line. */
typedef struct xtm_line xtm_line;

/* Library ABI version (XTM_ABI_VERSION). */
int32_t xtm_abi_version(void);

/* Message describing the last error of the calling thread ("" if none). */
const char *xtm_last_error(void);

/* Number of visible HIP devices (>= 0) or a negative status. */
int32_t xtm_device_count(void);

/* Build the device element table of a line.  Replaces TrackerData's element
 * buffer and ElementRefData (xtrack/tracker_data.py:67-172, 237-255) and the
 * per-class parameter derivation done at kernel entry (e.g.
 * elements_src/quadrupole.h:15-75).  `device` is the HIP device ordinal;
 * device = -1 builds a host-only handle that can be inspected
 * (xtm_line_num_records, xtm_line_count_ops) but not tracked.
 * The arrays are read once; the caller keeps ownership. */
int32_t xtm_line_create(const int32_t *types, const int64_t *offsets,
                        const double *params, int64_t num_elements,
                        double line_length, int32_t device, xtm_line **out);

/* Release a line and its device memory. */
int32_t xtm_line_destroy(xtm_line *line);

/* Number of elements / distinct parameter records of a created line. */
int64_t xtm_line_num_elements(const xtm_line *line);
int64_t xtm_line_num_records(const xtm_line *line);

/* Compile line code for a created line and use it for every later
 * xtm_track_line call.  Counterpart of the reference's per-line kernel
 * generation and compilation at build_tracker time (xtrack/tracker.py:
 * 544-751, compiled by xobjects): here the generated source has one
 * straight-line function per distinct element record with the record's
 * constants as literals, calling the same element maps as the program
 * interpreter, and is compiled with hiprtc for gfx950.  waves_per_eu: the
 * kernel's occupancy target (0: the library default).  Returns
 * XTM_ERR_UNSUPPORTED, leaving the interpreter in use, when the line has
 * more distinct records than code is generated for; XTM_ERR_HIP when
 * compilation or loading fails (message in xtm_last_error). */
int32_t xtm_line_specialize(xtm_line *line, int32_t waves_per_eu);

/* Options of xtm_line_specialize_ex (bit mask). */
#define XTM_SPEC_FAST_ARITH 1u

/* xtm_line_specialize with options (no counterpart in the reference, which
 * compiles one arithmetic).  XTM_SPEC_FAST_ARITH: the polar drifts of sector
 * bends drop their divisions -- the path-length quotient is re-associated
 * onto the factor the map already forms, and the two remaining reciprocals
 * (of the square root and of ca - sa px/pz) come from the hardware
 * estimates with two Newton steps -- so coordinates differ from the IEEE
 * evaluation by a few ulp per element (about 1e-15 relative after one turn
 * of a 28k-element ring); losses are decided by the same aperture tests.
 * xtm_line_specialize(line, w) is xtm_line_specialize_ex(line, w, 0). */
int32_t xtm_line_specialize_ex(xtm_line *line, int32_t waves_per_eu, uint32_t options);

/* The options of the line's compiled code, or -1 when it has none. */
int32_t xtm_line_specialized_options(const xtm_line *line);

/* 1 when the line tracks with generated code, else 0. */
int32_t xtm_line_is_specialized(const xtm_line *line);

/* Diagnostics (no counterpart in the reference): the source that
 * xtm_line_specialize compiles, written to buf (capacity cap, NUL
 * terminated, *len = full length), the record index of each element
 * (rec_of_elem, num_elements entries, may be NULL), and the element
 * program the generated code refers to (xtm_line_program_size doubles)
 * with its per-element entry offsets (num_elements int64).  Works on
 * host-only handles; the CPU tests compile the source for the host. */
int32_t xtm_line_generate_source(const xtm_line *line, int32_t waves_per_eu, char *buf,
                                 int64_t cap, int64_t *len, int32_t *rec_of_elem);
int32_t xtm_line_generate_source_ex(const xtm_line *line, int32_t waves_per_eu,
                                    uint32_t options, char *buf, int64_t cap, int64_t *len,
                                    int32_t *rec_of_elem);
int64_t xtm_line_program_size(const xtm_line *line);
int32_t xtm_line_copy_program(const xtm_line *line, double *out);
int32_t xtm_line_copy_entry(const xtm_line *line, int64_t *out);

/* Track particles in place, asynchronously on `stream`.  Same arguments and
 * semantics as the reference kernel `track_line` (xtrack/tracker.py:546-564):
 * turns [0, num_turns), elements [ele_start, ele_start+num_ele_track) each
 * turn, end-of-turn actions when flag_end_turn_actions > 0, monitor mode
 * flag_monitor (0 none, 1 turn-by-turn, 2 element-by-element).  `monitor`
 * may be NULL when flag_monitor == 0.  Particle and monitor pointers are
 * device pointers. */
int32_t xtm_track_line(xtm_line *line, const xtm_particles *particles,
                       int32_t num_turns, int32_t ele_start,
                       int32_t num_ele_track, int32_t flag_end_turn_actions,
                       int32_t flag_reset_s_at_end_turn, int32_t flag_monitor,
                       const xtm_monitor *monitor, uint64_t track_flags,
                       void *stream);

/* Update the feature accumulators of `features` (capacity equal to the
 * particle set's) with the present coordinates of every active particle
 * whose at_turn lies in [start_at_turn, stop_at_turn): one record at the
 * start of its current turn (no counterpart in the reference). */
int32_t xtm_features_record(const xtm_particles *particles, const xtm_features *features,
                            void *stream);

/* num_turns whole turns from element 0 (end-of-turn actions on, s reset,
 * no monitor) as one-turn xtm_track_line launches, each preceded by
 * xtm_features_record: the surrogate's feature turns without a
 * turn-by-turn record (no counterpart in the reference). */
int32_t xtm_track_line_features(xtm_line *line, const xtm_particles *particles,
                                int32_t num_turns, const xtm_features *features,
                                uint64_t track_flags, void *stream);

/* Operation count of one turn of the kernel's element maps for one
 * particle with 6D coordinates coords = (x, px, y, py, zeta, delta) and
 * chi = 1, obtained by running the same element code on the host with an
 * operation-counting scalar.  out[0..5]: add/sub, mul, div, sqrt,
 * elementary-function calls, total arithmetic.  No counterpart in the
 * reference; used for the algorithmic-FLOP roofline. */
int32_t xtm_line_count_ops(const xtm_line *line, const double *coords, double p0c,
                           double mass0, double q0, double *out);

/* ---- batched lattice variants ---------------------------------------------
 * Many lines (lattice variants with their own element counts and
 * strengths) tracked in one launch, each over its own contiguous range of
 * one particle set.  No counterpart in the reference, which tracks one
 * line per `Line.track` call (xtrack/tracker.py:1372, one `track_line`
 * launch per call); each variant's particles get exactly what that call
 * gives them for full turns from element 0 (end-of-turn actions on, s
 * reset at the end of each turn, no monitor).  The lines run on the
 * element-program interpreter. */
typedef struct xtm_batch xtm_batch;

/* A batch over `num_lines` line handles (all on one device, or all
 * host-only for layout inspection).  The lines must outlive the batch. */
int32_t xtm_batch_create(xtm_line *const *lines, int32_t num_lines, xtm_batch **out);
int32_t xtm_batch_destroy(xtm_batch *batch);

/* Particles of line k are [line_start[k], line_start[k+1]) of the particle
 * set passed to xtm_batch_track (line_start has num_lines + 1 entries,
 * non-decreasing, from 0).  block_size (64, 128 or 256 threads; the
 * per-thread LDS slots are sized by the launch) and waves_per_eu (3, 4 or
 * 5, the occupancy of the interpreter kernel instance; the compiled batch
 * code is built for 5) select the launch.  Builds the block table: each
 * block holds particles of one line only, and blocks of longer lines come
 * first. */
int32_t xtm_batch_set_layout(xtm_batch *batch, const int64_t *line_start,
                             int32_t block_size, int32_t waves_per_eu);

/* Number of blocks of the current layout; the table itself as
 * (line, first particle, count) triples (3 * num_blocks int64). */
int64_t xtm_batch_num_blocks(const xtm_batch *batch);
int32_t xtm_batch_copy_blocks(const xtm_batch *batch, int64_t *out);

/* Compile the batch's lines into one kernel (hiprtc, gfx950): one function
 * per distinct element structure, with the constants read from each line's
 * program, so that variants differing only in strengths share the code.
 * Particles with chi == 1 then track on it; the interpreter follows on the
 * same stream for the others.  Returns XTM_ERR_UNSUPPORTED when the lines
 * have more than 2,048 distinct structures. */
int32_t xtm_batch_specialize(xtm_batch *batch, int32_t waves_per_eu);
int32_t xtm_batch_is_specialized(const xtm_batch *batch);

/* Diagnostics: the source xtm_batch_specialize compiles (as
 * xtm_line_generate_source) and the structure index of every element of
 * every line, concatenated in line order (recs may be NULL).  Works on
 * host-only batches; the CPU tests compile the source for the host. */
int32_t xtm_batch_generate_source(const xtm_batch *batch, int32_t waves_per_eu, char *buf,
                                  int64_t cap, int64_t *len, int32_t *recs);

/* Track num_turns full turns, asynchronously on `stream`.  The particle set
 * must hold at least line_start[num_lines] particles.  Launches kernels
 * only (no allocation, no copy), so it may be captured in a HIP graph. */
int32_t xtm_batch_track(xtm_batch *batch, const xtm_particles *particles,
                        int32_t num_turns, uint64_t track_flags, void *stream);

/* xtm_batch_track as one-turn launches, each preceded by
 * xtm_features_record (as xtm_track_line_features; one normalisation for
 * all lines). */
int32_t xtm_batch_track_features(xtm_batch *batch, const xtm_particles *particles,
                                 int32_t num_turns, const xtm_features *features,
                                 uint64_t track_flags, void *stream);

/* ---- survival-turn surrogate network --------------------------------------
 * No counterpart in the reference.  C[M x N] = op(A) op(B) on the matrix
 * cores: BF16 operands (raw 16-bit patterns), FP32 accumulation, op(A)(m, k)
 * = A[m*sam + k*sak], op(B)(k, n) = B[k*sbk + n*sbn] (strides give plain or
 * transposed operands), then + bias[n] when bias is not NULL and ReLU when
 * relu != 0.  C (FP32, row stride ldc) always; Cb, when not NULL, the same
 * values rounded to BF16 (nearest even).  Device pointers; asynchronous on
 * `stream`.  M up to 65535 * 64 rows. */
int32_t xtm_gemm_bf16(const uint16_t *A, int64_t sam, int64_t sak, const uint16_t *B,
                      int64_t sbk, int64_t sbn, float *C, int64_t ldc, uint16_t *Cb,
                      const float *bias, int32_t relu, int64_t M, int64_t N, int64_t K,
                      void *stream);

/* xtm_gemm_bf16 without any allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same kernels
 * and results as xtm_gemm_bf16; the k-split partials go to a file-scope
 * device array of xtm_gemm_bf16_scratch_floats() floats (2^20, 4 MiB) per
 * GPU instead of a library-allocated workspace, and products whose
 * partials would not fit are rejected (XTM_ERR_ARG, message in
 * xtm_last_error).  Calls that overlap on one GPU share the array and
 * race. */
int32_t xtm_gemm_bf16_noalloc(const uint16_t *A, int64_t sam, int64_t sak, const uint16_t *B,
                              int64_t sbk, int64_t sbn, float *C, int64_t ldc, uint16_t *Cb,
                              const float *bias, int32_t relu, int64_t M, int64_t N, int64_t K,
                              void *stream);
int64_t xtm_gemm_bf16_scratch_floats(void);

/* Wait for all work queued on `stream`. */
int32_t xtm_synchronize(void *stream);

/* Diagnostics: evaluate the library's device elementary functions
 * (fn: 0 sin, 1 cos, 2 sinh, 3 cosh, 4 exp, 5 tan, 6 sqrt, 7 div by x[i+n];
 * the kernel's in-range sequences: 8 reciprocal, 9 sqrt, 10 div by x[i+n];
 * 11 atan2(x[i], x[i+n]), 12 asin)
 * on n device doubles.  No counterpart in the reference; used to check that
 * the device and host builds of the functions agree bit for bit and that the
 * in-range sequences return the IEEE results. */
int32_t xtm_math_eval(int32_t fn, const double *x, double *y, int64_t n,
                      void *stream);

#ifdef __cplusplus
}
#endif

#endif /* XTRACK_AMD_H */
