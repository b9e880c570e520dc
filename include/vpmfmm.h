/*
 * vpmfmm.h -- C ABI of the MI355X (gfx950) 2-D FMM Biot-Savart velocity
 * library (lsmps_vpm_v2_amd/libvpmfmm.so).
 *
 * Drop-in boundary for the velocity path of anggarapangestu/LSMPS_VPM_v2:
 *   VelocityCalc::get_velocity(Particle&, const GridNode&, int)
 *       src/velocity_calculation/velocity_calc.hpp:42, velocity_calc.cpp:149-251
 *   VelocityCalc::velocity_fmm_2d(Particle&, int)     velocity_calc.cpp:455-578
 *   TreeCell::initializeTree / updateTree              src/FMM/treeCell.hpp:152-153
 *   fmm2D::calcField + get_Field_x / get_Field_y       src/FMM/fmm2D.hpp:59-67
 * Plain pointers and sizes only; every function returns a vpmfmm_status and
 * never throws.  All work of one call runs on the handle's HIP stream and is
 * complete when the call returns.  One handle per host thread.
 */
#ifndef VPMFMM_H
#define VPMFMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VPMFMM_ABI_VERSION 1

typedef enum vpmfmm_status {
    VPMFMM_OK = 0,
    VPMFMM_BAD_ARG = 1,              /* invalid option, size or pointer          */
    VPMFMM_NONFINITE_INPUT = 2,      /* NaN/inf particle coordinate              */
    VPMFMM_DEGENERATE_DOMAIN = 3,    /* all particles at one point (root size 0) */
    VPMFMM_TREE_DEPTH_EXCEEDED = 4,  /* a cell still over threshold at level 16  */
    VPMFMM_OUT_OF_MEMORY = 5,
    VPMFMM_INTERNAL = 6,
    VPMFMM_DUPLICATE_POINTS = 7,     /* two particles of one list-1 pair coincide */
    VPMFMM_HIP_ERROR = 8,
    VPMFMM_OUTSIDE_BOXES = 9,        /* type-1: a particle outside every box of the
                                        reference's hierarchy (see vpmfmm_velocity_type1) */
    VPMFMM_LEVEL1_LEAF = 10          /* 3-D type 2: a level-1 cell is a leaf while the tree
                                        goes deeper (the reference's get_childID throws)   */
} vpmfmm_status;

typedef enum vpmfmm_precision {
    VPMFMM_FP64 = 0,   /* plain FP64 path, reference semantics                     */
    VPMFMM_MIXED = 1   /* FP32 P2P/P2M/L2P, FP32 MFMA M2L, FP64 geometry           */
} vpmfmm_precision;

/* Options.  Defaults (vpmfmm_default_opts) are the reference's values:
 * P_max 10, par_count_max 75, src_count_max 50, expTree 1.0 (global.cpp:303-313),
 * u_inf 1, v_inf 0 (global.cpp:164-165). */
typedef struct vpmfmm_opts {
    int32_t p;           /* expansion order, 1..40 (Pars::P_max)                       */
    int32_t max_par;     /* leaf rule, particles per cell, 1..4096 (Pars::par_count_max) */
    int32_t max_src;     /* leaf rule, active particles per cell, 0..4096 (src_count_max) */
    double exp_tree;     /* root expansion in percent, > 0 (Pars::expTree)              */
    int32_t min_level;   /* cells above this level always split (reference: 1)          */
    int32_t max_level;   /* cells at this level always leaves, 0 = none (max 16)        */
    int32_t precision;   /* vpmfmm_precision                                            */
    int32_t binom_mode;  /* M2L binomials: 1 = the reference's int32 helper
                            (fmm2D.cpp:28-58, parity), 0 = exact                         */
    int32_t list4_mode;  /* 1 = M2L from the list-4 leaf multipole (M2L_P2_METHOD_TYPE 1,
                            reference default), 2 = P2L from its particles              */
    double u_inf;        /* free stream added by the velocity entry points              */
    double v_inf;
    int32_t device;      /* HIP device ordinal                                          */
    int32_t timing;      /* 1 = record per-stage times into vpmfmm_stats                */
    int32_t stage_mask;  /* evaluation parts (diagnostics): 1 P2P, 2 M2P (list 3),
                            4 L2P (far field); 0 = all                                 */
    int32_t order_select;   /* mixed path: vpmfmm_order_select (fixed / learned)        */
    int32_t reselect_every; /* learned: re-choose the leaf threshold every k calls
                               (0 = on the first call and when n changes by > 10%)      */
    int32_t judge_every;    /* learned: the call that makes a fresh order selection (and,
                               for k > 0, every k-th call) also solves at the fixed order
                               p on the same tree and compares; a difference above 0.9 x
                               the model's target (the rest is left for the fixed order's
                               own error) returns the fixed-order result and keeps the
                               fixed order until the next re-selection; -1 = never      */
    int32_t icutoff;        /* near-field (P2P) kernel: 0 = singular, the type-2 FMM
                               (fmm2D.cpp:164-173); 1 / 2 / 3 = the high-order
                               algebraic / Gaussian / super-Gaussian regularization of
                               regul_func_2d (regularization_function_2d.cpp:11-37,
                               Pars::icutoff) at sij^2 = (si^2 + sj^2) / 4, as the
                               type-1 near field (subroutine_all_fmm_2d.cpp:594-627):
                               pairs at r = 0 add nothing.  The far field stays
                               singular.  > 0 needs the sizes: velocity entries only */
    int32_t reserved[3];
} vpmfmm_opts;

/* Expansion-order source of the mixed path. */
typedef enum vpmfmm_order_select {
    VPMFMM_ORDER_FIXED = 0,   /* opts.p, or the table of vpmfmm_set_level_orders          */
    VPMFMM_ORDER_LEARNED = 1  /* the MLP of vpmfmm_set_selector picks p per level and
                                 the leaf threshold; fixed order opts.p as the fallback  */
} vpmfmm_order_select;

#define VPMFMM_SEL_FEATURES 9
#define VPMFMM_SEL_HIDDEN 64
#define VPMFMM_SEL_PARAMS 4930  /* 9*64 + 64 + 64*64 + 64 + 64*2 + 2 */

/* Learned selection model: 9 features per tree level -> 64 tanh -> 64 tanh
 * -> (e, t) = ln(level's share of the squared relative L2 error),
 * ln(level's share of the solve time in ms).  Features of level L (raw, before
 * x_norm = (x - mu) / sd): L, log2 cells, log2(particles / cell), log2 side,
 * log10(sum|q| of the level / sum|q|), log10(max|q| N / sum|q|),
 * leaves / cells, log2 leaf threshold, p (0 for levels 0 and 1).
 * Parameters: W1[9][64], b1[64], W2[64][64], b2[64], W3[64][2], b3[2]. */
typedef struct vpmfmm_selector {
    int32_t n_features;      /* VPMFMM_SEL_FEATURES                                     */
    int32_t hidden;          /* VPMFMM_SEL_HIDDEN                                       */
    int32_t n_params;        /* VPMFMM_SEL_PARAMS                                       */
    int32_t p_min, p_max;    /* candidate orders: 1 <= p_min <= p_max <= 16, < 16 values */
    int32_t n_thresholds;    /* candidate leaf thresholds (max_par = max_src), 1..8     */
    int32_t thresholds[8];
    float target_err;        /* relative L2 velocity error to meet, e.g. 1e-4           */
    float margin;            /* feature excursion allowed beyond [lo, hi], in hi - lo   */
    float mu[9], sd[9], lo[9], hi[9];
} vpmfmm_selector;

/* What the last mixed-path call used. */
typedef struct vpmfmm_selection {
    int32_t source;          /* 0 fixed, 1 learned, 2 learned model fell back to opts.p,
                                3 the fixed-order judge rejected the learned orders      */
    int32_t threshold;       /* leaf threshold of the tree (max_par; max_src)           */
    int32_t p[17];           /* order per level 0..16                                  */
    int32_t reselected;      /* 1 if this call re-chose the leaf threshold              */
    float pred_err;          /* predicted relative L2 error (source 1), else -1          */
    float pred_ms;           /* predicted solve time (source 1), else -1                 */
    float ms_reselect;       /* time spent re-choosing the threshold in this call       */
    float judge_err;         /* rel-L2 difference learned vs fixed order (judged calls), else -1 */
} vpmfmm_selection;

typedef struct vpmfmm_stats {
    int32_t n, n_active, n_cells, n_leaves;
    int32_t min_leaf_level, max_leaf_level;
    int64_t n_list1, n_list2, n_list3, n_list4;  /* list entries summed over cells   */
    int64_t p2p_pairs;                             /* ordered (target, source) pairs  */
    int64_t m2p_evals;                             /* (particle, list-3 cell) pairs   */
    double ms_total;                               /* stage times (timing = 1)        */
    double ms_tree, ms_lists, ms_upward, ms_m2l, ms_downward, ms_eval;
    int64_t m2p_terms;   /* sum over (particle, list-3 cell) of (order of the cell + 1) */
    int64_t l2p_terms;   /* sum over particles of the order of their leaf             */
    int64_t scratch_bytes; /* device scratch the last call took from the array of
                              vpmfmm_velocity_device_noalloc (0 for the other calls)   */
} vpmfmm_stats;

typedef struct vpmfmm_handle_s vpmfmm_handle;

/* Fills *o with the reference defaults. */
void vpmfmm_default_opts(vpmfmm_opts *o);

/* Handle lifetime.  Replaces the solver objects held by VelocityCalc
 * (TreeCell treeData member, velocity_calc.hpp:19; fmm2D local,
 * velocity_calc.cpp:560).  Device buffers persist and grow on demand. */
int vpmfmm_create(const vpmfmm_opts *o, vpmfmm_handle **out);
int vpmfmm_destroy(vpmfmm_handle *h);
int vpmfmm_set_opts(vpmfmm_handle *h, const vpmfmm_opts *o);
int vpmfmm_get_opts(const vpmfmm_handle *h, vpmfmm_opts *o);
/* Use the caller's HIP stream (hipStream_t passed as void*); NULL restores
 * the handle's own stream. */
int vpmfmm_set_stream(vpmfmm_handle *h, void *hip_stream);

/* Field of the singular 2-D kernel, host pointers:
 *   ex_i = sum_j q_j (x_i - x_j)/r_ij^2,  ey_i = sum_j q_j (y_i - y_j)/r_ij^2
 * with the reference's FMM semantics (inactive particles are P2P sources only).
 * Replaces TreeCell::initializeTree(pos, active) (treeCell.cpp:1043) followed by
 * fmm2D::calcField(tree, pos, active, src) and get_Field_x/y (fmm2D.cpp:907,
 * 1146-1157).  active: one byte per particle, nonzero = active. */
int vpmfmm_field(vpmfmm_handle *h, int32_t n, const double *x, const double *y,
                 const double *q, const uint8_t *active, double *ex, double *ey);
/* Same with device pointers (inputs resident in HBM). */
int vpmfmm_field_device(vpmfmm_handle *h, int32_t n, const double *x, const double *y,
                        const double *q, const uint8_t *active, double *ex, double *ey);

/* Potential of the same kernel, host pointers:
 *   phi_i = sum_{j != i} q_j log r_ij
 * with the reference's FMM semantics: fmm2D::calcPotential + get_Potential
 * (fmm2D.cpp:681-896, 1137-1140; caller pressure_poisson.cpp:216-219).  FP64
 * whatever opts.precision (the order, leaf rule, binomials and list-4 mode of
 * the options apply); singular kernel only (icutoff 0).  Two particles of a
 * list-1 pair at one position give DUPLICATE_POINTS (log 0). */
int vpmfmm_potential(vpmfmm_handle *h, int32_t n, const double *x, const double *y,
                     const double *q, const uint8_t *active, double *phi);
/* Same with device pointers. */
int vpmfmm_potential_device(vpmfmm_handle *h, int32_t n, const double *x, const double *y,
                            const double *q, const uint8_t *active, double *phi);

/* Velocity of VelocityCalc::velocity_fmm_2d + get_velocity's free stream
 * (velocity_calc.cpp:465-481, 572-575, 226-230):
 *   q = -omega s^2 / (2 pi),  u = Ey + u_inf,  v = -Ex + v_inf. */
int vpmfmm_velocity(vpmfmm_handle *h, int32_t n, const double *x, const double *y,
                    const double *s, const double *omega, const uint8_t *active, double *u,
                    double *v);
int vpmfmm_velocity_device(vpmfmm_handle *h, int32_t n, const double *x, const double *y,
                           const double *s, const double *omega, const uint8_t *active,
                           double *u, double *v);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): vpmfmm_velocity_device with every scratch buffer (sorted
 * particles, cells, lists, expansions, tables, judge buffers) carved from one
 * static device array of VPMFMM_NOALLOC_SCRATCH_BYTES instead of the handle's
 * grow-only buffers.  The call makes no hipMalloc / hipFree (the handle, its
 * streams and events were made by vpmfmm_create, which also resolves the
 * array's address on the handle's device).  Limits: n <= VPMFMM_NOALLOC_MAX_N,
 * and a tree whose buffers exceed the array returns BAD_ARG (a uniform-like
 * field at the maximum n takes about 110 MiB; vpmfmm_stats.scratch_bytes
 * reports the amount).  Each device holds one array for the life of the
 * process, shared by every handle on it: calls on one device that overlap in
 * time (two host threads, or two handles on different streams) race on it.
 * The tree is not kept for the inspection functions after such a call. */
#define VPMFMM_NOALLOC_SCRATCH_BYTES (256ull << 20)
#define VPMFMM_NOALLOC_MAX_N (1 << 20)
int vpmfmm_velocity_device_noalloc(vpmfmm_handle *h, int32_t n, const double *x,
                                   const double *y, const double *s, const double *omega,
                                   const uint8_t *active, double *u, double *v);

/* Type-1 ("old") FMM options: Pars::n_s, lmax, ndp, icutoff of the
 * reference's global.cpp:316-328 (defaults 50, 7, 10, 2). */
typedef struct vpmfmm_type1_opts {
    int32_t n_s;       /* boxes holding more than n_s particles split, 1..4096        */
    int32_t lmax;      /* deepest level, 2..16                                        */
    int32_t ndp;       /* multipole terms (orders 0..ndp-1); the local expansion has
                          ndp + 1 terms (orders 0..ndp), 1..40                        */
    int32_t icutoff;   /* near field: 0 singular, 1 high-order algebraic, 2 Gaussian,
                          3 super-Gaussian (regul_func_2d); pairs at r = 0 add nothing */
    int32_t l2l_mode;  /* 1 = the L2L shift as the reference runs it: its bico helper
                          rewrites the k argument in place (subroutine_all_fmm_2d.cpp:
                          580-583), which changes the binomial and exponent of later
                          terms (biotsavart_fmm_2d.cpp:299-307) and holds the accuracy
                          near 4e-4 whatever ndp; 0 = the exact shift                  */
    int32_t reserved[3];
} vpmfmm_type1_opts;

void vpmfmm_type1_default_opts(vpmfmm_type1_opts *o);

/* Replaces VelocityCalc::velocity_old (velocity_calc.cpp:390-440,
 * VELOCITY_2D_TYPE 1) -> velocity_biot_savart::biotsavart_fmm_2d
 * (biotsavart_fmm_2d.cpp:4-501) with its box hierarchy and lists
 * (subroutine_all_fmm_2d.cpp:97-556), followed by get_velocity's free stream
 * (velocity_calc.cpp:226-230): u, v of every particle from its position, size
 * s (Particle::s, read for icutoff > 0; may be NULL for icutoff 0) and vortex
 * strength gz (Particle::gz = vorticity s^2), every particle a target and a
 * source:
 *   u = u_inf - sum_j f(r, s_ij) (y_i - y_j) gz_j / (2 pi r^2),
 *   v = v_inf + sum_j f(r, s_ij) (x_i - x_j) gz_j / (2 pi r^2)
 * over the near-field pairs at r > 0 (f of the icutoff kernel at
 * s_ij^2 = (s_i^2 + s_j^2) / 4) and the far field by the type-1 expansions.
 * The boxes, their bounds and every list decision are the reference's (its
 * rounded box bounds and absolute 1e-15 box tolerance).  The reference drops a
 * particle that lies on the upper bound of its level-1 boxes (all particles on
 * one horizontal or vertical line leave the padding at zero) or between a
 * split box's rounded child bounds; such input returns VPMFMM_OUTSIDE_BOXES.
 * All particles at one point give u_inf, v_inf (the reference's empty boxes).
 * FP64 throughout; u_inf, v_inf and device come from the handle's options,
 * everything else from *o.  Host pointers for vpmfmm_velocity_type1, device
 * pointers (results on the handle's stream) for the _device form. */
int vpmfmm_velocity_type1(vpmfmm_handle *h, const vpmfmm_type1_opts *o, int32_t n,
                          const double *x, const double *y, const double *s, const double *gz,
                          double *u, double *v);
int vpmfmm_velocity_type1_device(vpmfmm_handle *h, const vpmfmm_type1_opts *o, int32_t n,
                                 const double *x, const double *y, const double *s,
                                 const double *gz, double *u, double *v);

/* ------------------------------------------------------------------------
 * The 3-D vector FMM (the reference's default 3-D velocity, VELOCITY_3D_TYPE
 * 1: VelocityCalc::velocity_fmm_3d_fast, velocity_calc.cpp:2384-2525 ->
 * fmm3D::calcVelocityFast, src/FMM/fast3DFMM.cpp:32-92, helpers
 * src/FMM/fmm3D.cpp:84-210; and, with opts.type 2, VELOCITY_3D_TYPE 2:
 * VelocityCalc::velocity_fmm_3d, velocity_calc.cpp:2172-2383 ->
 * TreeCell::initializeTree + fmm3D::calcVelocity, treeCell.cpp:1043-1243,
 * fmm3D.cpp:585-700, 1064-1205).  FP64.
 *
 * Octree of FMMCell::generateTree (fast3DFMM.cpp:485-745); order-2 Cartesian
 * moments of the vortex strengths alpha = omega s^3 / (4 pi); every particle
 * is a target; a cell whose centre is farther than support * side from the
 * target adds its moments' term, a nearer leaf adds the direct sum
 *   u += alpha_j x (x_i - x_j) / |x_i - x_j|^3
 * over its active sources (the target itself excluded).  As in the
 * reference, the walk starts at the children of the level-1 cells, so the
 * sources of a level-1 leaf (an octant of the root holding fewer than
 * crit_num particles) are never summed.
 * ---------------------------------------------------------------------- */
typedef struct vpmfmm_fmm3d_opts {
    int32_t crit_num;   /* a cell with fewer particles is a leaf (Pars::src_count_max
                           of the DIM == 3 parameters, global.cpp:779: 30), 1..4096   */
    int32_t max_depth;  /* refuse deeper trees, 1..20 (the reference has no bound;
                           at least crit_num coincident particles never separate)    */
    double exp_tree;    /* root expansion in percent (Pars::expTree, global.cpp:769:
                           1.0)                                                       */
    double center[3];   /* a point the root box always holds (Pars::x/y/zcenter,
                           global.cpp:609-611: 0)                                     */
    double support;     /* opening factor (SUPPORT_RADIUS_FACTOR 2.0: fast3DFMM.cpp:4, fmm3D.cpp:3)*/
    double u_inf, v_inf, w_inf;  /* free stream added to every velocity (get_velocity,
                           velocity_calc.cpp:221-236; global.cpp:633-635: 1, 0, 0)   */
    int32_t type;       /* VELOCITY_3D_TYPE (velocity_calc.cpp:31): 1 (default)
                           velocity_fmm_3d_fast, the FMMCell octree; 2 velocity_fmm_3d,
                           fmm3D::calcVelocity over the 3-D TreeCell (root box of the
                           particles alone, truncated cell indices, the leaf rule below,
                           every particle of a near leaf summed, inactive ones too)   */
    int32_t max_par;    /* type 2: a cell is a leaf with at most max_par particles     */
    int32_t max_src;    /*   and at most max_src active ones (Pars::par_count_max 50,
                           src_count_max 30, global.cpp:778-779), 0..4096             */
    int32_t reserved;
} vpmfmm_fmm3d_opts;

void vpmfmm_fmm3d_default_opts(vpmfmm_fmm3d_opts *o);

/* (u, v, w) of every particle from positions (x, y, z), sizes s, vorticity
 * (wx, wy, wz) and active flags (inactive particles are targets; type 1
 * never sums them, type 2 sums them in the near field).  Host pointers for
 * vpmfmm_velocity3d, device pointers (on the handle's device) for
 * vpmfmm_velocity3d_device.  VPMFMM_NONFINITE_INPUT for a non-finite
 * coordinate, VPMFMM_TREE_DEPTH_EXCEEDED for a tree deeper than max_depth,
 * VPMFMM_DEGENERATE_DOMAIN (type 2) for every particle at one point,
 * VPMFMM_LEVEL1_LEAF (type 2) where the reference throws. */
int vpmfmm_velocity3d(vpmfmm_handle *h, const vpmfmm_fmm3d_opts *o, int32_t n, const double *x,
                      const double *y, const double *z, const double *s, const double *wx,
                      const double *wy, const double *wz, const uint8_t *active, double *u,
                      double *v, double *w);
int vpmfmm_velocity3d_device(vpmfmm_handle *h, const vpmfmm_fmm3d_opts *o, int32_t n,
                             const double *x, const double *y, const double *z, const double *s,
                             const double *wx, const double *wy, const double *wz,
                             const uint8_t *active, double *u, double *v, double *w);

/* The FMM tool's own entry (fmm3D::calcVelocityFast, fast3DFMM.cpp:32-92):
 * the rotational velocity alone (the free stream of *o is not added) from the
 * vortex strengths (ax, ay, az) themselves.  Host pointers. */
int vpmfmm_velocity3d_alpha(vpmfmm_handle *h, const vpmfmm_fmm3d_opts *o, int32_t n,
                            const double *x, const double *y, const double *z, const double *ax,
                            const double *ay, const double *az, const uint8_t *active, double *u,
                            double *v, double *w);

/* What the last 3-D solve built and evaluated. */
typedef struct vpmfmm_fmm3d_stats {
    int32_t n, n_active, n_cells, n_leaves, depth;  /* cells exclude the root      */
    int32_t sorted_levels;     /* octree levels the particle sort resolved (10, 20) */
    double root_centre[3], root_side;
    uint64_t pairs;            /* direct (target, source) pairs                     */
    uint64_t m2p;              /* (target, far cell) moment terms                   */
    uint64_t wave_pairs;       /* wavefront steps of the direct sums (64 lanes each;
                                  pairs / (64 wave_pairs) is the lane use)          */
    uint64_t wave_m2p;         /* wavefront steps of the moment terms               */
    double ms_tree;            /* stage times of a handle made with timing = 1 (HIP
                                  events; -1 otherwise): octree and compacted sources, */
    double ms_moments;         /* moments,                                           */
    double ms_eval;            /* evaluation kernel                                  */
} vpmfmm_fmm3d_stats;

int vpmfmm_fmm3d_last_stats(vpmfmm_handle *h, vpmfmm_fmm3d_stats *st);

/* The cells of the last 3-D solve (n_cells entries each, any pointer may be
 * NULL), depth-first: level, path key (3 bits per level, octant x + 2y + 4z,
 * level 1 highest), particle count, leaf, active, centre (3 per cell), side.
 * vpmfmm_fmm3d_perm: the particles in the tree's order (n entries). */
int vpmfmm_fmm3d_cells(vpmfmm_handle *h, int32_t *level, int64_t *key, int32_t *count,
                       uint8_t *leaf, uint8_t *active, double *centre3, double *side);
int vpmfmm_fmm3d_perm(vpmfmm_handle *h, int32_t *perm);

/* Advection schemes of advection::main_advection (main_advection.cpp:6-39). */
typedef enum vpmfmm_advect_scheme {
    VPMFMM_ADVECT_EULER = 0, /* advection_euler, :44-61 (the scheme the reference compiles) */
    VPMFMM_ADVECT_RK2 = 1    /* advection_rk2, :64-83 (scales the current velocity only)    */
} vpmfmm_advect_scheme;

/* Advection sub-step, device pointers, in place on x, y
 * (advection::main_advection, main_advection.cpp:6-83):
 *   EULER: x += u dt, y += v dt;
 *   RK2:   x += (u + 2 u/2 + 2 u/4 + u/4) dt / 6, the same for y with v.
 * Rounded as the reference's host loop (no FMA contraction). */
int vpmfmm_advect_device(vpmfmm_handle *h, int32_t n, double *x, double *y, const double *u,
                         const double *v, double dt, int32_t scheme);

/* One step of the reference's iteration loop restricted to the velocity path
 * (main.cpp:296-429 without penalization, diffusion and remeshing): the
 * advection sub-step with the current (u, v) (main.cpp:373), then the new
 * velocity at the moved particles into (u, v) (get_velocity, main.cpp:410).
 * Device pointers; x, y, u, v are updated in place. */
int vpmfmm_step_device(vpmfmm_handle *h, int32_t n, double *x, double *y, const double *s,
                       const double *omega, const uint8_t *active, double *u, double *v,
                       double dt, int32_t scheme);

/* FP64 O(N^2) direct sum of the same kernel on the GPU, device pointers: the
 * field at targets tidx[0..nt) (indices into x/y; j == i skipped), or at all
 * n particles when tidx is NULL.  src_mask (may be NULL) selects sources.
 * An accuracy reference, not part of the FMM path (the reference keeps a
 * regularised variant in biotsavart_direct_2d.cpp). */
int vpmfmm_direct_device(vpmfmm_handle *h, int32_t n, const double *x, const double *y,
                         const double *q, const uint8_t *src_mask, int32_t nt,
                         const int32_t *tidx, double *ex, double *ey);

/* Statistics of the last successful call. */
int vpmfmm_last_stats(const vpmfmm_handle *h, vpmfmm_stats *st);
/* Message of the last failed call on this handle ("" if none). */
const char *vpmfmm_last_error(const vpmfmm_handle *h);

/* ---- inspection of the last tree (parity tests; data copied to the host) ----
 * Cells are identified by the reference's cell ID
 *   ID = (4^L - 1)/3 + ix + iy 2^L   (treeCell.cpp:114-148, 223-250).       */

/* Builds the tree and lists only (TreeCell::initializeTree, treeCell.cpp:1043). */
int vpmfmm_build_tree(vpmfmm_handle *h, int32_t n, const double *x, const double *y,
                      const uint8_t *active);
int vpmfmm_tree_counts(vpmfmm_handle *h, int32_t *n_cells, int32_t *n_leaves,
                       int32_t *min_leaf_level, int32_t *max_leaf_level);
int vpmfmm_tree_geometry(vpmfmm_handle *h, double *root_len, double *pivot2);
/* All cells sorted by ID (arrays of n_cells). */
int vpmfmm_tree_cells(vpmfmm_handle *h, int64_t *id, int32_t *level, int32_t *ix, int32_t *iy,
                      int32_t *par_num, int32_t *src_num, uint8_t *is_leaf, uint8_t *is_active);
/* Leaves sorted by ID with their particles (caller indices, ascending):
 * leaf_id[n_leaves], poff[n_leaves + 1], pidx[n]. */
int vpmfmm_tree_leaves(vpmfmm_handle *h, int64_t *leaf_id, int32_t *poff, int32_t *pidx);
/* Lists of one cell (TreeCell::intList / extList, treeCell.cpp:636, 818):
 * lists 1 and 3 for leaves, lists 2 and 4 for cells of level >= 2.  Counts
 * are always written; entries only up to cap. */
int vpmfmm_tree_lists(vpmfmm_handle *h, int64_t id, int64_t *l1, int32_t *n1, int64_t *l2,
                      int32_t *n2, int64_t *l3, int32_t *n3, int64_t *l4, int32_t *n4,
                      int32_t cap);

/* Expansion order per tree level for the mixed-precision path: p[L] for
 * levels 0..nlev-1 (0 or a missing entry: opts.p), each in [1, 16]; nlev = 0
 * clears the table.  Translations between levels of different order truncate
 * or zero-pad.  No counterpart in the reference, which has one global P_max
 * (global.cpp:311); used by the learned order selection. */
int vpmfmm_set_level_orders(vpmfmm_handle *h, const int32_t *p, int32_t nlev);

/* Multipole (ak) and local (bk) coefficients of every cell after the last
 * FP64 field/velocity/potential call, cells sorted by ID: ids[n_cells],
 * ak/bk[n_cells * 2(p+1)] as (re, im) pairs.  b_0 is formed by the potential
 * calls only (it never reaches the field).  Stage-level parity hook (P2M,
 * M2M, M2L + L2L). */
int vpmfmm_tree_expansions(vpmfmm_handle *h, int64_t *ids, double *ak, double *bk);

/* ---- learned order / tree-level selection (no counterpart in the reference,
 * whose order is one global P_max, global.cpp:311, and whose tree depth
 * follows the fixed 75/50 leaf rule, global.cpp:303-313) ---- */

/* Installs the selection model (host pointers, copied; params has
 * sel->n_params floats).  sel = NULL removes it: learned calls then use the
 * fixed order. */
int vpmfmm_set_selector(vpmfmm_handle *h, const vpmfmm_selector *sel, const float *params);

/* The orders, threshold and predictions used by the last field/velocity call. */
int vpmfmm_last_selection(vpmfmm_handle *h, vpmfmm_selection *out);

/* Feature rows of the tree of the last call (order column 0), levels 0..16:
 * feats[17 * 9]; raw[17 * 5] = cells, particles, sum|q|, max|q|, leaves per
 * level; *n_levels = deepest level + 1.  Builds the training data. */
int vpmfmm_level_features(vpmfmm_handle *h, float *feats, double *raw, int32_t *n_levels);

/* MLP kernels on the handle's stream; every pointer is a device pointer.
 * x[rows * 9] are normalised feature rows, params[VPMFMM_SEL_PARAMS].
 * forward: out[rows * 2] = (e, t).
 * loss_grad: solves s = 0..nseg-1 own rows seg[s] .. seg[s+1]-1 (non-empty),
 *   y[2s] = ln(err^2), y[2s+1] = ln(ms); loss = mean_s (LSE_s(e) - y[2s])^2
 *   + w_time (LSE_s(t) - y[2s+1])^2 into *loss, its gradient into grad.
 * train: `steps` Adam steps (beta 0.9 / 0.999, eps 1e-8, cosine learning rate
 *   from lr to lr_min, bias correction counted from step0 + 1) updating params
 *   and the moments m, v in place; loss_hist[steps] receives each step's loss. */
int vpmfmm_mlp_forward(vpmfmm_handle *h, const float *params, const float *x, int32_t rows,
                       float *out);
int vpmfmm_mlp_loss_grad(vpmfmm_handle *h, const float *params, const float *x, int32_t rows,
                         const int32_t *seg, int32_t nseg, const float *y, float w_time,
                         float *grad, float *loss);
int vpmfmm_mlp_train(vpmfmm_handle *h, float *params, float *m, float *v, const float *x,
                     int32_t rows, const int32_t *seg, int32_t nseg, const float *y, float w_time,
                     int32_t steps, int32_t step0, float lr, float lr_min, float *loss_hist);

#ifdef __cplusplus
}
#endif
#endif
