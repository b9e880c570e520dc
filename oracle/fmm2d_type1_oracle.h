/*
 * fmm2d_type1_oracle.h -- CPU oracle for the reference's type-1 ("old") 2-D
 * FMM velocity.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker, never the product: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product library (lsmps_vpm_v2_amd) must never link or call it.
 *
 * It restates, in plain C with its own data structures, the path
 * VelocityCalc::velocity_old (src/velocity_calculation/velocity_calc.cpp:
 * 390-440, VELOCITY_2D_TYPE 1) ->
 * velocity_biot_savart::biotsavart_fmm_2d (biotsavart_fmm_2d.cpp:4-501):
 *   - box hierarchy          subroutine_all_fmm_2d.cpp:97-287, 635-678
 *   - lists 1-4              subroutine_all_fmm_2d.cpp:291-556
 *   - binomial helper bico   subroutine_all_fmm_2d.cpp:560-591
 *   - near-field pair sum    subroutine_all_fmm_2d.cpp:595-629
 *   - regularization         regularization_function_2d.cpp:11-37
 *   - P2M, M2M, M2L, P2L, L2L, M2P, L2P   biotsavart_fmm_2d.cpp:120-419
 *
 * Pinning (tests/test_oracle_type1.py):
 *   everything -> the reference's own three sources compiled into oracle/_ref
 *           (tree, lists and velocity; committed goldens written from it);
 *   box hierarchy and lists -> brute-force geometric classification and the
 *           exactly-once coverage of every (target, source) pair;
 *   velocity -> convergence to the regularized direct sum with l2l_mode 0,
 *           and the planted L2L of the reference (l2l_mode 1) told apart.
 */
#ifndef FMM2D_TYPE1_ORACLE_H
#define FMM2D_TYPE1_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Options.  Defaults are the reference's global.cpp:316-328. */
typedef struct {
    int n_s;      /* boxes holding more than n_s particles split (Pars::n_s = 50)    */
    int lmax;     /* deepest level, 2..16 (Pars::lmax = 7)                            */
    int ndp;      /* multipole terms (orders 0..ndp-1); the local expansion has
                     ndp + 1 terms (orders 0..ndp) (Pars::ndp = 10)                   */
    int icutoff;  /* near-field kernel of regul_func_2d: 0 singular, 1 high-order
                     algebraic, 2 Gaussian, 3 super-Gaussian (Pars::icutoff = 2)       */
    int l2l_mode; /* 1 = L2L as the reference runs it (bico rewrites its k argument
                     in place, which changes the shift's later binomials and
                     exponents, biotsavart_fmm_2d.cpp:299-307); 0 = the exact shift   */
} orc_t1_opts;

/* Pars::tol2 (global.cpp:325): the absolute tolerance of the box tests. */
#define ORC_T1_TOL2 1.0e-15

void orc_t1_default_opts(orc_t1_opts *o);

/* bico (subroutine_all_fmm_2d.cpp:560-591): nCk in double, rounding as the
 * reference does; *k is rewritten to n - k where the reference rewrites it. */
double orc_t1_bico(int n, int *k);

/* The L2L shift as run: child b_l += coef[l][kp] d^expo[l][kp] parent b_kp for
 * kp = l..ndp ((ndp + 1)^2 row-major entries; entries below the diagonal 0). */
void orc_t1_l2l_table(int ndp, int l2l_mode, double *coef, int *expo);

/* Tree (opaque). */
typedef struct orc_t1_tree orc_t1_tree;

/* Builds the box hierarchy of hierarchy_mesh over n particles.  Returns 0, or
 * 1 for bad arguments, 2 for non-finite input, 5 out of memory, 7 when the
 * domain is too small for the reference's absolute box tolerance (its finest
 * box side below 64 (tol2 + lmax |coord|max eps)). */
int  orc_t1_build(orc_t1_tree **out, int n, const double *x, const double *y,
                  const orc_t1_opts *o);
void orc_t1_free(orc_t1_tree *t);
int  orc_t1_nboxes(const orc_t1_tree *t);
int  orc_t1_depth(const orc_t1_tree *t);  /* lev: the deepest level with boxes */
/* particles that fell between the rounded bounds of a splitting box's
 * children (in no box below it: no velocity, no contribution) */
int  orc_t1_nlost(const orc_t1_tree *t);
/* root after padding and squaring: xmin, xmax, ymin, ymax, db (level-1 side) */
void orc_t1_root(const orc_t1_tree *t, double *geo5);
/* boxes in the reference's order (level by level, creation order inside a
 * level); any output may be NULL.  bounds: xmin, xmax, ymin, ymax per box. */
void orc_t1_boxes(const orc_t1_tree *t, int *level, int *ix, int *iy, int *count,
                  int *nchild, int *parent, double *bounds4);
/* particles of box b in the reference's order (original index order) */
int  orc_t1_box_particles(const orc_t1_tree *t, int b, int *out, int cap);
/* list which = 1..4 of box b (reference order; list 1 and 3 of leaves only,
 * list 2 and 4 of every box); returns its length (entries beyond cap dropped) */
int  orc_t1_list(const orc_t1_tree *t, int b, int which, int *out, int cap);

/* Rotational velocity (no free stream) of biotsavart_fmm_2d(pi, pi, ...):
 * u, v at every particle from strengths g (Particle::gz) and sizes s (used by
 * icutoff > 0).  Same return codes as orc_t1_build. */
int orc_t1_velocity(int n, const double *x, const double *y, const double *s, const double *g,
                    const orc_t1_opts *o, double *u, double *v);

/* Regularized direct sum of the same kernel over every pair at r > 0
 * (biotsavart_direct_2d's sum with the type-1 near-field factor), at targets
 * tidx[0..nt) (NULL = all particles).  OpenMP over targets. */
void orc_t1_direct(int n, const double *x, const double *y, const double *s, const double *g,
                   int icutoff, int nt, const int *tidx, double *u, double *v);

#ifdef __cplusplus
}
#endif

#endif
