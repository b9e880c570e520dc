/*
 * fmm2d_oracle.h -- CPU oracle for the 2-D FMM Biot-Savart velocity path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker, never the product: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product library (lsmps_vpm_v2_amd) must never link or call it.
 *
 * It restates, in plain C with its own data structures, the reference
 * algorithm of anggarapangestu/LSMPS_VPM_v2:
 *   - adaptive quadtree build        src/FMM/treeCell.cpp:1043-1243
 *   - same-level neighbours          src/FMM/treeCell.cpp:562-625
 *   - lists 1/3 (intList)            src/FMM/treeCell.cpp:636-804
 *   - lists 2/4 (extList)            src/FMM/treeCell.cpp:818-1032
 *   - P2M, M2M, M2L, L2L             src/FMM/fmm2D.cpp:237-284, 309-379, 409-556, 587-651
 *   - P2P, M2P (list 3), L2P         src/FMM/fmm2D.cpp:135-179, 1040-1066, 1078-1093
 *   - binomial helper (int32)        src/FMM/fmm2D.cpp:28-58
 *   - velocity packing / write-back  src/velocity_calculation/velocity_calc.cpp:465-481, 572-575
 * plus an O(N^2) FP64 singular direct sum used as the accuracy reference.
 *
 * Pinning (what checks each function, see tests/test_oracle_*.py):
 *   tree build, lists 1-4, binomial table, end-to-end field
 *        -> the reference's own sources compiled into oracle/_ref (bit-exact
 *           tree/list membership, field to 1e-12) and the committed golden
 *           fixtures written from it;
 *   lists -> brute-force O(C^2) geometric classification and the
 *           exactly-once coverage invariant;
 *   P2M/M2M/L2L/M2P/L2P -> exact series identities on a few particles;
 *   M2L   -> convergence to the direct sum with exact binomials;
 *   direct sum -> brute force on tiny inputs and the Perlman closed form.
 */
#ifndef FMM2D_ORACLE_H
#define FMM2D_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Option block of the oracle.  Defaults mirror the reference's global.cpp:303-313. */
typedef struct {
    int    p;          /* expansion order (Pars::P_max = 10)                        */
    int    max_par;    /* leaf rule: particles per cell (Pars::par_count_max = 75)   */
    int    max_src;    /* leaf rule: active particles per cell (Pars::src_count_max) */
    double exp_tree;   /* root expansion in percent (Pars::expTree = 1.0)           */
    int    min_level;  /* cells above this level are always split (reference: 1)    */
    int    max_level;  /* cells at this level are always leaves (0: no cap)         */
    int    binom_mode; /* 0: exact binomials, 1: the reference int32 helper (M2M,
                          M2L and L2L all call it)                                  */
    int    list4_mode; /* 1: M2L from the list-4 leaf multipole (reference),
                          2: P2L from the list-4 leaf's active particles            */
    int    stage_mask; /* evaluation parts: 1 P2P, 2 M2P (list 3), 4 L2P; 0 = all  */
    int    icutoff;    /* near-field (P2P) kernel, regul_func_2d
                          (regularization_function_2d.cpp:11-37): 0 singular,
                          1 high-order algebraic, 2 Gaussian, 3 super-Gaussian;
                          > 0 needs the particle sizes (orc_velocity)            */
} orc_opts;

/* Status codes (shared numbering with the product C-ABI). */
enum {
    ORC_OK = 0,
    ORC_BAD_ARG = 1,
    ORC_NONFINITE = 2,
    ORC_DEGENERATE = 3,      /* all particles at one point: root length zero      */
    ORC_DEPTH = 4,           /* tree deeper than ORC_MAX_DEPTH                    */
    ORC_NOMEM = 5,
    ORC_INTERNAL = 6
};

#define ORC_MAX_DEPTH 16

void orc_default_opts(orc_opts *o);

/* Tree handle (opaque to callers). */
typedef struct orc_tree orc_tree;

int  orc_tree_build(orc_tree **out, int n, const double *x, const double *y,
                    const uint8_t *active, const orc_opts *o);
void orc_tree_free(orc_tree *t);

/* Tree inspection. Cells are reported sorted by reference cell ID. */
int  orc_tree_counts(const orc_tree *t, int *n_cells, int *n_leaves,
                     int *min_level, int *max_level);
void orc_tree_geometry(const orc_tree *t, double *root_len, double *pivot2);
void orc_tree_cells(const orc_tree *t, int64_t *id, int *level, int *ix, int *iy,
                    int *par_num, int *src_num, uint8_t *is_leaf, uint8_t *is_active);
/* Leaf list in reference order and each leaf's particle indices (index order). */
void orc_tree_leaves(const orc_tree *t, int64_t *leaf_id, int *poff, int *pidx);

/* Lists of one cell (identified by reference ID).  Returns counts; the
 * output buffers must hold at least cap entries (counts beyond cap are
 * still reported, entries are not written). */
int  orc_int_list(const orc_tree *t, int64_t id, int64_t *l1, int *n1,
                  int64_t *l3, int *n3, int cap);
int  orc_ext_list(const orc_tree *t, int64_t id, int64_t *l2, int *n2,
                  int64_t *l4, int *n4, int cap);

/* Reference-semantics FMM field: ex, ey per particle (fmm2D::calcField).
 * q is the source value (-omega s^2 / 2pi in the reference packing). */
int  orc_fmm_field(int n, const double *x, const double *y, const double *q,
                   const uint8_t *active, const orc_opts *o,
                   double *ex, double *ey);

/* The same run, returning every cell's multipole (ak) and local (bk)
 * coefficients, cells sorted by ID: ids[n_cells], ak/bk[n_cells * 2(p+1)]
 * (re, im interleaved).  b_0 is not formed (it never reaches the field). */
int  orc_fmm_expansions(int n, const double *x, const double *y, const double *q,
                        const uint8_t *active, const orc_opts *o, int64_t *ids,
                        double *ak, double *bk);

/* Velocity as VelocityCalc::get_velocity with the FMM branch:
 * u = Ey + u_inf, v = -Ex + v_inf, q = -omega s^2/(2 pi). */
int  orc_velocity(int n, const double *x, const double *y, const double *s,
                  const double *omega, const uint8_t *active, const orc_opts *o,
                  double u_inf, double v_inf, double *u, double *v);

/* Reference-semantics FMM potential (fmm2D::calcPotential, fmm2D.cpp:681-896):
 * phi_i = sum_{j != i} q_j log r_ij, from P2P (potDirSum, :77-122), M2P over
 * list 3 and L2P with the order-0 local coefficient b_0 that the M2L of
 * setupFMM forms (:463-467, 504-508; P2L :540-548).  icutoff must be 0. */
int  orc_fmm_potential(int n, const double *x, const double *y, const double *q,
                       const uint8_t *active, const orc_opts *o, double *phi);

/* Direct potential sum phi_i = sum_j q_j log(r_ij^2) / 2 (j == i and, with
 * src_mask, masked sources skipped) at targets tidx or all n particles. */
void orc_direct_potential(int n, const double *x, const double *y, const double *q,
                          const uint8_t *src_mask, int nt, const int *tidx, double *phi);

/* Singular FP64 direct sum at targets tidx[0..nt) (indices into the source
 * arrays, so j == i is skipped), or at all n particles when tidx == NULL.
 * src_mask (may be NULL) restricts the sources.  OpenMP over targets. */
void orc_direct_field(int n, const double *x, const double *y, const double *q,
                      const uint8_t *src_mask, int nt, const int *tidx,
                      double *ex, double *ey);

/* Binomials: the reference int32 helper (fmm2D.cpp:28-58, two's-complement
 * wrap on overflow) and the exact value. */
int    orc_binom_ref(int n, int k);
double orc_binom_exact(int n, int k);

/* Stage-level entry points for the series-identity tests. */
void orc_p2m(int np, const double *x, const double *y, const double *q,
             double cx, double cy, int p, double *ak /* 2(p+1) re,im */);
void orc_m2m(const double *ak_child, double dx, double dy, int p, int binom_mode,
             double *ak_parent_acc);
void orc_m2l(const double *ak_src, double dx, double dy, int p, int binom_mode,
             double *bk_acc /* d = z_src - z_tgt */);
/* the order-0 local coefficient of the potential (b_0 += ..., d = z_src - z_tgt) */
void orc_m2l0(const double *ak_src, double dx, double dy, int p, double *bk_acc);
void orc_l2l(const double *bk_parent, double dx, double dy, int p, int binom_mode,
             double *bk_child_acc);
void orc_m2p(const double *ak, double dx, double dy, int p, double *ex, double *ey);
void orc_l2p(const double *bk, double dx, double dy, int p, double *ex, double *ey);

#ifdef __cplusplus
}
#endif
#endif
