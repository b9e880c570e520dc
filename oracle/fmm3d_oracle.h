/*
 * fmm3d_oracle.h -- CPU oracle for the reference's 3-D velocity paths.
 * Type 1 (the default): VelocityCalc::velocity_fmm_3d_fast
 * (velocity_calc.cpp:2384-2525, VELOCITY_3D_TYPE 1) ->
 * fmm3D::calcVelocityFast (fast3DFMM.cpp:32-92):
 *   - octree                FMMCell::generateTree  fast3DFMM.cpp:485-745
 *   - P2M, M2M              fastFMM3Dutils::multipoleCalc  :164-317,
 *                           fmm3D::P2M_calc / M2M_calc  fmm3D.cpp:84-145
 *   - evaluation            fastFMM3Dutils::evaluateFMM  :318-484,
 *                           fmm3D::M2P_mul_calc  fmm3D.cpp:147-210
 * Type 2: VelocityCalc::velocity_fmm_3d (velocity_calc.cpp:2172-2383,
 * VELOCITY_3D_TYPE 2) -> TreeCell::initializeTree in 3-D
 * (treeCell.cpp:1043-1243) and fmm3D::calcVelocity (fmm3D.cpp:1064-1205 with
 * setupVelocityFMM :585-700, VELOCITY_CALCULATION_TYPE 2, :24).
 *
 * TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/, never by the
 * product library.
 *
 * The reference's 3-D code cannot be compiled here as a parity reference:
 * global.hpp:85 fixes DIM at 2, and FMMCell::generateTree writes
 * minDom[2] into arrays of DIM entries.  This restatement is pinned by
 * mathematics instead (tests/test_oracle_fmm3d.py): the tree's invariants,
 * the M2M / P2M moment identity, the M2P multipliers against derivatives of
 * the kernel, the evaluation against the 3-D direct sum with the opening
 * test forced open, and the treecode's error with the reference's test.
 */
#ifndef FMM3D_ORACLE_H
#define FMM3D_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int crit_num;       /* type 1: a cell with fewer particles is a leaf (Pars::src_count_max
                           of the DIM == 3 parameters = 30)                                */
    double exp_tree;    /* root expansion in percent (Pars::expTree = 1.0)                 */
    double center[3];   /* the box always contains this point (Pars::x/y/zcenter = 0)      */
    double support;     /* opening factor: a cell is far when r > support * size
                           (SUPPORT_RADIUS_FACTOR = 2.0, fast3DFMM.cpp:4, fmm3D.cpp:3) */
    int max_depth;      /* refuse deeper trees (the reference has no bound)               */
    int type;           /* 1: fast3DFMM (FMMCell octree, VELOCITY_3D_TYPE 1, the default);
                           2: fmm3D::calcVelocity over the 3-D TreeCell (VELOCITY_3D_TYPE 2) */
    int max_par, max_src; /* type 2: a cell is a leaf when it holds at most max_par
                           particles and max_src active ones (Pars::par_count_max 50,
                           src_count_max 30)                                               */
} orc3_opts;

void orc3_default_opts(orc3_opts *o);

/* Tree (opaque). */
typedef struct orc3_tree orc3_tree;

/* Returns 0, 1 bad arguments, 2 non-finite input, 3 every particle at one
 * point (type 2), 4 depth beyond max_depth, 5 out of memory.  The velocity
 * calls also return 6 for a type-2 tree with a level-1 leaf, where the
 * reference's get_childID throws. */
int  orc3_build(orc3_tree **out, int n, const double *x, const double *y, const double *z,
                const uint8_t *active, const orc3_opts *o);
void orc3_free(orc3_tree *t);
int  orc3_ncells(const orc3_tree *t);
int  orc3_depth(const orc3_tree *t);   /* maxLevel */
/* cells in creation order: level, path key (3 bits per level, octant
 * x + 2y + 4z, level 1 in the highest used bits), particle count, leaf
 * (nPar < crit_num), active, centre (3) and size */
void orc3_cells(const orc3_tree *t, int *level, int64_t *key, int *count, uint8_t *leaf,
                uint8_t *active, double *centre3, double *size);
/* particles of a leaf in the reference's order (original index order) */
int  orc3_leaf_particles(const orc3_tree *t, int cell, int *out, int cap);

/* u, v, w of every particle (rotational part: no free stream) from the
 * vortex strengths alpha = omega s^3 / (4 pi), as calcVelocityFast. */
int orc3_velocity_alpha(int n, const double *x, const double *y, const double *z,
                        const double *ax, const double *ay, const double *az,
                        const uint8_t *active, const orc3_opts *o, double *u, double *v,
                        double *w);
/* The same from vorticity and size (velocity_fmm_3d_fast forms alpha,
 * velocity_calc.cpp:2429-2431), plus the free stream (get_velocity,
 * velocity_calc.cpp:221-236). */
int orc3_velocity(int n, const double *x, const double *y, const double *z, const double *s,
                  const double *wx, const double *wy, const double *wz, const uint8_t *active,
                  const orc3_opts *o, double u_inf, double v_inf, double w_inf, double *u,
                  double *v, double *w);

/* The 10 M2P multipliers per component at offset (dx, dy, dz) from the cell
 * centre (M2P_mul_calc). */
void orc3_m2p_mul(double dx, double dy, double dz, double *mx, double *my, double *mz);
/* P2M of one particle into a 10-term moment (P2M_calc), and M2M of a child
 * moment into a parent moment at offset parent - child (M2M_calc). */
void orc3_p2m(double *mp, double src, double dx, double dy, double dz);
void orc3_m2m(double *mp, const double *mpc, double dx, double dy, double dz);

/* OpenMP threads of the calling thread's later parallel regions (the
 * oracle's and any other library's sharing the OpenMP runtime). */
void orc_set_threads(int n);

/* Direct 3-D Biot-Savart sum u_i = sum_{j != i, active j} alpha_j x (x_i - x_j) / r^3
 * at targets tidx (NULL = all).  OpenMP over targets. */
void orc3_direct(int n, const double *x, const double *y, const double *z, const double *ax,
                 const double *ay, const double *az, const uint8_t *active, int nt,
                 const int *tidx, double *u, double *v, double *w);

#ifdef __cplusplus
}
#endif

#endif
