// ref_harness.cpp -- C entry points around the reference's own FMM path.
//
// TEST INFRASTRUCTURE ONLY.  Compiled (with the reference's TreeCell and
// fmm2D sources and its type-1 FMM sources biotsavart_fmm_2d.cpp,
// subroutine_all_fmm_2d.cpp and regularization_function_2d.cpp, straight
// from the reference tree) into
// oracle/_ref/libref_fmm2d.so by oracle/Makefile.  It replays the packing of
// VelocityCalc::velocity_fmm_2d (velocity_calc.cpp:465-481) and calls
// TreeCell::initializeTree (treeCell.cpp:1043) and fmm2D::calcField
// (fmm2D.cpp:907), exposing plain C functions for the golden-vector script,
// the oracle tests and bench.py's CPU baseline.
//
// fmm2D::binComb is private; the harness widens access for its own
// translation unit only, to read the helper's compiled values.  Every
// standard header the reference headers pull in is included first, so the
// access change touches only the reference's own declarations.
#include <fcntl.h>
#include <omp.h>
#include <time.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <exception>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#define private public
#include "src/FMM/fmm2D.hpp"
#undef private
#include "src/velocity_calculation/velocity_biot_savart.hpp"

extern "C" void ref_pars_set(int p, int max_par, int max_src, double exp_tree);
extern "C" void ref_pars_set_type1(int n_s, int lmax, int ndp, int icutoff);

namespace {

TreeCell *g_tree = nullptr;

// The reference prints timings and logs on stdout; the harness mutes fd 1
// while it runs so that callers keep a clean stdout.
struct MuteStdout {
    int saved = -1;
    MuteStdout() {
        std::cout.flush();
        std::fflush(stdout);
        saved = dup(1);
        int fd = open("/dev/null", O_WRONLY);
        if (fd >= 0) {
            dup2(fd, 1);
            close(fd);
        }
    }
    ~MuteStdout() {
        std::cout.flush();
        std::fflush(stdout);
        if (saved >= 0) {
            dup2(saved, 1);
            close(saved);
        }
    }
};

void pack(int n, const double *x, const double *y, const uint8_t *active,
          std::vector<std::vector<double>> &pos, std::vector<bool> &mark) {
    pos.assign(n, std::vector<double>(2, 0.0));
    mark.assign(n, false);
    for (int i = 0; i < n; ++i) {
        pos[i][0] = x[i];
        pos[i][1] = y[i];
        mark[i] = active[i] != 0;
    }
}

}  // namespace

extern "C" {

int ref_set_params(int p, int max_par, int max_src, double exp_tree) {
    ref_pars_set(p, max_par, max_src, exp_tree);
    return 0;
}

void ref_set_threads(int n) { omp_set_num_threads(n); }

int ref_binom(int n, int k) {
    fmm2D f;
    return f.binComb(n, k);
}

int ref_build_tree(int n, const double *x, const double *y, const uint8_t *active) {
    std::vector<std::vector<double>> pos;
    std::vector<bool> mark;
    pack(n, x, y, active, pos, mark);
    delete g_tree;
    g_tree = new TreeCell();
    try {
        MuteStdout mute;
        g_tree->initializeTree(pos, mark);
    } catch (const std::exception &) {
        delete g_tree;
        g_tree = nullptr;
        return 6;
    }
    return 0;
}

int ref_tree_counts(int *n_cells, int *n_leaves, int *min_level, int *max_level) {
    if (!g_tree) return 1;
    *n_cells = (int)g_tree->treeData.size();
    *n_leaves = (int)g_tree->leafList.size();
    *min_level = g_tree->min_level;
    *max_level = g_tree->max_level;
    return 0;
}

int ref_tree_geometry(double *root_len, double *pivot2) {
    if (!g_tree) return 1;
    *root_len = g_tree->rootLength;
    pivot2[0] = g_tree->pivotCoord[0];
    pivot2[1] = g_tree->pivotCoord[1];
    return 0;
}

// All cells sorted by ID.
int ref_tree_cells(int64_t *id, int *level, int *ix, int *iy, int *par_num, int *src_num,
                   uint8_t *is_leaf, uint8_t *is_active) {
    if (!g_tree) return 1;
    std::vector<int> ids;
    ids.reserve(g_tree->treeData.size());
    for (const auto &kv : g_tree->treeData) ids.push_back(kv.first);
    std::sort(ids.begin(), ids.end());
    for (size_t k = 0; k < ids.size(); ++k) {
        const Cell *c = g_tree->treeData.at(ids[k]);
        id[k] = c->ID;
        level[k] = c->level;
        ix[k] = c->index[0];
        iy[k] = c->index[1];
        par_num[k] = c->parNum;
        src_num[k] = c->srcParNum;
        is_leaf[k] = c->isLeaf ? 1 : 0;
        is_active[k] = c->isActive ? 1 : 0;
    }
    return 0;
}

// Leaf list in the reference order, with each leaf's particle list.
int ref_tree_leaves(int64_t *leaf_id, int *poff, int *pidx) {
    if (!g_tree) return 1;
    int off = 0;
    for (size_t k = 0; k < g_tree->leafList.size(); ++k) {
        const Cell *c = g_tree->treeData.at(g_tree->leafList[k]);
        leaf_id[k] = c->ID;
        poff[k] = off;
        for (int p : c->parIDList) pidx[off++] = p;
    }
    poff[g_tree->leafList.size()] = off;
    return 0;
}

int ref_int_list(int64_t id, int64_t *l1, int *n1, int64_t *l3, int *n3, int cap) {
    if (!g_tree) return 1;
    std::vector<int> a, b;
    try {
        MuteStdout mute;
        g_tree->intList((int)id, a, b);
    } catch (const std::exception &) {
        return 6;
    }
    *n1 = (int)a.size();
    *n3 = (int)b.size();
    for (int i = 0; i < *n1 && i < cap; ++i) l1[i] = a[i];
    for (int i = 0; i < *n3 && i < cap; ++i) l3[i] = b[i];
    return 0;
}

int ref_ext_list(int64_t id, int64_t *l2, int *n2, int64_t *l4, int *n4, int cap) {
    if (!g_tree) return 1;
    std::vector<int> a, b;
    try {
        MuteStdout mute;
        g_tree->extList((int)id, a, b);
    } catch (const std::exception &) {
        return 6;
    }
    *n2 = (int)a.size();
    *n4 = (int)b.size();
    for (int i = 0; i < *n2 && i < cap; ++i) l2[i] = a[i];
    for (int i = 0; i < *n4 && i < cap; ++i) l4[i] = b[i];
    return 0;
}

// velocity_fmm_2d's tree + calcField with the source values given directly.
// seconds[0..1] (optional): tree build and FMM wall times.
int ref_fmm_field(int n, const double *x, const double *y, const double *q,
                  const uint8_t *active, double *ex, double *ey, double *seconds) {
    std::vector<std::vector<double>> pos;
    std::vector<bool> mark;
    pack(n, x, y, active, pos, mark);
    std::vector<double> src(q, q + n);
    try {
        MuteStdout mute;
        TreeCell tree;
        double t0 = omp_get_wtime();
        tree.initializeTree(pos, mark);
        double t1 = omp_get_wtime();
        fmm2D fmm;
        fmm.calcField(tree, pos, mark, src);
        std::vector<double> fx = fmm.get_Field_x();
        std::vector<double> fy = fmm.get_Field_y();
        double t2 = omp_get_wtime();
        for (int i = 0; i < n; ++i) {
            ex[i] = fx[i];
            ey[i] = fy[i];
        }
        if (seconds) {
            seconds[0] = t1 - t0;
            seconds[1] = t2 - t1;
        }
    } catch (const std::exception &) {
        return 6;
    }
    return 0;
}

// The same tree + fmm2D::calcPotential and get_Potential
// (fmm2D.cpp:681-896, 1137-1140), the potential variant of the path.
int ref_fmm_potential(int n, const double *x, const double *y, const double *q,
                      const uint8_t *active, double *phi) {
    std::vector<std::vector<double>> pos;
    std::vector<bool> mark;
    pack(n, x, y, active, pos, mark);
    std::vector<double> src(q, q + n);
    try {
        MuteStdout mute;
        TreeCell tree;
        tree.initializeTree(pos, mark);
        fmm2D fmm;
        fmm.calcPotential(tree, pos, mark, src);
        std::vector<double> f = fmm.get_Potential();
        for (int i = 0; i < n; ++i) phi[i] = f[i];
    } catch (const std::exception &) {
        return 6;
    }
    return 0;
}

// The type-1 path: VelocityCalc::velocity_old (velocity_calc.cpp:390-440)
// copies the particles into two Particle sets (targets and sources, the same
// data) and calls velocity_biot_savart::biotsavart_fmm_2d with
// n_inter = 1; u, v are the rotational part (no free stream).
int ref_type1_velocity(int n, const double *x, const double *y, const double *s,
                       const double *g, int n_s, int lmax, int ndp, int icutoff, double *u,
                       double *v) {
    ref_pars_set_type1(n_s, lmax, ndp, icutoff);
    Particle pi, pj;
    pi.num = pj.num = n;
    pi.x.assign(x, x + n);
    pi.y.assign(y, y + n);
    pi.s.assign(s, s + n);
    pi.gz.assign(g, g + n);
    pi.u.assign(n, 0.0);
    pi.v.assign(n, 0.0);
    pj = pi;
    try {
        MuteStdout mute;
        velocity_biot_savart f;
        f.biotsavart_fmm_2d(pi, pj, icutoff, n_s, 1, ndp);
    } catch (const std::exception &) {
        return 6;
    }
    for (int i = 0; i < n; ++i) {
        u[i] = pi.u[i];
        v[i] = pi.v[i];
    }
    return 0;
}

}  // extern "C"
