/* sanitize_driver.c -- runs every entry point of the C oracle on small seeded
 * fields, for builds under AddressSanitizer + UndefinedBehaviorSanitizer and
 * ThreadSanitizer (test infrastructure only; see the Makefile targets asan and
 * tsan).  Exits non-zero on a wrong status or a non-finite result; the
 * sanitizers report (and, with halt_on_error, stop at) any memory, undefined
 * behaviour or data-race error. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "fmm2d_oracle.h"
#include "fmm2d_type1_oracle.h"
#include "fmm3d_oracle.h"

static uint64_t g_state = 0x243F6A8885A308D3ull;

static double urand(void) { /* splitmix64 -> [0, 1) */
    g_state += 0x9E3779B97F4A7C15ull;
    uint64_t z = g_state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static int finite_all(const double *a, int n) {
    for (int i = 0; i < n; ++i)
        if (!isfinite(a[i])) return 0;
    return 1;
}

#define CHECK(cond, what)                                            \
    do {                                                             \
        if (!(cond)) {                                               \
            fprintf(stderr, "sanitize_driver: %s failed\n", (what)); \
            return 1;                                                \
        }                                                            \
    } while (0)

/* One field: clustered points (several clusters of different spread), some
 * inactive, sizes for the regularized kernels. */
static int run_field(int n, int clusters, double inactive) {
    if (n < 2) { /* one particle: a degenerate domain (root length 0) */
        double x1 = 0.5, y1 = -0.5;
        uint8_t a1 = 1;
        orc_opts o1;
        orc_default_opts(&o1);
        orc_tree *t1 = NULL;
        CHECK(orc_tree_build(&t1, 1, &x1, &y1, &a1, &o1) == ORC_DEGENERATE, "one particle");
        if (t1) orc_tree_free(t1);
        return 0;
    }
    double *x = malloc(sizeof(double) * n), *y = malloc(sizeof(double) * n);
    double *q = malloc(sizeof(double) * n), *s = malloc(sizeof(double) * n);
    double *w = malloc(sizeof(double) * n);
    uint8_t *a = malloc((size_t)n);
    double *ex = malloc(sizeof(double) * n), *ey = malloc(sizeof(double) * n);
    int *tidx = malloc(sizeof(int) * n);
    for (int i = 0; i < n; ++i) {
        const int c = (int)(urand() * clusters);
        const double cx = -1.0 + 2.0 * (double)c / clusters, sp = 0.02 + 0.3 * (c % 3);
        x[i] = cx + sp * (urand() - 0.5);
        y[i] = 0.5 * cx + sp * (urand() - 0.5);
        s[i] = 0.01 + 0.01 * urand();
        w[i] = 2.0 * urand() - 1.0;
        q[i] = -w[i] * s[i] * s[i] / (2.0 * M_PI);
        a[i] = urand() >= inactive;
        tidx[i] = (int)(urand() * n);
    }
    orc_opts o;
    orc_default_opts(&o);
    for (int rule = 0; rule < 3; ++rule) {
        o.max_par = rule == 0 ? 75 : (rule == 1 ? 8 : 40);
        o.max_src = rule == 0 ? 50 : (rule == 1 ? 8 : 40);
        o.max_level = rule == 2 ? 5 : 0;
        orc_tree *t = NULL;
        CHECK(orc_tree_build(&t, n, x, y, a, &o) == ORC_OK, "orc_tree_build");
        int nc, nl, lo, hi;
        CHECK(orc_tree_counts(t, &nc, &nl, &lo, &hi) == ORC_OK, "orc_tree_counts");
        double R, piv[2];
        orc_tree_geometry(t, &R, piv);
        int64_t *id = malloc(sizeof(int64_t) * nc);
        int *lev = malloc(sizeof(int) * nc), *ix = malloc(sizeof(int) * nc);
        int *iy = malloc(sizeof(int) * nc), *pn = malloc(sizeof(int) * nc);
        int *sn = malloc(sizeof(int) * nc);
        uint8_t *leaf = malloc((size_t)nc), *act = malloc((size_t)nc);
        orc_tree_cells(t, id, lev, ix, iy, pn, sn, leaf, act);
        int64_t *lid = malloc(sizeof(int64_t) * nl);
        int *poff = malloc(sizeof(int) * (nl + 1)), *pidx = malloc(sizeof(int) * n);
        orc_tree_leaves(t, lid, poff, pidx);
        CHECK(poff[nl] == n, "leaf particle count");
        enum { CAP = 64 };
        int64_t l1[CAP], l2[CAP], l3[CAP], l4[CAP];
        for (int j = 0; j < nc; ++j) {
            int n1, n2, n3, n4;
            if (leaf[j]) orc_int_list(t, id[j], l1, &n1, l3, &n3, CAP);
            if (lev[j] >= 2) orc_ext_list(t, id[j], l2, &n2, l4, &n4, CAP);
        }
        free(id); free(lev); free(ix); free(iy); free(pn); free(sn); free(leaf); free(act);
        free(lid); free(poff); free(pidx);
        orc_tree_free(t);
        for (int p = 4; p <= 16; p += 6)
            for (int bm = 0; bm <= 1; ++bm)
                for (int l4m = 1; l4m <= 2; ++l4m) {
                    o.p = p;
                    o.binom_mode = bm;
                    o.list4_mode = l4m;
                    CHECK(orc_fmm_field(n, x, y, q, a, &o, ex, ey) == ORC_OK, "orc_fmm_field");
                    CHECK(finite_all(ex, n) && finite_all(ey, n), "finite field");
                }
        o.p = 10;
        o.binom_mode = 1;
        o.list4_mode = 1;
        for (int ic = 0; ic <= 3; ++ic) {
            o.icutoff = ic;
            CHECK(orc_velocity(n, x, y, s, w, a, &o, 1.0, 0.0, ex, ey) == ORC_OK, "orc_velocity");
            CHECK(finite_all(ex, n) && finite_all(ey, n), "finite velocity");
        }
        o.icutoff = 0;
        for (int l4m = 1; l4m <= 2; ++l4m) { /* the potential (calcPotential) */
            o.list4_mode = l4m;
            CHECK(orc_fmm_potential(n, x, y, q, a, &o, ex) == ORC_OK, "orc_fmm_potential");
            CHECK(finite_all(ex, n), "finite potential");
        }
        o.list4_mode = 1;
        /* the expansions of every cell (count from a tree of the same options) */
        orc_tree *t2 = NULL;
        CHECK(orc_tree_build(&t2, n, x, y, a, &o) == ORC_OK, "orc_tree_build (expansions)");
        orc_tree_counts(t2, &nc, &nl, &lo, &hi);
        orc_tree_free(t2);
        int64_t *ids = malloc(sizeof(int64_t) * nc);
        double *ak = malloc(sizeof(double) * 2 * (o.p + 1) * (size_t)nc);
        double *bk = malloc(sizeof(double) * 2 * (o.p + 1) * (size_t)nc);
        CHECK(orc_fmm_expansions(n, x, y, q, a, &o, ids, ak, bk) == ORC_OK, "orc_fmm_expansions");
        free(ids); free(ak); free(bk);
    }
    /* direct sum over all targets, a target list with repeats, and a source mask */
    orc_direct_field(n, x, y, q, NULL, n, NULL, ex, ey);
    CHECK(finite_all(ex, n), "direct (all)");
    orc_direct_field(n, x, y, q, a, n / 3, tidx, ex, ey);
    CHECK(finite_all(ex, n / 3), "direct (targets, mask)");
    orc_direct_potential(n, x, y, q, a, n / 3, tidx, ex);
    CHECK(finite_all(ex, n / 3), "direct potential");
    /* stage entry points */
    double ak[2 * 17] = {0}, bk[2 * 17] = {0}, fx = 0, fy = 0;
    orc_p2m(n < 50 ? n : 50, x, y, q, 0.1, -0.2, 16, ak);
    orc_m2m(ak, 0.05, -0.03, 16, 1, bk);
    orc_m2l0(ak, 3.0, 1.0, 16, bk);
    orc_m2l(ak, 3.0, 1.0, 16, 1, bk);
    orc_l2l(bk, 0.02, 0.01, 16, 0, ak);
    orc_m2p(ak, 2.0, 1.5, 16, &fx, &fy);
    orc_l2p(bk, 0.01, 0.02, 16, &fx, &fy);
    CHECK(isfinite(fx) && isfinite(fy), "stage entry points");
    for (int nn = 0; nn < 64; ++nn)
        for (int k = 0; k <= nn + 1; ++k) (void)orc_binom_ref(nn, k), (void)orc_binom_exact(nn, k);
    /* the type-1 path: velocity (every kernel, both L2L modes), tree and
     * lists, direct sum */
    for (int ic = 0; ic <= 3; ++ic) {
        orc_t1_opts t1;
        orc_t1_default_opts(&t1);
        t1.icutoff = ic;
        t1.n_s = 5 + 10 * ic;
        t1.lmax = 4 + ic;
        t1.ndp = 3 + 3 * ic;
        t1.l2l_mode = ic % 2;
        CHECK(orc_t1_velocity(n, x, y, s, q, &t1, ex, ey) == 0, "type-1 velocity");
        CHECK(finite_all(ex, n) && finite_all(ey, n), "type-1 velocity finite");
    }
    {
        orc_t1_opts t1;
        orc_t1_default_opts(&t1);
        t1.n_s = 4;
        orc_t1_tree *tt = NULL;
        CHECK(orc_t1_build(&tt, n, x, y, &t1) == 0, "type-1 tree");
        const int nb = orc_t1_nboxes(tt);
        int *lst = (int *)malloc(sizeof(int) * 4096);
        int *lv = (int *)malloc(sizeof(int) * (nb + 1));
        double *bd = (double *)malloc(sizeof(double) * 4 * (nb + 1));
        orc_t1_boxes(tt, lv, NULL, NULL, NULL, NULL, NULL, bd);
        for (int b = 0; b < nb; ++b)
            for (int w = 1; w <= 4; ++w) CHECK(orc_t1_list(tt, b, w, lst, 4096) >= 0, "type-1 list");
        if (nb > 0) CHECK(orc_t1_box_particles(tt, 0, lst, 4096) > 0, "type-1 box particles");
        free(lst); free(lv); free(bd);
        orc_t1_free(tt);
        double coef[121];
        int expo[121];
        orc_t1_l2l_table(10, 1, coef, expo);
        orc_t1_direct(n, x, y, s, q, 2, n / 3, tidx, ex, ey);
        CHECK(finite_all(ex, n / 3), "type-1 direct");
    }
    /* the 3-D path: the (x, y, q) field lifted to 3-D with z from the sizes,
     * tree, velocity at two leaf thresholds, direct sum */
    {
        double *z = (double *)malloc(sizeof(double) * n);
        double *ez = (double *)malloc(sizeof(double) * n);
        for (int i = 0; i < n; ++i) z[i] = 50.0 * s[i] - 0.5 + 0.3 * x[i];
        for (int crit = 3; crit <= 50; crit += 47) {
            orc3_opts o3;
            orc3_default_opts(&o3);
            o3.crit_num = crit;
            orc3_tree *t3 = NULL;
            CHECK(orc3_build(&t3, n, x, y, z, a, &o3) == 0, "3-D tree");
            const int nc = orc3_ncells(t3);
            int *lv = (int *)malloc(sizeof(int) * nc), *lst = (int *)malloc(sizeof(int) * n);
            int64_t *key = (int64_t *)malloc(sizeof(int64_t) * nc);
            double *cen = (double *)malloc(sizeof(double) * 3 * nc);
            orc3_cells(t3, lv, key, NULL, NULL, NULL, cen, NULL);
            for (int c = 0; c < nc; ++c) CHECK(orc3_leaf_particles(t3, c, lst, n) >= 0, "3-D leaf");
            free(lv); free(lst); free(key); free(cen);
            orc3_free(t3);
            CHECK(orc3_velocity(n, x, y, z, s, q, w, q, a, &o3, 1.0, 0.0, 0.5, ex, ey, ez) == 0,
                  "3-D velocity");
            CHECK(finite_all(ex, n) && finite_all(ey, n) && finite_all(ez, n), "3-D velocity finite");
        }
        {
            /* type 2: the TreeCell octree (a level-1 leaf beside deeper
             * cells is status 6, as the reference throws there) */
            orc3_opts o3;
            orc3_default_opts(&o3);
            o3.type = 2;
            orc3_tree *t3 = NULL;
            const int bs = orc3_build(&t3, n, x, y, z, a, &o3);
            CHECK(bs == 0 || bs == 3, "3-D type-2 tree");
            if (t3) orc3_free(t3);
            const int vs = orc3_velocity(n, x, y, z, s, q, w, q, a, &o3, 1.0, 0.0, 0.5, ex, ey, ez);
            CHECK(vs == 0 || vs == 3 || vs == 6, "3-D type-2 velocity");
            if (vs == 0)
                CHECK(finite_all(ex, n) && finite_all(ey, n) && finite_all(ez, n),
                      "3-D type-2 velocity finite");
        }
        orc3_direct(n, x, y, z, q, w, q, a, n / 3, tidx, ex, ey, ez);
        CHECK(finite_all(ez, n / 3), "3-D direct");
        free(z); free(ez);
    }
    free(x); free(y); free(q); free(s); free(w); free(a); free(ex); free(ey); free(tidx);
    return 0;
}

int main(void) {
    static const int sizes[][2] = {{1, 1}, {2, 1}, {3, 2}, {257, 3}, {1500, 5}, {4000, 7}};
    for (size_t i = 0; i < sizeof(sizes) / sizeof(sizes[0]); ++i)
        if (run_field(sizes[i][0], sizes[i][1], i % 2 ? 0.2 : 0.0)) return 1;
    /* invalid inputs take their error paths */
    double nanx[3] = {0.0, NAN, 1.0}, y3[3] = {0.0, 0.5, 1.0};
    uint8_t a3[3] = {1, 1, 1};
    orc_opts o;
    orc_default_opts(&o);
    orc_tree *t = NULL;
    CHECK(orc_tree_build(&t, 3, nanx, y3, a3, &o) == ORC_NONFINITE, "non-finite input status");
    double same[4] = {0.3, 0.3, 0.3, 0.3};
    uint8_t a4[4] = {1, 1, 1, 1};
    o.max_par = o.max_src = 1;
    t = NULL;
    CHECK(orc_tree_build(&t, 4, same, same, a4, &o) != ORC_OK, "coincident points status");
    if (t) orc_tree_free(t);
    printf("sanitize_driver: all entry points ran\n");
    return 0;
}
