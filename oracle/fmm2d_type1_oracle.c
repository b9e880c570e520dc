/*
 * fmm2d_type1_oracle.c -- CPU oracle for the reference's type-1 ("old") 2-D
 * FMM velocity (biotsavart_fmm_2d and its subroutines).
 *
 * TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/, smoke() and the
 * cpu_baseline leg of bench.py, never by the product library.
 *
 * The box tests and box bounds follow the reference's floating-point
 * operations exactly (bounds are formed by the same additions, the tests
 * compare against the absolute tolerance Pars::tol2 = 1e-15), so every box,
 * list and particle-membership decision is the reference's.  Complex
 * arithmetic follows libstdc++'s std::complex<double>: powers by
 * square-and-multiply (std::pow(complex, int)), products as (ac - bd,
 * ad + bc), quotients by Smith's scaled division (libgcc's __divdc3 in the
 * normal range).
 *
 * The reference scans every box of every level for each list; this file
 * visits only the boxes whose integer position is within one box of the
 * target box (at the coarser of the two levels).  That is the same set
 * whenever every box side is far above the rounding of the bounds, which
 * orc_t1_build checks (status 7 otherwise).
 */
#include "fmm2d_type1_oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define T1_MAX_LEVEL 16
#define T1_TOL ORC_T1_TOL2

/* ------------------------------------------------------------------------ */
/* complex arithmetic of libstdc++                                           */
/* ------------------------------------------------------------------------ */
typedef struct { double re, im; } t1c;

static inline t1c c_(double re, double im) { t1c r = {re, im}; return r; }
static inline t1c c_add(t1c a, t1c b) { return c_(a.re + b.re, a.im + b.im); }
static inline t1c c_sub(t1c a, t1c b) { return c_(a.re - b.re, a.im - b.im); }
static inline t1c c_mul(t1c a, t1c b) {
    return c_(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline t1c c_rmul(double s, t1c a) { return c_(a.re * s, a.im * s); }
/* Smith's division, as libgcc's __divdc3 for operands in the normal range */
static inline t1c c_div(t1c a, t1c b) {
    double ratio, denom;
    if (fabs(b.re) < fabs(b.im)) {
        ratio = b.re / b.im;
        denom = (b.re * ratio) + b.im;
        return c_(((a.re * ratio) + a.im) / denom, ((a.im * ratio) - a.re) / denom);
    }
    ratio = b.im / b.re;
    denom = (b.im * ratio) + b.re;
    return c_(((a.im * ratio) + a.re) / denom, (a.im - (a.re * ratio)) / denom);
}
/* std::pow(complex<double>, int): __complex_pow_unsigned, and 1 / z^-n */
static t1c c_powu(t1c x, unsigned n) {
    t1c y = (n % 2) ? x : c_(1.0, 0.0);
    while (n >>= 1) {
        x = c_mul(x, x);
        if (n % 2) y = c_mul(y, x);
    }
    return y;
}
static inline t1c c_powi(t1c z, int n) {
    return n < 0 ? c_div(c_(1.0, 0.0), c_powu(z, (unsigned)(-n))) : c_powu(z, (unsigned)n);
}
/* std::min as the reference calls it: (b < a) ? b : a */
static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }

/* ------------------------------------------------------------------------ */
/* options, bico, L2L table                                                  */
/* ------------------------------------------------------------------------ */
void orc_t1_default_opts(orc_t1_opts *o) {
    o->n_s = 50;      /* global.cpp:326 */
    o->lmax = 7;      /* global.cpp:316 */
    o->ndp = 10;      /* global.cpp:327 */
    o->icutoff = 2;   /* global.cpp:328 */
    o->l2l_mode = 1;  /* the reference as it runs */
}

static int t1_opts_ok(const orc_t1_opts *o) {
    return o && o->n_s >= 1 && o->lmax >= 2 && o->lmax <= T1_MAX_LEVEL && o->ndp >= 1 &&
           o->ndp <= 40 && o->icutoff >= 0 && o->icutoff <= 3 &&
           (o->l2l_mode == 0 || o->l2l_mode == 1);
}

/* subroutine_all_fmm_2d.cpp:560-591 (the out-of-range branch prints an error
 * there and yields 0) */
double orc_t1_bico(int n, int *k) {
    if ((*k < 0) || (*k > n) || (n < 0)) return 0.0;
    if ((*k == 0) || (*k == n) || (n == 0)) return 1.0;
    if ((*k == 1) || (*k == n - 1)) return (double)n;
    if (*k > n / 2) *k = n - *k;
    const int n_k = n - *k;
    double c = 1.0;
    for (int i = 1; i <= *k; i++) c = c * ((double)(n_k + i) / (double)(i));
    return c;
}

/* The coefficient and exponent of each term of the L2L shift
 * (biotsavart_fmm_2d.cpp:299-307): l = kl - 1 is passed to bico by reference
 * and stays rewritten for the rest of the kp sweep in mode 1. */
void orc_t1_l2l_table(int ndp, int l2l_mode, double *coef, int *expo) {
    const int m = ndp + 1;
    for (int i = 0; i < m * m; ++i) {
        coef[i] = 0.0;
        expo[i] = 0;
    }
    for (int l = 0; l <= ndp; ++l) {
        int ll = l;
        for (int kp = l; kp <= ndp; ++kp) {
            int kk = ll;
            const double cbi = orc_t1_bico(kp, l2l_mode ? &ll : &kk);
            coef[l * m + kp] = cbi;
            expo[l * m + kp] = kp - ll;
        }
    }
}

/* regularization_function_2d.cpp:11-37 (q of the pair, 1/(2 pi) included) */
static double t1_regul(int icutoff, double rij, double sij) {
    if (icutoff == 0) return 1.0e0 / (2.0e0 * M_PI);
    const double rij2 = rij * rij, sij2 = sij * sij;
    const double rho2 = rij2 / sij2;
    if (icutoff == 1) return ((rho2 * (rho2 + 2.0e0)) / ((rho2 + 1.0e0) * (rho2 + 1.0e0))) /
                             (2.0e0 * M_PI);
    const double qt = -rho2 / (2.0e0);
    if (icutoff == 2) return (1.0e0 - exp(qt)) / (2.0e0 * M_PI);
    const double qte = (1 - qt) * exp(qt);
    return (1.0e0 - qte) / (2.0e0 * M_PI);
}

/* One pair of direct_sum (subroutine_all_fmm_2d.cpp:613-626). */
static inline void t1_pair(double xi, double yi, double si, double xj, double yj, double sj,
                           double gj, int icutoff, double *u, double *v) {
    const double dxij = xi - xj, dyij = yi - yj;
    const double rij2 = dxij * dxij + dyij * dyij;
    if (!(rij2 > 0.0e0)) return;
    const double sij2 = 0.25e0 * (si * si + sj * sj);
    const double q = t1_regul(icutoff, sqrt(rij2), sqrt(sij2));
    *u = *u - q * dyij * gj / rij2;
    *v = *v + q * dxij * gj / rij2;
}

/* ------------------------------------------------------------------------ */
/* box hierarchy                                                             */
/* ------------------------------------------------------------------------ */
struct orc_t1_tree {
    orc_t1_opts o;
    int n, lev, nbox, nlost;
    double geo[5];                      /* xmin, xmax, ymin, ymax, db */
    int lev_off[T1_MAX_LEVEL + 2];      /* boxes of level k: [lev_off[k], lev_off[k+1]) */
    int *level, *ix, *iy, *parent, *nchild, *child, *pbeg, *pcnt;
    double *bnd;                        /* xmin, xmax, ymin, ymax per box */
    int *pool;
    size_t pool_len, pool_cap;
    int cap;
    uint64_t *hkey;                     /* (level, ix, iy) -> box, open addressing */
    int *hval;
    size_t hmask;
    const double *x, *y;
};

static uint64_t t1_hkey(int level, int ix, int iy) {
    return ((uint64_t)(level + 1) << 40) | ((uint64_t)ix << 20) | (uint64_t)iy;
}
static size_t t1_hash(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33;
    return (size_t)k;
}
static int t1_find(const orc_t1_tree *t, int level, int ix, int iy) {
    if (ix < 0 || iy < 0 || ix >= (1 << level) || iy >= (1 << level)) return -1;
    const uint64_t k = t1_hkey(level, ix, iy);
    for (size_t h = t1_hash(k) & t->hmask;; h = (h + 1) & t->hmask) {
        if (t->hkey[h] == k) return t->hval[h];
        if (t->hkey[h] == 0) return -1;
    }
}

static int t1_grow(orc_t1_tree *t) {
    const int nc = t->cap ? 2 * t->cap : 1024;
#define T1_RE(f, ty, m)                                                   \
    do {                                                                  \
        ty *p_ = (ty *)realloc(t->f, sizeof(ty) * (size_t)nc * (m));      \
        if (!p_) return 0;                                                \
        t->f = p_;                                                        \
    } while (0)
    T1_RE(level, int, 1);
    T1_RE(ix, int, 1);
    T1_RE(iy, int, 1);
    T1_RE(parent, int, 1);
    T1_RE(nchild, int, 1);
    T1_RE(child, int, 4);
    T1_RE(pbeg, int, 1);
    T1_RE(pcnt, int, 1);
    T1_RE(bnd, double, 4);
#undef T1_RE
    t->cap = nc;
    return 1;
}

static int t1_pool_reserve(orc_t1_tree *t, size_t more) {
    if (t->pool_len + more <= t->pool_cap) return 1;
    size_t nc = t->pool_cap ? t->pool_cap : 4096;
    while (nc < t->pool_len + more) nc *= 2;
    int *p = (int *)realloc(t->pool, sizeof(int) * nc);
    if (!p) return 0;
    t->pool = p;
    t->pool_cap = nc;
    return 1;
}

/* Box with bounds [xl, xh) x [yl, yh) over the given particles (in order);
 * created when it holds any (amount_inbox_new, subroutine_all_fmm_2d.cpp:
 * 635-678, and hierarchy_mesh's "npr_in > 0", :189, :248). */
static int t1_make_box(orc_t1_tree *t, int level, int ix, int iy, int par, double xl, double xh,
                       double yl, double yh, const int *src, int nsrc, int *made) {
    *made = -1;
    if (!t1_pool_reserve(t, (size_t)nsrc)) return 0;
    const size_t start = t->pool_len;
    int cnt = 0;
    for (int u = 0; u < nsrc; ++u) {
        const int i = src ? src[u] : u;
        const double xi = t->x[i], yi = t->y[i];
        if ((xi >= xl) && (xi < xh) && (yi >= yl) && (yi < yh)) t->pool[start + cnt++] = i;
    }
    if (cnt == 0) return 1;
    if (t->nbox == t->cap && !t1_grow(t)) return 0;
    const int b = t->nbox++;
    t->level[b] = level;
    t->ix[b] = ix;
    t->iy[b] = iy;
    t->parent[b] = par;
    t->nchild[b] = 0;
    t->child[4 * b + 0] = t->child[4 * b + 1] = t->child[4 * b + 2] = t->child[4 * b + 3] = -1;
    t->pbeg[b] = (int)start;
    t->pcnt[b] = cnt;
    t->bnd[4 * b + 0] = xl;
    t->bnd[4 * b + 1] = xh;
    t->bnd[4 * b + 2] = yl;
    t->bnd[4 * b + 3] = yh;
    t->pool_len += (size_t)cnt;
    *made = b;
    return 1;
}

void orc_t1_free(orc_t1_tree *t) {
    if (!t) return;
    free(t->level);
    free(t->ix);
    free(t->iy);
    free(t->parent);
    free(t->nchild);
    free(t->child);
    free(t->pbeg);
    free(t->pcnt);
    free(t->bnd);
    free(t->pool);
    free(t->hkey);
    free(t->hval);
    free(t);
}

/* hierarchy_mesh, subroutine_all_fmm_2d.cpp:97-287. */
int orc_t1_build(orc_t1_tree **out, int n, const double *x, const double *y,
                 const orc_t1_opts *o) {
    *out = NULL;
    if (!t1_opts_ok(o) || n < 0 || (n > 0 && (!x || !y))) return 1;
    for (int i = 0; i < n; ++i)
        if (!isfinite(x[i]) || !isfinite(y[i])) return 2;
    orc_t1_tree *t = (orc_t1_tree *)calloc(1, sizeof(orc_t1_tree));
    if (!t) return 5;
    t->o = *o;
    t->n = n;
    t->x = x;
    t->y = y;
    const int lmax = o->lmax;

    /* root: bounding box (:120-131), padding by dr 1e-5 (:143-148), square on
     * the longer side (:150-170) */
    double xmin = 1.0e0 / T1_TOL, xmax = -1.0e0 / T1_TOL;
    double ymin = 1.0e0 / T1_TOL, ymax = -1.0e0 / T1_TOL;
    for (int i = 0; i < n; i++) {
        xmin = smin(xmin, x[i]);
        xmax = smax(xmax, x[i]);
        ymin = smin(ymin, y[i]);
        ymax = smax(ymax, y[i]);
    }
    const double tlout = 1.0e-5;
    const double dr = smin(fabs(xmax - xmin), fabs(ymax - ymin)) / ((double)(1LL << lmax));
    xmin = xmin - dr * tlout;
    xmax = xmax + dr * tlout;
    ymin = ymin - dr * tlout;
    ymax = ymax + dr * tlout;
    const double rdx = fabs(xmax - xmin), rdy = fabs(ymax - ymin);
    double db = 0.0;
    if (rdy >= rdx) {
        const double cent = (xmin + xmax) / 2.0e0;
        xmin = cent - rdy / 2.0e0;
        xmax = cent + rdy / 2.0e0;
        db = rdy / 2.0e0;
    } else {
        const double cent = (ymin + ymax) / 2.0e0;
        ymin = cent - rdx / 2.0e0;
        ymax = cent + rdx / 2.0e0;
        db = rdx / 2.0e0;
    }
    t->geo[0] = xmin;
    t->geo[1] = xmax;
    t->geo[2] = ymin;
    t->geo[3] = ymax;
    t->geo[4] = db;
    if (n > 0 && db > 0.0) {
        /* the box tests must not see a box side near the bound rounding */
        const double m = fmax(fmax(fabs(xmin), fabs(xmax)), fmax(fabs(ymin), fabs(ymax)));
        const double finest = db / (double)(1LL << (lmax - 1));
        if (!(finest > 64.0 * (T1_TOL + lmax * m * DBL_EPSILON))) {
            orc_t1_free(t);
            return 7;
        }
    }

    /* level 1: the four quadrants, x outer, y inner (:175-206) */
    for (int k = 0; k <= T1_MAX_LEVEL + 1; ++k) t->lev_off[k] = 0;
    t->lev_off[1] = 0;
    int made;
    if (n > 0) {
        for (int ib1 = 1; ib1 <= 2; ib1++)
            for (int ib2 = 1; ib2 <= 2; ib2++) {
                const double xl = xmin + (ib1 - 1) * db, xh = xmin + ib1 * db;
                const double yl = ymin + (ib2 - 1) * db, yh = ymin + ib2 * db;
                if (!t1_make_box(t, 1, ib1 - 1, ib2 - 1, -1, xl, xh, yl, yh, NULL, n, &made)) {
                    orc_t1_free(t);
                    return 5;
                }
            }
    }
    /* particles outside every quadrant belong to no box */
    {
        long long inside = 0;
        for (int b = 0; b < t->nbox; ++b) inside += t->pcnt[b];
        t->nlost = n - (int)inside;
    }
    t->lev_off[2] = t->nbox;
    /* split while a box holds more than n_s, to lmax (:216-279) */
    int lev = 1;
    for (int k = 1; k <= lmax - 1; k++) {
        int n_stop = 0;
        const int b0 = t->lev_off[k], b1 = t->lev_off[k + 1];
        for (int b = b0; b < b1; ++b) {
            t->nchild[b] = 0;
            if (t->pcnt[b] > o->n_s) {
                n_stop = n_stop + 1;
                const double xl0 = t->bnd[4 * b + 0], xh0 = t->bnd[4 * b + 1];
                const double yl0 = t->bnd[4 * b + 2], yh0 = t->bnd[4 * b + 3];
                const double dbx = (xh0 - xl0) / 2.0e0, dby = (yh0 - yl0) / 2.0e0;
                int kept = 0;
                for (int ib1 = 1; ib1 <= 2; ib1++)
                    for (int ib2 = 1; ib2 <= 2; ib2++) {
                        const double xl = xl0 + (ib1 - 1) * dbx, xh = xl0 + ib1 * dbx;
                        const double yl = yl0 + (ib2 - 1) * dby, yh = yl0 + ib2 * dby;
                        /* the parent's particles are read from the pool, which
                         * t1_make_box may move: copy the range's base first */
                        const int pb = t->pbeg[b], pc = t->pcnt[b];
                        if (!t1_pool_reserve(t, (size_t)pc)) {
                            orc_t1_free(t);
                            return 5;
                        }
                        if (!t1_make_box(t, k + 1, 2 * t->ix[b] + (ib1 - 1),
                                         2 * t->iy[b] + (ib2 - 1), b, xl, xh, yl, yh,
                                         t->pool + pb, pc, &made)) {
                            orc_t1_free(t);
                            return 5;
                        }
                        if (made >= 0) {
                            t->child[4 * b + t->nchild[b]++] = made;
                            kept += t->pcnt[made];
                        }
                    }
                t->nlost += t->pcnt[b] - kept;
            }
        }
        t->lev_off[k + 2] = t->nbox;
        lev = k + 1;
        if (n_stop == 0) {
            lev = k;
            break;
        }
    }
    for (int k = lev + 1; k <= T1_MAX_LEVEL + 1; ++k) t->lev_off[k] = t->nbox;
    t->lev = lev;
    /* position hash */
    size_t hc = 16;
    while (hc < 2 * (size_t)t->nbox + 16) hc *= 2;
    t->hkey = (uint64_t *)calloc(hc, sizeof(uint64_t));
    t->hval = (int *)malloc(hc * sizeof(int));
    if (!t->hkey || !t->hval) {
        orc_t1_free(t);
        return 5;
    }
    t->hmask = hc - 1;
    for (int b = 0; b < t->nbox; ++b) {
        const uint64_t k = t1_hkey(t->level[b], t->ix[b], t->iy[b]);
        size_t h = t1_hash(k) & t->hmask;
        while (t->hkey[h] != 0) h = (h + 1) & t->hmask;
        t->hkey[h] = k;
        t->hval[h] = b;
    }
    *out = t;
    return 0;
}

int orc_t1_nboxes(const orc_t1_tree *t) { return t->nbox; }
int orc_t1_depth(const orc_t1_tree *t) { return t->lev; }
int orc_t1_nlost(const orc_t1_tree *t) { return t->nlost; }
void orc_t1_root(const orc_t1_tree *t, double *geo5) {
    for (int i = 0; i < 5; ++i) geo5[i] = t->geo[i];
}
void orc_t1_boxes(const orc_t1_tree *t, int *level, int *ix, int *iy, int *count, int *nchild,
                  int *parent, double *bounds4) {
    for (int b = 0; b < t->nbox; ++b) {
        if (level) level[b] = t->level[b];
        if (ix) ix[b] = t->ix[b];
        if (iy) iy[b] = t->iy[b];
        if (count) count[b] = t->pcnt[b];
        if (nchild) nchild[b] = t->nchild[b];
        if (parent) parent[b] = t->parent[b];
        if (bounds4)
            for (int d = 0; d < 4; ++d) bounds4[4 * b + d] = t->bnd[4 * b + d];
    }
}
int orc_t1_box_particles(const orc_t1_tree *t, int b, int *out, int cap) {
    if (b < 0 || b >= t->nbox) return -1;
    const int c = t->pcnt[b];
    for (int u = 0; u < c && u < cap; ++u) out[u] = t->pool[t->pbeg[b] + u];
    return c;
}

/* ------------------------------------------------------------------------ */
/* lists (subroutine_all_fmm_2d.cpp:291-556)                                 */
/* ------------------------------------------------------------------------ */
/* a touches b: an edge of one within tol2 of an edge of the other, with the
 * ranges overlapping in the other direction (the adjacency test of
 * list_one / list_two / list_three / list_four) */
static inline int t1_touch(const double *a, const double *b) {
    const double dx_xn = fabs(a[1] - b[0]), dx_nx = fabs(a[0] - b[1]);
    const double dy_xn = fabs(a[3] - b[2]), dy_nx = fabs(a[2] - b[3]);
    return (((dy_xn <= T1_TOL) || (dy_nx <= T1_TOL)) && (a[0] <= b[1] + T1_TOL) &&
            (a[1] >= b[0] - T1_TOL)) ||
           (((dx_xn <= T1_TOL) || (dx_nx <= T1_TOL)) && (a[2] <= b[3] + T1_TOL) &&
            (a[3] >= b[2] - T1_TOL));
}
/* a is well separated from b: a gap above tol2 in some direction */
static inline int t1_sep(const double *a, const double *b) {
    return (a[0] > b[1] + T1_TOL) || (a[1] < b[0] - T1_TOL) || (a[2] > b[3] + T1_TOL) ||
           (a[3] < b[2] - T1_TOL);
}

static int cmp_int(const void *a, const void *b) {
    const int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* Boxes of level k2 whose position is within one box of box tb (at the
 * coarser of the two levels), in the reference's order inside the level.
 * buf must hold 4 * 2^(k2 - level) + 9 entries. */
static int t1_near(const orc_t1_tree *t, int tb, int k2, int *buf) {
    const int lt = t->level[tb], ix = t->ix[tb], iy = t->iy[tb];
    int m = 0;
    if (k2 <= lt) {
        const int X = ix >> (lt - k2), Y = iy >> (lt - k2);
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int c = t1_find(t, k2, X + dx, Y + dy);
                if (c >= 0) buf[m++] = c;
            }
    } else {
        /* the ring of level-k2 positions around tb's footprint (boxes inside
         * the footprint are tb's descendants: neither touching nor separated) */
        const int d = k2 - lt;
        const int X0 = ix << d, X1 = (ix + 1) << d, Y0 = iy << d, Y1 = (iy + 1) << d;
        for (int X = X0 - 1; X <= X1; ++X) {
            const int edge_x = X == X0 - 1 || X == X1;
            if (edge_x) {
                for (int Y = Y0 - 1; Y <= Y1; ++Y) {
                    const int c = t1_find(t, k2, X, Y);
                    if (c >= 0) buf[m++] = c;
                }
            } else {
                int c = t1_find(t, k2, X, Y0 - 1);
                if (c >= 0) buf[m++] = c;
                c = t1_find(t, k2, X, Y1);
                if (c >= 0) buf[m++] = c;
            }
        }
    }
    qsort(buf, (size_t)m, sizeof(int), cmp_int);
    return m;
}

static size_t t1_near_cap(const orc_t1_tree *t) {
    return 4 * ((size_t)1 << (t->lev > 1 ? t->lev - 1 : 0)) + 16;
}

#define T1_PUSH(v)                     \
    do {                               \
        if (len < cap) out[len] = (v); \
        ++len;                         \
    } while (0)

/* list 1 of a leaf: itself, then every childless box of any level touching
 * it (list_one, :291-344) */
static int t1_list1(const orc_t1_tree *t, int b, int *out, int cap, int *buf) {
    int len = 0;
    if (t->nchild[b] > 0) return 0;
    T1_PUSH(b);
    const double *bb = t->bnd + 4 * b;
    for (int k2 = 1; k2 <= t->lev; ++k2) {
        const int m = t1_near(t, b, k2, buf);
        for (int u = 0; u < m; ++u) {
            const int c = buf[u];
            if (c != b && t->nchild[c] == 0 && t1_touch(t->bnd + 4 * c, bb)) T1_PUSH(c);
        }
    }
    return len;
}

/* list 2: children of the parent's touching colleagues (with children),
 * well separated from the box; at level 1 the other level-1 boxes well
 * separated from it (list_two, :348-426) */
static int t1_list2(const orc_t1_tree *t, int b, int *out, int cap, int *buf) {
    int len = 0;
    const int k = t->level[b];
    const double *bb = t->bnd + 4 * b;
    if (k == 1) {
        for (int c = t->lev_off[1]; c < t->lev_off[2]; ++c)
            if (c != b && t1_sep(t->bnd + 4 * c, bb)) T1_PUSH(c);
        return len;
    }
    const int p = t->parent[b];
    const double *pb = t->bnd + 4 * p;
    const int m = t1_near(t, p, k - 1, buf);
    for (int u = 0; u < m; ++u) {
        const int c = buf[u];
        if (t->nchild[c] > 0 && c != p && t1_touch(t->bnd + 4 * c, pb))
            for (int q = 0; q < t->nchild[c]; ++q) {
                const int e = t->child[4 * c + q];
                if (e != b && t1_sep(t->bnd + 4 * e, bb)) T1_PUSH(e);
            }
    }
    return len;
}

/* list 3 of a leaf: children, well separated from it, of the boxes of its
 * level and finer that touch it and have children (list_three, :430-495) */
static int t1_list3(const orc_t1_tree *t, int b, int *out, int cap, int *buf) {
    int len = 0;
    if (t->nchild[b] > 0) return 0;
    const double *bb = t->bnd + 4 * b;
    for (int k2 = t->level[b]; k2 <= t->lev; ++k2) {
        const int m = t1_near(t, b, k2, buf);
        for (int u = 0; u < m; ++u) {
            const int c = buf[u];
            if (c != b && t->nchild[c] > 0 && t1_touch(t->bnd + 4 * c, bb))
                for (int q = 0; q < t->nchild[c]; ++q) {
                    const int e = t->child[4 * c + q];
                    if (t1_sep(t->bnd + 4 * e, bb)) T1_PUSH(e);
                }
        }
    }
    return len;
}

/* list 4: childless boxes of coarser levels touching the parent and well
 * separated from the box (list_four, :499-556) */
static int t1_list4(const orc_t1_tree *t, int b, int *out, int cap, int *buf) {
    int len = 0;
    const int k = t->level[b];
    if (k == 1) return 0;
    const int p = t->parent[b];
    const double *bb = t->bnd + 4 * b, *pb = t->bnd + 4 * p;
    for (int k2 = 1; k2 <= k - 1; ++k2) {
        const int m = t1_near(t, p, k2, buf);
        for (int u = 0; u < m; ++u) {
            const int c = buf[u];
            if (t->nchild[c] == 0 && t1_touch(t->bnd + 4 * c, pb) && t1_sep(t->bnd + 4 * c, bb))
                T1_PUSH(c);
        }
    }
    return len;
}
#undef T1_PUSH

int orc_t1_list(const orc_t1_tree *t, int b, int which, int *out, int cap) {
    if (b < 0 || b >= t->nbox || which < 1 || which > 4) return -1;
    int *buf = (int *)malloc(sizeof(int) * t1_near_cap(t));
    if (!buf) return -1;
    int len = 0;
    if (which == 1) len = t1_list1(t, b, out, cap, buf);
    else if (which == 2) len = t1_list2(t, b, out, cap, buf);
    else if (which == 3) len = t1_list3(t, b, out, cap, buf);
    else len = t1_list4(t, b, out, cap, buf);
    free(buf);
    return len;
}

/* ------------------------------------------------------------------------ */
/* velocity (biotsavart_fmm_2d.cpp:93-446)                                   */
/* ------------------------------------------------------------------------ */
static inline t1c t1_cen(const orc_t1_tree *t, int b) {
    const double *bb = t->bnd + 4 * b;
    return c_((bb[1] + bb[0]) / 2.0e0, (bb[3] + bb[2]) / 2.0e0);
}

int orc_t1_velocity(int n, const double *x, const double *y, const double *s, const double *g,
                    const orc_t1_opts *o, double *u, double *v) {
    if (n > 0 && (!g || !u || !v || (o && o->icutoff > 0 && !s))) return 1;
    orc_t1_tree *t = NULL;
    const int st = orc_t1_build(&t, n, x, y, o);
    if (st) return st;
    for (int i = 0; i < n; ++i) u[i] = v[i] = 0.0;
    const int ndp = o->ndp, nb = t->nbox, lev = t->lev;
    const int ma = ndp, ml = ndp + 1;  /* terms of a (orders 0..ndp-1) and b (0..ndp) */
    t1c *ak = (t1c *)calloc((size_t)nb * ma + 1, sizeof(t1c));
    t1c *bk = (t1c *)calloc((size_t)nb * ml + 1, sizeof(t1c));
    double *l2c = (double *)malloc(sizeof(double) * ml * ml);
    int *l2e = (int *)malloc(sizeof(int) * ml * ml);
    if (!ak || !bk || !l2c || !l2e) {
        free(ak);
        free(bk);
        free(l2c);
        free(l2e);
        orc_t1_free(t);
        return 5;
    }
    orc_t1_l2l_table(ndp, o->l2l_mode, l2c, l2e);
    const size_t ncap = t1_near_cap(t);
    int fail = 0;

    /* P2M at childless boxes, order outer, particles inner (:123-153):
     * ak[kp] += g (zj - zm)^(kp - 1) */
#pragma omp parallel for schedule(dynamic, 16)
    for (int b = 0; b < nb; ++b) {
        if (t->nchild[b] != 0) continue;
        const t1c zm = t1_cen(t, b);
        const int *pp = t->pool + t->pbeg[b];
        for (int kp = 1; kp <= ndp; kp++)
            for (int q = 0; q < t->pcnt[b]; ++q) {
                const int jp = pp[q];
                const t1c zj = c_(x[jp], y[jp]);
                ak[(size_t)b * ma + kp - 1] =
                    c_add(ak[(size_t)b * ma + kp - 1], c_rmul(g[jp], c_powi(c_sub(zj, zm), kp - 1)));
            }
    }
    /* M2M, levels lev-1 .. 1 (:162-188):
     * ak_p[kl] += bico(kl-1, kp-1) (zc - zp)^(kl-kp) ak_c[kp] */
    for (int k = lev - 1; k >= 1; k--) {
#pragma omp parallel for schedule(dynamic, 16)
        for (int b = t->lev_off[k]; b < t->lev_off[k + 1]; ++b) {
            if (t->nchild[b] <= 0) continue;
            const t1c zp = t1_cen(t, b);
            for (int q = 0; q < t->nchild[b]; ++q) {
                const int c = t->child[4 * b + q];
                const t1c zc = t1_cen(t, c);
                for (int kl = 1; kl <= ndp; kl++)
                    for (int kp = 1; kp <= kl; kp++) {
                        int kpm1 = kp - 1;
                        const double cbi = orc_t1_bico(kl - 1, &kpm1);
                        t1c *dst = &ak[(size_t)b * ma + kl - 1];
                        *dst = c_add(*dst, c_mul(c_rmul(cbi, c_powi(c_sub(zc, zp), kl - kp)),
                                                 ak[(size_t)c * ma + kp - 1]));
                    }
            }
        }
    }
    /* M2L from list 2 and P2L from list 4, every box (:199-271):
     * bk[kl] += (-1)^kp bico(ll+kp-1, kp-1) ak[kp] / (zo - zg)^(kp+ll)
     * bk[kl] -= g / (zj - zg)^(1+ll),  ll = kl - 1 */
#pragma omp parallel
    {
        int *buf = (int *)malloc(sizeof(int) * ncap);
        int cap = 4096;
        int *lst = (int *)malloc(sizeof(int) * (size_t)cap);
        if (!buf || !lst) {
#pragma omp atomic write
            fail = 1;
        }
#pragma omp for schedule(dynamic, 16)
        for (int b = 0; b < nb; ++b) {
            if (!buf || !lst) continue;
            const t1c zg = t1_cen(t, b);
            t1c *bb = bk + (size_t)b * ml;
            for (int w = 2; w <= 4; w += 2) {
                int len = w == 2 ? t1_list2(t, b, lst, cap, buf) : t1_list4(t, b, lst, cap, buf);
                if (len > cap) {
                    int *nl = (int *)realloc(lst, sizeof(int) * (size_t)len);
                    if (!nl) {
#pragma omp atomic write
                        fail = 1;
                        continue;
                    }
                    lst = nl;
                    cap = len;
                    len = w == 2 ? t1_list2(t, b, lst, cap, buf) : t1_list4(t, b, lst, cap, buf);
                }
                for (int e = 0; e < len; ++e) {
                    const int c = lst[e];
                    if (w == 2) {
                        const t1c zo = t1_cen(t, c);
                        for (int kl = 1; kl <= ndp + 1; kl++) {
                            const int ll = kl - 1;
                            for (int kp = 1; kp <= ndp; kp++) {
                                int kpm1 = kp - 1;
                                const double cbi = orc_t1_bico(ll + kp - 1, &kpm1);
                                const double sg = (kp % 2) ? -1.0e0 : 1.0e0;
                                bb[kl - 1] = c_add(
                                    bb[kl - 1],
                                    c_div(c_rmul(sg * cbi, ak[(size_t)c * ma + kp - 1]),
                                          c_powi(c_sub(zo, zg), kp + ll)));
                            }
                        }
                    } else {
                        const int *pp = t->pool + t->pbeg[c];
                        for (int q = 0; q < t->pcnt[c]; ++q) {
                            const int jp = pp[q];
                            const t1c zo = c_(x[jp], y[jp]);
                            for (int kl = 1; kl <= ndp + 1; kl++) {
                                const int ll = kl - 1;
                                bb[kl - 1] = c_sub(bb[kl - 1],
                                                   c_div(c_(g[jp], 0.0), c_powi(c_sub(zo, zg), 1 + ll)));
                            }
                        }
                    }
                }
            }
        }
        free(buf);
        free(lst);
    }
    /* L2L, levels 1 .. lev-1 (:283-312), with the table of the shift as run */
    for (int k = 1; k <= lev - 1; k++) {
#pragma omp parallel for schedule(dynamic, 16)
        for (int b = t->lev_off[k]; b < t->lev_off[k + 1]; ++b) {
            if (t->nchild[b] <= 0) continue;
            const t1c zp = t1_cen(t, b);
            for (int q = 0; q < t->nchild[b]; ++q) {
                const int c = t->child[4 * b + q];
                const t1c zc = t1_cen(t, c);
                for (int l = 0; l <= ndp; ++l)
                    for (int kp = l; kp <= ndp; ++kp) {
                        t1c *dst = &bk[(size_t)c * ml + l];
                        *dst = c_add(*dst, c_mul(c_rmul(l2c[l * ml + kp],
                                                        c_powi(c_sub(zc, zp), l2e[l * ml + kp])),
                                                 bk[(size_t)b * ml + kp]));
                    }
            }
        }
    }
    /* evaluation at every childless box (:321-427): direct sums over list 1,
     * M2P over list 3, then L2P */
    const double twopi = 2.0e0 * M_PI;
#pragma omp parallel
    {
        int *buf = (int *)malloc(sizeof(int) * ncap);
        int cap = 4096;
        int *lst = (int *)malloc(sizeof(int) * (size_t)cap);
        if (!buf || !lst) {
#pragma omp atomic write
            fail = 1;
        }
#pragma omp for schedule(dynamic, 8)
        for (int b = 0; b < nb; ++b) {
            if (!buf || !lst || t->nchild[b] != 0) continue;
            const int nip = t->pcnt[b];
            const int *ipp = t->pool + t->pbeg[b];
            const t1c zg = t1_cen(t, b);
            for (int w = 1; w <= 3; w += 2) {
                int len = w == 1 ? t1_list1(t, b, lst, cap, buf) : t1_list3(t, b, lst, cap, buf);
                if (len > cap) {
                    int *nl = (int *)realloc(lst, sizeof(int) * (size_t)len);
                    if (!nl) {
#pragma omp atomic write
                        fail = 1;
                        continue;
                    }
                    lst = nl;
                    cap = len;
                    len = w == 1 ? t1_list1(t, b, lst, cap, buf) : t1_list3(t, b, lst, cap, buf);
                }
                if (w == 1) {
                    for (int e = 0; e < len; ++e) {
                        const int c = lst[e];
                        const int *jpp = t->pool + t->pbeg[c];
                        for (int i2 = 0; i2 < nip; ++i2) {
                            const int i = ipp[i2];
                            const double si = s ? s[i] : 0.0;
                            for (int j2 = 0; j2 < t->pcnt[c]; ++j2) {
                                const int j = jpp[j2];
                                t1_pair(x[i], y[i], si, x[j], y[j], s ? s[j] : 0.0, g[j],
                                        o->icutoff, &u[i], &v[i]);
                            }
                        }
                    }
                } else if (len > 0) {
                    for (int i2 = 0; i2 < nip; ++i2) {
                        const int ip = ipp[i2];
                        t1c cmw = c_(0.0, 0.0);
                        const t1c z = c_(x[ip], y[ip]);
                        for (int e = 0; e < len; ++e) {
                            const int c = lst[e];
                            const t1c zm = t1_cen(t, c);
                            for (int kp = 1; kp <= ndp; kp++)
                                cmw = c_add(cmw, c_div(ak[(size_t)c * ma + kp - 1],
                                                       c_powi(c_sub(z, zm), kp)));
                        }
                        const t1c cmw2 = c_mul(c_(-0.0, -1.0), cmw);
                        u[ip] = u[ip] + cmw2.re / twopi;
                        v[ip] = v[ip] - cmw2.im / twopi;
                    }
                }
            }
            for (int i2 = 0; i2 < nip; ++i2) {
                const int ip = ipp[i2];
                const t1c zi = c_(x[ip], y[ip]);
                t1c cmw = c_(0.0, 0.0);
                for (int kl = 1; kl <= ndp + 1; kl++)
                    cmw = c_add(cmw, c_mul(bk[(size_t)b * ml + kl - 1], c_powi(c_sub(zi, zg), kl - 1)));
                const t1c cmw2 = c_mul(c_(-0.0, -1.0), cmw);
                u[ip] = u[ip] + cmw2.re / twopi;
                v[ip] = v[ip] - cmw2.im / twopi;
            }
        }
        free(buf);
        free(lst);
    }
    free(ak);
    free(bk);
    free(l2c);
    free(l2e);
    orc_t1_free(t);
    return fail ? 5 : 0;
}

void orc_t1_direct(int n, const double *x, const double *y, const double *s, const double *g,
                   int icutoff, int nt, const int *tidx, double *u, double *v) {
    const int m = tidx ? nt : n;
#pragma omp parallel for schedule(static)
    for (int a = 0; a < m; ++a) {
        const int i = tidx ? tidx[a] : a;
        double ui = 0.0, vi = 0.0;
        const double si = s ? s[i] : 0.0;
        for (int j = 0; j < n; ++j)
            t1_pair(x[i], y[i], si, x[j], y[j], s ? s[j] : 0.0, g[j], icutoff, &ui, &vi);
        u[a] = ui;
        v[a] = vi;
    }
}
