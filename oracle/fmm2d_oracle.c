/*
 * fmm2d_oracle.c -- CPU oracle for the 2-D FMM Biot-Savart velocity path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/, smoke() and the
 * cpu_baseline leg of bench.py, never by the product library.
 *
 * A restatement of the reference semantics (anggarapangestu/LSMPS_VPM_v2,
 * src/FMM/treeCell.cpp and src/FMM/fmm2D.cpp) with flat arrays and an
 * open-addressing hash from reference cell ID to cell index.  File:line
 * citations are given at each function.  Complex powers are formed by
 * repeated multiplication; the reference uses std::pow, so results agree to
 * rounding (~1e-14 relative), not bit for bit.
 */
#include "fmm2d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* treeCell.cpp:5 -- "10e-10", i.e. 1e-9. */
#define ORC_EPS 10e-10

/* ------------------------------------------------------------------------ */
/* small complex helpers                                                     */
/* ------------------------------------------------------------------------ */
typedef struct { double re, im; } cx_t;

static inline cx_t cx(double re, double im) { cx_t r = {re, im}; return r; }
static inline cx_t cmul(cx_t a, cx_t b) {
    return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cx_t cadd(cx_t a, cx_t b) { return cx(a.re + b.re, a.im + b.im); }
static inline cx_t csub(cx_t a, cx_t b) { return cx(a.re - b.re, a.im - b.im); }
static inline cx_t cscale(cx_t a, double s) { return cx(a.re * s, a.im * s); }
static inline cx_t cinv(cx_t a) {
    double d = a.re * a.re + a.im * a.im;
    return cx(a.re / d, -a.im / d);
}

/* ------------------------------------------------------------------------ */
/* options                                                                   */
/* ------------------------------------------------------------------------ */
void orc_default_opts(orc_opts *o) {
    o->p = 10;           /* global.cpp:311 */
    o->max_par = 75;     /* global.cpp:312 */
    o->max_src = 50;     /* global.cpp:313 */
    o->exp_tree = 1.0;   /* global.cpp:303 */
    o->min_level = 1;    /* root never a leaf, treeCell.cpp:1129 */
    o->max_level = 0;
    o->binom_mode = 1;   /* reference helper */
    o->list4_mode = 1;   /* M2L_P2_METHOD_TYPE 1, fmm2D.cpp:7 */
    o->stage_mask = 0;
    o->icutoff = 0;      /* the type-2 FMM's P2P is singular, fmm2D.cpp:164-173 */
}

static int opts_ok(const orc_opts *o) {
    if (!o) return 0;
    if (o->p < 1 || o->p > 40) return 0;
    if (o->max_par < 1 || o->max_src < 0) return 0;
    if (!(o->exp_tree > 0.0)) return 0;
    if (o->min_level < 1 || o->min_level > ORC_MAX_DEPTH) return 0;
    if (o->max_level < 0 || o->max_level > ORC_MAX_DEPTH) return 0;
    if (o->max_level && o->max_level < o->min_level) return 0;
    if (o->binom_mode != 0 && o->binom_mode != 1) return 0;
    if (o->list4_mode != 1 && o->list4_mode != 2) return 0;
    if (o->stage_mask < 0 || o->stage_mask > 7) return 0;
    if (o->icutoff < 0 || o->icutoff > 3) return 0;
    return 1;
}

/* Regularization factor of a near-field pair: 2 pi q of regul_func_2d
 * (regularization_function_2d.cpp:11-37) for the pair distance rij and the
 * pair size sij (the reference's direct_sum passes sqrt of both,
 * subroutine_all_fmm_2d.cpp:614-623, and regul_func_2d squares them again). */
static double regul_factor(int icutoff, double rij, double sij) {
    const double rij2 = pow(rij, 2), sij2 = pow(sij, 2);
    const double rho2 = rij2 / sij2;
    if (icutoff == 1) return (rho2 * (rho2 + 2.0)) / pow(rho2 + 1.0, 2);
    const double qt = -rho2 / 2.0;
    if (icutoff == 2) return 1.0 - exp(qt);
    return 1.0 - (1 - qt) * exp(qt);
}

/* ------------------------------------------------------------------------ */
/* binomials                                                                 */
/* ------------------------------------------------------------------------ */
/* fmm2D.cpp:28-58: product n..max+1 then divide by min..2, in int.  The
 * product overflows int32 for the M2L arguments at P >= 10; g++ on x86-64
 * wraps (imul), which is emulated here with uint32 arithmetic. */
int orc_binom_ref(int n, int k) {
    if (k > n) return 0;
    int inv = n - k;
    int mx = inv > k ? inv : k;
    int mn = inv > k ? k : inv;
    uint32_t c = 1u;
    for (int i = n; i > mx; --i) c = c * (uint32_t)i;
    int32_t cs = (int32_t)c;
    for (int i = mn; i > 1; --i) cs = cs / i;
    return cs;
}

/* Exact in 128-bit integers (C(n-k+i, i) is an integer at every step),
 * rounded once to double; exact for every n <= 127. */
double orc_binom_exact(int n, int k) {
    if (k < 0 || k > n) return 0.0;
    if (k > n - k) k = n - k;
    unsigned __int128 c = 1;
    for (int i = 1; i <= k; ++i) c = c * (unsigned)(n - k + i) / (unsigned)i;
    return (double)c;
}

/* ------------------------------------------------------------------------ */
/* cell IDs (treeCell.cpp:114-148, 223-280)                                  */
/* ------------------------------------------------------------------------ */
static inline int64_t start_id(int level) {
    int64_t p = 1;
    for (int i = 0; i < level; ++i) p *= 4;
    return (p - 1) / 3;
}
static inline int64_t index2id(int level, int ix, int iy) {
    return start_id(level) + (int64_t)ix + (int64_t)iy * ((int64_t)1 << level);
}
static inline int id2level(int64_t id) {
    int l = 0;
    while (start_id(l + 1) <= id) ++l;
    return l;
}
static inline void id2index(int64_t id, int level, int *ix, int *iy) {
    int64_t local = id - start_id(level);
    int64_t cnt = (int64_t)1 << level;
    *ix = (int)(local % cnt);
    *iy = (int)(local / cnt);
}
/* treeCell.cpp:336-368 */
static inline int64_t parent_id(int64_t id) {
    int l = id2level(id);
    int ix, iy;
    id2index(id, l, &ix, &iy);
    return index2id(l - 1, ix / 2, iy / 2);
}

/* ------------------------------------------------------------------------ */
/* tree                                                                      */
/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t id;
    int level, ix, iy;
    double len, cx, cy;
    int par_num, src_num;
    int is_leaf, is_active;
    int poff, pcnt;        /* leaves: slice of plist */
} orc_cell;

struct orc_tree {
    int n_cells, cap_cells;
    orc_cell *cells;
    int64_t *hk; int *hv; int hcap;   /* id -> cell index */
    int *leaf_list; int n_leaves;     /* reference leafList order */
    int *plist;                       /* leaf particle lists, leaf order */
    int min_level, max_level;
    double root_len, pivot[2];
    int n;
};

static inline uint64_t hmix(uint64_t k) {
    k ^= k >> 33; k *= 0xff51afd7ed558ccdULL; k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ULL; k ^= k >> 33; return k;
}
static int h_find(const orc_tree *t, int64_t id) {
    uint64_t m = (uint64_t)t->hcap - 1, h = hmix((uint64_t)id) & m;
    while (t->hk[h] != -1) {
        if (t->hk[h] == id) return t->hv[h];
        h = (h + 1) & m;
    }
    return -1;
}
static int h_grow(orc_tree *t);
static int h_insert(orc_tree *t, int64_t id, int v) {
    if ((int64_t)(t->n_cells + 1) * 2 > t->hcap && h_grow(t)) return -1;
    uint64_t m = (uint64_t)t->hcap - 1, h = hmix((uint64_t)id) & m;
    while (t->hk[h] != -1) h = (h + 1) & m;
    t->hk[h] = id; t->hv[h] = v;
    return 0;
}
static int h_grow(orc_tree *t) {
    int ncap = t->hcap ? t->hcap * 2 : 1024;
    int64_t *nk = (int64_t *)malloc(sizeof(int64_t) * ncap);
    int *nv = (int *)malloc(sizeof(int) * ncap);
    if (!nk || !nv) { free(nk); free(nv); return -1; }
    for (int i = 0; i < ncap; ++i) nk[i] = -1;
    uint64_t m = (uint64_t)ncap - 1;
    for (int i = 0; i < t->hcap; ++i) {
        if (t->hk[i] == -1) continue;
        uint64_t h = hmix((uint64_t)t->hk[i]) & m;
        while (nk[h] != -1) h = (h + 1) & m;
        nk[h] = t->hk[i]; nv[h] = t->hv[i];
    }
    free(t->hk); free(t->hv);
    t->hk = nk; t->hv = nv; t->hcap = ncap;
    return 0;
}

static int add_cell(orc_tree *t, int level, int ix, int iy, double len) {
    if (t->n_cells == t->cap_cells) {
        int nc = t->cap_cells ? t->cap_cells * 2 : 1024;
        orc_cell *p = (orc_cell *)realloc(t->cells, sizeof(orc_cell) * nc);
        if (!p) return -1;
        t->cells = p; t->cap_cells = nc;
    }
    orc_cell *c = &t->cells[t->n_cells];
    memset(c, 0, sizeof(*c));
    c->id = index2id(level, ix, iy);
    c->level = level; c->ix = ix; c->iy = iy; c->len = len;
    /* treeCell.hpp:68: centre = pivot + (0.5 + index) * length */
    c->cx = t->pivot[0] + ((0.5 + ix) * len);
    c->cy = t->pivot[1] + ((0.5 + iy) * len);
    if (h_insert(t, c->id, t->n_cells)) return -1;
    return t->n_cells++;
}

void orc_tree_free(orc_tree *t) {
    if (!t) return;
    free(t->cells); free(t->hk); free(t->hv); free(t->leaf_list); free(t->plist);
    free(t);
}

/* treeCell.cpp:1043-1243 (initializeTree). */
int orc_tree_build(orc_tree **out, int n, const double *x, const double *y,
                   const uint8_t *active, const orc_opts *o) {
    *out = NULL;
    if (n < 1 || !opts_ok(o)) return ORC_BAD_ARG;
    for (int i = 0; i < n; ++i)
        if (!isfinite(x[i]) || !isfinite(y[i])) return ORC_NONFINITE;

    orc_tree *t = (orc_tree *)calloc(1, sizeof(orc_tree));
    if (!t) return ORC_NOMEM;
    t->n = n;
    if (h_grow(t)) { orc_tree_free(t); return ORC_NOMEM; }

    /* bounding box over every particle, seeded by particle 0 (:1064-1075) */
    double mn[2] = {x[0], y[0]}, mx[2] = {x[0], y[0]};
    for (int i = 1; i < n; ++i) {
        if (mn[0] > x[i]) mn[0] = x[i];
        if (mx[0] < x[i]) mx[0] = x[i];
        if (mn[1] > y[i]) mn[1] = y[i];
        if (mx[1] < y[i]) mx[1] = y[i];
    }
    /* root length: longest side times (1 + expTree/100) (:1079-1087) */
    double R = 0.0;
    for (int d = 0; d < 2; ++d) {
        double s = fabs(mx[d] - mn[d]);
        if (s > R) R = s;
    }
    R *= 1.0 + (o->exp_tree / 100.0);
    if (!(R > 0.0)) { orc_tree_free(t); return ORC_DEGENERATE; }
    t->root_len = R;
    /* pivot (:1092) */
    for (int d = 0; d < 2; ++d) t->pivot[d] = (mx[d] + mn[d] - R) / 2.0;

    /* root cell (:1117-1133) */
    int nact = 0;
    for (int i = 0; i < n; ++i) nact += active[i] ? 1 : 0;
    if (add_cell(t, 0, 0, 0, R) < 0) { orc_tree_free(t); return ORC_NOMEM; }
    t->cells[0].src_num = nact;
    t->cells[0].is_active = nact != 0;
    t->cells[0].par_num = n;
    t->cells[0].is_leaf = 0;

    uint8_t *settled = (uint8_t *)calloc(n, 1);
    int *cell_of = (int *)malloc(sizeof(int) * n);
    int *newc = NULL; int n_new = 0, cap_new = 0;
    t->leaf_list = (int *)malloc(sizeof(int) * n);
    t->plist = (int *)malloc(sizeof(int) * n);
    if (!settled || !cell_of || !t->leaf_list || !t->plist) {
        free(settled); free(cell_of); orc_tree_free(t); return ORC_NOMEM;
    }

    int status = ORC_OK;
    int unsettled = n, level = 0;
    while (unsettled > 0) {
        ++level;
        if (level > ORC_MAX_DEPTH) { status = ORC_DEPTH; break; }
        double size = R / pow(2.0, level);   /* get_cellSize, treeCell.cpp:101-105 */
        int cnt = 1 << level;
        n_new = 0;
        /* bin every unsettled particle at this level (:1152-1178) */
        for (int i = 0; i < n; ++i) {
            if (settled[i]) continue;
            int ix = (int)((x[i] - t->pivot[0]) / size);   /* coord2index :290-303 */
            int iy = (int)((y[i] - t->pivot[1]) / size);
            if (ix < 0 || ix >= cnt || iy < 0 || iy >= cnt) { status = ORC_INTERNAL; break; }
            int64_t id = index2id(level, ix, iy);
            int ci = h_find(t, id);
            if (ci < 0) {
                ci = add_cell(t, level, ix, iy, size);
                if (ci < 0) { status = ORC_NOMEM; break; }
                if (n_new == cap_new) {
                    cap_new = cap_new ? cap_new * 2 : 1024;
                    int *p = (int *)realloc(newc, sizeof(int) * cap_new);
                    if (!p) { status = ORC_NOMEM; break; }
                    newc = p;
                }
                newc[n_new++] = ci;
            }
            cell_of[i] = ci;
            t->cells[ci].par_num++;
            if (active[i]) t->cells[ci].src_num++;
        }
        if (status != ORC_OK) break;
        /* leaf and active marks for the cells created at this level (:1186-1211);
         * min_level / max_level extend the rule (reference: 1 / none) */
        for (int k = 0; k < n_new; ++k) {
            orc_cell *c = &t->cells[newc[k]];
            int leaf = (c->src_num <= o->max_src && c->par_num <= o->max_par &&
                        level >= o->min_level) ||
                       (o->max_level && level >= o->max_level);
            if (leaf) {
                c->is_leaf = 1;
                t->leaf_list[t->n_leaves++] = newc[k];
            }
            if (c->src_num != 0) c->is_active = 1;
        }
        /* settle particles of leaves, in particle-index order (:1215-1232) */
        for (int i = 0; i < n; ++i) {
            if (settled[i]) continue;
            if (t->cells[cell_of[i]].is_leaf) {
                settled[i] = 1;       /* cell_of[i] keeps the leaf index */
                --unsettled;
            }
        }
    }
    if (status == ORC_OK) {
        /* particle lists: leaf order, particle-index order inside a leaf */
        int off = 0;
        for (int k = 0; k < t->n_leaves; ++k) {
            orc_cell *c = &t->cells[t->leaf_list[k]];
            c->poff = off; c->pcnt = 0;
            off += c->par_num;
        }
        for (int i = 0; i < n; ++i) {
            orc_cell *c = &t->cells[cell_of[i]];
            t->plist[c->poff + c->pcnt++] = i;
        }
        t->min_level = t->cells[t->leaf_list[0]].level;               /* :1236-1240 */
        t->max_level = t->cells[t->leaf_list[t->n_leaves - 1]].level;
    }
    free(settled); free(cell_of); free(newc);
    if (status != ORC_OK) { orc_tree_free(t); return status; }
    *out = t;
    return ORC_OK;
}

int orc_tree_counts(const orc_tree *t, int *n_cells, int *n_leaves,
                    int *min_level, int *max_level) {
    *n_cells = t->n_cells; *n_leaves = t->n_leaves;
    *min_level = t->min_level; *max_level = t->max_level;
    return ORC_OK;
}

void orc_tree_geometry(const orc_tree *t, double *root_len, double *pivot2) {
    *root_len = t->root_len; pivot2[0] = t->pivot[0]; pivot2[1] = t->pivot[1];
}

static int cmp_cell_id(const void *a, const void *b, void *ctx) {
    const orc_cell *c = (const orc_cell *)ctx;
    int64_t ia = c[*(const int *)a].id, ib = c[*(const int *)b].id;
    return ia < ib ? -1 : ia > ib;
}
/* qsort_r is a GNU extension; keep a tiny insertion-free merge sort instead */
static void sort_by_id(const orc_tree *t, int *ord, int n) {
    if (n < 2) return;
    int *tmp = (int *)malloc(sizeof(int) * n);
    for (int w = 1; w < n; w *= 2) {
        for (int lo = 0; lo < n; lo += 2 * w) {
            int mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
            int a = lo, b = mid, k = lo;
            while (a < mid && b < hi)
                tmp[k++] = cmp_cell_id(&ord[a], &ord[b], t->cells) <= 0 ? ord[a++] : ord[b++];
            while (a < mid) tmp[k++] = ord[a++];
            while (b < hi) tmp[k++] = ord[b++];
        }
        memcpy(ord, tmp, sizeof(int) * n);
    }
    free(tmp);
}

void orc_tree_cells(const orc_tree *t, int64_t *id, int *level, int *ix, int *iy,
                    int *par_num, int *src_num, uint8_t *is_leaf, uint8_t *is_active) {
    int *ord = (int *)malloc(sizeof(int) * t->n_cells);
    for (int i = 0; i < t->n_cells; ++i) ord[i] = i;
    sort_by_id(t, ord, t->n_cells);
    for (int k = 0; k < t->n_cells; ++k) {
        const orc_cell *c = &t->cells[ord[k]];
        id[k] = c->id; level[k] = c->level; ix[k] = c->ix; iy[k] = c->iy;
        par_num[k] = c->par_num; src_num[k] = c->src_num;
        is_leaf[k] = (uint8_t)c->is_leaf; is_active[k] = (uint8_t)c->is_active;
    }
    free(ord);
}

void orc_tree_leaves(const orc_tree *t, int64_t *leaf_id, int *poff, int *pidx) {
    int off = 0;
    for (int k = 0; k < t->n_leaves; ++k) {
        const orc_cell *c = &t->cells[t->leaf_list[k]];
        leaf_id[k] = c->id;
        poff[k] = off;
        for (int j = 0; j < c->pcnt; ++j) pidx[off + j] = t->plist[c->poff + j];
        off += c->pcnt;
    }
    poff[t->n_leaves] = off;
}

/* ------------------------------------------------------------------------ */
/* lists                                                                     */
/* ------------------------------------------------------------------------ */
typedef struct { int64_t *v; int n, cap; } idvec;
static void iv_push(idvec *a, int64_t x) {
    if (a->n == a->cap) {
        a->cap = a->cap ? a->cap * 2 : 32;
        a->v = (int64_t *)realloc(a->v, sizeof(int64_t) * a->cap);
    }
    a->v[a->n++] = x;
}
static int iv_has(const idvec *a, int64_t x) {
    for (int i = 0; i < a->n; ++i) if (a->v[i] == x) return 1;
    return 0;
}

/* treeCell.cpp:562-625: 3x3 same-level candidates incl. self, x fastest. */
static int same_level_nbrs(const orc_cell *c, int64_t out[9]) {
    int cnt = 1 << c->level, m = 0;
    for (int i = 0; i < 9; ++i) {
        int dx = (i % 3) - 1, dy = ((i / 3) % 3) - 1;
        int jx = c->ix + dx, jy = c->iy + dy;
        if (jx < 0 || jx >= cnt || jy < 0 || jy >= cnt) continue;
        out[m++] = index2id(c->level, jx, jy);
    }
    return m;
}
/* treeCell.cpp:378-417: child i has index 2*idx + ((i >> d) & 1). */
static void child_ids(const orc_cell *c, int64_t out[4]) {
    for (int i = 0; i < 4; ++i)
        out[i] = index2id(c->level + 1, 2 * c->ix + (i & 1), 2 * c->iy + ((i >> 1) & 1));
}

/* treeCell.cpp:636-804 (intList): lists 1 and 3 of a cell. */
static void int_list(const orc_tree *t, const orc_cell *cur, idvec *l1, idvec *l3) {
    idvec q = {0}, flag = {0}, parflag = {0};
    int64_t nb[9];
    int m = same_level_nbrs(cur, nb);
    for (int i = 0; i < m; ++i) { iv_push(&q, nb[i]); iv_push(&flag, nb[i]); }
    for (int qi = 0; qi < q.n; ++qi) {
        int64_t id = q.v[qi];
        int ci = h_find(t, id);
        if (ci < 0) {                             /* :706-720 parent walk */
            int64_t pid = parent_id(id);
            if (!iv_has(&flag, pid)) {
                iv_push(&q, pid); iv_push(&flag, pid); iv_push(&parflag, pid);
            }
            continue;
        }
        const orc_cell *ng = &t->cells[ci];
        if (!ng->is_active) continue;             /* :726 */
        if (ng->is_leaf) { iv_push(l1, id); continue; }   /* :729-735 */
        if (iv_has(&parflag, id)) continue;       /* :738 */
        int64_t ch[4];
        child_ids(ng, ch);
        double adj = ((cur->len + (ng->len / 2.0)) / 2.0) + ORC_EPS;   /* :743 */
        for (int k = 0; k < 4; ++k) {
            int cc = h_find(t, ch[k]);
            if (cc < 0) continue;
            const orc_cell *c = &t->cells[cc];
            if (!c->is_active) continue;
            int is_adj = fabs(c->cx - cur->cx) <= adj && fabs(c->cy - cur->cy) <= adj;
            if (is_adj) {
                if (!iv_has(&flag, ch[k])) { iv_push(&q, ch[k]); iv_push(&flag, ch[k]); }
            } else {
                iv_push(l3, ch[k]);
            }
        }
    }
    free(q.v); free(flag.v); free(parflag.v);
}

/* treeCell.cpp:818-1032 (extList): lists 2 and 4 of a cell (level >= 2). */
static void ext_list(const orc_tree *t, const orc_cell *cur, idvec *l2, idvec *l4) {
    int pi = h_find(t, parent_id(cur->id));
    const orc_cell *par = &t->cells[pi];
    idvec q4 = {0}, flag4 = {0};
    int64_t nb[9];
    int m = same_level_nbrs(par, nb);
    for (int i = 0; i < m; ++i) {
        int64_t id = nb[i];
        if (id == par->id) continue;                          /* :872 */
        int ci = h_find(t, id);
        if (ci < 0) {                                         /* :875-891 */
            int64_t pid = parent_id(id);
            if (!iv_has(&flag4, pid)) { iv_push(&q4, pid); iv_push(&flag4, pid); }
            continue;
        }
        const orc_cell *ng = &t->cells[ci];
        if (!ng->is_active) continue;                         /* :897 */
        if (ng->is_leaf) { iv_push(&q4, id); iv_push(&flag4, id); continue; }  /* :900-904 */
        int64_t ch[4];
        child_ids(ng, ch);
        double md = cur->len + ORC_EPS;                       /* :910 */
        for (int k = 0; k < 4; ++k) {
            int cc = h_find(t, ch[k]);
            if (cc < 0) continue;
            const orc_cell *c = &t->cells[cc];
            if (!c->is_active) continue;
            int sep = fabs(c->cx - cur->cx) > md || fabs(c->cy - cur->cy) > md;
            if (sep) iv_push(l2, ch[k]);
        }
    }
    for (int qi = 0; qi < q4.n; ++qi) {                       /* :964-1029 */
        int64_t id = q4.v[qi];
        int ci = h_find(t, id);
        if (ci < 0) {
            int64_t pid = parent_id(id);
            if (!iv_has(&flag4, pid)) { iv_push(&q4, pid); iv_push(&flag4, pid); }
            continue;
        }
        const orc_cell *ng = &t->cells[ci];
        if (!ng->is_active) continue;
        if (!ng->is_leaf) continue;
        double md = ((ng->len + cur->len) / 2.0) + ORC_EPS;   /* :994 */
        int sep = fabs(ng->cx - cur->cx) > md || fabs(ng->cy - cur->cy) > md;
        if (sep) iv_push(l4, id);
    }
    free(q4.v); free(flag4.v);
}

int orc_int_list(const orc_tree *t, int64_t id, int64_t *l1, int *n1,
                 int64_t *l3, int *n3, int cap) {
    int ci = h_find(t, id);
    if (ci < 0) return ORC_BAD_ARG;
    idvec a = {0}, b = {0};
    int_list(t, &t->cells[ci], &a, &b);
    *n1 = a.n; *n3 = b.n;
    for (int i = 0; i < a.n && i < cap; ++i) l1[i] = a.v[i];
    for (int i = 0; i < b.n && i < cap; ++i) l3[i] = b.v[i];
    free(a.v); free(b.v);
    return ORC_OK;
}

int orc_ext_list(const orc_tree *t, int64_t id, int64_t *l2, int *n2,
                 int64_t *l4, int *n4, int cap) {
    int ci = h_find(t, id);
    if (ci < 0 || t->cells[ci].level < 2) return ORC_BAD_ARG;
    idvec a = {0}, b = {0};
    ext_list(t, &t->cells[ci], &a, &b);
    *n2 = a.n; *n4 = b.n;
    for (int i = 0; i < a.n && i < cap; ++i) l2[i] = a.v[i];
    for (int i = 0; i < b.n && i < cap; ++i) l4[i] = b.v[i];
    free(a.v); free(b.v);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* expansion operators (single translations)                                 */
/* ------------------------------------------------------------------------ */
/* fmm2D.cpp:260-279: a_p = sum q (z - zc)^p over the given particles. */
void orc_p2m(int np, const double *x, const double *y, const double *q,
             double cxx, double cyy, int p, double *ak) {
    for (int j = 0; j < np; ++j) {
        cx_t z = cx(x[j] - cxx, y[j] - cyy), pw = cx(1.0, 0.0);
        ak[0] += q[j];
        for (int k = 1; k <= p; ++k) {
            pw = cmul(pw, z);
            ak[2 * k] += q[j] * pw.re;
            ak[2 * k + 1] += q[j] * pw.im;
        }
    }
}

static double binom_sel(int binom_mode, int n, int k) {
    return binom_mode ? (double)orc_binom_ref(n, k) : orc_binom_exact(n, k);
}

/* fmm2D.cpp:361-376 with d = z_child - z_parent.  binComb(p-1, k-1) also
 * wraps in int32 once P >= 19, hence binom_mode here too. */
void orc_m2m(const double *akc, double dx, double dy, int p, int binom_mode, double *akp) {
    cx_t d = cx(dx, dy);
    cx_t *dp = (cx_t *)malloc(sizeof(cx_t) * (p + 1));
    dp[0] = cx(1.0, 0.0);
    for (int k = 1; k <= p; ++k) dp[k] = cmul(dp[k - 1], d);
    cx_t a0 = cx(akc[0], akc[1]);
    akp[0] += a0.re; akp[1] += a0.im;
    for (int m = 1; m <= p; ++m) {
        cx_t s = cmul(a0, dp[m]);
        for (int k = 1; k <= m; ++k) {
            double c = binom_sel(binom_mode, m - 1, k - 1) * ((double)m / (double)k);
            s = cadd(s, cscale(cmul(cx(akc[2 * k], akc[2 * k + 1]), dp[m - k]), c));
        }
        akp[2 * m] += s.re; akp[2 * m + 1] += s.im;
    }
    free(dp);
}

/* fmm2D.cpp:465-481 with d = z_src - z_tgt; b_0 is not formed (it never
 * reaches the field: L2L feeds b_p from b_k, k >= p, and L2P starts at 1). */
void orc_m2l(const double *aks, double dx, double dy, int p, int binom_mode, double *bk) {
    cx_t w = cinv(cx(dx, dy));
    cx_t *wp = (cx_t *)malloc(sizeof(cx_t) * (2 * p + 1));
    wp[0] = cx(1.0, 0.0);
    for (int k = 1; k <= 2 * p; ++k) wp[k] = cmul(wp[k - 1], w);
    cx_t a0 = cx(aks[0], aks[1]);
    for (int m = 1; m <= p; ++m) {
        cx_t s = cscale(cmul(a0, wp[m]), 1.0 / (double)m);
        for (int k = 1; k <= p; ++k) {
            double c = binom_sel(binom_mode, m + k - 1, k - 1);
            double sg = (k & 1) ? -1.0 : 1.0;
            s = cadd(s, cscale(cmul(cx(aks[2 * k], aks[2 * k + 1]), wp[m + k]), c * sg / (double)k));
        }
        bk[2 * m] -= s.re; bk[2 * m + 1] -= s.im;
    }
    free(wp);
}

/* fmm2D.cpp:463-467 (list 2) and 504-508 (list 4), the order-0 row that only
 * the potential reads: b_0 += a_0 log(z_t - z_s) - sum_k a_k/k (-1)^k d^-k,
 * d = z_s - z_t.  log is the principal branch of std::log: log|z| + i arg z,
 * with z_t - z_s = -d exactly. */
void orc_m2l0(const double *aks, double dx, double dy, int p, double *bk) {
    cx_t w = cinv(cx(dx, dy)), pw = w;
    cx_t s = cmul(cx(aks[0], aks[1]), cx(log(hypot(dx, dy)), atan2(-dy, -dx)));
    for (int k = 1; k <= p; ++k) {
        double sg = (k & 1) ? -1.0 : 1.0;
        s = csub(s, cscale(cmul(cx(aks[2 * k], aks[2 * k + 1]), pw), sg / (double)k));
        pw = cmul(pw, w);
    }
    bk[0] += s.re; bk[1] += s.im;
}

/* fmm2D.cpp:637-643 with d = z_child - z_parent (binComb(k, p) wraps from P = 18). */
void orc_l2l(const double *bkp, double dx, double dy, int p, int binom_mode, double *bkc) {
    cx_t d = cx(dx, dy);
    cx_t *dp = (cx_t *)malloc(sizeof(cx_t) * (p + 1));
    dp[0] = cx(1.0, 0.0);
    for (int k = 1; k <= p; ++k) dp[k] = cmul(dp[k - 1], d);
    for (int m = 0; m <= p; ++m) {
        cx_t s = cx(0.0, 0.0);
        for (int k = m; k <= p; ++k)
            s = cadd(s, cscale(cmul(cx(bkp[2 * k], bkp[2 * k + 1]), dp[k - m]),
                               binom_sel(binom_mode, k, m)));
        bkc[2 * m] += s.re; bkc[2 * m + 1] += s.im;
    }
    free(dp);
}

/* fmm2D.cpp:1052-1062: F = sum a_p (z - zc)^-(p+1); Ex += Re F, Ey -= Im F. */
void orc_m2p(const double *ak, double dx, double dy, int p, double *ex, double *ey) {
    cx_t w = cinv(cx(dx, dy)), pw = w, f = cx(0.0, 0.0);
    for (int k = 0; k <= p; ++k) {
        f = cadd(f, cmul(cx(ak[2 * k], ak[2 * k + 1]), pw));
        pw = cmul(pw, w);
    }
    *ex += f.re; *ey -= f.im;
}

/* fmm2D.cpp:1083-1093: F = sum_{p>=1} p b_p (z - z0)^(p-1). */
void orc_l2p(const double *bk, double dx, double dy, int p, double *ex, double *ey) {
    cx_t z = cx(dx, dy), pw = cx(1.0, 0.0), f = cx(0.0, 0.0);
    for (int k = 1; k <= p; ++k) {
        f = cadd(f, cscale(cmul(cx(bk[2 * k], bk[2 * k + 1]), pw), (double)k));
        pw = cmul(pw, z);
    }
    *ex += f.re; *ey -= f.im;
}

/* ------------------------------------------------------------------------ */
/* full FMM (fmm2D::setupFMM + calcField)                                    */
/* ------------------------------------------------------------------------ */
static int cmp_i64(const void *a, const void *b) {
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return x < y ? -1 : x > y;
}

/* Core of fmm2D::setupFMM + calcField.  With ids/ak_out/bk_out non-NULL the
 * per-cell expansions are also returned, cells sorted by ID. */
static int fmm_run(int n, const double *x, const double *y, const double *q,
                   const double *sz, const uint8_t *active, const orc_opts *o, double *ex,
                   double *ey, int64_t *ids, double *ak_out, double *bk_out, double *phi) {
    if (o && o->icutoff > 0 && !sz) return ORC_BAD_ARG;
    if (phi && o && o->icutoff > 0) return ORC_BAD_ARG;  /* no regularized potential */
    orc_tree *t = NULL;
    int st = orc_tree_build(&t, n, x, y, active, o);
    if (st) return st;
    const int P = o->p, W = 2 * (P + 1);
    double *ak = (double *)calloc((size_t)t->n_cells * W, sizeof(double));
    double *bk = (double *)calloc((size_t)t->n_cells * W, sizeof(double));
    if (!ak || !bk) { free(ak); free(bk); orc_tree_free(t); return ORC_NOMEM; }

    /* cells grouped by level, ascending ID inside a level */
    int maxL = t->max_level;
    int *ord = (int *)malloc(sizeof(int) * t->n_cells);
    for (int i = 0; i < t->n_cells; ++i) ord[i] = i;
    sort_by_id(t, ord, t->n_cells);
    int *lstart = (int *)calloc(ORC_MAX_DEPTH + 2, sizeof(int));
    for (int i = 0; i < t->n_cells; ++i) lstart[t->cells[ord[i]].level + 1]++;
    for (int l = 0; l <= ORC_MAX_DEPTH; ++l) lstart[l + 1] += lstart[l];

    double *sx = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    double *sy = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    double *sq = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));

    /* P2M over active leaves, active particles only (fmm2D.cpp:237-284) */
    for (int li = 0; li < t->n_leaves; ++li) {
        orc_cell *c = &t->cells[t->leaf_list[li]];
        if (!c->is_active) continue;
        int m = 0;
        for (int j = 0; j < c->pcnt; ++j) {
            int pj = t->plist[c->poff + j];
            if (!active[pj]) continue;
            sx[m] = x[pj]; sy[m] = y[pj]; sq[m] = q[pj]; ++m;
        }
        orc_p2m(m, sx, sy, sq, c->cx, c->cy, P, &ak[(size_t)(c - t->cells) * W]);
    }

    /* M2M, levels maxL-1 .. 1, non-leaf active cells (fmm2D.cpp:309-379) */
    for (int l = maxL - 1; l > 0; --l) {
        for (int k = lstart[l]; k < lstart[l + 1]; ++k) {
            orc_cell *c = &t->cells[ord[k]];
            if (c->is_leaf || !c->is_active) continue;
            int64_t ch[4];
            child_ids(c, ch);
            for (int i = 0; i < 4; ++i) {
                int cc = h_find(t, ch[i]);
                if (cc < 0) continue;
                orc_cell *d = &t->cells[cc];
                orc_m2m(&ak[(size_t)cc * W], d->cx - c->cx, d->cy - c->cy, P, o->binom_mode,
                        &ak[(size_t)(c - t->cells) * W]);
            }
        }
    }

    /* M2L for every cell of level >= 2 from lists 2 and 4 (fmm2D.cpp:409-556) */
    idvec l2 = {0}, l4 = {0};
    for (int k = lstart[2]; k < t->n_cells; ++k) {
        orc_cell *c = &t->cells[ord[k]];
        size_t cidx = (size_t)(c - t->cells);
        l2.n = 0; l4.n = 0;
        ext_list(t, c, &l2, &l4);
        for (int i = 0; i < l2.n; ++i) {
            int sidx = h_find(t, l2.v[i]);
            orc_cell *s = &t->cells[sidx];
            if (phi)
                orc_m2l0(&ak[(size_t)sidx * W], s->cx - c->cx, s->cy - c->cy, P, &bk[cidx * W]);
            orc_m2l(&ak[(size_t)sidx * W], s->cx - c->cx, s->cy - c->cy, P, o->binom_mode,
                    &bk[cidx * W]);
        }
        for (int i = 0; i < l4.n; ++i) {
            int sidx = h_find(t, l4.v[i]);
            orc_cell *s = &t->cells[sidx];
            if (o->list4_mode == 1) {
                if (phi)
                    orc_m2l0(&ak[(size_t)sidx * W], s->cx - c->cx, s->cy - c->cy, P,
                             &bk[cidx * W]);
                orc_m2l(&ak[(size_t)sidx * W], s->cx - c->cx, s->cy - c->cy, P,
                        o->binom_mode, &bk[cidx * W]);
            } else {
                /* fmm2D.cpp:526-553: P2L from the active particles */
                for (int j = 0; j < s->pcnt; ++j) {
                    int pj = t->plist[s->poff + j];
                    if (!active[pj]) continue;
                    if (phi) { /* b_0 += q log(z_t - z_j) (:540) */
                        double dxx = c->cx - x[pj], dyy = c->cy - y[pj];
                        bk[cidx * W] += q[pj] * log(hypot(dxx, dyy));
                        bk[cidx * W + 1] += q[pj] * atan2(dyy, dxx);
                    }
                    cx_t w = cinv(cx(x[pj] - c->cx, y[pj] - c->cy)), pw = w;
                    for (int m = 1; m <= P; ++m) {
                        bk[cidx * W + 2 * m] -= q[pj] * pw.re / (double)m;
                        bk[cidx * W + 2 * m + 1] -= q[pj] * pw.im / (double)m;
                        pw = cmul(pw, w);
                    }
                }
            }
        }
    }
    free(l2.v); free(l4.v);

    /* L2L, levels 2 .. maxL-1, every non-leaf cell (fmm2D.cpp:587-651) */
    for (int l = 2; l < maxL; ++l) {
        for (int k = lstart[l]; k < lstart[l + 1]; ++k) {
            orc_cell *c = &t->cells[ord[k]];
            if (c->is_leaf) continue;
            int64_t ch[4];
            child_ids(c, ch);
            for (int i = 0; i < 4; ++i) {
                int cc = h_find(t, ch[i]);
                if (cc < 0) continue;
                orc_cell *d = &t->cells[cc];
                orc_l2l(&bk[(size_t)(c - t->cells) * W], d->cx - c->cx, d->cy - c->cy, P, o->binom_mode,
                        &bk[(size_t)cc * W]);
            }
        }
    }

    /* evaluation per leaf (fmm2D.cpp:940-1106); stage_mask selects the parts */
    const int mask = o->stage_mask ? o->stage_mask : 7;
    const int do_p2p = mask & 1, do_m2p = mask & 2, do_l2p = mask & 4;
    for (int i = 0; i < n; ++i) { ex[i] = 0.0; ey[i] = 0.0; }
    if (phi)
        for (int i = 0; i < n; ++i) phi[i] = 0.0;
    idvec l1 = {0}, l3 = {0};
    for (int li = 0; li < t->n_leaves; ++li) {
        orc_cell *c = &t->cells[t->leaf_list[li]];
        l1.n = 0; l3.n = 0;
        int_list(t, c, &l1, &l3);
        /* P2P over every particle of the list-1 cells, skipping j == i (:135-179) */
        for (int a = 0; a < (do_p2p ? l1.n : 0); ++a) {
            orc_cell *s = &t->cells[h_find(t, l1.v[a])];
            for (int jj = 0; jj < s->pcnt; ++jj) {
                int j = t->plist[s->poff + jj];
                for (int ii = 0; ii < c->pcnt; ++ii) {
                    int i = t->plist[c->poff + ii];
                    if (i == j) continue;
                    double dx = x[i] - x[j], dy = y[i] - y[j];
                    double r2 = dx * dx + dy * dy;
                    if (phi) { /* potDirSum (fmm2D.cpp:113-115) */
                        phi[i] += q[j] * log(r2) / 2.0;
                        continue;
                    }
                    if (o->icutoff > 0) {
                        /* the type-1 near field, direct_sum
                           (subroutine_all_fmm_2d.cpp:594-627): sij^2 =
                           (si^2 + sj^2) / 4, pairs at r = 0 skipped */
                        if (!(r2 > 0.0)) continue;
                        double sij2 = 0.25 * (pow(sz[i], 2) + pow(sz[j], 2));
                        double f = regul_factor(o->icutoff, sqrt(r2), sqrt(sij2));
                        ex[i] += f * q[j] * dx / r2;
                        ey[i] += f * q[j] * dy / r2;
                        continue;
                    }
                    ex[i] += q[j] * dx / r2;
                    ey[i] += q[j] * dy / r2;
                }
            }
        }
        for (int ii = 0; phi && ii < c->pcnt; ++ii) {
            /* calcPotential per target (fmm2D.cpp:817-880): M2P over list 3,
               a_0 log(z - zc) - sum a_p (z - zc)^-p / p, then L2P, b_0 +
               sum b_p (z - z0)^p; the real parts */
            int i = t->plist[c->poff + ii];
            for (int a = 0; a < (do_m2p ? l3.n : 0); ++a) {
                int sidx = h_find(t, l3.v[a]);
                orc_cell *s = &t->cells[sidx];
                const double *as = &ak[(size_t)sidx * W];
                double dx = x[i] - s->cx, dy = y[i] - s->cy;
                cx_t w = cinv(cx(dx, dy)), pw = w;
                cx_t f = cmul(cx(as[0], as[1]), cx(log(hypot(dx, dy)), atan2(dy, dx)));
                for (int k = 1; k <= P; ++k) {
                    f = csub(f, cscale(cmul(cx(as[2 * k], as[2 * k + 1]), pw), 1.0 / (double)k));
                    pw = cmul(pw, w);
                }
                phi[i] += f.re;
            }
            if (do_l2p) {
                const double *bs = &bk[(size_t)(c - t->cells) * W];
                cx_t z = cx(x[i] - c->cx, y[i] - c->cy), pw = z, f = cx(bs[0], bs[1]);
                for (int k = 1; k <= P; ++k) {
                    f = cadd(f, cmul(cx(bs[2 * k], bs[2 * k + 1]), pw));
                    pw = cmul(pw, z);
                }
                phi[i] += f.re;
            }
        }
        for (int ii = 0; !phi && ii < c->pcnt; ++ii) {
            int i = t->plist[c->poff + ii];
            for (int a = 0; a < (do_m2p ? l3.n : 0); ++a) {       /* M2P over list 3 */
                int sidx = h_find(t, l3.v[a]);
                orc_cell *s = &t->cells[sidx];
                orc_m2p(&ak[(size_t)sidx * W], x[i] - s->cx, y[i] - s->cy, P, &ex[i], &ey[i]);
            }
            if (do_l2p)
                orc_l2p(&bk[(size_t)(c - t->cells) * W], x[i] - c->cx, y[i] - c->cy, P,
                        &ex[i], &ey[i]);
        }
    }
    free(l1.v); free(l3.v);
    if (ids)
        for (int k = 0; k < t->n_cells; ++k) {
            ids[k] = t->cells[ord[k]].id;
            memcpy(&ak_out[(size_t)k * W], &ak[(size_t)ord[k] * W], sizeof(double) * W);
            memcpy(&bk_out[(size_t)k * W], &bk[(size_t)ord[k] * W], sizeof(double) * W);
        }
    free(sx); free(sy); free(sq);
    free(ord); free(lstart); free(ak); free(bk);
    orc_tree_free(t);
    (void)cmp_i64;
    return ORC_OK;
}

int orc_fmm_field(int n, const double *x, const double *y, const double *q,
                  const uint8_t *active, const orc_opts *o, double *ex, double *ey) {
    return fmm_run(n, x, y, q, NULL, active, o, ex, ey, NULL, NULL, NULL, NULL);
}

int orc_fmm_potential(int n, const double *x, const double *y, const double *q,
                      const uint8_t *active, const orc_opts *o, double *phi) {
    double *ex = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    double *ey = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    if (!ex || !ey) { free(ex); free(ey); return ORC_NOMEM; }
    int st = fmm_run(n, x, y, q, NULL, active, o, ex, ey, NULL, NULL, NULL, phi);
    free(ex); free(ey);
    return st;
}

int orc_fmm_expansions(int n, const double *x, const double *y, const double *q,
                       const uint8_t *active, const orc_opts *o, int64_t *ids, double *ak,
                       double *bk) {
    double *ex = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    double *ey = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    if (!ex || !ey) { free(ex); free(ey); return ORC_NOMEM; }
    int st = fmm_run(n, x, y, q, NULL, active, o, ex, ey, ids, ak, bk, NULL);
    free(ex); free(ey);
    return st;
}

/* velocity_calc.cpp:465-481 and 572-575, plus get_velocity's free stream
 * (velocity_calc.cpp:226-230). */
int orc_velocity(int n, const double *x, const double *y, const double *s,
                 const double *omega, const uint8_t *active, const orc_opts *o,
                 double u_inf, double v_inf, double *u, double *v) {
    double *q = (double *)calloc(n > 0 ? n : 1, sizeof(double));
    double *ex = (double *)malloc(sizeof(double) * n);
    double *ey = (double *)malloc(sizeof(double) * n);
    if (!q || !ex || !ey) { free(q); free(ex); free(ey); return ORC_NOMEM; }
    for (int i = 0; i < n; ++i) q[i] = -omega[i] * pow(s[i], 2.0) / (2 * M_PI);
    int st = fmm_run(n, x, y, q, s, active, o, ex, ey, NULL, NULL, NULL, NULL);
    if (st == ORC_OK)
        for (int i = 0; i < n; ++i) { u[i] = ey[i] + u_inf; v[i] = -ex[i] + v_inf; }
    free(q); free(ex); free(ey);
    return st;
}

/* ------------------------------------------------------------------------ */
/* direct sum                                                                */
/* ------------------------------------------------------------------------ */
void orc_direct_field(int n, const double *x, const double *y, const double *q,
                      const uint8_t *src_mask, int nt, const int *tidx,
                      double *ex, double *ey) {
    int m = tidx ? nt : n;
#pragma omp parallel for schedule(dynamic, 16)
    for (int t = 0; t < m; ++t) {
        int i = tidx ? tidx[t] : t;
        double xi = x[i], yi = y[i], sx = 0.0, sy = 0.0;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            if (src_mask && !src_mask[j]) continue;
            double dx = xi - x[j], dy = yi - y[j];
            double r2 = dx * dx + dy * dy;
            sx += q[j] * dx / r2;
            sy += q[j] * dy / r2;
        }
        ex[t] = sx; ey[t] = sy;
    }
}

/* potDirSum's kernel summed over every source (fmm2D.cpp:113-115):
 * phi_i = sum_{j != i} q_j log(r^2) / 2.  OpenMP over targets. */
void orc_direct_potential(int n, const double *x, const double *y, const double *q,
                          const uint8_t *src_mask, int nt, const int *tidx, double *phi) {
    int m = tidx ? nt : n;
#pragma omp parallel for schedule(dynamic, 16)
    for (int t = 0; t < m; ++t) {
        int i = tidx ? tidx[t] : t;
        double s = 0.0;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            if (src_mask && !src_mask[j]) continue;
            double dx = x[i] - x[j], dy = y[i] - y[j];
            s += q[j] * log(dx * dx + dy * dy) / 2.0;
        }
        phi[t] = s;
    }
}
