/*
 * fmm3d_oracle.c -- CPU oracle for the reference's two 3-D velocity paths
 * (type 1: fast3DFMM.cpp; type 2: fmm3D::calcVelocity over the 3-D TreeCell;
 * both with the fmm3D.cpp helpers).  See fmm3d_oracle.h.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).
 *
 * Arithmetic follows the reference's expressions term by term (cell centres
 * by the same additions, the multipliers of M2P_mul_calc in the same
 * association), so results agree with it to rounding.  The evaluation walks
 * the same cells with the same decisions as its level queues; only the order
 * of the additions into a target differs (depth first here).
 */
#include "fmm3d_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define O3_TERMS 10

void orc3_default_opts(orc3_opts *o) {
    /* the reference's DIM == 3 parameter block (global.cpp:474-875) */
    o->crit_num = 30;      /* FMMCell::critNum = Pars::src_count_max (global.cpp:779) */
    o->exp_tree = 1.0;     /* global.cpp:769 */
    o->center[0] = o->center[1] = o->center[2] = 0.0;  /* global.cpp:609-611 */
    o->support = 2.0;      /* SUPPORT_RADIUS_FACTOR, fast3DFMM.cpp:4 and fmm3D.cpp:3 */
    o->max_depth = 21;
    o->type = 1;           /* VELOCITY_3D_TYPE, velocity_calc.cpp:31 */
    o->max_par = 50;       /* TreeCell::max_particle = Pars::par_count_max (global.cpp:778) */
    o->max_src = 30;       /* TreeCell::max_source = Pars::src_count_max (global.cpp:779) */
}

struct orc3_tree {
    int n, ncell, cap, depth;
    orc3_opts o;
    int *level, *parent, *npar, *chd, *flag, *pbeg, *pcnt;
    int64_t *key;
    uint8_t *active, *leaf;
    int *nsrc;       /* active particles (type 2) */
    double *xc, *yc, *zc, *size;
    int *lev_start;  /* cells of level L: [lev_start[L], lev_start[L + 1]) */
    int *pool;       /* leaf particle lists */
};

static int o3_grow(orc3_tree *t) {
    const int nc = t->cap ? 2 * t->cap : 1024;
#define O3_RE(f, ty, m)                                                   \
    do {                                                                  \
        ty *p_ = (ty *)realloc(t->f, sizeof(ty) * (size_t)nc * (m));      \
        if (!p_) return 0;                                                \
        t->f = p_;                                                        \
    } while (0)
    O3_RE(level, int, 1);
    O3_RE(parent, int, 1);
    O3_RE(npar, int, 1);
    O3_RE(chd, int, 8);
    O3_RE(flag, int, 1);
    O3_RE(pbeg, int, 1);
    O3_RE(pcnt, int, 1);
    O3_RE(key, int64_t, 1);
    O3_RE(active, uint8_t, 1);
    O3_RE(leaf, uint8_t, 1);
    O3_RE(nsrc, int, 1);
    O3_RE(xc, double, 1);
    O3_RE(yc, double, 1);
    O3_RE(zc, double, 1);
    O3_RE(size, double, 1);
#undef O3_RE
    t->cap = nc;
    return 1;
}

void orc3_free(orc3_tree *t) {
    if (!t) return;
    free(t->level);
    free(t->parent);
    free(t->npar);
    free(t->chd);
    free(t->flag);
    free(t->pbeg);
    free(t->pcnt);
    free(t->key);
    free(t->active);
    free(t->leaf);
    free(t->nsrc);
    free(t->xc);
    free(t->yc);
    free(t->zc);
    free(t->size);
    free(t->lev_start);
    free(t->pool);
    free(t);
}

static int o3_new_cell(orc3_tree *t, int lev, int par, int64_t key, double xc, double yc,
                       double zc, double size) {
    if (t->ncell == t->cap && !o3_grow(t)) return -1;
    const int c = t->ncell++;
    t->level[c] = lev;
    t->parent[c] = par;
    t->npar[c] = 0;
    for (int k = 0; k < 8; ++k) t->chd[8 * c + k] = -1;
    t->flag[c] = 0;
    t->pbeg[c] = 0;
    t->pcnt[c] = 0;
    t->key[c] = key;
    t->active[c] = 0;
    t->leaf[c] = 0;
    t->nsrc[c] = 0;
    t->xc[c] = xc;
    t->yc[c] = yc;
    t->zc[c] = zc;
    t->size[c] = size;
    return c;
}

/* FMMCell::generateTree, fast3DFMM.cpp:485-745. */
static int o3_build_t1(orc3_tree **out, int n, const double *x, const double *y,
                       const double *z, const uint8_t *active, const orc3_opts *o) {
    *out = NULL;
    if (o->crit_num < 1) return 1;
    for (int i = 0; i < n; ++i)
        if (!isfinite(x[i]) || !isfinite(y[i]) || !isfinite(z[i])) return 2;
    orc3_tree *t = (orc3_tree *)calloc(1, sizeof(orc3_tree));
    if (!t) return 5;
    t->n = n;
    t->o = *o;
    /* root (:507-548): the box of the particles and the domain centre */
    double mn[3], mx[3], len[3], cen[3];
    for (int d = 0; d < 3; ++d) mn[d] = mx[d] = o->center[d];
    const double *p[3] = {x, y, z};
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            if (p[d][i] < mn[d]) mn[d] = p[d][i];
            if (p[d][i] > mx[d]) mx[d] = p[d][i];
        }
    double root = 0.0;
    for (int d = 0; d < 3; ++d) {
        len[d] = mx[d] - mn[d];
        root = (root < len[d]) ? len[d] : root;  /* std::max(len, root) */
        cen[d] = 0.5 * (mx[d] + mn[d]);
    }
    root *= (1 + (o->exp_tree / 100));
    t->lev_start = (int *)malloc(sizeof(int) * (size_t)(o->max_depth + 3));
    t->pool = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *pcell = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    uint8_t *done = (uint8_t *)calloc((size_t)(n > 0 ? n : 1), 1);
    if (!t->lev_start || !t->pool || !pcell || !done ||
        o3_new_cell(t, 0, -1, 0, cen[0], cen[1], cen[2], root) < 0) {
        free(pcell);
        free(done);
        orc3_free(t);
        return 5;
    }
    t->npar[0] = n;
    t->active[0] = 1;  /* the root is active (:539) */
    t->lev_start[0] = 0;
    for (int i = 0; i < n; ++i) pcell[i] = 0;
    int unsettled = n, lev = 0, st = 0;
    /* level by level (:615-712) */
    while (unsettled > 0) {
        ++lev;
        if (lev > o->max_depth) {
            st = 4;
            break;
        }
        t->lev_start[lev] = t->ncell;
        for (int i = 0; i < n; ++i) {
            if (done[i]) continue;
            const int c = pcell[i];
            const int oct = (x[i] > t->xc[c]) * 1 + (y[i] > t->yc[c]) * 2 + (z[i] > t->zc[c]) * 4;
            if (t->chd[8 * c + oct] < 0) {
                const double cl = t->size[c] / 2.0;
                const int ch = o3_new_cell(t, lev, c, t->key[c] * 8 + oct,
                                           t->xc[c] + cl * (((oct >> 0) % 2) - 0.5),
                                           t->yc[c] + cl * (((oct >> 1) % 2) - 0.5),
                                           t->zc[c] + cl * (((oct >> 2) % 2) - 0.5), cl);
                if (ch < 0) {
                    st = 5;
                    break;
                }
                t->chd[8 * c + oct] = ch;
                t->flag[c] |= (1 << oct);
            }
            const int ch = t->chd[8 * c + oct];
            pcell[i] = ch;
            t->npar[ch]++;
        }
        if (st) break;
        /* settle the particles of cells with fewer than crit_num (:687-708) */
        for (int i = 0; i < n; ++i) {
            if (done[i]) continue;
            const int c = pcell[i];
            if (t->npar[c] < o->crit_num) {
                t->pcnt[c]++;
                if (active[i]) t->active[c] = 1;
                done[i] = 1;
                --unsettled;
            }
        }
    }
    if (!st) {
        t->depth = lev;
        t->lev_start[lev + 1] = t->ncell;
        for (int c = 1; c < t->ncell; ++c) t->leaf[c] = t->npar[c] < o->crit_num;
        /* leaf particle lists in index order */
        int off = 0;
        for (int c = 0; c < t->ncell; ++c) {
            t->pbeg[c] = off;
            off += t->pcnt[c];
            t->pcnt[c] = 0;
        }
        for (int i = 0; i < n; ++i) {
            const int c = pcell[i];
            t->pool[t->pbeg[c] + t->pcnt[c]++] = i;
        }
        /* active marks upward (:714-741) */
        for (int L = t->depth - 1; L > 0; --L)
            for (int c = t->lev_start[L]; c < t->lev_start[L + 1]; ++c) {
                if (t->npar[c] < o->crit_num) continue;
                for (int oc = 0; oc < 8; ++oc) {
                    if (((t->flag[c] >> oc) % 2) == 0) continue;
                    if (t->active[t->chd[8 * c + oc]]) {
                        t->active[c] = 1;
                        break;
                    }
                }
            }
    }
    free(pcell);
    free(done);
    if (st) {
        orc3_free(t);
        return st;
    }
    *out = t;
    return 0;
}

/* Cells of one level keyed by their index triple (open addressing). */
typedef struct {
    uint64_t *k;
    int *v;
    size_t cap;
} o3_hash;

static uint64_t o3_mix(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

static int o3_hash_init(o3_hash *h, size_t want) {
    size_t cap = 1024;
    while (cap < 2 * want) cap <<= 1;
    h->k = (uint64_t *)malloc(sizeof(uint64_t) * cap);
    h->v = (int *)malloc(sizeof(int) * cap);
    if (!h->k || !h->v) return 0;
    h->cap = cap;
    for (size_t i = 0; i < cap; ++i) h->v[i] = -1;
    return 1;
}

static int *o3_hash_slot(o3_hash *h, uint64_t key) {
    size_t i = o3_mix(key) & (h->cap - 1);
    while (h->v[i] >= 0 && h->k[i] != key) i = (i + 1) & (h->cap - 1);
    h->k[i] = key;
    return &h->v[i];
}

static void o3_hash_free(o3_hash *h) {
    free(h->k);
    free(h->v);
}

/* TreeCell::initializeTree (treeCell.cpp:1043-1243) in 3-D: the root box of
 * the particles alone (min / max from the first particle on, :1066-1077),
 * side = the longest extent * (1 + expTree / 100) (:1080-1088), pivot =
 * (max + min - side) / 2 (:1091); at each level a particle's cell index is
 * (int)((x - pivot) / cellSize) per axis (coord2index, :290-298; cellSize =
 * side / 2^level, :101-104), the cells created in particle order; a new cell
 * is a leaf when srcParNum <= max_source and parNum <= max_particle, active
 * when srcParNum != 0 (:1185-1199); the root is never a leaf.  The cell
 * centre is pivot + (0.5 + index) * length (Cell, treeCell.hpp:68).  Parent
 * and children are linked by index arithmetic (get_childID, :378-412). */
static int o3_build_t2(orc3_tree **out, int n, const double *x, const double *y,
                       const double *z, const uint8_t *active, const orc3_opts *o) {
    *out = NULL;
    if (o->max_par < 0 || o->max_src < 0 || n < 1) return 1;
    const double *p[3] = {x, y, z};
    double mn[3], mx[3], piv[3];
    for (int d = 0; d < 3; ++d) mn[d] = mx[d] = p[d][0];
    for (int i = 1; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            if (mn[d] > p[d][i]) mn[d] = p[d][i];
            if (mx[d] < p[d][i]) mx[d] = p[d][i];
        }
    double root = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double sz = fabs(mx[d] - mn[d]);
        if (sz > root) root = sz;
    }
    root *= 1.0 + (o->exp_tree / 100.0);
    if (!(root > 0.0)) return 3;  /* every particle at one point */
    for (int d = 0; d < 3; ++d) piv[d] = (mx[d] + mn[d] - root) / 2.0;
    orc3_tree *t = (orc3_tree *)calloc(1, sizeof(orc3_tree));
    if (!t) return 5;
    t->n = n;
    t->o = *o;
    t->lev_start = (int *)malloc(sizeof(int) * (size_t)(o->max_depth + 3));
    t->pool = (int *)malloc(sizeof(int) * (size_t)n);
    int *pcell = (int *)malloc(sizeof(int) * (size_t)n);
    int *ix = (int *)malloc(sizeof(int) * 3 * (size_t)n);
    uint8_t *done = (uint8_t *)calloc((size_t)n, 1);
    int nsrc_all = 0;
    for (int i = 0; i < n; ++i) nsrc_all += active[i] != 0;
    if (!t->lev_start || !t->pool || !pcell || !ix || !done ||
        o3_new_cell(t, 0, -1, 0, piv[0] + 0.5 * root, piv[1] + 0.5 * root, piv[2] + 0.5 * root,
                    root) < 0) {
        free(pcell);
        free(ix);
        free(done);
        orc3_free(t);
        return 5;
    }
    t->npar[0] = n;
    t->nsrc[0] = nsrc_all;
    t->active[0] = nsrc_all != 0;
    t->lev_start[0] = 0;
    int unsettled = n, lev = 0, st = 0;
    while (unsettled > 0) {
        ++lev;
        if (lev > o->max_depth) {
            st = 4;
            break;
        }
        t->lev_start[lev] = t->ncell;
        const double cs = root / pow(2.0, lev);
        o3_hash hsh;
        if (!o3_hash_init(&hsh, (size_t)unsettled)) {
            st = 5;
            break;
        }
        o3_hash hpar;
        const int pb = t->lev_start[lev - 1], pe = t->lev_start[lev];
        if (!o3_hash_init(&hpar, (size_t)(pe - pb))) {
            o3_hash_free(&hsh);
            st = 5;
            break;
        }
        /* the previous level's cells by index, for the parent links */
        for (int c = pb; c < pe; ++c) {
            const uint64_t k = (uint64_t)(uint32_t)ix[3 * (size_t)t->pbeg[c]] |
                               ((uint64_t)(uint32_t)ix[3 * (size_t)t->pbeg[c] + 1] << 21) |
                               ((uint64_t)(uint32_t)ix[3 * (size_t)t->pbeg[c] + 2] << 42);
            if (lev > 1) *o3_hash_slot(&hpar, k) = c;
        }
        for (int i = 0; i < n; ++i) {
            if (done[i]) continue;
            int id[3];
            for (int d = 0; d < 3; ++d) id[d] = (int)((p[d][i] - piv[d]) / cs);
            const uint64_t k = (uint64_t)(uint32_t)id[0] | ((uint64_t)(uint32_t)id[1] << 21) |
                               ((uint64_t)(uint32_t)id[2] << 42);
            int *slot = o3_hash_slot(&hsh, k);
            if (*slot < 0) {
                const double len = cs;
                int par = 0;
                if (lev > 1) {
                    const uint64_t kp = (uint64_t)(uint32_t)(id[0] / 2) |
                                        ((uint64_t)(uint32_t)(id[1] / 2) << 21) |
                                        ((uint64_t)(uint32_t)(id[2] / 2) << 42);
                    par = *o3_hash_slot(&hpar, kp);
                }
                const int oc = (id[0] & 1) | ((id[1] & 1) << 1) | ((id[2] & 1) << 2);
                const int c = par < 0 ? -1
                                      : o3_new_cell(t, lev, par, t->key[par] * 8 + oc,
                                                    piv[0] + ((0.5 + id[0]) * len),
                                                    piv[1] + ((0.5 + id[1]) * len),
                                                    piv[2] + ((0.5 + id[2]) * len), len);
                if (c < 0) {
                    st = par < 0 ? 6 : 5;
                    break;
                }
                t->chd[8 * par + oc] = c;
                t->flag[par] |= 1 << oc;
                t->pbeg[c] = i;  /* a particle of the cell, for its index triple */
                *slot = c;
            }
            const int c = *slot;
            pcell[i] = c;
            for (int d = 0; d < 3; ++d) ix[3 * (size_t)i + d] = id[d];
            t->npar[c]++;
            if (active[i]) t->nsrc[c]++;
        }
        o3_hash_free(&hsh);
        o3_hash_free(&hpar);
        if (st) break;
        for (int c = t->lev_start[lev]; c < t->ncell; ++c) {
            t->leaf[c] = t->nsrc[c] <= o->max_src && t->npar[c] <= o->max_par;
            t->active[c] = t->nsrc[c] != 0;
        }
        for (int i = 0; i < n; ++i) {
            if (done[i] || !t->leaf[pcell[i]]) continue;
            t->pcnt[pcell[i]]++;
            done[i] = 1;
            --unsettled;
        }
    }
    if (!st) {
        t->depth = lev;
        t->lev_start[lev + 1] = t->ncell;
        int off = 0;
        for (int c = 0; c < t->ncell; ++c) {
            t->pbeg[c] = off;
            off += t->pcnt[c];
            t->pcnt[c] = 0;
        }
        for (int i = 0; i < n; ++i) {
            const int c = pcell[i];
            t->pool[t->pbeg[c] + t->pcnt[c]++] = i;
        }
    }
    free(pcell);
    free(ix);
    free(done);
    if (st) {
        orc3_free(t);
        return st;
    }
    *out = t;
    return 0;
}

int orc3_build(orc3_tree **out, int n, const double *x, const double *y, const double *z,
               const uint8_t *active, const orc3_opts *o) {
    *out = NULL;
    if (!o || n < 0 || o->max_depth < 1 || o->max_depth > 21 || (o->type != 1 && o->type != 2) ||
        (n > 0 && (!x || !y || !z || !active)))
        return 1;
    for (int i = 0; i < n; ++i)
        if (!isfinite(x[i]) || !isfinite(y[i]) || !isfinite(z[i])) return 2;
    return o->type == 1 ? o3_build_t1(out, n, x, y, z, active, o)
                        : o3_build_t2(out, n, x, y, z, active, o);
}

int orc3_ncells(const orc3_tree *t) { return t->ncell; }
int orc3_depth(const orc3_tree *t) { return t->depth; }

void orc3_cells(const orc3_tree *t, int *level, int64_t *key, int *count, uint8_t *leaf,
                uint8_t *active, double *centre3, double *size) {
    for (int c = 0; c < t->ncell; ++c) {
        if (level) level[c] = t->level[c];
        if (key) key[c] = t->key[c];
        if (count) count[c] = t->npar[c];
        if (leaf) leaf[c] = t->leaf[c];
        if (active) active[c] = t->active[c];
        if (centre3) {
            centre3[3 * c] = t->xc[c];
            centre3[3 * c + 1] = t->yc[c];
            centre3[3 * c + 2] = t->zc[c];
        }
        if (size) size[c] = t->size[c];
    }
}

int orc3_leaf_particles(const orc3_tree *t, int c, int *out, int cap) {
    if (c < 0 || c >= t->ncell) return -1;
    for (int u = 0; u < t->pcnt[c] && u < cap; ++u) out[u] = t->pool[t->pbeg[c] + u];
    return t->pcnt[c];
}

/* fmm3D::P2M_calc (fmm3D.cpp:84-112) */
void orc3_p2m(double *mp, double src, double dx, double dy, double dz) {
    mp[0] += src;
    mp[1] += src * dx;
    mp[2] += src * dy;
    mp[3] += src * dz;
    mp[4] += src * dx * dx * 0.5;
    mp[5] += src * dy * dy * 0.5;
    mp[6] += src * dz * dz * 0.5;
    mp[7] += src * dx * dy;
    mp[8] += src * dy * dz;
    mp[9] += src * dx * dz;
}

/* fmm3D::M2M_calc (fmm3D.cpp:114-145) */
void orc3_m2m(double *mp, const double *c, double dx, double dy, double dz) {
    mp[0] += c[0];
    mp[1] += c[1] + c[0] * dx;
    mp[2] += c[2] + c[0] * dy;
    mp[3] += c[3] + c[0] * dz;
    mp[4] += c[4] + (c[0] * dx * dx * 0.5) + (c[1] * dx);
    mp[5] += c[5] + (c[0] * dy * dy * 0.5) + (c[2] * dy);
    mp[6] += c[6] + (c[0] * dz * dz * 0.5) + (c[3] * dz);
    mp[7] += c[7] + (c[0] * dx * dy) + (c[1] * dy) + (c[2] * dx);
    mp[8] += c[8] + (c[0] * dy * dz) + (c[2] * dz) + (c[3] * dy);
    mp[9] += c[9] + (c[0] * dx * dz) + (c[1] * dz) + (c[3] * dx);
}

/* fmm3D::M2P_mul_calc (fmm3D.cpp:147-210): the Taylor coefficients of
 * (dx, dy, dz) / R^3 in the moment order of P2M_calc. */
static void o3_m2p_mul(double R2, double dx, double dy, double dz, double *mx, double *my,
                       double *mz) {
    const double R = sqrt(R2), R3 = R * R2, R5 = R3 * R2, R7 = R5 * R2;
    mx[0] = dx / R3;
    mx[1] = -(3 * dx * dx / R5) + (1 / R3);
    mx[2] = -(3 * dx * dy / R5);
    mx[3] = -(3 * dx * dz / R5);
    mx[4] = (15 * dx * dx * dx / R7) - (9 * dx / R5);
    mx[5] = (15 * dx * dy * dy / R7) - (3 * dx / R5);
    mx[6] = (15 * dx * dz * dz / R7) - (3 * dx / R5);
    mx[7] = (15 * dx * dx * dy / R7) - (3 * dy / R5);
    mx[8] = (15 * dx * dy * dz / R7);
    mx[9] = (15 * dx * dx * dz / R7) - (3 * dz / R5);
    my[0] = dy / R3;
    my[1] = -(3 * dy * dx / R5);
    my[2] = -(3 * dy * dy / R5) + (1 / R3);
    my[3] = -(3 * dy * dz / R5);
    my[4] = (15 * dy * dx * dx / R7) - (3 * dy / R5);
    my[5] = (15 * dy * dy * dy / R7) - (9 * dy / R5);
    my[6] = (15 * dy * dz * dz / R7) - (3 * dy / R5);
    my[7] = (15 * dy * dx * dy / R7) - (3 * dx / R5);
    my[8] = (15 * dy * dy * dz / R7) - (3 * dz / R5);
    my[9] = (15 * dy * dx * dz / R7);
    mz[0] = dz / R3;
    mz[1] = -(3 * dz * dx / R5);
    mz[2] = -(3 * dz * dy / R5);
    mz[3] = -(3 * dz * dz / R5) + (1 / R3);
    mz[4] = (15 * dz * dx * dx / R7) - (3 * dz / R5);
    mz[5] = (15 * dz * dy * dy / R7) - (3 * dz / R5);
    mz[6] = (15 * dz * dz * dz / R7) - (9 * dz / R5);
    mz[7] = (15 * dz * dx * dy / R7);
    mz[8] = (15 * dz * dy * dz / R7) - (3 * dy / R5);
    mz[9] = (15 * dz * dx * dz / R7) - (3 * dx / R5);
}

void orc3_m2p_mul(double dx, double dy, double dz, double *mx, double *my, double *mz) {
    o3_m2p_mul(dx * dx + dy * dy + dz * dz, dx, dy, dz, mx, my, mz);
}

/* One direct pair (evaluateFMM, fast3DFMM.cpp:437-458). */
static inline void o3_pair(double xp, double yp, double zp, const double *x, const double *y,
                           const double *z, const double *ax, const double *ay,
                           const double *az, int j, double *u, double *v, double *w) {
    const double dx = xp - x[j], dy = yp - y[j], dz = zp - z[j];
    const double R2 = dx * dx + dy * dy + dz * dz;
    const double R = sqrt(R2), R3 = R * R2;
    *u += (ay[j] * dz - az[j] * dy) / R3;
    *v += (az[j] * dx - ax[j] * dz) / R3;
    *w += (ax[j] * dy - ay[j] * dx) / R3;
}

int orc3_velocity_alpha(int n, const double *x, const double *y, const double *z,
                        const double *ax, const double *ay, const double *az,
                        const uint8_t *active, const orc3_opts *o, double *u, double *v,
                        double *w) {
    if (n > 0 && (!ax || !ay || !az || !u || !v || !w)) return 1;
    orc3_tree *t = NULL;
    const int st = orc3_build(&t, n, x, y, z, active, o);
    if (st) return st;
    const int nc = t->ncell;
    double *m = (double *)calloc((size_t)nc * 3 * O3_TERMS + 1, sizeof(double));
    if (!m) {
        orc3_free(t);
        return 5;
    }
    const double *al[3] = {ax, ay, az};
    /* P2M at active leaves over their active particles (:187-219) */
    for (int c = 1; c < nc; ++c) {
        if (!t->leaf[c] || !t->active[c]) continue;
        for (int q = 0; q < t->pcnt[c]; ++q) {
            const int i = t->pool[t->pbeg[c] + q];
            if (!active[i]) continue;
            const double dx = t->xc[c] - x[i], dy = t->yc[c] - y[i], dz = t->zc[c] - z[i];
            for (int k = 0; k < 3; ++k) orc3_p2m(m + ((size_t)c * 3 + k) * O3_TERMS, al[k][i], dx, dy, dz);
        }
    }
    /* M2M, levels maxLevel-1 .. 1 (:245-287) */
    for (int L = t->depth - 1; L > 0; --L)
        for (int c = t->lev_start[L]; c < t->lev_start[L + 1]; ++c) {
            if (!t->active[c] || t->leaf[c]) continue;
            for (int oc = 0; oc < 8; ++oc) {
                if (((t->flag[c] >> oc) % 2) == 0) continue;
                const int ch = t->chd[8 * c + oc];
                const double dx = t->xc[c] - t->xc[ch], dy = t->yc[c] - t->yc[ch],
                             dz = t->zc[c] - t->zc[ch];
                for (int k = 0; k < 3; ++k)
                    orc3_m2m(m + ((size_t)c * 3 + k) * O3_TERMS, m + ((size_t)ch * 3 + k) * O3_TERMS,
                             dx, dy, dz);
            }
        }
    /* evaluation (type 1: evaluateFMM, fast3DFMM.cpp:318-484; type 2:
     * fmm3D::calcVelocity with VELOCITY_CALCULATION_TYPE 2, fmm3D.cpp:
     * 1136-1196): from the children of the level-1 cells, a cell is far
     * (M2P) beyond support * size, else opened (non-leaf) or summed directly
     * (leaf; the target itself skipped; type 1 sums its active particles,
     * type 2 all of them); inactive cells are skipped with their subtrees.
     * Type 2 asks the level-1 cells for their children with get_childID,
     * which throws for a leaf (treeCell.cpp:384-388): that is status 6. */
    const int depth = t->depth;
    const double sup = t->o.support;
    const int t2 = t->o.type == 2;
    if (t2 && depth >= 2)
        for (int oc = 0; oc < 8; ++oc)
            if (((t->flag[0] >> oc) % 2) && t->leaf[t->chd[oc]]) {
                free(m);
                orc3_free(t);
                return 6;
            }
#pragma omp parallel
    {
        int *stk = (int *)malloc(sizeof(int) * (size_t)(8 * (depth + 2) + 16));
#pragma omp for schedule(dynamic, 256)
        for (int i = 0; i < n; ++i) {
            double vu = 0.0, vv = 0.0, vw = 0.0;
            const double xp = x[i], yp = y[i], zp = z[i];
            int sp = 0;
            if (stk && depth >= 2) {
                /* level-1 cells in reverse octant order, so the stack pops
                 * them (and their children) in octant order */
                for (int oc = 7; oc >= 0; --oc)
                    if ((t->flag[0] >> oc) % 2) stk[sp++] = t->chd[oc];
            }
            while (sp > 0) {
                const int c = stk[--sp];
                if (t->level[c] >= depth) continue;
                /* the children of c, each tested */
                int open[8], no = 0;
                for (int oc = 0; oc < 8; ++oc) {
                    if (((t->flag[c] >> oc) % 2) == 0) continue;
                    const int ch = t->chd[8 * c + oc];
                    if (!t->active[ch]) continue;
                    const double dx = xp - t->xc[ch], dy = yp - t->yc[ch], dz = zp - t->zc[ch];
                    const double R2 = dx * dx + dy * dy + dz * dz;
                    const double Rc = t->size[ch] * sup;
                    if (R2 > Rc * Rc) {
                        double mx[O3_TERMS], my[O3_TERMS], mz[O3_TERMS];
                        o3_m2p_mul(R2, dx, dy, dz, mx, my, mz);
                        const double *Vx = m + ((size_t)ch * 3 + 0) * O3_TERMS;
                        const double *Vy = m + ((size_t)ch * 3 + 1) * O3_TERMS;
                        const double *Vz = m + ((size_t)ch * 3 + 2) * O3_TERMS;
                        for (int k = 0; k < O3_TERMS; ++k) {
                            vu += Vy[k] * mz[k] - Vz[k] * my[k];
                            vv += Vz[k] * mx[k] - Vx[k] * mz[k];
                            vw += Vx[k] * my[k] - Vy[k] * mx[k];
                        }
                    } else if (!t->leaf[ch]) {
                        open[no++] = ch;
                    } else {
                        for (int q = 0; q < t->pcnt[ch]; ++q) {
                            const int j = t->pool[t->pbeg[ch] + q];
                            if ((!t2 && !active[j]) || j == i) continue;
                            o3_pair(xp, yp, zp, x, y, z, ax, ay, az, j, &vu, &vv, &vw);
                        }
                    }
                }
                while (no > 0) stk[sp++] = open[--no];
            }
            u[i] = vu;
            v[i] = vv;
            w[i] = vw;
        }
        free(stk);
    }
    free(m);
    orc3_free(t);
    return 0;
}

int orc3_velocity(int n, const double *x, const double *y, const double *z, const double *s,
                  const double *wx, const double *wy, const double *wz, const uint8_t *active,
                  const orc3_opts *o, double u_inf, double v_inf, double w_inf, double *u,
                  double *v, double *w) {
    if (n > 0 && (!s || !wx || !wy || !wz)) return 1;
    double *a = (double *)calloc(3 * (size_t)(n > 0 ? n : 1), sizeof(double));
    if (!a) return 5;
    for (int i = 0; i < n; ++i) {
        const double s3 = pow(s[i], 3.0);
        a[i] = wx[i] * s3 / (4 * M_PI);
        a[n + i] = wy[i] * s3 / (4 * M_PI);
        a[2 * n + i] = wz[i] * s3 / (4 * M_PI);
    }
    const int st = orc3_velocity_alpha(n, x, y, z, a, a + n, a + 2 * n, active, o, u, v, w);
    free(a);
    if (st) return st;
    for (int i = 0; i < n; ++i) {
        u[i] += u_inf;
        v[i] += v_inf;
        w[i] += w_inf;
    }
    return 0;
}

void orc3_direct(int n, const double *x, const double *y, const double *z, const double *ax,
                 const double *ay, const double *az, const uint8_t *active, int nt,
                 const int *tidx, double *u, double *v, double *w) {
    const int m = tidx ? nt : n;
#pragma omp parallel for schedule(static)
    for (int a = 0; a < m; ++a) {
        const int i = tidx ? tidx[a] : a;
        double vu = 0.0, vv = 0.0, vw = 0.0;
        for (int j = 0; j < n; ++j) {
            if (j == i || (active && !active[j])) continue;
            o3_pair(x[i], y[i], z[i], x, y, z, ax, ay, az, j, &vu, &vv, &vw);
        }
        u[a] = vu;
        v[a] = vv;
        w[a] = vw;
    }
}

void orc_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
}
