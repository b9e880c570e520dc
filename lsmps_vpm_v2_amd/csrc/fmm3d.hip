// fmm3d.hip -- the 3-D vector treecodes of the reference on gfx950, FP64:
// the default 3-D path (VELOCITY_3D_TYPE 1, src/FMM/fast3DFMM.cpp with the
// fmm3D.cpp helpers) and the second one (VELOCITY_3D_TYPE 2,
// fmm3D::calcVelocity over the 3-D TreeCell, fmm3D.cpp:585-700, 1064-1205;
// treeCell.cpp:1043-1243), which shares the walk and the moments and differs
// in the octree (root box, cell index rule, leaf rule) and in summing every
// particle of a near leaf, inactive ones included.  Type 1 below unless said.
//
// Reference semantics kept exactly:
//   * octree (FMMCell::generateTree, fast3DFMM.cpp:485-745): root box = the
//     particles' box together with the domain centre, side = the longest
//     extent times (1 + expTree / 100), centre = the box midpoint; a particle
//     goes to octant (x > xc) + 2 (y > yc) + 4 (z > zc) of its cell; a child's
//     centre is the parent's plus (size / 2) (bit - 1/2); a particle settles
//     at the first level where its cell holds fewer than critNum particles;
//     a cell is active when a particle of its subtree is active.  The keys
//     below replay these comparisons and additions per particle, so every
//     cell, centre and size is the reference's, bit for bit.
//   * moments: the 10 order-2 Cartesian moments of each vorticity component
//     about the cell centre, d = centre - particle, over active particles
//     (P2M_calc, fmm3D.cpp:84-112).  The reference forms them at the leaves
//     and shifts them upward (M2M_calc :114-145); the order-2 shift is exact,
//     so here each cell shifts its leaves' moments directly (equal up to
//     rounding, one launch instead of one per level).
//   * evaluation (evaluateFMM, fast3DFMM.cpp:318-484): every particle is a
//     target; the traversal starts at the children of the level-1 cells (a
//     level-1 leaf's sources are never summed); an inactive cell is skipped
//     with its subtree; a cell is far when |r|^2 > (2 size)^2 (the same
//     rounded comparison) and then adds its M2P term (M2P_mul_calc,
//     fmm3D.cpp:147-210); a near non-leaf cell is opened; a near leaf adds
//     the direct Biot-Savart sum over its active sources, the target itself
//     excluded.
//
// Mapping: cells are emitted in depth-first pre-order with a skip pointer
// (the next cell after the subtree), so a traversal is a loop over one index.
// One wavefront takes 64 consecutive sorted targets and walks the tree once
// for all of them: a cell is visited when some lane has to test it, and each
// lane keeps one bit per level saying whether it opened the cell of that
// level on the current path (a lane tests a cell only when its bit of the
// parent's level is set).  Cell records, moments and source particles are
// wave-uniform loads (scalar cache).
//
// This file is compiled without FMA contraction: cell centres and the
// opening test round as the reference's; the arithmetic of the sums opts
// back in (contract(fast) blocks).
#include <cmath>

#include "fmm3d.hpp"
#include "prim.hpp"

namespace vpm {
namespace {

constexpr int kHalf = kFmm3Depth / 2;  // levels per key word
constexpr unsigned kF3NonFinite = 1u, kF3Deeper = 2u, kF3Depth = 4u, kF3Degenerate = 8u;
constexpr int kMomStride = 32;
constexpr int kEvalBlock = 256;

struct __align__(16) Par3 {
    double x, y, z, ax, ay, az, act, pad;
};

// meta: level (bits 0-7), leaf (bit 8), active (bit 9); [begin, end): the
// cell's particles in tree order, [abegin, aend): its active ones in the
// compacted source array
struct __align__(16) Cell3 {
    double cx, cy, cz, rc2;
    uint32_t begin, end, skip, meta;
    uint32_t abegin, aend, pad0, pad1;
};

// An active source in tree order: position, tree index (for the self test),
// strength.
struct __align__(16) Src3 {
    double x, y, z, idx, ax, ay, az, pad;
};

__device__ __forceinline__ uint64_t pfx(uint64_t k, int L) { return k >> (3 * (kFmm3Depth - L)); }

// -------------------------------------------------------------------------
// bounding box and root
// -------------------------------------------------------------------------
__device__ __forceinline__ double dmin(double a, double b) { return b < a ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return b > a ? b : a; }

__global__ void __launch_bounds__(256)
    k3_bbox(const double *__restrict__ x, const double *__restrict__ y,
            const double *__restrict__ z, int n, double *__restrict__ part) {
    __shared__ double sm[4][7];
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    double bad = 0.0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (long long)gridDim.x * 256) {
        const double p[3] = {x[i], y[i], z[i]};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (!isfinite(p[d])) {
                bad = 1.0;
                continue;
            }
            mn[d] = dmin(mn[d], p[d]);
            mx[d] = dmax(mx[d], p[d]);
        }
    }
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            mn[d] = dmin(mn[d], __shfl_xor(mn[d], m, kWave));
            mx[d] = dmax(mx[d], __shfl_xor(mx[d], m, kWave));
        }
        bad = dmax(bad, __shfl_xor(bad, m, kWave));
    }
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        for (int d = 0; d < 3; ++d) {
            sm[w][d] = mn[d];
            sm[w][3 + d] = mx[d];
        }
        sm[w][6] = bad;
    }
    __syncthreads();
    if (threadIdx.x < 7) {
        double r = sm[0][threadIdx.x];
        for (int q = 1; q < 4; ++q) {
            const double o = sm[q][threadIdx.x];
            r = (threadIdx.x < 3) ? dmin(r, o) : dmax(r, o);
        }
        part[(size_t)blockIdx.x * 8 + threadIdx.x] = r;
    }
}

// The root.  Type 1, FMMCell::generateTree (fast3DFMM.cpp:507-531): the
// box starts at the domain centre, side = max extent * (1 + expTree / 100),
// centre = the box midpoint.  Type 2, TreeCell::initializeTree
// (treeCell.cpp:1066-1091): the box of the particles alone, side = the
// longest extent * (1 + expTree / 100), pivot (lower corner) = (max + min -
// side) / 2.  Also clears the header's counters.
__global__ void __launch_bounds__(256)
    k3_root(const double *__restrict__ part, int nb, double c0, double c1, double c2,
            double exp_tree, int type, Fmm3Hdr *__restrict__ h) {
    __shared__ double sm[4][7];
    const bool t2 = type == 2;
    double mn[3] = {c0, c1, c2}, mx[3] = {c0, c1, c2}, bad = 0.0;
    if (t2)
        for (int d = 0; d < 3; ++d) {
            mn[d] = INFINITY;
            mx[d] = -INFINITY;
        }
    for (int b = threadIdx.x; b < nb; b += 256) {
        for (int d = 0; d < 3; ++d) {
            mn[d] = dmin(mn[d], part[(size_t)b * 8 + d]);
            mx[d] = dmax(mx[d], part[(size_t)b * 8 + 3 + d]);
        }
        bad = dmax(bad, part[(size_t)b * 8 + 6]);
    }
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            mn[d] = dmin(mn[d], __shfl_xor(mn[d], m, kWave));
            mx[d] = dmax(mx[d], __shfl_xor(mx[d], m, kWave));
        }
        bad = dmax(bad, __shfl_xor(bad, m, kWave));
    }
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        for (int d = 0; d < 3; ++d) {
            sm[w][d] = mn[d];
            sm[w][3 + d] = mx[d];
        }
        sm[w][6] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < 4; ++q) {
            for (int d = 0; d < 3; ++d) {
                sm[0][d] = dmin(sm[0][d], sm[q][d]);
                sm[0][3 + d] = dmax(sm[0][3 + d], sm[q][3 + d]);
            }
            sm[0][6] = dmax(sm[0][6], sm[q][6]);
        }
        double root = 0.0;
        unsigned f = sm[0][6] != 0.0 ? kF3NonFinite : 0u;
        if (!t2) {
            for (int d = 0; d < 3; ++d) {
                const double len = sm[0][3 + d] - sm[0][d];
                root = (len < root) ? root : len;  // std::max(len, root)
                h->root_c[d] = 0.5 * (sm[0][3 + d] + sm[0][d]);
                h->pivot[d] = 0.0;
            }
            root *= (1 + (exp_tree / 100));
        } else {
            for (int d = 0; d < 3; ++d) {
                const double side = fabs(sm[0][3 + d] - sm[0][d]);
                if (side > root) root = side;
            }
            root *= 1.0 + (exp_tree / 100.0);
            for (int d = 0; d < 3; ++d) {
                h->pivot[d] = (sm[0][3 + d] + sm[0][d] - root) / 2.0;
                h->root_c[d] = h->pivot[d] + (0.5 * root);  // the root cell's centre
            }
            if (!(root > 0.0)) f |= kF3Degenerate;
        }
        h->root_len = root;
        h->flags = f;
        h->min_settle = 0;
        h->depth = 0;
        h->n_leaves = 0;
        h->pairs = 0;
        h->m2p = 0;
        h->wave_pairs = 0;
        h->wave_m2p = 0;
    }
}

// -------------------------------------------------------------------------
// keys, sort, gather
// -------------------------------------------------------------------------
// The particle's octant path, level by level from the root (the octant from
// the parent centre, then the child centre from the parent's).  hi holds
// levels 1..10, lo levels 11..20; sk receives the word the first sort pass
// orders (hi for a 10-level sort, lo for the first of a 20-level one).
//
// Type 2 (TreeCell): the level-L index is (int)((x - pivot) / (side / 2^L))
// (coord2index, treeCell.cpp:290-298).  side / 2^L is exact and scaling by a
// power of two commutes with rounding, so that index is the top L bits of
// t = (x - pivot) / side taken to 20 binary places: one division per axis,
// and every level's index nests in the next.
__global__ void __launch_bounds__(256)
    k3_keys(const double *__restrict__ x, const double *__restrict__ y,
            const double *__restrict__ z, int n, const Fmm3Hdr *__restrict__ h,
            uint32_t *__restrict__ khi, uint32_t *__restrict__ klo, uint32_t *__restrict__ sk,
            uint32_t *__restrict__ sv, int deep, int type) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double px = x[i], py = y[i], pz = z[i];
    uint32_t hi = 0, lo = 0;
    if (type == 2) {
        const double R = h->root_len;
        uint32_t id[3];
        const double p[3] = {px, py, pz};
        for (int d = 0; d < 3; ++d) {
            const double t = (p[d] - h->pivot[d]) / R;
            const double sc = t * (double)(1u << kFmm3Depth);
            id[d] = sc >= 0.0 && sc < (double)(1u << kFmm3Depth) ? (uint32_t)sc : 0u;
        }
        for (int L = 1; L <= kFmm3Depth; ++L) {
            const int sh = kFmm3Depth - L;
            const uint32_t o = ((id[0] >> sh) & 1u) | (((id[1] >> sh) & 1u) << 1) |
                               (((id[2] >> sh) & 1u) << 2);
            if (L <= kHalf)
                hi = (hi << 3) | o;
            else
                lo = (lo << 3) | o;
        }
    } else {
        double cx = h->root_c[0], cy = h->root_c[1], cz = h->root_c[2], sz = h->root_len;
        for (int L = 1; L <= kFmm3Depth; ++L) {
            const uint32_t o = (px > cx ? 1u : 0u) | (py > cy ? 2u : 0u) | (pz > cz ? 4u : 0u);
            if (L <= kHalf)
                hi = (hi << 3) | o;
            else
                lo = (lo << 3) | o;
            const double cl = sz / 2.0;
            cx = cx + cl * ((double)(o & 1u) - 0.5);
            cy = cy + cl * ((double)((o >> 1) & 1u) - 0.5);
            cz = cz + cl * ((double)((o >> 2) & 1u) - 0.5);
            sz = cl;
        }
    }
    khi[i] = hi;
    klo[i] = lo;
    sk[i] = deep ? lo : hi;
    sv[i] = (uint32_t)i;
}

// Second pass of a 20-level sort: the high words in the order of the first.
__global__ void __launch_bounds__(256)
    k3_regather(const uint32_t *__restrict__ perm, const uint32_t *__restrict__ khi, int n,
                uint32_t *__restrict__ sk, uint32_t *__restrict__ sv) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const uint32_t i = perm[j];
    sk[j] = khi[i];
    sv[j] = i;
}

// Sorted particle records, 60-bit keys, active flags.  The vortex strength
// is formed as velocity_fmm_3d_fast does (velocity_calc.cpp:2429-2431):
// omega * s^3 / (4 pi); without sizes (s null) (wx, wy, wz) are the
// strengths.
__global__ void __launch_bounds__(256)
    k3_gather(const uint32_t *__restrict__ perm, const uint32_t *__restrict__ khi,
              const uint32_t *__restrict__ klo, const double *__restrict__ x,
              const double *__restrict__ y, const double *__restrict__ z,
              const double *__restrict__ s, const double *__restrict__ wx,
              const double *__restrict__ wy, const double *__restrict__ wz,
              const uint8_t *__restrict__ active, int n, Par3 *__restrict__ P,
              uint64_t *__restrict__ key, uint8_t *__restrict__ act8) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const uint32_t i = perm[j];
    key[j] = ((uint64_t)khi[i] << 30) | (uint64_t)klo[i];
    const bool a = active[i] != 0;
    Par3 r;
    r.x = x[i];
    r.y = y[i];
    r.z = z[i];
    if (s) {
        const double s3 = pow(s[i], 3.0);
        const double k4pi = 4 * M_PI;
        r.ax = wx[i] * s3 / k4pi;
        r.ay = wy[i] * s3 / k4pi;
        r.az = wz[i] * s3 / k4pi;
    } else {  // the strengths themselves (calcVelocityFast's alpha)
        r.ax = wx[i];
        r.ay = wy[i];
        r.az = wz[i];
    }
    r.act = a ? 1.0 : 0.0;
    r.pad = 0.0;
    P[j] = r;
    act8[j] = a ? 1 : 0;
}

// -------------------------------------------------------------------------
// leaf levels and cell emission
// -------------------------------------------------------------------------
__device__ __forceinline__ int common_levels(uint64_t a, uint64_t b, int D) {
    const uint64_t d = a ^ b;
    const int m = d ? (kFmm3Depth - 1) - (63 - __clzll((long long)d)) / 3 : D;
    return min(m, D);
}

// Per-block depth, leaf count and shallowest leaf level (k3_settle_stats
// combines them); the rare unsettled flag is the only atomic.
__device__ __forceinline__ void settle_block_stats(bool valid, int st, int mp, bool unsettled,
                                                   int D, uint4 *__restrict__ bstat,
                                                   Fmm3Hdr *__restrict__ h) {
    __shared__ uint32_t s_d[4], s_l[4], s_m[4];
    int dmax_ = st, dmin_ = valid && st > 0 ? st : 255;
    for (int m = 1; m < kWave; m <<= 1) {
        dmax_ = max(dmax_, __shfl_xor(dmax_, m, kWave));
        dmin_ = min(dmin_, __shfl_xor(dmin_, m, kWave));
    }
    const uint64_t leaves = __ballot(valid && st > mp);
    const bool any_unsettled = __ballot(unsettled) != 0;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        s_d[threadIdx.x / kWave] = (uint32_t)dmax_;
        s_l[threadIdx.x / kWave] = (uint32_t)__popcll(leaves);
        s_m[threadIdx.x / kWave] = (uint32_t)dmin_;
        if (any_unsettled) atomicOr(&h->flags, D < kFmm3Depth ? kF3Deeper : kF3Depth);
    }
    __syncthreads();
    if (threadIdx.x == 0)
        bstat[blockIdx.x] = make_uint4(max(max(s_d[0], s_d[1]), max(s_d[2], s_d[3])),
                                       s_l[0] + s_l[1] + s_l[2] + s_l[3],
                                       min(min(s_m[0], s_m[1]), min(s_m[2], s_m[3])), 0u);
}

// lcpw[j] (j <= n - crit): levels shared by the ends of the window of crit
// sorted particles starting at j.  The cell of particle i at level L holds
// crit particles or more exactly when some window containing i has
// lcpw >= L (the keys are sorted, so a window's ends bound it).
__global__ void __launch_bounds__(256)
    k3_lcpw(const uint64_t *__restrict__ key, int n, int D, int crit,
            uint8_t *__restrict__ lcpw) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j > n - crit) return;
    lcpw[j] = (uint8_t)common_levels(key[j], key[j + crit - 1], D);
}

// settle[i]: the first level whose cell around particle i holds fewer than
// crit particles (1 + the largest lcpw of the windows containing i);
// mprev[i]: levels shared with particle i - 1; starts[i]: the cells particle
// i opens in pre-order (levels mprev + 1 .. settle).
__global__ void __launch_bounds__(256)
    k3_settle(const uint64_t *__restrict__ key, const uint8_t *__restrict__ lcpw, int n, int D,
              int crit, uint8_t *__restrict__ settle, uint8_t *__restrict__ mprev,
              uint32_t *__restrict__ starts, uint4 *__restrict__ bstat,
              Fmm3Hdr *__restrict__ h) {
    extern __shared__ uint8_t s_w[];
    const int i0 = blockIdx.x * 256;
    const int i = i0 + threadIdx.x;
    // windows j in [i0 - crit + 1, i0 + 255] clipped to [0, n - crit]
    const int j_lo = max(0, i0 - crit + 1), j_hi = min(i0 + 255, n - crit);
    for (int j = j_lo + (int)threadIdx.x; j <= j_hi; j += 256) s_w[j - j_lo] = lcpw[j];
    __syncthreads();
    int st = 0, mp = 0;
    bool unsettled = false;
    if (i < n) {
        int mx = 0;
        const int a = max(j_lo, i - crit + 1), b = min(i, j_hi);
        for (int j = a; j <= b; ++j) mx = max(mx, (int)s_w[j - j_lo]);
        unsettled = mx >= D;
        st = unsettled ? 0 : mx + 1;
        if (i > 0) mp = common_levels(key[i - 1], key[i], D);
        settle[i] = (uint8_t)st;
        mprev[i] = (uint8_t)mp;
        starts[i] = st > mp ? (uint32_t)(st - mp) : 0u;
    }
    settle_block_stats(i < n, st, mp, unsettled, D, bstat, h);
}

// Type 2 (TreeCell): the first level whose cell around particle i holds at
// most max_par particles and max_src active ones (the cell's range from two
// binary searches inside the parent's range; its active count from the
// active scan A).
__global__ void __launch_bounds__(256)
    k3_settle_t2(const uint64_t *__restrict__ key, const uint32_t *__restrict__ A, int n, int D,
                 int max_par, int max_src, uint8_t *__restrict__ settle,
                 uint8_t *__restrict__ mprev, uint32_t *__restrict__ starts,
                 uint4 *__restrict__ bstat, Fmm3Hdr *__restrict__ h) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    int st = 0, mp = 0;
    bool unsettled = false;
    if (i < n) {
        const uint64_t ki = key[i];
        int lo = 0, hi = n;
        for (int L = 1; L <= D; ++L) {
            const uint64_t p = pfx(ki, L);
            int a = lo, b = i;
            while (a < b) {
                const int m = (a + b) >> 1;
                if (pfx(key[m], L) < p) a = m + 1;
                else b = m;
            }
            int e0 = i + 1, e1 = hi;
            while (e0 < e1) {
                const int m = (e0 + e1) >> 1;
                if (pfx(key[m], L) > p) e1 = m;
                else e0 = m + 1;
            }
            if (e0 - a <= max_par && (int)(A[e0] - A[a]) <= max_src) {
                st = L;
                break;
            }
            lo = a;
            hi = e0;
        }
        if (i > 0) mp = common_levels(key[i - 1], ki, D);
        unsettled = st == 0;
        settle[i] = (uint8_t)st;
        mprev[i] = (uint8_t)mp;
        starts[i] = st > mp ? (uint32_t)(st - mp) : 0u;
    }
    settle_block_stats(i < n, st, mp, unsettled, D, bstat, h);
}

__global__ void __launch_bounds__(1024)
    k3_settle_stats(const uint4 *__restrict__ bstat, int nb, Fmm3Hdr *__restrict__ h) {
    __shared__ uint32_t sd[16], sl[16], sm[16];
    uint32_t d = 0, l = 0, mn = 255;
    for (int q = threadIdx.x; q < nb; q += 1024) {
        d = max(d, bstat[q].x);
        l += bstat[q].y;
        mn = min(mn, bstat[q].z);
    }
    for (int m = 1; m < kWave; m <<= 1) {
        d = max(d, (uint32_t)__shfl_xor((int)d, m, kWave));
        l += (uint32_t)__shfl_xor((int)l, m, kWave);
        mn = min(mn, (uint32_t)__shfl_xor((int)mn, m, kWave));
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
        sd[threadIdx.x / kWave] = d;
        sl[threadIdx.x / kWave] = l;
        sm[threadIdx.x / kWave] = mn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t td = 0, tl = 0, tm = 255;
        for (int q = 0; q < 16; ++q) {
            td = max(td, sd[q]);
            tl += sl[q];
            tm = min(tm, sm[q]);
        }
        h->depth = td;
        h->n_leaves = tl;
        h->min_settle = tm;
    }
}

// Cells of particle i (levels mprev + 1 .. settle) at pre-order positions
// pos[i] ...: centre, (support size)^2, particle range, skip pointer (the
// first cell of the first particle past the range), level, leaf, active (an
// active particle in the range).  The centre is replayed from the root
// (type 1: parent + (size / 2)(bit - 1/2)) or formed from the level's index
// (type 2: pivot + (0.5 + index) size, Cell::Cell, treeCell.hpp:68).  The
// direct-sum sources of a cell are its active particles (type 1) or all of
// them (type 2, fmm3D.cpp:1172-1186 reads no active mark).
__global__ void __launch_bounds__(256)
    k3_emit(const uint64_t *__restrict__ key, const uint8_t *__restrict__ settle,
            const uint8_t *__restrict__ mprev, const uint32_t *__restrict__ pos,
            const uint32_t *__restrict__ A, int n, uint32_t ncell,
            const Fmm3Hdr *__restrict__ h, double support, int type,
            Cell3 *__restrict__ cells) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int st = settle[i], mp = mprev[i];
    if (st <= mp) return;
    const bool t2 = type == 2;
    const uint64_t ki = key[i];
    double cx = h->root_c[0], cy = h->root_c[1], cz = h->root_c[2], sz = h->root_len;
    uint32_t ix = 0, iy = 0, iz = 0;
    uint32_t p = pos[i];
    int hi = n;
    for (int L = 1; L <= st; ++L) {
        const uint32_t o = (uint32_t)(pfx(ki, L) & 7u);
        const double cl = sz / 2.0;
        if (t2) {
            ix = 2u * ix + (o & 1u);
            iy = 2u * iy + ((o >> 1) & 1u);
            iz = 2u * iz + ((o >> 2) & 1u);
            cx = h->pivot[0] + ((0.5 + (double)ix) * cl);
            cy = h->pivot[1] + ((0.5 + (double)iy) * cl);
            cz = h->pivot[2] + ((0.5 + (double)iz) * cl);
        } else {
            cx = cx + cl * ((double)(o & 1u) - 0.5);
            cy = cy + cl * ((double)((o >> 1) & 1u) - 0.5);
            cz = cz + cl * ((double)((o >> 2) & 1u) - 0.5);
        }
        sz = cl;
        if (L <= mp) continue;
        const uint64_t pf = pfx(ki, L);
        int e0 = i + 1, e1 = hi;
        while (e0 < e1) {
            const int m = (e0 + e1) >> 1;
            if (pfx(key[m], L) > pf) e1 = m;
            else e0 = m + 1;
        }
        hi = e0;
        Cell3 c;
        c.cx = cx;
        c.cy = cy;
        c.cz = cz;
        const double rck = sz * support;  // R_check (fast3DFMM.cpp:376)
        c.rc2 = rck * rck;
        c.begin = (uint32_t)i;
        c.end = (uint32_t)e0;
        c.skip = e0 < n ? pos[e0] : ncell;
        c.abegin = t2 ? (uint32_t)i : A[i];
        c.aend = t2 ? (uint32_t)e0 : A[e0];
        c.pad0 = c.pad1 = 0u;
        const bool act = A[e0] > A[i];
        c.meta = (uint32_t)L | (L == st ? 0x100u : 0u) | (act ? 0x200u : 0u);
        if (p < ncell) cells[p] = c;
        ++p;
    }
}

// The direct-sum sources in tree order: the active particles (A: exclusive
// scan of the flags) or, with all, every particle; pad holds the active
// flag (P2M reads active particles only).
__global__ void __launch_bounds__(256)
    k3_compact(const Par3 *__restrict__ P, const uint32_t *__restrict__ A, int n, int all,
               Src3 *__restrict__ S) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n || (!all && A[j + 1] == A[j])) return;
    const Par3 q = P[j];
    Src3 r;
    r.x = q.x;
    r.y = q.y;
    r.z = q.z;
    r.idx = (double)j;
    r.ax = q.ax;
    r.ay = q.ay;
    r.az = q.az;
    r.pad = q.act;
    S[all ? (uint32_t)j : A[j]] = r;
}

// -------------------------------------------------------------------------
// moments
// -------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

// One thread per leaf of level >= 2 (the level-1 cells are never tested):
// P2M of its active particles.
__global__ void __launch_bounds__(256)
    k3_p2m(const Cell3 *__restrict__ cells, int ncell, const Src3 *__restrict__ S,
           double *__restrict__ mom) {
#pragma clang fp contract(fast)
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= ncell) return;
    const Cell3 C = cells[c];
    const int L = (int)(C.meta & 0xffu);
    if (L < 2 || (C.meta & 0x300u) != 0x300u) return;
    double m[30];
#pragma unroll
    for (int k = 0; k < 30; ++k) m[k] = 0.0;
    for (uint32_t j = C.abegin; j < C.aend; ++j) {
        const Src3 q = S[j];
        if (q.pad == 0.0) continue;
        const double dx = C.cx - q.x, dy = C.cy - q.y, dz = C.cz - q.z;
        const double t[10] = {1.0,          dx,           dy,           dz,      dx * dx * 0.5,
                              dy * dy * 0.5, dz * dz * 0.5, dx * dy,      dy * dz, dx * dz};
        const double a[3] = {q.ax, q.ay, q.az};
#pragma unroll
        for (int cc = 0; cc < 3; ++cc)
#pragma unroll
            for (int k = 0; k < 10; ++k) m[cc * 10 + k] += a[cc] * t[k];
    }
    double *o = mom + (size_t)c * kMomStride;
#pragma unroll
    for (int k = 0; k < 30; k += 2) *reinterpret_cast<double2 *>(o + k) = make_double2(m[k], m[k + 1]);
}

// One wavefront per non-leaf cell of level >= 2: its leaves' moments shifted
// to its centre (the order-2 shift of M2M_calc, applied from each leaf).
__global__ void __launch_bounds__(256)
    k3_m2m(const Cell3 *__restrict__ cells, int ncell, double *__restrict__ mom) {
#pragma clang fp contract(fast)
    const int lane = threadIdx.x & (kWave - 1);
    const int c = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4 + threadIdx.x / kWave));
    if (c >= ncell) return;
    const Cell3 C = cells[c];
    const int L = (int)(C.meta & 0xffu);
    if (L < 2 || (C.meta & 0x300u) != 0x200u) return;
    double m[30];
#pragma unroll
    for (int k = 0; k < 30; ++k) m[k] = 0.0;
    for (uint32_t d = (uint32_t)c + 1 + lane; d < C.skip; d += kWave) {
        const Cell3 D = cells[d];
        if ((D.meta & 0x300u) != 0x300u) continue;
        const double dx = C.cx - D.cx, dy = C.cy - D.cy, dz = C.cz - D.cz;
        const double *mc = mom + (size_t)d * kMomStride;
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            const double *q = mc + cc * 10;
            double *o = m + cc * 10;
            const double q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
            o[0] += q0;
            o[1] += q1 + q0 * dx;
            o[2] += q2 + q0 * dy;
            o[3] += q3 + q0 * dz;
            o[4] += q[4] + q0 * dx * dx * 0.5 + q1 * dx;
            o[5] += q[5] + q0 * dy * dy * 0.5 + q2 * dy;
            o[6] += q[6] + q0 * dz * dz * 0.5 + q3 * dz;
            o[7] += q[7] + q0 * dx * dy + q1 * dy + q2 * dx;
            o[8] += q[8] + q0 * dy * dz + q2 * dz + q3 * dy;
            o[9] += q[9] + q0 * dx * dz + q1 * dz + q3 * dx;
        }
    }
#pragma unroll
    for (int k = 0; k < 30; ++k) m[k] = wave_sum(m[k]);
    if (lane < 30) {
        double v = m[0];
#pragma unroll
        for (int k = 1; k < 30; ++k) v = lane == k ? m[k] : v;
        mom[(size_t)c * kMomStride + lane] = v;
    }
}

// -------------------------------------------------------------------------
// evaluation
// -------------------------------------------------------------------------
constexpr int kEvalWaves = kEvalBlock / kWave;

// Lanes of one wavefront read what other lanes of it wrote to LDS.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// |r|^2 as the reference rounds it (no contraction: the opening test must
// take the same side).
__device__ __forceinline__ double r2_exact(double dx, double dy, double dz) {
    return dx * dx + dy * dy + dz * dz;
}

// 1 / sqrt(x) for finite x > 0: the hardware estimate y0 refined by the
// binomial series (1 - t)^(-1/2) = 1 + t/2 + 3 t^2/8 + ..., t = 1 - x y0^2.
__device__ __forceinline__ double rsqrt_pos(double x) {
#pragma clang fp contract(fast)
    const double y0 = __builtin_amdgcn_rsq(x);
    const double t = fma(-x * y0, y0, 1.0);
    return fma(y0 * t, fma(t, 0.375, 0.5), y0);
}

// The far-cell term: the moments against the Taylor coefficients of
// r / |r|^3 at the target (M2P_mul_calc), crossed into the velocity.
__device__ __forceinline__ void m2p_term(const double *__restrict__ M, double dx, double dy,
                                         double dz, double R2, double &vu, double &vv,
                                         double &vw) {
#pragma clang fp contract(fast)
    const double ir = rsqrt_pos(R2);
    const double i2 = ir * ir, i3 = i2 * ir, i5 = i3 * i2, i7 = i5 * i2;
    const double a = 3.0 * i5, b = 15.0 * i7;
    const double xx = dx * dx, yy = dy * dy, zz = dz * dz;
    const double xy = dx * dy, yz = dy * dz, xz = dx * dz, bxyz = b * xy * dz;
    const double mx[10] = {dx * i3,          i3 - a * xx,          -a * xy,
                           -a * xz,          (b * xx - 3.0 * a) * dx, (b * yy - a) * dx,
                           (b * zz - a) * dx, (b * xx - a) * dy,   bxyz,
                           (b * xx - a) * dz};
    const double my[10] = {dy * i3,          -a * xy,           i3 - a * yy,
                           -a * yz,          (b * xx - a) * dy, (b * yy - 3.0 * a) * dy,
                           (b * zz - a) * dy, (b * yy - a) * dx, (b * yy - a) * dz,
                           bxyz};
    const double mz[10] = {dz * i3,           -a * xz,           -a * yz,
                           i3 - a * zz,       (b * xx - a) * dz, (b * yy - a) * dz,
                           (b * zz - 3.0 * a) * dz, bxyz,        (b * zz - a) * dy,
                           (b * zz - a) * dx};
    // each cross term accumulated on its own (two FMAs per term)
    double su = 0.0, sv = 0.0, sw = 0.0;
#pragma unroll
    for (int k = 0; k < 10; ++k) {
        const double Vx = M[k], Vy = M[10 + k], Vz = M[20 + k];
        su = fma(Vy, mz[k], su);
        su = fma(-Vz, my[k], su);
        sv = fma(Vz, mx[k], sv);
        sv = fma(-Vx, mz[k], sv);
        sw = fma(Vx, my[k], sw);
        sw = fma(-Vy, mx[k], sw);
    }
    vu += su;
    vv += sv;
    vw += sw;
}

// Workgroups are dispatched to the 8 XCDs in turn (workgroup b to XCD
// b mod 8); this order hands each XCD one contiguous run of the targets, so
// the cells and sources around them stay in that XCD's L2 (7.22 -> 6.82 ms
// for the 780k-particle vortex ring).  A tail of fewer than 8 per XCD keeps
// the plain order.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int per = nb / kXcds;
    if (b >= per * kXcds) return b;
    return (b % kXcds) * per + b / kXcds;
}

// The direct sum of a near leaf's sources [ab, ae) for the lanes with near
// set, the sources read as wave-uniform loads (scalar cache), four per
// batch; SELF drops the pair of a target with itself.
template <bool SELF>
__device__ __forceinline__ void direct_pairs(const Src3 *__restrict__ S, uint32_t ab, uint32_t ae,
                                             double xp, double yp, double zp, double di,
                                             bool near, double &vu, double &vv, double &vw) {
#pragma clang fp contract(fast)
    if (!near) return;
#pragma unroll 4
    for (uint32_t j = ab; j < ae; ++j) {
        const Src3 q = S[j];
        const double ex = xp - q.x, ey = yp - q.y, ez = zp - q.z;
        const double r2 = ex * ex + ey * ey + ez * ez;
        const double ir = rsqrt_pos(r2);
        const double i3 = (!SELF || q.idx != di) ? ir * ir * ir : 0.0;
        vu += (q.ay * ez - q.az * ey) * i3;
        vv += (q.az * ex - q.ax * ez) * i3;
        vw += (q.ax * ey - q.ay * ex) * i3;
    }
}

// A near leaf's active sources are read as wave-uniform loads (scalar
// cache), four per batch.
__global__ void __launch_bounds__(kEvalBlock)
    k3_eval(const Cell3 *__restrict__ cells, int ncell, const double *__restrict__ mom,
            const Par3 *__restrict__ P, const Src3 *__restrict__ S,
            const uint32_t *__restrict__ perm, int n, double u_inf, double v_inf, double w_inf,
            double *__restrict__ u, double *__restrict__ v, double *__restrict__ w,
            int src_all, unsigned long long *__restrict__ wstat) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    const int blk = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int wave = __builtin_amdgcn_readfirstlane(blk * kEvalWaves + wv);
    const int i = wave * kWave + lane;
    if (wave * kWave >= n) return;
    const bool valid = i < n;
    double xp = 0.0, yp = 0.0, zp = 0.0;
    bool my_act = false;
    if (valid) {
        xp = P[i].x;
        yp = P[i].y;
        zp = P[i].z;
        my_act = P[i].act != 0.0;
    }
    const double di = (double)i;
    double vu = 0.0, vv = 0.0, vw = 0.0;
    uint32_t npair = 0, nfar = 0;
    uint32_t wfar = 0, wpair = 0;  // wave-level steps (uniform)
    uint32_t open = 0;  // bit L: this lane opened the level-L cell of the current path
    int c = 0;
    while (c < ncell) {
        const Cell3 C = cells[c];
        const int L = (int)(C.meta & 0xffu);
        if (L == 1) {
            // level-1 cells are always opened and never tested; a level-1
            // leaf is passed over
            if (C.meta & 0x100u) {
                c = (int)C.skip;
            } else {
                open = valid ? 2u : 0u;
                c = c + 1;
            }
            continue;
        }
        if (!(C.meta & 0x200u)) {
            c = (int)C.skip;
            continue;
        }
        const bool test = (open >> (L - 1)) & 1u;
        const double dx = xp - C.cx, dy = yp - C.cy, dz = zp - C.cz;
        const double R2 = r2_exact(dx, dy, dz);
        const bool far = test && R2 > C.rc2;
        const bool near = test && !(R2 > C.rc2);
        if (__ballot(far)) {
            ++wfar;
            if (far) {
                m2p_term(mom + (size_t)c * kMomStride, dx, dy, dz, R2, vu, vv, vw);
                ++nfar;
            }
        }
        const bool any_near = __ballot(near) != 0;
        if (C.meta & 0x100u) {
            if (any_near) {
                const uint32_t ab = C.abegin, ae = C.aend;
                wpair += ae - ab;
                // the target itself is a source of its own leaf (when active,
                // unless every particle is a source); only a leaf holding some
                // lane's own particle needs the self test
                const bool mine = (uint32_t)i >= C.begin && (uint32_t)i < C.end;
                if (__ballot(near && mine))
                    direct_pairs<true>(S, ab, ae, xp, yp, zp, di, near, vu, vv, vw);
                else
                    direct_pairs<false>(S, ab, ae, xp, yp, zp, di, near, vu, vv, vw);
                if (near) npair += ae - ab - ((src_all || my_act) && mine ? 1u : 0u);
            }
            c = (int)C.skip;
        } else {
            open = near ? (open | (1u << L)) : (open & ~(1u << L));
            c = any_near ? c + 1 : (int)C.skip;
        }
    }
    if (valid) {
        const uint32_t o = perm[i];
        u[o] = vu + u_inf;
        v[o] = vv + v_inf;
        w[o] = vw + w_inf;
    }
    // work counts per wavefront (summed by k3_eval_stats; same-address
    // atomics from every wavefront would serialise)
    unsigned long long sp = npair, sf = nfar;
    for (int m = 1; m < kWave; m <<= 1) {
        sp += __shfl_xor(sp, m, kWave);
        sf += __shfl_xor(sf, m, kWave);
    }
    if (lane == 0) {
        ulonglong4 r;
        r.x = sp;
        r.y = sf;
        r.z = wpair;
        r.w = wfar;
        reinterpret_cast<ulonglong4 *>(wstat)[wave] = r;
    }
}

// Sum of the per-wavefront work counts into the header.
__global__ void __launch_bounds__(1024)
    k3_eval_stats(const unsigned long long *__restrict__ wstat, int nw,
                  Fmm3Hdr *__restrict__ h) {
    __shared__ unsigned long long sm[16][4];
    unsigned long long a[4] = {0, 0, 0, 0};
    for (int q = threadIdx.x; q < nw; q += 1024)
        for (int k = 0; k < 4; ++k) a[k] += wstat[4 * (size_t)q + k];
    for (int m = 1; m < kWave; m <<= 1)
        for (int k = 0; k < 4; ++k) a[k] += __shfl_xor(a[k], m, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0)
        for (int k = 0; k < 4; ++k) sm[threadIdx.x / kWave][k] = a[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t[4] = {0, 0, 0, 0};
        for (int q = 0; q < 16; ++q)
            for (int k = 0; k < 4; ++k) t[k] += sm[q][k];
        h->pairs = t[0];
        h->m2p = t[1];
        h->wave_pairs = t[2];
        h->wave_m2p = t[3];
    }
}

}  // namespace

Ctx3::~Ctx3() {
    if (hdr_h) (void)hipHostFree(hdr_h);
}

void fmm3d_velocity(Ctx3 &c, hipStream_t s, int n, const double *x, const double *y,
                    const double *z, const double *sz, const double *wx, const double *wy,
                    const double *wz, const uint8_t *active, const Fmm3Params &p, double *u,
                    double *v, double *w, hipEvent_t *ev) {
    if (n < 1) throw Error(kBadArg, "n must be >= 1");
    if (!x || !y || !z || !wx || !wy || !wz || !active || !u || !v || !w)
        throw Error(kBadArg, "null pointer");
    if (p.type != 1 && p.type != 2) throw Error(kBadArg, "type must be 1 or 2");
    const bool t2 = p.type == 2;
    if (!t2 && (p.crit_num < 1 || p.crit_num > 4096))
        throw Error(kBadArg, "crit_num must be in [1, 4096]");
    if (t2 && (p.max_par < 0 || p.max_par > 4096 || p.max_src < 0 || p.max_src > 4096))
        throw Error(kBadArg, "max_par and max_src must be in [0, 4096]");
    if (p.max_depth < 1 || p.max_depth > kFmm3Depth)
        throw Error(kBadArg, "max_depth must be in [1, 20]");
    if (!(p.exp_tree >= 0.0) || !std::isfinite(p.exp_tree)) throw Error(kBadArg, "exp_tree must be >= 0");
    // with no expansion the type-2 box's upper face holds a particle whose
    // index is out of range (the reference's index2ID throws there)
    if (t2 && !(p.exp_tree > 0.0)) throw Error(kBadArg, "exp_tree must be > 0 for type 2");
    if (!(p.support > 0.0) || !std::isfinite(p.support)) throw Error(kBadArg, "support must be > 0");
    for (double q : {p.center[0], p.center[1], p.center[2], p.u_inf, p.v_inf, p.w_inf})
        if (!std::isfinite(q)) throw Error(kBadArg, "centre and free stream must be finite");
    if (!c.hdr_h) VPM_HIP(hipHostMalloc(&c.hdr_h, sizeof(Fmm3Hdr), hipHostMallocDefault));
    c.hdr_d = c.hdr_buf.as<Fmm3Hdr>(1);
    Fmm3Hdr *hd = c.hdr_d;
    c.last = p;

    const int nb = div_up(n, 256);
    const int nbb = std::min(1024, div_up(n, 256 * 8));
    double *part = c.bbox.as<double>(8 * (size_t)nbb);
    if (ev) VPM_HIP(hipEventRecord(ev[0], s));
    k3_bbox<<<nbb, 256, 0, s>>>(x, y, z, n, part);
    VPM_LAUNCH_CHECK();

    uint32_t *khi = c.khi.as<uint32_t>(n), *klo = c.klo.as<uint32_t>(n);
    uint32_t *ka = c.key_a.as<uint32_t>(n), *kb = c.key_b.as<uint32_t>(n);
    uint32_t *va = c.val_a.as<uint32_t>(n), *vb = c.val_b.as<uint32_t>(n);
    const size_t tmp_words = std::max(radix_tmp_words(n), scan_tmp_words((long long)n + 1)) + 64;
    uint32_t *tmp = c.tmp.as<uint32_t>(tmp_words);
    Par3 *P = c.par.as<Par3>(n);
    uint64_t *key = c.key64.as<uint64_t>(n);
    uint8_t *act8 = c.act8.as<uint8_t>(n);
    uint32_t *A = c.act_scan.as<uint32_t>((size_t)n + 1);
    uint8_t *settle = c.settle.as<uint8_t>(n), *mprev = c.mprev.as<uint8_t>(n);
    uint32_t *starts = c.starts.as<uint32_t>(n), *pos = c.pos.as<uint32_t>((size_t)n + 1);

    // The sort resolves 10 levels when the tree is that shallow, else 20; the
    // choice is a function of the tree alone (a build sorted the other way is
    // repeated), so the order inside a leaf, and every rounding after it,
    // depends on the input only.
    int D = c.levels_guess;
    for (int attempt = 0;; ++attempt) {
        const int deep = D > kHalf;
        // (the root kernel also clears the header's flags and counters)
        k3_root<<<1, 256, 0, s>>>(part, nbb, p.center[0], p.center[1], p.center[2], p.exp_tree,
                                  p.type, hd);
        k3_keys<<<nb, 256, 0, s>>>(x, y, z, n, hd, khi, klo, ka, va, deep, p.type);
        const bool alt = radix_sort_pairs(ka, va, kb, vb, n, 0, 3 * kHalf, tmp, s);
        uint32_t *perm = alt ? vb : va;
        if (deep) {
            uint32_t *k2 = alt ? ka : kb, *v2 = alt ? va : vb;
            uint32_t *k3 = alt ? kb : ka, *v3 = alt ? vb : va;
            k3_regather<<<nb, 256, 0, s>>>(perm, khi, n, k2, v2);
            const bool alt2 = radix_sort_pairs(k2, v2, k3, v3, n, 0, 3 * kHalf, tmp, s);
            perm = alt2 ? v3 : v2;
        }
        c.perm = perm;
        k3_gather<<<nb, 256, 0, s>>>(perm, khi, klo, x, y, z, sz, wx, wy, wz, active, n, P, key,
                                     act8);
        scan_exclusive_u8(act8, A, n, &hd->n_active, true, tmp, s);
        uint4 *bstat = c.bstat.as<uint4>(nb);
        if (t2) {
            k3_settle_t2<<<nb, 256, 0, s>>>(key, A, n, D, p.max_par, p.max_src, settle, mprev,
                                            starts, bstat, hd);
        } else {
            uint8_t *lcpw = c.lcpw.as<uint8_t>(n);
            if (n >= p.crit_num)
                k3_lcpw<<<div_up(n - p.crit_num + 1, 256), 256, 0, s>>>(key, n, D, p.crit_num,
                                                                        lcpw);
            k3_settle<<<nb, 256, (size_t)(256 + p.crit_num), s>>>(
                key, lcpw, n, D, p.crit_num, settle, mprev, starts, bstat, hd);
        }
        k3_settle_stats<<<1, 1024, 0, s>>>(bstat, nb, hd);
        scan_exclusive_u32(starts, pos, n, &hd->n_cells, true, tmp, s);
        VPM_LAUNCH_CHECK();
        VPM_HIP(hipMemcpyAsync(c.hdr_h, hd, sizeof(Fmm3Hdr), hipMemcpyDeviceToHost, s));
        VPM_HIP(hipStreamSynchronize(s));
        const Fmm3Hdr &hh = *c.hdr_h;
        if (hh.flags & kF3NonFinite) throw Error(kNonFinite, "non-finite particle coordinate");
        if (hh.flags & kF3Degenerate)
            throw Error(kDegenerate, "every particle at one point (the type-2 root has no size)");
        if (hh.flags & kF3Depth)
            throw Error(kDepth, "octree deeper than 20 levels (too many coincident particles?)");
        const int want = ((hh.flags & kF3Deeper) || (int)hh.depth > kHalf) ? kFmm3Depth : kHalf;
        if (want != D && attempt == 0) {
            D = want;
            c.levels_guess = want;
            continue;
        }
        if (want != D) throw Error(kInternal, "3-D sort depth did not settle");
        break;
    }
    const Fmm3Hdr hh = *c.hdr_h;
    if ((int)hh.depth > p.max_depth) throw Error(kDepth, "octree deeper than max_depth");
    // type 2 asks every level-1 cell for its children (fmm3D.cpp:1150-1152);
    // get_childID throws for a leaf (treeCell.cpp:384-388), so a level-1 leaf
    // beside deeper cells has no result in the reference
    if (t2 && hh.depth >= 2 && hh.min_settle == 1)
        throw Error(kLevel1Leaf, "type 2: a level-1 cell is a leaf while the tree goes deeper "
                                 "(the reference's get_childID throws there)");
    const uint32_t ncell = hh.n_cells;
    Fmm3Info &fi = c.info;
    fi.n = n;
    fi.n_active = (int)hh.n_active;
    fi.n_cells = (int)ncell;
    fi.n_leaves = (int)hh.n_leaves;
    fi.depth = (int)hh.depth;
    for (int d = 0; d < 3; ++d) fi.root_c[d] = hh.root_c[d];
    fi.root_len = hh.root_len;
    fi.sorted_levels = D;

    Cell3 *cells = c.cells.as<Cell3>(ncell);
    double *mom = c.mom.as<double>((size_t)ncell * kMomStride);
    Src3 *S = c.src.as<Src3>((size_t)std::max(1u, t2 ? (uint32_t)n : hh.n_active));
    k3_emit<<<nb, 256, 0, s>>>(key, settle, mprev, pos, A, n, ncell, hd, p.support, p.type,
                               cells);
    k3_compact<<<nb, 256, 0, s>>>(P, A, n, t2 ? 1 : 0, S);
    if (ev) VPM_HIP(hipEventRecord(ev[1], s));
    const int cb = div_up(ncell, 4);
    if (ncell > 0) {
        k3_p2m<<<div_up(ncell, 256), 256, 0, s>>>(cells, (int)ncell, S, mom);
        k3_m2m<<<cb, 256, 0, s>>>(cells, (int)ncell, mom);
    }
    if (ev) VPM_HIP(hipEventRecord(ev[2], s));
    const int waves = div_up(n, kWave);
    unsigned long long *wstat = c.wstat.as<unsigned long long>(4 * (size_t)waves);
    k3_eval<<<div_up(waves, kEvalWaves), kEvalBlock, 0, s>>>(
        cells, (int)ncell, mom, P, S, c.perm, n, p.u_inf, p.v_inf, p.w_inf, u, v, w, t2 ? 1 : 0,
        wstat);
    if (ev) VPM_HIP(hipEventRecord(ev[3], s));
    k3_eval_stats<<<1, 1024, 0, s>>>(wstat, waves, hd);
    VPM_LAUNCH_CHECK();
    VPM_HIP(hipMemcpyAsync(c.hdr_h, hd, sizeof(Fmm3Hdr), hipMemcpyDeviceToHost, s));
    VPM_HIP(hipStreamSynchronize(s));
    fi.pairs = c.hdr_h->pairs;
    fi.m2p = c.hdr_h->m2p;
    fi.wave_pairs = c.hdr_h->wave_pairs;
    fi.wave_m2p = c.hdr_h->wave_m2p;
    fi.ms_tree = fi.ms_moments = fi.ms_eval = -1.0;
    if (ev) {
        float a = 0.f, b = 0.f, e = 0.f;
        VPM_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
        VPM_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
        VPM_HIP(hipEventElapsedTime(&e, ev[2], ev[3]));
        fi.ms_tree = a;
        fi.ms_moments = b;
        fi.ms_eval = e;
    }
}

void fmm3d_cells(Ctx3 &c, hipStream_t s, int32_t *level, int64_t *key, int32_t *count,
                 uint8_t *leaf, uint8_t *active, double *centre3, double *size) {
    const int nc = c.info.n_cells, n = c.info.n;
    if (n == 0) throw Error(kBadArg, "no 3-D solve has run on this handle");
    std::vector<Cell3> cl((size_t)nc);
    std::vector<uint64_t> k((size_t)n);
    if (nc)
        VPM_HIP(hipMemcpyAsync(cl.data(), c.cells.p, sizeof(Cell3) * nc, hipMemcpyDeviceToHost, s));
    VPM_HIP(hipMemcpyAsync(k.data(), c.key64.p, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, s));
    VPM_HIP(hipStreamSynchronize(s));
    for (int j = 0; j < nc; ++j) {
        const Cell3 &q = cl[(size_t)j];
        const int L = (int)(q.meta & 0xffu);
        if (level) level[j] = L;
        if (key) key[j] = (int64_t)(k[q.begin] >> (3 * (kFmm3Depth - L)));
        if (count) count[j] = (int32_t)(q.end - q.begin);
        if (leaf) leaf[j] = (q.meta >> 8) & 1u;
        if (active) active[j] = (q.meta >> 9) & 1u;
        if (centre3) {
            centre3[3 * j] = q.cx;
            centre3[3 * j + 1] = q.cy;
            centre3[3 * j + 2] = q.cz;
        }
        if (size) size[j] = std::ldexp(c.info.root_len, -L);
    }
}

void fmm3d_perm(Ctx3 &c, hipStream_t s, int32_t *perm) {
    if (c.info.n == 0 || !c.perm) throw Error(kBadArg, "no 3-D solve has run on this handle");
    VPM_HIP(hipMemcpyAsync(perm, c.perm, sizeof(uint32_t) * c.info.n, hipMemcpyDeviceToHost, s));
    VPM_HIP(hipStreamSynchronize(s));
}

}  // namespace vpm
