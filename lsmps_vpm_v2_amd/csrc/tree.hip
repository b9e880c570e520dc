// tree.hip -- cell binning and interaction lists on gfx950.
//
// Reproduces the adaptive quadtree of TreeCell::initializeTree (reference
// src/FMM/treeCell.cpp:1043-1243) and the lists of TreeCell::intList /
// extList (treeCell.cpp:636-1032) with data-parallel kernels:
//
//  1. bounding box, root length and pivot exactly as treeCell.cpp:1064-1092;
//  2. a 32-bit Morton key per particle.  The reference cell index at level L
//     is (int)((x - pivot) / (R / 2^L)); since R / 2^L is exact and IEEE
//     division commutes with power-of-two scaling, that index equals the top
//     L bits of floor(((x - pivot) / R) * 2^16), so one key serves all
//     16 levels;
//  3. stable LSD radix sort of (key, index): every cell of every level is a
//     contiguous range (Morton order inside it, index order among equal keys);
//  4. the leaf level of every particle without walking levels: a cell at
//     level L holds more than W particles iff some run of W+1 consecutive
//     sorted particles that contains the particle shares its level-L prefix,
//     i.e. iff max over such runs of lcp(first, last) >= L (likewise for runs
//     of active particles with W = max_src);
//  5. cells emitted in pre-order (particle i starts the cells of levels
//     lcp(i-1, i)+1 .. leafLevel(i)), stable-partitioned by level, so cells of
//     one level are contiguous and Morton-ordered (binary-searchable);
//  6. per cell the 3x3 "deepest existing cell" table, from which lists 1-4 are
//     formed with the reference's rules and its 1e-9 tolerance.
//
// All geometry is FP64 with contraction disabled (built with
// -ffp-contract=off) so centres and distances round as in the reference.
#include "ctx.hpp"
#include "prim.hpp"

namespace vpm {
namespace {

constexpr unsigned kLeaf = 1u, kActive = 2u;

__device__ __forceinline__ uint32_t part1by1(uint32_t v) {
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}
__device__ __forceinline__ uint32_t compact1by1(uint32_t v) {
    v &= 0x55555555u;
    v = (v | (v >> 1)) & 0x33333333u;
    v = (v | (v >> 2)) & 0x0F0F0F0Fu;
    v = (v | (v >> 4)) & 0x00FF00FFu;
    v = (v | (v >> 8)) & 0x0000FFFFu;
    return v;
}
// Number of leading quadtree levels two keys share (0..16).
__device__ __forceinline__ int lcp_levels(uint32_t a, uint32_t b) {
    uint32_t x = a ^ b;
    return x == 0 ? kMaxDepth : (__clz(x) >> 1);
}
// Cell side at a level: R / 2^L (exact).
__device__ __forceinline__ double cell_len(double R, int level) {
    return R / (double)(1u << level);
}

// Sum of v over a 256-thread block, written to *dst by thread 0 (every thread
// of the block must call it: block sums of scan inputs, so that their scans
// skip the reduce pass).
__device__ __forceinline__ void block_sum_store(uint32_t v, uint32_t *dst) {
    __shared__ uint32_t sw[256 / kWave];
    for (int m = 1; m < kWave; m <<= 1) v += __shfl_xor(v, m, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) sw[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int u = 0; u < (int)blockDim.x / kWave; ++u) t += sw[u];
        *dst = t;
    }
}

// ---------------------------------------------------------------------------
// 1. bounding box and root geometry
// ---------------------------------------------------------------------------
// Block 0 also clears the device header, statistic rows included (every later
// kernel of the build runs after this one on the stream).
__global__ void __launch_bounds__(256)
    k_bbox_partial(const double *x, const double *y, int n, double *part, DevHdr *hdr) {
    __shared__ double s[4][256];
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    if (blockIdx.x == 0)
        for (int u = threadIdx.x; u < (int)(sizeof(DevHdr) / 4); u += blockDim.x)
            reinterpret_cast<uint32_t *>(hdr)[u] = 0u;
    double mnx = x[0], mxx = x[0], mny = y[0], mxy = y[0];
    int nf = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        double a = x[i], b = y[i];
        nf |= !isfinite(a) || !isfinite(b);
        mnx = fmin(mnx, a);
        mxx = fmax(mxx, a);
        mny = fmin(mny, b);
        mxy = fmax(mxy, b);
    }
    s[0][threadIdx.x] = mnx;
    s[1][threadIdx.x] = mxx;
    s[2][threadIdx.x] = mny;
    s[3][threadIdx.x] = mxy;
    __syncthreads();
    if (nf) atomicOr(&bad, 1);
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            s[0][threadIdx.x] = fmin(s[0][threadIdx.x], s[0][threadIdx.x + o]);
            s[1][threadIdx.x] = fmax(s[1][threadIdx.x], s[1][threadIdx.x + o]);
            s[2][threadIdx.x] = fmin(s[2][threadIdx.x], s[2][threadIdx.x + o]);
            s[3][threadIdx.x] = fmax(s[3][threadIdx.x], s[3][threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[5 * blockIdx.x + 0] = s[0][0];
        part[5 * blockIdx.x + 1] = s[1][0];
        part[5 * blockIdx.x + 2] = s[2][0];
        part[5 * blockIdx.x + 3] = s[3][0];
        part[5 * blockIdx.x + 4] = bad ? 1.0 : 0.0;
    }
}

// ---------------------------------------------------------------------------
// 2-3. root geometry and Morton keys, sort, gather
// ---------------------------------------------------------------------------
// Type-1 key bits of one coordinate (hierarchy_mesh and amount_inbox_new,
// subroutine_all_fmm_2d.cpp:175-263, 635-678): the level-1 box is
// [lo + b db, lo + (b + 1) db); below it a box [l, h) has the children
// [l + b d, l + (b + 1) d), d = (h - l) / 2, each half-open, so a coordinate
// takes child 0 below l + d and child 1 below l + 2 d.  The bounds are formed
// by the reference's own additions (contraction off), so the path is the
// reference's to the last bit.  Returns false if the coordinate falls outside
// the level-1 boxes or past the rounded upper bound of a split box's children.
__device__ __forceinline__ bool t1_path(double x, double lo, double db, int lmax,
                                        uint32_t &bits) {
#pragma clang fp contract(off)
    uint32_t b = x < lo + 1.0 * db ? 0u : 1u;
    double l = lo + (double)b * db, h = lo + (double)(b + 1) * db;
    bool ok = x >= lo && x < h;
    bits = b;
    for (int lv = 2; lv <= lmax; ++lv) {
        const double d = (h - l) / 2.0;
        b = x < l + 1.0 * d ? 0u : 1u;
        const double nl = l + (double)b * d, nh = l + (double)(b + 1) * d;
        ok = ok && x < nh;
        l = nl;
        h = nh;
        bits = (bits << 1) | b;
    }
    return ok;
}

// Root geometry, treeCell.cpp:1079-1092: R = max side * (1 + expTree/100),
// pivot = (max + min - R)/2, from the bounding-box partials; every block
// reduces them itself (a few KB, read from L2), so no launch separates the
// box from the keys.  Block 0 records the geometry and the flags in the
// header.  Then the Morton keys of one radix tile (kRsTile particles, the
// radix sort's element order) and the tile's counts of the first sort digit
// (key >> shift) & 255, so the sort's first pass needs no counting launch.
// With T1, the type-1 root (hierarchy_mesh, subroutine_all_fmm_2d.cpp:
// 120-170: the box padded by dr 1e-5 on each side, dr = shorter side /
// 2^lmax, then squared about its centre on the longer side; db = half that
// side) recorded as pivot = (xmin, ymin), root_len = 2 db, and the keys of
// the type-1 paths (t1_path) in the same bit layout.
template <bool T1>
__global__ void __launch_bounds__(kRsBlock)
    k_keys(const double *x, const double *y, int n, const double *part, int nb,
           double expansion, int lmax, DevHdr *hdr, uint32_t *key, uint32_t *val, int shift,
           uint32_t *table, int tiles) {
#pragma clang fp contract(off)
    __shared__ double s[4][kRsBlock / kWave];
    __shared__ double geo[3];
    __shared__ uint32_t h[256];
    __shared__ int sbad;
    if (threadIdx.x == 0) sbad = 0;
    h[threadIdx.x] = 0;
    double mn0 = part[0], mx0 = part[1], mn1 = part[2], mx1 = part[3];
    int bad = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        mn0 = fmin(mn0, part[5 * b + 0]);
        mx0 = fmax(mx0, part[5 * b + 1]);
        mn1 = fmin(mn1, part[5 * b + 2]);
        mx1 = fmax(mx1, part[5 * b + 3]);
        bad |= part[5 * b + 4] != 0.0;
    }
    for (int m = 1; m < kWave; m <<= 1) {
        mn0 = fmin(mn0, __shfl_xor(mn0, m, kWave));
        mx0 = fmax(mx0, __shfl_xor(mx0, m, kWave));
        mn1 = fmin(mn1, __shfl_xor(mn1, m, kWave));
        mx1 = fmax(mx1, __shfl_xor(mx1, m, kWave));
    }
    const int wv = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        s[0][wv] = mn0;
        s[1][wv] = mx0;
        s[2][wv] = mn1;
        s[3][wv] = mx1;
    }
    __syncthreads();
    if (bad) atomicOr(&sbad, 1);
    if (threadIdx.x == 0) {
        double mn[2] = {s[0][0], s[2][0]}, mx[2] = {s[1][0], s[3][0]};
        for (int u = 1; u < kRsBlock / kWave; ++u) {
            mn[0] = fmin(mn[0], s[0][u]);
            mx[0] = fmax(mx[0], s[1][u]);
            mn[1] = fmin(mn[1], s[2][u]);
            mx[1] = fmax(mx[1], s[3][u]);
        }
        if (T1) {
            double xmin = mn[0], xmax = mx[0], ymin = mn[1], ymax = mx[1];
            const double tlout = 1.0e-5;
            const double sx = fabs(xmax - xmin), sy = fabs(ymax - ymin);
            const double dr = (sy < sx ? sy : sx) / (double)(1u << lmax);
            xmin = xmin - dr * tlout;
            xmax = xmax + dr * tlout;
            ymin = ymin - dr * tlout;
            ymax = ymax + dr * tlout;
            const double rdx = fabs(xmax - xmin), rdy = fabs(ymax - ymin);
            double db;
            if (rdy >= rdx) {
                const double cent = (xmin + xmax) / 2.0;
                xmin = cent - rdy / 2.0;
                db = rdy / 2.0;
            } else {
                const double cent = (ymin + ymax) / 2.0;
                ymin = cent - rdx / 2.0;
                db = rdx / 2.0;
            }
            geo[0] = 2.0 * db;
            geo[1] = xmin;
            geo[2] = ymin;
        } else {
            double R = 0.0;
            for (int d = 0; d < 2; ++d) {
                double side = fabs(mx[d] - mn[d]);
                if (side > R) R = side;
            }
            R *= expansion;
            geo[0] = R;
            for (int d = 0; d < 2; ++d) geo[1 + d] = (mx[d] + mn[d] - R) / 2.0;
        }
        if (blockIdx.x == 0) {
            hdr->root_len = geo[0];
            for (int d = 0; d < 2; ++d) {
                hdr->mn[d] = mn[d];
                hdr->mx[d] = mx[d];
                hdr->pivot[d] = geo[1 + d];
            }
        }
    }
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        unsigned f = 0;
        if (sbad) f |= kFlagNonFinite;
        if (!(geo[0] > 0.0)) f |= kFlagRange;
        atomicOr(&hdr->flags, f);  // other blocks may raise kFlagT1Lost meanwhile
        hdr->leaf_lvl_min = kMaxDepth + 1;
        hdr->leaf_lvl_max = 0;
    }
    const double R = geo[0], px = geo[1], py = geo[2];
    const int base = blockIdx.x * kRsTile;
    bool lost = false;
#pragma unroll
    for (int j = 0; j < kRsItems; ++j) {
        const int i = base + j * kRsBlock + threadIdx.x;
        if (i >= n) break;
        uint32_t ux = 0, uy = 0;
        if (T1) {
            const double db = R / 2.0;
            const bool okx = t1_path(x[i], px, db, lmax, ux);
            const bool oky = t1_path(y[i], py, db, lmax, uy);
            lost |= !(okx && oky);
            ux <<= 16 - lmax;
            uy <<= 16 - lmax;
        } else {
            const double tx = (x[i] - px) / R, ty = (y[i] - py) / R;
            const double sx = tx * 65536.0, sy = ty * 65536.0;
            if (sx >= 0.0 && sx < 65536.0) ux = (uint32_t)sx;
            if (sy >= 0.0 && sy < 65536.0) uy = (uint32_t)sy;
        }
        const uint32_t k = part1by1(ux) | (part1by1(uy) << 1);
        key[i] = k;
        val[i] = (uint32_t)i;
        atomicAdd(&h[(k >> shift) & 255u], 1u);
    }
    if (T1 && lost) atomicOr(&hdr->flags, kFlagT1Lost);
    __syncthreads();
    table[(size_t)threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

// Particles in Morton order.  The source strength is q[o] qscale, or, for the
// velocity entry points (q == nullptr), -omega s^2 / (2 pi) formed here in the
// operation order of velocity_calc.cpp:477 (pow(s, 2) is the exact square).
// Also the level lcps of the sorted keys: d[i] = lcp(key[i-1], key[i])
// (d[0] = d[n] = 0) and w[i] = lcp(key[i], key[i+W]).
__global__ void __launch_bounds__(256)
    k_gather(const uint32_t *perm, const double *x, const double *y, const double *q,
             const double *ssz, const double *omega, const uint8_t *a, int n, double *xs,
             double *ys, double *qs, uint8_t *as, uint32_t *bs_act, const uint32_t *key, int W,
             uint8_t *d, uint8_t *w, double *ss, double qscale, DevHdr *hdr_w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t act = 0;
    if (i < n) {
        const uint32_t ki = key[i];
        d[i] = i == 0 ? 0 : (uint8_t)lcp_levels(key[i - 1], ki);
        if (i + W < n) w[i] = (uint8_t)lcp_levels(ki, key[i + W]);
        if (i == n - 1) d[n] = 0;
        const uint32_t o = perm[i];
        xs[i] = x[o];
        ys[i] = y[o];
        if (q) {
            qs[i] = q[o] * qscale;
        } else {
            const double sz = ssz[o];
            qs[i] = -omega[o] * (sz * sz) / (2 * M_PI);
        }
        if (ss) ss[i] = ssz[o];  // sizes, for a regularized near field
        act = a[o] ? 1u : 0u;
        as[i] = (uint8_t)act;
    }
    // inactive particles per wavefront into the block's statistic row (no
    // atomic at all when every particle is active)
    const uint64_t inact = __ballot(i < n && !act);
    if (inact && (threadIdx.x & (kWave - 1)) == 0)
        atomicAdd(&hdr_w->stat[blockIdx.x % kStatRows][kStInactive],
                  (unsigned long long)__popcll(inact));
    block_sum_store(act, bs_act + blockIdx.x);  // active particles per block
}

// ---------------------------------------------------------------------------
// 4. leaf levels from windowed LCPs
// ---------------------------------------------------------------------------
// Block sums of up to four 64-bit statistics, added by thread 0 to the
// header's statistic row of this block (every thread must call it).
__device__ __forceinline__ void block_stats(DevHdr *hdr, int col0, int ncol,
                                            unsigned long long v0, unsigned long long v1,
                                            unsigned long long v2 = 0, unsigned long long v3 = 0) {
    __shared__ unsigned long long sv[4][16];
    unsigned long long v[4] = {v0, v1, v2, v3};
#pragma unroll
    for (int c = 0; c < 4; ++c)
        for (int m = 1; m < kWave; m <<= 1) v[c] += __shfl_xor(v[c], m, kWave);
    const int w = threadIdx.x / kWave, nw = (blockDim.x + kWave - 1) / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0)
#pragma unroll
        for (int c = 0; c < 4; ++c) sv[c][w] = v[c];
    __syncthreads();
    if (threadIdx.x < ncol) {
        unsigned long long t = 0;
        for (int u = 0; u < nw; ++u) t += sv[threadIdx.x][u];
        if (t) atomicAdd(&hdr->stat[blockIdx.x % kStatRows][col0 + threadIdx.x], t);
    }
}

constexpr int kSplitBlock = 256;

// Sliding-window maxima by doubling (a sparse table): after step s the
// buffer holds the maximum over [x, x + 2^s); a window of length B, with
// 2^k <= B < 2^(k+1), is the larger of the two 2^k windows at its ends.
// Entries outside the valid range hold 0, which never raises a maximum of
// valid lcp values.  The source stays intact: the steps ping-pong between b
// and c; returns the buffer holding the 2^k maxima.  A query's window lies
// inside [0, len), so the entries the doubling leaves partial (x + 2^s
// beyond len) are never read.
__device__ __forceinline__ const uint8_t *block_window_max(const uint8_t *src, uint8_t *b,
                                                           uint8_t *c, int len, int k) {
    const uint8_t *a = src;
    for (int st = 0; st < k; ++st) {
        const int sft = 1 << st;
        for (int x = threadIdx.x; x < len; x += blockDim.x)
            b[x] = (x + sft < len) ? max(a[x], a[x + sft]) : a[x];
        __syncthreads();
        a = b;
        b = c;
        c = const_cast<uint8_t *>(a);
    }
    return a;
}

// floor(log2(v)) for v >= 1
__device__ __forceinline__ int ilog2(int v) { return 31 - __clz(v); }

// Per particle: the deepest level whose cell is over the particle (W) or
// source (S) threshold, hence the leaf level; and the number of cells whose
// first particle it is.
__global__ void __launch_bounds__(kSplitBlock)
    k_split(const uint32_t *key, const uint8_t *d, const uint8_t *w, const uint32_t *A,
            const uint32_t *apos, const DevHdr *hdr, int n, int W, int S,
            int min_level, int max_level, uint8_t *leaf_lvl, DevHdr *hdr_w, uint32_t *blk_lv,
            int nbs) {
    extern __shared__ uint8_t lds[];
    // cells started per level (a particle whose lcp with its predecessor is d
    // starts the cells of levels d+1 .. its leaf level, the last a leaf)
    __shared__ uint32_t lvc[kLvCols], lvl_leaf[kMaxDepth + 1];
    if (threadIdx.x < kLvCols) lvc[threadIdx.x] = 0;
    if (threadIdx.x <= kMaxDepth) lvl_leaf[threadIdx.x] = 0;
    const int i0 = blockIdx.x * kSplitBlock;
    const int i = i0 + threadIdx.x;
    const int ilast = min(i0 + kSplitBlock, n) - 1;
    const int na = (int)hdr->n_active;
    // every particle active: the active ranks and positions are the identity
    // (the scan then leaves A and apos unwritten)
    const bool all_act = na == n;
    auto rank = [&](int x) { return all_act ? x : (int)A[x]; };
    auto apos_at = [&](int m) { return all_act ? (uint32_t)m : apos[m]; };
    // runs of W+1 particles: start j in [i - W, i]; LDS index j - (i0 - W)
    const int wlen = W + kSplitBlock;
    const int wbase = i0 - W;
    const int wlast = n - 1 - W;  // last valid run start (< 0: no run of W+1 particles)
    uint8_t *sw = lds, *gw = sw + wlen, *hw = gw + wlen;
    for (int t = threadIdx.x; t < wlen; t += kSplitBlock) {
        const int j = wbase + t;
        sw[t] = (j >= 0 && j <= wlast) ? w[j] : 0;
    }
    // runs of S+1 active particles: start m in [r - S, r - 1]; LDS index m - vbase
    const int r_lo = rank(i0), r_hi = rank(ilast);
    const int vbase = r_lo - S - 1;
    const int vlast = na - 1 - S;
    const int vlen = r_hi - vbase + 1;
    uint8_t *sv = hw + wlen, *gv = sv + vlen, *hv = gv + vlen;
    const int kw = ilog2(W + 1), ks = S > 0 ? ilog2(S) : 0;
    // Every particle active and S == W: the active positions are the
    // identity, v[m] = w[m], so the active window is the particle window
    // shifted by one (sv[a] = sw[a - 1]) and, with equal doubling depths, its
    // maxima are those of the particle window: one table instead of two.
    const bool same = all_act && S == W && kw == ks;
    if (!same) {
        // v[m] = lcp(key of active m, key of active m + S), formed here from
        // the compacted active positions
        for (int t = threadIdx.x; t < vlen; t += kSplitBlock) {
            const int m = vbase + t;
            sv[t] = (m >= 0 && m <= vlast)
                        ? (uint8_t)lcp_levels(key[apos_at(m)], key[apos_at(m + S)])
                        : 0;
        }
    }
    __syncthreads();
    const uint8_t *mw = block_window_max(sw, gw, hw, wlen, kw);
    const uint8_t *mv = (S > 0 && !same) ? block_window_max(sv, gv, hv, vlen, ks) : sv;
    int my_min = kMaxDepth + 1, my_max = 0, my_depth = 0;
    if (i < n) {
        const uint32_t ki = key[i];
        const int t = threadIdx.x;
        int par_split = -1, src_split = -1;
        const int r = rank(i);
        const int a = r - S - vbase;  // active window [r - S, r - 1] in LDS coordinates
        // window [t, t + W] and active window [a, a + S - 1] from their end 2^k maxima
        if (wlast >= 0) par_split = (int)max(mw[t], mw[t + W + 1 - (1 << kw)]);
        if (S > 0)
            src_split = same ? (int)max(mw[a - 1], mw[a - 1 + S - (1 << ks)])
                             : (int)max(mv[a], mv[a + S - (1 << ks)]);
        if (r <= vlast) {  // run starting at the first active at or after i
            const int vr = same ? (int)sw[r - vbase - 1] : (int)sv[r - vbase];
            const int e = min(vr, lcp_levels(ki, key[apos_at(r)]));
            src_split = max(src_split, e);
        }
        const int m0 = r - S - 1;
        if (m0 >= 0 && m0 <= vlast) {  // run ending at the last active before i
            // (v[m0] for the block's first particle lies just before the window)
            const int vm = !same ? (int)sv[m0 - vbase]
                           : m0 - vbase >= 1 ? (int)sw[m0 - vbase - 1]
                                             : lcp_levels(key[m0], key[m0 + S]);
            const int e = min(vm, lcp_levels(key[apos_at(r - 1)], ki));
            src_split = max(src_split, e);
        }
        int lvl = max(max(par_split, src_split) + 1, max(min_level, 1));
        if (max_level > 0) lvl = min(lvl, max_level);
        if (lvl > kMaxDepth) {
            my_depth = 1;  // depth overflow
            lvl = kMaxDepth;
        }
        leaf_lvl[i] = (uint8_t)lvl;
        const int di = d[i];
        for (int L = di + 1; L <= lvl; ++L) atomicAdd(&lvc[L], 1u);
        if (lvl > di) {
            atomicAdd(&lvc[kLvCols - 1], 1u);
            atomicAdd(&lvl_leaf[lvl], 1u);
        }
        my_min = lvl;
        my_max = lvl;
    }
    // block min / max / overflow, then one row of atomics per block
    for (int m = 1; m < kWave; m <<= 1) {
        my_min = min(my_min, __shfl_xor(my_min, m, kWave));
        my_max = max(my_max, __shfl_xor(my_max, m, kWave));
        my_depth |= __shfl_xor(my_depth, m, kWave);
    }
    __shared__ int s_mm[kSplitBlock / kWave][3];
    const int wv = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        s_mm[wv][0] = my_min;
        s_mm[wv][1] = my_max;
        s_mm[wv][2] = my_depth;
    }
    __syncthreads();
    // cells and leaves this block starts per level, for the level-major emission
    if (threadIdx.x <= kMaxDepth) {
        blk_lv[(size_t)threadIdx.x * nbs + blockIdx.x] = lvc[threadIdx.x];
        blk_lv[(size_t)(kMaxDepth + 1 + threadIdx.x) * nbs + blockIdx.x] = lvl_leaf[threadIdx.x];
    }
    uint32_t *row = hdr_w->split_rows[blockIdx.x % kSlotRows];
    if (threadIdx.x < kLvCols) {
        const uint32_t c = lvc[threadIdx.x];
        if (c) atomicAdd(row + threadIdx.x, c);
    } else if (threadIdx.x == kLvCols) {
        int mn = s_mm[0][0], mx = s_mm[0][1], dp = s_mm[0][2];
        for (int u = 1; u < kSplitBlock / kWave; ++u) {
            mn = min(mn, s_mm[u][0]);
            mx = max(mx, s_mm[u][1]);
            dp |= s_mm[u][2];
        }
        atomicMax(row + kSlotMinInv, 31u - (uint32_t)mn);
        atomicMax(row + kSlotMax, (uint32_t)mx);
        if (dp) atomicOr(row + kSlotDepth, 1u);
    }
}

// ---------------------------------------------------------------------------
// 5. cells
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_lv_scan(uint32_t *blk_lv, int nbs, uint32_t *lv_tot) {
    // exclusive scan, in place, of one row of per-block counts: rows 0 ..
    // kMaxDepth cells per level, the next kMaxDepth + 1 rows leaves per level
    __shared__ uint32_t sw[256 / kWave];
    uint32_t *row = blk_lv + (size_t)blockIdx.x * nbs;
    const int per = (nbs + 255) / 256;
    const int b0 = threadIdx.x * per, b1 = min(b0 + per, nbs);
    uint32_t sum = 0;
    for (int b = b0; b < b1; ++b) sum += row[b];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    uint32_t inc = sum;
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o, kWave);
        if (lane >= o) inc += y;
    }
    if (lane == kWave - 1) sw[w] = inc;
    __syncthreads();
    uint32_t run = inc - sum;
    for (int u = 0; u < w; ++u) run += sw[u];
    for (int b = b0; b < b1; ++b) {
        const uint32_t v = row[b];
        row[b] = run;
        run += v;
    }
    if (threadIdx.x == 255) lv_tot[blockIdx.x] = run;
}

// Cells in level-major order, Morton order inside a level (the order the
// reference's level-by-level build gives): particle i starts the cells of
// levels d[i]+1 .. leaf_lvl[i], each with first particle i, the last a leaf.
// A level-L cell goes to lev_off[L] + (cells of level L started by earlier
// blocks, from k_lv_scan) + (by earlier particles of this block, from wave
// ballots); a leaf's cell index goes to leaf_idx in the same way (leaves in
// cell order).  The blocks are k_split's.  Slots at or beyond cap are dropped
// (the host sees the cell count and emits again into grown buffers); block 0
// writes the root and the cell and leaf totals.
__global__ void __launch_bounds__(kSplitBlock)
    k_emit(const uint8_t *d, const uint8_t *leaf_lvl, int n, const uint32_t *blk_lv, int nbs,
           const uint32_t *lv_tot, uint32_t *lev_level, uint32_t *lev_begin, uint32_t *leaf_idx,
           uint32_t cap, DevHdr *hdr_w) {
    constexpr int kW = kSplitBlock / kWave;
    constexpr int kL = kMaxDepth + 1;  // row offset of the leaf counts
    __shared__ uint32_t s_base[kMaxDepth + 1], s_lbase[kMaxDepth + 1];
    __shared__ uint32_t s_w[kMaxDepth + 1][kW], s_lw[kMaxDepth + 1][kW];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
    if (tid <= kMaxDepth) {
        uint32_t off = 1, loff = 0;  // the root is cell 0
        for (int l = 1; l < tid; ++l) {
            off += lv_tot[l];
            loff += lv_tot[kL + l];
        }
        s_base[tid] = off + blk_lv[(size_t)tid * nbs + blockIdx.x];
        s_lbase[tid] = loff + blk_lv[(size_t)(kL + tid) * nbs + blockIdx.x];
    }
    if (blockIdx.x == 0 && tid == 0) {
        lev_level[0] = 0;
        lev_begin[0] = 0;
        uint32_t tot = 0, ltot = 0;
        for (int l = 1; l <= kMaxDepth; ++l) {
            tot += lv_tot[l];
            ltot += lv_tot[kL + l];
        }
        hdr_w->n_cells_pre = tot;
        hdr_w->n_leaves = ltot;
    }
    const int i = blockIdx.x * kSplitBlock + tid;
    const int di = i < n ? (int)d[i] : 0, lv = i < n ? (int)leaf_lvl[i] : 0;
    const uint64_t below = lane ? (~0ull >> (kWave - lane)) : 0ull;
    // per-wave counts start at zero: a wave visits only the levels its own
    // particles start (a few of the 16)
    if (tid < (kMaxDepth + 1) * kW) {
        (&s_w[0][0])[tid] = 0;
        (&s_lw[0][0])[tid] = 0;
    }
    int lo = lv > di ? di + 1 : kMaxDepth + 1, hi = lv > di ? lv : 0;
    for (int m = 1; m < kWave; m <<= 1) {
        lo = min(lo, __shfl_xor(lo, m, kWave));
        hi = max(hi, __shfl_xor(hi, m, kWave));
    }
    lo = __builtin_amdgcn_readfirstlane(lo);
    hi = __builtin_amdgcn_readfirstlane(hi);
    __syncthreads();
    for (int L = lo; L <= hi; ++L) {
        const uint64_t m = __ballot(di < L && L <= lv);
        const uint64_t ml = __ballot(di < L && L == lv);
        if (lane == 0) {
            s_w[L][w] = (uint32_t)__popcll(m);
            s_lw[L][w] = (uint32_t)__popcll(ml);
        }
    }
    __syncthreads();
    for (int L = lo; L <= hi; ++L) {
        const bool mine = di < L && L <= lv;
        const uint64_t m = __ballot(mine);
        const uint64_t ml = __ballot(mine && L == lv);
        if (mine) {
            uint32_t pos = s_base[L] + (uint32_t)__popcll(m & below);
            for (int u = 0; u < w; ++u) pos += s_w[L][u];
            if (pos < cap) {
                lev_level[pos] = (uint32_t)L;
                lev_begin[pos] = (uint32_t)i;
                if (L == lv) {  // the leaf over particle i
                    uint32_t r = s_lbase[L] + (uint32_t)__popcll(ml & below);
                    for (int u = 0; u < w; ++u) r += s_lw[L][u];
                    leaf_idx[r] = pos;
                }
            }
        }
    }
}

// End of the particle range that starts at b and shares the key prefix pre
// (key >> sh): the first index after b whose prefix differs, or n.  The 64
// lanes of a wavefront search together, one load per lane per round: a
// 64-way gallop from b (strides 1, 64, 4096, ...) brackets the end, then
// 64-way refinement narrows the bracket to one index, about 2 log64 of the
// range in dependent loads.  b, sh and pre are wavefront-uniform and every
// lane of the wavefront is active.
__device__ __forceinline__ uint32_t range_end_wave(const uint32_t *key, int n, uint32_t b, int sh,
                                                   uint32_t pre, int lane) {
    // invariant: key[lo] has the prefix; the first index without it is <= hi
    long long lo = b, hi, stride = 1;
    for (;;) {
        const long long pos = lo + stride * (lane + 1);
        bool out = pos >= n;
        if (!out) out = (key[pos] >> sh) != pre;
        const uint64_t m = __ballot(out);
        if (m) {
            const int f = __ffsll((unsigned long long)m) - 1;
            hi = min((long long)n, lo + stride * (f + 1));
            lo += stride * f;
            break;
        }
        lo += stride * kWave;
        stride *= kWave;
    }
    while (hi - lo > 1) {
        const long long st = (hi - lo + kWave - 1) / kWave;  // lane 63 reaches hi
        const long long pos = lo + st * (lane + 1);
        bool out = pos >= hi;
        if (!out) out = (key[pos] >> sh) != pre;
        const uint64_t m = __ballot(out);
        const int f = __ffsll((unsigned long long)m) - 1;
        hi = min(hi, lo + st * (f + 1));
        lo += st * f;
    }
    return (uint32_t)hi;
}

// Cell attributes: range, counts, flags, index, centre (treeCell.hpp:68) of
// cell j at level L with particle range [b, e) and level prefix pre.
__device__ __forceinline__ void cell_attrs(
    int j, int L, uint32_t b, uint32_t e, uint32_t pre, const uint32_t *A,
    const uint8_t *leaf_lvl, int n, const DevHdr *hdr, uint8_t *c_level, uint8_t *c_flags,
    uint32_t *c_key, uint32_t *c_begin, uint32_t *c_end, uint32_t *c_src, double *c_cx,
    double *c_cy, int *c_child) {
#pragma clang fp contract(off)
    // every particle active: the scan left A unwritten (the identity)
    const uint32_t src = hdr->n_active == (uint32_t)n ? e - b : A[e] - A[b];
    unsigned f = 0;
    if (L > 0 && (int)leaf_lvl[b] == L) f |= kLeaf;
    if (src > 0) f |= kActive;
    const uint32_t ix = compact1by1(pre), iy = compact1by1(pre >> 1);
    const double R = hdr->root_len;
    const double len = cell_len(R, L);
    c_level[j] = (uint8_t)L;
    c_flags[j] = (uint8_t)f;
    c_key[j] = pre;
    c_begin[j] = b;
    c_end[j] = e;
    c_src[j] = src;
    c_cx[j] = hdr->pivot[0] + ((0.5 + (double)ix) * len);
    c_cy[j] = hdr->pivot[1] + ((0.5 + (double)iy) * len);
    reinterpret_cast<int4 *>(c_child)[j] = make_int4(-1, -1, -1, -1);  // filled by k_cells2
}

constexpr int kCells1Block = 256;

__global__ void __launch_bounds__(kCells1Block)
    k_cells1(const uint32_t *lev_level, const uint32_t *lev_begin, int nc, const uint32_t *key,
             const uint32_t *A, const uint8_t *leaf_lvl, int n, const DevHdr *hdr,
             uint8_t *c_level, uint8_t *c_flags, uint32_t *c_key, uint32_t *c_begin,
             uint32_t *c_end, uint32_t *c_src, double *c_cx, double *c_cy, int *c_child,
             SplitLevels sl, DevHdr *hdr_w) {
    // one wavefront per cell (j is wavefront-uniform)
    const int j = blockIdx.x * (kCells1Block / kWave) + (int)(threadIdx.x / kWave);
    const int lane = (int)(threadIdx.x & (kWave - 1));
    // level offsets / counts and the leaf-level range for the later kernels
    if (blockIdx.x == 0 && threadIdx.x < kMaxDepth + 2) {
        hdr_w->lev_off[threadIdx.x] = sl.lev_off[threadIdx.x];
        hdr_w->lev_cnt[threadIdx.x] = sl.lev_cnt[threadIdx.x];
        if (threadIdx.x == 0) {
            hdr_w->leaf_lvl_min = sl.leaf_lvl_min;
            hdr_w->leaf_lvl_max = sl.leaf_lvl_max;
        }
    }
    if (j >= nc) return;
    const int L = (int)lev_level[j];
    const uint32_t b = lev_begin[j];
    uint32_t e = (uint32_t)n, pre = 0;
    if (L > 0) {
        const int sh = 32 - 2 * L;
        pre = key[b] >> sh;
        e = range_end_wave(key, n, b, sh, pre, lane);
    }
    if (lane == 0)
        cell_attrs(j, L, b, e, pre, A, leaf_lvl, n, hdr, c_level, c_flags, c_key, c_begin, c_end,
                   c_src, c_cx, c_cy, c_child);
}

__device__ __forceinline__ int find_cell(const uint32_t *c_key, const uint32_t *lev_off, int level,
                                         uint32_t pre) {
    int lo = (int)lev_off[level], hi = (int)lev_off[level + 1];
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        uint32_t k = c_key[mid];
        if (k < pre) lo = mid + 1;
        else hi = mid;
    }
    return (lo < (int)lev_off[level + 1] && c_key[lo] == pre) ? lo : -1;
}

// Parent links (binary search one level up) and child slots (reference child
// order: i = (ix & 1) + 2 (iy & 1), treeCell.cpp:407-414).
__global__ void __launch_bounds__(256)
    k_cells2(const uint8_t *c_level, const uint32_t *c_key, int nc, const DevHdr *hdr,
             int *c_parent, int *c_child) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nc) return;
    const int L = c_level[j];
    if (L == 0) {
        c_parent[j] = -1;
        return;
    }
    const uint32_t pre = c_key[j];
    const int par = L == 1 ? 0 : find_cell(c_key, hdr->lev_off, L - 1, pre >> 2);
    c_parent[j] = par;
    const int qx = (int)(compact1by1(pre) & 1u), qy = (int)(compact1by1(pre >> 1) & 1u);
    if (par >= 0) c_child[4 * par + qx + 2 * qy] = j;
}

// Type-1 box bounds of every cell (hierarchy_mesh, subroutine_all_fmm_2d.cpp:
// 179-182, 229-240): from the root (pivot, side root_len = 2 db) down the
// cell's path with the reference's additions, in the order the reference
// forms them (t1_path); the centre (max + min) / 2 (:201-202, 258-259)
// replaces the pivot-based centre of cell_attrs.
__global__ void __launch_bounds__(256)
    k_t1_bounds(const uint8_t *c_level, const uint32_t *c_key, int nc, const DevHdr *hdr,
                double4 *bnd, double *c_cx, double *c_cy) {
#pragma clang fp contract(off)
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nc) return;
    const int L = c_level[j];
    const double db = hdr->root_len / 2.0;
    double lo[2] = {hdr->pivot[0], hdr->pivot[1]}, hi[2] = {lo[0] + 2.0 * db, lo[1] + 2.0 * db};
    if (L > 0) {
        const uint32_t pre = c_key[j];
        const uint32_t ix[2] = {compact1by1(pre), compact1by1(pre >> 1)};
        for (int d = 0; d < 2; ++d) {
            const double o = lo[d];
            uint32_t b = (ix[d] >> (L - 1)) & 1u;
            double l = o + (double)b * db, h = o + (double)(b + 1) * db;
            for (int lv = 2; lv <= L; ++lv) {
                b = (ix[d] >> (L - lv)) & 1u;
                const double dd = (h - l) / 2.0;
                const double nl = l + (double)b * dd, nh = l + (double)(b + 1) * dd;
                l = nl;
                h = nh;
            }
            lo[d] = l;
            hi[d] = h;
        }
    }
    bnd[j] = make_double4(lo[0], hi[0], lo[1], hi[1]);
    c_cx[j] = (hi[0] + lo[0]) / 2.0;
    c_cy[j] = (hi[1] + lo[1]) / 2.0;
}

// ---------------------------------------------------------------------------
// 6. neighbour table and lists
// ---------------------------------------------------------------------------
// The box tests of the type-1 lists (list_one .. list_four,
// subroutine_all_fmm_2d.cpp:291-556) on bounds (xmin, xmax, ymin, ymax),
// against the absolute tolerance Pars::tol2 = 1e-15 (global.cpp:325):
// a touches b when an edge of one lies within tol2 of an edge of the other
// and the ranges overlap (within tol2) in the other direction; a is well
// separated from b when a gap above tol2 parts them in some direction.
constexpr double kT1Tol = 1.0e-15;

__device__ __forceinline__ bool t1_touch(const double4 a, const double4 b) {
#pragma clang fp contract(off)
    const double dx_xn = fabs(a.y - b.x), dx_nx = fabs(a.x - b.y);
    const double dy_xn = fabs(a.w - b.z), dy_nx = fabs(a.z - b.w);
    return (((dy_xn <= kT1Tol) || (dy_nx <= kT1Tol)) && (a.x <= b.y + kT1Tol) &&
            (a.y >= b.x - kT1Tol)) ||
           (((dx_xn <= kT1Tol) || (dx_nx <= kT1Tol)) && (a.z <= b.w + kT1Tol) &&
            (a.w >= b.z - kT1Tol));
}
__device__ __forceinline__ bool t1_sep(const double4 a, const double4 b) {
#pragma clang fp contract(off)
    return (a.x > b.y + kT1Tol) || (a.y < b.x - kT1Tol) || (a.z > b.w + kT1Tol) ||
           (a.w < b.z - kT1Tol);
}

// nbr[9j + k], k = (dx+1) + 3 (dy+1) (treeCell.cpp:593-622 order): the
// deepest existing cell at level <= L containing the level-L cell at
// (ix+dx, iy+dy), or -1 outside [0, 2^L)^2.  One thread per (cell, k): up the
// parent links to the lowest common ancestor of the cell and that position,
// then down the child slots toward the position while the child exists (the
// cells containing a position form one chain from the root, so the walk
// ends at the deepest of them).
__global__ void __launch_bounds__(256)
    k_nbr(const uint8_t *c_level, const uint32_t *c_key, const int *c_parent, const int *c_child,
          int nc, int *nbr) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 9LL * nc) return;
    const int j = (int)(idx / 9), k = (int)(idx % 9);
    const int L = c_level[j];
    if (L == 0) {
        nbr[9 * j + k] = k == 4 ? 0 : -1;
        return;
    }
    if (k == 4) {
        nbr[9 * j + k] = j;
        return;
    }
    const uint32_t pre = c_key[j];
    const int ix = (int)compact1by1(pre), iy = (int)compact1by1(pre >> 1);
    const int cnt = 1 << L;
    const int X = ix + (k % 3) - 1, Y = iy + (k / 3) - 1;
    int res = -1;
    if (X >= 0 && X < cnt && Y >= 0 && Y < cnt) {
        int up = 1;  // (X, Y) != (ix, iy), and both are 0 at level 0
        while ((ix >> up) != (X >> up) || (iy >> up) != (Y >> up)) ++up;
        int cur = j;
        for (int u = 0; u < up && cur >= 0; ++u) cur = c_parent[cur];
        for (int sh = up - 1; sh >= 0 && cur >= 0; --sh) {
            const int ch = c_child[4 * cur + ((X >> sh) & 1) + 2 * ((Y >> sh) & 1)];
            if (ch < 0) break;
            cur = ch;
        }
        res = cur;
    }
    nbr[9 * j + k] = res;
}

// Lists 2 and 4 of every cell of level >= 2 (treeCell.cpp:818-1032).  Eight
// lanes per cell; lane u examines the parent's neighbour k = u (u < 4) or
// u + 1 (the parent itself, k = 4, is skipped as at treeCell.cpp:872):
//  * a non-leaf neighbour at the parent's level contributes its active
//    children that are well separated from the cell (list 2);
//  * an active leaf neighbour is a list-4 candidate, kept at its first
//    occurrence among the eight lanes and if separated from the cell.
// The lanes' entries are compacted with a prefix sum over the eight lanes, in
// lane order and child order within a lane.
// With T1 the candidates are the same (the cells a type-1 box test can pass
// lie within one cell of the parent) and the reference's type-1 tests decide
// (list_two, list_four): a non-leaf neighbour contributes only if it touches
// the parent, its children only if well separated from the cell; a list-4
// leaf must touch the parent and be well separated from the cell.
constexpr int kM2lLanes = 8;

template <bool T1>
__global__ void __launch_bounds__(256)
    k_m2l_lists(const uint8_t *c_level, const uint8_t *c_flags, const double *c_cx,
                const double *c_cy, const int *c_parent, const int *c_child, const int *nbr,
                const double4 *bnd, const DevHdr *hdr, int j0, int nc, int *l2, uint8_t *n2,
                int *l4, uint8_t *n4, DevHdr *hdr_w) {
#pragma clang fp contract(off)
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = j0 + (int)(gid / kM2lLanes);
    const int u = (int)(gid % kM2lLanes);
    const int k = u < 4 ? u : u + 1;
    const bool live = j < nc;
    int ch[4] = {-1, -1, -1, -1};
    bool keep[4] = {false, false, false, false};
    int cand = -1;
    double cx = 0.0, cy = 0.0, s = 0.0, R = 0.0;
    if (live) {
        const int L = c_level[j];
        R = hdr->root_len;
        s = cell_len(R, L);
        cx = c_cx[j];
        cy = c_cy[j];
        const int A = nbr[9 * c_parent[j] + k];
        if (A >= 0) {
            const unsigned fa = c_flags[A];
            if (fa & kActive) {
                if (c_level[A] == L - 1 && !(fa & kLeaf)) {
                    const double md2 = s + kListEps;
                    const bool adj = !T1 || t1_touch(bnd[A], bnd[c_parent[j]]);
#pragma unroll
                    for (int q = 0; q < 4; ++q) ch[q] = adj ? c_child[4 * A + q] : -1;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        if (ch[q] < 0 || !(c_flags[ch[q]] & kActive)) continue;
                        keep[q] = T1 ? t1_sep(bnd[ch[q]], bnd[j])
                                     : fabs(c_cx[ch[q]] - cx) > md2 || fabs(c_cy[ch[q]] - cy) > md2;
                    }
                } else if (fa & kLeaf) {
                    cand = A;
                }
            }
        }
    }
    // first occurrence of a list-4 candidate among the cell's lanes
    bool dup = false;
#pragma unroll
    for (int v = 0; v < kM2lLanes - 1; ++v) {
        const int o = __shfl(cand, v, kM2lLanes);
        dup |= v < u && o == cand;
    }
    bool sep4 = false;
    if (cand >= 0 && !dup) {
        if (T1) {
            sep4 = t1_touch(bnd[cand], bnd[c_parent[j]]) && t1_sep(bnd[cand], bnd[j]);
        } else {
            const double md = ((cell_len(R, c_level[cand]) + s) / 2.0) + kListEps;
            sep4 = fabs(c_cx[cand] - cx) > md || fabs(c_cy[cand] - cy) > md;
        }
    }
    const int m2 = (int)keep[0] + (int)keep[1] + (int)keep[2] + (int)keep[3];
    const int m4 = sep4 ? 1 : 0;
    int inc2 = m2, inc4 = m4;
#pragma unroll
    for (int o = 1; o < kM2lLanes; o <<= 1) {
        const int y2 = __shfl_up(inc2, o, kM2lLanes), y4 = __shfl_up(inc4, o, kM2lLanes);
        if (u >= o) {
            inc2 += y2;
            inc4 += y4;
        }
    }
    const int tot2 = __shfl(inc2, kM2lLanes - 1, kM2lLanes);
    const int tot4 = __shfl(inc4, kM2lLanes - 1, kM2lLanes);
    if (live) {
        int o2 = inc2 - m2;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (!keep[q]) continue;
            if (o2 < 27) l2[27 * (size_t)j + o2] = ch[q];
            ++o2;
        }
        if (sep4) l4[8 * (size_t)j + (inc4 - m4)] = cand;
        if (u == 0) {
            if (tot2 > 27) atomicOr(&hdr_w->flags, kFlagListOverflow);
            n2[j] = (uint8_t)min(tot2, 27);
            n4[j] = (uint8_t)tot4;
        }
    }
    block_stats(hdr_w, kStSum2, 2, (unsigned long long)(live ? m2 : 0),
                (unsigned long long)(live ? m4 : 0));
}

// Lists 1 and 3 of every leaf (treeCell.cpp:636-804), count or fill pass.
// Sixteen lanes per leaf; lane k < 9 handles the leaf's neighbour k:
//  * a coarser neighbour (reached by the parent walk, treeCell.cpp:706-720,
//    738) joins list 1 if it is an active leaf, at its first occurrence among
//    the nine lanes;
//  * an active leaf at the leaf's level joins list 1;
//  * an active non-leaf is walked depth first: children adjacent to the leaf
//    join list 1 (leaves) or are walked (non-leaves), the others join list 3.
// Each lane's walk only visits its own neighbour's subtree.  Along an edge a
// level adds at most two adjacent children and the walk pops one, so the
// per-lane stack (in LDS) grows by at most one entry per level.
// The count pass stores each lane's counts (cntk) and the leaf totals; the
// fill pass writes at the leaf offset plus the prefix of the lanes' counts and
// also counts the P2P pairs and M2P / L2P terms the lists imply (statistics),
// reduced per block into per-block partials.
// With T1 the walk visits the same cells (every cell a type-1 box test can
// pass is adjacent to the leaf or a child of an adjacent cell) and the
// reference's type-1 tests decide (list_one, list_three): a leaf joins list
// 1 if it touches the leaf (the leaf itself always, first); the children of a
// visited non-leaf that touches the leaf join list 3 if well separated from
// it.  Adjacent non-leaf children are walked whatever the tests say, since
// the reference tests every cell on its own.
constexpr int kLeafLanes = 16;
constexpr int kWalkDepth = 24;

template <bool FILL, bool T1>
__global__ void __launch_bounds__(256)
    k_leaf_lists(const uint32_t *leaf_idx, int nl, const uint8_t *c_level, const uint8_t *c_flags,
                 const double *c_cx, const double *c_cy, const int *c_child, const int *nbr,
                 const double4 *bnd, const DevHdr *hdr, const uint8_t *orders, uint32_t *cnt1,
                 uint32_t *cnt3,
                 uint2 *cntk, const uint32_t *off1, const uint32_t *off3, int *l1, int *l3,
                 uint32_t cap1, uint32_t cap3, uint32_t *bs1, uint32_t *bs3,
                 const uint32_t *c_begin, const uint32_t *c_end, DevHdr *hdr_w) {
#pragma clang fp contract(off)
    __shared__ int stk[kWalkDepth][256];
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int t = (int)(gid / kLeafLanes);
    const int k = (int)(gid % kLeafLanes);
    const bool leaf_live = t < nl;
    const bool live = leaf_live && k < 9;
    unsigned long long my_pairs = 0, my_m2p = 0, my_m2p_terms = 0, my_l2p_terms = 0;
    bool overflow = false;
    int j = 0, L = 0;
    double R = 0.0, s = 0.0, cx = 0.0, cy = 0.0;
    int A = -1, cand = -1;
    unsigned fa = 0;
    if (live) {
        j = (int)leaf_idx[t];
        L = c_level[j];
        R = hdr->root_len;
        s = cell_len(R, L);
        cx = c_cx[j];
        cy = c_cy[j];
        A = nbr[9 * j + k];
        if (A >= 0) {
            fa = c_flags[A];
            if (c_level[A] < L && (fa & kLeaf) && (fa & kActive)) cand = A;
        }
    }
    bool dup = false;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
        const int o = __shfl(cand, v, kLeafLanes);
        dup |= v < k && o == cand;
    }
    uint32_t o1 = 0, o3 = 0;
    if (FILL) {
        const uint2 mine = live ? cntk[9 * (size_t)t + k] : make_uint2(0u, 0u);
        uint32_t i1 = mine.x, i3 = mine.y;
#pragma unroll
        for (int o = 1; o < kLeafLanes; o <<= 1) {
            const uint32_t y1 = __shfl_up(i1, o, kLeafLanes), y3 = __shfl_up(i3, o, kLeafLanes);
            if (k >= o) {
                i1 += y1;
                i3 += y3;
            }
        }
        if (leaf_live) {
            o1 = off1[t] + i1 - mine.x;
            o3 = off3[t] + i3 - mine.y;
        }
    }
    uint32_t n1 = 0, n3 = 0;
    unsigned long long src_total = 0, m2p_w = 0;
    bool self = false;
    auto emit1 = [&](int c) {
        if (FILL) {
            if (o1 + n1 < cap1) l1[o1 + n1] = c;
            else overflow = true;
            src_total += c_end[c] - c_begin[c];
            self |= c == j;
        }
        ++n1;
    };
    auto emit3 = [&](int c) {
        if (FILL) {
            if (o3 + n3 < cap3) l3[o3 + n3] = c;
            else overflow = true;
            m2p_w += (unsigned long long)orders[c_level[c]] + 1ull;
        }
        ++n3;
    };
    const double4 bj = T1 && live ? bnd[j] : make_double4(0.0, 0.0, 0.0, 0.0);
    if (live && A >= 0) {
        if (c_level[A] < L) {
            if (cand >= 0 && !dup && (!T1 || t1_touch(bnd[A], bj))) emit1(A);
        } else if (fa & kActive) {
            if (fa & kLeaf) {
                if (!T1 || A == j || t1_touch(bnd[A], bj)) emit1(A);
            } else {
                const int me = threadIdx.x;
                int sp = 0;
                stk[sp++][me] = A;
                while (sp > 0) {
                    const int B = stk[--sp][me];
                    const double adj = ((s + (cell_len(R, c_level[B]) / 2.0)) / 2.0) + kListEps;
                    const bool tb = !T1 || t1_touch(bnd[B], bj);
                    int chq[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) chq[q] = c_child[4 * B + q];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int ch = chq[q];
                        if (ch < 0) continue;
                        const unsigned fc = c_flags[ch];
                        if (!(fc & kActive)) continue;
                        const bool is_adj =
                            fabs(c_cx[ch] - cx) <= adj && fabs(c_cy[ch] - cy) <= adj;
                        if (is_adj) {
                            if (fc & kLeaf) {
                                if (!T1 || t1_touch(bnd[ch], bj)) emit1(ch);
                            } else if (sp < kWalkDepth) {
                                stk[sp++][me] = ch;
                            } else {
                                overflow = true;
                            }
                            // bounds rounded apart by more than tol2
                            if (T1 && tb && t1_sep(bnd[ch], bj)) emit3(ch);
                        } else if (!T1 || (tb && t1_sep(bnd[ch], bj))) {
                            emit3(ch);
                        }
                    }
                }
            }
        }
    }
    if (!FILL) {
        // per-block sums of the list sizes (16 leaves per block) for the scans
        block_sum_store(live ? n1 : 0u, bs1 + blockIdx.x);
        __syncthreads();  // block_sum_store's LDS is reused
        block_sum_store(live ? n3 : 0u, bs3 + blockIdx.x);
        uint32_t s1 = n1, s3 = n3;
#pragma unroll
        for (int o = 1; o < kLeafLanes; o <<= 1) {
            s1 += __shfl_xor(s1, o, kLeafLanes);
            s3 += __shfl_xor(s3, o, kLeafLanes);
        }
        if (live) cntk[9 * (size_t)t + k] = make_uint2(n1, n3);
        if (leaf_live && k == 0) {
            cnt1[t] = s1;
            cnt3[t] = s3;
        }
    } else if (live) {
        const unsigned long long nt = c_end[j] - c_begin[j];
        my_pairs = src_total * nt - (self ? nt : 0ull);
        my_m2p = (unsigned long long)n3 * nt;
        my_m2p_terms = m2p_w * nt;
        if (k == 4) my_l2p_terms = (unsigned long long)orders[L] * nt;
    }
    if (overflow) atomicOr(&hdr_w->flags, kFlagListOverflow);
    if (FILL) {
        block_stats(hdr_w, kStPairs, 4, my_pairs, my_m2p, my_m2p_terms, my_l2p_terms);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------
void stream_mark(Ctx &c) {
    if (!c.ev_sync) VPM_HIP(hipEventCreateWithFlags(&c.ev_sync, hipEventDisableTiming));
    VPM_HIP(hipEventRecord(c.ev_sync, c.stream));
}

void stream_mark_wait(Ctx &c) {
    for (;;) {
        const hipError_t e = hipEventQuery(c.ev_sync);
        if (e == hipSuccess) return;
        if (e != hipErrorNotReady) VPM_HIP(e);
    }
}

void stream_wait(Ctx &c) {
    stream_mark(c);
    stream_mark_wait(c);
}

static void check_flags(const DevHdr &h) {
    if (h.flags & kFlagNonFinite) throw Error(kNonFinite, "non-finite particle coordinate");
    if (h.flags & kFlagRange) throw Error(kDegenerate, "all particles at one point (root length 0)");
    if (h.flags & kFlagT1Lost)
        throw Error(kOutsideBoxes,
                    "a particle lies outside every box of the type-1 hierarchy (on the upper "
                    "bound of the level-1 boxes or between a split box's rounded child bounds); "
                    "the reference drops such a particle");
    if (h.flags & kFlagDepth)
        throw Error(kDepth, "tree deeper than 16 levels (more coincident particles than the leaf thresholds)");
    if (h.flags & kFlagListOverflow) throw Error(kInternal, "list traversal stack overflow");
}

// Key bits sorted for a tree of the given depth: the levels down to the
// deepest leaf, in whole 8-bit radix passes.
static int canonical_sort_bits(int depth) { return std::max(8, 8 * div_up(2LL * depth, 8)); }

void tree_build(Ctx &c, int n, const double *x, const double *y, const Sources &src,
                const uint8_t *active, const TreeParams &tp) {
    hipStream_t s = c.stream;
    const int W = tp.max_par, S = tp.max_src;

    // 1. bounding box (the first kernel also clears the header)
    const int nbb = std::min(1024, div_up(n, 256 * 8));
    double *part = c.bbox_part.as<double>(5 * (size_t)nbb);
    k_bbox_partial<<<nbb, 256, 0, s>>>(x, y, n, part, c.hdr_d);
    const double expansion = 1.0 + (tp.exp_tree / 100.0);
    VPM_LAUNCH_CHECK();

    // 2-3. keys, sort, gather
    uint32_t *ka = c.key_a.as<uint32_t>(n), *kb = c.key_b.as<uint32_t>(n);
    uint32_t *va = c.val_a.as<uint32_t>(n), *vb = c.val_b.as<uint32_t>(n);
    const size_t tmp_words =
        std::max(radix_tmp_words(n), scan_tmp_words((long long)n + 1) + 64) + 64;
    uint32_t *tmp = c.tmp.as<uint32_t>(tmp_words);
    // the top sort_bits of the key (levels <= sort_bits / 2 resolved), a guess
    // from the previous build checked after the split
    const int bits = c.sort_bits;
    const int tiles = radix_tiles(n);
    c.t1 = tp.type1;
    (tp.type1 ? k_keys<true> : k_keys<false>)<<<tiles, kRsBlock, 0, s>>>(
        x, y, n, part, nbb, expansion, tp.max_level, c.hdr_d, ka, va, 32 - bits, tmp, tiles);
    const bool alt = radix_sort_pairs(ka, va, kb, vb, n, 32 - bits, 32, tmp, s, true);
    uint32_t *key = alt ? kb : ka;
    uint32_t *perm = alt ? vb : va;
    c.key_sorted = key;
    c.perm_sorted = perm;
    double *xs = c.xs.as<double>(n), *ys = c.ys.as<double>(n), *qs = c.qs.as<double>(n);
    uint8_t *as = c.acts.as<uint8_t>(n);
    uint32_t *bs_act = c.bs_act.as<uint32_t>((size_t)div_up(n, 256) + 1);
    uint8_t *d = c.lcp_d.as<uint8_t>(n + 1), *w = c.lcp_w.as<uint8_t>(n + 1);
    double *ss = (c.icut > 0 && src.s) ? c.ss.as<double>(n) : nullptr;
    k_gather<<<div_up(n, 256), 256, 0, s>>>(perm, x, y, src.q, src.s, src.omega, active, n, xs,
                                            ys, qs, as, bs_act, key, W, d, w, ss, src.qscale,
                                            c.hdr_d);

    // 4. leaf levels, cells per level and leaves
    uint32_t *A = c.act_scan.as<uint32_t>(n + 1);
    uint32_t *apos = c.act_pos.as<uint32_t>(n);
    // (every particle active: A and apos are the identity and stay unwritten)
    scan_exclusive_bs_u8(as, A, apos, n, bs_act, 256, &c.hdr_d->n_active, true, tmp, s,
                         &c.hdr_d->stat[0][kStInactive], kStatRows, kStCols);
    uint8_t *leaf_lvl = c.leaf_lvl.as<uint8_t>(n);
    const size_t lds = 3 * (size_t)(W + kSplitBlock) + 3 * (size_t)(S + kSplitBlock + 2) + 16;
    const int nbs = div_up(n, kSplitBlock);
    uint32_t *blk_lv = c.lv_blk.as<uint32_t>((size_t)2 * (kMaxDepth + 1) * nbs);
    uint32_t *lv_tot = c.lv_tot.as<uint32_t>(2 * (kMaxDepth + 1));
    k_split<<<nbs, kSplitBlock, lds, s>>>(key, d, w, A, apos, c.hdr_d, n, W, S, tp.min_level,
                                          tp.max_level, leaf_lvl, c.hdr_d, blk_lv, nbs);
    // The one synchronisation of the build: the header (split rows) is read
    // back now; the per-level scan and the cell emission need no host
    // decision and run while the host waits for the copy and decides.
    VPM_HIP(hipMemcpyAsync(c.hdr_h, c.hdr_d, sizeof(DevHdr), hipMemcpyDeviceToHost, s));
    stream_mark(c);
    k_lv_scan<<<2 * (kMaxDepth + 1), 256, 0, s>>>(blk_lv, nbs, lv_tot);
    // emission into buffers sized from the previous build (dropped beyond)
    const size_t cap = std::max(c.cell_cap, (size_t)n / 4 + 1024);
    uint32_t *lev_level = c.lev_level.as<uint32_t>(cap), *lev_begin = c.lev_begin.as<uint32_t>(cap);
    uint32_t *leaf_idx = c.leaf_idx.as<uint32_t>(cap);
    k_emit<<<nbs, kSplitBlock, 0, s>>>(d, leaf_lvl, n, blk_lv, nbs, lv_tot, lev_level, lev_begin,
                                       leaf_idx, (uint32_t)cap, c.hdr_d);
    VPM_LAUNCH_CHECK();
    stream_mark_wait(c);
    DevHdr &hh = *c.hdr_h;
    if (hh.flags & (kFlagNonFinite | kFlagRange | kFlagT1Lost)) check_flags(hh);
    // the split rows: leaf-level range, depth overflow, cells per level, leaves
    SplitLevels sl{};
    {
        uint32_t col[kLvCols] = {}, min_inv = 0, lmax = 0, depth = 0;
        for (int r = 0; r < kSlotRows; ++r) {
            const uint32_t *row = hh.split_rows[r];
            for (int cc = 0; cc < kLvCols; ++cc) col[cc] += row[cc];
            min_inv = std::max(min_inv, row[kSlotMinInv]);
            lmax = std::max(lmax, row[kSlotMax]);
            depth |= row[kSlotDepth];
        }
        if (depth) hh.flags |= kFlagDepth;
        hh.leaf_lvl_min = sl.leaf_lvl_min = 31u - min_inv;
        hh.leaf_lvl_max = sl.leaf_lvl_max = lmax;
        uint32_t off = 1;  // the root is cell 0
        sl.lev_off[0] = 0;
        sl.lev_cnt[0] = 1;
        for (int L = 1; L <= kMaxDepth + 1; ++L) {
            const uint32_t cnt_l = L <= kMaxDepth ? col[L] : 0u;
            sl.lev_off[L] = off;
            sl.lev_cnt[L] = cnt_l;
            off += cnt_l;
        }
        hh.n_leaves_split = col[kLvCols - 1];
        for (int L = 0; L < kMaxDepth + 2; ++L) hh.lev_off[L] = sl.lev_off[L];
    }
    // A split on the sorted top bits is exact while its deepest leaf level is
    // at most bits / 2: a window's lcp is taken from its two end keys, which
    // bounds the whole window only where the keys are sorted, so every lcp
    // below bits / 2 is exact and one at or beyond it (possibly overstated,
    // up to a false depth overflow) puts a leaf deeper than bits / 2.  Such a
    // build is repeated with the full key, whose depth is exact.  The sorted
    // bit count is a function of the depth alone (canonical_sort_bits), so the
    // particle order inside a leaf, and with it every rounding, depends on the
    // input only; a wrong guess builds again (at most twice).
    {
        const int depth = (int)hh.leaf_lvl_max;
        const bool partial_too_shallow =
            bits < 32 && ((hh.flags & kFlagDepth) || 2 * depth > bits);
        const int want = partial_too_shallow ? 32 : canonical_sort_bits(depth);
        if (!partial_too_shallow) check_flags(hh);  // a depth overflow of the full key is real
        if (want != bits) {
            c.sort_bits = want;
            tree_build(c, n, x, y, src, active, tp);
            return;
        }
    }

    // 5. cells (the cell count from the split rows; the scan's own total is
    // checked against it at the end of the solve)
    const int nc = (int)sl.lev_off[kMaxDepth + 1];
    TreeInfo &ti = c.info;
    ti.n = n;
    ti.n_active = (int)hh.n_active;
    ti.n_cells = nc;
    ti.max_level = (int)hh.leaf_lvl_max;
    ti.min_leaf_level = (int)hh.leaf_lvl_min;
    ti.root_len = hh.root_len;
    ti.pivot[0] = hh.pivot[0];
    ti.pivot[1] = hh.pivot[1];
    for (int l = 0; l < kMaxDepth + 2; ++l) ti.lev_off[l] = hh.lev_off[l];
    ti.n_leaves = (int)hh.n_leaves_split;
    ti.tot1 = ti.tot3 = 0;
    if (ti.n_leaves < 1 || ti.n_leaves > nc)
        throw Error(kInternal, "leaf count of the split out of range");
    if ((size_t)nc > cap) {
        // more cells than the emission had room for: let it finish, grow the
        // buffers and emit again
        VPM_HIP(hipStreamSynchronize(s));
        lev_level = c.lev_level.as<uint32_t>(nc);
        lev_begin = c.lev_begin.as<uint32_t>(nc);
        leaf_idx = c.leaf_idx.as<uint32_t>(nc);
        k_emit<<<nbs, kSplitBlock, 0, s>>>(d, leaf_lvl, n, blk_lv, nbs, lv_tot, lev_level,
                                           lev_begin, leaf_idx, (uint32_t)nc, c.hdr_d);
    }
    c.cell_cap = (size_t)nc + nc / 4;

    uint8_t *c_level = c.c_level.as<uint8_t>(nc), *c_flags = c.c_flags.as<uint8_t>(nc);
    uint32_t *c_key = c.c_key.as<uint32_t>(nc), *c_begin = c.c_begin.as<uint32_t>(nc);
    uint32_t *c_end = c.c_end.as<uint32_t>(nc), *c_src = c.c_src.as<uint32_t>(nc);
    double *c_cx = c.c_cx.as<double>(nc), *c_cy = c.c_cy.as<double>(nc);
    int *c_parent = c.c_parent.as<int>(nc), *c_child = c.c_child.as<int>(4 * (size_t)nc);
    k_cells1<<<div_up(nc, kCells1Block / kWave), kCells1Block, 0, s>>>(lev_level, lev_begin, nc, key, A, leaf_lvl, n,
                                             c.hdr_d, c_level, c_flags, c_key, c_begin, c_end,
                                             c_src, c_cx, c_cy, c_child, sl, c.hdr_d);
    k_cells2<<<div_up(nc, 256), 256, 0, s>>>(c_level, c_key, nc, c.hdr_d, c_parent, c_child);
    if (tp.type1)
        k_t1_bounds<<<div_up(nc, 256), 256, 0, s>>>(c_level, c_key, nc, c.hdr_d,
                                                    c.c_bnd.as<double4>(nc), c_cx, c_cy);
    VPM_LAUNCH_CHECK();
}

void tree_lists(Ctx &c) {
    tree_lists_begin(c);
    tree_lists_end(c);
}

void tree_lists_begin(Ctx &c) {
    hipStream_t s = c.stream;
    TreeInfo &ti = c.info;
    const int nc = ti.n_cells;
    if (!c.orders) throw Error(kInternal, "expansion orders not set before the list build");
    uint8_t *c_level = (uint8_t *)c.c_level.p, *c_flags = (uint8_t *)c.c_flags.p;
    uint32_t *c_key = (uint32_t *)c.c_key.p;
    double *c_cx = (double *)c.c_cx.p, *c_cy = (double *)c.c_cy.p;
    int *c_parent = (int *)c.c_parent.p, *c_child = (int *)c.c_child.p;
    int *nbr = c.c_nbr.as<int>(9 * (size_t)nc);
    k_nbr<<<div_up(9LL * nc, 256), 256, 0, s>>>(c_level, c_key, c_parent, c_child, nc, nbr);

    int *l2 = c.l2.as<int>(27 * (size_t)nc), *l4 = c.l4.as<int>(8 * (size_t)nc);
    uint8_t *n2 = c.n2.as<uint8_t>(nc), *n4 = c.n4.as<uint8_t>(nc);
    const int j0 = (int)ti.lev_off[2];
    // n2 / n4 of every cell of level >= 2 are written by k_m2l_lists; levels 0
    // and 1 have no lists 2 / 4 and nothing reads their entries
    const int nb2 = nc > j0 ? (int)div_up((long long)(nc - j0) * kM2lLanes, 256LL) : 0;
    const double4 *bnd = c.t1 ? (const double4 *)c.c_bnd.p : nullptr;
    if (nb2 > 0)
        (c.t1 ? k_m2l_lists<true> : k_m2l_lists<false>)<<<nb2, 256, 0, s>>>(
            c_level, c_flags, c_cx, c_cy, c_parent, c_child, nbr, bnd, c.hdr_d, j0, nc, l2, n2,
            l4, n4, c.hdr_d);
    VPM_LAUNCH_CHECK();
}

void tree_lists_end(Ctx &c) {
    hipStream_t s = c.stream;
    TreeInfo &ti = c.info;
    const int nc = ti.n_cells, nl = ti.n_leaves;
    uint8_t *c_level = (uint8_t *)c.c_level.p, *c_flags = (uint8_t *)c.c_flags.p;
    double *c_cx = (double *)c.c_cx.p, *c_cy = (double *)c.c_cy.p;
    int *c_child = (int *)c.c_child.p;
    const int *nbr = (const int *)c.c_nbr.p;
    uint32_t *tmp = (uint32_t *)c.tmp.p;
    const int nbl = (int)div_up((long long)nl * kLeafLanes, 256LL);
    const uint32_t *leaf_idx = (const uint32_t *)c.leaf_idx.p;
    uint32_t *cnt1 = c.cnt1.as<uint32_t>(nl + 1), *cnt3 = c.cnt3.as<uint32_t>(nl + 1);
    uint32_t *off1 = c.off1.as<uint32_t>(nl + 1), *off3 = c.off3.as<uint32_t>(nl + 1);
    const uint32_t *c_begin = (const uint32_t *)c.c_begin.p, *c_end = (const uint32_t *)c.c_end.p;
    uint2 *cntk = c.cntk.as<uint2>(9 * (size_t)nl + 1);
    uint32_t *bs1 = c.bs_l13.as<uint32_t>(2 * (size_t)nbl + 2), *bs3 = bs1 + nbl + 1;
    const double4 *bnd = c.t1 ? (const double4 *)c.c_bnd.p : nullptr;
    (c.t1 ? k_leaf_lists<false, true> : k_leaf_lists<false, false>)<<<nbl, 256, 0, s>>>(
        leaf_idx, nl, c_level, c_flags, c_cx, c_cy, c_child, nbr, bnd, c.hdr_d, c.orders, cnt1,
        cnt3, cntk, nullptr, nullptr, nullptr, nullptr, 0, 0, bs1, bs3, c_begin, c_end, c.hdr_d);
    constexpr int kLeavesPerBlock = 256 / kLeafLanes;
    scan_exclusive_bs_u32_pair(cnt1, off1, bs1, &c.hdr_d->tot1, cnt3, off3, bs3, &c.hdr_d->tot3, nl,
                               kLeavesPerBlock, true, tmp, s);
    // Capacities without reading the totals back.  A leaf's list 1 holds
    // itself, at most 8 adjacent leaves at its level or coarser (one per
    // direction) and the finer adjacent leaves; each finer leaf is adjacent to
    // at most 8 coarser leaves, so sum |list 1| <= 17 nl.  A cell is in the
    // list 3 of leaves coarser than it and adjacent to its parent: at most 8,
    // so sum |list 3| <= 8 nc.  The fill pass checks both anyway.
    ti.cap1 = 17 * (size_t)nl + 1;
    ti.cap3 = 8 * (size_t)nc + 1;
    int *l1 = c.l1.as<int>(ti.cap1), *l3 = c.l3.as<int>(ti.cap3);
    (c.t1 ? k_leaf_lists<true, true> : k_leaf_lists<true, false>)<<<nbl, 256, 0, s>>>(
        leaf_idx, nl, c_level, c_flags, c_cx, c_cy, c_child, nbr, bnd, c.hdr_d, c.orders, nullptr,
        nullptr, cntk, off1, off3, l1, l3, (uint32_t)ti.cap1, (uint32_t)ti.cap3, nullptr, nullptr,
        c_begin, c_end, c.hdr_d);
    VPM_LAUNCH_CHECK();
}

void tree_lists_absorb(Ctx &c) {
    const DevHdr &hh = *c.hdr_h;
    TreeInfo &ti = c.info;
    check_flags(hh);
    if ((int)hh.n_leaves != ti.n_leaves || (int)hh.n_cells_pre + 1 != ti.n_cells)
        throw Error(kInternal, "cell or leaf count of the cell build disagrees with the leaf split");
    if ((size_t)hh.tot1 >= ti.cap1 || (size_t)hh.tot3 >= ti.cap3)
        throw Error(kInternal, "interaction lists exceed their bounds");
    ti.tot1 = (int)hh.tot1;
    ti.tot3 = (int)hh.tot3;
}

}  // namespace vpm
