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
__device__ __forceinline__ uint32_t prefix_at(uint32_t key, int level) {
    return level == 0 ? 0u : key >> (32 - 2 * level);
}
// Cell side at a level: R / 2^L (exact).
__device__ __forceinline__ double cell_len(double R, int level) {
    return R / (double)(1u << level);
}

// ---------------------------------------------------------------------------
// 1. bounding box and root geometry
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
    k_bbox_partial(const double *x, const double *y, int n, double *part) {
    __shared__ double s[4][256];
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
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

// treeCell.cpp:1079-1092: R = max side * (1 + expTree/100), pivot = (max + min - R)/2.
// One block reduces the per-block partials; thread 0 forms the root geometry.
__global__ void __launch_bounds__(256)
    k_bbox_final(const double *part, int nb, double expansion, DevHdr *hdr) {
#pragma clang fp contract(off)
    __shared__ double s[4][256];
    __shared__ int sbad;
    if (threadIdx.x == 0) sbad = 0;
    double mn0 = part[0], mx0 = part[1], mn1 = part[2], mx1 = part[3];
    int bad = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        mn0 = fmin(mn0, part[5 * b + 0]);
        mx0 = fmax(mx0, part[5 * b + 1]);
        mn1 = fmin(mn1, part[5 * b + 2]);
        mx1 = fmax(mx1, part[5 * b + 3]);
        bad |= part[5 * b + 4] != 0.0;
    }
    s[0][threadIdx.x] = mn0;
    s[1][threadIdx.x] = mx0;
    s[2][threadIdx.x] = mn1;
    s[3][threadIdx.x] = mx1;
    __syncthreads();
    if (bad) atomicOr(&sbad, 1);
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s[0][threadIdx.x] = fmin(s[0][threadIdx.x], s[0][threadIdx.x + o]);
            s[1][threadIdx.x] = fmax(s[1][threadIdx.x], s[1][threadIdx.x + o]);
            s[2][threadIdx.x] = fmin(s[2][threadIdx.x], s[2][threadIdx.x + o]);
            s[3][threadIdx.x] = fmax(s[3][threadIdx.x], s[3][threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double mn[2] = {s[0][0], s[2][0]}, mx[2] = {s[1][0], s[3][0]};
    double R = 0.0;
    for (int d = 0; d < 2; ++d) {
        double side = fabs(mx[d] - mn[d]);
        if (side > R) R = side;
    }
    R *= expansion;
    hdr->root_len = R;
    for (int d = 0; d < 2; ++d) {
        hdr->mn[d] = mn[d];
        hdr->mx[d] = mx[d];
        hdr->pivot[d] = (mx[d] + mn[d] - R) / 2.0;
    }
    unsigned f = 0;
    if (sbad) f |= kFlagNonFinite;
    if (!(R > 0.0)) f |= kFlagRange;
    hdr->flags |= f;
    hdr->leaf_lvl_min = kMaxDepth + 1;
    hdr->leaf_lvl_max = 0;
}

// ---------------------------------------------------------------------------
// 2-3. Morton keys, sort, gather
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
    k_keys(const double *x, const double *y, int n, const DevHdr *hdr, uint32_t *key,
           uint32_t *val) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double R = hdr->root_len, px = hdr->pivot[0], py = hdr->pivot[1];
    const double tx = (x[i] - px) / R, ty = (y[i] - py) / R;
    const double sx = tx * 65536.0, sy = ty * 65536.0;
    uint32_t ux = 0, uy = 0;
    if (sx >= 0.0 && sx < 65536.0) ux = (uint32_t)sx;
    if (sy >= 0.0 && sy < 65536.0) uy = (uint32_t)sy;
    key[i] = part1by1(ux) | (part1by1(uy) << 1);
    val[i] = (uint32_t)i;
}

__global__ void __launch_bounds__(256)
    k_gather(const uint32_t *perm, const double *x, const double *y, const double *q,
             const uint8_t *a, int n, double *xs, double *ys, double *qs, uint8_t *as) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t o = perm[i];
    xs[i] = x[o];
    ys[i] = y[o];
    qs[i] = q[o];
    as[i] = a[o] ? 1 : 0;
}

// ---------------------------------------------------------------------------
// 4. leaf levels from windowed LCPs
// ---------------------------------------------------------------------------
// d[i] = lcp(key[i-1], key[i]) (d[0] = d[n] = 0); w[j] = lcp(key[j], key[j+W]).
__global__ void __launch_bounds__(256)
    k_lcp(const uint32_t *key, int n, int W, uint8_t *d, uint8_t *w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    d[i] = (i == 0 || i == n) ? 0 : (uint8_t)lcp_levels(key[i - 1], key[i]);
    if (i + W < n) w[i] = (uint8_t)lcp_levels(key[i], key[i + W]);
}

__global__ void __launch_bounds__(256)
    k_act_pos(const uint8_t *as, const uint32_t *A, int n, uint32_t *apos) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && as[i]) apos[A[i]] = (uint32_t)i;
}

// v[m] = lcp(key[a_m], key[a_{m+S}]) over the sorted active particles.
__global__ void __launch_bounds__(256)
    k_lcp_v(const uint32_t *key, const uint32_t *apos, const DevHdr *hdr, int S, int n,
            uint8_t *v) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const long long na = hdr->n_active;
    if (m >= n || (long long)m + S >= na) return;
    v[m] = (uint8_t)lcp_levels(key[apos[m]], key[apos[m + S]]);
}

// Per-block partial results (two 64-bit words per block) and their one-block
// reduction.  Global atomics on one header word from thousands of blocks
// serialize across the chiplets; partials plus one small reduction do not.
enum PartOp { kPartSplit = 0, kPartM2lLists = 1, kPartLeafLists = 2, kPartLeafTerms = 3 };

// Block-wide (sum or min/max) of two 64-bit values; thread 0 writes part[blk].
template <bool MINMAX>
__device__ __forceinline__ void block_part(unsigned long long a, unsigned long long b,
                                           ulonglong2 *part) {
    __shared__ unsigned long long sa[16], sb[16];
    for (int m = 1; m < kWave; m <<= 1) {
        const unsigned long long oa = __shfl_xor(a, m, kWave), ob = __shfl_xor(b, m, kWave);
        if (MINMAX) {
            a = min(a, oa);
            b = max(b, ob);
        } else {
            a += oa;
            b += ob;
        }
    }
    const int w = threadIdx.x / kWave, nw = (blockDim.x + kWave - 1) / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        sa[w] = a;
        sb[w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int u = 1; u < nw; ++u) {
            if (MINMAX) {
                a = min(a, sa[u]);
                b = max(b, sb[u]);
            } else {
                a += sa[u];
                b += sb[u];
            }
        }
        part[blockIdx.x] = make_ulonglong2(a, b);
    }
}

__global__ void __launch_bounds__(256)
    k_reduce_part(const ulonglong2 *part, int nblk, int op, DevHdr *hdr) {
    const bool mm = op == kPartSplit;
    unsigned long long a = mm ? ~0ull : 0ull, b = 0ull;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
        const ulonglong2 v = part[i];
        if (mm) {
            a = min(a, v.x);
            b = max(b, v.y);
        } else {
            a += v.x;
            b += v.y;
        }
    }
    __shared__ unsigned long long sa[4], sb[4];
    for (int m = 1; m < kWave; m <<= 1) {
        const unsigned long long oa = __shfl_xor(a, m, kWave), ob = __shfl_xor(b, m, kWave);
        a = mm ? min(a, oa) : a + oa;
        b = mm ? max(b, ob) : b + ob;
    }
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        sa[w] = a;
        sb[w] = b;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int u = 1; u < 4; ++u) {
        a = mm ? min(a, sa[u]) : a + sa[u];
        b = mm ? max(b, sb[u]) : b + sb[u];
    }
    if (op == kPartSplit) {
        // a = min leaf level; b = max leaf level | depth-overflow bit 32
        hdr->leaf_lvl_min = (uint32_t)a;
        hdr->leaf_lvl_max = (uint32_t)(b & 0xffffffffull);
        if (b >> 32) hdr->flags |= kFlagDepth;
    } else if (op == kPartM2lLists) {
        hdr->sum2 = (uint32_t)a;
        hdr->sum4 = (uint32_t)b;
    } else if (op == kPartLeafLists) {
        hdr->pairs = a;
        hdr->m2p = b;
    } else {
        hdr->m2p_terms = a;
        hdr->l2p_terms = b;
    }
}

constexpr int kSplitBlock = 256;

// Sliding-window maxima by the van Herk / Gil-Werman method: with blocks of
// length B aligned at 0, the max over [t, t + B - 1] is max(h[t], g[t + B - 1])
// where g / h are block-prefix / block-suffix maxima.  Entries outside the
// valid range hold 0, which never raises a maximum of valid lcp values.  The
// prefix / suffix maxima are segmented Hillis-Steele scans over the whole
// workgroup (log2 B steps, ping-pong buffers); g and h return pointing at the
// buffers that hold the results.
__device__ __forceinline__ void block_prefix_suffix(const uint8_t *a, uint8_t *&g, uint8_t *&h,
                                                    uint8_t *g2, uint8_t *h2, int len, int B) {
    for (int x = threadIdx.x; x < len; x += blockDim.x) {
        g[x] = a[x];
        h[x] = a[x];
    }
    __syncthreads();
    for (int dd = 1; dd < B; dd <<= 1) {
        for (int x = threadIdx.x; x < len; x += blockDim.x) {
            const int off = x % B;  // position inside its block
            uint8_t gv = g[x], hv = h[x];
            if (off >= dd) gv = max(gv, g[x - dd]);
            if (off + dd < B && x + dd < len) hv = max(hv, h[x + dd]);
            g2[x] = gv;
            h2[x] = hv;
        }
        __syncthreads();
        uint8_t *t = g;
        g = g2;
        g2 = t;
        t = h;
        h = h2;
        h2 = t;
    }
}

// Per particle: the deepest level whose cell is over the particle (W) or
// source (S) threshold, hence the leaf level; and the number of cells whose
// first particle it is.
__global__ void __launch_bounds__(kSplitBlock)
    k_split(const uint32_t *key, const uint8_t *d, const uint8_t *w, const uint32_t *A,
            const uint32_t *apos, const uint8_t *v, const DevHdr *hdr, int n, int W, int S,
            int min_level, int max_level, uint8_t *leaf_lvl, uint32_t *ccount, ulonglong2 *part) {
    extern __shared__ uint8_t lds[];
    const int i0 = blockIdx.x * kSplitBlock;
    const int i = i0 + threadIdx.x;
    const int ilast = min(i0 + kSplitBlock, n) - 1;
    const int na = (int)hdr->n_active;
    // runs of W+1 particles: start j in [i - W, i]; LDS index j - (i0 - W)
    const int wlen = W + kSplitBlock;
    const int wbase = i0 - W;
    const int wlast = n - 1 - W;  // last valid run start (< 0: no run of W+1 particles)
    uint8_t *sw = lds, *gw = sw + wlen, *hw = gw + wlen, *gw2 = hw + wlen, *hw2 = gw2 + wlen;
    for (int t = threadIdx.x; t < wlen; t += kSplitBlock) {
        const int j = wbase + t;
        sw[t] = (j >= 0 && j <= wlast) ? w[j] : 0;
    }
    // runs of S+1 active particles: start m in [r - S, r - 1]; LDS index m - vbase
    const int r_lo = (int)A[i0], r_hi = (int)A[ilast];
    const int vbase = r_lo - S - 1;
    const int vlast = na - 1 - S;
    const int vlen = r_hi - vbase + 1;
    uint8_t *sv = hw2 + wlen, *gv = sv + vlen, *hv = gv + vlen, *gv2 = hv + vlen, *hv2 = gv2 + vlen;
    for (int t = threadIdx.x; t < vlen; t += kSplitBlock) {
        const int m = vbase + t;
        sv[t] = (m >= 0 && m <= vlast) ? v[m] : 0;
    }
    __syncthreads();
    block_prefix_suffix(sw, gw, hw, gw2, hw2, wlen, W + 1);
    if (S > 0) block_prefix_suffix(sv, gv, hv, gv2, hv2, vlen, S);
    __syncthreads();
    unsigned long long my_min = kMaxDepth + 1, my_max = 0;
    if (i < n) {
        const uint32_t ki = key[i];
        const int t = threadIdx.x;
        int par_split = wlast >= 0 ? (int)max(hw[t], gw[t + W]) : -1;
        int src_split = -1;
        const int r = (int)A[i];
        if (S > 0) {
            const int a = r - S - vbase;  // window [r - S, r - 1] in LDS coordinates
            src_split = (int)max(hv[a], gv[a + S - 1]);
        }
        if (r <= vlast) {  // run starting at the first active at or after i
            const int e = min((int)sv[r - vbase], lcp_levels(ki, key[apos[r]]));
            src_split = max(src_split, e);
        }
        const int m0 = r - S - 1;
        if (m0 >= 0 && m0 <= vlast) {  // run ending at the last active before i
            const int e = min((int)sv[m0 - vbase], lcp_levels(key[apos[r - 1]], ki));
            src_split = max(src_split, e);
        }
        int lvl = max(max(par_split, src_split) + 1, max(min_level, 1));
        if (max_level > 0) lvl = min(lvl, max_level);
        if (lvl > kMaxDepth) {
            my_max |= 1ull << 32;  // depth overflow
            lvl = kMaxDepth;
        }
        leaf_lvl[i] = (uint8_t)lvl;
        const int di = d[i];
        ccount[i] = lvl > di ? (uint32_t)(lvl - di) : 0u;
        my_min = (unsigned long long)lvl;
        my_max |= (unsigned long long)lvl;
    }
    block_part<true>(my_min, my_max, part);
}

// ---------------------------------------------------------------------------
// 5. cells
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
    k_emit(const uint8_t *d, const uint8_t *leaf_lvl, const uint32_t *base, int n,
           uint32_t *pre_level, uint32_t *pre_begin) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        pre_level[0] = 0;
        pre_begin[0] = 0;
    }
    if (i >= n) return;
    const int di = d[i], lv = leaf_lvl[i];
    uint32_t slot = 1 + base[i];
    for (int L = di + 1; L <= lv; ++L, ++slot) {
        pre_level[slot] = (uint32_t)L;
        pre_begin[slot] = (uint32_t)i;
    }
}

__global__ void k_lev_off(const uint32_t *lev_level, int nc, DevHdr *hdr) {
    const int L = threadIdx.x;
    if (L > kMaxDepth + 1) return;
    int lo = 0, hi = nc;  // first index with level >= L
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if ((int)lev_level[mid] < L) lo = mid + 1;
        else hi = mid;
    }
    hdr->lev_off[L] = (uint32_t)lo;
}

// Cell attributes: range, counts, flags, index, centre (treeCell.hpp:68).
__global__ void __launch_bounds__(256)
    k_cells1(const uint32_t *lev_level, const uint32_t *lev_begin, int nc, const uint32_t *key,
             const uint32_t *A, const uint8_t *leaf_lvl, int n, const DevHdr *hdr,
             uint8_t *c_level, uint8_t *c_flags, uint32_t *c_key, uint32_t *c_begin,
             uint32_t *c_end, uint32_t *c_src, double *c_cx, double *c_cy) {
#pragma clang fp contract(off)
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nc) return;
    const int L = (int)lev_level[j];
    const uint32_t b = lev_begin[j];
    uint32_t e, pre = 0;
    if (L == 0) {
        e = (uint32_t)n;
    } else {
        const int sh = 32 - 2 * L;
        pre = key[b] >> sh;
        long long lo = b, hi, step = 1;
        for (;;) {
            long long nx = lo + step;
            if (nx >= n) {
                hi = n;
                break;
            }
            if ((key[nx] >> sh) != pre) {
                hi = nx;
                break;
            }
            lo = nx;
            step <<= 1;
        }
        while (hi - lo > 1) {
            long long mid = lo + ((hi - lo) >> 1);
            if ((key[mid] >> sh) == pre) lo = mid;
            else hi = mid;
        }
        e = (uint32_t)hi;
    }
    const uint32_t src = A[e] - A[b];
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

__global__ void __launch_bounds__(256)
    k_leaf01(const uint8_t *c_flags, int nc, uint8_t *out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < nc) out[j] = (c_flags[j] & kLeaf) ? 1 : 0;
}

__global__ void __launch_bounds__(256)
    k_leaf_idx(const uint8_t *c_flags, const uint32_t *scan, int nc, uint32_t *leaf_idx) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < nc && (c_flags[j] & kLeaf)) leaf_idx[scan[j]] = (uint32_t)j;
}

// ---------------------------------------------------------------------------
// 6. neighbour table and lists
// ---------------------------------------------------------------------------
// Index of the cell with key `pre` in the sorted range [lo, hi), or -1;
// galloping search from `start` (neighbours lie close in Morton order).
__device__ __forceinline__ int find_near(const uint32_t *key, int lo, int hi, int start,
                                         uint32_t pre) {
    const uint32_t ks = key[start];
    if (ks == pre) return start;
    int a, b;
    if (ks < pre) {
        int prev = start, step = 1;
        for (;;) {
            const int nx = start + step;
            if (nx >= hi) {
                a = prev + 1;
                b = hi;
                break;
            }
            if (key[nx] >= pre) {
                a = prev + 1;
                b = nx;
                break;
            }
            prev = nx;
            step <<= 1;
        }
    } else {
        int prev = start, step = 1;
        for (;;) {
            const int nx = start - step;
            if (nx < lo) {
                a = lo;
                b = prev;
                break;
            }
            if (key[nx] < pre) {
                a = nx + 1;
                b = prev;
                break;
            }
            prev = nx;
            step <<= 1;
        }
    }
    while (a < b) {  // lower bound in [a, b]; key[b] >= pre or b == hi
        const int m = (a + b) >> 1;
        if (key[m] < pre) a = m + 1;
        else b = m;
    }
    return (a < hi && key[a] == pre) ? a : -1;
}

// nbr[9j + k], k = (dx+1) + 3 (dy+1) (treeCell.cpp:593-622 order): the
// deepest existing cell at level <= L containing the level-L cell at
// (ix+dx, iy+dy), or -1 outside [0, 2^L)^2.  One thread per (cell, k).
__global__ void __launch_bounds__(256)
    k_nbr(const uint8_t *c_level, const uint32_t *c_key, int nc, const DevHdr *hdr, int *nbr) {
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
        const uint32_t p = part1by1((uint32_t)X) | (part1by1((uint32_t)Y) << 1);
        res = find_near(c_key, (int)hdr->lev_off[L], (int)hdr->lev_off[L + 1], j, p);
        for (int l = L - 1; res < 0 && l >= 0; --l) {
            const int sh = L - l;
            const uint32_t q = part1by1((uint32_t)X >> sh) | (part1by1((uint32_t)Y >> sh) << 1);
            res = l == 0 ? 0 : find_cell(c_key, hdr->lev_off, l, q);
        }
    }
    nbr[9 * j + k] = res;
}

// Lists 2 and 4 of every cell of level >= 2 (treeCell.cpp:818-1032).
__global__ void __launch_bounds__(256)
    k_m2l_lists(const uint8_t *c_level, const uint8_t *c_flags, const double *c_cx,
                const double *c_cy, const int *c_parent, const int *c_child, const int *nbr,
                const DevHdr *hdr, int j0, int nc, int *l2, uint8_t *n2, int *l4, uint8_t *n4,
                DevHdr *hdr_w, ulonglong2 *part) {
#pragma clang fp contract(off)
    const int j = j0 + blockIdx.x * blockDim.x + threadIdx.x;
    int m2 = 0, m4c = 0, m4 = 0;
    if (j < nc) {
    const int L = c_level[j];
    const double R = hdr->root_len, s = cell_len(R, L);
    const double cx = c_cx[j], cy = c_cy[j];
    const int P = c_parent[j];
    int cand[8];
    const double md2 = s + kListEps;
    for (int k = 0; k < 9; ++k) {
        if (k == 4) continue;  // the parent itself (treeCell.cpp:872)
        const int A = nbr[9 * P + k];
        if (A < 0) continue;
        const unsigned fa = c_flags[A];
        if (!(fa & kActive)) continue;
        if (c_level[A] == L - 1 && !(fa & kLeaf)) {
            for (int q = 0; q < 4; ++q) {
                const int ch = c_child[4 * A + q];
                if (ch < 0 || !(c_flags[ch] & kActive)) continue;
                const bool sep = fabs(c_cx[ch] - cx) > md2 || fabs(c_cy[ch] - cy) > md2;
                if (sep) {
                    if (m2 < 27) l2[27 * j + m2++] = ch;
                    else atomicOr(&hdr_w->flags, kFlagListOverflow);
                }
            }
        } else if (fa & kLeaf) {
            bool dup = false;
            for (int t = 0; t < m4c; ++t) dup |= cand[t] == A;
            if (!dup) cand[m4c++] = A;
        }
    }
    for (int t = 0; t < m4c; ++t) {
        const int A = cand[t];
        const double md = ((cell_len(R, c_level[A]) + s) / 2.0) + kListEps;
        const bool sep = fabs(c_cx[A] - cx) > md || fabs(c_cy[A] - cy) > md;
        if (sep) l4[8 * j + m4++] = A;
    }
    n2[j] = (uint8_t)m2;
    n4[j] = (uint8_t)m4;
    }
    block_part<false>((unsigned long long)m2, (unsigned long long)m4, part);
}

// Lists 1 and 3 of every leaf (treeCell.cpp:636-804), count or fill pass.
// The fill pass also counts the P2P pairs and M2P evaluations the lists imply
// (statistics), reduced per block into per-block partials.
template <bool FILL>
__global__ void __launch_bounds__(256)
    k_leaf_lists(const uint32_t *leaf_idx, int nl, const uint8_t *c_level, const uint8_t *c_flags,
                 const double *c_cx, const double *c_cy, const int *c_child, const int *nbr,
                 const DevHdr *hdr, uint32_t *cnt1, uint32_t *cnt3, const uint32_t *off1,
                 const uint32_t *off3, int *l1, int *l3, const uint32_t *c_begin,
                 const uint32_t *c_end, DevHdr *hdr_w, ulonglong2 *part, ulonglong2 *part_terms) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long my_pairs = 0, my_m2p = 0, my_m2p_terms = 0, my_l2p_terms = 0;
    unsigned long long m2p_w = 0;  // sum over list-3 cells of (order of the cell's level + 1)
    bool overflow = false;
    if (t < nl) {
        const int j = (int)leaf_idx[t];
        const int L = c_level[j];
        const double R = hdr->root_len, s = cell_len(R, L);
        const double cx = c_cx[j], cy = c_cy[j];
        uint32_t n1 = 0, n3 = 0;
        unsigned long long src_total = 0;
        bool self = false;
        const uint32_t o1 = FILL ? off1[t] : 0, o3 = FILL ? off3[t] : 0;
        int coarse[9];
        int nco = 0;
        int stack[64];
        int sp = 0;
        auto emit1 = [&](int c) {
            if (FILL) {
                l1[o1 + n1] = c;
                src_total += c_end[c] - c_begin[c];
                self |= c == j;
            }
            ++n1;
        };
        for (int k = 0; k < 9; ++k) {
            const int A = nbr[9 * j + k];
            if (A < 0) continue;
            const unsigned fa = c_flags[A];
            if (c_level[A] < L) {  // reached through the parent walk (treeCell.cpp:706-720, 738)
                if ((fa & kLeaf) && (fa & kActive)) {
                    bool dup = false;
                    for (int u = 0; u < nco; ++u) dup |= coarse[u] == A;
                    if (!dup) {
                        coarse[nco++] = A;
                        emit1(A);
                    }
                }
                continue;
            }
            if (!(fa & kActive)) continue;
            if (fa & kLeaf) {
                emit1(A);
                continue;
            }
            if (sp < 64) stack[sp++] = A;
            else overflow = true;
        }
        while (sp > 0) {
            const int B = stack[--sp];
            const double adj = ((s + (cell_len(R, c_level[B]) / 2.0)) / 2.0) + kListEps;
            for (int q = 0; q < 4; ++q) {
                const int ch = c_child[4 * B + q];
                if (ch < 0) continue;
                const unsigned fc = c_flags[ch];
                if (!(fc & kActive)) continue;
                const bool is_adj = fabs(c_cx[ch] - cx) <= adj && fabs(c_cy[ch] - cy) <= adj;
                if (is_adj) {
                    if (fc & kLeaf) emit1(ch);
                    else if (sp < 64) stack[sp++] = ch;
                    else overflow = true;
                } else {
                    if (FILL) {
                        l3[o3 + n3] = ch;
                        m2p_w += (unsigned long long)hdr->sel_p[c_level[ch]] + 1ull;
                    }
                    ++n3;
                }
            }
        }
        if (!FILL) {
            cnt1[t] = n1;
            cnt3[t] = n3;
        } else {
            const unsigned long long nt = c_end[j] - c_begin[j];
            my_pairs = src_total * nt - (self ? nt : 0ull);
            my_m2p = (unsigned long long)n3 * nt;
            my_m2p_terms = m2p_w * nt;
            my_l2p_terms = (unsigned long long)hdr->sel_p[L] * nt;
        }
    }
    if (overflow) atomicOr(&hdr_w->flags, kFlagListOverflow);
    if (FILL) {
        block_part<false>(my_pairs, my_m2p, part);
        __syncthreads();  // block_part's LDS is reused
        block_part<false>(my_m2p_terms, my_l2p_terms, part_terms);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------
static void sync_hdr(Ctx &c) {
    VPM_HIP(hipMemcpyAsync(c.hdr_h, c.hdr_d, sizeof(DevHdr), hipMemcpyDeviceToHost, c.stream));
    VPM_HIP(hipStreamSynchronize(c.stream));
}

static void check_flags(const DevHdr &h) {
    if (h.flags & kFlagNonFinite) throw Error(kNonFinite, "non-finite particle coordinate");
    if (h.flags & kFlagRange) throw Error(kDegenerate, "all particles at one point (root length 0)");
    if (h.flags & kFlagDepth)
        throw Error(kDepth, "tree deeper than 16 levels (more coincident particles than the leaf thresholds)");
    if (h.flags & kFlagListOverflow) throw Error(kInternal, "list traversal stack overflow");
}

void tree_build(Ctx &c, int n, const double *x, const double *y, const double *q,
                const uint8_t *active, const TreeParams &tp) {
    hipStream_t s = c.stream;
    const int W = tp.max_par, S = tp.max_src;
    VPM_HIP(hipMemsetAsync(c.hdr_d, 0, sizeof(DevHdr), s));

    // 1. bounding box
    const int nbb = std::min(1024, div_up(n, 256 * 8));
    double *part = c.bbox_part.as<double>(5 * (size_t)nbb);
    k_bbox_partial<<<nbb, 256, 0, s>>>(x, y, n, part);
    const double expansion = 1.0 + (tp.exp_tree / 100.0);
    k_bbox_final<<<1, 256, 0, s>>>(part, nbb, expansion, c.hdr_d);
    VPM_LAUNCH_CHECK();

    // 2-3. keys, sort, gather
    uint32_t *ka = c.key_a.as<uint32_t>(n), *kb = c.key_b.as<uint32_t>(n);
    uint32_t *va = c.val_a.as<uint32_t>(n), *vb = c.val_b.as<uint32_t>(n);
    const size_t tmp_words =
        std::max(radix_tmp_words(n), scan_tmp_words((long long)n + 1) + 64) + 64;
    uint32_t *tmp = c.tmp.as<uint32_t>(std::max(tmp_words, radix_tmp_words(16LL * n + 16)));
    k_keys<<<div_up(n, 256), 256, 0, s>>>(x, y, n, c.hdr_d, ka, va);
    const bool alt = radix_sort_pairs(ka, va, kb, vb, n, 0, 32, tmp, s);
    uint32_t *key = alt ? kb : ka;
    uint32_t *perm = alt ? vb : va;
    c.key_sorted = key;
    c.perm_sorted = perm;
    double *xs = c.xs.as<double>(n), *ys = c.ys.as<double>(n), *qs = c.qs.as<double>(n);
    uint8_t *as = c.acts.as<uint8_t>(n);
    k_gather<<<div_up(n, 256), 256, 0, s>>>(perm, x, y, q, active, n, xs, ys, qs, as);

    // 4. leaf levels
    uint8_t *d = c.lcp_d.as<uint8_t>(n + 1), *w = c.lcp_w.as<uint8_t>(n + 1);
    k_lcp<<<div_up(n + 1, 256), 256, 0, s>>>(key, n, W, d, w);
    uint32_t *A = c.act_scan.as<uint32_t>(n + 1);
    scan_exclusive_u8(as, A, n, &c.hdr_d->n_active, true, tmp, s);
    uint32_t *apos = c.act_pos.as<uint32_t>(n);
    k_act_pos<<<div_up(n, 256), 256, 0, s>>>(as, A, n, apos);
    uint8_t *v = c.lcp_v.as<uint8_t>(n + 1);
    k_lcp_v<<<div_up(n, 256), 256, 0, s>>>(key, apos, c.hdr_d, S, n, v);
    uint8_t *leaf_lvl = c.leaf_lvl.as<uint8_t>(n);
    uint32_t *cnt = c.ccount.as<uint32_t>(n);
    const size_t lds = 5 * (size_t)(W + kSplitBlock) + 5 * (size_t)(S + kSplitBlock + 2) + 16;
    const int nbs = div_up(n, kSplitBlock);
    ulonglong2 *spart = c.part.as<ulonglong2>(nbs);
    k_split<<<nbs, kSplitBlock, lds, s>>>(key, d, w, A, apos, v, c.hdr_d, n, W, S, tp.min_level,
                                          tp.max_level, leaf_lvl, cnt, spart);
    k_reduce_part<<<1, 256, 0, s>>>(spart, nbs, kPartSplit, c.hdr_d);
    uint32_t *base = c.cbase.as<uint32_t>(n + 1);
    scan_exclusive_u32(cnt, base, n, &c.hdr_d->n_cells_pre, true, tmp, s);
    VPM_LAUNCH_CHECK();
    sync_hdr(c);
    check_flags(*c.hdr_h);

    // 5. cells
    const int nc = 1 + (int)c.hdr_h->n_cells_pre;
    TreeInfo &ti = c.info;
    ti.n = n;
    ti.n_active = (int)c.hdr_h->n_active;
    ti.n_cells = nc;
    ti.max_level = (int)c.hdr_h->leaf_lvl_max;
    ti.min_leaf_level = (int)c.hdr_h->leaf_lvl_min;
    ti.root_len = c.hdr_h->root_len;
    ti.pivot[0] = c.hdr_h->pivot[0];
    ti.pivot[1] = c.hdr_h->pivot[1];

    uint32_t *pl = c.pre_level.as<uint32_t>(nc), *pb = c.pre_slot.as<uint32_t>(nc);
    uint32_t *ll = c.lev_level.as<uint32_t>(nc), *lb = c.lev2pre.as<uint32_t>(nc);
    k_emit<<<div_up(n, 256), 256, 0, s>>>(d, leaf_lvl, base, n, pl, pb);
    const bool alt2 = radix_sort_pairs(pl, pb, ll, lb, nc, 0, 8, tmp, s);
    const uint32_t *lev_level = alt2 ? ll : pl;
    const uint32_t *lev_begin = alt2 ? lb : pb;
    k_lev_off<<<1, 64, 0, s>>>(lev_level, nc, c.hdr_d);

    uint8_t *c_level = c.c_level.as<uint8_t>(nc), *c_flags = c.c_flags.as<uint8_t>(nc);
    uint32_t *c_key = c.c_key.as<uint32_t>(nc), *c_begin = c.c_begin.as<uint32_t>(nc);
    uint32_t *c_end = c.c_end.as<uint32_t>(nc), *c_src = c.c_src.as<uint32_t>(nc);
    double *c_cx = c.c_cx.as<double>(nc), *c_cy = c.c_cy.as<double>(nc);
    int *c_parent = c.c_parent.as<int>(nc), *c_child = c.c_child.as<int>(4 * (size_t)nc);
    k_cells1<<<div_up(nc, 256), 256, 0, s>>>(lev_level, lev_begin, nc, key, A, leaf_lvl, n,
                                             c.hdr_d, c_level, c_flags, c_key, c_begin, c_end,
                                             c_src, c_cx, c_cy);
    VPM_HIP(hipMemsetAsync(c_child, 0xFF, 4 * sizeof(int) * (size_t)nc, s));
    k_cells2<<<div_up(nc, 256), 256, 0, s>>>(c_level, c_key, nc, c.hdr_d, c_parent, c_child);

    uint8_t *leaf01 = c.lcp_w.as<uint8_t>(std::max(n + 1, nc));
    k_leaf01<<<div_up(nc, 256), 256, 0, s>>>(c_flags, nc, leaf01);
    uint32_t *lscan = c.leaf_flag_scan.as<uint32_t>(nc + 1);
    scan_exclusive_u8(leaf01, lscan, nc, &c.hdr_d->n_leaves, true, tmp, s);
    uint32_t *leaf_idx = c.leaf_idx.as<uint32_t>(nc);
    k_leaf_idx<<<div_up(nc, 256), 256, 0, s>>>(c_flags, lscan, nc, leaf_idx);
    VPM_LAUNCH_CHECK();
}

void tree_lists(Ctx &c) {
    hipStream_t s = c.stream;
    const int nc = c.info.n_cells, n = c.info.n;
    uint8_t *c_level = (uint8_t *)c.c_level.p, *c_flags = (uint8_t *)c.c_flags.p;
    uint32_t *c_key = (uint32_t *)c.c_key.p;
    double *c_cx = (double *)c.c_cx.p, *c_cy = (double *)c.c_cy.p;
    int *c_parent = (int *)c.c_parent.p, *c_child = (int *)c.c_child.p;
    int *nbr = c.c_nbr.as<int>(9 * (size_t)nc);
    uint32_t *tmp = (uint32_t *)c.tmp.p;
    k_nbr<<<div_up(9LL * nc, 256), 256, 0, s>>>(c_level, c_key, nc, c.hdr_d, nbr);

    // level offsets and leaf count are needed on the host for the launches below
    sync_hdr(c);
    TreeInfo &ti = c.info;
    for (int l = 0; l < kMaxDepth + 2; ++l) ti.lev_off[l] = c.hdr_h->lev_off[l];
    ti.n_leaves = (int)c.hdr_h->n_leaves;
    const int nl = ti.n_leaves;
    (void)n;

    int *l2 = c.l2.as<int>(27 * (size_t)nc), *l4 = c.l4.as<int>(8 * (size_t)nc);
    uint8_t *n2 = c.n2.as<uint8_t>(nc), *n4 = c.n4.as<uint8_t>(nc);
    const int j0 = (int)ti.lev_off[2];
    // n2 / n4 of every cell of level >= 2 are written by k_m2l_lists; levels 0
    // and 1 have no lists 2 / 4 and nothing reads their entries
    const int nb2 = nc > j0 ? div_up(nc - j0, 256) : 0;
    const int nbl = div_up(nl, 256);
    ulonglong2 *part = c.part.as<ulonglong2>(std::max(std::max(nb2, 2 * nbl), 1));
    if (nb2 > 0)
        k_m2l_lists<<<nb2, 256, 0, s>>>(c_level, c_flags, c_cx, c_cy, c_parent, c_child, nbr,
                                        c.hdr_d, j0, nc, l2, n2, l4, n4, c.hdr_d, part);
    k_reduce_part<<<1, 256, 0, s>>>(part, nb2, kPartM2lLists, c.hdr_d);
    const uint32_t *leaf_idx = (const uint32_t *)c.leaf_idx.p;
    uint32_t *cnt1 = c.cnt1.as<uint32_t>(nl + 1), *cnt3 = c.cnt3.as<uint32_t>(nl + 1);
    uint32_t *off1 = c.off1.as<uint32_t>(nl + 1), *off3 = c.off3.as<uint32_t>(nl + 1);
    const uint32_t *c_begin = (const uint32_t *)c.c_begin.p, *c_end = (const uint32_t *)c.c_end.p;
    k_leaf_lists<false><<<div_up(nl, 256), 256, 0, s>>>(leaf_idx, nl, c_level, c_flags, c_cx,
                                                         c_cy, c_child, nbr, c.hdr_d, cnt1, cnt3,
                                                         nullptr, nullptr, nullptr, nullptr,
                                                         c_begin, c_end, c.hdr_d, nullptr, nullptr);
    scan_exclusive_u32(cnt1, off1, nl, &c.hdr_d->tot1, true, tmp, s);
    scan_exclusive_u32(cnt3, off3, nl, &c.hdr_d->tot3, true, tmp, s);
    VPM_LAUNCH_CHECK();
    sync_hdr(c);
    check_flags(*c.hdr_h);
    ti.tot1 = (int)c.hdr_h->tot1;
    ti.tot3 = (int)c.hdr_h->tot3;
    int *l1 = c.l1.as<int>((size_t)ti.tot1 + 1), *l3 = c.l3.as<int>((size_t)ti.tot3 + 1);
    k_leaf_lists<true><<<div_up(nl, 256), 256, 0, s>>>(leaf_idx, nl, c_level, c_flags, c_cx,
                                                        c_cy, c_child, nbr, c.hdr_d, nullptr,
                                                        nullptr, off1, off3, l1, l3, c_begin, c_end,
                                                        c.hdr_d, part, part + nbl);
    k_reduce_part<<<1, 256, 0, s>>>(part, nbl, kPartLeafLists, c.hdr_d);
    k_reduce_part<<<1, 256, 0, s>>>(part + nbl, nbl, kPartLeafTerms, c.hdr_d);
    VPM_LAUNCH_CHECK();
}

}  // namespace vpm
