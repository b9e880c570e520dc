// fmm32.hip -- mixed-precision FMM stages on gfx950.
//
// Precision split: cell geometry, centre differences and particle-minus-centre
// displacements are formed in FP64 (tree.hip / k_p2m32) and rounded once;
// everything else is FP32 with FP32 accumulation.
//
// Scaled coefficients.  With s_L the side of a level-L cell,
//   a^_k = a_k / s^k,   b^_m = b_m s^m,
// every translation becomes level independent:
//   M2M  a^par_m = sum_k C(m,k) e^(m-k) 2^-m a^ch_k     (e = (zch - zpar)/s_ch)
//   M2L  b^_m    = sum_k T_mk delta^-(m+k) a^_k         (delta = integer offset)
//   L2L  b^ch_m  = sum_k C(k,m) e^(k-m) 2^-k b^par_k
// with T_m0 = -1/m, T_mk = -C(m+k-1,k-1)(-1)^k/k (the reference's M2L,
// fmm2D.cpp:465-481, with exact binomials).  The scaled magnitudes stay O(1)
// (|(z - zc)/s| <= 1/sqrt(2), |delta| >= 2), so FP32 neither under- nor
// overflows at any level.
//
// M2L over list 2 (same-level cells at integer offsets in [-3,3]^2 \ [-1,1]^2,
// 40 offsets, at most 27 per cell) runs on MFMA: one wavefront owns 16 target
// cells of one level; for each offset it multiplies the realified 2p x 2(p+1)
// matrix of that offset with the 16 sources' scaled coefficient vectors
// (v_mfma_f32_16x16x4f32, K = 4 per instruction, FP32 accumulate).  Offsets a
// target does not use contribute zero columns.  List 4 (coarser leaves) has
// irregular offsets and is done on the vector ALUs.
#include "ctx.hpp"
#include "fmm32.hpp"

namespace vpm {
namespace {

constexpr unsigned kLeaf = 1u, kActive = 2u;
typedef float v4f __attribute__((ext_vector_type(4)));

struct cf {
    float x, y;
};
__device__ __forceinline__ cf cmul(cf a, cf b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ cf cadd(cf a, cf b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cf cscale(cf a, float s) { return {a.x * s, a.y * s}; }
__device__ __forceinline__ cf cinv(cf a) {
    const float r = __builtin_amdgcn_rcpf(a.x * a.x + a.y * a.y);
    return {a.x * r, -a.y * r};
}
__device__ __forceinline__ cf cpowi(cf z, int p) {
    cf r = {1.f, 0.f};
    while (p > 0) {
        if (p & 1) r = cmul(r, z);
        z = cmul(z, z);
        p >>= 1;
    }
    return r;
}
__device__ __forceinline__ cf ldc(const float *a, size_t i) {
    const float2 v = reinterpret_cast<const float2 *>(a)[i];
    return {v.x, v.y};
}
__device__ __forceinline__ void stc(float *a, size_t i, cf v) {
    reinterpret_cast<float2 *>(a)[i] = make_float2(v.x, v.y);
}
__device__ __forceinline__ uint32_t compact1by1(uint32_t v) {
    v &= 0x55555555u;
    v = (v | (v >> 1)) & 0x33333333u;
    v = (v | (v >> 2)) & 0x0F0F0F0Fu;
    v = (v | (v >> 4)) & 0x00FF00FFu;
    v = (v | (v >> 8)) & 0x0000FFFFu;
    return v;
}

// Lanes of one wavefront read what other lanes of it wrote to LDS.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// P2M: one wavefront per leaf, lanes over its particles.  Each lane also
// writes its particles' FP32 offsets from the leaf centre (FP64 difference,
// rounded once) and FP32 strengths for the later stages; lane-private running
// powers accumulate a^_k = sum q ((z - zc)/s)^k (k <= p, active particles of
// active leaves; inactive leaves get zero).  The lanes' sums are reduced
// through LDS: lane l writes column l of a (order x lane) table, lane k sums
// row k (stride 65 keeps the rows on distinct banks), in two halves of 9
// orders, so a wavefront holds 4.7 KB of LDS and 7 wavefronts per SIMD fit
// (4 with the whole table; the upward pass 3-4 us shorter at 780k).
constexpr int kP2mWaves = 4;

__global__ void __launch_bounds__(kWave *kP2mWaves)
    k_p2m32(const uint32_t *leaf_idx, int nl, const uint8_t *c_level, const uint8_t *c_flags,
            const uint32_t *c_begin, const uint32_t *c_end, const double *c_cx,
            const double *c_cy, const double *xs, const double *ys, const double *qs,
            const uint8_t *as, const LevelTab *lt, int kst, float2 *oxy, float *qf,
            float *ahat) {
    __shared__ float2 red[kP2mWaves][9][kWave + 1];
    const int PM1 = kst / 2;
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const int t = blockIdx.x * kP2mWaves + w;
    if (t >= nl) return;
    const int j = (int)leaf_idx[t];
    const int L = c_level[j];
    const int p = lt->p[L];
    const float inv_s = lt->inv_s[L];
    const double cx = c_cx[j], cy = c_cy[j];
    const bool act = (c_flags[j] & kActive) != 0;  // inactive leaves: zero
    float ar[17], ai[17];
#pragma unroll
    for (int k = 0; k < 17; ++k) ar[k] = ai[k] = 0.f;
    const uint32_t e = c_end[j];
    for (uint32_t i = c_begin[j] + lane; i < e; i += kWave) {
        const float2 o = make_float2((float)(xs[i] - cx), (float)(ys[i] - cy));
        const float qi = (float)qs[i];
        oxy[i] = o;
        qf[i] = qi;
        if (!act || !as[i]) continue;
        const cf z = {o.x * inv_s, o.y * inv_s};
        cf v = {qi, 0.f};
#pragma unroll
        for (int k = 0; k < 17; ++k) {
            if (k <= p) {
                ar[k] += v.x;
                ai[k] += v.y;
                v = cmul(v, z);
            }
        }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int k0 = 9 * h;
#pragma unroll
        for (int kk = 0; kk < 9; ++kk) {
            const int k = k0 + kk;
            if (k < 17 && k <= p) red[w][kk][lane] = make_float2(ar[k], ai[k]);
        }
        wave_sync();
        if (lane >= k0 && lane < k0 + 9 && lane <= p) {
            const int row = lane - k0;
            float2 s0 = make_float2(0.f, 0.f), s1 = s0, s2 = s0, s3 = s0;
#pragma unroll 4
            for (int l = 0; l < kWave; l += 4) {
                const float2 a0 = red[w][row][l], a1 = red[w][row][l + 1];
                const float2 a2 = red[w][row][l + 2], a3 = red[w][row][l + 3];
                s0.x += a0.x; s0.y += a0.y;
                s1.x += a1.x; s1.y += a1.y;
                s2.x += a2.x; s2.y += a2.y;
                s3.x += a3.x; s3.y += a3.y;
            }
            stc(ahat, (size_t)j * PM1 + lane,
                cf{(s0.x + s1.x) + (s2.x + s3.x), (s0.y + s1.y) + (s2.y + s3.y)});
        }
        wave_sync();
    }
}

// M2M and L2L by one wavefront per parent cell.  The children's (M2M) or the
// parent's (L2L) coefficients are first loaded by all lanes at once into LDS,
// then each lane forms one output order from LDS and the level-independent
// table m2m[q][m][k] / l2l[q][m][k] (complex, stride PM1 x PM1, also in LDS).
// q is the child quadrant, c_child's order: x bit + 2 * y bit.
constexpr int kMaxPM1 = 17;

// Child q's index from the parent's four child slots.
__device__ __forceinline__ int child_of(int4 c4, int q) {
    return q == 0 ? c4.x : q == 1 ? c4.y : q == 2 ? c4.z : c4.w;
}

// M2M into parent j of order p from children of order pc (inactive parents
// get zero: their own parents read them).  The caller knows the level, so
// the parent's flags and child slots are the only loads before the children's
// coefficients.
__device__ __forceinline__ void m2m_wave(int j, int p, int pc, int lane, int PM1,
                                         const uint8_t *c_flags, const int *c_child,
                                         const float2 *tab, float2 *sa, float *ahat) {
    const int4 c4 = reinterpret_cast<const int4 *>(c_child)[j];
    const unsigned f = c_flags[j];
    if (f & kLeaf) return;
    const bool act = (f & kActive) != 0;
    for (int idx = lane; idx < 4 * PM1; idx += kWave) {
        const int q = idx / PM1, k = idx % PM1;
        const int ch = act ? child_of(c4, q) : -1;
        sa[idx] = (ch >= 0 && k <= pc)
                      ? reinterpret_cast<const float2 *>(ahat)[(size_t)ch * PM1 + k]
                      : make_float2(0.f, 0.f);
    }
    wave_sync();
    if (lane <= p) {
        // one accumulator per child and a fully unrolled order loop: the LDS
        // loads issue ahead of four independent multiply-add chains
        const int m = lane, kmax = min(m, pc);
        cf acc[4] = {{0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}};
#pragma unroll
        for (int k = 0; k < kMaxPM1; ++k) {
            if (k <= kmax) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float2 tv = tab[(q * PM1 + m) * PM1 + k], av = sa[q * PM1 + k];
                    acc[q] = cadd(acc[q], cmul(cf{tv.x, tv.y}, cf{av.x, av.y}));
                }
            }
        }
        stc(ahat, (size_t)j * PM1 + m, cadd(cadd(acc[0], acc[1]), cadd(acc[2], acc[3])));
    }
    wave_sync();
}

// L2L from parent j (order pp) into its children (order pc): lane = (q, m),
// q = lane / 16, m = lane % 16 + 1, adds sum_{k >= m} l2l[q][m][k] b^_k(parent).
__device__ __forceinline__ void l2l_wave(int j, int pp, int pc, int lane, int PM1,
                                         const uint8_t *c_flags, const int *c_child,
                                         const float2 *tab, float2 *sb, float *bhat) {
    const int4 c4 = reinterpret_cast<const int4 *>(c_child)[j];
    if (c_flags[j] & kLeaf) return;
    for (int k = lane; k < PM1; k += kWave)
        sb[k] = k <= pp ? reinterpret_cast<const float2 *>(bhat)[(size_t)j * PM1 + k]
                        : make_float2(0.f, 0.f);
    wave_sync();
    const int q = lane >> 4, m = (lane & 15) + 1;
    const int ch = child_of(c4, q);
    if (ch >= 0 && m <= pc) {
        // fully unrolled over k with two accumulators (even / odd k)
        const float2 *row = tab + (q * PM1 + m) * PM1;
        cf acc[2] = {{0.f, 0.f}, {0.f, 0.f}};
#pragma unroll
        for (int k = 1; k < kMaxPM1; ++k) {
            if (k >= m && k <= pp) {
                const float2 tv = row[k], bv = sb[k];
                acc[k & 1] = cadd(acc[k & 1], cmul(cf{tv.x, tv.y}, cf{bv.x, bv.y}));
            }
        }
        const size_t o = (size_t)ch * PM1 + m;
        stc(bhat, o, cadd(ldc(bhat, o), cadd(acc[0], acc[1])));
    }
    wave_sync();
}

// Table (4 PM1 x PM1 complex, stride PM1) and the order of every level into
// LDS.  Only the entries (q, m, k) with m, k up to the largest order of levels
// lo .. hi are loaded: the passes over those levels read no others.
__device__ __forceinline__ void load_tab(const float *g, int PM1, float2 *tab,
                                         const LevelTab *lt, int *s_lp, int lo, int hi) {
    if (threadIdx.x < kMaxDepth + 2) s_lp[threadIdx.x] = lt->p[threadIdx.x];
    __syncthreads();
    int pm = 0;
    for (int L = lo; L <= hi; ++L) pm = max(pm, s_lp[L]);
    const int P1 = min(pm, PM1 - 1) + 1;
    const float2 *src = reinterpret_cast<const float2 *>(g);
    for (int u = threadIdx.x; u < 4 * P1 * P1; u += blockDim.x) {
        const int q = u / (P1 * P1), r = u - q * P1 * P1, m = r / P1, k = r - m * P1;
        const int idx = (q * PM1 + m) * PM1 + k;
        tab[idx] = src[idx];
    }
    __syncthreads();
}

// Upward / downward passes in three tiers.  Cells are level-major and
// Morton-ordered inside a level, so the descendants of a level-K cell with key
// r form, at each level L > K, the contiguous range of keys
// [r 4^(L-K), (r+1) 4^(L-K)).  Below the split level K (the first level with
// kSubRoots cells) one workgroup owns one level-K cell and walks its subtree
// level by level with workgroup barriers (a wavefront per parent cell); the
// levels between the first level with kTopCells cells and K get one launch
// each (a wavefront per cell over the whole GPU); the few cells above go to
// one workgroup (k_*_top).
constexpr int kSubBlock = 256;
constexpr int kTopBlock = 1024;
constexpr int kSubRoots = 1024;  // subtree roots wanted at the split level K
constexpr int kTopCells = 64;    // levels with fewer cells go to the one-workgroup kernel

// Lower bound of v in c_key[lo, hi), found by a group of kKeyGroup consecutive
// lanes together: each round the group tests the last key of each of
// kKeyGroup equal chunks of the range (one load per lane) and keeps the chunk
// where the bound lies, so about log16 of the range in dependent loads.  All
// lanes of the wavefront call it; each group may search its own range.
constexpr int kKeyGroup = 16;

__device__ int lower_key_group(const uint32_t *c_key, int lo, int hi, unsigned long long v) {
    const int lane = (int)(threadIdx.x & (kWave - 1));
    const int gl = lane & (kKeyGroup - 1), gbase = lane - gl;
    for (;;) {
        const bool live = lo < hi;  // the bound lies in [lo, hi]
        if (!__ballot(live)) break;
        const int st = live ? (hi - lo + kKeyGroup - 1) / kKeyGroup : 0;
        const int p = lo + st * (gl + 1) - 1;
        bool ge = true;
        if (live && p < hi) ge = (unsigned long long)c_key[p] >= v;
        const uint64_t m = (__ballot(ge) >> gbase) & ((1ull << kKeyGroup) - 1);
        if (live) {
            if (m == 0) {  // every tested key is below v, the last one at hi - 1
                lo = hi;
            } else {
                const int f = __ffsll((unsigned long long)m) - 1;
                hi = min(hi, lo + st * (f + 1) - 1);
                lo += st * f;
            }
        }
    }
    return lo;
}

// Index range [s_lo[L], s_hi[L]) of the descendants at level L = K .. maxL of
// the level-K cell of this workgroup: two lower bounds per level, one
// lane group each.
__device__ __forceinline__ void subtree_ranges(int K, int maxL, const DevHdr *hdr,
                                               const uint32_t *c_key, int *s_lo, int *s_hi) {
    const int r = (int)hdr->lev_off[K] + blockIdx.x;
    const unsigned long long kr = c_key[r];
    const int nsearch = 2 * (maxL - K + 1);
    for (int s0 = 0; s0 < nsearch; s0 += kSubBlock / kKeyGroup) {
        const int si = s0 + (int)threadIdx.x / kKeyGroup;
        int lo = 0, hi = 0, L = K;
        unsigned long long v = 0;
        if (si < nsearch) {
            L = K + si / 2;
            lo = (int)hdr->lev_off[L];
            hi = (int)hdr->lev_off[L + 1];
            v = (kr + (unsigned long long)(si & 1)) << (2 * (L - K));
        }
        const int pos = lower_key_group(c_key, lo, hi, v);
        if (si < nsearch && (threadIdx.x & (kKeyGroup - 1)) == 0) {
            if (si & 1) s_hi[L] = pos;
            else s_lo[L] = pos;
        }
    }
    __syncthreads();
}

__global__ void __launch_bounds__(kSubBlock)
    k_m2m_subtree(int K, int maxL, const DevHdr *hdr, const uint32_t *c_key,
                  const uint8_t *c_flags, const int *c_child,
                  const LevelTab *lt, const float *m2m, int kst, float *ahat) {
    __shared__ int s_lo[kMaxDepth + 2], s_hi[kMaxDepth + 2];
    __shared__ int s_lp[kMaxDepth + 2];
    __shared__ float2 tab[4 * kMaxPM1 * kMaxPM1], sa[kSubBlock / kWave][4 * kMaxPM1];
    const int PM1 = kst / 2;
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    load_tab(m2m, PM1, tab, lt, s_lp, K, maxL);
    subtree_ranges(K, maxL, hdr, c_key, s_lo, s_hi);
    for (int L = maxL - 1; L >= K; --L) {
        for (int j = s_lo[L] + w; j < s_hi[L]; j += kSubBlock / kWave)
            m2m_wave(j, s_lp[L], s_lp[L + 1], lane, PM1, c_flags, c_child, tab, sa[w], ahat);
        __syncthreads();
    }
}

// M2M for the parents of levels Lhi .. 2 (top levels, few cells): one workgroup.
// Multipoles of levels 0 and 1 are never read (list-2 and list-3 cells are of
// level >= 2), so the upward pass stops at level 2.
__global__ void __launch_bounds__(kTopBlock)
    k_m2m_top(int Lhi, const DevHdr *hdr, const uint8_t *c_flags, const int *c_child,
              const LevelTab *lt, const float *m2m, int kst, float *ahat) {
    __shared__ int s_lp[kMaxDepth + 2], s_off[kMaxDepth + 2];
    if (threadIdx.x < kMaxDepth + 2) s_off[threadIdx.x] = (int)hdr->lev_off[threadIdx.x];
    __shared__ float2 tab[4 * kMaxPM1 * kMaxPM1], sa[kTopBlock / kWave][4 * kMaxPM1];
    const int PM1 = kst / 2;
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    load_tab(m2m, PM1, tab, lt, s_lp, 2, Lhi + 1);
    for (int L = Lhi; L >= 2; --L) {
        for (int j = s_off[L] + w; j < s_off[L + 1]; j += kTopBlock / kWave)
            m2m_wave(j, s_lp[L], s_lp[L + 1], lane, PM1, c_flags, c_child, tab, sa[w], ahat);
        __syncthreads();
    }
}

// M2M / L2L for the parent cells [j0, j1) of one level: wavefront per cell.
__global__ void __launch_bounds__(kSubBlock)
    k_m2m_level(int L, int j0, int j1, const uint8_t *c_flags, const int *c_child,
                const LevelTab *lt, const float *m2m, int kst, float *ahat) {
    __shared__ int s_lp[kMaxDepth + 2];
    __shared__ float2 tab[4 * kMaxPM1 * kMaxPM1], sa[kSubBlock / kWave][4 * kMaxPM1];
    const int PM1 = kst / 2;
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    load_tab(m2m, PM1, tab, lt, s_lp, L, L + 1);
    const int j = j0 + blockIdx.x * (kSubBlock / kWave) + w;
    if (j < j1) {
        m2m_wave(j, s_lp[L], s_lp[L + 1], lane, PM1, c_flags, c_child, tab, sa[w], ahat);
    }
}

__global__ void __launch_bounds__(kSubBlock)
    k_l2l_level(int L, int j0, int j1, const uint8_t *c_flags, const int *c_child,
                const LevelTab *lt, const float *l2l, int kst, float *bhat) {
    __shared__ int s_lp[kMaxDepth + 2];
    __shared__ float2 tab[4 * kMaxPM1 * kMaxPM1], sb[kSubBlock / kWave][kMaxPM1];
    const int PM1 = kst / 2;
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    load_tab(l2l, PM1, tab, lt, s_lp, L, L + 1);
    const int j = j0 + blockIdx.x * (kSubBlock / kWave) + w;
    if (j < j1) {
        l2l_wave(j, s_lp[L], s_lp[L + 1], lane, PM1, c_flags, c_child, tab, sb[w], bhat);
    }
}

// L2L parents at levels 2 .. Lhi (children 3 .. Lhi + 1): one workgroup.
__global__ void __launch_bounds__(kTopBlock)
    k_l2l_top(int Lhi, const DevHdr *hdr, const uint8_t *c_flags, const int *c_child,
              const LevelTab *lt, const float *l2l, int kst, float *bhat) {
    __shared__ int s_lp[kMaxDepth + 2], s_off[kMaxDepth + 2];
    if (threadIdx.x < kMaxDepth + 2) s_off[threadIdx.x] = (int)hdr->lev_off[threadIdx.x];
    __shared__ float2 tab[4 * kMaxPM1 * kMaxPM1], sb[kTopBlock / kWave][kMaxPM1];
    const int PM1 = kst / 2;
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    load_tab(l2l, PM1, tab, lt, s_lp, 2, Lhi + 1);
    for (int L = 2; L <= Lhi; ++L) {
        for (int j = s_off[L] + w; j < s_off[L + 1]; j += kTopBlock / kWave)
            l2l_wave(j, s_lp[L], s_lp[L + 1], lane, PM1, c_flags, c_child, tab, sb[w], bhat);
        __syncthreads();
    }
}

// L2L parents at levels max(K, 2) .. maxL-1 of one subtree (children K+1 ..)
__global__ void __launch_bounds__(kSubBlock)
    k_l2l_subtree(int K, int maxL, const DevHdr *hdr, const uint32_t *c_key,
                  const uint8_t *c_flags, const int *c_child,
                  const LevelTab *lt, const float *l2l, int kst, float *bhat) {
    __shared__ int s_lo[kMaxDepth + 2], s_hi[kMaxDepth + 2];
    __shared__ int s_lp[kMaxDepth + 2];
    __shared__ float2 tab[4 * kMaxPM1 * kMaxPM1], sb[kSubBlock / kWave][kMaxPM1];
    const int PM1 = kst / 2;
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    load_tab(l2l, PM1, tab, lt, s_lp, max(K, 2), maxL);
    subtree_ranges(K, maxL, hdr, c_key, s_lo, s_hi);
    for (int L = max(K, 2); L <= maxL - 1; ++L) {
        for (int j = s_lo[L] + w; j < s_hi[L]; j += kSubBlock / kWave)
            l2l_wave(j, s_lp[L], s_lp[L + 1], lane, PM1, c_flags, c_child, tab, sb[w], bhat);
        __syncthreads();
    }
}

// M2L over list 2 on MFMA.  Tile = 16 consecutive cells of one level (all of
// order P), one workgroup of kM2lWaves wavefronts per tile.  Lane l: column
// n = l % 16 (target), k-quarter kq = l / 16.  The realified translation
// D[2(m-1)+{0,1}][2k+{0,1}] of each of the 40 offsets (|dx|,|dy| <= 3, not
// both <= 1) is the MFMA A operand, read from a table in lane order; the B
// operand is the source cell's scaled multipole (a^_k re/im along
// K = 2(P+1)).  Wavefront w takes the offsets oc = w, w + kM2lWaves, ... (an
// offset no target of the tile uses is skipped), so the chain of source
// gathers per wavefront is kM2lWaves times shorter; the partial results are
// summed through LDS.
constexpr int kM2lWaves = 4;

template <int P>
__device__ __forceinline__ void m2l_tile(int j0, int cnt, int w, int lane, int (*smap)[49],
                                         float *red, const uint32_t *c_key, const int *l2,
                                         const uint8_t *n2, const float *R, const float *ahat,
                                         int kst, const uint8_t *n4, const float *b4,
                                         float *bhat) {
    constexpr int K = 2 * (P + 1), KS = (K + 3) / 4, MT = (2 * P + 15) / 16;
    const int n = lane % 16, kq = lane / 16;
    const int j = j0 + n;
    // source map of the tile: slot (dx+3) + 7 (dy+3) -> source cell or -1
    for (int t = threadIdx.x; t < 16 * 49; t += blockDim.x) smap[t / 49][t % 49] = -1;
    __syncthreads();
    {
        const int tn = threadIdx.x % 16, tu = threadIdx.x / 16;  // 16 entries of list 2 at a time
        const int tj = j0 + tn;
        if (tn < cnt) {
            const uint32_t kj = c_key[tj];
            const int jx = (int)compact1by1(kj), jy = (int)compact1by1(kj >> 1);
            const int m2 = n2[tj];
            for (int u = tu; u < m2; u += (int)blockDim.x / 16) {
                const int sidx = l2[27 * tj + u];
                const uint32_t ks = c_key[sidx];
                const int dx = (int)compact1by1(ks) - jx, dy = (int)compact1by1(ks >> 1) - jy;
                smap[tn][(dx + 3) + 7 * (dy + 3)] = sidx;
            }
        }
    }
    __syncthreads();
    v4f acc[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) acc[mt] = v4f{0.f, 0.f, 0.f, 0.f};
    int o = 0;
    for (int slot = 0; slot < 49; ++slot) {
        const int dx = slot % 7 - 3, dy = slot / 7 - 3;
        if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1) continue;
        const int oc = o++;
        if (oc % kM2lWaves != w) continue;
        const int sidx = smap[n][slot];
        if (__ballot(sidx >= 0) == 0) continue;  // no target of the tile uses this offset
        const float *src = ahat + (size_t)(sidx < 0 ? 0 : sidx) * kst;
        float b[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int k = ks * 4 + kq;
            b[ks] = (sidx >= 0 && k < K) ? src[k] : 0.f;
        }
        const float *Ro = R + (size_t)oc * KS * MT * kWave + lane;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
                acc[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(Ro[(ks * MT + mt) * kWave], b[ks],
                                                               acc[mt], 0, 0, 0);
    }
    // sum of the wavefronts' partial results: red[w][(mt * 4 + v) * 64 + lane]
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int v = 0; v < 4; ++v) red[(w * 8 + mt * 4 + v) * kWave + lane] = acc[mt][v];
    __syncthreads();
    if (w != 0) return;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            float sum = 0.f;
#pragma unroll
            for (int u = 0; u < kM2lWaves; ++u) sum += red[(u * 8 + mt * 4 + v) * kWave + lane];
            const int row = mt * 16 + 4 * kq + v;  // row of (Re b^_1, Im b^_1, Re b^_2, ...)
            if (n < cnt && row < 2 * P) {
                const size_t idx = (size_t)j * kst + 2 + row;
                // list 4 by P2L, formed beforehand into b4 (cells with a list 4)
                bhat[idx] = (b4 && n4[j]) ? sum + b4[idx] : sum;
            }
        }
    }
}

// Table addresses per order (kernel argument).
struct M2lTabs {
    const float *r[17];
};

// All tiles of every order in one launch: a workgroup per tile, dispatched on
// the tile's order (tiles are grouped by order).
__global__ void __launch_bounds__(kWave *kM2lWaves)
    k_m2l_mfma(const int2 *tiles, const uint8_t *tile_p, int ntiles, const uint32_t *c_key,
               const int *l2, const uint8_t *n2, M2lTabs tabs, const float *ahat, int kst,
               const uint8_t *n4, const float *b4, float *bhat) {
    __shared__ int smap[16][49];
    __shared__ float red[kM2lWaves * 8 * kWave];
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const int tile = blockIdx.x;
    if (tile >= ntiles) return;
    const int2 td = tiles[tile];
    const int p = tile_p[tile];
#define M2L_CASE(P)                                                                          \
    case P:                                                                                  \
        m2l_tile<P>(td.x, td.y, w, lane, smap, red, c_key, l2, n2, tabs.r[P], ahat, kst, n4,  \
                    b4, bhat);                                                               \
        break;
    switch (p) {
        M2L_CASE(1) M2L_CASE(2) M2L_CASE(3) M2L_CASE(4) M2L_CASE(5) M2L_CASE(6) M2L_CASE(7)
        M2L_CASE(8) M2L_CASE(9) M2L_CASE(10) M2L_CASE(11) M2L_CASE(12) M2L_CASE(13)
        M2L_CASE(14) M2L_CASE(15) M2L_CASE(16)
        default: break;
    }
#undef M2L_CASE
}

// Per-level table (order clamped to 1..16, side R / 2^L, its inverse) and the
// 16-cell M2L tiles of levels 2..maxL, grouped by order (ascending) and then
// by level, from the device-resident orders: one workgroup.
__global__ void __launch_bounds__(256)
    k_mixed_setup(const uint8_t *orders, const DevHdr *hdr, int maxL, LevelTab *lt, int2 *tiles,
                  uint8_t *tile_p) {
    __shared__ int s_first[kMaxDepth + 2], s_p[kMaxDepth + 2], s_off[kMaxDepth + 2],
        s_cnt[kMaxDepth + 2];
    const int tid = threadIdx.x;
    if (tid < kMaxDepth + 2) {
        const int p = min(max((int)orders[tid], 1), 16);
        s_p[tid] = p;
        s_off[tid] = (int)hdr->lev_off[tid];
        s_cnt[tid] = (int)hdr->lev_cnt[tid];
        const double side = hdr->root_len / ldexp(1.0, tid);
        lt->p[tid] = p;
        lt->s[tid] = (float)side;
        lt->inv_s[tid] = (float)(1.0 / side);
    }
    __syncthreads();
    if (tid >= 2 && tid <= maxL) {
        // tiles of the levels ordered before this one: lower order, or the
        // same order and a lower level
        const int L = tid, pL = s_p[L];
        int off = 0;
#pragma unroll
        for (int u = 2; u < kMaxDepth + 2; ++u)
            if (u <= maxL && (s_p[u] < pL || (s_p[u] == pL && u < L))) off += (s_cnt[u] + 15) / 16;
        s_first[L] = off;
    }
    __syncthreads();
    for (int L = 2; L <= maxL; ++L) {
        const int j0 = s_off[L], cnt = s_cnt[L];
        const int nt = (cnt + 15) / 16;
        for (int u = tid; u < nt; u += blockDim.x) {
            tiles[s_first[L] + u] = make_int2(j0 + 16 * u, min(16, cnt - 16 * u));
            tile_p[s_first[L] + u] = (uint8_t)s_p[L];
        }
    }
}

// List 4, mode 2: P2L from the active particles of coarser leaves,
//   b4_m = -q/m (s_t / (z_j - z_t))^m  (cells with a list 4 only),
// into its own buffer: it needs only the particles and list 4, so it runs
// before M2L (beside the upward pass), and the M2L epilogue adds it.
// One wavefront per target cell; lanes stride over the source particles and
// keep the running sums of q w^m for every order (<= 16) in registers; the
// -1/m factor is applied once per order after the lanes' sums are reduced
// through LDS (lane l writes column l, lane m sums row m; stride 65), in two
// halves of 8 orders so that 8 wavefronts per SIMD fit (4 with the whole
// table).
constexpr int kP2lWaves = 4;

__global__ void __launch_bounds__(kWave *kP2lWaves)
    k_p2l_32(int j0, int nc, const uint8_t *c_level, const int *l4, const uint8_t *n4,
             const double *c_cx, const double *c_cy, const uint32_t *c_begin,
             const uint32_t *c_end, const float2 *oxy, const float *qf, const uint8_t *as,
             const LevelTab *lt, int kst, float *b4) {
    __shared__ float2 red[kP2lWaves][8][kWave + 1];
    const int PM1 = kst / 2;
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const int j = j0 + blockIdx.x * kP2lWaves + w;
    if (j >= nc) return;
    const int m4 = n4[j];
    if (m4 == 0) return;
    const int L = c_level[j];
    const int p = lt->p[L];
    const double cx = c_cx[j], cy = c_cy[j];
    const float inv_st = lt->inv_s[L];
    float ar[16], ai[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) ar[m] = ai[m] = 0.f;
    for (int u = 0; u < m4; ++u) {
        const int sc = l4[8 * j + u];
        const float dcx = (float)(c_cx[sc] - cx), dcy = (float)(c_cy[sc] - cy);
        for (uint32_t i = c_begin[sc] + lane; i < c_end[sc]; i += kWave) {
            if (!as[i]) continue;
            const float2 o = oxy[i];
            const cf wz = cinv(cf{(dcx + o.x) * inv_st, (dcy + o.y) * inv_st});
            cf pw = cscale(wz, qf[i]);
#pragma unroll
            for (int m = 1; m <= 16; ++m) {
                if (m <= p) {
                    ar[m - 1] += pw.x;
                    ai[m - 1] += pw.y;
                    pw = cmul(pw, wz);
                }
            }
        }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int mm = 1; mm <= 8; ++mm) {
            const int mo = 8 * h + mm;
            if (mo <= p) red[w][mm - 1][lane] = make_float2(ar[mo - 1], ai[mo - 1]);
        }
        wave_sync();
        const int m = lane + 1;
        if (m > 8 * h && m <= 8 * h + 8 && m <= p) {
            const int row = m - 1 - 8 * h;
            float2 s0 = make_float2(0.f, 0.f), s1 = s0, s2 = s0, s3 = s0;
#pragma unroll 4
            for (int l = 0; l < kWave; l += 4) {
                const float2 a0 = red[w][row][l], a1 = red[w][row][l + 1];
                const float2 a2 = red[w][row][l + 2], a3 = red[w][row][l + 3];
                s0.x += a0.x; s0.y += a0.y;
                s1.x += a1.x; s1.y += a1.y;
                s2.x += a2.x; s2.y += a2.y;
                s3.x += a3.x; s3.y += a3.y;
            }
            const float c = -1.f / (float)m;
            const cf sum = {((s0.x + s1.x) + (s2.x + s3.x)) * c, ((s0.y + s1.y) + (s2.y + s3.y)) * c};
            stc(b4, (size_t)j * PM1 + m, sum);
        }
        wave_sync();
    }
}

// List 4, mode 1: M2L from the scaled multipoles of the coarser leaves;
// thread per (target cell, m):
//   b^_m += sum_k T_mk a^_k (s_s w)^k (s_t w)^m,  w = 1/(z_s - z_t).
__global__ void __launch_bounds__(256)
    k_m2l4_32(int j0, int nc, const uint8_t *c_level, const int *l4, const uint8_t *n4,
              const double *c_cx, const double *c_cy, const LevelTab *lt, const float *tm2l,
              const float *ahat, int kst, float *bhat) {
    const int PM1 = kst / 2;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)(nc - j0) * (PM1 - 1)) return;
    const int j = j0 + (int)(idx / (PM1 - 1)), m = 1 + (int)(idx % (PM1 - 1));
    const int m4 = n4[j];
    const int L = c_level[j];
    if (m4 == 0 || m > lt->p[L]) return;
    const double cx = c_cx[j], cy = c_cy[j];
    const float st = lt->s[L];
    cf acc = {0.f, 0.f};
    for (int u = 0; u < m4; ++u) {
        const int s = l4[8 * j + u];
        const int Ls = c_level[s];
        const int ps = lt->p[Ls];
        const cf dz = {(float)((c_cx[s] - cx) / st), (float)((c_cy[s] - cy) / st)};
        const cf wt = cinv(dz);                       // s_t w
        const cf ws = cscale(wt, lt->s[Ls] / st);     // s_s w
        const float *row = tm2l + (size_t)m * PM1;
        cf sacc = cscale(ldc(ahat, (size_t)s * PM1 + ps), row[ps]);
        for (int k = ps - 1; k >= 0; --k)
            sacc = cadd(cmul(sacc, ws), cscale(ldc(ahat, (size_t)s * PM1 + k), row[k]));
        acc = cadd(acc, cmul(sacc, cpowi(wt, m)));
    }
    const size_t o = (size_t)j * PM1 + m;
    stc(bhat, o, cadd(ldc(bhat, o), acc));
}

// Evaluation: one wavefront per leaf.  Sources are staged in LDS in rounds of
// kStage as (x - c_t, y - c_t, q, index), the centre shift done in FP64 and
// rounded once.  Lanes split into G groups of 64/G (G = 8, 4, 2, 1 for leaves
// of up to 8, 16, 32, more targets): a group owns the targets, the groups
// split the sources and list-3 cells and are reduced with cross-lane shuffles.
// Every round is padded to a multiple of G with far-away sources of zero
// strength, so all lanes run one trip count with a compile-time stride.  Two
// particles at one position always share a leaf, so only the leaf's own
// particles (a phase of their own) need the self test, by index.
constexpr int kEvalWaves = 4;
constexpr int kStage = 256;  // sources staged per wavefront per round

// Padding source: r^2 ~ 1e36 stays finite and q = 0 contributes exactly 0.
__device__ __forceinline__ float4 far_source() {
    return make_float4(1e18f, 1e18f, 0.f, __uint_as_float(0xffffffffu));
}

// P2P of one target over the staged sources g, g + G, ... (st = stage + g,
// nu = padded round length / G).
// Regularization factor of a near-field pair in FP32 (regul_func_2d,
// regularization_function_2d.cpp:11-37) at rho^2 = r^2 / sij^2.  Where the
// reference's 1 - exp form cancels (small h = rho^2 / 2), the Taylor series
// of the same function is used: Gaussian 1 - e^-h below h = 1/8, super-
// Gaussian 1 - (1 + h) e^-h below h = 1/2 (truncation under 5e-7 relative).
// Such close pairs are rare, so the series is only evaluated by a wavefront
// that has one (a wave-uniform branch).  Far padding sources can give
// rho^2 = inf: the Gaussian takes it (e = 0, factor 1); the other two clamp
// it, so that (x + 1)^2 and (1 + h) e stay finite.
template <int ICUT>
__device__ __forceinline__ float regul32(float rho2) {
    if (ICUT == 1) {
        const float x = fminf(rho2, 1e15f);
        return x * (x + 2.f) * __builtin_amdgcn_rcpf((x + 1.f) * (x + 1.f));
    }
    const float h = 0.5f * (ICUT == 3 ? fminf(rho2, 1e15f) : rho2);
    const float e = __expf(-h);
    if (ICUT == 2) {
        float f = 1.f - e;
        const bool near = h < 0.125f;
        if (__ballot(near))
            f = near ? h * (1.f - h * (0.5f - h * (1.f / 6 - h * (1.f / 24 - h * (1.f / 120)))))
                     : f;
        return f;
    }
    float f = 1.f - (1.f + h) * e;
    const bool near = h < 0.5f;
    if (__ballot(near))
        f = near ? h * h * (0.5f - h * (1.f / 3 - h * (1.f / 8 - h * (1.f / 30 -
                   h * (1.f / 144 - h * (1.f / 840 - h * (1.f / 5760)))))))
                 : f;
    return f;
}

template <int G, bool SELF, int ICUT>
__device__ __forceinline__ void p2p_round(const float4 *st, const float *sst, int nu, float2 ti,
                                          float ts4, uint32_t i, float &ex, float &ey) {
    // (the regularized kernels' wave-uniform branch keeps their loop rolled)
    constexpr int kUnroll = ICUT ? 1 : 4;
#pragma unroll kUnroll
    for (int u = 0; u < nu; ++u) {
        const float4 sv = st[u * G];
        const float dx = ti.x - sv.x, dy = ti.y - sv.y;
        const float r2 = dx * dx + dy * dy;
        float inv = __builtin_amdgcn_rcpf(r2);
        if (ICUT == 0) {
            if (SELF) inv = (__float_as_uint(sv.w) == i) ? 0.f : inv;
        } else {
            // sij^2 = (si^2 + sj^2) / 4 (both terms staged as s^2 / 4);
            // pairs at r = 0 (the target itself included) add nothing
            const float rho2 = r2 * __builtin_amdgcn_rcpf(ts4 + sst[u * G]);
            inv = r2 > 0.f ? regul32<ICUT>(rho2) * inv : 0.f;
        }
        const float qr = sv.z * inv;
        ex += qr * dx;
        ey += qr * dy;
    }
}

struct EvalArgs {
    const uint32_t *leaf_idx;
    const uint8_t *c_level, *c_flags;
    const uint32_t *c_begin, *c_end;
    const double *c_cx, *c_cy;
    const uint32_t *off1;
    const int *l1;
    const uint32_t *off3;
    const int *l3;
    const float2 *oxy;
    const float *qf;
    const double *ss;  // sorted sizes (regularized near field only)
    const uint32_t *perm;
    const LevelTab *lt;
    const float *ahat, *bhat;
    int kst, mask, velocity;
    double u_inf, v_inf;
    double *out_a, *out_b;
    unsigned *flags;  // header flags (DUPLICATE_POINTS)
};

template <int G, int ICUT>
__device__ __forceinline__ void eval_leaf(const EvalArgs &A, int t, int lane, float4 *stg,
                                          float *sst, uint32_t *s_off, uint32_t *s_beg,
                                          float *s_dx, float *s_dy) {
    constexpr int LG = kWave / G;
    const int PM1 = A.kst / 2;
    const int g = lane / LG, tl = lane % LG;
    const int j = __builtin_amdgcn_readfirstlane((int)A.leaf_idx[t]);
    const uint32_t b = (uint32_t)__builtin_amdgcn_readfirstlane((int)A.c_begin[j]);
    const uint32_t e = (uint32_t)__builtin_amdgcn_readfirstlane((int)A.c_end[j]);
    const double tcx = A.c_cx[j], tcy = A.c_cy[j];
    const bool own_src = (A.c_flags[j] & kActive) != 0;  // the leaf is in its own list 1
    const uint32_t a1 = A.off1[t], z1 = A.off1[t + 1], a3 = A.off3[t], z3 = A.off3[t + 1];
    const int L = A.c_level[j];
    const int p = A.lt->p[L];
    const float inv_s = A.lt->inv_s[L];
    for (uint32_t tb = b; tb < e; tb += LG) {
        const uint32_t i = tb + tl;
        const bool valid = i < e;
        const float2 ti = valid ? A.oxy[i] : make_float2(0.f, 0.f);
        const float ts4 = (ICUT && valid) ? (float)(0.25 * A.ss[i] * A.ss[i]) : 0.f;
        // each staging round sums in FP32; the round totals add in FP64, so
        // leaves with very many sources do not lose the FP32 accumulation
        double exd = 0.0, eyd = 0.0;
        bool dup = false;
        if ((A.mask & 1) && own_src) {  // P2P from the leaf's own particles
            for (uint32_t base = b; base < e; base += kStage) {
                const int fill = (int)min<uint32_t>((uint32_t)kStage, e - base);
                const int fillp = (fill + G - 1) / G * G;
                for (int r = lane; r < fillp; r += kWave) {
                    if (r < fill) {
                        const uint32_t src = base + (uint32_t)r;
                        const float2 o = A.oxy[src];
                        stg[r] = make_float4(o.x, o.y, A.qf[src], __uint_as_float(src));
                        if (ICUT) sst[r] = (float)(0.25 * A.ss[src] * A.ss[src]);
                    } else {
                        stg[r] = far_source();
                        if (ICUT) sst[r] = 0.f;
                    }
                }
                wave_sync();
                float ex = 0.f, ey = 0.f;
                p2p_round<G, true, ICUT>(stg + g, sst + g, fillp / G, ti, ts4, i, ex, ey);
                // two particles at one position (only ever in one leaf) give
                // r = 0 and a non-finite sum: the singular kernel's error
                if (ICUT == 0) dup |= valid && !(isfinite(ex) && isfinite(ey));
                exd += (double)ex;
                eyd += (double)ey;
                wave_sync();
            }
        }
        // P2P from the other list-1 cells.  Up to 64 cells at a time: lane k
        // fetches cell k's range and centre shift (one round of loads), a wave
        // scan gives each cell its offset in the concatenated source list, and
        // the sources are staged in rounds of kStage, each lane locating its
        // source's cell by binary search over the offsets.
        for (uint32_t c0 = a1; (A.mask & 1) && c0 < z1; c0 += kWave) {
            const int ncell = (int)min<uint32_t>(kWave, z1 - c0);
            uint32_t cnt = 0, beg = 0;
            float dcx = 0.f, dcy = 0.f;
            if (lane < ncell) {
                const int sc = A.l1[c0 + lane];
                if (sc != j) {  // the own leaf was done above
                    beg = A.c_begin[sc];
                    cnt = A.c_end[sc] - beg;
                    dcx = (float)(A.c_cx[sc] - tcx);
                    dcy = (float)(A.c_cy[sc] - tcy);
                }
            }
            uint32_t inc = cnt;
            for (int o = 1; o < kWave; o <<= 1) {
                const uint32_t y = __shfl_up(inc, o, kWave);
                if (lane >= o) inc += y;
            }
            // wave-uniform: the round loops run on the scalar unit
            const uint32_t total =
                (uint32_t)__builtin_amdgcn_readfirstlane((int)__shfl(inc, kWave - 1, kWave));
            s_off[lane] = inc - cnt;
            s_beg[lane] = beg;
            s_dx[lane] = dcx;
            s_dy[lane] = dcy;
            wave_sync();
            for (uint32_t base = 0; base < total; base += kStage) {
                const int fill = (int)min<uint32_t>((uint32_t)kStage, total - base);
                const int fillp = (fill + G - 1) / G * G;
                for (int r = lane; r < fillp; r += kWave) {
                    if (r >= fill) {
                        stg[r] = far_source();
                        if (ICUT) sst[r] = 0.f;
                        continue;
                    }
                    const uint32_t gi = base + (uint32_t)r;
                    // last cell whose offset is <= gi (empty cells share the
                    // offset of the next one and are never the last)
                    int lo = 0, hi = ncell - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (s_off[mid] <= gi) lo = mid;
                        else hi = mid - 1;
                    }
                    const uint32_t src = s_beg[lo] + (gi - s_off[lo]);
                    const float2 o = A.oxy[src];
                    stg[r] = make_float4(o.x + s_dx[lo], o.y + s_dy[lo], A.qf[src],
                                         __uint_as_float(src));
                    if (ICUT) sst[r] = (float)(0.25 * A.ss[src] * A.ss[src]);
                }
                wave_sync();
                float ex = 0.f, ey = 0.f;
                p2p_round<G, false, ICUT>(stg + g, sst + g, fillp / G, ti, ts4, i, ex, ey);
                exd += (double)ex;
                eyd += (double)ey;
                wave_sync();
            }
        }
        // M2P over list 3: F = (1/s) sum a^_k xi^-(k+1), xi = (z - zc)/s
        for (uint32_t v = a3 + g; (A.mask & 2) && v < z3; v += G) {
            const int sc = A.l3[v];
            const int Ls = A.c_level[sc];
            const int ps = A.lt->p[Ls];
            const float is = A.lt->inv_s[Ls];
            const float dcx = (float)(tcx - A.c_cx[sc]), dcy = (float)(tcy - A.c_cy[sc]);
            const cf wv = cinv(cf{(ti.x + dcx) * is, (ti.y + dcy) * is});
            const size_t cb = (size_t)sc * PM1;
            cf f = ldc(A.ahat, cb + ps);
            for (int k = ps - 1; k >= 0; --k) f = cadd(cmul(f, wv), ldc(A.ahat, cb + k));
            f = cscale(cmul(f, wv), is);
            if (valid) {
                exd += (double)f.x;
                eyd -= (double)f.y;
            }
        }
#pragma unroll
        for (int m = LG; m < kWave; m <<= 1) {
            exd += __shfl_xor(exd, m, kWave);
            eyd += __shfl_xor(eyd, m, kWave);
        }
        if (valid && g == 0) {
            // local expansions exist from level 2 down (M2L and L2L form no
            // other; b^ of a shallower leaf is never written): a leaf of level
            // 0 or 1 -- a tree one level deep -- has no far field
            if ((A.mask & 4) && L >= 2) {
                // L2P: F = (1/s) sum_m m b^_m zeta^(m-1)
                const cf z = {ti.x * inv_s, ti.y * inv_s};
                const size_t cb = (size_t)j * PM1;
                cf f = cscale(ldc(A.bhat, cb + p), (float)p);
                for (int m = p - 1; m >= 1; --m)
                    f = cadd(cmul(f, z), cscale(ldc(A.bhat, cb + m), (float)m));
                f = cscale(f, inv_s);
                exd += (double)f.x;
                eyd -= (double)f.y;
            }
            const uint32_t o = A.perm[i];
            if (A.velocity) {
                A.out_a[o] = eyd + A.u_inf;
                A.out_b[o] = -exd + A.v_inf;
            } else {
                A.out_a[o] = exd;
                A.out_b[o] = eyd;
            }
        }
        if (dup) atomicOr(A.flags, kFlagDuplicate);
    }
}

template <int ICUT>
__global__ void __launch_bounds__(kWave *kEvalWaves) k_eval32(EvalArgs A, int nl) {
    __shared__ float4 stage[kEvalWaves][kStage];
    __shared__ float stage_s[ICUT ? kEvalWaves : 1][ICUT ? kStage : 1];
    __shared__ uint32_t s_off[kEvalWaves][kWave], s_beg[kEvalWaves][kWave];
    __shared__ float s_dx[kEvalWaves][kWave], s_dy[kEvalWaves][kWave];
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    // (an XCD-contiguous leaf order cut the L2 misses but unbalanced the XCDs:
    // the dense band's leaves landed on a few of them, 0.129 -> 0.140 ms)
    const int t = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * kEvalWaves + w));
    if (t >= nl) return;
    const int j = (int)A.leaf_idx[t];
    const int nt = (int)(A.c_end[j] - A.c_begin[j]);
    float *sst = stage_s[ICUT ? w : 0];
    if (nt <= 8)
        eval_leaf<8, ICUT>(A, t, lane, stage[w], sst, s_off[w], s_beg[w], s_dx[w], s_dy[w]);
    else if (nt <= 16)
        eval_leaf<4, ICUT>(A, t, lane, stage[w], sst, s_off[w], s_beg[w], s_dx[w], s_dy[w]);
    else if (nt <= 32)
        eval_leaf<2, ICUT>(A, t, lane, stage[w], sst, s_off[w], s_beg[w], s_dx[w], s_dy[w]);
    else
        eval_leaf<1, ICUT>(A, t, lane, stage[w], sst, s_off[w], s_beg[w], s_dx[w], s_dy[w]);
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static const int kOffsets[40][2] = {
#define O(dx, dy) {dx, dy}
    O(-3, -3), O(-2, -3), O(-1, -3), O(0, -3), O(1, -3), O(2, -3), O(3, -3),
    O(-3, -2), O(-2, -2), O(-1, -2), O(0, -2), O(1, -2), O(2, -2), O(3, -2),
    O(-3, -1), O(-2, -1), O(2, -1), O(3, -1),
    O(-3, 0), O(-2, 0), O(2, 0), O(3, 0),
    O(-3, 1), O(-2, 1), O(2, 1), O(3, 1),
    O(-3, 2), O(-2, 2), O(-1, 2), O(0, 2), O(1, 2), O(2, 2), O(3, 2),
    O(-3, 3), O(-2, 3), O(-1, 3), O(0, 3), O(1, 3), O(2, 3), O(3, 3),
#undef O
};

void mixed_tables(int pmax, std::vector<float> &m2m, std::vector<float> &l2l,
                  std::vector<float> &tm2l, std::vector<std::vector<float>> &rlane) {
    const int PM1 = pmax + 1;
    // child centre minus parent centre over the child side: (+-1/2, +-1/2)
    const double ex[4] = {-0.5, 0.5, -0.5, 0.5}, ey[4] = {-0.5, -0.5, 0.5, 0.5};
    m2m.assign((size_t)4 * PM1 * PM1 * 2, 0.f);
    l2l.assign((size_t)4 * PM1 * PM1 * 2, 0.f);
    for (int q = 0; q < 4; ++q) {
        for (int m = 0; m <= pmax; ++m) {
            for (int k = 0; k <= pmax; ++k) {
                // M2M: C(m,k) e^(m-k) 2^-m, k <= m ; L2L: C(k,m) e^(k-m) 2^-k, k >= m
                const int d = k <= m ? m - k : k - m;
                double pr = 1.0, pi = 0.0;
                for (int t = 0; t < d; ++t) {
                    const double nr = pr * ex[q] - pi * ey[q], ni = pr * ey[q] + pi * ex[q];
                    pr = nr;
                    pi = ni;
                }
                const size_t o = (((size_t)q * PM1 + m) * PM1 + k) * 2;
                if (k <= m) {
                    const double c = exact_binom(m, k) * std::ldexp(1.0, -m);
                    m2m[o] = (float)(c * pr);
                    m2m[o + 1] = (float)(c * pi);
                }
                if (k >= m) {
                    const double c = exact_binom(k, m) * std::ldexp(1.0, -k);
                    l2l[o] = (float)(c * pr);
                    l2l[o + 1] = (float)(c * pi);
                }
            }
        }
    }
    // T_mk (exact binomials) for the vector list-4 M2L
    tm2l.assign((size_t)PM1 * PM1, 0.f);
    for (int m = 1; m <= pmax; ++m) {
        tm2l[(size_t)m * PM1] = (float)(-1.0 / m);
        for (int k = 1; k <= pmax; ++k)
            tm2l[(size_t)m * PM1 + k] =
                (float)(-(exact_binom(m + k - 1, k - 1) * ((k & 1) ? -1.0 : 1.0) / k));
    }
    // realified M2L matrices per offset, in MFMA lane order, one table per p
    rlane.assign(pmax + 1, {});
    for (int p = 1; p <= pmax; ++p) {
        const int K = 2 * (p + 1), KS = (K + 3) / 4, MT = (2 * p + 15) / 16;
        std::vector<float> &R = rlane[p];
        R.assign((size_t)40 * KS * MT * kWave, 0.f);
        for (int o = 0; o < 40; ++o) {
            const double dx = kOffsets[o][0], dy = kOffsets[o][1];
            const double r2 = dx * dx + dy * dy;
            const double wr = dx / r2, wi = -dy / r2;  // 1/delta
            std::vector<double> pwr(2 * p + 2), pwi(2 * p + 2);
            pwr[0] = 1.0;
            pwi[0] = 0.0;
            for (int t = 1; t <= 2 * p + 1; ++t) {
                pwr[t] = pwr[t - 1] * wr - pwi[t - 1] * wi;
                pwi[t] = pwr[t - 1] * wi + pwi[t - 1] * wr;
            }
            for (int m = 1; m <= p; ++m) {
                for (int k = 0; k <= p; ++k) {
                    const double T = k == 0 ? -1.0 / m
                                            : -(exact_binom(m + k - 1, k - 1) *
                                                ((k & 1) ? -1.0 : 1.0) / k);
                    const double Mr = T * pwr[m + k], Mi = T * pwi[m + k];
                    // rows 2(m-1) (Re) and 2(m-1)+1 (Im); columns 2k (Re a^), 2k+1 (Im a^)
                    const int r0 = 2 * (m - 1), r1 = r0 + 1, c0 = 2 * k, c1 = c0 + 1;
                    const double val[2][2] = {{Mr, -Mi}, {Mi, Mr}};
                    const int rows[2] = {r0, r1}, cols[2] = {c0, c1};
                    for (int a = 0; a < 2; ++a)
                        for (int bq = 0; bq < 2; ++bq) {
                            const int row = rows[a], col = cols[bq];
                            const int mt = row / 16, ks = col / 4;
                            // lane holds A[row % 16][col % 4] with lane = (row%16) + 16 (col%4)
                            const int lane = (row % 16) + 16 * (col % 4);
                            R[(((size_t)o * KS + ks) * MT + mt) * kWave + lane] = (float)val[a][bq];
                        }
                }
            }
        }
    }
}

int mixed_tile_count(const Ctx &c) {
    const TreeInfo &ti = c.info;
    int ntiles = 0;
    for (int L = 2; L <= ti.max_level; ++L)
        ntiles += div_up((long long)(ti.lev_off[L + 1] - ti.lev_off[L]), 16);
    return ntiles;
}

int mixed_level_setup(Ctx &c, const uint8_t *orders, LevelTab *lt, int2 *tiles,
                      uint8_t *tile_p) {
    const TreeInfo &ti = c.info;
    const int ntiles = mixed_tile_count(c);
    k_mixed_setup<<<1, 256, 0, c.stream>>>(orders, c.hdr_d, ti.max_level, lt, tiles, tile_p);
    VPM_LAUNCH_CHECK();
    return ntiles;
}

// Split levels of the three-tier passes: K = first level with >= kSubRoots
// cells (subtrees below), T = first level with >= kTopCells cells (one
// workgroup above T).
static void tier_levels(const TreeInfo &ti, int &K, int &T) {
    const int maxL = ti.max_level;
    K = T = maxL;
    for (int L = maxL; L >= 1; --L) {
        const int cnt = (int)(ti.lev_off[L + 1] - ti.lev_off[L]);
        if (cnt >= kSubRoots) K = L;
        if (cnt >= kTopCells) T = L;
    }
    T = std::min(T, K);
}

// Upward pass (P2M, M2M) on stream s; the buffers come from fmm32_prepare.
// No clearing is needed: P2M / M2M write every coefficient a cell's readers
// use (zero for inactive cells), the list-2 M2L writes b^_1..p of every cell
// of level >= 2 before list 4 and L2L add to them, and the evaluation reads
// b^ only of leaves of level >= 2.
void fmm32_p2m(Ctx &c, const Mixed32 &mx, hipStream_t s) {
    const int nl = c.info.n_leaves;
    k_p2m32<<<div_up(nl, kP2mWaves), kWave * kP2mWaves, 0, s>>>(
        (const uint32_t *)c.leaf_idx.p, nl, (const uint8_t *)c.c_level.p,
        (const uint8_t *)c.c_flags.p, (const uint32_t *)c.c_begin.p, (const uint32_t *)c.c_end.p,
        (const double *)c.c_cx.p, (const double *)c.c_cy.p, (const double *)c.xs.p,
        (const double *)c.ys.p, (const double *)c.qs.p, (const uint8_t *)c.acts.p, mx.lt,
        2 * (mx.pmax + 1), (float2 *)c.oxy.p, (float *)c.qf.p, (float *)c.ak.p);
    VPM_LAUNCH_CHECK();
}

void fmm32_m2m(Ctx &c, const Mixed32 &mx, hipStream_t s, const FmmMarks &mk) {
    const TreeInfo &ti = c.info;
    const int kst = 2 * (mx.pmax + 1);
    float *ahat = (float *)c.ak.p;
    const uint8_t *c_flags = (const uint8_t *)c.c_flags.p;
    const uint32_t *c_key = (const uint32_t *)c.c_key.p;
    const int *c_child = (const int *)c.c_child.p;
    const int maxL = ti.max_level;
    int K, T;
    tier_levels(ti, K, T);
    const int nK = (int)(ti.lev_off[K + 1] - ti.lev_off[K]);
    if (K < maxL)
        k_m2m_subtree<<<nK, kSubBlock, 0, s>>>(K, maxL, c.hdr_d, c_key, c_flags, c_child, mx.lt,
                                               mx.m2m, kst, ahat);
    for (int L = K - 1; L >= std::max(T, 2); --L) {
        const int j0 = (int)ti.lev_off[L], j1 = (int)ti.lev_off[L + 1];
        k_m2m_level<<<div_up(j1 - j0, kSubBlock / kWave), kSubBlock, 0, s>>>(
            L, j0, j1, c_flags, c_child, mx.lt, mx.m2m, kst, ahat);
    }
    if (std::min(T, K) - 1 >= 2)
        k_m2m_top<<<1, kTopBlock, 0, s>>>(std::min(T, K) - 1, c.hdr_d, c_flags, c_child, mx.lt,
                                          mx.m2m, kst, ahat);
    mk.mark(3, s);
    VPM_LAUNCH_CHECK();
}

void fmm32_upward(Ctx &c, const Mixed32 &mx, hipStream_t s, const FmmMarks &mk) {
    fmm32_p2m(c, mx, s);
    fmm32_m2m(c, mx, s, mk);
}

void fmm32_p2l(Ctx &c, const Mixed32 &mx, hipStream_t s) {
    const TreeInfo &ti = c.info;
    const int nc = ti.n_cells, j2 = (int)ti.lev_off[2];
    if (nc <= j2) return;
    k_p2l_32<<<div_up(nc - j2, kP2lWaves), kWave * kP2lWaves, 0, s>>>(
        j2, nc, (const uint8_t *)c.c_level.p, (const int *)c.l4.p, (const uint8_t *)c.n4.p,
        (const double *)c.c_cx.p, (const double *)c.c_cy.p, (const uint32_t *)c.c_begin.p,
        (const uint32_t *)c.c_end.p, (const float2 *)c.oxy.p, (const float *)c.qf.p,
        (const uint8_t *)c.acts.p, mx.lt, 2 * (mx.pmax + 1), (float *)c.b4.p);
    VPM_LAUNCH_CHECK();
}

void fmm32_rest(Ctx &c, const Mixed32 &mx, int list4_mode, int stage_mask, double *out_a,
                double *out_b, bool velocity, double u_inf, double v_inf, const FmmMarks &mk) {
    hipStream_t s = c.stream;
    const TreeInfo &ti = c.info;
    const int nc = ti.n_cells, nl = ti.n_leaves, n = ti.n;
    const int kst = 2 * (mx.pmax + 1), PM1 = mx.pmax + 1;
    float *ahat = (float *)c.ak.p, *bhat = (float *)c.bk.p;
    const float2 *oxy = (const float2 *)c.oxy.p;
    const float *qf = (const float *)c.qf.p;
    const uint8_t *c_level = (const uint8_t *)c.c_level.p, *c_flags = (const uint8_t *)c.c_flags.p;
    const uint32_t *c_key = (const uint32_t *)c.c_key.p;
    const uint32_t *c_begin = (const uint32_t *)c.c_begin.p, *c_end = (const uint32_t *)c.c_end.p;
    const double *c_cx = (const double *)c.c_cx.p, *c_cy = (const double *)c.c_cy.p;
    const int *c_child = (const int *)c.c_child.p;
    const uint32_t *leaf_idx = (const uint32_t *)c.leaf_idx.p;
    const int maxL = ti.max_level;
    int K, T;
    tier_levels(ti, K, T);
    const int nK = (int)(ti.lev_off[K + 1] - ti.lev_off[K]);
    (void)n;
    // M2L list 2 on MFMA: every tile in one launch
    if (mx.ntiles > 0) {
        M2lTabs tabs;
        for (int p = 0; p <= 16; ++p) tabs.r[p] = mx.rptr_host[p];
        k_m2l_mfma<<<mx.ntiles, kWave * kM2lWaves, 0, s>>>(
            mx.tiles, mx.tile_p, mx.ntiles, c_key, (const int *)c.l2.p, (const uint8_t *)c.n2.p,
            tabs, ahat, kst, (const uint8_t *)c.n4.p,
            list4_mode == 2 ? (const float *)c.b4.p : nullptr, bhat);
    }
    const int j2 = (int)ti.lev_off[2];
    if (nc > j2 && list4_mode != 2)
        k_m2l4_32<<<div_up((long long)(nc - j2) * (PM1 - 1), 256), 256, 0, s>>>(
            j2, nc, c_level, (const int *)c.l4.p, (const uint8_t *)c.n4.p, c_cx, c_cy, mx.lt,
            mx.tm2l, ahat, kst, bhat);
    mk.mark(4, s);
    if (maxL >= 3) {
        // parents 2 .. T-1 (one workgroup), T .. K-1 (a launch per level), K .. (subtrees)
        if (T - 1 >= 2)
            k_l2l_top<<<1, kTopBlock, 0, s>>>(T - 1, c.hdr_d, c_flags, c_child, mx.lt, mx.l2l, kst,
                                              bhat);
        for (int L = std::max(T, 2); L <= K - 1; ++L) {
            const int j0 = (int)ti.lev_off[L], j1 = (int)ti.lev_off[L + 1];
            k_l2l_level<<<div_up(j1 - j0, kSubBlock / kWave), kSubBlock, 0, s>>>(
                L, j0, j1, c_flags, c_child, mx.lt, mx.l2l, kst, bhat);
        }
        if (K < maxL)
            k_l2l_subtree<<<nK, kSubBlock, 0, s>>>(K, maxL, c.hdr_d, c_key, c_flags, c_child,
                                                   mx.lt, mx.l2l, kst, bhat);
    }
    mk.mark(5, s);
    EvalArgs ea{leaf_idx, c_level, c_flags, c_begin, c_end, c_cx, c_cy,
                (const uint32_t *)c.off1.p, (const int *)c.l1.p, (const uint32_t *)c.off3.p,
                (const int *)c.l3.p, oxy, qf, (const double *)c.ss.p, c.perm_sorted, mx.lt,
                ahat, bhat, kst, stage_mask, velocity ? 1 : 0, u_inf, v_inf, out_a, out_b,
                &c.hdr_d->flags};
    const int nb = div_up(nl, kEvalWaves), nth = kWave * kEvalWaves;
    switch (c.icut) {
        case 1: k_eval32<1><<<nb, nth, 0, s>>>(ea, nl); break;
        case 2: k_eval32<2><<<nb, nth, 0, s>>>(ea, nl); break;
        case 3: k_eval32<3><<<nb, nth, 0, s>>>(ea, nl); break;
        default: k_eval32<0><<<nb, nth, 0, s>>>(ea, nl); break;
    }
    mk.mark(6, s);
    VPM_LAUNCH_CHECK();
}

void fmm32_prepare(Ctx &c, const Mixed32 &mx) {
    const TreeInfo &ti = c.info;
    const size_t kst = 2 * (size_t)(mx.pmax + 1);
    c.ak.as<float>((size_t)ti.n_cells * kst);
    c.bk.as<float>((size_t)ti.n_cells * kst);
    c.b4.as<float>((size_t)ti.n_cells * kst);
    c.oxy.as<float2>(ti.n);
    c.qf.as<float>(ti.n);
}

void fmm32_run(Ctx &c, const Mixed32 &mx, int list4_mode, int stage_mask, double *out_a,
               double *out_b, bool velocity, double u_inf, double v_inf, const FmmMarks &mk) {
    fmm32_prepare(c, mx);
    fmm32_upward(c, mx, c.stream, mk);
    if (list4_mode == 2) fmm32_p2l(c, mx, c.stream);
    fmm32_rest(c, mx, list4_mode, stage_mask, out_a, out_b, velocity, u_inf, v_inf, mk);
}

}  // namespace vpm
