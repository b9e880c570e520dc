// fmm64.hip -- plain FP64 FMM stages on gfx950.
//
// Expansion conventions follow fmm2D (reference src/FMM/fmm2D.cpp):
//   multipole a_p = sum q (z - zc)^p                 (P2M, :237-284)
//   M2M  a_p += a'_0 d^p + sum_k C(p-1,k-1)(p/k) a'_k d^(p-k),  d = zc' - zc   (:361-376)
//   M2L  b_p -= a_0/p d^-p + sum_k C(p+k-1,k-1)(-1)^k/k a_k d^-(p+k), d = zs - zt (:465-481)
//   L2L  b'_p += sum_{k>=p} C(k,p) b_k d^(k-p),  d = zc' - zc      (:637-643)
//   M2P  F = sum a_p (z - zc)^-(p+1)                  (list 3, :1052-1062)
//   L2P  F = sum_{p>=1} p b_p (z - zc)^(p-1)          (:1083-1093)
//   Ex += Re F, Ey -= Im F; P2P Ex += q dx/r^2, Ey += q dy/r^2 (:164-173).
// b_0 is never formed: L2L feeds b_p only from b_k with k >= p and L2P starts
// at p = 1, so b_0 cannot reach the field.
//
// The per-order coefficient tables (M2L including the reference's int32
// binomials in parity mode) are built on the host (api.cpp) and read from
// global memory.  Thread mappings: one thread per (cell, order) for the
// translations (Horner in the translation distance, no per-thread arrays),
// one wavefront per leaf for the evaluation.
#include <cmath>
#include <cstdlib>
#include <vector>

#include "ctx.hpp"

namespace vpm {
namespace {

constexpr unsigned kLeaf = 1u, kActive = 2u;

struct cd {
    double x, y;
};
__device__ __forceinline__ cd cmul(cd a, cd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cd cadd(cd a, cd b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cd csub(cd a, cd b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cd cscale(cd a, double s) { return {a.x * s, a.y * s}; }
__device__ __forceinline__ cd cinv(cd a) {
    const double r = 1.0 / (a.x * a.x + a.y * a.y);
    return {a.x * r, -a.y * r};
}
__device__ __forceinline__ cd cpowi(cd z, int p) {
    cd r = {1.0, 0.0};
    while (p > 0) {
        if (p & 1) r = cmul(r, z);
        z = cmul(z, z);
        p >>= 1;
    }
    return r;
}
__device__ __forceinline__ cd ld(const double2 *a, size_t i) {
    double2 v = a[i];
    return {v.x, v.y};
}
__device__ __forceinline__ void st(double2 *a, size_t i, cd v) { a[i] = make_double2(v.x, v.y); }

__device__ __forceinline__ double shfl_xor_d(double v, int m) {
    return __shfl_xor(v, m, kWave);
}

// Lanes of one wavefront read what other lanes of it wrote to LDS.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// P2M (fmm2D.cpp:237-284, with the sign and 1/p folded into the translation
// tables): four leaves per wavefront, 16 lanes each.  A leaf's lanes stage 16
// of its particles at a time in LDS as (z - zc, q), zero strength for
// inactive particles and leaves; lane r then adds q (z - zc)^p for its own
// order p = k0 + r over the staged particles, in particle order (the
// reference's loop order), with no cross-lane reduction.  Orders beyond 15
// take further chunks of 16.
constexpr int kP2m64Waves = 4;
constexpr int kP2mLanes = 16;                    // lanes per leaf
constexpr int kP2mLeaves = kWave / kP2mLanes;    // leaves per wavefront

__global__ void __launch_bounds__(kWave *kP2m64Waves)
    k_p2m64(const uint32_t *leaf_idx, int nl, int P, const uint8_t *c_flags,
            const uint32_t *c_begin, const uint32_t *c_end, const double *c_cx,
            const double *c_cy, const double *xs, const double *ys, const double *qs,
            const uint8_t *as, double2 *ak) {
    __shared__ double4 stg[kP2m64Waves][kWave];
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const int g = lane / kP2mLanes, r = lane % kP2mLanes;
    const long long t = ((long long)blockIdx.x * kP2m64Waves + w) * kP2mLeaves + g;
    const bool has = t < nl;
    const int P1 = P + 1;
    const int j = has ? (int)leaf_idx[t] : 0;
    const bool act = has && (c_flags[j] & kActive) != 0;
    const uint32_t b = has ? c_begin[j] : 0u, cnt = has ? c_end[j] - b : 0u;
    const double cx = has ? c_cx[j] : 0.0, cy = has ? c_cy[j] : 0.0;
    // rounds of 16 particles: the largest of the wavefront's leaves (uniform)
    uint32_t mx = cnt;
    for (int m = kP2mLanes; m < kWave; m <<= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, m, kWave));
    mx = (uint32_t)__builtin_amdgcn_readfirstlane((int)mx);
    for (int k0 = 0; k0 < P1; k0 += kP2mLanes) {
        const int p = k0 + r;
        cd acc = {0.0, 0.0};
        for (uint32_t base = 0; base < mx; base += kP2mLanes) {
            double4 v = make_double4(0.0, 0.0, 0.0, 0.0);
            const uint32_t i = b + base + (uint32_t)r;
            if (act && base + (uint32_t)r < cnt && as[i])
                v = make_double4(xs[i] - cx, ys[i] - cy, qs[i], 0.0);
            stg[w][lane] = v;
            wave_sync();
            if (p < P1) {
                for (int u = 0; u < kP2mLanes; ++u) {
                    const double4 sv = stg[w][g * kP2mLanes + u];
                    if (sv.z != 0.0) acc = cadd(acc, cscale(cpowi(cd{sv.x, sv.y}, p), sv.z));
                }
            }
            wave_sync();
        }
        if (has && p < P1) st(ak, (size_t)j * P1 + p, acc);
    }
}

// M2M for the non-leaf cells of one level: thread per (parent, p),
//   a_p(parent) = sum_children sum_{k=0..p} c_pk a_k(child) d^(p-k),
// d = z_child - z_parent, c_p0 = 1 (fmm2D.cpp:309-379).
__global__ void __launch_bounds__(256)
    k_m2m64(int j0, int j1, int P, const double *tab, const uint8_t *c_flags, const double *c_cx,
            const double *c_cy, const int *c_child, double2 *ak) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int P1 = P + 1;
    if (idx >= (long long)(j1 - j0) * P1) return;
    const int j = j0 + (int)(idx / P1), p = (int)(idx % P1);
    if (c_flags[j] & kLeaf) return;
    const double cx = c_cx[j], cy = c_cy[j];
    cd acc = {0.0, 0.0};
    for (int q = 0; q < 4; ++q) {
        const int ch = c_child[4 * j + q];
        if (ch < 0) continue;
        const cd d = {c_cx[ch] - cx, c_cy[ch] - cy};
        const size_t b = (size_t)ch * P1;
        // Horner from k = 0, which carries the power d^p
        cd sacc = ld(ak, b);
        for (int k = 1; k <= p; ++k) sacc = cadd(cmul(sacc, d), cscale(ld(ak, b + k), tab[p * P1 + k]));
        acc = cadd(acc, sacc);
    }
    st(ak, (size_t)j * P1 + p, acc);
}

// The same M2M with G parent cells per wavefront, 64 / G lanes each (G = 4
// for orders P <= 15, 1 up to 16): a cell's four children's coefficients and
// the table are staged in LDS in one round of loads, then lane p of the cell
// runs k_m2m64's Horner sums from LDS (the same operations in the same
// order), so a level costs one dependent load, not one per order.
constexpr int kTrWaves = 4;    // wavefronts per block of the per-cell translations
constexpr int kTrP1 = 17;      // orders up to 16

template <int G>
__global__ void __launch_bounds__(kWave *kTrWaves)
    k_m2m64w(int j0, int j1, int P, const double *tab, const uint8_t *c_flags, const double *c_cx,
             const double *c_cy, const int *c_child, double2 *ak) {
    constexpr int LN = kWave / G;
    __shared__ double s_tab[kTrP1 * kTrP1];
    __shared__ double2 sa[kTrWaves][G][4 * kTrP1];
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const int g = lane / LN, r = lane % LN;
    const int P1 = P + 1;
    for (int t = threadIdx.x; t < P1 * P1; t += blockDim.x) s_tab[t] = tab[t];
    const int j = j0 + ((int)blockIdx.x * kTrWaves + w) * G + g;
    const bool live = j < j1 && !(c_flags[j] & kLeaf);  // uniform over the cell's lanes
    int4 c4 = make_int4(-1, -1, -1, -1);
    if (live) {
        c4 = reinterpret_cast<const int4 *>(c_child)[j];
        for (int idx = r; idx < 4 * P1; idx += LN) {
            const int q = idx / P1, k = idx - q * P1;
            const int ch = q == 0 ? c4.x : q == 1 ? c4.y : q == 2 ? c4.z : c4.w;
            sa[w][g][q * kTrP1 + k] = ch >= 0 ? ak[(size_t)ch * P1 + k] : make_double2(0.0, 0.0);
        }
    }
    __syncthreads();
    if (!live || r > P) return;
    const int p = r;
    const double cx = c_cx[j], cy = c_cy[j];
    cd acc = {0.0, 0.0};
    for (int q = 0; q < 4; ++q) {
        const int ch = q == 0 ? c4.x : q == 1 ? c4.y : q == 2 ? c4.z : c4.w;
        if (ch < 0) continue;
        const cd d = {c_cx[ch] - cx, c_cy[ch] - cy};
        const double2 *a = &sa[w][g][q * kTrP1];
        cd sacc = ld(a, 0);
        for (int k = 1; k <= p; ++k) sacc = cadd(cmul(sacc, d), cscale(ld(a, k), s_tab[p * P1 + k]));
        acc = cadd(acc, sacc);
    }
    st(ak, (size_t)j * P1 + p, acc);
}

// M2M for the parents of levels Lhi .. 1 (each holding at most kTopCells64
// cells, orders P <= 15) in one workgroup: k_m2m64w's four-cells-per-wavefront
// body, a level at a time with a barrier between levels, in place of a launch
// per level.
constexpr int kTopWaves64 = 8;
constexpr int kTopCells64 = 64;

__global__ void __launch_bounds__(kWave *kTopWaves64)
    k_m2m64_top(int Lhi, const DevHdr *hdr, int P, const double *tab, const uint8_t *c_flags,
                const double *c_cx, const double *c_cy, const int *c_child, double2 *ak) {
    constexpr int G = 4, LN = kWave / G, S = kTrP1 - 1;
    __shared__ double s_tab[kTrP1 * kTrP1];
    __shared__ double2 sa[kTopWaves64][G][4 * S];
    __shared__ int s_off[kMaxDepth + 2];
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const int g = lane / LN, r = lane % LN;
    const int P1 = P + 1;
    for (int t = threadIdx.x; t < P1 * P1; t += blockDim.x) s_tab[t] = tab[t];
    if (threadIdx.x < kMaxDepth + 2) s_off[threadIdx.x] = (int)hdr->lev_off[threadIdx.x];
    __syncthreads();
    for (int L = Lhi; L >= 1; --L) {
        const int j0 = s_off[L], nj = s_off[L + 1] - j0;
        for (int u0 = 0; u0 < nj; u0 += kTopWaves64 * G) {
            const int u = u0 + w * G + g, j = j0 + u;
            const bool live = u < nj && !(c_flags[j] & kLeaf);
            int4 c4 = make_int4(-1, -1, -1, -1);
            if (live) {
                c4 = reinterpret_cast<const int4 *>(c_child)[j];
                for (int idx = r; idx < 4 * P1; idx += LN) {
                    const int q = idx / P1, k = idx - q * P1;
                    const int ch = q == 0 ? c4.x : q == 1 ? c4.y : q == 2 ? c4.z : c4.w;
                    sa[w][g][q * S + k] = ch >= 0 ? ak[(size_t)ch * P1 + k] : make_double2(0.0, 0.0);
                }
            }
            wave_sync();
            if (live && r <= P) {
                const int p = r;
                const double cx = c_cx[j], cy = c_cy[j];
                cd acc = {0.0, 0.0};
                for (int q = 0; q < 4; ++q) {
                    const int ch = q == 0 ? c4.x : q == 1 ? c4.y : q == 2 ? c4.z : c4.w;
                    if (ch < 0) continue;
                    const cd d = {c_cx[ch] - cx, c_cy[ch] - cy};
                    const double2 *a = &sa[w][g][q * S];
                    cd sacc = ld(a, 0);
                    for (int k = 1; k <= p; ++k)
                        sacc = cadd(cmul(sacc, d), cscale(ld(a, k), s_tab[p * P1 + k]));
                    acc = cadd(acc, sacc);
                }
                st(ak, (size_t)j * P1 + p, acc);
            }
            wave_sync();
        }
        __syncthreads();  // the next level up reads these multipoles
    }
}

// Principal complex logarithm, as std::log: log|z| + i arg z.
__device__ __forceinline__ cd clog(cd z) { return {log(hypot(z.x, z.y)), atan2(z.y, z.x)}; }

// M2L for every cell of level >= 2 from lists 2 and 4 (fmm2D.cpp:409-556):
// thread per (cell, order).  tab[p][k] holds the coefficient of a_k w^(p+k),
// w = 1/(zs - zt), so b_p = w^p sum_k tab[p][k] a_k w^k (Horner in w).  With
// POT the order-0 row of the potential is formed too (fmm2D.cpp:463-467,
// 504-508): b_0 = a_0 log(z_t - z_s) - sum_{k>=1} a_k (-1)^k / k w^k, and by
// P2L b_0 += q log(z_t - z_i) (:540).  With PT > 0 (orders P <= PT) a
// source's coefficients are loaded into registers at once before the Horner
// sweep, rather than one dependent load per step.
template <bool POT, int PT>
__global__ void __launch_bounds__(256)
    k_m2l64(int j0, int nc, int P, int list4_mode, const double *tab, const int *l2,
            const uint8_t *n2, const int *l4, const uint8_t *n4, const double *c_cx,
            const double *c_cy, const uint32_t *c_begin, const uint32_t *c_end, const double *xs,
            const double *ys, const double *qs, const uint8_t *as, const double2 *ak, double2 *bk) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int P1 = P + 1, NP = POT ? P1 : P;  // orders formed: 0..P (potential) or 1..P
    if (idx >= (long long)(nc - j0) * NP) return;
    const int j = j0 + (int)(idx / NP), p = (POT ? 0 : 1) + (int)(idx % NP);
    const double cx = c_cx[j], cy = c_cy[j];
    const double *row = tab + (size_t)p * P1;
    cd acc = {0.0, 0.0};
    const int m2 = n2[j], m4 = n4[j];
    for (int u = 0; u < m2 + (list4_mode == 1 ? m4 : 0); ++u) {
        const int sidx = u < m2 ? l2[27 * j + u] : l4[8 * j + (u - m2)];
        const cd w = cinv(cd{c_cx[sidx] - cx, c_cy[sidx] - cy});
        const size_t b = (size_t)sidx * P1;
        if (POT && p == 0) {
            cd sw = {0.0, 0.0};
            for (int k = P; k >= 1; --k)
                sw = cmul(cadd(sw, cscale(ld(ak, b + k), ((k & 1) ? 1.0 : -1.0) / k)), w);
            acc = cadd(acc, cadd(sw, cmul(ld(ak, b), clog(cd{cx - c_cx[sidx], cy - c_cy[sidx]}))));
            continue;
        }
        cd s;
        if constexpr (PT > 0) {
            cd a[PT + 1];
#pragma unroll
            for (int k = 0; k <= PT; ++k) a[k] = k <= P ? ld(ak, b + k) : cd{0.0, 0.0};
            s = cscale(a[PT], PT <= P ? row[PT] : 0.0);
#pragma unroll
            for (int k = PT - 1; k >= 0; --k)
                if (k <= P) s = cadd(cmul(s, w), cscale(a[k], row[k]));
        } else {
            s = cscale(ld(ak, b + P), row[P]);
            for (int k = P - 1; k >= 0; --k) s = cadd(cmul(s, w), cscale(ld(ak, b + k), row[k]));
        }
        acc = cadd(acc, cmul(s, cpowi(w, p)));
    }
    if (list4_mode == 2) {
        // P2L from the active particles of each list-4 leaf (fmm2D.cpp:526-553)
        const double inv_p = p > 0 ? 1.0 / (double)p : 0.0;
        for (int u = 0; u < m4; ++u) {
            const int sidx = l4[8 * j + u];
            for (uint32_t i = c_begin[sidx]; i < c_end[sidx]; ++i) {
                if (!as[i]) continue;
                if (POT && p == 0) {
                    acc = cadd(acc, cscale(clog(cd{cx - xs[i], cy - ys[i]}), qs[i]));
                    continue;
                }
                const cd w = cinv(cd{xs[i] - cx, ys[i] - cy});
                acc = cadd(acc, cscale(cpowi(w, p), -qs[i] * inv_p));
            }
        }
    }
    st(bk, (size_t)j * P1 + p, acc);
}

// M2L by one wavefront per target cell, for orders P <= 16 (the same sums as
// k_m2l64).  The wave stages its M2L sources' coefficients a_k and
// w = 1/(zs - zt) in LDS once (the list-2 entries, then the list-4 ones
// with list4_mode 1); lane l = g NP + i (NP = P orders, P + 1 for the
// potential) forms order p = i (+1 without the potential) over the sources
// u = g, g + G, ... (G = 64 / NP groups), and the groups' partial sums are
// added in group order.  Every lane works, and a source's coefficients are
// read from HBM / L2 once per target instead of once per (target, order).
// P2L (list4_mode 2) splits the list-4 particles over the groups the same way.
// With L4ONLY the wave takes list 4 alone (M2L or P2L by list4_mode), skips a
// cell without one and adds its sum to b_p, which the matrix-core list-2
// kernel has written.
constexpr int kM2lwWaves = 4;
constexpr int kM2lwSrc = 27 + 8;  // list 2 <= 27, list 4 <= 8
constexpr int kM2lwP1 = 17;       // orders up to 16

template <bool POT, bool L4ONLY>
__global__ void __launch_bounds__(kWave *kM2lwWaves)
    k_m2l64w(int j0, int nc, int P, int list4_mode, const double *tab, const int *l2,
             const uint8_t *n2, const int *l4, const uint8_t *n4, const double *c_cx,
             const double *c_cy, const uint32_t *c_begin, const uint32_t *c_end,
             const double *xs, const double *ys, const double *qs, const uint8_t *as,
             const double2 *ak, double2 *bk) {
    // sources staged per wavefront: lists 2 and 4, or list 4 alone (a
    // smaller LDS footprint, so more wavefronts reside)
    constexpr int NS = L4ONLY ? 8 : kM2lwSrc;
    __shared__ double2 sak[kM2lwWaves][NS * kM2lwP1];
    __shared__ double2 sw[kM2lwWaves][NS], slg[kM2lwWaves][NS];
    __shared__ double2 red[kM2lwWaves][kWave];
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const int j = j0 + blockIdx.x * kM2lwWaves + w;
    if (j >= nc) return;  // uniform over the wavefront
    const int P1 = P + 1, NP = POT ? P1 : P, G = kWave / NP;
    const int g = lane / NP, i = lane % NP, p = (POT ? 0 : 1) + i;
    const double cx = c_cx[j], cy = c_cy[j];
    const int m2 = L4ONLY ? 0 : n2[j], m4 = n4[j];
    if (L4ONLY && m4 == 0) return;  // uniform over the wavefront
    const int ns = m2 + (list4_mode == 1 ? m4 : 0);
    // staging: w and (potential) log(z_t - z_s) per source, then the a_k rows
    if (lane < ns) {
        const int sidx = lane < m2 ? l2[27 * j + lane] : l4[8 * j + (lane - m2)];
        const cd wv = cinv(cd{c_cx[sidx] - cx, c_cy[sidx] - cy});
        sw[w][lane] = make_double2(wv.x, wv.y);
        if (POT) {
            const cd lg = clog(cd{cx - c_cx[sidx], cy - c_cy[sidx]});
            slg[w][lane] = make_double2(lg.x, lg.y);
        }
    }
    for (int r = lane; r < ns * P1; r += kWave) {
        const int u = r / P1, k = r % P1;
        const int sidx = u < m2 ? l2[27 * j + u] : l4[8 * j + (u - m2)];
        sak[w][u * kM2lwP1 + k] = ak[(size_t)sidx * P1 + k];
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    cd acc = {0.0, 0.0};
    if (g < G) {
        const double *row = tab + (size_t)p * P1;
        for (int u = g; u < ns; u += G) {
            const double2 w2 = sw[w][u];
            const cd wv = {w2.x, w2.y};
            const double2 *a = &sak[w][u * kM2lwP1];
            if (POT && p == 0) {
                cd sm = {0.0, 0.0};
                for (int k = P; k >= 1; --k)
                    sm = cmul(cadd(sm, cscale(ld(a, k), ((k & 1) ? 1.0 : -1.0) / k)), wv);
                const double2 lg = slg[w][u];
                acc = cadd(acc, cadd(sm, cmul(ld(a, 0), cd{lg.x, lg.y})));
                continue;
            }
            cd sm = cscale(ld(a, P), row[P]);
            for (int k = P - 1; k >= 0; --k) sm = cadd(cmul(sm, wv), cscale(ld(a, k), row[k]));
            acc = cadd(acc, cmul(sm, cpowi(wv, p)));
        }
        if (list4_mode == 2) {
            // P2L (fmm2D.cpp:526-553): the list-4 leaves' active particles,
            // taken in turn by the groups
            const double inv_p = p > 0 ? 1.0 / (double)p : 0.0;
            int q = 0;
            for (int u = 0; u < m4; ++u) {
                const int sidx = l4[8 * j + u];
                for (uint32_t pi = c_begin[sidx]; pi < c_end[sidx]; ++pi, ++q) {
                    if (q % G != g || !as[pi]) continue;
                    if (POT && p == 0) {
                        acc = cadd(acc, cscale(clog(cd{cx - xs[pi], cy - ys[pi]}), qs[pi]));
                        continue;
                    }
                    const cd wv = cinv(cd{xs[pi] - cx, ys[pi] - cy});
                    acc = cadd(acc, cscale(cpowi(wv, p), -qs[pi] * inv_p));
                }
            }
        }
    }
    red[w][lane] = make_double2(acc.x, acc.y);
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane < NP) {
        cd sum = {0.0, 0.0};
        for (int gg = 0; gg < G; ++gg) sum = cadd(sum, ld(red[w], gg * NP + lane));
        const size_t o = (size_t)j * P1 + (POT ? 0 : 1) + lane;
        st(bk, o, L4ONLY ? cadd(ld(bk, o), sum) : sum);
    }
}

// List-2 M2L of the FP64 path on the matrix cores (orders P <= 16, field and
// velocity; the potential keeps k_m2l64w).  A workgroup takes a tile of 16
// consecutive cells of one level; the tiles of levels 2.. follow each other
// in the grid.  With len the level's cell side and delta = (zs - zt) / len
// (|dx|, |dy| <= 3, not both <= 1: 40 offsets), the list-2 sum
//   b_p = sum_k tab[p][k] a_k (zs - zt)^-(p+k)
// is, in the scaled coefficients a^_k = a_k len^-k and b^_p = b_p len^p,
//   b^ = sum over offsets of M_delta a^,  M_delta[p][k] = tab[p][k] delta^-(p+k),
// a realified 2P x 2(P+1) matrix per offset (fmm64_mfma_table), the A operand
// of v_mfma_f64_16x16x4f64; the B operand is a source cell's a^ along
// K = 2(P+1).  Lane l: target column n = l % 16, k-quarter kq = l / 16;
// wavefront w takes the offsets w, w + kM2l64Waves, ... that a target of the
// tile uses, and the wavefronts' accumulators are summed through LDS.  The
// output rows follow the f64 instruction's layout D[4v + l / 16][l % 16].
// The offset is exact here, where the reference forms zs - zt from rounded
// centres: the two agree to the rounding of the centres (about 1e-16 of the
// root box over the cell side).  List 4 is added afterwards by k_m2l64w in
// its list-4 mode.
constexpr int kM2l64Waves = 4;
typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t compact1by1(uint32_t v) {
    v &= 0x55555555u;
    v = (v | (v >> 1)) & 0x33333333u;
    v = (v | (v >> 2)) & 0x0F0F0F0Fu;
    v = (v | (v >> 4)) & 0x00FF00FFu;
    v = (v | (v >> 8)) & 0x0000FFFFu;
    return v;
}

// Eight resident wavefronts per SIMD (the register budget that allows them:
// 64 per lane, accumulators included): the kernel waits on its gathers, and
// 8 instead of 5 resident wavefronts cut it from 170 to 146 us at 780k (p 10).
template <int P>
__global__ void __launch_bounds__(kWave *kM2l64Waves) __attribute__((amdgpu_waves_per_eu(8, 8)))
    k_m2l64_mfma(const DevHdr *hdr, int maxL, const uint32_t *c_key, const int *l2,
                 const uint8_t *n2, const double *R, const double2 *ak, double2 *bk) {
    constexpr int P1 = P + 1, K = 2 * P1, KS = (K + 3) / 4, MT = (2 * P + 15) / 16;
    __shared__ int smap[16][49];
    __shared__ double red[kM2l64Waves][MT * 4][kWave];
    __shared__ double s_sc[P1];
    __shared__ int s_tile[3];
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    if (threadIdx.x == 0) {
        // this block's tile: level L, first cell, cells (up to 16)
        int b = (int)blockIdx.x, L = 2, j0 = -1, cnt = 0;
        for (; L <= maxL; ++L) {
            const int c0 = (int)hdr->lev_off[L], cn = (int)hdr->lev_off[L + 1] - c0;
            const int nt = (cn + 15) / 16;
            if (b < nt) {
                j0 = c0 + 16 * b;
                cnt = min(16, cn - 16 * b);
                break;
            }
            b -= nt;
        }
        s_tile[0] = L;
        s_tile[1] = j0;
        s_tile[2] = cnt;
    }
    for (int t = threadIdx.x; t < 16 * 49; t += blockDim.x) smap[t / 49][t % 49] = -1;
    __syncthreads();
    const int L = s_tile[0], j0 = s_tile[1], cnt = s_tile[2];
    if (j0 < 0) return;  // (uniform over the block)
    if (threadIdx.x < P1) {
        // len^-k, k = 0..P
        const double inv = ldexp(1.0 / hdr->root_len, L);
        double sc = 1.0;
        for (int k = 0; k < (int)threadIdx.x; ++k) sc *= inv;
        s_sc[threadIdx.x] = sc;
    }
    {
        // source map of the tile: slot (dx+3) + 7 (dy+3) -> source cell or -1
        const int tn = threadIdx.x % 16, tu = threadIdx.x / 16;
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
    const int n = lane % 16, kq = lane / 16;
    const double *a1 = reinterpret_cast<const double *>(ak);
    // orders 9 and 10: rows 16.. (at most 4) on v_mfma_f64_4x4x4f64, whose
    // A operand holds row 16 + l % 4, column l / 16 of a 4-step, whose B
    // operand is laid out as the 16x16x4 one (target l % 16, step row l / 16)
    // and whose result is row 16 + l / 16 of target l % 16
    constexpr bool LEFT = 2 * P > 16 && 2 * P <= 20;
    v4d acc[MT];
    double acc4 = 0.0;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) acc[mt] = v4d{0.0, 0.0, 0.0, 0.0};
    int o = 0;
    for (int slot = 0; slot < 49; ++slot) {
        const int dx = slot % 7 - 3, dy = slot / 7 - 3;
        if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1) continue;
        const int oc = o++;
        if (oc % kM2l64Waves != w) continue;
        const int sidx = smap[n][slot];
        if (__ballot(sidx >= 0) == 0) continue;  // no target of the tile uses this offset
        double b[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int col = ks * 4 + kq;  // 2k (Re a_k) or 2k + 1 (Im a_k)
            b[ks] = (sidx >= 0 && col < K) ? a1[(size_t)sidx * K + col] * s_sc[col >> 1] : 0.0;
        }
        const double *Ro = R + (size_t)oc * KS * MT * kWave + lane;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
            for (int mt = 0; mt < (LEFT ? 1 : MT); ++mt)
                acc[mt] = __builtin_amdgcn_mfma_f64_16x16x4f64(Ro[(ks * MT + mt) * kWave], b[ks],
                                                               acc[mt], 0, 0, 0);
            if (LEFT)
                acc4 = __builtin_amdgcn_mfma_f64_4x4x4f64(Ro[(ks * MT + 1) * kWave], b[ks], acc4,
                                                          0, 0, 0);
        }
    }
#pragma unroll
    for (int mt = 0; mt < (LEFT ? 1 : MT); ++mt)
#pragma unroll
        for (int v = 0; v < 4; ++v) red[w][mt * 4 + v][lane] = acc[mt][v];
    if (LEFT) red[w][4][lane] = acc4;
    __syncthreads();
    if (w != 0) return;
    double *bo = reinterpret_cast<double *>(bk);
    if (LEFT) {
        double sum = 0.0;
#pragma unroll
        for (int u = 0; u < kM2l64Waves; ++u) sum += red[u][4][lane];
        const int row = 16 + kq;
        if (n < cnt && row < 2 * P) {
            const int p = row / 2 + 1;
            bo[((size_t)(j0 + n) * P1 + p) * 2 + (row & 1)] = sum * s_sc[p];
        }
    }
#pragma unroll
    for (int mt = 0; mt < (LEFT ? 1 : MT); ++mt) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            double sum = 0.0;
#pragma unroll
            for (int u = 0; u < kM2l64Waves; ++u) sum += red[u][mt * 4 + v][lane];
            const int row = mt * 16 + 4 * v + kq;  // Re b^_1, Im b^_1, Re b^_2, ...
            if (n < cnt && row < 2 * P) {
                const int p = row / 2 + 1;
                bo[((size_t)(j0 + n) * P1 + p) * 2 + (row & 1)] = sum * s_sc[p];
            }
        }
    }
}

// L2L for the cells of one level (>= 3): each child adds its parent's shifted
// local expansion.  tab[p][k] = C(k, p).  Orders 1..P, and 0 for the
// potential (fmm2D.cpp:637-643 runs p from 0).
template <bool POT>
__global__ void __launch_bounds__(256)
    k_l2l64(int j0, int j1, int P, const double *tab, const int *c_parent, const double *c_cx,
            const double *c_cy, double2 *bk) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int P1 = P + 1, NP = POT ? P1 : P;
    if (idx >= (long long)(j1 - j0) * NP) return;
    const int j = j0 + (int)(idx / NP), p = (POT ? 0 : 1) + (int)(idx % NP);
    const int par = c_parent[j];
    const cd d = {c_cx[j] - c_cx[par], c_cy[j] - c_cy[par]};
    const size_t b = (size_t)par * P1;
    const double *row = tab + (size_t)p * P1;
    cd s = cscale(ld(bk, b + P), row[P]);
    for (int k = P - 1; k >= p; --k) s = cadd(cmul(s, d), cscale(ld(bk, b + k), row[k]));
    const size_t o = (size_t)j * P1 + p;
    st(bk, o, cadd(ld(bk, o), s));
}

// The same L2L with G child cells per wavefront, 64 / G lanes each (as
// k_m2m64w): the parent's coefficients and the table are staged in LDS in
// one round of loads, then lane i of the cell runs k_l2l64's Horner sum for
// order p from LDS (the same operations in the same order).
template <bool POT, int G>
__global__ void __launch_bounds__(kWave *kTrWaves)
    k_l2l64w(int j0, int j1, int P, const double *tab, const int *c_parent, const double *c_cx,
             const double *c_cy, double2 *bk) {
    constexpr int LN = kWave / G;
    __shared__ double s_tab[kTrP1 * kTrP1];
    __shared__ double2 sb[kTrWaves][G][kTrP1];
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const int g = lane / LN, r = lane % LN;
    const int P1 = P + 1, NP = POT ? P1 : P;
    for (int t = threadIdx.x; t < P1 * P1; t += blockDim.x) s_tab[t] = tab[t];
    const int j = j0 + ((int)blockIdx.x * kTrWaves + w) * G + g;
    const bool live = j < j1;  // uniform over the cell's lanes
    int par = 0;
    if (live) {
        par = c_parent[j];
        for (int k = r; k < P1; k += LN) sb[w][g][k] = bk[(size_t)par * P1 + k];
    }
    __syncthreads();
    if (!live || r >= NP) return;
    const int p = (POT ? 0 : 1) + r;
    const cd d = {c_cx[j] - c_cx[par], c_cy[j] - c_cy[par]};
    const double *row = s_tab + p * P1;
    const double2 *b = sb[w][g];
    cd s = cscale(ld(b, P), row[P]);
    for (int k = P - 1; k >= p; --k) s = cadd(cmul(s, d), cscale(ld(b, k), row[k]));
    const size_t o = (size_t)j * P1 + p;
    st(bk, o, cadd(ld(bk, o), s));
}

// L2L into the cells of levels 3 .. Lhi (each holding at most kTopCells64
// cells, orders P <= 15) in one workgroup: k_l2l64w's body, a level at a
// time with a barrier between levels, in place of a launch per level.
template <bool POT>
__global__ void __launch_bounds__(kWave *kTopWaves64)
    k_l2l64_top(int Lhi, const DevHdr *hdr, int P, const double *tab, const int *c_parent,
                const double *c_cx, const double *c_cy, double2 *bk) {
    constexpr int G = 4, LN = kWave / G;
    __shared__ double s_tab[kTrP1 * kTrP1];
    __shared__ double2 sb[kTopWaves64][G][kTrP1];
    __shared__ int s_off[kMaxDepth + 2];
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const int g = lane / LN, r = lane % LN;
    const int P1 = P + 1, NP = POT ? P1 : P;
    for (int t = threadIdx.x; t < P1 * P1; t += blockDim.x) s_tab[t] = tab[t];
    if (threadIdx.x < kMaxDepth + 2) s_off[threadIdx.x] = (int)hdr->lev_off[threadIdx.x];
    __syncthreads();
    for (int L = 3; L <= Lhi; ++L) {
        const int j0 = s_off[L], nj = s_off[L + 1] - j0;
        for (int u0 = 0; u0 < nj; u0 += kTopWaves64 * G) {
            const int u = u0 + w * G + g, j = j0 + u;
            const bool live = u < nj;
            int par = 0;
            if (live) {
                par = c_parent[j];
                for (int k = r; k < P1; k += LN) sb[w][g][k] = bk[(size_t)par * P1 + k];
            }
            wave_sync();
            if (live && r < NP) {
                const int p = (POT ? 0 : 1) + r;
                const cd d = {c_cx[j] - c_cx[par], c_cy[j] - c_cy[par]};
                const double *row = s_tab + p * P1;
                const double2 *b = sb[w][g];
                cd sv = cscale(ld(b, P), row[P]);
                for (int k = P - 1; k >= p; --k) sv = cadd(cmul(sv, d), cscale(ld(b, k), row[k]));
                const size_t o = (size_t)j * P1 + p;
                st(bk, o, cadd(ld(bk, o), sv));
            }
            wave_sync();
        }
        __syncthreads();  // the next level down reads these coefficients
    }
}

// Evaluation: one wavefront per leaf.  Lanes are split into G groups of
// 64/G lanes; a group owns the targets, the G groups split the sources and are
// reduced with cross-lane shuffles.  List-1 sources are staged in LDS (x, y, q,
// index) by the whole wavefront, then read by every lane.
constexpr int kEvalWaves = 4;
// Sources staged per wavefront per round: 128 (4 KB of LDS per wavefront)
// leaves room for 7 wavefronts per SIMD where 256 allowed 5; the FP64 pair
// loop is latency bound, and the larger count of resident wavefronts cut the
// 780k evaluation by about 50 us.  A target's sum does not depend on it.
constexpr int kStage64 = 128;



// 1/x to full double precision: hardware estimate and two Newton steps.
__device__ __forceinline__ double rcp64(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

// Regularization factor of a near-field pair, 2 pi q of regul_func_2d
// (regularization_function_2d.cpp:11-37) at rho^2 = r^2 / sij^2, in the
// reference's form: 1 high-order algebraic, 2 Gaussian, 3 super-Gaussian.
template <int ICUT>
__device__ __forceinline__ double regul64(double rho2) {
    if (ICUT == 1) return (rho2 * (rho2 + 2.0)) / ((rho2 + 1.0) * (rho2 + 1.0));
    const double qt = -0.5 * rho2;
    if (ICUT == 2) return 1.0 - exp(qt);
    return 1.0 - (1.0 - qt) * exp(qt);
}

// ICUT 0: the singular P2P of fmm2D (fmm2D.cpp:164-173); a pair at r = 0
// raises DUPLICATE_POINTS.  ICUT 1..3: the type-1 near field of direct_sum
// (subroutine_all_fmm_2d.cpp:594-627): each pair term times regul64 at
// sij^2 = (si^2 + sj^2) / 4 (ss: the sorted sizes); pairs at r = 0 add
// nothing.
// T1: the type-1 FMM (biotsavart_fmm_2d.cpp:321-419): pairs at r = 0 add
// nothing whatever the kernel; the multipoles hold orders 0..P-1 (stride P)
// and the local expansions orders 0..P (stride P + 1) in the Cauchy form
// F = sum_l b_l (z - z0)^l.
// One leaf's targets in groups of 64 / G lanes, G a compile-time constant so
// that the source stride and the group reduction are fixed and the pair loop
// unrolls (four pairs in flight per lane); stage / stage_s are the
// wavefront's LDS rows.  Returns whether a pair at r = 0 was seen.
template <int ICUT, bool T1, int G>
__device__ __forceinline__ bool eval64_leaf(int t, int j, int lane, double4 *stage,
                                            double *stage_s, int P, const uint32_t *c_begin,
                                            const uint32_t *c_end, const double *c_cx,
                                            const double *c_cy, const uint32_t *off1,
                                            const int *l1, const uint32_t *off3, const int *l3,
                                            const double *xs, const double *ys, const double *qs,
                                            const double *ss, const uint32_t *perm,
                                            const double2 *ak, const double2 *bk, int velocity,
                                            double u_inf, double v_inf, int mask, double *out_a,
                                            double *out_b) {
    const uint32_t b = c_begin[j], e = c_end[j];
    constexpr int LG = kWave / G;
    const int g = lane / LG, tl = lane % LG;
    const int P1 = P + 1;
    const double lcx = c_cx[j], lcy = c_cy[j];
    const uint32_t a1 = off1[t], z1 = off1[t + 1], a3 = off3[t], z3 = off3[t + 1];
    bool dup = false;
    for (uint32_t tb = b; tb < e; tb += LG) {
        const uint32_t i = tb + tl;
        const bool valid = i < e;
        const double xi = valid ? xs[i] : 0.0, yi = valid ? ys[i] : 0.0;
        const double si2 = (ICUT && valid) ? ss[i] * ss[i] : 0.0;
        double ex = 0.0, ey = 0.0;
        // P2P over every particle of the list-1 cells, j != i (fmm2D.cpp:135-179)
        uint32_t u = a1, pos = a1 < z1 ? c_begin[l1[a1]] : 0u;
        while ((mask & 1) && u < z1) {
            int fill = 0;
            while (u < z1 && fill < kStage64) {
                const uint32_t se = c_end[l1[u]];
                const int take = (int)min<uint32_t>(se - pos, (uint32_t)(kStage64 - fill));
                for (int r = lane; r < take; r += kWave) {
                    const uint32_t src = pos + r;
                    stage[fill + r] = make_double4(xs[src], ys[src], qs[src], (double)src);
                    if (ICUT) stage_s[fill + r] = ss[src] * ss[src];
                }
                fill += take;
                pos += take;
                if (pos >= se && ++u < z1) pos = c_begin[l1[u]];
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
#pragma unroll 4
            for (int r = g; r < fill; r += G) {
                const double4 sv = stage[r];
                const double dx = xi - sv.x, dy = yi - sv.y;
                const double r2 = dx * dx + dy * dy;
                if (ICUT == 0 && !T1 && valid && (uint32_t)sv.w != i) {
                    dup |= r2 == 0.0;
                    const double qr = sv.z * rcp64(r2);
                    ex += qr * dx;
                    ey += qr * dy;
                }
                if (ICUT == 0 && T1 && valid && r2 > 0.0) {
                    const double qr = sv.z * rcp64(r2);
                    ex += qr * dx;
                    ey += qr * dy;
                }
                if (ICUT != 0 && valid && r2 > 0.0) {
                    const double sij2 = 0.25 * (si2 + stage_s[r]);
                    const double qr = sv.z * regul64<ICUT>(r2 / sij2) * rcp64(r2);
                    ex += qr * dx;
                    ey += qr * dy;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
        // M2P over list 3, split across groups
        const int PM = T1 ? P - 1 : P;  // highest multipole order
        for (uint32_t v = a3 + g; (mask & 2) && v < z3; v += G) {
            const int sc = l3[v];
            const cd wv = cinv(cd{xi - c_cx[sc], yi - c_cy[sc]});
            const size_t sbk = (size_t)sc * (PM + 1);
            cd f = ld(ak, sbk + PM);
            for (int p = PM - 1; p >= 0; --p) f = cadd(cmul(f, wv), ld(ak, sbk + p));
            f = cmul(f, wv);
            if (valid) {
                ex += f.x;
                ey -= f.y;
            }
        }
        for (int m = LG; m < kWave; m <<= 1) {
            ex += shfl_xor_d(ex, m);
            ey += shfl_xor_d(ey, m);
        }
        if (valid && g == 0) {
            if (mask & 4) {
                // L2P
                const cd z = {xi - lcx, yi - lcy};
                const size_t lb = (size_t)j * P1;
                cd f;
                if (T1) {
                    f = ld(bk, lb + P);
                    for (int p = P - 1; p >= 0; --p) f = cadd(cmul(f, z), ld(bk, lb + p));
                } else {
                    f = cscale(ld(bk, lb + P), (double)P);
                    for (int p = P - 1; p >= 1; --p)
                        f = cadd(cmul(f, z), cscale(ld(bk, lb + p), (double)p));
                }
                ex += f.x;
                ey -= f.y;
            }
            const uint32_t o = perm[i];
            if (velocity) {
                out_a[o] = ey + u_inf;
                out_b[o] = -ex + v_inf;
            } else {
                out_a[o] = ex;
                out_b[o] = ey;
            }
        }
    }
    return dup;
}

// The singular kernel is held to 64 registers per lane so that 8 wavefronts
// per SIMD reside (7 otherwise; 27 us of 188 at 780k); the regularized ones
// would spill under that limit and keep the compiler's choice.
template <int ICUT, bool T1 = false>
__global__ void __launch_bounds__(kWave *kEvalWaves)
    __attribute__((amdgpu_waves_per_eu(ICUT == 0 ? 8 : 1, 8)))
    k_eval64(const uint32_t *leaf_idx, int nl, int P, const uint32_t *c_begin, const uint32_t *c_end,
             const double *c_cx, const double *c_cy, const uint32_t *off1, const int *l1,
             const uint32_t *off3, const int *l3, const double *xs, const double *ys,
             const double *qs, const double *ss, const uint32_t *perm, const double2 *ak,
             const double2 *bk, int velocity, double u_inf, double v_inf, int mask,
             double *out_a, double *out_b, unsigned *flags) {
    __shared__ double4 stage[kEvalWaves][kStage64];
    __shared__ double stage_s[ICUT ? kEvalWaves : 1][ICUT ? kStage64 : 1];
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const int t = blockIdx.x * kEvalWaves + w;
    if (t >= nl) return;
    const int j = (int)leaf_idx[t];
    const int nt = (int)(c_end[j] - c_begin[j]);
    double *sst = stage_s[ICUT ? w : 0];
    bool dup;
#define EVAL64_G(GG)                                                                          \
    eval64_leaf<ICUT, T1, GG>(t, j, lane, stage[w], sst, P, c_begin, c_end, c_cx, c_cy, off1, \
                              l1, off3, l3, xs, ys, qs, ss, perm, ak, bk, velocity, u_inf,   \
                              v_inf, mask, out_a, out_b)
    if (nt <= 8) dup = EVAL64_G(8);
    else if (nt <= 16) dup = EVAL64_G(4);
    else if (nt <= 32) dup = EVAL64_G(2);
    else dup = EVAL64_G(1);
#undef EVAL64_G
    if (dup) atomicOr(flags, kFlagDuplicate);
}

// Potential evaluation (fmm2D::calcPotential, fmm2D.cpp:681-896), one
// wavefront per leaf as k_eval64: P2P phi += q log(r^2) / 2 over every
// particle of the list-1 cells, j != i (potDirSum, :77-122); M2P over list 3,
// Re[a_0 log(z - zc) - sum_p a_p (z - zc)^-p / p] (:817-834); L2P
// Re[b_0 + sum_p b_p (z - z0)^p] (:854-866).  A pair at r = 0 raises
// DUPLICATE_POINTS (log 0).
__global__ void __launch_bounds__(kWave *kEvalWaves)
    k_pot64(const uint32_t *leaf_idx, int nl, int P, const uint32_t *c_begin, const uint32_t *c_end,
            const double *c_cx, const double *c_cy, const uint32_t *off1, const int *l1,
            const uint32_t *off3, const int *l3, const double *xs, const double *ys,
            const double *qs, const uint32_t *perm, const double2 *ak, const double2 *bk,
            int mask, double *phi, unsigned *flags) {
    __shared__ double4 stage[kEvalWaves][kStage64];
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const int t = blockIdx.x * kEvalWaves + w;
    if (t >= nl) return;
    const int j = (int)leaf_idx[t];
    const uint32_t b = c_begin[j], e = c_end[j];
    const int nt = (int)(e - b);
    const int G = nt <= 8 ? 8 : nt <= 16 ? 4 : nt <= 32 ? 2 : 1;
    const int LG = kWave / G;
    const int g = lane / LG, tl = lane % LG;
    const int P1 = P + 1;
    const double lcx = c_cx[j], lcy = c_cy[j];
    const uint32_t a1 = off1[t], z1 = off1[t + 1], a3 = off3[t], z3 = off3[t + 1];
    bool dup = false;
    for (uint32_t tb = b; tb < e; tb += LG) {
        const uint32_t i = tb + tl;
        const bool valid = i < e;
        const double xi = valid ? xs[i] : 0.0, yi = valid ? ys[i] : 0.0;
        double ph = 0.0;
        uint32_t u = a1, pos = a1 < z1 ? c_begin[l1[a1]] : 0u;
        while ((mask & 1) && u < z1) {
            int fill = 0;
            while (u < z1 && fill < kStage64) {
                const uint32_t se = c_end[l1[u]];
                const int take = (int)min<uint32_t>(se - pos, (uint32_t)(kStage64 - fill));
                for (int r = lane; r < take; r += kWave) {
                    const uint32_t src = pos + r;
                    stage[w][fill + r] = make_double4(xs[src], ys[src], qs[src], (double)src);
                }
                fill += take;
                pos += take;
                if (pos >= se && ++u < z1) pos = c_begin[l1[u]];
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            for (int r = g; r < fill; r += G) {
                const double4 sv = stage[w][r];
                const double dx = xi - sv.x, dy = yi - sv.y;
                const double r2 = dx * dx + dy * dy;
                if (valid && (uint32_t)sv.w != i) {
                    dup |= r2 == 0.0;
                    ph += sv.z * log(r2) / 2.0;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
        for (uint32_t v = a3 + g; (mask & 2) && v < z3; v += G) {
            const int sc = l3[v];
            const cd zr = {xi - c_cx[sc], yi - c_cy[sc]};
            const cd wv = cinv(zr);
            const size_t sbk = (size_t)sc * P1;
            // sum_{p>=1} a_p w^p / p by Horner, then a_0 log(z - zc) minus it
            cd f = {0.0, 0.0};
            for (int p = P; p >= 1; --p) f = cmul(cadd(f, cscale(ld(ak, sbk + p), 1.0 / p)), wv);
            const cd a0l = cmul(ld(ak, sbk), clog(zr));
            if (valid) ph += a0l.x - f.x;
        }
        for (int m = LG; m < kWave; m <<= 1) ph += shfl_xor_d(ph, m);
        if (valid && g == 0) {
            if (mask & 4) {
                const cd z = {xi - lcx, yi - lcy};
                const size_t lb = (size_t)j * P1;
                cd f = ld(bk, lb + P);
                for (int p = P - 1; p >= 0; --p) f = cadd(cmul(f, z), ld(bk, lb + p));
                ph += f.x;
            }
            phi[perm[i]] = ph;
        }
    }
    if (dup) atomicOr(flags, kFlagDuplicate);
}

// Type-1 M2L and P2L (biotsavart_fmm_2d.cpp:199-271) for every cell of level
// >= 1, thread per (cell, order l = 0..ND), with multipoles a_0..a_{ND-1}
// (stride ND) of strengths q = -g / (2 pi):
//   b_l = sum_{list 2} w^(l+1) sum_j tab[l][j] a_j w^j,   w = 1 / (zs - zt),
//         tab[l][j] = (-1)^(j+1) C(l + j, j)              (Horner in w)
//       - sum_{list 4} sum_i q_i w_i^(l+1),               w_i = 1 / (z_i - zt)
// (the reference's list 4 is always P2L from the leaf's particles).  Cells of
// level 1 have no lists (their locals are zero, as in the reference).
__global__ void __launch_bounds__(256)
    k_m2l_t1(int j0, int nc, int ND, const double *tab, const uint8_t *c_level, const int *l2,
             const uint8_t *n2, const int *l4, const uint8_t *n4, const double *c_cx,
             const double *c_cy, const uint32_t *c_begin, const uint32_t *c_end, const double *xs,
             const double *ys, const double *qs, const double2 *ak, double2 *bk) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int NL = ND + 1;
    if (idx >= (long long)(nc - j0) * NL) return;
    const int j = j0 + (int)(idx / NL), l = (int)(idx % NL);
    cd acc = {0.0, 0.0};
    if (c_level[j] >= 2) {
        const double cx = c_cx[j], cy = c_cy[j];
        const double *row = tab + (size_t)l * ND;
        const int m2 = n2[j], m4 = n4[j];
        for (int u = 0; u < m2; ++u) {
            const int sidx = l2[27 * j + u];
            const cd w = cinv(cd{c_cx[sidx] - cx, c_cy[sidx] - cy});
            const size_t b = (size_t)sidx * ND;
            cd sm = cscale(ld(ak, b + ND - 1), row[ND - 1]);
            for (int k = ND - 2; k >= 0; --k) sm = cadd(cmul(sm, w), cscale(ld(ak, b + k), row[k]));
            acc = cadd(acc, cmul(sm, cpowi(w, l + 1)));
        }
        for (int u = 0; u < m4; ++u) {
            const int sidx = l4[8 * j + u];
            for (uint32_t i = c_begin[sidx]; i < c_end[sidx]; ++i) {
                const cd w = cinv(cd{xs[i] - cx, ys[i] - cy});
                acc = csub(acc, cscale(cpowi(w, l + 1), qs[i]));
            }
        }
    }
    st(bk, (size_t)j * NL + l, acc);
}

// Type-1 L2L (biotsavart_fmm_2d.cpp:283-312) for the cells of one level
// (>= 3; level-1 locals are zero): thread per (cell, l),
//   b_l += sum_{k = l..ND} coef[l][k] d^expo[l][k] b_k(parent),  d = zc - zp,
// with the table of the shift as the reference runs it (or the exact one).
// Along a row the exponent rises by one per term except at most once, where
// it is formed afresh.
__global__ void __launch_bounds__(256)
    k_l2l_t1(int j0, int j1, int ND, const double *coef, const int *expo, const int *c_parent,
             const double *c_cx, const double *c_cy, double2 *bk) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int NL = ND + 1;
    if (idx >= (long long)(j1 - j0) * NL) return;
    const int j = j0 + (int)(idx / NL), l = (int)(idx % NL);
    const int par = c_parent[j];
    const cd d = {c_cx[j] - c_cx[par], c_cy[j] - c_cy[par]};
    const size_t b = (size_t)par * NL;
    cd s = {0.0, 0.0}, pw = {1.0, 0.0};
    int e_prev = 0;
    for (int k = l; k <= ND; ++k) {
        const int e = expo[l * NL + k];
        if (e == e_prev + 1) pw = cmul(pw, d);
        else if (e != e_prev) pw = cpowi(d, e);
        e_prev = e;
        s = cadd(s, cscale(cmul(pw, ld(bk, b + k)), coef[l * NL + k]));
    }
    const size_t o = (size_t)j * NL + l;
    st(bk, o, cadd(ld(bk, o), s));
}

// FP64 direct sum with LDS source tiles (accuracy reference only).
__global__ void __launch_bounds__(256)
    k_direct64(int n, const double *x, const double *y, const double *q, const uint8_t *mask,
               int nt, const int *tidx, double *ex, double *ey) {
    __shared__ double sx[256], sy[256], sq[256];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int i = t < nt ? (tidx ? tidx[t] : t) : -1;
    if (i >= n) i = -1;
    const bool valid = i >= 0;
    const double xi = valid ? x[i] : 0.0, yi = valid ? y[i] : 0.0;
    double ax = 0.0, ay = 0.0;
    for (int base = 0; base < n; base += 256) {
        const int jj = base + threadIdx.x;
        if (jj < n) {
            sx[threadIdx.x] = x[jj];
            sy[threadIdx.x] = y[jj];
            sq[threadIdx.x] = (mask && !mask[jj]) ? 0.0 : q[jj];
        }
        __syncthreads();
        const int m = min(256, n - base);
        for (int k = 0; k < m; ++k) {
            const int src = base + k;
            if (src == i) continue;
            const double dx = xi - sx[k], dy = yi - sy[k];
            const double qr = sq[k] / (dx * dx + dy * dy);
            ax += qr * dx;
            ay += qr * dy;
        }
        __syncthreads();
    }
    if (t < nt) {  // an out-of-range target index yields NaN
        ex[t] = valid ? ax : __builtin_nan("");
        ey[t] = valid ? ay : __builtin_nan("");
    }
}

}  // namespace

void fmm64_run(Ctx &c, int P, int list4_mode, int stage_mask, double *out_a, double *out_b,
               bool velocity, double u_inf, double v_inf, const FmmMarks &mk, bool potential) {
    hipStream_t s = c.stream;
    const TreeInfo &ti = c.info;
    const int nc = ti.n_cells, nl = ti.n_leaves, P1 = P + 1;
    double2 *ak = c.ak.as<double2>((size_t)nc * P1), *bk = c.bk.as<double2>((size_t)nc * P1);
    VPM_HIP(hipMemsetAsync(ak, 0, sizeof(double2) * (size_t)nc * P1, s));
    VPM_HIP(hipMemsetAsync(bk, 0, sizeof(double2) * (size_t)nc * P1, s));
    const uint8_t *c_flags = (const uint8_t *)c.c_flags.p;
    const uint32_t *c_begin = (const uint32_t *)c.c_begin.p, *c_end = (const uint32_t *)c.c_end.p;
    const double *c_cx = (const double *)c.c_cx.p, *c_cy = (const double *)c.c_cy.p;
    const int *c_child = (const int *)c.c_child.p, *c_parent = (const int *)c.c_parent.p;
    const double *xs = (const double *)c.xs.p, *ys = (const double *)c.ys.p,
                 *qs = (const double *)c.qs.p;
    const uint8_t *as = (const uint8_t *)c.acts.p;
    const uint32_t *leaf_idx = (const uint32_t *)c.leaf_idx.p;
    const double *t_m2l = (const double *)c.tab_m2l.p, *t_m2m = (const double *)c.tab_m2m.p,
                 *t_l2l = (const double *)c.tab_l2l.p;

    k_p2m64<<<div_up(nl, kP2m64Waves * kP2mLeaves), kWave * kP2m64Waves, 0, s>>>(
        leaf_idx, nl, P, c_flags, c_begin, c_end, c_cx, c_cy, xs, ys, qs, as, ak);
    // the top levels (parents 1 .. Lt, each of at most kTopCells64 cells) go to
    // one workgroup after the per-level launches below them
    auto lev_cnt = [&](int L) { return (int)(ti.lev_off[L + 1] - ti.lev_off[L]); };
    int Lt = 0;
    if (P1 <= kTrP1 - 1)
        while (Lt + 1 <= ti.max_level - 1 && lev_cnt(Lt + 1) <= kTopCells64) ++Lt;
    for (int L = ti.max_level - 1; L > Lt; --L) {
        const int j0 = (int)ti.lev_off[L], j1 = (int)ti.lev_off[L + 1];
        if (j1 <= j0) continue;
        if (P1 <= kTrP1 - 1)
            k_m2m64w<4><<<div_up(j1 - j0, kTrWaves * 4), kWave * kTrWaves, 0, s>>>(
                j0, j1, P, t_m2m, c_flags, c_cx, c_cy, c_child, ak);
        else if (P1 == kTrP1)
            k_m2m64w<1><<<div_up(j1 - j0, kTrWaves), kWave * kTrWaves, 0, s>>>(
                j0, j1, P, t_m2m, c_flags, c_cx, c_cy, c_child, ak);
        else
            k_m2m64<<<div_up((long long)(j1 - j0) * P1, 256), 256, 0, s>>>(
                j0, j1, P, t_m2m, c_flags, c_cx, c_cy, c_child, ak);
    }
    if (Lt >= 1)
        k_m2m64_top<<<1, kWave * kTopWaves64, 0, s>>>(Lt, c.hdr_d, P, t_m2m, c_flags, c_cx, c_cy,
                                                     c_child, ak);
    mk.mark(3, s);
    // M2L: every cell of level >= 2
    const int j2 = (int)ti.lev_off[2];
    const int NP = potential ? P1 : P;  // orders formed per cell
    const double *t_r64 = (const double *)c.tab_r64.p;
    if (nc > j2 && P >= 1 && P < kM2lwP1 && !potential && t_r64 && c.tab_r64_p == P) {
        // list 2 on the matrix cores, then list 4 added by the wavefront kernel
        int ntiles = 0;
        for (int L = 2; L <= ti.max_level; ++L)
            ntiles += (int)div_up((long long)(ti.lev_off[L + 1] - ti.lev_off[L]), 16LL);
        const uint32_t *c_key = (const uint32_t *)c.c_key.p;
        const int *l2 = (const int *)c.l2.p;
        const uint8_t *n2 = (const uint8_t *)c.n2.p;
        switch (P) {
#define M2L64_CASE(PP)                                                                          \
    case PP:                                                                                    \
        k_m2l64_mfma<PP><<<ntiles, kWave * kM2l64Waves, 0, s>>>(c.hdr_d, ti.max_level, c_key, l2, \
                                                              n2, t_r64, ak, bk);               \
        break;
            M2L64_CASE(1) M2L64_CASE(2) M2L64_CASE(3) M2L64_CASE(4) M2L64_CASE(5) M2L64_CASE(6)
            M2L64_CASE(7) M2L64_CASE(8) M2L64_CASE(9) M2L64_CASE(10) M2L64_CASE(11)
            M2L64_CASE(12) M2L64_CASE(13) M2L64_CASE(14) M2L64_CASE(15) M2L64_CASE(16)
#undef M2L64_CASE
            default: break;
        }
        k_m2l64w<false, true><<<div_up(nc - j2, kM2lwWaves), kWave * kM2lwWaves, 0, s>>>(
            j2, nc, P, list4_mode, t_m2l, l2, n2, (const int *)c.l4.p,
            (const uint8_t *)c.n4.p, c_cx, c_cy, c_begin, c_end, xs, ys, qs, as, ak, bk);
    } else if (nc > j2 && P < kM2lwP1)
        (potential ? k_m2l64w<true, false> : k_m2l64w<false, false>)<<<
            div_up(nc - j2, kM2lwWaves), kWave * kM2lwWaves, 0, s>>>(
            j2, nc, P, list4_mode, t_m2l, (const int *)c.l2.p, (const uint8_t *)c.n2.p,
            (const int *)c.l4.p, (const uint8_t *)c.n4.p, c_cx, c_cy, c_begin, c_end, xs, ys, qs,
            as, ak, bk);
    else if (nc > j2)
        (potential ? k_m2l64<true, 0> : k_m2l64<false, 0>)<<<
            div_up((long long)(nc - j2) * NP, 256), 256, 0, s>>>(
            j2, nc, P, list4_mode, t_m2l, (const int *)c.l2.p, (const uint8_t *)c.n2.p,
            (const int *)c.l4.p, (const uint8_t *)c.n4.p, c_cx, c_cy, c_begin, c_end, xs, ys, qs,
            as, ak, bk);
    mk.mark(4, s);
    // L2L: children at levels 3 .. maxL (parents 2 .. maxL-1, fmm2D.cpp:587);
    // the top children levels (3 .. Lc, each of at most kTopCells64 cells) in
    // one workgroup first
    int Lc = 2;
    if (P1 <= kTrP1 - 1)
        while (Lc + 1 <= ti.max_level && lev_cnt(Lc + 1) <= kTopCells64) ++Lc;
    if (Lc >= 3)
        (potential ? k_l2l64_top<true> : k_l2l64_top<false>)<<<1, kWave * kTopWaves64, 0, s>>>(
            Lc, c.hdr_d, P, t_l2l, c_parent, c_cx, c_cy, bk);
    for (int L = std::max(3, Lc + 1); L <= ti.max_level; ++L) {
        const int j0 = (int)ti.lev_off[L], j1 = (int)ti.lev_off[L + 1];
        if (j1 <= j0) continue;
        if (P1 <= kTrP1 - 1)
            (potential ? k_l2l64w<true, 4> : k_l2l64w<false, 4>)<<<div_up(j1 - j0, kTrWaves * 4),
                                                                  kWave * kTrWaves, 0, s>>>(
                j0, j1, P, t_l2l, c_parent, c_cx, c_cy, bk);
        else if (P1 == kTrP1)
            (potential ? k_l2l64w<true, 1> : k_l2l64w<false, 1>)<<<div_up(j1 - j0, kTrWaves),
                                                                  kWave * kTrWaves, 0, s>>>(
                j0, j1, P, t_l2l, c_parent, c_cx, c_cy, bk);
        else
            (potential ? k_l2l64<true> : k_l2l64<false>)<<<div_up((long long)(j1 - j0) * NP, 256),
                                                            256, 0, s>>>(j0, j1, P, t_l2l, c_parent,
                                                                         c_cx, c_cy, bk);
    }
    mk.mark(5, s);
    if (potential) {
        k_pot64<<<div_up(nl, kEvalWaves), kWave * kEvalWaves, 0, s>>>(
            leaf_idx, nl, P, c_begin, c_end, c_cx, c_cy, (const uint32_t *)c.off1.p,
            (const int *)c.l1.p, (const uint32_t *)c.off3.p, (const int *)c.l3.p, xs, ys, qs,
            c.perm_sorted, ak, bk, stage_mask, out_a, &c.hdr_d->flags);
        mk.mark(6, s);
        VPM_LAUNCH_CHECK();
        return;
    }
    const double *ss = (const double *)c.ss.p;
    auto eval = [&](auto kern) {
        kern<<<div_up(nl, kEvalWaves), kWave * kEvalWaves, 0, s>>>(
            leaf_idx, nl, P, c_begin, c_end, c_cx, c_cy, (const uint32_t *)c.off1.p,
            (const int *)c.l1.p, (const uint32_t *)c.off3.p, (const int *)c.l3.p, xs, ys, qs, ss,
            c.perm_sorted, ak, bk, velocity ? 1 : 0, u_inf, v_inf, stage_mask, out_a, out_b,
            &c.hdr_d->flags);
    };
    switch (c.icut) {
        case 1: eval(k_eval64<1>); break;
        case 2: eval(k_eval64<2>); break;
        case 3: eval(k_eval64<3>); break;
        default: eval(k_eval64<0>); break;
    }
    mk.mark(6, s);
    VPM_LAUNCH_CHECK();
}

void fmm_t1_run(Ctx &c, int ND, double *u, double *v, double u_inf, double v_inf,
                const FmmMarks &mk) {
    hipStream_t s = c.stream;
    const TreeInfo &ti = c.info;
    const int nc = ti.n_cells, nl = ti.n_leaves, NL = ND + 1, PM = ND - 1;
    double2 *ak = c.ak.as<double2>((size_t)nc * ND), *bk = c.bk.as<double2>((size_t)nc * NL);
    VPM_HIP(hipMemsetAsync(ak, 0, sizeof(double2) * (size_t)nc * ND, s));
    const uint8_t *c_flags = (const uint8_t *)c.c_flags.p, *c_level = (const uint8_t *)c.c_level.p;
    const uint32_t *c_begin = (const uint32_t *)c.c_begin.p, *c_end = (const uint32_t *)c.c_end.p;
    const double *c_cx = (const double *)c.c_cx.p, *c_cy = (const double *)c.c_cy.p;
    const int *c_child = (const int *)c.c_child.p, *c_parent = (const int *)c.c_parent.p;
    const double *xs = (const double *)c.xs.p, *ys = (const double *)c.ys.p,
                 *qs = (const double *)c.qs.p;
    const uint8_t *as = (const uint8_t *)c.acts.p;
    const uint32_t *leaf_idx = (const uint32_t *)c.leaf_idx.p;

    // upward: P2M and M2M at order PM = ND - 1 (biotsavart_fmm_2d.cpp:123-188;
    // the exact shift C(p, k), the reference's bico values to rounding)
    k_p2m64<<<div_up(nl, kP2m64Waves * kP2mLeaves), kWave * kP2m64Waves, 0, s>>>(
        leaf_idx, nl, PM, c_flags, c_begin, c_end, c_cx, c_cy, xs, ys, qs, as, ak);
    for (int L = ti.max_level - 1; L >= 1; --L) {
        const int j0 = (int)ti.lev_off[L], j1 = (int)ti.lev_off[L + 1];
        if (j1 <= j0) continue;
        k_m2m64<<<div_up((long long)(j1 - j0) * ND, 256), 256, 0, s>>>(
            j0, j1, PM, (const double *)c.t1_m2m.p, c_flags, c_cx, c_cy, c_child, ak);
    }
    mk.mark(3, s);
    // M2L and P2L: every cell of level >= 1 (all of bk written)
    const int j1 = (int)ti.lev_off[1];
    if (nc > j1)
        k_m2l_t1<<<div_up((long long)(nc - j1) * NL, 256), 256, 0, s>>>(
            j1, nc, ND, (const double *)c.t1_m2l.p, c_level, (const int *)c.l2.p,
            (const uint8_t *)c.n2.p, (const int *)c.l4.p, (const uint8_t *)c.n4.p, c_cx, c_cy,
            c_begin, c_end, xs, ys, qs, ak, bk);
    mk.mark(4, s);
    for (int L = 3; L <= ti.max_level; ++L) {
        const int j0 = (int)ti.lev_off[L], jz = (int)ti.lev_off[L + 1];
        if (jz <= j0) continue;
        k_l2l_t1<<<div_up((long long)(jz - j0) * NL, 256), 256, 0, s>>>(
            j0, jz, ND, (const double *)c.t1_l2l_c.p, (const int *)c.t1_l2l_e.p, c_parent, c_cx,
            c_cy, bk);
    }
    mk.mark(5, s);
    const double *ss = (const double *)c.ss.p;
    auto eval = [&](auto kern) {
        kern<<<div_up(nl, kEvalWaves), kWave * kEvalWaves, 0, s>>>(
            leaf_idx, nl, ND, c_begin, c_end, c_cx, c_cy, (const uint32_t *)c.off1.p,
            (const int *)c.l1.p, (const uint32_t *)c.off3.p, (const int *)c.l3.p, xs, ys, qs, ss,
            c.perm_sorted, ak, bk, 1, u_inf, v_inf, 7, u, v, &c.hdr_d->flags);
    };
    switch (c.icut) {
        case 1: eval(k_eval64<1, true>); break;
        case 2: eval(k_eval64<2, true>); break;
        case 3: eval(k_eval64<3, true>); break;
        default: eval(k_eval64<0, true>); break;
    }
    mk.mark(6, s);
    VPM_LAUNCH_CHECK();
}

// Partial sums of |(a, b) - (ja, jb)|^2 and |(ja - a0, jb - b0)|^2 per block.
__global__ void __launch_bounds__(256)
    k_diff_norms(int n, const double *a, const double *b, const double *ja, const double *jb,
                 double a0, double b0, double *part) {
    double num = 0.0, den = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double da = a[i] - ja[i], db = b[i] - jb[i];
        const double ra = ja[i] - a0, rb = jb[i] - b0;
        num += da * da + db * db;
        den += ra * ra + rb * rb;
    }
    for (int m = 1; m < kWave; m <<= 1) {
        num += __shfl_xor(num, m, kWave);
        den += __shfl_xor(den, m, kWave);
    }
    __shared__ double sn[4], sd[4];
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        sn[w] = num;
        sd[w] = den;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sn[0] + sn[1] + sn[2] + sn[3];
        part[2 * blockIdx.x + 1] = sd[0] + sd[1] + sd[2] + sd[3];
    }
}

// Relative L2 difference of (a, b) against the judge (ja, jb), whose
// reference level (a0, b0) is removed from the norm (the free stream of a
// velocity): sqrt(sum |d|^2 / sum |j - (a0, b0)|^2).  Synchronises the stream.
double rel_diff64(int n, const double *a, const double *b, const double *ja, const double *jb,
                  double a0, double b0, double *part_dev, hipStream_t s) {
    constexpr int kBlocks = 512;
    k_diff_norms<<<kBlocks, 256, 0, s>>>(n, a, b, ja, jb, a0, b0, part_dev);
    VPM_LAUNCH_CHECK();
    std::vector<double> h(2 * kBlocks);
    VPM_HIP(hipMemcpyAsync(h.data(), part_dev, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                           s));
    VPM_HIP(hipStreamSynchronize(s));
    double num = 0.0, den = 0.0;
    for (int k = 0; k < kBlocks; ++k) {
        num += h[2 * k];
        den += h[2 * k + 1];
    }
    return den > 0.0 ? std::sqrt(num / den) : (num > 0.0 ? INFINITY : 0.0);
}

__global__ void __launch_bounds__(256)
    k_fill_pair(double *a, double *b, int n, double va, double vb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        a[i] = va;
        b[i] = vb;
    }
}

void fill_pair(double *a, double *b, int n, double va, double vb, hipStream_t s) {
    k_fill_pair<<<div_up(n, 256), 256, 0, s>>>(a, b, n, va, vb);
    VPM_LAUNCH_CHECK();
}

void direct64(int n, const double *x, const double *y, const double *q, const uint8_t *src_mask,
              int nt, const int *tidx, double *ex, double *ey, hipStream_t s) {
    k_direct64<<<div_up(nt, 256), 256, 0, s>>>(n, x, y, q, src_mask, nt, tidx, ex, ey);
    VPM_LAUNCH_CHECK();
}

// Realified list-2 M2L matrices of order P for k_m2l64_mfma, one per offset
// delta = (dx, dy) (source minus target centre over the cell side; |dx|,
// |dy| <= 3, not both <= 1; slot order, dy major): row 2(p-1) (Re b^_p) and
// 2(p-1)+1 (Im), p = 1..P; column 2k (Re a^_k) and 2k+1 (Im), k = 0..P;
// entry M[p][k] = tab[p][k] delta^-(p+k), tab being the path's M2L table
// (tab[p][0] = -1/p, tab[p][k] = -C(p+k-1, k-1) (-1)^k / k with its
// binomials).  Stored in the A-operand lane order of v_mfma_f64_16x16x4f64:
// R[((o KS + ks) MT + mt) 64 + l] = M[16 mt + l % 16][4 ks + l / 16].
void fmm64_mfma_table(int P, const double *tab, std::vector<double> &R) {
    const int P1 = P + 1, K = 2 * P1, KS = (K + 3) / 4, MT = (2 * P + 15) / 16;
    const bool left = 2 * P > 16 && 2 * P <= 20;  // rows 16.. for the 4x4x4 instruction
    R.assign((size_t)40 * KS * MT * kWave, 0.0);
    std::vector<double> pwr(2 * P + 2), pwi(2 * P + 2);
    int o = 0;
    for (int dy = -3; dy <= 3; ++dy) {
        for (int dx = -3; dx <= 3; ++dx) {
            if (std::abs(dx) <= 1 && std::abs(dy) <= 1) continue;
            const double r2 = (double)(dx * dx + dy * dy);
            const double wr = dx / r2, wi = -dy / r2;  // 1 / delta
            pwr[0] = 1.0;
            pwi[0] = 0.0;
            for (int t = 1; t <= 2 * P + 1; ++t) {
                pwr[t] = pwr[t - 1] * wr - pwi[t - 1] * wi;
                pwi[t] = pwr[t - 1] * wi + pwi[t - 1] * wr;
            }
            for (int p = 1; p <= P; ++p) {
                for (int k = 0; k <= P; ++k) {
                    const double T = tab[(size_t)p * P1 + k];
                    const double Mr = T * pwr[p + k], Mi = T * pwi[p + k];
                    const int r0 = 2 * (p - 1), c0 = 2 * k;
                    const double val[2][2] = {{Mr, -Mi}, {Mi, Mr}};
                    for (int a = 0; a < 2; ++a) {
                        for (int bq = 0; bq < 2; ++bq) {
                            const int row = r0 + a, col = c0 + bq;
                            const int mt = row / 16, ks = col / 4;
                            const size_t base = (((size_t)o * KS + ks) * MT + mt) * kWave;
                            if (left && row >= 16) {
                                // v_mfma_f64_4x4x4f64 A operand: row 16 + l % 4,
                                // step column l / 16, the same for the four blocks
                                for (int q = 0; q < 4; ++q)
                                    R[base + (row - 16) + 4 * q + 16 * (col % 4)] = val[a][bq];
                            } else {
                                const int lane = (row % 16) + 16 * (col % 4);
                                R[base + lane] = val[a][bq];
                            }
                        }
                    }
                }
            }
            ++o;
        }
    }
}

}  // namespace vpm
