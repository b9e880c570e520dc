// prim.hip -- exclusive scans and a stable LSD radix sort for gfx950.
//
// Scans: tile reduce -> down-sweep over 2048-element tiles (256 threads x
// 8 items), wave64 shuffles inside a tile; the down-sweep block sums the tile
// totals before it itself (up to kFusedSpine tiles), larger scans add a spine
// launch.  When the kernel that produced the input already wrote the sums of
// its runs of `per` elements (block sums), the reduce pass is skipped and the
// scan is one launch.  Radix sort: 8-bit digits over 1024-element tiles;
// per-block digit histograms, one launch that scans each digit's row of the
// digit-major histogram table, then a stable scatter that adds the digit
// bases (a block scan of the 256 digit totals) and whose in-tile ranks come
// from 64-lane ballot "peer masks" (lanes of a wavefront holding the same
// digit).
#include "prim.hpp"

namespace vpm {
namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

constexpr int kFusedSpine = 2048;  // tiles up to which the down-sweep forms its own prefix

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        uint32_t y = __shfl_up(x, o, kWave);
        if (lane >= o) x += y;
    }
    return x;
}

// Exclusive scan across a block of BLOCK threads.  `sm` holds BLOCK/64 words.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *sm, uint32_t &total) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    uint32_t inc = wave_incl_scan(v);
    if (lane == kWave - 1) sm[w] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int i = 0; i < BLOCK / kWave; ++i) {
        uint32_t s = sm[i];
        before += (i < w) ? s : 0u;
        all += s;
    }
    __syncthreads();
    total = all;
    return before + inc - v;
}

template <typename T>
__global__ void __launch_bounds__(kScanBlock) k_scan_reduce(const T *in, int n, uint32_t *bsum) {
    __shared__ uint32_t sm[kScanBlock / kWave];
    const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        long long i = base + j;
        if (i < n) s += (uint32_t)in[i];
    }
    uint32_t tot;
    block_excl_scan<kScanBlock>(s, sm, tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(1024) k_scan_spine(uint32_t *bsum, int nb, uint32_t *total) {
    __shared__ uint32_t sm[1024 / kWave];
    uint32_t carry = 0;
    for (int base = 0; base < nb; base += 1024) {
        int i = base + threadIdx.x;
        uint32_t v = i < nb ? bsum[i] : 0u;
        uint32_t tot;
        uint32_t ex = block_excl_scan<1024>(v, sm, tot);
        if (i < nb) bsum[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

template <typename T>
__global__ void __launch_bounds__(kScanBlock)
    k_scan_down(const T *in, uint32_t *out, int n, const uint32_t *bsum, const uint32_t *total,
                int write_end) {
    __shared__ uint32_t sm[kScanBlock / kWave];
    const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        long long i = base + j;
        v[j] = i < n ? (uint32_t)in[i] : 0u;
        s += v[j];
    }
    uint32_t tot;
    uint32_t run = bsum[blockIdx.x] + block_excl_scan<kScanBlock>(s, sm, tot);
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        long long i = base + j;
        if (i < n) out[i] = run;
        run += v[j];
    }
    if (write_end && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = *total;
}

// Down-sweep that forms its tile prefix itself: block b sums bsum[0..b).
template <typename T>
__global__ void __launch_bounds__(kScanBlock)
    k_scan_down_fused(const T *in, uint32_t *out, int n, const uint32_t *bsum, uint32_t *total,
                      int write_end) {
    __shared__ uint32_t sm[kScanBlock / kWave];
    uint32_t pre = 0;
    for (int b = threadIdx.x; b < (int)blockIdx.x; b += kScanBlock) pre += bsum[b];
    uint32_t pre_all;
    block_excl_scan<kScanBlock>(pre, sm, pre_all);
    const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        long long i = base + j;
        v[j] = i < n ? (uint32_t)in[i] : 0u;
        s += v[j];
    }
    uint32_t tot;
    uint32_t run = pre_all + block_excl_scan<kScanBlock>(s, sm, tot);
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        long long i = base + j;
        if (i < n) out[i] = run;
        run += v[j];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
        if (total) *total = pre_all + tot;
        if (write_end) out[n] = pre_all + tot;
    }
}

// Down-sweep from block sums of the producer: bsum[k] is the sum of
// in[k per, (k+1) per); tile b (kScanTile elements, a multiple of per) adds
// pre(b) = bsum[0 .. b kScanTile / per), summed by the block itself, to its
// own scan.  out (if not null) receives the exclusive scan; compact (if not
// null, for 0/1 inputs) receives i at position out[i] for every in[i] = 1.
template <typename T>
struct ScanBs {
    const T *in;
    uint32_t *out, *compact;
    const uint32_t *bsum;
    uint32_t *total;
};

// kScanItems consecutive elements from base (aligned: base is a multiple of
// kScanItems); vector loads when the whole run is in range.
__device__ __forceinline__ void load_items(const uint8_t *in, long long base, int n,
                                           uint32_t (&v)[kScanItems]) {
    if (base + kScanItems <= n) {
        const uint2 w = *reinterpret_cast<const uint2 *>(in + base);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            v[j] = (w.x >> (8 * j)) & 255u;
            v[j + 4] = (w.y >> (8 * j)) & 255u;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) v[j] = base + j < n ? (uint32_t)in[base + j] : 0u;
    }
}
__device__ __forceinline__ void load_items(const uint32_t *in, long long base, int n,
                                           uint32_t (&v)[kScanItems]) {
    if (base + kScanItems <= n) {
        const uint4 a = *reinterpret_cast<const uint4 *>(in + base);
        const uint4 b = *reinterpret_cast<const uint4 *>(in + base + 4);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) v[j] = base + j < n ? in[base + j] : 0u;
    }
}

// blockIdx.y selects one of up to two independent scans of the same length.
template <typename T>
__global__ void __launch_bounds__(kScanBlock)
    k_scan_down_bs(ScanBs<T> a0, ScanBs<T> a1, int n, int per, int nbs, int write_end,
                   const unsigned long long *skip, int skip_n, int skip_stride) {
    __shared__ uint32_t sm[kScanBlock / kWave];
    if (skip) {
        // the producer counted the zeros of a 0/1 input: none means out and
        // compact are the identity, which the caller then uses unwritten
        __shared__ int s_skip;
        if (threadIdx.x == 0) {
            unsigned long long z = 0;
            for (int k = 0; k < skip_n; ++k) z += skip[(size_t)k * skip_stride];
            s_skip = z == 0;
        }
        __syncthreads();
        if (s_skip) {
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                if (a0.total) *a0.total = (uint32_t)n;
                if (write_end && a0.out) a0.out[n] = (uint32_t)n;
            }
            return;
        }
    }
    const ScanBs<T> &a = blockIdx.y == 0 ? a0 : a1;
    const T *in = a.in;
    uint32_t *out = a.out, *compact = a.compact, *total = a.total;
    const uint32_t *bsum = a.bsum;
    const int kend = (int)min((long long)nbs, (long long)blockIdx.x * kScanTile / per);
    uint32_t pre = 0, pre_all;
    for (int k = threadIdx.x; k < kend; k += kScanBlock) pre += bsum[k];
    block_excl_scan<kScanBlock>(pre, sm, pre_all);
    const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    load_items(in, base, n, v);
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) s += v[j];
    uint32_t tot;
    uint32_t run = pre_all + block_excl_scan<kScanBlock>(s, sm, tot);
    uint32_t r[kScanItems];
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        r[j] = run;
        if (compact && v[j] && base + j < n) compact[run] = (uint32_t)(base + j);
        run += v[j];
    }
    if (out) {
        if (base + kScanItems <= n) {
            *reinterpret_cast<uint4 *>(out + base) = make_uint4(r[0], r[1], r[2], r[3]);
            *reinterpret_cast<uint4 *>(out + base + 4) = make_uint4(r[4], r[5], r[6], r[7]);
        } else {
#pragma unroll
            for (int j = 0; j < kScanItems; ++j)
                if (base + j < n) out[base + j] = r[j];
        }
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
        if (total) *total = pre_all + tot;
        if (write_end && out) out[n] = pre_all + tot;
    }
}

__global__ void k_write_zero_total(uint32_t *out, int n, uint32_t *total, int write_end) {
    if (total) *total = 0;
    if (write_end) out[n] = 0;
}

template <typename T>
void scan_impl(const T *in, uint32_t *out, int n, uint32_t *total, bool write_end, uint32_t *tmp,
               hipStream_t s) {
    if (n <= 0) {
        k_write_zero_total<<<1, 1, 0, s>>>(out, 0, total, write_end ? 1 : 0);
        VPM_LAUNCH_CHECK();
        return;
    }
    const int nb = div_up(n, kScanTile);
    uint32_t *bsum = tmp;
    k_scan_reduce<T><<<nb, kScanBlock, 0, s>>>(in, n, bsum);
    if (nb <= kFusedSpine) {
        k_scan_down_fused<T><<<nb, kScanBlock, 0, s>>>(in, out, n, bsum, total, write_end ? 1 : 0);
        VPM_LAUNCH_CHECK();
        return;
    }
    uint32_t *tot = total ? total : tmp + nb;
    k_scan_spine<<<1, 1024, 0, s>>>(bsum, nb, tot);
    k_scan_down<T><<<nb, kScanBlock, 0, s>>>(in, out, n, bsum, tot, write_end ? 1 : 0);
    VPM_LAUNCH_CHECK();
}

// Exclusive scan of each digit's row of the histogram table (block = digit),
// in place; the digit's total goes to totals[digit].
__global__ void __launch_bounds__(kRsBlock)
    k_rs_rowscan(uint32_t *table, int nb, uint32_t *totals) {
    __shared__ uint32_t sm[kRsBlock / kWave];
    uint32_t *row = table + (size_t)blockIdx.x * nb;
    constexpr int kItems = 4;
    uint32_t carry = 0;
    for (int b0 = 0; b0 < nb; b0 += kRsBlock * kItems) {
        const int i0 = b0 + threadIdx.x * kItems;
        uint32_t v[kItems], sum = 0;
#pragma unroll
        for (int u = 0; u < kItems; ++u) {
            v[u] = i0 + u < nb ? row[i0 + u] : 0u;
            sum += v[u];
        }
        uint32_t tot;
        uint32_t run = carry + block_excl_scan<kRsBlock>(sum, sm, tot);
#pragma unroll
        for (int u = 0; u < kItems; ++u) {
            if (i0 + u < nb) row[i0 + u] = run;
            run += v[u];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

__global__ void __launch_bounds__(kRsBlock)
    k_rs_hist(const uint32_t *keys, int n, int shift, uint32_t *table, int nb) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const long long base = (long long)blockIdx.x * kRsTile;
#pragma unroll 4
    for (int j = 0; j < kRsItems; ++j) {
        long long i = base + (long long)j * kRsBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    table[(size_t)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: element order inside a tile is (item j, wave w, lane l),
// which is the index order.  Each wavefront ranks its keys by digit with
// 64-lane ballot peer masks (lanes holding the same digit); the per-digit
// counts of every (item, wave) slot go to LDS, one exclusive scan per digit
// over the slots (in index order) gives each key its rank inside the tile's
// digit run.  The tile is then sorted by digit in LDS and written out in that
// order, so consecutive threads store consecutive positions of a digit's run
// in the output (the digit base is the scan of the digit totals plus this
// tile's row prefix).
__global__ void __launch_bounds__(kRsBlock)
    k_rs_scatter(const uint32_t *kin, const uint32_t *vin, uint32_t *kout, uint32_t *vout, int n,
                 int shift, const uint32_t *table, const uint32_t *totals, int nb) {
    constexpr int kW = kRsBlock / kWave;
    constexpr int kSlots = kRsItems * kW;  // (item, wave) slots in index order
    __shared__ uint32_t cnt[kSlots][256];
    __shared__ uint32_t goff[256], lbase[256];
    __shared__ uint32_t sk[kRsTile], sv[kRsTile];
    __shared__ uint32_t sm[kRsBlock / kWave];
    const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
#pragma unroll
    for (int q = 0; q < kSlots; ++q) cnt[q][t] = 0;
    // digit base in the output: exclusive scan of the digit totals + row prefix
    uint32_t all;
    const uint32_t dbase = block_excl_scan<kRsBlock>(totals[t], sm, all);  // synchronises
    const uint32_t gbase = dbase + table[(size_t)t * nb + blockIdx.x];
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (kWave - lane));
    const long long base = (long long)blockIdx.x * kRsTile;
    const int tile_n = (int)min((long long)kRsTile, (long long)n - base);
    uint32_t k[kRsItems], v[kRsItems], d[kRsItems], rank[kRsItems];
    bool valid[kRsItems];
#pragma unroll
    for (int j = 0; j < kRsItems; ++j) {
        const long long i = base + (long long)j * kRsBlock + t;
        valid[j] = i < n;
        k[j] = valid[j] ? kin[i] : 0u;
        v[j] = valid[j] ? vin[i] : 0u;
        d[j] = (k[j] >> shift) & 255u;
        uint64_t peers = __ballot(valid[j]);
#pragma unroll
        for (int bb = 0; bb < 8; ++bb) {
            const uint32_t bit = (d[j] >> bb) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        rank[j] = (uint32_t)__popcll(peers & below);
        // the lowest lane of each digit group writes the group's count
        if (valid[j] && (peers & below) == 0) cnt[j * kW + w][d[j]] = (uint32_t)__popcll(peers);
    }
    __syncthreads();
    // thread t = digit: exclusive scan over the slots, then the tile-local
    // digit bases (exclusive scan of the tile's digit counts)
    uint32_t run = 0;
#pragma unroll
    for (int q = 0; q < kSlots; ++q) {
        const uint32_t c = cnt[q][t];
        cnt[q][t] = run;
        run += c;
    }
    uint32_t tile_all;
    const uint32_t lb = block_excl_scan<kRsBlock>(run, sm, tile_all);  // synchronises
    lbase[t] = lb;
    goff[t] = gbase;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kRsItems; ++j) {
        if (!valid[j]) continue;
        const uint32_t pos = lbase[d[j]] + cnt[j * kW + w][d[j]] + rank[j];
        sk[pos] = k[j];
        sv[pos] = v[j];
    }
    __syncthreads();
    for (int p = t; p < tile_n; p += kRsBlock) {
        const uint32_t key = sk[p];
        const uint32_t dd = (key >> shift) & 255u;
        const uint32_t dst = goff[dd] + ((uint32_t)p - lbase[dd]);
        kout[dst] = key;
        vout[dst] = sv[p];
    }
}

}  // namespace

size_t scan_tmp_words(long long n) { return (size_t)div_up(n > 0 ? n : 1, kScanTile) + 2; }

void scan_exclusive_u32(const uint32_t *in, uint32_t *out, int n, uint32_t *total,
                        bool write_total_at_end, uint32_t *tmp, hipStream_t s) {
    scan_impl<uint32_t>(in, out, n, total, write_total_at_end, tmp, s);
}

void scan_exclusive_u8(const uint8_t *in, uint32_t *out, int n, uint32_t *total,
                       bool write_total_at_end, uint32_t *tmp, hipStream_t s) {
    scan_impl<uint8_t>(in, out, n, total, write_total_at_end, tmp, s);
}

// Compaction of 0/1 flags into the positions of the ones (helper of the
// long-input fallback): compact[scan[i]] = i where flag[i].
template <typename T>
__global__ void __launch_bounds__(256)
    k_compact(const T *flag, const uint32_t *scan, int n, uint32_t *compact) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i]) compact[scan[i]] = (uint32_t)i;
}

template <typename T>
void scan_bs_impl(const T *in, uint32_t *out, uint32_t *compact, int n, uint32_t *bsum, int per,
                  uint32_t *total, bool write_end, uint32_t *tmp, hipStream_t s,
                  const unsigned long long *skip = nullptr, int skip_n = 0, int skip_stride = 0) {
    if (n <= 0) {
        k_write_zero_total<<<1, 1, 0, s>>>(out, 0, total, (write_end && out) ? 1 : 0);
        VPM_LAUNCH_CHECK();
        return;
    }
    if (per <= 0 || kScanTile % per != 0) throw Error(kInternal, "scan block sums: bad run length");
    if (!out && !compact) throw Error(kInternal, "scan block sums: no output");
    const int nb = div_up(n, kScanTile), nbs = div_up(n, per);
    if (nbs > kFusedSpine * 4) {
        // very long inputs: the tile reduce (input read once more) beats a
        // one-block spine over this many block sums
        if (!out) throw Error(kInternal, "scan block sums: long compaction needs the scan output");
        scan_impl<T>(in, out, n, total, write_end, tmp, s);
        if (compact) k_compact<T><<<div_up(n, 256), 256, 0, s>>>(in, out, n, compact);
        VPM_LAUNCH_CHECK();
        return;
    }
    const ScanBs<T> a{in, out, compact, bsum, total};
    k_scan_down_bs<T><<<nb, kScanBlock, 0, s>>>(a, a, n, per, nbs, write_end ? 1 : 0, skip,
                                                skip_n, skip_stride);
    VPM_LAUNCH_CHECK();
}

void scan_exclusive_bs_u32(const uint32_t *in, uint32_t *out, int n, uint32_t *bsum, int per,
                           uint32_t *total, bool write_total_at_end, uint32_t *tmp,
                           hipStream_t s) {
    scan_bs_impl<uint32_t>(in, out, nullptr, n, bsum, per, total, write_total_at_end, tmp, s);
}

void scan_exclusive_bs_u32_pair(const uint32_t *in0, uint32_t *out0, uint32_t *bsum0,
                                uint32_t *total0, const uint32_t *in1, uint32_t *out1,
                                uint32_t *bsum1, uint32_t *total1, int n, int per,
                                bool write_total_at_end, uint32_t *tmp, hipStream_t s) {
    const int nbs = n > 0 ? div_up(n, per) : 0;
    if (n <= 0 || per <= 0 || kScanTile % per != 0 || nbs > kFusedSpine * 4) {
        scan_bs_impl<uint32_t>(in0, out0, nullptr, n, bsum0, per, total0, write_total_at_end, tmp,
                               s);
        scan_bs_impl<uint32_t>(in1, out1, nullptr, n, bsum1, per, total1, write_total_at_end, tmp,
                               s);
        return;
    }
    const ScanBs<uint32_t> a0{in0, out0, nullptr, bsum0, total0};
    const ScanBs<uint32_t> a1{in1, out1, nullptr, bsum1, total1};
    k_scan_down_bs<uint32_t><<<dim3(div_up(n, kScanTile), 2), kScanBlock, 0, s>>>(
        a0, a1, n, per, nbs, write_total_at_end ? 1 : 0, nullptr, 0, 0);
    VPM_LAUNCH_CHECK();
}

void scan_exclusive_bs_u8(const uint8_t *in, uint32_t *out, uint32_t *compact, int n,
                          uint32_t *bsum, int per, uint32_t *total, bool write_total_at_end,
                          uint32_t *tmp, hipStream_t s, const unsigned long long *zeros,
                          int zeros_n, int zeros_stride) {
    scan_bs_impl<uint8_t>(in, out, compact, n, bsum, per, total, write_total_at_end, tmp, s,
                          zeros, zeros_n, zeros_stride);
}

size_t radix_tmp_words(long long n) {
    const long long nb = div_up(n > 0 ? n : 1, kRsTile);
    return (size_t)(256 * nb) + 256;
}

bool radix_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t *keys_alt, uint32_t *vals_alt,
                      int n, int begin_bit, int end_bit, uint32_t *tmp, hipStream_t s,
                      bool first_counts) {
    if (n <= 1) return false;
    const int nb = div_up(n, kRsTile);
    uint32_t *table = tmp;
    uint32_t *totals = tmp + (size_t)256 * nb;
    bool alt = false;
    for (int shift = begin_bit; shift < end_bit; shift += 8) {
        const uint32_t *ki = alt ? keys_alt : keys;
        const uint32_t *vi = alt ? vals_alt : vals;
        uint32_t *ko = alt ? keys : keys_alt;
        uint32_t *vo = alt ? vals : vals_alt;
        if (!(first_counts && shift == begin_bit))
            k_rs_hist<<<nb, kRsBlock, 0, s>>>(ki, n, shift, table, nb);
        k_rs_rowscan<<<256, kRsBlock, 0, s>>>(table, nb, totals);
        k_rs_scatter<<<nb, kRsBlock, 0, s>>>(ki, vi, ko, vo, n, shift, table, totals, nb);
        VPM_LAUNCH_CHECK();
        alt = !alt;
    }
    return alt;
}

}  // namespace vpm
