// prim.hpp -- device-wide primitives: exclusive scans and an LSD radix sort.
#pragma once

#include "common.hpp"

namespace vpm {

// Scratch words needed by the scans for n elements.
size_t scan_tmp_words(long long n);

// out[i] = sum_{j<i} in[j]; if total != nullptr, *total (device) = sum of all.
// out may have room for n+1 entries: when write_total_at_end, out[n] = total.
void scan_exclusive_u32(const uint32_t *in, uint32_t *out, int n, uint32_t *total,
                        bool write_total_at_end, uint32_t *tmp, hipStream_t s);
void scan_exclusive_u8(const uint8_t *in, uint32_t *out, int n, uint32_t *total,
                       bool write_total_at_end, uint32_t *tmp, hipStream_t s);

// The same, for an input whose producer wrote bsum[k] = sum of
// in[k per, (k+1) per) (per divides the 2048-element scan tile): one launch
// (the two-launch scan for inputs of more than 8192 runs).
void scan_exclusive_bs_u32(const uint32_t *in, uint32_t *out, int n, uint32_t *bsum, int per,
                           uint32_t *total, bool write_total_at_end, uint32_t *tmp,
                           hipStream_t s);
// Two independent such scans of the same length in one launch.
void scan_exclusive_bs_u32_pair(const uint32_t *in0, uint32_t *out0, uint32_t *bsum0,
                                uint32_t *total0, const uint32_t *in1, uint32_t *out1,
                                uint32_t *bsum1, uint32_t *total1, int n, int per,
                                bool write_total_at_end, uint32_t *tmp, hipStream_t s);
// For 0/1 flags: out (may be null) gets the scan, compact (may be null) the
// positions of the ones in order (compact[out[i]] = i where in[i] = 1).
// zeros (may be null): zeros_n device counters, zeros_stride apart, whose sum
// is the number of zero flags; when it is 0, only the total (n, and out[n])
// may be written and the caller reads out and compact as the identity.
void scan_exclusive_bs_u8(const uint8_t *in, uint32_t *out, uint32_t *compact, int n,
                          uint32_t *bsum, int per, uint32_t *total, bool write_total_at_end,
                          uint32_t *tmp, hipStream_t s, const unsigned long long *zeros = nullptr,
                          int zeros_n = 0, int zeros_stride = 0);

// Scratch words needed by the radix sort for n elements.
size_t radix_tmp_words(long long n);

// Radix tiles: kRsTile keys per workgroup of kRsBlock threads, element
// base + j kRsBlock + thread for j < kRsItems.  The per-tile digit counts of a
// pass live in tmp[digit * tiles + tile] (tiles = radix_tiles(n)).
constexpr int kRsBlock = 256;  // = number of digit values
constexpr int kRsItems = 4;
constexpr int kRsTile = kRsBlock * kRsItems;
inline int radix_tiles(long long n) { return (int)((n + kRsTile - 1) / kRsTile); }

// Stable LSD radix sort of (key, val) pairs on key bits [begin_bit, end_bit),
// 8 bits per pass.  Buffers ping-pong between (keys, vals) and
// (keys_alt, vals_alt); the return value tells where the result lies
// (false: keys/vals, true: keys_alt/vals_alt).  first_counts: the kernel that
// wrote the keys also wrote the first pass's per-tile digit counts into tmp.
bool radix_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t *keys_alt, uint32_t *vals_alt,
                      int n, int begin_bit, int end_bit, uint32_t *tmp, hipStream_t s,
                      bool first_counts = false);

}  // namespace vpm
