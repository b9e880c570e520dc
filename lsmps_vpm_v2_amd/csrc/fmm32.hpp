// fmm32.hpp -- mixed-precision path: per-level table and launch context.
#pragma once

#include <vector>

#include "ctx.hpp"

namespace vpm {

// Per-level expansion order and cell side (levels 0..16), device resident.
struct LevelTab {
    int p[kMaxDepth + 2];
    float s[kMaxDepth + 2];
    float inv_s[kMaxDepth + 2];
};

// Device pointers and sizes for one mixed-precision run.
struct Mixed32 {
    int pmax = 0;                      // coefficient storage is 2 (pmax + 1) floats per cell
    const LevelTab *lt = nullptr;
    const float *m2m = nullptr, *l2l = nullptr, *tm2l = nullptr;
    const float *rptr_host[17] = {};   // device pointer of the MFMA-ordered M2L matrices per p
    const int2 *tiles = nullptr;       // (first cell, count) of 16-cell tiles, grouped by order
    const uint8_t *tile_p = nullptr;   // order of each tile
    int ntiles = 0;
};

// Host builders of the level-independent tables (see fmm32.hip).
void mixed_tables(int pmax, std::vector<float> &m2m, std::vector<float> &l2l,
                  std::vector<float> &tm2l, std::vector<std::vector<float>> &rlane);

// Level table and M2L tiles on the device from the device-resident orders
// (levels 0..17) and the built tree; returns the number of tiles, which
// mixed_tile_count gives beforehand for sizing tiles and tile_p.
int mixed_tile_count(const Ctx &c);
int mixed_level_setup(Ctx &c, const uint8_t *orders, LevelTab *lt, int2 *tiles,
                      uint8_t *tile_p);

// The mixed FMM of the tree in c: fmm32_prepare (host: sizes the expansion
// and particle buffers), fmm32_upward (P2M, M2M; on any stream, it needs the
// tree only), fmm32_rest (M2L, list 4, L2L, evaluation on c.stream; needs
// the lists and the upward pass, and, for list4_mode 2, fmm32_p2l).
// fmm32_run does all of them on c.stream.
void fmm32_prepare(Ctx &c, const Mixed32 &mx);
void fmm32_upward(Ctx &c, const Mixed32 &mx, hipStream_t s, const FmmMarks &mk);
// its two parts: P2M (one launch), then M2M up the tree
void fmm32_p2m(Ctx &c, const Mixed32 &mx, hipStream_t s);
void fmm32_m2m(Ctx &c, const Mixed32 &mx, hipStream_t s, const FmmMarks &mk);
// List 4 by P2L (list4_mode 2) into c.b4: after the lists and P2M (it needs
// the particle offsets), before fmm32_rest, whose M2L adds it.
void fmm32_p2l(Ctx &c, const Mixed32 &mx, hipStream_t s);
void fmm32_rest(Ctx &c, const Mixed32 &mx, int list4_mode, int stage_mask, double *out_a,
                double *out_b, bool velocity, double u_inf, double v_inf, const FmmMarks &mk);
void fmm32_run(Ctx &c, const Mixed32 &mx, int list4_mode, int stage_mask, double *out_a,
               double *out_b, bool velocity, double u_inf, double v_inf, const FmmMarks &mk);

}  // namespace vpm
