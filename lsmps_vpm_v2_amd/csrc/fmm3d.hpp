// fmm3d.hpp -- the 3-D vector treecode (reference VELOCITY_3D_TYPE 1:
// VelocityCalc::velocity_fmm_3d_fast -> fmm3D::calcVelocityFast,
// src/FMM/fast3DFMM.cpp) on gfx950, FP64.
//
// Layout in HBM:
//   particles, octree-key sorted:  Par3 {x, y, z, alpha x/y/z, active, pad}
//                                  (64 B), key (u64: 3 bits per level, level 1
//                                  in bits 57..59), perm (u32 sorted -> caller)
//   cells (levels >= 1, the root is implicit), depth-first pre-order:
//                                  Cell3 {centre, (support size)^2, particle
//                                  range, skip = next cell after the subtree,
//                                  level | leaf | active} (48 B)
//   moments: 30 FP64 per cell (the 10 order-2 Cartesian moments of each
//            vorticity component), stride 32
#pragma once

#include "ctx.hpp"

namespace vpm {

// Deepest level the particle keys resolve (two 30-bit words of 10 levels).
constexpr int kFmm3Depth = 20;

struct Fmm3Params {
    int crit_num = 30;       // a cell with fewer particles is a leaf
    int max_depth = kFmm3Depth;
    double exp_tree = 1.0;   // root expansion in percent
    double center[3] = {0.0, 0.0, 0.0};  // a point the root box always holds
    double support = 2.0;    // a cell is far when |r|^2 > (support size)^2
    double u_inf = 0.0, v_inf = 0.0, w_inf = 0.0;
    int type = 1;            // 1: fast3DFMM (FMMCell); 2: fmm3D::calcVelocity (TreeCell)
    int max_par = 50, max_src = 30;  // type 2 leaf rule: at most these many (active) particles
};

// Device header of a 3-D solve, read back by the host once per build.
struct Fmm3Hdr {
    double root_c[3];
    double root_len;
    unsigned flags;
    uint32_t n_cells;   // total of the cell-emission scan
    uint32_t n_active;  // total of the active scan
    uint32_t depth;     // deepest leaf level
    uint32_t n_leaves;
    uint32_t min_settle;  // shallowest leaf level
    double pivot[3];      // type 2: the root box's lower corner
    unsigned long long pairs, m2p;  // evaluation work: direct pairs, far-cell terms
    // wavefront steps: sources staged for near lanes, far cells with a far lane
    unsigned long long wave_pairs, wave_m2p;
};

struct Fmm3Info {
    int n = 0, n_active = 0, n_cells = 0, n_leaves = 0, depth = 0;
    double root_c[3] = {0, 0, 0}, root_len = 0;
    unsigned long long pairs = 0, m2p = 0, wave_pairs = 0, wave_m2p = 0;
    // stage times (HIP events; -1 when the solve ran without them): the
    // octree with the compacted sources, the moments, the evaluation kernel
    double ms_tree = -1, ms_moments = -1, ms_eval = -1;
    int sorted_levels = 0;  // levels the particle sort resolved (10 or 20)
};

struct Ctx3 {
    Fmm3Hdr *hdr_d = nullptr;
    Fmm3Hdr *hdr_h = nullptr;  // pinned mirror
    DBuf hdr_buf, bbox, khi, klo, key_a, key_b, val_a, val_b, tmp;
    DBuf par, src, key64, act8, act_scan, lcpw, settle, mprev, starts, pos, cells, mom;
    DBuf bstat, wstat;
    uint32_t *perm = nullptr;  // sorted slot -> caller index (val_a or val_b)
    int levels_guess = 10;     // the previous build's sort depth
    Fmm3Info info;
    Fmm3Params last;
    ~Ctx3();
};

// u, v, w (device) of every particle: the treecode's rotational velocity from
// the vorticity (wx, wy, wz) and size s (alpha = omega s^3 / (4 pi)), plus the
// free stream.  With s null, (wx, wy, wz) are the strengths alpha.  Throws
// Error on bad input (non-finite coordinates, a tree deeper than max_depth).
// ev (4 events, may be null) times the stages.
void fmm3d_velocity(Ctx3 &c, hipStream_t s, int n, const double *x, const double *y,
                    const double *z, const double *sz, const double *wx, const double *wy,
                    const double *wz, const uint8_t *active, const Fmm3Params &p, double *u,
                    double *v, double *w, hipEvent_t *ev = nullptr);

// The cells of the last solve (host arrays of n_cells entries, any may be
// null): level, path key (octant x + 2y + 4z per level, level 1 highest),
// particle count, leaf, active, centre (3 per cell), size.
void fmm3d_cells(Ctx3 &c, hipStream_t s, int32_t *level, int64_t *key, int32_t *count,
                 uint8_t *leaf, uint8_t *active, double *centre3, double *size);

// Sorted slot -> caller index of the last solve (host array of n entries).
void fmm3d_perm(Ctx3 &c, hipStream_t s, int32_t *perm);

}  // namespace vpm
