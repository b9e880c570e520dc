// ctx.hpp -- solver context: device buffers and the stage entry points.
//
// Layout in HBM (all arrays are structure-of-arrays):
//   particles, Morton-sorted:  xs, ys, qs (f64), acts (u8), perm (u32: sorted
//                              slot -> caller index), key (u32 Morton key)
//   cells, level-major and Morton-ordered inside a level (root at index 0):
//                              level (u8), flags (u8: leaf, active), key (u32
//                              level prefix), begin/end (u32 particle range),
//                              src (u32 active count), cx, cy (f64 centre),
//                              parent (i32), child[4] (i32), nbr[9] (i32)
//   lists:  M2L sources per cell, fixed stride (list 2: 27, list 4: 8);
//           per leaf CSR of list-1 cells and list-3 cells
//   expansions: ak, bk as (P+1) complex per cell, cell-major
#pragma once

#include <vector>

#include "common.hpp"

namespace vpm {

// Leaf-split statistics, accumulated by atomics into kSlotRows rows of one
// cache line each (row = block mod kSlotRows) in the header and summed on the
// host at the build's synchronisation: cells per level 0..16, leaves,
// then 31 - min leaf level, max leaf level and the depth-overflow flag.
constexpr int kLvCols = kMaxDepth + 2;
constexpr int kSlotRows = 16;
constexpr int kSlotStride = 32;
constexpr int kSlotMinInv = kLvCols, kSlotMax = kLvCols + 1, kSlotDepth = kLvCols + 2;

// List statistics, accumulated by per-block atomics into kStatRows rows (row =
// block mod kStatRows) and summed on the host: list-2 / list-4 entries, P2P
// (target, source) pairs, (target, list-3 cell) evaluations, and the M2P /
// L2P terms (sum of the order + 1 of each list-3 cell per target; the order
// of each particle's leaf).
enum StatCol {
    kStSum2, kStSum4, kStPairs, kStM2p, kStM2pTerms, kStL2pTerms,
    kStInactive,  // particles with active = 0 (the gather counts them)
    kStCols = 8
};
constexpr int kStatRows = 16;

// Header written by device kernels and read back by the host.
struct DevHdr {
    double root_len;
    double pivot[2];
    double mn[2], mx[2];
    unsigned flags;
    uint32_t n_active;
    uint32_t n_cells_pre;        // cells excluding the root
    uint32_t n_leaves;
    uint32_t lev_off[kMaxDepth + 2];
    uint32_t tot1, tot3;
    uint32_t leaf_lvl_min, leaf_lvl_max;
    uint8_t sel_p[kMaxDepth + 2];  // selected order per level
    uint8_t sel_pad[2];
    uint32_t sel_source;           // 0 fixed, 1 learned, 2 fallback to the fixed order
    float sel_pred_err, sel_pred_ms;
    uint32_t lev_cnt[kMaxDepth + 2];  // cells per level, counted by the leaf split
    uint32_t n_leaves_split;          // leaves, counted by the leaf split
    uint32_t pad[3];
    unsigned long long stat[kStatRows][kStCols];
    uint32_t split_rows[kSlotRows][kSlotStride];
};

// What the build's synchronisation derives from the split rows, handed to the
// device by value (the first kernel after it writes it into the header).
struct SplitLevels {
    uint32_t lev_off[kMaxDepth + 2];
    uint32_t lev_cnt[kMaxDepth + 2];
    uint32_t leaf_lvl_min, leaf_lvl_max;
};

// Host sum of one statistic column.
inline unsigned long long stat_total(const DevHdr &h, int col) {
    unsigned long long s = 0;
    for (int r = 0; r < kStatRows; ++r) s += h.stat[r][col];
    return s;
}


// Grow-only device buffer.  While a scratch arena is active on the calling
// thread (the no-allocation entry point), get() carves from the arena instead
// of allocating, and the buffer's own allocation (heap_p, heap_cap) is put
// back when the arena closes.
struct DBuf {
    void *p = nullptr;
    size_t cap = 0;
    bool in_arena = false;
    void *heap_p = nullptr;
    size_t heap_cap = 0;
    void *get(size_t bytes);
    template <class T>
    T *as(size_t count) {
        return static_cast<T *>(get(count * sizeof(T) + 16));
    }
    void release();
    ~DBuf() { release(); }
};

// Tree-build parameters (reference knobs plus the level extensions).  With
// type1 the root and the particle keys are those of the type-1 box hierarchy
// (hierarchy_mesh, subroutine_all_fmm_2d.cpp:97-287): exp_tree is unused and
// max_level is its lmax.
struct TreeParams {
    int max_par, max_src;
    double exp_tree;
    int min_level, max_level;
    bool type1 = false;
};

// Host copy of what the tree build produced.
struct TreeInfo {
    int n = 0;
    int n_active = 0;
    int n_cells = 0;         // including the root
    int n_leaves = 0;
    int max_level = 0;       // deepest leaf level
    int min_leaf_level = 0;  // shallowest leaf level
    uint32_t lev_off[kMaxDepth + 2] = {};
    int tot1 = 0, tot3 = 0;
    size_t cap1 = 0, cap3 = 0;  // capacities of the list-1 / list-3 arrays
    double root_len = 0, pivot[2] = {0, 0};
};

struct Ctx {
    hipStream_t stream = nullptr;
    hipEvent_t ev_sync = nullptr;  // completion marker of the per-solve synchronisations
    // second stream (default priority) for the upward pass, which runs
    // concurrently with the list build; fork / join events
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_p2m = nullptr;
    DevHdr *hdr_d = nullptr;  // device header
    DevHdr *hdr_h = nullptr;  // pinned host mirror

    // particles
    DBuf x_in, y_in, q_in, a_in, ex_out, ey_out;  // staging for host-pointer calls
    DBuf key_a, key_b, val_a, val_b;
    DBuf xs, ys, qs, acts;
    // type-1 tree (TreeParams::type1): box bounds per cell as the reference
    // rounds them (xmin, xmax, ymin, ymax); the list kernels then apply the
    // reference's box tests to them
    bool t1 = false;
    DBuf c_bnd;
    // near-field kernel of this call (0 singular, 1..3 the regularized
    // kernels of regul_func_2d) and, when regularized, the sorted sizes
    int icut = 0;
    DBuf ss;
    uint32_t *key_sorted = nullptr;   // points into key_a or key_b after the sort
    uint32_t *perm_sorted = nullptr;  // sorted slot -> caller index (val_a or val_b)
    DBuf lcp_d, lcp_w, act_scan, act_pos, leaf_lvl, lv_blk, lv_tot, leaf_of;
    DBuf tmp;  // scan / sort scratch
    DBuf bbox_part;
    DBuf bs_act, bs_l13;  // producer block sums of scan inputs
    // key bits the particle sort covers (top bits; levels <= sort_bits / 2 are
    // resolved): the previous build's choice as the first guess of the next
    int sort_bits = 32;
    size_t cell_cap = 0;  // cells the emission has room for before the count is known

    // cells
    DBuf lev_level, lev_begin;
    DBuf c_level, c_flags, c_key, c_begin, c_end, c_src, c_cx, c_cy, c_parent, c_child, c_nbr;
    DBuf leaf_idx;
    DBuf l2, n2, l4, n4;
    DBuf cnt1, cnt3, cntk, off1, off3, l1, l3;

    // expansions
    DBuf ak, bk, tab_m2l, tab_m2m, tab_l2l;
    // FP64 list-2 M2L matrices for the matrix cores (fmm64_mfma_table) and
    // the order they were built for (-1: none)
    DBuf tab_r64;
    int tab_r64_p = -1;
    // type-1 tables: M2M (order ndp - 1, as tab_m2m), M2L (ndp + 1) x ndp,
    // L2L coefficients and exponents (ndp + 1)^2, and the all-active flags
    // of a type-1 call
    DBuf t1_m2m, t1_m2l, t1_l2l_c, t1_l2l_e, t1_ones;

    // mixed path: FP32 particle offsets and sources, tables, tiles
    DBuf oxy, qf;
    DBuf b4;  // list-4 P2L contributions (mixed path, list4_mode 2)
    DBuf sel_part, sel_feats, sel_raw;  // level statistics of the selection
    // expansion order per level (levels 0..17), device resident: slot 0 holds
    // the learned selection (written by the selection kernel), slot 1 the
    // fixed orders (uploaded when they change); `orders` points at the slot
    // the list build and the mixed path use
    DBuf orders_buf;
    const uint8_t *orders = nullptr;
    DBuf mx_lt, mx_m2m, mx_l2l, mx_tm2l, mx_rdata, mx_tiles, mx_tile_p;

    TreeInfo info;
    std::vector<double> stage_ms;
};

// Precomputed coefficient tables (host-side builders).
void fill_tables(int p, int binom_mode, std::vector<double> &m2l, std::vector<double> &m2m,
                 std::vector<double> &l2l);
int ref_binom_int32(int n, int k);
double exact_binom(int n, int k);

// Source strengths of a build: q per particle, or (q == nullptr) the particle
// size s and vorticity omega, from which the build forms q = -omega s^2 / (2 pi).
struct Sources {
    const double *q = nullptr;
    const double *s = nullptr;
    const double *omega = nullptr;
    double qscale = 1.0;  // the build forms q[i] * qscale from q
};

// Stage entry points (tree.hip).  Inputs are device pointers.
void tree_build(Ctx &c, int n, const double *x, const double *y, const Sources &src,
                const uint8_t *active, const TreeParams &tp);
// Waits until the work queued on c.stream so far has completed by polling an
// event (the host returns as soon as the GPU reaches it, without the
// scheduler wake-up of a blocking synchronisation).
void stream_wait(Ctx &c);
// The same in two steps: mark the point on the stream, queue more work, then
// wait for the mark.
void stream_mark(Ctx &c);
void stream_mark_wait(Ctx &c);

// Lists of the tree in c, with no host synchronisation: sizes come from the
// counts of tree_build and the list arrays are sized by proven bounds.
void tree_lists(Ctx &c);
// The same in two parts (neighbour table and lists 2 / 4; lists 1 / 3), so
// that other work can be enqueued between them.
void tree_lists_begin(Ctx &c);
void tree_lists_end(Ctx &c);
// Host copy of the list totals and the checks that need the list build's
// header (call after the header has been read back).
void tree_lists_absorb(Ctx &c);

// Optional stage markers: events 3..6 are recorded after the upward pass,
// M2L, the downward pass and the evaluation.
struct FmmMarks {
    hipEvent_t *ev = nullptr;
    void mark(int i, hipStream_t s) const {
        if (ev) VPM_HIP(hipEventRecord(ev[i], s));
    }
};

// FP64 FMM (fmm64.hip): expansions and evaluation.  Output either the field
// (ex, ey) or the velocity (u = ey + u_inf, v = -ex + v_inf), scattered to the
// caller's particle order; with `potential`, the potential phi = sum q log r
// into out_a (fmm2D::calcPotential: the local coefficient b_0 is formed and
// carried down, out_b and velocity are unused).
void fmm64_run(Ctx &c, int p, int list4_mode, int stage_mask, double *out_a, double *out_b,
               bool velocity, double u_inf, double v_inf, const FmmMarks &mk,
               bool potential = false);

// The FP64 path's list-2 M2L matrices for the matrix cores, order P <= 16,
// from its M2L table tab ((P+1) x (P+1)): 40 offsets x 2P x 2(P+1) in
// v_mfma_f64_16x16x4f64 A-operand lane order.
void fmm64_mfma_table(int P, const double *tab, std::vector<double> &R);

// Type-1 FMM (fmm64.hip, biotsavart_fmm_2d.cpp:93-446) on a type-1 tree:
// multipoles of orders 0..ndp-1 (P2M, M2M), locals of orders 0..ndp from
// M2L (list 2) and P2L (list 4), the L2L shift of the tables (coefficient,
// exponent) in t1_l2l_c / t1_l2l_e, evaluation with the near-field kernel
// c.icut; u = Ey + u_inf, v = -Ex + v_inf.
void fmm_t1_run(Ctx &c, int ndp, double *u, double *v, double u_inf, double v_inf,
                const FmmMarks &mk);

// Particle advection sub-step (advect.hip): Euler x += u dt, or the
// reference's velocity-scaling "rk2" update.
constexpr int kAdvectEuler = 0, kAdvectRk2 = 1;
void advect(int n, double *x, double *y, const double *u, const double *v, double dt, int scheme,
            hipStream_t s);

// a[i] = va, b[i] = vb for i < n.
void fill_pair(double *a, double *b, int n, double va, double vb, hipStream_t s);

// FP64 direct sum on the GPU (accuracy reference, not part of the FMM).
void direct64(int n, const double *x, const double *y, const double *q, const uint8_t *src_mask,
              int nt, const int *tidx, double *ex, double *ey, hipStream_t s);

}  // namespace vpm
