// api.cpp -- C ABI (include/vpmfmm.h): handle, option checks, stage
// orchestration, host staging and tree inspection.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vpmfmm.h"
#include "ctx.hpp"
#include "fmm32.hpp"
#include "fmm3d.hpp"
#include "select.hpp"

namespace vpm {

// Buffers replaced by larger ones during a call.  hipFree waits for the
// device, so freeing them at once would stall the launches already queued;
// they are freed when the call returns (free_retired, from guarded()).
static thread_local std::vector<void *> g_retired;

static void free_retired() {
    for (void *q : g_retired) (void)hipFree(q);
    g_retired.clear();
}

// Bump allocator over the device scratch array of the no-allocation entry
// point.  Each buffer asked for during the call takes a fresh 256-byte
// aligned block (a buffer that grows takes another); nothing is freed until
// the arena closes, when every buffer gets its own allocation back.
struct Arena {
    char *base = nullptr;
    size_t size = 0, used = 0;
    std::vector<DBuf *> touched;
};
static thread_local Arena *g_arena = nullptr;

// The scratch array of the no-allocation entry point.  Host code never uses
// its name as an address: each device holds its own copy, whose address is
// resolved once per device (scratch_address) and kept.
static __device__ __align__(256) unsigned char scratch[VPMFMM_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on device `dev`, which must be current.
static void *scratch_address(int dev) {
    if (dev < 0 || dev >= kMaxDevices) throw Error(kBadArg, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        VPM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw Error(kHipError, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

void *DBuf::get(size_t bytes) {
    if (g_arena) {
        Arena &a = *g_arena;
        if (in_arena && bytes <= cap) return p;
        const size_t nb = (bytes + 255) & ~(size_t)255;
        if (nb > a.size - a.used)
            throw Error(kBadArg, "the no-allocation scratch (" + std::to_string(a.size >> 20) +
                                     " MiB) is too small for this tree: " +
                                     std::to_string((a.used + nb) >> 20) +
                                     " MiB asked; use vpmfmm_velocity_device");
        if (!in_arena) {
            heap_p = p;
            heap_cap = cap;
            in_arena = true;
            a.touched.push_back(this);
        }
        p = a.base + a.used;
        cap = nb;
        a.used += nb;
        return p;
    }
    if (bytes <= cap && p) return p;
    // headroom, so that sizes that creep up from call to call (a moving
    // particle field) regrow rarely
    size_t nb = std::max(bytes + bytes / 2, cap + cap / 2);
    nb = (nb + 255) & ~(size_t)255;
    if (p) g_retired.push_back(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc(&p, nb) != hipSuccess) {
        p = nullptr;
        cap = 0;
        (void)hipGetLastError();
        throw Error(kNoMem, "device allocation of " + std::to_string(nb) + " bytes failed");
    }
    cap = nb;
    return p;
}

void DBuf::release() {
    void *own = in_arena ? heap_p : p;  // never an address inside the scratch array
    if (own) (void)hipFree(own);
    p = heap_p = nullptr;
    cap = heap_cap = 0;
    in_arena = false;
}

// fmm2D::binComb (reference src/FMM/fmm2D.cpp:28-58): product n..max+1, then
// division by min..2, in 32-bit int.  The product wraps for the M2L arguments
// at P >= 10 (and for M2M/L2L from P = 18); g++ on x86-64 wraps two's
// complement, emulated here with unsigned arithmetic.
int ref_binom_int32(int n, int k) {
    if (k > n) return 0;
    const int inv = n - k;
    const int mx = inv > k ? inv : k, mn = inv > k ? k : inv;
    uint32_t c = 1u;
    for (int i = n; i > mx; --i) c *= (uint32_t)i;
    int32_t cs = (int32_t)c;
    for (int i = mn; i > 1; --i) cs /= i;
    return cs;
}

// Exact in 128-bit integers (C(n-k+i, i) is an integer after every step),
// rounded once to double.
double exact_binom(int n, int k) {
    if (k < 0 || k > n) return 0.0;
    if (k > n - k) k = n - k;
    unsigned __int128 c = 1;
    for (int i = 1; i <= k; ++i) c = c * (unsigned)(n - k + i) / (unsigned)i;
    return (double)c;
}

// Row-major (P+1) x (P+1) tables:
//   m2l[p][0] = -1/p, m2l[p][k] = -C(p+k-1, k-1) (-1)^k / k   (fmm2D.cpp:465-481)
//   m2m[p][k] = C(p-1, k-1) p / k                             (fmm2D.cpp:371-374)
//   l2l[p][k] = C(k, p)                                       (fmm2D.cpp:639-641)
void fill_tables(int P, int binom_mode, std::vector<double> &m2l, std::vector<double> &m2m,
                 std::vector<double> &l2l) {
    const int P1 = P + 1;
    auto B = [&](int n, int k) {
        return binom_mode ? (double)ref_binom_int32(n, k) : exact_binom(n, k);
    };
    m2l.assign((size_t)P1 * P1, 0.0);
    m2m.assign((size_t)P1 * P1, 0.0);
    l2l.assign((size_t)P1 * P1, 0.0);
    for (int p = 1; p <= P; ++p) {
        m2l[p * P1 + 0] = -1.0 / (double)p;
        for (int k = 1; k <= P; ++k) {
            const double sg = (k & 1) ? -1.0 : 1.0;
            m2l[p * P1 + k] = -(B(p + k - 1, k - 1) * sg / (double)k);
        }
        for (int k = 1; k <= p; ++k) m2m[p * P1 + k] = B(p - 1, k - 1) * ((double)p / (double)k);
    }
    for (int p = 0; p <= P; ++p)
        for (int k = p; k <= P; ++k) l2l[p * P1 + k] = B(k, p);
}

double rel_diff64(int n, const double *a, const double *b, const double *ja, const double *jb,
                  double a0, double b0, double *part_dev, hipStream_t s);

}  // namespace vpm

using namespace vpm;

struct vpmfmm_handle_s {
    vpmfmm_opts o{};
    Ctx ctx;
    hipStream_t own_stream = nullptr;
    std::string err;
    vpmfmm_stats st{};
    int tab_p = -1, tab_binom = -1;
    int exp_p = -1;  // order of the expansions held on the device (-1: none)
    std::vector<int> level_p;  // per-level orders of the mixed path (empty: opts.p everywhere)
    int mx_pmax = -1;          // order the mixed tables were built for
    const float *mx_rptr_host[17] = {};  // device addresses of the M2L tables per order
    // fixed orders per level as last uploaded to the device slot (valid once
    // fix_valid), so they are copied only when they change
    uint8_t fix_orders[kMaxDepth + 2] = {};
    bool fix_valid = false;
    hipEvent_t ev[8] = {};
    // learned selection
    bool has_sel = false;
    vpmfmm_selector sel{};
    DBuf sel_model_d, sel_params_d;
    int sel_threshold = 0;   // leaf threshold chosen by the last re-selection (0: none)
    int sel_n = 0;           // particle count at the last re-selection
    int sel_calls = 0;       // calls since the last re-selection
    int judge_calls = 0;     // learned calls since the last re-selection (judge cadence)
    bool judge_reject = false;  // the judge rejected the learned orders: fixed order until re-selection
    // orders chosen by the last selection, reused while the tree depth is unchanged
    bool sel_cached = false;
    bool sel_fresh = false;  // this call selected its orders (judged by default)
    int sel_cached_maxl = -1;
    DevHdr sel_cache{};  // sel_p, sel_source, sel_pred_* of that selection
    DBuf judge_a, judge_b, judge_part;
    vpmfmm_selection last_sel{};
    bool judged_reject_now = false;  // the judge of this call rejected the learned orders
    bool concurrent_marks = false;   // the last run built the lists beside the upward pass
    MlpWork mlp;
    DBuf mlp_grad;
    void *scratch = nullptr;  // this device's no-allocation scratch array
    // inspection cache of the last tree (host copies)
    bool insp = false;
    std::vector<uint8_t> h_level, h_flags;
    std::vector<uint32_t> h_key, h_begin, h_end, h_src, h_perm, h_leaf_idx, h_off1, h_off3;
    std::vector<int> h_l1, h_l3, h_l2, h_l4;
    std::vector<uint8_t> h_n2, h_n4;
    std::vector<int64_t> h_id;
    // type-1 tables on the device (ndp, l2l_mode they were built for) and the
    // per-level orders of a type-1 tree (list statistics)
    int t1_nd = -1, t1_mode = -1;
    uint8_t t1_orders[kMaxDepth + 2] = {};
    // the 3-D treecode's buffers (made by its first call)
    std::unique_ptr<Ctx3> c3;
};

namespace {

constexpr int kMaxP = 40;
constexpr float kJudgeShare = 0.9f;  // share of the target the judged difference may take
constexpr int kMaxThreshold = 4096;

int opts_check(const vpmfmm_opts &o, std::string &msg) {
    if (o.p < 1 || o.p > kMaxP) msg = "p must be in [1, 40]";
    else if (o.max_par < 1 || o.max_par > kMaxThreshold) msg = "max_par must be in [1, 4096]";
    else if (o.max_src < 0 || o.max_src > kMaxThreshold) msg = "max_src must be in [0, 4096]";
    else if (!(o.exp_tree > 0.0) || !std::isfinite(o.exp_tree)) msg = "exp_tree must be > 0";
    else if (o.min_level < 1 || o.min_level > kMaxDepth) msg = "min_level must be in [1, 16]";
    else if (o.max_level < 0 || o.max_level > kMaxDepth) msg = "max_level must be in [0, 16]";
    else if (o.max_level && o.max_level < o.min_level) msg = "max_level < min_level";
    else if (o.precision != VPMFMM_FP64 && o.precision != VPMFMM_MIXED) msg = "bad precision";
    else if (o.binom_mode != 0 && o.binom_mode != 1) msg = "binom_mode must be 0 or 1";
    else if (o.list4_mode != 1 && o.list4_mode != 2) msg = "list4_mode must be 1 or 2";
    else if (o.stage_mask < 0 || o.stage_mask > 7) msg = "stage_mask must be in [0, 7]";
    else if (o.order_select != VPMFMM_ORDER_FIXED && o.order_select != VPMFMM_ORDER_LEARNED)
        msg = "order_select must be 0 (fixed) or 1 (learned)";
    else if (o.reselect_every < 0) msg = "reselect_every must be >= 0";
    else if (o.judge_every < -1) msg = "judge_every must be >= -1";
    else if (o.icutoff < 0 || o.icutoff > 3) msg = "icutoff must be in [0, 3]";
    else return VPMFMM_OK;
    return VPMFMM_BAD_ARG;
}

template <class F>
int guarded(vpmfmm_handle *h, F &&f) {
    if (!h) return VPMFMM_BAD_ARG;
    int st = VPMFMM_OK;
    try {
        h->err.clear();
        f();
    } catch (const Error &e) {
        h->err = e.what();
        st = e.code;
    } catch (const std::exception &e) {
        h->err = e.what();
        st = VPMFMM_INTERNAL;
    }
    free_retired();
    return st;
}

void ensure_tables(vpmfmm_handle *h) {
    const int P = h->o.p;
    if (h->tab_p == P && h->tab_binom == h->o.binom_mode) return;
    std::vector<double> m2l, m2m, l2l;
    fill_tables(P, h->o.binom_mode, m2l, m2m, l2l);
    Ctx &c = h->ctx;
    const size_t nb = m2l.size() * sizeof(double);
    VPM_HIP(hipMemcpyAsync(c.tab_m2l.get(nb), m2l.data(), nb, hipMemcpyHostToDevice, c.stream));
    VPM_HIP(hipMemcpyAsync(c.tab_m2m.get(nb), m2m.data(), nb, hipMemcpyHostToDevice, c.stream));
    VPM_HIP(hipMemcpyAsync(c.tab_l2l.get(nb), l2l.data(), nb, hipMemcpyHostToDevice, c.stream));
    // the list-2 M2L matrices for the matrix cores (orders up to 16)
    std::vector<double> r64;
    c.tab_r64_p = -1;
    if (P >= 1 && P <= 16) {
        fmm64_mfma_table(P, m2l.data(), r64);
        const size_t nr = r64.size() * sizeof(double);
        VPM_HIP(hipMemcpyAsync(c.tab_r64.get(nr), r64.data(), nr, hipMemcpyHostToDevice,
                               c.stream));
        c.tab_r64_p = P;
    }
    VPM_HIP(hipStreamSynchronize(c.stream));
    h->tab_p = P;
    h->tab_binom = h->o.binom_mode;
}

void mark(vpmfmm_handle *h, int i) {
    if (h->o.timing) VPM_HIP(hipEventRecord(h->ev[i], h->ctx.stream));
}

float span(vpmfmm_handle *h, int a, int b) {
    float ms = 0.f;
    VPM_HIP(hipEventElapsedTime(&ms, h->ev[a], h->ev[b]));
    return ms;
}

bool learned(const vpmfmm_handle *h) {
    return h->o.precision == VPMFMM_MIXED && h->o.order_select == VPMFMM_ORDER_LEARNED &&
           h->has_sel;
}

TreeParams tree_params(const vpmfmm_handle *h, int threshold) {
    TreeParams tp{h->o.max_par, h->o.max_src, h->o.exp_tree, h->o.min_level, h->o.max_level};
    if (threshold > 0) tp.max_par = tp.max_src = threshold;
    return tp;
}

// Device slots of the per-level orders: 0 the learned selection, 1 the fixed
// orders of the options.
uint8_t *learned_slot(vpmfmm_handle *h) { return h->ctx.orders_buf.as<uint8_t>(64); }

// Fixed orders per level (opts.p, or the per-level orders of the mixed path),
// uploaded to slot 1 when they differ from the last upload.
const uint8_t *fixed_slot(vpmfmm_handle *h) {
    uint8_t o[kMaxDepth + 2];
    for (int L = 0; L < kMaxDepth + 2; ++L) {
        int p = h->o.p;
        if (h->o.precision == VPMFMM_MIXED && !learned(h) && L < (int)h->level_p.size() &&
            h->level_p[L] > 0)
            p = h->level_p[L];
        o[L] = (uint8_t)std::min(p, 255);
    }
    uint8_t *slot = h->ctx.orders_buf.as<uint8_t>(64) + 32;
    if (!h->fix_valid || std::memcmp(o, h->fix_orders, sizeof(o)) != 0) {
        std::memcpy(h->fix_orders, o, sizeof(o));
        VPM_HIP(hipMemcpyAsync(slot, h->fix_orders, sizeof(o), hipMemcpyHostToDevice,
                               h->ctx.stream));
        h->fix_valid = true;
    }
    return slot;
}

// Features of the tree just built and the selected orders, into the header
// (read back at the next header synchronisation).
void select_on_tree(vpmfmm_handle *h, int threshold) {
    Ctx &c = h->ctx;
    float *feats = c.sel_feats.as<float>((kMaxDepth + 1) * kSelFeatures);
    float *raw = c.sel_raw.as<float>((kMaxDepth + 1) * kSelRaw);
    level_features(c, threshold, feats, raw);
    select_orders(c, (const SelModel *)h->sel_model_d.p, (const float *)h->sel_params_d.p, feats,
                  h->o.p, learned_slot(h));
}

// Learned path: choose the leaf threshold among the model's candidates by the
// predicted solve time of each candidate tree (orders selected per tree).
void reselect_threshold(vpmfmm_handle *h, int n, const double *x, const double *y,
                        const Sources &q, const uint8_t *a) {
    Ctx &c = h->ctx;
    float best = INFINITY;
    int best_t = 0;
    for (int k = 0; k < h->sel.n_thresholds; ++k) {
        const int t = h->sel.thresholds[k];
        tree_build(c, n, x, y, q, a, tree_params(h, t));
        select_on_tree(h, t);
        VPM_HIP(hipMemcpyAsync(c.hdr_h, c.hdr_d, sizeof(DevHdr), hipMemcpyDeviceToHost, c.stream));
        VPM_HIP(hipStreamSynchronize(c.stream));
        if (c.hdr_h->sel_source == 1 && c.hdr_h->sel_pred_ms < best) {
            best = c.hdr_h->sel_pred_ms;
            best_t = t;
        }
    }
    h->sel_threshold = best_t;  // 0: no candidate met the target (fixed options)
    h->sel_n = n;
    h->sel_calls = 0;
    h->judge_calls = 0;
    h->judge_reject = false;
    h->sel_cached = false;
}

// Tree and (learned path) orders; the lists too unless the caller builds them
// itself (run() builds them concurrently with the upward pass).
void build(vpmfmm_handle *h, int n, const double *x, const double *y, const Sources &q,
           const uint8_t *a, bool lists = true) {
    std::string msg;
    if (opts_check(h->o, msg)) throw Error(kBadArg, msg);
    if (n < 1) throw Error(kBadArg, "n must be >= 1");
    if (!x || !y || !a) throw Error(kBadArg, "null pointer");
    VPM_HIP(hipSetDevice(h->o.device));
    h->insp = false;
    h->exp_p = -1;
    const bool lrn = learned(h);
    h->last_sel = vpmfmm_selection{};
    if (lrn) {
        const bool due = h->sel_n == 0 || std::abs(n - h->sel_n) * 10 > h->sel_n ||
                         (h->o.reselect_every > 0 && h->sel_calls >= h->o.reselect_every);
        if (due) {
            const auto t0 = std::chrono::steady_clock::now();
            reselect_threshold(h, n, x, y, q, a);
            h->last_sel.reselected = 1;
            h->last_sel.ms_reselect = (float)std::chrono::duration<double, std::milli>(
                                          std::chrono::steady_clock::now() - t0)
                                          .count();
        }
        ++h->sel_calls;
    }
    const int thr = lrn ? h->sel_threshold : 0;
    const TreeParams tp = tree_params(h, thr);
    mark(h, 0);
    tree_build(h->ctx, n, x, y, q, a, tp);
    // orders: a fresh selection after a threshold re-selection or when the
    // tree depth changed, else the orders of the last selection (still in
    // the learned slot); the fixed orders when not learning or after the
    // judge rejected the learned ones
    const bool reuse = lrn && h->sel_cached && h->sel_cached_maxl == h->ctx.info.max_level;
    if (lrn && !reuse) select_on_tree(h, tp.max_par);
    h->sel_fresh = lrn && !reuse;
    h->ctx.orders = (lrn && !h->judge_reject) ? learned_slot(h) : fixed_slot(h);
    mark(h, 1);
    h->concurrent_marks = false;
    if (lists) {
        tree_lists(h->ctx);
        mark(h, 2);
    }
    h->last_sel.threshold = tp.max_par;
}

// The mixed upward pass on the side stream while the list build runs on the
// main stream; the main stream waits for it before M2L.
void upward_with_lists(vpmfmm_handle *h, const Mixed32 &mx, const FmmMarks &mk) {
    Ctx &c = h->ctx;
    if (!c.side) VPM_HIP(hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking));
    if (!c.ev_fork) VPM_HIP(hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming));
    if (!c.ev_join) VPM_HIP(hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming));
    if (!c.ev_p2m) VPM_HIP(hipEventCreateWithFlags(&c.ev_p2m, hipEventDisableTiming));
    fmm32_prepare(c, mx);
    VPM_HIP(hipEventRecord(c.ev_fork, c.stream));
    VPM_HIP(hipStreamWaitEvent(c.side, c.ev_fork, 0));
    // launches interleaved so that both streams have work early, the upward
    // pass (the longer chain) first
    fmm32_p2m(c, mx, c.side);
    VPM_HIP(hipEventRecord(c.ev_p2m, c.side));
    tree_lists_begin(c);
    fmm32_m2m(c, mx, c.side, mk);
    VPM_HIP(hipEventRecord(c.ev_join, c.side));
    tree_lists_end(c);
    mark(h, 2);
    // list 4 by P2L on the main stream once P2M has formed the offsets, beside
    // the rest of the upward pass
    if (h->o.list4_mode == 2) {
        VPM_HIP(hipStreamWaitEvent(c.stream, c.ev_p2m, 0));
        fmm32_p2l(c, mx, c.stream);
    }
    VPM_HIP(hipStreamWaitEvent(c.stream, c.ev_join, 0));
    mark(h, 7);
    h->concurrent_marks = true;
}

void read_hdr(Ctx &c) {
    VPM_HIP(hipMemcpyAsync(c.hdr_h, c.hdr_d, sizeof(DevHdr), hipMemcpyDeviceToHost, c.stream));
    stream_wait(c);
}

// Final synchronisation of a call: list totals and checks, the selection
// cache of a fresh selection, statistics and stage times.
void finish_stats(vpmfmm_handle *h, bool fmm) {
    Ctx &c = h->ctx;
    read_hdr(c);
    tree_lists_absorb(c);
    const DevHdr &hh = *c.hdr_h;
    if (h->sel_fresh) {
        std::memcpy(h->sel_cache.sel_p, hh.sel_p, sizeof(hh.sel_p));
        h->sel_cache.sel_source = hh.sel_source;
        h->sel_cache.sel_pred_err = hh.sel_pred_err;
        h->sel_cache.sel_pred_ms = hh.sel_pred_ms;
        h->sel_cached = true;
        h->sel_cached_maxl = c.info.max_level;
    }
    const TreeInfo &ti = c.info;
    vpmfmm_stats &s = h->st;
    s = vpmfmm_stats{};
    s.n = ti.n;
    s.n_active = ti.n_active;
    s.n_cells = ti.n_cells;
    s.n_leaves = ti.n_leaves;
    s.min_leaf_level = ti.min_leaf_level;
    s.max_leaf_level = ti.max_level;
    s.n_list1 = ti.tot1;
    s.n_list3 = ti.tot3;
    const DevHdr &hs = *c.hdr_h;
    s.n_list2 = (int64_t)stat_total(hs, kStSum2);
    s.n_list4 = (int64_t)stat_total(hs, kStSum4);
    s.p2p_pairs = (int64_t)stat_total(hs, kStPairs);
    s.m2p_evals = (int64_t)stat_total(hs, kStM2p);
    s.m2p_terms = (int64_t)stat_total(hs, kStM2pTerms);
    s.l2p_terms = (int64_t)stat_total(hs, kStL2pTerms);
    if (h->o.timing) {
        s.ms_tree = span(h, 0, 1);
        s.ms_lists = span(h, 1, 2);
        if (fmm) {
            // concurrent list build and upward pass: both from the fork
            s.ms_upward = h->concurrent_marks ? span(h, 1, 3) : span(h, 2, 3);
            s.ms_m2l = h->concurrent_marks ? span(h, 7, 4) : span(h, 3, 4);
            s.ms_downward = span(h, 4, 5);
            s.ms_eval = span(h, 5, 6);
            s.ms_total = span(h, 0, 6);
        } else {
            s.ms_total = span(h, 0, 2);
        }
    }
    if (c.hdr_h->flags & kFlagDuplicate)
        throw Error(kDuplicate, "two particles of a list-1 pair share a position (r = 0)");
}

// Mixed path set-up: level-independent tables (rebuilt only when the largest
// order grows), then the per-level table and the 16-cell M2L tiles built on
// the device from the device-resident orders (no host round trip).
Mixed32 mixed_setup(vpmfmm_handle *h, bool force_fixed = false) {
    Ctx &c = h->ctx;
    const TreeInfo &ti = c.info;
    const bool lrn = learned(h) && !h->judge_reject && !force_fixed;
    const uint8_t *fix = fixed_slot(h);
    const uint8_t *ord = lrn ? learned_slot(h) : fix;
    // the fixed orders also serve the learned path (fallback and judge)
    int pmax = learned(h) ? h->sel.p_max : 0;
    for (int L = 0; L <= ti.max_level; ++L) pmax = std::max(pmax, (int)h->fix_orders[L]);
    if (pmax > 16) throw Error(kBadArg, "the mixed path supports expansion orders up to 16");
    if (h->mx_pmax < pmax) {
        std::vector<float> m2m, l2l, tm2l;
        std::vector<std::vector<float>> rl;
        mixed_tables(pmax, m2m, l2l, tm2l, rl);
        auto up = [&](DBuf &b, const std::vector<float> &v) {
            VPM_HIP(hipMemcpyAsync(b.get(v.size() * 4 + 16), v.data(), v.size() * 4,
                                   hipMemcpyHostToDevice, c.stream));
        };
        up(c.mx_m2m, m2m);
        up(c.mx_l2l, l2l);
        up(c.mx_tm2l, tm2l);
        size_t tot = 0;
        for (auto &r : rl) tot += r.size();
        float *rd = c.mx_rdata.as<float>(tot + 1);
        std::vector<const float *> ptrs(rl.size(), nullptr);
        size_t off = 0;
        for (size_t p = 0; p < rl.size(); ++p) {
            ptrs[p] = rd + off;
            if (!rl[p].empty())
                VPM_HIP(hipMemcpyAsync(rd + off, rl[p].data(), rl[p].size() * 4,
                                       hipMemcpyHostToDevice, c.stream));
            off += rl[p].size();
        }
        for (size_t p = 0; p < 17; ++p) h->mx_rptr_host[p] = p < ptrs.size() ? ptrs[p] : nullptr;
        VPM_HIP(hipStreamSynchronize(c.stream));
        h->mx_pmax = pmax;
    }
    const int ntiles = mixed_tile_count(c);
    int2 *td = c.mx_tiles.as<int2>((size_t)ntiles + 1);
    uint8_t *tpd = c.mx_tile_p.as<uint8_t>((size_t)ntiles + 1);
    LevelTab *ltd = c.mx_lt.as<LevelTab>(1);
    mixed_level_setup(c, ord, ltd, td, tpd);
    Mixed32 mx;
    mx.pmax = h->mx_pmax;  // tables (and coefficient stride) of the largest order built so far
    mx.lt = ltd;
    mx.m2m = (const float *)c.mx_m2m.p;
    mx.l2l = (const float *)c.mx_l2l.p;
    mx.tm2l = (const float *)c.mx_tm2l.p;
    for (int p = 0; p <= 16; ++p) mx.rptr_host[p] = h->mx_rptr_host[p];
    mx.tiles = td;
    mx.tile_p = tpd;
    mx.ntiles = ntiles;
    return mx;
}

// Selection report of the last call (after its final synchronisation).
void report_selection(vpmfmm_handle *h, bool lrn_run) {
    vpmfmm_selection &ls = h->last_sel;
    const DevHdr &sel = h->sel_fresh ? *h->ctx.hdr_h : h->sel_cache;
    if (h->o.precision != VPMFMM_MIXED) {
        ls.pred_err = ls.pred_ms = -1.f;
        for (int L = 0; L <= kMaxDepth; ++L) ls.p[L] = h->o.p;
        return;
    }
    if (learned(h)) {
        ls.source = (!lrn_run || h->judged_reject_now) ? 3 : (int)sel.sel_source;
        ls.pred_err = lrn_run ? sel.sel_pred_err : -1.f;
        ls.pred_ms = lrn_run ? sel.sel_pred_ms : -1.f;
    } else {
        ls.pred_err = ls.pred_ms = -1.f;
    }
    for (int L = 0; L <= kMaxDepth; ++L)
        ls.p[L] = (lrn_run && !h->judged_reject_now) ? sel.sel_p[L] : h->fix_orders[L];
}

void run(vpmfmm_handle *h, int n, const double *x, const double *y, const Sources &q,
         const uint8_t *a, double *oa, double *ob, bool velocity) {
    if (!(q.q || (q.s && q.omega)) || !oa || !ob) throw Error(kBadArg, "null pointer");
    if (h->o.icutoff > 0 && !q.s)
        throw Error(kBadArg, "the regularized near field (icutoff > 0) needs the particle "
                             "sizes: use a velocity entry point");
    h->ctx.icut = h->o.icutoff;
    const bool mixed = h->o.precision == VPMFMM_MIXED;
    build(h, n, x, y, q, a, !mixed);
    Ctx &c = h->ctx;
    FmmMarks mk{h->o.timing ? h->ev : nullptr};
    const int mask = h->o.stage_mask ? h->o.stage_mask : 7;
    vpmfmm_selection &ls = h->last_sel;
    ls.judge_err = -1.f;
    h->judged_reject_now = false;
    bool lrn = false;
    if (h->o.precision == VPMFMM_MIXED) {
        lrn = learned(h) && !h->judge_reject;
        Mixed32 mx = mixed_setup(h);
        upward_with_lists(h, mx, mk);
        fmm32_rest(c, mx, h->o.list4_mode, mask, oa, ob, velocity, h->o.u_inf, h->o.v_inf, mk);
        // the fixed-order judge (same tree): every fresh selection, and every
        // judge_every-th learned call
        const bool judge_due =
            h->o.judge_every >= 0 &&
            (h->sel_fresh || (h->o.judge_every > 0 && (h->judge_calls % h->o.judge_every) == 0));
        ++h->judge_calls;
        if (lrn && judge_due) {
            // the selection's source: read back for a fresh selection (this also
            // lets the learned solve finish before buffers may grow below)
            read_hdr(c);
            const uint32_t src = h->sel_fresh ? c.hdr_h->sel_source : h->sel_cache.sel_source;
            if (src == 1) {
                const int nn = c.info.n;
                double *ja = h->judge_a.as<double>(nn), *jb = h->judge_b.as<double>(nn);
                Mixed32 mf = mixed_setup(h, true);
                fmm32_run(c, mf, h->o.list4_mode, mask, ja, jb, velocity, h->o.u_inf,
                          h->o.v_inf, FmmMarks{nullptr});
                const double a0 = velocity ? h->o.u_inf : 0.0, b0 = velocity ? h->o.v_inf : 0.0;
                ls.judge_err = (float)rel_diff64(nn, oa, ob, ja, jb, a0, b0,
                                                 h->judge_part.as<double>(1024), c.stream);
                // The learned error against the exact sum is at most the judged
                // difference plus the fixed order's own error; the judge keeps
                // kJudgeShare of the target for the first, leaving the rest for
                // the second (p = 12 measures 1.5e-6, 0.015 of a 1e-4 target,
                // on the 780k flat plate)
                if (!(ls.judge_err <= kJudgeShare * h->sel.target_err)) {
                    h->judge_reject = true;
                    h->judged_reject_now = true;
                    VPM_HIP(hipMemcpyAsync(oa, ja, sizeof(double) * nn, hipMemcpyDeviceToDevice,
                                           c.stream));
                    VPM_HIP(hipMemcpyAsync(ob, jb, sizeof(double) * nn, hipMemcpyDeviceToDevice,
                                           c.stream));
                }
            }
        }
    } else {
        ensure_tables(h);
        fmm64_run(c, h->o.p, h->o.list4_mode, mask, oa, ob, velocity, h->o.u_inf, h->o.v_inf,
                  mk);
        h->exp_p = h->o.p;
    }
    finish_stats(h, true);
    report_selection(h, lrn);
}

// The no-allocation call: every buffer the call asks for is carved from the
// device's scratch array.  The handle's caches of device-resident data (the
// FP64 and mixed tables, the fixed-order slot, the learned orders) refer to
// its own buffers, so they are re-uploaded or re-selected inside the arena;
// afterwards the buffers and those caches are put back.  The learned orders
// are selected afresh by the next call (the cache then describes this call's
// tree), and the tree is not kept for inspection.
class ArenaScope {
public:
    explicit ArenaScope(vpmfmm_handle *h) : h_(h) {
        a_.base = static_cast<char *>(h->scratch);
        a_.size = VPMFMM_NOALLOC_SCRATCH_BYTES;
        a_.touched.reserve(96);
        tab_p_ = h->tab_p;
        tab_binom_ = h->tab_binom;
        tab_r64_p_ = h->ctx.tab_r64_p;
        mx_pmax_ = h->mx_pmax;
        std::memcpy(mx_rptr_, h->mx_rptr_host, sizeof(mx_rptr_));
        fix_valid_ = h->fix_valid;
        std::memcpy(fix_orders_, h->fix_orders, sizeof(fix_orders_));
        h->tab_p = h->tab_binom = -1;
        h->mx_pmax = -1;
        h->fix_valid = false;
        h->sel_cached = false;
        g_arena = &a_;
    }
    ~ArenaScope() {
        g_arena = nullptr;
        for (DBuf *b : a_.touched) {
            b->p = b->heap_p;
            b->cap = b->heap_cap;
            b->heap_p = nullptr;
            b->heap_cap = 0;
            b->in_arena = false;
        }
        h_->tab_p = tab_p_;
        h_->tab_binom = tab_binom_;
        h_->ctx.tab_r64_p = tab_r64_p_;  // the order of the table in the handle's buffer
        h_->mx_pmax = mx_pmax_;
        std::memcpy(h_->mx_rptr_host, mx_rptr_, sizeof(mx_rptr_));
        h_->fix_valid = fix_valid_;
        std::memcpy(h_->fix_orders, fix_orders_, sizeof(fix_orders_));
        h_->sel_cached = false;
        h_->insp = false;
        h_->exp_p = -1;
        h_->ctx.info.n_cells = 0;  // inspection: no tree in the handle's buffers
        h_->ctx.key_sorted = h_->ctx.perm_sorted = nullptr;
        h_->ctx.orders = nullptr;
    }
    size_t used() const { return a_.used; }

private:
    vpmfmm_handle *h_;
    Arena a_;
    int tab_p_, tab_binom_, tab_r64_p_, mx_pmax_;
    const float *mx_rptr_[17];
    bool fix_valid_;
    uint8_t fix_orders_[kMaxDepth + 2];
};

// The potential (fmm2D::calcPotential, fmm2D.cpp:681-896): the FP64 FMM with
// the handle's order, leaf rule, binomials and list-4 mode whatever its
// precision (the reference computes it in FP64 only), singular kernel only.
void run_potential(vpmfmm_handle *h, int n, const double *x, const double *y, const double *q,
                   const uint8_t *a, double *phi) {
    if (!q || !phi) throw Error(kBadArg, "null pointer");
    if (h->o.icutoff != 0)
        throw Error(kBadArg, "the potential has the singular kernel only (icutoff 0)");
    struct Restore {
        vpmfmm_handle *h;
        vpmfmm_opts o;
        ~Restore() { h->o = o; }
    } restore{h, h->o};
    h->o.precision = VPMFMM_FP64;
    h->ctx.icut = 0;
    build(h, n, x, y, Sources{q}, a, true);
    ensure_tables(h);
    const int mask = h->o.stage_mask ? h->o.stage_mask : 7;
    fmm64_run(h->ctx, h->o.p, h->o.list4_mode, mask, phi, nullptr, false, 0.0, 0.0,
              FmmMarks{h->o.timing ? h->ev : nullptr}, true);
    h->exp_p = h->o.p;
    finish_stats(h, true);
    report_selection(h, false);
}

// bico (subroutine_all_fmm_2d.cpp:560-591): nCk in double, formed as the
// reference forms it; k is rewritten to n - k where the reference rewrites it
// (its k is a reference parameter).
double t1_bico(int n, int &k) {
    if (k < 0 || k > n || n < 0) return 0.0;
    if (k == 0 || k == n || n == 0) return 1.0;
    if (k == 1 || k == n - 1) return (double)n;
    if (k > n / 2) k = n - k;
    const int n_k = n - k;
    double c = 1.0;
    for (int i = 1; i <= k; ++i) c = c * ((double)(n_k + i) / (double)i);
    return c;
}

int t1_opts_check(const vpmfmm_type1_opts &t, std::string &msg) {
    if (t.n_s < 1 || t.n_s > kMaxThreshold) msg = "n_s must be in [1, 4096]";
    else if (t.lmax < 2 || t.lmax > kMaxDepth) msg = "lmax must be in [2, 16]";
    else if (t.ndp < 1 || t.ndp > kMaxP) msg = "ndp must be in [1, 40]";
    else if (t.icutoff < 0 || t.icutoff > 3) msg = "icutoff must be in [0, 3]";
    else if (t.l2l_mode != 0 && t.l2l_mode != 1) msg = "l2l_mode must be 0 or 1";
    else return VPMFMM_OK;
    return VPMFMM_BAD_ARG;
}

// Type-1 tables: M2M at order ndp - 1 (exact binomials), M2L
// tab[l][j] = (-1)^(j+1) C(l + j, j) (biotsavart_fmm_2d.cpp:222-231), and
// the L2L shift as run (biotsavart_fmm_2d.cpp:299-307: l is handed to bico by
// reference, so after a rewrite the sweep goes on with the new value in both
// the binomial and the exponent) or exact.
void ensure_t1_tables(vpmfmm_handle *h, int nd, int mode) {
    if (h->t1_nd == nd && h->t1_mode == mode) return;
    Ctx &c = h->ctx;
    std::vector<double> m2l, m2m, l2l;
    fill_tables(nd - 1, 0, m2l, m2m, l2l);
    const int nl = nd + 1;
    std::vector<double> tm2l((size_t)nl * nd), coef((size_t)nl * nl, 0.0);
    std::vector<int> expo((size_t)nl * nl, 0);
    for (int l = 0; l <= nd; ++l)
        for (int j = 0; j < nd; ++j) tm2l[(size_t)l * nd + j] = ((j & 1) ? 1.0 : -1.0) * exact_binom(l + j, j);
    for (int l = 0; l <= nd; ++l) {
        int ll = l;
        for (int kp = l; kp <= nd; ++kp) {
            double cb;
            if (mode == 1) {
                cb = t1_bico(kp, ll);
            } else {
                cb = exact_binom(kp, l);
            }
            coef[(size_t)l * nl + kp] = cb;
            expo[(size_t)l * nl + kp] = kp - ll;
        }
    }
    auto up = [&](DBuf &b, const void *src, size_t bytes) {
        VPM_HIP(hipMemcpyAsync(b.get(bytes), src, bytes, hipMemcpyHostToDevice, c.stream));
    };
    up(c.t1_m2m, m2m.data(), m2m.size() * sizeof(double));
    up(c.t1_m2l, tm2l.data(), tm2l.size() * sizeof(double));
    up(c.t1_l2l_c, coef.data(), coef.size() * sizeof(double));
    up(c.t1_l2l_e, expo.data(), expo.size() * sizeof(int));
    VPM_HIP(hipStreamSynchronize(c.stream));
    h->t1_nd = nd;
    h->t1_mode = mode;
}

// The type-1 velocity (VelocityCalc::velocity_old -> biotsavart_fmm_2d) on
// device pointers: the type-1 tree (every particle active), its lists with
// the reference's box tests, the FP64 type-1 expansions.
void run_type1(vpmfmm_handle *h, const vpmfmm_type1_opts &t, int n, const double *x,
               const double *y, const double *sz, const double *g, double *u, double *v) {
    std::string msg;
    if (t1_opts_check(t, msg)) throw Error(kBadArg, msg);
    if (n < 1) throw Error(kBadArg, "n must be >= 1");
    if (!x || !y || !g || !u || !v || (t.icutoff > 0 && !sz)) throw Error(kBadArg, "null pointer");
    VPM_HIP(hipSetDevice(h->o.device));
    Ctx &c = h->ctx;
    h->insp = false;
    h->exp_p = -1;
    h->last_sel = vpmfmm_selection{};
    c.icut = t.icutoff;
    uint8_t *ones = c.t1_ones.as<uint8_t>(n);
    VPM_HIP(hipMemsetAsync(ones, 1, (size_t)n, c.stream));
    Sources src;
    src.q = g;
    src.s = t.icutoff > 0 ? sz : nullptr;
    src.qscale = -1.0 / (2.0 * M_PI);  // u = Ey, v = -Ex of q = -gz / (2 pi)
    const TreeParams tp{t.n_s, t.n_s, 1.0, 1, t.lmax, true};
    mark(h, 0);
    try {
        tree_build(c, n, x, y, src, ones, tp);
    } catch (const Error &e) {
        if (e.code != kDegenerate) throw;
        // every particle at one point: the reference's quadrants are empty
        // (db = 0), so the rotational part is zero
        c.info.n_cells = 0;
        fill_pair(u, v, n, h->o.u_inf, h->o.v_inf, c.stream);
        VPM_HIP(hipStreamSynchronize(c.stream));
        h->st = vpmfmm_stats{};
        h->st.n = h->st.n_active = n;
        return;
    }
    for (int L = 0; L < kMaxDepth + 2; ++L) h->t1_orders[L] = (uint8_t)t.ndp;
    uint8_t *slot = c.orders_buf.as<uint8_t>(64) + 32;
    VPM_HIP(hipMemcpyAsync(slot, h->t1_orders, sizeof(h->t1_orders), hipMemcpyHostToDevice,
                           c.stream));
    h->fix_valid = false;  // the fixed-order slot now holds ndp
    c.orders = slot;
    mark(h, 1);
    h->concurrent_marks = false;
    tree_lists(c);
    mark(h, 2);
    ensure_t1_tables(h, t.ndp, t.l2l_mode);
    fmm_t1_run(c, t.ndp, u, v, h->o.u_inf, h->o.v_inf, FmmMarks{h->o.timing ? h->ev : nullptr});
    h->sel_fresh = false;
    finish_stats(h, true);
    vpmfmm_selection &ls = h->last_sel;
    ls.pred_err = ls.pred_ms = ls.judge_err = -1.f;
    for (int L = 0; L <= kMaxDepth; ++L) ls.p[L] = t.ndp;
    ls.threshold = t.n_s;
}

// Host-pointer staging.
struct Staged {
    const double *x, *y, *q;
    const uint8_t *a;
};

Staged stage_in(vpmfmm_handle *h, int n, const double *x, const double *y, const double *q,
                const uint8_t *a) {
    if (n < 1) throw Error(kBadArg, "n must be >= 1");
    if (!x || !y || !q || !a) throw Error(kBadArg, "null pointer");
    VPM_HIP(hipSetDevice(h->o.device));
    Ctx &c = h->ctx;
    double *dx = c.x_in.as<double>(n), *dy = c.y_in.as<double>(n), *dq = c.q_in.as<double>(n);
    uint8_t *da = c.a_in.as<uint8_t>(n);
    VPM_HIP(hipMemcpyAsync(dx, x, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    VPM_HIP(hipMemcpyAsync(dy, y, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    VPM_HIP(hipMemcpyAsync(dq, q, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    VPM_HIP(hipMemcpyAsync(da, a, n, hipMemcpyHostToDevice, c.stream));
    return {dx, dy, dq, da};
}

void stage_out(vpmfmm_handle *h, int n, double *a, double *b) {
    Ctx &c = h->ctx;
    VPM_HIP(hipMemcpyAsync(a, c.ex_out.p, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    VPM_HIP(hipMemcpyAsync(b, c.ey_out.p, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    VPM_HIP(hipStreamSynchronize(c.stream));
}

int64_t ref_id(int L, uint32_t ix, uint32_t iy) {
    int64_t p4 = 1;
    for (int i = 0; i < L; ++i) p4 *= 4;
    return (p4 - 1) / 3 + (int64_t)ix + (int64_t)iy * ((int64_t)1 << L);
}

uint32_t compact(uint32_t v) {
    v &= 0x55555555u;
    v = (v | (v >> 1)) & 0x33333333u;
    v = (v | (v >> 2)) & 0x0F0F0F0Fu;
    v = (v | (v >> 4)) & 0x00FF00FFu;
    v = (v | (v >> 8)) & 0x0000FFFFu;
    return v;
}

template <class T>
void d2h(std::vector<T> &dst, const void *src, size_t count, hipStream_t s) {
    dst.resize(count);
    if (count) VPM_HIP(hipMemcpyAsync(dst.data(), src, count * sizeof(T), hipMemcpyDeviceToHost, s));
}

void inspect(vpmfmm_handle *h) {
    if (h->insp) return;
    Ctx &c = h->ctx;
    const TreeInfo &ti = c.info;
    if (ti.n_cells == 0) throw Error(kBadArg, "no tree has been built on this handle");
    const int nc = ti.n_cells, nl = ti.n_leaves;
    hipStream_t s = c.stream;
    d2h(h->h_level, c.c_level.p, nc, s);
    d2h(h->h_flags, c.c_flags.p, nc, s);
    d2h(h->h_key, c.c_key.p, nc, s);
    d2h(h->h_begin, c.c_begin.p, nc, s);
    d2h(h->h_end, c.c_end.p, nc, s);
    d2h(h->h_src, c.c_src.p, nc, s);
    d2h(h->h_perm, c.perm_sorted, ti.n, s);
    d2h(h->h_leaf_idx, c.leaf_idx.p, nl, s);
    d2h(h->h_off1, c.off1.p, nl + 1, s);
    d2h(h->h_off3, c.off3.p, nl + 1, s);
    d2h(h->h_l1, c.l1.p, ti.tot1, s);
    d2h(h->h_l3, c.l3.p, ti.tot3, s);
    d2h(h->h_l2, c.l2.p, 27 * (size_t)nc, s);
    d2h(h->h_l4, c.l4.p, 8 * (size_t)nc, s);
    d2h(h->h_n2, c.n2.p, nc, s);
    d2h(h->h_n4, c.n4.p, nc, s);
    VPM_HIP(hipStreamSynchronize(s));
    h->h_id.resize(nc);
    for (int j = 0; j < nc; ++j) {
        const uint32_t k = h->h_key[j];
        h->h_id[j] = ref_id(h->h_level[j], compact(k), compact(k >> 1));
    }
    h->insp = true;
}

}  // namespace

// ---------------------------------------------------------------------------
// exported functions
// ---------------------------------------------------------------------------
extern "C" {

void vpmfmm_default_opts(vpmfmm_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->p = 10;            // global.cpp:311
    o->max_par = 75;      // global.cpp:312
    o->max_src = 50;      // global.cpp:313
    o->exp_tree = 1.0;    // global.cpp:303
    o->min_level = 1;
    o->max_level = 0;
    o->precision = VPMFMM_FP64;
    o->binom_mode = 1;    // reference helper (parity)
    o->list4_mode = 1;    // M2L_P2_METHOD_TYPE 1 (fmm2D.cpp:7)
    o->u_inf = 1.0;       // global.cpp:164
    o->v_inf = 0.0;       // global.cpp:165
    o->device = 0;
    o->timing = 0;
}

int vpmfmm_create(const vpmfmm_opts *o, vpmfmm_handle **out) {
    if (!out) return VPMFMM_BAD_ARG;
    *out = nullptr;
    vpmfmm_opts oo;
    if (o) oo = *o;
    else vpmfmm_default_opts(&oo);
    std::string msg;
    if (int st = opts_check(oo, msg)) return st;
    vpmfmm_handle *h = new (std::nothrow) vpmfmm_handle();
    if (!h) return VPMFMM_OUT_OF_MEMORY;
    h->o = oo;
    int st = guarded(h, [&] {
        VPM_HIP(hipSetDevice(oo.device));
        VPM_HIP(hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking));
        h->ctx.stream = h->own_stream;
        VPM_HIP(hipMalloc(&h->ctx.hdr_d, sizeof(DevHdr)));
        VPM_HIP(hipHostMalloc(&h->ctx.hdr_h, sizeof(DevHdr), hipHostMallocDefault));
        for (auto &e : h->ev) VPM_HIP(hipEventCreate(&e));
        // streams and events of every path, made here so that no call makes any
        Ctx &c = h->ctx;
        VPM_HIP(hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking));
        VPM_HIP(hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming));
        VPM_HIP(hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming));
        VPM_HIP(hipEventCreateWithFlags(&c.ev_p2m, hipEventDisableTiming));
        VPM_HIP(hipEventCreateWithFlags(&c.ev_sync, hipEventDisableTiming));
        h->scratch = scratch_address(oo.device);
    });
    if (st != VPMFMM_OK) {
        vpmfmm_destroy(h);
        return st;
    }
    *out = h;
    return VPMFMM_OK;
}

int vpmfmm_destroy(vpmfmm_handle *h) {
    if (!h) return VPMFMM_OK;
    (void)hipSetDevice(h->o.device);
    if (h->ctx.stream) (void)hipStreamSynchronize(h->ctx.stream);
    for (auto &e : h->ev)
        if (e) (void)hipEventDestroy(e);
    if (h->ctx.ev_sync) (void)hipEventDestroy(h->ctx.ev_sync);
    if (h->ctx.ev_fork) (void)hipEventDestroy(h->ctx.ev_fork);
    if (h->ctx.ev_join) (void)hipEventDestroy(h->ctx.ev_join);
    if (h->ctx.ev_p2m) (void)hipEventDestroy(h->ctx.ev_p2m);
    if (h->ctx.side) {
        (void)hipStreamSynchronize(h->ctx.side);
        (void)hipStreamDestroy(h->ctx.side);
    }
    if (h->ctx.hdr_d) (void)hipFree(h->ctx.hdr_d);
    if (h->ctx.hdr_h) (void)hipHostFree(h->ctx.hdr_h);
    h->ctx.hdr_d = nullptr;
    h->ctx.hdr_h = nullptr;
    hipStream_t own = h->own_stream;
    delete h;  // DBuf destructors free device buffers
    if (own) (void)hipStreamDestroy(own);
    return VPMFMM_OK;
}

int vpmfmm_set_opts(vpmfmm_handle *h, const vpmfmm_opts *o) {
    if (!h || !o) return VPMFMM_BAD_ARG;
    std::string msg;
    if (int st = opts_check(*o, msg)) {
        h->err = msg;
        return st;
    }
    if (o->device != h->o.device) {
        h->err = "the device of a handle is fixed at creation";
        return VPMFMM_BAD_ARG;
    }
    h->o = *o;
    return VPMFMM_OK;
}

int vpmfmm_get_opts(const vpmfmm_handle *h, vpmfmm_opts *o) {
    if (!h || !o) return VPMFMM_BAD_ARG;
    *o = h->o;
    return VPMFMM_OK;
}

int vpmfmm_set_stream(vpmfmm_handle *h, void *stream) {
    if (!h) return VPMFMM_BAD_ARG;
    h->ctx.stream = stream ? (hipStream_t)stream : h->own_stream;
    return VPMFMM_OK;
}

int vpmfmm_field_device(vpmfmm_handle *h, int32_t n, const double *x, const double *y,
                        const double *q, const uint8_t *active, double *ex, double *ey) {
    return guarded(h, [&] { run(h, n, x, y, Sources{q}, active, ex, ey, false); });
}

int vpmfmm_field(vpmfmm_handle *h, int32_t n, const double *x, const double *y, const double *q,
                 const uint8_t *active, double *ex, double *ey) {
    return guarded(h, [&] {
        if (!ex || !ey) throw Error(kBadArg, "null pointer");
        Staged s = stage_in(h, n, x, y, q, active);
        Ctx &c = h->ctx;
        double *oa = c.ex_out.as<double>(n), *ob = c.ey_out.as<double>(n);
        run(h, n, s.x, s.y, Sources{s.q}, s.a, oa, ob, false);
        stage_out(h, n, ex, ey);
    });
}

int vpmfmm_potential_device(vpmfmm_handle *h, int32_t n, const double *x, const double *y,
                            const double *q, const uint8_t *active, double *phi) {
    return guarded(h, [&] { run_potential(h, n, x, y, q, active, phi); });
}

int vpmfmm_potential(vpmfmm_handle *h, int32_t n, const double *x, const double *y,
                     const double *q, const uint8_t *active, double *phi) {
    return guarded(h, [&] {
        if (!phi) throw Error(kBadArg, "null pointer");
        Staged st = stage_in(h, n, x, y, q, active);
        Ctx &c = h->ctx;
        double *dphi = c.ex_out.as<double>(n);
        run_potential(h, n, st.x, st.y, st.q, st.a, dphi);
        VPM_HIP(hipMemcpyAsync(phi, dphi, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
        VPM_HIP(hipStreamSynchronize(c.stream));
    });
}

int vpmfmm_velocity_device(vpmfmm_handle *h, int32_t n, const double *x, const double *y,
                           const double *s, const double *omega, const uint8_t *active, double *u,
                           double *v) {
    return guarded(h, [&] {
        if (n < 1 || !s || !omega) throw Error(kBadArg, "bad size or null pointer");
        VPM_HIP(hipSetDevice(h->o.device));
        run(h, n, x, y, Sources{nullptr, s, omega}, active, u, v, true);
    });
}

int vpmfmm_velocity_device_noalloc(vpmfmm_handle *h, int32_t n, const double *x,
                                   const double *y, const double *s, const double *omega,
                                   const uint8_t *active, double *u, double *v) {
    return guarded(h, [&] {
        if (n < 1 || !s || !omega) throw Error(kBadArg, "bad size or null pointer");
        if (n > VPMFMM_NOALLOC_MAX_N)
            throw Error(kBadArg, "n = " + std::to_string(n) +
                                     " exceeds the no-allocation limit of " +
                                     std::to_string(VPMFMM_NOALLOC_MAX_N) +
                                     " particles; use vpmfmm_velocity_device");
        if (!h->scratch) throw Error(kInternal, "no scratch address on this handle");
        VPM_HIP(hipSetDevice(h->o.device));
        size_t used = 0;
        try {
            ArenaScope arena(h);
            try {
                run(h, n, x, y, Sources{nullptr, s, omega}, active, u, v, true);
            } catch (...) {
                // launches of a failed call may still read or write the
                // scratch: let them finish before the array is used again
                (void)hipStreamSynchronize(h->ctx.stream);
                (void)hipStreamSynchronize(h->ctx.side);
                throw;
            }
            used = arena.used();
        } catch (...) {
            h->st = vpmfmm_stats{};
            throw;
        }
        h->st.scratch_bytes = (int64_t)used;
    });
}

int vpmfmm_velocity(vpmfmm_handle *h, int32_t n, const double *x, const double *y, const double *s,
                    const double *omega, const uint8_t *active, double *u, double *v) {
    return guarded(h, [&] {
        if (n < 1 || !x || !y || !s || !omega || !active || !u || !v)
            throw Error(kBadArg, "bad size or null pointer");
        VPM_HIP(hipSetDevice(h->o.device));
        Ctx &c = h->ctx;
        double *dx = c.x_in.as<double>(n), *dy = c.y_in.as<double>(n);
        double *ds = c.ex_out.as<double>(n), *dw = c.ey_out.as<double>(n);
        uint8_t *da = c.a_in.as<uint8_t>(n);
        VPM_HIP(hipMemcpyAsync(dx, x, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
        VPM_HIP(hipMemcpyAsync(dy, y, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
        VPM_HIP(hipMemcpyAsync(ds, s, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
        VPM_HIP(hipMemcpyAsync(dw, omega, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
        VPM_HIP(hipMemcpyAsync(da, active, n, hipMemcpyHostToDevice, c.stream));
        // the outputs reuse the s / omega staging (read by the tree build first)
        run(h, n, dx, dy, Sources{nullptr, ds, dw}, da, ds, dw, true);
        VPM_HIP(hipMemcpyAsync(u, ds, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
        VPM_HIP(hipMemcpyAsync(v, dw, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
        VPM_HIP(hipStreamSynchronize(c.stream));
    });
}

void vpmfmm_type1_default_opts(vpmfmm_type1_opts *o) {
    if (!o) return;
    *o = vpmfmm_type1_opts{};
    o->n_s = 50;      // global.cpp:326
    o->lmax = 7;      // global.cpp:316
    o->ndp = 10;      // global.cpp:327
    o->icutoff = 2;   // global.cpp:328
    o->l2l_mode = 1;  // the reference as it runs
}

int vpmfmm_velocity_type1_device(vpmfmm_handle *h, const vpmfmm_type1_opts *o, int32_t n,
                                 const double *x, const double *y, const double *s,
                                 const double *gz, double *u, double *v) {
    return guarded(h, [&] {
        if (!o) throw Error(kBadArg, "null options");
        run_type1(h, *o, n, x, y, s, gz, u, v);
    });
}

int vpmfmm_velocity_type1(vpmfmm_handle *h, const vpmfmm_type1_opts *o, int32_t n,
                          const double *x, const double *y, const double *s, const double *gz,
                          double *u, double *v) {
    return guarded(h, [&] {
        if (!o) throw Error(kBadArg, "null options");
        if (n < 1 || !x || !y || !gz || !u || !v || (o->icutoff > 0 && !s))
            throw Error(kBadArg, "bad size or null pointer");
        VPM_HIP(hipSetDevice(h->o.device));
        Ctx &c = h->ctx;
        double *dx = c.x_in.as<double>(n), *dy = c.y_in.as<double>(n);
        double *ds = c.ex_out.as<double>(n), *dg = c.ey_out.as<double>(n);
        VPM_HIP(hipMemcpyAsync(dx, x, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
        VPM_HIP(hipMemcpyAsync(dy, y, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
        if (o->icutoff > 0)
            VPM_HIP(hipMemcpyAsync(ds, s, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
        VPM_HIP(hipMemcpyAsync(dg, gz, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
        // the outputs reuse the s / gz staging (read by the tree build first)
        run_type1(h, *o, n, dx, dy, o->icutoff > 0 ? ds : nullptr, dg, ds, dg);
        VPM_HIP(hipMemcpyAsync(u, ds, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
        VPM_HIP(hipMemcpyAsync(v, dg, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
        VPM_HIP(hipStreamSynchronize(c.stream));
    });
}

void vpmfmm_fmm3d_default_opts(vpmfmm_fmm3d_opts *o) {
    if (!o) return;
    *o = vpmfmm_fmm3d_opts{};
    // the reference's DIM == 3 parameter block (global.cpp:474-875)
    o->crit_num = 30;        // FMMCell::critNum = Pars::src_count_max (global.cpp:779)
    o->max_depth = kFmm3Depth;
    o->exp_tree = 1.0;       // Pars::expTree (global.cpp:769)
    o->support = 2.0;        // SUPPORT_RADIUS_FACTOR (fast3DFMM.cpp:4, fmm3D.cpp:3)
    o->u_inf = 1.0;          // Pars::u_inf, v_inf, w_inf (global.cpp:633-635)
    o->type = 1;             // VELOCITY_3D_TYPE (velocity_calc.cpp:31)
    o->max_par = 50;         // TreeCell: Pars::par_count_max (global.cpp:778)
    o->max_src = 30;         //           Pars::src_count_max (global.cpp:779)
}

namespace {
Fmm3Params fmm3_params(const vpmfmm_fmm3d_opts &o) {
    Fmm3Params p;
    p.crit_num = o.crit_num;
    p.max_depth = o.max_depth;
    p.exp_tree = o.exp_tree;
    for (int d = 0; d < 3; ++d) p.center[d] = o.center[d];
    p.support = o.support;
    p.u_inf = o.u_inf;
    p.v_inf = o.v_inf;
    p.w_inf = o.w_inf;
    p.type = o.type;
    p.max_par = o.max_par;
    p.max_src = o.max_src;
    return p;
}

Ctx3 &ctx3(vpmfmm_handle *h) {
    if (!h->c3) h->c3.reset(new Ctx3());
    return *h->c3;
}
}  // namespace

int vpmfmm_velocity3d_device(vpmfmm_handle *h, const vpmfmm_fmm3d_opts *o, int32_t n,
                             const double *x, const double *y, const double *z, const double *s,
                             const double *wx, const double *wy, const double *wz,
                             const uint8_t *active, double *u, double *v, double *w) {
    return guarded(h, [&] {
        if (!o) throw Error(kBadArg, "null options");
        VPM_HIP(hipSetDevice(h->o.device));
        if (!s) throw Error(kBadArg, "null pointer");
        fmm3d_velocity(ctx3(h), h->ctx.stream, n, x, y, z, s, wx, wy, wz, active, fmm3_params(*o),
                       u, v, w, h->o.timing ? h->ev : nullptr);
    });
}

namespace {
// Host-pointer 3-D solve: s null passes (wx, wy, wz) as the strengths and
// adds no free stream.
void velocity3d_host(vpmfmm_handle *h, const vpmfmm_fmm3d_opts &o, int32_t n, const double *x,
                     const double *y, const double *z, const double *s, const double *wx,
                     const double *wy, const double *wz, const uint8_t *active, double *u,
                     double *v, double *w) {
    {
        if (n < 1 || !x || !y || !z || !wx || !wy || !wz || !active || !u || !v || !w)
            throw Error(kBadArg, "bad size or null pointer");
        VPM_HIP(hipSetDevice(h->o.device));
        Ctx &c = h->ctx;
        const size_t b = sizeof(double) * (size_t)n;
        // staging: seven inputs, then the three outputs in the first three slots
        double *st = c.x_in.as<double>(7 * (size_t)n);
        uint8_t *da = c.a_in.as<uint8_t>(n);
        const double *src[7] = {x, y, z, s, wx, wy, wz};
        for (int k = 0; k < 7; ++k)
            if (src[k])
                VPM_HIP(hipMemcpyAsync(st + (size_t)k * n, src[k], b, hipMemcpyHostToDevice,
                                       c.stream));
        VPM_HIP(hipMemcpyAsync(da, active, n, hipMemcpyHostToDevice, c.stream));
        double *out = c.ex_out.as<double>(3 * (size_t)n);
        Fmm3Params p = fmm3_params(o);
        if (!s) p.u_inf = p.v_inf = p.w_inf = 0.0;
        fmm3d_velocity(ctx3(h), c.stream, n, st, st + n, st + 2 * (size_t)n,
                       s ? st + 3 * (size_t)n : nullptr, st + 4 * (size_t)n, st + 5 * (size_t)n,
                       st + 6 * (size_t)n, da, p, out, out + n, out + 2 * (size_t)n,
                       h->o.timing ? h->ev : nullptr);
        VPM_HIP(hipMemcpyAsync(u, out, b, hipMemcpyDeviceToHost, c.stream));
        VPM_HIP(hipMemcpyAsync(v, out + n, b, hipMemcpyDeviceToHost, c.stream));
        VPM_HIP(hipMemcpyAsync(w, out + 2 * (size_t)n, b, hipMemcpyDeviceToHost, c.stream));
        VPM_HIP(hipStreamSynchronize(c.stream));
    }
}
}  // namespace

int vpmfmm_velocity3d(vpmfmm_handle *h, const vpmfmm_fmm3d_opts *o, int32_t n, const double *x,
                      const double *y, const double *z, const double *s, const double *wx,
                      const double *wy, const double *wz, const uint8_t *active, double *u,
                      double *v, double *w) {
    return guarded(h, [&] {
        if (!o) throw Error(kBadArg, "null options");
        if (!s) throw Error(kBadArg, "null pointer");
        velocity3d_host(h, *o, n, x, y, z, s, wx, wy, wz, active, u, v, w);
    });
}

int vpmfmm_velocity3d_alpha(vpmfmm_handle *h, const vpmfmm_fmm3d_opts *o, int32_t n,
                            const double *x, const double *y, const double *z, const double *ax,
                            const double *ay, const double *az, const uint8_t *active, double *u,
                            double *v, double *w) {
    return guarded(h, [&] {
        if (!o) throw Error(kBadArg, "null options");
        velocity3d_host(h, *o, n, x, y, z, nullptr, ax, ay, az, active, u, v, w);
    });
}

int vpmfmm_fmm3d_last_stats(vpmfmm_handle *h, vpmfmm_fmm3d_stats *st) {
    return guarded(h, [&] {
        if (!st) throw Error(kBadArg, "null pointer");
        *st = vpmfmm_fmm3d_stats{};
        if (!h->c3) return;
        const Fmm3Info &fi = h->c3->info;
        st->n = fi.n;
        st->n_active = fi.n_active;
        st->n_cells = fi.n_cells;
        st->n_leaves = fi.n_leaves;
        st->depth = fi.depth;
        st->sorted_levels = fi.sorted_levels;
        for (int d = 0; d < 3; ++d) st->root_centre[d] = fi.root_c[d];
        st->root_side = fi.root_len;
        st->pairs = fi.pairs;
        st->m2p = fi.m2p;
        st->wave_pairs = fi.wave_pairs;
        st->wave_m2p = fi.wave_m2p;
        st->ms_tree = fi.ms_tree;
        st->ms_moments = fi.ms_moments;
        st->ms_eval = fi.ms_eval;
    });
}

int vpmfmm_fmm3d_cells(vpmfmm_handle *h, int32_t *level, int64_t *key, int32_t *count,
                       uint8_t *leaf, uint8_t *active, double *centre3, double *side) {
    return guarded(h, [&] {
        if (!h->c3) throw Error(kBadArg, "no 3-D solve has run on this handle");
        VPM_HIP(hipSetDevice(h->o.device));
        fmm3d_cells(*h->c3, h->ctx.stream, level, key, count, leaf, active, centre3, side);
    });
}

int vpmfmm_fmm3d_perm(vpmfmm_handle *h, int32_t *perm) {
    return guarded(h, [&] {
        if (!h->c3) throw Error(kBadArg, "no 3-D solve has run on this handle");
        if (!perm) throw Error(kBadArg, "null pointer");
        VPM_HIP(hipSetDevice(h->o.device));
        fmm3d_perm(*h->c3, h->ctx.stream, perm);
    });
}

static_assert(kAdvectEuler == VPMFMM_ADVECT_EULER && kAdvectRk2 == VPMFMM_ADVECT_RK2,
              "advection scheme codes");

int vpmfmm_advect_device(vpmfmm_handle *h, int32_t n, double *x, double *y, const double *u,
                         const double *v, double dt, int32_t scheme) {
    return guarded(h, [&] {
        if (n < 1 || !x || !y || !u || !v) throw Error(kBadArg, "bad size or null pointer");
        if (scheme != VPMFMM_ADVECT_EULER && scheme != VPMFMM_ADVECT_RK2)
            throw Error(kBadArg, "unknown advection scheme");
        VPM_HIP(hipSetDevice(h->o.device));
        advect(n, x, y, u, v, dt, scheme, h->ctx.stream);
        VPM_HIP(hipStreamSynchronize(h->ctx.stream));
    });
}

int vpmfmm_step_device(vpmfmm_handle *h, int32_t n, double *x, double *y, const double *s,
                       const double *omega, const uint8_t *active, double *u, double *v,
                       double dt, int32_t scheme) {
    return guarded(h, [&] {
        if (n < 1 || !x || !y || !s || !omega || !u || !v)
            throw Error(kBadArg, "bad size or null pointer");
        if (scheme != VPMFMM_ADVECT_EULER && scheme != VPMFMM_ADVECT_RK2)
            throw Error(kBadArg, "unknown advection scheme");
        VPM_HIP(hipSetDevice(h->o.device));
        // same stream: the tree build reads the moved positions
        advect(n, x, y, u, v, dt, scheme, h->ctx.stream);
        run(h, n, x, y, Sources{nullptr, s, omega}, active, u, v, true);
    });
}

int vpmfmm_direct_device(vpmfmm_handle *h, int32_t n, const double *x, const double *y,
                         const double *q, const uint8_t *src_mask, int32_t nt, const int32_t *tidx,
                         double *ex, double *ey) {
    return guarded(h, [&] {
        if (n < 1 || !x || !y || !q || !ex || !ey) throw Error(kBadArg, "bad size or null pointer");
        if (!tidx) nt = n;
        if (nt < 1) throw Error(kBadArg, "nt must be >= 1");
        VPM_HIP(hipSetDevice(h->o.device));
        direct64(n, x, y, q, src_mask, nt, tidx, ex, ey, h->ctx.stream);
        VPM_HIP(hipStreamSynchronize(h->ctx.stream));
    });
}

int vpmfmm_last_stats(const vpmfmm_handle *h, vpmfmm_stats *st) {
    if (!h || !st) return VPMFMM_BAD_ARG;
    *st = h->st;
    return VPMFMM_OK;
}

const char *vpmfmm_last_error(const vpmfmm_handle *h) { return h ? h->err.c_str() : "null handle"; }

int vpmfmm_build_tree(vpmfmm_handle *h, int32_t n, const double *x, const double *y,
                      const uint8_t *active) {
    return guarded(h, [&] {
        std::vector<double> zeros((size_t)std::max(n, 1), 0.0);
        Staged s = stage_in(h, n, x, y, zeros.data(), active);
        build(h, n, s.x, s.y, Sources{s.q}, s.a);
        finish_stats(h, false);
    });
}

int vpmfmm_tree_counts(vpmfmm_handle *h, int32_t *n_cells, int32_t *n_leaves,
                       int32_t *min_leaf_level, int32_t *max_leaf_level) {
    return guarded(h, [&] {
        const TreeInfo &ti = h->ctx.info;
        if (ti.n_cells == 0) throw Error(kBadArg, "no tree has been built on this handle");
        if (n_cells) *n_cells = ti.n_cells;
        if (n_leaves) *n_leaves = ti.n_leaves;
        if (min_leaf_level) *min_leaf_level = ti.min_leaf_level;
        if (max_leaf_level) *max_leaf_level = ti.max_level;
    });
}

int vpmfmm_tree_geometry(vpmfmm_handle *h, double *root_len, double *pivot2) {
    return guarded(h, [&] {
        const TreeInfo &ti = h->ctx.info;
        if (ti.n_cells == 0) throw Error(kBadArg, "no tree has been built on this handle");
        if (root_len) *root_len = ti.root_len;
        if (pivot2) {
            pivot2[0] = ti.pivot[0];
            pivot2[1] = ti.pivot[1];
        }
    });
}

int vpmfmm_tree_cells(vpmfmm_handle *h, int64_t *id, int32_t *level, int32_t *ix, int32_t *iy,
                      int32_t *par_num, int32_t *src_num, uint8_t *is_leaf, uint8_t *is_active) {
    return guarded(h, [&] {
        inspect(h);
        const int nc = h->ctx.info.n_cells;
        std::vector<int> ord(nc);
        for (int j = 0; j < nc; ++j) ord[j] = j;
        std::sort(ord.begin(), ord.end(), [&](int a, int b) { return h->h_id[a] < h->h_id[b]; });
        for (int k = 0; k < nc; ++k) {
            const int j = ord[k];
            const uint32_t key = h->h_key[j];
            if (id) id[k] = h->h_id[j];
            if (level) level[k] = h->h_level[j];
            if (ix) ix[k] = (int32_t)compact(key);
            if (iy) iy[k] = (int32_t)compact(key >> 1);
            if (par_num) par_num[k] = (int32_t)(h->h_end[j] - h->h_begin[j]);
            if (src_num) src_num[k] = (int32_t)h->h_src[j];
            if (is_leaf) is_leaf[k] = (h->h_flags[j] & 1u) ? 1 : 0;
            if (is_active) is_active[k] = (h->h_flags[j] & 2u) ? 1 : 0;
        }
    });
}

int vpmfmm_tree_leaves(vpmfmm_handle *h, int64_t *leaf_id, int32_t *poff, int32_t *pidx) {
    return guarded(h, [&] {
        inspect(h);
        const int nl = h->ctx.info.n_leaves;
        std::vector<int> ord(h->h_leaf_idx.begin(), h->h_leaf_idx.end());
        std::sort(ord.begin(), ord.end(), [&](int a, int b) { return h->h_id[a] < h->h_id[b]; });
        int off = 0;
        for (int k = 0; k < nl; ++k) {
            const int j = ord[k];
            if (leaf_id) leaf_id[k] = h->h_id[j];
            if (poff) poff[k] = off;
            const int b = off;
            for (uint32_t i = h->h_begin[j]; i < h->h_end[j]; ++i, ++off)
                if (pidx) pidx[off] = (int32_t)h->h_perm[i];
            // inside a leaf the device keeps Morton order; report index order
            // (the order of the reference's parIDList, treeCell.cpp:1215-1232)
            if (pidx) std::sort(pidx + b, pidx + off);
        }
        if (poff) poff[nl] = off;
    });
}

int vpmfmm_tree_lists(vpmfmm_handle *h, int64_t id, int64_t *l1, int32_t *n1, int64_t *l2,
                      int32_t *n2, int64_t *l3, int32_t *n3, int64_t *l4, int32_t *n4,
                      int32_t cap) {
    return guarded(h, [&] {
        inspect(h);
        const int nc = h->ctx.info.n_cells;
        int j = -1;
        for (int k = 0; k < nc; ++k)
            if (h->h_id[k] == id) {
                j = k;
                break;
            }
        if (j < 0) throw Error(kBadArg, "no cell with this ID");
        auto put = [&](const int *src, int cnt, int64_t *dst, int32_t *ncnt) {
            if (ncnt) *ncnt = cnt;
            if (dst)
                for (int u = 0; u < cnt && u < cap; ++u) dst[u] = h->h_id[src[u]];
        };
        int c1 = 0, c3 = 0;
        const int *p1 = nullptr, *p3 = nullptr;
        if (h->h_flags[j] & 1u) {
            const auto it = std::find(h->h_leaf_idx.begin(), h->h_leaf_idx.end(), (uint32_t)j);
            const size_t t = (size_t)(it - h->h_leaf_idx.begin());
            c1 = (int)(h->h_off1[t + 1] - h->h_off1[t]);
            c3 = (int)(h->h_off3[t + 1] - h->h_off3[t]);
            p1 = h->h_l1.data() + h->h_off1[t];
            p3 = h->h_l3.data() + h->h_off3[t];
        }
        put(p1, c1, l1, n1);
        put(p3, c3, l3, n3);
        const bool m2l = h->h_level[j] >= 2;
        put(h->h_l2.data() + 27 * (size_t)j, m2l ? h->h_n2[j] : 0, l2, n2);
        put(h->h_l4.data() + 8 * (size_t)j, m2l ? h->h_n4[j] : 0, l4, n4);
    });
}

int vpmfmm_set_level_orders(vpmfmm_handle *h, const int32_t *p, int32_t nlev) {
    return guarded(h, [&] {
        if (nlev < 0 || nlev > kMaxDepth + 1 || (nlev > 0 && !p))
            throw Error(kBadArg, "nlev must be in [0, 17]");
        std::vector<int> v(p ? p : nullptr, p ? p + nlev : nullptr);
        for (int x : v)
            if (x < 0 || x > 16) throw Error(kBadArg, "per-level orders must be in [0, 16]");
        h->level_p = v;
    });
}

int vpmfmm_set_selector(vpmfmm_handle *h, const vpmfmm_selector *sel, const float *params) {
    return guarded(h, [&] {
        if (!sel) {
            h->has_sel = false;
            h->sel_threshold = h->sel_n = h->sel_calls = h->judge_calls = 0;
            h->judge_reject = false;
            h->sel_cached = false;
            return;
        }
        const vpmfmm_selector &m = *sel;
        if (m.n_features != kSelFeatures || m.hidden != kSelHidden || m.n_params != kSelParams)
            throw Error(kBadArg, "selector shape must be 9 features, 64 hidden, 4930 parameters");
        if (m.p_min < 1 || m.p_max > 16 || m.p_min > m.p_max || m.p_max - m.p_min >= 16)
            throw Error(kBadArg, "selector orders must satisfy 1 <= p_min <= p_max <= 16");
        if (m.n_thresholds < 1 || m.n_thresholds > 8)
            throw Error(kBadArg, "selector needs 1..8 leaf thresholds");
        for (int k = 0; k < m.n_thresholds; ++k)
            if (m.thresholds[k] < 1 || m.thresholds[k] > kMaxThreshold)
                throw Error(kBadArg, "selector thresholds must be in [1, 4096]");
        if (!(m.target_err > 0.f) || !std::isfinite(m.target_err) || !(m.margin >= 0.f))
            throw Error(kBadArg, "selector target_err must be > 0 and margin >= 0");
        if (!params) throw Error(kBadArg, "null pointer");
        for (int i = 0; i < kSelFeatures; ++i)
            if (!(m.sd[i] > 0.f) || !std::isfinite(m.mu[i]) || !std::isfinite(m.lo[i]) ||
                !std::isfinite(m.hi[i]) || m.lo[i] > m.hi[i])
                throw Error(kBadArg, "selector normalisation / range is invalid");
        for (int k = 0; k < kSelParams; ++k)
            if (!std::isfinite(params[k])) throw Error(kNonFinite, "non-finite selector parameter");
        SelModel sm{};
        for (int i = 0; i < kSelFeatures; ++i) {
            sm.mu[i] = m.mu[i];
            sm.sd[i] = m.sd[i];
            sm.lo[i] = m.lo[i];
            sm.hi[i] = m.hi[i];
        }
        sm.margin = m.margin;
        sm.target_err = m.target_err;
        sm.p_min = m.p_min;
        sm.p_max = m.p_max;
        VPM_HIP(hipSetDevice(h->o.device));
        Ctx &c = h->ctx;
        VPM_HIP(hipMemcpyAsync(h->sel_model_d.get(sizeof(SelModel)), &sm, sizeof(SelModel),
                               hipMemcpyHostToDevice, c.stream));
        VPM_HIP(hipMemcpyAsync(h->sel_params_d.get(sizeof(float) * kSelParams), params,
                               sizeof(float) * kSelParams, hipMemcpyHostToDevice, c.stream));
        VPM_HIP(hipStreamSynchronize(c.stream));
        h->sel = m;
        h->has_sel = true;
        h->sel_threshold = h->sel_n = h->sel_calls = h->judge_calls = 0;
        h->judge_reject = false;
        h->sel_cached = false;
    });
}

int vpmfmm_last_selection(vpmfmm_handle *h, vpmfmm_selection *out) {
    return guarded(h, [&] {
        if (!out) throw Error(kBadArg, "null pointer");
        *out = h->last_sel;
    });
}

int vpmfmm_level_features(vpmfmm_handle *h, float *feats, double *raw, int32_t *n_levels) {
    return guarded(h, [&] {
        Ctx &c = h->ctx;
        if (c.info.n_cells == 0) throw Error(kBadArg, "no tree has been built on this handle");
        VPM_HIP(hipSetDevice(h->o.device));
        const int thr = h->last_sel.threshold > 0 ? h->last_sel.threshold : h->o.max_par;
        float *fd = c.sel_feats.as<float>((kMaxDepth + 1) * kSelFeatures);
        float *rd = c.sel_raw.as<float>((kMaxDepth + 1) * kSelRaw);
        level_features(c, thr, fd, rd);
        std::vector<float> f((kMaxDepth + 1) * kSelFeatures), r((kMaxDepth + 1) * kSelRaw);
        VPM_HIP(hipMemcpyAsync(f.data(), fd, f.size() * 4, hipMemcpyDeviceToHost, c.stream));
        VPM_HIP(hipMemcpyAsync(r.data(), rd, r.size() * 4, hipMemcpyDeviceToHost, c.stream));
        VPM_HIP(hipStreamSynchronize(c.stream));
        if (feats) std::memcpy(feats, f.data(), f.size() * 4);
        if (raw)
            for (size_t k = 0; k < r.size(); ++k) raw[k] = r[k];
        if (n_levels) *n_levels = c.info.max_level + 1;
    });
}

namespace {
void check_segments(vpmfmm_handle *h, const int32_t *seg, int32_t nseg, int32_t rows) {
    if (nseg < 1 || rows < 1 || !seg) throw Error(kBadArg, "need nseg >= 1 and rows >= 1");
    std::vector<int32_t> hs(nseg + 1);
    VPM_HIP(hipMemcpyAsync(hs.data(), seg, hs.size() * 4, hipMemcpyDeviceToHost, h->ctx.stream));
    VPM_HIP(hipStreamSynchronize(h->ctx.stream));
    if (hs[0] != 0 || hs[nseg] != rows) throw Error(kBadArg, "seg must run from 0 to rows");
    for (int s = 0; s < nseg; ++s)
        if (hs[s + 1] <= hs[s]) throw Error(kBadArg, "every solve needs at least one row");
}
}  // namespace

int vpmfmm_mlp_forward(vpmfmm_handle *h, const float *params, const float *x, int32_t rows,
                       float *out) {
    return guarded(h, [&] {
        if (!params || !x || !out || rows < 0) throw Error(kBadArg, "null pointer or rows < 0");
        VPM_HIP(hipSetDevice(h->o.device));
        mlp_forward(h->ctx.stream, params, x, rows, out, nullptr, nullptr);
        VPM_HIP(hipStreamSynchronize(h->ctx.stream));
    });
}

int vpmfmm_mlp_loss_grad(vpmfmm_handle *h, const float *params, const float *x, int32_t rows,
                         const int32_t *seg, int32_t nseg, const float *y, float w_time,
                         float *grad, float *loss) {
    return guarded(h, [&] {
        if (!params || !x || !y || !grad || !loss) throw Error(kBadArg, "null pointer");
        VPM_HIP(hipSetDevice(h->o.device));
        check_segments(h, seg, nseg, rows);
        mlp_loss_grad(h->ctx.stream, params, x, rows, seg, nseg, y, w_time, h->mlp, grad, loss);
        VPM_HIP(hipStreamSynchronize(h->ctx.stream));
    });
}

int vpmfmm_mlp_train(vpmfmm_handle *h, float *params, float *m, float *v, const float *x,
                     int32_t rows, const int32_t *seg, int32_t nseg, const float *y, float w_time,
                     int32_t steps, int32_t step0, float lr, float lr_min, float *loss_hist) {
    return guarded(h, [&] {
        if (!params || !m || !v || !x || !y || !loss_hist) throw Error(kBadArg, "null pointer");
        if (steps < 1 || step0 < 0 || !(lr > 0.f) || !(lr_min >= 0.f))
            throw Error(kBadArg, "need steps >= 1, step0 >= 0, lr > 0, lr_min >= 0");
        VPM_HIP(hipSetDevice(h->o.device));
        check_segments(h, seg, nseg, rows);
        hipStream_t s = h->ctx.stream;
        float *grad = h->mlp_grad.as<float>(kSelParams);
        const double pi = 3.14159265358979323846;
        for (int k = 0; k < steps; ++k) {
            const float lr_k = lr_min + 0.5f * (lr - lr_min) *
                                            (float)(1.0 + std::cos(pi * (double)k / (double)steps));
            mlp_loss_grad(s, params, x, rows, seg, nseg, y, w_time, h->mlp, grad, loss_hist + k);
            mlp_adam(s, params, grad, m, v, kSelParams, step0 + k + 1, lr_k, 0.9f, 0.999f, 1e-8f);
        }
        VPM_HIP(hipStreamSynchronize(s));
    });
}

int vpmfmm_tree_expansions(vpmfmm_handle *h, int64_t *ids, double *ak, double *bk) {
    return guarded(h, [&] {
        inspect(h);
        if (h->exp_p < 0) throw Error(kBadArg, "no FP64 expansions from the last call");
        const int nc = h->ctx.info.n_cells, P1 = h->exp_p + 1;
        std::vector<double> a((size_t)nc * 2 * P1), b((size_t)nc * 2 * P1);
        Ctx &c = h->ctx;
        VPM_HIP(hipMemcpyAsync(a.data(), c.ak.p, a.size() * sizeof(double), hipMemcpyDeviceToHost,
                               c.stream));
        VPM_HIP(hipMemcpyAsync(b.data(), c.bk.p, b.size() * sizeof(double), hipMemcpyDeviceToHost,
                               c.stream));
        VPM_HIP(hipStreamSynchronize(c.stream));
        std::vector<int> ord(nc);
        for (int j = 0; j < nc; ++j) ord[j] = j;
        std::sort(ord.begin(), ord.end(), [&](int x, int y) { return h->h_id[x] < h->h_id[y]; });
        for (int k = 0; k < nc; ++k) {
            const int j = ord[k];
            if (ids) ids[k] = h->h_id[j];
            if (ak) std::memcpy(ak + (size_t)k * 2 * P1, a.data() + (size_t)j * 2 * P1, 16 * P1);
            if (bk) std::memcpy(bk + (size_t)k * 2 * P1, b.data() + (size_t)j * 2 * P1, 16 * P1);
        }
    });
}

}  // extern "C"
