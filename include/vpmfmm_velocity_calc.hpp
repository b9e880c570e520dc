/* vpmfmm_velocity_calc.hpp -- header-only C++ drop-in for the FMM branch of
 * the reference's velocity manager.
 *
 * Replaces class VelocityCalc (src/velocity_calculation/velocity_calc.hpp:15-47)
 * for VELOCITY_2D_TYPE = FMM (2) and the type-1 FMM (1, set_default_velocity_type):
 * get_velocity(particle, base_grid, step) and
 * get_velocity(particle, step) (velocity_calc.cpp:149-251) compute, on the FMM branch,
 *   u = E_y + u_inf, v = -E_x + v_inf,  E_i = sum_{j != i} q_j (x_i - x_j) / r_ij^2,
 *   q_j = -omega_j s_j^2 / (2 pi)
 * (velocity_calc.cpp:455-578, fmm2D.cpp) on the GPU through include/vpmfmm.h,
 * reading particle.num, x, y, s, vorticity, isActive and writing particle.u, v
 * (resized to num).  The GridNode argument is ignored, as on the reference's
 * FMM branch.  Errors become the exception types the reference throws:
 * std::invalid_argument (bad options or arguments), std::range_error
 * (non-finite input, degenerate domain, tree deeper than 16 levels, two
 * particles at one position), std::runtime_error (device or internal errors).
 *
 * Link with lsmps_vpm_v2_amd/libvpmfmm.so.  The particle type is a template
 * parameter so the header compiles against the reference's struct Particle
 * (Utils.hpp:77-159) or any type with the same members.  Define
 * VPMFMM_NO_GLOBAL_VELOCITYCALC to keep the class inside namespace vpmfmm.
 * With DIM defined as 3 before this header (the reference's global.hpp:85),
 * get_velocity takes the reference's 3-D branch instead: VELOCITY_3D_TYPE 1,
 * velocity_fmm_3d_fast (velocity_calc.cpp:201-236, 2384-2525), reading
 * particle.num, x, y, z, s, vortx, vorty, vortz, isActive and writing u, v, w
 * (free stream and treecode knobs from set_default_fmm3d_options).
 * The whole reference program must see this one class: include it through
 * vpmfmm_reference_velocity_calc.hpp, which takes the place of the
 * reference's src/velocity_calculation/velocity_calc.hpp, the header both of
 * its callers include (main.cpp:44, penalization_calc.cpp:3).
 */
#ifndef VPMFMM_VELOCITY_CALC_HPP
#define VPMFMM_VELOCITY_CALC_HPP

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "vpmfmm.h"

// The dimension get_velocity serves: the reference's DIM (global.hpp:85) when
// it is defined before this header, else 2.
#if defined(DIM)
#define VPMFMM_DIM DIM
#else
#define VPMFMM_DIM 2
#endif

namespace vpmfmm {

inline void throw_status(int st, const char *where, vpmfmm_handle *h) {
    if (st == VPMFMM_OK) return;
    std::string msg = std::string(where) + ": ";
    const char *e = h ? vpmfmm_last_error(h) : nullptr;
    msg += (e && *e) ? e : "error";
    switch (st) {
    case VPMFMM_BAD_ARG:
        throw std::invalid_argument(msg);
    case VPMFMM_NONFINITE_INPUT:
    case VPMFMM_DEGENERATE_DOMAIN:
    case VPMFMM_TREE_DEPTH_EXCEEDED:
    case VPMFMM_DUPLICATE_POINTS:
    case VPMFMM_OUTSIDE_BOXES:
    case VPMFMM_LEVEL1_LEAF:
        throw std::range_error(msg);
    default:
        throw std::runtime_error(msg);
    }
}

// Owning wrapper of one solver handle.
class Solver {
public:
    explicit Solver(const vpmfmm_opts *o = nullptr) {
        const int st = vpmfmm_create(o, &h_);
        if (st != VPMFMM_OK) throw_status(st, "vpmfmm_create", nullptr);
    }
    ~Solver() { vpmfmm_destroy(h_); }
    Solver(const Solver &) = delete;
    Solver &operator=(const Solver &) = delete;

    vpmfmm_handle *handle() const { return h_; }

    vpmfmm_opts opts() const {
        vpmfmm_opts o;
        throw_status(vpmfmm_get_opts(h_, &o), "vpmfmm_get_opts", h_);
        return o;
    }
    void set_opts(const vpmfmm_opts &o) { throw_status(vpmfmm_set_opts(h_, &o), "vpmfmm_set_opts", h_); }

    // Host arrays in, host arrays out (u, v include the free stream).
    void velocity(int32_t n, const double *x, const double *y, const double *s,
                  const double *omega, const uint8_t *active, double *u, double *v) {
        throw_status(vpmfmm_velocity(h_, n, x, y, s, omega, active, u, v), "vpmfmm_velocity", h_);
    }
    void field(int32_t n, const double *x, const double *y, const double *q,
               const uint8_t *active, double *ex, double *ey) {
        throw_status(vpmfmm_field(h_, n, x, y, q, active, ex, ey), "vpmfmm_field", h_);
    }
    // The type-1 FMM (u, v include the free stream).
    void velocity_type1(const vpmfmm_type1_opts &t, int32_t n, const double *x, const double *y,
                        const double *s, const double *gz, double *u, double *v) {
        throw_status(vpmfmm_velocity_type1(h_, &t, n, x, y, s, gz, u, v),
                     "vpmfmm_velocity_type1", h_);
    }
    // The 3-D treecode (u, v, w include the free stream of o).
    void velocity3d(const vpmfmm_fmm3d_opts &o, int32_t n, const double *x, const double *y,
                    const double *z, const double *s, const double *wx, const double *wy,
                    const double *wz, const uint8_t *active, double *u, double *v, double *w) {
        throw_status(vpmfmm_velocity3d(h_, &o, n, x, y, z, s, wx, wy, wz, active, u, v, w),
                     "vpmfmm_velocity3d", h_);
    }
    vpmfmm_stats stats() const {
        vpmfmm_stats s;
        throw_status(vpmfmm_last_stats(h_, &s), "vpmfmm_last_stats", h_);
        return s;
    }

private:
    vpmfmm_handle *h_ = nullptr;
};

inline bool same_opts(const vpmfmm_opts &a, const vpmfmm_opts &b) {
    return a.p == b.p && a.max_par == b.max_par && a.max_src == b.max_src &&
           a.exp_tree == b.exp_tree && a.min_level == b.min_level &&
           a.max_level == b.max_level && a.precision == b.precision &&
           a.binom_mode == b.binom_mode && a.list4_mode == b.list4_mode &&
           a.u_inf == b.u_inf && a.v_inf == b.v_inf && a.device == b.device &&
           a.timing == b.timing && a.stage_mask == b.stage_mask &&
           a.order_select == b.order_select && a.reselect_every == b.reselect_every &&
           a.judge_every == b.judge_every && a.icutoff == b.icutoff;
}

// Options of every default-constructed VelocityCalc: the reference defaults
// (P_max 10, 75/50 leaf rule, expTree 1 %, global.cpp:303-313; u_inf 1,
// v_inf 0, global.cpp:164-165) until set_default_options replaces them.  They
// are read at an object's first velocity call, so the reference's
// `VelocityCalc velocity_tool;` (main.cpp:72) and the one that
// penalization::no_slip_iterative makes per call (penalization_calc.cpp:363)
// both follow one set_default_options call made before the time loop.
inline vpmfmm_opts &default_options_slot() {
    static vpmfmm_opts o = [] {
        vpmfmm_opts d;
        vpmfmm_default_opts(&d);
        return d;
    }();
    return o;
}
inline void set_default_options(const vpmfmm_opts &o) { default_options_slot() = o; }
inline vpmfmm_opts default_options() { return default_options_slot(); }

// The branch of get_velocity (the reference's VELOCITY_2D_TYPE,
// velocity_calc.cpp:25, 170-198): 2 = the FMM branch (velocity_fmm_2d, the
// default here), 1 = the type-1 FMM (velocity_old) with the type-1 options
// (Pars::n_s, lmax, ndp, icutoff; the reference's global.cpp:316-328 unless
// given).  Read at each get_velocity call of a default-constructed object.
struct VelocityBranch {
    int type = 2;
    vpmfmm_type1_opts type1{};
};
inline VelocityBranch &default_branch_slot() {
    static VelocityBranch b = [] {
        VelocityBranch d;
        vpmfmm_type1_default_opts(&d.type1);
        return d;
    }();
    return b;
}
inline void set_default_velocity_type(int type, const vpmfmm_type1_opts *type1 = nullptr) {
    if (type != 1 && type != 2)
        throw std::invalid_argument("set_default_velocity_type: type must be 1 or 2");
    default_branch_slot().type = type;
    if (type1) default_branch_slot().type1 = *type1;
}

// Options of the 3-D branch: the reference's DIM == 3 parameters (critNum =
// src_count_max 30, expTree 1 %, centre 0, free stream (1, 0, 0):
// global.cpp:609-611, 633-635, 769, 779; SUPPORT_RADIUS_FACTOR 2, fast3DFMM.cpp:4)
// until set_default_fmm3d_options replaces them.  Read at each 3-D call.
inline vpmfmm_fmm3d_opts &default_fmm3d_slot() {
    static vpmfmm_fmm3d_opts o = [] {
        vpmfmm_fmm3d_opts d;
        vpmfmm_fmm3d_default_opts(&d);
        return d;
    }();
    return o;
}
inline void set_default_fmm3d_options(const vpmfmm_fmm3d_opts &o) { default_fmm3d_slot() = o; }

// One solver per host thread and option set, made at its first use and kept
// for the life of the process (never destroyed: the HIP runtime may already
// be gone when static destructors run).  Every VelocityCalc of that thread
// with those options shares it, so an object made per call reuses the
// device buffers, tables and learned selection of the first.
inline Solver &shared_solver(const vpmfmm_opts &o) {
    thread_local std::vector<std::pair<vpmfmm_opts, Solver *>> *cache =
        new std::vector<std::pair<vpmfmm_opts, Solver *>>();
    for (auto &e : *cache)
        if (same_opts(e.first, o)) return *e.second;
    Solver *s = new Solver(&o);
    cache->emplace_back(o, s);
    return *s;
}

// Drop-in velocity manager (FMM branch).
class VelocityCalc {
public:
    // The options of set_default_options (the reference defaults if never set).
    VelocityCalc() = default;
    explicit VelocityCalc(const vpmfmm_opts &o) : opts_(o), explicit_opts_(true) {}

    template <class ParticleT, class GridT>
    void get_velocity(ParticleT &particle, const GridT & /*base_grid*/, const int step) {
        get_velocity(particle, step);
    }

    template <class ParticleT>
    void get_velocity(ParticleT &particle, const int step) {
#if VPMFMM_DIM == 3
        velocity_fmm_3d_fast(particle, step);
#else
        (void)step;
        get_velocity_2d(particle);
#endif
    }

    // The 3-D branch (VELOCITY_3D_TYPE 1): velocity_fmm_3d_fast plus the free
    // stream (velocity_calc.cpp:201-236); every particle is a target, the
    // active ones are the sources.
    template <class ParticleT>
    void velocity_fmm_3d_fast(ParticleT &particle, const int /*step*/) {
        const int n = particle.num;
        if (n < 1) throw std::invalid_argument("VelocityCalc::get_velocity: no particles");
        if ((int)particle.x.size() < n || (int)particle.y.size() < n ||
            (int)particle.z.size() < n || (int)particle.s.size() < n ||
            (int)particle.vortx.size() < n || (int)particle.vorty.size() < n ||
            (int)particle.vortz.size() < n || (int)particle.isActive.size() < n)
            throw std::invalid_argument("VelocityCalc::get_velocity: particle arrays shorter than num");
        active_.resize(n);
        for (int i = 0; i < n; ++i) active_[i] = particle.isActive[i] ? 1 : 0;
        particle.u.clear();
        particle.v.clear();
        particle.w.clear();
        particle.u.resize(n, 0.0);
        particle.v.resize(n, 0.0);
        particle.w.resize(n, 0.0);
        solver().velocity3d(default_fmm3d_slot(), n, particle.x.data(), particle.y.data(),
                            particle.z.data(), particle.s.data(), particle.vortx.data(),
                            particle.vorty.data(), particle.vortz.data(), active_.data(),
                            particle.u.data(), particle.v.data(), particle.w.data());
    }

private:
    template <class ParticleT>
    void get_velocity_2d(ParticleT &particle) {
        const VelocityBranch br = default_branch_slot();
        if (br.type == 1) {
            velocity_old(particle, br.type1);
            return;
        }
        const int n = particle.num;
        if (n < 1) throw std::invalid_argument("VelocityCalc::get_velocity: no particles");
        if ((int)particle.x.size() < n || (int)particle.y.size() < n ||
            (int)particle.s.size() < n || (int)particle.vorticity.size() < n ||
            (int)particle.isActive.size() < n)
            throw std::invalid_argument("VelocityCalc::get_velocity: particle arrays shorter than num");
        active_.resize(n);
        for (int i = 0; i < n; ++i) active_[i] = particle.isActive[i] ? 1 : 0;
        particle.u.clear();
        particle.v.clear();
        particle.u.resize(n, 0.0);
        particle.v.resize(n, 0.0);
        solver().velocity(n, particle.x.data(), particle.y.data(), particle.s.data(),
                          particle.vorticity.data(), active_.data(), particle.u.data(),
                          particle.v.data());
    }

public:
    // VelocityCalc::velocity_old (velocity_calc.cpp:390-440) and get_velocity's
    // free stream: every particle (isActive is not read) with its size s and
    // vortex strength gz (Particle::gz) through the type-1 FMM.
    template <class ParticleT>
    void velocity_old(ParticleT &particle, const vpmfmm_type1_opts &t) {
        const int n = particle.num;
        if (n < 1) throw std::invalid_argument("VelocityCalc::get_velocity: no particles");
        if ((int)particle.x.size() < n || (int)particle.y.size() < n ||
            (int)particle.s.size() < n || (int)particle.gz.size() < n)
            throw std::invalid_argument("VelocityCalc::get_velocity: particle arrays shorter than num");
        particle.u.clear();
        particle.v.clear();
        particle.u.resize(n, 0.0);
        particle.v.resize(n, 0.0);
        solver().velocity_type1(t, n, particle.x.data(), particle.y.data(), particle.s.data(),
                                particle.gz.data(), particle.u.data(), particle.v.data());
    }

    // The reference's LSMPS Poisson test entry (velocity_calc.hpp:46) is not
    // part of this library.
    template <class ParticleT>
    void LSMPS_poisson_2d(ParticleT &, const int) {
        throw std::logic_error("VelocityCalc::LSMPS_poisson_2d is not provided by vpmfmm");
    }

    // The shared solver of this object's options (resolved at first use).
    Solver &solver() {
        if (!solver_) solver_ = &shared_solver(explicit_opts_ ? opts_ : default_options());
        return *solver_;
    }

private:
    vpmfmm_opts opts_{};
    bool explicit_opts_ = false;
    Solver *solver_ = nullptr;
    std::vector<uint8_t> active_;
};

}  // namespace vpmfmm

#ifndef VPMFMM_NO_GLOBAL_VELOCITYCALC
using vpmfmm::VelocityCalc;
#endif

#endif
