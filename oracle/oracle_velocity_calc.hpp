/* oracle_velocity_calc.hpp -- a VelocityCalc backed by the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  It takes the place of the
 * reference's src/velocity_calculation/velocity_calc.hpp in the oracle
 * build of the reference time-loop driver (oracle/build_loop.py): the
 * reference's own loop then runs with the FMM branch of get_velocity
 * (velocity_calc.cpp:149-251, 455-578) as the oracle restates it
 * (orc_velocity: the reference's tree, lists and expansions, P_max 10, the
 * 75/50 leaf rule, its int32 binomials), so the same loop with the GPU
 * drop-in can be compared with it.  The include guard and the headers are
 * those of the replaced header, as in include/vpmfmm_reference_velocity_calc.hpp.
 */
#ifndef INCLUDED_VELOCITY_POISSON
#define INCLUDED_VELOCITY_POISSON

#include "../../Utils.hpp"
#include "../FMM/treeCell.hpp"
#include "../grid_block/gridNode.hpp"
#include "../LSMPS_poisson/poissonLSMPS.hpp"

#include <cstdint>
#include <stdexcept>
#include <vector>

#include "fmm2d_oracle.h"

class VelocityCalc {
public:
    template <class GridT>
    void get_velocity(Particle &p, const GridT &, const int step) {
        get_velocity(p, step);
    }
    void get_velocity(Particle &p, const int) {
        const int n = p.num;
        std::vector<uint8_t> a(n);
        for (int i = 0; i < n; ++i) a[i] = p.isActive[i] ? 1 : 0;
        p.u.assign(n, 0.0);
        p.v.assign(n, 0.0);
        orc_opts o;
        orc_default_opts(&o);
        const int st = orc_velocity(n, p.x.data(), p.y.data(), p.s.data(), p.vorticity.data(),
                                    a.data(), &o, Pars::u_inf, Pars::v_inf, p.u.data(),
                                    p.v.data());
        if (st) throw std::runtime_error("oracle velocity failed");
    }
};

#endif
