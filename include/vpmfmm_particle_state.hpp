// vpmfmm_particle_state.hpp -- the reference's 2-D particle-state CSV files,
// read and written (header-only C++17, no dependency on the GPU library).
//
// Format (reference src/save_data/state_data.cpp:335-392, type 0): a header
// line, then per particle
//     xp,yp,gpz,vor,up,vp,sp,active,chi[,pressure]
// with 12 significant digits (precision(12)), the velocity relative to the
// body (u - ubody, v - vbody) and the active flag as 0 / 1.  The neighbour
// columns of the LSMPS neighbour search (ngh_num, ngh_ID) are not handled.
//
// read_2d_particle follows initialization::read_2d_particle
// (src/initialization/initialization.cpp:561-691): header skipped, every
// field parsed with std::stod, column 9 kept as the pressure when asked for,
// the level of each particle from its size (:660-666).  Two departures, both
// on inputs the reference cannot handle: a size below 0.99 sigma (for which
// the reference's level loop never ends) throws std::invalid_argument, and so
// does a line with fewer than 8 fields.  chi (column 8), which the reference
// does not restore, is returned in its own vector.
#pragma once

#include <cstdint>
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace vpmfmm {

struct ParticleState2D {
    int num = 0;
    std::vector<double> x, y, s, gz, vorticity, u, v, chi, P;
    std::vector<uint8_t> isActive;
    std::vector<int> level;
};

// initialization.cpp:660-666, with the non-terminating case rejected.
inline int particle_level(double sp, int max_level, double sigma) {
    int mul = (int)((sp + sigma / 100) / sigma);
    if (mul < 1) throw std::invalid_argument("particle size below 0.99 sigma");
    int lvl = max_level;
    while (mul != 1) {
        --lvl;
        mul /= 2;
    }
    return lvl;
}

// initialization.cpp:561-691.
inline ParticleState2D read_2d_particle(const std::string &file, int max_level, double sigma,
                                        bool pressure = false) {
    std::ifstream reader(file);
    if (!reader.is_open()) throw std::runtime_error("Error: The intended file is not existed!");
    ParticleState2D p;
    std::string line;
    std::getline(reader, line);  // header
    std::vector<double> vals;
    while (reader.peek() != EOF) {
        std::getline(reader, line);
        vals.clear();
        std::string entry;
        for (char c : line) {
            if (c == ',') {
                vals.push_back(std::stod(entry));
                entry.clear();
                continue;
            }
            entry += c;
        }
        vals.push_back(std::stod(entry));
        if (vals.size() < 8) throw std::invalid_argument("particle line with fewer than 8 fields");
        const double sp = vals[6];
        p.x.push_back(vals[0]);
        p.y.push_back(vals[1]);
        p.gz.push_back(vals[2]);
        p.vorticity.push_back(vals[3]);
        p.u.push_back(vals[4]);
        p.v.push_back(vals[5]);
        p.s.push_back(sp);
        p.isActive.push_back(vals[7] != 0.0 ? 1 : 0);
        p.chi.push_back(vals.size() > 8 ? vals[8] : 0.0);
        if (pressure) {
            if (vals.size() < 10) throw std::invalid_argument("no pressure column");
            p.P.push_back(vals[9]);
        }
        p.level.push_back(particle_level(sp, max_level, sigma));
        ++p.num;
    }
    return p;
}

// state_data.cpp:335-392 (type 0, the whole domain).
inline void save_par_state_2D(const ParticleState2D &p, const std::string &file,
                              double ubody = 0.0, double vbody = 0.0, bool pressure = false) {
    std::ofstream w(file);
    if (!w.is_open()) throw std::runtime_error("cannot open " + file);
    w.precision(12);
    w << "xp,yp,gpz,vor,up,vp,sp,active,chi";
    if (pressure) w << ",pressure";
    w << "\n";
    for (int i = 0; i < p.num; ++i) {
        w << p.x[i] << "," << p.y[i] << "," << p.gz[i] << "," << p.vorticity[i] << ","
          << p.u[i] - ubody << "," << p.v[i] - vbody << "," << p.s[i] << ","
          << (p.isActive[i] ? 1 : 0) << "," << (i < (int)p.chi.size() ? p.chi[i] : 0.0);
        if (pressure) w << "," << p.P[i];
        w << "\n";
    }
}

}  // namespace vpmfmm
