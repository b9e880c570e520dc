// Drop-in check: the reference's call `velocity_tool.get_velocity(particle,
// nodeGridList, step)` (main.cpp:410) against vpmfmm_velocity_calc.hpp, with a
// particle struct that has the reference's member names and types
// (Utils.hpp:77-159: std::vector<double> x, y, s, vorticity, u, v;
// std::vector<bool> isActive; int num).
//
//   drop_in_velocity <n> <out.bin>
// writes n, then x, y, s, omega, active (as double), u, v, and the type-1
// branch's gz, u, v (n doubles each).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "vpmfmm_velocity_calc.hpp"

struct Particle {
    std::vector<double> x, y, u, v, gz, s, vorticity;
    std::vector<bool> isActive;
    int num = 0;
};
struct GridNode {};

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s n out.bin\n", argv[0]);
        return 2;
    }
    const int n = std::atoi(argv[1]);
    Particle p;
    p.num = n;
    uint64_t st = 12345;
    auto rnd = [&]() {  // splitmix64 -> [0, 1)
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int i = 0; i < n; ++i) {
        p.x.push_back(-2.0 + 4.0 * rnd());
        p.y.push_back(-1.0 + 2.0 * rnd());
        p.s.push_back(0.01);
        p.vorticity.push_back(std::sin(3.0 * p.x.back()) * std::cos(2.0 * p.y.back()) * 50.0);
        p.isActive.push_back(rnd() < 0.9);
        p.gz.push_back(p.vorticity.back() * p.s.back() * p.s.back());
    }
    GridNode grid;
    VelocityCalc velocity_tool;
    try {
        velocity_tool.get_velocity(p, grid, 0);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "get_velocity failed: %s\n", e.what());
        return 1;
    }
    // invalid input must surface as the reference's exception type
    Particle bad = p;
    bad.x[0] = NAN;
    bool threw = false;
    try {
        velocity_tool.get_velocity(bad, 0);
    } catch (const std::range_error &) {
        threw = true;
    }
    if (!threw) {
        std::fprintf(stderr, "non-finite input did not raise std::range_error\n");
        return 1;
    }
    // a VelocityCalc made per call, as penalization::no_slip_iterative does
    // (penalization_calc.cpp:363, 402), shares the first one's solver (and
    // its device buffers) and gives the same velocity
    {
        VelocityCalc per_call;
        Particle q = p;
        per_call.get_velocity(q, 0);
        if (&per_call.solver() != &velocity_tool.solver()) {
            std::fprintf(stderr, "a per-call VelocityCalc made a solver of its own\n");
            return 1;
        }
        for (int i = 0; i < n; ++i)
            if (q.u[i] != p.u[i] || q.v[i] != p.v[i]) {
                std::fprintf(stderr, "per-call velocity differs at %d\n", i);
                return 1;
            }
        vpmfmm_opts o;
        vpmfmm_default_opts(&o);
        o.p = 12;
        VelocityCalc other(o);
        if (&other.solver() == &velocity_tool.solver()) {
            std::fprintf(stderr, "different options share a solver\n");
            return 1;
        }
    }
    // the type-1 branch (VELOCITY_2D_TYPE 1, velocity_old): every particle,
    // the reference's type-1 options, gz as the strength
    Particle t1 = p;
    vpmfmm::set_default_velocity_type(1);
    try {
        VelocityCalc old_tool;
        old_tool.get_velocity(t1, grid, 0);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "type-1 get_velocity failed: %s\n", e.what());
        return 1;
    }
    vpmfmm::set_default_velocity_type(2);
    FILE *f = std::fopen(argv[2], "wb");
    if (!f) return 1;
    const double nd = n;
    std::fwrite(&nd, 8, 1, f);
    std::vector<double> a(n);
    for (int i = 0; i < n; ++i) a[i] = p.isActive[i] ? 1.0 : 0.0;
    for (const std::vector<double> *v : {&p.x, &p.y, &p.s, &p.vorticity, &a, &p.u, &p.v,
                                         &p.gz, &t1.u, &t1.v})
        std::fwrite(v->data(), 8, n, f);
    std::fclose(f);
    std::printf("drop_in_velocity: n=%d u[0]=%.17g v[0]=%.17g\n", n, p.u[0], p.v[0]);
    return 0;
}
