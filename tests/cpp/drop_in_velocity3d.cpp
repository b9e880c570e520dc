// Drop-in check of the 3-D branch: the reference's call
// `velocity_tool.get_velocity(particle, nodeGridList, step)` (main.cpp:410)
// compiled with DIM 3 against vpmfmm_velocity_calc.hpp, with a particle
// struct that has the reference's member names and types (Utils.hpp:77-159:
// std::vector<double> x, y, z, u, v, w, s, vortx, vorty, vortz;
// std::vector<bool> isActive; int num).
//
//   drop_in_velocity3d <n> <out.bin>
// writes n, then x, y, z, s, vortx, vorty, vortz, active (as double), u, v, w
// of the default branch (VELOCITY_3D_TYPE 1) and u, v, w of VELOCITY_3D_TYPE 2
// (n doubles each).
#define DIM 3
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "vpmfmm_velocity_calc.hpp"

struct Particle {
    std::vector<double> x, y, z, u, v, w, s, vortx, vorty, vortz, vorticity;
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
    uint64_t st = 777;
    auto rnd = [&]() {  // splitmix64 -> [0, 1)
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int i = 0; i < n; ++i) {
        p.x.push_back(-1.0 + 2.0 * rnd());
        p.y.push_back(-1.0 + 2.0 * rnd());
        p.z.push_back(-0.5 + 1.0 * rnd());
        p.s.push_back(0.02);
        p.vortx.push_back(std::sin(3.0 * p.y.back()) * 20.0);
        p.vorty.push_back(std::cos(2.0 * p.z.back()) * 20.0);
        p.vortz.push_back(std::sin(p.x.back() + p.y.back()) * 20.0);
        p.isActive.push_back(rnd() < 0.9);
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
    bad.z[0] = NAN;
    try {
        velocity_tool.get_velocity(bad, grid, 0);
        std::fprintf(stderr, "non-finite input was accepted\n");
        return 1;
    } catch (const std::range_error &) {
    }
    if ((int)p.w.size() != n) return 1;
    // VELOCITY_3D_TYPE 2 (velocity_fmm_3d: the TreeCell octree)
    std::vector<double> u1 = p.u, v1 = p.v, w1 = p.w;
    vpmfmm_fmm3d_opts o3;
    vpmfmm_fmm3d_default_opts(&o3);
    o3.type = 2;
    vpmfmm::set_default_fmm3d_options(o3);
    try {
        velocity_tool.get_velocity(p, grid, 0);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "get_velocity (type 2) failed: %s\n", e.what());
        return 1;
    }
    FILE *f = std::fopen(argv[2], "wb");
    if (!f) return 1;
    const double nn = n;
    std::fwrite(&nn, sizeof(double), 1, f);
    for (const std::vector<double> *a : {&p.x, &p.y, &p.z, &p.s, &p.vortx, &p.vorty, &p.vortz})
        std::fwrite(a->data(), sizeof(double), n, f);
    for (int i = 0; i < n; ++i) {
        const double a = p.isActive[i] ? 1.0 : 0.0;
        std::fwrite(&a, sizeof(double), 1, f);
    }
    for (const std::vector<double> *a : {&u1, &v1, &w1, &p.u, &p.v, &p.w})
        std::fwrite(a->data(), sizeof(double), n, f);
    std::fclose(f);
    std::printf("3-D drop-in: %d particles\n", n);
    return 0;
}
