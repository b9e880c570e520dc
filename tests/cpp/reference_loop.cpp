// reference_loop.cpp -- the reference's own time loop around whichever
// VelocityCalc the build's velocity_calc.hpp provides (the GPU drop-in, or
// the oracle for comparison; oracle/build_loop.py builds both).
//
// It restates the order of main.cpp (main.cpp:51-429) without the two
// unconditional exits (:137, :276) and the interactive prompt (:262-275):
// bodies (N_BODY = 0 in the reference's configuration, so no penalization
// or force step is compiled, main.cpp:339-343), the initial particles and
// the Perlman vorticity with its velocity solution (opt_init_vorticity 1),
// the initial remeshing and neighbour set-up (:134, :139), then per step:
// advection (:373), diffusion (:376), remeshing every rmsh_inv steps
// (:395-397), the domain-boundary update (:403) and the velocity
// (get_velocity, :410).  The reference's sources are compiled unchanged;
// only velocity_calc.hpp is replaced.
//
//   reference_loop <steps> <out.bin>
// writes the particle count, then x, y, s, vorticity, u, v (double each)
// after the given number of steps.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "global.hpp"
#include "Utils.hpp"
#include "src/grid_block/gridNode.hpp"
#include "src/geometry/geometry.hpp"
#include "src/initialization/initialization.hpp"
#include "src/remeshing/remeshing.hpp"
#include "src/advection/advection.hpp"
#include "src/diffusion/diffusion.hpp"
#include "src/velocity_calculation/velocity_calc.hpp"

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s steps out.bin\n", argv[0]);
        return 2;
    }
    const int steps = std::atoi(argv[1]);
    std::vector<Body> bodyList(N_BODY);
    Particle particle;
    GridNode nodeGridList;
    geometry geom_tool;
    initialization init_tool;
    remeshing remesh_tool;
    VelocityCalc velocity_tool;
    advection advection_tool;
    diffusion diffusion_tool;

    geom_tool.generateBody(bodyList);
    init_tool.initialize_particle(particle, bodyList, nodeGridList);
    init_tool.initialize_vorticity(particle);
    // main.cpp remeshes before its neighbour set-up (:134, :139); the first
    // remeshing's adaptation then reads the empty neighbour lists
    // (gridNodeAdapt.cpp:1201) and faults, so the lists are set up first
    remesh_tool.set_neighbor(particle, nodeGridList);
    remesh_tool.get_remeshing(particle, nodeGridList, 0, bodyList);
    remesh_tool.set_neighbor(particle, nodeGridList);
    for (int step = 0; step < steps; ++step) {
        advection_tool.main_advection(particle);
        diffusion_tool.main_diffusion(particle);
        if ((step % Pars::rmsh_inv) == 0)
            remesh_tool.get_remeshing(particle, nodeGridList, step, bodyList);
        init_tool.update_domain_boundary(particle);
        velocity_tool.get_velocity(particle, nodeGridList, step);
        std::fprintf(stderr, "reference_loop: step %d done, %d particles\n", step, particle.num);
    }
    FILE *f = std::fopen(argv[2], "wb");
    if (!f) return 1;
    const double n = particle.num;
    std::fwrite(&n, 8, 1, f);
    for (const std::vector<double> *v : {&particle.x, &particle.y, &particle.s,
                                         &particle.vorticity, &particle.u, &particle.v})
        std::fwrite(v->data(), 8, (size_t)particle.num, f);
    std::fclose(f);
    return 0;
}
