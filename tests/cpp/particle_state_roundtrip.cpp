// Reads a particle-state CSV with include/vpmfmm_particle_state.hpp, writes it
// back, and prints each particle's level (one per line).
//   particle_state_roundtrip IN.csv OUT.csv MAX_LEVEL SIGMA
// Exit status 2 and the exception text on stderr when reading fails.
#include <cstdio>
#include <cstdlib>
#include <exception>

#include "vpmfmm_particle_state.hpp"

int main(int argc, char **argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s IN.csv OUT.csv MAX_LEVEL SIGMA\n", argv[0]);
        return 1;
    }
    try {
        auto p = vpmfmm::read_2d_particle(argv[1], std::atoi(argv[3]), std::atof(argv[4]));
        vpmfmm::save_par_state_2D(p, argv[2]);
        for (int l : p.level) std::printf("%d\n", l);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 2;
    }
    return 0;
}
