// ref_harness.cpp — drives the reference's own BoidSimulation (compiled from
// the reference sources in place, see oracle/Makefile.ref) through its public
// API: setup -> init -> update x T, dumping chosen states.  TEST
// INFRASTRUCTURE ONLY: used to pin the oracle and to write golden fixtures.
//
// usage: ref_harness OUT N T F sep_r sep_a sep_w ali_r ali_a ali_w coh_r coh_a
//                    coh_w vmax vmin INIT SEED DUMP_STEPS(comma list)
// OUT receives, for every dumped step t (state after t updates):
//   uint32 t, then N*3 f32 positions, N*3 f32 velocities, N*9 f32 get_force().
// Built twice: -DREF_GRID links src/boid_simulation_grid.cpp (boidsim_grid),
// otherwise src/boid_simulation.cpp (boidsim).
#include "boid_simulation.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#ifdef REF_GRID
// The grid engine's file init never registers boids in the cell lists
// (src/boid_simulation_grid.cpp:137-160); register them in index order the
// way its random init does (:120,:135) so update() has valid lists.
struct Harness : BoidSimulation {
    void register_all() { for (unsigned i = 0; i < N; ++i) update_list_grid(boids[i]); }
};
#else
struct Harness : BoidSimulation {
    void register_all() {}
};
#endif

int main(int argc, char** argv)
{
    if (argc != 19) {
        std::fprintf(stderr, "usage: %s OUT N T F sep_r sep_a sep_w ali_r ali_a ali_w coh_r coh_a coh_w vmax vmin INIT SEED DUMP\n", argv[0]);
        return 2;
    }
    const char* out = argv[1];
    unsigned N = (unsigned)std::strtoul(argv[2], nullptr, 10);
    unsigned T = (unsigned)std::strtoul(argv[3], nullptr, 10);
    double v[12];
    for (int k = 0; k < 12; ++k) v[k] = std::strtod(argv[4 + k], nullptr);
    std::string init = argv[16];
    int seed = std::atoi(argv[17]);
    std::set<unsigned> dump;
    {
        std::stringstream ss(argv[18]);
        std::string tok;
        while (std::getline(ss, tok, ',')) if (!tok.empty()) dump.insert((unsigned)std::strtoul(tok.c_str(), nullptr, 10));
    }
    Harness* sim = new Harness();
    sim->setup(N, v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8], v[9], v[10], v[11], init, seed);
    sim->init();
    if (init != "test" && init != "random_uniform") sim->register_all();
    std::ofstream f(out, std::ios::binary | std::ios::trunc);
    if (!f) { std::fprintf(stderr, "cannot open %s\n", out); return 1; }
    std::vector<float> buf(15 * (size_t)N);
    for (unsigned t = 0; t <= T; ++t) {
        if (dump.count(t)) {
            for (unsigned i = 0; i < N; ++i) {
                sim->get(i, &buf[3 * i], &buf[3 * i + 1], &buf[3 * i + 2]);
                sim->get_v(i, &buf[3 * N + 3 * i], &buf[3 * N + 3 * i + 1], &buf[3 * N + 3 * i + 2]);
                float* g = &buf[6 * (size_t)N + 9 * i];
                sim->get_force(i, g, g + 1, g + 2, g + 3, g + 4, g + 5, g + 6, g + 7, g + 8);
            }
            f.write((const char*)&t, 4);
            f.write((const char*)buf.data(), sizeof(float) * buf.size());
        }
        if (t < T) sim->update();
    }
    f.close();
    // no delete: the reference destructor frees new[] arrays with delete
    // (src/boid_simulation.cpp:39-46), which is undefined behaviour.
    return 0;
}
