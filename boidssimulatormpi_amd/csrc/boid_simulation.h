// boid_simulation.h — drop-in BoidSimulation over the MI355X engine's C ABI.
//
// Public surface of the reference class (src/boid_simulation.h:15-71): the
// same member functions with the same signatures and argument order, and
// the same public fields, so the reference's driver (src/main.cpp) and any
// code written against that class compile unchanged.  The engine variant is
// fixed when the shim is compiled, as the reference picks it at link time
// (CMakeLists.txt:85-103):
//   boidsim       -> BS_MODE_ALLPAIRS_F32REF (bit-identical to the reference)
//   boidsim_grid  -> BS_MODE_GRID_F64 (uniform grid, FP64 state)
// The environment variable BOIDS_ENGINE_MODE (allpairs_f32ref, allpairs_f64,
// grid_f64, grid_mixed) overrides the variant at run time.
//
// Not carried over: the `Boid* boids` array (the GPU owns the state; read it
// with get()/get_v()) and the protected std::list grid members.
#ifndef BOIDS_AMD_BOID_SIMULATION_H
#define BOIDS_AMD_BOID_SIMULATION_H

#include <iostream>
#include <string>
#include <vector>

#include "boids_engine.h"
#include "ptcl.h"

class BoidSimulation {
public:
    BoidSimulation();
    virtual ~BoidSimulation();
    void setup(unsigned int number_of_agents,
               double field_size,
               double separation_sight_distance,
               double separation_sight_angle,
               double separation_force_coefficient,
               double alignment_sight_distance,
               double alignment_sight_angle,
               double alignment_force_coefficient,
               double cohesion_sight_distance,
               double cohesion_sight_angle,
               double cohesion_force_coefficient,
               double velocity_max,
               double velocity_min,
               std::string initialization,
               int rand_seed);
    virtual void init();
    virtual void update();
    virtual int get(unsigned int id, float* x, float* y, float* z);
    virtual int get_v(unsigned int id, float* vx, float* vy, float* vz);
    virtual int get(unsigned int id, float* x, float* y, float* z, float* vx, float* vy, float* vz);
    bool is_openmp_enabled() const;
    int get_max_threads() const;

    int get_force(unsigned int id, float* coh_x, float* coh_y, float* coh_z, float* sep_x, float* sep_y,
                  float* sep_z, float* ali_x, float* ali_y, float* ali_z);

    unsigned int N;
    unsigned int time_step;
    double field_size;
    double field_size_X, field_size_Y, field_size_Z;
    interaction_parameters separation, alignment, cohesion;
    velocity_parameters velocity;

    unsigned int GRID_NUM;     // cells per axis of the grid variant (0: all-pairs)

    // Engine mode asked for (BS_MODE_*, BOIDS_ENGINE_MODE or the variant's
    // default), for diagnostics; a grid mode may run all-pairs on a field
    // with fewer than 3 cells per axis (GRID_NUM 0 then).
    int engine_mode() const { return mode_; }
    // Ask the engine to keep per-boid forces for get_force (call before setup()).
    void enable_force_output(bool on) { want_forces_ = on; }

protected:
    std::string initialization_type;
    int rand_seed;

private:
    void fatal(const std::string& what) const;
    void refresh() const;
    bool grid_mode() const;

    bs_engine* engine_ = nullptr;
    int mode_;
    bool want_forces_ = false;
    mutable bool dirty_ = true;
    mutable std::vector<float> pos_, vel_, frc_;
};

std::ostream& operator<<(std::ostream& stream, const BoidSimulation& boidsim);

#endif
