// boid_simulation.cpp — BoidSimulation forwarded to the engine's C ABI.
// Compiled twice (see Makefile.tools): -DBS_VARIANT_GRID for boidsim_grid.
// Host C++ only; it links libboids_engine.so and sees no HIP types.
#include "boid_simulation.h"

#include <cstdlib>
#include <cstring>

#ifdef BS_VARIANT_GRID
static const int kDefaultMode = BS_MODE_GRID_F32REF;   // the reference's float arithmetic, bit for bit
#else
static const int kDefaultMode = BS_MODE_ALLPAIRS_F32REF;
#endif

static int mode_from_env(int fallback)
{
    const char* m = std::getenv("BOIDS_ENGINE_MODE");
    if (!m || !*m) return fallback;
    if (!std::strcmp(m, "allpairs_f32ref")) return BS_MODE_ALLPAIRS_F32REF;
    if (!std::strcmp(m, "allpairs_f64")) return BS_MODE_ALLPAIRS_F64;
    if (!std::strcmp(m, "grid_f64")) return BS_MODE_GRID_F64;
    if (!std::strcmp(m, "grid_mixed")) return BS_MODE_GRID_MIXED;
    if (!std::strcmp(m, "grid_f32ref")) return BS_MODE_GRID_F32REF;
    std::cerr << "unknown BOIDS_ENGINE_MODE: " << m << std::endl;
    std::exit(-1);
}

BoidSimulation::BoidSimulation() : N(0), time_step(0), field_size(0), field_size_X(0), field_size_Y(0),
                                   field_size_Z(0), separation{}, alignment{}, cohesion{}, velocity{},
                                   GRID_NUM(0), rand_seed(0), mode_(mode_from_env(kDefaultMode)) {}

BoidSimulation::~BoidSimulation() { bs_destroy(engine_); }

// Errors end the process the way the reference does (std::cerr, exit(-1)).
void BoidSimulation::fatal(const std::string& what) const
{
    std::cerr << what << ": " << bs_last_error(engine_) << std::endl;
    std::exit(-1);
}

// src/boid_simulation.cpp:46-77
void BoidSimulation::setup(unsigned int number_of_agents, double field_size, double separation_sight_distance,
                           double separation_sight_angle, double separation_force_coefficient,
                           double alignment_sight_distance, double alignment_sight_angle,
                           double alignment_force_coefficient, double cohesion_sight_distance,
                           double cohesion_sight_angle, double cohesion_force_coefficient, double velocity_max,
                           double velocity_min, std::string initialization, int rand_seed)
{
    field_size_X = field_size_Y = field_size_Z = field_size;
    this->field_size = field_size;
    separation = {separation_sight_distance, separation_sight_angle, separation_force_coefficient};
    alignment = {alignment_sight_distance, alignment_sight_angle, alignment_force_coefficient};
    cohesion = {cohesion_sight_distance, cohesion_sight_angle, cohesion_force_coefficient};
    velocity = {velocity_max, velocity_min};
    N = number_of_agents;
    initialization_type = initialization;
    this->rand_seed = rand_seed;
    time_step = 0;

    bs_params p;
    std::memset(&p, 0, sizeof(p));
    p.n = N;
    p.flags = want_forces_ ? BS_FLAG_FORCES : 0;
    p.field = field_size;
    p.sep_r = separation_sight_distance; p.sep_a = separation_sight_angle; p.sep_w = separation_force_coefficient;
    p.ali_r = alignment_sight_distance; p.ali_a = alignment_sight_angle; p.ali_w = alignment_force_coefficient;
    p.coh_r = cohesion_sight_distance; p.coh_a = cohesion_sight_angle; p.coh_w = cohesion_force_coefficient;
    p.vmax = velocity_max;
    p.vmin = velocity_min;
    p.mode = mode_;
    bs_destroy(engine_);
    engine_ = nullptr;
    if (bs_create(&p, &engine_) != BS_OK) fatal("engine setup failed");
    bs_info info;
    bs_get_info(engine_, &info);
    GRID_NUM = (unsigned)info.cells_per_axis;
    pos_.assign(3 * (size_t)N, 0.f);
    vel_.assign(3 * (size_t)N, 0.f);
    frc_.assign(9 * (size_t)N, 0.f);
    if (grid_mode())
        std::cout << "# Grid List Optimization enabled" << std::endl
                  << "  grid num:" << GRID_NUM << std::endl;
}

// The grid variants behave as the reference's grid engine (srand48(seed) at
// init, src/boid_simulation_grid.cpp:98), whichever enumeration the engine
// falls back to on a small field.
bool BoidSimulation::grid_mode() const
{
    return mode_ == BS_MODE_GRID_F64 || mode_ == BS_MODE_GRID_MIXED || mode_ == BS_MODE_GRID_F32REF;
}

// src/boid_simulation.cpp:79-137 (all-pairs: drand48 never seeded) and
// src/boid_simulation_grid.cpp:95-161 (grid: srand48(seed)).
void BoidSimulation::init()
{
    std::vector<float> st(6 * (size_t)N);
    const bool grid = grid_mode();
    int rc;
    if (initialization_type == "test" || initialization_type == "random_uniform") {
        const int kind = initialization_type == "test" ? BS_INIT_TEST : BS_INIT_RANDOM_UNIFORM;
        rc = bs_generate_init(kind, grid ? 1 : 0, rand_seed, N, field_size, velocity.max, velocity.min, st.data());
    } else {
        rc = bs_read_init_file(initialization_type.c_str(), N, st.data());
        if (rc != BS_OK) {
            std::cerr << "Invalid initial placement type or file name: " << initialization_type << std::endl;
            std::exit(-1);
        }
    }
    if (rc != BS_OK) fatal("init failed");
    if (bs_set_state_f32(engine_, st.data()) != BS_OK) fatal("init failed");
    dirty_ = true;
}

void BoidSimulation::update()
{
    if (bs_step(engine_, 1) != BS_OK) fatal("update failed");
    dirty_ = true;
}

// One bulk device-to-host copy per state; the per-id getters then read it.
void BoidSimulation::refresh() const
{
    if (!dirty_) return;
    if (bs_get_state_f32(engine_, pos_.data(), vel_.data()) != BS_OK) fatal("state read failed");
    if (want_forces_ && bs_get_forces_f32(engine_, frc_.data()) != BS_OK) fatal("force read failed");
    dirty_ = false;
}

int BoidSimulation::get(unsigned int id, float* x, float* y, float* z)
{
    refresh();
    *x = pos_[3 * (size_t)id]; *y = pos_[3 * (size_t)id + 1]; *z = pos_[3 * (size_t)id + 2];
    return 0;
}

int BoidSimulation::get_v(unsigned int id, float* vx, float* vy, float* vz)
{
    refresh();
    *vx = vel_[3 * (size_t)id]; *vy = vel_[3 * (size_t)id + 1]; *vz = vel_[3 * (size_t)id + 2];
    return 0;
}

int BoidSimulation::get(unsigned int id, float* x, float* y, float* z, float* vx, float* vy, float* vz)
{
    get(id, x, y, z);
    get_v(id, vx, vy, vz);
    return 0;
}

// src/boid_simulation.cpp:272-290: (float)w * dv of the last update.
int BoidSimulation::get_force(unsigned int id, float* coh_x, float* coh_y, float* coh_z, float* sep_x,
                              float* sep_y, float* sep_z, float* ali_x, float* ali_y, float* ali_z)
{
    refresh();
    const float* f = &frc_[9 * (size_t)id];
    const float wc = (float)cohesion.force_coefficient, ws = (float)separation.force_coefficient,
                wa = (float)alignment.force_coefficient;
    *coh_x = wc * f[0]; *coh_y = wc * f[1]; *coh_z = wc * f[2];
    *sep_x = ws * f[3]; *sep_y = ws * f[4]; *sep_z = ws * f[5];
    *ali_x = wa * f[6]; *ali_y = wa * f[7]; *ali_z = wa * f[8];
    return 0;
}

bool BoidSimulation::is_openmp_enabled() const { return false; }

int BoidSimulation::get_max_threads() const { return -1; }

std::ostream& operator<<(std::ostream& stream, const BoidSimulation& b)
{
    stream << "N: " << b.N << std::endl
           << "field size: " << b.field_size_X << "," << b.field_size_Y << "," << b.field_size_Z << std::endl
           << "#Separation" << std::endl
           << "force: " << b.separation.force_coefficient << std::endl
           << "area distance: " << b.separation.sight_distance << std::endl
           << "area angle: " << b.separation.sight_agnle << std::endl
           << "#Alignment" << std::endl
           << "force: " << b.alignment.force_coefficient << std::endl
           << "area distance: " << b.alignment.sight_distance << std::endl
           << "area angle: " << b.alignment.sight_agnle << std::endl
           << "#Cohesion" << std::endl
           << "force: " << b.cohesion.force_coefficient << std::endl
           << "area distance: " << b.cohesion.sight_distance << std::endl
           << "area angle: " << b.cohesion.sight_agnle << std::endl
           << "#Velocity" << std::endl
           << "min: " << b.velocity.min << std::endl
           << "max: " << b.velocity.max << std::endl
           << "#Grid List Optimization: " << b.GRID_NUM << std::endl
           << "#OpenMP: Disabled" << std::endl;
    return stream;
}
