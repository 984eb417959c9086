/*
 * boids_oracle.h — CPU restatement of the reference's per-timestep boid update.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (libboids_engine.so, the BoidSimulation shim and the
 * executables) never links or calls it.
 *
 * What it restates (reference files, relative to the reference root):
 *   src/vector3D.cpp:13-126        float 3-vector arithmetic (norm, normalized, dot)
 *   src/boid.cpp:27-42             Boid::isInsideArea (radius test, then view cone via acos)
 *   src/boid_simulation.cpp:79-137 init ("test", "random_uniform", file)
 *   src/boid_simulation.cpp:143-249 update(): phase A (forces) + phase B (integrate)
 *   src/boid_simulation.cpp:252-290 get/get_v/get_force
 *   src/boid_simulation_grid.cpp:48-93,383-451 GRID_NUM sizing and cell registration
 *
 * Two instantiations of one generic body (boids_oracle_step.inc):
 *   *_f32 : float state, double parameters/compares, glibc sqrtf/acosf,
 *           no FMA contraction — the reference's own arithmetic.  Pinned
 *           bit-for-bit against the compiled reference (oracle/_ref) and
 *           the golden fixtures under tests/golden/.
 *   *_f64 : the same text with every float replaced by double
 *           (sqrt/acos).  No reference exists at FP64; it is pinned by the
 *           float instantiation sharing its text, by closed-form cases and
 *           by all-pairs == cell-list brute force.
 *
 * Two candidate enumerations with identical per-pair arithmetic and
 * identical (ascending-index) summation order:
 *   BO_ENUM_ALLPAIRS : for j = 0..N-1 (src/boid_simulation.cpp:161)
 *   BO_ENUM_CELLS    : 27-cell window of a uniform grid, candidates sorted
 *                      ascending before the pair loop; bit-identical to
 *                      all-pairs whenever the grid has >= 3 cells per axis.
 */
#ifndef BOIDS_ORACLE_H
#define BOIDS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Rule order follows BoidSimulation::setup (src/boid_simulation.h:19-33):
 * separation, alignment, cohesion; each {sight_distance, sight_angle,
 * force_coefficient}; then velocity max before min. */
typedef struct {
    double field;         /* F: cubic field edge */
    double sep_r, sep_a, sep_w;
    double ali_r, ali_a, ali_w;
    double coh_r, coh_a, coh_w;
    double vmax, vmin;
} bo_params;

enum { BO_ENUM_ALLPAIRS = 0, BO_ENUM_CELLS = 1 };

/* Per-boid neighbour digest: 4 x uint64 per boid:
 *   [0] = n_coh | n_sep << 21 | n_ali << 42   (counts, 21 bits each)
 *   [1] = sum over cohesion neighbours j of bo_mix64(j)   (mod 2^64)
 *   [2] = same for separation, [3] = same for alignment.
 * Order-independent, so any enumeration order gives the same digest. */
uint64_t bo_mix64(uint64_t j);

/* One update() on n boids.  pos/vel are n*3 (x,y,z per boid), updated in
 * place.  forces (optional, n*9) receives dv_coh, dv_sep, dv_ali after
 * finalisation (src/boid_simulation.cpp:203-211), before the weights;
 * get_force multiplies by (float)w.  digest (optional, n*4) as above.
 * nthreads <= 0 means the OpenMP default.  Returns 0, or -1 on bad args. */
int bo_step_f32(const bo_params* p, uint32_t n, float* pos, float* vel,
                float* forces, uint64_t* digest, int enumeration, int nthreads);
int bo_step_f64(const bo_params* p, uint32_t n, double* pos, double* vel,
                double* forces, uint64_t* digest, int enumeration, int nthreads);

/* CPU-baseline timing helper (bench.py's cpu_baseline leg): the parts of
 * one cell-enumeration update, timed without changing the state --
 * seconds[0] = cell-list build, seconds[1] = phase A of owners 0, stride,
 * 2 stride, ..., seconds[2] = phase B of all boids (on a scratch copy).
 * Returns the number of sampled owners, or -1 (bad arguments, or a field
 * with fewer than 3 cells per axis). */
long bo_time_step_sample_f32(const bo_params* p, uint32_t n, const float* pos, const float* vel,
                             uint32_t stride, int nthreads, double* seconds);
long bo_time_step_sample_f64(const bo_params* p, uint32_t n, const double* pos, const double* vel,
                             uint32_t stride, int nthreads, double* seconds);

/* Cells per axis the oracle's cell enumeration uses (0 => all-pairs). */
int bo_cells_per_axis(const bo_params* p);

/* POSIX drand48 family restated (48-bit LCG a=0x5DEECE66D, c=0xB). */
typedef struct { uint64_t x; } bo_rng;
void   bo_srand48(bo_rng* r, long seed);   /* X = (seed & 0xffffffff) << 16 | 0x330E */
void   bo_rng_unseeded(bo_rng* r);         /* glibc's state before any srand48: X = 0 */
double bo_drand48(bo_rng* r);

/* Reference init kinds (src/boid_simulation.cpp:82-112): kind 0 = "test",
 * 1 = "random_uniform".  Draws from *r (the caller chooses the seeding:
 * the all-pairs engine never seeds drand48, the grid engine calls
 * srand48(seed), src/boid_simulation_grid.cpp:98). */
int bo_init_reference(int kind, bo_rng* r, uint32_t n, double field,
                      double vmax, double vmin, float* pos, float* vel);

/* Flock metrics for long-run statistical comparisons.  Distances use the
 * periodic metric of tools/calc_distance.pyx:6-14 (per axis
 * min(|dx|, F - |dx|)).
 *   polarisation = |sum_i v_i/|v_i|| / N   (boids with |v| = 0 add nothing)
 *   mean_nn      = mean_i min_{j != i} d(i, j)
 *   clusters     = connected components of the graph d(i, j) < eps with at
 *                  least min_size members; largest = size of the biggest.
 * Brute force O(N^2) (OpenMP). */
typedef struct {
    double polarisation, mean_nn;
    uint64_t clusters, largest;
} bo_metrics_out;
int bo_metrics(uint32_t n, double field, const double* pos, const double* vel, double eps,
               uint32_t min_size, int nthreads, bo_metrics_out* out);

#ifdef __cplusplus
}
#endif
#endif
