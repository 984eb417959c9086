/*
 * boids_engine.h — C ABI of the MI355X boid-step engine (libboids_engine.so).
 *
 * The reference's engine boundary is the C++ class BoidSimulation
 * (src/boid_simulation.h:15-71).  Its public surface is re-exposed on top of
 * this ABI by boidssimulatormpi_amd/csrc/boid_simulation.{h,cpp}; each entry
 * point below names the reference member it replaces.  Plain pointers and
 * sizes only; the engine owns its device buffers and one HIP stream.
 *
 * Conventions
 *   - Every int-returning call returns 0 (BS_OK) or a negative BS_E* code;
 *     bs_last_error() then describes the failure.  (The reference prints to
 *     std::cerr and calls exit(-1); the C++ shim maps errors back to that.)
 *   - One host thread per engine; calls on one engine must not overlap.
 *   - Host arrays are in boid-id order (the order get(id, ...) uses).
 *   - Modes: see BS_MODE_*.  All GPU work is hand-written HIP for gfx950.
 */
#ifndef BOIDS_ENGINE_H
#define BOIDS_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bs_engine bs_engine;

enum {
    BS_OK = 0,
    BS_E_ARG = -1,       /* invalid argument */
    BS_E_HIP = -2,       /* HIP runtime error (no device, launch failure, ...) */
    BS_E_NOMEM = -3,     /* device or host allocation failed */
    BS_E_STATE = -4,     /* call not valid in the engine's current state */
    BS_E_MATH = -5       /* host libm breaks an assumption the exact thresholds need */
};

/* Arithmetic / enumeration variants of the step. */
enum {
    /* float state, reference arithmetic, ascending-j sums: bit-identical to
     * the reference boidsim (src/boid_simulation.cpp:143-249). */
    BS_MODE_ALLPAIRS_F32REF = 0,
    /* the same algorithm promoted to double (all-pairs). */
    BS_MODE_ALLPAIRS_F64 = 1,
    /* double, uniform grid of cells with edge >= the largest radius (x
     * binned at half the edge), every cell within reach searched
     * (boidsim_grid's role, src/boid_simulation_grid.cpp:165-293).  Falls
     * back to all-pairs when the field holds fewer than 3 cells per axis. */
    BS_MODE_GRID_F64 = 2,
    /* mixed precision grid: FP64 master state and position update, FP32
     * pair tests on exact 32-bit lattice differences, FP64 recheck of pairs
     * inside a guard band around every threshold, force sums of
     * integer-valued terms in FP64 (exact, so independent of order). */
    BS_MODE_GRID_MIXED = 3,
    /* float state, the reference's arithmetic, uniform grid of cells with
     * edge >= the largest radius; every owner visits the in-range boids of
     * its 27 cells in ascending id, so the float sums are the reference's:
     * bit-identical to boidsim_grid (src/boid_simulation_grid.cpp:165-293),
     * which equals boidsim.  Falls back to ALLPAIRS_F32REF (the same bits)
     * below 3 cells per axis. */
    BS_MODE_GRID_F32REF = 4
};

/* Init kinds (src/boid_simulation.cpp:79-137 for the first two). */
enum {
    BS_INIT_TEST = 0,            /* "test" */
    BS_INIT_RANDOM_UNIFORM = 1,  /* "random_uniform" (theta1 ~ U[0,pi], as the reference) */
    BS_INIT_ISOTROPIC = 2,       /* uniform positions, isotropic directions, |v| ~ U[vmin,vmax] */
    BS_INIT_CLUSTERED = 3        /* 64 wrapped Gaussian blobs (sigma 0.05 F), per-blob heading + N(0,0.2) jitter */
};

/* Creation flags. */
enum {
    BS_FLAG_FORCES = 1,   /* keep per-boid dv_coh/dv_sep/dv_ali for get_forces (get_force) */
    BS_FLAG_DIGEST = 2,   /* keep per-boid neighbour digests (parity checks) */
    BS_FLAG_TIMING = 4,   /* record HIP events around every kernel launch */
    BS_FLAG_GRAPH = 8     /* grid modes: capture one update as a HIP graph on the
                           * first bs_step and replay it (timing then covers
                           * whole updates only, as BS_K_STEP) */
};

/* Parameters of BoidSimulation::setup (src/boid_simulation.h:19-33), same
 * order: separation, alignment, cohesion {distance, angle, coefficient},
 * velocity max before min.  Angles are fractions of pi (a >= 1: distance
 * test only, src/boid.cpp:29-35). */
typedef struct {
    uint32_t n;
    uint32_t flags;              /* BS_FLAG_* */
    double field;
    double sep_r, sep_a, sep_w;
    double ali_r, ali_a, ali_w;
    double coh_r, coh_a, coh_w;
    double vmax, vmin;
    int32_t mode;                /* BS_MODE_* */
    int32_t cells_per_axis;      /* grid modes: 0 = automatic */
    int32_t device;              /* HIP device ordinal */
    int32_t block;               /* grid modes: neighbour-kernel workgroup size, 64, 128 or 256; 0 = automatic (64) */
    int32_t tile;                /* grid modes: LDS tile depth, candidates staged per thread and pass:
                                  * 5, 6, 9 or 12; 0 = automatic (5) */
    int32_t fill;                /* grid modes: most boids per chunk, percent of the workgroup size
                                  * (25..100); 0 = automatic (100, less for small flocks: about 2048
                                  * chunks at least) */
} bs_params;

/* Read-only facts about a created engine. */
typedef struct {
    int32_t mode;                /* effective mode (grid modes may fall back to all-pairs) */
    int32_t cells_per_axis;      /* 0 for all-pairs */
    double cell_edge;
    uint64_t device_bytes;       /* device memory the engine holds */
    uint64_t steps;              /* updates done since the last set_state/init */
    int32_t block;               /* neighbour-kernel workgroup size (grid modes), else 0 */
    int32_t tile;                /* LDS tile depth (grid modes), else 0 */
    int32_t fill;                /* chunk size limit in percent of block (grid modes), else 0 */
    int32_t reserved;
} bs_info;

/* Kernel ids for bs_get_timing. */
enum {
    BS_K_PAIR = 0,       /* fused neighbour + integrate kernel (all modes) */
    BS_K_SCAN = 1,       /* cell-count scan + block table */
    BS_K_SCATTER = 2,    /* counting-sort scatter */
    BS_K_GATHER = 3,     /* in-cell rank + state gather */
    BS_K_OTHER = 4,
    BS_K_STEP = 5,       /* whole graph-replayed updates (BS_FLAG_GRAPH) */
    BS_K_COUNT = 6
};

/* ctor + setup() (src/boid_simulation.h:17-33).  *out is NULL on failure;
 * the message is then available from bs_last_error(NULL). */
int bs_create(const bs_params* p, bs_engine** out);
/* Parameter sweep: `count` independent flocks in one engine.  The
 * reference runs a sweep as one process per parameter set
 * (settings/parameter_search_vel_2/gen.py:32-42 writes the .ini files,
 * tools/gen_jobscript.py:37-65 one job per file); here every flock keeps its
 * own n, field, rules and velocity limits, and one launch per kernel steps
 * them all.  Grid modes only; mode, flags and device must agree; every
 * flock needs >= 3 cells per axis.  Boids are numbered flock after flock:
 * flock k owns ids [boid_base, boid_base + n) of every state, force and
 * digest array (bs_get_flock_info). */
int bs_create_batch(const bs_params* flocks, uint32_t count, bs_engine** out);
typedef struct {
    uint32_t boid_base, n;
    int32_t cells_per_axis;
    uint32_t cell_base;          /* first global cell key of the flock */
    double field, cell_edge;
} bs_flock_info;
int bs_get_flock_count(const bs_engine* e, uint32_t* count);
int bs_get_flock_info(const bs_engine* e, uint32_t k, bs_flock_info* out);

/* ~BoidSimulation() */
void bs_destroy(bs_engine* e);
/* Message of the last failed call on e (e == NULL: last bs_create failure). */
const char* bs_last_error(const bs_engine* e);

/* Host-side init generator (init(), src/boid_simulation.cpp:79-112).
 * Writes n*6 floats in the reference's init-file layout (x y z vx vy vz per
 * boid).  seeded = 0 reproduces the all-pairs engine (drand48 never
 * seeded); seeded = 1 calls srand48(seed) first as the grid engine does
 * (src/boid_simulation_grid.cpp:98).  Needs no GPU. */
int bs_generate_init(int32_t kind, int32_t seeded, int32_t seed, uint32_t n, double field,
                     double vmax, double vmin, float* xyzv6);

/* init() for a file: n*6 float32 little-endian, no header
 * (src/boid_simulation.cpp:113-136).  Returns BS_E_ARG if the file is
 * missing or short. */
int bs_read_init_file(const char* path, uint32_t n, float* xyzv6);

/* Upload a state (id order).  f32: n*6 interleaved as the init file;
 * f64: n*3 positions and n*3 velocities. */
int bs_set_state_f32(bs_engine* e, const float* xyzv6);
int bs_set_state_f64(bs_engine* e, const double* pos3, const double* vel3);

/* update() x nsteps (src/boid_simulation.cpp:143-249), stream-ordered. */
int bs_step(bs_engine* e, uint32_t nsteps);
/* Wait for queued work. */
int bs_sync(bs_engine* e);

/* get()/get_v() for all ids at once (src/boid_simulation.cpp:252-270). */
int bs_get_state_f32(bs_engine* e, float* pos3, float* vel3);
int bs_get_state_f64(bs_engine* e, double* pos3, double* vel3);
/* get_force() for all ids (src/boid_simulation.cpp:272-290), before the
 * weights: n*9 = dv_coh xyz, dv_sep xyz, dv_ali xyz of the last update
 * (zeros before the first).  Needs BS_FLAG_FORCES. */
int bs_get_forces_f32(bs_engine* e, float* f9);
int bs_get_forces_f64(bs_engine* e, double* f9);
/* Neighbour digests of the last update: n*4 uint64 per boid,
 * [0] = n_coh | n_sep << 21 | n_ali << 42, [1..3] = sum over neighbours j
 * of mix64(j) for cohesion, separation, alignment (mix64 = splitmix64
 * finaliser of j + 0x9E3779B97F4A7C15).  j is the neighbour's id within its
 * flock (in a batch, id - boid_base), so a flock's digests equal those of
 * the same flock in an engine of its own.  Needs BS_FLAG_DIGEST. */
int bs_get_digest(bs_engine* e, uint64_t* d4);

int bs_get_info(const bs_engine* e, bs_info* out);
/* Accumulated HIP-event time of one kernel class since the last reset
 * (BS_FLAG_TIMING).  ms = total milliseconds, launches = launch count. */
int bs_get_timing(bs_engine* e, int32_t kernel, double* ms, uint64_t* launches);
int bs_reset_timing(bs_engine* e);

/* Exact decision thresholds the kernels use in place of sqrt/acos
 * (exposed for tests).  out[0..2] = squared-distance thresholds for
 * sep, ali, coh; out[3..5] = cosine thresholds (2.0 = never, -2.0 = no
 * cone test) for the state precision of the given mode. */
int bs_debug_thresholds(const bs_params* p, double* out6);

/* Mixed mode: number of candidate pairs decided by the FP64 recheck since
 * creation (guard-band hits).  0 for other modes. */
int bs_debug_rechecks(bs_engine* e, uint64_t* out);

/* Grid built by the last counting sort (grid modes), for exactness tests of
 * the build that replaces the reference's cell-list registration
 * (update_list_grid / clear_list_grid, src/boid_simulation_grid.cpp:383-451).
 * key_sorted[n]: global cell key of every sorted slot; id_sorted[n]: the
 * boid id in that slot; start[ncells + 1]: first slot of every cell;
 * table4[4 * nchunks]: the neighbour kernels' chunk table, entries (y,z row
 * within the flock, first slot, end slot, flock).  Any array may be NULL;
 * info is filled in every case.  Cell keys: x column (x_columns_per_cell
 * columns per cell edge) + columns * (y cell + G * z cell), plus the flock's
 * cell_base. */
typedef struct {
    uint32_t n, ncells, nchunks, max_chunks;
    uint32_t overflow;           /* chunk count beyond max_chunks (0: none; else results are invalid) */
    int32_t x_columns_per_cell;  /* x columns per cell edge (1 or 2) */
    int32_t chunk_boids;         /* most boids per chunk */
    int32_t chunk_span;          /* most x columns a chunk covers */
} bs_grid_info;
int bs_debug_grid(bs_engine* e, uint32_t* key_sorted, uint32_t* id_sorted, uint32_t* start, uint32_t* table4,
                  bs_grid_info* info);

/* Flock metrics of the current state, for long-run statistical comparison.
 * Periodic distance as tools/calc_distance.pyx:6-14.  polarisation =
 * |sum v/|v|| / N; mean_nn = mean nearest-neighbour distance; clusters =
 * connected components of the graph d < eps with >= min_size members;
 * largest = size of the biggest component.  No reference counterpart in the
 * engine (the reference computes these offline in tools/). */
typedef struct {
    double polarisation, mean_nn;
    uint64_t clusters, largest;
} bs_metrics_out;
int bs_metrics(bs_engine* e, double eps, uint32_t min_size, bs_metrics_out* out);
/* The same for flock k of a batch (bs_metrics fails on a batch of more than one flock). */
int bs_metrics_flock(bs_engine* e, uint32_t flock, double eps, uint32_t min_size, bs_metrics_out* out);
/* bs_metrics_flock without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (analysis inside tight
 * stepping loops, harnesses that forbid hipMalloc).  Its workspace is a
 * static device array of 64 MiB on every GPU the library runs on (held
 * until the process ends) instead of temporary memory; a flock whose
 * workspace exceeds it is rejected with BS_E_ARG (the workspace takes at
 * most ~128 bytes per boid: flocks of up to ~5 x 10^5 boids always fit).
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * race.  Results equal bs_metrics_flock's. */
int bs_metrics_flock_noalloc(bs_engine* e, uint32_t flock, double eps, uint32_t min_size, bs_metrics_out* out);

/* Measurement helper: `reps` launches of a plain streaming copy of `bytes`
 * bytes on the current device (reference point for HBM byte counters).
 * *ms_per_copy receives the average HIP-event time per copy. */
int bs_debug_copy(uint64_t bytes, int32_t reps, double* ms_per_copy);

/* is_openmp_enabled()/get_max_threads() have no device counterpart; the
 * shim answers them.  Library version string. */
const char* bs_version(void);

#ifdef __cplusplus
}
#endif
#endif
