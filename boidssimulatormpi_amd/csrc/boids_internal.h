// boids_internal.h — types shared by the engine host code and the gfx950
// kernels.  Not part of the public ABI (see include/boids_engine.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace boids {

// Rule slots, in BoidSimulation::setup order (src/boid_simulation.h:19-33).
enum { SEP = 0, ALI = 1, COH = 2 };

constexpr int kBlock = 256;        // threads per block of the streaming kernels
constexpr int kMaxCells = 1024;    // cells per axis cap of the grid modes
// Grid modes bin x into DevParams::sx columns per cell edge (1 or 2;
// columns of edge/sx, y and z keep the cell edge >= the largest radius): a
// boid's neighbours lie within sx columns on either side, so its window is
// (2 sx + 1) columns x 3 x 3 rows -- with sx = 2, 2.5 cell edges along x
// instead of 3.  kSubXMax bounds sx.
constexpr int kSubXMax = 2;
// Neighbour kernels: a workgroup of NB threads (64, 128 or 256) owns a
// chunk of at most NB sorted boids inside one (y,z) row of cells; rows are
// cut into span groups of at most kSpanMax x columns, and a chunk covers at
// most the engine's chunk span (<= kChunkSpanMax) columns.
constexpr int kSpanMax = 62;
constexpr int kChunkSpanMax = 16;
#ifndef BS_NBR_WAVES
#define BS_NBR_WAVES 0             // >0: ask the compiler for at least this many waves per SIMD
#endif
#if BS_NBR_WAVES > 0
#define BS_NBR_OCCUPANCY __attribute__((amdgpu_waves_per_eu(BS_NBR_WAVES)))
#else
#define BS_NBR_OCCUPANCY
#endif

// Everything a kernel needs about the rules, by value in the kernarg segment.
struct DevParams {
    uint32_t n;
    int32_t G;                 // cells per axis (grid modes); x has sx * G columns
    int32_t sx;                // x columns per cell edge (grid modes): 1 or 2
    double F, half;            // field edge, F/2
    double inv_edge;           // G / F (x columns: sx * G / F)
    // double-precision exact thresholds (FP64 modes and mixed rechecks)
    double s_thr[3];           // d < r  <=>  s < s_thr   (s = squared distance, no FMA)
    double c_thr[3];           // acos(c) <= pi*a  <=>  c >= c_thr && c <= 1
    double s_max;              // max of s_thr
    double w[3];
    double vmax, vmin;
    // float exact thresholds (reference-faithful float mode)
    float s_thr_f[3], c_thr_f[3], s_max_f;
    float w_f[3];
    float vmax_f, vmin_f;
    int32_t cone[3];           // 1: the rule has a view cone (a < 1); 0: distance only
    int32_t never[3];          // 1: the cone accepts nothing (a < 0 or NaN)
    // mixed precision: positions as 32-bit fixed-point lattice coordinates
    // (unit = F / 2^32), FP32 pair tests with guard bands
    double lat_scale;          // 2^32 / F
    double unit;               // F / 2^32
    float s_lo[3], s_hi[3];    // squared lattice distance: < s_lo surely in, >= s_hi surely out
    float s_hi_max;
    float c_lo[3], c_hi[3];    // c32 <= c_lo - k/d surely out, c32 >= c_hi + k/d surely in
                               // (a rule without view cone: -inf, the cosine decides nothing)
    float c_top[3];            // c32 >= c_top: too close to 1, recheck (no view cone: +inf)
    float band_k;              // k: cosine band growth with 1/d (lattice units)
    int32_t same_cone;         // 1: c_lo, c_hi and c_top agree for the three rules (one shared cone test)
    int32_t coarse_thr;        // squared distance bound on the top 16 lattice bits
    // velocity quantum of the mixed lattice records: a power of two with
    // |v| / vq < 2^31 for |v| <= vmax, so quantised velocities are integers
    // and their sums are exact in any order
    double vq, inv_vq;
};

struct StateF64 {              // SoA, sorted (grid) or id order (all-pairs)
    double *x, *y, *z, *vx, *vy, *vz;
    uint32_t* id;
};

struct StateF32 {
    float *x, *y, *z, *vx, *vy, *vz;
};

struct Lattice {               // mixed mode companion of the sorted FP64 state
    uint4* p;                  // x, y, z lattice coordinates; w: packed sub-lattice residuals (bit 31: NaN state)
    float4* v;                 // velocity in quanta of DevParams::vq, integer-valued FP32 (w unused)
};

// One flock of a grid engine.  A grid engine simulates one or more
// independent flocks (a parameter sweep); flock f owns the boid ids/slots
// [boid_base, boid_base + P.n) and the global cell keys
// [cell_base, cell_base + sx G^3), so one counting sort over global keys keeps
// every flock contiguous.  (y,z) rows are numbered globally from row_base.
struct FlockDesc {
    DevParams P;
    uint32_t boid_base, cell_base, row_base, nrows;
    uint32_t span;             // x columns per neighbour-kernel chunk group (<= kSpanMax)
    uint32_t pad[3];
};

// Kernel view of the flock table.
struct Flocks {
    const FlockDesc* fl;       // [count] in device memory
    const uint32_t* flock_of;  // flock of each slot (nullptr: one flock)
    const uint32_t* row_flock; // flock of each (y,z) row (nullptr: one flock)
    uint32_t count;            // number of flocks
    uint32_t n;                // boids over all flocks
    uint32_t rows;             // (y,z) rows over all flocks
};

// Launchers (boids_kernels.hip).  All launch on `stream`.
hipError_t launch_allpairs_f32ref(const DevParams& P, StateF32 in, StateF32 out, float* forces,
                                  uint64_t* digest, hipStream_t stream);
hipError_t launch_allpairs_f64(const DevParams& P, StateF64 in, StateF64 out, double* forces,
                               uint64_t* digest, hipStream_t stream);
hipError_t launch_hash_count(Flocks F, StateF64 st, uint32_t* key, uint32_t* rank, uint32_t* count,
                             hipStream_t stream);
hipError_t launch_scan(uint32_t* count, uint32_t* start, uint32_t n, uint32_t* partials,
                       hipStream_t stream);
// chunk = most boids per chunk, cspan = most x columns a chunk covers,
// max_cols = x columns of the longest row (sx * G over the flocks), cap =
// entries `table` holds (a count beyond it is clamped and left in *overflow).
hipError_t launch_rows(Flocks F, const uint32_t* start, uint32_t* row_count, uint32_t* row_off,
                       uint32_t* partials, uint4* table, uint32_t* nchunks, uint32_t cap, uint32_t* overflow,
                       int chunk, int cspan, int max_cols, hipStream_t stream);
hipError_t launch_scatter(uint32_t n, const uint32_t* key, const uint32_t* start, uint32_t* count,
                          const uint32_t* id, uint32_t* perm, uint32_t* permid, hipStream_t stream);
// Scatter by the arrival ranks the counting pass recorded (no atomics),
// then sort every run of 64 slots by (cell, id) (the gather merges a cell's
// sorted parts into its in-cell id order).
hipError_t launch_scatter_ranked(uint32_t n, const uint32_t* key, const uint32_t* rank, const uint32_t* start,
                                 const uint32_t* id, uint32_t* perm, uint32_t* permid, hipStream_t stream);
// The run sort alone: every run of 64 slots of perm / permid (boid index /
// id, grouped by cell key[perm]) sorted by (cell, id) in place.  perm and
// permid may be one array where a boid's index is its id.
hipError_t launch_sort_runs(uint32_t n, const uint32_t* key, uint32_t* perm, uint32_t* permid, hipStream_t stream);
// lat (mixed) receives the lattice copy; rec (FP64 grid) the records the
// neighbour kernel fetches, three planes of n double2: (x, y), (z, vx), (vy, vz).
hipError_t launch_gather(Flocks F, const uint32_t* perm, const uint32_t* permid, const uint32_t* key,
                         const uint32_t* start, uint32_t* count, StateF64 src, StateF64 dst,
                         uint32_t* key_sorted, Lattice* lat, double2* rec, hipStream_t stream);
// Neighbour kernels (boids_nbr.hip).  nb = workgroup size (64, 128 or 256),
// tile = LDS candidates per thread and pass (5, 6, 9 or 12), max_cols = most x
// columns of a chunk's tile (columns a chunk covers + 2 sx), max_chunks =
// capacity of `table`.
hipError_t launch_nbr_f64(Flocks F, int nb, int tile, int max_cols, StateF64 in, const double2* rec,
                          StateF64 out, const uint32_t* key_sorted, const uint32_t* start, const uint4* table,
                          const uint32_t* nchunks, uint32_t max_chunks, uint32_t* key_out, uint32_t* rank,
                          uint32_t* count, double* forces, uint64_t* digest, hipStream_t stream);
hipError_t launch_nbr_mixed(Flocks F, int nb, int tile, int max_cols, StateF64 in, StateF64 out, Lattice lat,
                            const uint32_t* key_sorted, const uint32_t* start, const uint4* table,
                            const uint32_t* nchunks, uint32_t max_chunks, uint32_t* key_out, uint32_t* rank,
                            uint32_t* count, double* forces, uint64_t* digest, uint32_t* recheck,
                            hipStream_t stream);
hipError_t launch_copy(const void* src, void* dst, uint64_t bytes, hipStream_t stream);
// Reference-faithful float grid (boids_f32grid.hip): the cell list of the
// float state (keys, counting sort by (cell, id), sorted records), then the
// step from it (id order in, id order out).
hipError_t launch_f32_grid_sort(const DevParams& P, StateF32 st, uint32_t* key, uint32_t* rank, uint32_t* count,
                                uint32_t* start, uint32_t* partials, uint32_t* perm, float4* rec, float4* vrec,
                                hipStream_t stream);
hipError_t launch_f32_grid_step(const DevParams& P, StateF32 in, StateF32 out, const float4* rec,
                                const float4* vrec, const uint32_t* key, const uint32_t* start, float* forces,
                                uint64_t* digest, hipStream_t stream);
// flock metrics (boids_metrics.hip)
size_t metrics_workspace_bytes(uint32_t n, int G);
// Workspace regions are carved from `ws` (metrics_workspace_bytes long);
// *out4_dev receives the device address of the 4 result doubles.
hipError_t launch_metrics(uint32_t n, double F, int G, double eps, uint32_t min_size, const double* pos3,
                          const double* vel3, void* ws, size_t ws_bytes, double** out4_dev,
                          hipStream_t stream);
hipError_t launch_f32_to_f64(uint32_t n, StateF32 st, double* pos3, double* vel3, hipStream_t stream);
hipError_t launch_unsort_f64(uint32_t n, StateF64 src, double* pos3, double* vel3,
                             hipStream_t stream);

}  // namespace boids
