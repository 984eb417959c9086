/*
 * boids_oracle.c — CPU restatement of the reference boid step.
 * TEST INFRASTRUCTURE ONLY (see boids_oracle.h for scope and citations).
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off, no -ffast-math).
 */
#define _GNU_SOURCE
#include "boids_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

uint64_t bo_mix64(uint64_t j)
{
    /* splitmix64 finaliser of j+1 */
    uint64_t z = j + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* ---- uniform cell list used only to enumerate candidates ---------------- */

int bo_cells_per_axis(const bo_params* p)
{
    double rmax = p->coh_r;
    if (p->ali_r > rmax) rmax = p->ali_r;
    if (p->sep_r > rmax) rmax = p->sep_r;
    if (!(rmax > 0.0) || !(p->field > 0.0)) return 0;
    /* edge >= rmax * (1 + 1e-4): a pair that passes d < r always lies in
     * adjacent cells, with room for the rounding of the float state. */
    double g = floor(p->field / (rmax * 1.0001));
    if (g < 3.0) return 0;
    if (g > 1024.0) g = 1024.0;
    return (int)g;
}

typedef struct {
    int G;
    double edge;
    uint32_t n;
    uint32_t* cell_of;    /* n */
    uint32_t* start;      /* G^3 + 1 */
    uint32_t* items;      /* n, ascending boid index inside each cell */
} bo_cell_list;

static uint32_t bo_axis_cell(double x, double edge, int G)
{
    double c = floor(x / edge);
    if (!(c >= 0.0)) c = 0.0;           /* also NaN */
    if (c > (double)(G - 1)) c = (double)(G - 1);
    return (uint32_t)c;
}

static int bo_cell_list_build(bo_cell_list* cl, double F, int G, uint32_t n, const double* pos)
{
    size_t ncell = (size_t)G * G * G;
    cl->G = G;
    cl->edge = F / G;
    cl->n = n;
    cl->cell_of = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    cl->start = (uint32_t*)calloc(ncell + 1, sizeof(uint32_t));
    cl->items = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    if (!cl->cell_of || !cl->start || !cl->items) return -1;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t cx = bo_axis_cell(pos[3 * (size_t)i + 0], cl->edge, G);
        uint32_t cy = bo_axis_cell(pos[3 * (size_t)i + 1], cl->edge, G);
        uint32_t cz = bo_axis_cell(pos[3 * (size_t)i + 2], cl->edge, G);
        uint32_t c = cx + (uint32_t)G * (cy + (uint32_t)G * cz);
        cl->cell_of[i] = c;
        cl->start[c + 1]++;
    }
    for (size_t c = 0; c < ncell; ++c) cl->start[c + 1] += cl->start[c];
    uint32_t* fill = (uint32_t*)malloc(sizeof(uint32_t) * ncell);
    if (!fill) return -1;
    memcpy(fill, cl->start, sizeof(uint32_t) * ncell);
    for (uint32_t i = 0; i < n; ++i) cl->items[fill[cl->cell_of[i]]++] = i;   /* stable */
    free(fill);
    return 0;
}

static void bo_cell_list_free(bo_cell_list* cl)
{
    free(cl->cell_of); free(cl->start); free(cl->items);
    memset(cl, 0, sizeof(*cl));
}

static int bo_cmp_u32(const void* a, const void* b)
{
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return (x > y) - (x < y);
}

/* Candidates of boid i: every boid of the 27 cells around i's cell, in
 * ascending index order (the reference sums in ascending j).  Returns the
 * count, or UINT32_MAX on allocation failure. */
static uint32_t bo_cell_list_gather(const bo_cell_list* cl, uint32_t i, uint32_t** buf, uint32_t* cap)
{
    const int G = cl->G;
    uint32_t c = cl->cell_of[i];
    int cx = (int)(c % (uint32_t)G), cy = (int)((c / (uint32_t)G) % (uint32_t)G), cz = (int)(c / ((uint32_t)G * G));
    uint32_t total = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                int x = (cx + dx + G) % G, y = (cy + dy + G) % G, z = (cz + dz + G) % G;
                uint32_t cc = (uint32_t)x + (uint32_t)G * ((uint32_t)y + (uint32_t)G * (uint32_t)z);
                total += cl->start[cc + 1] - cl->start[cc];
            }
    if (total > *cap) {
        uint32_t ncap = total + total / 2 + 64;
        uint32_t* nb = (uint32_t*)realloc(*buf, sizeof(uint32_t) * ncap);
        if (!nb) return UINT32_MAX;
        *buf = nb; *cap = ncap;
    }
    uint32_t k = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                int x = (cx + dx + G) % G, y = (cy + dy + G) % G, z = (cz + dz + G) % G;
                uint32_t cc = (uint32_t)x + (uint32_t)G * ((uint32_t)y + (uint32_t)G * (uint32_t)z);
                for (uint32_t t = cl->start[cc]; t < cl->start[cc + 1]; ++t) (*buf)[k++] = cl->items[t];
            }
    qsort(*buf, k, sizeof(uint32_t), bo_cmp_u32);
    return k;
}

/* ---- the two instantiations -------------------------------------------- */

#define R float
#define SQRT_R sqrtf
#define ACOS_R acosf
#define FN(x) x##_f32
#include "boids_oracle_step.inc"
#undef R
#undef SQRT_R
#undef ACOS_R
#undef FN

#define R double
#define SQRT_R sqrt
#define ACOS_R acos
#define FN(x) x##_f64
#include "boids_oracle_step.inc"
#undef R
#undef SQRT_R
#undef ACOS_R
#undef FN

/* ---- drand48 and the reference inits ------------------------------------ */

void bo_srand48(bo_rng* r, long seed)
{
    r->x = (((uint64_t)(uint32_t)seed) << 16) | 0x330Eu;
}

void bo_rng_unseeded(bo_rng* r) { r->x = 0; }

double bo_drand48(bo_rng* r)
{
    r->x = (0x5DEECE66Dull * r->x + 0xBull) & ((1ull << 48) - 1);
    return (double)r->x * (1.0 / 281474976710656.0);   /* X / 2^48, exact */
}

int bo_init_reference(int kind, bo_rng* r, uint32_t n, double F,
                      double vmax, double vmin, float* pos, float* vel)
{
    if (kind != 0 && kind != 1) return -1;
    for (uint32_t i = 0; i < n; ++i) {
        float* p = pos + 3 * (size_t)i;
        float* v = vel + 3 * (size_t)i;
        if (kind == 0) {                 /* "test" (src/boid_simulation.cpp:82-99) */
            p[0] = (float)(F * 3 / 8 + bo_drand48(r) * F / 4);
            p[1] = (float)(F * 3 / 8 + bo_drand48(r) * F / 4);
            if (bo_drand48(r) > 0.5) {
                p[2] = (float)(bo_drand48(r) * F / 4 + F * 3 / 4);
            } else {
                p[2] = (float)(bo_drand48(r) * F / 4);
            }
        } else {                         /* "random_uniform" (:100-105) */
            p[0] = (float)(F * bo_drand48(r));
            p[1] = (float)(F * bo_drand48(r));
            p[2] = (float)(F * bo_drand48(r));
        }
        double s = bo_drand48(r) * (vmax - vmin) + vmin;
        double th1 = bo_drand48(r) * M_PI;
        double th2 = bo_drand48(r) * 2.0 * M_PI;
        v[0] = (float)(s * sin(th1) * cos(th2));
        v[1] = (float)(s * sin(th1) * sin(th2));
        v[2] = (float)(s * cos(th1));
    }
    return 0;
}

/* ---- flock metrics (brute force) ---------------------------------------- */

static double bo_pdist(const double* a, const double* b, double F)
{
    double s = 0.0;
    for (int k = 0; k < 3; ++k) {
        double d = fabs(a[k] - b[k]);
        if (F - d < d) d = F - d;
        s += d * d;
    }
    return sqrt(s);
}

static uint32_t bo_find(uint32_t* parent, uint32_t x)
{
    while (parent[x] != x) {
        parent[x] = parent[parent[x]];
        x = parent[x];
    }
    return x;
}

int bo_metrics(uint32_t n, double F, const double* pos, const double* vel, double eps,
               uint32_t min_size, int nthreads, bo_metrics_out* out)
{
    if (!out || (n && (!pos || !vel))) return -1;
    memset(out, 0, sizeof(*out));
    if (n == 0) return 0;
    omp_set_num_threads(nthreads > 0 ? nthreads : omp_get_num_procs());   /* 0: every host thread */
    double sx = 0, sy = 0, sz = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const double* v = vel + 3 * (size_t)i;
        double s = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (s > 0) { sx += v[0] / s; sy += v[1] / s; sz += v[2] / s; }
    }
    out->polarisation = sqrt(sx * sx + sy * sy + sz * sz) / n;
    double* nn = (double*)malloc(sizeof(double) * n);
    uint32_t* parent = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* size = (uint32_t*)calloc(n, sizeof(uint32_t));
    if (!nn || !parent || !size) { free(nn); free(parent); free(size); return -1; }
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        double best = INFINITY;
        for (uint32_t j = 0; j < n; ++j) {
            if (j == (uint32_t)ii) continue;
            double d = bo_pdist(pos + 3 * (size_t)ii, pos + 3 * (size_t)j, F);
            if (d < best) best = d;
        }
        nn[ii] = best;
    }
    double acc = 0;
    for (uint32_t i = 0; i < n; ++i) acc += nn[i];
    out->mean_nn = n > 1 ? acc / n : INFINITY;
    for (uint32_t i = 0; i < n; ++i) parent[i] = i;
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = i + 1; j < n; ++j)
            if (bo_pdist(pos + 3 * (size_t)i, pos + 3 * (size_t)j, F) < eps) {
                uint32_t a = bo_find(parent, i), b = bo_find(parent, j);
                if (a != b) parent[a > b ? a : b] = a < b ? a : b;
            }
    for (uint32_t i = 0; i < n; ++i) size[bo_find(parent, i)]++;
    for (uint32_t i = 0; i < n; ++i) {
        if (size[i] >= min_size && size[i] > 0) out->clusters++;
        if (size[i] > out->largest) out->largest = size[i];
    }
    free(nn); free(parent); free(size);
    return 0;
}
