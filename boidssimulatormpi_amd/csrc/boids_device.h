// boids_device.h — device helpers shared by the gfx950 kernel files.
//
// Compiled with -ffp-contract=off: every multiply and add that feeds a
// neighbour decision is rounded separately, exactly as the oracle (and, in
// float mode, the reference's x86-64 build) computes it.
#pragma once

#include "boids_internal.h"

#include <math.h>

namespace boids {

typedef short v2s __attribute__((ext_vector_type(2)));   // packed int16 pair

// Ids below `me` in the ascending run ids[0, len).
__device__ __forceinline__ uint32_t count_below(const uint32_t* ids, uint32_t len, uint32_t me)
{
    uint32_t lo = 0, hi = len;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (ids[mid] < me) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint64_t mix64(uint64_t j)
{
    uint64_t z = j + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <typename R>
__device__ __forceinline__ R norm3(R x, R y, R z)
{
    R s = x * x + y * y + z * z;   // (x*x + y*y) + z*z, unfused
    return sqrt(s);
}

// Per-axis minimum image of src/boid_simulation.cpp:164-178: the float
// difference is compared in double against F/2; the shifted value is
// rounded back to R.
template <typename R>
__device__ __forceinline__ R min_image(R q, R p, double F, double half)
{
    if ((double)(q - p) > half) return (R)((double)q - F);
    if ((double)(p - q) > half) return (R)((double)q + F);
    return q;
}

template <typename R> struct Th;
template <> struct Th<float> {
    __device__ static float s(const DevParams& P, int r) { return P.s_thr_f[r]; }
    __device__ static float c(const DevParams& P, int r) { return P.c_thr_f[r]; }
    __device__ static float smax(const DevParams& P) { return P.s_max_f; }
    __device__ static float w(const DevParams& P, int r) { return P.w_f[r]; }
    __device__ static float vmax(const DevParams& P) { return P.vmax_f; }
    __device__ static float vmin(const DevParams& P) { return P.vmin_f; }
};
template <> struct Th<double> {
    __device__ static double s(const DevParams& P, int r) { return P.s_thr[r]; }
    __device__ static double c(const DevParams& P, int r) { return P.c_thr[r]; }
    __device__ static double smax(const DevParams& P) { return P.s_max; }
    __device__ static double w(const DevParams& P, int r) { return P.w[r]; }
    __device__ static double vmax(const DevParams& P) { return P.vmax; }
    __device__ static double vmin(const DevParams& P) { return P.vmin; }
};

// Phase B of one boid (src/boid_simulation.cpp:216-248).
template <typename R>
__device__ __forceinline__ void integrate(const DevParams& P, R& px, R& py, R& pz, R& vx, R& vy,
                                          R& vz, R dx, R dy, R dz)
{
    vx += dx; vy += dy; vz += dz;
    R s = norm3(vx, vy, vz);
    if ((double)s > 0. && (double)s > P.vmax) {
        R t = norm3(vx, vy, vz);
        R k = Th<R>::vmax(P);
        vx = (vx / t) * k; vy = (vy / t) * k; vz = (vz / t) * k;
    } else if ((double)s > 0. && (double)s < P.vmin) {
        R t = norm3(vx, vy, vz);
        R k = Th<R>::vmin(P);
        vx = (vx / t) * k; vy = (vy / t) * k; vz = (vz / t) * k;
    }
    px += vx; py += vy; pz += vz;
    const double F = P.F;
    if ((double)px < 0.0) px = (R)(F + (double)px); else if ((double)px > F) px = (R)((double)px - F);
    if ((double)py < 0.0) py = (R)(F + (double)py); else if ((double)py > F) py = (R)((double)py - F);
    if ((double)pz < 0.0) pz = (R)(F + (double)pz); else if ((double)pz > F) pz = (R)((double)pz - F);
}

__device__ __forceinline__ uint32_t axis_cell(double x, double inv_edge, int G)
{
    double c = floor(x * inv_edge);
    if (!(c >= 0.0)) c = 0.0;
    if (c > (double)(G - 1)) c = (double)(G - 1);
    return (uint32_t)c;
}

// Cell key: x column (sx per cell edge; sx = 1 or 2, so sx * inv_edge is
// exactly sx G / F) + sx G * (y cell + G * z cell).
__device__ __forceinline__ uint32_t cell_of(const DevParams& P, double x, double y, double z)
{
    const uint32_t G = (uint32_t)P.G, Gx = (uint32_t)P.sx * G;
    return axis_cell(x, (double)P.sx * P.inv_edge, (int)Gx) +
           Gx * (axis_cell(y, P.inv_edge, P.G) + G * axis_cell(z, P.inv_edge, P.G));
}

// Exclusive scan of one value per thread across the block (blockDim a
// multiple of 64, <= 1024): wave64 inclusive scans by lane shuffles, then a
// scan of the wave totals in LDS.  *total receives the block sum.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* wsum, uint32_t* total)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    if (wid == 0) {
        uint32_t w = lane < nw ? wsum[lane] : 0u;
        uint32_t u = w;
#pragma unroll
        for (int d = 1; d < 16; d <<= 1) {
            uint32_t y = __shfl_up(u, d, 64);
            if (lane >= d) u += y;
        }
        if (lane < nw) wsum[lane] = u - w;          // exclusive wave offsets
        if (lane == nw - 1) wsum[16] = u;            // block total
    }
    __syncthreads();
    uint32_t r = x - v + wsum[wid];
    *total = wsum[16];
    __syncthreads();
    return r;
}

__device__ __forceinline__ uint32_t flock_of(const Flocks& F, uint32_t slot)
{
    return F.flock_of ? F.flock_of[slot] : 0u;
}

__device__ __forceinline__ uint32_t to_lattice(double x, double scale)
{
    return (uint32_t)(int64_t)rint(x * scale);   // two's-complement wrap: x = F maps to 0
}

// Sub-lattice residuals: x * scale - rint(x * scale) lies in [-1/2, 1/2]
// lattice units; rounded to kSubBits fractional bits it fits a signed
// 10-bit field ([-256, 256]).  Three fields pack into the lattice record's
// spare word (x in bits 0-9, y in 10-19, z in 20-29; bit 31 marks a boid
// whose position or velocity is NaN), so a displacement
// refined to 2^-kSubBits units costs no extra bytes.  It matters where a
// term is normalised: the separation unit vector e/|e| of a close pair
// would otherwise carry the 1/2-unit rounding of both ends.
constexpr int kSubBits = 9;

__device__ __forceinline__ uint32_t sub_lattice(double x, double scale)
{
    const double t = x * scale;
    const int32_t r = (int32_t)rint((t - rint(t)) * (double)(1 << kSubBits));
    return (uint32_t)r & 0x3FFu;
}

__device__ __forceinline__ uint32_t pack_sub(double x, double y, double z, double scale)
{
    return sub_lattice(x, scale) | (sub_lattice(y, scale) << 10) | (sub_lattice(z, scale) << 20);
}

// Field a (0 = x, 1 = y, 2 = z) of pack_sub, sign-extended.
template <int A>
__device__ __forceinline__ int32_t sub_field(uint32_t w)
{
    return (int32_t)(w << (22 - 10 * A)) >> 22;
}

// Phase-B + outputs shared by the FP64 and mixed neighbour kernels.
__device__ __forceinline__ void finish_boid(const DevParams& P, uint32_t cbase, uint32_t s, uint32_t id,
                                            double px, double py, double pz, double vx, double vy,
                                            double vz, double cx_, double cy_, double cz_, double sx_,
                                            double sy_, double sz_, double ax_, double ay_,
                                            double az_, StateF64 out, uint32_t* key_out, uint32_t* rank,
                                            uint32_t* count, double* forces)
{
    const double dx = (cx_ * P.w[COH] + sx_ * P.w[SEP]) + ax_ * P.w[ALI];
    const double dy = (cy_ * P.w[COH] + sy_ * P.w[SEP]) + ay_ * P.w[ALI];
    const double dz = (cz_ * P.w[COH] + sz_ * P.w[SEP]) + az_ * P.w[ALI];
    if (forces) {
        double* f = forces + 9 * (size_t)id;
        f[0] = cx_; f[1] = cy_; f[2] = cz_; f[3] = sx_; f[4] = sy_; f[5] = sz_; f[6] = ax_; f[7] = ay_; f[8] = az_;
    }
    integrate<double>(P, px, py, pz, vx, vy, vz, dx, dy, dz);
    out.x[s] = px; out.y[s] = py; out.z[s] = pz;
    out.vx[s] = vx; out.vy[s] = vy; out.vz[s] = vz;
    out.id[s] = id;
    const uint32_t c = cbase + cell_of(P, px, py, pz);
    key_out[s] = c;
    rank[s] = atomicAdd(&count[c], 1u);   // arrival rank in the cell: the scatter's offset
}

// ---------------------------------------------------------------------------
// Neighbour kernels: chunk scheduling.
// ---------------------------------------------------------------------------
// Chunk of workgroup b, or nchunks (no work) for the grid's spare
// workgroups.  Consecutive workgroups land on different XCDs (8 on
// MI355X): each XCD gets a contiguous run of ceil(nchunks / 8) chunks, so
// neighbouring rows share its L2 and every XCD gets an eighth of the work
// (the grid is sized for the chunk table's capacity, which the actual
// chunk count stays below).  Needs gridDim >= 8 * ceil(nchunks / 8).
__device__ __forceinline__ uint32_t xcd_chunk(uint32_t b, uint32_t nchunks)
{
    const uint32_t per = (nchunks + 7u) >> 3;
    const uint32_t x = b & 7u, y = b >> 3;
    if (y >= per) return nchunks;
    const uint32_t c = x * per + y;
    return c < nchunks ? c : nchunks;
}

// The oracle's FP64 decision for the pair (i, j): the minimum image is
// taken with the reference rule from the FP64 master positions.
__device__ __forceinline__ void exact_decide(const DevParams& P, double px, double py, double pz,
                                             double vx, double vy, double vz, double vn, double qx,
                                             double qy, double qz, bool& in_s, bool& in_a, bool& in_c)
{
    qx = min_image(qx, px, P.F, P.half);
    qy = min_image(qy, py, P.F, P.half);
    qz = min_image(qz, pz, P.F, P.half);
    const double ex = px - qx, ey = py - qy, ez = pz - qz;
    const double s2 = ex * ex + ey * ey + ez * ez;
    in_c = s2 < P.s_thr[COH]; in_s = s2 < P.s_thr[SEP]; in_a = s2 < P.s_thr[ALI];
    if ((in_c && P.cone[COH]) || (in_s && P.cone[SEP]) || (in_a && P.cone[ALI])) {
        const double gx = qx - px, gy = qy - py, gz = qz - pz;
        const double dot = vx * gx + vy * gy + vz * gz;
        const double c = dot / (vn * sqrt(s2));
        if (P.cone[COH]) in_c = in_c && c >= P.c_thr[COH] && c <= 1.0;
        if (P.cone[SEP]) in_s = in_s && c >= P.c_thr[SEP] && c <= 1.0;
        if (P.cone[ALI]) in_a = in_a && c >= P.c_thr[ALI] && c <= 1.0;
    }
}


}  // namespace boids
