// boids_kernels.hip — gfx950 kernels of the boid step.
//
// Compiled with -ffp-contract=off: every multiply and add that feeds a
// neighbour decision is rounded separately, exactly as the oracle (and, in
// float mode, the reference's x86-64 build) computes it.  Multiply-adds are
// written explicitly (fmaf) only where a result feeds no exact decision.
//
// Per-pair semantics (all modes) follow Boid::isInsideArea
// (src/boid.cpp:27-42) and the phase-A loop of src/boid_simulation.cpp:153-212;
// sqrt/acos are replaced by exact thresholds computed on the host
// (boids_engine.hip: thresholds()), so the kernels compare
//   s < s_thr        <=>  sqrt(s) < r                 (sqrt monotone, correctly rounded)
//   c >= c_thr, c<=1 <=>  acos(c) <= pi*a              (acos monotone; NaN/|c|>1 reject)
#include "boids_internal.h"

#include <math.h>

namespace boids {

typedef short v2s __attribute__((ext_vector_type(2)));   // packed int16 pair

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

__device__ __forceinline__ uint32_t cell_of(const DevParams& P, double x, double y, double z)
{
    uint32_t G = (uint32_t)P.G;
    return axis_cell(x, P.inv_edge, P.G) + G * (axis_cell(y, P.inv_edge, P.G) + G * axis_cell(z, P.inv_edge, P.G));
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

// ---------------------------------------------------------------------------
// All-pairs step: one thread per boid i, candidates j = 0..n-1 ascending,
// staged through LDS tiles.  Same per-pair arithmetic and summation order
// as the reference (src/boid_simulation.cpp:153-248), fused with phase B.
// R = float: bit-identical to boidsim.  R = double: the FP64 promotion.
// ---------------------------------------------------------------------------
template <typename R, typename St>
__global__ __launch_bounds__(kBlock) void k_allpairs(DevParams P, St in, St out, R* forces,
                                                     uint64_t* digest)
{
    __shared__ R tx[kBlock], ty[kBlock], tz[kBlock], tvx[kBlock], tvy[kBlock], tvz[kBlock];
    const uint32_t n = P.n;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const bool active = i < n;
    R px = 0, py = 0, pz = 0, vx = 0, vy = 0, vz = 0;
    if (active) { px = in.x[i]; py = in.y[i]; pz = in.z[i]; vx = in.vx[i]; vy = in.vy[i]; vz = in.vz[i]; }
    const R vn = norm3(vx, vy, vz);
    const R smax = Th<R>::smax(P);
    const R s_c = Th<R>::s(P, COH), s_s = Th<R>::s(P, SEP), s_a = Th<R>::s(P, ALI);
    const R c_c = Th<R>::c(P, COH), c_s = Th<R>::c(P, SEP), c_a = Th<R>::c(P, ALI);
    const bool k_c = P.cone[COH], k_s = P.cone[SEP], k_a = P.cone[ALI];
    R scx = 0, scy = 0, scz = 0, ssx = 0, ssy = 0, ssz = 0, sax = 0, say = 0, saz = 0;
    int nc = 0, ns = 0, na = 0;
    uint64_t hc = 0, hs = 0, ha = 0;

    for (uint32_t base = 0; base < n; base += kBlock) {
        __syncthreads();
        const uint32_t j = base + threadIdx.x;
        if (j < n) {
            tx[threadIdx.x] = in.x[j]; ty[threadIdx.x] = in.y[j]; tz[threadIdx.x] = in.z[j];
            tvx[threadIdx.x] = in.vx[j]; tvy[threadIdx.x] = in.vy[j]; tvz[threadIdx.x] = in.vz[j];
        }
        __syncthreads();
        const uint32_t m = min((uint32_t)kBlock, n - base);
        if (!active) continue;
        for (uint32_t k = 0; k < m; ++k) {
            const R qx = min_image(tx[k], px, P.F, P.half);
            const R qy = min_image(ty[k], py, P.F, P.half);
            const R qz = min_image(tz[k], pz, P.F, P.half);
            const R ex = px - qx, ey = py - qy, ez = pz - qz;
            const R s = ex * ex + ey * ey + ez * ez;
            if (!(s < smax)) continue;
            if (base + k == i) continue;
            bool in_c = s < s_c, in_s = s < s_s, in_a = s < s_a;
            if ((in_c && k_c) || (in_s && k_s) || (in_a && k_a)) {
                const R gx = qx - px, gy = qy - py, gz = qz - pz;
                const R dot = vx * gx + vy * gy + vz * gz;
                const R c = dot / (vn * sqrt(s));
                if (k_c) in_c = in_c && c >= c_c && c <= (R)1;
                if (k_s) in_s = in_s && c >= c_s && c <= (R)1;
                if (k_a) in_a = in_a && c >= c_a && c <= (R)1;
            }
            if (in_c) { nc++; scx += qx; scy += qy; scz += qz; hc += mix64(base + k); }
            if (in_s) {
                const R t = sqrt(s);
                ns++; ssx += ex / t; ssy += ey / t; ssz += ez / t; hs += mix64(base + k);
            }
            if (in_a) { na++; sax += tvx[k]; say += tvy[k]; saz += tvz[k]; ha += mix64(base + k); }
        }
    }
    if (!active) return;
    if (nc != 0) { const R k = (R)nc; scx = scx / k - px; scy = scy / k - py; scz = scz / k - pz; }
    if (na != 0) { const R k = (R)na; sax = sax / k - vx; say = say / k - vy; saz = saz / k - vz; }
    const R wc = Th<R>::w(P, COH), ws = Th<R>::w(P, SEP), wa = Th<R>::w(P, ALI);
    const R dx = (scx * wc + ssx * ws) + sax * wa;
    const R dy = (scy * wc + ssy * ws) + say * wa;
    const R dz = (scz * wc + ssz * ws) + saz * wa;
    if (forces) {
        R* f = forces + 9 * (size_t)i;
        f[0] = scx; f[1] = scy; f[2] = scz; f[3] = ssx; f[4] = ssy; f[5] = ssz; f[6] = sax; f[7] = say; f[8] = saz;
    }
    if (digest) {
        uint64_t* d = digest + 4 * (size_t)i;
        d[0] = (uint64_t)nc | ((uint64_t)ns << 21) | ((uint64_t)na << 42);
        d[1] = hc; d[2] = hs; d[3] = ha;
    }
    integrate<R>(P, px, py, pz, vx, vy, vz, dx, dy, dz);
    out.x[i] = px; out.y[i] = py; out.z[i] = pz; out.vx[i] = vx; out.vy[i] = vy; out.vz[i] = vz;
}

// ---------------------------------------------------------------------------
// Grid build: cell hash + count, exclusive scan, scatter, in-cell rank +
// gather.  Sorted order is (cell, boid id), independent of history.  Keys
// are global over the flock table (flock cell_base + local cell).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t flock_of(const Flocks& F, uint32_t slot)
{
    return F.flock_of ? F.flock_of[slot] : 0u;
}

__global__ __launch_bounds__(kBlock) void k_hash_count(Flocks F, StateF64 st, uint32_t* key,
                                                        uint32_t* count)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= F.n) return;
    const FlockDesc& fd = F.fl[flock_of(F, i)];
    const uint32_t c = fd.cell_base + cell_of(fd.P, st.x[i], st.y[i], st.z[i]);
    key[i] = c;
    atomicAdd(&count[c], 1u);
}

constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

// Tile-local exclusive scan of count[] into start[]; tile totals into
// partials[]; count[] is cleared for reuse as the scatter's fill counter.
__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(uint32_t* count, uint32_t* start,
                                                              uint32_t m, uint32_t* partials)
{
    __shared__ uint32_t wsum[17];
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t idx = base + k;
        v[k] = idx < m ? count[idx] : 0u;
        sum += v[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(sum, wsum, &total);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t idx = base + k;
        if (idx < m) { start[idx] = run; count[idx] = 0u; }
        run += v[k];
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_partials(uint32_t* partials, uint32_t nt)
{
    __shared__ uint32_t wsum[17];
    uint32_t carry = 0;
    for (uint32_t b = 0; b < nt; b += kScanThreads) {
        const uint32_t idx = b + threadIdx.x;
        const uint32_t v = idx < nt ? partials[idx] : 0u;
        uint32_t total;
        const uint32_t e = block_exclusive_scan(v, wsum, &total);
        if (idx < nt) partials[idx] = e + carry;
        carry += total;
    }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(uint32_t* start, uint32_t m,
                                                            const uint32_t* partials)
{
    const uint32_t add = partials[blockIdx.x];
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t idx = base + k;
        if (idx < m) start[idx] += add;
    }
}

// Block table of the neighbour kernels: every (y,z) row of cells of every
// flock is split into ceil(n_row/kNbr) chunks of consecutive sorted boids
// of near-equal size, so that no block spans two rows.  Entries are
// (local row, first slot, end slot, flock).  One workgroup.
__global__ __launch_bounds__(kScanThreads) void k_rows(Flocks F, const uint32_t* start,
                                                        uint4* table, uint32_t* nchunks)
{
    __shared__ uint32_t wsum[17];
    uint32_t carry = 0;
    for (uint32_t b = 0; b < F.rows; b += kScanThreads) {
        const uint32_t r = b + threadIdx.x;
        uint32_t lo = 0, nr = 0, f = 0, rl = 0;
        if (r < F.rows) {
            uint32_t a = 0, z = F.count - 1;              // last flock with row_base <= r
            while (a < z) {
                const uint32_t m = (a + z + 1) >> 1;
                if (F.fl[m].row_base <= r) a = m; else z = m - 1;
            }
            f = a;
            const FlockDesc& fd = F.fl[f];
            const uint32_t G = (uint32_t)fd.P.G;
            rl = r - fd.row_base;
            const uint32_t c0 = fd.cell_base + rl * G;
            lo = start[c0];
            nr = start[c0 + G] - lo;
        }
        const uint32_t kr = (nr + kNbr - 1) / kNbr;
        uint32_t total;
        const uint32_t off = block_exclusive_scan(kr, wsum, &total) + carry;
        if (kr) {
            const uint32_t csz = (nr + kr - 1) / kr;
            for (uint32_t k = 0; k < kr; ++k) {
                const uint32_t t0 = lo + k * csz;
                const uint32_t t1 = min(lo + nr, t0 + csz);
                table[off + k] = make_uint4(rl, t0, t1, f);
            }
        }
        carry += total;
    }
    if (threadIdx.x == 0) *nchunks = carry;
}

__global__ __launch_bounds__(kBlock) void k_scatter(uint32_t n, const uint32_t* key,
                                                    const uint32_t* start, uint32_t* count,
                                                    const uint32_t* id, uint32_t* perm,
                                                    uint32_t* permid)
{
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const uint32_t c = key[s];
    const uint32_t t = start[c] + atomicAdd(&count[c], 1u);
    perm[t] = s;
    permid[t] = id[s];
}

__device__ __forceinline__ uint32_t to_lattice(double x, double scale)
{
    return (uint32_t)(int64_t)rint(x * scale);   // two's-complement wrap: x = F maps to 0
}

// Final slot = cell start + rank of the boid id among its cell's ids, so
// the order inside a cell is ascending id whatever the atomics did.
__global__ __launch_bounds__(kBlock) void k_gather(Flocks F, const uint32_t* perm,
                                                   const uint32_t* permid, const uint32_t* key,
                                                   const uint32_t* start, uint32_t* count,
                                                   StateF64 src, StateF64 dst,
                                                   uint32_t* key_sorted, Lattice lat, int with_lat)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= F.n) return;
    const uint32_t s = perm[t];
    const uint32_t me = permid[t];
    const uint32_t c = key[s];
    const uint32_t lo = start[c], hi = start[c + 1];
    uint32_t rank = 0;
    for (uint32_t u = lo; u < hi; ++u) rank += permid[u] < me ? 1u : 0u;
    const uint32_t f = lo + rank;
    const double x = src.x[s], y = src.y[s], z = src.z[s];
    const double vx = src.vx[s], vy = src.vy[s], vz = src.vz[s];
    dst.x[f] = x; dst.y[f] = y; dst.z[f] = z;
    dst.vx[f] = vx; dst.vy[f] = vy; dst.vz[f] = vz;
    dst.id[f] = me;
    key_sorted[f] = c;
    if (rank == 0) count[c] = 0u;
    if (with_lat) {
        const double scale = F.fl[flock_of(F, f)].P.lat_scale;   // slots stay inside their flock
        lat.p[f] = make_uint4(to_lattice(x, scale), to_lattice(y, scale), to_lattice(z, scale), f);
        lat.vx[f] = (float)vx; lat.vy[f] = (float)vy; lat.vz[f] = (float)vz;
    }
}

__global__ __launch_bounds__(kBlock) void k_unsort_f64(uint32_t n, StateF64 src, double* pos3,
                                                       double* vel3)
{
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const size_t i = src.id[s];
    pos3[3 * i] = src.x[s]; pos3[3 * i + 1] = src.y[s]; pos3[3 * i + 2] = src.z[s];
    vel3[3 * i] = src.vx[s]; vel3[3 * i + 1] = src.vy[s]; vel3[3 * i + 2] = src.vz[s];
}

// ---------------------------------------------------------------------------
// Neighbour pass shared machinery.  A block owns a chunk [t0,t1) of sorted
// boids inside one (y,z) row of cells, spanning cells cxa..cxb.  For each of
// the 9 neighbour rows it stages the x-segment of cells cxa-1..cxb+1
// (periodically wrapped) in LDS; each boid then scans its own 3-cell window
// cx-1..cx+1 of that segment.  Segments larger than the LDS capacity are
// processed in passes.
// ---------------------------------------------------------------------------
// Segment tables live in dynamic LDS sized for G + 2 cells (segment_bytes).
struct Segment {
    uint32_t* segc;    // [G+3] LDS-index prefix over segment cells
    uint32_t* segg;    // [G+2] global sorted start of each segment cell
    int8_t* segsh;     // [G+2] x periodic shift of each segment cell: -1, 0, +1 (x F)
    uint32_t* wsum;    // [17] scan scratch
};

__host__ __device__ constexpr size_t segment_bytes(int G)
{
    return ((size_t)(G + 3) * 4 + (size_t)(G + 2) * 4 + (size_t)(G + 2) + 15) & ~(size_t)15;
}

__device__ __forceinline__ Segment segment_view(unsigned char* dyn, int G, uint32_t* wsum)
{
    Segment S;
    S.segc = (uint32_t*)dyn;
    S.segg = S.segc + (G + 3);
    S.segsh = (int8_t*)(S.segg + (G + 2));
    S.wsum = wsum;
    return S;
}

// Builds the segment of neighbour row `nrow` (of the flock whose cells start
// at global key cbase) for cells cxa-1 .. cxa+nseg-2.
__device__ __forceinline__ void build_segment(Segment& S, const uint32_t* start, uint32_t cbase, int G,
                                              int cxa, int nseg, uint32_t nrow)
{
    // counts, then a block-wide exclusive scan (each thread owns up to
    // ceil(nseg/kNbr) consecutive cells)
    const int per = (nseg + kNbr - 1) / kNbr;
    const int k0 = threadIdx.x * per;
    uint32_t local = 0;
    for (int k = k0; k < min(nseg, k0 + per); ++k) {
        int cx = cxa - 1 + k;
        int sh = 0;
        if (cx < 0) { cx += G; sh = -1; } else if (cx >= G) { cx -= G; sh = 1; }
        const uint32_t c = cbase + (uint32_t)cx + (uint32_t)G * nrow;
        const uint32_t a = start[c], b = start[c + 1];
        S.segg[k] = a;
        S.segsh[k] = (int8_t)sh;
        S.segc[k] = b - a;           // temporarily the count
        local += b - a;
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(local, S.wsum, &total);
    for (int k = k0; k < min(nseg, k0 + per); ++k) {
        const uint32_t cnt = S.segc[k];
        S.segc[k] = run;
        run += cnt;
    }
    if (threadIdx.x == 0) S.segc[nseg] = total;
    __syncthreads();
}

// Segment cell holding segment index g (largest k with segc[k] <= g).
__device__ __forceinline__ int seg_cell(const Segment& S, int nseg, uint32_t g)
{
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (S.segc[mid] <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Phase-B + outputs shared by the FP64 and mixed neighbour kernels.
__device__ __forceinline__ void finish_boid(const DevParams& P, uint32_t cbase, uint32_t s, uint32_t id,
                                            double px, double py, double pz, double vx, double vy,
                                            double vz, double cx_, double cy_, double cz_, double sx_,
                                            double sy_, double sz_, double ax_, double ay_,
                                            double az_, StateF64 out, uint32_t* key_out,
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
    atomicAdd(&count[c], 1u);
}

// ---------------------------------------------------------------------------
// FP64 plain 27-cell kernel.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t grid)
{
    // consecutive workgroups land on different XCDs (8 on MI355X); give
    // each XCD a contiguous run of chunks so neighbouring rows share its L2.
    const uint32_t per = grid >> 3;
    return (b & 7u) * per + (b >> 3);
}

template <bool DIGEST>
__global__ __launch_bounds__(kNbr) BS_NBR_OCCUPANCY void k_nbr_f64(Flocks FL, StateF64 in, StateF64 out,
                                                  const uint32_t* __restrict__ key_sorted,
                                                  const uint32_t* __restrict__ start,
                                                  const uint4* __restrict__ table,
                                                  const uint32_t* __restrict__ nchunks,
                                                  uint32_t* key_out, uint32_t* count,
                                                  double* forces, uint64_t* digest)
{
    __shared__ double sx[kCapF64], sy[kCapF64], sz[kCapF64], svx[kCapF64], svy[kCapF64], svz[kCapF64];
    __shared__ uint32_t sid[DIGEST ? kCapF64 : 1];
    __shared__ uint16_t sq[kQueue * kNbr];   // per-lane queues, lane-interleaved
    __shared__ uint32_t wsum[17];
    extern __shared__ __align__(16) unsigned char dyn[];

    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
    if (b >= *nchunks) return;
    const uint4 ch = table[b];
    const FlockDesc& fd = FL.fl[ch.w];
    const DevParams& P = fd.P;
    const uint32_t cbase = fd.cell_base;
    const int G = P.G;
    Segment S = segment_view(dyn, G, wsum);
    const uint32_t row = ch.x, t0 = ch.y, t1 = ch.z;
    const int cy = (int)(row % (uint32_t)G), cz = (int)(row / (uint32_t)G);
    const uint32_t s = t0 + threadIdx.x;
    const bool active = s < t1;
    const int cxa = (int)((key_sorted[t0] - cbase) % (uint32_t)G);
    const int cxb = (int)((key_sorted[t1 - 1] - cbase) % (uint32_t)G);
    const int nseg = cxb - cxa + 3;

    double px = 0, py = 0, pz = 0, vx = 0, vy = 0, vz = 0;
    uint32_t myid = 0, mykey = 0, mystart = 0;
    int L = 1;
    if (active) {
        px = in.x[s]; py = in.y[s]; pz = in.z[s];
        vx = in.vx[s]; vy = in.vy[s]; vz = in.vz[s];
        myid = in.id[s];
        mykey = key_sorted[s];
        mystart = start[mykey];
        L = (int)((mykey - cbase) % (uint32_t)G) - cxa + 1;
    }
    const double vn = norm3(vx, vy, vz);
    const double F = P.F;
    // phase-1 bound: a fused s2 differs from the unfused one by < 1e-15 relative
    const double s_filter = P.s_max * (1.0 + 1e-12);
    double scx = 0, scy = 0, scz = 0, ssx = 0, ssy = 0, ssz = 0, sax = 0, say = 0, saz = 0;
    int nc = 0, ns = 0, na = 0;
    uint64_t hc = 0, hs = 0, ha = 0;
    const bool k_c = P.cone[COH], k_s = P.cone[SEP], k_a = P.cone[ALI];

    for (int rr = 0; rr < 9; ++rr) {
        const int dz = rr / 3 - 1, dy = rr % 3 - 1;
        int ny = cy + dy, nz = cz + dz;
        double shy = 0.0, shz = 0.0;
        if (ny < 0) { ny += G; shy = -F; } else if (ny >= G) { ny -= G; shy = F; }
        if (nz < 0) { nz += G; shz = -F; } else if (nz >= G) { nz -= G; shz = F; }
        const uint32_t nrow = (uint32_t)ny + (uint32_t)G * (uint32_t)nz;
        __syncthreads();                      // previous readers of S and the tile are done
        build_segment(S, start, cbase, G, cxa, nseg, nrow);
        const uint32_t total = S.segc[nseg];
        const uint32_t wlo = S.segc[L - 1], whi = S.segc[L + 2];
        const uint32_t selfk = (rr == 4) ? S.segc[L] + (s - mystart) : 0xFFFFFFFFu;
        for (uint32_t base = 0; base < total; base += kCapF64) {
            const uint32_t here = min((uint32_t)kCapF64, total - base);
            __syncthreads();
            for (uint32_t idx = threadIdx.x; idx < here; idx += kNbr) {
                const uint32_t g = base + idx;
                const int k = seg_cell(S, nseg, g);
                const uint32_t slot = S.segg[k] + (g - S.segc[k]);
                const int shx = S.segsh[k];
                const double x = in.x[slot], y = in.y[slot], z = in.z[slot];
                sx[idx] = shx == 0 ? x : (shx < 0 ? x - F : x + F);
                sy[idx] = shy == 0.0 ? y : y + shy;
                sz[idx] = shz == 0.0 ? z : z + shz;
                svx[idx] = in.vx[slot]; svy[idx] = in.vy[slot]; svz[idx] = in.vz[slot];
                if (DIGEST) sid[idx] = in.id[slot];
            }
            __syncthreads();
            if (!active) continue;
            // Full rule evaluation of one candidate inside the largest radius.
            // The decisions are the oracle's expressions; the sums are
            // branch-free (adding +0.0 leaves a sum unchanged) so that
            // consecutive pairs overlap.
            auto pair = [&](uint32_t kk) {
                    const double qx = sx[kk], qy = sy[kk], qz = sz[kk];
                    const double wx = svx[kk], wy = svy[kk], wz = svz[kk];
                    const double ex = px - qx, ey = py - qy, ez = pz - qz;
                    const double s2 = ex * ex + ey * ey + ez * ez;
                    const double gx = qx - px, gy = qy - py, gz = qz - pz;
                    const double dot = vx * gx + vy * gy + vz * gz;
                    const double t = sqrt(s2);
                    const double c = dot / (vn * t);
                    const bool cin = c <= 1.0;
                    const bool in_c = s2 < P.s_thr[COH] && (!k_c || (c >= P.c_thr[COH] && cin));
                    const bool in_s = s2 < P.s_thr[SEP] && (!k_s || (c >= P.c_thr[SEP] && cin));
                    const bool in_a = s2 < P.s_thr[ALI] && (!k_a || (c >= P.c_thr[ALI] && cin));
                    nc += in_c; scx += in_c ? qx : 0.0; scy += in_c ? qy : 0.0; scz += in_c ? qz : 0.0;
                    na += in_a; sax += in_a ? wx : 0.0; say += in_a ? wy : 0.0; saz += in_a ? wz : 0.0;
                    if (in_s) {
                        ns++; ssx += ex / t; ssy += ey / t; ssz += ez / t;
                    }
                    if (DIGEST) {
                        const uint64_t h = mix64(sid[kk]);
                        hc += in_c ? h : 0; hs += in_s ? h : 0; ha += in_a ? h : 0;
                    }
            };
            // Phase 1: squared distance only (any rounding may be used here:
            // the filter keeps a superset, the pair body re-decides exactly).
            const uint32_t kb = max(wlo, base), ke = min(whi, base + here);
#if BS_NBR_QUEUE
            // Candidates inside the largest radius go to this lane's queue,
            // drained (phase 2, densely packed) when any lane fills up.
            // Four candidates per iteration keep four LDS->distance chains in
            // flight; a slot is always written and kept only on a hit.
            auto drain = [&](uint32_t qn) {
                uint32_t j = 0;
                for (; j + 2 <= qn; j += 2) {
                    const uint32_t k0 = sq[j * kNbr + threadIdx.x], k1 = sq[(j + 1) * kNbr + threadIdx.x];
                    pair(k0);
                    pair(k1);
                }
                if (j < qn) pair(sq[j * kNbr + threadIdx.x]);
            };
            auto near = [&](uint32_t k) {
                const uint32_t kk = k - base;
                const double ex = px - sx[kk], ey = py - sy[kk], ez = pz - sz[kk];
                const double s2 = fma(ex, ex, fma(ey, ey, ez * ez));
                return s2 < s_filter && k != selfk;
            };
            uint32_t qn = 0;
            uint32_t k = kb;
            for (; k + 4 <= ke; k += 4) {
                if (__any(qn > (uint32_t)kQueue - 4)) { drain(qn); qn = 0; }
                const bool h0 = near(k), h1 = near(k + 1), h2 = near(k + 2), h3 = near(k + 3);
                const uint32_t kk = k - base;
                sq[qn * kNbr + threadIdx.x] = (uint16_t)kk; qn += h0;
                sq[qn * kNbr + threadIdx.x] = (uint16_t)(kk + 1); qn += h1;
                sq[qn * kNbr + threadIdx.x] = (uint16_t)(kk + 2); qn += h2;
                sq[qn * kNbr + threadIdx.x] = (uint16_t)(kk + 3); qn += h3;
            }
            for (; k < ke; ++k) {
                if (__any(qn == (uint32_t)kQueue)) { drain(qn); qn = 0; }
                sq[qn * kNbr + threadIdx.x] = (uint16_t)(k - base);
                qn += near(k);
            }
            drain(qn);
#else
            for (uint32_t k = kb; k < ke; ++k) {
                const uint32_t kk = k - base;
                const double ex = px - sx[kk], ey = py - sy[kk], ez = pz - sz[kk];
                const double s2 = fma(ex, ex, fma(ey, ey, ez * ez));
                if (s2 < s_filter && k != selfk) pair(kk);
            }
#endif
        }
    }
    if (!active) return;
    if (nc != 0) { const double k = (double)nc; scx = scx / k - px; scy = scy / k - py; scz = scz / k - pz; }
    if (na != 0) { const double k = (double)na; sax = sax / k - vx; say = say / k - vy; saz = saz / k - vz; }
    if (DIGEST) {
        uint64_t* d = digest + 4 * (size_t)myid;
        d[0] = (uint64_t)nc | ((uint64_t)ns << 21) | ((uint64_t)na << 42);
        d[1] = hc; d[2] = hs; d[3] = ha;
    }
    finish_boid(P, cbase, s, myid, px, py, pz, vx, vy, vz, scx, scy, scz, ssx, ssy, ssz, sax, say, saz,
                out, key_out, count, forces);
}

// ---------------------------------------------------------------------------
// Mixed-precision 27-cell kernel.  Candidates are staged as 32-bit lattice
// coordinates (unit F/2^32): the difference of two lattice coordinates,
// read as int32, is the minimum-image displacement exactly, so the pair
// test is FP32 arithmetic on exact integers.  Each decision is taken in
// FP32 when it is outside a guard band around its threshold and re-taken
// with the oracle's FP64 expression otherwise.  Force sums are FP32 with
// Kahan compensation; phase B runs in FP64.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void kahan_add(float& s, float& c, float x)
{
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
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

// FP32 verdict of one rule: 1 = in, 0 = out, -1 = inside a guard band.
__device__ __forceinline__ int rule_verdict(const DevParams& P, int r, float s, float c, float band)
{
    if (!(s < P.s_hi[r])) return 0;
    int v = (s < P.s_lo[r]) ? 1 : -1;
    if (P.cone[r]) {
        if (!(c < P.c_top)) return -1;
        if (c >= P.c_hi[r] + band) return v;
        if (c <= P.c_lo[r] - band) return 0;
        return -1;
    }
    return v;
}

template <bool DIGEST>
__global__ __launch_bounds__(kNbr) BS_NBR_OCCUPANCY void k_nbr_mixed(Flocks FL, StateF64 in, StateF64 out,
                                                    Lattice lat,
                                                    const uint32_t* __restrict__ key_sorted,
                                                    const uint32_t* __restrict__ start,
                                                    const uint4* __restrict__ table,
                                                    const uint32_t* __restrict__ nchunks,
                                                    uint32_t* key_out, uint32_t* count,
                                                    double* forces, uint64_t* digest,
                                                    uint32_t* recheck)
{
    __shared__ uint4 sl[kCapMixed];          // lattice x, y, z, sorted slot
    __shared__ uint2 sc[kCapMixed];          // top 16 bits of x|y, z: the coarse lattice
    __shared__ float svx[kCapMixed], svy[kCapMixed], svz[kCapMixed];
    __shared__ uint32_t sid[DIGEST ? kCapMixed : 1];
    __shared__ uint16_t sq[kQueue * kNbr];
    __shared__ uint32_t wsum[17];
    extern __shared__ __align__(16) unsigned char dyn[];

    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
    if (b >= *nchunks) return;
    const uint4 ch = table[b];
    const FlockDesc& fd = FL.fl[ch.w];
    const DevParams& P = fd.P;
    const uint32_t cbase = fd.cell_base;
    const int G = P.G;
    Segment S = segment_view(dyn, G, wsum);
    const uint32_t row = ch.x, t0 = ch.y, t1 = ch.z;
    const int cy = (int)(row % (uint32_t)G), cz = (int)(row / (uint32_t)G);
    const uint32_t s = t0 + threadIdx.x;
    const bool active = s < t1;
    const int cxa = (int)((key_sorted[t0] - cbase) % (uint32_t)G);
    const int cxb = (int)((key_sorted[t1 - 1] - cbase) % (uint32_t)G);
    const int nseg = cxb - cxa + 3;

    double px = 0, py = 0, pz = 0, vx = 0, vy = 0, vz = 0;
    uint32_t myid = 0, mystart = 0;
    uint4 me = make_uint4(0, 0, 0, 0xFFFFFFFFu);
    float fvx = 0.f, fvy = 0.f, fvz = 0.f;
    int L = 1;
    if (active) {
        px = in.x[s]; py = in.y[s]; pz = in.z[s];
        vx = in.vx[s]; vy = in.vy[s]; vz = in.vz[s];
        myid = in.id[s];
        me = lat.p[s];
        fvx = lat.vx[s]; fvy = lat.vy[s]; fvz = lat.vz[s];
        const uint32_t mykey = key_sorted[s];
        mystart = start[mykey];
        L = (int)((mykey - cbase) % (uint32_t)G) - cxa + 1;
    }
    const v2s my_xy = __builtin_bit_cast(v2s, (me.x >> 16) | (me.y & 0xFFFF0000u));
    const v2s my_z = __builtin_bit_cast(v2s, me.z >> 16);
    const double vn = norm3(vx, vy, vz);
    const float fvv = fvx * fvx + fvy * fvy + fvz * fvz;
    // a vanishing FP32 velocity makes the FP32 cosine meaningless: recheck
    const bool vbad = !(fvv > 1e-30f);
    const float inv_vn = vbad ? 0.f : __builtin_amdgcn_rsqf(fvv);
    // cohesion: exact integer sum of lattice displacements; separation and
    // alignment: FP32 with Kahan compensation
    int64_t scx = 0, scy = 0, scz = 0;
    float ssx = 0, ssy = 0, ssz = 0, csx = 0, csy = 0, csz = 0;
    float sax = 0, say = 0, saz = 0, cax = 0, cay = 0, caz = 0;
    int nc = 0, ns = 0, na = 0;
    uint32_t nre = 0;
    uint64_t hc = 0, hs = 0, ha = 0;

    for (int rr = 0; rr < 9; ++rr) {
        const int dz = rr / 3 - 1, dy = rr % 3 - 1;
        int ny = cy + dy, nz = cz + dz;
        if (ny < 0) ny += G; else if (ny >= G) ny -= G;
        if (nz < 0) nz += G; else if (nz >= G) nz -= G;
        const uint32_t nrow = (uint32_t)ny + (uint32_t)G * (uint32_t)nz;
        __syncthreads();
        build_segment(S, start, cbase, G, cxa, nseg, nrow);
        const uint32_t total = S.segc[nseg];
        const uint32_t wlo = S.segc[L - 1], whi = S.segc[L + 2];
        const uint32_t selfk = (rr == 4) ? S.segc[L] + (s - mystart) : 0xFFFFFFFFu;
        for (uint32_t base = 0; base < total; base += kCapMixed) {
            const uint32_t here = min((uint32_t)kCapMixed, total - base);
            __syncthreads();
            for (uint32_t idx = threadIdx.x; idx < here; idx += kNbr) {
                const uint32_t g = base + idx;
                const int k = seg_cell(S, nseg, g);
                const uint32_t slot = S.segg[k] + (g - S.segc[k]);
                const uint4 p4 = lat.p[slot];
                sl[idx] = p4;
                sc[idx] = make_uint2((p4.x >> 16) | (p4.y & 0xFFFF0000u), p4.z >> 16);
                svx[idx] = lat.vx[slot]; svy[idx] = lat.vy[slot]; svz[idx] = lat.vz[slot];
                if (DIGEST) sid[idx] = in.id[slot];
            }
            __syncthreads();
            if (!active) continue;
            // FP32 rule evaluation of one coarse-filtered candidate.
            auto pair = [&](uint32_t kk) {
                const uint4 q = sl[kk];
                const float wx = svx[kk], wy = svy[kk], wz = svz[kk];
                const int32_t ix = (int32_t)(q.x - me.x), iy = (int32_t)(q.y - me.y), iz = (int32_t)(q.z - me.z);
                const float gx = (float)ix, gy = (float)iy, gz = (float)iz;
                const float s2 = fmaf(gz, gz, fmaf(gy, gy, gx * gx));
                const float inv_d = __builtin_amdgcn_rsqf(s2);
                const float c = fmaf(fvz, gz, fmaf(fvy, gy, fvx * gx)) * inv_d * inv_vn;
                const float band = P.band_k * inv_d;
                const int vc = rule_verdict(P, COH, s2, c, band);
                const int vs = rule_verdict(P, SEP, s2, c, band);
                const int va = rule_verdict(P, ALI, s2, c, band);
                bool in_c = vc > 0, in_s = vs > 0, in_a = va > 0;
                if (vc < 0 || vs < 0 || va < 0 || (vbad && (vc | vs | va))) {
                    const uint32_t j = q.w;
                    exact_decide(P, px, py, pz, vx, vy, vz, vn, in.x[j], in.y[j], in.z[j], in_s, in_a, in_c);
                    ++nre;
                }
                nc += in_c;
                scx += in_c ? (int64_t)ix : 0; scy += in_c ? (int64_t)iy : 0; scz += in_c ? (int64_t)iz : 0;
                ns += in_s;
                kahan_add(ssx, csx, in_s ? -gx * inv_d : 0.f);
                kahan_add(ssy, csy, in_s ? -gy * inv_d : 0.f);
                kahan_add(ssz, csz, in_s ? -gz * inv_d : 0.f);
                na += in_a;
                kahan_add(sax, cax, in_a ? wx : 0.f); kahan_add(say, cay, in_a ? wy : 0.f); kahan_add(saz, caz, in_a ? wz : 0.f);
                if (DIGEST) {
                    const uint64_t h = mix64(sid[kk]);
                    hc += in_c ? h : 0; hs += in_s ? h : 0; ha += in_a ? h : 0;
                }
            };
            // Phase 1: coarse squared distance on the top 16 lattice bits
            // (packed int16 subtract + dot2, saturating); it under-estimates
            // no distance by more than sqrt(3) coarse units, and coarse_thr
            // carries that margin, so no candidate of any rule is dropped.
            const uint32_t kb = max(wlo, base), ke = min(whi, base + here);
            auto coarse = [&](uint32_t kk) {
                const uint2 cq = sc[kk];
                const v2s dxy = __builtin_bit_cast(v2s, cq.x) - my_xy;
                const v2s dzz = __builtin_bit_cast(v2s, cq.y) - my_z;
                return __builtin_amdgcn_sdot2(dxy, dxy, __builtin_amdgcn_sdot2(dzz, dzz, 0, true), true);
            };
#if BS_NBR_QUEUE
            auto drain = [&](uint32_t qn) {
                uint32_t j = 0;
                for (; j + 2 <= qn; j += 2) {
                    const uint32_t k0 = sq[j * kNbr + threadIdx.x], k1 = sq[(j + 1) * kNbr + threadIdx.x];
                    pair(k0);
                    pair(k1);
                }
                if (j < qn) pair(sq[j * kNbr + threadIdx.x]);
            };
            auto near = [&](uint32_t k) { return coarse(k - base) <= P.coarse_thr && k != selfk; };
            uint32_t qn = 0;
            uint32_t k = kb;
            for (; k + 4 <= ke; k += 4) {
                if (__any(qn > (uint32_t)kQueue - 4)) { drain(qn); qn = 0; }
                const bool h0 = near(k), h1 = near(k + 1), h2 = near(k + 2), h3 = near(k + 3);
                const uint32_t kk = k - base;
                sq[qn * kNbr + threadIdx.x] = (uint16_t)kk; qn += h0;
                sq[qn * kNbr + threadIdx.x] = (uint16_t)(kk + 1); qn += h1;
                sq[qn * kNbr + threadIdx.x] = (uint16_t)(kk + 2); qn += h2;
                sq[qn * kNbr + threadIdx.x] = (uint16_t)(kk + 3); qn += h3;
            }
            for (; k < ke; ++k) {
                if (__any(qn == (uint32_t)kQueue)) { drain(qn); qn = 0; }
                sq[qn * kNbr + threadIdx.x] = (uint16_t)(k - base);
                qn += near(k);
            }
            drain(qn);
#else
            for (uint32_t k = kb; k < ke; ++k) {
                const uint32_t kk = k - base;
                if (coarse(kk) <= P.coarse_thr && k != selfk) pair(kk);
            }
#endif
        }
    }
    if (nre) atomicAdd(recheck, nre);
    if (!active) return;
    // dv_coh = mean(q) - p = mean(q - p): the lattice sums are already relative
    double dcx = 0, dcy = 0, dcz = 0, dax = 0, day = 0, daz = 0;
    if (nc != 0) {
        const double k = P.unit / (double)nc;
        dcx = (double)scx * k; dcy = (double)scy * k; dcz = (double)scz * k;
    }
    if (na != 0) {
        const double k = (double)na;
        dax = ((double)sax - (double)cax) / k - vx; day = ((double)say - (double)cay) / k - vy; daz = ((double)saz - (double)caz) / k - vz;
    }
    const double dsx = (double)ssx - (double)csx, dsy = (double)ssy - (double)csy, dsz = (double)ssz - (double)csz;
    if (DIGEST) {
        uint64_t* d = digest + 4 * (size_t)myid;
        d[0] = (uint64_t)nc | ((uint64_t)ns << 21) | ((uint64_t)na << 42);
        d[1] = hc; d[2] = hs; d[3] = ha;
    }
    finish_boid(P, cbase, s, myid, px, py, pz, vx, vy, vz, dcx, dcy, dcz, dsx, dsy, dsz, dax, day, daz,
                out, key_out, count, forces);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline uint32_t blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

hipError_t launch_allpairs_f32ref(const DevParams& P, StateF32 in, StateF32 out, float* forces,
                                  uint64_t* digest, hipStream_t stream)
{
    if (P.n == 0) return hipSuccess;
    hipLaunchKernelGGL((k_allpairs<float, StateF32>), dim3(blocks_for(P.n)), dim3(kBlock), 0, stream,
                       P, in, out, forces, digest);
    return hipGetLastError();
}

hipError_t launch_allpairs_f64(const DevParams& P, StateF64 in, StateF64 out, double* forces,
                               uint64_t* digest, hipStream_t stream)
{
    if (P.n == 0) return hipSuccess;
    hipLaunchKernelGGL((k_allpairs<double, StateF64>), dim3(blocks_for(P.n)), dim3(kBlock), 0, stream,
                       P, in, out, forces, digest);
    return hipGetLastError();
}

hipError_t launch_hash_count(Flocks F, StateF64 st, uint32_t* key, uint32_t* count, hipStream_t stream)
{
    if (F.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_hash_count, dim3(blocks_for(F.n)), dim3(kBlock), 0, stream, F, st, key, count);
    return hipGetLastError();
}

hipError_t launch_scan(uint32_t* count, uint32_t* start, uint32_t m, uint32_t* partials,
                       hipStream_t stream)
{
    const uint32_t nt = (m + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(kScanThreads), 0, stream, count, start, m, partials);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kScanThreads), 0, stream, partials, nt);
    hipLaunchKernelGGL(k_scan_add, dim3(nt), dim3(kScanThreads), 0, stream, start, m,
                       (const uint32_t*)partials);
    return hipGetLastError();
}

hipError_t launch_rows(Flocks F, const uint32_t* start, uint4* table, uint32_t* nchunks,
                       hipStream_t stream)
{
    if (F.count == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_rows, dim3(1), dim3(kScanThreads), 0, stream, F, start, table, nchunks);
    return hipGetLastError();
}

hipError_t launch_scatter(uint32_t n, const uint32_t* key, const uint32_t* start, uint32_t* count,
                          const uint32_t* id, uint32_t* perm, uint32_t* permid, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter, dim3(blocks_for(n)), dim3(kBlock), 0, stream, n, key, start,
                       count, id, perm, permid);
    return hipGetLastError();
}

hipError_t launch_gather(Flocks F, const uint32_t* perm, const uint32_t* permid, const uint32_t* key,
                         const uint32_t* start, uint32_t* count, StateF64 src, StateF64 dst,
                         uint32_t* key_sorted, Lattice* lat, hipStream_t stream)
{
    if (F.n == 0) return hipSuccess;
    Lattice l = lat ? *lat : Lattice{nullptr, nullptr, nullptr, nullptr};
    hipLaunchKernelGGL(k_gather, dim3(blocks_for(F.n)), dim3(kBlock), 0, stream, F, perm, permid,
                       key, start, count, src, dst, key_sorted, l, lat ? 1 : 0);
    return hipGetLastError();
}

static inline uint32_t round8(uint32_t g) { return (g + 7u) & ~7u; }

hipError_t launch_nbr_f64(Flocks F, int max_g, StateF64 in, StateF64 out, const uint32_t* key_sorted,
                          const uint32_t* start, const uint4* table, const uint32_t* nchunks,
                          uint32_t max_chunks, uint32_t* key_out, uint32_t* count, double* forces,
                          uint64_t* digest, hipStream_t stream)
{
    if (F.n == 0) return hipSuccess;
    if (max_g < 3 || max_g > kMaxCells) return hipErrorInvalidValue;
    const dim3 grid(round8(max_chunks));
    const size_t dyn = segment_bytes(max_g);
    if (digest)
        hipLaunchKernelGGL(k_nbr_f64<true>, grid, dim3(kNbr), dyn, stream, F, in, out, key_sorted,
                           start, table, nchunks, key_out, count, forces, digest);
    else
        hipLaunchKernelGGL(k_nbr_f64<false>, grid, dim3(kNbr), dyn, stream, F, in, out, key_sorted,
                           start, table, nchunks, key_out, count, forces, digest);
    return hipGetLastError();
}

hipError_t launch_nbr_mixed(Flocks F, int max_g, StateF64 in, StateF64 out, Lattice lat,
                            const uint32_t* key_sorted, const uint32_t* start, const uint4* table,
                            const uint32_t* nchunks, uint32_t max_chunks, uint32_t* key_out,
                            uint32_t* count, double* forces, uint64_t* digest, uint32_t* recheck,
                            hipStream_t stream)
{
    if (F.n == 0) return hipSuccess;
    if (max_g < 3 || max_g > kMaxCells) return hipErrorInvalidValue;
    const dim3 grid(round8(max_chunks));
    const size_t dyn = segment_bytes(max_g);
    if (digest)
        hipLaunchKernelGGL(k_nbr_mixed<true>, grid, dim3(kNbr), dyn, stream, F, in, out, lat,
                           key_sorted, start, table, nchunks, key_out, count, forces, digest, recheck);
    else
        hipLaunchKernelGGL(k_nbr_mixed<false>, grid, dim3(kNbr), dyn, stream, F, in, out, lat,
                           key_sorted, start, table, nchunks, key_out, count, forces, digest, recheck);
    return hipGetLastError();
}

// Plain streaming copy (16 B per lane per iteration): the reference point
// for HBM byte counters (a copy of B bytes moves exactly B in and B out).
__global__ __launch_bounds__(kBlock) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                 uint64_t n16)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * kBlock)
        dst[i] = src[i];
}

hipError_t launch_copy(const void* src, void* dst, uint64_t bytes, hipStream_t stream)
{
    const uint64_t n16 = bytes / 16;
    if (n16 == 0) return hipSuccess;
    const uint64_t want = (n16 + kBlock - 1) / kBlock;
    const uint32_t grid = (uint32_t)(want < 65536 ? want : 65536);
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(kBlock), 0, stream, (const uint4*)src, (uint4*)dst, n16);
    return hipGetLastError();
}

hipError_t launch_unsort_f64(uint32_t n, StateF64 src, double* pos3, double* vel3,
                             hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_unsort_f64, dim3(blocks_for(n)), dim3(kBlock), 0, stream, n, src, pos3, vel3);
    return hipGetLastError();
}

}  // namespace boids
