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
#include "boids_device.h"

namespace boids {


// ---------------------------------------------------------------------------
// All-pairs step: one wave per boid i.  The 64 lanes take the decisions for
// candidates j = base + lane (the reference's per-pair arithmetic,
// src/boid_simulation.cpp:153-212), then the wave adds the in-range pairs'
// terms in ascending j -- the reference's summation order -- one pair at a
// time, every lane holding the same sums, and lane 0 runs phase B
// (:216-248).  R = float: bit-identical to boidsim.  R = double: the FP64
// promotion.  A flock of N boids takes N waves, so small flocks (cfg0's
// N = 1024) still fill the chip; the per-pair decisions, not the ordered
// sums, are the work.
// ---------------------------------------------------------------------------
constexpr int kWaveOwners = kBlock / 64;     // owners (waves) per workgroup

template <typename R>
__device__ __forceinline__ R lane_value(R v, uint32_t l)
{
    if constexpr (sizeof(R) == 4) {
        return __builtin_bit_cast(R, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), (int)l));
    } else {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, (int)l);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), (int)l);
        return __builtin_bit_cast(R, ((uint64_t)hi << 32) | lo);
    }
}

template <typename R, typename St>
__global__ __launch_bounds__(kBlock) void k_allpairs_wave(DevParams P, St in, St out, R* forces,
                                                          uint64_t* digest)
{
    const uint32_t n = P.n;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t i = blockIdx.x * kWaveOwners + (threadIdx.x >> 6);
    if (i >= n) return;                                  // the whole wave
    R px = in.x[i], py = in.y[i], pz = in.z[i], vx = in.vx[i], vy = in.vy[i], vz = in.vz[i];
    const R vn = norm3(vx, vy, vz);
    const R smax = Th<R>::smax(P);
    const R s_c = Th<R>::s(P, COH), s_s = Th<R>::s(P, SEP), s_a = Th<R>::s(P, ALI);
    const R c_c = Th<R>::c(P, COH), c_s = Th<R>::c(P, SEP), c_a = Th<R>::c(P, ALI);
    const bool k_c = P.cone[COH], k_s = P.cone[SEP], k_a = P.cone[ALI];
    R scx = 0, scy = 0, scz = 0, ssx = 0, ssy = 0, ssz = 0, sax = 0, say = 0, saz = 0;
    int nc = 0, ns = 0, na = 0;
    uint64_t hc = 0, hs = 0, ha = 0;

    // candidate positions two batches ahead of the one being tested
    auto fetch = [&](uint32_t base, R& x, R& y, R& z) {
        const uint32_t j = min(base + lane, n - 1);
        x = in.x[j]; y = in.y[j]; z = in.z[j];
    };
    R ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
    fetch(0, ax, ay, az);
    if (64 < n) fetch(64, bx, by, bz);
    for (uint32_t base = 0; base < n; base += 64) {
        const R cx = ax, cy = ay, cz = az;
        ax = bx; ay = by; az = bz;
        if (base + 128 < n) fetch(base + 128, bx, by, bz);
        const uint32_t j = base + lane;
        bool in_c = false, in_s = false, in_a = false;
        R qx = 0, qy = 0, qz = 0, tx = 0, ty = 0, tz = 0, wx = 0, wy = 0, wz = 0;
        if (j < n && j != i) {
            qx = min_image(cx, px, P.F, P.half);
            qy = min_image(cy, py, P.F, P.half);
            qz = min_image(cz, pz, P.F, P.half);
            const R ex = px - qx, ey = py - qy, ez = pz - qz;
            const R s = ex * ex + ey * ey + ez * ez;
            if (s < smax) {
                in_c = s < s_c; in_s = s < s_s; in_a = s < s_a;
                if ((in_c && k_c) || (in_s && k_s) || (in_a && k_a)) {
                    const R gx = qx - px, gy = qy - py, gz = qz - pz;
                    const R dot = vx * gx + vy * gy + vz * gz;
                    const R c = dot / (vn * sqrt(s));
                    if (k_c) in_c = in_c && c >= c_c && c <= (R)1;
                    if (k_s) in_s = in_s && c >= c_s && c <= (R)1;
                    if (k_a) in_a = in_a && c >= c_a && c <= (R)1;
                }
                if (in_s) {
                    const R t = sqrt(s);
                    tx = ex / t; ty = ey / t; tz = ez / t;
                }
                if (in_a) { wx = in.vx[j]; wy = in.vy[j]; wz = in.vz[j]; }
            }
        }
        const int flags = (in_c ? 1 : 0) | (in_s ? 2 : 0) | (in_a ? 4 : 0);
        uint64_t m = __builtin_amdgcn_ballot_w64(flags != 0);
        while (m) {                                      // ascending j: the reference's order
            const uint32_t l = (uint32_t)__builtin_ctzll(m);
            m &= m - 1;
            const int f = __builtin_amdgcn_readlane(flags, (int)l);
            const uint64_t h = mix64(base + l);
            if (f & 1) {
                nc++; scx += lane_value(qx, l); scy += lane_value(qy, l); scz += lane_value(qz, l); hc += h;
            }
            if (f & 2) {
                ns++; ssx += lane_value(tx, l); ssy += lane_value(ty, l); ssz += lane_value(tz, l); hs += h;
            }
            if (f & 4) {
                na++; sax += lane_value(wx, l); say += lane_value(wy, l); saz += lane_value(wz, l); ha += h;
            }
        }
    }
    if (lane != 0) return;
    if (nc != 0) { const R k = (R)nc; scx = scx / k - px; scy = scy / k - py; scz = scz / k - pz; }
    if (na != 0) { const R k = (R)na; sax = sax / k - vx; say = say / k - vy; saz = saz / k - vz; }
    const R wc = Th<R>::w(P, COH), ws = Th<R>::w(P, SEP), wa = Th<R>::w(P, ALI);
    const R dx = (scx * wc + ssx * ws) + sax * wa;
    const R dy = (scy * wc + ssy * ws) + say * wa;
    const R dz = (scz * wc + ssz * ws) + saz * wa;
    if (forces) {
        R* fo = forces + 9 * (size_t)i;
        fo[0] = scx; fo[1] = scy; fo[2] = scz; fo[3] = ssx; fo[4] = ssy; fo[5] = ssz;
        fo[6] = sax; fo[7] = say; fo[8] = saz;
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

__global__ __launch_bounds__(kBlock) void k_hash_count(Flocks F, StateF64 st, uint32_t* key,
                                                        uint32_t* rank, uint32_t* count)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= F.n) return;
    const FlockDesc& fd = F.fl[flock_of(F, i)];
    const uint32_t c = fd.cell_base + cell_of(fd.P, st.x[i], st.y[i], st.z[i]);
    key[i] = c;
    rank[i] = atomicAdd(&count[c], 1u);
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

// Chunk table of the neighbour kernels, entries (local row, first slot, end
// slot, flock).  Every (y,z) row of cells of every flock (sx G x columns)
// is cut into groups of `span` consecutive columns, and every group into
// runs of `chunk` consecutive sorted boids, cut short where a run would cover
// more than `cspan` columns (sparse stretches), so a chunk lies in one
// row, holds at most `chunk` boids and its tile tables stay small.  A short
// cut moves the run's first column on by at least cspan, so a group has at
// most (span - 1) / cspan of them: the table holds at most n / chunk +
// groups * (1 + (span - 1) / cspan) entries.  One wave per row: the
// row's column starts are read into LDS (coalesced) and one lane walks them.
// Pass 1 counts the row's chunks; after a scan of the counts, pass 2 writes
// them at the row's offset (rows in order: neighbouring rows' chunks stay
// adjacent in the table).
constexpr int kRowWaves = 4;

//
// The bound is the table's capacity (`cap`); pass 2 writes no entry at or
// beyond it, and should the count ever exceed it, *nchunks is clamped to
// the capacity and the full count is left in *overflow (non-zero marks the
// table as cut short; the host reports it as an error).
template <bool WRITE>
__global__ __launch_bounds__(64 * kRowWaves) void k_rows_pass(Flocks F, const uint32_t* __restrict__ start,
                                                              uint32_t* row_count, const uint32_t* row_off,
                                                              uint4* table, uint32_t* nchunks, uint32_t chunk,
                                                              uint32_t cspan, uint32_t stride, uint32_t cap,
                                                              uint32_t* overflow)
{
    extern __shared__ uint32_t rs[];            // kRowWaves x stride (>= the longest row's columns + 1)
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t r = blockIdx.x * kRowWaves + w;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (WRITE) {
            const uint32_t total = row_off[F.rows];
            *nchunks = min(total, cap);
            if (total > cap) *overflow = total;
        } else {
            row_count[F.rows] = 0u;             // the scan's total lands in row_off[rows]
        }
    }
    if (r >= F.rows) return;                    // whole waves only: no barrier follows
    const uint32_t f = F.row_flock ? F.row_flock[r] : 0u;
    const FlockDesc& fd = F.fl[f];
    const uint32_t G = (uint32_t)fd.P.sx * (uint32_t)fd.P.G, span = fd.span;   // x columns of the row
    const uint32_t rl = r - fd.row_base;
    const uint32_t c0 = fd.cell_base + rl * G;
    uint32_t* cs = rs + w * stride;
    for (uint32_t x = lane; x <= G; x += 64) cs[x] = start[c0 + x];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane != 0) return;
    uint32_t k = WRITE ? row_off[r] : 0u;
    for (uint32_t x0 = 0; x0 < G; x0 += span) {
        const uint32_t x1 = min(G, x0 + span);
        const uint32_t hi = cs[x1];
        uint32_t t0 = cs[x0], xa = x0;
        while (t0 < hi) {
            while (cs[xa + 1] <= t0) ++xa;          // cell of boid t0
            uint32_t t1 = min(hi, t0 + chunk);
            if (xa + cspan < x1) t1 = min(t1, cs[xa + cspan]);
            if (WRITE && k < cap) table[k] = make_uint4(rl, t0, t1, f);
            ++k;
            t0 = t1;
        }
    }
    if (!WRITE) row_count[r] = k;
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

__global__ __launch_bounds__(kBlock) void k_scatter_ranked(uint32_t n, const uint32_t* __restrict__ key,
                                                           const uint32_t* __restrict__ rank,
                                                           const uint32_t* __restrict__ start,
                                                           const uint32_t* __restrict__ id, uint32_t* perm,
                                                           uint32_t* permid)
{
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    const uint32_t t = start[key[s]] + rank[s];
    perm[t] = s;
    permid[t] = id[s];
}

// In-cell order, step 1: every run of 64 consecutive slots (one wave) is
// sorted by (cell, id) in registers, in place: a bitonic network across the
// wave's lanes over (cell part within the run, id) -- the slots are grouped
// by cell already, so this sorts each cell's part inside the run.  A cell
// spanning runs leaves one ascending part per run, which k_gather merges.
// Keys are 32-bit (6-bit part index, 26-bit id) when every id is below 2^26,
// else 64-bit.  A run already in order (the common case: boids arriving from
// one old cell are ranked in lane order) is left alone.
template <typename K>
__device__ __forceinline__ K shfl_xor_key(K v, uint32_t m)
{
    if constexpr (sizeof(K) == 4) {
        return (K)__shfl_xor((uint32_t)v, m, 64);
    } else {
        const uint32_t lo = __shfl_xor((uint32_t)v, m, 64), hi = __shfl_xor((uint32_t)(v >> 32), m, 64);
        return ((K)hi << 32) | lo;
    }
}

template <typename K>
__global__ __launch_bounds__(kBlock) void k_sort_runs(uint32_t n, const uint32_t* __restrict__ key,
                                                      uint32_t* perm, uint32_t* permid)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    const bool in = t < n;
    uint32_t p = 0u, id = 0u, c = 0xFFFFFFFFu;
    if (in) {
        p = perm[t];
        id = permid[t];
        c = key[p];
    }
    const uint32_t cp = __shfl_up(c, 1, 64), ip = __shfl_up(id, 1, 64);
    const bool fresh = lane == 0 || cp != c;         // first slot of a cell's part
    if (__builtin_amdgcn_ballot_w64(in && !fresh && ip > id) == 0) return;   // wave-uniform
    const uint64_t starts = __builtin_amdgcn_ballot_w64(fresh);
    const uint32_t part = __builtin_amdgcn_mbcnt_hi((uint32_t)(starts >> 32),
                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)starts, 0u)) +
                          (fresh ? 1u : 0u) - 1u;     // parts begun at or before this lane, minus one
    constexpr uint32_t kShift = sizeof(K) == 4 ? 26u : 32u;
    K k = in ? (((K)part << kShift) | (K)id) : ~(K)0;   // past the end: sorts last, never written
    for (uint32_t size = 2; size <= 64; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            const K pk = shfl_xor_key(k, stride);
            const uint32_t pp = __shfl_xor(p, stride, 64);
            // ascending blocks of `size` where lane & size == 0: the lower
            // lane of a pair keeps the smaller key there, the larger elsewhere
            const bool keep_min = ((lane & stride) == 0) == ((lane & size) == 0);
            if (keep_min ? pk < k : pk > k) { k = pk; p = pp; }
        }
    }
    if (in) {
        perm[t] = p;
        permid[t] = (uint32_t)(k & (((K)1 << kShift) - 1));
    }
}

// In-cell order, step 2: final slot = cell start + rank of the boid id among
// its cell's ids, so the order inside a cell is ascending id whatever the
// atomics did.  The rank is the slot's position in its own sorted part plus
// a binary search in each other part of the cell (k_sort_runs): O(1) for a
// cell inside one run, O(occupancy / 64 * log 64) for a dense one, instead
// of a pass over the whole cell per boid (a scan of short parts instead of
// the search measured slower: its loop is as long as the part).
__global__ __launch_bounds__(kBlock) void k_gather(Flocks F, const uint32_t* perm,
                                                   const uint32_t* permid, const uint32_t* key,
                                                   const uint32_t* start, uint32_t* count,
                                                   StateF64 src, StateF64 dst,
                                                   uint32_t* key_sorted, Lattice lat, int with_lat,
                                                   double2* rec)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= F.n) return;
    const uint32_t s = perm[t];
    const uint32_t me = permid[t];
    const uint32_t c = key[s];
    const uint32_t lo = start[c], hi = start[c + 1];
    const uint32_t run = t >> 6;
    uint32_t rank = t - max(lo, run << 6);
    for (uint32_t v = lo >> 6; (v << 6) < hi; ++v) {
        if (v == run) continue;
        const uint32_t a = max(lo, v << 6), b = min(hi, (v + 1) << 6);
        rank += count_below(permid + a, b - a, me);
    }
    const uint32_t f = lo + rank;
    const double x = src.x[s], y = src.y[s], z = src.z[s];
    const double vx = src.vx[s], vy = src.vy[s], vz = src.vz[s];
    dst.x[f] = x; dst.y[f] = y; dst.z[f] = z;
    dst.vx[f] = vx; dst.vy[f] = vy; dst.vz[f] = vz;
    dst.id[f] = me;
    key_sorted[f] = c;
    if (rank == 0) count[c] = 0u;
    if (with_lat) {
        const DevParams& P = F.fl[flock_of(F, f)].P;   // slots stay inside their flock
        const double scale = P.lat_scale;
        const bool nan_state = x != x || y != y || z != z || vx != vx || vy != vy || vz != vz;
        const uint32_t nanv = nan_state ? 0x80000000u : 0u;   // NaN state mark
        lat.p[f] = make_uint4(to_lattice(x, scale), to_lattice(y, scale), to_lattice(z, scale),
                              pack_sub(x, y, z, scale) | nanv);
        const double iq = P.inv_vq;
        lat.v[f] = make_float4((float)rint(vx * iq), (float)rint(vy * iq), (float)rint(vz * iq), 0.f);
    }
    if (rec) {                                 // three planes of n records: (x, y), (z, vx), (vy, vz)
        rec[f] = make_double2(x, y);
        rec[(size_t)F.n + f] = make_double2(z, vx);
        rec[2 * (size_t)F.n + f] = make_double2(vy, vz);
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

// Launchers
// ---------------------------------------------------------------------------
static inline uint32_t blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

hipError_t launch_allpairs_f32ref(const DevParams& P, StateF32 in, StateF32 out, float* forces,
                                  uint64_t* digest, hipStream_t stream)
{
    if (P.n == 0) return hipSuccess;
    hipLaunchKernelGGL((k_allpairs_wave<float, StateF32>), dim3((P.n + kWaveOwners - 1) / kWaveOwners),
                       dim3(kBlock), 0, stream, P, in, out, forces, digest);
    return hipGetLastError();
}

hipError_t launch_allpairs_f64(const DevParams& P, StateF64 in, StateF64 out, double* forces,
                               uint64_t* digest, hipStream_t stream)
{
    if (P.n == 0) return hipSuccess;
    hipLaunchKernelGGL((k_allpairs_wave<double, StateF64>), dim3((P.n + kWaveOwners - 1) / kWaveOwners),
                       dim3(kBlock), 0, stream, P, in, out, forces, digest);
    return hipGetLastError();
}

hipError_t launch_hash_count(Flocks F, StateF64 st, uint32_t* key, uint32_t* rank, uint32_t* count,
                             hipStream_t stream)
{
    if (F.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_hash_count, dim3(blocks_for(F.n)), dim3(kBlock), 0, stream, F, st, key, rank, count);
    return hipGetLastError();
}

hipError_t launch_scan(uint32_t* count, uint32_t* start, uint32_t m, uint32_t* partials,
                       hipStream_t stream)
{
    const uint32_t nt = (m + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(kScanThreads), 0, stream, count, start, m, partials);
    if (nt <= 1) return hipGetLastError();      // one tile: its local scan is the scan
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kScanThreads), 0, stream, partials, nt);
    hipLaunchKernelGGL(k_scan_add, dim3(nt), dim3(kScanThreads), 0, stream, start, m,
                       (const uint32_t*)partials);
    return hipGetLastError();
}

hipError_t launch_rows(Flocks F, const uint32_t* start, uint32_t* row_count, uint32_t* row_off,
                       uint32_t* partials, uint4* table, uint32_t* nchunks, uint32_t cap, uint32_t* overflow,
                       int chunk, int cspan, int max_cols, hipStream_t stream)
{
    if (F.count == 0 || chunk <= 0 || F.rows == 0 || cspan < 1 || cspan > kChunkSpanMax || max_cols < 1 ||
        max_cols > kSubXMax * kMaxCells)
        return hipErrorInvalidValue;
    const dim3 grid((F.rows + kRowWaves - 1) / kRowWaves);
    const uint32_t stride = (uint32_t)max_cols + 1;
    const size_t dyn = sizeof(uint32_t) * kRowWaves * stride;
    hipLaunchKernelGGL(k_rows_pass<false>, grid, dim3(64 * kRowWaves), dyn, stream, F, start, row_count,
                       (const uint32_t*)row_off, table, nchunks, (uint32_t)chunk, (uint32_t)cspan, stride, cap,
                       overflow);
    hipError_t st = launch_scan(row_count, row_off, F.rows + 1, partials, stream);
    if (st != hipSuccess) return st;
    hipLaunchKernelGGL(k_rows_pass<true>, grid, dim3(64 * kRowWaves), dyn, stream, F, start, row_count,
                       (const uint32_t*)row_off, table, nchunks, (uint32_t)chunk, (uint32_t)cspan, stride, cap,
                       overflow);
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

hipError_t launch_scatter_ranked(uint32_t n, const uint32_t* key, const uint32_t* rank, const uint32_t* start,
                                 const uint32_t* id, uint32_t* perm, uint32_t* permid, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_ranked, dim3(blocks_for(n)), dim3(kBlock), 0, stream, n, key, rank, start, id,
                       perm, permid);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    return launch_sort_runs(n, key, perm, permid, stream);
}

hipError_t launch_sort_runs(uint32_t n, const uint32_t* key, uint32_t* perm, uint32_t* permid, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    if (n <= (1u << 26))
        hipLaunchKernelGGL(k_sort_runs<uint32_t>, dim3(blocks_for(n)), dim3(kBlock), 0, stream, n, key, perm, permid);
    else
        hipLaunchKernelGGL(k_sort_runs<uint64_t>, dim3(blocks_for(n)), dim3(kBlock), 0, stream, n, key, perm, permid);
    return hipGetLastError();
}

hipError_t launch_gather(Flocks F, const uint32_t* perm, const uint32_t* permid, const uint32_t* key,
                         const uint32_t* start, uint32_t* count, StateF64 src, StateF64 dst,
                         uint32_t* key_sorted, Lattice* lat, double2* rec, hipStream_t stream)
{
    if (F.n == 0) return hipSuccess;
    Lattice l = lat ? *lat : Lattice{nullptr, nullptr};
    hipLaunchKernelGGL(k_gather, dim3(blocks_for(F.n)), dim3(kBlock), 0, stream, F, perm, permid,
                       key, start, count, src, dst, key_sorted, l, lat ? 1 : 0, rec);
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
