// boids_nbr.hip — fused neighbour + integrate kernels of the grid modes
// (the cell-list update of src/boid_simulation_grid.cpp:165-293, phase A of
// src/boid_simulation.cpp:153-212 per pair, phase B :216-248).
//
// A workgroup of NB threads owns a chunk [t0, t1) of sorted boids inside
// one (y,z) row of cells, spanning x columns cxa..cxb of that row (at most
// the engine's chunk span; a row has sx columns per cell edge).  Its tile is the 3 x 3
// neighbour rows over the columns cxa-sx .. cxb+sx, ordered column-major:
// entry e = 9 * column + rr, with rr = 3 * (dz + 1) + (dy + 1), each entry
// the boids of one (column, row) cell in sorted order.  A boid in column L
// therefore finds its whole window, columns L-sx .. L+sx of the 9 rows, in
// the one contiguous tile range [off(L-sx, 0), off(L+sx+1, 0)).
//
// Per candidate the tile holds a coarse position (top 16 bits of the 32-bit
// lattice coordinate per axis) and the index of its tile entry (cell); the
// slot is the entry's sorted start + offset.  Phase 1 tests every candidate
// of the window with a packed 16-bit squared distance that never drops a
// pair inside the largest radius, lanes of one cell reading the same tile
// entry at a time, and appends hits to the wave's list by ballot.  Phase 2
// evaluates the list 64 pairs at a time, every lane busy, fetching each
// candidate's record from L2 by slot, takes the rule decisions
//   FP64 plain: the oracle's expressions (exact thresholds, s < s_thr and
//               c >= c_thr && c <= 1, on the minimum image);
//   mixed:      FP32 on exact lattice differences outside guard bands,
//               the FP64 expression inside them,
// and adds the pair to its owner's sums with LDS atomics.
// Tiles larger than the LDS capacity are processed in passes.
#include "boids_device.h"

#include <type_traits>

namespace boids {

template <int NB, int TD> struct NbrCfg {
    static constexpr int kCap = NB * TD;    // staged candidates per pass
};

// Entry tables of a tile in dynamic LDS: off[E + 1] (prefix of the entry
// sizes) and sbase[E] (slot of the entry's first boid minus off[e], so a tile
// position p holds slot sbase[e] + p), E = 9 * nseg
// for a tile of nseg <= max_cols columns.
__host__ __device__ constexpr size_t tile_table_bytes(int max_cols)
{
    return ((size_t)(9 * max_cols + 1) * 4 + (size_t)(9 * max_cols) * 4 + 15) & ~(size_t)15;
}

// Builds the entry tables of the chunk's tile (counts, then a block-wide
// exclusive scan; each thread owns consecutive entries).
template <int NB>
__device__ __forceinline__ void build_tile_tables(const uint32_t* __restrict__ start, uint32_t cbase, int G,
                                                  int sx, int cxa, int nseg, int cy, int cz, uint32_t* off,
                                                  uint32_t* sbase, uint32_t* wsum)
{
    const int E = 9 * nseg;
    const int per = (E + NB - 1) / NB;
    const int e0 = threadIdx.x * per;
    const int e1 = min(E, e0 + per);
    uint32_t local = 0;
    for (int e = e0; e < e1; ++e) {
        const int xc = e / 9, rr = e - 9 * xc;
        const int Gx = sx * G;
        int cx = cxa - sx + xc;             // in [-sx, Gx - 1 + sx]: one wrap at most
        if (cx < 0) cx += Gx; else if (cx >= Gx) cx -= Gx;
        int ny = cy + rr % 3 - 1, nz = cz + rr / 3 - 1;
        if (ny < 0) ny += G; else if (ny >= G) ny -= G;
        if (nz < 0) nz += G; else if (nz >= G) nz -= G;
        const uint32_t c = cbase + (uint32_t)cx + (uint32_t)Gx * ((uint32_t)ny + (uint32_t)G * (uint32_t)nz);
        const uint32_t a = start[c], b = start[c + 1];
        sbase[e] = a;                     // the cell's sorted start, until the scan below
        off[e] = b - a;                 // the count, until the scan below
        local += b - a;
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(local, wsum, &total);
    for (int e = e0; e < e1; ++e) {
        const uint32_t cnt = off[e];
        off[e] = run;
        sbase[e] -= run;                  // slot = sbase[e] + tile position (mod 2^32)
        run += cnt;
    }
    if (threadIdx.x == 0) off[E] = total;
    __syncthreads();
}

// Coarse position of a lattice point: top 16 bits of x|y, z.
__device__ __forceinline__ uint2 coarse_of(uint32_t lx, uint32_t ly, uint32_t lz)
{
    return make_uint2((lx >> 16) | (ly & 0xFFFF0000u), lz >> 16);
}

// Squared coarse distance (int16 minimum-image differences, the x|y pair
// packed; the dot product saturates instead of wrapping).  z is the low
// half of q.y (the high half holds the tile entry): its difference is taken
// mod 2^16 and squared as a sign-extended 16-bit value, dz^2 <= 2^30.
__device__ __forceinline__ int32_t coarse_d2(uint2 q, v2s my_xy, uint32_t my_z)
{
    const v2s dxy = __builtin_bit_cast(v2s, q.x) - my_xy;
    const int32_t dz = (int16_t)(uint16_t)(q.y - my_z);   // v_sub + v_mul_i32_i24 on the sign-extended words
    return __builtin_amdgcn_sdot2(dxy, dxy, dz * dz, true);
}

// Padding entry of a chunk's tile, far from every owner of the chunk: the
// owners lie in one (y, z) cell row, whose coarse y and z span 65536 / G
// units each; the padding sits half the 16-bit range away from the row's
// centre on both axes, so its coarse distance to an owner is at least
// sqrt(2) (32768 - 32768 / G) >= 30893 units, above the coarse bound of any
// radius below the cell edge (at most 65536 / G <= 21846 units).
__device__ __forceinline__ uint2 far_entry(int cy, int cz, int G)
{
    const uint32_t yc = (uint32_t)(2 * cy + 1) * 32768u / (uint32_t)G;
    const uint32_t zc = (uint32_t)(2 * cz + 1) * 32768u / (uint32_t)G;
    return make_uint2(((yc + 32768u) & 0xFFFFu) << 16, (zc + 32768u) & 0xFFFFu);
}

// Stages tile positions [base, base + here); each thread copies one run of
// consecutive positions, calling STORE(position - base, slot, entry).
template <int NB, typename STORE>
__device__ __forceinline__ void stage_tile(const uint32_t* off, const uint32_t* sbase, int E, uint32_t base,
                                           uint32_t here, STORE store)
{
    const uint32_t per = (here + NB - 1) / NB;
    const uint32_t p0 = base + threadIdx.x * per;
    const uint32_t p1 = min(base + here, p0 + per);
    if (p0 >= p1) return;
    int lo = 0, hi = E - 1;                  // last entry with off <= p0
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= p0) lo = mid; else hi = mid - 1;
    }
    int e = lo;
    uint32_t ee = off[e + 1], g = sbase[e];
    for (uint32_t p = p0; p < p1; ++p) {
        while (p >= ee) { ++e; ee = off[e + 1]; g = sbase[e]; }
        store(p - base, g + p, (uint32_t)e);
    }
}

// Element i of a 16-byte record array at a wave-uniform base, addressed as
// base + 32-bit byte offset (16 i < 2^32: engines hold at most 2^28 boids),
// so a load takes one shift where 64-bit index arithmetic takes three.
template <typename T>
__device__ __forceinline__ T at32(const T* base, uint32_t i)
{
    static_assert(sizeof(T) == 16, "16-byte records");
    return *(const T*)((const char*)base + (size_t)(i << 4));
}

// Value of v in lane src of this wave (all lanes active).
__device__ __forceinline__ uint32_t lane_read(uint32_t v, uint32_t src)
{
    return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)v);
}
__device__ __forceinline__ float lane_read(float v, uint32_t src)
{
    return __builtin_bit_cast(float, lane_read(__builtin_bit_cast(uint32_t, v), src));
}
__device__ __forceinline__ double lane_read(double v, uint32_t src)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint64_t lo = lane_read((uint32_t)u, src), hi = lane_read((uint32_t)(u >> 32), src);
    return __builtin_bit_cast(double, (hi << 32) | lo);
}

// Mask of the wave's lanes where c holds, taken from the compare's own lane
// mask (__ballot and __any take an int, which costs a select and a second
// compare per call).
__device__ __forceinline__ uint64_t wave_mask(bool c)
{
    return __builtin_amdgcn_ballot_w64(c);
}

// Lanes of this wave below the calling lane whose bit is set in m.
__device__ __forceinline__ uint32_t lanes_below(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Per-wave hit list of the neighbour kernels, entries (tile position << 6)
// | owner lane, appended by ballot (each hit stored under its own lane
// mask) and drained in full rounds of 64 (kRounds rounds per step) once it
// holds kDrain entries.  A filter step appends at most 4 x 64 entries to a
// list holding fewer than kDrain, so kCap bounds every index.  The entries
// are 16-bit where the tile positions fit 10 bits (NB x TD <= 1024, the
// default shapes): half the list's LDS, one more workgroup per CU.
#ifndef BS_LIST_DRAIN
#define BS_LIST_DRAIN 128
#endif
#ifndef BS_MIXED_ROUNDS
#define BS_MIXED_ROUNDS 4
#endif
#ifndef BS_F64_ROUNDS
#define BS_F64_ROUNDS 4
#endif
template <uint32_t R> struct ListCfg {
    static constexpr uint32_t kRounds = R;                // rounds of 64 pairs evaluated per drain step
    static constexpr uint32_t kStep = 64 * R;
    static constexpr uint32_t kDrain = BS_LIST_DRAIN > kStep ? BS_LIST_DRAIN : kStep;
    static constexpr uint32_t kCap = kDrain + 4 * 64;     // + one filter step (4 candidates per lane)
    static_assert(kDrain % kStep == 0, "list drain threshold");
};
template <int NB, int TD>
using ListEntry = typename std::conditional<(NB * TD <= 1024), uint16_t, uint32_t>::type;

// Phase 1 of the neighbour kernels over one lane's window [k, k1) of the
// staged tile (positions relative to the pass's first): four candidates per
// lane and step, the next four read ahead (positions stay <= k1 + 3, inside
// the tile's 4 entries of padding); the list is flushed whenever it holds
// kDrain entries, so a step always starts below kDrain and appends at most
// 4 x 64 more.  A candidate within the coarse bound is appended as (tile
// position << 6) | lane, in (step, candidate, lane) order (phase 2 reads the
// candidate's tile entry from the tile); the owner itself is a hit (coarse
// distance 0) and is skipped by the evaluation.  flush(tail) evaluates the
// list's full rounds and returns the number of entries it leaves at the
// front.  Returns the final tail.  Every lane of the wave takes part.
template <uint32_t kDrain, typename LE, typename FLUSH>
__device__ __forceinline__ uint32_t filter_window(const uint2* tco, LE* list, uint32_t k, uint32_t k1,
                                                  uint32_t lane, v2s my_xy, uint32_t my_z, int32_t thr, FLUSH&& flush)
{
    // A finished lane tests nothing (its bound is below every distance); a
    // lane with 1-3 candidates left also tests the next positions up to 3
    // past its window end: boids outside the window (farther than the
    // largest radius, rejected exactly by phase 2) or the tile's padding
    // (far from every owner of the chunk).
    uint32_t tail = 0;
    uint2 a0 = tco[k], a1 = tco[k + 1], a2 = tco[k + 2], a3 = tco[k + 3];
    while (wave_mask(k < k1) != 0) {
        const uint32_t kn = min(k + 4, k1);
        const uint2 b0 = tco[kn], b1 = tco[kn + 1], b2 = tco[kn + 2], b3 = tco[kn + 3];
        const int32_t te = k < k1 ? thr : -1;
        const bool h0 = coarse_d2(a0, my_xy, my_z) <= te, h1 = coarse_d2(a1, my_xy, my_z) <= te;
        const bool h2 = coarse_d2(a2, my_xy, my_z) <= te, h3 = coarse_d2(a3, my_xy, my_z) <= te;
        const uint64_t m0 = wave_mask(h0), m1 = wave_mask(h1), m2 = wave_mask(h2), m3 = wave_mask(h3);
        const uint32_t e = (k << 6) | lane;
        const uint32_t t1 = tail + __popcll(m0), t2 = t1 + __popcll(m1), t3 = t2 + __popcll(m2);
        // the wave-uniform list offsets go into the (scalar) base address
        if (h0) (list + tail)[lanes_below(m0)] = (LE)e;
        if (h1) (list + t1)[lanes_below(m1)] = (LE)(e + (1u << 6));
        if (h2) (list + t2)[lanes_below(m2)] = (LE)(e + (2u << 6));
        if (h3) (list + t3)[lanes_below(m3)] = (LE)(e + (3u << 6));
        tail = t3 + __popcll(m3);
        if (tail >= kDrain) tail = flush(tail);
        k = kn;
        a0 = b0; a1 = b1; a2 = b2; a3 = b3;
    }
    return tail;
}

// ---------------------------------------------------------------------------
// FP64 plain.
// ---------------------------------------------------------------------------
// Phase 1 as in the mixed kernel (coarse test, per-wave hit list); phase 2
// evaluates the list 64 pairs at a time with the oracle's expressions on
// the minimum image and adds each pair to its owner's sums with LDS
// atomics.  The sums are FP64 in an order that
// depends on the list, not the oracle's ascending order (deterministic for
// a given state; they differ from the oracle's in the last bits only).
template <int NB, int TD, bool DIGEST>
__global__ __launch_bounds__(NB) BS_NBR_OCCUPANCY void k_nbr_f64(Flocks FL, StateF64 in,
                                                    const double2* __restrict__ rec, StateF64 out,
                                                    const uint32_t* __restrict__ key_sorted,
                                                    const uint32_t* __restrict__ start,
                                                    const uint4* __restrict__ table,
                                                    const uint32_t* __restrict__ nchunks,
                                                    uint32_t* key_out, uint32_t* rank, uint32_t* count,
                                                    double* forces, uint64_t* digest)
{
    using Cfg = NbrCfg<NB, TD>;
    using LC = ListCfg<BS_F64_ROUNDS>;
    constexpr uint32_t kRounds = LC::kRounds, kStep = LC::kStep, kListDrain = LC::kDrain, kListCap = LC::kCap;
    constexpr int NW = NB / 64;
    using LE = ListEntry<NB, TD>;
    // The tile first, at LDS address 0 (the largest variable is placed
    // first), so phase 1's reads need no base-address additions.  tco:
    // coarse positions, tile entry in bits 22-31 of .y; + 4: a finished
    // lane's look-ahead reads (positions stay <= base + here).  Before the
    // first pass stages it, the tile's first words serve as the tile-table
    // scan's scratch.
    __shared__ struct {
        uint2 tco[Cfg::kCap + 4];
        double acc[9][NB];                      // sums of q (cohesion), e/|e| (separation), v (alignment)
    } sh;
    auto& tco = sh.tco;
    auto& acc = sh.acc;
    __shared__ LE wl[NW][kListCap];
    __shared__ uint64_t acc_n[NB];              // nc | ns << 21 | na << 42
    __shared__ uint64_t acc_h[DIGEST ? 3 * NB : 1];
    extern __shared__ __align__(16) unsigned char dyn[];
    static_assert(sizeof(sh.tco) >= 17 * sizeof(uint32_t), "scan scratch inside the tile");

    const uint32_t nch = *nchunks;
    const uint32_t b = xcd_chunk(blockIdx.x, nch);
    if (b >= nch) return;
    const uint4 ch = table[b];
    const uint32_t row = __builtin_amdgcn_readfirstlane(ch.x);
    const uint32_t t0 = __builtin_amdgcn_readfirstlane(ch.y);
    const uint32_t t1 = __builtin_amdgcn_readfirstlane(ch.z);
    const uint32_t f = __builtin_amdgcn_readfirstlane(ch.w);
    const DevParams P = FL.fl[f].P;
    const uint32_t cbase = FL.fl[f].cell_base;
    const uint32_t idbase = DIGEST ? FL.fl[f].boid_base : 0u;   // digests hash flock-local ids
    const int G = P.G;
    const int sx = P.sx;                              // x columns per cell edge
    const uint32_t Gx = (uint32_t)(sx * G);           // x columns of a row
    const int cy = (int)(row % (uint32_t)G), cz = (int)(row / (uint32_t)G);
    const int cxa = (int)((key_sorted[t0] - cbase) % Gx);
    const int cxb = (int)((key_sorted[t1 - 1] - cbase) % Gx);
    const int nseg = cxb - cxa + 1 + 2 * sx;
    const int E = 9 * nseg;
    uint32_t* off = (uint32_t*)dyn;
    uint32_t* sbase = off + (E + 1);

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u, wbase = tid - lane;
    const uint32_t s = t0 + tid;
    const bool active = s < t1;
    double px = 0, py = 0, pz = 0, vx = 0, vy = 0, vz = 0;
    uint32_t myid = 0;
    int L = sx;
    if (active) {
        px = in.x[s]; py = in.y[s]; pz = in.z[s];
        vx = in.vx[s]; vy = in.vy[s]; vz = in.vz[s];
        myid = in.id[s];
        L = (int)((key_sorted[s] - cbase) % Gx) - cxa + sx;
    }
    const double sc = P.lat_scale;
    const uint2 mc = coarse_of(to_lattice(px, sc), to_lattice(py, sc), to_lattice(pz, sc));
    const v2s my_xy = __builtin_bit_cast(v2s, mc.x);
    const uint32_t my_z = mc.y;
    const double vn = norm3(vx, vy, vz);            // owner data stays in registers (read by ds_bpermute)
#pragma unroll
    for (int q = 0; q < 9; ++q) acc[q][tid] = 0.0;
    acc_n[tid] = 0;
    if (DIGEST) { acc_h[tid] = 0; acc_h[NB + tid] = 0; acc_h[2 * NB + tid] = 0; }
    const double F = P.F, half = P.half;
    const bool k_c = P.cone[COH], k_s = P.cone[SEP], k_a = P.cone[ALI];

    build_tile_tables<NB>(start, cbase, G, sx, cxa, nseg, cy, cz, off, sbase, (uint32_t*)tco);
    const uint32_t total = off[E];
    const uint32_t wlo = off[9 * (L - sx)], whi = off[9 * (L + sx + 1)];
    LE* list = wl[__builtin_amdgcn_readfirstlane(tid >> 6)];   // wave-uniform: a scalar base

    // One (owner, candidate) pair per lane: the candidate's record from L2
    // by slot, read ahead of the evaluation.  The records lie in three
    // planes of FL.n double2: (x, y), (z, vx), (vy, vz).
    struct Rec { double2 a, b, c; uint32_t e, o, j; bool valid; };
    const double2* rec_b = rec + FL.n;
    const double2* rec_c = rec + 2 * (size_t)FL.n;
    // sorted slot of a list entry's candidate (pass base `base`): the tile
    // entry (cell) of its tile position gives the cell's slot base
    auto slot_of = [&](uint32_t ent, uint32_t base) {
        const uint32_t pos = ent >> 6;
        return sbase[tco[pos].y >> 22] + base + pos;
    };
    auto load = [&](uint32_t idx, uint32_t lim, uint32_t base) {
        Rec r;
        r.valid = idx < lim;
        r.e = list[min(idx, lim - 1)];
        r.o = wbase + (r.e & 63u);
        r.j = slot_of(r.e, base);
        r.a = at32(rec, r.j);
        r.b = at32(rec_b, r.j);
        r.c = at32(rec_c, r.j);
        return r;
    };
    // The oracle's decisions on the minimum image (src/boid.cpp:27-42,
    // src/boid_simulation.cpp:153-212), the pair added to its owner's sums.
    auto eval = [&](const Rec& r) {
        // every lane of the wave is active here: each reads its owner's
        // registers (the owner is a lane of the same wave)
        const uint32_t o = r.o, ol = r.e & 63u;
        const double ox = lane_read(px, ol), oy = lane_read(py, ol), oz = lane_read(pz, ol);
        const double ovx = lane_read(vx, ol), ovy = lane_read(vy, ol), ovz = lane_read(vz, ol);
        const double on = lane_read(vn, ol);
        const double qx = min_image(r.a.x, ox, F, half);
        const double qy = min_image(r.a.y, oy, F, half);
        const double qz = min_image(r.b.x, oz, F, half);
        const double ex = ox - qx, ey = oy - qy, ez = oz - qz;
        const double s2 = ex * ex + ey * ey + ez * ez;
        const double gx = qx - ox, gy = qy - oy, gz = qz - oz;
        const double dot = ovx * gx + ovy * gy + ovz * gz;
        const double t = sqrt(s2);
        const double c = dot / (on * t);
        const bool cin = c <= 1.0;
        const bool skip = !r.valid || r.j == t0 + o;      // padding, or the owner itself
        const bool in_c = !skip && s2 < P.s_thr[COH] && (!k_c || (c >= P.c_thr[COH] && cin));
        const bool in_s = !skip && s2 < P.s_thr[SEP] && (!k_s || (c >= P.c_thr[SEP] && cin));
        const bool in_a = !skip && s2 < P.s_thr[ALI] && (!k_a || (c >= P.c_thr[ALI] && cin));
        if (in_c) { atomicAdd(&acc[0][o], qx); atomicAdd(&acc[1][o], qy); atomicAdd(&acc[2][o], qz); }
        if (in_s) { atomicAdd(&acc[3][o], ex / t); atomicAdd(&acc[4][o], ey / t); atomicAdd(&acc[5][o], ez / t); }
        if (in_a) { atomicAdd(&acc[6][o], r.b.y); atomicAdd(&acc[7][o], r.c.x); atomicAdd(&acc[8][o], r.c.y); }
        if (in_c | in_s | in_a)
            atomicAdd(&acc_n[o], (uint64_t)in_c | ((uint64_t)in_s << 21) | ((uint64_t)in_a << 42));
        if (DIGEST) {
            const uint64_t h = mix64(in.id[r.j] - idbase);
            if (in_c) atomicAdd(&acc_h[o], h);
            if (in_s) atomicAdd(&acc_h[NB + o], h);
            if (in_a) atomicAdd(&acc_h[2 * NB + o], h);
        }
    };
    auto drain = [&](uint32_t lim, uint32_t base) __attribute__((always_inline)) {
        Rec c[kRounds];
#pragma unroll
        for (uint32_t i = 0; i < kRounds; ++i) c[i] = load(64 * i + lane, lim, base);
        for (uint32_t r0 = 0; r0 < lim; r0 += kStep) {
            Rec nx[kRounds];
#pragma unroll
            for (uint32_t i = 0; i < kRounds; ++i) nx[i] = load(r0 + kStep + 64 * i + lane, lim, base);
#pragma unroll
            for (uint32_t i = 0; i < kRounds; ++i) eval(c[i]);
#pragma unroll
            for (uint32_t i = 0; i < kRounds; ++i) c[i] = nx[i];
        }
    };

    for (uint32_t base = 0; base < total; base += Cfg::kCap) {
        const uint32_t here = min((uint32_t)Cfg::kCap, total - base);
        __syncthreads();
        stage_tile<NB>(off, sbase, E, base, here, [&](uint32_t kk, uint32_t slot, uint32_t e) {
            const uint2 c = coarse_of(to_lattice(in.x[slot], sc), to_lattice(in.y[slot], sc), to_lattice(in.z[slot], sc));
            tco[kk] = make_uint2(c.x, c.y | (e << 22));
        });
        if (tid < 4) tco[here + tid] = far_entry(cy, cz, G);   // the look-ahead's padding
        __syncthreads();
        const uint32_t kb = active ? min(max(wlo, base), base + here) : base;
        const uint32_t ke = active ? min(max(whi, kb), base + here) : base;
        const uint32_t tail = filter_window<kListDrain>(tco, list, kb - base, ke - base, lane, my_xy, my_z, P.coarse_thr,
                                                        [&](uint32_t t) __attribute__((always_inline)) {
            // phase 2 on every full step of rounds; the rest moves to the front
            const uint32_t lim = t - t % kStep, rem = t - lim;
            drain(lim, base);
            LE mv[kRounds];
#pragma unroll
            for (uint32_t i = 0; i < kRounds; ++i) mv[i] = list[lim + min(64 * i + lane, rem)];   // <= t < kListCap
#pragma unroll
            for (uint32_t i = 0; i < kRounds; ++i)
                if (64 * i + lane < rem) list[64 * i + lane] = mv[i];
            return rem;
        });
        if (tail) drain(tail, base);
    }
    if (!active) return;
    // this wave's atomics are complete: LDS operations of a wave execute in order
    const uint64_t cnt = acc_n[tid];
    const uint32_t nc = (uint32_t)(cnt & 0x1FFFFFu), na = (uint32_t)(cnt >> 42);
    double scx = acc[0][tid], scy = acc[1][tid], scz = acc[2][tid];
    double sax = acc[6][tid], say = acc[7][tid], saz = acc[8][tid];
    if (nc != 0) { const double k = (double)nc; scx = scx / k - px; scy = scy / k - py; scz = scz / k - pz; }
    if (na != 0) { const double k = (double)na; sax = sax / k - vx; say = say / k - vy; saz = saz / k - vz; }
    if (DIGEST) {
        uint64_t* d = digest + 4 * (size_t)myid;
        d[0] = cnt;
        d[1] = acc_h[tid]; d[2] = acc_h[NB + tid]; d[3] = acc_h[2 * NB + tid];
    }
    finish_boid(P, cbase, s, myid, px, py, pz, vx, vy, vz, scx, scy, scz, acc[3][tid], acc[4][tid], acc[5][tid],
                sax, say, saz, out, key_out, rank, count, forces);
}

// ---------------------------------------------------------------------------
// Mixed precision.  The difference of two lattice coordinates (unit F/2^32),
// read as int32, is the minimum-image displacement exactly, so the FP32 pair
// test works on exact integers.  A decision is taken in FP32 when it lies
// outside its guard band and re-taken with the oracle's FP64 expression
// otherwise.  Sums are FP64 (cohesion: exact integer lattice sums); phase B
// runs in FP64.
// ---------------------------------------------------------------------------
// FP32 verdict of one rule: in = surely inside, out = surely outside
// (neither: inside a guard band, decided in FP64).  c_dn / c_up are the
// FP32 cosine c moved down / up by the displacement band; c < c_top keeps
// it off 1.  A rule without a view cone has c_lo = c_hi = -inf and c_top =
// +inf, so the cosine decides nothing there (a NaN cosine falls to the FP64
// decision).  Only compares and their lane masks: no per-lane selects.  A
// skipped candidate is surely outside every rule.
__device__ __forceinline__ void rule_bands(const DevParams& P, int r, float s2, float c, float c_dn, float c_up,
                                           bool skip, bool& in, bool& out)
{
    in = (s2 < P.s_lo[r]) & (c_dn >= P.c_hi[r]) & (c < P.c_top[r]) & !skip;
    out = !(s2 < P.s_hi[r]) | (c_up <= P.c_lo[r]) | skip;
}

constexpr double kSepQ = 1073741824.0;                // separation terms in quanta of 2^-30
constexpr double kSepQInv = 1.0 / 1073741824.0;

// Adds a 32-bit integer term to a 64-bit LDS sum: two's-complement sums are
// exact and independent of the order terms arrive in, with an integer
// atomic and no conversion to double per term.
__device__ __forceinline__ void add_i64(int64_t& sum, int32_t v)
{
    atomicAdd((unsigned long long*)&sum, (unsigned long long)(int64_t)v);
}

// A boid's neighbour counts share one 64-bit LDS word, nc | ns << 21 |
// na << 42 (the digest's layout).  Counts stay below 2^20, so bit 41 is
// free: it marks a separation term that is NaN -- the reference's 0/0 unit
// vector of a coincident pair -- which an integer sum cannot carry; the
// owner's separation sum is then NaN, as the reference's is.
constexpr uint64_t kNanSep = 1ull << 41;
constexpr uint64_t kCounts = ~kNanSep;

template <int NB, int TD, bool DIGEST>
__global__ __launch_bounds__(NB) BS_NBR_OCCUPANCY void k_nbr_mixed(Flocks FL, StateF64 in, StateF64 out,
                                                      Lattice lat,
                                                      const uint32_t* __restrict__ key_sorted,
                                                      const uint32_t* __restrict__ start,
                                                      const uint4* __restrict__ table,
                                                      const uint32_t* __restrict__ nchunks,
                                                      uint32_t* key_out, uint32_t* rank, uint32_t* count,
                                                      double* forces, uint64_t* digest, uint32_t* recheck)
{
    using Cfg = NbrCfg<NB, TD>;
    using LC = ListCfg<BS_MIXED_ROUNDS>;
    constexpr uint32_t kRounds = LC::kRounds, kStep = LC::kStep, kListDrain = LC::kDrain, kListCap = LC::kCap;
    constexpr int NW = NB / 64;
    using LE = ListEntry<NB, TD>;
    // The tile first, at LDS address 0 (the largest variable is placed
    // first), so phase 1's reads need no base-address additions.  tco:
    // coarse positions, tile entry e in bits 22-31 of .y (z in bits 0-15);
    // + 4: a finished lane's look-ahead reads.  Before the first pass stages
    // it, the tile's first words serve as the tile-table scan's scratch.
    __shared__ struct {
        uint2 tco[Cfg::kCap + 4];
        int64_t acc[9][NB];                     // exact integer sums: cohesion, separation, alignment
    } sh;
    auto& tco = sh.tco;
    auto& acc = sh.acc;
    __shared__ LE wl[NW][kListCap];
    __shared__ uint64_t acc_n[NB];              // nc | ns << 21 | na << 42
    __shared__ uint64_t acc_h[DIGEST ? 3 * NB : 1];
    extern __shared__ __align__(16) unsigned char dyn[];
    static_assert(sizeof(sh.tco) >= 17 * sizeof(uint32_t), "scan scratch inside the tile");

    const uint32_t nch = *nchunks;
    const uint32_t b = xcd_chunk(blockIdx.x, nch);
    if (b >= nch) return;
    const uint4 ch = table[b];
    const uint32_t row = __builtin_amdgcn_readfirstlane(ch.x);
    const uint32_t t0 = __builtin_amdgcn_readfirstlane(ch.y);
    const uint32_t t1 = __builtin_amdgcn_readfirstlane(ch.z);
    const uint32_t f = __builtin_amdgcn_readfirstlane(ch.w);
    const DevParams P = FL.fl[f].P;
    const uint32_t cbase = FL.fl[f].cell_base;
    const uint32_t idbase = DIGEST ? FL.fl[f].boid_base : 0u;   // digests hash flock-local ids
    const int G = P.G;
    const int sx = P.sx;                              // x columns per cell edge
    const uint32_t Gx = (uint32_t)(sx * G);           // x columns of a row
    const int cy = (int)(row % (uint32_t)G), cz = (int)(row / (uint32_t)G);
    const int cxa = (int)((key_sorted[t0] - cbase) % Gx);
    const int cxb = (int)((key_sorted[t1 - 1] - cbase) % Gx);
    const int nseg = cxb - cxa + 1 + 2 * sx;
    const int E = 9 * nseg;
    uint32_t* off = (uint32_t*)dyn;
    uint32_t* sbase = off + (E + 1);

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u, wbase = tid - lane;
    const uint32_t s = t0 + tid;
    const bool active = s < t1;
    uint4 me = make_uint4(0, 0, 0, 0);
    int L = sx;
    // Owner data that evaluating lanes read from this lane's registers: the
    // lattice record and the FP32 unit velocity.  A velocity whose FP32
    // square vanishes or is not finite (or a NaN state) has no FP32
    // direction: its unit vector is NaN, which makes every cosine of the
    // owner NaN, so no cone verdict is taken in FP32 and each pair within
    // reach of a rule is decided by the FP64 expression.  The FP64 state is
    // read again for phase B instead of being held through both phases.
    float3 fu = make_float3(0.f, 0.f, 0.f);
    if (active) {
        me = lat.p[s];
        L = (int)((key_sorted[s] - cbase) % Gx) - cxa + sx;
        const float fvx = (float)in.vx[s], fvy = (float)in.vy[s], fvz = (float)in.vz[s];
        const float fvv = fvx * fvx + fvy * fvy + fvz * fvz;
        const float rs = (fvv > 1e-30f) ? __builtin_amdgcn_rsqf(fvv) : __builtin_nanf("");
        fu = make_float3(fvx * rs, fvy * rs, fvz * rs);
    }
    const uint2 mc = coarse_of(me.x, me.y, me.z);
    const v2s my_xy = __builtin_bit_cast(v2s, mc.x);
    const uint32_t my_z = mc.y;
#pragma unroll
    for (int q = 0; q < 9; ++q) acc[q][tid] = 0;
    acc_n[tid] = 0;
    if (DIGEST) { acc_h[tid] = 0; acc_h[NB + tid] = 0; acc_h[2 * NB + tid] = 0; }
    uint32_t nre = 0;

    build_tile_tables<NB>(start, cbase, G, sx, cxa, nseg, cy, cz, off, sbase, (uint32_t*)tco);
    const uint32_t total = off[E];
    const uint32_t wlo = off[9 * (L - sx)], whi = off[9 * (L + sx + 1)];
    LE* list = wl[__builtin_amdgcn_readfirstlane(tid >> 6)];   // wave-uniform: a scalar base
    constexpr double kSub = 1.0 / (double)(1 << kSubBits);

    // The separation term of a pair: the unit vector e/|e| in quanta of
    // 2^-30, in FP64 (an FP32 unit vector carries ~1e-7 per term, and dense
    // clusters sum tens of them), on the displacement refined by the
    // sub-lattice residuals (a close pair's direction would otherwise carry
    // the 1/2-unit rounding of both ends).
    auto sep_term = [&](uint32_t o, const uint4& m, const uint4& q) {
        const int32_t ix = (int32_t)(q.x - m.x), iy = (int32_t)(q.y - m.y), iz = (int32_t)(q.z - m.z);
        const double ex = fma((double)(sub_field<0>(q.w) - sub_field<0>(m.w)), kSub, (double)ix);
        const double ey = fma((double)(sub_field<1>(q.w) - sub_field<1>(m.w)), kSub, (double)iy);
        const double ez = fma((double)(sub_field<2>(q.w) - sub_field<2>(m.w)), kSub, (double)iz);
        const double e2 = fma(ez, ez, fma(ey, ey, ex * ex));
        double y = __builtin_amdgcn_rsq(e2);           // estimate, then one Newton step
        y = y * fma(-0.5 * e2, y * y, 1.5);
        const double t = -y * kSepQ;
        // |e/|e|| * 2^30 <= 2^30; a coincident pair (e = 0) adds 0 and marks
        // the owner's separation sum NaN
        add_i64(acc[3][o], (int32_t)rint(ex * t));
        add_i64(acc[4][o], (int32_t)rint(ey * t));
        add_i64(acc[5][o], (int32_t)rint(ez * t));
        if (!(e2 > 0.0)) atomicOr((unsigned long long*)&acc_n[o], (unsigned long long)kNanSep);
    };
    // Adds the pair (owner o at lattice m, candidate record q, w, slot j)
    // to the owner's sums.  Every term is an integer, so the 64-bit LDS
    // sums are exact and independent of the order pairs arrive in:
    // cohesion adds lattice displacements, alignment quantised velocities
    // (lattice record), separation the quantised unit vector.
    auto accumulate = [&](uint32_t o, const uint4& m, const uint4& q, const float4& w, uint32_t j, bool in_c,
                          bool in_s, bool in_a) {
        const int32_t ix = (int32_t)(q.x - m.x), iy = (int32_t)(q.y - m.y), iz = (int32_t)(q.z - m.z);
        if (in_c) {
            add_i64(acc[0][o], ix); add_i64(acc[1][o], iy); add_i64(acc[2][o], iz);
        }
        if (in_s) sep_term(o, m, q);
        if (in_a) {
            add_i64(acc[6][o], (int32_t)w.x); add_i64(acc[7][o], (int32_t)w.y); add_i64(acc[8][o], (int32_t)w.z);
        }
        if (in_c | in_s | in_a)
            atomicAdd(&acc_n[o], (uint64_t)in_c | ((uint64_t)in_s << 21) | ((uint64_t)in_a << 42));
        if (DIGEST) {
            const uint64_t h = mix64(in.id[j] - idbase);
            if (in_c) atomicAdd(&acc_h[o], h);
            if (in_s) atomicAdd(&acc_h[NB + o], h);
            if (in_a) atomicAdd(&acc_h[2 * NB + o], h);
        }
    };

    // One (owner, candidate) pair per lane: the owner's data from LDS, the
    // candidate's lattice record from L2 by slot; all reads issued ahead
    // of the evaluation.
    struct Rec { uint4 q, m; float4 w; float3 ou; uint32_t e, o, j; bool valid; };
    // sorted slot of a list entry's candidate (pass base `base`): the tile
    // entry (cell) of its tile position gives the cell's slot base
    auto slot_of = [&](uint32_t ent, uint32_t base) {
        const uint32_t pos = ent >> 6;
        return sbase[tco[pos].y >> 22] + base + pos;
    };
    auto load = [&](const LE* lst, uint32_t idx, uint32_t lim, uint32_t base) {
        Rec r;
        r.valid = idx < lim;
        r.e = lst[min(idx, lim - 1)];
        r.o = wbase + (r.e & 63u);
        r.j = slot_of(r.e, base);
        r.q = at32(lat.p, r.j);
        r.w = at32(lat.v, r.j);
        return r;
    };
    // FP32 verdicts outside the guard bands, the oracle's FP64 ones inside.
    auto eval = [&](Rec r) {
        // every lane of the wave is active here: each reads its owner's registers
        const uint32_t ol = r.e & 63u;
        r.m = make_uint4(lane_read(me.x, ol), lane_read(me.y, ol), lane_read(me.z, ol), lane_read(me.w, ol));
        r.ou = make_float3(lane_read(fu.x, ol), lane_read(fu.y, ol), lane_read(fu.z, ol));
        const int32_t ix = (int32_t)(r.q.x - r.m.x), iy = (int32_t)(r.q.y - r.m.y), iz = (int32_t)(r.q.z - r.m.z);
        const float gx = (float)ix, gy = (float)iy, gz = (float)iz;
        const float s2 = fmaf(gz, gz, fmaf(gy, gy, gx * gx));
        const float inv_d = __builtin_amdgcn_rsqf(s2);
        const float c = fmaf(r.ou.z, gz, fmaf(r.ou.y, gy, r.ou.x * gx)) * inv_d;
        const float band = P.band_k * inv_d;
        const float c_dn = c - band, c_up = c + band;
        // padding, the owner itself, or a boid whose state is NaN (bit 31 of
        // the residual word: its lattice coordinates mean nothing, and the
        // reference's NaN distance makes it nobody's neighbour)
        const bool skip = !r.valid | (r.j == t0 + r.o) | ((int32_t)(r.q.w | r.m.w) < 0);
        bool in_c, out_c, in_s, out_s, in_a, out_a;
        if (P.same_cone) {
            // one view cone for all three rules (a flock-uniform branch): its
            // compares are shared
            const bool cin = (c_dn >= P.c_hi[0]) & (c < P.c_top[0]);
            const bool cout = c_up <= P.c_lo[0];
            in_c = (s2 < P.s_lo[COH]) & cin & !skip;
            out_c = !(s2 < P.s_hi[COH]) | cout | skip;
            in_s = (s2 < P.s_lo[SEP]) & cin & !skip;
            out_s = !(s2 < P.s_hi[SEP]) | cout | skip;
            in_a = (s2 < P.s_lo[ALI]) & cin & !skip;
            out_a = !(s2 < P.s_hi[ALI]) | cout | skip;
        } else {
            rule_bands(P, COH, s2, c, c_dn, c_up, skip, in_c, out_c);
            rule_bands(P, SEP, s2, c, c_dn, c_up, skip, in_s, out_s);
            rule_bands(P, ALI, s2, c, c_dn, c_up, skip, in_a, out_a);
        }
        const bool sure = (in_c | out_c) & (in_s | out_s) & (in_a | out_a);
        if (!sure) {
            const uint32_t so = t0 + r.o;
            const double ovx = in.vx[so], ovy = in.vy[so], ovz = in.vz[so];
            exact_decide(P, in.x[so], in.y[so], in.z[so], ovx, ovy, ovz, norm3(ovx, ovy, ovz), in.x[r.j], in.y[r.j],
                         in.z[r.j], in_s, in_a, in_c);
            ++nre;
        }
        accumulate(r.o, r.m, r.q, r.w, r.j, in_c, in_s, in_a);
    };
    // Evaluates list[0, lim) in rounds of 64 lanes, kRounds rounds per
    // step, the next kRounds fetched while these are evaluated.
    auto drain = [&](uint32_t lim, uint32_t base) __attribute__((always_inline)) {
        Rec c[kRounds];
#pragma unroll
        for (uint32_t i = 0; i < kRounds; ++i) c[i] = load(list, 64 * i + lane, lim, base);
        for (uint32_t r0 = 0; r0 < lim; r0 += kStep) {
            Rec nx[kRounds];
#pragma unroll
            for (uint32_t i = 0; i < kRounds; ++i) nx[i] = load(list, r0 + kStep + 64 * i + lane, lim, base);
#pragma unroll
            for (uint32_t i = 0; i < kRounds; ++i) eval(c[i]);
#pragma unroll
            for (uint32_t i = 0; i < kRounds; ++i) c[i] = nx[i];
        }
    };

    for (uint32_t base = 0; base < total; base += Cfg::kCap) {
        const uint32_t here = min((uint32_t)Cfg::kCap, total - base);
        __syncthreads();
        stage_tile<NB>(off, sbase, E, base, here, [&](uint32_t kk, uint32_t slot, uint32_t e) {
            const uint4 q = lat.p[slot];
            const uint2 c = coarse_of(q.x, q.y, q.z);
            tco[kk] = make_uint2(c.x, c.y | (e << 22));
        });
        if (tid < 4) tco[here + tid] = far_entry(cy, cz, G);   // the look-ahead's padding
        __syncthreads();
        // this lane's window in the staged range [base, base + here) (empty
        // for idle lanes); positions stay <= base + here, so the look-ahead
        // reads stay inside tco
        const uint32_t kb = active ? min(max(wlo, base), base + here) : base;
        const uint32_t ke = active ? min(max(whi, kb), base + here) : base;
        // phase 1 (filter_window), phase 2 whenever the list is full
        const uint32_t tail = filter_window<kListDrain>(tco, list, kb - base, ke - base, lane, my_xy, my_z, P.coarse_thr,
                                                        [&](uint32_t t) __attribute__((always_inline)) {
            // phase 2 on every full step of rounds; the rest moves to the front
            const uint32_t lim = t - t % kStep, rem = t - lim;
            drain(lim, base);
            LE mv[kRounds];
#pragma unroll
            for (uint32_t i = 0; i < kRounds; ++i) mv[i] = list[lim + min(64 * i + lane, rem)];   // <= t < kListCap
#pragma unroll
            for (uint32_t i = 0; i < kRounds; ++i)
                if (64 * i + lane < rem) list[64 * i + lane] = mv[i];
            return rem;
        });
        if (tail) drain(tail, base);
    }
    if (nre) atomicAdd(recheck, nre);
    if (!active) return;
    const double px = in.x[s], py = in.y[s], pz = in.z[s];
    const double vx = in.vx[s], vy = in.vy[s], vz = in.vz[s];
    const uint32_t myid = in.id[s];
    // this wave's atomics are complete: LDS operations of a wave execute in order
    const uint64_t word = acc_n[tid], cnt = word & kCounts;
    const uint32_t nc = (uint32_t)(cnt & 0x1FFFFFu), na = (uint32_t)(cnt >> 42);
    const double qnan = __builtin_nan("");
    // dv_coh = mean(q) - p = mean(q - p): the lattice sums are already relative
    double dcx = 0, dcy = 0, dcz = 0, dax = 0, day = 0, daz = 0;
    if (nc != 0) {
        const double k = P.unit / (double)nc;
        dcx = (double)acc[0][tid] * k; dcy = (double)acc[1][tid] * k; dcz = (double)acc[2][tid] * k;
    }
    if (na != 0) {
        const double k = (double)na;
        dax = (double)acc[6][tid] * P.vq / k - vx; day = (double)acc[7][tid] * P.vq / k - vy;
        daz = (double)acc[8][tid] * P.vq / k - vz;
    }
    double dsx = (double)acc[3][tid] * kSepQInv, dsy = (double)acc[4][tid] * kSepQInv;
    double dsz = (double)acc[5][tid] * kSepQInv;
    if (word & kNanSep) dsx = dsy = dsz = qnan;
    if (DIGEST) {
        uint64_t* d = digest + 4 * (size_t)myid;
        d[0] = cnt;
        d[1] = acc_h[tid]; d[2] = acc_h[NB + tid]; d[3] = acc_h[2 * NB + tid];
    }
    finish_boid(P, cbase, s, myid, px, py, pz, vx, vy, vz, dcx, dcy, dcz, dsx, dsy, dsz, dax, day, daz, out, key_out,
                rank, count, forces);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline uint32_t round8(uint32_t g) { return (g + 7u) & ~7u; }

template <int NB, int TD>
static void launch_f64_nb(Flocks F, size_t dyn, dim3 grid, StateF64 in, const double2* rec, StateF64 out,
                          const uint32_t* key_sorted,
                          const uint32_t* start, const uint4* table, const uint32_t* nchunks, uint32_t* key_out,
                          uint32_t* rank, uint32_t* count, double* forces, uint64_t* digest, hipStream_t stream)
{
    if (digest)
        hipLaunchKernelGGL((k_nbr_f64<NB, TD, true>), grid, dim3(NB), dyn, stream, F, in, rec, out, key_sorted, start,
                           table, nchunks, key_out, rank, count, forces, digest);
    else
        hipLaunchKernelGGL((k_nbr_f64<NB, TD, false>), grid, dim3(NB), dyn, stream, F, in, rec, out, key_sorted, start,
                           table, nchunks, key_out, rank, count, forces, digest);
}

template <int NB, int TD>
static void launch_mixed_nb(Flocks F, size_t dyn, dim3 grid, StateF64 in, StateF64 out, Lattice lat,
                            const uint32_t* key_sorted, const uint32_t* start, const uint4* table,
                            const uint32_t* nchunks, uint32_t* key_out, uint32_t* rank, uint32_t* count,
                            double* forces, uint64_t* digest, uint32_t* recheck, hipStream_t stream)
{
    if (digest)
        hipLaunchKernelGGL((k_nbr_mixed<NB, TD, true>), grid, dim3(NB), dyn, stream, F, in, out, lat, key_sorted,
                           start, table, nchunks, key_out, rank, count, forces, digest, recheck);
    else
        hipLaunchKernelGGL((k_nbr_mixed<NB, TD, false>), grid, dim3(NB), dyn, stream, F, in, out, lat, key_sorted,
                           start, table, nchunks, key_out, rank, count, forces, digest, recheck);
}

// Instantiates LAUNCH<NB, TD> for the supported (workgroup size, tile depth)
// pairs; false for any other pair.
#define BS_NBR_DISPATCH(nb, tile, CALL)                                        \
    do {                                                                      \
        if (nb == 64 && tile == 5) { CALL(64, 5); break; }                    \
        if (nb == 64 && tile == 6) { CALL(64, 6); break; }                    \
        if (nb == 64 && tile == 9) { CALL(64, 9); break; }                    \
        if (nb == 64 && tile == 12) { CALL(64, 12); break; }                  \
        if (nb == 128 && tile == 5) { CALL(128, 5); break; }                  \
        if (nb == 128 && tile == 6) { CALL(128, 6); break; }                  \
        if (nb == 128 && tile == 9) { CALL(128, 9); break; }                  \
        if (nb == 128 && tile == 12) { CALL(128, 12); break; }                \
        if (nb == 256 && tile == 5) { CALL(256, 5); break; }                  \
        if (nb == 256 && tile == 6) { CALL(256, 6); break; }                  \
        if (nb == 256 && tile == 9) { CALL(256, 9); break; }                  \
        if (nb == 256 && tile == 12) { CALL(256, 12); break; }                \
        return hipErrorInvalidValue;                                          \
    } while (0)

hipError_t launch_nbr_f64(Flocks F, int nb, int tile, int max_cols, StateF64 in, const double2* rec,
                          StateF64 out, const uint32_t* key_sorted, const uint32_t* start, const uint4* table,
                          const uint32_t* nchunks, uint32_t max_chunks, uint32_t* key_out, uint32_t* rank,
                          uint32_t* count, double* forces, uint64_t* digest, hipStream_t stream)
{
    if (F.n == 0) return hipSuccess;
    if (max_cols < 3 || max_cols > kChunkSpanMax + 2 * kSubXMax) return hipErrorInvalidValue;
    const dim3 grid(round8(max_chunks));
    const size_t dyn = tile_table_bytes(max_cols);
#define BS_CALL(NB_, TD_) launch_f64_nb<NB_, TD_>(F, dyn, grid, in, rec, out, key_sorted, start, table, nchunks, \
                                                   key_out, rank, count, forces, digest, stream)
    BS_NBR_DISPATCH(nb, tile, BS_CALL);
#undef BS_CALL
    return hipGetLastError();
}

hipError_t launch_nbr_mixed(Flocks F, int nb, int tile, int max_cols, StateF64 in, StateF64 out, Lattice lat,
                            const uint32_t* key_sorted, const uint32_t* start, const uint4* table,
                            const uint32_t* nchunks, uint32_t max_chunks, uint32_t* key_out, uint32_t* rank,
                            uint32_t* count, double* forces, uint64_t* digest, uint32_t* recheck,
                            hipStream_t stream)
{
    if (F.n == 0) return hipSuccess;
    if (max_cols < 3 || max_cols > kChunkSpanMax + 2 * kSubXMax) return hipErrorInvalidValue;
    const dim3 grid(round8(max_chunks));
    const size_t dyn = tile_table_bytes(max_cols);
#define BS_CALL(NB_, TD_) launch_mixed_nb<NB_, TD_>(F, dyn, grid, in, out, lat, key_sorted, start, table, nchunks, \
                                                     key_out, rank, count, forces, digest, recheck, stream)
    BS_NBR_DISPATCH(nb, tile, BS_CALL);
#undef BS_CALL
    return hipGetLastError();
}

}  // namespace boids
