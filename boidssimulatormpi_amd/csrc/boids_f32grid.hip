// boids_f32grid.hip — reference-faithful float grid step (BS_MODE_GRID_F32REF):
// what boidsim_grid computes (src/boid_simulation_grid.cpp:165-293), bit for
// bit.  The reference's grid engine equals its all-pairs engine exactly
// because each boid's candidates are visited in ascending index order
// (:186 copies the primary cell's list; phase A of :153-212 sums in float in
// visit order).  Here the cell list is a counting sort of the float state by
// (cell, id); one wave per owner collects the boids of its 27 cells within
// the largest radius, sorts them by id and adds their terms in ascending id
// (a serial 27-way merge by id serves owners with more boids in reach than
// the wave's list holds), so every float sum adds the same terms in the same
// order as the reference.  The state stays in id order (as boidsim keeps
// it).
//
// Compiled with -ffp-contract=off: every multiply and add is rounded on its
// own, as the reference's x86-64 build computes it.
#include "boids_device.h"

namespace boids {

// The three rule decisions of src/boid.cpp:27-42 for owner p (velocity v,
// |v| = vn) and candidate position t (raw, before the minimum image), with
// the exact float thresholds (s < s_thr <=> sqrtf(s) < r; c >= c_thr &&
// c <= 1 <=> acosf(c) <= pi a).  q receives the shifted candidate position.
struct PairF32 {
    float qx, qy, qz, ex, ey, ez, s;
    bool in_c, in_s, in_a;
};

__device__ __forceinline__ PairF32 decide_f32(const DevParams& P, float px, float py, float pz, float vx, float vy,
                                              float vz, float vn, float tx, float ty, float tz)
{
    PairF32 r;
    r.qx = min_image(tx, px, P.F, P.half);
    r.qy = min_image(ty, py, P.F, P.half);
    r.qz = min_image(tz, pz, P.F, P.half);
    r.ex = px - r.qx; r.ey = py - r.qy; r.ez = pz - r.qz;
    r.s = r.ex * r.ex + r.ey * r.ey + r.ez * r.ez;
    r.in_c = r.s < P.s_thr_f[COH];
    r.in_s = r.s < P.s_thr_f[SEP];
    r.in_a = r.s < P.s_thr_f[ALI];
    if ((r.in_c && P.cone[COH]) || (r.in_s && P.cone[SEP]) || (r.in_a && P.cone[ALI])) {
        const float gx = r.qx - px, gy = r.qy - py, gz = r.qz - pz;
        const float dot = vx * gx + vy * gy + vz * gz;
        const float c = dot / (vn * sqrtf(r.s));
        if (P.cone[COH]) r.in_c = r.in_c && c >= P.c_thr_f[COH] && c <= 1.0f;
        if (P.cone[SEP]) r.in_s = r.in_s && c >= P.c_thr_f[SEP] && c <= 1.0f;
        if (P.cone[ALI]) r.in_a = r.in_a && c >= P.c_thr_f[ALI] && c <= 1.0f;
    }
    return r;
}

__device__ __forceinline__ uint32_t cell_f32(const DevParams& P, float x, float y, float z)
{
    const uint32_t G = (uint32_t)P.G;
    return axis_cell((double)x, P.inv_edge, P.G) +
           G * (axis_cell((double)y, P.inv_edge, P.G) + G * axis_cell((double)z, P.inv_edge, P.G));
}

// Cell key of every boid (id order) and its arrival rank in the cell.
__global__ __launch_bounds__(kBlock) void k_f32_hash(DevParams P, StateF32 st, uint32_t* key, uint32_t* rank,
                                                     uint32_t* count)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t c = cell_f32(P, st.x[i], st.y[i], st.z[i]);
    key[i] = c;
    rank[i] = atomicAdd(&count[c], 1u);
}

// Boids grouped by cell in arrival order: perm[start[c] + rank] = id.
__global__ __launch_bounds__(kBlock) void k_f32_scatter(uint32_t n, const uint32_t* __restrict__ key,
                                                        const uint32_t* __restrict__ rank,
                                                        const uint32_t* __restrict__ start, uint32_t* perm)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    perm[start[key[i]] + rank[i]] = i;
}

// Final slot = cell start + the id's rank among the cell's ids, so ids
// ascend inside every cell; the slot's record is the candidate position
// with the id in .w, and the (unshifted) velocity.  perm's runs of 64 slots
// are sorted by (cell, id) (launch_sort_runs): the rank is the slot's place
// in its own part of the cell plus a binary search in each other part.
__global__ __launch_bounds__(kBlock) void k_f32_cells(uint32_t n, StateF32 st, const uint32_t* __restrict__ key,
                                                      const uint32_t* __restrict__ start,
                                                      const uint32_t* __restrict__ perm, float4* rec, float4* vrec)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const uint32_t me = perm[t];
    const uint32_t c = key[me];
    const uint32_t lo = start[c], hi = start[c + 1];
    const uint32_t run = t >> 6;
    uint32_t r = t - max(lo, run << 6);
    for (uint32_t v = lo >> 6; (v << 6) < hi; ++v) {
        if (v == run) continue;
        const uint32_t a = max(lo, v << 6), b = min(hi, (v + 1) << 6);
        r += count_below(perm + a, b - a, me);
    }
    const uint32_t f = lo + r;
    rec[f] = make_float4(st.x[me], st.y[me], st.z[me], __builtin_bit_cast(float, me));
    vrec[f] = make_float4(st.vx[me], st.vy[me], st.vz[me], 0.f);
}

// Sums of one owner in the reference's order, the update's outputs.
struct SumsF32 {
    float scx = 0, scy = 0, scz = 0, ssx = 0, ssy = 0, ssz = 0, sax = 0, say = 0, saz = 0;
    int nc = 0, ns = 0, na = 0;
    uint64_t hc = 0, hs = 0, ha = 0;
};

// Serial form: the owner's in-range boids visited in ascending id by a lazy
// 27-way merge.  Per neighbour cell k: cur[k] = the cell's next unvisited
// slot, endp[k] its end, head[k] = the id at cur[k] once cur[k] has been
// moved to the next in-range candidate (~0u: the cell is exhausted).  Each
// merge step takes the smallest head, adds that boid to the sums and moves
// its cell on.  cur/endp/head: 27 words each of scratch.
__device__ SumsF32 merge_serial(const DevParams& P, const float4* __restrict__ rec, const float4* __restrict__ vrec,
                                const uint32_t* __restrict__ start, uint32_t i, float px, float py, float pz,
                                float vx, float vy, float vz, float vn, int cx, int cy, int cz, uint32_t* cur,
                                uint32_t* endp, uint32_t* head)
{
    const int G = P.G;
    auto advance = [&](int k) {
        uint32_t u = cur[k];
        const uint32_t e = endp[k];
        uint32_t h = ~0u;
        for (; u < e; ++u) {
            const float4 q = rec[u];
            const uint32_t j = __builtin_bit_cast(uint32_t, q.w);
            if (j == i) continue;
            const PairF32 r = decide_f32(P, px, py, pz, vx, vy, vz, vn, q.x, q.y, q.z);
            if (r.in_c | r.in_s | r.in_a) { h = j; break; }
        }
        cur[k] = u;
        head[k] = h;
    };
    for (int k = 0; k < 27; ++k) {
        int x = cx + k % 3 - 1, y = cy + (k / 3) % 3 - 1, z = cz + k / 9 - 1;
        x += x < 0 ? G : (x >= G ? -G : 0);
        y += y < 0 ? G : (y >= G ? -G : 0);
        z += z < 0 ? G : (z >= G ? -G : 0);
        const uint32_t cc = (uint32_t)x + (uint32_t)G * ((uint32_t)y + (uint32_t)G * (uint32_t)z);
        cur[k] = start[cc];
        endp[k] = start[cc + 1];
        advance(k);
    }
    SumsF32 S;
    for (;;) {
        uint32_t best = ~0u;
        int kb = 0;
        for (int k = 0; k < 27; ++k) {
            const uint32_t h = head[k];
            if (h < best) { best = h; kb = k; }
        }
        if (best == ~0u) break;
        const uint32_t u = cur[kb];
        const float4 q = rec[u];
        const PairF32 r = decide_f32(P, px, py, pz, vx, vy, vz, vn, q.x, q.y, q.z);
        if (r.in_c) { S.nc++; S.scx += r.qx; S.scy += r.qy; S.scz += r.qz; S.hc += mix64(best); }
        if (r.in_s) {
            const float t = sqrtf(r.s);
            S.ns++; S.ssx += r.ex / t; S.ssy += r.ey / t; S.ssz += r.ez / t; S.hs += mix64(best);
        }
        if (r.in_a) {
            const float4 w = vrec[u];
            S.na++; S.sax += w.x; S.say += w.y; S.saz += w.z; S.ha += mix64(best);
        }
        cur[kb] = u + 1;
        advance(kb);
    }
    return S;
}

// Wave form: one one-wave workgroup per owner (id order).  (1) The lanes
// test the candidates of the 27 cells, 64 at a time in (cell, slot) order,
// against the largest radius and append the ids of those within it to an
// LDS list.  (2) The list is sorted (bitonic network in LDS).  (3) For 64
// sorted boids at a time the lanes read their states by id, take the rule
// decisions and write the pairs' nine float terms to an LDS table, 0 where
// a rule does not hold; lane c < 9 then adds column c of the table in
// ascending id, so
// each float sum adds the reference's terms in the reference's order (a
// running sum that starts at +0 is never -0, so adding +0 leaves it
// unchanged bit for bit).  An owner with more boids within the largest
// radius than the list holds takes the serial merge on lane 0.
constexpr uint32_t kF32HitCap = 1024;        // boids per owner sorted in LDS

// Orders this wave's LDS writes before its later LDS reads (a wave's LDS
// operations execute in order; this keeps the compiler from moving them).
__device__ __forceinline__ void wave_lds_order()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <typename T>
__device__ __forceinline__ T lane_of(T v, uint32_t l)
{
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), (int)l));
}

// 64-bit sum over the wave (butterfly), in every lane.
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v)
{
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, d, 64);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), d, 64);
        v += ((uint64_t)hi << 32) | lo;
    }
    return v;
}

__global__ __launch_bounds__(64) void k_f32_grid_wave(DevParams P, StateF32 in, StateF32 out,
                                                      const float4* __restrict__ rec,
                                                      const float4* __restrict__ vrec,
                                                      const uint32_t* __restrict__ key,
                                                      const uint32_t* __restrict__ start, float* forces,
                                                      uint64_t* digest)
{
    __shared__ uint32_t hb[kF32HitCap];
    __shared__ float terms[9][64];                     // term c of the block's pair l: terms[c][l]
    const uint32_t lane = threadIdx.x;
    const uint32_t i = blockIdx.x;                     // the grid has one workgroup per boid
    const float px = in.x[i], py = in.y[i], pz = in.z[i];
    const float vx = in.vx[i], vy = in.vy[i], vz = in.vz[i];
    const float vn = norm3(vx, vy, vz);
    const int G = P.G;
    const uint32_t c = key[i];
    const int cx = (int)(c % (uint32_t)G), cy = (int)((c / (uint32_t)G) % (uint32_t)G),
              cz = (int)(c / ((uint32_t)G * (uint32_t)G));

    // lane k < 27: neighbour cell k's slot range; exclusive prefix of the sizes
    uint32_t ca = 0, cn = 0;
    if (lane < 27) {
        const int k = (int)lane;
        int x = cx + k % 3 - 1, y = cy + (k / 3) % 3 - 1, z = cz + k / 9 - 1;
        x += x < 0 ? G : (x >= G ? -G : 0);
        y += y < 0 ? G : (y >= G ? -G : 0);
        z += z < 0 ? G : (z >= G ? -G : 0);
        const uint32_t cc = (uint32_t)x + (uint32_t)G * ((uint32_t)y + (uint32_t)G * (uint32_t)z);
        ca = start[cc];
        cn = start[cc + 1] - ca;
    }
    uint32_t incl = cn;
    for (uint32_t d = 1; d < 32; d <<= 1) {
        const uint32_t t = __shfl_up(incl, d, 64);
        if (lane >= d) incl += t;
    }
    const uint32_t excl = incl - cn;                   // non-decreasing over lanes 0..26
    const uint32_t M = (uint32_t)__builtin_amdgcn_readlane((int)incl, 26);

    // (1) candidates within the largest radius, appended in (cell, slot) order;
    // each batch's records are fetched while the previous batch is tested
    auto record_of = [&](uint32_t m0) {
        const uint32_t m = min(m0 + lane, M - 1);
        // the last cell k in 0..26 with excl[k] <= m (binary lifting over the lanes)
        uint32_t kk = 0;
#pragma unroll
        for (uint32_t st = 16; st > 0; st >>= 1) {
            const uint32_t pr = kk + st;
            const uint32_t ex = (uint32_t)__shfl((int)excl, (int)min(pr, 26u), 64);
            kk = (pr <= 26u && ex <= m) ? pr : kk;
        }
        const uint32_t slot = (uint32_t)__shfl((int)ca, (int)kk, 64) + (m - (uint32_t)__shfl((int)excl, (int)kk, 64));
        return rec[slot];
    };
    uint32_t H = 0;
    float4 qn = M ? record_of(0) : make_float4(0.f, 0.f, 0.f, 0.f);
    for (uint32_t m0 = 0; m0 < M; m0 += 64) {
        const float4 q = qn;
        if (m0 + 64 < M) qn = record_of(m0 + 64);
        const uint32_t j = __builtin_bit_cast(uint32_t, q.w);
        const float ex = px - min_image(q.x, px, P.F, P.half);
        const float ey = py - min_image(q.y, py, P.F, P.half);
        const float ez = pz - min_image(q.z, pz, P.F, P.half);
        const bool hit = (m0 + lane < M) && (j != i) && (ex * ex + ey * ey + ez * ez < P.s_max_f);
        const uint64_t mask = __builtin_amdgcn_ballot_w64(hit);
        const uint32_t pos = H + __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                           __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
        if (hit && pos < kF32HitCap) hb[pos] = j;
        H += (uint32_t)__popcll(mask);
    }

    SumsF32 S;
    if (H > kF32HitCap) {
        // more boids within reach than the list holds: the serial merge (lane 0)
        if (lane != 0) return;
        S = merge_serial(P, rec, vrec, start, i, px, py, pz, vx, vy, vz, vn, cx, cy, cz, hb, hb + 27, hb + 54);
    } else if (H > 0) {
        // (3), one block: the terms of the sorted boids h0 .. h0 + 63 (lane
        // l holds the id of boid h0 + l in j); lane c < 9 sums column c
        float sum = 0.f;
        uint64_t hc = 0, hs = 0, ha = 0;                // this lane's pairs' id hashes
        auto block = [&](uint32_t h0, uint32_t j) __attribute__((always_inline)) {
            float t[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            bool fc = false, fs = false, fa = false;
            if (h0 + lane < H) {
                const PairF32 r = decide_f32(P, px, py, pz, vx, vy, vz, vn, in.x[j], in.y[j], in.z[j]);
                fc = r.in_c; fs = r.in_s; fa = r.in_a;
                if (fc) { t[0] = r.qx; t[1] = r.qy; t[2] = r.qz; }
                if (fs) {
                    const float d = sqrtf(r.s);
                    t[3] = r.ex / d; t[4] = r.ey / d; t[5] = r.ez / d;
                }
                if (fa) { t[6] = in.vx[j]; t[7] = in.vy[j]; t[8] = in.vz[j]; }
                if (digest) {
                    const uint64_t hh = mix64(j);
                    hc += fc ? hh : 0; hs += fs ? hh : 0; ha += fa ? hh : 0;
                }
            }
            S.nc += __popcll(__builtin_amdgcn_ballot_w64(fc));
            S.ns += __popcll(__builtin_amdgcn_ballot_w64(fs));
            S.na += __popcll(__builtin_amdgcn_ballot_w64(fa));
#pragma unroll
            for (int k = 0; k < 9; ++k) terms[k][lane] = t[k];
            wave_lds_order();
            if (lane < 9) {
                const float* col = terms[lane];
                const uint32_t cnt = min(64u, H - h0);
                for (uint32_t l = 0; l < cnt; ++l) sum += col[l];   // ascending id: the reference's order
            }
            wave_lds_order();
        };
        if (H <= 4 * 64) {
            // (2) bitonic sort in registers: element e = 64 r + lane in
            // register r, at least 64 elements, padded with ~0u
            constexpr uint32_t R = 4;
            uint32_t P2 = 64;
            while (P2 < H) P2 <<= 1;
            uint32_t kr[R];
#pragma unroll
            for (uint32_t r = 0; r < R; ++r) kr[r] = 64 * r + lane < H ? hb[64 * r + lane] : ~0u;
            for (uint32_t size = 2; size <= P2; size <<= 1) {
                for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                    if (stride >= 64) {                      // partner in register r ^ (stride / 64), same lane
                        auto cx = [&](uint32_t& a, uint32_t& b, uint32_t r) {
                            const bool asc = ((64 * r + lane) & size) == 0;
                            const uint32_t lo = min(a, b), hi = max(a, b);
                            a = asc ? lo : hi;
                            b = asc ? hi : lo;
                        };
                        if (stride == 64) {
                            cx(kr[0], kr[1], 0);
                            if (P2 > 128) cx(kr[2], kr[3], 2);
                        } else {                             // stride 128: P2 = 256
                            cx(kr[0], kr[2], 0);
                            cx(kr[1], kr[3], 1);
                        }
                    } else {                                 // partner in lane ^ stride, same register
#pragma unroll
                        for (uint32_t r = 0; r < R; ++r) {
                            if (64 * r >= P2) continue;
                            const uint32_t pk = (uint32_t)__shfl_xor((int)kr[r], (int)stride, 64);
                            const bool keep_min = ((lane & stride) == 0) == (((64 * r + lane) & size) == 0);
                            kr[r] = keep_min ? min(kr[r], pk) : max(kr[r], pk);
                        }
                    }
                }
            }
            block(0, kr[0]);
            if (H > 64) block(64, kr[1]);
            if (H > 128) block(128, kr[2]);
            if (H > 192) block(192, kr[3]);
        } else {
            // (2) bitonic sort of the ids in LDS, padded to a power of two
            uint32_t P2 = 1;
            while (P2 < H) P2 <<= 1;
            for (uint32_t t = H + lane; t < P2; t += 64) hb[t] = ~0u;
            wave_lds_order();
            for (uint32_t size = 2; size <= P2; size <<= 1) {
                for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                    for (uint32_t t = lane; t < P2 / 2; t += 64) {
                        const uint32_t a = 2 * t - (t & (stride - 1)), b = a + stride;
                        const uint32_t ka = hb[a], kb = hb[b];
                        if ((ka > kb) == ((a & size) == 0)) { hb[a] = kb; hb[b] = ka; }
                    }
                    wave_lds_order();
                }
            }
            for (uint32_t h0 = 0; h0 < H; h0 += 64) block(h0, h0 + lane < H ? hb[h0 + lane] : 0u);
        }
        S.scx = lane_of(sum, 0); S.scy = lane_of(sum, 1); S.scz = lane_of(sum, 2);
        S.ssx = lane_of(sum, 3); S.ssy = lane_of(sum, 4); S.ssz = lane_of(sum, 5);
        S.sax = lane_of(sum, 6); S.say = lane_of(sum, 7); S.saz = lane_of(sum, 8);
        if (digest) { S.hc = wave_sum_u64(hc); S.hs = wave_sum_u64(hs); S.ha = wave_sum_u64(ha); }
        if (lane != 0) return;
    } else if (lane != 0) {
        return;
    }
    float ox = px, oy = py, oz = pz, ovx = vx, ovy = vy, ovz = vz;
    float scx = S.scx, scy = S.scy, scz = S.scz, sax = S.sax, say = S.say, saz = S.saz;
    if (S.nc != 0) { const float k = (float)S.nc; scx = scx / k - px; scy = scy / k - py; scz = scz / k - pz; }
    if (S.na != 0) { const float k = (float)S.na; sax = sax / k - vx; say = say / k - vy; saz = saz / k - vz; }
    const float wc = P.w_f[COH], ws = P.w_f[SEP], wa = P.w_f[ALI];
    const float dx = (scx * wc + S.ssx * ws) + sax * wa;
    const float dy = (scy * wc + S.ssy * ws) + say * wa;
    const float dz = (scz * wc + S.ssz * ws) + saz * wa;
    if (forces) {
        float* f = forces + 9 * (size_t)i;
        f[0] = scx; f[1] = scy; f[2] = scz; f[3] = S.ssx; f[4] = S.ssy; f[5] = S.ssz; f[6] = sax; f[7] = say; f[8] = saz;
    }
    if (digest) {
        uint64_t* d = digest + 4 * (size_t)i;
        d[0] = (uint64_t)S.nc | ((uint64_t)S.ns << 21) | ((uint64_t)S.na << 42);
        d[1] = S.hc; d[2] = S.hs; d[3] = S.ha;
    }
    integrate<float>(P, ox, oy, oz, ovx, ovy, ovz, dx, dy, dz);
    out.x[i] = ox; out.y[i] = oy; out.z[i] = oz; out.vx[i] = ovx; out.vy[i] = ovy; out.vz[i] = ovz;
}

static inline uint32_t blocks_of(uint32_t n, uint32_t b) { return (n + b - 1) / b; }

hipError_t launch_f32_grid_sort(const DevParams& P, StateF32 st, uint32_t* key, uint32_t* rank, uint32_t* count,
                                uint32_t* start, uint32_t* partials, uint32_t* perm, float4* rec, float4* vrec,
                                hipStream_t stream)
{
    if (P.n == 0) return hipSuccess;
    const uint32_t ncell = (uint32_t)P.G * (uint32_t)P.G * (uint32_t)P.G;
    hipLaunchKernelGGL(k_f32_hash, dim3(blocks_of(P.n, kBlock)), dim3(kBlock), 0, stream, P, st, key, rank, count);
    hipError_t e = launch_scan(count, start, ncell + 1, partials, stream);    // clears count for the next step
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_f32_scatter, dim3(blocks_of(P.n, kBlock)), dim3(kBlock), 0, stream, P.n,
                       (const uint32_t*)key, (const uint32_t*)rank, (const uint32_t*)start, perm);
    e = launch_sort_runs(P.n, key, perm, perm, stream);     // ids in id order: index = id
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_f32_cells, dim3(blocks_of(P.n, kBlock)), dim3(kBlock), 0, stream, P.n, st,
                       (const uint32_t*)key, (const uint32_t*)start, (const uint32_t*)perm, rec, vrec);
    return hipGetLastError();
}

hipError_t launch_f32_grid_step(const DevParams& P, StateF32 in, StateF32 out, const float4* rec,
                                const float4* vrec, const uint32_t* key, const uint32_t* start, float* forces,
                                uint64_t* digest, hipStream_t stream)
{
    if (P.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_f32_grid_wave, dim3(P.n), dim3(64), 0, stream, P, in, out, rec, vrec, key, start, forces,
                       digest);
    return hipGetLastError();
}

}  // namespace boids
