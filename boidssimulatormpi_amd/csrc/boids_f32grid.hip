// boids_f32grid.hip — reference-faithful float grid step (BS_MODE_GRID_F32REF):
// what boidsim_grid computes (src/boid_simulation_grid.cpp:165-293), bit for
// bit.  The reference's grid engine equals its all-pairs engine exactly
// because each boid's candidates are visited in ascending index order
// (:186 copies the primary cell's list; phase A of :153-212 sums in float in
// visit order).  Here the cell list is a counting sort of the float state by
// (cell, id); an owner visits the in-range boids of its 27 cells in ascending
// id by a lazy 27-way merge (each cell's ids ascend; the smallest head is
// taken next), so every float sum adds the same terms in the same order as
// the reference.  The state stays in id order (as boidsim keeps it).
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
// with the id in .w, and the (unshifted) velocity.
__global__ __launch_bounds__(kBlock) void k_f32_cells(uint32_t n, StateF32 st, const uint32_t* __restrict__ key,
                                                      const uint32_t* __restrict__ start,
                                                      const uint32_t* __restrict__ perm, float4* rec, float4* vrec)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= n) return;
    const uint32_t me = perm[t];
    const uint32_t c = key[me];
    const uint32_t lo = start[c], hi = start[c + 1];
    uint32_t r = 0;
    for (uint32_t u = lo; u < hi; ++u) r += perm[u] < me ? 1u : 0u;
    const uint32_t f = lo + r;
    rec[f] = make_float4(st.x[me], st.y[me], st.z[me], __builtin_bit_cast(float, me));
    vrec[f] = make_float4(st.vx[me], st.vy[me], st.vz[me], 0.f);
}

// One owner per thread (id order).  Per neighbour cell k of the 27: cur[k]
// = the cell's next unvisited slot, end[k] its end, head[k] = the id at
// cur[k] once cur[k] has been moved to the next in-range candidate (~0u:
// the cell is exhausted).  Each merge step takes the smallest head, adds
// that boid to the sums (the reference's order), and moves its cell on.
constexpr int kF32GridBlock = 64;

template <int NB>
__global__ __launch_bounds__(NB) void k_f32_grid(DevParams P, StateF32 in, StateF32 out,
                                                 const float4* __restrict__ rec, const float4* __restrict__ vrec,
                                                 const uint32_t* __restrict__ key,
                                                 const uint32_t* __restrict__ start, float* forces,
                                                 uint64_t* digest)
{
    __shared__ uint32_t cur[27][NB], endp[27][NB], head[27][NB];
    const uint32_t tid = threadIdx.x;
    const uint32_t i = blockIdx.x * NB + tid;
    if (i >= P.n) return;                          // no barrier below: every thread is on its own
    const float px = in.x[i], py = in.y[i], pz = in.z[i];
    const float vx = in.vx[i], vy = in.vy[i], vz = in.vz[i];
    const float vn = norm3(vx, vy, vz);
    const int G = P.G;
    const uint32_t c = key[i];
    const int cx = (int)(c % (uint32_t)G), cy = (int)((c / (uint32_t)G) % (uint32_t)G),
              cz = (int)(c / ((uint32_t)G * (uint32_t)G));

    // moves cell k's cursor to its next in-range candidate (not the owner)
    auto advance = [&](int k) {
        uint32_t u = cur[k][tid];
        const uint32_t e = endp[k][tid];
        uint32_t h = ~0u;
        for (; u < e; ++u) {
            const float4 q = rec[u];
            const uint32_t j = __builtin_bit_cast(uint32_t, q.w);
            if (j == i) continue;
            const PairF32 r = decide_f32(P, px, py, pz, vx, vy, vz, vn, q.x, q.y, q.z);
            if (r.in_c | r.in_s | r.in_a) { h = j; break; }
        }
        cur[k][tid] = u;
        head[k][tid] = h;
    };
    for (int k = 0; k < 27; ++k) {
        int x = cx + k % 3 - 1, y = cy + (k / 3) % 3 - 1, z = cz + k / 9 - 1;
        x += x < 0 ? G : (x >= G ? -G : 0);
        y += y < 0 ? G : (y >= G ? -G : 0);
        z += z < 0 ? G : (z >= G ? -G : 0);
        const uint32_t cc = (uint32_t)x + (uint32_t)G * ((uint32_t)y + (uint32_t)G * (uint32_t)z);
        cur[k][tid] = start[cc];
        endp[k][tid] = start[cc + 1];
        advance(k);
    }
    float scx = 0, scy = 0, scz = 0, ssx = 0, ssy = 0, ssz = 0, sax = 0, say = 0, saz = 0;
    int nc = 0, ns = 0, na = 0;
    uint64_t hc = 0, hs = 0, ha = 0;
    for (;;) {
        uint32_t best = ~0u;
        int kb = 0;
        for (int k = 0; k < 27; ++k) {
            const uint32_t h = head[k][tid];
            if (h < best) { best = h; kb = k; }
        }
        if (best == ~0u) break;
        const uint32_t u = cur[kb][tid];
        const float4 q = rec[u];
        const PairF32 r = decide_f32(P, px, py, pz, vx, vy, vz, vn, q.x, q.y, q.z);
        if (r.in_c) { nc++; scx += r.qx; scy += r.qy; scz += r.qz; hc += mix64(best); }
        if (r.in_s) {
            const float t = sqrtf(r.s);
            ns++; ssx += r.ex / t; ssy += r.ey / t; ssz += r.ez / t; hs += mix64(best);
        }
        if (r.in_a) {
            const float4 w = vrec[u];
            na++; sax += w.x; say += w.y; saz += w.z; ha += mix64(best);
        }
        cur[kb][tid] = u + 1;
        advance(kb);
    }
    float ox = px, oy = py, oz = pz, ovx = vx, ovy = vy, ovz = vz;
    if (nc != 0) { const float k = (float)nc; scx = scx / k - px; scy = scy / k - py; scz = scz / k - pz; }
    if (na != 0) { const float k = (float)na; sax = sax / k - vx; say = say / k - vy; saz = saz / k - vz; }
    const float wc = P.w_f[COH], ws = P.w_f[SEP], wa = P.w_f[ALI];
    const float dx = (scx * wc + ssx * ws) + sax * wa;
    const float dy = (scy * wc + ssy * ws) + say * wa;
    const float dz = (scz * wc + ssz * ws) + saz * wa;
    if (forces) {
        float* f = forces + 9 * (size_t)i;
        f[0] = scx; f[1] = scy; f[2] = scz; f[3] = ssx; f[4] = ssy; f[5] = ssz; f[6] = sax; f[7] = say; f[8] = saz;
    }
    if (digest) {
        uint64_t* d = digest + 4 * (size_t)i;
        d[0] = (uint64_t)nc | ((uint64_t)ns << 21) | ((uint64_t)na << 42);
        d[1] = hc; d[2] = hs; d[3] = ha;
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
    hipLaunchKernelGGL(k_f32_cells, dim3(blocks_of(P.n, kBlock)), dim3(kBlock), 0, stream, P.n, st,
                       (const uint32_t*)key, (const uint32_t*)start, (const uint32_t*)perm, rec, vrec);
    return hipGetLastError();
}

hipError_t launch_f32_grid_step(const DevParams& P, StateF32 in, StateF32 out, const float4* rec,
                                const float4* vrec, const uint32_t* key, const uint32_t* start, float* forces,
                                uint64_t* digest, hipStream_t stream)
{
    if (P.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_f32_grid<kF32GridBlock>, dim3(blocks_of(P.n, kF32GridBlock)), dim3(kF32GridBlock), 0,
                       stream, P, in, out, rec, vrec, key, start, forces, digest);
    return hipGetLastError();
}

}  // namespace boids
