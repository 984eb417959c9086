// boids_metrics.hip — flock metrics on the GPU (polarisation, mean
// nearest-neighbour distance, cluster count), for long-run statistical
// comparison of engine modes.  Distances use the periodic metric of the
// reference's analysis tools (tools/calc_distance.pyx:6-14: per axis
// min(|dx|, F - |dx|)).  Not on the update path; clarity over speed.
//
// Input: n positions / velocities in id order (double, AoS x,y,z).  A private
// uniform grid (edge >= eps, >= 3 cells per axis when possible) is built
// with the engine's scan and scatter kernels.
#include "boids_internal.h"

#include <math.h>

namespace boids {

struct MGrid {
    int G;                    // cells per axis (>= 1)
    double F, edge, inv_edge;
    uint32_t n;
};

__device__ __forceinline__ uint32_t m_axis(double x, const MGrid& g)
{
    double c = floor(x * g.inv_edge);
    if (!(c >= 0.0)) c = 0.0;
    if (c > (double)(g.G - 1)) c = (double)(g.G - 1);
    return (uint32_t)c;
}

__device__ __forceinline__ double m_dist2(const double* a, double bx, double by, double bz, double F)
{
    double s = 0.0;
    const double b[3] = {bx, by, bz};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double d = fabs(a[k] - b[k]);
        if (F - d < d) d = F - d;
        s += d * d;
    }
    return s;
}

__global__ __launch_bounds__(kBlock) void k_m_hash(MGrid g, const double* pos3, uint32_t* key, uint32_t* count)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.n) return;
    const double* p = pos3 + 3 * (size_t)i;
    const uint32_t G = (uint32_t)g.G;
    const uint32_t c = m_axis(p[0], g) + G * (m_axis(p[1], g) + G * m_axis(p[2], g));
    key[i] = c;
    atomicAdd(&count[c], 1u);
}

// cell-ordered copy of the positions: cpos[t] = pos[perm[t]], cid[t] = perm[t]
__global__ __launch_bounds__(kBlock) void k_m_order(MGrid g, const double* pos3, const uint32_t* perm,
                                                    double4* cpos)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= g.n) return;
    const uint32_t i = perm[t];
    const double* p = pos3 + 3 * (size_t)i;
    cpos[t] = make_double4(p[0], p[1], p[2], __longlong_as_double((long long)i));   // w carries the id bits
}

__device__ __forceinline__ uint32_t m_id(const double4& q) { return (uint32_t)__double_as_longlong(q.w); }

// float SoA state (reference-faithful mode) -> double AoS
__global__ __launch_bounds__(kBlock) void k_m_from_f32(uint32_t n, StateF32 st, double* pos3, double* vel3)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    pos3[3 * (size_t)i] = st.x[i]; pos3[3 * (size_t)i + 1] = st.y[i]; pos3[3 * (size_t)i + 2] = st.z[i];
    vel3[3 * (size_t)i] = st.vx[i]; vel3[3 * (size_t)i + 1] = st.vy[i]; vel3[3 * (size_t)i + 2] = st.vz[i];
}

__global__ __launch_bounds__(kBlock) void k_m_fill(uint32_t* a, uint32_t n, uint32_t v)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_m_iota(uint32_t* a, uint32_t n)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) a[i] = i;
}

// Visits every distinct cell with Chebyshev ring distance <= r around
// (cx,cy,cz) exactly once (wrapped); fn(cell) per cell.  Only ring == r
// when `only_ring`.
template <typename Fn>
__device__ __forceinline__ void m_ring(const MGrid& g, int cx, int cy, int cz, int r, bool only_ring, Fn fn)
{
    const int G = g.G;
    const int span = 2 * r + 1;
    const int lo = span >= G ? 0 : -r, hi = span >= G ? G - 1 : r;   // distinct offsets per axis
    for (int dz = lo; dz <= hi; ++dz)
        for (int dy = lo; dy <= hi; ++dy)
            for (int dx = lo; dx <= hi; ++dx) {
                if (only_ring && span < G) {
                    const int m = max(abs(dx), max(abs(dy), abs(dz)));
                    if (m != r) continue;
                }
                int x = cx + dx, y = cy + dy, z = cz + dz;
                x = ((x % G) + G) % G; y = ((y % G) + G) % G; z = ((z % G) + G) % G;
                fn((uint32_t)x + (uint32_t)G * ((uint32_t)y + (uint32_t)G * (uint32_t)z));
            }
}

// Nearest-neighbour distance of every boid (id order), by growing rings:
// after ring r every point farther than r*edge from the boid's cell has been
// excluded, so the search stops once the best distance is <= r*edge.
__global__ __launch_bounds__(kBlock) void k_m_nn(MGrid g, const double* pos3, const uint32_t* key,
                                                 const uint32_t* start, const double4* cpos, double* nn)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.n) return;
    const double* p = pos3 + 3 * (size_t)i;
    const uint32_t G = (uint32_t)g.G, c = key[i];
    const int cx = (int)(c % G), cy = (int)((c / G) % G), cz = (int)(c / (G * G));
    double best = INFINITY;
    auto scan_cell = [&](uint32_t cc) {
        for (uint32_t t = start[cc]; t < start[cc + 1]; ++t) {
            const double4 q = cpos[t];
            if (m_id(q) == i) continue;
            const double d = m_dist2(p, q.x, q.y, q.z, g.F);
            if (d < best) best = d;
        }
    };
    const int rmax = (g.G + 1) / 2;
    for (int r = 0; r <= rmax; ++r) {
        m_ring(g, cx, cy, cz, r, true, scan_cell);
        if (2 * r + 1 >= g.G) break;                              // every cell visited
        const double reach = r * g.edge;
        if (best <= reach * reach) break;
    }
    nn[i] = sqrt(best);
}

__device__ __forceinline__ uint32_t m_find(uint32_t* parent, uint32_t x)
{
    while (true) {
        const uint32_t p = parent[x];
        if (p == x) return x;
        const uint32_t gp = parent[p];
        if (gp != p) parent[x] = gp;                              // path halving: pointers only move rootward
        x = p;
    }
}

__device__ __forceinline__ void m_union(uint32_t* parent, uint32_t a, uint32_t b)
{
    while (true) {
        a = m_find(parent, a);
        b = m_find(parent, b);
        if (a == b) return;
        if (a > b) { const uint32_t t = a; a = b; b = t; }
        const uint32_t old = atomicCAS(&parent[b], b, a);         // hook the larger root under the smaller
        if (old == b) return;
        b = old;
    }
}

// Links every pair closer than eps (edge >= eps, so ring 1 suffices).
__global__ __launch_bounds__(kBlock) void k_m_link(MGrid g, const double* pos3, const uint32_t* key,
                                                   const uint32_t* start, const double4* cpos, double eps2,
                                                   uint32_t* parent)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.n) return;
    const double* p = pos3 + 3 * (size_t)i;
    const uint32_t G = (uint32_t)g.G, c = key[i];
    const int cx = (int)(c % G), cy = (int)((c / G) % G), cz = (int)(c / (G * G));
    m_ring(g, cx, cy, cz, 1, false, [&](uint32_t cc) {
        for (uint32_t t = start[cc]; t < start[cc + 1]; ++t) {
            const double4 q = cpos[t];
            const uint32_t j = m_id(q);
            if (j <= i) continue;
            if (m_dist2(p, q.x, q.y, q.z, g.F) < eps2) m_union(parent, i, j);
        }
    });
}

__global__ __launch_bounds__(kBlock) void k_m_sizes(uint32_t n, uint32_t* parent, uint32_t* size)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    atomicAdd(&size[m_find(parent, i)], 1u);
}

// One workgroup: polarisation sums, NN sum, cluster count and largest size,
// each accumulated in a fixed order (deterministic).
__global__ __launch_bounds__(1024) void k_m_reduce(uint32_t n, const double* vel3, const double* nn,
                                                   const uint32_t* parent, const uint32_t* size,
                                                   uint32_t min_size, double* out5)
{
    __shared__ double sh[4][1024];
    __shared__ unsigned long long sc[1024], sl[1024];
    double vx = 0, vy = 0, vz = 0, s = 0;
    unsigned long long cl = 0, lg = 0;
    for (uint32_t i = threadIdx.x; i < n; i += 1024) {
        const double* v = vel3 + 3 * (size_t)i;
        const double m = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (m > 0) { vx += v[0] / m; vy += v[1] / m; vz += v[2] / m; }
        s += nn[i];
        if (parent[i] == i) {
            if (size[i] >= min_size && size[i] > 0) cl++;
            if (size[i] > lg) lg = size[i];
        }
    }
    sh[0][threadIdx.x] = vx; sh[1][threadIdx.x] = vy; sh[2][threadIdx.x] = vz; sh[3][threadIdx.x] = s;
    sc[threadIdx.x] = cl; sl[threadIdx.x] = lg;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
            sc[threadIdx.x] += sc[threadIdx.x + w];
            if (sl[threadIdx.x + w] > sl[threadIdx.x]) sl[threadIdx.x] = sl[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out5[0] = sqrt(sh[0][0] * sh[0][0] + sh[1][0] * sh[1][0] + sh[2][0] * sh[2][0]) / (double)n;
        out5[1] = n > 1 ? sh[3][0] / (double)n : INFINITY;
        out5[2] = (double)sc[0];
        out5[3] = (double)sl[0];
    }
}

static inline uint32_t mblocks(uint32_t n) { return (n + kBlock - 1) / kBlock; }

hipError_t launch_f32_to_f64(uint32_t n, StateF32 st, double* pos3, double* vel3, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_m_from_f32, dim3(mblocks(n)), dim3(kBlock), 0, stream, n, st, pos3, vel3);
    return hipGetLastError();
}

// Workspace layout: one function carves the regions, both for sizing
// (base == nullptr) and for the launch, so the two cannot disagree.
struct MetricsWs {
    double4* cpos;      // n    cell-ordered positions (id bits in w)
    double* nn;         // n
    uint32_t *key, *perm, *permid, *ids, *parent, *size;   // n each
    uint32_t *count, *start;                               // cells each
    uint32_t* partials; // scan tiles
    double* out;        // 4
    size_t bytes;       // total, from the (unaligned) base
};

static_assert(sizeof(double4) == 32, "double4 is 32 bytes");

static MetricsWs metrics_layout(char* base, uint32_t n, int G)
{
    const size_t cells = (size_t)G * G * G + 1;
    const size_t ntiles = (cells + 4095) / 4096;
    MetricsWs m;
    uintptr_t p = ((uintptr_t)base + 255) & ~(uintptr_t)255;
    auto take = [&](size_t bytes) {
        const uintptr_t at = p;
        p = (p + bytes + 255) & ~(uintptr_t)255;
        return at;
    };
    m.cpos = (double4*)take(sizeof(double4) * (size_t)n);
    m.nn = (double*)take(sizeof(double) * (size_t)n);
    m.key = (uint32_t*)take(4 * (size_t)n);
    m.perm = (uint32_t*)take(4 * (size_t)n);
    m.permid = (uint32_t*)take(4 * (size_t)n);
    m.ids = (uint32_t*)take(4 * (size_t)n);
    m.parent = (uint32_t*)take(4 * (size_t)n);
    m.size = (uint32_t*)take(4 * (size_t)n);
    m.count = (uint32_t*)take(4 * cells);
    m.start = (uint32_t*)take(4 * cells);
    m.partials = (uint32_t*)take(4 * ntiles);
    m.out = (double*)take(sizeof(double) * 4);
    m.bytes = (size_t)(p - (uintptr_t)base);
    return m;
}

size_t metrics_workspace_bytes(uint32_t n, int G)
{
    return metrics_layout(nullptr, n, G).bytes + 256;   // + worst-case alignment of a real base
}

hipError_t launch_metrics(uint32_t n, double F, int G, double eps, uint32_t min_size, const double* pos3,
                          const double* vel3, void* ws, size_t ws_bytes, double** out4_dev, hipStream_t stream)
{
    const MetricsWs m = metrics_layout((char*)ws, n, G);
    if ((char*)(m.out + 4) > (char*)ws + ws_bytes) return hipErrorInvalidValue;   // layout must fit
    *out4_dev = m.out;
    if (n == 0) return hipSuccess;
    MGrid g;
    g.G = G; g.F = F; g.edge = F / G; g.inv_edge = G / F; g.n = n;
    const uint32_t cells = (uint32_t)((size_t)G * G * G + 1);
    double4* cpos = m.cpos;
    double* nn = m.nn;
    uint32_t *key = m.key, *perm = m.perm, *permid = m.permid, *ids = m.ids, *parent = m.parent, *size = m.size;
    uint32_t *count = m.count, *start = m.start, *partials = m.partials;
    double* out5_dev = m.out;
    hipLaunchKernelGGL(k_m_fill, dim3(mblocks(cells)), dim3(kBlock), 0, stream, count, cells, 0u);
    hipLaunchKernelGGL(k_m_hash, dim3(mblocks(n)), dim3(kBlock), 0, stream, g, pos3, key, count);
    hipError_t st = launch_scan(count, start, cells, partials, stream);
    if (st != hipSuccess) return st;
    hipLaunchKernelGGL(k_m_iota, dim3(mblocks(n)), dim3(kBlock), 0, stream, ids, n);
    st = launch_scatter(n, key, start, count, ids, perm, permid, stream);
    if (st != hipSuccess) return st;
    hipLaunchKernelGGL(k_m_order, dim3(mblocks(n)), dim3(kBlock), 0, stream, g, pos3, (const uint32_t*)permid, cpos);
    hipLaunchKernelGGL(k_m_nn, dim3(mblocks(n)), dim3(kBlock), 0, stream, g, pos3, (const uint32_t*)key,
                       (const uint32_t*)start, (const double4*)cpos, nn);
    hipLaunchKernelGGL(k_m_iota, dim3(mblocks(n)), dim3(kBlock), 0, stream, parent, n);
    hipLaunchKernelGGL(k_m_fill, dim3(mblocks(n)), dim3(kBlock), 0, stream, size, n, 0u);
    hipLaunchKernelGGL(k_m_link, dim3(mblocks(n)), dim3(kBlock), 0, stream, g, pos3, (const uint32_t*)key,
                       (const uint32_t*)start, (const double4*)cpos, eps * eps, parent);
    hipLaunchKernelGGL(k_m_sizes, dim3(mblocks(n)), dim3(kBlock), 0, stream, n, parent, size);
    hipLaunchKernelGGL(k_m_reduce, dim3(1), dim3(1024), 0, stream, n, vel3, (const double*)nn,
                       (const uint32_t*)parent, (const uint32_t*)size, min_size, out5_dev);
    return hipGetLastError();
}

}  // namespace boids
