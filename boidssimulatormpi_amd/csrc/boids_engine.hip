// boids_engine.hip — host side of libboids_engine.so: the C ABI declared in
// include/boids_engine.h, exact-threshold derivation, init generators and
// the per-step launch sequence.  Kernels live in boids_kernels.hip.
#include "boids_engine.h"
#include "boids_internal.h"

// Automatic LDS tile depth of the neighbour kernels (candidates per thread
// and pass): 5, 6, 9 or 12.
#ifndef BS_MIXED_TILE
#define BS_MIXED_TILE 5
#endif
#ifndef BS_F64_TILE
#define BS_F64_TILE 5
#endif
// x columns per cell edge of the grid modes (DevParams::sx): both paths bin
// x at half the cell edge, a 2.5-edge window along x instead of 3 (with
// 12-column chunks; with 16-column chunks the FP64 kernel was faster on the
// reference's cells).
#ifndef BS_MIXED_SUBX
#define BS_MIXED_SUBX 2
#endif
#ifndef BS_F64_SUBX
#define BS_F64_SUBX 2
#endif
static_assert(BS_MIXED_SUBX >= 1 && BS_MIXED_SUBX <= boids::kSubXMax && BS_F64_SUBX >= 1 &&
                  BS_F64_SUBX <= boids::kSubXMax, "x columns per cell edge");
// Most x columns a neighbour-kernel chunk covers (k_rows_pass cuts runs of
// boids short there): narrow chunks give the lanes of a wave windows of more
// equal length (phase 1 runs as long as the longest), wide ones fewer tile
// restagings; measured best for both paths: 12 half-columns (6 cell edges).
#ifndef BS_MIXED_CHUNK_SPAN
#define BS_MIXED_CHUNK_SPAN 12
#endif
#ifndef BS_F64_CHUNK_SPAN
#define BS_F64_CHUNK_SPAN 12
#endif
static_assert(BS_MIXED_CHUNK_SPAN >= 1 && BS_MIXED_CHUNK_SPAN <= boids::kChunkSpanMax && BS_F64_CHUNK_SPAN >= 1 &&
                  BS_F64_CHUNK_SPAN <= boids::kChunkSpanMax, "chunk span");

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cstdint>
#include <fstream>
#include <map>
#include <mutex>
#include <string>
#include <vector>

using namespace boids;

static thread_local std::string g_create_error;

// Static device scratch of the allocation-free entry point
// (bs_metrics_flock_noalloc): kScratchBytes on every GPU the library runs
// on, held until the process ends.  On the host, `scratch` names a host-side
// stand-in, not a device address: its address on each device is looked up
// once (hipGetSymbolAddress, with that device current) and only that address
// is handed to kernels.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static std::map<int, void*> g_scratch_addr;   // device -> address of `scratch` there

// Address of `scratch` on `device`, which must be the current device;
// nullptr (and *why set) if the lookup fails.
static void* scratch_address(int device, std::string* why)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_addr.find(device);
    if (it != g_scratch_addr.end()) return it->second;
    void* ptr = nullptr;
    const hipError_t st = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (st != hipSuccess) {
        *why = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(st);
        return nullptr;
    }
    g_scratch_addr[device] = ptr;
    return ptr;
}

struct bs_engine {
    bs_params prm{};
    int mode = 0;          // effective mode
    int G = 0;
    double edge = 0.0;
    DevParams P{};
    int device = 0;
    hipStream_t stream = nullptr;
    bool has_state = false;
    uint64_t steps = 0;
    uint64_t device_bytes = 0;
    std::string err;
    std::vector<void*> allocs;

    // state buffers (A = current, B = next)
    StateF32 fa{}, fb{};
    StateF64 da{}, db{};
    // grid
    uint32_t *key = nullptr, *key_sorted = nullptr, *count = nullptr, *start = nullptr;
    uint32_t *partials = nullptr, *nchunks = nullptr, *perm = nullptr, *permid = nullptr;
    uint32_t* rank = nullptr;          // arrival rank of each slot in its new cell
    uint32_t *row_count = nullptr, *row_off = nullptr;   // chunks per (y,z) row, and their offsets
    uint4* table = nullptr;
    uint32_t max_chunks = 0, ncell1 = 0;
    // flock table (one entry for a single-flock engine)
    std::vector<FlockDesc> flocks;
    std::vector<double> flock_field;
    FlockDesc* flocks_dev = nullptr;
    uint32_t* flock_of = nullptr;      // slot -> flock (batches only)
    uint32_t* row_flock = nullptr;     // (y,z) row -> flock (batches only)
    uint32_t rows = 0;
    uint64_t groups = 0;               // (row, span group) pairs over all flocks
    int max_g = 0, max_span = 0;
    int sx = 1;                        // x columns per cell edge of every flock (grid modes)
    int cspan = kChunkSpanMax;         // most x columns a neighbour-kernel chunk covers
    int nb = 64;                       // neighbour-kernel workgroup size
    int tile = 12;                     // LDS tile depth (candidates per thread and pass)
    int fill = 100;                    // chunk size target, percent of nb
    uint32_t chunk = 256;              // boids per chunk at most (fill% of nb)
    bool batch = false;
    // one captured step (BS_FLAG_GRAPH)
    hipGraphExec_t graph = nullptr;
    bool capturing = false;
    Lattice lat{};
    double2* rec = nullptr;            // FP64 grid: interleaved pair records
    float4 *rec4 = nullptr, *vrec4 = nullptr;   // float grid: sorted positions (+ id) and velocities
    uint32_t* recheck = nullptr;
    uint32_t* overflow = nullptr;      // chunk-table count beyond its capacity (0: none)
    void* scratch = nullptr;           // this device's static scratch (allocation-free entry points)
    std::string scratch_why;           // why it is unavailable
    // outputs
    void* forces = nullptr;   // float or double, 9n, id order
    uint64_t* digest = nullptr;
    double *tmp_pos = nullptr, *tmp_vel = nullptr;

    // timing
    struct Pending { int kernel; hipEvent_t a, b; };
    std::vector<hipEvent_t> free_events;
    std::vector<Pending> pending;
    double ms[BS_K_COUNT] = {0};
    uint64_t launches[BS_K_COUNT] = {0};
};

// ----------------------------------------------------------------------------
// small helpers
// ----------------------------------------------------------------------------
static int fail(bs_engine* e, int code, const std::string& msg)
{
    if (e) e->err = msg; else g_create_error = msg;
    return code;
}

static int hip_fail(bs_engine* e, hipError_t st, const char* what)
{
    return fail(e, BS_E_HIP, std::string(what) + ": " + hipGetErrorString(st));
}

#define HIP_TRY(e, expr)                                               \
    do {                                                               \
        hipError_t st_ = (expr);                                       \
        if (st_ != hipSuccess) return hip_fail((e), st_, #expr);       \
    } while (0)

template <typename T>
static int dev_alloc(bs_engine* e, T** p, size_t count)
{
    size_t bytes = std::max<size_t>(1, count * sizeof(T));
    void* ptr = nullptr;
    hipError_t st = hipMalloc(&ptr, bytes);
    if (st != hipSuccess) return fail(e, BS_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    e->allocs.push_back(ptr);
    e->device_bytes += bytes;
    *p = (T*)ptr;
    return BS_OK;
}

static hipEvent_t take_event(bs_engine* e)
{
    if (!e->free_events.empty()) {
        hipEvent_t ev = e->free_events.back();
        e->free_events.pop_back();
        return ev;
    }
    hipEvent_t ev = nullptr;
    if (hipEventCreate(&ev) != hipSuccess) return nullptr;
    return ev;
}

// Brackets one launch group with HIP events on the engine's stream.
struct Timed {
    bs_engine* e;
    int kernel;
    hipEvent_t a = nullptr;
    Timed(bs_engine* e_, int k) : e(e_), kernel(k)
    {
        if ((e->prm.flags & BS_FLAG_TIMING) && !e->capturing) {
            a = take_event(e);
            if (a) (void)hipEventRecord(a, e->stream);
        }
    }
    void stop()
    {
        if (!a) return;
        hipEvent_t b = take_event(e);
        if (!b) { e->free_events.push_back(a); a = nullptr; return; }
        (void)hipEventRecord(b, e->stream);
        e->pending.push_back({kernel, a, b});
        a = nullptr;
    }
};

static int resolve_timing(bs_engine* e)
{
    if (e->pending.empty()) return BS_OK;
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    for (auto& p : e->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            e->ms[p.kernel] += ms;
            e->launches[p.kernel] += 1;
        }
        e->free_events.push_back(p.a);
        e->free_events.push_back(p.b);
    }
    e->pending.clear();
    return BS_OK;
}

// ----------------------------------------------------------------------------
// exact thresholds: the kernels decide `sqrt(s) < r` and `acos(c) <= pi*a`
// by comparing s and c with the boundary values below.  The rounded sqrt is
// monotone, so {s : sqrt(s) < r} = [0, s_thr).  The host acos is checked to be
// monotone around the boundary, so {c : acos(c) <= pi*a} = [c_thr, 1].
// ----------------------------------------------------------------------------
static double sthr_f64(double r)
{
    if (isnan(r)) return -1.0;
    if (!(r > 0.0)) return 0.0;
    if (isinf(r)) return INFINITY;
    double s = r * r;
    if (isinf(s)) s = DBL_MAX;
    while (s > 0.0 && sqrt(s) >= r) s = nextafter(s, 0.0);
    while (!(sqrt(s) >= r)) {
        if (isinf(s)) break;
        s = nextafter(s, INFINITY);
    }
    return s;
}

static float sthr_f32(double r)
{
    if (isnan(r)) return -1.0f;
    if (!(r > 0.0)) return 0.0f;
    float s = (float)(r * r);
    if (isinf(s)) s = FLT_MAX;
    while (s > 0.0f && (double)sqrtf(s) >= r) s = nextafterf(s, 0.0f);
    while (!((double)sqrtf(s) >= r)) {
        if (isinf(s)) break;
        s = nextafterf(s, INFINITY);
    }
    return s;
}

// ordinal <-> float over [-1, 1]
static int64_t ord32(float c) { uint32_t u; memcpy(&u, &c, 4); return (u & 0x80000000u) ? -(int64_t)(u & 0x7fffffffu) : (int64_t)u; }
static float from_ord32(int64_t o)
{
    uint32_t u = o < 0 ? (0x80000000u | (uint32_t)(-o)) : (uint32_t)o;
    float c; memcpy(&c, &u, 4); return c;
}
static int64_t ord64(double c) { uint64_t u; memcpy(&u, &c, 8); return (u >> 63) ? -(int64_t)(u & 0x7fffffffffffffffull) : (int64_t)u; }
static double from_ord64(int64_t o)
{
    uint64_t u = o < 0 ? (0x8000000000000000ull | (uint64_t)(-o)) : (uint64_t)o;
    double c; memcpy(&c, &u, 8); return c;
}

// Smallest c in [-1,1] with acos(c) <= theta (ACOS = acosf on float, acos on
// double).  Returns 2.0 when none.  *ok = 0 if the predicate is not a clean
// false..true step within +-window ordinals of the boundary.
template <typename T, typename FN>
static double cone_threshold(double theta, FN acos_fn, int64_t (*ord)(T), T (*from)(int64_t),
                             int64_t window, int* ok)
{
    auto pred = [&](int64_t o) { return (double)acos_fn(from(o)) <= theta; };
    const int64_t lo0 = ord((T)-1.0), hi0 = ord((T)1.0);
    *ok = 1;
    if (!pred(hi0)) return 2.0;
    if (pred(lo0)) return -1.0;
    int64_t lo = lo0, hi = hi0;      // pred(lo) false, pred(hi) true
    while (hi - lo > 1) {
        int64_t mid = lo + (hi - lo) / 2;
        if (pred(mid)) hi = mid; else lo = mid;
    }
    for (int64_t o = std::max(lo0, hi - window); o < hi; ++o)
        if (pred(o)) { *ok = 0; break; }
    for (int64_t o = hi; o <= std::min(hi0, hi + window); ++o)
        if (!pred(o)) { *ok = 0; break; }
    return (double)from(hi);
}

static float round_down_f(double x) { float f = (float)x; if ((double)f > x) f = nextafterf(f, -INFINITY); return f; }
static float round_up_f(double x) { float f = (float)x; if ((double)f < x) f = nextafterf(f, INFINITY); return f; }

static int derive_thresholds(const bs_params* p, DevParams* P, std::string* why)
{
    const double r[3] = {p->sep_r, p->ali_r, p->coh_r};
    const double a[3] = {p->sep_a, p->ali_a, p->coh_a};
    const double w[3] = {p->sep_w, p->ali_w, p->coh_w};
    P->F = p->field;
    P->half = p->field / 2;
    P->s_max = -1.0;
    P->s_max_f = -1.0f;
    P->unit = p->field / 4294967296.0;
    P->lat_scale = 4294967296.0 / p->field;
    P->s_hi_max = 0.0f;
    const double beta_c = 1e-5;
    for (int k = 0; k < 3; ++k) {
        P->s_thr[k] = sthr_f64(r[k]);
        P->s_thr_f[k] = sthr_f32(r[k]);
        P->s_max = std::max(P->s_max, P->s_thr[k]);
        P->s_max_f = std::max(P->s_max_f, P->s_thr_f[k]);
        P->w[k] = w[k];
        P->w_f[k] = (float)w[k];
        P->cone[k] = !(a[k] >= 1.0);   // src/boid.cpp:29: a >= 1 skips the cone (NaN keeps it)
        P->never[k] = 0;
        P->c_thr[k] = -2.0;
        P->c_thr_f[k] = -2.0f;
        if (P->cone[k]) {
            const double theta = M_PI * a[k];
            int ok32 = 1, ok64 = 1;
            double c32 = cone_threshold<float>(theta, [](float c) { return acosf(c); }, ord32, from_ord32, 1 << 16, &ok32);
            double c64 = cone_threshold<double>(theta, [](double c) { return acos(c); }, ord64, from_ord64, 1 << 14, &ok64);
            if (!ok32 || !ok64) {
                *why = "host acos/acosf is not monotone around the view-cone boundary of rule " + std::to_string(k);
                return BS_E_MATH;
            }
            P->c_thr_f[k] = (float)c32;
            P->c_thr[k] = c64;
            P->never[k] = c64 > 1.0;
        }
        // mixed guard bands (lattice units); beta_d covers the FP32 distance
        // error (relative ~3e-7) and the lattice rounding (sqrt(3) units).
        const double beta_d = 1e-5 + 2e-6 * (isfinite(r[k]) ? fabs(r[k]) : 0.0);
        const double lo = r[k] - beta_d, hi = r[k] + beta_d;
        P->s_lo[k] = (lo > 0.0) ? round_down_f((lo / P->unit) * (lo / P->unit)) : 0.0f;
        P->s_hi[k] = (hi > 0.0) ? round_up_f((hi / P->unit) * (hi / P->unit)) : 0.0f;
        if (isnan(r[k])) { P->s_lo[k] = 0.0f; P->s_hi[k] = 0.0f; }
        P->s_hi_max = std::max(P->s_hi_max, P->s_hi[k]);
        const double cth = P->c_thr[k];
        P->c_lo[k] = P->cone[k] ? round_down_f(cth - beta_c) : -INFINITY;
        P->c_hi[k] = P->cone[k] ? round_up_f(cth + beta_c) : -INFINITY;
        P->c_top[k] = P->cone[k] ? round_down_f(1.0 - beta_c) : INFINITY;
    }
    P->band_k = 8.0f;   // >= 2*sqrt(3) lattice units of displacement error, x2 margin
    P->same_cone = 1;
    for (int k = 1; k < 3; ++k)
        if (!(P->c_lo[k] == P->c_lo[0] && P->c_hi[k] == P->c_hi[0] && P->c_top[k] == P->c_top[0])) P->same_cone = 0;
    // Coarse filter on the top 16 lattice bits (unit F/2^16): a coarse
    // difference is off by at most one unit per axis, so every pair with
    // d < max(r) + beta_d has coarse squared distance < (r16 + 2)^2.
    {
        double rhi = 0.0;
        bool unbounded = false;
        for (int k = 0; k < 3; ++k) {
            if (isnan(r[k])) continue;
            if (isinf(r[k])) { unbounded = true; continue; }
            rhi = std::max(rhi, r[k] + 1e-5 + 2e-6 * fabs(r[k]));
        }
        const double r16 = rhi / (p->field / 65536.0);
        const double thr = ceil((r16 + 2.0) * (r16 + 2.0));
        P->coarse_thr = (unbounded || !(thr < 2147483647.0)) ? 2147483647 : (int32_t)thr;
    }
    P->vmax = p->vmax;
    P->vmin = p->vmin;
    P->vmax_f = (float)p->vmax;
    P->vmin_f = (float)p->vmin;
    // vmax in [2^(ex-1), 2^ex): quantum 2^(ex-31) <= vmax * 2^-30
    int ex = -9;
    if (isfinite(p->vmax) && p->vmax > 0.0) frexp(p->vmax, &ex);
    P->vq = ldexp(1.0, ex - 31);
    P->inv_vq = ldexp(1.0, 31 - ex);
    return BS_OK;
}

// Largest g with g^3 <= n (exact; floor(cbrt(n)) can land one below a
// perfect cube).
static uint32_t icbrt(uint32_t n)
{
    uint64_t g = (uint64_t)cbrt((double)n);
    while (g * g * g > n) --g;
    while ((g + 1) * (g + 1) * (g + 1) <= n) ++g;
    return (uint32_t)g;
}

static int cells_per_axis(const bs_params* p)
{
    if (p->cells_per_axis > 0) {
        double rmax = std::max(p->coh_r, std::max(p->ali_r, p->sep_r));
        double edge = p->field / p->cells_per_axis;
        if (p->cells_per_axis < 3 || !(edge >= rmax * 1.0001)) return -1;
        return std::min(p->cells_per_axis, kMaxCells);
    }
    double rmax = std::max(p->coh_r, std::max(p->ali_r, p->sep_r));
    if (!(rmax > 0.0) || !(p->field > 0.0)) return 0;
    double g = floor(p->field / (rmax * 1.0001));
    if (g < 3.0) return 0;
    // no more cells than boids: finer grids only add empty-cell overhead
    double cap = std::max(3.0, (double)icbrt(std::max<uint32_t>(p->n, 27)));
    g = std::min(g, std::max(cap, 3.0));
    g = std::min(g, (double)kMaxCells);
    return (int)g;
}

// ----------------------------------------------------------------------------
// init generators (host): drand48 restated, reference kinds + workloads
// ----------------------------------------------------------------------------
namespace {
struct Drand48 {
    uint64_t x = 0;                       // glibc's state before any srand48
    void seed(long s) { x = (((uint64_t)(uint32_t)s) << 16) | 0x330Eu; }
    double next()
    {
        x = (0x5DEECE66Dull * x + 0xBull) & ((1ull << 48) - 1);
        return (double)x * (1.0 / 281474976710656.0);
    }
};

double gauss(Drand48& r)
{
    const double u1 = 1.0 - r.next();     // (0, 1]
    const double u2 = r.next();
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

double wrap_field(double x, double F)
{
    x -= F * floor(x / F);
    if (x >= F) x -= F;
    if (x < 0.0) x = 0.0;
    return x;
}
}  // namespace

extern "C" int bs_generate_init(int32_t kind, int32_t seeded, int32_t seed, uint32_t n, double F,
                                double vmax, double vmin, float* out)
{
    if ((n && !out) || kind < 0 || kind > 3) return fail(nullptr, BS_E_ARG, "bs_generate_init: bad arguments");
    Drand48 r;
    if (seeded) r.seed(seed);
    if (kind == BS_INIT_TEST || kind == BS_INIT_RANDOM_UNIFORM) {
        // src/boid_simulation.cpp:82-112, draw for draw
        for (uint32_t i = 0; i < n; ++i) {
            float* o = out + 6 * (size_t)i;
            if (kind == BS_INIT_TEST) {
                o[0] = (float)(F * 3 / 8 + r.next() * F / 4);
                o[1] = (float)(F * 3 / 8 + r.next() * F / 4);
                if (r.next() > 0.5) o[2] = (float)(r.next() * F / 4 + F * 3 / 4);
                else o[2] = (float)(r.next() * F / 4);
            } else {
                o[0] = (float)(F * r.next());
                o[1] = (float)(F * r.next());
                o[2] = (float)(F * r.next());
            }
            double s = r.next() * (vmax - vmin) + vmin;
            double th1 = r.next() * M_PI;
            double th2 = r.next() * 2.0 * M_PI;
            o[3] = (float)(s * sin(th1) * cos(th2));
            o[4] = (float)(s * sin(th1) * sin(th2));
            o[5] = (float)(s * cos(th1));
        }
        return BS_OK;
    }
    if (kind == BS_INIT_ISOTROPIC) {
        // per boid: x, y, z, speed, cos(theta) = 2u-1, phi = 2 pi u
        for (uint32_t i = 0; i < n; ++i) {
            float* o = out + 6 * (size_t)i;
            o[0] = (float)(F * r.next());
            o[1] = (float)(F * r.next());
            o[2] = (float)(F * r.next());
            double s = vmin + (vmax - vmin) * r.next();
            double ct = 2.0 * r.next() - 1.0;
            double ph = 2.0 * M_PI * r.next();
            double st = sqrt(std::max(0.0, 1.0 - ct * ct));
            o[3] = (float)(s * st * cos(ph));
            o[4] = (float)(s * st * sin(ph));
            o[5] = (float)(s * ct);
        }
        return BS_OK;
    }
    // BS_INIT_CLUSTERED: 64 blobs, centres uniform, sigma = 0.05 F, positions
    // wrapped; heading per blob isotropic, boid direction = heading rotated
    // by N(0, 0.2) rad about a uniformly random perpendicular axis.
    const int nb = 64;
    double cen[64][3], hd[64][3];
    for (int b = 0; b < nb; ++b) {
        for (int k = 0; k < 3; ++k) cen[b][k] = F * r.next();
        double ct = 2.0 * r.next() - 1.0, ph = 2.0 * M_PI * r.next();
        double st = sqrt(std::max(0.0, 1.0 - ct * ct));
        hd[b][0] = st * cos(ph); hd[b][1] = st * sin(ph); hd[b][2] = ct;
    }
    const double sigma = 0.05 * F;
    for (uint32_t i = 0; i < n; ++i) {
        float* o = out + 6 * (size_t)i;
        int b = (int)(nb * r.next());
        if (b >= nb) b = nb - 1;
        for (int k = 0; k < 3; ++k) o[k] = (float)wrap_field(cen[b][k] + sigma * gauss(r), F);
        const double* h = hd[b];
        double ax[3] = {1.0, 0.0, 0.0};
        if (fabs(h[0]) > 0.9) { ax[0] = 0.0; ax[1] = 1.0; }
        double e1[3] = {ax[1] * h[2] - ax[2] * h[1], ax[2] * h[0] - ax[0] * h[2], ax[0] * h[1] - ax[1] * h[0]};
        double l1 = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
        for (int k = 0; k < 3; ++k) e1[k] /= l1;
        double e2[3] = {h[1] * e1[2] - h[2] * e1[1], h[2] * e1[0] - h[0] * e1[2], h[0] * e1[1] - h[1] * e1[0]};
        double th = 0.2 * gauss(r), ph = 2.0 * M_PI * r.next();
        double s = vmin + (vmax - vmin) * r.next();
        for (int k = 0; k < 3; ++k) {
            double d = cos(th) * h[k] + sin(th) * (cos(ph) * e1[k] + sin(ph) * e2[k]);
            o[3 + k] = (float)(s * d);
        }
    }
    return BS_OK;
}

extern "C" int bs_read_init_file(const char* path, uint32_t n, float* out)
{
    if (!path || (n && !out)) return fail(nullptr, BS_E_ARG, "bs_read_init_file: bad arguments");
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail(nullptr, BS_E_ARG, std::string("Invalid initial placement type or file name: ") + path);
    f.read((char*)out, (std::streamsize)(sizeof(float) * 6 * (size_t)n));
    if ((size_t)f.gcount() != sizeof(float) * 6 * (size_t)n)
        return fail(nullptr, BS_E_ARG, std::string("init file too short: ") + path);
    return BS_OK;   // little-endian host: the file layout is the memory layout
}

// ----------------------------------------------------------------------------
// engine lifetime
// ----------------------------------------------------------------------------
static void release(bs_engine* e)
{
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (auto& p : e->pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    for (auto ev : e->free_events) (void)hipEventDestroy(ev);
    if (e->graph) (void)hipGraphExecDestroy(e->graph);
    for (void* p : e->allocs) (void)hipFree(p);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

static int alloc_f64(bs_engine* e, StateF64* s, uint32_t n)
{
    int rc;
    if ((rc = dev_alloc(e, &s->x, n)) || (rc = dev_alloc(e, &s->y, n)) || (rc = dev_alloc(e, &s->z, n)) ||
        (rc = dev_alloc(e, &s->vx, n)) || (rc = dev_alloc(e, &s->vy, n)) || (rc = dev_alloc(e, &s->vz, n)) ||
        (rc = dev_alloc(e, &s->id, n)))
        return rc;
    return BS_OK;
}

static int alloc_f32(bs_engine* e, StateF32* s, uint32_t n)
{
    int rc;
    if ((rc = dev_alloc(e, &s->x, n)) || (rc = dev_alloc(e, &s->y, n)) || (rc = dev_alloc(e, &s->z, n)) ||
        (rc = dev_alloc(e, &s->vx, n)) || (rc = dev_alloc(e, &s->vy, n)) || (rc = dev_alloc(e, &s->vz, n)))
        return rc;
    return BS_OK;
}

static bool is_grid(int mode) { return mode == BS_MODE_GRID_F64 || mode == BS_MODE_GRID_MIXED; }
// float state in id order (the reference-faithful modes)
static bool is_f32(int mode) { return mode == BS_MODE_ALLPAIRS_F32REF || mode == BS_MODE_GRID_F32REF; }

static Flocks flock_view(const bs_engine* e)
{
    Flocks F;
    F.fl = e->flocks_dev;
    F.flock_of = e->flock_of;
    F.row_flock = e->row_flock;
    F.count = (uint32_t)e->flocks.size();
    F.n = e->prm.n;
    F.rows = e->rows;
    return F;
}

// Shared by bs_create (one flock) and bs_create_batch (a sweep): validates
// every flock, derives its thresholds and grid, lays the flocks out
// back to back (boids, cells and rows) and allocates the device buffers.
static int create_engine(const bs_params* fp, uint32_t count, bool batch, bs_engine** out)
{
    const bs_params* p = &fp[0];
    if (!(p->field > 0.0) || !isfinite(p->field)) return fail(nullptr, BS_E_ARG, "field size must be positive and finite");
    if (p->mode < BS_MODE_ALLPAIRS_F32REF || p->mode > BS_MODE_GRID_F32REF) return fail(nullptr, BS_E_ARG, "unknown mode");
    if (batch && !is_grid(p->mode)) return fail(nullptr, BS_E_ARG, "bs_create_batch: grid modes only");
    uint64_t ntot = 0;
    for (uint32_t k = 0; k < count; ++k) {
        const bs_params& q = fp[k];
        const std::string at = batch ? "flock " + std::to_string(k) + ": " : std::string();
        if (!(q.field > 0.0) || !isfinite(q.field)) return fail(nullptr, BS_E_ARG, at + "field size must be positive and finite");
        if (q.mode != p->mode || q.device != p->device || q.flags != p->flags || q.block != p->block ||
            q.tile != p->tile || q.fill != p->fill)
            return fail(nullptr, BS_E_ARG, at + "mode, device, flags, block, tile and fill must be the same for every flock");
        ntot += q.n;
    }
    if (ntot > (1u << 28)) return fail(nullptr, BS_E_ARG, "n too large (max 2^28 boids per engine)");

    bs_engine* e = new bs_engine();
    e->prm = *p;
    e->prm.n = (uint32_t)ntot;
    e->mode = p->mode;
    e->batch = batch;
    std::string why;
    int rc = derive_thresholds(p, &e->P, &why);
    if (rc) { g_create_error = why; release(e); return rc; }
    e->P.n = p->n;

    if (is_grid(e->mode)) {
        if (p->block != 0 && p->block != 64 && p->block != 128 && p->block != 256) {
            release(e);
            return fail(nullptr, BS_E_ARG, "block must be 0, 64, 128 or 256");
        }
        if (p->tile != 0 && p->tile != 5 && p->tile != 6 && p->tile != 9 && p->tile != 12) { release(e); return fail(nullptr, BS_E_ARG, "tile must be 0, 5, 6, 9 or 12"); }
        if (p->fill != 0 && (p->fill < 25 || p->fill > 100)) { release(e); return fail(nullptr, BS_E_ARG, "fill must be 0 or 25..100"); }
        e->tile = p->tile ? p->tile : (p->mode == BS_MODE_GRID_MIXED ? BS_MIXED_TILE : BS_F64_TILE);
        // Automatic fill: full chunks, unless that leaves fewer than ~2048
        // of them (8 workgroups per CU) -- a small flock then runs in
        // smaller chunks, more workgroups in flight (down to 25 %).
        {
            uint64_t nsum = 0;
            for (uint32_t k = 0; k < count; ++k) nsum += fp[k].n;
            const int nb_auto = p->block ? p->block : 64;
            const uint64_t want = 100 * nsum / ((uint64_t)nb_auto * 2048);
            e->fill = p->fill ? p->fill : (int)std::max<uint64_t>(25, std::min<uint64_t>(100, want));
        }
        uint64_t cells = 0, rows = 0;
        bool coarse = false;
        // grid geometry of the path (every flock of an engine shares the mode)
        e->sx = p->mode == BS_MODE_GRID_MIXED ? BS_MIXED_SUBX : BS_F64_SUBX;
        e->cspan = p->mode == BS_MODE_GRID_MIXED ? BS_MIXED_CHUNK_SPAN : BS_F64_CHUNK_SPAN;
        for (uint32_t k = 0; k < count; ++k) {
            const bs_params& q = fp[k];
            const std::string at = batch ? "flock " + std::to_string(k) + ": " : std::string();
            FlockDesc d{};
            rc = derive_thresholds(&q, &d.P, &why);
            if (rc) { release(e); return fail(nullptr, rc, at + why); }
            d.P.n = q.n;
            const int G = cells_per_axis(&q);
            if (G < 0) { release(e); return fail(nullptr, BS_E_ARG, at + "cells_per_axis must be >= 3 with edge >= max radius"); }
            if (G < 3) {
                if (batch) { release(e); return fail(nullptr, BS_E_ARG, at + "fewer than 3 cells per axis (field too small for its radii); step it in its own engine"); }
                e->mode = BS_MODE_ALLPAIRS_F64;   // windows (3 cells in y and z) would overlap
                break;
            }
            d.P.G = G;
            d.P.sx = e->sx;
            d.P.inv_edge = G / q.field;
            d.cell_base = (uint32_t)cells;
            d.row_base = (uint32_t)rows;
            d.nrows = (uint32_t)G * (uint32_t)G;
            // rows (sx G x columns) are cut into the fewest equal groups of
            // <= kSpanMax columns
            const uint32_t Gx = (uint32_t)d.P.sx * (uint32_t)G;
            const uint32_t ng = (Gx + kSpanMax - 1) / kSpanMax;
            d.span = (Gx + ng - 1) / ng;
            // chunk-table slack per span group: its last chunk, plus one chunk
            // cut short per cspan columns (k_rows_pass)
            e->groups += (uint64_t)G * G * ((Gx + d.span - 1) / d.span) * (1 + (d.span - 1) / e->cspan);
            e->max_span = std::max(e->max_span, (int)d.span);
            cells += (uint64_t)Gx * G * G;
            rows += (uint64_t)G * G;
            if (!(d.P.unit * 8.0 < 1e-6)) coarse = true;   // lattice too coarse for the mixed guard bands
            e->flocks.push_back(d);
            e->flock_field.push_back(q.field);
            e->max_g = std::max(e->max_g, G);
        }
        // Automatic workgroup size 64: a chunk never leaves its (y,z) row,
        // so short rows (small flocks) fill 64 lanes better, and at the
        // smaller LDS footprint more workgroups fit a CU (measured faster
        // than 128 for long rows too).
        e->nb = p->block ? p->block : 64;
        e->chunk = std::max(1u, (uint32_t)(e->fill * e->nb / 100));
        if (e->mode != BS_MODE_ALLPAIRS_F64) {
            if (cells + 1 > 0x7fffffffull) { release(e); return fail(nullptr, BS_E_ARG, "too many grid cells over all flocks"); }
            uint32_t base = 0;
            for (uint32_t k = 0; k < count; ++k) { e->flocks[k].boid_base = base; base += fp[k].n; }
            e->rows = (uint32_t)rows;
            e->ncell1 = (uint32_t)(cells + 1);
            e->G = e->flocks[0].P.G;
            e->edge = fp[0].field / e->G;
            e->P.G = e->G;
            e->P.sx = e->sx;
            e->P.inv_edge = e->flocks[0].P.inv_edge;
            if (e->mode == BS_MODE_GRID_MIXED && coarse) e->mode = BS_MODE_GRID_F64;
        } else {
            e->flocks.clear();
            e->flock_field.clear();
        }
    } else if (e->mode == BS_MODE_GRID_F32REF) {
        // plain cells (edge >= the largest radius), >= 3 per axis so the 27
        // neighbour cells are distinct; else all-pairs, which gives the same
        // bits (the reference's grid engine equals its all-pairs engine)
        const int G = cells_per_axis(p);
        if (G < 0) { release(e); return fail(nullptr, BS_E_ARG, "cells_per_axis must be >= 3 with edge >= max radius"); }
        if (G < 3) {
            e->mode = BS_MODE_ALLPAIRS_F32REF;
        } else {
            e->G = G;
            e->edge = p->field / G;
            e->P.G = G;
            e->P.sx = 1;
            e->P.inv_edge = G / p->field;
            e->ncell1 = (uint32_t)((uint64_t)G * G * G + 1);
        }
    }

    hipError_t st = hipSetDevice(p->device);
    if (st != hipSuccess) { std::string m = std::string("hipSetDevice: ") + hipGetErrorString(st); release(e); g_create_error = m; return BS_E_HIP; }
    e->device = p->device;
    st = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (st != hipSuccess) { std::string m = std::string("hipStreamCreate: ") + hipGetErrorString(st); e->stream = nullptr; release(e); g_create_error = m; return BS_E_HIP; }
    e->scratch = scratch_address(p->device, &e->scratch_why);   // once per device, this device current

    const uint32_t n = e->prm.n;
    auto bail = [&](int code) { std::string m = e->err; release(e); g_create_error = m; return code; };
    if (is_f32(e->mode)) {
        if ((rc = alloc_f32(e, &e->fa, n)) || (rc = alloc_f32(e, &e->fb, n))) return bail(rc);
        if (p->flags & BS_FLAG_FORCES) { float* f; if ((rc = dev_alloc(e, &f, 9 * (size_t)n))) return bail(rc); e->forces = f; }
    } else {
        if ((rc = alloc_f64(e, &e->da, n)) || (rc = alloc_f64(e, &e->db, n))) return bail(rc);
        if (e->mode == BS_MODE_ALLPAIRS_F64) e->db.id = e->da.id;   // all-pairs stays in id order
        if (p->flags & BS_FLAG_FORCES) { double* f; if ((rc = dev_alloc(e, &f, 9 * (size_t)n))) return bail(rc); e->forces = f; }
    }
    if ((rc = dev_alloc(e, &e->tmp_pos, 3 * (size_t)n)) || (rc = dev_alloc(e, &e->tmp_vel, 3 * (size_t)n))) return bail(rc);
    if (p->flags & BS_FLAG_DIGEST) { if ((rc = dev_alloc(e, &e->digest, 4 * (size_t)n))) return bail(rc); }
    if (e->mode == BS_MODE_GRID_F32REF) {
        const uint32_t ntiles = (e->ncell1 + 4095) / 4096;
        if ((rc = dev_alloc(e, &e->key, n)) || (rc = dev_alloc(e, &e->rank, n)) || (rc = dev_alloc(e, &e->perm, n)) ||
            (rc = dev_alloc(e, &e->count, e->ncell1)) || (rc = dev_alloc(e, &e->start, e->ncell1)) ||
            (rc = dev_alloc(e, &e->partials, ntiles)) || (rc = dev_alloc(e, &e->rec4, n)) ||
            (rc = dev_alloc(e, &e->vrec4, n)))
            return bail(rc);
        st = hipMemsetAsync(e->count, 0, sizeof(uint32_t) * e->ncell1, e->stream);
        if (st != hipSuccess) { hip_fail(e, st, "hipMemsetAsync"); return bail(BS_E_HIP); }
    }
    if (is_grid(e->mode)) {
        // full chunks <= n / chunk; groups counts each (row, span group)'s last chunk and short cuts (k_rows_pass)
        if ((uint64_t)(n + e->chunk - 1) / e->chunk + e->groups + 1 > 0x7fffffffull) return bail(fail(e, BS_E_ARG, "chunk table too large"));
        e->max_chunks = (uint32_t)((n + e->chunk - 1) / e->chunk + e->groups + 1);
        const uint32_t ntiles = (e->ncell1 + 4095) / 4096;
        if ((rc = dev_alloc(e, &e->key, n)) || (rc = dev_alloc(e, &e->key_sorted, n)) ||
            (rc = dev_alloc(e, &e->count, e->ncell1)) || (rc = dev_alloc(e, &e->start, e->ncell1)) ||
            (rc = dev_alloc(e, &e->partials, ntiles)) || (rc = dev_alloc(e, &e->nchunks, 1)) ||
            (rc = dev_alloc(e, &e->perm, n)) || (rc = dev_alloc(e, &e->permid, n)) || (rc = dev_alloc(e, &e->rank, n)) ||
            (rc = dev_alloc(e, &e->table, e->max_chunks)) || (rc = dev_alloc(e, &e->recheck, 1)) ||
            (rc = dev_alloc(e, &e->overflow, 1)) ||
            (rc = dev_alloc(e, &e->row_count, (size_t)e->rows + 1)) ||
            (rc = dev_alloc(e, &e->row_off, (size_t)e->rows + 1)) ||
            (rc = dev_alloc(e, &e->flocks_dev, e->flocks.size())))
            return bail(rc);
        if (e->flocks.size() > 1) {
            if ((rc = dev_alloc(e, &e->flock_of, n))) return bail(rc);
            std::vector<uint32_t> fo(n);
            for (uint32_t k = 0; k < e->flocks.size(); ++k)
                std::fill(fo.begin() + e->flocks[k].boid_base, fo.begin() + e->flocks[k].boid_base + e->flocks[k].P.n, k);
            st = hipMemcpy(e->flock_of, fo.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice);
            if (st != hipSuccess) { hip_fail(e, st, "hipMemcpy"); return bail(BS_E_HIP); }
            if ((rc = dev_alloc(e, &e->row_flock, e->rows))) return bail(rc);
            std::vector<uint32_t> rf(e->rows);
            for (uint32_t k = 0; k < e->flocks.size(); ++k)
                std::fill(rf.begin() + e->flocks[k].row_base, rf.begin() + e->flocks[k].row_base + e->flocks[k].nrows, k);
            st = hipMemcpy(e->row_flock, rf.data(), sizeof(uint32_t) * e->rows, hipMemcpyHostToDevice);
            if (st != hipSuccess) { hip_fail(e, st, "hipMemcpy"); return bail(BS_E_HIP); }
        }
        st = hipMemcpy(e->flocks_dev, e->flocks.data(), sizeof(FlockDesc) * e->flocks.size(), hipMemcpyHostToDevice);
        if (st != hipSuccess) { hip_fail(e, st, "hipMemcpy"); return bail(BS_E_HIP); }
        if (e->mode == BS_MODE_GRID_MIXED) {
            if ((rc = dev_alloc(e, &e->lat.p, n)) || (rc = dev_alloc(e, &e->lat.v, n)))
                return bail(rc);
        } else if ((rc = dev_alloc(e, &e->rec, 3 * (size_t)n))) {
            return bail(rc);
        }
        st = hipMemsetAsync(e->count, 0, sizeof(uint32_t) * e->ncell1, e->stream);
        if (st == hipSuccess) st = hipMemsetAsync(e->recheck, 0, sizeof(uint32_t), e->stream);
        if (st == hipSuccess) st = hipMemsetAsync(e->overflow, 0, sizeof(uint32_t), e->stream);
        if (st != hipSuccess) { hip_fail(e, st, "hipMemsetAsync"); return bail(BS_E_HIP); }
    }
    if (e->forces) {
        size_t bytes = 9 * (size_t)n * (is_f32(e->mode) ? sizeof(float) : sizeof(double));
        st = hipMemsetAsync(e->forces, 0, bytes, e->stream);
        if (st != hipSuccess) { hip_fail(e, st, "hipMemsetAsync"); return bail(BS_E_HIP); }
    }
    if (e->digest) {
        st = hipMemsetAsync(e->digest, 0, 32 * (size_t)n, e->stream);
        if (st != hipSuccess) { hip_fail(e, st, "hipMemsetAsync"); return bail(BS_E_HIP); }
    }
    st = hipStreamSynchronize(e->stream);
    if (st != hipSuccess) { hip_fail(e, st, "hipStreamSynchronize"); return bail(BS_E_HIP); }
    *out = e;
    return BS_OK;
}

extern "C" int bs_create(const bs_params* p, bs_engine** out)
{
    if (!out) return fail(nullptr, BS_E_ARG, "bs_create: out is NULL");
    *out = nullptr;
    if (!p) return fail(nullptr, BS_E_ARG, "bs_create: params is NULL");
    return create_engine(p, 1, false, out);
}

extern "C" int bs_create_batch(const bs_params* flocks, uint32_t count, bs_engine** out)
{
    if (!out) return fail(nullptr, BS_E_ARG, "bs_create_batch: out is NULL");
    *out = nullptr;
    if (!flocks || count == 0) return fail(nullptr, BS_E_ARG, "bs_create_batch: no flocks");
    return create_engine(flocks, count, true, out);
}

extern "C" int bs_get_flock_count(const bs_engine* e, uint32_t* count)
{
    if (!e || !count) return BS_E_ARG;
    *count = e->flocks.empty() ? 1u : (uint32_t)e->flocks.size();
    return BS_OK;
}

extern "C" int bs_get_flock_info(const bs_engine* e, uint32_t k, bs_flock_info* out)
{
    if (!e || !out) return BS_E_ARG;
    memset(out, 0, sizeof(*out));
    if (e->flocks.empty()) {
        if (k != 0) return BS_E_ARG;
        out->n = e->prm.n;
        out->field = e->prm.field;
        return BS_OK;
    }
    if (k >= e->flocks.size()) return BS_E_ARG;
    const FlockDesc& d = e->flocks[k];
    out->boid_base = d.boid_base;
    out->n = d.P.n;
    out->cells_per_axis = d.P.G;
    out->cell_base = d.cell_base;
    out->field = e->flock_field[k];
    out->cell_edge = e->flock_field[k] / d.P.G;
    return BS_OK;
}

extern "C" void bs_destroy(bs_engine* e) { release(e); }

extern "C" const char* bs_last_error(const bs_engine* e)
{
    return e ? e->err.c_str() : g_create_error.c_str();
}

extern "C" const char* bs_version(void) { return "boids_engine 0.1 (gfx950)"; }

// ----------------------------------------------------------------------------
// state in / out
// ----------------------------------------------------------------------------
// Sort the B-side (id order or previous-slot order) state into A by (cell, id).
static int grid_sort_into_a(bs_engine* e)
{
    Timed t(e, BS_K_SCAN);
    HIP_TRY(e, launch_scan(e->count, e->start, e->ncell1, e->partials, e->stream));
    HIP_TRY(e, launch_rows(flock_view(e), e->start, e->row_count, e->row_off, e->partials, e->table, e->nchunks,
                           e->max_chunks, e->overflow,
                           (int)e->chunk, e->cspan, e->sx * e->max_g, e->stream));
    t.stop();
    Timed t2(e, BS_K_SCATTER);
    HIP_TRY(e, launch_scatter_ranked(e->prm.n, e->key, e->rank, e->start, e->db.id, e->perm, e->permid, e->stream));
    t2.stop();
    Timed t3(e, BS_K_GATHER);
    HIP_TRY(e, launch_gather(flock_view(e), e->perm, e->permid, e->key, e->start, e->count, e->db, e->da,
                             e->key_sorted, e->mode == BS_MODE_GRID_MIXED ? &e->lat : nullptr,
                             e->mode == BS_MODE_GRID_F64 ? e->rec : nullptr, e->stream));
    t3.stop();
    return BS_OK;
}

static int upload_f64(bs_engine* e, const double* pos3, const double* vel3)
{
    const uint32_t n = e->prm.n;
    std::vector<double> buf(7 * (size_t)n);
    double *x = buf.data(), *y = x + n, *z = y + n, *vx = z + n, *vy = vx + n, *vz = vy + n;
    uint32_t* ids = (uint32_t*)(vz + n);
    for (uint32_t i = 0; i < n; ++i) {
        x[i] = pos3[3 * (size_t)i]; y[i] = pos3[3 * (size_t)i + 1]; z[i] = pos3[3 * (size_t)i + 2];
        vx[i] = vel3[3 * (size_t)i]; vy[i] = vel3[3 * (size_t)i + 1]; vz[i] = vel3[3 * (size_t)i + 2];
        ids[i] = i;
    }
    const bool grid = is_grid(e->mode);
    StateF64 dst = grid ? e->db : e->da;
    const size_t b = sizeof(double) * n;
    HIP_TRY(e, hipMemcpyAsync(dst.x, x, b, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(e, hipMemcpyAsync(dst.y, y, b, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(e, hipMemcpyAsync(dst.z, z, b, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(e, hipMemcpyAsync(dst.vx, vx, b, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(e, hipMemcpyAsync(dst.vy, vy, b, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(e, hipMemcpyAsync(dst.vz, vz, b, hipMemcpyHostToDevice, e->stream));
    HIP_TRY(e, hipMemcpyAsync(dst.id, ids, sizeof(uint32_t) * n, hipMemcpyHostToDevice, e->stream));
    if (grid) {
        HIP_TRY(e, launch_hash_count(flock_view(e), e->db, e->key, e->rank, e->count, e->stream));
        int rc = grid_sort_into_a(e);
        if (rc) return rc;
    }
    HIP_TRY(e, hipStreamSynchronize(e->stream));   // the host staging buffer dies here
    e->has_state = true;
    e->steps = 0;
    return BS_OK;
}

extern "C" int bs_set_state_f64(bs_engine* e, const double* pos3, const double* vel3)
{
    if (!e) return BS_E_ARG;
    if (e->prm.n && (!pos3 || !vel3)) return fail(e, BS_E_ARG, "bs_set_state_f64: NULL array");
    if (is_f32(e->mode)) return fail(e, BS_E_STATE, "float-state mode: use bs_set_state_f32");
    return upload_f64(e, pos3, vel3);
}

extern "C" int bs_set_state_f32(bs_engine* e, const float* xyzv6)
{
    if (!e) return BS_E_ARG;
    const uint32_t n = e->prm.n;
    if (n && !xyzv6) return fail(e, BS_E_ARG, "bs_set_state_f32: NULL array");
    if (is_f32(e->mode)) {
        std::vector<float> buf(6 * (size_t)n);
        for (uint32_t i = 0; i < n; ++i)
            for (int k = 0; k < 6; ++k) buf[(size_t)k * n + i] = xyzv6[6 * (size_t)i + k];
        float* dst[6] = {e->fa.x, e->fa.y, e->fa.z, e->fa.vx, e->fa.vy, e->fa.vz};
        for (int k = 0; k < 6; ++k)
            HIP_TRY(e, hipMemcpyAsync(dst[k], buf.data() + (size_t)k * n, sizeof(float) * n, hipMemcpyHostToDevice, e->stream));
        HIP_TRY(e, hipStreamSynchronize(e->stream));
        e->has_state = true;
        e->steps = 0;
        return BS_OK;
    }
    std::vector<double> pos(3 * (size_t)n), vel(3 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            pos[3 * (size_t)i + k] = xyzv6[6 * (size_t)i + k];
            vel[3 * (size_t)i + k] = xyzv6[6 * (size_t)i + 3 + k];
        }
    return upload_f64(e, pos.data(), vel.data());
}

// Reports a chunk table cut short by its capacity (k_rows_pass clamps the
// count and leaves the full one in e->overflow): the update then skipped
// boids, so the state is not a valid result.  The capacity bound makes this
// unreachable; the check runs wherever the host reads results back.
static int check_grid(bs_engine* e)
{
    if (!e->overflow) return BS_OK;
    uint32_t v = 0;
    HIP_TRY(e, hipMemcpyAsync(&v, e->overflow, sizeof(uint32_t), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    if (v)
        return fail(e, BS_E_STATE, "chunk table overflow: " + std::to_string(v) + " chunks for a capacity of " +
                                       std::to_string(e->max_chunks) + " (results invalid)");
    return BS_OK;
}

extern "C" int bs_get_state_f64(bs_engine* e, double* pos3, double* vel3)
{
    if (!e) return BS_E_ARG;
    const uint32_t n = e->prm.n;
    if (n && (!pos3 || !vel3)) return fail(e, BS_E_ARG, "bs_get_state_f64: NULL array");
    if (!e->has_state) return fail(e, BS_E_STATE, "no state: call bs_set_state_* first");
    if (is_f32(e->mode)) {
        std::vector<float> p(3 * (size_t)n), v(3 * (size_t)n);
        int rc = bs_get_state_f32(e, p.data(), v.data());
        if (rc) return rc;
        for (size_t k = 0; k < 3 * (size_t)n; ++k) { pos3[k] = p[k]; vel3[k] = v[k]; }
        return BS_OK;
    }
    HIP_TRY(e, launch_unsort_f64(n, e->da, e->tmp_pos, e->tmp_vel, e->stream));
    HIP_TRY(e, hipMemcpyAsync(pos3, e->tmp_pos, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(e, hipMemcpyAsync(vel3, e->tmp_vel, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    return check_grid(e);
}

extern "C" int bs_get_state_f32(bs_engine* e, float* pos3, float* vel3)
{
    if (!e) return BS_E_ARG;
    const uint32_t n = e->prm.n;
    if (n && (!pos3 || !vel3)) return fail(e, BS_E_ARG, "bs_get_state_f32: NULL array");
    if (!e->has_state) return fail(e, BS_E_STATE, "no state: call bs_set_state_* first");
    if (is_f32(e->mode)) {
        std::vector<float> buf(6 * (size_t)n);
        float* src[6] = {e->fa.x, e->fa.y, e->fa.z, e->fa.vx, e->fa.vy, e->fa.vz};
        for (int k = 0; k < 6; ++k)
            HIP_TRY(e, hipMemcpyAsync(buf.data() + (size_t)k * n, src[k], sizeof(float) * n, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(e, hipStreamSynchronize(e->stream));
        for (uint32_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) {
                pos3[3 * (size_t)i + k] = buf[(size_t)k * n + i];
                vel3[3 * (size_t)i + k] = buf[(size_t)(3 + k) * n + i];
            }
        return BS_OK;
    }
    std::vector<double> p(3 * (size_t)n), v(3 * (size_t)n);
    int rc = bs_get_state_f64(e, p.data(), v.data());
    if (rc) return rc;
    for (size_t k = 0; k < 3 * (size_t)n; ++k) { pos3[k] = (float)p[k]; vel3[k] = (float)v[k]; }
    return BS_OK;
}

extern "C" int bs_get_forces_f64(bs_engine* e, double* f9)
{
    if (!e) return BS_E_ARG;
    const uint32_t n = e->prm.n;
    if (!e->forces) return fail(e, BS_E_STATE, "engine created without BS_FLAG_FORCES");
    if (n && !f9) return fail(e, BS_E_ARG, "bs_get_forces_f64: NULL array");
    if (is_f32(e->mode)) {
        std::vector<float> tmp(9 * (size_t)n);
        HIP_TRY(e, hipMemcpyAsync(tmp.data(), e->forces, sizeof(float) * 9 * n, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(e, hipStreamSynchronize(e->stream));
        for (size_t k = 0; k < tmp.size(); ++k) f9[k] = tmp[k];
        return BS_OK;
    }
    HIP_TRY(e, hipMemcpyAsync(f9, e->forces, sizeof(double) * 9 * n, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    return BS_OK;
}

extern "C" int bs_get_forces_f32(bs_engine* e, float* f9)
{
    if (!e) return BS_E_ARG;
    const uint32_t n = e->prm.n;
    if (!e->forces) return fail(e, BS_E_STATE, "engine created without BS_FLAG_FORCES");
    if (n && !f9) return fail(e, BS_E_ARG, "bs_get_forces_f32: NULL array");
    if (is_f32(e->mode)) {
        HIP_TRY(e, hipMemcpyAsync(f9, e->forces, sizeof(float) * 9 * n, hipMemcpyDeviceToHost, e->stream));
        HIP_TRY(e, hipStreamSynchronize(e->stream));
        return BS_OK;
    }
    std::vector<double> tmp(9 * (size_t)n);
    int rc = bs_get_forces_f64(e, tmp.data());
    if (rc) return rc;
    for (size_t k = 0; k < tmp.size(); ++k) f9[k] = (float)tmp[k];
    return BS_OK;
}

extern "C" int bs_get_digest(bs_engine* e, uint64_t* d4)
{
    if (!e) return BS_E_ARG;
    if (!e->digest) return fail(e, BS_E_STATE, "engine created without BS_FLAG_DIGEST");
    if (e->prm.n && !d4) return fail(e, BS_E_ARG, "bs_get_digest: NULL array");
    HIP_TRY(e, hipMemcpyAsync(d4, e->digest, 32 * (size_t)e->prm.n, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    return check_grid(e);
}

// Grid built by the last sort (grid modes): sorted cell keys and boid ids,
// cell starts and the chunk table, for exactness tests of the build.
extern "C" int bs_debug_grid(bs_engine* e, uint32_t* key_sorted, uint32_t* id_sorted, uint32_t* start,
                             uint32_t* table4, bs_grid_info* info)
{
    if (!e || !info) return BS_E_ARG;
    memset(info, 0, sizeof(*info));
    if (!is_grid(e->mode)) return fail(e, BS_E_STATE, "bs_debug_grid: not a grid engine");
    if (!e->has_state) return fail(e, BS_E_STATE, "no state: call bs_set_state_* first");
    const uint32_t n = e->prm.n;
    info->n = n;
    info->ncells = e->ncell1 - 1;
    info->max_chunks = e->max_chunks;
    info->x_columns_per_cell = e->sx;
    info->chunk_boids = (int32_t)e->chunk;
    info->chunk_span = e->cspan;
    HIP_TRY(e, hipMemcpyAsync(&info->nchunks, e->nchunks, sizeof(uint32_t), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(e, hipMemcpyAsync(&info->overflow, e->overflow, sizeof(uint32_t), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    if (key_sorted)
        HIP_TRY(e, hipMemcpyAsync(key_sorted, e->key_sorted, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, e->stream));
    if (id_sorted)
        HIP_TRY(e, hipMemcpyAsync(id_sorted, e->da.id, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, e->stream));
    if (start)
        HIP_TRY(e, hipMemcpyAsync(start, e->start, sizeof(uint32_t) * e->ncell1, hipMemcpyDeviceToHost, e->stream));
    if (table4 && info->nchunks)
        HIP_TRY(e, hipMemcpyAsync(table4, e->table, sizeof(uint4) * std::min(info->nchunks, e->max_chunks),
                                  hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    return BS_OK;
}

// ----------------------------------------------------------------------------
// the step
// ----------------------------------------------------------------------------
// One grid update: fused neighbour pass (A -> B, with the next cell keys
// and counts), then the counting sort of B back into A.  Every launch
// reads and writes the same buffers each step, so the sequence can be
// captured once and replayed.
static int grid_step(bs_engine* e)
{
    Timed t(e, BS_K_PAIR);
    const Flocks F = flock_view(e);
    if (e->mode == BS_MODE_GRID_F64)
        HIP_TRY(e, launch_nbr_f64(F, e->nb, e->tile, std::min(e->max_span, e->cspan) + 2 * e->sx, e->da, e->rec, e->db, e->key_sorted, e->start, e->table, e->nchunks,
                                  e->max_chunks, e->key, e->rank, e->count, (double*)e->forces, e->digest, e->stream));
    else
        HIP_TRY(e, launch_nbr_mixed(F, e->nb, e->tile, std::min(e->max_span, e->cspan) + 2 * e->sx, e->da, e->db, e->lat, e->key_sorted, e->start, e->table,
                                    e->nchunks, e->max_chunks, e->key, e->rank, e->count, (double*)e->forces,
                                    e->digest, e->recheck, e->stream));
    t.stop();
    return grid_sort_into_a(e);
}

// Captures grid_step into e->graph (once per engine).  The step holds no
// memset node: the scan clears the counters the next step accumulates.
static int capture_step(bs_engine* e)
{
    if (e->graph) return BS_OK;
    HIP_TRY(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    e->capturing = true;
    int rc = grid_step(e);
    e->capturing = false;
    hipGraph_t g = nullptr;
    hipError_t st = hipStreamEndCapture(e->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (st != hipSuccess) return hip_fail(e, st, "hipStreamEndCapture");
    st = hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (st != hipSuccess) { e->graph = nullptr; return hip_fail(e, st, "hipGraphInstantiate"); }
    return BS_OK;
}

extern "C" int bs_step(bs_engine* e, uint32_t nsteps)
{
    if (!e) return BS_E_ARG;
    if (!e->has_state) return fail(e, BS_E_STATE, "no state: call bs_set_state_* first");
    if (is_grid(e->mode) && (e->prm.flags & BS_FLAG_GRAPH) && nsteps) {
        int rc = capture_step(e);
        if (rc) return rc;
        Timed t(e, BS_K_STEP);
        for (uint32_t k = 0; k < nsteps; ++k) {
            HIP_TRY(e, hipGraphLaunch(e->graph, e->stream));
            e->steps++;
        }
        t.stop();
        return BS_OK;
    }
    for (uint32_t k = 0; k < nsteps; ++k) {
        switch (e->mode) {
        case BS_MODE_ALLPAIRS_F32REF: {
            Timed t(e, BS_K_PAIR);
            HIP_TRY(e, launch_allpairs_f32ref(e->P, e->fa, e->fb, (float*)e->forces, e->digest, e->stream));
            t.stop();
            std::swap(e->fa, e->fb);
            break;
        }
        case BS_MODE_GRID_F32REF: {
            Timed ts(e, BS_K_SCAN);
            HIP_TRY(e, launch_f32_grid_sort(e->P, e->fa, e->key, e->rank, e->count, e->start, e->partials, e->perm,
                                            e->rec4, e->vrec4, e->stream));
            ts.stop();
            Timed t(e, BS_K_PAIR);
            HIP_TRY(e, launch_f32_grid_step(e->P, e->fa, e->fb, e->rec4, e->vrec4, e->key, e->start,
                                            (float*)e->forces, e->digest, e->stream));
            t.stop();
            std::swap(e->fa, e->fb);
            break;
        }
        case BS_MODE_ALLPAIRS_F64: {
            Timed t(e, BS_K_PAIR);
            HIP_TRY(e, launch_allpairs_f64(e->P, e->da, e->db, (double*)e->forces, e->digest, e->stream));
            t.stop();
            std::swap(e->da, e->db);   // both sides share one identity id array
            break;
        }
        case BS_MODE_GRID_F64:
        case BS_MODE_GRID_MIXED: {
            int rc = grid_step(e);
            if (rc) return rc;
            break;
        }
        default:
            return fail(e, BS_E_STATE, "bad mode");
        }
        e->steps++;
    }
    return BS_OK;
}

extern "C" int bs_sync(bs_engine* e)
{
    if (!e) return BS_E_ARG;
    HIP_TRY(e, hipStreamSynchronize(e->stream));
#ifdef BS_DEBUG_CHECKS
    // debug builds check the chunk table after every synchronisation
    int rc = check_grid(e);
    if (rc) return rc;
#endif
    return resolve_timing(e);
}

extern "C" int bs_get_info(const bs_engine* e, bs_info* out)
{
    if (!e || !out) return BS_E_ARG;
    out->mode = e->mode;
    out->cells_per_axis = e->G;
    out->cell_edge = e->edge;
    out->device_bytes = e->device_bytes;
    out->steps = e->steps;
    out->block = is_grid(e->mode) ? e->nb : 0;
    out->tile = is_grid(e->mode) ? e->tile : 0;
    out->fill = is_grid(e->mode) ? e->fill : 0;
    out->reserved = 0;
    return BS_OK;
}

extern "C" int bs_get_timing(bs_engine* e, int32_t kernel, double* ms, uint64_t* launches)
{
    if (!e || kernel < 0 || kernel >= BS_K_COUNT) return BS_E_ARG;
    int rc = resolve_timing(e);
    if (rc) return rc;
    if (ms) *ms = e->ms[kernel];
    if (launches) *launches = e->launches[kernel];
    return BS_OK;
}

extern "C" int bs_reset_timing(bs_engine* e)
{
    if (!e) return BS_E_ARG;
    int rc = resolve_timing(e);
    if (rc) return rc;
    for (int k = 0; k < BS_K_COUNT; ++k) { e->ms[k] = 0.0; e->launches[k] = 0; }
    return BS_OK;
}

extern "C" int bs_debug_thresholds(const bs_params* p, double* out6)
{
    if (!p || !out6) return fail(nullptr, BS_E_ARG, "bs_debug_thresholds: NULL");
    DevParams P{};
    std::string why;
    int rc = derive_thresholds(p, &P, &why);
    if (rc) return fail(nullptr, rc, why);
    const bool f32 = is_f32(p->mode);
    for (int k = 0; k < 3; ++k) {
        out6[k] = f32 ? (double)P.s_thr_f[k] : P.s_thr[k];
        out6[3 + k] = P.cone[k] ? (f32 ? (double)P.c_thr_f[k] : P.c_thr[k]) : -2.0;
    }
    return BS_OK;
}

// Recheck counter of the mixed kernel (pairs decided in FP64), for reports.
extern "C" int bs_debug_rechecks(bs_engine* e, uint64_t* out)
{
    if (!e || !out) return BS_E_ARG;
    *out = 0;
    if (!e->recheck) return BS_OK;
    uint32_t v = 0;
    HIP_TRY(e, hipMemcpyAsync(&v, e->recheck, sizeof(uint32_t), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(e, hipStreamSynchronize(e->stream));
    *out = v;
    return BS_OK;
}

// Known-size streaming copy on the default device, for calibrating HBM byte
// counters: `reps` launches each moving `bytes` in and `bytes` out.
extern "C" int bs_debug_copy(uint64_t bytes, int32_t reps, double* ms_per_copy)
{
    if (bytes < 16 || reps <= 0) return fail(nullptr, BS_E_ARG, "bs_debug_copy: bad arguments");
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = BS_OK;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) rc = fail(nullptr, BS_E_NOMEM, "hipMalloc failed");
    if (!rc && (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
                hipEventCreate(&e1) != hipSuccess))
        rc = fail(nullptr, BS_E_HIP, "stream/event creation failed");
    if (!rc && hipMemsetAsync(a, 1, bytes, st) != hipSuccess) rc = fail(nullptr, BS_E_HIP, "memset failed");
    if (!rc) {
        (void)hipEventRecord(e0, st);
        for (int k = 0; k < reps && !rc; ++k)
            if (launch_copy(a, b, bytes, st) != hipSuccess) rc = fail(nullptr, BS_E_HIP, "copy launch failed");
        (void)hipEventRecord(e1, st);
        if (hipStreamSynchronize(st) != hipSuccess) rc = fail(nullptr, BS_E_HIP, "copy failed");
        float ms = 0.f;
        if (!rc && ms_per_copy && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) *ms_per_copy = ms / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

// Flock metrics of the current state (see boids_metrics.hip).  The
// workspace is temporary device memory allocated and released inside the
// call, or (static_ws) the device's static scratch array.  Flock k of a
// batch is the id range [boid_base, boid_base + n) of the unsorted state.
static int metrics_impl(bs_engine* e, uint32_t flock, double eps, uint32_t min_size, bs_metrics_out* out,
                        bool static_ws)
{
    if (!e || !out) return BS_E_ARG;
    if (!e->has_state) return fail(e, BS_E_STATE, "no state: call bs_set_state_* first");
    if (!(eps >= 0.0) || !isfinite(eps)) return fail(e, BS_E_ARG, "eps must be finite and >= 0");
    const uint32_t nflocks = e->flocks.empty() ? 1u : (uint32_t)e->flocks.size();
    if (flock >= nflocks) return fail(e, BS_E_ARG, "flock index out of range");
    const uint32_t ntot = e->prm.n;
    const uint32_t base = e->flocks.empty() ? 0u : e->flocks[flock].boid_base;
    const uint32_t n = e->flocks.empty() ? ntot : e->flocks[flock].P.n;
    const double F = e->flocks.empty() ? e->prm.field : e->flock_field[flock];
    memset(out, 0, sizeof(*out));
    if (n == 0) return BS_OK;
    // edge >= eps (links need only the 27 neighbouring cells); no more cells than ~8 per boid
    double g = eps > 0.0 ? floor(F / (eps * 1.0001)) : 1024.0;
    const double cap = std::max(1.0, 2.0 * floor(cbrt((double)n)));
    g = std::min(std::min(g, cap), 1024.0);
    const int G = std::max(1, (int)g);
    const size_t bytes = metrics_workspace_bytes(n, G);
    if (static_ws && bytes > kScratchBytes)
        return fail(e, BS_E_ARG, "metrics workspace of " + std::to_string(bytes) + " bytes exceeds the static scratch (" +
                                     std::to_string(kScratchBytes) + " bytes): use bs_metrics_flock");
    if (static_ws && !e->scratch) return fail(e, BS_E_STATE, "static scratch unavailable: " + e->scratch_why);
    if (is_f32(e->mode))
        HIP_TRY(e, launch_f32_to_f64(ntot, e->fa, e->tmp_pos, e->tmp_vel, e->stream));
    else
        HIP_TRY(e, launch_unsort_f64(ntot, e->da, e->tmp_pos, e->tmp_vel, e->stream));
    void* ws = nullptr;
    if (static_ws) ws = e->scratch;
    else if (hipMalloc(&ws, bytes) != hipSuccess) return fail(e, BS_E_NOMEM, "metrics workspace allocation failed");
    double* res = nullptr;
    hipError_t st = launch_metrics(n, F, G, eps, min_size, e->tmp_pos + 3 * (size_t)base,
                                   e->tmp_vel + 3 * (size_t)base, ws, static_ws ? kScratchBytes : bytes, &res,
                                   e->stream);
    double host[4] = {0, 0, 0, 0};
    if (st == hipSuccess) st = hipMemcpyAsync(host, res, sizeof(host), hipMemcpyDeviceToHost, e->stream);
    if (st == hipSuccess) st = hipStreamSynchronize(e->stream);
    if (!static_ws) (void)hipFree(ws);
    if (st != hipSuccess) return hip_fail(e, st, "metrics");
    out->polarisation = host[0];
    out->mean_nn = host[1];
    out->clusters = (uint64_t)host[2];
    out->largest = (uint64_t)host[3];
    return BS_OK;
}

extern "C" int bs_metrics_flock(bs_engine* e, uint32_t flock, double eps, uint32_t min_size, bs_metrics_out* out)
{
    return metrics_impl(e, flock, eps, min_size, out, false);
}

// The same without any device allocation: the workspace is this device's
// static scratch array (its address looked up once per device at engine
// creation).  Calls that overlap on one GPU share the array and race.
extern "C" int bs_metrics_flock_noalloc(bs_engine* e, uint32_t flock, double eps, uint32_t min_size,
                                        bs_metrics_out* out)
{
    return metrics_impl(e, flock, eps, min_size, out, true);
}

extern "C" int bs_metrics(bs_engine* e, double eps, uint32_t min_size, bs_metrics_out* out)
{
    if (!e || !out) return BS_E_ARG;
    if (e->flocks.size() > 1) return fail(e, BS_E_STATE, "batched engine: use bs_metrics_flock");
    return bs_metrics_flock(e, 0, eps, min_size, out);
}
