// First-hit tracer: seeded Lambertian rays + FP64 Amanatides-Woo traversal
// over the bit-packed occupancy grid, and per-patch integer counts.
//
// Reference: generate_rays, _initial_cell, _axis_state, _trace_kernel,
// trace_first_hit, first_hit_counts, pool_counts
// (porerad/view_factors.py:135-300).  The traversal repeats the reference's
// FP64 operations one for one (no contraction), so the integer counts are
// bit-identical to the reference CPU implementation on the same seeds.
//
// Work decomposition of the fused kernel (k_trace_fused):
//   * ray g of a launch is (seed s, patch p, ray r) with
//     g = s*(n_p*R) + (p-p0)*R + r; within one seed consecutive g are
//     consecutive positions of that seed's PCG64 stream (4 draws per ray);
//   * every lane owns a contiguous run of rays, so it jumps its stream once
//     (2^b jump tables) and then only draws; its receiver writes are
//     sequential, which keeps partial cache lines merged in L2;
//   * a wave keeps tracing while lanes are busy and refills idle lanes in
//     one batch once at least REFILL_MIN of them wait, so the ray set-up
//     (PCG64, sincos, sqrt, 9 FP64 divides) runs once per batch instead of
//     once per finishing lane, and lanes do not wait for the longest ray.
#include <climits>

#include "prh_device.h"
#include "prh_internal.h"

namespace prh {

__constant__ JumpTables c_jump;

namespace {

constexpr int TRACE_THREADS = 256;
constexpr int COUNT_THREADS = 256;
constexpr int MAX_RAYS_PER_PATCH_CALL = 16384;

struct GeoView {
    const uint32_t *occ;
    const uint32_t *cell_code;
    const int32_t *hot_map;
    const int32_t *cold_map;
    const double *patch_origin;
    const int8_t *patch_face;
    const int8_t *patch_class;
    int nx, ny, nz, nzw;
};

GeoView view_of(const Geometry &g)
{
    return GeoView{g.occ, g.cell_code, g.hot_map, g.cold_map, g.patch_origin, g.patch_face, g.patch_class,
                   g.nx, g.ny, g.nz, g.nzw};
}

enum : int { RUNNING = 0, AT_WALL = 1, AT_SOLID = 2, FAILED = 3 };

// Per-lane traversal state (the loop variables of _trace_kernel).
struct Ray {
    double oy, oz, dy, dz;
    double tx, ty, tz, ddx, ddy, ddz, time;
    int i, jj, kk;
    int sx, sy, sz;
    int face, fx, fy, fz;
    uint32_t it, max_it;
    int term;
};

__device__ inline int normal_code_of(const GeoView &G, int64_t p)
{
    const int cls = G.patch_class[p];
    return cls == 2 ? (int)G.patch_face[p] : cls;
}

// view_factors.py:211-224.  Returns false for dx == 0 (failed ray).
__device__ inline bool ray_setup(Ray &R, const double o[3], const double d[3], const GeoView &G)
{
    const double ox = o[0], dx = d[0];
    R.oy = o[1];
    R.oz = o[2];
    R.dy = d[1];
    R.dz = d[2];
    if (dx == 0.0) return false;
    R.i = initial_cell(ox, dx);
    const int j = initial_cell(R.oy, R.dy);
    const int k = initial_cell(R.oz, R.dz);
    axis_state(ox, dx, R.i, R.sx, R.tx, R.ddx);
    axis_state(R.oy, R.dy, j, R.sy, R.ty, R.ddy);
    axis_state(R.oz, R.dz, k, R.sz, R.tz, R.ddz);
    R.jj = (int)py_imod64(j, G.ny);
    R.kk = (int)py_imod64(k, G.nz);
    R.fx = R.sx > 0 ? 1 : 0;
    R.fy = R.sy > 0 ? 3 : 2;
    R.fz = R.sz > 0 ? 5 : 4;
    R.face = -1;
    R.time = 0.0;
    R.it = 0;
    R.max_it = max_steps_of(dx, R.dy, R.dz, G.nx, G.ny, G.nz);
    R.term = RUNNING;
    return true;
}

// One iteration of the traversal loop, view_factors.py:225-255.
__device__ inline void ray_step(Ray &R, const GeoView &G, uint64_t &reads)
{
    if ((unsigned)R.i >= (unsigned)G.nx) {
        R.term = AT_WALL;
        return;
    }
    const uint32_t word = G.occ[((unsigned)R.i * (unsigned)G.ny + (unsigned)R.jj) * (unsigned)G.nzw + ((unsigned)R.kk >> 5)];
    ++reads;
    if ((word >> (R.kk & 31)) & 1u) {
        R.term = AT_SOLID;
        return;
    }
    double nt = R.tx;
    if (R.ty < nt) nt = R.ty;
    if (R.tz < nt) nt = R.tz;
    const bool cx = fabs(R.tx - nt) <= TIE_TOLERANCE;
    const bool cy = fabs(R.ty - nt) <= TIE_TOLERANCE;
    const bool cz = fabs(R.tz - nt) <= TIE_TOLERANCE;
    if (cx) {
        R.i += R.sx;
        R.tx = R.tx + R.ddx;
        R.face = R.fx;
    }
    if (cy) {
        int jn = R.jj + R.sy;
        R.jj = jn == G.ny ? 0 : (jn < 0 ? G.ny - 1 : jn);
        R.ty = R.ty + R.ddy;
        if (!cx) R.face = R.fy;
    }
    if (cz) {
        int kn = R.kk + R.sz;
        R.kk = kn == G.nz ? 0 : (kn < 0 ? G.nz - 1 : kn);
        R.tz = R.tz + R.ddz;
        if (!cx && !cy) R.face = R.fz;
    }
    R.time = nt;
    if (++R.it >= R.max_it) R.term = FAILED;
}

// Receiving patch of a finished traversal (view_factors.py:226-236), -1 on failure.
__device__ inline int ray_receiver(const Ray &R, const GeoView &G)
{
    if (R.term == AT_WALL) {
        const double yw = R.oy + R.dy * R.time;
        const double zw = R.oz + R.dz * R.time;
        const int64_t jw = py_imod64((int64_t)floor(yw), G.ny);
        const int64_t kw = py_imod64((int64_t)floor(zw), G.nz);
        const int64_t c = jw * G.nz + kw;
        return R.i < 0 ? G.hot_map[c] : G.cold_map[c];
    }
    if (R.term == AT_SOLID) {
        if (R.face < 0) return -1;
        const int64_t cell = ((int64_t)R.i * G.ny + R.jj) * G.nz + R.kk;
        const uint32_t code = G.cell_code[cell];
        const uint32_t m = code & 63u;
        if (!((m >> R.face) & 1u)) return -1;
        return (int)(code >> 6) + __popc(m & ((1u << R.face) - 1u));
    }
    return -1;
}

struct FusedArgs {
    GeoView G;
    int64_t p0, n_p;
    int R, S;
    const uint64_t *seeds;  // S x {state_hi, state_lo, inc_hi, inc_lo}
    int64_t total, per_lane;
    int32_t *receiver;
    unsigned long long *reads;
    unsigned long long *failed;
};

__device__ inline void load_seed(const uint64_t *seeds, int s, u128 &st, u128 &inc)
{
    st = make_u128(seeds[4 * s + 0], seeds[4 * s + 1]);
    inc = make_u128(seeds[4 * s + 2], seeds[4 * s + 3]);
}

__device__ inline uint64_t wave_sum_u64(uint64_t v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

template <int REFILL_MIN>
__global__ __launch_bounds__(TRACE_THREADS) void k_trace_fused(FusedArgs a)
{
    const GeoView &G = a.G;
    const int64_t lane_id = (int64_t)blockIdx.x * TRACE_THREADS + threadIdx.x;
    int64_t g = lane_id * a.per_lane;
    int64_t g_end = g + a.per_lane;
    if (g_end > a.total) g_end = a.total;
    const int64_t seg = a.n_p * a.R;
    const double fny = (double)G.ny, fnz = (double)G.nz;

    int s = 0, r = 0;
    int64_t p = 0;
    u128 st = make_u128(0, 0), inc = make_u128(0, 0);
    if (g < g_end) {
        s = (int)(g / seg);
        const int64_t rem = g - (int64_t)s * seg;
        const int64_t q = rem / a.R;
        r = (int)(rem - q * a.R);
        p = a.p0 + q;
        u128 st0;
        load_seed(a.seeds, s, st0, inc);
        st = pcg_jump(st0, inc, 4ull * (uint64_t)(p * a.R + r), c_jump);
    }

    Ray R;
    R.term = RUNNING;
    R.i = 0;
    bool active = false;
    int64_t g_cur = -1;
    uint64_t reads = 0, failed = 0;

    while (true) {
        if (!active) {
            if (g_cur >= 0) {
                const int rec = ray_receiver(R, G);
                failed += rec < 0 ? 1u : 0u;
                a.receiver[g_cur] = rec;
                g_cur = -1;
            }
            while (g < g_end) {
                const double *po = G.patch_origin + 3 * p;
                double o[3], d[3];
                make_ray(st, inc, po[0], po[1], po[2], normal_code_of(G, p), fny, fnz, o, d);
                g_cur = g;
                ++g;
                if (++r == a.R) {
                    r = 0;
                    if (++p == a.p0 + a.n_p) {
                        p = a.p0;
                        ++s;
                        if (g < g_end) {
                            u128 st0;
                            load_seed(a.seeds, s, st0, inc);
                            st = pcg_jump(st0, inc, 4ull * (uint64_t)(a.p0 * a.R), c_jump);
                        }
                    }
                }
                if (!ray_setup(R, o, d, G)) {
                    a.receiver[g_cur] = -1;
                    ++failed;
                    g_cur = -1;
                    continue;
                }
                active = true;
                break;
            }
        }
        if (__ballot(active) == 0) break;
        while (true) {
            if (active) {
                ray_step(R, G, reads);
                if (R.term != RUNNING) active = false;
            }
            if (__ballot(active) == 0) break;
            if (__popcll(__ballot(!active && g < g_end)) >= REFILL_MIN) break;
        }
    }
    reads = wave_sum_u64(reads);
    failed = wave_sum_u64(failed);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(a.reads, (unsigned long long)reads);
        if (failed) atomicAdd(a.failed, (unsigned long long)failed);
    }
}

// Per-patch sorted unique receivers and their counts (the np.unique of
// first_hit_counts, view_factors.py:284-286, pooled over seeds as
// pool_counts does, view_factors.py:290-300).  One workgroup per patch,
// bitonic sort of the patch's S*R receivers in LDS.
__global__ __launch_bounds__(COUNT_THREADS) void k_count_patches(const int32_t *receiver, int64_t n_p, int R, int S,
                                                                 int SRp2, int32_t *uniq_recv, int32_t *uniq_cnt,
                                                                 int32_t *n_unique)
{
    extern __shared__ int32_t lds[];
    int32_t *key = lds;
    int32_t *part = lds + SRp2;
    const int SR = S * R;
    const int chunk = SRp2 >= COUNT_THREADS ? SRp2 / COUNT_THREADS : 1;
    for (int64_t q = blockIdx.x; q < n_p; q += gridDim.x) {
        for (int t = threadIdx.x; t < SRp2; t += COUNT_THREADS) {
            int32_t v = INT_MAX;
            if (t < SR) {
                const int s = t / R, r = t - s * R;
                v = receiver[(int64_t)s * n_p * R + q * R + r];
            }
            key[t] = v;
        }
        __syncthreads();
        for (int k = 2; k <= SRp2; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int t = threadIdx.x; t < SRp2; t += COUNT_THREADS) {
                    const int u = t ^ j;
                    if (u > t) {
                        const int32_t x = key[t], y = key[u];
                        const bool up = (t & k) == 0;
                        if ((x > y) == up) {
                            key[t] = y;
                            key[u] = x;
                        }
                    }
                }
                __syncthreads();
            }
        }
        // Run starts per thread chunk, block scan for ranks.
        const int b0 = threadIdx.x * chunk;
        int starts = 0;
        if (b0 < SRp2) {
            for (int e = b0; e < b0 + chunk && e < SR; ++e) starts += (e == 0 || key[e] != key[e - 1]) ? 1 : 0;
        }
        part[threadIdx.x] = starts;
        __syncthreads();
        for (int off = 1; off < COUNT_THREADS; off <<= 1) {
            const int add = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
            __syncthreads();
            part[threadIdx.x] += add;
            __syncthreads();
        }
        int rank = part[threadIdx.x] - starts;
        const int total = part[COUNT_THREADS - 1];
        if (b0 < SRp2) {
            for (int e = b0; e < b0 + chunk && e < SR; ++e) {
                if (e == 0 || key[e] != key[e - 1]) {
                    int end = e + 1;
                    while (end < SR && key[end] == key[e]) ++end;
                    uniq_recv[q * SR + rank] = key[e];
                    uniq_cnt[q * SR + rank] = end - e;
                    ++rank;
                }
            }
        }
        if (threadIdx.x == 0) n_unique[q] = total;
        __syncthreads();
    }
}

__global__ __launch_bounds__(COUNT_THREADS) void k_compact_edges(const int32_t *uniq_recv, const int32_t *uniq_cnt,
                                                                 const int32_t *n_unique, const int64_t *offsets,
                                                                 int64_t n_p, int SR, int64_t p0, int32_t *origin,
                                                                 int32_t *receiver, int32_t *count)
{
    for (int64_t q = blockIdx.x; q < n_p; q += gridDim.x) {
        const int64_t off = offsets[q];
        const int nu = n_unique[q];
        for (int u = threadIdx.x; u < nu; u += COUNT_THREADS) {
            origin[off + u] = (int32_t)(p0 + q);
            receiver[off + u] = uniq_recv[q * SR + u];
            count[off + u] = uniq_cnt[q * SR + u];
        }
    }
}

// generate_rays for the API (one thread per ray, full jump per ray).
__global__ __launch_bounds__(TRACE_THREADS) void k_generate(GeoView G, int64_t p0, int64_t n_p, int R, u128 st0,
                                                            u128 inc, double *origins, double *directions)
{
    const int64_t m = (int64_t)blockIdx.x * TRACE_THREADS + threadIdx.x;
    if (m >= n_p * R) return;
    const int64_t q = m / R;
    const int64_t r = m - q * R;
    const int64_t p = p0 + q;
    u128 st = pcg_jump(st0, inc, 4ull * (uint64_t)(p * R + r), c_jump);
    const double *po = G.patch_origin + 3 * p;
    double o[3], d[3];
    make_ray(st, inc, po[0], po[1], po[2], normal_code_of(G, p), (double)G.ny, (double)G.nz, o, d);
    for (int c = 0; c < 3; ++c) {
        origins[3 * m + c] = o[c];
        directions[3 * m + c] = d[c];
    }
}

// trace_first_hit for given rays (one thread per ray).
__global__ __launch_bounds__(TRACE_THREADS) void k_trace_given(GeoView G, const double *origins,
                                                               const double *directions, int64_t n_rays,
                                                               int32_t *receiver, unsigned long long *failed)
{
    const int64_t m = (int64_t)blockIdx.x * TRACE_THREADS + threadIdx.x;
    if (m >= n_rays) return;
    double o[3] = {origins[3 * m], origins[3 * m + 1], origins[3 * m + 2]};
    double d[3] = {directions[3 * m], directions[3 * m + 1], directions[3 * m + 2]};
    Ray R;
    int rec = -1;
    if (ray_setup(R, o, d, G)) {
        uint64_t reads = 0;
        while (R.term == RUNNING) ray_step(R, G, reads);
        rec = ray_receiver(R, G);
    }
    receiver[m] = rec;
    if (rec < 0) atomicAdd(failed, 1ull);
}

// ------------------------------------------------------------------ host side
bool g_tables_ready[64] = {};

int ensure_tables()
{
    int dev = 0;
    PRH_HIP(hipGetDevice(&dev));
    if (dev >= 0 && dev < 64 && g_tables_ready[dev]) return PRH_OK;
    JumpTables t;
    host_jump_tables(&t);
    PRH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_jump), &t, sizeof(t)));
    if (dev >= 0 && dev < 64) g_tables_ready[dev] = true;
    return PRH_OK;
}

struct Workspace {
    void *seeds = nullptr;
    size_t seeds_cap = 0;
    void *receiver = nullptr;
    size_t receiver_cap = 0;
    void *uniq_recv = nullptr;
    size_t uniq_recv_cap = 0;
    void *uniq_cnt = nullptr;
    size_t uniq_cnt_cap = 0;
    void *n_unique = nullptr;
    size_t n_unique_cap = 0;
    void *offsets = nullptr;
    size_t offsets_cap = 0;
    void *counters = nullptr;
    size_t counters_cap = 0;
    hipEvent_t ev[4] = {};
    bool events = false;
};

Workspace g_ws[64];
double g_last_ms[3] = {0, 0, 0};

int resident_blocks(const void *kernel, int threads, int *blocks)
{
    int dev = 0, cus = 0, per_cu = 0;
    PRH_HIP(hipGetDevice(&dev));
    PRH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    PRH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0));
    if (per_cu < 1) per_cu = 1;
    *blocks = cus * per_cu;
    return PRH_OK;
}

}  // namespace

}  // namespace prh

struct prh_edges {
    int device;
    int64_t nnz;
    int32_t *origin;
    int32_t *receiver;
    int32_t *count;
};

using namespace prh;

extern "C" {

int prh_last_timings(double *ms, int32_t n)
{
    PRH_CHECK_ARG(ms && n >= 0, "prh_last_timings: null output");
    for (int i = 0; i < n && i < 3; ++i) ms[i] = g_last_ms[i];
    return PRH_OK;
}

int prh_generate_rays(const prh_geometry *geo, int64_t p0, int64_t p1, int32_t rays_per_patch,
                      const uint64_t state_inc[4], double *origins, double *directions)
{
    PRH_CHECK_ARG(geo && state_inc && origins && directions, "prh_generate_rays: null pointer");
    const Geometry &g = geo->g;
    PRH_CHECK_ARG(0 <= p0 && p0 <= p1 && p1 <= g.n_patches, "prh_generate_rays: patch range out of bounds");
    PRH_CHECK_ARG(rays_per_patch > 0, "prh_generate_rays: rays_per_patch must be positive");
    const int64_t m = (p1 - p0) * (int64_t)rays_per_patch;
    if (m == 0) return PRH_OK;
    PRH_HIP(hipSetDevice(g.device));
    int rc = ensure_tables();
    if (rc) return rc;
    hipStream_t st;
    if ((rc = current_stream(&st))) return rc;
    double *d_o = nullptr, *d_d = nullptr;
    PRH_HIP(hipMallocAsync((void **)&d_o, m * 3 * sizeof(double), st));
    PRH_HIP(hipMallocAsync((void **)&d_d, m * 3 * sizeof(double), st));
    const unsigned blocks = (unsigned)((m + TRACE_THREADS - 1) / TRACE_THREADS);
    k_generate<<<blocks, TRACE_THREADS, 0, st>>>(view_of(g), p0, p1 - p0, rays_per_patch,
                                                 make_u128(state_inc[0], state_inc[1]),
                                                 make_u128(state_inc[2], state_inc[3]), d_o, d_d);
    PRH_HIP(hipGetLastError());
    PRH_HIP(hipMemcpyAsync(origins, d_o, m * 3 * sizeof(double), hipMemcpyDeviceToHost, st));
    PRH_HIP(hipMemcpyAsync(directions, d_d, m * 3 * sizeof(double), hipMemcpyDeviceToHost, st));
    PRH_HIP(hipFreeAsync(d_o, st));
    PRH_HIP(hipFreeAsync(d_d, st));
    PRH_HIP(hipStreamSynchronize(st));
    return PRH_OK;
}

int prh_trace_rays(const prh_geometry *geo, const double *origins, const double *directions, int64_t n_rays,
                   int32_t *receiver, int64_t *n_failed)
{
    PRH_CHECK_ARG(geo && receiver && (n_rays == 0 || (origins && directions)), "prh_trace_rays: null pointer");
    PRH_CHECK_ARG(n_rays >= 0, "prh_trace_rays: negative ray count");
    if (n_failed) *n_failed = 0;
    if (n_rays == 0) return PRH_OK;
    const Geometry &g = geo->g;
    PRH_HIP(hipSetDevice(g.device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    double *d_o = nullptr, *d_d = nullptr;
    int32_t *d_r = nullptr;
    unsigned long long *d_f = nullptr;
    PRH_HIP(hipMallocAsync((void **)&d_o, n_rays * 3 * sizeof(double), st));
    PRH_HIP(hipMallocAsync((void **)&d_d, n_rays * 3 * sizeof(double), st));
    PRH_HIP(hipMallocAsync((void **)&d_r, n_rays * sizeof(int32_t), st));
    PRH_HIP(hipMallocAsync((void **)&d_f, sizeof(unsigned long long), st));
    PRH_HIP(hipMemcpyAsync(d_o, origins, n_rays * 3 * sizeof(double), hipMemcpyHostToDevice, st));
    PRH_HIP(hipMemcpyAsync(d_d, directions, n_rays * 3 * sizeof(double), hipMemcpyHostToDevice, st));
    PRH_HIP(hipMemsetAsync(d_f, 0, sizeof(unsigned long long), st));
    const unsigned blocks = (unsigned)((n_rays + TRACE_THREADS - 1) / TRACE_THREADS);
    k_trace_given<<<blocks, TRACE_THREADS, 0, st>>>(view_of(g), d_o, d_d, n_rays, d_r, d_f);
    PRH_HIP(hipGetLastError());
    unsigned long long failed = 0;
    PRH_HIP(hipMemcpyAsync(receiver, d_r, n_rays * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    PRH_HIP(hipMemcpyAsync(&failed, d_f, sizeof(failed), hipMemcpyDeviceToHost, st));
    PRH_HIP(hipFreeAsync(d_o, st));
    PRH_HIP(hipFreeAsync(d_d, st));
    PRH_HIP(hipFreeAsync(d_r, st));
    PRH_HIP(hipFreeAsync(d_f, st));
    PRH_HIP(hipStreamSynchronize(st));
    if (n_failed) *n_failed = (int64_t)failed;
    if (failed) {
        set_error("%lld rays did not resolve a receiving patch", (long long)failed);
        return PRH_ERR_UNRESOLVED;
    }
    return PRH_OK;
}

int prh_trace_counts(const prh_geometry *geo, int64_t p0, int64_t p1, int32_t rays_per_patch, int32_t n_seeds,
                     const uint64_t *state_inc, prh_edges **out, int64_t *nnz_out, int64_t *occupancy_reads)
{
    PRH_CHECK_ARG(geo && state_inc && out, "prh_trace_counts: null pointer");
    const Geometry &g = geo->g;
    PRH_CHECK_ARG(0 <= p0 && p0 <= p1 && p1 <= g.n_patches, "prh_trace_counts: patch range out of bounds");
    PRH_CHECK_ARG(rays_per_patch > 0 && n_seeds > 0, "prh_trace_counts: rays_per_patch and n_seeds must be positive");
    const int SR = rays_per_patch * n_seeds;
    if ((int64_t)rays_per_patch * n_seeds > MAX_RAYS_PER_PATCH_CALL) {
        set_error("prh_trace_counts: rays_per_patch * n_seeds = %lld exceeds %d",
                  (long long)rays_per_patch * n_seeds, MAX_RAYS_PER_PATCH_CALL);
        return PRH_ERR_CAPACITY;
    }
    *out = nullptr;
    PRH_HIP(hipSetDevice(g.device));
    int dev = g.device;
    int rc = ensure_tables();
    if (rc) return rc;
    hipStream_t st;
    if ((rc = current_stream(&st))) return rc;
    Workspace &W = g_ws[dev];
    if (!W.events) {
        for (auto &e : W.ev) PRH_HIP(hipEventCreate(&e));
        W.events = true;
    }
    const int64_t n_p = p1 - p0;
    const int64_t total = n_p * SR;
    int SRp2 = 1;
    while (SRp2 < SR) SRp2 <<= 1;

    if ((rc = ensure_capacity(&W.seeds, &W.seeds_cap, sizeof(uint64_t) * 4 * n_seeds))) return rc;
    if ((rc = ensure_capacity(&W.receiver, &W.receiver_cap, sizeof(int32_t) * (total > 0 ? total : 1)))) return rc;
    if ((rc = ensure_capacity(&W.uniq_recv, &W.uniq_recv_cap, sizeof(int32_t) * (total > 0 ? total : 1)))) return rc;
    if ((rc = ensure_capacity(&W.uniq_cnt, &W.uniq_cnt_cap, sizeof(int32_t) * (total > 0 ? total : 1)))) return rc;
    if ((rc = ensure_capacity(&W.n_unique, &W.n_unique_cap, sizeof(int32_t) * (n_p > 0 ? n_p : 1)))) return rc;
    if ((rc = ensure_capacity(&W.offsets, &W.offsets_cap, sizeof(int64_t) * (n_p > 0 ? n_p : 1)))) return rc;
    if ((rc = ensure_capacity(&W.counters, &W.counters_cap, sizeof(unsigned long long) * 2))) return rc;
    PRH_HIP(hipMemcpyAsync(W.seeds, state_inc, sizeof(uint64_t) * 4 * n_seeds, hipMemcpyHostToDevice, st));
    PRH_HIP(hipMemsetAsync(W.counters, 0, sizeof(unsigned long long) * 2, st));
    unsigned long long *d_reads = (unsigned long long *)W.counters;
    unsigned long long *d_failed = d_reads + 1;

    int64_t nnz = 0;
    int32_t *d_origin = nullptr, *d_receiver = nullptr, *d_count = nullptr;
    unsigned long long counters[2] = {0, 0};
    if (total > 0) {
        constexpr int REFILL = 16;
        int blocks = 0;
        if ((rc = resident_blocks((const void *)k_trace_fused<REFILL>, TRACE_THREADS, &blocks))) return rc;
        const int64_t lanes_max = (int64_t)blocks * TRACE_THREADS;
        const int64_t per_lane = (total + lanes_max - 1) / lanes_max;
        const int64_t lanes = (total + per_lane - 1) / per_lane;
        const unsigned grid = (unsigned)((lanes + TRACE_THREADS - 1) / TRACE_THREADS);
        FusedArgs a{view_of(g), p0, n_p, rays_per_patch, n_seeds, (const uint64_t *)W.seeds, total, per_lane,
                    (int32_t *)W.receiver, d_reads, d_failed};
        PRH_HIP(hipEventRecord(W.ev[0], st));
        k_trace_fused<REFILL><<<grid, TRACE_THREADS, 0, st>>>(a);
        PRH_HIP(hipGetLastError());
        PRH_HIP(hipEventRecord(W.ev[1], st));
        const unsigned cgrid = (unsigned)(n_p < 65536 ? n_p : 65536);
        const size_t lds = sizeof(int32_t) * (SRp2 + COUNT_THREADS);
        k_count_patches<<<cgrid, COUNT_THREADS, lds, st>>>((const int32_t *)W.receiver, n_p, rays_per_patch,
                                                           n_seeds, SRp2, (int32_t *)W.uniq_recv,
                                                           (int32_t *)W.uniq_cnt, (int32_t *)W.n_unique);
        PRH_HIP(hipGetLastError());
        PRH_HIP(hipEventRecord(W.ev[2], st));
        PRH_HIP(hipMemcpyAsync(counters, W.counters, sizeof(counters), hipMemcpyDeviceToHost, st));
        PRH_HIP(hipStreamSynchronize(st));
        if (counters[1]) {
            set_error("%llu rays did not resolve a receiving patch", counters[1]);
            return PRH_ERR_UNRESOLVED;
        }
        if ((rc = exclusive_scan_i32_to_i64((const int32_t *)W.n_unique, (int64_t *)W.offsets, n_p, &nnz, st)))
            return rc;
        const int64_t alloc = nnz > 0 ? nnz : 1;
        PRH_HIP(hipMalloc((void **)&d_origin, alloc * sizeof(int32_t)));
        PRH_HIP(hipMalloc((void **)&d_receiver, alloc * sizeof(int32_t)));
        PRH_HIP(hipMalloc((void **)&d_count, alloc * sizeof(int32_t)));
        k_compact_edges<<<cgrid, COUNT_THREADS, 0, st>>>((const int32_t *)W.uniq_recv, (const int32_t *)W.uniq_cnt,
                                                         (const int32_t *)W.n_unique, (const int64_t *)W.offsets,
                                                         n_p, SR, p0, d_origin, d_receiver, d_count);
        PRH_HIP(hipGetLastError());
        PRH_HIP(hipEventRecord(W.ev[3], st));
        PRH_HIP(hipStreamSynchronize(st));
        float ms = 0.f;
        PRH_HIP(hipEventElapsedTime(&ms, W.ev[0], W.ev[1]));
        g_last_ms[0] = ms;
        PRH_HIP(hipEventElapsedTime(&ms, W.ev[1], W.ev[2]));
        g_last_ms[1] = ms;
        PRH_HIP(hipEventElapsedTime(&ms, W.ev[2], W.ev[3]));
        g_last_ms[2] = ms;
    }
    prh_edges *e = new prh_edges{dev, nnz, d_origin, d_receiver, d_count};
    *out = e;
    if (nnz_out) *nnz_out = nnz;
    if (occupancy_reads) *occupancy_reads = (int64_t)counters[0];
    return PRH_OK;
}

int prh_edges_copy(const prh_edges *e, int32_t *origin, int32_t *receiver, int64_t *count)
{
    PRH_CHECK_ARG(e, "prh_edges_copy: null edge table");
    if (e->nnz == 0) return PRH_OK;
    PRH_HIP(hipSetDevice(e->device));
    if (origin) PRH_HIP(hipMemcpy(origin, e->origin, e->nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (receiver) PRH_HIP(hipMemcpy(receiver, e->receiver, e->nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (count) {
        // Widen the int32 device counts into the int64 host array in place:
        // copy into the upper half of the output buffer, then expand upwards.
        int32_t *tmp = reinterpret_cast<int32_t *>(count) + e->nnz;
        PRH_HIP(hipMemcpy(tmp, e->count, e->nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
        for (int64_t k = 0; k < e->nnz; ++k) count[k] = (int64_t)tmp[k];
    }
    return PRH_OK;
}

void prh_edges_destroy(prh_edges *e)
{
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->origin) (void)hipFree(e->origin);
    if (e->receiver) (void)hipFree(e->receiver);
    if (e->count) (void)hipFree(e->count);
    delete e;
}

}  // extern "C"
