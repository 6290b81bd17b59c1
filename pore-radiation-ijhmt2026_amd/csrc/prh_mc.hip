// Multi-bounce Monte Carlo ray statistics of a slab of microstructure.
//
// Reference: trace_properties (porerad/properties.py:54-307) and the ray
// kernel of trace_penetration (porerad/penetration.py:54-305), written there
// in Taichi.  Rays enter through x = 0 at uniformly distributed transverse
// positions (periodic in y and z) and are followed through the pore space;
// at a solid surface a ray is absorbed with probability eps, otherwise it is
// reflected into a Lambertian direction about the surface normal, for at
// most max_bounces reflections.  Two restart rules for a reflected ray:
// "article" (the second and later reflections restart at the hit point
// advanced once more along the preceding segment) and "hit_point"; both
// displace the restart point by 1e-9 cells along the normal.
//
// One thread per ray; the random stream of a ray is the reference's 32-bit
// hash chain (_hash_u32 / _rand01), so every ray follows the same sequence of
// draws.  The traversal (_trace) repeats the reference's FP64 operations one
// for one (compiled without contraction): initial cell with the
// |o - floor(o + 1/2)| <= 1e-12 rule, 1e300 for axes with d = 0, the 1e-12
// tie tolerance, the face normal of the last crossing, the step cap
// 10 nx ny nz + 10.  Per-ray results are written to arrays and reduced in a
// fixed order (the reference accumulates with atomics in no fixed order).
//
// The Lambertian azimuth is phi = 2 pi m / 2^24 for the 24-bit integer m of
// the second draw; cos(phi) and sin(phi) come from prh_libm.h, which gives
// the host C library's bits (the functions the reference's CPU kernel calls;
// checked on all 2^24 values of phi).  The rays then follow the reference's
// paths bit for bit.
#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "prh_internal.h"
#include "prh_libm.h"

namespace prh {
namespace {

constexpr int MC_THREADS = 256;
constexpr double TIE = 1.0e-12;
constexpr double RESTART_OFFSET = 1.0e-9;

struct McView {
    const uint32_t *occ;
    int nx, ny, nz;
    int max_steps;
};

__device__ __forceinline__ uint32_t hash_u32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

// _rand01: the next state and a uniform in [0, 1) with 24 bits.
__device__ __forceinline__ double rand01(uint32_t &state)
{
    state = hash_u32(state + 0x9E3779B9u);
    return (double)(state & 0x00FFFFFFu) / (double)0x01000000u;
}

constexpr double AZIMUTH_STEPS = 16777216.0;  // 2^24
constexpr double TWO_PI = 6.283185307179586;  // 2.0 * math.pi, exact

__device__ __forceinline__ int floor_mod(int a, int n)
{
    const int r = a % n;
    return r < 0 ? r + n : r;
}

__device__ __forceinline__ bool solid_at(const McView &V, int i, int jj, int kk)
{
    const uint32_t c = ((uint32_t)i * (uint32_t)V.ny + (uint32_t)jj) * (uint32_t)V.nz + (uint32_t)kk;
    return (V.occ[c >> 5] >> (c & 31u)) & 1u;
}

struct Hit {
    int fate;  // 0 solid, 1 exit at x = 0, 2 exit at x = nx, 3 step cap
    double t, nx, ny, nz;
};

// _trace: cell-by-cell traversal from o along d (normalised here).
__device__ Hit mc_trace(const McView &V, double ox, double oy, double oz, double dx0, double dy0, double dz0)
{
    const double norm = sqrt(dx0 * dx0 + dy0 * dy0 + dz0 * dz0);
    const double dx = dx0 / norm, dy = dy0 / norm, dz = dz0 / norm;
    int i = (int)floor(ox), j = (int)floor(oy), k = (int)floor(oz);
    if (dx < 0.0 && fabs(ox - floor(ox + 0.5)) <= TIE) i -= 1;
    if (dy < 0.0 && fabs(oy - floor(oy + 0.5)) <= TIE) j -= 1;
    if (dz < 0.0 && fabs(oz - floor(oz + 0.5)) <= TIE) k -= 1;
    int sx = 0, sy = 0, sz = 0;
    double tx = 1.0e300, ty = 1.0e300, tz = 1.0e300, ddx = 1.0e300, ddy = 1.0e300, ddz = 1.0e300;
    const double safe_dx = fabs(dx) > 0.0 ? dx : 1.0;
    const double safe_dy = fabs(dy) > 0.0 ? dy : 1.0;
    const double safe_dz = fabs(dz) > 0.0 ? dz : 1.0;
    if (dx > 0.0) {
        sx = 1;
        tx = ((double)(i + 1) - ox) / safe_dx;
        ddx = 1.0 / safe_dx;
    } else if (dx < 0.0) {
        sx = -1;
        tx = ((double)i - ox) / safe_dx;
        ddx = -1.0 / safe_dx;
    }
    if (dy > 0.0) {
        sy = 1;
        ty = ((double)(j + 1) - oy) / safe_dy;
        ddy = 1.0 / safe_dy;
    } else if (dy < 0.0) {
        sy = -1;
        ty = ((double)j - oy) / safe_dy;
        ddy = -1.0 / safe_dy;
    }
    if (dz > 0.0) {
        sz = 1;
        tz = ((double)(k + 1) - oz) / safe_dz;
        ddz = 1.0 / safe_dz;
    } else if (dz < 0.0) {
        sz = -1;
        tz = ((double)k - oz) / safe_dz;
        ddz = -1.0 / safe_dz;
    }
    Hit h{-1, 0.0, 0.0, 0.0, 0.0};
    int steps = 0;
    while (h.fate == -1 && steps < V.max_steps) {
        steps += 1;
        if (i < 0) {
            h.fate = 1;
        } else if (i >= V.nx) {
            h.fate = 2;
        } else if (solid_at(V, i, floor_mod(j, V.ny), floor_mod(k, V.nz))) {
            h.fate = 0;
        } else {
            const double t_next = fmin(tx, fmin(ty, tz));
            const bool cx = fabs(tx - t_next) <= TIE;
            const bool cy = fabs(ty - t_next) <= TIE;
            const bool cz = fabs(tz - t_next) <= TIE;
            if (cx) {
                i += sx;
                tx += ddx;
                h.nx = -(double)sx;
                h.ny = 0.0;
                h.nz = 0.0;
            }
            if (cy) {
                j += sy;
                ty += ddy;
                if (!cx) {
                    h.nx = 0.0;
                    h.ny = -(double)sy;
                    h.nz = 0.0;
                }
            }
            if (cz) {
                k += sz;
                tz += ddz;
                if (!cx && !cy) {
                    h.nx = 0.0;
                    h.ny = 0.0;
                    h.nz = -(double)sz;
                }
            }
            h.t = t_next;
        }
    }
    if (h.fate == -1) h.fate = 3;
    return h;
}

// _lambertian: a cosine-weighted direction about the unit normal n.
__device__ __forceinline__ void lambertian(const McView &V, uint32_t &state, double nxn, double nyn, double nzn,
                                           double &dx, double &dy, double &dz)
{
    const double xi1 = rand01(state);
    (void)rand01(state);  // xi2 = m / 2^24 with m the low 24 bits of the state
    const double phi = TWO_PI * ((double)(state & 0x00FFFFFFu) / AZIMUTH_STEPS);
    double sin_phi, cos_phi;
    libm::sincos(phi, sin_phi, cos_phi);
    double hx = 1.0, hy = 0.0, hz = 0.0;
    if (fabs(nxn) > 0.9) {
        hx = 0.0;
        hy = 1.0;
        hz = 0.0;
    }
    double ux = hy * nzn - hz * nyn;
    double uy = hz * nxn - hx * nzn;
    double uz = hx * nyn - hy * nxn;
    const double un = sqrt(ux * ux + uy * uy + uz * uz);
    ux /= un;
    uy /= un;
    uz /= un;
    const double vx = nyn * uz - nzn * uy;
    const double vy = nzn * ux - nxn * uz;
    const double vz = nxn * uy - nyn * ux;
    const double radius = sqrt(xi1);
    const double cx = radius * cos_phi;
    const double cy = radius * sin_phi;
    const double cz = sqrt(fmax(0.0, 1.0 - xi1));
    dx = cx * ux + cy * vx + cz * nxn;
    dy = cx * uy + cy * vy + cz * nyn;
    dz = cx * uz + cy * vz + cz * nzn;
}

// The reflection loop shared by both kernels, from a solid hit with normal n
// at distance dist along d from o.  Returns the final fate: 0 absorbed,
// 1/2 exit through x = 0 / x = nx, 3 stopped (step cap or reflection limit).
// mu_sum / mu_sq / scattered / interactions accumulate the properties'
// statistics.
template <bool HIT_POINT>
__device__ int reflect_loop(const McView &V, uint32_t &state, double eps, int max_bounces, double &ox, double &oy,
                            double &oz, double &dx, double &dy, double &dz, Hit h, int64_t &interactions,
                            int64_t &scattered, double &mu_sum, double &mu_sq)
{
    double nxn = h.nx, nyn = h.ny, nzn = h.nz, dist = h.t;
    int bounces = 0;
    while (true) {
        interactions += 1;
        const double u = rand01(state);
        if (u < eps) return 0;
        const double pdx = dx, pdy = dy, pdz = dz;
        lambertian(V, state, nxn, nyn, nzn, dx, dy, dz);
        const double mu = pdx * dx + pdy * dy + pdz * dz;
        scattered += 1;
        mu_sum += mu;
        mu_sq += mu * mu;
        bounces += 1;
        if (bounces >= max_bounces) return 3;
        if (HIT_POINT) {
            ox = ox + nxn * RESTART_OFFSET;
            oy = oy + nyn * RESTART_OFFSET;
            oz = oz + nzn * RESTART_OFFSET;
        } else {
            ox = ox + pdx * dist + nxn * RESTART_OFFSET;
            oy = oy + pdy * dist + nyn * RESTART_OFFSET;
            oz = oz + pdz * dist + nzn * RESTART_OFFSET;
        }
        const Hit g = mc_trace(V, ox, oy, oz, dx, dy, dz);
        if (g.fate != 0) return g.fate;
        dist = g.t;
        nxn = g.nx;
        nyn = g.ny;
        nzn = g.nz;
        if (nxn * nxn + nyn * nyn + nzn * nzn <= 1.0e-24) {
            nxn = -dx;
            nyn = -dy;
            nzn = -dz;
        }
        ox = ox + dx * dist;
        oy = oy + dy * dist;
        oz = oz + dz * dist;
    }
}

__device__ __forceinline__ uint32_t ray_state(uint32_t seed, int64_t ray)
{
    return hash_u32(seed + (uint32_t)(ray + 1) * 747796405u);
}

// trace_properties's kernel: beam rays along +x.  Per ray: the first path
// length as an event (solid hit) or censored length (else; -1 in the other
// array), the interactions and reflections, and the sums of mu and mu^2.
template <bool HIT_POINT>
__global__ __launch_bounds__(MC_THREADS) void k_mc_properties(McView V, int64_t n, uint32_t seed, double eps,
                                                             int max_bounces, double *event_len, double *censored_len,
                                                             int64_t *inter, int64_t *scat, double *mu1, double *mu2)
{
    const int64_t ray = (int64_t)blockIdx.x * MC_THREADS + threadIdx.x;
    if (ray >= n) return;
    uint32_t state = ray_state(seed, ray);
    const double ry = rand01(state);
    const double rz = rand01(state);
    double ox = 0.0, oy = ry * (double)V.ny, oz = rz * (double)V.nz;
    double dx = 1.0, dy = 0.0, dz = 0.0;
    Hit h = mc_trace(V, ox, oy, oz, dx, dy, dz);
    int64_t interactions = 0, scattered = 0;
    double s1 = 0.0, s2 = 0.0;
    if (h.fate == 0) {
        if (h.nx * h.nx + h.ny * h.ny + h.nz * h.nz <= 1.0e-24) {
            h.nx = -dx;
            h.ny = -dy;
            h.nz = -dz;
        }
        event_len[ray] = h.t;
        censored_len[ray] = -1.0;
        if (HIT_POINT) {
            ox = ox + dx * h.t;
            oy = oy + dy * h.t;
            oz = oz + dz * h.t;
        }
        (void)reflect_loop<HIT_POINT>(V, state, eps, max_bounces, ox, oy, oz, dx, dy, dz, h, interactions, scattered,
                                      s1, s2);
    } else {
        event_len[ray] = -1.0;
        censored_len[ray] = h.t;
    }
    inter[ray] = interactions;
    scat[ray] = scattered;
    mu1[ray] = s1;
    mu2[ray] = s2;
}

// trace_penetration's kernel: Lambertian rays about +x; the fate of each ray
// (0 reflected, 1 transmitted, 2 absorbed, 3 stopped).
template <bool HIT_POINT>
__global__ __launch_bounds__(MC_THREADS) void k_mc_penetration(McView V, int64_t n, uint32_t seed, double eps,
                                                              int max_bounces, uint8_t *fate)
{
    const int64_t ray = (int64_t)blockIdx.x * MC_THREADS + threadIdx.x;
    if (ray >= n) return;
    uint32_t state = ray_state(seed, ray);
    const double ry = rand01(state);
    const double rz = rand01(state);
    double ox = 0.0, oy = ry * (double)V.ny, oz = rz * (double)V.nz;
    double dx, dy, dz;
    lambertian(V, state, 1.0, 0.0, 0.0, dx, dy, dz);
    Hit h = mc_trace(V, ox, oy, oz, dx, dy, dz);
    int code;
    if (h.fate != 0) {
        code = h.fate == 1 ? 0 : (h.fate == 2 ? 1 : 3);
    } else {
        if (h.nx * h.nx + h.ny * h.ny + h.nz * h.nz <= 1.0e-24) {
            h.nx = -dx;
            h.ny = -dy;
            h.nz = -dz;
        }
        if (HIT_POINT) {
            ox = ox + dx * h.t;
            oy = oy + dy * h.t;
            oz = oz + dz * h.t;
        }
        int64_t inter = 0, scat = 0;
        double s1 = 0.0, s2 = 0.0;
        const int f = reflect_loop<HIT_POINT>(V, state, eps, max_bounces, ox, oy, oz, dx, dy, dz, h, inter, scat, s1,
                                              s2);
        code = f == 0 ? 2 : (f == 1 ? 0 : (f == 2 ? 1 : 3));
    }
    fate[ray] = (uint8_t)code;
}

// Fixed-order block sums of four per-ray quantities (two integer, two FP64),
// then their totals in block order.
__global__ __launch_bounds__(MC_THREADS) void k_mc_block_sums(const int64_t *a, const int64_t *b, const double *c,
                                                             const double *d, const uint8_t *fate, int64_t n,
                                                             double *part)
{
    __shared__ double sh[MC_THREADS][6];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const int64_t i = (int64_t)blockIdx.x * MC_THREADS + threadIdx.x;
    if (i < n) {
        if (fate) {
            acc[fate[i]] = 1.0;
        } else {
            acc[0] = (double)a[i];
            acc[1] = (double)b[i];
            acc[4] = c[i];
            acc[5] = d[i];
        }
    }
    for (int q = 0; q < 6; ++q) sh[threadIdx.x][q] = acc[q];
    __syncthreads();
    for (int off = MC_THREADS / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            for (int q = 0; q < 6; ++q) sh[threadIdx.x][q] += sh[threadIdx.x + off][q];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 6; ++q) part[(size_t)q * gridDim.x + blockIdx.x] = sh[0][q];
}

__global__ __launch_bounds__(MC_THREADS) void k_mc_totals(const double *part, int64_t n_part, double *out)
{
    __shared__ double sh[MC_THREADS];
    const int q = blockIdx.x;
    double s = 0.0;
    for (int64_t b = threadIdx.x; b < n_part; b += MC_THREADS) s += part[(size_t)q * n_part + b];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = MC_THREADS / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = sh[0];
}

McView view_of(const Geometry &g)
{
    const int64_t cap = 10 * (int64_t)g.nx * g.ny * g.nz + 10;
    return McView{g.occ, g.nx, g.ny, g.nz, (int)std::min<int64_t>(cap, INT32_MAX)};
}

// Device buffers of one call, freed on every return.
struct McBuffers {
    std::vector<void *> ptrs;
    ~McBuffers()
    {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    T *alloc(int64_t count)
    {
        void *p = nullptr;
        if (hipMalloc(&p, sizeof(T) * (size_t)std::max<int64_t>(count, 1)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (T *)p;
    }
};

}  // namespace
}  // namespace prh

using namespace prh;

extern "C" {

int prh_mc_properties(const prh_geometry *geo, int64_t n_rays, uint32_t seed, double epsilon, int32_t max_bounces,
                      int32_t restart_hit_point, double *event_len, double *censored_len, int64_t counters[2],
                      double sums[2])
{
    PRH_CHECK_ARG(geo && event_len && censored_len && counters && sums, "prh_mc_properties: null pointer");
    PRH_CHECK_ARG(n_rays > 0 && n_rays < (int64_t(1) << 31) && max_bounces >= 1, "prh_mc_properties: bad sizes");
    const Geometry &g = geo->g;
    PRH_HIP(hipSetDevice(g.device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    McBuffers B;
    double *ev = B.alloc<double>(n_rays), *ce = B.alloc<double>(n_rays), *m1 = B.alloc<double>(n_rays),
           *m2 = B.alloc<double>(n_rays);
    int64_t *in = B.alloc<int64_t>(n_rays), *sc = B.alloc<int64_t>(n_rays);
    const unsigned grid = (unsigned)((n_rays + MC_THREADS - 1) / MC_THREADS);
    double *part = B.alloc<double>(6 * (int64_t)grid), *tot = B.alloc<double>(6);
    if (!ev || !ce || !m1 || !m2 || !in || !sc || !part || !tot) {
        set_error("prh_mc_properties: out of device memory");
        return PRH_ERR_HIP;
    }
    const McView V = view_of(g);
    if (restart_hit_point)
        k_mc_properties<true><<<grid, MC_THREADS, 0, st>>>(V, n_rays, seed, epsilon, max_bounces, ev, ce, in, sc, m1, m2);
    else
        k_mc_properties<false><<<grid, MC_THREADS, 0, st>>>(V, n_rays, seed, epsilon, max_bounces, ev, ce, in, sc, m1,
                                                            m2);
    k_mc_block_sums<<<grid, MC_THREADS, 0, st>>>(in, sc, m1, m2, nullptr, n_rays, part);
    k_mc_totals<<<6, MC_THREADS, 0, st>>>(part, (int64_t)grid, tot);
    PRH_HIP(hipGetLastError());
    double host[6];
    PRH_HIP(hipMemcpyAsync(host, tot, sizeof(host), hipMemcpyDeviceToHost, st));
    PRH_HIP(hipMemcpyAsync(event_len, ev, sizeof(double) * n_rays, hipMemcpyDeviceToHost, st));
    PRH_HIP(hipMemcpyAsync(censored_len, ce, sizeof(double) * n_rays, hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));
    counters[0] = (int64_t)host[0];
    counters[1] = (int64_t)host[1];
    sums[0] = host[4];
    sums[1] = host[5];
    return PRH_OK;
}

int prh_mc_penetration(const prh_geometry *geo, int64_t n_rays, uint32_t seed, double epsilon, int32_t max_bounces,
                       int32_t restart_hit_point, int64_t counts[4])
{
    PRH_CHECK_ARG(geo && counts, "prh_mc_penetration: null pointer");
    PRH_CHECK_ARG(n_rays > 0 && n_rays < (int64_t(1) << 31) && max_bounces >= 1, "prh_mc_penetration: bad sizes");
    const Geometry &g = geo->g;
    PRH_HIP(hipSetDevice(g.device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    McBuffers B;
    uint8_t *fate = B.alloc<uint8_t>(n_rays);
    const unsigned grid = (unsigned)((n_rays + MC_THREADS - 1) / MC_THREADS);
    double *part = B.alloc<double>(6 * (int64_t)grid), *tot = B.alloc<double>(6);
    if (!fate || !part || !tot) {
        set_error("prh_mc_penetration: out of device memory");
        return PRH_ERR_HIP;
    }
    const McView V = view_of(g);
    if (restart_hit_point)
        k_mc_penetration<true><<<grid, MC_THREADS, 0, st>>>(V, n_rays, seed, epsilon, max_bounces, fate);
    else
        k_mc_penetration<false><<<grid, MC_THREADS, 0, st>>>(V, n_rays, seed, epsilon, max_bounces, fate);
    k_mc_block_sums<<<grid, MC_THREADS, 0, st>>>(nullptr, nullptr, nullptr, nullptr, fate, n_rays, part);
    k_mc_totals<<<4, MC_THREADS, 0, st>>>(part, (int64_t)grid, tot);
    PRH_HIP(hipGetLastError());
    double host[4];
    PRH_HIP(hipMemcpyAsync(host, tot, sizeof(host), hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));
    for (int q = 0; q < 4; ++q) counts[q] = (int64_t)host[q];
    return PRH_OK;
}

}  // extern "C"
