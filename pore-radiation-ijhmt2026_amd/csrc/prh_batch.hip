// Batched mixed-precision radiosity engine with continuous batching.
//
// Solves (I - diag(rho) F_p^T) J = E, rho = 1 - eps, for up to 64 emissions
// ("slots") at once on one device-resident exchange operator, and returns the
// absorbed power eps * (F_p^T J) -- the quantity the reference's reflection
// series sums reflection by reflection (porerad/exchange.py:217-245).
//
// Method.  F_p^T x = d * S(d x / A) is similar to a symmetric matrix of
// spectral radius 1 (F_p is row-stochastic and A F_p symmetric), so the
// eigenvalues of I - diag(rho) F_p^T lie in [1 - rho_max, 1 + rho_max].  A
// slot alternates epochs of m Chebyshev iterations on that interval
// (residual form) with one FP64 refinement step that adds the epoch's
// correction c to the FP64 solution x and recomputes the FP64 residual
// r = E - (x - rho * F_p^T x), until ||r||_1 <= rtol ||E||_1.  m is the
// Chebyshev count whose error bound covers what is left to reach the
// tolerance, at most a 1e-3 reduction per epoch.
//
// Every sweep ("tick") reads the operator once for all 64 slots.  The
// vector a tick gathers (z, the scaled search direction) is held in FP16:
// a sweep costs one random cache-line request per nonzero, and 64 FP16
// slots fill the 128 B line that 32 FP32 slots filled, so a tick serves
// twice the slots in about the same time.  The products are accumulated in
// FP64 from the FP16 inputs: K = C + C^T is stored exactly as uint16 when
// all patches share area and ray count (S = A/(2R) K + A delta I), else S in
// FP64.  Each tick takes the direction it actually applied, z / d, as the
// step of x and of the residual recursion (FP32), so rounding z to FP16
// changes the search direction but never makes x and r disagree; FP16
// limits what one epoch can resolve to about 1e-3, hence the epoch target.
// The refinement step of a slot is just another tick: the slot's gathered
// vector is its correction c, whose FP64 product F_p^T c is exact up to FP64
// rounding; it updates x += s c and y = F_p^T x += s F_p^T c and then starts
// the next epoch in the same pass.  Each slot runs its own schedule
// (Chebyshev tick, last tick, refinement tick, idle) independently of the
// others, and a slot may be admitted or retired at any tick.  Slot vectors
// are interleaved [row][slot] so one 128 B line serves all 64 slots;
// workgroups are mapped so that each XCD sweeps a contiguous range of rows.
//
// Host scheduling is pipelined: the host launches tick t and only then
// waits for tick t-1, so the device always has the next tick queued.  A
// slot's decisions (epoch lengths, scales, convergence) depend only on its
// own residual norms, read with a fixed lag of one tick; its arithmetic
// never mixes with another slot's and every reduction has a fixed order, so
// a solve returns the same bits alone or in any mix.
#include <algorithm>
#include <type_traits>
#include <cmath>
#include <vector>

#include "prh_internal.h"
#include "prh_operator_impl.h"

namespace prh {
namespace {

constexpr int SLOTS = 64;
constexpr int PER = 8;                 // FP16 slots per 16 B gather: lane h serves slots 8h .. 8h+7
constexpr double ENGINE_REDUCTION = 1.0e-3;  // finest reduction one epoch aims for (FP16 directions)
constexpr float Z_MAX = 65504.f;       // largest finite FP16
constexpr int BT = 256;
constexpr int ROWS = BT / 64;
constexpr int RING = 8;
constexpr double SIGMA = 5.670374419e-8;  // porerad/constants.py:3

enum : int32_t { MODE_IDLE = 0, MODE_CHEB = 1, MODE_LAST = 2, MODE_REFINE = 3 };

// Per-slot instructions for one tick.
struct SlotParam {
    int32_t mode;
    float alpha, beta;  // CHEB: dir' = alpha dir + beta r'
    int32_t pad;
    double s;           // REFINE: scale of the epoch that ends
    double inv_s_next;  // REFINE: 1 / scale of the epoch that starts
};

struct MatView {
    int64_t n;
    const int64_t *rowptr;
    const int32_t *col;
    const uint16_t *kval;
    const double *val64;
    const float *d32;
    const double *d64;
    const double *area;
    double inv_2R, delta;
    // gathered form of a slot vector v: z = d v (equal areas) or d v a_ref / A
    double aref, inv_aref;
    float aref32;
    const int32_t *row_order;  // sweep position -> row, or null for the stored order
    uint32_t blocks_per_xcd;
};

struct SlotVecs {
    float *x32, *r32, *rho32;
    double *x64, *y64, *r64, *E64, *eps64;
};

// Workgroups are dealt round-robin to the 8 XCDs; give XCD x the logical
// blocks [x * blocks_per_xcd, (x+1) * blocks_per_xcd) so that its L2 serves a
// contiguous (spatially coherent) range of patches.
__device__ __forceinline__ int64_t first_row(const MatView &M)
{
    const uint32_t b = blockIdx.x;
    return ((int64_t)(b % 8u) * M.blocks_per_xcd + (b / 8u)) * ROWS;
}

// The gathered form of a slot vector v in FP16: d v (equal areas) or
// d v a_ref / A, saturated to the finite FP16 range (a saturated entry only
// changes the search direction: every tick steps by what it gathered).
__device__ __forceinline__ _Float16 scaled(float di, float v, float ai, const MatView &M, bool uniform)
{
    const float z = uniform ? di * v : di * v * M.aref32 / ai;
    return (_Float16)fminf(fmaxf(z, -Z_MAX), Z_MAX);
}

// The slot vector a gathered entry stands for: z / d or z A / (d a_ref).
__device__ __forceinline__ float unscaled(float di, float z, float ai, const MatView &M, bool uniform)
{
    return uniform ? z / di : z * ai / (di * M.aref32);
}

struct alignas(16) Z8 {
    _Float16 v[PER];
};

// Slots 8h .. 8h+7 of row j of an interleaved slot vector.  The byte offset
// j * 128 + 16 h stays below 2^32 (at most 2^25 patches), so the load takes a
// scalar base and one 32-bit offset register.
__device__ __forceinline__ Z8 gather8(const _Float16 *z, int32_t j, int h)
{
    const uint32_t off = (uint32_t)j * (uint32_t)(SLOTS * sizeof(_Float16)) + 16u * (uint32_t)h;
    return *reinterpret_cast<const Z8 *>(reinterpret_cast<const char *>(z) + off);
}

enum : int32_t { FREE = 0, RUNNING = 1, FINISHING = 2, DONE = 3 };

// Schedule of one slot, kept on the device and moved on by k_schedule before
// every tick, so that ticks can be replayed from graphs without the host.
struct DevSlot {
    int32_t state;     // RUNNING until the device finds it converged (FINISHING)
    int32_t refining;  // the next tick is the slot's refinement
    int32_t refined;   // the tick just run was its refinement: its norm is in the totals
    int32_t k, m;      // tick index within the epoch, epoch length
    int32_t epochs;    // refinements made
    int32_t ticks;     // Chebyshev ticks run
    int32_t failed;    // stopped at MAX_EPOCHS without converging
    double rtol, rho_max, e_norm, s, tau;
};

// What the next tick runs, written by k_schedule.
struct TickCtl {
    uint32_t group_mask;  // bit h: one of the slots 8h .. 8h+7 runs
    int32_t any_refine;   // a slot refines (the tick writes norm partials)
    int32_t running;      // slots that run
    int32_t pad;
    unsigned long long ticks, slot_ticks;  // ticks that ran a slot, and their running slots summed
};

struct TickArgs {
    MatView M;
    SlotVecs V;
    const _Float16 *z_in;
    _Float16 *z_out;
    const SlotParam *param;  // [SLOTS]
    double *part_abs;        // [SLOTS][n_blocks], written when a slot refines
    int64_t n_blocks;
    const TickCtl *ctl;
};

// One tick: y = F_p^T z for every slot (FP64 accumulation of FP16 inputs),
// then each slot's own update (Chebyshev step, last step, or refinement).
// Eight waves per SIMD (64 VGPRs, no spills) keep the most gathers in flight.
template <bool UNIFORM>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_tick(TickArgs a)
{
    __shared__ double s_abs[ROWS][SLOTS];
    const MatView &M = a.M;
    const uint32_t group_mask = a.ctl->group_mask;
    const int wave = threadIdx.x >> 6;
    // one row per wavefront: the row (and its CSR range) is wave-uniform
    // the wave's sweep position, and its row (row_order keeps the rows an XCD
    // sweeps together close in space, so their gathers share L2 lines)
    const int64_t pos = first_row(M) + wave;
    const int64_t row = (int64_t)__builtin_amdgcn_readfirstlane(
        (int)(pos < M.n && M.row_order ? (int64_t)M.row_order[pos] : pos));
    const int lane = threadIdx.x & 63;
    const int h = lane & 7;   // gather: slots 8h .. 8h+7
    const int g = lane >> 3;  // nonzero group
    double acc[PER];
#pragma unroll
    for (int p = 0; p < PER; ++p) acc[p] = 0.0;
    if (row < M.n && group_mask) {
        // A tick makes one random cache-line request per nonzero: the 64 slots
        // of a row are one 128 B line, fetched by the 8 lanes of a group, 16 B
        // (slots 8h .. 8h+7) each.  Every lane loads a distinct column and
        // value (64 nonzeros per load), the lanes exchange them through the LDS
        // crossbar, and each lane then has 8 gathers in flight.  Entries past
        // the row's end repeat its last entry with weight 0, and idle slot
        // groups are gathered too (their sums are never used), so the loop is
        // branch-free.
        const int64_t e0 = M.rowptr[row], e1 = M.rowptr[row + 1];
        const uint32_t len = (uint32_t)(e1 - e0);  // at most one entry per patch: < 2^25
        const int32_t *col = M.col + e0;
        const uint16_t *kval = UNIFORM ? M.kval + e0 : nullptr;
        const double *val = UNIFORM ? nullptr : M.val64 + e0;
        for (uint32_t base = 0; base < len; base += 64u) {
            const uint32_t e = base + (uint32_t)lane;
            const bool ok = e < len;
            const uint32_t ec = ok ? e : len - 1;
            const int32_t jl = col[ec];
            const double w = UNIFORM ? (double)kval[ec] : val[ec];
            const double vl = ok ? w : 0.0;
            int32_t j[8];
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                j[u] = __shfl(jl, g + 8 * u);
                v[u] = __shfl(vl, g + 8 * u);
            }
            Z8 z[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) z[u] = gather8(a.z_in, j[u], h);
#pragma unroll
            for (int u = 0; u < 8; ++u)
#pragma unroll
                for (int p = 0; p < PER; ++p) acc[p] = fma(v[u], (double)(float)z[u].v[p], acc[p]);
        }
    }
#pragma unroll
    for (int off = 8; off < 64; off <<= 1)
#pragma unroll
        for (int p = 0; p < PER; ++p) acc[p] += __shfl_xor(acc[p], off);
    // Every lane now holds the sums of its 8 slots; lane (h, g) updates slot 8h + g.
    double wsum = acc[0];
#pragma unroll
    for (int p = 1; p < PER; ++p) wsum = g == p ? acc[p] : wsum;
    double r_abs = 0.0;
    const int q = PER * h + g;
    if (row < M.n) {
        const SlotParam P = a.param[q];
        if (P.mode != MODE_IDLE) {
            const SlotVecs &V = a.V;
            const size_t at = (size_t)row * SLOTS + q;
            const float zi = (float)a.z_in[at];
            const double di = M.d64[row];
            const double y = UNIFORM ? di * (wsum * M.inv_2R + M.delta * (double)zi) : di * wsum * M.inv_aref;
            const float di32 = M.d32[row];
            const float ai32 = UNIFORM ? 1.f : (float)M.area[row];
            if (P.mode == MODE_REFINE) {
                const double ai = UNIFORM ? 1.0 : M.area[row];
                const double c = UNIFORM ? (double)zi / di : (double)zi * ai / (di * M.aref);
                const double x = V.x64[at] + P.s * c;
                const double yy = V.y64[at] + P.s * y;
                const double r = V.E64[at] - (x - (1.0 - V.eps64[at]) * yy);
                V.x64[at] = x;
                V.y64[at] = yy;
                V.r64[at] = r;
                r_abs = fabs(r);
                // next epoch: x32 = 0, r32 = dir = r / s
                const float rn = (float)(r * P.inv_s_next);
                V.x32[at] = 0.f;
                V.r32[at] = rn;
                a.z_out[at] = scaled(di32, rn, ai32, M, UNIFORM);
            } else {
                // the direction this tick applied, as gathered
                const float dk = unscaled(di32, zi, ai32, M, UNIFORM);
                const float Ad = dk - V.rho32[at] * (float)y;
                const float rn = V.r32[at] - Ad;
                const float x = V.x32[at] + dk;
                V.x32[at] = x;
                if (P.mode == MODE_CHEB) {
                    const float dn = P.alpha * dk + P.beta * rn;
                    V.r32[at] = rn;
                    a.z_out[at] = scaled(di32, dn, ai32, M, UNIFORM);
                } else {  // MODE_LAST: the epoch's correction is gathered next
                    a.z_out[at] = scaled(di32, x, ai32, M, UNIFORM);
                }
            }
        }
    }
    if (!a.ctl->any_refine) return;  // the same for every thread of the grid
    s_abs[wave][q] = r_abs;
    __syncthreads();
    if (threadIdx.x < SLOTS) {
        double sa = 0.0;
        for (int w = 0; w < ROWS; ++w) sa += s_abs[w][threadIdx.x];
        a.part_abs[(size_t)threadIdx.x * a.n_blocks + blockIdx.x] = sa;
    }
}

// Before every tick, one thread per slot: take in the norm of a refinement
// the last tick made (stop when ||r||_1 <= rtol ||E||_1, else size the next
// epoch), then write the slot's instructions for the next tick.  A slot's
// decisions read only its own state and norms.
__global__ __launch_bounds__(SLOTS) void k_schedule(DevSlot *S, const double *tot, SlotParam *param, TickCtl *ctl)
{
    __shared__ uint32_t gm;
    __shared__ int32_t ar, run;
    if (threadIdx.x == 0) {
        gm = 0u;
        ar = 0;
        run = 0;
    }
    __syncthreads();
    const int q = threadIdx.x;
    DevSlot st = S[q];
    SlotParam p{MODE_IDLE, 0.f, 1.f, 0, 0.0, 0.0};
    if (st.state == RUNNING && st.refined) {
        const double r_norm = tot[q];
        st.refined = 0;
        st.epochs += 1;
        const bool converged = r_norm <= st.rtol * st.e_norm;
        if (converged || st.epochs >= MAX_EPOCHS) {
            st.failed = converged ? 0 : 1;
            st.state = FINISHING;
        } else {
            st.tau = epoch_target(st.rtol, st.e_norm, r_norm, ENGINE_REDUCTION);
            st.m = ticks_for(st.rho_max, st.tau);
        }
    }
    if (st.state == RUNNING) {
        if (st.refining) {
            // the epoch's correction is gathered now: refine, then the next epoch
            const double s_next = st.s * st.tau;
            p.mode = MODE_REFINE;
            p.s = st.s;
            st.s = s_next > 0.0 ? s_next : 1.0;
            p.inv_s_next = 1.0 / st.s;
            st.refining = 0;
            st.refined = 1;
            st.k = 0;
            atomicOr(&ar, 1);
        } else {
            const bool last = st.m > 0 && st.k >= st.m - 1;
            p.mode = last ? MODE_LAST : MODE_CHEB;
            if (!last) cheb_coefficient(st.rho_max, st.k, &p.alpha, &p.beta);
            st.k += 1;
            st.ticks += 1;
            st.refining = last ? 1 : 0;
        }
        atomicOr(&gm, 1u << (q / PER));
        atomicAdd(&run, 1);
    }
    param[q] = p;
    S[q] = st;
    __syncthreads();
    if (q == 0) {
        ctl->group_mask = gm;
        ctl->any_refine = ar;
        ctl->running = run;
        if (run) {
            ctl->ticks += 1;
            ctl->slot_ticks += (unsigned long long)run;
        }
    }
}

// Fixed-order totals of columns of partial sums: out[c] = sum_b part[c][b].
__global__ __launch_bounds__(256) void k_column_totals(const double *part, int64_t n_blocks, double *out)
{
    __shared__ double sh[256];
    const int c = blockIdx.x;
    double a = 0.0;
    for (int64_t b = threadIdx.x; b < n_blocks; b += 256) a += part[(size_t)c * n_blocks + b];
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = sh[0];
}

// The refinement norms of a tick (column totals of its partials), when a slot refined.
__global__ __launch_bounds__(256) void k_refine_totals(const double *part, int64_t n_blocks, const TickCtl *ctl,
                                                       double *out)
{
    if (!ctl->any_refine) return;
    __shared__ double sh[256];
    const int c = blockIdx.x;
    double a = 0.0;
    for (int64_t b = threadIdx.x; b < n_blocks; b += 256) a += part[(size_t)c * n_blocks + b];
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = sh[0];
}

// Sum of v over the 256 threads of the block, written to *part_at.
__device__ __forceinline__ void block_sum_to(double v, double *sh, double *part_at)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *part_at = sh[0];
    __syncthreads();
}

// Admission state of slot q in row i: x = y = 0, r = E, and the first
// epoch's FP32 state r32 = dir = E / s with its gathered vector.
__device__ __forceinline__ void admit_row(const MatView &M, const SlotVecs &V, _Float16 *z_next, int64_t i, int q,
                                          double E, double e, double inv_s, bool uniform)
{
    const size_t at = (size_t)i * SLOTS + q;
    V.E64[at] = E;
    V.eps64[at] = e;
    V.rho32[at] = (float)(1.0 - e);
    V.x64[at] = 0.0;
    V.y64[at] = 0.0;
    V.r64[at] = E;
    const float r = (float)(E * inv_s);
    V.x32[at] = 0.f;
    V.r32[at] = r;
    z_next[at] = scaled(M.d32[i], r, uniform ? 1.f : (float)M.area[i], M, uniform);
}

__global__ __launch_bounds__(256) void k_admit(MatView M, SlotVecs V, _Float16 *z_next, int q, const double *E,
                                               const double *eps, int eps_per_patch, double inv_s, int uniform)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M.n) return;
    admit_row(M, V, z_next, i, q, E[i], eps_per_patch ? eps[i] : eps[0], inv_s, uniform != 0);
}

struct FieldView {
    const int8_t *cls;      // 0 hot wall, 1 cold wall, 2 internal
    const int32_t *cell_x;  // x index of the owning cell (internal patches)
    double nx;
};

// Emission of the prescribed linear field between T_hot and T_cold with
// uniform emissivity: walls at T_m +/- dT/2, internal patches at
// (T_m + dT/2) - (i + 1/2) dT / nx, minus the isothermal state at T_m
// (porerad/coupled.py:138-143, 180-184; porerad/exchange.py:265-270).
// Also the block sums of |E|.
__global__ __launch_bounds__(256) void k_admit_linear(MatView M, SlotVecs V, FieldView F, _Float16 *z_next, int q,
                                                      double eps, double eps_sigma, double T_m, double dT,
                                                      double inv_s, int uniform, double *part)
{
    __shared__ double sh[256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double E = 0.0;
    if (i < M.n) {
        const int c = F.cls[i];
        double T;
        if (c == 0) T = T_m + dT / 2.0;
        else if (c == 1) T = T_m - dT / 2.0;
        else T = (T_m + dT / 2.0) - ((double)F.cell_x[i] + 0.5) * dT / F.nx;
        E = eps_sigma * M.area[i] * ((T - T_m) * (T + T_m) * (T * T + T_m * T_m));
        admit_row(M, V, z_next, i, q, E, eps, inv_s, uniform != 0);
    }
    block_sum_to(fabs(E), sh, part + blockIdx.x);
}

// Results of a finished slot, as block sums: per class sums of
// net = eps y - E (hot, cold, internal; zero without a field) and the power
// still in flight sum(r).
__global__ __launch_bounds__(256) void k_finish(int64_t n, int q, FieldView F, SlotVecs V, int64_t n_part,
                                                double *part)
{
    __shared__ double sh[256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    if (i < n) {
        const size_t at = (size_t)i * SLOTS + q;
        const double net = V.eps64[at] * V.y64[at] - V.E64[at];
        if (F.cls) {
            const int c = F.cls[i];
            s0 = c == 0 ? net : 0.0;
            s1 = c == 1 ? net : 0.0;
            s2 = c == 2 ? net : 0.0;
        }
        s3 = V.r64[at];
    }
    block_sum_to(s0, sh, part + blockIdx.x);
    block_sum_to(s1, sh, part + n_part + blockIdx.x);
    block_sum_to(s2, sh, part + 2 * n_part + blockIdx.x);
    block_sum_to(s3, sh, part + 3 * n_part + blockIdx.x);
}

// Absorbed power eps * y of slot q into a packed vector.
__global__ void k_collect(int64_t n, int q, const double *eps64, const double *y64, double *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t at = (size_t)i * SLOTS + q;
    out[i] = eps64[at] * y64[at];
}

unsigned grid_of(int64_t n, int b) { return (unsigned)std::max<int64_t>(1, (n + b - 1) / b); }

// Host view of a slot.  The host owns admission and retirement (FREE ->
// RUNNING, FINISHING -> DONE -> FREE); the device moves RUNNING to FINISHING.
struct SlotHost {
    int32_t state = FREE;
    bool field = false;
    bool failed = false;
    int32_t iterations = 0;
    int64_t admit_launch = 0;   // first launch that runs the slot
    bool finish_queued = false; // k_finish queued, slot_ev recorded after it
};

struct Pinned {
    DevSlot init[SLOTS];          // admission states, copied to the device
    DevSlot mirror[RING][SLOTS];  // device states after a launch
    TickCtl ctl[RING];
    double fin[SLOTS][4];
};

constexpr int GRAPH_TICKS = 8;  // ticks of a long replay (even: keeps the z-buffer parity)

}  // namespace
}  // namespace prh

using namespace prh;

struct prh_batch {
    prh_operator *op = nullptr;
    int device = 0;
    int64_t n = 0, n_blocks = 0;
    uint32_t blocks_per_xcd = 0;
    float *x32 = nullptr, *r32 = nullptr, *rho32 = nullptr;
    _Float16 *zbuf[2] = {nullptr, nullptr};
    double *x64 = nullptr, *y64 = nullptr, *r64 = nullptr, *E64 = nullptr, *eps64 = nullptr;
    double *part = nullptr, *tot = nullptr, *stage = nullptr;
    SlotParam *param_dev = nullptr;
    DevSlot *dslots = nullptr;
    TickCtl *ctl = nullptr;
    int8_t *cls = nullptr;
    int32_t *cell_x = nullptr;
    int32_t field_nx = 0;
    double cross_section = 0.0;
    double mean_area = 0.0, max_area = 0.0;  // a_ref of the gathered form (unequal areas), largest area
    Pinned *host = nullptr;
    // replays: [long][parity] = GRAPH_TICKS or 1 tick starting on z buffer `parity`
    hipGraph_t graph[2][2] = {};
    hipGraphExec_t exec[2][2] = {};
    const int32_t *graph_order = nullptr;  // row order the graphs were captured with
    bool graphs_ready = false;
    // launches (replays): completion events, their tick counts and device time
    hipEvent_t ev[RING] = {};
    hipEvent_t tick_ev[RING][2] = {};
    int launch_ticks[RING] = {};
    int64_t launches = 0;  // replays launched
    int64_t read = 0;      // replays whose results have been read
    hipEvent_t slot_ev[SLOTS] = {};
    int64_t timed_ticks = 0;
    unsigned long long slot_ticks = 0;
    double tick_ms = 0.0;
    int64_t ticks = 0;  // ticks launched (their parity selects the z buffer)
    SlotHost slot[SLOTS];
    int64_t bytes = 0;
};

namespace {

MatView view_of(const prh_batch *b)
{
    const prh_operator *op = b->op;
    MatView M;
    M.n = op->n;
    M.rowptr = op->rowptr;
    M.col = op->col;
    M.kval = op->kval;
    M.val64 = op->val;
    M.d32 = op->scale32;
    M.d64 = op->scale;
    M.area = op->area;
    M.inv_2R = op->inv_2R;
    M.delta = op->delta;
    M.aref = op->uniform ? 1.0 : b->mean_area;
    M.inv_aref = 1.0 / M.aref;
    M.aref32 = (float)M.aref;
    M.row_order = op->row_order;
    M.blocks_per_xcd = b->blocks_per_xcd;
    return M;
}

SlotVecs vecs_of(const prh_batch *b)
{
    return SlotVecs{b->x32, b->r32, b->rho32, b->x64, b->y64, b->r64, b->E64, b->eps64};
}

FieldView field_of(const prh_batch *b) { return FieldView{b->cls, b->cell_x, (double)b->field_nx}; }

void free_graphs(prh_batch *b)
{
    for (int l = 0; l < 2; ++l)
        for (int p = 0; p < 2; ++p) {
            if (b->exec[l][p]) (void)hipGraphExecDestroy(b->exec[l][p]);
            if (b->graph[l][p]) (void)hipGraphDestroy(b->graph[l][p]);
            b->exec[l][p] = nullptr;
            b->graph[l][p] = nullptr;
        }
    b->graphs_ready = false;
}

void free_batch(prh_batch *b)
{
    free_graphs(b);
    void *ptrs[] = {b->x32,    b->r32,   b->rho32,  b->zbuf[0], b->zbuf[1], b->x64,
                    b->y64,    b->r64,   b->E64,    b->eps64,   b->part,    b->tot,
                    b->stage,  b->cls,   b->cell_x, b->param_dev, b->dslots, b->ctl};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (b->host) (void)hipHostFree(b->host);
    for (int k = 0; k < RING; ++k) {
        if (b->ev[k]) (void)hipEventDestroy(b->ev[k]);
        for (hipEvent_t e : b->tick_ev[k])
            if (e) (void)hipEventDestroy(e);
    }
    for (hipEvent_t e : b->slot_ev)
        if (e) (void)hipEventDestroy(e);
}

// The z buffer the next tick gathers.
_Float16 *z_next(prh_batch *b) { return b->zbuf[b->ticks & 1]; }

// Enqueue `count` ticks whose first gathers z buffer `parity`: per tick the
// schedule, the sweep and the refinement norms.
int enqueue_ticks(prh_batch *b, int count, int parity, hipStream_t st)
{
    TickArgs a;
    a.M = view_of(b);
    a.V = vecs_of(b);
    a.param = b->param_dev;
    a.part_abs = b->part;
    a.n_blocks = b->n_blocks;
    a.ctl = b->ctl;
    for (int t = 0; t < count; ++t) {
        a.z_in = b->zbuf[(parity + t) & 1];
        a.z_out = b->zbuf[(parity + t + 1) & 1];
        k_schedule<<<1, SLOTS, 0, st>>>(b->dslots, b->tot, b->param_dev, b->ctl);
        if (b->op->uniform) k_tick<true><<<(unsigned)b->n_blocks, BT, 0, st>>>(a);
        else k_tick<false><<<(unsigned)b->n_blocks, BT, 0, st>>>(a);
        k_refine_totals<<<SLOTS, 256, 0, st>>>(b->part, b->n_blocks, b->ctl, b->tot);
    }
    PRH_HIP(hipGetLastError());
    return PRH_OK;
}

// Capture the four replays (GRAPH_TICKS or 1 tick, from either z buffer).
int capture_graphs(prh_batch *b, hipStream_t st)
{
    free_graphs(b);
    PRH_HIP(hipStreamSynchronize(st));
    for (int l = 0; l < 2; ++l)
        for (int p = 0; p < 2; ++p) {
            if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                set_error("prh_batch: stream capture did not start");
                return PRH_ERR_HIP;
            }
            const int rc = enqueue_ticks(b, l ? GRAPH_TICKS : 1, p, st);
            hipGraph_t g = nullptr;
            const hipError_t end = hipStreamEndCapture(st, &g);
            b->graph[l][p] = g;
            if (rc) return rc;
            if (end != hipSuccess || !g) {
                set_error("prh_batch: graph capture failed: %s", hipGetErrorString(end));
                return PRH_ERR_HIP;
            }
            if (hipGraphInstantiate(&b->exec[l][p], g, nullptr, nullptr, 0) != hipSuccess) {
                set_error("prh_batch: graph instantiation failed");
                return PRH_ERR_HIP;
            }
        }
    b->graph_order = b->op->row_order;
    b->graphs_ready = true;
    return PRH_OK;
}

// Replay `count` (GRAPH_TICKS or 1) ticks, then copy the slot states back.
int launch(prh_batch *b, int count, hipStream_t st)
{
    if (!b->graphs_ready || b->graph_order != b->op->row_order) {
        int rc = capture_graphs(b, st);
        if (rc) return rc;
    }
    const int r = (int)(b->launches % RING);
    PRH_HIP(hipEventRecord(b->tick_ev[r][0], st));
    PRH_HIP(hipGraphLaunch(b->exec[count == GRAPH_TICKS ? 1 : 0][b->ticks & 1], st));
    PRH_HIP(hipEventRecord(b->tick_ev[r][1], st));
    PRH_HIP(hipMemcpyAsync(b->host->mirror[r], b->dslots, sizeof(DevSlot) * SLOTS, hipMemcpyDeviceToHost, st));
    PRH_HIP(hipMemcpyAsync(&b->host->ctl[r], b->ctl, sizeof(TickCtl), hipMemcpyDeviceToHost, st));
    PRH_HIP(hipEventRecord(b->ev[r], st));
    b->launch_ticks[r] = count;
    b->ticks += count;
    b->launches += 1;
    return PRH_OK;
}

int queue_finish(prh_batch *b, int q, hipStream_t st)
{
    const int64_t n = b->n;
    const unsigned grid = grid_of(n, 256);
    // tot[0, SLOTS) holds refinement norms a later schedule still reads: the sums go above
    k_finish<<<grid, 256, 0, st>>>(n, q, field_of(b), vecs_of(b), (int64_t)grid, b->part);
    k_column_totals<<<4, 256, 0, st>>>(b->part, (int64_t)grid, b->tot + SLOTS);
    PRH_HIP(hipGetLastError());
    PRH_HIP(hipMemcpyAsync(b->host->fin[q], b->tot + SLOTS, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
    PRH_HIP(hipEventRecord(b->slot_ev[q], st));
    b->slot[q].finish_queued = true;
    return PRH_OK;
}

// Read the results of launch L (in order): device time, and the slots the
// device found converged, whose results are then summed.
int read_launch(prh_batch *b, int64_t L, hipStream_t st)
{
    const int r = (int)(L % RING);
    PRH_HIP(hipEventSynchronize(b->ev[r]));
    float ms = 0.f;
    PRH_HIP(hipEventElapsedTime(&ms, b->tick_ev[r][0], b->tick_ev[r][1]));
    b->tick_ms += ms;
    b->timed_ticks += b->launch_ticks[r];
    b->slot_ticks = b->host->ctl[r].slot_ticks;
    for (int q = 0; q < SLOTS; ++q) {
        SlotHost &S = b->slot[q];
        if (S.state != RUNNING || S.admit_launch > L) continue;
        const DevSlot &D = b->host->mirror[r][q];
        if (D.state != FINISHING) continue;
        S.state = FINISHING;
        S.failed = D.failed != 0;
        S.iterations = D.ticks + D.epochs;
        int rc = queue_finish(b, q, st);
        if (rc) return rc;
    }
    b->read = L + 1;
    return PRH_OK;
}

// Finished slots whose sums have arrived become DONE (waiting for them if `wait`).
int settle(prh_batch *b, bool wait)
{
    for (int q = 0; q < SLOTS; ++q) {
        SlotHost &S = b->slot[q];
        if (S.state != FINISHING || !S.finish_queued) continue;
        if (wait) PRH_HIP(hipEventSynchronize(b->slot_ev[q]));
        const hipError_t e = hipEventQuery(b->slot_ev[q]);
        if (e == hipSuccess) {
            S.state = DONE;
        } else if (e != hipErrorNotReady) {
            set_error("prh_batch: %s", hipGetErrorString(e));
            return PRH_ERR_HIP;
        }
    }
    return PRH_OK;
}

// Wait for everything launched so far and read it.
int batch_drain(prh_batch *b, hipStream_t st)
{
    for (int64_t L = b->read; L < b->launches; ++L) {
        int rc = read_launch(b, L, st);
        if (rc) return rc;
    }
    return settle(b, true);
}

int first_free(const prh_batch *b)
{
    for (int k = 0; k < SLOTS; ++k)
        if (b->slot[k].state == FREE) return k;
    return -1;
}

void report(const prh_batch *b, uint64_t *done_mask, uint64_t *active_mask)
{
    uint64_t done = 0, active = 0;
    for (int q = 0; q < SLOTS; ++q) {
        if (b->slot[q].state != FREE) active |= uint64_t(1) << q;
        if (b->slot[q].state == DONE) done |= uint64_t(1) << q;
    }
    if (done_mask) *done_mask = done;
    if (active_mask) *active_mask = active;
}

// Admission state of a slot: RUNNING from the next launch on.
void admit_slot(prh_batch *b, int q, double rtol, double rho_max, double e_norm, double s0, bool field,
                hipStream_t st, int *rc)
{
    DevSlot &D = b->host->init[q];
    D = DevSlot{};
    D.state = RUNNING;
    D.rtol = rtol;
    D.rho_max = rho_max;
    D.e_norm = e_norm;
    D.s = s0;
    D.tau = epoch_target(rtol, 1.0, 1.0, ENGINE_REDUCTION);
    D.m = ticks_for(rho_max, D.tau);
    *rc = PRH_OK;
    if (hipMemcpyAsync(&b->dslots[q], &D, sizeof(DevSlot), hipMemcpyHostToDevice, st) != hipSuccess) {
        set_error("prh_batch: slot admission copy failed");
        *rc = PRH_ERR_HIP;
        return;
    }
    SlotHost &S = b->slot[q];
    S = SlotHost();
    S.state = RUNNING;
    S.field = field;
    S.admit_launch = b->launches;
}

}  // namespace

extern "C" {

int prh_batch_create(prh_operator *op, prh_batch **out)
{
    PRH_CHECK_ARG(op && out, "prh_batch_create: null pointer");
    *out = nullptr;
    if (op->n > (int64_t(1) << 25)) {
        set_error("prh_batch_create: %lld patches exceed the engine's 2^25-patch capacity", (long long)op->n);
        return PRH_ERR_CAPACITY;
    }
    PRH_HIP(hipSetDevice(op->device));
    prh_batch *b = new prh_batch();
    b->op = op;
    b->device = op->device;
    b->n = op->n;
    const int64_t nb_logical = (op->n + ROWS - 1) / ROWS;
    b->blocks_per_xcd = (uint32_t)((nb_logical + 7) / 8);
    b->n_blocks = (int64_t)b->blocks_per_xcd * 8;
    const size_t ns = (size_t)op->n * SLOTS;
    const int64_t row_blocks = grid_of(op->n, 256);
    bool ok = true;
    auto alloc = [&](void **p, size_t bytes) {
        if (ok && hipMalloc(p, bytes) != hipSuccess) ok = false;
        b->bytes += (int64_t)bytes;
    };
    alloc((void **)&b->x32, ns * 4);
    alloc((void **)&b->r32, ns * 4);
    alloc((void **)&b->rho32, ns * 4);
    alloc((void **)&b->zbuf[0], ns * 2);
    alloc((void **)&b->zbuf[1], ns * 2);
    alloc((void **)&b->x64, ns * 8);
    alloc((void **)&b->y64, ns * 8);
    alloc((void **)&b->r64, ns * 8);
    alloc((void **)&b->E64, ns * 8);
    alloc((void **)&b->eps64, ns * 8);
    // partial sums: refinement norms [SLOTS][n_blocks], and the row-block
    // sums of admissions (|E|, tail of the array) and finishes ([4][row_blocks])
    alloc((void **)&b->part, (size_t)std::max<int64_t>(SLOTS * b->n_blocks, 4 * row_blocks) * 8);
    alloc((void **)&b->tot, 2 * SLOTS * 8);
    alloc((void **)&b->stage, (size_t)op->n * 2 * 8);
    alloc((void **)&b->param_dev, SLOTS * sizeof(SlotParam));
    alloc((void **)&b->dslots, SLOTS * sizeof(DevSlot));
    alloc((void **)&b->ctl, sizeof(TickCtl));
    if (ok && hipHostMalloc((void **)&b->host, sizeof(Pinned), hipHostMallocDefault) != hipSuccess) ok = false;
    for (int k = 0; ok && k < RING; ++k) {
        if (hipEventCreateWithFlags(&b->ev[k], hipEventDisableTiming) != hipSuccess) ok = false;
        for (hipEvent_t &e : b->tick_ev[k])
            if (ok && hipEventCreate(&e) != hipSuccess) ok = false;
    }
    for (int q = 0; ok && q < SLOTS; ++q)
        if (hipEventCreateWithFlags(&b->slot_ev[q], hipEventDisableTiming) != hipSuccess) ok = false;
    if (!ok) {
        free_batch(b);
        delete b;
        set_error("prh_batch_create: out of device memory");
        return PRH_ERR_HIP;
    }
    // Zero the slot state so idle slots carry finite values through every sweep.
    for (void *p : {(void *)b->x32, (void *)b->r32, (void *)b->rho32}) PRH_HIP(hipMemset(p, 0, ns * 4));
    for (void *p : {(void *)b->zbuf[0], (void *)b->zbuf[1]}) PRH_HIP(hipMemset(p, 0, ns * 2));
    PRH_HIP(hipMemset(b->dslots, 0, SLOTS * sizeof(DevSlot)));  // every slot FREE
    PRH_HIP(hipMemset(b->ctl, 0, sizeof(TickCtl)));
    PRH_HIP(hipMemset(b->tot, 0, 2 * SLOTS * 8));
    // a_ref of the gathered form (mean area) and the largest area
    {
        std::vector<double> area(op->n);
        PRH_HIP(hipMemcpy(area.data(), op->area, op->n * 8, hipMemcpyDeviceToHost));
        double sum = 0.0, largest = 0.0;
        for (double a : area) {
            sum += a;
            largest = std::max(largest, a);
        }
        b->mean_area = op->n ? sum / (double)op->n : 1.0;
        b->max_area = largest;
        if (!(b->mean_area > 0.0)) b->mean_area = 1.0;
    }
    *out = b;
    return PRH_OK;
}

int prh_batch_bytes(const prh_batch *b, int64_t *bytes)
{
    PRH_CHECK_ARG(b && bytes, "prh_batch_bytes: null pointer");
    *bytes = b->bytes;
    return PRH_OK;
}

int prh_batch_stats(const prh_batch *b, double *out, int32_t n)
{
    PRH_CHECK_ARG(b && out && n >= 0, "prh_batch_stats: null pointer");
    const double v[] = {(double)b->timed_ticks, b->tick_ms, (double)b->slot_ticks, (double)b->op->nnz_s,
                        (double)b->n, b->op->uniform ? 2.0 : 8.0};
    for (int i = 0; i < n && i < (int)(sizeof(v) / sizeof(v[0])); ++i) out[i] = v[i];
    return PRH_OK;
}

int prh_batch_free_slots(const prh_batch *b, uint64_t *free_mask)
{
    PRH_CHECK_ARG(b && free_mask, "prh_batch_free_slots: null pointer");
    uint64_t m = 0;
    for (int q = 0; q < SLOTS; ++q)
        if (b->slot[q].state == FREE) m |= uint64_t(1) << q;
    *free_mask = m;
    return PRH_OK;
}

int prh_batch_set_field(prh_batch *b, const int8_t *patch_class, const int32_t *cell_x, int32_t nx,
                        double cross_section_m2)
{
    PRH_CHECK_ARG(b && patch_class && cell_x, "prh_batch_set_field: null pointer");
    PRH_CHECK_ARG(nx > 0 && cross_section_m2 > 0.0, "prh_batch_set_field: nx and cross section must be positive");
    const int64_t n = b->n;
    for (int64_t i = 0; i < n; ++i) {
        PRH_CHECK_ARG(patch_class[i] >= 0 && patch_class[i] <= 2, "prh_batch_set_field: patch class out of range");
        PRH_CHECK_ARG(patch_class[i] != 2 || (cell_x[i] >= 0 && cell_x[i] < nx),
                      "prh_batch_set_field: cell x index out of range");
    }
    PRH_HIP(hipSetDevice(b->device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    if ((rc = batch_drain(b, st))) return rc;
    if (!b->cls) {
        PRH_HIP(hipMalloc((void **)&b->cls, n));
        PRH_HIP(hipMalloc((void **)&b->cell_x, n * 4));
        b->bytes += n * 5;
    }
    PRH_HIP(hipMemcpyAsync(b->cls, patch_class, n, hipMemcpyHostToDevice, st));
    PRH_HIP(hipMemcpyAsync(b->cell_x, cell_x, n * 4, hipMemcpyHostToDevice, st));
    PRH_HIP(hipStreamSynchronize(st));
    b->field_nx = nx;
    b->cross_section = cross_section_m2;
    return PRH_OK;
}

int prh_batch_admit(prh_batch *b, const double *emitted, const double *emissivity, int32_t eps_per_patch,
                    double relative_tolerance, int32_t *slot_out)
{
    PRH_CHECK_ARG(b && emitted && emissivity && slot_out, "prh_batch_admit: null pointer");
    PRH_CHECK_ARG(relative_tolerance > 0.0, "prh_batch_admit: relative_tolerance must be positive");
    const int q = first_free(b);
    if (q < 0) {
        set_error("prh_batch_admit: all %d slots are busy", SLOTS);
        return PRH_ERR_CAPACITY;
    }
    const int64_t n = b->n;
    double e_norm = 0.0, e_max = 0.0, rho_max = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        e_norm += std::fabs(emitted[i]);
        e_max = std::max(e_max, std::fabs(emitted[i]));
        const double e = eps_per_patch ? emissivity[i] : emissivity[0];
        PRH_CHECK_ARG(e >= 0.0 && e <= 1.0, "prh_batch_admit: emissivity must lie in [0, 1]");
        rho_max = std::max(rho_max, 1.0 - e);
    }
    PRH_CHECK_ARG(rho_max < 1.0, "prh_batch_admit: an emissivity of 0 everywhere has no absorbing patch");
    PRH_HIP(hipSetDevice(b->device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    // the first epoch's scale: |E| / s <= 1 keeps the FP16 gathered vector in range
    const double s0 = e_max > 0.0 ? e_max : 1.0;
    PRH_HIP(hipMemcpyAsync(b->stage, emitted, n * 8, hipMemcpyHostToDevice, st));
    PRH_HIP(hipMemcpyAsync(b->stage + n, emissivity, (eps_per_patch ? n : 1) * 8, hipMemcpyHostToDevice, st));
    k_admit<<<grid_of(n, 256), 256, 0, st>>>(view_of(b), vecs_of(b), z_next(b), q, b->stage, b->stage + n,
                                             eps_per_patch, 1.0 / s0, b->op->uniform ? 1 : 0);
    PRH_HIP(hipGetLastError());
    if (e_norm == 0.0) {
        // nothing to iterate: the zero solution is exact (the slot never runs on the device)
        SlotHost &S = b->slot[q];
        S = SlotHost();
        S.state = DONE;
        for (int c = 0; c < 4; ++c) b->host->fin[q][c] = 0.0;
    } else {
        admit_slot(b, q, relative_tolerance, rho_max, e_norm, s0, false, st, &rc);
        if (rc) return rc;
    }
    // the caller's arrays and the staging buffer are free again once this returns
    PRH_HIP(hipStreamSynchronize(st));
    *slot_out = q;
    return PRH_OK;
}

int prh_batch_admit_linear(prh_batch *b, double T_hot_K, double T_cold_K, double emissivity,
                           double relative_tolerance, int32_t *slot_out)
{
    PRH_CHECK_ARG(b && slot_out, "prh_batch_admit_linear: null pointer");
    PRH_CHECK_ARG(b->cls, "prh_batch_admit_linear: no field set (prh_batch_set_field)");
    PRH_CHECK_ARG(relative_tolerance > 0.0, "prh_batch_admit_linear: relative_tolerance must be positive");
    PRH_CHECK_ARG(emissivity > 0.0 && emissivity <= 1.0, "prh_batch_admit_linear: emissivity must lie in (0, 1]");
    PRH_CHECK_ARG(std::isfinite(T_hot_K) && std::isfinite(T_cold_K),
                  "prh_batch_admit_linear: temperatures must be finite");
    const int q = first_free(b);
    if (q < 0) {
        set_error("prh_batch_admit_linear: all %d slots are busy", SLOTS);
        return PRH_ERR_CAPACITY;
    }
    PRH_HIP(hipSetDevice(b->device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    const double T_m = 0.5 * (T_hot_K + T_cold_K);
    const double dT = T_hot_K - T_cold_K;
    // scale of the first epoch: the largest |E| of the linear profile (walls
    // at T_m +/- dT/2, interior in between), so |E| / s <= 1
    const double Th4 = std::fabs(std::pow(T_hot_K, 4) - std::pow(T_m, 4));
    const double Tc4 = std::fabs(std::pow(T_cold_K, 4) - std::pow(T_m, 4));
    double s0 = emissivity * SIGMA * b->max_area * std::max(Th4, Tc4);
    if (!(s0 > 0.0) || !std::isfinite(s0)) s0 = 1.0;
    const int64_t n = b->n;
    const unsigned grid = grid_of(n, 256);
    // the slot's schedule first; ||E||_1, summed on the device, then lands in it
    admit_slot(b, q, relative_tolerance, 1.0 - emissivity, 0.0, s0, true, st, &rc);
    if (rc) return rc;
    double *part = b->part + (size_t)SLOTS * b->n_blocks - grid;
    k_admit_linear<<<grid, 256, 0, st>>>(view_of(b), vecs_of(b), field_of(b), z_next(b), q, emissivity,
                                         emissivity * SIGMA, T_m, dT, 1.0 / s0, b->op->uniform ? 1 : 0, part);
    k_column_totals<<<1, 256, 0, st>>>(part, (int64_t)grid, &b->dslots[q].e_norm);
    PRH_HIP(hipGetLastError());
    *slot_out = q;
    return PRH_OK;
}

int prh_batch_step(prh_batch *b, int32_t max_ticks, uint64_t *done_mask, uint64_t *active_mask)
{
    PRH_CHECK_ARG(b, "prh_batch_step: null batch");
    PRH_CHECK_ARG(max_ticks >= 1, "prh_batch_step: max_ticks must be positive");
    PRH_HIP(hipSetDevice(b->device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    // Replays are pipelined: replay L+1 is queued before the results of
    // replay L are read, so the device always has the next ticks.
    int32_t t = 0;
    for (;;) {
        bool running = false;
        for (int q = 0; q < SLOTS; ++q) running |= b->slot[q].state == RUNNING;
        if (!running) {
            // nothing runs: read everything launched and let the finishes land
            if ((rc = batch_drain(b, st))) return rc;
            break;
        }
        if (t >= max_ticks) break;  // the last replay stays in flight
        const int count = max_ticks - t >= GRAPH_TICKS ? GRAPH_TICKS : 1;
        if ((rc = launch(b, count, st))) return rc;
        t += count;
        for (int64_t L = b->read; L < b->launches - 1; ++L)
            if ((rc = read_launch(b, L, st))) return rc;
        if ((rc = settle(b, false))) return rc;
        uint64_t done = 0;
        report(b, &done, nullptr);
        if (done) break;
    }
    report(b, done_mask, active_mask);
    return PRH_OK;
}

int prh_batch_channels(prh_batch *b, int32_t slot, double *q_W_m2, double *residual, int32_t *iterations)
{
    PRH_CHECK_ARG(b && q_W_m2 && residual && iterations, "prh_batch_channels: null pointer");
    PRH_CHECK_ARG(slot >= 0 && slot < SLOTS && b->slot[slot].state == DONE,
                  "prh_batch_channels: slot %d has no finished solve", slot);
    PRH_CHECK_ARG(b->cls, "prh_batch_channels: no field set (prh_batch_set_field)");
    SlotHost &S = b->slot[slot];
    if (S.failed) {
        set_error("radiosity reflection series did not converge");
        return PRH_ERR_SERIES;
    }
    const double *f = b->host->fin[slot];
    for (int c = 0; c < 3; ++c) q_W_m2[c] = -f[c] / b->cross_section;
    *residual = f[3];
    *iterations = S.iterations;
    return PRH_OK;
}

int prh_batch_result(prh_batch *b, int32_t slot, double *absorbed, double *residual, int32_t *iterations)
{
    PRH_CHECK_ARG(b && absorbed && residual && iterations, "prh_batch_result: null pointer");
    PRH_CHECK_ARG(slot >= 0 && slot < SLOTS && b->slot[slot].state == DONE,
                  "prh_batch_result: slot %d has no finished solve", slot);
    SlotHost &S = b->slot[slot];
    if (S.failed) {
        S = SlotHost();
        set_error("radiosity reflection series did not converge");
        return PRH_ERR_SERIES;
    }
    PRH_HIP(hipSetDevice(b->device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    k_collect<<<grid_of(b->n, 256), 256, 0, st>>>(b->n, slot, b->eps64, b->y64, b->stage);
    PRH_HIP(hipGetLastError());
    PRH_HIP(hipMemcpyAsync(absorbed, b->stage, b->n * 8, hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));
    *residual = b->host->fin[slot][3];
    *iterations = S.iterations;
    S = SlotHost();
    return PRH_OK;
}

int prh_batch_release(prh_batch *b, int32_t slot)
{
    PRH_CHECK_ARG(b, "prh_batch_release: null batch");
    PRH_CHECK_ARG(slot >= 0 && slot < SLOTS && b->slot[slot].state == DONE,
                  "prh_batch_release: slot %d has no finished solve", slot);
    b->slot[slot] = SlotHost();
    return PRH_OK;
}

// Mixed-precision path of prh_radiosity: continuous batching of n_rhs
// emissions through one engine (slots refill as solves finish).
int prh_radiosity_mixed(prh_operator *op, int32_t n_rhs, const double *emitted, const double *emissivity,
                        int32_t emissivity_per_rhs, double relative_tolerance, double *absorbed, double *residual,
                        int32_t *iterations)
{
    if (!op->batch) {
        int rc = prh_batch_create(op, &op->batch);
        if (rc) return rc;
    }
    prh_batch *b = op->batch;
    const int64_t n = op->n;
    int rhs_of_slot[SLOTS];
    for (int q = 0; q < SLOTS; ++q) rhs_of_slot[q] = -1;
    int next = 0, finished = 0, status = PRH_OK;
    while (finished < n_rhs) {
        uint64_t free_mask = 0;
        int rc = prh_batch_free_slots(b, &free_mask);
        if (rc) return rc;
        while (next < n_rhs && free_mask) {
            int32_t q = -1;
            const double *eps = emissivity + (emissivity_per_rhs ? (int64_t)next * n : 0);
            if ((rc = prh_batch_admit(b, emitted + (int64_t)next * n, eps, 1, relative_tolerance, &q))) return rc;
            rhs_of_slot[q] = next++;
            free_mask &= ~(uint64_t(1) << q);
        }
        // one tick per step while emissions wait for slots (a finished slot is
        // refilled at once); long replays once every emission is in
        uint64_t done = 0, active = 0;
        if ((rc = prh_batch_step(b, next < n_rhs ? 1 : MAX_TICKS * MAX_EPOCHS, &done, &active))) return rc;
        for (int q = 0; q < SLOTS; ++q) {
            if (!((done >> q) & 1u)) continue;
            const int k = rhs_of_slot[q];
            rc = prh_batch_result(b, q, absorbed + (int64_t)k * n, residual + k, iterations + k);
            if (rc == PRH_ERR_SERIES) status = rc;
            else if (rc) return rc;
            rhs_of_slot[q] = -1;
            ++finished;
        }
    }
    if (status) set_error("radiosity reflection series did not converge");
    return status;
}

void prh_batch_destroy(prh_batch *b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    hipStream_t st;
    if (current_stream(&st) == PRH_OK) (void)hipStreamSynchronize(st);
    free_batch(b);
    delete b;
}

}  // extern "C"
