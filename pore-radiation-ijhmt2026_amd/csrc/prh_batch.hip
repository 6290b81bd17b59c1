// Batched mixed-precision radiosity engine (continuous batching of solves).
//
// Solves (I - diag(rho) F_p^T) J = E, rho = 1 - eps, for up to 16 emissions
// ("slots") at once on one device-resident exchange operator, and returns the
// absorbed power eps * (F_p^T J) -- the quantity the reference's reflection
// series sums reflection by reflection (porerad/exchange.py:217-245).
//
// Method.  F_p^T x = d * S(d x / A) is similar to a symmetric matrix of
// spectral radius 1 (F_p is row-stochastic and A F_p symmetric), so the
// eigenvalues of I - diag(rho) F_p^T lie in [1 - rho_max, 1 + rho_max].
// Each epoch runs m Chebyshev iterations on that interval in FP32 (residual
// form), adds the correction to the FP64 solution, and recomputes the FP64
// residual r = E - (x - rho * F_p^T x) with an FP64 sweep; epochs repeat
// until ||r||_1 <= rtol ||E||_1 (iterative refinement).  m is the Chebyshev
// count whose error bound, for the slot's rho_max, covers what is left to
// reach the tolerance (at most a 1e-6 reduction, what FP32 sweeps resolve).
//
// Sweeps read the operator once for all 16 slots: the sparsity of S with the
// exact integer K = C + C^T as uint16 when all patches share area and ray
// count (S = A/(2R) K + A delta I), else FP32/FP64 values of S; the slot
// vectors are interleaved [row][slot] so one gather serves every slot.
// Workgroups are mapped so that each XCD sweeps a contiguous range of rows.
//
// Continuous batching: a free slot may be filled at any epoch boundary while
// the other slots keep iterating.  A slot's arithmetic never mixes with
// another slot's, every reduction has a fixed order, and a slot stops its
// own Chebyshev iterations after its own count m (its direction is set to
// zero while longer slots finish the epoch), so a solve returns the same
// bits alone or in any mix.
#include <algorithm>
#include <cmath>
#include <vector>

#include "prh_internal.h"
#include "prh_operator_impl.h"

namespace prh {
namespace {

constexpr int SLOTS = 16;
constexpr int BT = 256;
constexpr int ROWS = BT / 64;
constexpr int MAX_TICKS = 64;
constexpr int MAX_EPOCHS = 30;
constexpr double INNER_REDUCTION = 1.0e-6;

struct MatView {
    int64_t n;
    const int64_t *rowptr;
    const int32_t *col;
    const uint16_t *kval;
    const float *val32;
    const double *val64;
    const float *d32;
    const double *d64;
    const double *area;
    float inv_2R_f, delta_f;
    double inv_2R, delta;
    uint32_t blocks_per_xcd;
};

// Workgroups are dealt round-robin to the 8 XCDs; give XCD x the logical
// blocks [x * blocks_per_xcd, (x+1) * blocks_per_xcd) so that its L2 serves a
// contiguous (spatially coherent) range of patches.
__device__ __forceinline__ int64_t first_row(const MatView &M)
{
    const uint32_t b = blockIdx.x;
    return ((int64_t)(b % 8u) * M.blocks_per_xcd + (b / 8u)) * ROWS;
}

struct TickArgs {
    MatView M;
    const float *z_in;
    float *z_out;
    float *x32, *r32, *dir32;
    const float *rho32;
    const float *coef;  // [SLOTS][2] alpha, beta of this tick
};

// One FP32 Chebyshev iteration for all slots: y = F_p^T dir, then
// x += dir, r -= dir - rho y, dir = alpha dir + beta r, z = scaled dir.
template <bool UNIFORM>
__global__ __launch_bounds__(BT) void k_cheb_tick(TickArgs a)
{
    const MatView &M = a.M;
    const int64_t row = first_row(M) + (threadIdx.x >> 6);
    if (row >= M.n) return;
    const int lane = threadIdx.x & 63;
    const int h = lane & 3;   // slots 4h .. 4h+3
    const int g = lane >> 2;  // nonzero group
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    const int64_t e1 = M.rowptr[row + 1];
    for (int64_t e = M.rowptr[row] + g; e < e1; e += 16) {
        const int32_t j = M.col[e];
        const float v = UNIFORM ? (float)M.kval[e] : M.val32[e];
        const float4 z = *reinterpret_cast<const float4 *>(a.z_in + (size_t)j * SLOTS + 4 * h);
        acc.x += v * z.x;
        acc.y += v * z.y;
        acc.z += v * z.z;
        acc.w += v * z.w;
    }
    for (int off = 4; off < 64; off <<= 1) {
        acc.x += __shfl_xor(acc.x, off);
        acc.y += __shfl_xor(acc.y, off);
        acc.z += __shfl_xor(acc.z, off);
        acc.w += __shfl_xor(acc.w, off);
    }
    if (g != 0) return;
    const float di = M.d32[row];
    const float ai = UNIFORM ? 1.f : (float)M.area[row];
    const float w[4] = {acc.x, acc.y, acc.z, acc.w};
    for (int c = 0; c < 4; ++c) {
        const int q = 4 * h + c;
        const size_t at = (size_t)row * SLOTS + q;
        const float y = UNIFORM ? di * (w[c] * M.inv_2R_f + M.delta_f * a.z_in[at]) : di * w[c];
        const float dk = a.dir32[at];
        const float Ad = dk - a.rho32[at] * y;
        const float rn = a.r32[at] - Ad;
        const float dn = a.coef[2 * q] * dk + a.coef[2 * q + 1] * rn;
        a.x32[at] = a.x32[at] + dk;
        a.r32[at] = rn;
        a.dir32[at] = dn;
        a.z_out[at] = UNIFORM ? di * dn : di * dn / ai;
    }
}

struct ResArgs {
    MatView M;
    const double *z64;
    const double *x64, *E64, *eps64;
    double *y64, *r64;
    double *part_abs, *part_sum;  // [blocks][SLOTS]
};

// FP64 residual sweep: y = F_p^T x, r = E - (x - (1 - eps) y); per-block
// sums of |r| and r per slot.
template <bool UNIFORM>
__global__ __launch_bounds__(BT) void k_residual64(ResArgs a)
{
    __shared__ double s_abs[ROWS][SLOTS], s_sum[ROWS][SLOTS];
    const MatView &M = a.M;
    const int wave = threadIdx.x >> 6;
    const int64_t row = first_row(M) + wave;
    const int lane = threadIdx.x & 63;
    const int h = lane & 7;   // slots 2h, 2h+1
    const int g = lane >> 3;  // nonzero group
    double2 acc = make_double2(0.0, 0.0);
    if (row < M.n) {
        const int64_t e1 = M.rowptr[row + 1];
        for (int64_t e = M.rowptr[row] + g; e < e1; e += 8) {
            const int32_t j = M.col[e];
            const double v = UNIFORM ? (double)M.kval[e] : M.val64[e];
            const double2 z = *reinterpret_cast<const double2 *>(a.z64 + (size_t)j * SLOTS + 2 * h);
            acc.x += v * z.x;
            acc.y += v * z.y;
        }
    }
    for (int off = 8; off < 64; off <<= 1) {
        acc.x += __shfl_xor(acc.x, off);
        acc.y += __shfl_xor(acc.y, off);
    }
    if (g == 0) {
        const double w[2] = {acc.x, acc.y};
        for (int c = 0; c < 2; ++c) {
            const int q = 2 * h + c;
            double r = 0.0;
            if (row < M.n) {
                const size_t at = (size_t)row * SLOTS + q;
                const double di = M.d64[row];
                const double y = UNIFORM ? di * (w[c] * M.inv_2R + M.delta * a.z64[at]) : di * w[c];
                r = a.E64[at] - (a.x64[at] - (1.0 - a.eps64[at]) * y);
                a.y64[at] = y;
                a.r64[at] = r;
            }
            s_abs[wave][q] = fabs(r);
            s_sum[wave][q] = r;
        }
    }
    __syncthreads();
    if (threadIdx.x < SLOTS) {
        double sa = 0.0, ss = 0.0;
        for (int w = 0; w < ROWS; ++w) {
            sa += s_abs[w][threadIdx.x];
            ss += s_sum[w][threadIdx.x];
        }
        a.part_abs[(size_t)blockIdx.x * SLOTS + threadIdx.x] = sa;
        a.part_sum[(size_t)blockIdx.x * SLOTS + threadIdx.x] = ss;
    }
}

// Fixed-order per-slot totals of the block partial sums.
__global__ __launch_bounds__(256) void k_slot_totals(const double *part_abs, const double *part_sum, int64_t n_blocks,
                                                     double *tot_abs, double *tot_sum)
{
    __shared__ double sa[256], ss[256];
    const int q = blockIdx.x;
    double a = 0.0, s = 0.0;
    for (int64_t b = threadIdx.x; b < n_blocks; b += 256) {
        a += part_abs[b * SLOTS + q];
        s += part_sum[b * SLOTS + q];
    }
    sa[threadIdx.x] = a;
    ss[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            sa[threadIdx.x] += sa[threadIdx.x + off];
            ss[threadIdx.x] += ss[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        tot_abs[q] = sa[0];
        tot_sum[q] = ss[0];
    }
}

// Epoch start for the slots in `mask`: r32 = r64 / s, x32 = 0, dir = r32,
// z = scaled dir.  Other slots get zero direction (they stay put).
__global__ void k_epoch_begin(MatView M, uint32_t mask, const double *inv_s, const double *r64, float *x32, float *r32,
                              float *dir32, float *z32, int uniform)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M.n * SLOTS) return;
    const int64_t row = t / SLOTS;
    const int q = (int)(t % SLOTS);
    const bool on = (mask >> q) & 1u;
    const float r = on ? (float)(r64[t] * inv_s[q]) : 0.f;
    x32[t] = 0.f;
    r32[t] = r;
    dir32[t] = r;
    const float di = M.d32[row];
    z32[t] = uniform ? di * r : di * r / (float)M.area[row];
}

// Epoch end: x64 += s * x32 and z64 = scaled x64 for the slots in mask.
__global__ void k_epoch_accumulate(MatView M, uint32_t mask, const double *s, const float *x32, double *x64,
                                   double *z64, int uniform)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M.n * SLOTS) return;
    const int64_t row = t / SLOTS;
    const int q = (int)(t % SLOTS);
    double x = x64[t];
    if ((mask >> q) & 1u) x = x + s[q] * (double)x32[t];
    x64[t] = x;
    const double di = M.d64[row];
    z64[t] = uniform ? di * x : di * x / M.area[row];
}

// Admission of one slot: E, eps in; x64 = 0, r64 = E.
__global__ void k_admit(int64_t n, int q, const double *E, const double *eps, int eps_per_patch, double *E64,
                        double *eps64, float *rho32, double *x64, double *r64)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t at = (size_t)i * SLOTS + q;
    const double e = eps_per_patch ? eps[i] : eps[0];
    E64[at] = E[i];
    eps64[at] = e;
    rho32[at] = (float)(1.0 - e);
    x64[at] = 0.0;
    r64[at] = E[i];
}

// Retirement of one slot: absorbed = eps * y (column q) into a packed vector.
__global__ void k_collect(int64_t n, int q, const double *eps64, const double *y64, double *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t at = (size_t)i * SLOTS + q;
    out[i] = eps64[at] * y64[at];
}

unsigned grid_of(int64_t n, int b) { return (unsigned)std::max<int64_t>(1, (n + b - 1) / b); }

struct SlotHost {
    bool active = false, done = false, failed = false;
    double rtol = 0, e_norm = 0, rho_max = 0, r_norm = 0, r_sum = 0;
    int epochs = 0, ticks = 0, epoch_ticks = 0;
};

}  // namespace
}  // namespace prh

using namespace prh;

struct prh_batch {
    prh_operator *op = nullptr;
    int device = 0;
    int64_t n = 0, n_blocks = 0;
    uint32_t blocks_per_xcd = 0;
    float *x32 = nullptr, *r32 = nullptr, *dir32 = nullptr, *za = nullptr, *zb = nullptr, *rho32 = nullptr;
    double *x64 = nullptr, *E64 = nullptr, *eps64 = nullptr, *y64 = nullptr, *r64 = nullptr, *z64 = nullptr;
    double *part_abs = nullptr, *part_sum = nullptr, *tot = nullptr, *svec = nullptr, *stage = nullptr;
    float *coef = nullptr;
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
    M.val32 = op->val32;
    M.val64 = op->val;
    M.d32 = op->scale32;
    M.d64 = op->scale;
    M.area = op->area;
    M.inv_2R = op->inv_2R;
    M.delta = op->delta;
    M.inv_2R_f = (float)op->inv_2R;
    M.delta_f = (float)op->delta;
    M.blocks_per_xcd = b->blocks_per_xcd;
    return M;
}

void free_batch(prh_batch *b)
{
    void *ptrs[] = {b->x32, b->r32, b->dir32, b->za, b->zb, b->rho32, b->x64, b->E64, b->eps64, b->y64,
                    b->r64, b->z64, b->part_abs, b->part_sum, b->tot, b->svec, b->stage, b->coef};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
}

// Chebyshev iterations that shrink the error bound 2 c^k / (1 + c^2k) below tau.
int ticks_for(double rho_max, double tau)
{
    if (rho_max <= 0.0) return 1;
    const double kappa = (1.0 + rho_max) / (1.0 - rho_max);
    const double c = (std::sqrt(kappa) - 1.0) / (std::sqrt(kappa) + 1.0);
    const int m = (int)std::ceil(std::log(tau / 2.0) / std::log(c));
    return std::min(std::max(m, 1), MAX_TICKS);
}

// Reduction an epoch aims for: what is left to reach the tolerance, with a
// factor 10 of margin, but no finer than FP32 sweeps resolve.
double epoch_target(const SlotHost &S)
{
    const double left = 0.1 * S.rtol * S.e_norm / std::max(S.r_norm, 1e-300);
    return std::min(std::max(left, INNER_REDUCTION), 0.5);
}

// alpha_k, beta_k of Chebyshev acceleration on [1 - delta, 1 + delta]
// (theta = 1): rho_0 = delta, rho_{k+1} = 1 / (2/delta - rho_k),
// dir_{k+1} = rho_{k+1} rho_k dir_k + (2 rho_{k+1} / delta) r_{k+1}.
void cheb_coefficients(double delta, int m, std::vector<float> &alpha, std::vector<float> &beta)
{
    alpha.assign(m, 0.f);
    beta.assign(m, 1.f);
    if (delta <= 0.0) return;
    double rho = delta;
    for (int k = 0; k < m; ++k) {
        const double next = 1.0 / (2.0 / delta - rho);
        alpha[k] = (float)(next * rho);
        beta[k] = (float)(2.0 * next / delta);
        rho = next;
    }
}

}  // namespace

extern "C" {

int prh_batch_create(prh_operator *op, prh_batch **out)
{
    PRH_CHECK_ARG(op && out, "prh_batch_create: null pointer");
    *out = nullptr;
    PRH_HIP(hipSetDevice(op->device));
    prh_batch *b = new prh_batch();
    b->op = op;
    b->device = op->device;
    b->n = op->n;
    const int64_t nb_logical = (op->n + ROWS - 1) / ROWS;
    b->blocks_per_xcd = (uint32_t)((nb_logical + 7) / 8);
    b->n_blocks = (int64_t)b->blocks_per_xcd * 8;
    const size_t ns = (size_t)op->n * SLOTS;
    bool ok = true;
    auto alloc = [&](void **p, size_t bytes) {
        if (ok && hipMalloc(p, bytes) != hipSuccess) ok = false;
        b->bytes += (int64_t)bytes;
    };
    alloc((void **)&b->x32, ns * 4);
    alloc((void **)&b->r32, ns * 4);
    alloc((void **)&b->dir32, ns * 4);
    alloc((void **)&b->za, ns * 4);
    alloc((void **)&b->zb, ns * 4);
    alloc((void **)&b->rho32, ns * 4);
    alloc((void **)&b->x64, ns * 8);
    alloc((void **)&b->E64, ns * 8);
    alloc((void **)&b->eps64, ns * 8);
    alloc((void **)&b->y64, ns * 8);
    alloc((void **)&b->r64, ns * 8);
    alloc((void **)&b->z64, ns * 8);
    alloc((void **)&b->part_abs, (size_t)b->n_blocks * SLOTS * 8);
    alloc((void **)&b->part_sum, (size_t)b->n_blocks * SLOTS * 8);
    alloc((void **)&b->tot, 2 * SLOTS * 8);
    alloc((void **)&b->svec, 2 * SLOTS * 8);
    alloc((void **)&b->stage, (size_t)op->n * 2 * 8);
    alloc((void **)&b->coef, (size_t)MAX_TICKS * SLOTS * 2 * 4);
    if (!ok) {
        free_batch(b);
        delete b;
        set_error("prh_batch_create: out of device memory");
        return PRH_ERR_HIP;
    }
    // Zero the slot state so idle slots carry zeros through every sweep.
    for (void *p : {(void *)b->x32, (void *)b->r32, (void *)b->dir32, (void *)b->za, (void *)b->zb,
                    (void *)b->rho32})
        PRH_HIP(hipMemset(p, 0, ns * 4));
    for (void *p : {(void *)b->x64, (void *)b->E64, (void *)b->eps64, (void *)b->y64, (void *)b->r64,
                    (void *)b->z64})
        PRH_HIP(hipMemset(p, 0, ns * 8));
    *out = b;
    return PRH_OK;
}

int prh_batch_bytes(const prh_batch *b, int64_t *bytes)
{
    PRH_CHECK_ARG(b && bytes, "prh_batch_bytes: null pointer");
    *bytes = b->bytes;
    return PRH_OK;
}

int prh_batch_free_slots(const prh_batch *b, uint32_t *free_mask)
{
    PRH_CHECK_ARG(b && free_mask, "prh_batch_free_slots: null pointer");
    uint32_t m = 0;
    for (int q = 0; q < SLOTS; ++q)
        if (!b->slot[q].active) m |= 1u << q;
    *free_mask = m;
    return PRH_OK;
}

int prh_batch_admit(prh_batch *b, const double *emitted, const double *emissivity, int32_t eps_per_patch,
                    double relative_tolerance, int32_t *slot_out)
{
    PRH_CHECK_ARG(b && emitted && emissivity && slot_out, "prh_batch_admit: null pointer");
    PRH_CHECK_ARG(relative_tolerance > 0.0, "prh_batch_admit: relative_tolerance must be positive");
    int q = -1;
    for (int k = 0; k < SLOTS; ++k)
        if (!b->slot[k].active) {
            q = k;
            break;
        }
    if (q < 0) {
        set_error("prh_batch_admit: all %d slots are busy", SLOTS);
        return PRH_ERR_CAPACITY;
    }
    const int64_t n = b->n;
    double e_norm = 0.0, rho_max = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        e_norm += std::fabs(emitted[i]);
        const double e = eps_per_patch ? emissivity[i] : emissivity[0];
        PRH_CHECK_ARG(e >= 0.0 && e <= 1.0, "prh_batch_admit: emissivity must lie in [0, 1]");
        rho_max = std::max(rho_max, 1.0 - e);
    }
    PRH_CHECK_ARG(rho_max < 1.0, "prh_batch_admit: an emissivity of 0 everywhere has no absorbing patch");
    PRH_HIP(hipSetDevice(b->device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    PRH_HIP(hipMemcpyAsync(b->stage, emitted, n * 8, hipMemcpyHostToDevice, st));
    PRH_HIP(hipMemcpyAsync(b->stage + n, emissivity, (eps_per_patch ? n : 1) * 8, hipMemcpyHostToDevice, st));
    k_admit<<<grid_of(n, 256), 256, 0, st>>>(n, q, b->stage, b->stage + n, eps_per_patch, b->E64, b->eps64,
                                             b->rho32, b->x64, b->r64);
    PRH_HIP(hipGetLastError());
    PRH_HIP(hipStreamSynchronize(st));
    SlotHost &S = b->slot[q];
    S = SlotHost();
    S.active = true;
    S.rtol = relative_tolerance;
    S.e_norm = std::max(e_norm, 2.2250738585072014e-308);
    S.rho_max = rho_max;
    S.r_norm = e_norm;
    S.done = e_norm == 0.0;
    *slot_out = q;
    return PRH_OK;
}

int prh_batch_epoch(prh_batch *b, uint32_t *done_mask, uint32_t *active_mask)
{
    PRH_CHECK_ARG(b, "prh_batch_epoch: null batch");
    PRH_HIP(hipSetDevice(b->device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    const prh_operator *op = b->op;
    const MatView M = view_of(b);
    uint32_t mask = 0;
    int m = 0;
    double s[SLOTS], inv_s[SLOTS];
    for (int q = 0; q < SLOTS; ++q) {
        SlotHost &S = b->slot[q];
        s[q] = 0.0;
        inv_s[q] = 0.0;
        if (!S.active || S.done) continue;
        if (S.epochs >= MAX_EPOCHS) {
            S.done = true;
            S.failed = true;
            continue;
        }
        mask |= 1u << q;
        S.epoch_ticks = ticks_for(S.rho_max, epoch_target(S));
        m = std::max(m, S.epoch_ticks);
        const double scale = S.r_norm / (double)b->n;
        s[q] = scale > 0.0 ? scale : 1.0;
        inv_s[q] = 1.0 / s[q];
    }
    if (mask) {
        // Per-tick coefficients of every running slot.
        std::vector<float> coef((size_t)m * SLOTS * 2, 0.f), al, be;
        for (int q = 0; q < SLOTS; ++q) {
            if (!((mask >> q) & 1u)) continue;
            const int mq = b->slot[q].epoch_ticks;
            cheb_coefficients(b->slot[q].rho_max, mq, al, be);
            // The direction after the slot's last own tick is zero, so longer
            // slots in the same epoch leave this slot untouched.
            for (int k = 0; k + 1 < mq; ++k) {
                coef[((size_t)k * SLOTS + q) * 2] = al[k];
                coef[((size_t)k * SLOTS + q) * 2 + 1] = be[k];
            }
        }
        PRH_HIP(hipMemcpyAsync(b->coef, coef.data(), coef.size() * 4, hipMemcpyHostToDevice, st));
        PRH_HIP(hipMemcpyAsync(b->svec, s, SLOTS * 8, hipMemcpyHostToDevice, st));
        PRH_HIP(hipMemcpyAsync(b->svec + SLOTS, inv_s, SLOTS * 8, hipMemcpyHostToDevice, st));
        const int64_t ns = b->n * SLOTS;
        k_epoch_begin<<<grid_of(ns, 256), 256, 0, st>>>(M, mask, b->svec + SLOTS, b->r64, b->x32, b->r32, b->dir32,
                                                        b->za, op->uniform ? 1 : 0);
        float *zin = b->za, *zout = b->zb;
        const unsigned grid = (unsigned)b->n_blocks;
        for (int k = 0; k < m; ++k) {
            TickArgs a{M, zin, zout, b->x32, b->r32, b->dir32, b->rho32, b->coef + (size_t)k * SLOTS * 2};
            if (op->uniform) k_cheb_tick<true><<<grid, BT, 0, st>>>(a);
            else k_cheb_tick<false><<<grid, BT, 0, st>>>(a);
            std::swap(zin, zout);
        }
        k_epoch_accumulate<<<grid_of(ns, 256), 256, 0, st>>>(M, mask, b->svec, b->x32, b->x64, b->z64,
                                                             op->uniform ? 1 : 0);
        ResArgs r{M, b->z64, b->x64, b->E64, b->eps64, b->y64, b->r64, b->part_abs, b->part_sum};
        if (op->uniform) k_residual64<true><<<grid, BT, 0, st>>>(r);
        else k_residual64<false><<<grid, BT, 0, st>>>(r);
        k_slot_totals<<<SLOTS, 256, 0, st>>>(b->part_abs, b->part_sum, b->n_blocks, b->tot, b->tot + SLOTS);
        PRH_HIP(hipGetLastError());
        double tot[2 * SLOTS];
        PRH_HIP(hipMemcpyAsync(tot, b->tot, sizeof(tot), hipMemcpyDeviceToHost, st));
        PRH_HIP(hipStreamSynchronize(st));
        for (int q = 0; q < SLOTS; ++q) {
            if (!((mask >> q) & 1u)) continue;
            SlotHost &S = b->slot[q];
            S.epochs += 1;
            S.ticks += S.epoch_ticks;
            S.r_norm = tot[q];
            S.r_sum = tot[SLOTS + q];
            if (S.r_norm <= S.rtol * S.e_norm) S.done = true;
        }
    }
    uint32_t done = 0, active = 0;
    for (int q = 0; q < SLOTS; ++q) {
        if (b->slot[q].active) active |= 1u << q;
        if (b->slot[q].active && b->slot[q].done) done |= 1u << q;
    }
    if (done_mask) *done_mask = done;
    if (active_mask) *active_mask = active;
    return PRH_OK;
}

int prh_batch_result(prh_batch *b, int32_t slot, double *absorbed, double *residual, int32_t *iterations)
{
    PRH_CHECK_ARG(b && absorbed && residual && iterations, "prh_batch_result: null pointer");
    PRH_CHECK_ARG(slot >= 0 && slot < SLOTS && b->slot[slot].active && b->slot[slot].done,
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
    if (S.epochs == 0) {
        // zero emission: nothing was iterated
        std::fill(absorbed, absorbed + b->n, 0.0);
    } else {
        k_collect<<<grid_of(b->n, 256), 256, 0, st>>>(b->n, slot, b->eps64, b->y64, b->stage);
        PRH_HIP(hipGetLastError());
        PRH_HIP(hipMemcpyAsync(absorbed, b->stage, b->n * 8, hipMemcpyDeviceToHost, st));
        PRH_HIP(hipStreamSynchronize(st));
    }
    *residual = S.r_sum;
    *iterations = S.ticks + S.epochs;
    S = SlotHost();
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
    int slot_of_rhs[SLOTS];
    int rhs_of_slot[SLOTS];
    for (int q = 0; q < SLOTS; ++q) rhs_of_slot[q] = -1;
    (void)slot_of_rhs;
    int next = 0, finished = 0, status = PRH_OK;
    while (finished < n_rhs) {
        uint32_t free_mask = 0;
        int rc = prh_batch_free_slots(b, &free_mask);
        if (rc) return rc;
        while (next < n_rhs && free_mask) {
            int32_t q = -1;
            const double *eps = emissivity + (emissivity_per_rhs ? (int64_t)next * n : 0);
            if ((rc = prh_batch_admit(b, emitted + (int64_t)next * n, eps, 1, relative_tolerance, &q))) return rc;
            rhs_of_slot[q] = next++;
            free_mask &= ~(1u << q);
        }
        uint32_t done = 0, active = 0;
        if ((rc = prh_batch_epoch(b, &done, &active))) return rc;
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
    free_batch(b);
    delete b;
}

}  // extern "C"
