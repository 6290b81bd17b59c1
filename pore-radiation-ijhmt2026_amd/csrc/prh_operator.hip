// Exchange operator and FP64 reflection series.
//
// Reference: ExchangeOperator (porerad/exchange.py:37-122) and
// reflection_series (porerad/exchange.py:217-245).
//
// The symmetric matrix S = 0.5 (diag(A) F + F^T diag(A)) + delta diag(A) is
// assembled once per operator as an FP64 CSR over the union of the sparsity
// of C and C^T plus the diagonal; F_p is never stored: F_p x = (d/A) S(d x)
// and F_p^T x = d S(d x / A), as in the reference.  The raw count matrix is
// kept as two integer CSRs (C and C^T) for raw_matvec / raw_rmatvec.
//
// Every row is reduced by one wavefront in a fixed lane/tree order, so the
// results do not depend on scheduling or on which other right-hand sides are
// in flight: the same inputs always give the same bits.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "prh_internal.h"
#include "prh_operator_impl.h"

namespace prh {
namespace {

constexpr int ROW_THREADS = 256;  // 4 rows (one per wavefront) per workgroup
constexpr int ROWS_PER_BLOCK = ROW_THREADS / 64;
constexpr int RED_THREADS = 256;

__device__ inline double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// S in one of its two stored forms: FP64 values, or (every patch with the
// same area a and ray count R) the integer K = C + C^T with
// S = a (K / (2R) + delta I), read as 2 B instead of 8 B per nonzero.  Both
// are summed in FP64; the integer form only changes the rounding of S_ij.
struct SView {
    const int64_t *rowptr;
    const int32_t *col;
    const double *val;
    const uint16_t *kval;
    const double *area;
    double inv_2R, delta;
};

// sum_j v_j x_j over row `row` (v = integer K or FP64 S), FP64 accumulation
// of x in its own type, from the lane partials of one wavefront.
template <bool UNIFORM, typename V>
__device__ __forceinline__ double row_sum(const SView &S, int64_t row, int lane, const V *x)
{
    // four independent gathers in flight per lane (rows hold ~10^3 nonzeros)
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    const int64_t e1 = S.rowptr[row + 1];
    int64_t e = S.rowptr[row] + lane;
    for (; e + 192 < e1; e += 256) {
        const int32_t c0 = S.col[e], c1 = S.col[e + 64], c2 = S.col[e + 128], c3 = S.col[e + 192];
        const double v0 = UNIFORM ? (double)S.kval[e] : S.val[e];
        const double v1 = UNIFORM ? (double)S.kval[e + 64] : S.val[e + 64];
        const double v2 = UNIFORM ? (double)S.kval[e + 128] : S.val[e + 128];
        const double v3 = UNIFORM ? (double)S.kval[e + 192] : S.val[e + 192];
        const double x0 = (double)x[c0], x1 = (double)x[c1], x2 = (double)x[c2], x3 = (double)x[c3];
        a0 += v0 * x0;
        a1 += v1 * x1;
        a2 += v2 * x2;
        a3 += v3 * x3;
    }
    for (; e < e1; e += 64) a0 += (UNIFORM ? (double)S.kval[e] : S.val[e]) * (double)x[S.col[e]];
    return wave_sum((a0 + a1) + (a2 + a3));
}

// (S x)_row from the lane partials of one wavefront.
template <bool UNIFORM>
__device__ __forceinline__ double s_row(const SView &S, int64_t row, int lane, const double *x)
{
    const double acc = row_sum<UNIFORM, double>(S, row, lane, x);
    return UNIFORM ? S.area[row] * (acc * S.inv_2R + S.delta * x[row]) : acc;
}

// y = S x (one wavefront per row).
template <bool UNIFORM>
__global__ __launch_bounds__(ROW_THREADS) void k_spmv(SView S, int64_t n, const double *x, double *y)
{
    const int64_t row = (int64_t)blockIdx.x * ROWS_PER_BLOCK + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const double sx = s_row<UNIFORM>(S, row, lane, x);
    if (lane == 0) y[row] = sx;
}

// Raw count rows: y_i = sum_j (c_ij / R_{who}) x_j, with who = i for F and
// who = j for F^T (row of C^T holds origins).
__global__ __launch_bounds__(ROW_THREADS) void k_spmv_counts(const int64_t *rowptr, const int32_t *col,
                                                             const int32_t *cnt, const double *rays, int transpose,
                                                             int64_t n, const double *x, double *y)
{
    const int64_t row = (int64_t)blockIdx.x * ROWS_PER_BLOCK + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    double acc = 0.0;
    for (int64_t e = rowptr[row] + lane; e < rowptr[row + 1]; e += 64) {
        const int32_t j = col[e];
        const double w = (double)cnt[e] / rays[transpose ? j : row];
        acc += w * x[j];
    }
    acc = wave_sum(acc);
    if (lane == 0) y[row] = acc;
}

// Elementwise pre/post scalings of matvec / rmatvec.
__global__ void k_scale_in(const double *x, const double *d, const double *area, int mode, int64_t n, double *out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = mode == PRH_APPLY_MATVEC ? d[i] * x[i] : d[i] * (x[i] / area[i]);
}

__global__ void k_scale_out(const double *sx, const double *d, const double *area, int mode, int64_t n, double *y)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    y[i] = mode == PRH_APPLY_MATVEC ? (d[i] / area[i]) * sx[i] : d[i] * sx[i];
}

// flight = E, absorbed = 0, w = d (E / A), per-block sums of |E|.
__global__ __launch_bounds__(RED_THREADS) void k_series_init(const double *E, const double *d, const double *area,
                                                             int64_t n, double *flight, double *absorbed, double *w,
                                                             double *part_abs)
{
    __shared__ double red[RED_THREADS];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * RED_THREADS) {
        const double e = E[i];
        flight[i] = e;
        absorbed[i] = 0.0;
        w[i] = d[i] * (e / area[i]);
        s += fabs(e);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = RED_THREADS / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part_abs[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(RED_THREADS) void k_series_start(const double *part_abs, int n_part, SeriesState *S)
{
    __shared__ double red[RED_THREADS];
    double s = 0.0;
    for (int b = threadIdx.x; b < n_part; b += RED_THREADS) s += part_abs[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = RED_THREADS / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double tiny = 2.2250738585072014e-308;
        S->initial = red[0] > tiny ? red[0] : tiny;
        S->l1 = red[0];
        S->residual = 0.0;
        S->done = 0;
        S->iterations = 0;
    }
}

// One reflection: incident = d * (S w); absorbed += eps * incident;
// flight = (1 - eps) * incident; w_next = d * (flight / A).  Per-block
// partial sums of |flight| and flight for the stopping rule and residual.
template <bool UNIFORM>
__global__ __launch_bounds__(ROW_THREADS) void k_series_step(SView M, int64_t n, const double *d, const double *area,
                                                             const double *eps, const double *w, double *w_next,
                                                             double *absorbed, double *flight, double *part_abs,
                                                             double *part_sum, const SeriesState *S)
{
    if (S->done) return;
    __shared__ double sa[ROWS_PER_BLOCK], ss[ROWS_PER_BLOCK];
    const int wave = threadIdx.x >> 6;
    const int64_t row = (int64_t)blockIdx.x * ROWS_PER_BLOCK + wave;
    const int lane = threadIdx.x & 63;
    double fl = 0.0;
    if (row < n) {
        const double incident = d[row] * s_row<UNIFORM>(M, row, lane, w);
        const double ep = eps[row];
        fl = (1.0 - ep) * incident;
        if (lane == 0) {
            absorbed[row] = absorbed[row] + ep * incident;
            flight[row] = fl;
            w_next[row] = d[row] * (fl / area[row]);
        }
    }
    if (lane == 0) {
        sa[wave] = fabs(fl);
        ss[wave] = fl;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, s = 0.0;
        for (int k = 0; k < ROWS_PER_BLOCK; ++k) {
            a += sa[k];
            s += ss[k];
        }
        part_abs[blockIdx.x] = a;
        part_sum[blockIdx.x] = s;
    }
}

// ---------------------------------------------------------------- single-emission mixed solve
// One emission of (I - rho F_p^T) J = E with scalar rho = 1 - eps: epochs of
// FP32 Chebyshev sweeps in residual form, each ended by an FP64 refinement
// sweep (the batched engine's iteration for one slot).  z holds the gathered
// form of the FP32 vector of the sweep: d v (equal areas) or d v / A.
struct Mix1 {
    float *x32, *r32, *dir32;
    double *x64, *y64, *r64;
};

enum : int { MIX_CHEB = 0, MIX_LAST = 1, MIX_REFINE = 2 };

__device__ __forceinline__ float mix_scaled(float di, float v, float ai, bool uniform)
{
    return uniform ? di * v : di * v / ai;
}

// One sweep: y = F_p^T z (FP64 accumulation of the FP32 z), then the row's
// update.  CHEB: dir' = alpha dir + beta r'; LAST: the epoch's correction x32
// is gathered next; REFINE: x64 += s c, y64 += s y with c the correction,
// r64 = E - (x64 - rho y64), and the block sums of |r64|.
template <bool UNIFORM>
__global__ __launch_bounds__(ROW_THREADS) void k_mix1_sweep(SView M, int64_t n, const double *d, const float *d32,
                                                            Mix1 V, const float *z_in, float *z_out, int mode,
                                                            float alpha, float beta, float rho32, double rho,
                                                            double s, const double *E, double *part_abs)
{
    __shared__ double sa[ROWS_PER_BLOCK];
    const int wave = threadIdx.x >> 6;
    const int64_t row = (int64_t)blockIdx.x * ROWS_PER_BLOCK + wave;
    const int lane = threadIdx.x & 63;
    double ra = 0.0;
    if (row < n) {
        const double w = row_sum<UNIFORM, float>(M, row, lane, z_in);
        if (lane == 0) {
            const float zi = z_in[row];
            const double di = d[row];
            const double y = UNIFORM ? di * (w * M.inv_2R + M.delta * (double)zi) : di * w;
            if (mode == MIX_REFINE) {
                const double c = UNIFORM ? (double)zi / di : (double)zi * M.area[row] / di;
                const double x = V.x64[row] + s * c;
                const double yy = V.y64[row] + s * y;
                const double r = E[row] - (x - rho * yy);
                V.x64[row] = x;
                V.y64[row] = yy;
                V.r64[row] = r;
                ra = fabs(r);
            } else {
                const float ai32 = UNIFORM ? 1.f : (float)M.area[row];
                const float dk = V.dir32[row];
                const float Ad = dk - rho32 * (float)y;
                const float rn = V.r32[row] - Ad;
                const float x = V.x32[row] + dk;
                V.x32[row] = x;
                if (mode == MIX_CHEB) {
                    const float dn = alpha * dk + beta * rn;
                    V.r32[row] = rn;
                    V.dir32[row] = dn;
                    z_out[row] = mix_scaled(d32[row], dn, ai32, UNIFORM);
                } else {
                    z_out[row] = mix_scaled(d32[row], x, ai32, UNIFORM);
                }
            }
        }
    }
    if (mode != MIX_REFINE) return;  // the same for every thread of the grid
    if (lane == 0) sa[wave] = ra;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0;
        for (int k = 0; k < ROWS_PER_BLOCK; ++k) a += sa[k];
        part_abs[blockIdx.x] = a;
    }
}

// Start: x64 = y64 = 0, r64 = E, and the block sums of |E|.
__global__ __launch_bounds__(RED_THREADS) void k_mix1_start(const double *E, int64_t n, Mix1 V, double *part_abs)
{
    __shared__ double red[RED_THREADS];
    double a = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * RED_THREADS) {
        const double e = E[i];
        V.x64[i] = 0.0;
        V.y64[i] = 0.0;
        V.r64[i] = e;
        a += fabs(e);
    }
    red[threadIdx.x] = a;
    __syncthreads();
    for (int off = RED_THREADS / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part_abs[blockIdx.x] = red[0];
}

// A new epoch on the scale s: x32 = 0, r32 = dir32 = r64 / s, z = its gathered form.
__global__ void k_mix1_epoch(int64_t n, const float *d32, const double *area, int uniform, Mix1 V, double inv_s,
                             float *z)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float r = (float)(V.r64[i] * inv_s);
    V.x32[i] = 0.f;
    V.r32[i] = r;
    V.dir32[i] = r;
    z[i] = mix_scaled(d32[i], r, uniform ? 1.f : (float)area[i], uniform != 0);
}

// Fixed-order total of n_part partial sums into *out.
__global__ __launch_bounds__(RED_THREADS) void k_mix1_total(const double *part, int64_t n_part, double *out)
{
    __shared__ double red[RED_THREADS];
    double a = 0.0;
    for (int64_t b = threadIdx.x; b < n_part; b += RED_THREADS) a += part[b];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int off = RED_THREADS / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// absorbed = eps y64 (zero for the zero emission).
__global__ void k_mix1_collect(int64_t n, double eps, const double *y64, double *absorbed)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) absorbed[i] = y64 ? eps * y64[i] : 0.0;
}

__global__ __launch_bounds__(RED_THREADS) void k_series_check(const double *part_abs, const double *part_sum,
                                                              int64_t n_part, double rtol, SeriesState *S)
{
    if (S->done) return;
    __shared__ double ra[RED_THREADS], rs[RED_THREADS];
    double a = 0.0, s = 0.0;
    for (int64_t b = threadIdx.x; b < n_part; b += RED_THREADS) {
        a += part_abs[b];
        s += part_sum[b];
    }
    ra[threadIdx.x] = a;
    rs[threadIdx.x] = s;
    __syncthreads();
    for (int off = RED_THREADS / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            ra[threadIdx.x] += ra[threadIdx.x + off];
            rs[threadIdx.x] += rs[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        S->iterations += 1;
        S->l1 = ra[0];
        if (ra[0] <= rtol * S->initial) {
            S->done = 1;
            S->residual = rs[0];
        }
    }
}

// ---------------------------------------------------------------- batched FP64 series
// The reflection series of up to 32 emissions in one pass over the operator:
// slot vectors interleaved [row][FS] in FP64 (FS = 16 or 32), so one gathered
// column brings all slots in FS * 8 bytes (one 128 B line at FS = 16, the
// mixed engine's line with FP64 slots).  Each slot follows the reference's
// recursion (exchange.py:217-245) and stops on its own when
// ||flight||_1 <= rtol ||E||_1; the sums run in another order than
// k_series_step, so results agree to rounding.
constexpr int FSLOTS_MAX = 32;

struct SeriesSlots {
    double initial[FSLOTS_MAX];  // max(||E||_1, tiny)
    int32_t done[FSLOTS_MAX];
    int32_t iterations[FSLOTS_MAX];
};

// One wave per row: FS / 2 lanes cover a column's slots with 16 B loads, the
// wave's 64 / (FS / 2) lane groups take interleaved nonzeros, every lane has
// FS / 2 gathers in flight.  Workgroups are dealt round-robin to the 8 XCDs;
// XCD x sweeps the positions [x * blocks_per_xcd, (x + 1) * blocks_per_xcd)
// (in blocks) of the row order, so its L2 serves a contiguous range of it.
template <int FS, bool UNIFORM>
__global__ __launch_bounds__(ROW_THREADS) void k_series_batch(SView M, int64_t n, const double *d, const double *area,
                                                              const double *eps, const double *w, double *w_next,
                                                              double *absorbed, double *flight, double *part_abs,
                                                              uint32_t blocks_per_xcd, const int32_t *row_order,
                                                              const SeriesSlots *S)
{
    constexpr int HL = FS / 2;   // lanes per column
    constexpr int NG = 64 / HL;  // nonzeros in flight per wave
    constexpr int U = 64 / NG;   // gathers per lane per 64 nonzeros
    __shared__ double sa[ROWS_PER_BLOCK][FS];
    const int wave = threadIdx.x >> 6;
    const uint32_t b = blockIdx.x;
    const int64_t pos = ((int64_t)(b % 8u) * blocks_per_xcd + (b / 8u)) * ROWS_PER_BLOCK + wave;
    // the engine's sweep order when the operator has one (prh_operator_set_row_order)
    const int64_t row =
        (int64_t)__builtin_amdgcn_readfirstlane((int)(pos < n && row_order ? (int64_t)row_order[pos] : pos));
    const int lane = threadIdx.x & 63;
    const int h = lane % HL;  // slots 2h, 2h + 1
    const int g = lane / HL;  // nonzero group
    double acc0 = 0.0, acc1 = 0.0;
    if (row < n) {
        const int64_t e0 = M.rowptr[row], e1 = M.rowptr[row + 1];
        const uint32_t len = (uint32_t)(e1 - e0);
        const int32_t *col = M.col + e0;
        for (uint32_t base = 0; base < len; base += 64u) {
            const uint32_t e = base + (uint32_t)lane;
            const bool ok = e < len;
            const uint32_t ec = ok ? e : len - 1;
            const int32_t jl = col[ec];
            const double vl = ok ? (UNIFORM ? (double)M.kval[e0 + ec] : M.val[e0 + ec]) : 0.0;
            int32_t j[U];
            double v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                j[u] = __shfl(jl, g + NG * u);
                v[u] = __shfl(vl, g + NG * u);
            }
            double2 z[U];
#pragma unroll
            for (int u = 0; u < U; ++u) z[u] = *reinterpret_cast<const double2 *>(w + (size_t)j[u] * FS + 2 * h);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                acc0 = fma(v[u], z[u].x, acc0);
                acc1 = fma(v[u], z[u].y, acc1);
            }
        }
    }
    for (int off = HL; off < 64; off <<= 1) {
        acc0 += __shfl_xor(acc0, off);
        acc1 += __shfl_xor(acc1, off);
    }
    const int q = 2 * h + (g & 1);
    double fl = 0.0;
    if (g < 2 && row < n) {
        const size_t at = (size_t)row * FS + q;
        if (!S->done[q]) {
            const double wsum = g == 0 ? acc0 : acc1;
            const double sw = UNIFORM ? M.area[row] * (wsum * M.inv_2R + M.delta * w[at]) : wsum;
            const double incident = d[row] * sw;
            const double ep = eps[at];
            fl = (1.0 - ep) * incident;
            absorbed[at] = absorbed[at] + ep * incident;
            flight[at] = fl;
            w_next[at] = d[row] * (fl / area[row]);
        } else {
            w_next[at] = 0.0;  // a finished slot adds nothing to later gathers
        }
    }
    if (g < 2) sa[wave][q] = fabs(fl);
    __syncthreads();
    if (threadIdx.x < FS) {
        double a = 0.0;
        for (int k = 0; k < ROWS_PER_BLOCK; ++k) a += sa[k][threadIdx.x];
        part_abs[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = a;
    }
}

// Per slot (one block each): ||flight||_1 from the block partials in a fixed
// order; a slot that meets its tolerance is done after this reflection.
__global__ __launch_bounds__(RED_THREADS) void k_series_batch_check(const double *part_abs, int64_t n_blocks,
                                                                    double rtol, int32_t active, SeriesSlots *S)
{
    __shared__ double ra[RED_THREADS];
    const int q = blockIdx.x;
    if (q >= active || S->done[q]) return;
    double a = 0.0;
    for (int64_t b = threadIdx.x; b < n_blocks; b += RED_THREADS) a += part_abs[(size_t)q * n_blocks + b];
    ra[threadIdx.x] = a;
    __syncthreads();
    for (int off = RED_THREADS / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) ra[threadIdx.x] += ra[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        S->iterations[q] += 1;
        if (ra[0] <= rtol * S->initial[q]) S->done[q] = 1;
    }
}

SView sview_of(const prh_operator *op)
{
    return SView{op->rowptr, op->col, op->val, op->kval, op->area, op->inv_2R, op->delta};
}

// Position of column c in the sorted columns [lo, hi) of a CSR row, or -1.
__device__ __forceinline__ int64_t find_col(const int32_t *col, int64_t lo, int64_t hi, int32_t c)
{
    const int64_t end = hi;
    while (lo < hi) {
        const int64_t mid = lo + ((hi - lo) >> 1);
        if (col[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    return (lo < end && col[lo] == c) ? lo : -1;
}

// Reciprocity defect of the projected view factors over the count table's
// pairs (i, j), C_ij > 0 (exchange.py:124-140): A_i F_p,ij = A_i ((d_i / A_i)
// S_ij d_j) against A_j F_p,ji with the mirrored entry S_ji of the symmetric
// matrix.  Per wavefront the largest |difference| and |A_i F_p,ij|, kept as
// the bit patterns of non-negative doubles (which order as integers).
__global__ __launch_bounds__(ROW_THREADS) void k_reciprocity(const int64_t *c_rowptr, const int32_t *c_col,
                                                             const int64_t *rowptr, const int32_t *col,
                                                             const double *val, const double *d, const double *area,
                                                             int64_t n, unsigned long long *worst_out,
                                                             unsigned long long *largest_out)
{
    const int64_t i = (int64_t)blockIdx.x * ROWS_PER_BLOCK + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    double worst = 0.0, largest = 0.0;
    if (i < n) {
        const int64_t s0 = rowptr[i], s1 = rowptr[i + 1];
        for (int64_t e = c_rowptr[i] + lane; e < c_rowptr[i + 1]; e += 64) {
            const int32_t j = c_col[e];
            const int64_t ij = find_col(col, s0, s1, j);
            const int64_t ji = find_col(col, rowptr[j], rowptr[j + 1], (int32_t)i);
            const double s_ij = ij >= 0 ? val[ij] : 0.0, s_ji = ji >= 0 ? val[ji] : 0.0;
            const double joint_ij = area[i] * ((d[i] / area[i]) * s_ij * d[j]);
            const double joint_ji = area[j] * ((d[j] / area[j]) * s_ji * d[i]);
            worst = fmax(worst, fabs(joint_ij - joint_ji));
            largest = fmax(largest, fabs(joint_ij));
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        worst = fmax(worst, __shfl_xor(worst, off));
        largest = fmax(largest, __shfl_xor(largest, off));
    }
    if (lane == 0 && i < n) {
        atomicMax(worst_out, (unsigned long long)__double_as_longlong(worst));
        atomicMax(largest_out, (unsigned long long)__double_as_longlong(largest));
    }
}

unsigned blocks_for(int64_t n, int per_block)
{
    int64_t b = (n + per_block - 1) / per_block;
    return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace
}  // namespace prh

using namespace prh;

namespace {

void free_operator(prh_operator *op)
{
    void *ptrs[] = {op->rowptr, op->col, op->val, op->c_rowptr, op->t_rowptr, op->c_col, op->c_cnt,
                    op->t_col, op->t_cnt, op->rays, op->area, op->scale, op->vx, op->vy, op->vt,
                    op->flight, op->absorbed, op->w0, op->w1, op->eps, op->emit, op->part_abs,
                    op->part_sum, op->state, op->kval, op->val32, op->scale32, op->mix32, op->mix64,
                    op->row_order};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : op->series_ev)
        if (e) (void)hipEventDestroy(e);
}

// Host assembly of S over union(C, C^T, diagonal), rows merged in parallel.
struct HostCSR {
    std::vector<int64_t> c_rowptr, t_rowptr, s_rowptr;
    std::vector<int32_t> t_col, t_cnt, s_col;
    std::vector<double> s_val;
    std::vector<int64_t> s_k;  // K_ij = C_ij + C_ji on the sparsity of S
};

void assemble(int64_t n, const int32_t *origin, const int32_t *receiver, const int64_t *count, int64_t nnz,
              const int64_t *rays, const double *area, double delta, HostCSR &H)
{
    H.c_rowptr.assign(n + 1, 0);
    H.t_rowptr.assign(n + 1, 0);
    for (int64_t e = 0; e < nnz; ++e) {
        H.c_rowptr[origin[e] + 1]++;
        H.t_rowptr[receiver[e] + 1]++;
    }
    for (int64_t i = 0; i < n; ++i) {
        H.c_rowptr[i + 1] += H.c_rowptr[i];
        H.t_rowptr[i + 1] += H.t_rowptr[i];
    }
    // Stable counting sort by receiver: rows of C^T come out sorted by origin.
    H.t_col.resize(nnz);
    H.t_cnt.resize(nnz);
    {
        std::vector<int64_t> cursor(H.t_rowptr.begin(), H.t_rowptr.end() - 1);
        for (int64_t e = 0; e < nnz; ++e) {
            const int64_t at = cursor[receiver[e]]++;
            H.t_col[at] = origin[e];
            H.t_cnt[at] = (int32_t)count[e];
        }
    }
    // Row lengths of the union, then values.
    const int n_threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<int64_t> len(n, 0);
    auto row_len = [&](int64_t i) {
        int64_t a = H.c_rowptr[i], ae = H.c_rowptr[i + 1], b = H.t_rowptr[i], be = H.t_rowptr[i + 1];
        int64_t m = 0;
        bool diag = false;
        while (a < ae || b < be) {
            int32_t ja = a < ae ? receiver[a] : INT32_MAX, jb = b < be ? H.t_col[b] : INT32_MAX;
            int32_t j = std::min(ja, jb);
            if (ja == j) ++a;
            if (jb == j) ++b;
            diag |= j == i;
            ++m;
        }
        return m + (diag ? 0 : 1);
    };
    auto par = [&](auto fn) {
        std::vector<std::thread> th;
        for (int t = 0; t < n_threads; ++t)
            th.emplace_back([&, t]() {
                for (int64_t i = (n * t) / n_threads; i < (n * (t + 1)) / n_threads; ++i) fn(i);
            });
        for (auto &x : th) x.join();
    };
    par([&](int64_t i) { len[i] = row_len(i); });
    H.s_rowptr.assign(n + 1, 0);
    for (int64_t i = 0; i < n; ++i) H.s_rowptr[i + 1] = H.s_rowptr[i] + len[i];
    const int64_t nnz_s = H.s_rowptr[n];
    H.s_col.resize(nnz_s);
    H.s_val.resize(nnz_s);
    H.s_k.resize(nnz_s);
    par([&](int64_t i) {
        int64_t a = H.c_rowptr[i], ae = H.c_rowptr[i + 1], b = H.t_rowptr[i], be = H.t_rowptr[i + 1];
        int64_t out = H.s_rowptr[i];
        bool diag_done = false;
        const double Ai = area[i], Ri = (double)rays[i];
        auto emit_entry = [&](int32_t j, int64_t cij, int64_t cji) {
            // 0.5 (A_i F_ij + A_j F_ji) + delta A_i [i == j]
            const double fij = (double)cij / Ri;
            const double fji = (double)cji / (double)rays[j];
            double v = 0.5 * (Ai * fij + area[j] * fji);
            if (j == i) v += delta * Ai;
            H.s_col[out] = j;
            H.s_val[out] = v;
            H.s_k[out] = cij + cji;
            ++out;
        };
        while (a < ae || b < be) {
            int32_t ja = a < ae ? receiver[a] : INT32_MAX, jb = b < be ? H.t_col[b] : INT32_MAX;
            int32_t j = std::min(ja, jb);
            if (!diag_done && i < j) {
                emit_entry((int32_t)i, 0, 0);
                diag_done = true;
            }
            int64_t cij = 0, cji = 0;
            if (ja == j) cij = count[a++];
            if (jb == j) cji = H.t_cnt[b++];
            if (j == i) diag_done = true;
            emit_entry(j, cij, cji);
        }
        if (!diag_done) emit_entry((int32_t)i, 0, 0);
    });
}

template <typename T>
int upload(T **dst, const T *src, int64_t count, hipStream_t st)
{
    const int64_t c = count > 0 ? count : 1;
    PRH_HIP(hipMalloc((void **)dst, sizeof(T) * c));
    if (count > 0) PRH_HIP(hipMemcpyAsync(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice, st));
    return PRH_OK;
}

int apply_device(prh_operator *op, int mode, const double *dx, double *dy, hipStream_t st)
{
    const int64_t n = op->n;
    const unsigned rb = blocks_for(n, ROWS_PER_BLOCK), eb = blocks_for(n, 256);
    switch (mode) {
    case PRH_APPLY_SYMMETRIC:
        if (op->uniform) k_spmv<true><<<rb, ROW_THREADS, 0, st>>>(sview_of(op), n, dx, dy);
        else k_spmv<false><<<rb, ROW_THREADS, 0, st>>>(sview_of(op), n, dx, dy);
        break;
    case PRH_APPLY_MATVEC:
    case PRH_APPLY_RMATVEC:
        k_scale_in<<<eb, 256, 0, st>>>(dx, op->scale, op->area, mode, n, op->vt);
        if (op->uniform) k_spmv<true><<<rb, ROW_THREADS, 0, st>>>(sview_of(op), n, op->vt, dy);
        else k_spmv<false><<<rb, ROW_THREADS, 0, st>>>(sview_of(op), n, op->vt, dy);
        k_scale_out<<<eb, 256, 0, st>>>(dy, op->scale, op->area, mode, n, dy);
        break;
    case PRH_APPLY_RAW:
        k_spmv_counts<<<rb, ROW_THREADS, 0, st>>>(op->c_rowptr, op->c_col, op->c_cnt, op->rays, 0, n, dx, dy);
        break;
    case PRH_APPLY_RAW_T:
        k_spmv_counts<<<rb, ROW_THREADS, 0, st>>>(op->t_rowptr, op->t_col, op->t_cnt, op->rays, 1, n, dx, dy);
        break;
    default:
        set_error("prh_operator_apply: unknown mode %d", mode);
        return PRH_ERR_ARG;
    }
    PRH_HIP(hipGetLastError());
    return PRH_OK;
}

}  // namespace

namespace prh {

int operator_alloc_work(prh_operator *op)
{
    const int64_t n = op->n;
    double **work[] = {&op->vx, &op->vy, &op->vt, &op->flight, &op->absorbed, &op->w0, &op->w1, &op->eps, &op->emit};
    for (double **p : work)
        if (hipMalloc((void **)p, sizeof(double) * n) != hipSuccess) {
            set_error("prh_operator_create: out of device memory");
            return PRH_ERR_HIP;
        }
    const int64_t n_part = blocks_for(n, ROWS_PER_BLOCK);
    if (hipMalloc((void **)&op->part_abs, sizeof(double) * n_part) != hipSuccess ||
        hipMalloc((void **)&op->part_sum, sizeof(double) * n_part) != hipSuccess ||
        hipMalloc((void **)&op->state, sizeof(SeriesState)) != hipSuccess) {
        set_error("prh_operator_create: out of device memory");
        return PRH_ERR_HIP;
    }
    return PRH_OK;
}

void operator_free(prh_operator *op) { free_operator(op); }

int series_fp64(prh_operator *op, double relative_tolerance, int32_t max_reflections, hipStream_t st,
                double *residual, int32_t *iterations)
{
    const int64_t n = op->n;
    const unsigned rb = blocks_for(n, ROWS_PER_BLOCK);
    const int64_t n_part = rb;
    const unsigned ib = (unsigned)std::min<int64_t>(blocks_for(n, RED_THREADS), 1024);
    if (!op->series_ev[0]) {
        PRH_HIP(hipEventCreate(&op->series_ev[0]));
        PRH_HIP(hipEventCreate(&op->series_ev[1]));
    }
    PRH_HIP(hipEventRecord(op->series_ev[0], st));
    k_series_init<<<ib, RED_THREADS, 0, st>>>(op->emit, op->scale, op->area, n, op->flight, op->absorbed, op->w0,
                                              op->part_abs);
    k_series_start<<<1, RED_THREADS, 0, st>>>(op->part_abs, (int)ib, op->state);
    PRH_HIP(hipGetLastError());
    int32_t it = 0;
    SeriesState hs{};
    double *w_in = op->w0, *w_out = op->w1;
    while (it < max_reflections) {
        const int batch = std::min<int32_t>(16, max_reflections - it);
        for (int b = 0; b < batch; ++b) {
            auto step = op->uniform ? k_series_step<true> : k_series_step<false>;
            step<<<rb, ROW_THREADS, 0, st>>>(sview_of(op), n, op->scale, op->area, op->eps, w_in, w_out,
                                             op->absorbed, op->flight, op->part_abs, op->part_sum, op->state);
            k_series_check<<<1, RED_THREADS, 0, st>>>(op->part_abs, op->part_sum, n_part, relative_tolerance,
                                                      op->state);
            std::swap(w_in, w_out);
        }
        PRH_HIP(hipGetLastError());
        it += batch;
        PRH_HIP(hipEventRecord(op->series_ev[1], st));
        PRH_HIP(hipMemcpyAsync(&hs, op->state, sizeof(hs), hipMemcpyDeviceToHost, st));
        PRH_HIP(hipStreamSynchronize(st));
        if (hs.done) break;
    }
    {
        // steps launched after convergence return at once; count the reflections made
        float ms = 0.f;
        PRH_HIP(hipEventElapsedTime(&ms, op->series_ev[0], op->series_ev[1]));
        op->series_ms += ms;
        op->series_count += 1;
        op->series_steps += hs.iterations;
    }
    if (!hs.done) {
        set_error("radiosity reflection series did not converge");
        return PRH_ERR_SERIES;
    }
    *residual = hs.residual;
    *iterations = hs.iterations;
    return PRH_OK;
}

int symmetric_product(const prh_operator *op, const double *x, double *y, hipStream_t st)
{
    const unsigned rb = blocks_for(op->n, ROWS_PER_BLOCK);
    if (op->uniform) k_spmv<true><<<rb, ROW_THREADS, 0, st>>>(sview_of(op), op->n, x, y);
    else k_spmv<false><<<rb, ROW_THREADS, 0, st>>>(sview_of(op), op->n, x, y);
    PRH_HIP(hipGetLastError());
    return PRH_OK;
}

int mixed_solve_single(prh_operator *op, const double *E, double emissivity, double rtol, hipStream_t st,
                       double *absorbed, int32_t *iterations)
{
    PRH_CHECK_ARG(op && E && absorbed, "mixed_solve_single: null pointer");
    PRH_CHECK_ARG(emissivity > 0.0 && emissivity <= 1.0 && rtol > 0.0, "mixed_solve_single: bad emissivity or tolerance");
    const int64_t n = op->n;
    if (!op->mix32 || !op->mix64) {
        // both buffers or neither: the handle never keeps half a pair
        float *m32 = nullptr;
        double *m64 = nullptr;
        if (hipMalloc((void **)&m32, sizeof(float) * 5 * n) != hipSuccess ||
            hipMalloc((void **)&m64, sizeof(double) * (3 * n + 1)) != hipSuccess) {
            if (m32) (void)hipFree(m32);
            (void)hipGetLastError();
            set_error("mixed_solve_single: out of device memory");
            return PRH_ERR_HIP;
        }
        if (op->mix32) (void)hipFree(op->mix32);
        if (op->mix64) (void)hipFree(op->mix64);
        op->mix32 = m32;
        op->mix64 = m64;
    }
    const Mix1 V{op->mix32, op->mix32 + n, op->mix32 + 2 * n, op->mix64, op->mix64 + n, op->mix64 + 2 * n};
    float *z[2] = {op->mix32 + 3 * n, op->mix32 + 4 * n};
    double *total = op->mix64 + 3 * n;
    const unsigned rb = blocks_for(n, ROWS_PER_BLOCK), eb = blocks_for(n, 256);
    const unsigned ib = (unsigned)std::min<int64_t>(blocks_for(n, RED_THREADS), 1024);
    const int uniform = op->uniform ? 1 : 0;
    double e_norm = 0.0;
    k_mix1_start<<<ib, RED_THREADS, 0, st>>>(E, n, V, op->part_abs);
    k_mix1_total<<<1, RED_THREADS, 0, st>>>(op->part_abs, (int64_t)ib, total);
    PRH_HIP(hipGetLastError());
    PRH_HIP(hipMemcpyAsync(&e_norm, total, sizeof(double), hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));
    if (!(e_norm > 0.0)) {
        k_mix1_collect<<<eb, 256, 0, st>>>(n, emissivity, nullptr, absorbed);  // the zero emission
        PRH_HIP(hipGetLastError());
        if (iterations) *iterations = 0;
        return PRH_OK;
    }
    const double rho = 1.0 - emissivity;
    const float rho32 = (float)rho;
    double s = e_norm / (double)n, r_norm = e_norm;
    int m = ticks_for(rho, epoch_target(rtol, 1.0, 1.0));
    int sweeps = 0, epochs = 0;
    auto sweep = op->uniform ? k_mix1_sweep<true> : k_mix1_sweep<false>;
    k_mix1_epoch<<<eb, 256, 0, st>>>(n, op->scale32, op->area, uniform, V, 1.0 / s, z[0]);
    while (true) {
        int in = 0;
        for (int k = 0; k < m; ++k) {
            float alpha = 0.f, beta = 1.f;
            const bool last = k == m - 1;
            if (!last) cheb_coefficient(rho, k, &alpha, &beta);
            sweep<<<rb, ROW_THREADS, 0, st>>>(sview_of(op), n, op->scale, op->scale32, V, z[in], z[1 - in],
                                              last ? MIX_LAST : MIX_CHEB, alpha, beta, rho32, rho, s, E,
                                              op->part_abs);
            in = 1 - in;
            ++sweeps;
        }
        // the refinement gathers the epoch's correction x32 (left in z[in] by the last sweep)
        sweep<<<rb, ROW_THREADS, 0, st>>>(sview_of(op), n, op->scale, op->scale32, V, z[in], z[1 - in], MIX_REFINE,
                                          0.f, 1.f, rho32, rho, s, E, op->part_abs);
        k_mix1_total<<<1, RED_THREADS, 0, st>>>(op->part_abs, (int64_t)rb, total);
        PRH_HIP(hipGetLastError());
        PRH_HIP(hipMemcpyAsync(&r_norm, total, sizeof(double), hipMemcpyDeviceToHost, st));
        PRH_HIP(hipStreamSynchronize(st));
        ++epochs;
        if (r_norm <= rtol * e_norm) break;
        if (epochs >= MAX_EPOCHS || !(r_norm < HUGE_VAL)) {
            set_error("radiosity reflection series did not converge");
            return PRH_ERR_SERIES;
        }
        // the next epoch works on the residual scaled to mean magnitude 1
        m = ticks_for(rho, epoch_target(rtol, e_norm, r_norm));
        s = r_norm / (double)n;
        k_mix1_epoch<<<eb, 256, 0, st>>>(n, op->scale32, op->area, uniform, V, 1.0 / s, z[0]);
    }
    k_mix1_collect<<<eb, 256, 0, st>>>(n, emissivity, V.y64, absorbed);
    PRH_HIP(hipGetLastError());
    if (iterations) *iterations = sweeps + epochs;
    return PRH_OK;
}

}  // namespace prh

extern "C" {

int prh_operator_create(int64_t n, const int32_t *origin, const int32_t *receiver, const int64_t *count,
                        int64_t nnz, const int64_t *rays_per_origin, const double *area,
                        double support_diagonal_fraction, prh_operator **out)
{
    PRH_CHECK_ARG(out && rays_per_origin && area && (nnz == 0 || (origin && receiver && count)),
                  "prh_operator_create: null pointer");
    PRH_CHECK_ARG(n > 0 && n < INT32_MAX && nnz >= 0, "prh_operator_create: bad sizes");
    *out = nullptr;
    // Validate order, ranges and row totals (exchange.py:59-71).
    {
        std::vector<int64_t> observed(n, 0);
        for (int64_t e = 0; e < nnz; ++e) {
            const int32_t o = origin[e], r = receiver[e];
            PRH_CHECK_ARG(o >= 0 && o < n && r >= 0 && r < n && count[e] >= 0,
                          "prh_operator_create: edge %lld out of range", (long long)e);
            if (e > 0)
                PRH_CHECK_ARG(origin[e - 1] < o || (origin[e - 1] == o && receiver[e - 1] < r),
                              "prh_operator_create: edges must be sorted by (origin, receiver) and unique");
            observed[o] += count[e];
        }
        for (int64_t i = 0; i < n; ++i) {
            if (observed[i] != rays_per_origin[i]) {
                set_error("first-hit counts do not account for every emitted ray");
                return PRH_ERR_COUNTS;
            }
            PRH_CHECK_ARG(rays_per_origin[i] > 0, "prh_operator_create: rays_per_origin must be positive");
        }
    }
    HostCSR H;
    assemble(n, origin, receiver, count, nnz, rays_per_origin, area, support_diagonal_fraction, H);

    prh_operator *op = new prh_operator();
    PRH_HIP(hipGetDevice(&op->device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) {
        delete op;
        return rc;
    }
    op->n = n;
    op->nnz_s = H.s_rowptr[n];
    op->nnz_c = nnz;
    op->delta = support_diagonal_fraction;
    std::vector<int32_t> c_cnt(nnz);
    for (int64_t e = 0; e < nnz; ++e) c_cnt[e] = (int32_t)count[e];
    std::vector<double> rays_d(n), ones(n, 1.0);
    for (int64_t i = 0; i < n; ++i) rays_d[i] = (double)rays_per_origin[i];
    auto fail = [&](int code) {
        (void)hipStreamSynchronize(st);
        free_operator(op);
        delete op;
        return code;
    };
    if ((rc = upload(&op->rowptr, H.s_rowptr.data(), n + 1, st))) return fail(rc);
    if ((rc = upload(&op->col, H.s_col.data(), op->nnz_s, st))) return fail(rc);
    if ((rc = upload(&op->val, H.s_val.data(), op->nnz_s, st))) return fail(rc);
    if ((rc = upload(&op->c_rowptr, H.c_rowptr.data(), n + 1, st))) return fail(rc);
    if ((rc = upload(&op->c_col, receiver, nnz, st))) return fail(rc);
    if ((rc = upload(&op->c_cnt, c_cnt.data(), nnz, st))) return fail(rc);
    if ((rc = upload(&op->t_rowptr, H.t_rowptr.data(), n + 1, st))) return fail(rc);
    if ((rc = upload(&op->t_col, H.t_col.data(), nnz, st))) return fail(rc);
    if ((rc = upload(&op->t_cnt, H.t_cnt.data(), nnz, st))) return fail(rc);
    if ((rc = upload(&op->rays, rays_d.data(), n, st))) return fail(rc);
    if ((rc = upload(&op->area, area, n, st))) return fail(rc);
    if ((rc = upload(&op->scale, ones.data(), n, st))) return fail(rc);
    // Compact forms for the batched engine: exact integer K when every patch
    // has the same area and the same ray count (S = A/(2R) K + A delta I),
    // otherwise an FP32 copy of S.
    {
        bool uniform = true;
        for (int64_t i = 1; i < n && uniform; ++i)
            uniform = area[i] == area[0] && rays_per_origin[i] == rays_per_origin[0];
        int64_t kmax = 0;
        for (int64_t e = 0; e < op->nnz_s; ++e) kmax = std::max(kmax, H.s_k[e]);
        op->uniform = uniform && kmax <= 65535;
        if (op->uniform) {
            op->inv_2R = 1.0 / (2.0 * (double)rays_per_origin[0]);
            std::vector<uint16_t> k16(op->nnz_s);
            for (int64_t e = 0; e < op->nnz_s; ++e) k16[e] = (uint16_t)H.s_k[e];
            if ((rc = upload(&op->kval, k16.data(), op->nnz_s, st))) return fail(rc);
            if (hipStreamSynchronize(st) != hipSuccess) return fail(PRH_ERR_HIP);
        } else {
            std::vector<float> v32(op->nnz_s);
            for (int64_t e = 0; e < op->nnz_s; ++e) v32[e] = (float)H.s_val[e];
            if ((rc = upload(&op->val32, v32.data(), op->nnz_s, st))) return fail(rc);
            if (hipStreamSynchronize(st) != hipSuccess) return fail(PRH_ERR_HIP);
        }
        std::vector<float> ones32(n, 1.0f);
        if ((rc = upload(&op->scale32, ones32.data(), n, st))) return fail(rc);
        if (hipStreamSynchronize(st) != hipSuccess) return fail(PRH_ERR_HIP);
    }
    if ((rc = operator_alloc_work(op))) return fail(rc);
    if (hipStreamSynchronize(st) != hipSuccess) {
        set_error("prh_operator_create: upload failed");
        return fail(PRH_ERR_HIP);
    }
    *out = op;
    return PRH_OK;
}

int prh_operator_nnz(const prh_operator *op, int64_t *nnz_symmetric, int64_t *nnz_raw)
{
    PRH_CHECK_ARG(op, "prh_operator_nnz: null operator");
    if (nnz_symmetric) *nnz_symmetric = op->nnz_s;
    if (nnz_raw) *nnz_raw = op->nnz_c;
    return PRH_OK;
}

int prh_operator_csr(const prh_operator *op, int64_t *rowptr, int32_t *col, double *val)
{
    PRH_CHECK_ARG(op, "prh_operator_csr: null operator");
    PRH_HIP(hipSetDevice(op->device));
    if (rowptr) PRH_HIP(hipMemcpy(rowptr, op->rowptr, sizeof(int64_t) * (op->n + 1), hipMemcpyDeviceToHost));
    if (op->nnz_s == 0) return PRH_OK;
    if (col) PRH_HIP(hipMemcpy(col, op->col, sizeof(int32_t) * op->nnz_s, hipMemcpyDeviceToHost));
    if (val) PRH_HIP(hipMemcpy(val, op->val, sizeof(double) * op->nnz_s, hipMemcpyDeviceToHost));
    return PRH_OK;
}

int prh_operator_set_scale(prh_operator *op, const double *scale)
{
    PRH_CHECK_ARG(op && scale, "prh_operator_set_scale: null pointer");
    PRH_HIP(hipSetDevice(op->device));
    PRH_HIP(hipMemcpy(op->scale, scale, sizeof(double) * op->n, hipMemcpyHostToDevice));
    std::vector<float> s32(op->n);
    for (int64_t i = 0; i < op->n; ++i) s32[i] = (float)scale[i];
    PRH_HIP(hipMemcpy(op->scale32, s32.data(), sizeof(float) * op->n, hipMemcpyHostToDevice));
    return PRH_OK;
}

int prh_operator_set_row_order(prh_operator *op, const int32_t *order)
{
    PRH_CHECK_ARG(op, "prh_operator_set_row_order: null operator");
    PRH_HIP(hipSetDevice(op->device));
    {
        // sweeps already queued read the current order: let them finish first
        hipStream_t st;
        int rc = current_stream(&st);
        if (rc) return rc;
        PRH_HIP(hipStreamSynchronize(st));
    }
    if (!order) {
        if (op->row_order) (void)hipFree(op->row_order);
        op->row_order = nullptr;
        return PRH_OK;
    }
    std::vector<uint8_t> seen(op->n, 0);
    for (int64_t p = 0; p < op->n; ++p) {
        PRH_CHECK_ARG(order[p] >= 0 && order[p] < op->n && !seen[order[p]],
                      "prh_operator_set_row_order: not a permutation of the patches");
        seen[order[p]] = 1;
    }
    if (!op->row_order) PRH_HIP(hipMalloc((void **)&op->row_order, sizeof(int32_t) * std::max<int64_t>(op->n, 1)));
    PRH_HIP(hipMemcpy(op->row_order, order, sizeof(int32_t) * op->n, hipMemcpyHostToDevice));
    return PRH_OK;
}

int prh_operator_reciprocity(prh_operator *op, double *error)
{
    PRH_CHECK_ARG(op && error, "prh_operator_reciprocity: null pointer");
    PRH_HIP(hipSetDevice(op->device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    *error = 0.0;
    if (op->n == 0) return PRH_OK;
    // the first word of two work vectors holds the two maxima
    unsigned long long *worst = reinterpret_cast<unsigned long long *>(op->vx);
    unsigned long long *largest = reinterpret_cast<unsigned long long *>(op->vy);
    PRH_HIP(hipMemsetAsync(worst, 0, sizeof(unsigned long long), st));
    PRH_HIP(hipMemsetAsync(largest, 0, sizeof(unsigned long long), st));
    k_reciprocity<<<blocks_for(op->n, ROWS_PER_BLOCK), ROW_THREADS, 0, st>>>(
        op->c_rowptr, op->c_col, op->rowptr, op->col, op->val, op->scale, op->area, op->n, worst, largest);
    PRH_HIP(hipGetLastError());
    unsigned long long host[2] = {0, 0};
    PRH_HIP(hipMemcpyAsync(&host[0], worst, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    PRH_HIP(hipMemcpyAsync(&host[1], largest, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));
    double worst_v, largest_v;
    memcpy(&worst_v, &host[0], sizeof(double));
    memcpy(&largest_v, &host[1], sizeof(double));
    *error = worst_v / fmax(largest_v, 2.2250738585072014e-308);
    return PRH_OK;
}

int prh_operator_apply(prh_operator *op, int32_t mode, const double *x, double *y)
{
    PRH_CHECK_ARG(op && x && y, "prh_operator_apply: null pointer");
    PRH_HIP(hipSetDevice(op->device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    PRH_HIP(hipMemcpyAsync(op->vx, x, sizeof(double) * op->n, hipMemcpyHostToDevice, st));
    if ((rc = apply_device(op, mode, op->vx, op->vy, st))) return rc;
    PRH_HIP(hipMemcpyAsync(y, op->vy, sizeof(double) * op->n, hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));
    return PRH_OK;
}

namespace {

// The FP64 series of n_rhs emissions in groups of up to 32 slots; a group of
// at most 16 emissions runs on the 16-slot layout.
int radiosity_fp64_batched(prh_operator *op, int32_t n_rhs, const double *emitted, const double *emissivity,
                           int32_t emissivity_per_rhs, double rtol, int32_t max_reflections, hipStream_t st,
                           double *absorbed, double *residual, int32_t *iterations)
{
    const int64_t n = op->n;
    if (n_rhs == 0) return PRH_OK;
    const uint32_t bpx = (uint32_t)((blocks_for(n, ROWS_PER_BLOCK) + 7) / 8);
    const unsigned rb = 8u * bpx;  // every XCD the same number of blocks
    const size_t vec = (size_t)n * FSLOTS_MAX;
    std::vector<double> d(n), area(n), host(vec), host_eps(vec), host_w(vec), host_fl(vec);
    PRH_HIP(hipMemcpy(d.data(), op->scale, sizeof(double) * n, hipMemcpyDeviceToHost));
    PRH_HIP(hipMemcpy(area.data(), op->area, sizeof(double) * n, hipMemcpyDeviceToHost));
    double *buf = nullptr;
    SeriesSlots *S = nullptr;
    const size_t bytes = sizeof(double) * (5 * vec + (size_t)FSLOTS_MAX * rb);
    if (hipMalloc((void **)&buf, bytes) != hipSuccess || hipMalloc((void **)&S, sizeof(SeriesSlots)) != hipSuccess) {
        if (buf) (void)hipFree(buf);
        (void)hipGetLastError();
        set_error("prh_radiosity: out of device memory for the batched FP64 series");
        return PRH_ERR_HIP;
    }
    struct Free {
        double *b;
        SeriesSlots *s;
        ~Free()
        {
            (void)hipFree(b);
            (void)hipFree(s);
        }
    } release{buf, S};
    double *w0 = buf, *w1 = buf + vec, *abs_d = buf + 2 * vec, *fl_d = buf + 3 * vec, *eps_d = buf + 4 * vec;
    double *part = buf + 5 * vec;
    const SView M = sview_of(op);
    for (int32_t q0 = 0; q0 < n_rhs; q0 += FSLOTS_MAX) {
        const int active = std::min<int32_t>(FSLOTS_MAX, n_rhs - q0);
        const int FS = active <= 16 ? 16 : 32;
        const size_t v = (size_t)n * FS;
        SeriesSlots hs{};
        std::fill(host.begin(), host.begin() + v, 0.0);
        std::fill(host_eps.begin(), host_eps.begin() + v, 0.0);
        std::fill(host_w.begin(), host_w.begin() + v, 0.0);
        for (int q = 0; q < FSLOTS_MAX; ++q) {
            hs.done[q] = q < active ? 0 : 1;
            if (q >= active) continue;
            const double *E = emitted + (int64_t)(q0 + q) * n;
            const double *ep = emissivity + (emissivity_per_rhs ? (int64_t)(q0 + q) * n : 0);
            double l1 = 0.0;
            for (int64_t i = 0; i < n; ++i) {
                host[(size_t)i * FS + q] = E[i];
                host_eps[(size_t)i * FS + q] = ep[i];
                host_w[(size_t)i * FS + q] = d[i] * (E[i] / area[i]);
                l1 += fabs(E[i]);
            }
            hs.initial[q] = l1 > 2.2250738585072014e-308 ? l1 : 2.2250738585072014e-308;
        }
        PRH_HIP(hipMemcpyAsync(fl_d, host.data(), sizeof(double) * v, hipMemcpyHostToDevice, st));
        PRH_HIP(hipMemcpyAsync(eps_d, host_eps.data(), sizeof(double) * v, hipMemcpyHostToDevice, st));
        PRH_HIP(hipMemcpyAsync(w0, host_w.data(), sizeof(double) * v, hipMemcpyHostToDevice, st));
        PRH_HIP(hipMemsetAsync(abs_d, 0, sizeof(double) * v, st));
        PRH_HIP(hipMemcpyAsync(S, &hs, sizeof(hs), hipMemcpyHostToDevice, st));
        if (!op->series_ev[0]) {
            PRH_HIP(hipEventCreate(&op->series_ev[0]));
            PRH_HIP(hipEventCreate(&op->series_ev[1]));
        }
        PRH_HIP(hipEventRecord(op->series_ev[0], st));
        double *w_in = w0, *w_out = w1;
        int32_t ticks = 0;
        bool all_done = false;
        while (!all_done && ticks < max_reflections) {
            const int batch = std::min<int32_t>(8, max_reflections - ticks);
            for (int b = 0; b < batch; ++b) {
                auto step = FS == 16 ? (op->uniform ? k_series_batch<16, true> : k_series_batch<16, false>)
                                     : (op->uniform ? k_series_batch<32, true> : k_series_batch<32, false>);
                step<<<rb, ROW_THREADS, 0, st>>>(M, n, op->scale, op->area, eps_d, w_in, w_out, abs_d, fl_d, part,
                                                 bpx, op->row_order, S);
                k_series_batch_check<<<FS, RED_THREADS, 0, st>>>(part, (int64_t)rb, rtol, active, S);
                std::swap(w_in, w_out);
            }
            PRH_HIP(hipGetLastError());
            ticks += batch;
            PRH_HIP(hipEventRecord(op->series_ev[1], st));
            PRH_HIP(hipMemcpyAsync(&hs, S, sizeof(hs), hipMemcpyDeviceToHost, st));
            PRH_HIP(hipStreamSynchronize(st));
            all_done = true;
            for (int q = 0; q < active; ++q) all_done = all_done && hs.done[q];
        }
        float ms = 0.f;
        PRH_HIP(hipEventElapsedTime(&ms, op->series_ev[0], op->series_ev[1]));
        op->batched_ms += ms;
        op->batched_passes += ticks;
        op->batched_slot_passes += (int64_t)ticks * FS;
        if (!all_done) {
            set_error("radiosity reflection series did not converge");
            return PRH_ERR_SERIES;
        }
        PRH_HIP(hipMemcpyAsync(host.data(), abs_d, sizeof(double) * v, hipMemcpyDeviceToHost, st));
        PRH_HIP(hipMemcpyAsync(host_fl.data(), fl_d, sizeof(double) * v, hipMemcpyDeviceToHost, st));
        PRH_HIP(hipStreamSynchronize(st));
        for (int q = 0; q < active; ++q) {
            double *out = absorbed + (int64_t)(q0 + q) * n;
            double sum = 0.0;
            for (int64_t i = 0; i < n; ++i) {
                out[i] = host[(size_t)i * FS + q];
                sum += host_fl[(size_t)i * FS + q];
            }
            residual[q0 + q] = sum;
            iterations[q0 + q] = hs.iterations[q];
        }
    }
    return PRH_OK;
}

}  // namespace

int prh_radiosity(prh_operator *op, int32_t n_rhs, const double *emitted, const double *emissivity,
                  int32_t emissivity_per_rhs, double relative_tolerance, int32_t max_reflections,
                  int32_t precision, double *absorbed, double *residual, int32_t *iterations)
{
    PRH_CHECK_ARG(op && emitted && emissivity && absorbed && residual && iterations,
                  "prh_radiosity: null pointer");
    PRH_CHECK_ARG(n_rhs >= 0 && max_reflections >= 1, "prh_radiosity: bad sizes");
    PRH_CHECK_ARG(precision == PRH_FP64 || precision == PRH_MIXED || precision == PRH_FP64_BATCHED,
                  "prh_radiosity: unsupported precision %d", precision);
    PRH_HIP(hipSetDevice(op->device));
    if (precision == PRH_MIXED)
        return prh_radiosity_mixed(op, n_rhs, emitted, emissivity, emissivity_per_rhs, relative_tolerance, absorbed,
                                   residual, iterations);
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    if (precision == PRH_FP64_BATCHED)
        return radiosity_fp64_batched(op, n_rhs, emitted, emissivity, emissivity_per_rhs, relative_tolerance,
                                      max_reflections, st, absorbed, residual, iterations);
    const int64_t n = op->n;
    for (int32_t q = 0; q < n_rhs; ++q) {
        PRH_HIP(hipMemcpyAsync(op->emit, emitted + (int64_t)q * n, sizeof(double) * n, hipMemcpyHostToDevice, st));
        if (q == 0 || emissivity_per_rhs)
            PRH_HIP(hipMemcpyAsync(op->eps, emissivity + (emissivity_per_rhs ? (int64_t)q * n : 0),
                                   sizeof(double) * n, hipMemcpyHostToDevice, st));
        double res = 0.0;
        int32_t its = 0;
        if ((rc = prh::series_fp64(op, relative_tolerance, max_reflections, st, &res, &its))) return rc;
        PRH_HIP(hipMemcpyAsync(absorbed + (int64_t)q * n, op->absorbed, sizeof(double) * n, hipMemcpyDeviceToHost,
                               st));
        PRH_HIP(hipStreamSynchronize(st));
        residual[q] = res;
        iterations[q] = its;
    }
    return PRH_OK;
}

int prh_operator_stats(const prh_operator *op, double *out, int32_t n)
{
    PRH_CHECK_ARG(op && out && n >= 0, "prh_operator_stats: null pointer");
    const double v[] = {(double)op->series_count,  (double)op->series_steps,        op->series_ms,
                        (double)op->nnz_s,         (double)op->n,                   op->uniform ? 1.0 : 0.0,
                        (double)op->batched_passes, (double)op->batched_slot_passes, op->batched_ms};
    for (int i = 0; i < n && i < (int)(sizeof(v) / sizeof(v[0])); ++i) out[i] = v[i];
    return PRH_OK;
}

void prh_operator_destroy(prh_operator *op)
{
    if (!op) return;
    (void)hipSetDevice(op->device);
    if (op->batch) prh_batch_destroy(op->batch);
    free_operator(op);
    delete op;
}

}  // extern "C"
