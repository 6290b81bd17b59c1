// Device side of the strictly coupled radiation-conduction solve.
//
// The Newton-Krylov iteration of porerad/coupled.py:232-447 runs on cell
// vectors of the whole grid (n_cells = nx ny nz doubles).  This module keeps
// those vectors in HBM (a pool of slots addressed by index) together with
// the conduction matrix A (CSR as assembled by porerad/conduction.py:57-97),
// and provides the operations the iteration is made of:
//   - vector algebra with fixed-order reductions (results do not depend on
//     the launch, so a solve is reproducible run to run);
//   - y = A x with each row summed in column order, as a CSR product does;
//   - the radiation of a cell field: patch temperatures (walls fixed),
//     emission eps sigma A T^4, the FP64 reflection series on the device
//     exchange operator, and the net power of every internal patch gathered
//     into its owning cell in patch order (coupled.py:45-105);
//   - the Jacobian-vector product A v - dQ/dT v at fixed patch temperatures
//     (coupled.py:108-126, exchange.py:274-287);
//   - the fixed-step Jacobi-PCG preconditioner on A + diag(emission
//     derivative), its scalars kept on the device (coupled.py:292-315);
//   - the two modified Gram-Schmidt passes of one FGMRES column and the
//     linear combinations of basis vectors (coupled.py:232-289).
// The control flow (Newton steps, line search, small least-squares
// problems) stays with the caller.
#include <algorithm>
#include <cstring>
#include <cmath>
#include <vector>

#include "prh_internal.h"
#include "prh_operator_impl.h"

namespace prh {
namespace {

constexpr int CT = 256;          // threads per block
constexpr int RED_BLOCKS = 1024;  // blocks of the first reduction stage
constexpr double SIGMA = 5.670374419e-8;  // porerad/constants.py:3

unsigned blocks_of(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + CT - 1) / CT, 1 << 30)); }
unsigned red_blocks_of(int64_t n) { return (unsigned)std::min<int64_t>(RED_BLOCKS, std::max<int64_t>(1, (n + CT - 1) / CT)); }

__device__ __forceinline__ double block_sum(double v, double *sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = CT / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------- vector algebra
__global__ __launch_bounds__(CT) void k_fill(double *y, int64_t n, double v)
{
    for (int64_t i = (int64_t)blockIdx.x * CT + threadIdx.x; i < n; i += (int64_t)gridDim.x * CT) y[i] = v;
}

// y = a x + b y; with b == 0 the old y is not read (it may hold anything)
__global__ __launch_bounds__(CT) void k_axpby(double *y, const double *x, int64_t n, double a, double b)
{
    for (int64_t i = (int64_t)blockIdx.x * CT + threadIdx.x; i < n; i += (int64_t)gridDim.x * CT)
        y[i] = b == 0.0 ? a * x[i] : a * x[i] + b * y[i];
}

// y = x / d
__global__ __launch_bounds__(CT) void k_divide(double *y, const double *x, int64_t n, double d)
{
    for (int64_t i = (int64_t)blockIdx.x * CT + threadIdx.x; i < n; i += (int64_t)gridDim.x * CT) y[i] = x[i] / d;
}

// y -= s[k] x with the scalar read on the device
__global__ __launch_bounds__(CT) void k_sub_scaled(double *y, const double *x, int64_t n, const double *s, int k)
{
    const double a = s[k];
    for (int64_t i = (int64_t)blockIdx.x * CT + threadIdx.x; i < n; i += (int64_t)gridDim.x * CT) y[i] -= a * x[i];
}

// Block partial sums of x . y (grid-stride, fixed grid).
__global__ __launch_bounds__(CT) void k_dot_part(const double *x, const double *y, int64_t n, double *part)
{
    __shared__ double sh[CT];
    double a = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * CT + threadIdx.x; i < n; i += (int64_t)gridDim.x * CT) a += x[i] * y[i];
    a = block_sum(a, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = a;
}

// out[k] = sum of the n_part partials (one block, fixed order).
__global__ __launch_bounds__(CT) void k_finish_sum(const double *part, int n_part, double *out, int k)
{
    __shared__ double sh[CT];
    double a = 0.0;
    for (int b = threadIdx.x; b < n_part; b += CT) a += part[b];
    a = block_sum(a, sh);
    if (threadIdx.x == 0) out[k] = a;
}

// Partial statistics of x: sum x^2, max |x|, min, max, count of non-finite.
__global__ __launch_bounds__(CT) void k_stats_part(const double *x, int64_t n, double *part)
{
    __shared__ double sh[CT];
    __shared__ double mx[CT], mn[CT], mxa[CT];
    double sq = 0.0, amax = 0.0, lo = INFINITY, hi = -INFINITY, bad = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * CT + threadIdx.x; i < n; i += (int64_t)gridDim.x * CT) {
        const double v = x[i];
        if (!isfinite(v)) {
            bad += 1.0;
            continue;
        }
        sq += v * v;
        amax = fmax(amax, fabs(v));
        lo = fmin(lo, v);
        hi = fmax(hi, v);
    }
    mxa[threadIdx.x] = amax;
    mn[threadIdx.x] = lo;
    mx[threadIdx.x] = hi;
    __syncthreads();
    for (int off = CT / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            mxa[threadIdx.x] = fmax(mxa[threadIdx.x], mxa[threadIdx.x + off]);
            mn[threadIdx.x] = fmin(mn[threadIdx.x], mn[threadIdx.x + off]);
            mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + off]);
        }
        __syncthreads();
    }
    sq = block_sum(sq, sh);
    bad = block_sum(bad, sh);
    if (threadIdx.x == 0) {
        part[5 * blockIdx.x + 0] = sq;
        part[5 * blockIdx.x + 1] = mxa[0];
        part[5 * blockIdx.x + 2] = mn[0];
        part[5 * blockIdx.x + 3] = mx[0];
        part[5 * blockIdx.x + 4] = bad;
    }
}

__global__ __launch_bounds__(CT) void k_stats_finish(const double *part, int n_part, double *out)
{
    __shared__ double sh[CT];
    __shared__ double mx[CT], mn[CT], mxa[CT];
    double sq = 0.0, amax = 0.0, lo = INFINITY, hi = -INFINITY, bad = 0.0;
    for (int b = threadIdx.x; b < n_part; b += CT) {
        sq += part[5 * b];
        amax = fmax(amax, part[5 * b + 1]);
        lo = fmin(lo, part[5 * b + 2]);
        hi = fmax(hi, part[5 * b + 3]);
        bad += part[5 * b + 4];
    }
    mxa[threadIdx.x] = amax;
    mn[threadIdx.x] = lo;
    mx[threadIdx.x] = hi;
    __syncthreads();
    for (int off = CT / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            mxa[threadIdx.x] = fmax(mxa[threadIdx.x], mxa[threadIdx.x + off]);
            mn[threadIdx.x] = fmin(mn[threadIdx.x], mn[threadIdx.x + off]);
            mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + off]);
        }
        __syncthreads();
    }
    sq = block_sum(sq, sh);
    bad = block_sum(bad, sh);
    if (threadIdx.x == 0) {
        out[0] = sqrt(sq);
        out[1] = mxa[0];
        out[2] = mn[0];
        out[3] = mx[0];
        out[4] = bad;
    }
}

// out = x + sum_i y[i] v_i over `count` basis vectors stored back to back.
__global__ __launch_bounds__(CT) void k_combine(double *out, const double *x, const double *basis, int64_t stride,
                                                int count, const double *y, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * CT + threadIdx.x; i < n; i += (int64_t)gridDim.x * CT) {
        double a = 0.0;
        for (int k = 0; k < count; ++k) a += y[k] * basis[(int64_t)k * stride + i];
        out[i] = x[i] + a;
    }
}

// ---------------------------------------------------------------- conduction
// y = A x + (shift ? shift * x : 0), each row summed in column order.
__global__ __launch_bounds__(CT) void k_csr(const int64_t *rowptr, const int32_t *col, const double *val,
                                            const double *x, const double *shift, double *y, int64_t n)
{
    for (int64_t r = (int64_t)blockIdx.x * CT + threadIdx.x; r < n; r += (int64_t)gridDim.x * CT) {
        double a = 0.0;
        for (int64_t e = rowptr[r]; e < rowptr[r + 1]; ++e) a += val[e] * x[col[e]];
        y[r] = shift ? a + shift[r] * x[r] : a;
    }
}

// The same product when A is the conduction stencil of an nx x ny x nz grid
// (ny, nz >= 3): neighbours along x (not periodic) and along y, z (periodic),
// symmetric, so A holds per cell the diagonal and the coefficients of its +x,
// +y, +z faces (cx, cy, cz; zero at the last x plane) and the -x/-y/-z
// coefficients are the neighbours' +faces.  Each row is summed in ascending
// column order, as the CSR row is: the -x term first, the +x term last and
// the five terms of the cell's own plane (-y, -z, self, +z, +y, whose order
// changes where y or z wraps) sorted by a 9-comparator network.
__device__ __forceinline__ void cswap(int64_t &ca, double &va, int64_t &cb, double &vb)
{
    const bool s = cb < ca;
    const int64_t c0 = s ? cb : ca, c1 = s ? ca : cb;
    const double v0 = s ? vb : va, v1 = s ? va : vb;
    ca = c0;
    cb = c1;
    va = v0;
    vb = v1;
}

__global__ __launch_bounds__(CT) void k_stencil7(const double *diag, const double *cx, const double *cy,
                                                 const double *cz, int nx, int ny, int nz, const double *x,
                                                 const double *shift, double *y, int64_t n)
{
    const int64_t sxy = (int64_t)ny * nz;
    for (int64_t r = (int64_t)blockIdx.x * CT + threadIdx.x; r < n; r += (int64_t)gridDim.x * CT) {
        const int k = (int)(r % nz);
        const int j = (int)((r / nz) % ny);
        const int i = (int)(r / sxy);
        int64_t c0 = j > 0 ? r - nz : r + (int64_t)(ny - 1) * nz;  // -y
        int64_t c1 = k > 0 ? r - 1 : r + (nz - 1);                   // -z
        int64_t c2 = r;
        int64_t c3 = k + 1 < nz ? r + 1 : r - (nz - 1);              // +z
        int64_t c4 = j + 1 < ny ? r + nz : r - (int64_t)(ny - 1) * nz;  // +y
        double v0 = cy[c0], v1 = cz[c1], v2 = diag[r], v3 = cz[r], v4 = cy[r];
        cswap(c0, v0, c1, v1);
        cswap(c3, v3, c4, v4);
        cswap(c2, v2, c4, v4);
        cswap(c2, v2, c3, v3);
        cswap(c1, v1, c4, v4);
        cswap(c0, v0, c3, v3);
        cswap(c0, v0, c2, v2);
        cswap(c1, v1, c3, v3);
        cswap(c1, v1, c2, v2);
        double a = 0.0;
        if (i > 0) a += cx[r - sxy] * x[r - sxy];
        a += v0 * x[c0];
        a += v1 * x[c1];
        a += v2 * x[c2];
        a += v3 * x[c3];
        a += v4 * x[c4];
        if (i + 1 < nx) a += cx[r] * x[r + sxy];
        y[r] = shift ? a + shift[r] * x[r] : a;
    }
}

// R = (A T - b) - Q with b = hot T_hot + cold T_cold.
__global__ __launch_bounds__(CT) void k_residual(const double *AT, const double *hot, const double *cold,
                                                 const double *Q, double T_hot, double T_cold, double *R, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * CT + threadIdx.x; i < n; i += (int64_t)gridDim.x * CT) {
        const double b = hot[i] * T_hot + cold[i] * T_cold;
        R[i] = (AT[i] - b) - Q[i];
    }
}

// ---------------------------------------------------------------- patches <-> cells
// Patch temperatures: the walls at T_hot / T_cold, internal patches at the
// temperature of their cell (coupled.py:45-52).
__global__ __launch_bounds__(CT) void k_patch_temperature(const int8_t *cls, const int32_t *cell, const double *T,
                                                          double T_hot, double T_cold, double *Tp, int64_t n)
{
    for (int64_t p = (int64_t)blockIdx.x * CT + threadIdx.x; p < n; p += (int64_t)gridDim.x * CT) {
        const int c = cls[p];
        const double t = T[c == 2 ? cell[p] : 0];  // walls have no cell
        Tp[p] = c == 0 ? T_hot : (c == 1 ? T_cold : t);
    }
}

// E = eps sigma A T^4 (exchange.py:265-270 without a reference temperature).
__global__ __launch_bounds__(CT) void k_emission(const double *Tp, const double *area, double eps, double *E,
                                                 double *eps_vec, int64_t n)
{
    const double es = eps * SIGMA;
    for (int64_t p = (int64_t)blockIdx.x * CT + threadIdx.x; p < n; p += (int64_t)gridDim.x * CT) {
        const double t = Tp[p];
        const double t2 = t * t;
        E[p] = es * area[p] * (t2 * t2);
        eps_vec[p] = eps;
    }
}

// dE = 4 eps sigma A T^3 dT on internal patches (walls fixed), exchange.py:274-287.
__global__ __launch_bounds__(CT) void k_emission_derivative(const int8_t *cls, const int32_t *cell, const double *Tp,
                                                            const double *area, const double *dT, double eps,
                                                            double *dE, double *eps_vec, int64_t n)
{
    const double f = 4.0 * eps * SIGMA;
    for (int64_t p = (int64_t)blockIdx.x * CT + threadIdx.x; p < n; p += (int64_t)gridDim.x * CT) {
        const double t = Tp[p];
        const bool inner = cls[p] == 2;
        const double dv = dT[inner ? cell[p] : 0];  // walls have no cell
        const double d = inner ? dv : 0.0;
        dE[p] = f * area[p] * (t * t * t) * d;
        eps_vec[p] = eps;
    }
}

// Per cell, the sum over its internal patches (in patch order) of
// absorbed - emitted: the patch ids of a cell are its first id plus the rank
// of the face among the exposed faces (cell code = first id << 6 | mask).
// With `minus` set: out = Ax - that sum (the Jacobian product).
__global__ __launch_bounds__(CT) void k_cell_net(const uint32_t *code, const double *absorbed, const double *emitted,
                                                 const double *Ax, double *out, int64_t n_cells)
{
    for (int64_t c = (int64_t)blockIdx.x * CT + threadIdx.x; c < n_cells; c += (int64_t)gridDim.x * CT) {
        const uint32_t k = code[c];
        uint32_t m = k & 63u;
        uint32_t p = k >> 6;
        double s = 0.0;
        while (m) {
            s += absorbed[p] - emitted[p];
            ++p;
            m &= m - 1u;
        }
        out[c] = Ax ? Ax[c] - s : s;
    }
}

// Emission derivative summed per cell: sum_p 4 eps sigma A_p T_c^3 (coupled.py:62-71).
__global__ __launch_bounds__(CT) void k_emission_diagonal(const uint32_t *code, const double *area, const double *T,
                                                          double eps, double *diag, int64_t n_cells)
{
    const double f = 4.0 * eps * SIGMA;
    for (int64_t c = (int64_t)blockIdx.x * CT + threadIdx.x; c < n_cells; c += (int64_t)gridDim.x * CT) {
        const uint32_t k = code[c];
        uint32_t m = k & 63u;
        uint32_t p = k >> 6;
        const double t = T[c];
        const double t3 = t * t * t;
        double s = 0.0;
        while (m) {
            s += f * area[p] * t3;
            ++p;
            m &= m - 1u;
        }
        diag[c] = s;
    }
}

// ---------------------------------------------------------------- Jacobi-PCG
// scalars: [0] r.z, [1] p.Ap, [2] r.z of the next step
__global__ __launch_bounds__(CT) void k_pcg_start(const double *rhs, const double *inv_diag, double *x, double *r,
                                                  double *z, double *p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * CT + threadIdx.x; i < n; i += (int64_t)gridDim.x * CT) {
        const double v = rhs[i];
        x[i] = 0.0;
        r[i] = v;
        const double zz = inv_diag[i] * v;
        z[i] = zz;
        p[i] = zz;
    }
}

// x += alpha p, r -= alpha Ap, z = inv_diag r, alpha = s[0] / s[1]
__global__ __launch_bounds__(CT) void k_pcg_update(const double *s, const double *p, const double *Ap,
                                                   const double *inv_diag, double *x, double *r, double *z, int64_t n)
{
    const double alpha = s[0] / s[1];
    for (int64_t i = (int64_t)blockIdx.x * CT + threadIdx.x; i < n; i += (int64_t)gridDim.x * CT) {
        x[i] += alpha * p[i];
        r[i] -= alpha * Ap[i];
        z[i] = inv_diag[i] * r[i];
    }
}

// p = z + (s[2] / s[0]) p; then s[0] = s[2]
__global__ __launch_bounds__(CT) void k_pcg_direction(double *s, const double *z, double *p, int64_t n)
{
    const double beta = s[2] / s[0];
    for (int64_t i = (int64_t)blockIdx.x * CT + threadIdx.x; i < n; i += (int64_t)gridDim.x * CT)
        p[i] = z[i] + beta * p[i];
}

// r = b - A x, z = inv_diag r, p = z
__global__ __launch_bounds__(CT) void k_cg_start(const double *b, const double *Ax, const double *inv_diag, double *r,
                                                 double *z, double *p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * CT + threadIdx.x; i < n; i += (int64_t)gridDim.x * CT) {
        const double v = b[i] - Ax[i];
        r[i] = v;
        const double zz = inv_diag[i] * v;
        z[i] = zz;
        p[i] = zz;
    }
}

__global__ void k_copy_scalar(double *s, int to, int from)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) s[to] = s[from];
}

}  // namespace
}  // namespace prh

using namespace prh;

struct prh_coupled {
    int device = 0;
    hipStream_t stream = nullptr;  // its own stream: runs beside the library stream
    const prh_geometry *geo = nullptr;
    prh_operator *op = nullptr;
    int64_t n_cells = 0, n_patches = 0;
    int32_t n_vectors = 0;
    double *vec = nullptr;  // n_vectors x n_cells
    int64_t *rowptr = nullptr;
    int32_t *col = nullptr;
    double *val = nullptr, *diag = nullptr, *hot = nullptr, *cold = nullptr;
    // scratch cell vectors: A x, PCG r z p Ap, shift, inverse diagonal
    double *ax = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *ap = nullptr, *shift = nullptr, *inv = nullptr;
    // patch vectors: temperatures of the last radiation, linearisation point, derivative
    double *Tp = nullptr, *Tlin = nullptr, *dE = nullptr, *emitted = nullptr;
    double *part = nullptr, *scal = nullptr, *ycoef = nullptr;
    // the conduction stencil (k_stencil7), when A has exactly that form
    bool stencil = false;
    int32_t nx = 0, ny = 0, nz = 0;
    double *cx = nullptr, *cy = nullptr, *cz = nullptr;
    int32_t pcg_steps = 0;
    bool have_lin = false, have_pcg = false;
    // Jacobian-vector products: FP64 reflection series (PRH_FP64) or FP32
    // Chebyshev sweeps with FP64 refinement (PRH_MIXED, mixed_solve_single)
    int precision = PRH_FP64;
};

namespace {

double *slot(prh_coupled *c, int32_t v) { return c->vec + (size_t)v * c->n_cells; }

bool slot_ok(const prh_coupled *c, int32_t v) { return v >= 0 && v < c->n_vectors; }

void free_coupled(prh_coupled *c)
{
    void *ptrs[] = {c->vec, c->rowptr, c->col,  c->val,  c->diag, c->hot,     c->cold,  c->ax,
                    c->r,   c->z,      c->p,    c->ap,   c->shift, c->inv,    c->Tp,    c->Tlin,
                    c->dE,  c->emitted, c->part, c->scal, c->ycoef, c->cx,    c->cy,    c->cz};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

int stream_of(prh_coupled *c, hipStream_t *st)
{
    PRH_HIP(hipSetDevice(c->device));
    *st = c->stream;
    return PRH_OK;
}

// y = A x (+ shift * x): the stencil form when A has it, else the CSR rows.
int apply_conduction(prh_coupled *c, const double *x, const double *shift, double *y, hipStream_t st)
{
    if (c->stencil)
        k_stencil7<<<blocks_of(c->n_cells), CT, 0, st>>>(c->diag, c->cx, c->cy, c->cz, c->nx, c->ny, c->nz, x, shift,
                                                         y, c->n_cells);
    else
        k_csr<<<blocks_of(c->n_cells), CT, 0, st>>>(c->rowptr, c->col, c->val, x, shift, y, c->n_cells);
    PRH_HIP(hipGetLastError());
    return PRH_OK;
}

bool same_bits(double a, double b) { return std::memcmp(&a, &b, sizeof a) == 0; }

// Whether the CSR matrix is the conduction stencil of the grid (exact
// column sets per row, exact symmetry); fills the face coefficients if so.
bool detect_stencil(const int64_t *rp, const int32_t *col, const double *val, int nx, int ny, int nz,
                    std::vector<double> &cx, std::vector<double> &cy, std::vector<double> &cz)
{
    const int64_t sxy = (int64_t)ny * nz, n = (int64_t)nx * sxy;
    if (ny < 3 || nz < 3) return false;
    cx.assign(n, 0.0);
    cy.assign(n, 0.0);
    cz.assign(n, 0.0);
    for (int pass = 0; pass < 2; ++pass)
        for (int64_t r = 0; r < n; ++r) {
            const int k = (int)(r % nz), j = (int)((r / nz) % ny), i = (int)(r / sxy);
            const int64_t ym = j > 0 ? r - nz : r + (int64_t)(ny - 1) * nz, yp = j + 1 < ny ? r + nz : r - (int64_t)(ny - 1) * nz;
            const int64_t zm = k > 0 ? r - 1 : r + (nz - 1), zp = k + 1 < nz ? r + 1 : r - (nz - 1);
            int64_t want[7];
            int m = 0;
            if (i > 0) want[m++] = r - sxy;
            want[m++] = ym;
            want[m++] = zm;
            want[m++] = r;
            want[m++] = zp;
            want[m++] = yp;
            if (i + 1 < nx) want[m++] = r + sxy;
            std::sort(want, want + m);
            if (rp[r + 1] - rp[r] != m) return false;
            for (int q = 0; q < m; ++q) {
                const int64_t e = rp[r] + q, s = col[e];
                if (s != want[q]) return false;
                if (pass == 0) {
                    if (i + 1 < nx && s == r + sxy) cx[r] = val[e];
                    else if (s == yp) cy[r] = val[e];
                    else if (s == zp) cz[r] = val[e];
                } else {
                    // the -faces must equal the neighbours' +faces bit for bit
                    if (i > 0 && s == r - sxy && !same_bits(val[e], cx[s])) return false;
                    if (s == ym && !same_bits(val[e], cy[s])) return false;
                    if (s == zm && !same_bits(val[e], cz[s])) return false;
                }
            }
        }
    return true;
}

// scal[k] = x . y
int dot_into(prh_coupled *c, const double *x, const double *y, int k, hipStream_t st)
{
    const unsigned rb = red_blocks_of(c->n_cells);
    k_dot_part<<<rb, CT, 0, st>>>(x, y, c->n_cells, c->part);
    k_finish_sum<<<1, CT, 0, st>>>(c->part, (int)rb, c->scal, k);
    PRH_HIP(hipGetLastError());
    return PRH_OK;
}

// Radiation of the patch temperatures in c->Tp: emission into op->emit,
// series into op->absorbed.
int radiate(prh_coupled *c, double eps, double rtol, hipStream_t st, double *res, int32_t *its)
{
    prh_operator *op = c->op;
    const int64_t np = c->n_patches;
    k_emission<<<blocks_of(np), CT, 0, st>>>(c->Tp, op->area, eps, op->emit, op->eps, np);
    PRH_HIP(hipGetLastError());
    return series_fp64(op, rtol, 10000, st, res, its);
}

}  // namespace

extern "C" {

int prh_coupled_create(const prh_geometry *geo, prh_operator *op, int64_t n_cells, const int64_t *A_rowptr,
                       const int32_t *A_col, const double *A_val, const double *hot_rhs, const double *cold_rhs,
                       int32_t n_vectors, prh_coupled **out)
{
    PRH_CHECK_ARG(A_rowptr && A_col && A_val && hot_rhs && cold_rhs && out, "prh_coupled_create: null pointer");
    PRH_CHECK_ARG(n_vectors >= 1 && n_cells >= 1, "prh_coupled_create: sizes must be positive");
    PRH_CHECK_ARG((geo == nullptr) == (op == nullptr), "prh_coupled_create: give both geometry and operator, or neither");
    int device = 0;
    int64_t n_patches = 0;
    if (geo) {
        const Geometry &g = geo->g;
        PRH_CHECK_ARG(op->n == g.n_patches, "prh_coupled_create: operator has %lld patches, geometry %lld",
                      (long long)op->n, (long long)g.n_patches);
        PRH_CHECK_ARG(op->device == g.device, "prh_coupled_create: operator and geometry on different devices");
        PRH_CHECK_ARG((int64_t)g.nx * g.ny * g.nz == n_cells, "prh_coupled_create: geometry has %lld cells, not %lld",
                      (long long)g.nx * g.ny * g.nz, (long long)n_cells);
        device = g.device;
        n_patches = g.n_patches;
    } else {
        PRH_HIP(hipGetDevice(&device));
    }
    const int64_t n = n_cells;
    PRH_CHECK_ARG(A_rowptr[0] == 0 && A_rowptr[n] >= n, "prh_coupled_create: bad matrix row pointers");
    const int64_t nnz = A_rowptr[n];
    for (int64_t r = 0; r < n; ++r)
        PRH_CHECK_ARG(A_rowptr[r + 1] >= A_rowptr[r], "prh_coupled_create: row pointers not monotone");
    for (int64_t e = 0; e < nnz; ++e)
        PRH_CHECK_ARG(A_col[e] >= 0 && A_col[e] < n, "prh_coupled_create: column index out of range");
    *out = nullptr;
    PRH_HIP(hipSetDevice(device));
    prh_coupled *c = new prh_coupled();
    c->device = device;
    c->geo = geo;
    c->op = op;
    c->n_cells = n;
    c->n_patches = n_patches;
    c->n_vectors = n_vectors;
    const int64_t np = std::max<int64_t>(n_patches, 1);
    bool ok = true;
    auto alloc = [&](void **q, size_t bytes) {
        if (ok && hipMalloc(q, std::max<size_t>(bytes, 8)) != hipSuccess) ok = false;
    };
    alloc((void **)&c->vec, (size_t)n_vectors * n * 8);
    alloc((void **)&c->rowptr, (size_t)(n + 1) * 8);
    alloc((void **)&c->col, (size_t)nnz * 4);
    alloc((void **)&c->val, (size_t)nnz * 8);
    for (double **q : {&c->diag, &c->hot, &c->cold, &c->ax, &c->r, &c->z, &c->p, &c->ap, &c->shift, &c->inv})
        alloc((void **)q, (size_t)n * 8);
    for (double **q : {&c->Tp, &c->Tlin, &c->dE, &c->emitted}) alloc((void **)q, (size_t)np * 8);
    alloc((void **)&c->part, (size_t)RED_BLOCKS * 5 * 8);
    alloc((void **)&c->scal, 128 * 8);
    alloc((void **)&c->ycoef, 256 * 8);
    if (ok && hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) ok = false;
    if (!ok) {
        free_coupled(c);
        delete c;
        set_error("prh_coupled_create: out of device memory");
        return PRH_ERR_HIP;
    }
    std::vector<double> diag(n, 0.0);
    for (int64_t r = 0; r < n; ++r)
        for (int64_t e = A_rowptr[r]; e < A_rowptr[r + 1]; ++e)
            if (A_col[e] == r) diag[r] += A_val[e];
    PRH_HIP(hipMemcpy(c->rowptr, A_rowptr, (size_t)(n + 1) * 8, hipMemcpyHostToDevice));
    PRH_HIP(hipMemcpy(c->col, A_col, (size_t)nnz * 4, hipMemcpyHostToDevice));
    PRH_HIP(hipMemcpy(c->val, A_val, (size_t)nnz * 8, hipMemcpyHostToDevice));
    PRH_HIP(hipMemcpy(c->diag, diag.data(), (size_t)n * 8, hipMemcpyHostToDevice));
    if (geo) {  // a handle without geometry keeps the CSR rows
        const Geometry &g = geo->g;
        std::vector<double> fx, fy, fz;
        if ((int64_t)g.nx * g.ny * g.nz == n && detect_stencil(A_rowptr, A_col, A_val, g.nx, g.ny, g.nz, fx, fy, fz)) {
            alloc((void **)&c->cx, (size_t)n * 8);
            alloc((void **)&c->cy, (size_t)n * 8);
            alloc((void **)&c->cz, (size_t)n * 8);
            if (ok) {
                PRH_HIP(hipMemcpy(c->cx, fx.data(), (size_t)n * 8, hipMemcpyHostToDevice));
                PRH_HIP(hipMemcpy(c->cy, fy.data(), (size_t)n * 8, hipMemcpyHostToDevice));
                PRH_HIP(hipMemcpy(c->cz, fz.data(), (size_t)n * 8, hipMemcpyHostToDevice));
                c->stencil = true;
                c->nx = g.nx;
                c->ny = g.ny;
                c->nz = g.nz;
            }
        }
    }
    PRH_HIP(hipMemcpy(c->hot, hot_rhs, (size_t)n * 8, hipMemcpyHostToDevice));
    PRH_HIP(hipMemcpy(c->cold, cold_rhs, (size_t)n * 8, hipMemcpyHostToDevice));
    PRH_HIP(hipMemsetAsync(c->vec, 0, (size_t)n_vectors * n * 8, c->stream));
    PRH_HIP(hipStreamSynchronize(c->stream));
    *out = c;
    return PRH_OK;
}

void prh_coupled_destroy(prh_coupled *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_coupled(c);
    delete c;
}

int prh_coupled_set(prh_coupled *c, int32_t v, const double *host)
{
    PRH_CHECK_ARG(c && host && slot_ok(c, v), "prh_coupled_set: bad argument");
    hipStream_t st;
    int rc = stream_of(c, &st);
    if (rc) return rc;
    PRH_HIP(hipMemcpyAsync(slot(c, v), host, (size_t)c->n_cells * 8, hipMemcpyHostToDevice, st));
    PRH_HIP(hipStreamSynchronize(st));
    return PRH_OK;
}

int prh_coupled_get(prh_coupled *c, int32_t v, double *host)
{
    PRH_CHECK_ARG(c && host && slot_ok(c, v), "prh_coupled_get: bad argument");
    hipStream_t st;
    int rc = stream_of(c, &st);
    if (rc) return rc;
    PRH_HIP(hipMemcpyAsync(host, slot(c, v), (size_t)c->n_cells * 8, hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));
    return PRH_OK;
}

int prh_coupled_fill(prh_coupled *c, int32_t v, double value)
{
    PRH_CHECK_ARG(c && slot_ok(c, v), "prh_coupled_fill: bad argument");
    hipStream_t st;
    int rc = stream_of(c, &st);
    if (rc) return rc;
    k_fill<<<red_blocks_of(c->n_cells), CT, 0, st>>>(slot(c, v), c->n_cells, value);
    PRH_HIP(hipGetLastError());
    return PRH_OK;
}

int prh_coupled_axpby(prh_coupled *c, int32_t y, double a, int32_t x, double b)
{
    PRH_CHECK_ARG(c && slot_ok(c, y) && slot_ok(c, x), "prh_coupled_axpby: bad slot");
    hipStream_t st;
    int rc = stream_of(c, &st);
    if (rc) return rc;
    k_axpby<<<red_blocks_of(c->n_cells), CT, 0, st>>>(slot(c, y), slot(c, x), c->n_cells, a, b);
    PRH_HIP(hipGetLastError());
    return PRH_OK;
}

int prh_coupled_divide(prh_coupled *c, int32_t y, int32_t x, double d)
{
    PRH_CHECK_ARG(c && slot_ok(c, y) && slot_ok(c, x), "prh_coupled_divide: bad slot");
    hipStream_t st;
    int rc = stream_of(c, &st);
    if (rc) return rc;
    k_divide<<<red_blocks_of(c->n_cells), CT, 0, st>>>(slot(c, y), slot(c, x), c->n_cells, d);
    PRH_HIP(hipGetLastError());
    return PRH_OK;
}

int prh_coupled_dot(prh_coupled *c, int32_t x, int32_t y, double *out)
{
    PRH_CHECK_ARG(c && out && slot_ok(c, x) && slot_ok(c, y), "prh_coupled_dot: bad argument");
    hipStream_t st;
    int rc = stream_of(c, &st);
    if (rc) return rc;
    if ((rc = dot_into(c, slot(c, x), slot(c, y), 0, st))) return rc;
    PRH_HIP(hipMemcpyAsync(out, c->scal, 8, hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));
    return PRH_OK;
}

int prh_coupled_stats(prh_coupled *c, int32_t x, double *out)
{
    PRH_CHECK_ARG(c && out && slot_ok(c, x), "prh_coupled_stats: bad argument");
    hipStream_t st;
    int rc = stream_of(c, &st);
    if (rc) return rc;
    const unsigned rb = red_blocks_of(c->n_cells);
    k_stats_part<<<rb, CT, 0, st>>>(slot(c, x), c->n_cells, c->part);
    k_stats_finish<<<1, CT, 0, st>>>(c->part, (int)rb, c->scal);
    PRH_HIP(hipGetLastError());
    PRH_HIP(hipMemcpyAsync(out, c->scal, 5 * 8, hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));
    return PRH_OK;
}

int prh_coupled_matvec(prh_coupled *c, int32_t x, int32_t y)
{
    PRH_CHECK_ARG(c && slot_ok(c, x) && slot_ok(c, y) && x != y, "prh_coupled_matvec: bad slot");
    hipStream_t st;
    int rc = stream_of(c, &st);
    if (rc) return rc;
    return apply_conduction(c, slot(c, x), nullptr, slot(c, y), st);
}

int prh_coupled_radiation(prh_coupled *c, int32_t T, double T_hot, double T_cold, double emissivity, double rtol,
                          int32_t Q_out, double *series_residual, int32_t *series_iterations)
{
    PRH_CHECK_ARG(c && slot_ok(c, T) && slot_ok(c, Q_out) && series_residual && series_iterations,
                  "prh_coupled_radiation: bad argument");
    PRH_CHECK_ARG(c->geo, "prh_coupled_radiation: no geometry and operator in this handle");
    hipStream_t st;
    int rc = stream_of(c, &st);
    if (rc) return rc;
    const Geometry &g = c->geo->g;
    const int64_t np = c->n_patches;
    k_patch_temperature<<<blocks_of(np), CT, 0, st>>>(g.patch_class, g.patch_cell, slot(c, T), T_hot, T_cold, c->Tp,
                                                      np);
    PRH_HIP(hipGetLastError());
    if ((rc = radiate(c, emissivity, rtol, st, series_residual, series_iterations))) return rc;
    PRH_HIP(hipMemcpyAsync(c->emitted, c->op->emit, (size_t)np * 8, hipMemcpyDeviceToDevice, st));
    k_cell_net<<<blocks_of(c->n_cells), CT, 0, st>>>(g.cell_code, c->op->absorbed, c->op->emit, nullptr,
                                                     slot(c, Q_out), c->n_cells);
    PRH_HIP(hipGetLastError());
    return PRH_OK;
}

int prh_coupled_residual(prh_coupled *c, int32_t T, int32_t Q, double T_hot, double T_cold, int32_t R_out)
{
    PRH_CHECK_ARG(c && slot_ok(c, T) && slot_ok(c, Q) && slot_ok(c, R_out), "prh_coupled_residual: bad slot");
    hipStream_t st;
    int rc = stream_of(c, &st);
    if (rc) return rc;
    if ((rc = apply_conduction(c, slot(c, T), nullptr, c->ax, st))) return rc;
    k_residual<<<red_blocks_of(c->n_cells), CT, 0, st>>>(c->ax, c->hot, c->cold, slot(c, Q), T_hot, T_cold,
                                                         slot(c, R_out), c->n_cells);
    PRH_HIP(hipGetLastError());
    return PRH_OK;
}

int prh_coupled_linearize(prh_coupled *c, int32_t T, double T_hot, double T_cold)
{
    PRH_CHECK_ARG(c && slot_ok(c, T), "prh_coupled_linearize: bad slot");
    PRH_CHECK_ARG(c->geo, "prh_coupled_linearize: no geometry and operator in this handle");
    hipStream_t st;
    int rc = stream_of(c, &st);
    if (rc) return rc;
    const Geometry &g = c->geo->g;
    k_patch_temperature<<<blocks_of(c->n_patches), CT, 0, st>>>(g.patch_class, g.patch_cell, slot(c, T), T_hot,
                                                                T_cold, c->Tlin, c->n_patches);
    PRH_HIP(hipGetLastError());
    c->have_lin = true;
    return PRH_OK;
}

int prh_coupled_jacobian(prh_coupled *c, int32_t v, double emissivity, double rtol, int32_t out,
                         int32_t *series_iterations)
{
    PRH_CHECK_ARG(c && slot_ok(c, v) && slot_ok(c, out) && v != out, "prh_coupled_jacobian: bad slot");
    PRH_CHECK_ARG(c->geo, "prh_coupled_jacobian: no geometry and operator in this handle");
    PRH_CHECK_ARG(c->have_lin, "prh_coupled_jacobian: no linearisation point (prh_coupled_linearize)");
    hipStream_t st;
    int rc = stream_of(c, &st);
    if (rc) return rc;
    const Geometry &g = c->geo->g;
    prh_operator *op = c->op;
    const int64_t np = c->n_patches;
    k_emission_derivative<<<blocks_of(np), CT, 0, st>>>(g.patch_class, g.patch_cell, c->Tlin, op->area, slot(c, v),
                                                        emissivity, op->emit, op->eps, np);
    PRH_HIP(hipGetLastError());
    if ((rc = apply_conduction(c, slot(c, v), nullptr, c->ax, st))) return rc;
    double res = 0.0;
    int32_t its = 0;
    if (c->precision == PRH_MIXED) rc = mixed_solve_single(op, op->emit, emissivity, rtol, st, op->absorbed, &its);
    else rc = series_fp64(op, rtol, 10000, st, &res, &its);
    if (rc) return rc;
    k_cell_net<<<blocks_of(c->n_cells), CT, 0, st>>>(g.cell_code, op->absorbed, op->emit, c->ax, slot(c, out),
                                                     c->n_cells);
    PRH_HIP(hipGetLastError());
    if (series_iterations) *series_iterations = its;
    return PRH_OK;
}

int prh_coupled_set_precision(prh_coupled *c, int32_t precision)
{
    PRH_CHECK_ARG(c && (precision == PRH_FP64 || precision == PRH_MIXED), "prh_coupled_set_precision: bad argument");
    PRH_CHECK_ARG(c->op || precision == PRH_FP64, "prh_coupled_set_precision: no operator in this handle");
    c->precision = precision;
    return PRH_OK;
}

int prh_coupled_set_preconditioner(prh_coupled *c, int32_t T, double emissivity, int32_t steps)
{
    PRH_CHECK_ARG(c && slot_ok(c, T) && steps >= 1, "prh_coupled_set_preconditioner: bad argument");
    PRH_CHECK_ARG(c->geo, "prh_coupled_set_preconditioner: no geometry and operator in this handle");
    hipStream_t st;
    int rc = stream_of(c, &st);
    if (rc) return rc;
    const Geometry &g = c->geo->g;
    k_emission_diagonal<<<blocks_of(c->n_cells), CT, 0, st>>>(g.cell_code, c->op->area, slot(c, T), emissivity,
                                                              c->shift, c->n_cells);
    // inverse of diag(A) + shift
    k_fill<<<red_blocks_of(c->n_cells), CT, 0, st>>>(c->inv, c->n_cells, 1.0);
    PRH_HIP(hipGetLastError());
    std::vector<double> d(c->n_cells), s(c->n_cells);
    PRH_HIP(hipMemcpyAsync(d.data(), c->diag, (size_t)c->n_cells * 8, hipMemcpyDeviceToHost, st));
    PRH_HIP(hipMemcpyAsync(s.data(), c->shift, (size_t)c->n_cells * 8, hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));
    for (int64_t i = 0; i < c->n_cells; ++i) d[i] = 1.0 / (d[i] + s[i]);
    PRH_HIP(hipMemcpyAsync(c->inv, d.data(), (size_t)c->n_cells * 8, hipMemcpyHostToDevice, st));
    PRH_HIP(hipStreamSynchronize(st));
    c->pcg_steps = steps;
    c->have_pcg = true;
    return PRH_OK;
}

int prh_coupled_precondition(prh_coupled *c, int32_t rhs, int32_t out)
{
    PRH_CHECK_ARG(c && slot_ok(c, rhs) && slot_ok(c, out) && rhs != out, "prh_coupled_precondition: bad slot");
    PRH_CHECK_ARG(c->have_pcg, "prh_coupled_precondition: no preconditioner (prh_coupled_set_preconditioner)");
    hipStream_t st;
    int rc = stream_of(c, &st);
    if (rc) return rc;
    const int64_t n = c->n_cells;
    const unsigned eb = red_blocks_of(n);
    double *x = slot(c, out);
    k_pcg_start<<<eb, CT, 0, st>>>(slot(c, rhs), c->inv, x, c->r, c->z, c->p, n);
    PRH_HIP(hipGetLastError());
    if ((rc = dot_into(c, c->r, c->z, 0, st))) return rc;
    for (int step = 0; step < c->pcg_steps; ++step) {
        if ((rc = apply_conduction(c, c->p, c->shift, c->ap, st))) return rc;
        if ((rc = dot_into(c, c->p, c->ap, 1, st))) return rc;
        k_pcg_update<<<eb, CT, 0, st>>>(c->scal, c->p, c->ap, c->inv, x, c->r, c->z, n);
        PRH_HIP(hipGetLastError());
        if (step + 1 < c->pcg_steps) {
            if ((rc = dot_into(c, c->r, c->z, 2, st))) return rc;
            k_pcg_direction<<<eb, CT, 0, st>>>(c->scal, c->z, c->p, n);
            k_copy_scalar<<<1, 64, 0, st>>>(c->scal, 0, 2);
            PRH_HIP(hipGetLastError());
        }
    }
    return PRH_OK;
}

int prh_coupled_orthogonalize(prh_coupled *c, int32_t basis, int32_t count, int32_t w, double *h)
{
    PRH_CHECK_ARG(c && h && count >= 1 && count <= 60 && slot_ok(c, basis) && slot_ok(c, basis + count - 1) &&
                      slot_ok(c, w) && (w < basis || w >= basis + count),
                  "prh_coupled_orthogonalize: bad argument");
    hipStream_t st;
    int rc = stream_of(c, &st);
    if (rc) return rc;
    const int64_t n = c->n_cells;
    const unsigned eb = red_blocks_of(n);
    double *wv = slot(c, w);
    // two modified Gram-Schmidt passes; pass k keeps its coefficients in scal[64 k + i]
    for (int pass = 0; pass < 2; ++pass) {
        for (int i = 0; i < count; ++i) {
            const double *vi = slot(c, basis + i);
            if ((rc = dot_into(c, vi, wv, 64 * pass + i, st))) return rc;
            k_sub_scaled<<<eb, CT, 0, st>>>(wv, vi, n, c->scal, 64 * pass + i);
            PRH_HIP(hipGetLastError());
        }
    }
    if ((rc = dot_into(c, wv, wv, 127, st))) return rc;
    std::vector<double> s(128);
    PRH_HIP(hipMemcpyAsync(s.data(), c->scal, 128 * 8, hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < count; ++i) h[i] = (0.0 + s[i]) + s[64 + i];
    h[count] = std::sqrt(s[127]);
    return PRH_OK;
}

int prh_coupled_combine(prh_coupled *c, int32_t out, int32_t x, int32_t basis, int32_t count, const double *y)
{
    PRH_CHECK_ARG(c && y && count >= 1 && count <= 256 && slot_ok(c, out) && slot_ok(c, x) && slot_ok(c, basis) &&
                      slot_ok(c, basis + count - 1) && (out < basis || out >= basis + count),
                  "prh_coupled_combine: bad argument");
    hipStream_t st;
    int rc = stream_of(c, &st);
    if (rc) return rc;
    PRH_HIP(hipMemcpyAsync(c->ycoef, y, (size_t)count * 8, hipMemcpyHostToDevice, st));
    k_combine<<<red_blocks_of(c->n_cells), CT, 0, st>>>(slot(c, out), slot(c, x), slot(c, basis), c->n_cells, count,
                                                        c->ycoef, c->n_cells);
    PRH_HIP(hipGetLastError());
    PRH_HIP(hipStreamSynchronize(st));
    return PRH_OK;
}

int prh_coupled_cg(prh_coupled *c, int32_t b, int32_t x, double rtol, int32_t max_iterations, int32_t *iterations,
                   double *relative_residual)
{
    PRH_CHECK_ARG(c && slot_ok(c, b) && slot_ok(c, x) && b != x && iterations && relative_residual,
                  "prh_coupled_cg: bad argument");
    hipStream_t st;
    int rc = stream_of(c, &st);
    if (rc) return rc;
    const int64_t n = c->n_cells;
    const unsigned eb = red_blocks_of(n);
    double *xv = slot(c, x);
    const double *bv = slot(c, b);
    // Jacobi preconditioner of A alone
    std::vector<double> d(n);
    PRH_HIP(hipMemcpyAsync(d.data(), c->diag, (size_t)n * 8, hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));
    for (int64_t i = 0; i < n; ++i) d[i] = 1.0 / d[i];
    PRH_HIP(hipMemcpyAsync(c->inv, d.data(), (size_t)n * 8, hipMemcpyHostToDevice, st));
    c->have_pcg = false;  // the preconditioner inverse now belongs to this solve
    if ((rc = dot_into(c, bv, bv, 3, st))) return rc;
    if ((rc = apply_conduction(c, xv, nullptr, c->ax, st))) return rc;
    k_cg_start<<<eb, CT, 0, st>>>(bv, c->ax, c->inv, c->r, c->z, c->p, n);
    PRH_HIP(hipGetLastError());
    if ((rc = dot_into(c, c->r, c->z, 0, st))) return rc;
    if ((rc = dot_into(c, c->r, c->r, 4, st))) return rc;
    double h[5];
    PRH_HIP(hipMemcpyAsync(h, c->scal, sizeof(h), hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));
    const double b_norm = std::sqrt(h[3]);
    const double goal = rtol * b_norm;
    int32_t it = 0;
    double r_norm = std::sqrt(h[4]);
    const int CHECK = 16;
    while (r_norm > goal && it < max_iterations) {
        const int batch = std::min(CHECK, max_iterations - it);
        for (int k = 0; k < batch; ++k) {
            if ((rc = apply_conduction(c, c->p, nullptr, c->ap, st))) return rc;
            if ((rc = dot_into(c, c->p, c->ap, 1, st))) return rc;
            k_pcg_update<<<eb, CT, 0, st>>>(c->scal, c->p, c->ap, c->inv, xv, c->r, c->z, n);
            PRH_HIP(hipGetLastError());
            if ((rc = dot_into(c, c->r, c->z, 2, st))) return rc;
            k_pcg_direction<<<eb, CT, 0, st>>>(c->scal, c->z, c->p, n);
            k_copy_scalar<<<1, 64, 0, st>>>(c->scal, 0, 2);
            PRH_HIP(hipGetLastError());
        }
        it += batch;
        if ((rc = dot_into(c, c->r, c->r, 4, st))) return rc;
        PRH_HIP(hipMemcpyAsync(h, c->scal, sizeof(h), hipMemcpyDeviceToHost, st));
        PRH_HIP(hipStreamSynchronize(st));
        r_norm = std::sqrt(h[4]);
    }
    // true residual of the returned field
    if ((rc = apply_conduction(c, xv, nullptr, c->ax, st))) return rc;
    k_axpby<<<eb, CT, 0, st>>>(c->ax, bv, n, -1.0, 1.0);
    PRH_HIP(hipGetLastError());
    if ((rc = dot_into(c, c->ax, c->ax, 4, st))) return rc;
    PRH_HIP(hipMemcpyAsync(h, c->scal, sizeof(h), hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));
    *iterations = it;
    *relative_residual = b_norm > 0.0 ? std::sqrt(h[4]) / b_norm : std::sqrt(h[4]);
    return PRH_OK;
}

int prh_coupled_state(prh_coupled *c, double *patch_temperature, double *emitted, double *absorbed)
{
    PRH_CHECK_ARG(c, "prh_coupled_state: null handle");
    hipStream_t st;
    int rc = stream_of(c, &st);
    if (rc) return rc;
    const size_t bytes = (size_t)c->n_patches * 8;
    PRH_CHECK_ARG(c->geo, "prh_coupled_state: no geometry and operator in this handle");
    if (patch_temperature) PRH_HIP(hipMemcpyAsync(patch_temperature, c->Tp, bytes, hipMemcpyDeviceToHost, st));
    if (emitted) PRH_HIP(hipMemcpyAsync(emitted, c->emitted, bytes, hipMemcpyDeviceToHost, st));
    if (absorbed) PRH_HIP(hipMemcpyAsync(absorbed, c->op->absorbed, bytes, hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));
    return PRH_OK;
}

}  // extern "C"
