// Diagonal balancing of an exchange operator on the device.
//
// Reference: ExchangeOperator._closure_error and _balance,
// porerad/exchange.py:142-197.  The factor d > 0 solves d (S d) = A through
// y = log d: F(y) = y + log(S exp(y)) - log(A) = 0 by scipy.optimize.anderson
// (M = 10, w0 = 0.01, maxiter = 500, f_tol = tolerance), accepted when the
// closure max |d (S d) / A - 1| is within min(max(10 tol, 10 eps), 1e-10);
// otherwise the damped multiplicative iteration y += (log A - log(d S d)) / 2,
// at most 20000 steps.
//
// The Anderson solve restates SciPy 1.15.3 (scipy/optimize/_nonlin.py:
// nonlin_solve, GenericBroyden, Anderson; scipy/optimize/_linesearch.py:
// scalar_search_armijo), which porerad pins in requirements.txt:
//   * nonlin_solve: stop when max|F| <= f_tol (the first check never stops,
//     its dx norm being inf / inf), step dx = -J^-1 F, Armijo line search on
//     phi(s) = ||F(x + s dx)||_2^2 with phi0 = ||F||^2, derphi0 = -phi0,
//     amin = 1e-2 (the full step when it fails), then the Jacobian update;
//   * Anderson.solve: dx = -alpha f + sum_m gamma_m (dX_m + alpha dF_m),
//     gamma = a^-1 (dF_k . f), a_ij = (1 + w0^2 [i = j]) dF_i . dF_j over the
//     last M differences; a singular a clears the history;
//   * alpha = 0.5 max(||x0||, 1) / ||F(x0)||;
//   * a non-finite F, or the 2-norm of a non-finite vector, raises, which
//     _balance turns into the multiplicative fallback.
// Every vector stays on the device; products, dot products and norms are
// reductions in a fixed order, and the host runs the scalar logic and the
// <= 10 x 10 solve (Gaussian elimination with partial pivoting, as LAPACK
// dgesv).  The reductions differ from NumPy's in rounding only, so d agrees
// with the host computation to the balancing tolerance.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <limits>
#include <vector>

#include "prh_internal.h"
#include "prh_operator_impl.h"

namespace prh {
namespace {

constexpr int BT = 256;
constexpr int MAX_BLOCKS = 1024;
constexpr int HIST = 10;        // Anderson M
constexpr double W0 = 0.01;     // Anderson w0
constexpr int MAX_ITER = 500;   // anderson(maxiter=500)
constexpr int MULT_STEPS = 20000;
constexpr double CLOSURE_TOLERANCE = 1.0e-10;  // porerad/exchange.py:31

unsigned blocks_of(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + BT - 1) / BT, MAX_BLOCKS)); }

// Block partials of (max |v|, sum v^2, non-finite count) over a grid-stride loop.
__device__ __forceinline__ void block_stats(double m, double s, unsigned bad, double *part)
{
    __shared__ double sm[BT], ss[BT];
    __shared__ unsigned sb[BT];
    sm[threadIdx.x] = m;
    ss[threadIdx.x] = s;
    sb[threadIdx.x] = bad;
    __syncthreads();
    for (int off = BT / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + off]);
            ss[threadIdx.x] += ss[threadIdx.x + off];
            sb[threadIdx.x] += sb[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x + 0] = sm[0];
        part[3 * blockIdx.x + 1] = ss[0];
        part[3 * blockIdx.x + 2] = (double)sb[0];
    }
}

// out = (max, sum, bad) of the block partials, in block order.
__global__ __launch_bounds__(BT) void k_stats_total(const double *part, int n_part, double *out)
{
    __shared__ double sm[BT], ss[BT], sb[BT];
    double m = 0.0, s = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < n_part; i += BT) {
        m = fmax(m, part[3 * i]);
        s += part[3 * i + 1];
        b += part[3 * i + 2];
    }
    sm[threadIdx.x] = m;
    ss[threadIdx.x] = s;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int off = BT / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + off]);
            ss[threadIdx.x] += ss[threadIdx.x + off];
            sb[threadIdx.x] += sb[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sm[0];
        out[1] = ss[0];
        out[2] = sb[0];
    }
}

// d = exp(y); counts the entries with y non-finite or outside (log_min, log_max).
__global__ __launch_bounds__(BT) void k_exp(const double *y, int64_t n, double log_min, double log_max, double *d,
                                           double *part)
{
    unsigned bad = 0;
    for (int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT) {
        const double v = y[i];
        bad += (v < log_max && v > log_min) ? 0u : 1u;  // false for NaN
        d[i] = exp(v);
    }
    block_stats(0.0, 0.0, bad, part);
}

// f = y + log(prod) - log_area, with (max |f|, sum f^2, entries whose product
// is not finite and positive or whose f is not finite).
__global__ __launch_bounds__(BT) void k_log_residual(const double *y, const double *prod, const double *log_area,
                                                    int64_t n, double *f, double *part)
{
    double m = 0.0, s = 0.0;
    unsigned bad = 0;
    for (int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT) {
        const double p = prod[i];
        const bool ok = isfinite(p) && p > 0.0;
        const double v = (y[i] + log(p)) - log_area[i];
        bad += (ok && isfinite(v)) ? 0u : 1u;
        f[i] = v;
        m = fmax(m, fabs(v));
        s += v * v;
    }
    block_stats(m, s, bad, part);
}

// Closure: row = d prod must be finite and positive; max |row / A - 1|.
__global__ __launch_bounds__(BT) void k_closure(const double *d, const double *prod, const double *area, int64_t n,
                                               double *part)
{
    double m = 0.0;
    unsigned bad = 0;
    for (int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT) {
        const double row = d[i] * prod[i];
        bad += (isfinite(row) && row > 0.0) ? 0u : 1u;
        m = fmax(m, fabs(row / area[i] - 1.0));
    }
    block_stats(m, 0.0, bad, part);
}

// Multiplicative step: row = d prod (checked), y += 0.5 (log A - log row).
__global__ __launch_bounds__(BT) void k_mult_step(const double *d, const double *prod, const double *log_area,
                                                 int64_t n, double *y, double *part)
{
    unsigned bad = 0;
    for (int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT) {
        const double row = d[i] * prod[i];
        const bool ok = isfinite(row) && row > 0.0;
        bad += ok ? 0u : 1u;
        if (ok) y[i] = y[i] + 0.5 * (log_area[i] - log(row));
    }
    block_stats(0.0, 0.0, bad, part);
}

// (max |x|, sum x^2) of a vector.
__global__ __launch_bounds__(BT) void k_norms(const double *x, int64_t n, double *part)
{
    double m = 0.0, s = 0.0;
    unsigned bad = 0;
    for (int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT) {
        const double v = x[i];
        bad += isfinite(v) ? 0u : 1u;
        m = fmax(m, fabs(v));
        s += v * v;
    }
    block_stats(m, s, bad, part);
}

struct DotArgs {
    const double *v[HIST];
    int k;
};

// Block partials of v_j . w for j < k: part[j * n_blocks + block].
__global__ __launch_bounds__(BT) void k_dots(DotArgs a, const double *w, int64_t n, double *part)
{
    __shared__ double sh[BT];
    double acc[HIST];
    for (int j = 0; j < HIST; ++j) acc[j] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT) {
        const double wi = w[i];
        for (int j = 0; j < a.k; ++j) acc[j] += a.v[j][i] * wi;
    }
    for (int j = 0; j < a.k; ++j) {
        sh[threadIdx.x] = acc[j];
        __syncthreads();
        for (int off = BT / 2; off > 0; off >>= 1) {
            if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(size_t)j * gridDim.x + blockIdx.x] = sh[0];
        __syncthreads();
    }
}

// out[j] = sum of the n_part partials of dot j, in block order.
__global__ __launch_bounds__(BT) void k_dot_totals(const double *part, int n_part, double *out)
{
    __shared__ double sh[BT];
    const int j = blockIdx.x;
    double s = 0.0;
    for (int i = threadIdx.x; i < n_part; i += BT) s += part[(size_t)j * n_part + i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = BT / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = sh[0];
}

struct StepArgs {
    const double *dx[HIST];
    const double *df[HIST];
    double gamma[HIST];
    int k;
};

// step = -(-alpha f + sum_m gamma_m (dX_m + alpha dF_m)), in Anderson.solve's order.
__global__ __launch_bounds__(BT) void k_anderson_step(StepArgs a, const double *f, double alpha, int64_t n,
                                                     double *step)
{
    for (int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT) {
        double v = -alpha * f[i];
        for (int m = 0; m < a.k; ++m) v += a.gamma[m] * (a.dx[m][i] + alpha * a.df[m][i]);
        step[i] = -v;
    }
}

// out = x + s step.
__global__ __launch_bounds__(BT) void k_axpy(const double *x, double s, const double *step, int64_t n, double *out)
{
    for (int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT)
        out[i] = x[i] + s * step[i];
}

// Broyden differences: dX = x - last_x, dF = f - last_f; then last_x = x, last_f = f.
__global__ __launch_bounds__(BT) void k_differences(const double *x, const double *f, double *last_x, double *last_f,
                                                   double *dx, double *df, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT) {
        const double xi = x[i], fi = f[i];
        dx[i] = xi - last_x[i];
        df[i] = fi - last_f[i];
        last_x[i] = xi;
        last_f[i] = fi;
    }
}

__global__ __launch_bounds__(BT) void k_log_area(const double *area, int64_t n, double *out)
{
    for (int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT)
        out[i] = log(area[i]);
}

__global__ __launch_bounds__(BT) void k_fill(double *x, int64_t n, double v)
{
    for (int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT) x[i] = v;
}

__global__ __launch_bounds__(BT) void k_scale_out(const double *y, int64_t n, double *scale, float *scale32)
{
    for (int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT) {
        const double d = exp(y[i]);
        scale[i] = d;
        scale32[i] = (float)d;
    }
}

// Thrown-exception stand-in of the Python code: the Anderson attempt failed.
struct AndersonFailed {};

struct Balancer {
    prh_operator *op;
    hipStream_t st;
    int64_t n;
    unsigned grid;
    double log_min, log_max;
    std::vector<double *> vecs;
    double *part = nullptr, *tot = nullptr;
    double *log_area, *tmp_d, *tmp_p;
    double host[3 * HIST + 3];

    ~Balancer()
    {
        for (double *p : vecs) (void)hipFree(p);
        if (part) (void)hipFree(part);
        if (tot) (void)hipFree(tot);
    }
    double *vec()
    {
        double *p = nullptr;
        if (hipMalloc((void **)&p, sizeof(double) * n) != hipSuccess) return nullptr;
        vecs.push_back(p);
        return p;
    }
    int stats(double out[3])
    {
        k_stats_total<<<1, BT, 0, st>>>(part, (int)grid, tot);
        PRH_HIP(hipGetLastError());
        PRH_HIP(hipMemcpyAsync(out, tot, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
        PRH_HIP(hipStreamSynchronize(st));
        return PRH_OK;
    }
    // d = exp(y) into tmp_d, S d into tmp_p; false when y is out of range.
    int exp_product(const double *y, bool *ok)
    {
        double s[3];
        k_exp<<<grid, BT, 0, st>>>(y, n, log_min, log_max, tmp_d, part);
        PRH_HIP(hipGetLastError());
        int rc = stats(s);
        if (rc) return rc;
        *ok = s[2] == 0.0;
        if (!*ok) return PRH_OK;
        return symmetric_product(op, tmp_d, tmp_p, st);
    }
    // F(y) into f: sets *finite and the 2-norm and max-norm of f (inf when not finite).
    int residual(const double *y, double *f, bool *finite, double *norm2, double *normmax)
    {
        bool ok = false;
        int rc = exp_product(y, &ok);
        if (rc) return rc;
        if (ok) {
            double s[3];
            k_log_residual<<<grid, BT, 0, st>>>(y, tmp_p, log_area, n, f, part);
            PRH_HIP(hipGetLastError());
            if ((rc = stats(s))) return rc;
            ok = s[2] == 0.0;
            if (ok) {
                *finite = true;
                *norm2 = std::sqrt(s[1]);
                *normmax = s[0];
                return PRH_OK;
            }
        }
        // the reference returns np.full(n, np.inf)
        k_fill<<<grid, BT, 0, st>>>(f, n, HUGE_VAL);
        PRH_HIP(hipGetLastError());
        *finite = false;
        *norm2 = *normmax = HUGE_VAL;
        return PRH_OK;
    }
    int closure(const double *y, double *err)
    {
        bool ok = false;
        int rc = exp_product(y, &ok);
        if (rc) return rc;
        if (!ok) {
            *err = HUGE_VAL;
            return PRH_OK;
        }
        double s[3];
        k_closure<<<grid, BT, 0, st>>>(tmp_d, tmp_p, op->area, n, part);
        PRH_HIP(hipGetLastError());
        if ((rc = stats(s))) return rc;
        *err = s[2] == 0.0 ? s[0] : HUGE_VAL;
        return PRH_OK;
    }
    int norms(const double *x, double *norm2, double *normmax, bool *finite)
    {
        double s[3];
        k_norms<<<grid, BT, 0, st>>>(x, n, part);
        PRH_HIP(hipGetLastError());
        int rc = stats(s);
        if (rc) return rc;
        *finite = s[2] == 0.0;
        *norm2 = std::sqrt(s[1]);
        *normmax = s[0];
        return PRH_OK;
    }
    int dots(const std::vector<double *> &vs, const double *w, double *out)
    {
        if (vs.empty()) return PRH_OK;
        DotArgs a{};
        a.k = (int)vs.size();
        for (int j = 0; j < a.k; ++j) a.v[j] = vs[j];
        k_dots<<<grid, BT, 0, st>>>(a, w, n, part);
        k_dot_totals<<<a.k, BT, 0, st>>>(part, (int)grid, tot);
        PRH_HIP(hipGetLastError());
        PRH_HIP(hipMemcpyAsync(out, tot, sizeof(double) * a.k, hipMemcpyDeviceToHost, st));
        PRH_HIP(hipStreamSynchronize(st));
        return PRH_OK;
    }
};

// Solves a g = b (k <= HIST) by Gaussian elimination with partial pivoting;
// false for an exactly singular a (LAPACK dgesv's info > 0).
bool small_solve(int k, const double *a_in, const double *b_in, double *g)
{
    double a[HIST][HIST], b[HIST];
    for (int i = 0; i < k; ++i) {
        b[i] = b_in[i];
        for (int j = 0; j < k; ++j) a[i][j] = a_in[i * HIST + j];
    }
    for (int c = 0; c < k; ++c) {
        int p = c;
        for (int r = c + 1; r < k; ++r)
            if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        if (a[p][c] == 0.0) return false;
        if (p != c) {
            for (int j = 0; j < k; ++j) std::swap(a[c][j], a[p][j]);
            std::swap(b[c], b[p]);
        }
        for (int r = c + 1; r < k; ++r) {
            const double l = a[r][c] / a[c][c];
            for (int j = c + 1; j < k; ++j) a[r][j] -= l * a[c][j];
            b[r] -= l * b[c];
        }
    }
    for (int r = k - 1; r >= 0; --r) {
        double s = b[r];
        for (int j = r + 1; j < k; ++j) s -= a[r][j] * g[j];
        g[r] = s / a[r][r];
    }
    for (int i = 0; i < k; ++i)
        if (!std::isfinite(g[i])) return false;
    return true;
}

// The Anderson attempt (nonlin_solve with Anderson and the Armijo search).
// On success y holds the solution and *iterations the callback count;
// AndersonFailed stands for the exceptions _balance catches.
int anderson_solve(Balancer &B, double tol, double *y, int *iterations, bool *failed)
{
    const int64_t n = B.n;
    const unsigned grid = B.grid;
    hipStream_t st = B.st;
    *failed = true;
    double *f = B.vec(), *step = B.vec(), *xt = B.vec(), *ft = B.vec(), *last_x = B.vec(), *last_f = B.vec();
    std::vector<double *> ring_x, ring_f;
    for (int m = 0; m < HIST; ++m) {
        ring_x.push_back(B.vec());
        ring_f.push_back(B.vec());
    }
    for (double *p : ring_x)
        if (!p) return PRH_ERR_HIP;
    for (double *p : ring_f)
        if (!p) return PRH_ERR_HIP;
    if (!f || !step || !xt || !ft || !last_x || !last_f) {
        set_error("prh_operator_balance: out of device memory");
        return PRH_ERR_HIP;
    }
    // x = y0 = 0 (already in y); Fx = F(x); Fx_norm = norm(Fx)
    bool finite = false;
    double f_norm2 = 0.0, f_max = 0.0;
    int rc = B.residual(y, f, &finite, &f_norm2, &f_max);
    if (rc) return rc;
    if (!finite) return PRH_OK;  // scipy.linalg.norm raises on a non-finite F(x0)
    // Jacobian setup: alpha = 0.5 max(||x0||, 1) / ||F0|| with ||x0|| = 0
    double x_norm2 = 0.0, x_max = 0.0;  // norms of the iterate (x0 = 0)
    const double alpha = f_norm2 != 0.0 ? 0.5 * std::max(x_norm2, 1.0) / f_norm2 : 1.0;
    PRH_HIP(hipMemcpyAsync(last_x, y, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    PRH_HIP(hipMemcpyAsync(last_f, f, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    std::vector<double *> hx, hf;  // history, oldest first
    double a[HIST * HIST] = {};
    bool first = true, converged = false;
    int calls = 0;
    for (int it = 0; it < MAX_ITER; ++it) {
        // TerminationCondition.check: max|F| <= f_tol and |dx| / inf <= |x|,
        // which is NaN (false) while dx is the initial inf vector
        if (f_max == 0.0 || (f_max <= tol && !first && x_max >= 0.0)) {
            converged = true;
            break;
        }
        first = false;
        // step = -Anderson.solve(F)
        const int k = (int)hx.size();
        StepArgs sa{};
        sa.k = 0;
        if (k > 0) {
            double df_f[HIST], gamma[HIST];
            if ((rc = B.dots(hf, f, df_f))) return rc;
            if (small_solve(k, a, df_f, gamma)) {
                sa.k = k;
                for (int m = 0; m < k; ++m) {
                    sa.dx[m] = hx[m];
                    sa.df[m] = hf[m];
                    sa.gamma[m] = gamma[m];
                }
            } else {
                // singular: reset the Jacobian approximation
                for (int m = 0; m < k; ++m) {
                    ring_x.push_back(hx[m]);
                    ring_f.push_back(hf[m]);
                }
                hx.clear();
                hf.clear();
            }
        }
        k_anderson_step<<<grid, BT, 0, st>>>(sa, f, alpha, n, step);
        PRH_HIP(hipGetLastError());
        double step_norm2 = 0.0, step_max = 0.0;  // step_max: unused by the check (x_tol = inf)
        bool step_finite = false;
        if ((rc = B.norms(step, &step_norm2, &step_max, &step_finite))) return rc;
        if (!step_finite || step_norm2 == 0.0) return PRH_OK;  // scipy raises (norm / zero step)
        // Armijo search on phi(s) = ||F(x + s step)||^2
        const double phi0 = f_norm2 * f_norm2, derphi0 = -phi0;
        double tmp_s = 0.0, tmp_phi = phi0;
        bool tmp_in_ft = false;  // tmp_Fx is ft (else it is f)
        auto phi = [&](double s, double *out) -> int {
            if (s == tmp_s) {
                *out = tmp_phi;
                return PRH_OK;
            }
            k_axpy<<<grid, BT, 0, st>>>(y, s, step, n, xt);
            PRH_HIP(hipGetLastError());
            bool fin = false;
            double nrm = 0.0, mx = 0.0;
            int r = B.residual(xt, ft, &fin, &nrm, &mx);
            if (r) return r;
            const double p = fin ? nrm * nrm : HUGE_VAL;
            tmp_s = s;
            tmp_phi = p;
            tmp_in_ft = true;
            *out = p;
            return PRH_OK;
        };
        const double c1 = 1e-4, amin = 1e-2;
        double s = 1.0;
        bool found = false;
        {
            double alpha0 = 1.0, phi_a0 = 0.0;
            if ((rc = phi(alpha0, &phi_a0))) return rc;
            if (phi_a0 <= phi0 + c1 * alpha0 * derphi0) {
                s = alpha0;
                found = true;
            } else {
                double alpha1 = -(derphi0)*alpha0 * alpha0 / 2.0 / (phi_a0 - phi0 - derphi0 * alpha0);
                double phi_a1 = 0.0;
                if ((rc = phi(alpha1, &phi_a1))) return rc;
                if (phi_a1 <= phi0 + c1 * alpha1 * derphi0) {
                    s = alpha1;
                    found = true;
                } else {
                    while (alpha1 > amin) {
                        const double factor = alpha0 * alpha0 * alpha1 * alpha1 * (alpha1 - alpha0);
                        double aa = alpha0 * alpha0 * (phi_a1 - phi0 - derphi0 * alpha1) -
                                    alpha1 * alpha1 * (phi_a0 - phi0 - derphi0 * alpha0);
                        aa = aa / factor;
                        double bb = -alpha0 * alpha0 * alpha0 * (phi_a1 - phi0 - derphi0 * alpha1) +
                                    alpha1 * alpha1 * alpha1 * (phi_a0 - phi0 - derphi0 * alpha0);
                        bb = bb / factor;
                        double alpha2 = (-bb + std::sqrt(std::fabs(bb * bb - 3 * aa * derphi0))) / (3.0 * aa);
                        double phi_a2 = 0.0;
                        if ((rc = phi(alpha2, &phi_a2))) return rc;
                        if (phi_a2 <= phi0 + c1 * alpha2 * derphi0) {
                            s = alpha2;
                            found = true;
                            break;
                        }
                        if ((alpha1 - alpha2) > alpha1 / 2.0 || (1 - alpha2 / alpha1) < 0.96) alpha2 = alpha1 / 2.0;
                        alpha0 = alpha1;
                        alpha1 = alpha2;
                        phi_a0 = phi_a1;
                        phi_a1 = phi_a2;
                    }
                }
            }
        }
        if (!found) s = 1.0;  // no step length found: the full step
        // x = x + s step; Fx from the search when it evaluated this s
        const bool have = tmp_in_ft && s == tmp_s;
        k_axpy<<<grid, BT, 0, st>>>(y, s, step, n, xt);
        PRH_HIP(hipGetLastError());
        if (have) {
            std::swap(y, xt);  // y <- x + s step, ft holds its F
            std::swap(f, ft);
        } else {
            std::swap(y, xt);
            if ((rc = B.residual(y, f, &finite, &f_norm2, &f_max))) return rc;
        }
        {
            bool fin = false;
            if ((rc = B.norms(f, &f_norm2, &f_max, &fin))) return rc;
            if (!fin) return PRH_OK;  // Fx_norm = norm(Fx) raises on a non-finite F
            bool xfin = false;
            if ((rc = B.norms(y, &x_norm2, &x_max, &xfin))) return rc;
        }
        // Jacobian update: append (dX, dF), drop the oldest beyond M
        double *nx_ = nullptr, *nf_ = nullptr;
        if (hx.size() == (size_t)HIST) {
            nx_ = hx.front();
            nf_ = hf.front();
            hx.erase(hx.begin());
            hf.erase(hf.begin());
            for (int i = 0; i + 1 < HIST; ++i)
                for (int j = 0; j + 1 < HIST; ++j) a[i * HIST + j] = a[(i + 1) * HIST + (j + 1)];
        } else {
            nx_ = ring_x.back();
            nf_ = ring_f.back();
            ring_x.pop_back();
            ring_f.pop_back();
        }
        k_differences<<<grid, BT, 0, st>>>(y, f, last_x, last_f, nx_, nf_, n);
        PRH_HIP(hipGetLastError());
        hx.push_back(nx_);
        hf.push_back(nf_);
        const int kk = (int)hf.size();
        double g[HIST];
        if ((rc = B.dots(hf, nf_, g))) return rc;
        for (int i = 0; i < kk; ++i) {
            const double v = (i == kk - 1 ? 1.0 + W0 * W0 : 1.0) * g[i];
            a[i * HIST + (kk - 1)] = v;
            a[(kk - 1) * HIST + i] = v;
        }
        calls += 1;  // the callback
    }
    if (!converged) return PRH_OK;  // NoConvergence after maxiter
    *iterations = calls;
    *failed = false;
    // the iterate may live in a swapped buffer: copy it to the caller's y
    if (y != B.vecs[0]) PRH_HIP(hipMemcpyAsync(B.vecs[0], y, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    return PRH_OK;
}

}  // namespace
}  // namespace prh

using namespace prh;

extern "C" int prh_operator_balance(prh_operator *op, double tolerance, int32_t *method, int32_t *iterations)
{
    PRH_CHECK_ARG(op && method && iterations && tolerance > 0.0, "prh_operator_balance: bad argument");
    PRH_HIP(hipSetDevice(op->device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    Balancer B{};
    B.op = op;
    B.st = st;
    B.n = op->n;
    B.grid = blocks_of(op->n);
    B.log_min = std::log(std::numeric_limits<double>::denorm_min());
    B.log_max = std::log(DBL_MAX);
    double *y = B.vec();  // vecs[0]: the iterate
    B.log_area = B.vec();
    B.tmp_d = B.vec();
    B.tmp_p = B.vec();
    if (!y || !B.log_area || !B.tmp_d || !B.tmp_p || hipMalloc((void **)&B.part, sizeof(double) * 3 * HIST * B.grid) != hipSuccess ||
        hipMalloc((void **)&B.tot, sizeof(double) * (3 * HIST + 3)) != hipSuccess) {
        set_error("prh_operator_balance: out of device memory");
        return PRH_ERR_HIP;
    }
    const int64_t n = op->n;
    k_log_area<<<B.grid, BT, 0, st>>>(op->area, n, B.log_area);
    k_fill<<<B.grid, BT, 0, st>>>(y, n, 0.0);
    PRH_HIP(hipGetLastError());
    double err = 0.0;
    if ((rc = B.closure(y, &err))) return rc;
    if (err <= tolerance) {
        *method = 0;
        *iterations = 0;
    } else {
        int its = 0;
        bool failed = true;
        if ((rc = anderson_solve(B, tolerance, y, &its, &failed))) return rc;
        bool done = false;
        if (!failed) {
            const double accept = std::min(std::max(10.0 * tolerance, 10.0 * DBL_EPSILON), CLOSURE_TOLERANCE);
            if ((rc = B.closure(y, &err))) return rc;
            if (err <= accept) {
                *method = 1;
                *iterations = its;
                done = true;
            }
        }
        if (!done) {
            // the damped multiplicative iteration from y0 = 0
            k_fill<<<B.grid, BT, 0, st>>>(y, n, 0.0);
            PRH_HIP(hipGetLastError());
            for (int step = 1; step <= MULT_STEPS && !done; ++step) {
                bool ok = false;
                if ((rc = B.exp_product(y, &ok))) return rc;
                if (!ok) break;
                double s[3];
                k_mult_step<<<B.grid, BT, 0, st>>>(B.tmp_d, B.tmp_p, B.log_area, n, y, B.part);
                PRH_HIP(hipGetLastError());
                if ((rc = B.stats(s))) return rc;
                if (s[2] != 0.0) break;
                if ((rc = B.closure(y, &err))) return rc;
                if (!std::isfinite(err)) break;
                if (err <= std::min(tolerance, CLOSURE_TOLERANCE)) {
                    *method = 2;
                    *iterations = step;
                    done = true;
                }
            }
            if (!done) {
                set_error("diagonal balancing of the exchange matrix did not converge");
                return PRH_ERR_BALANCE;
            }
        }
    }
    k_scale_out<<<B.grid, BT, 0, st>>>(y, n, op->scale, op->scale32);
    PRH_HIP(hipGetLastError());
    PRH_HIP(hipStreamSynchronize(st));
    return PRH_OK;
}

extern "C" int prh_operator_get_scale(const prh_operator *op, double *scale)
{
    PRH_CHECK_ARG(op && scale, "prh_operator_get_scale: null pointer");
    PRH_HIP(hipSetDevice(op->device));
    PRH_HIP(hipMemcpy(scale, op->scale, sizeof(double) * op->n, hipMemcpyDeviceToHost));
    return PRH_OK;
}
