// Device-side layout of an exchange operator (shared by the operator and the
// batched radiosity engine).
#pragma once

#include "prh_internal.h"

#include <algorithm>
#include <cmath>

namespace prh {
// Schedule of the mixed-precision iteration, shared by the batched engine and
// the single-emission solve: epochs of FP32 Chebyshev sweeps on
// [1 - rho_max, 1 + rho_max], each ended by an FP64 refinement.
constexpr int MAX_TICKS = 64;
constexpr int MAX_EPOCHS = 40;
constexpr double INNER_REDUCTION = 1.0e-6;

// Chebyshev iterations that shrink the error bound 2 c^k / (1 + c^2k) below
// tau (at least 2).  Host and device (the batched engine schedules its slots
// on the device).
__host__ __device__ inline int ticks_for(double rho_max, double tau)
{
    if (rho_max <= 0.0) return 2;
    const double kappa = (1.0 + rho_max) / (1.0 - rho_max);
    const double c = (sqrt(kappa) - 1.0) / (sqrt(kappa) + 1.0);
    const int m = (int)ceil(log(tau / 2.0) / log(c));
    return m < 2 ? 2 : (m > MAX_TICKS ? MAX_TICKS : m);
}

// Reduction an epoch aims for: what is left to reach the tolerance, with a
// factor 10 of margin, but no finer than the iterations resolve (inner:
// INNER_REDUCTION with FP32 directions).
__host__ __device__ inline double epoch_target(double rtol, double e_norm, double r_norm,
                                               double inner = INNER_REDUCTION)
{
    const double left = 0.1 * rtol * e_norm / (r_norm > 1e-300 ? r_norm : 1e-300);
    const double t = left > inner ? left : inner;
    return t < 0.5 ? t : 0.5;
}

// alpha_k, beta_k of Chebyshev acceleration on [1 - delta, 1 + delta]:
// rho_0 = delta, rho_{k+1} = 1 / (2/delta - rho_k),
// dir_{k+1} = rho_{k+1} rho_k dir_k + (2 rho_{k+1} / delta) r_{k+1}.
__host__ __device__ inline void cheb_coefficient(double delta, int k, float *alpha, float *beta)
{
    if (delta <= 0.0) {
        *alpha = 0.f;
        *beta = 1.f;
        return;
    }
    double rho = delta, next = 0.0;
    for (int i = 0; i <= k; ++i) {
        next = 1.0 / (2.0 / delta - rho);
        if (i < k) rho = next;
    }
    *alpha = (float)(next * rho);
    *beta = (float)(2.0 * next / delta);
}

}  // namespace prh

namespace prh {

struct SeriesState {
    double initial;   // max(||E||_1, tiny)
    double l1;        // ||flight||_1 of the last reflection
    double residual;  // sum(flight) at convergence
    int32_t done;
    int32_t iterations;
};

}  // namespace prh

// HBM layout of one operator:
//   rowptr/col/val   symmetric S = 0.5 (diag(A) F + F^T diag(A)) + delta diag(A), FP64 CSR
//                    over union(C, C^T, diagonal)
//   kval             same sparsity, exact integer K_ij = C_ij + C_ji (uint16) when all
//                    areas and all ray counts are equal, so S = A/(2R) K + A delta I
//   val32            FP32 copy of S when kval is not available
//   c_* / t_*        raw counts by origin (C) and by receiver (C^T), int32 CSRs
//   scale / scale32  balancing factor d (FP64 and FP32)
struct prh_operator {
    int device = 0;
    int64_t n = 0;
    // symmetric S (FP64 CSR)
    int64_t nnz_s = 0;
    int64_t *rowptr = nullptr;
    int32_t *col = nullptr;
    double *val = nullptr;
    // raw counts: C by origin and C^T by receiver
    int64_t nnz_c = 0;
    int64_t *c_rowptr = nullptr, *t_rowptr = nullptr;
    int32_t *c_col = nullptr, *c_cnt = nullptr, *t_col = nullptr, *t_cnt = nullptr;
    double *rays = nullptr, *area = nullptr, *scale = nullptr;
    double delta = 0.0;
    // work vectors (n each)
    double *vx = nullptr, *vy = nullptr, *vt = nullptr;
    // series work
    double *flight = nullptr, *absorbed = nullptr, *w0 = nullptr, *w1 = nullptr, *eps = nullptr, *emit = nullptr;
    double *part_abs = nullptr, *part_sum = nullptr;
    prh::SeriesState *state = nullptr;
    // compact forms for the batched engine
    bool uniform = false;      // equal areas and equal ray counts
    double inv_2R = 0.0;       // 1 / (2 R) when uniform
    uint16_t *kval = nullptr;  // K = C + C^T on the sparsity of S (uniform only)
    float *val32 = nullptr;    // FP32 S (non-uniform only)
    float *scale32 = nullptr;
    // order in which the batched engine sweeps the rows (a permutation of the
    // patches, spatially coherent), or null for the stored order
    int32_t *row_order = nullptr;
    prh_batch *batch = nullptr;  // engine of the mixed-precision path, created on first use
    // single-emission mixed-precision solve (mixed_solve_single), allocated on first use:
    // FP32 x, r, dir, z[2] (n each) and FP64 x, y, r (n each) + one total
    float *mix32 = nullptr;
    double *mix64 = nullptr;
    // device time of the FP64 series (HIP events around each series)
    hipEvent_t series_ev[2] = {nullptr, nullptr};
    int64_t series_count = 0, series_steps = 0;
    double series_ms = 0.0;
    // device time of the batched FP64 series: passes, slot-weighted passes (sum of FS), ms
    int64_t batched_passes = 0, batched_slot_passes = 0;
    double batched_ms = 0.0;
};

extern "C" int prh_radiosity_mixed(prh_operator *op, int32_t n_rhs, const double *emitted, const double *emissivity,
                                   int32_t emissivity_per_rhs, double relative_tolerance, double *absorbed,
                                   double *residual, int32_t *iterations);


namespace prh {
// FP64 reflection series of the emission in op->emit with emissivity op->eps
// (device buffers, n each); the absorbed power is left in op->absorbed.
// PRH_ERR_SERIES when more than max_reflections are needed.
int series_fp64(prh_operator *op, double relative_tolerance, int32_t max_reflections, hipStream_t st,
                double *residual, int32_t *iterations);
// Work vectors and series state of an operator whose matrices are in place.
int operator_alloc_work(prh_operator *op);
// Every device buffer of an operator (not the struct itself).
void operator_free(prh_operator *op);
// The same contract by the batched engine's mixed-precision iteration for one
// device-resident emission E (n doubles, read on stream st) with a single
// emissivity: FP32 Chebyshev sweeps with FP64 refinement until
// ||E - (I - (1 - eps) F_p^T) J||_1 <= rtol ||E||_1; eps F_p^T J is written to
// `absorbed` (device).  The engine must not be driven from another thread
// meanwhile.
int batch_solve_device(prh_batch *b, const double *E, double emissivity, double rtol, hipStream_t st,
                       double *absorbed, int32_t *iterations);
// The same iteration for one emission without the engine's slot interleaving:
// the FP32 vectors are contiguous (4 B gathers from an L2-resident vector),
// one wavefront per row as in the FP64 series, and the host reads one norm
// per epoch.  Same contract as batch_solve_device.
int mixed_solve_single(prh_operator *op, const double *E, double emissivity, double rtol, hipStream_t st,
                       double *absorbed, int32_t *iterations);
// y = S x (symmetric_matvec) of device vectors, queued on st.
int symmetric_product(const prh_operator *op, const double *x, double *y, hipStream_t st);
}  // namespace prh
