// Device-side layout of an exchange operator (shared by the operator and the
// batched radiosity engine).
#pragma once

#include "prh_internal.h"

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
    prh_batch *batch = nullptr;  // engine of the mixed-precision path, created on first use
    // device time of the FP64 series (HIP events around each series)
    hipEvent_t series_ev[2] = {nullptr, nullptr};
    int64_t series_count = 0, series_steps = 0;
    double series_ms = 0.0;
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
}  // namespace prh
