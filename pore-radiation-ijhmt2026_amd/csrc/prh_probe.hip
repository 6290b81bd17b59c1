// Measurement probes of the batched radiosity sweep (not part of any solve).
//
// prh_probe_gather  times one sweep y = S z over every slot of an interleaved
//                   slot vector, for several slot formats: what one tick of a
//                   batched engine costs with FP32 x 32, FP16 x 64, FP16 x 32,
//                   FP64 x 16 or FP64 x 32 slots per patch.  The sweep has
//                   the structure of the engine's tick (one wave per row, 16 B
//                   gathers, XCD-contiguous rows, FP64 accumulation) and writes
//                   the scaled product back in the slot format.
// prh_probe_tiles   the matrix-core alternative: the 16 x 4 blocks of S that
//                   hold at least min_fill nonzeros become dense tiles,
//                   multiplied into FP32 x 32 slot vectors with
//                   v_mfma_f32_16x16x4f32, and the rest stays a CSR gathered
//                   as before; times both parts against the plain gather and
//                   checks that they add up to the same product.
//
// Both take a device-resident operator in the integer form K (equal areas and
// ray counts).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "prh_internal.h"
#include "prh_operator_impl.h"

namespace prh {
namespace {

constexpr int PT = 256;  // 4 rows (one per wave) per workgroup
constexpr int PROWS = PT / 64;

template <typename T, int PER>
struct alignas(16) Pack {
    T v[PER];
};

__device__ __forceinline__ double to_f64(float x) { return (double)x; }
__device__ __forceinline__ double to_f64(double x) { return x; }
__device__ __forceinline__ double to_f64(_Float16 x) { return (double)(float)x; }
template <typename T>
__device__ __forceinline__ T from_f64(double x)
{
    return (T)x;
}
template <>
__device__ __forceinline__ _Float16 from_f64<_Float16>(double x)
{
    return (_Float16)(float)x;
}

struct ProbeView {
    int64_t n;
    const int64_t *rowptr;
    const int32_t *col;
    const uint16_t *kval;
    const double *d;
    double inv_2R, delta;
    uint32_t blocks_per_xcd;
};

__device__ __forceinline__ int64_t probe_row(uint32_t blocks_per_xcd)
{
    const uint32_t b = blockIdx.x;
    const int64_t first = ((int64_t)(b % 8u) * blocks_per_xcd + (b / 8u)) * PROWS;
    return (int64_t)__builtin_amdgcn_readfirstlane((int)(first + (threadIdx.x >> 6)));
}

// z_out[row][s] = (F_p^T z_in)[row][s] in the slot format T (bounded: F_p^T
// has spectral radius 1), as the engine's tick forms it.  STREAM: the matrix
// (columns, values) is read with non-temporal loads, so its stream does not
// displace lines of the gathered vector from the L2.
template <typename T, int SL, bool STREAM = false>
__global__ __launch_bounds__(PT) void k_probe_gather(ProbeView M, const T *z_in, T *z_out)
{
    constexpr int PER = 16 / (int)sizeof(T);  // slots per 16 B load
    constexpr int HL = SL / PER;               // lanes per column
    constexpr int NG = 64 / HL;                // nonzeros in flight per wave
    constexpr int U = 64 / NG;                 // gathers per lane per 64 nonzeros
    static_assert(NG >= PER, "every slot needs a writing lane");
    const int64_t row = probe_row(M.blocks_per_xcd);
    const int lane = threadIdx.x & 63;
    const int h = lane % HL;
    const int g = lane / HL;
    double acc[PER];
#pragma unroll
    for (int p = 0; p < PER; ++p) acc[p] = 0.0;
    if (row < M.n) {
        const int64_t e0 = M.rowptr[row];
        const uint32_t len = (uint32_t)(M.rowptr[row + 1] - e0);
        const int32_t *col = M.col + e0;
        const uint16_t *kv = M.kval + e0;
        for (uint32_t base = 0; base < len; base += 64u) {
            const uint32_t e = base + (uint32_t)lane;
            const bool ok = e < len;
            const uint32_t ec = ok ? e : len - 1;
            const int32_t jl = STREAM ? __builtin_nontemporal_load(col + ec) : col[ec];
            const uint16_t kl = STREAM ? __builtin_nontemporal_load(kv + ec) : kv[ec];
            const double vl = ok ? (double)kl : 0.0;
            int32_t j[U];
            double v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                j[u] = __shfl(jl, g + NG * u);
                v[u] = __shfl(vl, g + NG * u);
            }
            Pack<T, PER> z[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t off = (uint32_t)j[u] * (uint32_t)(SL * sizeof(T)) + 16u * (uint32_t)h;
                z[u] = *reinterpret_cast<const Pack<T, PER> *>(reinterpret_cast<const char *>(z_in) + off);
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int p = 0; p < PER; ++p) acc[p] = fma(v[u], to_f64(z[u].v[p]), acc[p]);
        }
    }
#pragma unroll
    for (int off = HL; off < 64; off <<= 1)
#pragma unroll
        for (int p = 0; p < PER; ++p) acc[p] += __shfl_xor(acc[p], off);
    if (g < PER && row < M.n) {
        double w = acc[0];
#pragma unroll
        for (int p = 1; p < PER; ++p) w = g == p ? acc[p] : w;
        const size_t at = (size_t)row * SL + (size_t)(PER * h + g);
        const double di = M.d[row];
        const double y = di * (w * M.inv_2R + M.delta * to_f64(z_in[at]));
        z_out[at] = from_f64<T>(y);
    }
}

// Values of order one in every slot.
template <typename T>
__global__ void k_probe_fill(T *z, size_t count)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) z[i] = from_f64<T>(0.25 + (double)((uint32_t)(i * 2654435761u) >> 8) * 0x1p-25);
}

// ------------------------------------------------------------------ tiles
constexpr int TR = 16;  // tile rows
constexpr int TC = 4;   // tile columns
constexpr int TS = 32;  // slots (FP32)

typedef float floatx4 __attribute__((ext_vector_type(4)));

struct TileView {
    int64_t n, n_row_blocks;
    const int64_t *tile_ptr;  // [n_row_blocks + 1]
    const int32_t *tile_col;  // first column / TC of each tile
    const uint16_t *tile_val; // [tile][k][i], k < TC, i < TR
    uint32_t blocks_per_xcd;
};

// y32[row][s] = sum over the row block's tiles of K_tile z (FP32 MFMA
// accumulation).  One wave per row block of 16 rows.  With
// v_mfma_f32_16x16x4f32 lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]
// and holds D[4 (l / 16) + r][l % 16], r < 4.
__global__ __launch_bounds__(PT) void k_probe_tiles(TileView T, const float *z, float *y32)
{
    const int64_t rb = probe_row(T.blocks_per_xcd);
    const int lane = threadIdx.x & 63;
    const int i = lane & 15;
    const int k = lane >> 4;
    floatx4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
    if (rb >= T.n_row_blocks) return;
    const int64_t t0 = T.tile_ptr[rb], t1 = T.tile_ptr[rb + 1];
    for (int64_t t = t0; t < t1; ++t) {
        const float a = (float)T.tile_val[(size_t)t * (TR * TC) + (size_t)(k * TR + i)];
        const int64_t c = (int64_t)T.tile_col[t] * TC + k;
        float b0 = 0.f, b1 = 0.f;
        if (c < T.n) {
            b0 = z[(size_t)c * TS + i];
            b1 = z[(size_t)c * TS + 16 + i];
        }
        acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b1, acc1, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int64_t row = rb * TR + 4 * k + r;
        if (row < T.n) {
            y32[(size_t)row * TS + i] = acc0[r];
            y32[(size_t)row * TS + 16 + i] = acc1[r];
        }
    }
}

// y64[row][s] = (K z)[row][s] over a CSR with uint16 values (FP64 accumulation),
// the 32-slot FP32 gather: 8 lanes x 16 B per column, 8 gathers in flight per lane.
__global__ __launch_bounds__(PT) void k_probe_csr(ProbeView M, const float *z, double *y64)
{
    const int64_t row = probe_row(M.blocks_per_xcd);
    const int lane = threadIdx.x & 63;
    const int h = lane & 7;
    const int g = lane >> 3;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (row < M.n) {
        const int64_t e0 = M.rowptr[row];
        const uint32_t len = (uint32_t)(M.rowptr[row + 1] - e0);
        for (uint32_t base = 0; base < len; base += 64u) {
            const uint32_t e = base + (uint32_t)lane;
            const bool ok = e < len;
            const uint32_t ec = ok ? e : len - 1;
            const int32_t jl = M.col[e0 + ec];
            const double vl = ok ? (double)M.kval[e0 + ec] : 0.0;
            int32_t j[8];
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                j[u] = __shfl(jl, g + 8 * u);
                v[u] = __shfl(vl, g + 8 * u);
            }
            float4 zz[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                zz[u] = *reinterpret_cast<const float4 *>(z + (size_t)j[u] * TS + 4 * h);
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                acc[0] = fma(v[u], (double)zz[u].x, acc[0]);
                acc[1] = fma(v[u], (double)zz[u].y, acc[1]);
                acc[2] = fma(v[u], (double)zz[u].z, acc[2]);
                acc[3] = fma(v[u], (double)zz[u].w, acc[3]);
            }
        }
    }
#pragma unroll
    for (int off = 8; off < 64; off <<= 1)
#pragma unroll
        for (int p = 0; p < 4; ++p) acc[p] += __shfl_xor(acc[p], off);
    if (g < 4 && row < M.n) {
        const double w = g == 0 ? acc[0] : g == 1 ? acc[1] : g == 2 ? acc[2] : acc[3];
        y64[(size_t)row * TS + 4 * h + g] = w;
    }
}

uint32_t xcd_blocks(int64_t units) { return (uint32_t)(((units + PROWS - 1) / PROWS + 7) / 8); }

struct DeviceBuffers {
    std::vector<void *> p;
    ~DeviceBuffers()
    {
        for (void *q : p) (void)hipFree(q);
    }
    template <typename X>
    int get(X **out, size_t count)
    {
        void *q = nullptr;
        if (hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(X)) != hipSuccess) {
            (void)hipGetLastError();
            set_error("probe: out of device memory (%zu bytes)", count * sizeof(X));
            return PRH_ERR_HIP;
        }
        p.push_back(q);
        *out = static_cast<X *>(q);
        return PRH_OK;
    }
};

template <typename T, int SL>
int time_gather(const prh_operator *op, int32_t reps, hipStream_t st, double *ms_per_sweep)
{
    PRH_CHECK_ARG((uint64_t)op->n * SL * sizeof(T) < ((uint64_t)1 << 32),
                  "prh_probe_gather: slot vectors beyond 4 GiB");
    DeviceBuffers buf;
    T *z0 = nullptr, *z1 = nullptr;
    const size_t count = (size_t)op->n * SL;
    int rc = buf.get(&z0, count);
    if (!rc) rc = buf.get(&z1, count);
    if (rc) return rc;
    k_probe_fill<T><<<(unsigned)((count + 255) / 256), 256, 0, st>>>(z0, count);
    const uint32_t bpx = xcd_blocks(op->n);
    const ProbeView M{op->n, op->rowptr, op->col, op->kval, op->scale, op->inv_2R, op->delta, bpx};
    k_probe_gather<T, SL><<<8u * bpx, PT, 0, st>>>(M, z0, z1);  // warm-up
    hipEvent_t ev[2];
    PRH_HIP(hipEventCreate(&ev[0]));
    PRH_HIP(hipEventCreate(&ev[1]));
    PRH_HIP(hipEventRecord(ev[0], st));
    for (int r = 0; r < reps; ++r) {
        k_probe_gather<T, SL><<<8u * bpx, PT, 0, st>>>(M, z0, z1);
        std::swap(z0, z1);
    }
    PRH_HIP(hipEventRecord(ev[1], st));
    PRH_HIP(hipGetLastError());
    PRH_HIP(hipEventSynchronize(ev[1]));
    float ms = 0.f;
    PRH_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    *ms_per_sweep = ms / reps;
    return PRH_OK;
}

// Device ms per sweep of the FP16 x 64 gather over the CSR (rowptr, col, kval)
// with balancing factor d, plain (out[0]) and with streamed matrix loads
// (out[1]); the product of one sweep from z0 is left in y.
int time_fp16x64(int64_t n, const int64_t *rowptr, const int32_t *col, const uint16_t *kval, const double *d,
                 double inv_2R, double delta, const _Float16 *z0, _Float16 *y, _Float16 *z_work, int32_t reps,
                 hipStream_t st, double *out)
{
    const uint32_t bpx = xcd_blocks(n);
    const ProbeView M{n, rowptr, col, kval, d, inv_2R, delta, bpx};
    hipEvent_t ev[3];
    for (auto &e : ev) PRH_HIP(hipEventCreate(&e));
    k_probe_gather<_Float16, 64, false><<<8u * bpx, PT, 0, st>>>(M, z0, y);  // warm-up and the product
    PRH_HIP(hipEventRecord(ev[0], st));
    for (int r = 0; r < reps; ++r) k_probe_gather<_Float16, 64, false><<<8u * bpx, PT, 0, st>>>(M, z0, z_work);
    PRH_HIP(hipEventRecord(ev[1], st));
    for (int r = 0; r < reps; ++r) k_probe_gather<_Float16, 64, true><<<8u * bpx, PT, 0, st>>>(M, z0, z_work);
    PRH_HIP(hipEventRecord(ev[2], st));
    PRH_HIP(hipGetLastError());
    PRH_HIP(hipEventSynchronize(ev[2]));
    for (int k = 0; k < 2; ++k) {
        float ms = 0.f;
        PRH_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
        out[k] = ms / reps;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    return PRH_OK;
}

// Column-panel sweep, phase `panel` of n_panels: every row gathers only its
// entries with columns in the panel (pp[row][panel] .. pp[row][panel + 1]),
// adding to FP64 accumulators kept in HBM between phases; the last phase
// writes the FP16 product.  All rows of a phase read one slab of the gathered
// vector, small enough for each XCD's L2.  Rows are taken in `order` (null:
// stored order).
__global__ __launch_bounds__(PT) void k_probe_panel(ProbeView M, const int64_t *pp, int n_panels, int panel,
                                                    const int32_t *order, const _Float16 *z_in, double *acc,
                                                    _Float16 *z_out)
{
    const uint32_t b = blockIdx.x;
    const int64_t pos = ((int64_t)(b % 8u) * M.blocks_per_xcd + (b / 8u)) * PROWS + (threadIdx.x >> 6);
    const int64_t row =
        (int64_t)__builtin_amdgcn_readfirstlane((int)(pos < M.n && order ? (int64_t)order[pos] : pos));
    const int lane = threadIdx.x & 63;
    const int h = lane & 7;
    const int g = lane >> 3;
    double a[8];
#pragma unroll
    for (int p = 0; p < 8; ++p) a[p] = 0.0;
    if (row < M.n) {
        const int64_t e0 = pp[row * (n_panels + 1) + panel];
        const uint32_t len = (uint32_t)(pp[row * (n_panels + 1) + panel + 1] - e0);
        for (uint32_t base = 0; base < len; base += 64u) {
            const uint32_t e = base + (uint32_t)lane;
            const bool ok = e < len;
            const uint32_t ec = ok ? e : len - 1;
            const int32_t jl = M.col[e0 + ec];
            const double vl = ok ? (double)M.kval[e0 + ec] : 0.0;
            int32_t j[8];
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                j[u] = __shfl(jl, g + 8 * u);
                v[u] = __shfl(vl, g + 8 * u);
            }
            Pack<_Float16, 8> z[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const uint32_t off = (uint32_t)j[u] * 128u + 16u * (uint32_t)h;
                z[u] = *reinterpret_cast<const Pack<_Float16, 8> *>(reinterpret_cast<const char *>(z_in) + off);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
#pragma unroll
                for (int p = 0; p < 8; ++p) a[p] = fma(v[u], (double)(float)z[u].v[p], a[p]);
        }
    }
#pragma unroll
    for (int off = 8; off < 64; off <<= 1)
#pragma unroll
        for (int p = 0; p < 8; ++p) a[p] += __shfl_xor(a[p], off);
    if (row >= M.n) return;
    double w = a[0];
#pragma unroll
    for (int p = 1; p < 8; ++p) w = g == p ? a[p] : w;
    const size_t at = (size_t)row * 64 + (size_t)(8 * h + g);
    const double total = (panel == 0 ? 0.0 : acc[at]) + w;
    if (panel + 1 < n_panels) {
        acc[at] = total;
    } else {
        const double di = M.d[row];
        z_out[at] = from_f64<_Float16>(di * (total * M.inv_2R + M.delta * (double)(float)z_in[at]));
    }
}

// Values of order one in every slot, as a function of the patch's original index.
__global__ void k_probe_fill_perm(_Float16 *z, const int32_t *order, int64_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * 64) return;
    const size_t p = i / 64, s = i % 64;
    const size_t src = order ? (size_t)order[p] * 64 + s : i;
    z[i] = from_f64<_Float16>(0.25 + (double)((uint32_t)(src * 2654435761u) >> 8) * 0x1p-25);
}

}  // namespace
}  // namespace prh

using namespace prh;

extern "C" {

int prh_probe_gather(const prh_operator *op, int32_t format, int32_t reps, double *ms_per_sweep)
{
    PRH_CHECK_ARG(op && ms_per_sweep, "prh_probe_gather: null pointer");
    PRH_CHECK_ARG(op->uniform && op->kval, "prh_probe_gather: the operator has no integer form K");
    PRH_CHECK_ARG(reps >= 1 && reps <= 10000, "prh_probe_gather: reps must be in [1, 10000]");
    PRH_HIP(hipSetDevice(op->device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    switch (format) {
    case PRH_PROBE_F32X32: return time_gather<float, 32>(op, reps, st, ms_per_sweep);
    case PRH_PROBE_F16X64: return time_gather<_Float16, 64>(op, reps, st, ms_per_sweep);
    case PRH_PROBE_F16X32: return time_gather<_Float16, 32>(op, reps, st, ms_per_sweep);
    case PRH_PROBE_F64X16: return time_gather<double, 16>(op, reps, st, ms_per_sweep);
    case PRH_PROBE_F64X32: return time_gather<double, 32>(op, reps, st, ms_per_sweep);
    default: break;
    }
    set_error("prh_probe_gather: unknown slot format %d", format);
    return PRH_ERR_ARG;
}

int prh_probe_order(const prh_operator *op, const int32_t *order, int32_t reps, double *out, int32_t n_out)
{
    PRH_CHECK_ARG(op && order && out, "prh_probe_order: null pointer");
    PRH_CHECK_ARG(op->uniform && op->kval, "prh_probe_order: the operator has no integer form K");
    PRH_CHECK_ARG(reps >= 1 && reps <= 10000, "prh_probe_order: reps must be in [1, 10000]");
    const int64_t n = op->n, nnz = op->nnz_s;
    PRH_CHECK_ARG((uint64_t)n * 128 < ((uint64_t)1 << 32), "prh_probe_order: slot vectors beyond 4 GiB");
    std::vector<int32_t> inv(n, -1);
    for (int64_t p = 0; p < n; ++p) {
        PRH_CHECK_ARG(order[p] >= 0 && order[p] < n && inv[order[p]] < 0, "prh_probe_order: not a permutation");
        inv[order[p]] = (int32_t)p;
    }
    PRH_HIP(hipSetDevice(op->device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    std::vector<int64_t> rowptr(n + 1), prow(n + 1, 0);
    std::vector<int32_t> col(nnz), pcol(nnz);
    std::vector<uint16_t> kval(nnz), pval(nnz);
    std::vector<double> d(n), pd(n);
    PRH_HIP(hipMemcpy(rowptr.data(), op->rowptr, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost));
    PRH_HIP(hipMemcpy(col.data(), op->col, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost));
    PRH_HIP(hipMemcpy(kval.data(), op->kval, sizeof(uint16_t) * nnz, hipMemcpyDeviceToHost));
    PRH_HIP(hipMemcpy(d.data(), op->scale, sizeof(double) * n, hipMemcpyDeviceToHost));
    // row p of the reordered matrix is row order[p], columns renumbered by inv and sorted
    std::vector<std::pair<int32_t, uint16_t>> row;
    for (int64_t p = 0; p < n; ++p) {
        const int64_t i = order[p];
        row.clear();
        for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) row.push_back({inv[col[e]], kval[e]});
        std::sort(row.begin(), row.end());
        prow[p + 1] = prow[p] + (int64_t)row.size();
        for (size_t k = 0; k < row.size(); ++k) {
            pcol[prow[p] + (int64_t)k] = row[k].first;
            pval[prow[p] + (int64_t)k] = row[k].second;
        }
        pd[p] = d[i];
    }
    DeviceBuffers buf;
    int64_t *d_rowptr = nullptr;
    int32_t *d_col = nullptr, *d_order = nullptr;
    uint16_t *d_val = nullptr;
    double *d_d = nullptr;
    _Float16 *z = nullptr, *zp = nullptr, *y = nullptr, *yp = nullptr, *work = nullptr;
    const size_t vec = (size_t)n * 64;
    if ((rc = buf.get(&d_rowptr, n + 1)) || (rc = buf.get(&d_col, nnz)) || (rc = buf.get(&d_val, nnz)) ||
        (rc = buf.get(&d_d, n)) || (rc = buf.get(&d_order, n)) || (rc = buf.get(&z, vec)) ||
        (rc = buf.get(&zp, vec)) || (rc = buf.get(&y, vec)) || (rc = buf.get(&yp, vec)) || (rc = buf.get(&work, vec)))
        return rc;
    PRH_HIP(hipMemcpy(d_rowptr, prow.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice));
    PRH_HIP(hipMemcpy(d_col, pcol.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
    PRH_HIP(hipMemcpy(d_val, pval.data(), sizeof(uint16_t) * nnz, hipMemcpyHostToDevice));
    PRH_HIP(hipMemcpy(d_d, pd.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    PRH_HIP(hipMemcpy(d_order, order, sizeof(int32_t) * n, hipMemcpyHostToDevice));
    const unsigned fill_grid = (unsigned)((vec + 255) / 256);
    k_probe_fill_perm<<<fill_grid, 256, 0, st>>>(z, nullptr, n);
    k_probe_fill_perm<<<fill_grid, 256, 0, st>>>(zp, d_order, n);
    double t[4];
    if ((rc = time_fp16x64(n, op->rowptr, op->col, op->kval, op->scale, op->inv_2R, op->delta, z, y, work, reps, st,
                           t)))
        return rc;
    if ((rc = time_fp16x64(n, d_rowptr, d_col, d_val, d_d, op->inv_2R, op->delta, zp, yp, work, reps, st, t + 2)))
        return rc;
    // the reordered product is the original one, reordered (up to FP16 rounding of the outputs)
    std::vector<_Float16> hy(vec), hyp(vec);
    PRH_HIP(hipMemcpy(hy.data(), y, sizeof(_Float16) * vec, hipMemcpyDeviceToHost));
    PRH_HIP(hipMemcpy(hyp.data(), yp, sizeof(_Float16) * vec, hipMemcpyDeviceToHost));
    double worst = 0.0, largest = 0.0;
    for (int64_t p = 0; p < n; ++p)
        for (int s = 0; s < 64; ++s) {
            const double a = (double)(float)hy[(size_t)order[p] * 64 + s], b = (double)(float)hyp[(size_t)p * 64 + s];
            worst = std::max(worst, std::fabs(a - b));
            largest = std::max(largest, std::fabs(a));
        }
    const double v[] = {t[0], t[1], t[2], t[3], largest > 0.0 ? worst / largest : worst};
    for (int k = 0; k < n_out && k < 5; ++k) out[k] = v[k];
    return PRH_OK;
}

int prh_probe_panels(const prh_operator *op, int32_t n_panels, int32_t reps, double *out, int32_t n_out)
{
    PRH_CHECK_ARG(op && out, "prh_probe_panels: null pointer");
    PRH_CHECK_ARG(op->uniform && op->kval, "prh_probe_panels: the operator has no integer form K");
    PRH_CHECK_ARG(n_panels >= 1 && n_panels <= 64, "prh_probe_panels: n_panels must be in [1, 64]");
    PRH_CHECK_ARG(reps >= 1 && reps <= 10000, "prh_probe_panels: reps must be in [1, 10000]");
    const int64_t n = op->n, nnz = op->nnz_s;
    PRH_CHECK_ARG((uint64_t)n * 128 < ((uint64_t)1 << 32), "prh_probe_panels: slot vectors beyond 4 GiB");
    PRH_HIP(hipSetDevice(op->device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    std::vector<int64_t> rowptr(n + 1);
    std::vector<int32_t> col(nnz);
    PRH_HIP(hipMemcpy(rowptr.data(), op->rowptr, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost));
    PRH_HIP(hipMemcpy(col.data(), op->col, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost));
    // panel k holds the columns [k W, (k + 1) W); rows' columns are sorted
    const int64_t W = (n + n_panels - 1) / n_panels;
    std::vector<int64_t> pp((size_t)n * (n_panels + 1));
    for (int64_t r = 0; r < n; ++r) {
        int64_t e = rowptr[r];
        for (int k = 0; k <= n_panels; ++k) {
            const int64_t lim = std::min<int64_t>((int64_t)k * W, n);
            while (k > 0 && e < rowptr[r + 1] && col[e] < lim) ++e;
            pp[(size_t)r * (n_panels + 1) + k] = k == 0 ? rowptr[r] : (k == n_panels ? rowptr[r + 1] : e);
        }
    }
    DeviceBuffers buf;
    int64_t *d_pp = nullptr;
    double *acc = nullptr;
    _Float16 *z = nullptr, *y_plain = nullptr, *y_panel = nullptr;
    const size_t vec = (size_t)n * 64;
    if ((rc = buf.get(&d_pp, pp.size())) || (rc = buf.get(&acc, vec)) || (rc = buf.get(&z, vec)) ||
        (rc = buf.get(&y_plain, vec)) || (rc = buf.get(&y_panel, vec)))
        return rc;
    PRH_HIP(hipMemcpy(d_pp, pp.data(), sizeof(int64_t) * pp.size(), hipMemcpyHostToDevice));
    k_probe_fill_perm<<<(unsigned)((vec + 255) / 256), 256, 0, st>>>(z, nullptr, n);
    const uint32_t bpx = xcd_blocks(n);
    const ProbeView M{n, op->rowptr, op->col, op->kval, op->scale, op->inv_2R, op->delta, bpx};
    hipEvent_t ev[5];
    for (auto &e : ev) PRH_HIP(hipEventCreate(&e));
    // plain sweep, stored and in the operator's row order; panel sweeps, the same two orders
    const int32_t *orders[2] = {nullptr, op->row_order};
    k_probe_gather<_Float16, 64, false><<<8u * bpx, PT, 0, st>>>(M, z, y_plain);  // warm-up and the product
    PRH_HIP(hipEventRecord(ev[0], st));
    for (int r = 0; r < reps; ++r) k_probe_gather<_Float16, 64, false><<<8u * bpx, PT, 0, st>>>(M, z, y_plain);
    PRH_HIP(hipEventRecord(ev[1], st));
    for (int o = 0; o < 2; ++o) {
        for (int r = 0; r < reps; ++r)
            for (int k = 0; k < n_panels; ++k)
                k_probe_panel<<<8u * bpx, PT, 0, st>>>(M, d_pp, n_panels, k, orders[o], z, acc, y_panel);
        PRH_HIP(hipEventRecord(ev[2 + o], st));
    }
    PRH_HIP(hipGetLastError());
    PRH_HIP(hipEventSynchronize(ev[3]));
    float ms[3];
    for (int k = 0; k < 3; ++k) PRH_HIP(hipEventElapsedTime(&ms[k], ev[k], ev[k + 1]));
    for (auto &e : ev) (void)hipEventDestroy(e);
    std::vector<_Float16> a(vec), bb(vec);
    PRH_HIP(hipMemcpy(a.data(), y_plain, sizeof(_Float16) * vec, hipMemcpyDeviceToHost));
    PRH_HIP(hipMemcpy(bb.data(), y_panel, sizeof(_Float16) * vec, hipMemcpyDeviceToHost));
    double worst = 0.0, largest = 0.0;
    for (size_t q = 0; q < vec; ++q) {
        worst = std::max(worst, std::fabs((double)(float)a[q] - (double)(float)bb[q]));
        largest = std::max(largest, std::fabs((double)(float)a[q]));
    }
    const double v[] = {ms[0] / reps, ms[1] / reps, ms[2] / reps, largest > 0.0 ? worst / largest : worst};
    for (int k = 0; k < n_out && k < 4; ++k) out[k] = v[k];
    return PRH_OK;
}

int prh_probe_tiles(const prh_operator *op, int32_t min_fill, int32_t reps, double *out, int32_t n_out)
{
    PRH_CHECK_ARG(op && out, "prh_probe_tiles: null pointer");
    PRH_CHECK_ARG(op->uniform && op->kval, "prh_probe_tiles: the operator has no integer form K");
    PRH_CHECK_ARG(min_fill >= 1 && min_fill <= TR * TC, "prh_probe_tiles: min_fill must be in [1, 64]");
    PRH_CHECK_ARG(reps >= 1 && reps <= 10000, "prh_probe_tiles: reps must be in [1, 10000]");
    PRH_HIP(hipSetDevice(op->device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    const int64_t n = op->n, nnz = op->nnz_s;
    PRH_CHECK_ARG(n < ((int64_t)1 << 25) && nnz < ((int64_t)1 << 32), "prh_probe_tiles: operator too large");
    std::vector<int64_t> rowptr(n + 1);
    std::vector<int32_t> col(nnz);
    std::vector<uint16_t> kval(nnz);
    PRH_HIP(hipMemcpy(rowptr.data(), op->rowptr, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost));
    PRH_HIP(hipMemcpy(col.data(), op->col, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost));
    PRH_HIP(hipMemcpy(kval.data(), op->kval, sizeof(uint16_t) * nnz, hipMemcpyDeviceToHost));

    // split S into dense 16 x 4 tiles (blocks with >= min_fill nonzeros) and a remainder CSR
    const int64_t n_rb = (n + TR - 1) / TR;
    std::vector<int64_t> tile_ptr(n_rb + 1, 0), rem_ptr(n + 1, 0);
    std::vector<int32_t> tile_col, rem_col;
    std::vector<uint16_t> tile_val, rem_val;
    rem_col.reserve(nnz);
    rem_val.reserve(nnz);
    std::vector<uint64_t> keys;
    std::vector<std::vector<std::pair<int32_t, uint16_t>>> rem_rows(TR);
    for (int64_t rb = 0; rb < n_rb; ++rb) {
        const int64_t r0 = rb * TR, r1 = std::min<int64_t>(r0 + TR, n);
        keys.clear();
        for (int64_t r = r0; r < r1; ++r)
            for (int64_t e = rowptr[r]; e < rowptr[r + 1]; ++e)
                keys.push_back(((uint64_t)(uint32_t)col[e] << 36) | ((uint64_t)(r - r0) << 32) | (uint64_t)e);
        std::sort(keys.begin(), keys.end());
        for (auto &v : rem_rows) v.clear();
        size_t a = 0;
        while (a < keys.size()) {
            const uint64_t cb = (keys[a] >> 36) / TC;
            size_t b = a;
            while (b < keys.size() && (keys[b] >> 36) / TC == cb) ++b;
            if ((int64_t)(b - a) >= min_fill) {
                const size_t base = tile_val.size();
                tile_val.resize(base + TR * TC, 0);
                tile_col.push_back((int32_t)cb);
                for (size_t q = a; q < b; ++q) {
                    const int kk = (int)((keys[q] >> 36) % TC);
                    const int ii = (int)((keys[q] >> 32) & 0xF);
                    tile_val[base + kk * TR + ii] = kval[keys[q] & 0xFFFFFFFFull];
                }
            } else {
                for (size_t q = a; q < b; ++q) {
                    const int ii = (int)((keys[q] >> 32) & 0xF);
                    const uint64_t e = keys[q] & 0xFFFFFFFFull;
                    rem_rows[ii].push_back({col[e], kval[e]});
                }
            }
            a = b;
        }
        tile_ptr[rb + 1] = (int64_t)tile_col.size();
        for (int64_t r = r0; r < r1; ++r) {
            for (const auto &cv : rem_rows[r - r0]) {
                rem_col.push_back(cv.first);
                rem_val.push_back(cv.second);
            }
            rem_ptr[r + 1] = (int64_t)rem_col.size();
        }
    }
    const int64_t n_tiles = (int64_t)tile_col.size(), n_rem = (int64_t)rem_col.size();

    DeviceBuffers buf;
    int64_t *d_tile_ptr = nullptr, *d_rem_ptr = nullptr;
    int32_t *d_tile_col = nullptr, *d_rem_col = nullptr;
    uint16_t *d_tile_val = nullptr, *d_rem_val = nullptr;
    float *z = nullptr, *y32 = nullptr;
    double *y_full = nullptr, *y_rem = nullptr;
    const size_t vec = (size_t)n * TS;
    if ((rc = buf.get(&d_tile_ptr, n_rb + 1)) || (rc = buf.get(&d_rem_ptr, n + 1)) ||
        (rc = buf.get(&d_tile_col, n_tiles)) || (rc = buf.get(&d_rem_col, n_rem)) ||
        (rc = buf.get(&d_tile_val, (size_t)n_tiles * TR * TC)) || (rc = buf.get(&d_rem_val, n_rem)) ||
        (rc = buf.get(&z, vec)) || (rc = buf.get(&y32, vec)) || (rc = buf.get(&y_full, vec)) ||
        (rc = buf.get(&y_rem, vec)))
        return rc;
    PRH_HIP(hipMemcpy(d_tile_ptr, tile_ptr.data(), sizeof(int64_t) * (n_rb + 1), hipMemcpyHostToDevice));
    PRH_HIP(hipMemcpy(d_rem_ptr, rem_ptr.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice));
    if (n_tiles) {
        PRH_HIP(hipMemcpy(d_tile_col, tile_col.data(), sizeof(int32_t) * n_tiles, hipMemcpyHostToDevice));
        PRH_HIP(hipMemcpy(d_tile_val, tile_val.data(), sizeof(uint16_t) * tile_val.size(), hipMemcpyHostToDevice));
    }
    if (n_rem) {
        PRH_HIP(hipMemcpy(d_rem_col, rem_col.data(), sizeof(int32_t) * n_rem, hipMemcpyHostToDevice));
        PRH_HIP(hipMemcpy(d_rem_val, rem_val.data(), sizeof(uint16_t) * n_rem, hipMemcpyHostToDevice));
    }
    k_probe_fill<float><<<(unsigned)((vec + 255) / 256), 256, 0, st>>>(z, vec);

    const uint32_t bpx = xcd_blocks(n), bpx_t = xcd_blocks(n_rb);
    const ProbeView full{n, op->rowptr, op->col, op->kval, op->scale, op->inv_2R, op->delta, bpx};
    const ProbeView rem{n, d_rem_ptr, d_rem_col, d_rem_val, op->scale, op->inv_2R, op->delta, bpx};
    const TileView tiles{n, n_rb, d_tile_ptr, d_tile_col, d_tile_val, bpx_t};
    hipEvent_t ev[4];
    for (auto &e : ev) PRH_HIP(hipEventCreate(&e));
    k_probe_csr<<<8u * bpx, PT, 0, st>>>(full, z, y_full);  // warm-up
    PRH_HIP(hipEventRecord(ev[0], st));
    for (int r = 0; r < reps; ++r) k_probe_csr<<<8u * bpx, PT, 0, st>>>(full, z, y_full);
    PRH_HIP(hipEventRecord(ev[1], st));
    for (int r = 0; r < reps; ++r) k_probe_tiles<<<8u * bpx_t, PT, 0, st>>>(tiles, z, y32);
    PRH_HIP(hipEventRecord(ev[2], st));
    for (int r = 0; r < reps; ++r) k_probe_csr<<<8u * bpx, PT, 0, st>>>(rem, z, y_rem);
    PRH_HIP(hipEventRecord(ev[3], st));
    PRH_HIP(hipGetLastError());
    PRH_HIP(hipEventSynchronize(ev[3]));
    float ms[3];
    for (int k = 0; k < 3; ++k) PRH_HIP(hipEventElapsedTime(&ms[k], ev[k], ev[k + 1]));
    for (auto &e : ev) (void)hipEventDestroy(e);

    // the two parts add up to the plain product (FP32 tile accumulation)
    std::vector<double> hf(vec), hr(vec);
    std::vector<float> ht(vec);
    PRH_HIP(hipMemcpy(hf.data(), y_full, sizeof(double) * vec, hipMemcpyDeviceToHost));
    PRH_HIP(hipMemcpy(hr.data(), y_rem, sizeof(double) * vec, hipMemcpyDeviceToHost));
    PRH_HIP(hipMemcpy(ht.data(), y32, sizeof(float) * vec, hipMemcpyDeviceToHost));
    double worst = 0.0, largest = 0.0;
    for (size_t q = 0; q < vec; ++q) {
        worst = std::max(worst, std::fabs(hf[q] - (hr[q] + (double)ht[q])));
        largest = std::max(largest, std::fabs(hf[q]));
    }
    const double v[] = {ms[0] / reps, ms[1] / reps, ms[2] / reps, largest > 0.0 ? worst / largest : worst,
                        (double)n_tiles, (double)(nnz - n_rem) / (double)std::max<int64_t>(nnz, 1),
                        (double)n_rem};
    for (int k = 0; k < n_out && k < (int)(sizeof(v) / sizeof(v[0])); ++k) out[k] = v[k];
    return PRH_OK;
}

}  // extern "C"
