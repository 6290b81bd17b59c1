// Exchange-operator assembly on the device, from a device edge table.
//
// The same operator as prh_operator_create builds on the host
// (ExchangeOperator.__init__, porerad/exchange.py:47-76): the row totals of
// the counts are checked against the ray counts, C^T is formed, C and C^T are
// merged into the symmetric CSR of
//     S = 0.5 (diag(A) F + F^T diag(A)) + delta diag(A),   F_ij = C_ij / R_i,
// over union(C, C^T, diagonal), with the integer K = C + C^T beside it.  Every
// S_ij is the host's FP64 expression evaluated in the same order (no
// contraction), so both builds hold the same bits.
//
// Steps (one stream, rocPRIM device primitives):
//   1. validate the edges (ranges, strictly increasing (origin, receiver)
//      keys) and histogram origins and receivers; sum the counts per origin;
//   2. row pointers of C and C^T by exclusive scans of the histograms;
//   3. C^T by a radix sort of (receiver << 32 | origin) keys carrying the
//      counts;
//   4. S's sparsity by merging the C keys, the C^T keys and the diagonal
//      keys (three sorted sequences) and reducing equal keys: a C entry
//      carries C_ij in the low 32 bits of its value, a C^T entry carries C_ji
//      in the high 32 bits, so each reduced value holds both counts;
//   5. row pointers of S from the row boundaries of the reduced keys (every
//      row holds its diagonal), then columns, FP64 values and K per entry.
#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include <rocprim/device/device_merge.hpp>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce_by_key.hpp>
#include <rocprim/device/device_scan.hpp>

#include "prh_internal.h"
#include "prh_operator_impl.h"

namespace prh {
namespace {

constexpr int AT = 256;

unsigned grid_of(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + AT - 1) / AT, 1 << 20)); }

// Validation flags (bits) and histograms.
enum : uint32_t { BAD_RANGE = 1, BAD_ORDER = 2 };

__global__ __launch_bounds__(AT) void k_edges_check(const int32_t *origin, const int32_t *receiver,
                                                    const int32_t *count, int64_t nnz, int64_t n, uint32_t *flags,
                                                    uint32_t *hist_c, uint32_t *hist_t,
                                                    unsigned long long *row_total)
{
    for (int64_t e = (int64_t)blockIdx.x * AT + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * AT) {
        const int32_t o = origin[e], r = receiver[e], c = count[e];
        if (o < 0 || o >= n || r < 0 || r >= n || c < 0) {
            atomicOr(flags, BAD_RANGE);
            continue;
        }
        if (e > 0) {
            const int32_t po = origin[e - 1], pr = receiver[e - 1];
            if (!(po < o || (po == o && pr < r))) atomicOr(flags, BAD_ORDER);
        }
        atomicAdd(hist_c + o, 1u);
        atomicAdd(hist_t + r, 1u);
        atomicAdd(row_total + o, (unsigned long long)c);
    }
}

// Row totals against the ray counts; positive ray counts.
__global__ __launch_bounds__(AT) void k_rows_check(const unsigned long long *row_total, const int64_t *rays, int64_t n,
                                                   uint32_t *mismatch, uint32_t *nonpositive)
{
    for (int64_t i = (int64_t)blockIdx.x * AT + threadIdx.x; i < n; i += (int64_t)gridDim.x * AT) {
        if ((long long)row_total[i] != (long long)rays[i]) atomicOr(mismatch, 1u);
        if (rays[i] <= 0) atomicOr(nonpositive, 1u);
    }
}

// Keys and values of C (forward counts) and of the C^T sort input.
__global__ __launch_bounds__(AT) void k_edge_keys(const int32_t *origin, const int32_t *receiver, const int32_t *count,
                                                  int64_t nnz, uint64_t *key_c, uint64_t *val_c, uint64_t *key_t,
                                                  uint32_t *val_t)
{
    for (int64_t e = (int64_t)blockIdx.x * AT + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * AT) {
        const uint64_t o = (uint32_t)origin[e], r = (uint32_t)receiver[e];
        key_c[e] = o << 32 | r;
        val_c[e] = (uint32_t)count[e];
        key_t[e] = r << 32 | o;
        val_t[e] = (uint32_t)count[e];
    }
}

// Columns and counts of C^T, and its merge values (C_ji in the high half).
__global__ __launch_bounds__(AT) void k_transpose_out(const uint64_t *key_t, const uint32_t *val_t, int64_t nnz,
                                                      int32_t *t_col, int32_t *t_cnt, uint64_t *val_tm)
{
    for (int64_t e = (int64_t)blockIdx.x * AT + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * AT) {
        t_col[e] = (int32_t)(uint32_t)key_t[e];
        t_cnt[e] = (int32_t)val_t[e];
        val_tm[e] = (uint64_t)val_t[e] << 32;
    }
}

__global__ __launch_bounds__(AT) void k_diagonal_keys(int64_t n, uint64_t *key, uint64_t *val)
{
    for (int64_t i = (int64_t)blockIdx.x * AT + threadIdx.x; i < n; i += (int64_t)gridDim.x * AT) {
        key[i] = (uint64_t)i << 32 | (uint64_t)i;
        val[i] = 0;
    }
}

// Row starts of S: entry u opens row (key >> 32) when the previous entry is
// in another row; s_rowptr[n] = nnz_s.
__global__ __launch_bounds__(AT) void k_row_starts(const uint64_t *key, int64_t nnz_s, int64_t n, int64_t *rowptr)
{
    for (int64_t u = (int64_t)blockIdx.x * AT + threadIdx.x; u < nnz_s; u += (int64_t)gridDim.x * AT) {
        const uint32_t row = (uint32_t)(key[u] >> 32);
        if (u == 0 || (uint32_t)(key[u - 1] >> 32) != row) rowptr[row] = u;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) rowptr[n] = nnz_s;
}

// Columns, FP64 values and K of S (the host's expression, operation for operation).
__global__ __launch_bounds__(AT) void k_s_values(const uint64_t *key, const uint64_t *cnt, int64_t nnz_s,
                                                 const double *area, const double *rays, double delta, int32_t *col,
                                                 double *val, uint32_t *k, unsigned int *kmax)
{
    uint32_t m = 0;
    for (int64_t u = (int64_t)blockIdx.x * AT + threadIdx.x; u < nnz_s; u += (int64_t)gridDim.x * AT) {
        const int64_t i = (int64_t)(key[u] >> 32);
        const int32_t j = (int32_t)(uint32_t)key[u];
        const uint32_t cij = (uint32_t)cnt[u], cji = (uint32_t)(cnt[u] >> 32);
        // 0.5 (A_i F_ij + A_j F_ji) + delta A_i [i == j]
        const double fij = (double)cij / rays[i];
        const double fji = (double)cji / rays[j];
        double v = 0.5 * (area[i] * fij + area[j] * fji);
        if (j == i) v += delta * area[i];
        col[u] = j;
        val[u] = v;
        k[u] = cij + cji;
        m = std::max(m, cij + cji);
    }
    atomicMax(kmax, m);
}

__global__ __launch_bounds__(AT) void k_narrow(const uint32_t *k, const double *val, int64_t nnz_s, int uniform,
                                               uint16_t *k16, float *v32)
{
    for (int64_t u = (int64_t)blockIdx.x * AT + threadIdx.x; u < nnz_s; u += (int64_t)gridDim.x * AT) {
        if (uniform) k16[u] = (uint16_t)k[u];
        else v32[u] = (float)val[u];
    }
}

__global__ __launch_bounds__(AT) void k_fill_f64(double *x, int64_t n, double v)
{
    for (int64_t i = (int64_t)blockIdx.x * AT + threadIdx.x; i < n; i += (int64_t)gridDim.x * AT) x[i] = v;
}

__global__ __launch_bounds__(AT) void k_fill_f32(float *x, int64_t n, float v)
{
    for (int64_t i = (int64_t)blockIdx.x * AT + threadIdx.x; i < n; i += (int64_t)gridDim.x * AT) x[i] = v;
}

__global__ __launch_bounds__(AT) void k_i64_to_f64(const int64_t *x, int64_t n, double *y)
{
    for (int64_t i = (int64_t)blockIdx.x * AT + threadIdx.x; i < n; i += (int64_t)gridDim.x * AT) y[i] = (double)x[i];
}

// Device buffers released together when the build ends, whatever happens.
struct Scratch {
    std::vector<void *> held;
    ~Scratch()
    {
        for (void *p : held) (void)hipFree(p);
    }
    template <typename T>
    int alloc(T **p, int64_t count)
    {
        *p = nullptr;
        PRH_HIP(hipMalloc((void **)p, sizeof(T) * (size_t)std::max<int64_t>(count, 1)));
        held.push_back(*p);
        return PRH_OK;
    }
};

template <typename T>
int alloc_into(T **p, int64_t count)
{
    PRH_HIP(hipMalloc((void **)p, sizeof(T) * (size_t)std::max<int64_t>(count, 1)));
    return PRH_OK;
}

int build(prh_operator *op, const prh_edges *e, const int64_t *rays_host, const double *area_host, hipStream_t st)
{
    const int64_t n = op->n, nnz = e->nnz;
    Scratch S;
    int rc;
    // inputs of the per-patch tables
    int64_t *rays_i = nullptr;
    if ((rc = S.alloc(&rays_i, n))) return rc;
    if ((rc = alloc_into(&op->rays, n)) || (rc = alloc_into(&op->area, n)) || (rc = alloc_into(&op->scale, n)))
        return rc;
    PRH_HIP(hipMemcpyAsync(rays_i, rays_host, sizeof(int64_t) * n, hipMemcpyHostToDevice, st));
    PRH_HIP(hipMemcpyAsync(op->area, area_host, sizeof(double) * n, hipMemcpyHostToDevice, st));
    k_i64_to_f64<<<grid_of(n), AT, 0, st>>>(rays_i, n, op->rays);
    k_fill_f64<<<grid_of(n), AT, 0, st>>>(op->scale, n, 1.0);

    // 1. validation, histograms, row totals
    uint32_t *flags = nullptr, *hist_c = nullptr, *hist_t = nullptr;
    unsigned long long *row_total = nullptr;
    if ((rc = S.alloc(&flags, 4)) || (rc = S.alloc(&hist_c, n + 1)) || (rc = S.alloc(&hist_t, n + 1)) ||
        (rc = S.alloc(&row_total, n)))
        return rc;
    PRH_HIP(hipMemsetAsync(flags, 0, 4 * sizeof(uint32_t), st));
    PRH_HIP(hipMemsetAsync(hist_c, 0, (n + 1) * sizeof(uint32_t), st));
    PRH_HIP(hipMemsetAsync(hist_t, 0, (n + 1) * sizeof(uint32_t), st));
    PRH_HIP(hipMemsetAsync(row_total, 0, n * sizeof(unsigned long long), st));
    if (nnz > 0)
        k_edges_check<<<grid_of(nnz), AT, 0, st>>>(e->origin, e->receiver, e->count, nnz, n, flags, hist_c, hist_t,
                                                   row_total);
    k_rows_check<<<grid_of(n), AT, 0, st>>>(row_total, rays_i, n, flags + 1, flags + 2);
    PRH_HIP(hipGetLastError());
    uint32_t hflags[4] = {0, 0, 0, 0};
    PRH_HIP(hipMemcpyAsync(hflags, flags, sizeof(hflags), hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));
    PRH_CHECK_ARG(!(hflags[0] & BAD_RANGE), "prh_operator_create_device: an edge is out of range");
    PRH_CHECK_ARG(!(hflags[0] & BAD_ORDER),
                  "prh_operator_create_device: edges must be sorted by (origin, receiver) and unique");
    if (hflags[1]) {
        set_error("first-hit counts do not account for every emitted ray");
        return PRH_ERR_COUNTS;
    }
    PRH_CHECK_ARG(!hflags[2], "prh_operator_create_device: rays_per_origin must be positive");

    // 2. row pointers of C and C^T
    size_t tmp_bytes = 0, need = 0;
    if ((rc = alloc_into(&op->c_rowptr, n + 1)) || (rc = alloc_into(&op->t_rowptr, n + 1))) return rc;
    PRH_HIP(rocprim::exclusive_scan(nullptr, need, hist_c, op->c_rowptr, (int64_t)0, (size_t)(n + 1),
                                    rocprim::plus<int64_t>(), st));
    tmp_bytes = std::max(tmp_bytes, need);

    // 3. C^T: radix sort of (receiver, origin) keys with their counts
    uint64_t *key_c = nullptr, *val_c = nullptr, *key_t = nullptr, *key_ts = nullptr, *val_tm = nullptr;
    uint32_t *val_t = nullptr, *val_ts = nullptr;
    if ((rc = S.alloc(&key_c, nnz)) || (rc = S.alloc(&val_c, nnz)) || (rc = S.alloc(&key_t, nnz)) ||
        (rc = S.alloc(&key_ts, nnz)) || (rc = S.alloc(&val_t, nnz)) || (rc = S.alloc(&val_ts, nnz)))
        return rc;
    if (nnz > 0)
        k_edge_keys<<<grid_of(nnz), AT, 0, st>>>(e->origin, e->receiver, e->count, nnz, key_c, val_c, key_t, val_t);
    int bits = 1;
    while ((int64_t(1) << bits) < n) ++bits;
    PRH_HIP(rocprim::radix_sort_pairs(nullptr, need, key_t, key_ts, val_t, val_ts, (size_t)nnz, 0u,
                                      (unsigned)(32 + bits), st));
    tmp_bytes = std::max(tmp_bytes, need);

    // 4. merge C, C^T and the diagonal; reduce equal keys
    const int64_t n_cat = 2 * nnz + n;
    uint64_t *key_d = nullptr, *val_d = nullptr, *key_m1 = nullptr, *val_m1 = nullptr, *key_m = nullptr,
             *val_m = nullptr, *key_s = nullptr, *cnt_s = nullptr;
    int64_t *n_unique = nullptr;
    if ((rc = S.alloc(&key_d, n)) || (rc = S.alloc(&val_d, n)) || (rc = S.alloc(&key_m1, 2 * nnz)) ||
        (rc = S.alloc(&val_m1, 2 * nnz)) || (rc = S.alloc(&key_m, n_cat)) || (rc = S.alloc(&val_m, n_cat)) ||
        (rc = S.alloc(&key_s, n_cat)) || (rc = S.alloc(&cnt_s, n_cat)) || (rc = S.alloc(&n_unique, 1)) ||
        (rc = S.alloc(&val_tm, nnz)))
        return rc;
    PRH_HIP(rocprim::merge(nullptr, need, key_c, key_ts, key_m1, val_c, val_tm, val_m1, (size_t)nnz, (size_t)nnz,
                           rocprim::less<uint64_t>(), st));
    tmp_bytes = std::max(tmp_bytes, need);
    PRH_HIP(rocprim::merge(nullptr, need, key_m1, key_d, key_m, val_m1, val_d, val_m, (size_t)(2 * nnz), (size_t)n,
                           rocprim::less<uint64_t>(), st));
    tmp_bytes = std::max(tmp_bytes, need);
    PRH_HIP(rocprim::reduce_by_key(nullptr, need, key_m, val_m, (size_t)n_cat, key_s, cnt_s, n_unique,
                                   rocprim::plus<uint64_t>(), rocprim::equal_to<uint64_t>(), st));
    tmp_bytes = std::max(tmp_bytes, need);
    void *tmp = nullptr;
    if ((rc = S.alloc((char **)&tmp, (int64_t)tmp_bytes))) return rc;

    need = tmp_bytes;
    PRH_HIP(rocprim::exclusive_scan(tmp, need, hist_c, op->c_rowptr, (int64_t)0, (size_t)(n + 1),
                                    rocprim::plus<int64_t>(), st));
    need = tmp_bytes;
    PRH_HIP(rocprim::exclusive_scan(tmp, need, hist_t, op->t_rowptr, (int64_t)0, (size_t)(n + 1),
                                    rocprim::plus<int64_t>(), st));
    need = tmp_bytes;
    PRH_HIP(rocprim::radix_sort_pairs(tmp, need, key_t, key_ts, val_t, val_ts, (size_t)nnz, 0u, (unsigned)(32 + bits),
                                      st));
    op->nnz_c = nnz;
    if ((rc = alloc_into(&op->c_col, nnz)) || (rc = alloc_into(&op->c_cnt, nnz)) ||
        (rc = alloc_into(&op->t_col, nnz)) || (rc = alloc_into(&op->t_cnt, nnz)))
        return rc;
    if (nnz > 0) {
        PRH_HIP(hipMemcpyAsync(op->c_col, e->receiver, sizeof(int32_t) * nnz, hipMemcpyDeviceToDevice, st));
        PRH_HIP(hipMemcpyAsync(op->c_cnt, e->count, sizeof(int32_t) * nnz, hipMemcpyDeviceToDevice, st));
        k_transpose_out<<<grid_of(nnz), AT, 0, st>>>(key_ts, val_ts, nnz, op->t_col, op->t_cnt, val_tm);
    }
    k_diagonal_keys<<<grid_of(n), AT, 0, st>>>(n, key_d, val_d);
    PRH_HIP(hipGetLastError());
    need = tmp_bytes;
    PRH_HIP(rocprim::merge(tmp, need, key_c, key_ts, key_m1, val_c, val_tm, val_m1, (size_t)nnz, (size_t)nnz,
                           rocprim::less<uint64_t>(), st));
    need = tmp_bytes;
    PRH_HIP(rocprim::merge(tmp, need, key_m1, key_d, key_m, val_m1, val_d, val_m, (size_t)(2 * nnz), (size_t)n,
                           rocprim::less<uint64_t>(), st));
    need = tmp_bytes;
    PRH_HIP(rocprim::reduce_by_key(tmp, need, key_m, val_m, (size_t)n_cat, key_s, cnt_s, n_unique,
                                   rocprim::plus<uint64_t>(), rocprim::equal_to<uint64_t>(), st));
    int64_t nnz_s = 0;
    PRH_HIP(hipMemcpyAsync(&nnz_s, n_unique, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));
    op->nnz_s = nnz_s;

    // 5. S: row pointers, columns, values, K
    uint32_t *k32 = nullptr;
    unsigned int *kmax = nullptr;
    if ((rc = S.alloc(&k32, nnz_s)) || (rc = S.alloc(&kmax, 1))) return rc;
    if ((rc = alloc_into(&op->rowptr, n + 1)) || (rc = alloc_into(&op->col, nnz_s)) ||
        (rc = alloc_into(&op->val, nnz_s)))
        return rc;
    PRH_HIP(hipMemsetAsync(kmax, 0, sizeof(unsigned int), st));
    k_row_starts<<<grid_of(nnz_s), AT, 0, st>>>(key_s, nnz_s, n, op->rowptr);
    k_s_values<<<grid_of(nnz_s), AT, 0, st>>>(key_s, cnt_s, nnz_s, op->area, op->rays, op->delta, op->col, op->val,
                                              k32, kmax);
    PRH_HIP(hipGetLastError());
    unsigned int hkmax = 0;
    PRH_HIP(hipMemcpyAsync(&hkmax, kmax, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));

    // compact forms of the batched engine (as prh_operator_create)
    bool uniform = true;
    for (int64_t i = 1; i < n && uniform; ++i)
        uniform = area_host[i] == area_host[0] && rays_host[i] == rays_host[0];
    op->uniform = uniform && hkmax <= 65535u;
    if (op->uniform) {
        op->inv_2R = 1.0 / (2.0 * (double)rays_host[0]);
        if ((rc = alloc_into(&op->kval, nnz_s))) return rc;
    } else {
        if ((rc = alloc_into(&op->val32, nnz_s))) return rc;
    }
    k_narrow<<<grid_of(nnz_s), AT, 0, st>>>(k32, op->val, nnz_s, op->uniform ? 1 : 0, op->kval, op->val32);
    if ((rc = alloc_into(&op->scale32, n))) return rc;
    k_fill_f32<<<grid_of(n), AT, 0, st>>>(op->scale32, n, 1.0f);
    PRH_HIP(hipGetLastError());
    if ((rc = operator_alloc_work(op))) return rc;
    PRH_HIP(hipStreamSynchronize(st));
    return PRH_OK;
}

// ------------------------------------------------------------------ pooling
// pool_counts (porerad/view_factors.py:290-300) of device edge tables: the
// (origin, receiver) keys of every table radix-sorted together and equal keys'
// counts summed (integer sums, so the order among equal keys is immaterial).
__global__ void k_pool_keys(const int32_t *origin, const int32_t *receiver, const int32_t *count, int64_t nnz,
                            uint64_t *key, int64_t *val)
{
    for (int64_t e = (int64_t)blockIdx.x * AT + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * AT) {
        key[e] = (uint64_t)(uint32_t)origin[e] << 32 | (uint32_t)receiver[e];
        val[e] = count[e];
    }
}

__global__ void k_pool_split(const uint64_t *key, const int64_t *val, const int64_t *n_unique, int32_t *origin,
                             int32_t *receiver, int32_t *count, int32_t *overflow)
{
    const int64_t n = *n_unique;
    for (int64_t e = (int64_t)blockIdx.x * AT + threadIdx.x; e < n; e += (int64_t)gridDim.x * AT) {
        origin[e] = (int32_t)(key[e] >> 32);
        receiver[e] = (int32_t)(uint32_t)key[e];
        if (val[e] > INT32_MAX) *overflow = 1;
        count[e] = (int32_t)val[e];
    }
}

}  // namespace

int pool_edge_tables(const prh_edges *const *parts, int n_parts, int64_t n_patches, hipStream_t st, prh_edges **out)
{
    *out = nullptr;
    int64_t total = 0;
    int dev = 0;
    PRH_HIP(hipGetDevice(&dev));
    for (int k = 0; k < n_parts; ++k) total += parts[k]->nnz;
    Scratch S;
    int rc;
    uint64_t *key = nullptr, *key_s = nullptr, *key_u = nullptr;
    int64_t *val = nullptr, *val_s = nullptr, *val_u = nullptr, *n_unique = nullptr;
    int32_t *overflow = nullptr;
    if ((rc = S.alloc(&key, total)) || (rc = S.alloc(&val, total)) || (rc = S.alloc(&key_s, total)) ||
        (rc = S.alloc(&val_s, total)) || (rc = S.alloc(&key_u, total)) || (rc = S.alloc(&val_u, total)) ||
        (rc = S.alloc(&n_unique, 1)) || (rc = S.alloc(&overflow, 1)))
        return rc;
    int64_t at = 0;
    for (int k = 0; k < n_parts; ++k) {
        const prh_edges *e = parts[k];
        if (e->nnz > 0)
            k_pool_keys<<<grid_of(e->nnz), AT, 0, st>>>(e->origin, e->receiver, e->count, e->nnz, key + at, val + at);
        at += e->nnz;
    }
    PRH_HIP(hipGetLastError());
    int bits = 1;
    while ((int64_t(1) << bits) < n_patches) ++bits;
    size_t need_sort = 0, need_reduce = 0;
    PRH_HIP(rocprim::radix_sort_pairs(nullptr, need_sort, key, key_s, val, val_s, (size_t)total, 0u,
                                      (unsigned)(32 + bits), st));
    PRH_HIP(rocprim::reduce_by_key(nullptr, need_reduce, key_s, val_s, (size_t)total, key_u, val_u, n_unique,
                                   rocprim::plus<int64_t>(), rocprim::equal_to<uint64_t>(), st));
    char *tmp = nullptr;
    const size_t tmp_bytes = std::max(need_sort, need_reduce);
    if ((rc = S.alloc(&tmp, (int64_t)tmp_bytes))) return rc;
    size_t need = tmp_bytes;
    PRH_HIP(rocprim::radix_sort_pairs((void *)tmp, need, key, key_s, val, val_s, (size_t)total, 0u,
                                      (unsigned)(32 + bits), st));
    need = tmp_bytes;
    PRH_HIP(rocprim::reduce_by_key((void *)tmp, need, key_s, val_s, (size_t)total, key_u, val_u, n_unique,
                                   rocprim::plus<int64_t>(), rocprim::equal_to<uint64_t>(), st));
    int64_t nnz = 0;
    PRH_HIP(hipMemcpyAsync(&nnz, n_unique, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    PRH_HIP(hipMemsetAsync(overflow, 0, sizeof(int32_t), st));
    PRH_HIP(hipStreamSynchronize(st));
    prh_edges *res = new prh_edges{dev, nnz, nullptr, nullptr, nullptr};
    const int64_t alloc = nnz > 0 ? nnz : 1;
    if (hipMalloc((void **)&res->origin, alloc * sizeof(int32_t)) != hipSuccess ||
        hipMalloc((void **)&res->receiver, alloc * sizeof(int32_t)) != hipSuccess ||
        hipMalloc((void **)&res->count, alloc * sizeof(int32_t)) != hipSuccess) {
        prh_edges_destroy(res);
        (void)hipGetLastError();
        set_error("pooling the seed blocks' edges: out of device memory (%lld edges)", (long long)nnz);
        return PRH_ERR_HIP;
    }
    int32_t oflow = 0;
    if (nnz > 0) k_pool_split<<<grid_of(nnz), AT, 0, st>>>(key_u, val_u, n_unique, res->origin, res->receiver,
                                                           res->count, overflow);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&oflow, overflow, sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess || oflow) {
        prh_edges_destroy(res);
        if (e != hipSuccess) {
            set_error("pooling the seed blocks' edges: %s", hipGetErrorString(e));
            return PRH_ERR_HIP;
        }
        set_error("pooled count above 2^31 - 1 for one edge");
        return PRH_ERR_CAPACITY;
    }
    *out = res;
    return PRH_OK;
}

}  // namespace prh

using namespace prh;

extern "C" {

int prh_operator_create_device(const prh_edges *edges, int64_t n, const int64_t *rays_per_origin, const double *area,
                               double support_diagonal_fraction, prh_operator **out)
{
    PRH_CHECK_ARG(edges && rays_per_origin && area && out, "prh_operator_create_device: null pointer");
    PRH_CHECK_ARG(n > 0 && n < INT32_MAX, "prh_operator_create_device: bad sizes");
    *out = nullptr;
    PRH_HIP(hipSetDevice(edges->device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    prh_operator *op = new prh_operator();
    op->device = edges->device;
    op->n = n;
    op->delta = support_diagonal_fraction;
    if ((rc = build(op, edges, rays_per_origin, area, st))) {
        (void)hipStreamSynchronize(st);
        operator_free(op);
        delete op;
        return rc;
    }
    *out = op;
    return PRH_OK;
}

int prh_operator_edges(const prh_operator *op, int32_t *origin, int32_t *receiver, int64_t *count)
{
    PRH_CHECK_ARG(op, "prh_operator_edges: null operator");
    const int64_t n = op->n, nnz = op->nnz_c;
    PRH_HIP(hipSetDevice(op->device));
    if (origin) {
        std::vector<int64_t> rowptr(n + 1);
        PRH_HIP(hipMemcpy(rowptr.data(), op->c_rowptr, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; ++i)
            for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) origin[e] = (int32_t)i;
    }
    if (receiver && nnz > 0)
        PRH_HIP(hipMemcpy(receiver, op->c_col, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost));
    if (count && nnz > 0) {
        std::vector<int32_t> c(nnz);
        PRH_HIP(hipMemcpy(c.data(), op->c_cnt, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost));
        for (int64_t k = 0; k < nnz; ++k) count[k] = c[k];
    }
    return PRH_OK;
}

}  // extern "C"
