// Device exclusive prefix sums (three passes: tile reduce, scan of tile sums
// in one workgroup, tile scan + carry).  Used for patch numbering and edge
// offsets.  Deterministic: integer sums.
#include "prh_internal.h"

namespace prh {
namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

template <typename T>
__device__ int64_t block_exclusive_scan(int64_t v, int64_t *lds, int64_t *total)
{
    // Hillis-Steele over 256 partial sums in LDS.
    const int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_THREADS; off <<= 1) {
        int64_t add = t >= off ? lds[t - off] : 0;
        __syncthreads();
        lds[t] += add;
        __syncthreads();
    }
    int64_t incl = lds[t];
    *total = lds[SCAN_THREADS - 1];
    __syncthreads();
    return incl - v;
}

template <typename T>
__global__ __launch_bounds__(SCAN_THREADS) void k_tile_sums(const T *in, int64_t n, int64_t *tile_sums)
{
    __shared__ int64_t lds[SCAN_THREADS];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    int64_t s = 0;
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        int64_t idx = base + (int64_t)it * SCAN_THREADS + threadIdx.x;
        if (idx < n) s += (int64_t)in[idx];
    }
    int64_t total;
    block_exclusive_scan<T>(s, lds, &total);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tile_sums(int64_t *tile_sums, int64_t n_tiles,
                                                                 int64_t *grand_total)
{
    __shared__ int64_t lds[SCAN_THREADS];
    int64_t carry = 0;
    for (int64_t base = 0; base < n_tiles; base += SCAN_THREADS) {
        int64_t idx = base + threadIdx.x;
        int64_t v = idx < n_tiles ? tile_sums[idx] : 0;
        int64_t total;
        int64_t ex = block_exclusive_scan<int64_t>(v, lds, &total);
        if (idx < n_tiles) tile_sums[idx] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) *grand_total = carry;
}

// k_scan_tile_sums continuing from a running total: tile offsets start at
// *carry, the previous total goes to *start and the new one to *carry.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tile_sums_carry(int64_t *tile_sums, int64_t n_tiles,
                                                                       int64_t *carry, int64_t *start)
{
    __shared__ int64_t lds[SCAN_THREADS];
    const int64_t first = *carry;
    int64_t c = first;
    for (int64_t base = 0; base < n_tiles; base += SCAN_THREADS) {
        int64_t idx = base + threadIdx.x;
        int64_t v = idx < n_tiles ? tile_sums[idx] : 0;
        int64_t total;
        int64_t ex = block_exclusive_scan<int64_t>(v, lds, &total);
        if (idx < n_tiles) tile_sums[idx] = c + ex;
        c += total;
    }
    __syncthreads();  // every thread has read *carry
    if (threadIdx.x == 0) {
        *start = first;
        *carry = c;
    }
}

template <typename T>
__global__ __launch_bounds__(SCAN_THREADS) void k_tile_scan(const T *in, int64_t *out, int64_t n,
                                                            const int64_t *tile_offsets)
{
    __shared__ int64_t lds[SCAN_THREADS];
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int64_t v[SCAN_ITEMS];
    int64_t s = 0;
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        int64_t idx = base + it;
        v[it] = idx < n ? (int64_t)in[idx] : 0;
        s += v[it];
    }
    int64_t total;
    int64_t ex = block_exclusive_scan<T>(s, lds, &total) + tile_offsets[blockIdx.x];
    for (int it = 0; it < SCAN_ITEMS; ++it) {
        int64_t idx = base + it;
        if (idx < n) out[idx] = ex;
        ex += v[it];
    }
}

// The tile sums pass reads items strided by SCAN_THREADS while the tile scan
// pass reads them contiguously per thread; both cover the same SCAN_TILE
// elements of one tile, so the per-tile totals agree.
template <typename T>
int scan_impl(const T *in, int64_t *out, int64_t n, int64_t *total, hipStream_t stream)
{
    if (n <= 0) {
        *total = 0;
        return PRH_OK;
    }
    const int64_t n_tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    int64_t *tiles = nullptr;
    PRH_HIP(hipMallocAsync((void **)&tiles, sizeof(int64_t) * (n_tiles + 1), stream));
    k_tile_sums<T><<<dim3((unsigned)n_tiles), dim3(SCAN_THREADS), 0, stream>>>(in, n, tiles);
    k_scan_tile_sums<<<1, SCAN_THREADS, 0, stream>>>(tiles, n_tiles, tiles + n_tiles);
    k_tile_scan<T><<<dim3((unsigned)n_tiles), dim3(SCAN_THREADS), 0, stream>>>(in, out, n, tiles);
    PRH_HIP(hipGetLastError());
    PRH_HIP(hipMemcpyAsync(total, tiles + n_tiles, sizeof(int64_t), hipMemcpyDeviceToHost, stream));
    PRH_HIP(hipFreeAsync(tiles, stream));
    PRH_HIP(hipStreamSynchronize(stream));
    return PRH_OK;
}

}  // namespace

int64_t scan_scratch_elems(int64_t n)
{
    return n > 0 ? (n + SCAN_TILE - 1) / SCAN_TILE : 1;
}

int exclusive_scan_i32_to_i64_carry_async(const int32_t *in, int64_t *out, int64_t n, int64_t *scratch,
                                          int64_t *carry, int64_t *start, hipStream_t stream)
{
    const int64_t n_tiles = n > 0 ? (n + SCAN_TILE - 1) / SCAN_TILE : 0;
    if (n_tiles > 0)
        k_tile_sums<int32_t><<<dim3((unsigned)n_tiles), dim3(SCAN_THREADS), 0, stream>>>(in, n, scratch);
    k_scan_tile_sums_carry<<<1, SCAN_THREADS, 0, stream>>>(scratch, n_tiles, carry, start);
    if (n_tiles > 0)
        k_tile_scan<int32_t><<<dim3((unsigned)n_tiles), dim3(SCAN_THREADS), 0, stream>>>(in, out, n, scratch);
    PRH_HIP(hipGetLastError());
    return PRH_OK;
}

int exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, int64_t *total, hipStream_t stream)
{
    return scan_impl<int64_t>(in, out, n, total, stream);
}

int exclusive_scan_i32_to_i64(const int32_t *in, int64_t *out, int64_t n, int64_t *total, hipStream_t stream)
{
    return scan_impl<int32_t>(in, out, n, total, stream);
}

}  // namespace prh
