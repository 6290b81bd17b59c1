// Morphology descriptors of a two-phase digital microstructure on the device.
//
// Reference: porerad/descriptors.py.  A phase field is a uint8 array
// (nx, ny, nz), x-major (cell (i, j, k) at (i*ny + j)*nz + k), with 1 = solid
// and 0 = pore.  The library computes the integer quantities the reference's
// descriptors are made of; the floating-point summaries (ratios, quantiles,
// means of the chord list, the lineal-path function) are formed from these
// integers by the host, as the reference forms them with NumPy.
//
//   scalar counts   (scalar_descriptors, descriptors.py:63-79)
//     cells, solid cells (phase 1), pore cells (phase 0);
//     box interface faces per axis: neighbouring cells inside the box whose
//       solid indicators differ (interface_area, :38-43);
//     pore runs per axis with the box faces ending the runs
//       (mean_pore_chords, :46-48: the mean run length is pore / runs);
//     periodic interface faces per axis: pore indicator differs from the
//       next cell with wrap-around (interface_face_counts, :120-123);
//     26-connected pore clusters (scipy.ndimage.label with a full 3x3x3
//       structure, :51-60): the largest cluster's size, the cluster count and
//       the axes along which one cluster touches both box faces.
//   periodic chords (periodic_chords, :82-109)
//     the pore runs of every line along one axis of the periodic cell, in
//     the reference's order: lines in row-major order of the other two axes,
//     runs in order along each line, except that on a line whose first and
//     last cells are pore (and that is not all pore) the first and last runs
//     are joined into one chord, and those joined chords follow all the
//     others, in line order.  Also the chord-count histogram truncated to
//     hist_len bins and the number of all-pore lines.
//
// Connected components use a lock-free union-find over cell indices: every
// label points at a smaller or equal index of the same component; unions
// lower a root and finds halve paths, both with atomic minima, so the labels
// only decrease and every union ends.  The
// component of a cell is identified by its root, the smallest index in the
// component.  Sizes, face contacts and counts are integers reduced with
// atomics, so the results do not depend on the order of the threads.
#include <algorithm>
#include <vector>

#include "prh_internal.h"

namespace prh {
namespace {

constexpr int DESC_THREADS = 256;
constexpr int LINE_CHUNK = 8;
constexpr int LDS_HIST_MAX = 4096;

__device__ __forceinline__ int64_t wave_sum(int64_t v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// ------------------------------------------------------------------ scalar counts

// counts (int64): 0 solid, 1 pore, 2..4 box faces x/y/z, 5..7 pore runs x/y/z,
// 8..10 periodic faces x/y/z.
constexpr int N_CELL_COUNTS = 11;

__global__ void __launch_bounds__(DESC_THREADS) k_cell_counts(const uint8_t *__restrict__ p, int nx, int ny, int nz,
                                                              int64_t cells, unsigned long long *__restrict__ out)
{
    int32_t acc[N_CELL_COUNTS];
#pragma unroll
    for (int q = 0; q < N_CELL_COUNTS; ++q) acc[q] = 0;
    const int64_t sx = (int64_t)ny * nz, sy = nz;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(c % nz);
        const int64_t ij = c / nz;
        const int j = (int)(ij % ny);
        const int i = (int)(ij / ny);
        const uint8_t v = p[c];
        const bool solid = v == 1, pore = v == 0;
        acc[0] += solid;
        acc[1] += pore;
        // next and previous cell along x, y, z (periodic)
        const uint8_t nxp = p[i + 1 < nx ? c + sx : c - (int64_t)(nx - 1) * sx];
        const uint8_t nyp = p[j + 1 < ny ? c + sy : c - (int64_t)(ny - 1) * sy];
        const uint8_t nzp = p[k + 1 < nz ? c + 1 : c - (nz - 1)];
        const uint8_t pxm = p[i > 0 ? c - sx : c + (int64_t)(nx - 1) * sx];
        const uint8_t pym = p[j > 0 ? c - sy : c + (int64_t)(ny - 1) * sy];
        const uint8_t pzm = p[k > 0 ? c - 1 : c + (nz - 1)];
        acc[2] += (i + 1 < nx) && (solid != (nxp == 1));
        acc[3] += (j + 1 < ny) && (solid != (nyp == 1));
        acc[4] += (k + 1 < nz) && (solid != (nzp == 1));
        acc[5] += pore && (i == 0 || pxm != 0);
        acc[6] += pore && (j == 0 || pym != 0);
        acc[7] += pore && (k == 0 || pzm != 0);
        acc[8] += pore != (nxp == 0);
        acc[9] += pore != (nyp == 0);
        acc[10] += pore != (nzp == 0);
    }
    __shared__ int64_t part[DESC_THREADS / 64][N_CELL_COUNTS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < N_CELL_COUNTS; ++q) {
        const int64_t s = wave_sum((int64_t)acc[q]);
        if (lane == 0) part[wave][q] = s;
    }
    __syncthreads();
    if (threadIdx.x < N_CELL_COUNTS) {
        int64_t s = 0;
        for (int w = 0; w < DESC_THREADS / 64; ++w) s += part[w][threadIdx.x];
        if (s) atomicAdd(&out[threadIdx.x], (unsigned long long)s);
    }
}

// ------------------------------------------------------------------ 26-connected pore clusters

__device__ __forceinline__ int32_t load_label(const int32_t *L, int32_t x)
{
    return __hip_atomic_load(&L[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving: each visited label is lowered to its
// grandparent by an atomic minimum (a smaller index in the same component),
// so paths shorten as they are walked and labels still only decrease.
__device__ __forceinline__ int32_t find_root(int32_t *L, int32_t x)
{
    while (true) {
        const int32_t p = load_label(L, x);
        if (p == x) return x;
        const int32_t g = load_label(L, p);
        if (g == p) return p;
        atomicMin(&L[x], g);
        x = g;
    }
}

// Join the components of a and b: the larger root is lowered to the smaller
// one with an atomic minimum; if another thread lowered it first, continue
// from the value it holds now.  Labels only decrease, so this ends.
__device__ void unite(int32_t *L, int32_t a, int32_t b)
{
    while (true) {
        a = find_root(L, a);
        b = find_root(L, b);
        if (a == b) return;
        if (a < b) {
            const int32_t old = atomicMin(&L[b], a);
            if (old == b) return;
            b = old;
        } else {
            const int32_t old = atomicMin(&L[a], b);
            if (old == a) return;
            a = old;
        }
    }
}

// The 13 neighbours of a cell that precede it in index order.
__device__ __constant__ int8_t PRE_OFFSETS[13][3] = {
    {-1, -1, -1}, {-1, -1, 0}, {-1, -1, 1}, {-1, 0, -1}, {-1, 0, 0}, {-1, 0, 1}, {-1, 1, -1},
    {-1, 1, 0},   {-1, 1, 1},  {0, -1, -1}, {0, -1, 0},  {0, -1, 1}, {0, 0, -1}};

struct Cell {
    int i, j, k;
};

__device__ __forceinline__ Cell cell_of(int32_t c, int ny, int nz)
{
    const int k = c % nz;
    const int ij = c / nz;
    return Cell{ij / ny, ij % ny, k};
}

// Initial labels: a pore cell points at its first preceding pore neighbour
// (or at itself), a solid cell holds -1.
__global__ void __launch_bounds__(DESC_THREADS) k_label_init(const uint8_t *__restrict__ p, int nx, int ny, int nz,
                                                             int32_t cells, int32_t *__restrict__ L)
{
    const int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= cells) return;
    if (p[c] != 0) {
        L[c] = -1;
        return;
    }
    const Cell x = cell_of(c, ny, nz);
    int32_t label = c;
    for (int q = 0; q < 13; ++q) {
        const int i = x.i + PRE_OFFSETS[q][0], j = x.j + PRE_OFFSETS[q][1], k = x.k + PRE_OFFSETS[q][2];
        if (i < 0 || j < 0 || k < 0 || j >= ny || k >= nz) continue;
        const int32_t d = (i * ny + j) * nz + k;
        if (p[d] == 0) {
            label = d;
            break;
        }
    }
    L[c] = label;
}

__global__ void __launch_bounds__(DESC_THREADS) k_label_merge(const uint8_t *__restrict__ p, int nx, int ny, int nz,
                                                              int32_t cells, int32_t *L)
{
    const int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= cells || p[c] != 0) return;
    const Cell x = cell_of(c, ny, nz);
    for (int q = 0; q < 13; ++q) {
        const int i = x.i + PRE_OFFSETS[q][0], j = x.j + PRE_OFFSETS[q][1], k = x.k + PRE_OFFSETS[q][2];
        if (i < 0 || j < 0 || k < 0 || j >= ny || k >= nz) continue;
        const int32_t d = (i * ny + j) * nz + k;
        if (p[d] == 0) unite(L, c, d);
    }
}

// After the merge: every pore cell's label becomes its root; add the cell to
// the root's size and record the box faces it touches (bit 2a: face 0 of
// axis a, bit 2a+1: the last face).  Each wave walks cells in wave-uniform
// rounds and keeps a private count and face mask for one root (that of its
// first pore cell, nearly always the spanning cluster), added with one
// atomic each at the end; cells of other roots add their own atomics.
__global__ void __launch_bounds__(DESC_THREADS) k_label_stats(const uint8_t *__restrict__ p, int nx, int ny, int nz,
                                                              int32_t cells, int32_t *__restrict__ L,
                                                              int32_t *__restrict__ size, uint32_t *__restrict__ faces)
{
    const int lane = threadIdx.x & 63;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int32_t wave_root = -1, mine = 0;
    uint32_t mine_faces = 0;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < cells; base += stride) {
        const int32_t c = (int32_t)(base + lane);
        const bool pore = c < cells && p[c] == 0;
        int32_t root = -1;
        if (pore) {
            root = find_root(L, c);
            L[c] = root;
        }
        if (wave_root < 0) {
            const uint64_t active = __ballot(pore);
            if (active) wave_root = __shfl(root, __ffsll((unsigned long long)active) - 1);
        }
        if (!pore) continue;
        const Cell x = cell_of(c, ny, nz);
        uint32_t bits = 0;
        if (x.i == 0) bits |= 1u;
        if (x.i == nx - 1) bits |= 2u;
        if (x.j == 0) bits |= 4u;
        if (x.j == ny - 1) bits |= 8u;
        if (x.k == 0) bits |= 16u;
        if (x.k == nz - 1) bits |= 32u;
        if (root == wave_root) {
            ++mine;
            mine_faces |= bits;
        } else {
            atomicAdd(&size[root], 1);
            if (bits) atomicOr(&faces[root], bits);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        mine += __shfl_xor(mine, off);
        mine_faces |= __shfl_xor(mine_faces, off);
    }
    if (lane == 0 && mine) atomicAdd(&size[wave_root], mine);
    if (lane == 0 && mine_faces) atomicOr(&faces[wave_root], mine_faces);
}

// out: 0 largest cluster size, 1 cluster count, 2 percolation bits (bit a: a
// cluster touches both faces of axis a).
__global__ void __launch_bounds__(DESC_THREADS) k_label_reduce(const int32_t *__restrict__ L,
                                                               const int32_t *__restrict__ size,
                                                               const uint32_t *__restrict__ faces, int32_t cells,
                                                               unsigned long long *__restrict__ out)
{
    const int32_t c = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    const bool root = c < cells && L[c] == c;
    int32_t s = 0;
    uint32_t perc = 0;
    if (root) {
        s = size[c];
        const uint32_t f = faces[c];
        for (int a = 0; a < 3; ++a)
            if (((f >> (2 * a)) & 3u) == 3u) perc |= 1u << a;
    }
    for (int off = 32; off > 0; off >>= 1) {
        s = max(s, __shfl_xor(s, off));
        perc |= __shfl_xor(perc, off);
    }
    const uint64_t roots = __ballot(root);
    if ((threadIdx.x & 63) == 0 && roots) {
        atomicMax(&out[0], (unsigned long long)s);
        atomicAdd(&out[1], (unsigned long long)__popcll(roots));
        if (perc) atomicOr(&out[2], (unsigned long long)perc);
    }
}

// ------------------------------------------------------------------ periodic chords

struct LineGeom {
    int64_t stride;  // cell step along the axis
    int length;      // cells per line
    int inner;       // extent of the faster of the two other axes
    int64_t outer_step, inner_step;
};

__device__ __forceinline__ int64_t line_base(const LineGeom &G, int64_t m)
{
    return (m / G.inner) * G.outer_step + (m % G.inner) * G.inner_step;
}

// Walk one line in chunks of LINE_CHUNK loads and call emit(start, length)
// for each maximal pore run, in order.
template <typename Emit>
__device__ __forceinline__ void walk_runs(const uint8_t *__restrict__ p, int64_t base, const LineGeom &G, Emit emit)
{
    int start = -1;
    for (int t0 = 0; t0 < G.length; t0 += LINE_CHUNK) {
        uint8_t v[LINE_CHUNK];
#pragma unroll
        for (int u = 0; u < LINE_CHUNK; ++u) v[u] = t0 + u < G.length ? p[base + (int64_t)(t0 + u) * G.stride] : 1;
#pragma unroll
        for (int u = 0; u < LINE_CHUNK; ++u) {
            const int t = t0 + u;
            if (t >= G.length) break;
            if (v[u] == 0) {
                if (start < 0) start = t;
            } else if (start >= 0) {
                emit(start, t - start);
                start = -1;
            }
        }
    }
    if (start >= 0) emit(start, G.length - start);
}

__device__ __forceinline__ void hist_add(unsigned *lds, unsigned long long *hist, int hist_len, bool use_lds, int len)
{
    if (len >= hist_len) return;
    if (use_lds) atomicAdd(&lds[len], 1u);
    else atomicAdd(&hist[len], 1ull);
}

// Pass 1: per line the number of chords kept in place, whether its end runs
// are joined, the histogram and the number of all-pore lines.
__global__ void __launch_bounds__(DESC_THREADS) k_chord_count(const uint8_t *__restrict__ p, LineGeom G, int64_t lines,
                                                              int32_t *__restrict__ kept, int32_t *__restrict__ joined,
                                                              unsigned long long *__restrict__ hist, int hist_len,
                                                              unsigned long long *__restrict__ all_pore)
{
    __shared__ unsigned lds[LDS_HIST_MAX];
    const bool use_lds = hist_len <= LDS_HIST_MAX;
    if (use_lds)
        for (int b = threadIdx.x; b < hist_len; b += blockDim.x) lds[b] = 0u;
    __syncthreads();
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool full = false;
    if (m < lines) {
        const int64_t base = line_base(G, m);
        int runs = 0, first_len = 0, last_len = 0, last_start = -1;
        walk_runs(p, base, G, [&](int start, int len) {
            if (runs == 0 && start == 0) {
                first_len = len;  // held: may be joined with the last run
            } else if (start + len == G.length) {
                last_len = len;  // held likewise
                last_start = start;
            } else {
                hist_add(lds, hist, hist_len, use_lds, len);
            }
            ++runs;
        });
        const bool first_pore = first_len > 0, last_pore = last_start >= 0;
        full = first_len == G.length;
        const bool wrap = first_pore && last_pore;  // distinct end runs: the line is not all pore
        if (wrap) {
            hist_add(lds, hist, hist_len, use_lds, first_len + last_len);
        } else {
            if (first_pore) hist_add(lds, hist, hist_len, use_lds, first_len);
            if (last_pore) hist_add(lds, hist, hist_len, use_lds, last_len);
        }
        kept[m] = runs - (wrap ? 2 : 0);
        joined[m] = wrap ? 1 : 0;
    }
    const uint64_t fulls = __ballot(full);
    if ((threadIdx.x & 63) == 0 && fulls) atomicAdd(all_pore, (unsigned long long)__popcll(fulls));
    __syncthreads();
    if (use_lds)
        for (int b = threadIdx.x; b < hist_len; b += blockDim.x)
            if (lds[b]) atomicAdd(&hist[b], (unsigned long long)lds[b]);
}

// Pass 2: write the chords of every line at its offsets.
__global__ void __launch_bounds__(DESC_THREADS) k_chord_write(const uint8_t *__restrict__ p, LineGeom G, int64_t lines,
                                                              const int64_t *__restrict__ kept_off,
                                                              const int32_t *__restrict__ joined,
                                                              const int64_t *__restrict__ joined_off, int64_t n_kept,
                                                              int32_t *__restrict__ chords)
{
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= lines) return;
    const int64_t base = line_base(G, m);
    const bool wrap = joined[m] != 0;
    int64_t pos = kept_off[m];
    int first_len = 0;
    walk_runs(p, base, G, [&](int start, int len) {
        if (wrap && start == 0) {
            first_len = len;
        } else if (wrap && start + len == G.length) {
            chords[n_kept + joined_off[m]] = first_len + len;
        } else {
            chords[pos++] = len;
        }
    });
}

struct DescBuffers {
    std::vector<void *> ptrs;
    ~DescBuffers()
    {
        for (void *q : ptrs) (void)hipFree(q);
    }
    template <typename T>
    T *alloc(int64_t count)
    {
        void *q = nullptr;
        if (hipMalloc(&q, sizeof(T) * (size_t)std::max<int64_t>(count, 1)) != hipSuccess) return nullptr;
        ptrs.push_back(q);
        return (T *)q;
    }
};

unsigned grid_for(int64_t n) { return (unsigned)((n + DESC_THREADS - 1) / DESC_THREADS); }

LineGeom line_geom(const prh_phase_field *f, int axis)
{
    const int64_t sx = (int64_t)f->ny * f->nz, sy = f->nz;
    LineGeom G{};
    if (axis == 0) {  // lines (j, k): m = j*nz + k
        G = LineGeom{sx, f->nx, f->nz, sy, 1};
    } else if (axis == 1) {  // lines (i, k): m = i*nz + k
        G = LineGeom{sy, f->ny, f->nz, sx, 1};
    } else {  // lines (i, j): m = i*ny + j
        G = LineGeom{1, f->nz, f->ny, sx, sy};
    }
    return G;
}

}  // namespace
}  // namespace prh

using namespace prh;

extern "C" {

int prh_phase_field_create(const uint8_t *phases, int32_t nx, int32_t ny, int32_t nz, prh_phase_field **out)
{
    PRH_CHECK_ARG(out, "prh_phase_field_create: null pointer");
    PRH_CHECK_ARG(nx > 0 && ny > 0 && nz > 0, "prh_phase_field_create: extents must be positive");
    const int64_t cells = (int64_t)nx * ny * nz;
    PRH_CHECK_ARG(cells < (int64_t(1) << 31), "prh_phase_field_create: at most 2^31 - 1 cells");
    int device = 0;
    PRH_HIP(hipGetDevice(&device));
    prh_phase_field *f = new prh_phase_field{device, nx, ny, nz, cells, nullptr};
    if (hipMalloc((void **)&f->phases, (size_t)cells) != hipSuccess) {
        delete f;
        set_error("prh_phase_field_create: out of device memory");
        return PRH_ERR_HIP;
    }
    const hipError_t e = phases ? hipMemcpy(f->phases, phases, (size_t)cells, hipMemcpyHostToDevice)
                                : hipMemset(f->phases, 0, (size_t)cells);
    if (e != hipSuccess) {
        (void)hipFree(f->phases);
        delete f;
        set_error("prh_phase_field_create: upload failed");
        return PRH_ERR_HIP;
    }
    *out = f;
    return PRH_OK;
}

int prh_phase_field_read(const prh_phase_field *f, uint8_t *phases)
{
    PRH_CHECK_ARG(f && phases, "prh_phase_field_read: null pointer");
    PRH_HIP(hipSetDevice(f->device));
    PRH_HIP(hipMemcpy(phases, f->phases, (size_t)f->cells, hipMemcpyDeviceToHost));
    return PRH_OK;
}

void prh_phase_field_destroy(prh_phase_field *f)
{
    if (!f) return;
    (void)hipSetDevice(f->device);
    (void)hipFree(f->phases);
    delete f;
}

int prh_scalar_counts(const prh_phase_field *f, int32_t connectivity, int64_t counts[PRH_DESC_COUNT])
{
    PRH_CHECK_ARG(f && counts, "prh_scalar_counts: null pointer");
    PRH_HIP(hipSetDevice(f->device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    const int32_t cells = (int32_t)f->cells;
    DescBuffers B;
    unsigned long long *acc = B.alloc<unsigned long long>(N_CELL_COUNTS + 3);
    int32_t *L = nullptr, *size = nullptr;
    uint32_t *faces = nullptr;
    if (connectivity) {
        L = B.alloc<int32_t>(cells);
        size = B.alloc<int32_t>(cells);
        faces = B.alloc<uint32_t>(cells);
    }
    if (!acc || (connectivity && (!L || !size || !faces))) {
        set_error("prh_scalar_counts: out of device memory");
        return PRH_ERR_HIP;
    }
    PRH_HIP(hipMemsetAsync(acc, 0, sizeof(unsigned long long) * (N_CELL_COUNTS + 3), st));
    const unsigned cgrid = (unsigned)std::min<int64_t>(grid_for(cells), 2048);
    k_cell_counts<<<cgrid, DESC_THREADS, 0, st>>>(f->phases, f->nx, f->ny, f->nz, cells, acc);
    if (connectivity) {
        PRH_HIP(hipMemsetAsync(size, 0, sizeof(int32_t) * (size_t)cells, st));
        PRH_HIP(hipMemsetAsync(faces, 0, sizeof(uint32_t) * (size_t)cells, st));
        const unsigned g = grid_for(cells);
        k_label_init<<<g, DESC_THREADS, 0, st>>>(f->phases, f->nx, f->ny, f->nz, cells, L);
        k_label_merge<<<g, DESC_THREADS, 0, st>>>(f->phases, f->nx, f->ny, f->nz, cells, L);
        k_label_stats<<<std::min<unsigned>(g, 1024), DESC_THREADS, 0, st>>>(f->phases, f->nx, f->ny, f->nz, cells, L,
                                                                          size, faces);
        k_label_reduce<<<g, DESC_THREADS, 0, st>>>(L, size, faces, cells, acc + N_CELL_COUNTS);
    }
    PRH_HIP(hipGetLastError());
    unsigned long long h[N_CELL_COUNTS + 3];
    PRH_HIP(hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));
    counts[PRH_DESC_CELLS] = f->cells;
    counts[PRH_DESC_SOLID] = (int64_t)h[0];
    counts[PRH_DESC_PORE] = (int64_t)h[1];
    for (int a = 0; a < 3; ++a) {
        counts[PRH_DESC_BOX_FACES_X + a] = (int64_t)h[2 + a];
        counts[PRH_DESC_PORE_RUNS_X + a] = (int64_t)h[5 + a];
        counts[PRH_DESC_PERIODIC_FACES_X + a] = (int64_t)h[8 + a];
    }
    counts[PRH_DESC_LARGEST_CLUSTER] = connectivity ? (int64_t)h[N_CELL_COUNTS] : -1;
    counts[PRH_DESC_CLUSTERS] = connectivity ? (int64_t)h[N_CELL_COUNTS + 1] : -1;
    counts[PRH_DESC_PERCOLATES] = connectivity ? (int64_t)h[N_CELL_COUNTS + 2] : -1;
    return PRH_OK;
}

int prh_periodic_chords(const prh_phase_field *f, int32_t axis, int32_t *chords, int64_t capacity, int64_t *n_chords,
                        int64_t *n_all_pore, int64_t *histogram, int32_t hist_len)
{
    PRH_CHECK_ARG(f && n_chords && n_all_pore && histogram, "prh_periodic_chords: null pointer");
    PRH_CHECK_ARG(axis >= 0 && axis < 3, "prh_periodic_chords: axis must be 0, 1 or 2");
    PRH_CHECK_ARG(hist_len > 0 && capacity >= 0 && (chords || capacity == 0), "prh_periodic_chords: bad sizes");
    PRH_HIP(hipSetDevice(f->device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    const LineGeom G = line_geom(f, axis);
    const int64_t lines = f->cells / G.length;
    DescBuffers B;
    int32_t *kept = B.alloc<int32_t>(lines), *joined = B.alloc<int32_t>(lines);
    int64_t *kept_off = B.alloc<int64_t>(lines), *joined_off = B.alloc<int64_t>(lines);
    unsigned long long *hist = B.alloc<unsigned long long>(hist_len + 1);
    if (!kept || !joined || !kept_off || !joined_off || !hist) {
        set_error("prh_periodic_chords: out of device memory");
        return PRH_ERR_HIP;
    }
    PRH_HIP(hipMemsetAsync(hist, 0, sizeof(unsigned long long) * ((size_t)hist_len + 1), st));
    k_chord_count<<<grid_for(lines), DESC_THREADS, 0, st>>>(f->phases, G, lines, kept, joined, hist, hist_len,
                                                            hist + hist_len);
    PRH_HIP(hipGetLastError());
    int64_t n_kept = 0, n_joined = 0;
    if ((rc = exclusive_scan_i32_to_i64(kept, kept_off, lines, &n_kept, st))) return rc;
    if ((rc = exclusive_scan_i32_to_i64(joined, joined_off, lines, &n_joined, st))) return rc;
    std::vector<unsigned long long> h((size_t)hist_len + 1);
    PRH_HIP(hipMemcpyAsync(h.data(), hist, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));
    for (int b = 0; b < hist_len; ++b) histogram[b] = (int64_t)h[b];
    *n_all_pore = (int64_t)h[hist_len];
    *n_chords = n_kept + n_joined;
    if (!chords) return PRH_OK;
    if (n_kept + n_joined > capacity) {
        set_error("prh_periodic_chords: %lld chords exceed the capacity %lld", (long long)(n_kept + n_joined),
                  (long long)capacity);
        return PRH_ERR_CAPACITY;
    }
    int32_t *d_chords = B.alloc<int32_t>(n_kept + n_joined);
    if (!d_chords) {
        set_error("prh_periodic_chords: out of device memory");
        return PRH_ERR_HIP;
    }
    k_chord_write<<<grid_for(lines), DESC_THREADS, 0, st>>>(f->phases, G, lines, kept_off, joined, joined_off, n_kept,
                                                            d_chords);
    PRH_HIP(hipGetLastError());
    PRH_HIP(hipMemcpyAsync(chords, d_chords, sizeof(int32_t) * (size_t)(n_kept + n_joined), hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));
    return PRH_OK;
}

}  // extern "C"
