// Digital microstructure generators on the device.
//
// Reference: porerad/geometry.py.  The shape parameters (random draws, the
// periodic neighbour graph of the foam, fibre frames, long-cylinder node
// lists and disk offsets, the layered field's NumPy values) are formed by
// the host with the reference's own NumPy operations; the library does the
// per-cell work on a cubic phase field:
//
//   prh_stamp_shapes    the N^3 membership tests of shape_mask's short shapes
//                       (ellipsoids and cylinders, geometry.py:106-116) for a
//                       whole list of shapes in one pass, OR-ed into the field
//   prh_stamp_long_cylinders
//                       every long cylinder of a render in one launch
//                       (_long_cylinder, geometry.py:63-86, with the whole-cell
//                       roll of integer centres, :100-105): the disk test of
//                       each candidate offset on the device, then the scatter
//   prh_stamp_offsets   the same scatter from host-formed disk offsets
//   prh_threshold_field threshold (geometry.py:235-238) of a field, optionally
//                       after lamellar_field's seven-point noise smoothing
//                       (:229-232): the two order statistics of np.quantile's
//                       linear rule by radix selection, then field >= q
//
// Arithmetic of the membership test, in the reference's order and rounding:
//   x = ((i - c) + N/2) mod N - N/2   NumPy's float remainder (fmod, plus N
//                                     when negative), per axis
//   l_r = fma(Rt[r][2], z, fma(Rt[r][1], y, Rt[r][0] * x))
//                                     R^T x as OpenBLAS's dgemm forms it: a
//                                     rounded product, then two FMAs
//   ellipsoid  ((l0/rx)^2 + (l1/ry)^2) + (l2/rz)^2 <= 1
//   cylinder   (l0/rx)^2 + (l1/ry)^2 <= 1  and  |l2| <= rz
// The disk test of a long cylinder forms l = Rq^T o with the same FMA chain
// (NumPy's small product rounds so as well).  Everything else is compiled
// without contraction.  A cell whose offsets lie beyond a shape's reach on
// some axis is not tested against it (the shape's points lie within its
// reach of the centre; rounding of the rotation is far below the one-cell
// margin of the reach; the host gives a matrix that is not a rotation a
// reach that covers every cell).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "prh_internal.h"

namespace prh {
namespace {

constexpr int GEN_THREADS = 256;
constexpr int SHAPE_DOUBLES = 16;  // cx cy cz rx ry rz Rt[9] reach

__device__ __forceinline__ double min_image(int i, double c, double Nf, double half)
{
    const double a = ((double)i - c) + half;
    double m;
    if (a >= 0.0 && a < Nf) {
        m = a;
    } else if (a >= Nf && a < 2.0 * Nf) {
        m = a - Nf;  // exact: N <= a < 2N
    } else if (a < 0.0 && a > -Nf) {
        m = a + Nf;
    } else {
        m = fmod(a, Nf);
        if (m < 0.0) m += Nf;
    }
    return m - half;
}

__global__ void __launch_bounds__(GEN_THREADS) k_stamp_shapes(uint8_t *__restrict__ phases, int N,
                                                              const double *__restrict__ shapes,
                                                              const int32_t *__restrict__ kinds, int n_shapes)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t cells = (int64_t)N * N * N;
    if (c >= cells || phases[c] == 1) return;
    const int k = (int)(c % N);
    const int j = (int)((c / N) % N);
    const int i = (int)(c / ((int64_t)N * N));
    const double Nf = (double)N, half = Nf / 2.0;
    for (int s = 0; s < n_shapes; ++s) {
        const double *S = shapes + (int64_t)SHAPE_DOUBLES * s;
        const double reach = S[15];
        const double x = min_image(i, S[0], Nf, half);
        if (fabs(x) > reach) continue;
        const double y = min_image(j, S[1], Nf, half);
        if (fabs(y) > reach) continue;
        const double z = min_image(k, S[2], Nf, half);
        if (fabs(z) > reach) continue;
        const double *Rt = S + 6;
        const double l0 = __builtin_fma(Rt[2], z, __builtin_fma(Rt[1], y, Rt[0] * x));
        const double l1 = __builtin_fma(Rt[5], z, __builtin_fma(Rt[4], y, Rt[3] * x));
        const double l2 = __builtin_fma(Rt[8], z, __builtin_fma(Rt[7], y, Rt[6] * x));
        const double a = l0 / S[3], b = l1 / S[4];
        const double ab = a * a + b * b;
        bool inside;
        if (kinds[s] == 0) {
            const double e = l2 / S[5];
            inside = ab + e * e <= 1.0;
        } else {
            inside = ab <= 1.0 && fabs(l2) <= S[5];
        }
        if (inside) {
            phases[c] = 1;
            return;
        }
    }
}

__global__ void __launch_bounds__(GEN_THREADS) k_stamp_offsets(uint8_t *__restrict__ phases, int N,
                                                               const int32_t *__restrict__ nodes, int n_nodes,
                                                               const int32_t *__restrict__ offsets, int n_offsets,
                                                               int sx, int sy, int sz)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n_nodes * n_offsets) return;
    const int a = (int)(t / n_offsets), b = (int)(t % n_offsets);
    int i = (nodes[3 * a] + offsets[3 * b] + sx) % N;
    int j = (nodes[3 * a + 1] + offsets[3 * b + 1] + sy) % N;
    int k = (nodes[3 * a + 2] + offsets[3 * b + 2] + sz) % N;
    i += i < 0 ? N : 0;
    j += j < 0 ? N : 0;
    k += k < 0 ? N : 0;
    phases[((int64_t)i * N + j) * N + k] = 1;
}

// Long cylinders of one render (_long_cylinder for each, with the roll of
// integer centres): every (cylinder, node, candidate offset) of the
// (2 pad + 1)^3 box is one work item; the offset belongs to the disk when
// (l0/r)^2 + (l1/r)^2 <= 1 + 1e-12 and |l2| <= 0.5 + 1e-12 with l = Rq^T off
// formed as the FMA chain NumPy's product rounds to.
struct LongCylinders {
    int n_cyl;
    const int64_t *work_start;  // n_cyl + 1: first work item of each cylinder
    const int32_t *node_start;  // n_cyl + 1: first node of each cylinder
    const int32_t *nodes;       // x3, in [0, N)
    const double *rq_t;         // x9: Rq^T row-major
    const double *radius;
    const int32_t *pad;
    const int32_t *shift;  // x3, in [0, N)
};

__global__ void __launch_bounds__(GEN_THREADS) k_stamp_long_cylinders(uint8_t *__restrict__ phases, int N,
                                                                      LongCylinders A)
{
    const int64_t total = A.work_start[A.n_cyl];
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        int c = 0;
        while (c + 1 < A.n_cyl && A.work_start[c + 1] <= t) ++c;
        const int pad = A.pad[c], S = 2 * pad + 1;
        const int64_t K = (int64_t)S * S * S;
        const int64_t rem = t - A.work_start[c];
        const int node = A.node_start[c] + (int)(rem / K);
        const int cand = (int)(rem % K);
        const int ox = cand / (S * S) - pad, oy = (cand / S) % S - pad, oz = cand % S - pad;
        const double x = (double)ox, y = (double)oy, z = (double)oz;
        const double *Rt = A.rq_t + 9 * c;
        const double l0 = __builtin_fma(Rt[2], z, __builtin_fma(Rt[1], y, Rt[0] * x));
        const double l1 = __builtin_fma(Rt[5], z, __builtin_fma(Rt[4], y, Rt[3] * x));
        const double l2 = __builtin_fma(Rt[8], z, __builtin_fma(Rt[7], y, Rt[6] * x));
        const double r = A.radius[c];
        const double a = l0 / r, b = l1 / r;
        if (!(a * a + b * b <= 1.0 + 1e-12 && fabs(l2) <= 0.5 + 1e-12)) continue;
        const int32_t *nd = A.nodes + 3 * node, *sh = A.shift + 3 * c;
        int i = (nd[0] + ox + sh[0]) % N, j = (nd[1] + oy + sh[1]) % N, k = (nd[2] + oz + sh[2]) % N;
        i += i < 0 ? N : 0;
        j += j < 0 ? N : 0;
        k += k < 0 ? N : 0;
        phases[((int64_t)i * N + j) * N + k] = 1;
    }
}

__global__ void k_fill(uint8_t *__restrict__ p, int64_t n, uint8_t v)
{
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x)
        p[c] = v;
}

__global__ void __launch_bounds__(GEN_THREADS) k_count_value(const uint8_t *__restrict__ p, int64_t n, uint8_t v,
                                                             unsigned long long *__restrict__ out)
{
    int64_t s = 0;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x)
        s += p[c] == v;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(out, (unsigned long long)s);
}

// value = field + (noise + n[i-1] + n[i+1] + n[j-1] + n[j+1] + n[k-1] + n[k+1]) / 7, the terms added left
// to right (np.roll(noise, 1, a) holds the preceding cell along a, periodic).
__global__ void __launch_bounds__(GEN_THREADS) k_smooth_add(const double *__restrict__ field,
                                                            const double *__restrict__ noise, int N,
                                                            double *__restrict__ out)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t cells = (int64_t)N * N * N;
    if (c >= cells) return;
    const int k = (int)(c % N);
    const int j = (int)((c / N) % N);
    const int i = (int)(c / ((int64_t)N * N));
    const int64_t sx = (int64_t)N * N, sy = N;
    double s = noise[c];
    s = s + noise[c + (i > 0 ? -sx : (int64_t)(N - 1) * sx)];
    s = s + noise[c + (i + 1 < N ? sx : -(int64_t)(N - 1) * sx)];
    s = s + noise[c + (j > 0 ? -sy : (int64_t)(N - 1) * sy)];
    s = s + noise[c + (j + 1 < N ? sy : -(int64_t)(N - 1) * sy)];
    s = s + noise[c + (k > 0 ? -1 : N - 1)];
    s = s + noise[c + (k + 1 < N ? 1 : -(N - 1))];
    out[c] = field[c] + s / 7.0;
}

// Order-preserving map of doubles to unsigned 64-bit keys.
__device__ __forceinline__ uint64_t order_key(double v)
{
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
}

double key_value(uint64_t key)
{
    const uint64_t b = (key >> 63) ? (key & 0x7FFFFFFFFFFFFFFFULL) : ~key;
    double v;
    std::memcpy(&v, &b, sizeof v);
    return v;
}

// Histogram of the 8-bit digit at `shift` of the keys whose bits above the
// digit equal `prefix` (under `mask`).
__global__ void __launch_bounds__(GEN_THREADS) k_digit_histogram(const double *__restrict__ v, int64_t n,
                                                                 uint64_t prefix, uint64_t mask, int shift,
                                                                 unsigned long long *__restrict__ hist)
{
    __shared__ unsigned lds[256];
    lds[threadIdx.x] = 0u;
    __syncthreads();
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = order_key(v[c]);
        if ((key & mask) == prefix) atomicAdd(&lds[(key >> shift) & 0xFFu], 1u);
    }
    __syncthreads();
    if (lds[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)lds[threadIdx.x]);
}

__global__ void __launch_bounds__(GEN_THREADS) k_threshold(const double *__restrict__ v, int64_t n, double q,
                                                           uint8_t *__restrict__ phases)
{
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x)
        phases[c] = v[c] >= q ? 1 : 0;
}

unsigned grid_of(int64_t n) { return (unsigned)((n + GEN_THREADS - 1) / GEN_THREADS); }
unsigned stride_grid(int64_t n) { return (unsigned)std::min<int64_t>(std::max<int64_t>(grid_of(n), 1), 4096); }

struct GenBuffers {
    std::vector<void *> ptrs;
    ~GenBuffers()
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

// The k-th smallest of v[0..n) (0-based) by eight passes of 8-bit radix selection.
int select_kth(const double *v, int64_t n, int64_t k, unsigned long long *hist, hipStream_t st, double *out)
{
    uint64_t prefix = 0, mask = 0;
    unsigned long long h[256];
    for (int shift = 56; shift >= 0; shift -= 8) {
        PRH_HIP(hipMemsetAsync(hist, 0, sizeof(h), st));
        k_digit_histogram<<<stride_grid(n), GEN_THREADS, 0, st>>>(v, n, prefix, mask, shift, hist);
        PRH_HIP(hipGetLastError());
        PRH_HIP(hipMemcpyAsync(h, hist, sizeof(h), hipMemcpyDeviceToHost, st));
        PRH_HIP(hipStreamSynchronize(st));
        int b = 0;
        for (; b < 256; ++b) {
            if ((unsigned long long)k < h[b]) break;
            k -= (int64_t)h[b];
        }
        if (b == 256) {
            set_error("radix selection lost its rank (non-finite values?)");
            return PRH_ERR_ARG;
        }
        prefix |= (uint64_t)b << shift;
        mask |= (uint64_t)0xFF << shift;
    }
    *out = key_value(prefix);
    return PRH_OK;
}

int cubic(const prh_phase_field *f, const char *what)
{
    if (f->nx != f->ny || f->nx != f->nz) {
        set_error("%s: the field must be cubic (N x N x N)", what);
        return PRH_ERR_ARG;
    }
    return PRH_OK;
}

}  // namespace
}  // namespace prh

using namespace prh;

extern "C" {

int prh_phase_field_fill(prh_phase_field *f, uint8_t value)
{
    PRH_CHECK_ARG(f, "prh_phase_field_fill: null pointer");
    PRH_HIP(hipSetDevice(f->device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    k_fill<<<stride_grid(f->cells), GEN_THREADS, 0, st>>>(f->phases, f->cells, value);
    PRH_HIP(hipGetLastError());
    PRH_HIP(hipStreamSynchronize(st));
    return PRH_OK;
}

int prh_phase_field_copy(prh_phase_field *dst, const prh_phase_field *src)
{
    PRH_CHECK_ARG(dst && src, "prh_phase_field_copy: null pointer");
    PRH_CHECK_ARG(dst->nx == src->nx && dst->ny == src->ny && dst->nz == src->nz && dst->device == src->device,
                  "prh_phase_field_copy: fields differ in shape or device");
    PRH_HIP(hipSetDevice(dst->device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    PRH_HIP(hipMemcpyAsync(dst->phases, src->phases, (size_t)src->cells, hipMemcpyDeviceToDevice, st));
    PRH_HIP(hipStreamSynchronize(st));
    return PRH_OK;
}

int prh_phase_field_count(const prh_phase_field *f, uint8_t value, int64_t *count)
{
    PRH_CHECK_ARG(f && count, "prh_phase_field_count: null pointer");
    PRH_HIP(hipSetDevice(f->device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    GenBuffers B;
    unsigned long long *acc = B.alloc<unsigned long long>(1);
    if (!acc) {
        set_error("prh_phase_field_count: out of device memory");
        return PRH_ERR_HIP;
    }
    PRH_HIP(hipMemsetAsync(acc, 0, sizeof(unsigned long long), st));
    k_count_value<<<std::min<unsigned>(stride_grid(f->cells), 1024), GEN_THREADS, 0, st>>>(f->phases, f->cells,
                                                                                           value, acc);
    PRH_HIP(hipGetLastError());
    unsigned long long h = 0;
    PRH_HIP(hipMemcpyAsync(&h, acc, sizeof h, hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));
    *count = (int64_t)h;
    return PRH_OK;
}

int prh_stamp_shapes(prh_phase_field *f, int32_t n_shapes, const double *shapes, const int32_t *kinds)
{
    PRH_CHECK_ARG(f && (n_shapes == 0 || (shapes && kinds)), "prh_stamp_shapes: null pointer");
    PRH_CHECK_ARG(n_shapes >= 0, "prh_stamp_shapes: negative shape count");
    int rc = cubic(f, "prh_stamp_shapes");
    if (rc) return rc;
    for (int32_t s = 0; s < n_shapes; ++s) {
        PRH_CHECK_ARG(kinds[s] == 0 || kinds[s] == 1, "prh_stamp_shapes: shape %d has kind %d", s, kinds[s]);
        const double *S = shapes + (int64_t)SHAPE_DOUBLES * s;
        PRH_CHECK_ARG(S[3] > 0 && S[4] > 0 && S[5] > 0 && S[15] >= 0,
                      "prh_stamp_shapes: shape %d needs positive radii and reach", s);
    }
    if (n_shapes == 0) return PRH_OK;
    PRH_HIP(hipSetDevice(f->device));
    hipStream_t st;
    if ((rc = current_stream(&st))) return rc;
    GenBuffers B;
    double *d_shapes = B.alloc<double>((int64_t)SHAPE_DOUBLES * n_shapes);
    int32_t *d_kinds = B.alloc<int32_t>(n_shapes);
    if (!d_shapes || !d_kinds) {
        set_error("prh_stamp_shapes: out of device memory");
        return PRH_ERR_HIP;
    }
    PRH_HIP(hipMemcpyAsync(d_shapes, shapes, sizeof(double) * SHAPE_DOUBLES * n_shapes, hipMemcpyHostToDevice, st));
    PRH_HIP(hipMemcpyAsync(d_kinds, kinds, sizeof(int32_t) * n_shapes, hipMemcpyHostToDevice, st));
    k_stamp_shapes<<<grid_of(f->cells), GEN_THREADS, 0, st>>>(f->phases, f->nx, d_shapes, d_kinds, n_shapes);
    PRH_HIP(hipGetLastError());
    PRH_HIP(hipStreamSynchronize(st));
    return PRH_OK;
}

int prh_stamp_offsets(prh_phase_field *f, int32_t n_nodes, const int32_t *nodes, int32_t n_offsets,
                      const int32_t *offsets, const int32_t shift[3])
{
    PRH_CHECK_ARG(f && shift && (n_nodes == 0 || nodes) && (n_offsets == 0 || offsets),
                  "prh_stamp_offsets: null pointer");
    PRH_CHECK_ARG(n_nodes >= 0 && n_offsets >= 0, "prh_stamp_offsets: negative count");
    int rc = cubic(f, "prh_stamp_offsets");
    if (rc) return rc;
    const int N = f->nx;
    // Node + offset + shift must stay within int32 and above -N (one wrap).
    for (int d = 0; d < 3; ++d)
        PRH_CHECK_ARG(shift[d] >= 0 && shift[d] < N, "prh_stamp_offsets: shift outside [0, N)");
    for (int32_t a = 0; a < 3 * n_nodes; ++a)
        PRH_CHECK_ARG(nodes[a] >= 0 && nodes[a] < N, "prh_stamp_offsets: node outside [0, N)");
    for (int32_t b = 0; b < 3 * n_offsets; ++b)
        PRH_CHECK_ARG(offsets[b] > -N && offsets[b] < N, "prh_stamp_offsets: offset outside (-N, N)");
    if (n_nodes == 0 || n_offsets == 0) return PRH_OK;
    PRH_HIP(hipSetDevice(f->device));
    hipStream_t st;
    if ((rc = current_stream(&st))) return rc;
    GenBuffers B;
    int32_t *d_nodes = B.alloc<int32_t>(3 * (int64_t)n_nodes), *d_off = B.alloc<int32_t>(3 * (int64_t)n_offsets);
    if (!d_nodes || !d_off) {
        set_error("prh_stamp_offsets: out of device memory");
        return PRH_ERR_HIP;
    }
    PRH_HIP(hipMemcpyAsync(d_nodes, nodes, sizeof(int32_t) * 3 * n_nodes, hipMemcpyHostToDevice, st));
    PRH_HIP(hipMemcpyAsync(d_off, offsets, sizeof(int32_t) * 3 * n_offsets, hipMemcpyHostToDevice, st));
    k_stamp_offsets<<<grid_of((int64_t)n_nodes * n_offsets), GEN_THREADS, 0, st>>>(
        f->phases, N, d_nodes, n_nodes, d_off, n_offsets, shift[0], shift[1], shift[2]);
    PRH_HIP(hipGetLastError());
    PRH_HIP(hipStreamSynchronize(st));
    return PRH_OK;
}

int prh_stamp_long_cylinders(prh_phase_field *f, int32_t n_cyl, const int32_t *node_start, const int32_t *nodes,
                             const double *rq_t, const double *radius, const int32_t *shift)
{
    PRH_CHECK_ARG(f && (n_cyl == 0 || (node_start && nodes && rq_t && radius && shift)),
                  "prh_stamp_long_cylinders: null pointer");
    PRH_CHECK_ARG(n_cyl >= 0, "prh_stamp_long_cylinders: negative count");
    int rc = cubic(f, "prh_stamp_long_cylinders");
    if (rc) return rc;
    if (n_cyl == 0) return PRH_OK;
    const int N = f->nx;
    PRH_CHECK_ARG(node_start[0] == 0, "prh_stamp_long_cylinders: node_start[0] must be 0");
    std::vector<int64_t> work(n_cyl + 1, 0);
    std::vector<int32_t> pad(n_cyl);
    for (int32_t c = 0; c < n_cyl; ++c) {
        PRH_CHECK_ARG(node_start[c + 1] >= node_start[c], "prh_stamp_long_cylinders: node ranges must ascend");
        PRH_CHECK_ARG(radius[c] > 0.0 && radius[c] <= 0.25 * N,
                      "prh_stamp_long_cylinders: radius of cylinder %d outside (0, N/4]", c);
        pad[c] = (int32_t)std::ceil(radius[c]) + 2;
        const int64_t S = 2 * (int64_t)pad[c] + 1;
        work[c + 1] = work[c] + (int64_t)(node_start[c + 1] - node_start[c]) * S * S * S;
        for (int d = 0; d < 3; ++d)
            PRH_CHECK_ARG(shift[3 * c + d] >= 0 && shift[3 * c + d] < N, "prh_stamp_long_cylinders: shift outside [0, N)");
    }
    const int32_t n_nodes = node_start[n_cyl];
    for (int32_t a = 0; a < 3 * n_nodes; ++a)
        PRH_CHECK_ARG(nodes[a] >= 0 && nodes[a] < N, "prh_stamp_long_cylinders: node outside [0, N)");
    if (work[n_cyl] == 0) return PRH_OK;
    PRH_HIP(hipSetDevice(f->device));
    hipStream_t st;
    if ((rc = current_stream(&st))) return rc;
    GenBuffers B;
    int64_t *d_work = B.alloc<int64_t>(n_cyl + 1);
    int32_t *d_start = B.alloc<int32_t>(n_cyl + 1), *d_nodes = B.alloc<int32_t>(3 * (int64_t)n_nodes);
    int32_t *d_pad = B.alloc<int32_t>(n_cyl), *d_shift = B.alloc<int32_t>(3 * (int64_t)n_cyl);
    double *d_rq = B.alloc<double>(9 * (int64_t)n_cyl), *d_r = B.alloc<double>(n_cyl);
    if (!d_work || !d_start || !d_nodes || !d_pad || !d_shift || !d_rq || !d_r) {
        set_error("prh_stamp_long_cylinders: out of device memory");
        return PRH_ERR_HIP;
    }
    PRH_HIP(hipMemcpyAsync(d_work, work.data(), sizeof(int64_t) * (n_cyl + 1), hipMemcpyHostToDevice, st));
    PRH_HIP(hipMemcpyAsync(d_start, node_start, sizeof(int32_t) * (n_cyl + 1), hipMemcpyHostToDevice, st));
    PRH_HIP(hipMemcpyAsync(d_nodes, nodes, sizeof(int32_t) * 3 * (size_t)n_nodes, hipMemcpyHostToDevice, st));
    PRH_HIP(hipMemcpyAsync(d_pad, pad.data(), sizeof(int32_t) * n_cyl, hipMemcpyHostToDevice, st));
    PRH_HIP(hipMemcpyAsync(d_shift, shift, sizeof(int32_t) * 3 * (size_t)n_cyl, hipMemcpyHostToDevice, st));
    PRH_HIP(hipMemcpyAsync(d_rq, rq_t, sizeof(double) * 9 * (size_t)n_cyl, hipMemcpyHostToDevice, st));
    PRH_HIP(hipMemcpyAsync(d_r, radius, sizeof(double) * n_cyl, hipMemcpyHostToDevice, st));
    const LongCylinders A{n_cyl, d_work, d_start, d_nodes, d_rq, d_r, d_pad, d_shift};
    k_stamp_long_cylinders<<<stride_grid(work[n_cyl]), GEN_THREADS, 0, st>>>(f->phases, N, A);
    PRH_HIP(hipGetLastError());
    PRH_HIP(hipStreamSynchronize(st));
    return PRH_OK;
}

int prh_threshold_field(prh_phase_field *f, const double *field, const double *noise, double q, double *quantile)
{
    PRH_CHECK_ARG(f && field && quantile, "prh_threshold_field: null pointer");
    PRH_CHECK_ARG(q >= 0.0 && q <= 1.0, "prh_threshold_field: quantile outside [0, 1]");
    int rc = noise ? cubic(f, "prh_threshold_field (with noise)") : PRH_OK;
    if (rc) return rc;
    const int64_t n = f->cells;
    PRH_HIP(hipSetDevice(f->device));
    hipStream_t st;
    if ((rc = current_stream(&st))) return rc;
    GenBuffers B;
    double *d_field = B.alloc<double>(n), *d_noise = noise ? B.alloc<double>(n) : nullptr;
    double *values = noise ? B.alloc<double>(n) : d_field;
    unsigned long long *hist = B.alloc<unsigned long long>(256);
    if (!d_field || (noise && (!d_noise || !values)) || !hist) {
        set_error("prh_threshold_field: out of device memory");
        return PRH_ERR_HIP;
    }
    PRH_HIP(hipMemcpyAsync(d_field, field, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, st));
    if (noise) {
        PRH_HIP(hipMemcpyAsync(d_noise, noise, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, st));
        k_smooth_add<<<grid_of(n), GEN_THREADS, 0, st>>>(d_field, d_noise, f->nx, values);
        PRH_HIP(hipGetLastError());
    }
    // np.quantile's linear rule: virtual index v = (n - 1) q, order statistics
    // floor(v) and floor(v) + 1 (both the last one when v >= n - 1), gamma =
    // v - floor(v), and NumPy's _lerp (computed from the upper value when
    // gamma >= 1/2).
    const double v = (double)(n - 1) * q;
    int64_t lo = (int64_t)std::floor(v), hi = lo + 1;
    if (v >= (double)(n - 1)) lo = hi = n - 1;
    const double gamma = v - (double)(v >= (double)(n - 1) ? -1 : lo);
    double a = 0.0, b = 0.0;
    if ((rc = select_kth(values, n, lo, hist, st, &a))) return rc;
    if (hi == lo) b = a;
    else if ((rc = select_kth(values, n, hi, hist, st, &b))) return rc;
    const double diff = b - a;
    double qv = a + diff * gamma;
    if (gamma >= 0.5) qv = b - diff * (1.0 - gamma);
    *quantile = qv;
    k_threshold<<<stride_grid(n), GEN_THREADS, 0, st>>>(values, n, qv, f->phases);
    PRH_HIP(hipGetLastError());
    PRH_HIP(hipStreamSynchronize(st));
    return PRH_OK;
}

}  // extern "C"
