// Device patch table of an x-oriented microstructure.
//
// Reference: build_surface_patches, porerad/view_factors.py:70-132.
//   * wall patches come first, numbered in (j, k, hot/cold) order, and exist
//     only where the boundary cell is pore (view_factors.py:85-92);
//   * internal patches are the exposed solid faces in (i, j, k, face) order,
//     face 0..5 = +x, -x, +y, -y, +z, -z, periodic in y and z, with no solid
//     faces on the x boundaries (view_factors.py:94-103);
//   * origins sit 1e-9 off the face and are wrapped with float modulo in y, z
//     (view_factors.py:88-90, 105-109).
//
// Layout in HBM (see prh_internal.h): a flat occupancy bitset for the
// traversal, one 32-bit code per cell (first patch id << 6 | face mask) to
// turn a hit face into a patch id, and per-patch origin/face/class arrays.
#include <algorithm>

#include "prh_internal.h"
#include "prh_device.h"

namespace prh {
namespace {

constexpr int GEO_THREADS = 256;

__global__ __launch_bounds__(GEO_THREADS) void k_cell_masks(const uint8_t *solid, int nx, int ny, int nz,
                                                             uint8_t *mask, int32_t *face_count)
{
    const int64_t n = (int64_t)nx * ny * nz;
    for (int64_t c = (int64_t)blockIdx.x * GEO_THREADS + threadIdx.x; c < n; c += (int64_t)gridDim.x * GEO_THREADS) {
        const int k = (int)(c % nz);
        const int64_t ij = c / nz;
        const int j = (int)(ij % ny);
        const int i = (int)(ij / ny);
        uint32_t m = 0;
        if (solid[c]) {
            const int64_t row = (int64_t)i * ny;
            const int jp = j + 1 == ny ? 0 : j + 1, jm = j == 0 ? ny - 1 : j - 1;
            const int kp = k + 1 == nz ? 0 : k + 1, km = k == 0 ? nz - 1 : k - 1;
            if (i < nx - 1 && !solid[((row + ny) + j) * nz + k]) m |= 1u;
            if (i > 0 && !solid[((row - ny) + j) * nz + k]) m |= 2u;
            if (!solid[(row + jp) * nz + k]) m |= 4u;
            if (!solid[(row + jm) * nz + k]) m |= 8u;
            if (!solid[(row + j) * nz + kp]) m |= 16u;
            if (!solid[(row + j) * nz + km]) m |= 32u;
        }
        mask[c] = (uint8_t)m;
        face_count[c] = __popc(m);
    }
}

// Flat bitset: bit c of the occupancy is cell c = (i*ny + j)*nz + k.
__global__ __launch_bounds__(GEO_THREADS) void k_pack_occupancy(const uint8_t *solid, int64_t n_cells, uint32_t *occ)
{
    const int64_t n_words = (n_cells + 31) / 32;
    for (int64_t w = (int64_t)blockIdx.x * GEO_THREADS + threadIdx.x; w < n_words;
         w += (int64_t)gridDim.x * GEO_THREADS) {
        const int64_t c0 = w * 32;
        uint32_t bits = 0;
        for (int b = 0; b < 32 && c0 + b < n_cells; ++b)
            if (solid[c0 + b]) bits |= 1u << b;
        occ[w] = bits;
    }
}

// Brick bitset: bit b is set when brick b (2^shift cells per side, clipped
// at the grid end) holds a solid cell.
__global__ __launch_bounds__(GEO_THREADS) void k_pack_coarse(const uint8_t *solid, int nx, int ny, int nz, int shift,
                                                              int cnx, int cny, int cnz, uint32_t *coarse)
{
    const int64_t n_bricks = (int64_t)cnx * cny * cnz;
    const int64_t n_words = (n_bricks + 31) / 32;
    const int B = 1 << shift;
    for (int64_t w = (int64_t)blockIdx.x * GEO_THREADS + threadIdx.x; w < n_words;
         w += (int64_t)gridDim.x * GEO_THREADS) {
        uint32_t bits = 0;
        for (int t = 0; t < 32; ++t) {
            const int64_t b = w * 32 + t;
            if (b >= n_bricks) break;
            const int bk = (int)(b % cnz);
            const int bj = (int)((b / cnz) % cny);
            const int bi = (int)(b / ((int64_t)cnz * cny));
            bool any = false;
            for (int i = bi * B; i < (bi + 1) * B && i < nx && !any; ++i)
                for (int j = bj * B; j < (bj + 1) * B && j < ny && !any; ++j)
                    for (int k = bk * B; k < (bk + 1) * B && k < nz; ++k)
                        if (solid[((int64_t)i * ny + j) * nz + k]) {
                            any = true;
                            break;
                        }
            if (any) bits |= 1u << t;
        }
        coarse[w] = bits;
    }
}

// Walled bitset (Geometry::wocc): word w holds the 4x4x4 block of cells whose
// index P has P >> 6 == w; bit 63 - (P & 63) is set for a solid cell and for
// every cell of the planes i = -1 and i = nx.  Indices beyond ny, nz (the padding
// up to a power of two) stay clear.
__global__ __launch_bounds__(GEO_THREADS) void k_pack_walled(const uint8_t *solid, int nx, int ny, int nz, int lx,
                                                              int ly, int lz, int64_t n_words, uint64_t *wocc)
{
    const LeanLayout L = lean_layout(lx, ly, lz);
    for (int64_t w = (int64_t)blockIdx.x * GEO_THREADS + threadIdx.x; w < n_words;
         w += (int64_t)gridDim.x * GEO_THREADS) {
        uint64_t bits = 0;
        for (uint32_t t = 0; t < 64; ++t) {
            const uint32_t P = (uint32_t)w << (6 + LEAN_GAP) | t;
            const uint32_t ip = lean_compact(L, 0, P & L.mask[0]);
            const uint32_t j = lean_compact(L, 1, P & L.mask[1]);
            const uint32_t k = lean_compact(L, 2, P & L.mask[2]);
            bool set = false;
            if (j < (uint32_t)ny && k < (uint32_t)nz) {
                if (ip == 0u || ip == (uint32_t)nx + 1u) set = true;
                else if (ip <= (uint32_t)nx) set = solid[(((int64_t)ip - 1) * ny + j) * nz + k] != 0;
            }
            if (set) bits |= 1ull << (63u - t);  // mirrored: a left shift by t brings the cell to bit 63
        }
        wocc[w] = bits;
    }
}

// Geometry::wcode: entry c is the cell_code of the cell whose walled-bitset
// index is P = (c >> 6) << 9 | (c & 63), or 0 for the wall planes and the
// padding beyond nx, ny, nz.
__global__ __launch_bounds__(GEO_THREADS) void k_pack_wcode(const uint32_t *cell_code, int nx, int ny, int nz, int lx,
                                                             int ly, int lz, int64_t n, uint32_t *wcode)
{
    const LeanLayout L = lean_layout(lx, ly, lz);
    for (int64_t c = (int64_t)blockIdx.x * GEO_THREADS + threadIdx.x; c < n; c += (int64_t)gridDim.x * GEO_THREADS) {
        const uint32_t P = (uint32_t)(c >> 6) << (6 + LEAN_GAP) | (uint32_t)(c & 63);
        const uint32_t ip = lean_compact(L, 0, P & L.mask[0]);
        const uint32_t j = lean_compact(L, 1, P & L.mask[1]);
        const uint32_t k = lean_compact(L, 2, P & L.mask[2]);
        uint32_t code = 0u;
        if (ip >= 1u && ip <= (uint32_t)nx && j < (uint32_t)ny && k < (uint32_t)nz)
            code = cell_code[(((int64_t)ip - 1) * ny + j) * nz + k];
        wcode[c] = code;
    }
}

int ceil_log2(int n)
{
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

__global__ __launch_bounds__(GEO_THREADS) void k_wall_counts(const uint8_t *solid, int nx, int ny, int nz,
                                                              int32_t *wall_count)
{
    const int64_t n = (int64_t)ny * nz;
    for (int64_t c = (int64_t)blockIdx.x * GEO_THREADS + threadIdx.x; c < n; c += (int64_t)gridDim.x * GEO_THREADS) {
        const int64_t last = (int64_t)(nx - 1) * ny * nz;
        wall_count[c] = (solid[c] ? 0 : 1) + (solid[last + c] ? 0 : 1);
    }
}

__global__ __launch_bounds__(GEO_THREADS) void k_fill_walls(const uint8_t *solid, int nx, int ny, int nz,
                                                             const int64_t *wall_base, double *origin,
                                                             int8_t *face, int32_t *cell, int8_t *cls,
                                                             int32_t *hot_map, int32_t *cold_map)
{
    const int64_t n = (int64_t)ny * nz;
    for (int64_t c = (int64_t)blockIdx.x * GEO_THREADS + threadIdx.x; c < n; c += (int64_t)gridDim.x * GEO_THREADS) {
        const int j = (int)(c / nz), k = (int)(c % nz);
        const int64_t last = (int64_t)(nx - 1) * ny * nz;
        int64_t p = wall_base[c];
        hot_map[c] = -1;
        cold_map[c] = -1;
        for (int w = 0; w < 2; ++w) {
            if (solid[(w == 0 ? 0 : last) + c]) continue;
            origin[3 * p + 0] = w == 0 ? 1.0e-9 : (double)nx - 1.0e-9;
            origin[3 * p + 1] = (double)j + 0.5;
            origin[3 * p + 2] = (double)k + 0.5;
            face[p] = -1;
            cell[p] = -1;
            cls[p] = (int8_t)w;
            if (w == 0) hot_map[c] = (int32_t)p;
            else cold_map[c] = (int32_t)p;
            ++p;
        }
    }
}

__global__ __launch_bounds__(GEO_THREADS) void k_fill_internal(const uint8_t *mask, const int64_t *face_base,
                                                                int nx, int ny, int nz, int64_t n_wall,
                                                                double *origin, int8_t *face, int32_t *cell,
                                                                int8_t *cls, uint32_t *cell_code)
{
    const int64_t n = (int64_t)nx * ny * nz;
    const double offset = 0.5 + 1.0e-9;
    const double fny = (double)ny, fnz = (double)nz;
    for (int64_t c = (int64_t)blockIdx.x * GEO_THREADS + threadIdx.x; c < n; c += (int64_t)gridDim.x * GEO_THREADS) {
        const uint32_t m = mask[c];
        const int64_t first = n_wall + face_base[c];
        cell_code[c] = m ? (uint32_t)(first << 6) | m : 0u;
        if (!m) continue;
        const int k = (int)(c % nz);
        const int64_t ij = c / nz;
        const int j = (int)(ij % ny);
        const int i = (int)(ij / ny);
        int64_t p = first;
        for (int f = 0; f < 6; ++f) {
            if (!((m >> f) & 1u)) continue;
            double o0 = (double)i + 0.5, o1 = (double)j + 0.5, o2 = (double)k + 0.5;
            const double delta = (f & 1) ? -offset : offset;
            if (f < 2) o0 = o0 + delta;
            else if (f < 4) o1 = o1 + delta;
            else o2 = o2 + delta;
            origin[3 * p + 0] = o0;
            origin[3 * p + 1] = py_mod(o1, fny);
            origin[3 * p + 2] = py_mod(o2, fnz);
            face[p] = (int8_t)f;
            cell[p] = (int32_t)c;
            cls[p] = 2;
            ++p;
        }
    }
}

unsigned grid_for(int64_t n)
{
    int64_t b = (n + GEO_THREADS - 1) / GEO_THREADS;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return (unsigned)b;
}

void free_geometry(Geometry &g)
{
    void *ptrs[] = {g.occ, g.coarse, g.cell_code, g.patch_origin, g.patch_face, g.patch_cell, g.patch_class, g.hot_map,
                    g.cold_map, g.wocc, g.wcode};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    g.wocc = nullptr;
    g.wcode = nullptr;
    g.occ = nullptr;
    g.coarse = nullptr;
    g.cell_code = nullptr;
    g.patch_origin = nullptr;
    g.patch_face = nullptr;
    g.patch_cell = nullptr;
    g.patch_class = nullptr;
    g.hot_map = g.cold_map = nullptr;
}

int build(const uint8_t *solid_host, int nx, int ny, int nz, double H_m, Geometry &g)
{
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    PRH_HIP(hipGetDevice(&g.device));
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.occ_words = ((int64_t)nx * ny * nz + 31) / 32;
    // Smallest brick whose bitset stays within 72 KB, so that two 1024-lane
    // workgroups of the traversal kernel each hold a copy in one CU's LDS.
    for (g.coarse_shift = 1;; ++g.coarse_shift) {
        const int B = 1 << g.coarse_shift;
        g.cnx = (nx + B - 1) / B;
        g.cny = (ny + B - 1) / B;
        g.cnz = (nz + B - 1) / B;
        g.coarse_words = ((int64_t)g.cnx * g.cny * g.cnz + 31) / 32;
        if (g.coarse_words * 4 <= 72 * 1024) break;
    }
    g.H_m = H_m;
    const int64_t n_cells = (int64_t)nx * ny * nz;
    const int64_t n_lat = (int64_t)ny * nz;
    // The walled bitset exists when its index (with the LEAN_GAP zero bits)
    // fits 31 bits and both lateral sizes are at least 3 (each axis has two
    // interleaved bits, and a lateral step always changes j or k).
    g.wlx = std::max(2, ceil_log2(nx + 2));
    g.wly = std::max(2, ceil_log2(ny));
    g.wlz = std::max(2, ceil_log2(nz));
    const int walled_bits = g.wlx + g.wly + g.wlz;
    g.wocc_words =
        (ny >= 3 && nz >= 3 && walled_bits + (int)LEAN_GAP <= 31) ? (int64_t(1) << (walled_bits - 6)) : 0;

    uint8_t *solid = nullptr, *mask = nullptr;
    int32_t *face_count = nullptr, *wall_count = nullptr;
    int64_t *face_base = nullptr, *wall_base = nullptr;
    auto cleanup = [&]() {
        if (solid) (void)hipFree(solid);
        if (mask) (void)hipFree(mask);
        if (face_count) (void)hipFree(face_count);
        if (wall_count) (void)hipFree(wall_count);
        if (face_base) (void)hipFree(face_base);
        if (wall_base) (void)hipFree(wall_base);
    };
    struct Guard {
        decltype(cleanup) &f;
        ~Guard() { f(); }
    } guard{cleanup};

    PRH_HIP(hipMalloc((void **)&solid, n_cells));
    PRH_HIP(hipMalloc((void **)&mask, n_cells));
    PRH_HIP(hipMalloc((void **)&face_count, n_cells * sizeof(int32_t)));
    PRH_HIP(hipMalloc((void **)&face_base, n_cells * sizeof(int64_t)));
    PRH_HIP(hipMalloc((void **)&wall_count, n_lat * sizeof(int32_t)));
    PRH_HIP(hipMalloc((void **)&wall_base, n_lat * sizeof(int64_t)));
    PRH_HIP(hipMalloc((void **)&g.occ, (size_t)g.occ_words * sizeof(uint32_t)));
    PRH_HIP(hipMalloc((void **)&g.coarse, (size_t)g.coarse_words * sizeof(uint32_t)));
    PRH_HIP(hipMalloc((void **)&g.cell_code, n_cells * sizeof(uint32_t)));
    PRH_HIP(hipMalloc((void **)&g.hot_map, n_lat * sizeof(int32_t)));
    PRH_HIP(hipMalloc((void **)&g.cold_map, n_lat * sizeof(int32_t)));
    if (g.wocc_words) PRH_HIP(hipMalloc((void **)&g.wocc, (size_t)g.wocc_words * sizeof(uint64_t)));
    const int64_t wcode_n = g.wocc_words * 64;
    if (g.wocc_words) PRH_HIP(hipMalloc((void **)&g.wcode, (size_t)wcode_n * sizeof(uint32_t)));
    PRH_HIP(hipMemcpyAsync(solid, solid_host, n_cells, hipMemcpyHostToDevice, st));

    k_cell_masks<<<grid_for(n_cells), GEO_THREADS, 0, st>>>(solid, nx, ny, nz, mask, face_count);
    k_pack_occupancy<<<grid_for(g.occ_words), GEO_THREADS, 0, st>>>(solid, n_cells, g.occ);
    k_pack_coarse<<<grid_for(g.coarse_words), GEO_THREADS, 0, st>>>(solid, nx, ny, nz, g.coarse_shift, g.cnx, g.cny,
                                                                    g.cnz, g.coarse);
    if (g.wocc)
        k_pack_walled<<<grid_for(g.wocc_words), GEO_THREADS, 0, st>>>(solid, nx, ny, nz, g.wlx, g.wly, g.wlz,
                                                                      g.wocc_words, g.wocc);
    k_wall_counts<<<grid_for(n_lat), GEO_THREADS, 0, st>>>(solid, nx, ny, nz, wall_count);
    PRH_HIP(hipGetLastError());
    int64_t n_internal = 0, n_wall = 0;
    if ((rc = exclusive_scan_i32_to_i64(face_count, face_base, n_cells, &n_internal, st))) return rc;
    if ((rc = exclusive_scan_i32_to_i64(wall_count, wall_base, n_lat, &n_wall, st))) return rc;
    g.n_wall = n_wall;
    g.n_patches = n_wall + n_internal;
    if (g.n_patches >= (int64_t(1) << 26)) {
        set_error("prh_geometry_create: %lld patches exceed the 2^26 patch-id capacity",
                  (long long)g.n_patches);
        return PRH_ERR_CAPACITY;
    }
    const int64_t np = g.n_patches > 0 ? g.n_patches : 1;
    PRH_HIP(hipMalloc((void **)&g.patch_origin, np * 3 * sizeof(double)));
    PRH_HIP(hipMalloc((void **)&g.patch_face, np));
    PRH_HIP(hipMalloc((void **)&g.patch_cell, np * sizeof(int32_t)));
    PRH_HIP(hipMalloc((void **)&g.patch_class, np));
    k_fill_walls<<<grid_for(n_lat), GEO_THREADS, 0, st>>>(solid, nx, ny, nz, wall_base, g.patch_origin,
                                                          g.patch_face, g.patch_cell, g.patch_class, g.hot_map,
                                                          g.cold_map);
    k_fill_internal<<<grid_for(n_cells), GEO_THREADS, 0, st>>>(mask, face_base, nx, ny, nz, n_wall,
                                                               g.patch_origin, g.patch_face, g.patch_cell,
                                                               g.patch_class, g.cell_code);
    if (g.wcode)
        k_pack_wcode<<<grid_for(wcode_n), GEO_THREADS, 0, st>>>(g.cell_code, nx, ny, nz, g.wlx, g.wly, g.wlz, wcode_n,
                                                                 g.wcode);
    PRH_HIP(hipGetLastError());
    PRH_HIP(hipStreamSynchronize(st));
    return PRH_OK;
}

// The occupancy bitsets only (no patch tables): what the Monte Carlo
// kernels read.
int build_occupancy(const uint8_t *solid_host, int nx, int ny, int nz, Geometry &g)
{
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    PRH_HIP(hipGetDevice(&g.device));
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.H_m = 1.0;
    g.n_patches = g.n_wall = 0;
    g.occ_words = ((int64_t)nx * ny * nz + 31) / 32;
    for (g.coarse_shift = 1;; ++g.coarse_shift) {
        const int B = 1 << g.coarse_shift;
        g.cnx = (nx + B - 1) / B;
        g.cny = (ny + B - 1) / B;
        g.cnz = (nz + B - 1) / B;
        g.coarse_words = ((int64_t)g.cnx * g.cny * g.cnz + 31) / 32;
        if (g.coarse_words * 4 <= 72 * 1024) break;
    }
    const int64_t n_cells = (int64_t)nx * ny * nz;
    uint8_t *solid = nullptr;
    PRH_HIP(hipMalloc((void **)&solid, n_cells));
    const hipError_t e1 = hipMalloc((void **)&g.occ, (size_t)g.occ_words * sizeof(uint32_t));
    const hipError_t e2 = hipMalloc((void **)&g.coarse, (size_t)g.coarse_words * sizeof(uint32_t));
    hipError_t e3 = hipSuccess;
    if (e1 == hipSuccess && e2 == hipSuccess) {
        e3 = hipMemcpyAsync(solid, solid_host, n_cells, hipMemcpyHostToDevice, st);
        k_pack_occupancy<<<grid_for(g.occ_words), GEO_THREADS, 0, st>>>(solid, n_cells, g.occ);
        k_pack_coarse<<<grid_for(g.coarse_words), GEO_THREADS, 0, st>>>(solid, nx, ny, nz, g.coarse_shift, g.cnx,
                                                                        g.cny, g.cnz, g.coarse);
        if (e3 == hipSuccess) e3 = hipGetLastError();
        if (e3 == hipSuccess) e3 = hipStreamSynchronize(st);
    }
    (void)hipFree(solid);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
        set_error("prh_geometry_create_occupancy: %s",
                  hipGetErrorString(e1 != hipSuccess ? e1 : (e2 != hipSuccess ? e2 : e3)));
        return PRH_ERR_HIP;
    }
    return PRH_OK;
}

}  // namespace
}  // namespace prh

extern "C" {

int prh_geometry_create_occupancy(const uint8_t *solid_x, int32_t nx, int32_t ny, int32_t nz, prh_geometry **out)
{
    PRH_CHECK_ARG(out && solid_x, "prh_geometry_create_occupancy: null pointer");
    PRH_CHECK_ARG(nx > 0 && ny > 0 && nz > 0, "prh_geometry_create_occupancy: grid sizes must be positive");
    PRH_CHECK_ARG((int64_t)nx * ny * nz < (int64_t(1) << 31), "prh_geometry_create_occupancy: grid too large");
    *out = nullptr;
    prh_geometry *geo = new prh_geometry();
    int rc = prh::build_occupancy(solid_x, nx, ny, nz, geo->g);
    if (rc) {
        prh::free_geometry(geo->g);
        delete geo;
        return rc;
    }
    *out = geo;
    return PRH_OK;
}

int prh_geometry_create(const uint8_t *solid_x, int32_t nx, int32_t ny, int32_t nz, double H_m, prh_geometry **out)
{
    PRH_CHECK_ARG(out && solid_x, "prh_geometry_create: null pointer");
    PRH_CHECK_ARG(nx > 0 && ny > 0 && nz > 0, "prh_geometry_create: grid sizes must be positive");
    PRH_CHECK_ARG((int64_t)nx * ny * nz < (int64_t(1) << 31), "prh_geometry_create: grid too large");
    // the traversal forms cell indices with 24-bit multiplies ((i ny + j) <
    // 2^24) and packs i + 1, j, k into 16-bit fields of its ray records
    PRH_CHECK_ARG((int64_t)nx * ny < (int64_t(1) << 24) && nx < 65535 && ny < 65536 && nz < 65536,
                  "prh_geometry_create: grid too large");
    *out = nullptr;
    prh_geometry *geo = new prh_geometry();  // value-initialised: all pointers null
    int rc = prh::build(solid_x, nx, ny, nz, H_m, geo->g);
    if (rc) {
        prh::free_geometry(geo->g);
        delete geo;
        return rc;
    }
    *out = geo;
    return PRH_OK;
}

int prh_geometry_sizes(const prh_geometry *g, int64_t *n_patches, int64_t *n_wall)
{
    PRH_CHECK_ARG(g, "prh_geometry_sizes: null geometry");
    if (n_patches) *n_patches = g->g.n_patches;
    if (n_wall) *n_wall = g->g.n_wall;
    return PRH_OK;
}

int prh_geometry_patch_table(const prh_geometry *g, int8_t *patch_class, int32_t *patch_cell, int8_t *patch_face,
                             double *patch_origin)
{
    PRH_CHECK_ARG(g, "prh_geometry_patch_table: null geometry");
    const int64_t n = g->g.n_patches;
    if (n == 0) return PRH_OK;
    PRH_HIP(hipSetDevice(g->g.device));
    if (patch_class) PRH_HIP(hipMemcpy(patch_class, g->g.patch_class, n, hipMemcpyDeviceToHost));
    if (patch_cell) PRH_HIP(hipMemcpy(patch_cell, g->g.patch_cell, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (patch_face) PRH_HIP(hipMemcpy(patch_face, g->g.patch_face, n, hipMemcpyDeviceToHost));
    if (patch_origin)
        PRH_HIP(hipMemcpy(patch_origin, g->g.patch_origin, n * 3 * sizeof(double), hipMemcpyDeviceToHost));
    return PRH_OK;
}

void prh_geometry_destroy(prh_geometry *g)
{
    if (!g) return;
    (void)hipSetDevice(g->g.device);
    prh::free_geometry(g->g);
    delete g;
}

}  // extern "C"
