// First-hit tracer: seeded Lambertian rays + FP64 Amanatides-Woo traversal
// over the bit-packed occupancy grid, and per-patch integer counts.
//
// Reference: generate_rays, _initial_cell, _axis_state, _trace_kernel,
// trace_first_hit, first_hit_counts, pool_counts
// (porerad/view_factors.py:135-300).  The traversal repeats the reference's
// FP64 operations one for one (no contraction), so the integer counts are
// bit-identical to the reference CPU implementation on the same seeds.
//
// Pipeline of one prh_trace_counts call (all on the library stream):
//   k_gen_rays       seeded Lambertian rays and their DDA set-up (the FP64
//                    divisions of _axis_state) -> a 64 B start record and a
//                    32 B wall record per ray in HBM; fully converged waves,
//                    each thread jumps its PCG64 stream once and then only
//                    draws;
//   k_trace_lean     FP64 DDA, one ray per lane, over the walled bitset in
//                    4x4x4-block words (the current block's word held per
//                    lane); wavefronts take 512-record grains from a queue,
//                    stage 64 start records at a time in LDS and refill idle
//                    lanes from them; final states (a solid cell's index and
//                    face, or a wall crossing time) leave through LDS in
//                    512 B stores.  k_trace_stream is the earlier variant over
//                    the flat bitset with an LDS brick test;
//   k_count_wave     one wavefront per patch: receivers resolved from the
//                    final states, keys sorted in registers -> unique counts
//                    (k_count_patches: the LDS sort for more keys);
//   k_compact_ranges prefix-summed edge table (origin, receiver, count) in
//                    1 KB output ranges (k_compact_edges where the total stays
//                    on the device).
#include <climits>
#include <mutex>

#include "prh_device.h"
#include "prh_internal.h"

namespace prh {

__constant__ JumpTables c_jump;

// Wave-steps (stepping iterations x steps per iteration, counted per
// wavefront) and refills of the lean traversal since the last reset
// (prh_tracer_stats): with the occupancy reads they give the share of lanes
// that step.  Concurrent launches on one device add into the same counters.
__device__ unsigned long long g_trace_stats[2];

namespace {

constexpr int TRACE_THREADS = 256;
constexpr int STREAM_THREADS = 1024;  // traversal: 16 waves share one LDS brick bitset
constexpr int COUNT_THREADS = 256;
// Rays per patch (and call) that one wavefront counts with its keys in
// registers (32 per lane); beyond, a workgroup sorts them in LDS.
constexpr int COUNT_WAVE_MAX_KEYS = 2048;
constexpr int MAX_RAYS_PER_PATCH_CALL = 16384;

struct GeoView {
    const uint32_t *occ;
    const uint32_t *coarse;
    const uint32_t *cell_code;
    const int32_t *hot_map;
    const int32_t *cold_map;
    const double *patch_origin;
    const int8_t *patch_face;
    const int8_t *patch_class;
    int nx, ny, nz;
    int coarse_shift, cny, cnz;
    uint32_t coarse_words;
    double inv_ny, inv_nz;
    const uint64_t *wocc;  // walled bitset (Geometry::wocc), or null
    LeanLayout wl;         // its index layout
    uint32_t wl_cold;      // dilated i + 1 of the cold wall plane (nx + 1)
    const uint32_t *wcode;  // cell codes in walled-bitset order (Geometry::wcode), or null
    uint32_t wcode_n;       // its entries
};

GeoView view_of(const Geometry &g)
{
    return GeoView{g.occ,    g.coarse, g.cell_code, g.hot_map, g.cold_map, g.patch_origin, g.patch_face,
                   g.patch_class, g.nx, g.ny,     g.nz,      g.coarse_shift, g.cny,    g.cnz,
                   (uint32_t)g.coarse_words,     1.0 / (double)g.ny,   1.0 / (double)g.nz,
                   g.wocc,   g.wocc ? lean_layout(g.wlx, g.wly, g.wlz) : LeanLayout{},
                   g.wocc ? lean_dilate(lean_layout(g.wlx, g.wly, g.wlz), 0, (uint32_t)g.nx + 1u) : 0u,
                   g.wcode,  g.wcode ? (uint32_t)(g.wocc_words * 64) : 0u};
}

// DDA start state of one ray (the set-up of _trace_kernel,
// view_factors.py:211-224), written by the generation or set-up kernel and
// read once by the traversal: 64 B, so a refill is four 16 B loads.
struct alignas(16) DdaRec {
    double tx, ty, tz, ddx, ddy, ddz;
    uint32_t ij;   // (i + 1) | jj << 16
    uint32_t ks;   // kk | (sx + 1) << 16 | (sy + 1) << 18 | (sz + 1) << 20
    uint32_t cap;  // step cap; 0 for a ray with dx == 0 (it fails)
    uint32_t P;    // the start cell's walled-bitset index (0 without a walled bitset)
};

// What a wall receiver needs (view_factors.py:226-231), read only for rays
// that reach a wall: 32 B.
struct alignas(16) WallRec {
    double oy, oz, dy, dz;
};

// Final state of a traversed ray, 8 B.  A ray that the lean traversal ends
// on a wall keeps the time of its last crossing as a double, negated for the
// cold wall; the wall cell is worked out from it (and the ray's wall record)
// when the receiver is resolved.  Every other final state is a NaN: high word
// FIN_BOX | kind << 16 | low 16 bits, low word a 32-bit payload:
//   FIN_FAILED    nothing;
//   FIN_SOLID     payload the cell index, low bits the face;
//   FIN_HOT/COLD  payload the wall cell's j, low bits its k;
//   FIN_SOLID_P   payload the cell's walled-bitset index P (Geometry::wocc),
//                 low bits the face.
enum : uint32_t { FIN_FAILED = 0, FIN_SOLID = 1, FIN_HOT = 2, FIN_COLD = 3, FIN_SOLID_P = 4 };
constexpr uint32_t FIN_BOX = 0x7FF00000u;

__device__ __forceinline__ uint2 fin_box(uint32_t kind, uint32_t payload, uint32_t low16)
{
    return make_uint2(payload, FIN_BOX | kind << 16 | low16);
}

// Per-lane traversal state (the loop variables of _trace_kernel).  The face
// of the last crossing is kept as its axis (priority x > y > z) and turned
// into the face index 0..5 only when the ray ends.  The step cap is a
// countdown; why a ray stopped follows from the state: left == 0 is the step
// cap, i outside [0, nx) the wall, anything else a solid cell.  The number
// of occupancy reads follows from the steps taken.
struct Ray {
    double tx, ty, tz, ddx, ddy, ddz, time;
    int i, jj, kk;
    int sx, sy, sz;
    int axis;       // axis of the last crossing, -1 before the first
    uint32_t left;  // iterations left before the step cap
    uint32_t cap;   // step cap of this ray
};

__device__ __forceinline__ int normal_code_of(const GeoView &G, int64_t p)
{
    const int cls = G.patch_class[p];
    return cls == 2 ? (int)G.patch_face[p] : cls;
}

// v_min_f64 without the NaN canonicalisation the compiler adds around fmin
// for loop-carried values; the crossing times are never NaN (finite or the
// 1e300 sentinel), and for equal values either operand is the same number.
__device__ __forceinline__ double vmin_f64(double a, double b)
{
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// a[23:0] * b[23:0] + c as one v_mad_u32_u24 (b uniform); written out because the
// compiler folds 24-bit multiply + add into v_mad_u64_u32, a quarter-rate
// instruction.
__device__ __forceinline__ uint32_t mad_u24(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
    return r;
}

// a * b + c rounded once, as one v_fma_f64 (the compiler would turn an fma
// with a 0/1 factor back into an addition and two selects).
__device__ __forceinline__ double fma_f64(double a, double b, double c)
{
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// Launch constants of the traversal.
struct TraversalConst {
    uint32_t min_cap;    // 10 nx ny nz + 10 (view_factors.py:224)
    double grazing_dx;   // below this |dx| the crossing-scale cap may exceed min_cap
};

TraversalConst traversal_const(const Geometry &g)
{
    TraversalConst c;
    const int64_t min_cap = 10 * (int64_t)g.nx * g.ny * g.nz + 10;
    c.min_cap = (uint32_t)(min_cap < 0xFFFFFFFFll ? min_cap : 0xFFFFFFFFll);
    // crossing_scale = 8 nx (1 + (|dy|+|dz|)/|dx|) + 100 <= 8 nx + 16 nx/|dx| + 100,
    // which is <= min_cap once |dx| >= 16 nx / (min_cap - 8 nx - 100); a factor 2
    // keeps the test conservative against rounding.
    const double room = (double)min_cap - 8.0 * g.nx - 100.0;
    c.grazing_dx = room > 0.0 ? 2.0 * 16.0 * g.nx / room : 2.0;
    return c;
}

// The DDA set-up of one ray from its origin and direction
// (view_factors.py:211-224): initial cell (_initial_cell), per-axis step,
// crossing time and increment (_axis_state), the periodic wrap of the
// lateral indices and the step cap.  A ray with dx == 0 gets cap 0.
// spread: a table for lean_dilate_table (in LDS), or null.  GENERATED: the
// ray comes from make_ray (unit direction, origin a multiple of 2^-53), so
// the set-up's divisions need no IEEE special-case handling.
template <bool GENERATED = false>
__device__ __forceinline__ void make_dda(const double o[3], const double d[3], const GeoView &G,
                                         const TraversalConst &C, DdaRec &rec, WallRec &w,
                                         const uint32_t *spread = nullptr)
{
    w.oy = o[1];
    w.oz = o[2];
    w.dy = d[1];
    w.dz = d[2];
    if (d[0] == 0.0) {
        rec.tx = rec.ty = rec.tz = rec.ddx = rec.ddy = rec.ddz = 0.0;
        rec.ij = rec.ks = rec.cap = rec.P = 0u;
        return;
    }
    const double i = initial_cell(o[0], d[0]);
    const double j = initial_cell(o[1], d[1]);
    const double k = initial_cell(o[2], d[2]);
    int sx, sy, sz;
    axis_state<GENERATED>(o[0], d[0], i, sx, rec.tx, rec.ddx);
    axis_state<GENERATED>(o[1], d[1], j, sy, rec.ty, rec.ddy);
    axis_state<GENERATED>(o[2], d[2], k, sz, rec.tz, rec.ddz);
    // The reference keeps j, k unbounded and reads j % ny, k % nz; the
    // traversal carries the wrapped indices from the start.  A ray whose
    // x index starts outside [0, nx) ends at once on the wall of that side
    // (hot below, cold above), so i is clamped to [-1, nx].
    const int jj = floor_mod(j, G.ny, G.inv_ny);
    const int kk = floor_mod(k, G.nz, G.inv_nz);
    const int ic = i < 0.0 ? -1 : (i >= (double)G.nx ? G.nx : (int)i);
    rec.ij = (uint32_t)(ic + 1) | (uint32_t)jj << 16;
    rec.ks = (uint32_t)kk | (uint32_t)(sx + 1) << 16 | (uint32_t)(sy + 1) << 18 | (uint32_t)(sz + 1) << 20;
    uint32_t cap = C.min_cap;
    if (fabs(d[0]) < C.grazing_dx) cap = max_steps_of(d[0], d[1], d[2], G.nx, G.ny, G.nz);
    rec.cap = cap;
    // the start cell's walled-bitset index, for the lean traversal
    rec.P = !G.wocc ? 0u
            : spread ? lean_dilate_table(G.wl, 0, (uint32_t)(ic + 1), spread) |
                           lean_dilate_table(G.wl, 1, (uint32_t)jj, spread) |
                           lean_dilate_table(G.wl, 2, (uint32_t)kk, spread)
                     : lean_dilate(G.wl, 0, (uint32_t)(ic + 1)) | lean_dilate(G.wl, 1, (uint32_t)jj) |
                           lean_dilate(G.wl, 2, (uint32_t)kk);
}

// Traversal state of ray g from its record; false for a ray that fails at once.
// false for a ray that fails (dx == 0).
__device__ __forceinline__ bool ray_load(Ray &R, const DdaRec *recs, uint32_t g)
{
    const double2 *p = reinterpret_cast<const double2 *>(recs + g);
    const double2 a = p[0], b = p[1], c = p[2];
    const uint4 w = reinterpret_cast<const uint4 *>(recs + g)[3];
    R.tx = a.x;
    R.ty = a.y;
    R.tz = b.x;
    R.ddx = b.y;
    R.ddy = c.x;
    R.ddz = c.y;
    R.i = (int)(w.x & 0xFFFFu) - 1;
    R.jj = (int)(w.x >> 16);
    R.kk = (int)(w.y & 0xFFFFu);
    R.sx = (int)((w.y >> 16) & 3u) - 1;
    R.sy = (int)((w.y >> 18) & 3u) - 1;
    R.sz = (int)((w.y >> 20) & 3u) - 1;
    R.cap = w.z;
    R.left = w.z;
    R.axis = -1;
    R.time = 0.0;
    return w.z != 0u;
}

// How the traversal tests a cell's occupancy.
enum OccMode : int {
    OCC_GLOBAL = 0,  // read the cell's bit from the global bitset (L2-resident)
    OCC_BRICK = 1,   // the brick bitset held in LDS first; the global bit only
                     // for cells in bricks that hold solid
    OCC_AHEAD = 2,   // the global bitset word loaded ahead: the step is computed
                     // while the load is in flight and undone on a solid cell
    OCC_LEAN = 3,    // the walled bitset indexed by one packed cell index (lean_step)
};

// Flag added to an occupancy mode when ny and nz are powers of two: the
// periodic wrap of the lateral cell indices is then one AND with n - 1.
constexpr int OCC_POW2 = 4;
constexpr int occ_mode(int occ) { return occ & 3; }

// Index of the cell (i, jj, kk) in the flat bitset.  24-bit multiply-adds:
// i ny + j and nz stay below 2^24 and the index below 2^31 (checked at
// geometry creation).
__device__ __forceinline__ uint32_t cell_index(const Ray &R, const GeoView &G)
{
    return mad_u24(mad_u24((uint32_t)R.i, (uint32_t)G.ny, (uint32_t)R.jj), (uint32_t)G.nz, (uint32_t)R.kk);
}

__device__ __forceinline__ uint32_t occupancy_word(const GeoView &G, uint32_t cell)
{
    return *(const uint32_t *)((const char *)G.occ + ((cell >> 3) & ~3u));
}

// Occupancy test of the current cell (OCC_GLOBAL, OCC_BRICK).
template <int OCC>
__device__ __forceinline__ bool cell_is_solid(const Ray &R, const GeoView &G, const uint32_t *lds_coarse)
{
    if (occ_mode(OCC) == OCC_BRICK) {
        const unsigned s = (unsigned)G.coarse_shift;
        const uint32_t b = mad_u24(mad_u24((uint32_t)R.i >> s, (uint32_t)G.cny, (uint32_t)R.jj >> s),
                                   (uint32_t)G.cnz, (uint32_t)R.kk >> s);
        if (!((lds_coarse[b >> 5] >> (b & 31u)) & 1u)) return false;
    }
    const uint32_t cell = cell_index(R, G);
    return (occupancy_word(G, cell) >> (cell & 31u)) & 1u;
}

// The crossing of view_factors.py:237-255: every axis whose crossing time is
// within the tie tolerance of the minimum advances; the face is the first
// of them in x > y > z order.  The minimum is the value of one of the three
// times, so at least one axis crosses.  Updates are selects (the wave stays
// converged); a lateral index leaving [0, n) is wrapped by subtracting
// step * n (POW2: by masking with n - 1).
template <bool POW2>
__device__ __forceinline__ void ray_advance(Ray &R, const GeoView &G)
{
    const double nt = vmin_f64(vmin_f64(R.tx, R.ty), R.tz);
    const bool cx = fabs(R.tx - nt) <= TIE_TOLERANCE;
    const bool cy = fabs(R.ty - nt) <= TIE_TOLERANCE;
    const bool cz = fabs(R.tz - nt) <= TIE_TOLERANCE;
    // t += dd on a crossing as fma(c, dd, t) with c in {0, 1}: fma(1, dd, t)
    // rounds t + dd exactly as the addition does, and fma(0, dd, t) is t
    // (dd is finite; a -0 start time may become +0, which no comparison or
    // product distinguishes).
    R.i += cx ? R.sx : 0;
    R.tx = fma_f64(cx ? 1.0 : 0.0, R.ddx, R.tx);
    const int jn = R.jj + (cy ? R.sy : 0);
    if (POW2) R.jj = jn & (G.ny - 1);
    else R.jj = (unsigned)jn < (unsigned)G.ny ? jn : jn - __mul24(R.sy, G.ny);
    R.ty = fma_f64(cy ? 1.0 : 0.0, R.ddy, R.ty);
    const int kn = R.kk + (cz ? R.sz : 0);
    if (POW2) R.kk = kn & (G.nz - 1);
    else R.kk = (unsigned)kn < (unsigned)G.nz ? kn : kn - __mul24(R.sz, G.nz);
    R.tz = fma_f64(cz ? 1.0 : 0.0, R.ddz, R.tz);
    R.axis = cx ? 0 : (cy ? 1 : 2);
    R.time = nt;
}

// One iteration of the traversal loop, view_factors.py:225-255; returns
// whether the ray goes on.
template <int OCC>
__device__ __forceinline__ bool ray_step(Ray &R, const GeoView &G, const uint32_t *lds_coarse)
{
    constexpr bool POW2 = (OCC & OCC_POW2) != 0;
    if ((unsigned)R.i >= (unsigned)G.nx) return false;
    if (occ_mode(OCC) == OCC_AHEAD) {
        const uint32_t cell = cell_index(R, G);
        const uint32_t word = occupancy_word(G, cell);
        const int i0 = R.i, j0 = R.jj, k0 = R.kk, axis0 = R.axis;
        ray_advance<POW2>(R, G);
        if ((word >> (cell & 31u)) & 1u) {
            // the current cell was solid: the ray ends here, as if it had not moved
            R.i = i0;
            R.jj = j0;
            R.kk = k0;
            R.axis = axis0;
            return false;
        }
        return --R.left != 0u;
    }
    if (cell_is_solid<OCC>(R, G, lds_coarse)) return false;
    ray_advance<POW2>(R, G);
    return --R.left != 0u;
}

// Occupancy reads of a finished traversal: one per step plus the read that
// found the solid cell.
__device__ __forceinline__ uint32_t ray_reads(const Ray &R, const GeoView &G)
{
    const bool at_solid = R.left != 0u && (unsigned)R.i < (unsigned)G.nx;
    return (R.cap - R.left) + (at_solid ? 1u : 0u);
}

// Python's floor(v) % n for the wall cell index (view_factors.py:226-231),
// with |floor(v)| far below 2^52: r = f - n floor(f / n) from an inexact
// quotient, then one correction either way (all values are exact integers).
__device__ __forceinline__ uint32_t wrap_floor(double v, int n, double inv_n)
{
    const double f = floor(v);
    const double dn = (double)n;
    double r = f - dn * floor(f * inv_n);
    r = r < 0.0 ? r + dn : r;
    r = r >= dn ? r - dn : r;
    return (uint32_t)r;
}

// Final state of a finished traversal (see FIN_*).
__device__ __forceinline__ uint2 ray_final(const Ray &R, const GeoView &G, const WallRec *walls, uint32_t g)
{
    if (R.left == 0u) return fin_box(FIN_FAILED, 0u, 0u);
    if ((unsigned)R.i >= (unsigned)G.nx) {
        const double2 *p = reinterpret_cast<const double2 *>(walls + g);
        const double2 o = p[0], d = p[1];
        const double yw = o.x + d.x * R.time;
        const double zw = o.y + d.y * R.time;
        const uint32_t jw = wrap_floor(yw, G.ny, G.inv_ny);
        const uint32_t kw = wrap_floor(zw, G.nz, G.inv_nz);
        return fin_box(R.i < 0 ? FIN_HOT : FIN_COLD, jw, kw);
    }
    if (R.axis < 0) return fin_box(FIN_FAILED, 0u, 0u);
    // Direction bits as values (a select between struct fields would be
    // lowered to an indexed stack access).
    const uint32_t up = (R.sx > 0 ? 1u : 0u) | (R.sy > 0 ? 2u : 0u) | (R.sz > 0 ? 4u : 0u);
    const uint32_t face = 2u * (uint32_t)R.axis + ((up >> R.axis) & 1u);
    return fin_box(FIN_SOLID, cell_index(R, G), face);
}

// Receiving patch of a final state (view_factors.py:226-236), -1 on failure:
// the wall maps (the wall cell given, or worked out from the crossing time
// as ray_final does), or the patch of (cell, face) from the cell code (first
// patch id of the cell << 6 | mask of its exposed faces).  w is the ray's
// wall record.
__device__ __forceinline__ int patch_of_code(uint32_t code, uint32_t face)
{
    const uint32_t m = code & 63u;
    if (!((m >> face) & 1u)) return -1;
    return (int)(code >> 6) + __popc(m & ((1u << face) - 1u));
}

__device__ __forceinline__ int patch_of_face(uint32_t cell, uint32_t face, const GeoView &G)
{
    const uint32_t code = G.cell_code[cell];
    const uint32_t m = code & 63u;
    if (!((m >> face) & 1u)) return -1;
    return (int)(code >> 6) + __popc(m & ((1u << face) - 1u));
}

__device__ __forceinline__ bool fin_is_time(uint2 f)
{
    return (f.y & FIN_BOX) != FIN_BOX;
}

// The wall patch of a crossing time: floor(o + d t) mod n as ray_final forms it.
__device__ __forceinline__ int resolve_wall_time(uint2 f, const GeoView &G, const WallRec &w)
{
    const double time = __hiloint2double((int)(f.y & 0x7FFFFFFFu), (int)f.x);
    const double yw = w.oy + w.dy * time;
    const double zw = w.oz + w.dz * time;
    const uint32_t jw = wrap_floor(yw, G.ny, G.inv_ny);
    const uint32_t kw = wrap_floor(zw, G.nz, G.inv_nz);
    const size_t c = (size_t)jw * (size_t)G.nz + kw;
    return (f.y >> 31) ? G.cold_map[c] : G.hot_map[c];
}

// The receiver of a boxed final state (any but a crossing time).
__device__ __forceinline__ uint32_t fin_kind(uint2 f) { return (f.y >> 16) & 15u; }

// Entry of Geometry::wcode for a FIN_SOLID_P final state whose index is in
// range (bits 6-8 of P zero, entry below wcode_n), else 0.
__device__ __forceinline__ uint32_t fin_wcode_index(uint2 f, const GeoView &G)
{
    const uint32_t c = (f.x >> (6 + LEAN_GAP)) << 6 | (f.x & 63u);
    const bool ok = !fin_is_time(f) && fin_kind(f) == FIN_SOLID_P && ((f.x >> 6) & ((1u << LEAN_GAP) - 1u)) == 0u &&
                    c < G.wcode_n;
    return ok ? c : 0u;
}

// Receiver of a FIN_SOLID_P final state given its cell code (loaded from
// Geometry::wcode at fin_wcode_index), or -1.
__device__ __forceinline__ int resolve_solid_p(uint2 f, uint32_t code, const GeoView &G)
{
    const uint32_t low = f.y & 0xFFFFu;
    if (low > 5u || fin_wcode_index(f, G) == 0u) return -1;
    return patch_of_code(code, low);
}

__device__ __forceinline__ int resolve_boxed(uint2 f, const GeoView &G)
{
    const uint32_t kind = (f.y >> 16) & 15u;
    const uint32_t low = f.y & 0xFFFFu;
    if (kind == FIN_SOLID) {
        if ((uint64_t)f.x >= (uint64_t)G.nx * (uint64_t)G.ny * (uint64_t)G.nz || low > 5u) return -1;
        return patch_of_face(f.x, low, G);
    }
    if (kind == FIN_SOLID_P) {
        // bits 6-8 of P are zero; the table is dense in the other bits
        const uint32_t c = (f.x >> (6 + LEAN_GAP)) << 6 | (f.x & 63u);
        if (!G.wcode || low > 5u || (f.x >> 6) & ((1u << LEAN_GAP) - 1u) || c >= G.wcode_n) return -1;
        return patch_of_code(G.wcode[c], low);
    }
    if (kind == FIN_HOT || kind == FIN_COLD) {
        if (f.x >= (uint32_t)G.ny || low >= (uint32_t)G.nz) return -1;
        const size_t c = (size_t)f.x * (size_t)G.nz + low;
        return kind == FIN_HOT ? G.hot_map[c] : G.cold_map[c];
    }
    return -1;
}

__device__ __forceinline__ int resolve_receiver(uint2 f, const GeoView &G, const WallRec &w)
{
    return fin_is_time(f) ? resolve_wall_time(f, G, w) : resolve_boxed(f, G);
}

__device__ __forceinline__ void load_seed(const uint64_t *seeds, int s, u128 &st, u128 &inc)
{
    st = make_u128(seeds[4 * s + 0], seeds[4 * s + 1]);
    inc = make_u128(seeds[4 * s + 2], seeds[4 * s + 3]);
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// ------------------------------------------------------------------ ray generation
// Ray g of a launch is (seed s, patch p, ray r), g = s*(n_p*R) + (p-p0)*R + r,
// and uses draws 4 (p R + r) .. 4 (p R + r) + 3 of seed s's PCG64 stream.  A
// wavefront generates GEN_ROUNDS x 64 consecutive rays; lane l takes rays
// base + l + 64 k, so every round stores 64 adjacent records (coalesced) and
// a lane's next ray is 256 draws further on its stream: one jump with the
// fixed 2^8 table entry instead of a full jump-ahead, which each lane makes
// once (and again when its rays cross into the next seed's block).
constexpr uint32_t GEN_ROUNDS = 64;

struct GenArgs {
    GeoView G;
    TraversalConst C;
    uint32_t p0, n_p, R;
    // A patch's rays r = 0 .. R-1 of a launch are rays r0 + r of a block of
    // R_stream rays per patch: stream position 4 (p R_stream + r0 + r).  A
    // whole block is r0 = 0, R_stream = R.
    uint32_t R_stream, r0;
    const uint64_t *seeds;  // S x {state_hi, state_lo, inc_hi, inc_lo}
    uint32_t total;         // rays of this launch (or segment)
    uint32_t offset;        // index of its first record in the output arrays
    double *origins;     // total x 3 (rays only)
    double *directions;  // total x 3 (rays only)
    DdaRec *dda;         // records only
    WallRec *walls;      // records only
};

// The GEN_ROUNDS rays of one lane: rays base, base + 64, ... of the launch.
// RECORDS: write the traversal records (DDA start state, wall inputs)
// instead of the origins and directions, so the divisions of the set-up run
// here in converged waves rather than in the traversal's refills.
// The table of lean_dilate_table in the workgroup's LDS (every thread calls
// this before any returns).
__device__ __forceinline__ void fill_spread_table(uint32_t *t)
{
    for (int i = threadIdx.x; i < LEAN_SPREAD_ENTRIES; i += blockDim.x) t[i] = lean_open_gap(morton_spread((uint32_t)i));
    __syncthreads();
}

template <bool RECORDS>
__device__ __forceinline__ void gen_rays_body(const GenArgs &a, uint64_t base, const uint32_t *spread)
{
    const GeoView &G = a.G;
    const uint32_t seg = a.n_p * a.R;
    const double fny = (double)G.ny, fnz = (double)G.nz;
    const uint32_t lane = threadIdx.x & 63u;
    // records leave through LDS: each round's 64 records (4 KB + 2 KB of wall
    // inputs) are stored by consecutive lanes in consecutive 16 B pieces
    __shared__ uint4 s_dda[TRACE_THREADS / 64][64 * 4];
    __shared__ uint4 s_wall[TRACE_THREADS / 64][64 * 2];
    uint4 *sd = s_dda[threadIdx.x >> 6], *sw = s_wall[threadIdx.x >> 6];
    const uint64_t wave_base = base - lane;
    if (wave_base >= a.total) return;
    // (seed, patch, ray) of the lane's first ray; later rays are 64 further on
    uint32_t s = (uint32_t)base / seg;
    uint32_t q = ((uint32_t)base - s * seg) / a.R;
    uint32_t r = (uint32_t)base - s * seg - q * a.R;
    uint32_t s_cur = 0xFFFFFFFFu, q_cur = 0xFFFFFFFFu;
    // the stream runs on from one patch's rays to the next only for whole blocks
    const bool contiguous = a.R == a.R_stream;
    u128 st = make_u128(0, 0), inc = make_u128(0, 0), step_inc = make_u128(0, 0);
    for (uint32_t k = 0; k < GEN_ROUNDS; ++k) {
        const uint64_t round0 = wave_base + 64ull * k;  // first ray of the round (lane 0's)
        if (round0 >= a.total) break;                   // the same for every lane of the wave
        const uint64_t g64 = round0 + lane;
        const bool valid = g64 < a.total;
        const uint32_t g = (uint32_t)g64 + a.offset;
        if (k > 0) {
            r += 64u;
            while (r >= a.R) {
                r -= a.R;
                if (++q == a.n_p) {
                    q = 0;
                    ++s;
                }
            }
        }
        DdaRec rec;
        WallRec w;
        double o[3], d[3];
        if (valid) {
            const uint32_t p = a.p0 + q;
            if (s != s_cur || (!contiguous && q != q_cur)) {
                // first ray of this lane, or the stride crossed into the next
                // seed (or, for part of a block, into the next patch)
                u128 st0;
                load_seed(a.seeds, (int)s, st0, inc);
                st = pcg_jump(st0, inc, 4ull * ((uint64_t)p * a.R_stream + a.r0 + r), c_jump);
                step_inc = mul128(c_jump.G[8], inc);
                s_cur = s;
                q_cur = q;
            }
            u128 next = add128(mul128(c_jump.A[8], st), step_inc);  // 64 rays = 256 draws on
            const double *po = G.patch_origin + 3 * (size_t)p;
            make_ray(st, inc, po[0], po[1], po[2], normal_code_of(G, p), fny, fnz, o, d);
            st = next;
            if (RECORDS) make_dda<true>(o, d, G, a.C, rec, w, spread);
        }
        if (RECORDS) {
            if (valid) {
                const uint4 *rv = reinterpret_cast<const uint4 *>(&rec);
                const uint4 *wv = reinterpret_cast<const uint4 *>(&w);
                sd[4 * lane + 0] = rv[0];
                sd[4 * lane + 1] = rv[1];
                sd[4 * lane + 2] = rv[2];
                sd[4 * lane + 3] = rv[3];
                sw[2 * lane + 0] = wv[0];
                sw[2 * lane + 1] = wv[1];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            const uint64_t n_round = a.total - round0 < 64 ? a.total - round0 : 64;  // records of this round
            uint4 *dd = reinterpret_cast<uint4 *>(a.dda + (uint32_t)round0 + a.offset);
            uint4 *wd = reinterpret_cast<uint4 *>(a.walls + (uint32_t)round0 + a.offset);
            if (n_round == 64) {  // every round but a launch's last: no per-piece checks
#pragma unroll
                for (int m = 0; m < 4; ++m) dd[lane + 64u * m] = sd[lane + 64u * m];
#pragma unroll
                for (int m = 0; m < 2; ++m) wd[lane + 64u * m] = sw[lane + 64u * m];
            } else {
                const uint32_t nr = (uint32_t)n_round;
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const uint32_t piece = lane + 64u * m;
                    if ((piece >> 2) < nr) dd[piece] = sd[piece];
                }
#pragma unroll
                for (int m = 0; m < 2; ++m) {
                    const uint32_t piece = lane + 64u * m;
                    if ((piece >> 1) < nr) wd[piece] = sw[piece];
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        } else if (valid) {
            double *oo = a.origins + 3 * (size_t)g;
            double *dd = a.directions + 3 * (size_t)g;
            oo[0] = o[0];
            oo[1] = o[1];
            oo[2] = o[2];
            dd[0] = d[0];
            dd[1] = d[1];
            dd[2] = d[2];
        }
    }
}

template <bool RECORDS>
__global__ __launch_bounds__(TRACE_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_gen_rays(GenArgs a)
{
    __shared__ uint32_t s_spread[LEAN_SPREAD_ENTRIES];
    fill_spread_table(s_spread);
    const uint64_t wave = ((uint64_t)blockIdx.x * TRACE_THREADS + threadIdx.x) >> 6;
    gen_rays_body<RECORDS>(a, wave * (64ull * GEN_ROUNDS) + (threadIdx.x & 63u), s_spread);
}

// One segment of a multi-geometry batch: a patch range of one geometry,
// traced with one seed.  Its rays occupy records [offset, offset + total).
struct Segment {
    GenArgs gen;
    uint32_t gen_wave_begin;    // first generation wavefront of the segment
    uint32_t trace_block_begin;  // first traversal workgroup of the segment
    uint32_t grain;              // records per queue grab
    unsigned long long *queue;   // traversal queue of the segment (zero at launch)
    unsigned long long *reads;   // occupancy-read counter of the segment
};

__device__ __forceinline__ int segment_of(const Segment *segs, int n_seg, uint32_t x, bool trace)
{
    int s = 0;
    for (int k = 1; k < n_seg; ++k)
        if ((trace ? segs[k].trace_block_begin : segs[k].gen_wave_begin) <= x) s = k;
    return s;
}

// Records of every segment of a batch in one launch (each wavefront works
// inside one segment).
__global__ __launch_bounds__(TRACE_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_gen_rays_batch(
    const Segment *segs, int n_seg)
{
    __shared__ uint32_t s_spread[LEAN_SPREAD_ENTRIES];
    fill_spread_table(s_spread);
    // wave-uniform, so the segment's arguments load into scalar registers
    const uint32_t wave = __builtin_amdgcn_readfirstlane((blockIdx.x * TRACE_THREADS + threadIdx.x) >> 6);
    const int k = segment_of(segs, n_seg, wave, false);
    const GenArgs a = segs[k].gen;
    gen_rays_body<true>(a, (uint64_t)(wave - segs[k].gen_wave_begin) * (64ull * GEN_ROUNDS) + (threadIdx.x & 63u),
                        s_spread);
}

// Traversal records of given rays (prh_trace_rays).
__global__ __launch_bounds__(TRACE_THREADS) void k_setup_rays(GeoView G, TraversalConst C, const double *origins,
                                                              const double *directions, uint32_t total, DdaRec *dda,
                                                              WallRec *walls)
{
    const uint32_t g = blockIdx.x * TRACE_THREADS + threadIdx.x;
    if (g >= total) return;
    const double o[3] = {origins[3 * (size_t)g], origins[3 * (size_t)g + 1], origins[3 * (size_t)g + 2]};
    const double d[3] = {directions[3 * (size_t)g], directions[3 * (size_t)g + 1], directions[3 * (size_t)g + 2]};
    DdaRec rec;
    WallRec w;
    make_dda(o, d, G, C, rec, w);
    dda[g] = rec;
    walls[g] = w;
}

// ------------------------------------------------------------------ traversal
// Each wavefront takes grains of consecutive records from a shared counter
// (the queue) and streams them through its lanes: it keeps stepping while
// its lanes are busy; once at least REFILL_MIN lanes are idle, those lanes
// write their final states and load the next records in lane order
// (adjacent 64 B records, so the refill loads coalesce).  Work taken from a
// queue rather than a fixed range keeps every wavefront busy until the
// records run out, whatever the spread of ray lengths, so the launch ends
// within about one grain of work.
struct TraceArgs {
    GeoView G;
    uint32_t total, grain;
    const DdaRec *dda;
    const WallRec *walls;
    uint2 *fin;
    unsigned long long *reads;
    unsigned long long *queue;  // records handed out so far; zero at launch
};

// Records [base, base + total) streamed by the wavefronts that share queue.
template <int REFILL_MIN, int OCC>
__device__ __forceinline__ void trace_body(const GeoView &G, uint32_t base, uint32_t total, uint32_t grain,
                                           unsigned long long *queue, const DdaRec *dda, const WallRec *walls,
                                           uint2 *fin, unsigned long long *reads_out, const uint32_t *lds_coarse)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t cursor = 0, w_end = 0;  // the wavefront's grain [cursor, w_end)
    bool more = true;                // the queue may still hold records

    Ray R;
    R.i = 0;
    R.left = 1;
    bool active = false, pending = false;
    uint32_t g_cur = 0;
    uint64_t reads = 0;

    while (true) {
        if (!active && pending) {
            reads += ray_reads(R, G);
            fin[g_cur] = ray_final(R, G, walls, g_cur);
            pending = false;
        }
        if (more && cursor >= w_end) {
            // next grain (the counter is 64-bit: it passes total without wrapping)
            unsigned long long first = 0;
            if (lane == 0) first = atomicAdd(queue, (unsigned long long)grain);
            first = __shfl(first, 0);
            if (first >= total) {
                more = false;
            } else {
                cursor = base + (uint32_t)first;
                const unsigned long long last = first + grain;
                w_end = base + (last < total ? (uint32_t)last : total);
            }
        }
        if (cursor < w_end) {
            const uint64_t idle = __ballot(!active);
            if (!active) {
                const uint32_t g = cursor + (uint32_t)__popcll(idle & below);
                if (g < w_end) {
                    if (ray_load(R, dda, g)) {
                        active = true;
                        pending = true;
                        g_cur = g;
                    } else {
                        fin[g] = fin_box(FIN_FAILED, 0u, 0u);
                    }
                }
            }
            cursor += (uint32_t)__popcll(idle);
        }
        if (__ballot(active) == 0) {
            if (!more) break;  // no grain left (cursor >= w_end)
            continue;
        }
        while (true) {
            // two steps per wave-level check
            if (active) active = ray_step<OCC>(R, G, lds_coarse);
            if (active) active = ray_step<OCC>(R, G, lds_coarse);
            const uint64_t act = __ballot(active);
            if (act == 0) break;
            if ((more || cursor < w_end) && __popcll(~act) >= REFILL_MIN) break;
        }
    }
    reads = wave_sum_u64(reads);
    if (lane == 0 && reads) atomicAdd(reads_out, (unsigned long long)reads);
}

template <int REFILL_MIN, int OCC>
__global__ __launch_bounds__(STREAM_THREADS) __attribute__((amdgpu_waves_per_eu(8)))
void k_trace_stream(TraceArgs a)
{
    const GeoView &G = a.G;
    extern __shared__ uint32_t lds_coarse[];
    if (occ_mode(OCC) == OCC_BRICK) {
        for (uint32_t w = threadIdx.x; w < G.coarse_words; w += STREAM_THREADS) lds_coarse[w] = G.coarse[w];
        __syncthreads();
    }
    trace_body<REFILL_MIN, OCC>(G, 0u, a.total, a.grain, a.queue, a.dda, a.walls, a.fin, a.reads, lds_coarse);
}

// Every piece of a chunk in one launch: each workgroup works inside one piece
// (its LDS holds that geometry's brick bitset) and its wavefronts take grains
// from the piece's queue.  Each piece has up to a chip's worth of
// workgroups, dispatched in piece order, so the workgroups of the next piece
// move onto the compute units as the previous piece drains.
template <int REFILL_MIN, int OCC>
__global__ __launch_bounds__(STREAM_THREADS) __attribute__((amdgpu_waves_per_eu(8)))
void k_trace_batch(const Segment *segs, int n_seg, const DdaRec *dda, const WallRec *walls, uint2 *fin)
{
    const int k = segment_of(segs, n_seg, blockIdx.x, true);
    const GeoView G = segs[k].gen.G;
    extern __shared__ uint32_t lds_coarse[];
    if (occ_mode(OCC) == OCC_BRICK) {
        for (uint32_t w = threadIdx.x; w < G.coarse_words; w += STREAM_THREADS) lds_coarse[w] = G.coarse[w];
        __syncthreads();
    }
    trace_body<REFILL_MIN, OCC>(G, segs[k].gen.offset, segs[k].gen.total, segs[k].grain, segs[k].queue, dda, walls,
                                fin, segs[k].reads, lds_coarse);
}

// ------------------------------------------------------------------ lean traversal
// The same traversal over the walled bitset (Geometry::wocc), which holds the
// cells in 4x4x4 blocks of one 64-bit word and sets the two wall planes, so
// one bit test ends a ray at a solid cell or at either wall.  Each lane keeps
// the word of its current block in registers and loads the next one only when
// the ray leaves the block (about one step in four), which keeps the requests
// to L2 near those of the brick test without its index arithmetic.  The cell
// is held as three dilated values (each axis's bits of the index P, see
// LeanLayout): a crossing is (d - s) & mask with s = mask for +1 and 1 for -1,
// and P is their OR.  A lateral size that is not a power of two wraps by one
// compare and select.  No face is kept per step: the dilated steps of the
// last crossing are (x and y; z crossed when neither did), so the face
// follows from them when the ray ends.
struct LeanRay {
    double tx, ty, tz;
    // The smallest crossing time of each step, written by even steps of the
    // ray (counted from its first) to nt0 and by odd steps to nt1, before the
    // step finds out whether the ray goes on: the time of the last crossing
    // is in the register of the step before the one that stopped the ray.
    double nt0, nt1;
    // The crossing increments of t: low words of ddx, ddy, ddz with their
    // high words set for the axes that cross (0 otherwise: the value is then
    // a subnormal or zero, and t, which is always a positive normal number,
    // does not change when it is added).
    double incx, incy, incz;
    uint32_t hix, hiy, hiz;  // high words of ddx, ddy, ddz
    uint32_t xd, yd, zd;     // dilated i + 1, j, k
    uint32_t sx, sy, sz;     // dilated steps: the axis mask (+1) or its lowest bit (-1)
    uint32_t last_x, last_y; // the x and y steps of the last crossing (0: not crossed)
    uint32_t from_y, to_y, from_z, to_z;  // lateral wraps (sizes that are not powers of two)
    uint32_t P;              // xd | yd | zd: the cell's bit in the walled bitset
    uint64_t blk;            // the bitset word of the cell's block
    // P >> 6 (the byte offset of the block's word) after each step, written
    // by even steps to at0 and by odd steps to at1, so a step compares its
    // block with the one before without a copy; the word in blk is always
    // that of the block the last step reached
    uint32_t at0, at1;
    uint32_t left1;          // iterations left before the step cap, minus one (wraps when the cap is reached)
    uint32_t cap1;           // left1 at the start: the ray has not crossed a face while left1 == cap1
};

// The ray of start record *rec (staged in LDS).  A ray with dx == 0 (cap 0,
// P 0: a hot-wall cell) is loaded like any other: its first step finds the
// wall bit, and with its cap already spent it ends as FIN_FAILED.
// last_x and last_y are not reset: both are written by every crossing, and a
// ray's final state reads them only after it has crossed a face.
//
// Dilated steps: subtracting the axis mask adds one to the axis's value (the
// borrow runs through the gaps, which the mask then clears); subtracting the
// mask's lowest bit takes one away.  With at least two interleaved bits per
// axis the lowest bits are 4 (x), 2 (y) and 1 (z), and every +1 step has a
// bit above bit 2 set.  A step sign is taken from one bit of its code (+1 or
// not): an axis whose direction component is 0 never crosses, so its step
// value is never used.
template <bool POW2>
__device__ __forceinline__ void lean_load(LeanRay &R, const DdaRec *rec, const GeoView &G)
{
    const double2 *p = reinterpret_cast<const double2 *>(rec);
    const double2 a = p[0], b = p[1], c = p[2];
    const uint4 w = reinterpret_cast<const uint4 *>(rec)[3];
    R.tx = a.x;
    R.ty = a.y;
    R.tz = b.x;
    R.incx = b.y;
    R.incy = c.x;
    R.incz = c.y;
    R.hix = (uint32_t)__double2hiint(b.y);
    R.hiy = (uint32_t)__double2hiint(c.x);
    R.hiz = (uint32_t)__double2hiint(c.y);
    const LeanLayout &L = G.wl;
    // the record holds the start cell's index P and the step signs
    R.P = w.w;
    R.xd = w.w & L.mask[0];
    R.yd = w.w & L.mask[1];
    R.zd = w.w & L.mask[2];
    // (mask & (all ones for a +1 step, the high bit of the 2-bit code step + 1))
    // | lowest bit, two ALU operations each (the compiler otherwise turns the
    // expression into a compare and a select)
    asm("v_bfe_i32 %0, %1, 17, 1\n\tv_and_or_b32 %0, %0, %2, 4" : "=&v"(R.sx) : "v"(w.y), "s"(L.mask[0]));
    asm("v_bfe_i32 %0, %1, 19, 1\n\tv_and_or_b32 %0, %0, %2, 2" : "=&v"(R.sy) : "v"(w.y), "s"(L.mask[1]));
    asm("v_bfe_i32 %0, %1, 21, 1\n\tv_and_or_b32 %0, %0, %2, 1" : "=&v"(R.sz) : "v"(w.y), "s"(L.mask[2]));
    if (!POW2) {
        const bool py = (w.y >> 19) & 1u, pz = (w.y >> 21) & 1u;  // +1 steps
        // +1 from n - 1 reaches n, which wraps to 0; -1 from 0 leaves all the
        // axis's bits set, which wraps to n - 1
        R.from_y = py ? lean_dilate(L, 1, (uint32_t)G.ny) : L.mask[1];
        R.to_y = py ? 0u : lean_dilate(L, 1, (uint32_t)G.ny - 1u);
        R.from_z = pz ? lean_dilate(L, 2, (uint32_t)G.nz) : L.mask[2];
        R.to_z = pz ? 0u : lean_dilate(L, 2, (uint32_t)G.nz - 1u);
    }
    R.left1 = w.z - 1u;
    R.cap1 = R.left1;
    R.nt1 = 0.0;  // a ray given with x outside [0, nx) is on a wall before any crossing
    // the block of the first cell, read by the first step (an even step,
    // which compares with at1)
    R.at1 = R.P >> 6;
    R.blk = *reinterpret_cast<const uint64_t *>(reinterpret_cast<const char *>(G.wocc) + R.at1);
}

// One iteration of view_factors.py:225-255 on the walled bitset; returns
// whether the ray goes on.  ODD: the step's parity within the ray.  The
// crossing times are compared and advanced before the cell's bitset word is
// tested (they do not depend on it), so that work overlaps the word's load;
// the cell index moves only for the lanes whose cell is empty.  The word of
// the next cell's block is requested right after the crossing and first
// tested by the next step, so its latency overlaps that step's FP64 work.
template <bool POW2, bool ODD>
__device__ __forceinline__ bool lean_step(LeanRay &R, const GeoView &G)
{
    double &nt = ODD ? R.nt1 : R.nt0;
    nt = vmin_f64(vmin_f64(R.tx, R.ty), R.tz);
    const bool cx = fabs(R.tx - nt) <= TIE_TOLERANCE;
    const bool cy = fabs(R.ty - nt) <= TIE_TOLERANCE;
    const bool cz = fabs(R.tz - nt) <= TIE_TOLERANCE;
    // t += dd for a crossed axis (the reference's addition); for the others
    // t plus a subnormal, which rounds back to t.  Made before the cell's
    // bit is known: a ray that stops here never reads its crossing times
    // again (its final state is formed from the cell, the last crossing and
    // the step's minimum), so this work overlaps the wait for the block word
    // (the compiler would otherwise sink it behind the test).
    R.incx = __hiloint2double(cx ? (int)R.hix : 0, __double2loint(R.incx));
    R.incy = __hiloint2double(cy ? (int)R.hiy : 0, __double2loint(R.incy));
    R.incz = __hiloint2double(cz ? (int)R.hiz : 0, __double2loint(R.incz));
    R.tx = R.tx + R.incx;
    R.ty = R.ty + R.incy;
    R.tz = R.tz + R.incz;
    const uint32_t step_z = cz ? R.sz : 0u;
    asm volatile("" : "+v"(R.tx), "+v"(R.ty), "+v"(R.tz), "+v"(R.incx), "+v"(R.incy), "+v"(R.incz));
    __builtin_amdgcn_sched_barrier(0);
    // the cell's bit, shifted to bit 63: a solid cell or a wall ends the ray here
    if ((int64_t)(R.blk << (R.P & 63u)) < 0) return false;
    R.last_x = cx ? R.sx : 0u;
    R.last_y = cy ? R.sy : 0u;
    R.xd = (R.xd - R.last_x) & G.wl.mask[0];
    R.yd = (R.yd - R.last_y) & G.wl.mask[1];
    R.zd = (R.zd - step_z) & G.wl.mask[2];
    if (!POW2) {
        R.yd = R.yd == R.from_y ? R.to_y : R.yd;
        R.zd = R.zd == R.from_z ? R.to_z : R.zd;
    }
    R.P = R.xd | R.yd | R.zd;
    // byte offset of the next cell's block word (bits 6-8 of P are zero)
    uint32_t &at = ODD ? R.at1 : R.at0;
    at = R.P >> 6;
    if (at != (ODD ? R.at0 : R.at1))
        R.blk = *reinterpret_cast<const uint64_t *>(reinterpret_cast<const char *>(G.wocc) + at);
    uint32_t left1;
    const bool capped = __builtin_usub_overflow(R.left1, 1u, &left1);  // no iteration left
    R.left1 = left1;
    return !capped;
}

// Whether a finished lean ray stands on a wall plane (dilated i + 1 of 0 or nx + 1).
__device__ __forceinline__ bool lean_at_wall(const LeanRay &R, const GeoView &G)
{
    return R.xd == 0u || R.xd == G.wl_cold;
}

// Whether a finished lean ray stopped at its step cap.
__device__ __forceinline__ bool lean_capped(const LeanRay &R) { return R.left1 == 0xFFFFFFFFu; }

// Occupancy reads of a finished lean traversal (as ray_reads): one per
// crossing, plus the read that found a solid cell (a wall is not read, and a
// ray stopped by its cap read before each of its crossings).
__device__ __forceinline__ uint32_t lean_reads(const LeanRay &R, bool wall, bool capped)
{
    return (R.cap1 - R.left1) + (!capped && !wall ? 1u : 0u);
}

// Final state of a finished lean traversal: FIN_FAILED, the crossing time of
// a wall (negated for the cold wall), or FIN_SOLID_P with the cell's index P.
// Formed with selects, so a refill runs one path for every finishing lane.
__device__ __forceinline__ uint2 lean_final(const LeanRay &R, bool wall, bool capped)
{
    // failed: the step cap was reached, or the ray started in a solid cell (no face)
    const bool failed = capped || (!wall && R.left1 == R.cap1);
    // the face of the last crossing: the first axis crossed in x > y > z order;
    // a +1 step has a bit above bit 2 set, a -1 step is 4, 2 or 1
    const uint32_t axis = R.last_x != 0u ? 0u : (R.last_y != 0u ? 1u : 2u);
    const uint32_t step = R.last_x != 0u ? R.last_x : (R.last_y != 0u ? R.last_y : R.sz);
    const uint32_t face = 2u * axis + (step > 7u ? 1u : 0u);
    // the time of the last crossing: the two registers hold it and the
    // smallest crossing time of the step that stopped the ray, which is
    // larger (each step's smallest time exceeds the one before: every axis
    // within the tie tolerance of it crosses and its time grows by its
    // increment, the others are beyond the tolerance); a ray that stopped
    // before any crossing has nt1 = 0
    const double time = vmin_f64(R.nt0, R.nt1);
    const unsigned long long t = (unsigned long long)__double_as_longlong(time);
    const uint32_t x = wall ? (uint32_t)t : R.P;
    const uint32_t y = wall ? ((uint32_t)(t >> 32) | (R.xd == 0u ? 0u : 0x80000000u))
                            : (FIN_BOX | FIN_SOLID_P << 16 | face);
    return make_uint2(failed ? 0u : x, failed ? (FIN_BOX | FIN_FAILED << 16) : y);
}

// Start records staged per wavefront: the wavefront copies up to 64
// consecutive records of its grain (4 KB) from HBM straight into its LDS
// slice with four direct loads of 16 B per lane, and waits for them once;
// refills then read the records from LDS.  The wait on HBM comes once per 64
// records instead of at every refill.
constexpr uint32_t LEAN_STAGE = 64;  // records per wavefront slice

__device__ __forceinline__ void stage_records(DdaRec *slice, const DdaRec *src, uint32_t cnt, uint32_t lane)
{
    const uint4 *from = reinterpret_cast<const uint4 *>(src);
    uint4 *to = reinterpret_cast<uint4 *>(slice);
#pragma unroll
    for (uint32_t m = 0; m < 4; ++m) {
        const uint32_t piece = 64u * m + lane;  // 16 B pieces of the records, in order
        if (piece < 4u * cnt)
            __builtin_amdgcn_global_load_lds(from + piece, (__attribute__((address_space(3))) void *)(to + 64u * m),
                                             16, 0, 0);
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the records are in LDS
}

// Final states leave through LDS: each wavefront holds two buffers of
// LEAN_STAGE entries, one for the staged block of start records that refills
// take from and one for the block before it.  A ray that ends writes its
// final state into its block's buffer (or, when its block is older than
// both, straight to memory); when a new block is staged the older buffer is
// written out, its finished entries by one store instruction of 512 B in
// whole lines, and becomes the new block's buffer.
constexpr uint32_t FIN_EMPTY = 0xFFFFFFFFu;  // high word of an empty entry (no final state has it)

__device__ __forceinline__ void flush_finals(uint2 *buf, uint32_t first, uint32_t n, uint2 *fin, uint32_t lane)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    const uint2 v = buf[lane];
    if (lane < n && v.y != FIN_EMPTY) fin[first + lane] = v;
    buf[lane] = make_uint2(0u, FIN_EMPTY);
}

// trace_body with the lean step: the same queue and final-state order, in one
// loop whose refill part runs only when REFILL_MIN lanes are idle (or none is
// active), so the stepping is not a loop of its own.  Idle lanes take the
// staged records in lane order; the slice is refilled when it runs out.
constexpr int LEAN_STEPS = 6;  // steps between the wavefront's refill checks

template <int REFILL_MIN, bool POW2>
__device__ __forceinline__ void lean_body(const GeoView &G, uint32_t base, uint32_t total, uint32_t grain,
                                          unsigned long long *queue, const DdaRec *dda, uint2 *fin,
                                          unsigned long long *reads_out, DdaRec *slice, uint2 *fbuf)
{
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t cursor = 0, w_end = 0;  // the wavefront's grain [cursor, w_end), beyond the staged records
    bool more = true;                // the queue may still hold records
    uint32_t s_first = 0, s_n = 0, s_taken = 0;  // staged records: [s_first, s_first + s_n), s_taken handed out
    uint32_t p_first = 0, p_n = 0;               // the block staged before them
    uint32_t cur = 0;                            // buffer of the staged block (fbuf + LEAN_STAGE * cur)
    fbuf[lane] = make_uint2(0u, FIN_EMPTY);
    fbuf[LEAN_STAGE + lane] = make_uint2(0u, FIN_EMPTY);

    LeanRay R;
    R.xd = R.yd = R.zd = 0u;
    R.left1 = R.cap1 = 0u;
    R.last_x = R.last_y = 0u;
    R.nt0 = R.nt1 = 0.0;
    bool active = false, pending = false;
    uint32_t g_cur = 0;
    uint64_t reads = 0;
    int n_active = 0;       // lanes of the wavefront that are tracing
    bool finished = false;  // every record of the queue is traced and written
    uint32_t wave_steps = 0, refills = 0;

    do {
        const bool waiting = more || cursor < w_end || s_taken < s_n;  // records not yet handed out
        if (n_active == 0 || (waiting && n_active <= 64 - REFILL_MIN)) {
            ++refills;
            if (!active && pending) {
                const bool wall = lean_at_wall(R, G), capped = lean_capped(R);
                reads += lean_reads(R, wall, capped);
                const uint2 v = lean_final(R, wall, capped);
                const uint32_t dc = g_cur - s_first, dp = g_cur - p_first;
                if (dc < s_n || dp < p_n) {
                    fbuf[dc < s_n ? LEAN_STAGE * cur + dc : LEAN_STAGE * (cur ^ 1u) + dp] = v;
                } else {
                    fin[g_cur] = v;
                    // kept apart from the LDS store above (the empty asm ends
                    // this branch): one store through a selected address would
                    // be a flat store, which the next step's wait for its
                    // block word would wait for too
                    asm volatile("" ::: "memory");
                }
                pending = false;
            }
            for (int round = 0; round < 2; ++round) {
                if (s_taken == s_n) {
                    if (more && cursor >= w_end) {
                        unsigned long long first = 0;
                        if (lane == 0) first = atomicAdd(queue, (unsigned long long)grain);
                        // lane 0's value, held as a wave-uniform (scalar) value
                        const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)first);
                        const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(first >> 32));
                        first = (unsigned long long)hi << 32 | lo;
                        if (first >= total) {
                            more = false;
                        } else {
                            cursor = base + (uint32_t)first;
                            const unsigned long long last = first + grain;
                            w_end = base + (last < total ? (uint32_t)last : total);
                        }
                    }
                    if (cursor < w_end) {
                        const uint32_t cnt = w_end - cursor < LEAN_STAGE ? w_end - cursor : LEAN_STAGE;
                        // the older block's finals go out; its buffer takes the new block's
                        flush_finals(fbuf + LEAN_STAGE * (cur ^ 1u), p_first, p_n, fin, lane);
                        p_first = s_first;
                        p_n = s_n;
                        cur ^= 1u;
                        stage_records(slice, dda + cursor, cnt, lane);
                        s_first = cursor;
                        s_n = cnt;
                        s_taken = 0;
                        cursor += cnt;
                    }
                }
                const uint32_t avail = s_n - s_taken;
                const uint64_t idle = __ballot(!active);
                const uint32_t n_idle = (uint32_t)__popcll(idle);
                if (avail == 0u || n_idle == 0u) break;
                if (!active) {
                    const uint32_t rank = (uint32_t)__popcll(idle & below);
                    if (rank < avail) {
                        const uint32_t e = s_taken + rank;
                        lean_load<POW2>(R, slice + e, G);
                        active = true;
                        pending = true;
                        g_cur = s_first + e;
                    }
                }
                s_taken += n_idle < avail ? n_idle : avail;
            }
            n_active = __popcll(__ballot(active));
            finished = n_active == 0 && !more && cursor >= w_end && s_taken == s_n;
        }
        if (n_active > 0) wave_steps += LEAN_STEPS;
        static_assert(LEAN_STEPS % 2 == 0, "a ray's first step is an even step");
#pragma unroll
        for (int k = 0; k < LEAN_STEPS; k += 2) {
            if (active) active = lean_step<POW2, false>(R, G);
            if (active) active = lean_step<POW2, true>(R, G);
        }
        n_active = __popcll(__ballot(active));
    } while (!finished);
    flush_finals(fbuf + LEAN_STAGE * (cur ^ 1u), p_first, p_n, fin, lane);
    flush_finals(fbuf + LEAN_STAGE * cur, s_first, s_n, fin, lane);
    reads = wave_sum_u64(reads);
    if (lane == 0 && reads) atomicAdd(reads_out, (unsigned long long)reads);
    if (lane == 0) {
        atomicAdd(&g_trace_stats[0], (unsigned long long)wave_steps);
        atomicAdd(&g_trace_stats[1], (unsigned long long)refills);
    }
}

template <int REFILL_MIN, bool POW2>
__global__ __launch_bounds__(STREAM_THREADS) __attribute__((amdgpu_waves_per_eu(8)))
void k_trace_lean(TraceArgs a)
{
    __shared__ DdaRec s_stage[STREAM_THREADS / 64][LEAN_STAGE];
    __shared__ uint2 s_fin[STREAM_THREADS / 64][2 * LEAN_STAGE];
    lean_body<REFILL_MIN, POW2>(a.G, 0u, a.total, a.grain, a.queue, a.dda, a.fin, a.reads,
                                s_stage[threadIdx.x >> 6], s_fin[threadIdx.x >> 6]);
}

template <int REFILL_MIN, bool POW2>
__global__ __launch_bounds__(STREAM_THREADS) __attribute__((amdgpu_waves_per_eu(8)))
void k_trace_batch_lean(const Segment *segs, int n_seg, const DdaRec *dda, const WallRec *walls, uint2 *fin)
{
    __shared__ DdaRec s_stage[STREAM_THREADS / 64][LEAN_STAGE];
    __shared__ uint2 s_fin[STREAM_THREADS / 64][2 * LEAN_STAGE];
    const int k = segment_of(segs, n_seg, blockIdx.x, true);
    const GeoView G = segs[k].gen.G;
    lean_body<REFILL_MIN, POW2>(G, segs[k].gen.offset, segs[k].gen.total, segs[k].grain, segs[k].queue, dda, fin,
                                segs[k].reads, s_stage[threadIdx.x >> 6], s_fin[threadIdx.x >> 6]);
}

__global__ void k_zero_u64(unsigned long long *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// Receivers of final states (prh_trace_rays), with the count of failures.
__global__ __launch_bounds__(TRACE_THREADS) void k_resolve(const uint2 *fin, const WallRec *walls, GeoView G,
                                                           uint32_t total, int32_t *receiver,
                                                           unsigned long long *failed)
{
    const uint32_t g = blockIdx.x * TRACE_THREADS + threadIdx.x;
    int rec = 0;
    if (g < total) {
        rec = resolve_receiver(fin[g], G, walls[g]);
        receiver[g] = rec;
    }
    const uint64_t f = wave_sum_u64(g < total && rec < 0 ? 1ull : 0ull);
    if ((threadIdx.x & 63u) == 0 && f) atomicAdd(failed, (unsigned long long)f);
}

// Per-patch sorted unique receivers and their counts (the np.unique of
// first_hit_counts, view_factors.py:284-286, pooled over seeds as
// pool_counts does, view_factors.py:290-300).  One workgroup per patch: the
// receivers of the patch's S*R final states are resolved on load (failures
// counted), then bitonic-sorted in LDS.
__global__ __launch_bounds__(COUNT_THREADS) void k_count_patches(const uint2 *fin, const WallRec *walls, GeoView G,
                                                                 int64_t n_p, int R, int S,
                                                                 int SRp2, int32_t *uniq_recv, int32_t *uniq_cnt,
                                                                 int32_t *n_unique, unsigned long long *failed)
{
    extern __shared__ int32_t lds[];
    int32_t *key = lds;
    int32_t *part = lds + SRp2;
    const int SR = S * R;
    const int chunk = SRp2 >= COUNT_THREADS ? SRp2 / COUNT_THREADS : 1;
    for (int64_t q = blockIdx.x; q < n_p; q += gridDim.x) {
        uint32_t bad = 0;
        for (int t = threadIdx.x; t < SRp2; t += COUNT_THREADS) {
            int32_t v = INT_MAX;
            if (t < SR) {
                const int s = t / R, r = t - s * R;
                const int64_t at = (int64_t)s * n_p * R + q * R + r;
                v = resolve_receiver(fin[at], G, walls[at]);
                if (v < 0) {
                    ++bad;
                    v = INT_MAX;
                }
            }
            key[t] = v;
        }
        if (bad) atomicAdd(failed, (unsigned long long)bad);
        __syncthreads();
        for (int k = 2; k <= SRp2; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int t = threadIdx.x; t < SRp2; t += COUNT_THREADS) {
                    const int u = t ^ j;
                    if (u > t) {
                        const int32_t x = key[t], y = key[u];
                        const bool up = (t & k) == 0;
                        if ((x > y) == up) {
                            key[t] = y;
                            key[u] = x;
                        }
                    }
                }
                __syncthreads();
            }
        }
        // Run starts per thread chunk, block scan for ranks.
        const int b0 = threadIdx.x * chunk;
        int starts = 0;
        if (b0 < SRp2) {
            for (int e = b0; e < b0 + chunk && e < SR; ++e) starts += (e == 0 || key[e] != key[e - 1]) ? 1 : 0;
        }
        part[threadIdx.x] = starts;
        __syncthreads();
        for (int off = 1; off < COUNT_THREADS; off <<= 1) {
            const int add = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
            __syncthreads();
            part[threadIdx.x] += add;
            __syncthreads();
        }
        int rank = part[threadIdx.x] - starts;
        const int total = part[COUNT_THREADS - 1];
        if (b0 < SRp2) {
            for (int e = b0; e < b0 + chunk && e < SR; ++e) {
                if (e == 0 || key[e] != key[e - 1]) {
                    int end = e + 1;
                    while (end < SR && key[end] == key[e]) ++end;
                    uniq_recv[q * SR + rank] = key[e];
                    uniq_cnt[q * SR + rank] = end - e;
                    ++rank;
                }
            }
        }
        if (threadIdx.x == 0) n_unique[q] = total;
        __syncthreads();
    }
}

// The same for S*R <= 1024, one wavefront per patch with the keys in
// registers: element e = lane * EPL + r (EPL = keys per lane, a power of 2),
// so bitonic stages with partner distance j < EPL are register swaps and
// the others one __shfl_xor per key.  Runs are then found with one
// neighbour shuffle and a wave prefix sum; run starts go through LDS to
// give each run its length.
// Value of lane (lane ^ M) for a compile-time M: DPP within a quad, a
// swizzle within 32 lanes, a permute across the halves.
template <int M>
__device__ __forceinline__ int32_t xor_lanes(int32_t v)
{
    if constexpr (M == 1) return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);  // quad_perm 1,0,3,2
    if constexpr (M == 2) return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);  // quad_perm 2,3,0,1
    if constexpr (M == 3) return __builtin_amdgcn_mov_dpp(v, 0x1B, 0xF, 0xF, false);  // quad_perm 3,2,1,0
    if constexpr (M < 32) return __builtin_amdgcn_ds_swizzle(v, 0x1F | M << 10);     // and 31, xor M
    return __shfl_xor(v, M);
}

// The middle of a, b and c: with c = INT_MIN the smaller of a and b, with
// c = INT_MAX the larger, so one instruction keeps either side of an
// exchange.
__device__ __forceinline__ int32_t med3_i32(int32_t a, int32_t b, int32_t c)
{
    int32_t r;
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// One cross-lane compare-exchange of register r with lane (lane ^ M) at
// register r ^ FLIP: the lane whose bit M is clear keeps the smaller key.
template <int M, int FLIP, int EPL>
__device__ __forceinline__ void exchange_lanes(int32_t (&key)[EPL], bool lower)
{
    const int32_t keep = lower ? INT_MIN : INT_MAX;  // med3 with it keeps this lane's side
    if constexpr (FLIP == 0) {
#pragma unroll
        for (int r = 0; r < EPL; ++r) key[r] = med3_i32(key[r], xor_lanes<M>(key[r]), keep);
    } else {
#pragma unroll
        for (int r = 0; r < EPL / 2; ++r) {
            const int32_t a = xor_lanes<M>(key[r ^ FLIP]), b = xor_lanes<M>(key[r]);
            key[r] = med3_i32(key[r], a, keep);
            key[r ^ FLIP] = med3_i32(key[r ^ FLIP], b, keep);
        }
    }
}

// Half-cleaners J, J / 2, ..., 1 of the sort below: position i against i + J
// in each block of 2J, the smaller key to the lower position.
template <int J, int EPL>
__device__ __forceinline__ void sort_clean(int32_t (&key)[EPL], int lane)
{
    if constexpr (J >= EPL) {
        exchange_lanes<J / EPL, 0>(key, (lane & (J / EPL)) == 0);
    } else {
#pragma unroll
        for (int r = 0; r < EPL; ++r)
            if ((r & J) == 0) {
                const int32_t a = key[r], b = key[r | J];
                key[r] = a < b ? a : b;
                key[r | J] = a < b ? b : a;
            }
    }
    if constexpr (J > 1) sort_clean<J / 2, EPL>(key, lane);
}

// Ascending sort of the 64 * EPL keys of a wavefront, key r of lane l at
// position l * EPL + r: a bitonic network in the form whose merges start by
// comparing mirrored positions (i with k - 1 - i in each block of k), so
// every compare-exchange puts the smaller key at the lower position and no
// direction is selected per block.
template <int K, int EPL>
__device__ __forceinline__ void sort_merge_level(int32_t (&key)[EPL], int lane)
{
    // the mirrored compare of blocks of K
    if constexpr (K <= EPL) {
#pragma unroll
        for (int r = 0; r < EPL; ++r)
            if ((r & (K / 2)) == 0) {
                const int q = r ^ (K - 1);
                const int32_t a = key[r], b = key[q];
                key[r] = a < b ? a : b;
                key[q] = a < b ? b : a;
            }
    } else {
        constexpr int M = (K - 1) / EPL;
        exchange_lanes<M, EPL - 1>(key, (lane & (K / 2 / EPL)) == 0);
    }
    // then half-cleaners of blocks of J * 2, J = K / 4 ... 1
    if constexpr (K >= 4) sort_clean<K / 4, EPL>(key, lane);
}

template <int EPL>
__device__ __forceinline__ void sort_keys(int32_t (&key)[EPL], int lane)
{
    sort_merge_level<2, EPL>(key, lane);
    if constexpr (64 * EPL >= 4) sort_merge_level<4, EPL>(key, lane);
    if constexpr (64 * EPL >= 8) sort_merge_level<8, EPL>(key, lane);
    if constexpr (64 * EPL >= 16) sort_merge_level<16, EPL>(key, lane);
    if constexpr (64 * EPL >= 32) sort_merge_level<32, EPL>(key, lane);
    if constexpr (64 * EPL >= 64) sort_merge_level<64, EPL>(key, lane);
    if constexpr (64 * EPL >= 128) sort_merge_level<128, EPL>(key, lane);
    if constexpr (64 * EPL >= 256) sort_merge_level<256, EPL>(key, lane);
    if constexpr (64 * EPL >= 512) sort_merge_level<512, EPL>(key, lane);
    if constexpr (64 * EPL >= 1024) sort_merge_level<1024, EPL>(key, lane);
    if constexpr (64 * EPL >= 2048) sort_merge_level<2048, EPL>(key, lane);
    static_assert(64 * EPL <= 2048, "at most 2048 keys per wavefront");
}

template <int EPL, bool ONE_SEED>
__device__ __forceinline__ void count_wave_body(const uint2 *fin, const WallRec *walls, GeoView G,
                                                              int64_t n_p, int R, int S,
                                                              int32_t *uniq_recv, int32_t *uniq_cnt,
                                                              int32_t *n_unique, unsigned long long *failed)
{
    constexpr int N = 64 * EPL;
    constexpr int WAVES = COUNT_THREADS / 64;
    __shared__ int32_t s_pos[WAVES][N + 1];
    __shared__ int32_t s_key[WAVES][N];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    int32_t *pos = s_pos[wave];
    int32_t *ukey = s_key[wave];
    const int SR0 = S * R;
    const uint64_t below = (1ull << lane) - 1ull;
    for (int64_t q = (int64_t)blockIdx.x * WAVES + wave; q < n_p; q += (int64_t)gridDim.x * WAVES) {
        // S * R, opaque to the compiler in each iteration: the per-key
        // comparisons with it are then made where they are used instead of
        // being hoisted out of the loop as lane masks that do not fit the
        // scalar registers
        int SR = SR0;
        asm volatile("" : "+s"(SR));
        int32_t key[EPL];
        uint32_t bad = 0;
        uint64_t timed = 0;  // bit r: key r is a wall crossing time, resolved below
        // key r of a lane is the patch's ray lane + 64 r, so each load
        // instruction reads 64 consecutive final states (the sort below does
        // not depend on where a key starts).  All final states are loaded
        // first, then the walled-bitset cell codes of the FIN_SOLID_P ones
        // (entry 0 for the others), so the loads of a patch are in flight
        // together; then the receivers are worked out.
        // (in batches of at most 16 keys, which bounds the registers)
        constexpr int BATCH = EPL < 16 ? EPL : 16;
        // key t = lane + 64 r is ray u of seed s (t = s R + u); with several
        // seeds (s, u) advance by 64 keys at a time instead of a division per key
        int ks = 0, ku = lane;
        if (!ONE_SEED)
            while (ku >= R) {
                ku -= R;
                ++ks;
            }
#pragma unroll
        for (int b0 = 0; b0 < EPL; b0 += BATCH) {
            uint2 f[BATCH];
#pragma unroll
            for (int i = 0; i < BATCH; ++i) {
                const int t = lane + 64 * (b0 + i);
                f[i] = make_uint2(0u, FIN_BOX | FIN_FAILED << 16);
                if (t < SR) f[i] = fin[(int64_t)ks * n_p * R + q * R + (ONE_SEED ? t : ku)];
                if (!ONE_SEED) {
                    ku += 64;
                    while (ku >= R) {
                        ku -= R;
                        ++ks;
                    }
                }
            }
            uint32_t code[BATCH];
            if (G.wcode) {
#pragma unroll
                for (int i = 0; i < BATCH; ++i) code[i] = G.wcode[fin_wcode_index(f[i], G)];
            }
#pragma unroll
            for (int i = 0; i < BATCH; ++i) {
                const int r = b0 + i, t = lane + 64 * r;
                int32_t v = INT_MAX;
                if (t < SR) {
                    if (fin_is_time(f[i])) {
                        timed |= 1ull << r;
                    } else {
                        v = fin_kind(f[i]) == FIN_SOLID_P && G.wcode ? resolve_solid_p(f[i], code[i], G)
                                                                       : resolve_boxed(f[i], G);
                        if (v < 0) {
                            ++bad;
                            v = INT_MAX;
                        }
                    }
                }
                key[r] = v;
            }
        }
        // Wall crossing times: listed in LDS and resolved one per lane, so the
        // wall arithmetic and its record loads run once per 64 such rays.
        if (__ballot(timed != 0ull)) {
            int32_t *list = pos, *res = ukey;
            int n = 0;
#pragma unroll
            for (int r = 0; r < EPL; ++r) {
                const bool w = (timed >> r) & 1ull;
                const uint64_t m = __ballot(w);
                if (w) list[n + __popcll(m & below)] = lane + 64 * r;
                n += __popcll(m);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            for (int i = lane; i < n; i += 64) {
                const int t = list[i];
                const int s = ONE_SEED ? 0 : t / R, u = t - s * R;
                const int64_t at = (int64_t)s * n_p * R + q * R + u;
                res[i] = resolve_wall_time(fin[at], G, walls[at]);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            n = 0;
#pragma unroll
            for (int r = 0; r < EPL; ++r) {
                const bool w = (timed >> r) & 1ull;
                const uint64_t m = __ballot(w);
                if (w) {
                    int32_t v = res[n + __popcll(m & below)];
                    if (v < 0) {
                        ++bad;
                        v = INT_MAX;
                    }
                    key[r] = v;
                }
                n += __popcll(m);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        }
        bad = (uint32_t)wave_sum_u64(bad);
        if (lane == 0 && bad) atomicAdd(failed, (unsigned long long)bad);
        sort_keys<EPL>(key, lane);
        // run starts among the first SR elements
        const int32_t prev_last = __shfl_up(key[EPL - 1], 1);
        int starts = 0;
#pragma unroll
        for (int r = 0; r < EPL; ++r) {
            const int e = lane * EPL + r;
            const int32_t prev = r == 0 ? prev_last : key[r - 1];
            starts += (e < SR && (e == 0 || key[r] != prev)) ? 1 : 0;
        }
        int incl = starts;
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(incl, off);
            if (lane >= off) incl += v;
        }
        const int total = __shfl(incl, 63);
        int rank = incl - starts;
#pragma unroll
        for (int r = 0; r < EPL; ++r) {
            const int e = lane * EPL + r;
            const int32_t prev = r == 0 ? prev_last : key[r - 1];
            if (e < SR && (e == 0 || key[r] != prev)) {
                ukey[rank] = key[r];
                pos[rank] = e;
                ++rank;
            }
        }
        if (lane == 0) {
            pos[total] = SR;
            n_unique[q] = total;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        // receivers and counts leave LDS in consecutive runs of 64
        for (int u = lane; u < total; u += 64) {
            uniq_recv[q * SR + u] = ukey[u];
            uniq_cnt[q * SR + u] = pos[u + 1] - pos[u];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
}

// One-seed calls (the common case) hold the sort at four waves per SIMD;
// several seeds keep the compiler's register count.
template <int EPL>
__global__ __launch_bounds__(COUNT_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_count_wave(
    const uint2 *fin, const WallRec *walls, GeoView G, int64_t n_p, int R, int S, int32_t *uniq_recv,
    int32_t *uniq_cnt, int32_t *n_unique, unsigned long long *failed)
{
    count_wave_body<EPL, true>(fin, walls, G, n_p, R, S, uniq_recv, uniq_cnt, n_unique, failed);
}

// 32 keys per lane hold about 200 VGPRs (two waves per SIMD) either way.
template <int EPL>
__global__ __launch_bounds__(COUNT_THREADS) void k_count_wave_wide(const uint2 *fin, const WallRec *walls, GeoView G,
                                                                   int64_t n_p, int R, int S, int32_t *uniq_recv,
                                                                   int32_t *uniq_cnt, int32_t *n_unique,
                                                                   unsigned long long *failed)
{
    count_wave_body<EPL, true>(fin, walls, G, n_p, R, S, uniq_recv, uniq_cnt, n_unique, failed);
}

template <int EPL>
__global__ __launch_bounds__(COUNT_THREADS) void k_count_wave_seeds(const uint2 *fin, const WallRec *walls, GeoView G,
                                                                    int64_t n_p, int R, int S, int32_t *uniq_recv,
                                                                    int32_t *uniq_cnt, int32_t *n_unique,
                                                                    unsigned long long *failed)
{
    count_wave_body<EPL, false>(fin, walls, G, n_p, R, S, uniq_recv, uniq_cnt, n_unique, failed);
}

// The k_count_wave instance for EPL keys per lane and S seeds per call.
using CountWaveKernel = void (*)(const uint2 *, const WallRec *, GeoView, int64_t, int, int, int32_t *, int32_t *,
                                 int32_t *, unsigned long long *);
CountWaveKernel count_wave_kernel(int epl, int S)
{
    if (S == 1)
        return epl == 1 ? k_count_wave<1>
             : epl == 2 ? k_count_wave<2>
             : epl == 4 ? k_count_wave<4>
             : epl == 8 ? k_count_wave<8>
             : epl == 16 ? k_count_wave<16>
                        : k_count_wave_wide<32>;
    return epl == 1 ? k_count_wave_seeds<1>
         : epl == 2 ? k_count_wave_seeds<2>
         : epl == 4 ? k_count_wave_seeds<4>
         : epl == 8 ? k_count_wave_seeds<8>
         : epl == 16 ? k_count_wave_seeds<16>
                    : k_count_wave_seeds<32>;
}

// Compaction by output ranges, for a pass whose edge total is known: block
// r writes edges [r * COMPACT_RANGE, (r + 1) * COMPACT_RANGE) of the table,
// so every store instruction of the three arrays writes 1 KB in whole lines.
// range_q[r] is the patch holding the range's first edge (k_range_starts);
// the range's later edges reach at most a few patches further.
constexpr int COMPACT_RANGE = 1024;

__global__ __launch_bounds__(256) void k_range_starts(const int64_t *offsets, int64_t n_p, int64_t nnz,
                                                      int32_t *range_q)
{
    for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < n_p; q += (int64_t)gridDim.x * 256) {
        const int64_t lo = offsets[q], hi = q + 1 < n_p ? offsets[q + 1] : nnz;
        for (int64_t r = (lo + COMPACT_RANGE - 1) / COMPACT_RANGE; r * COMPACT_RANGE < hi; ++r) range_q[r] = (int32_t)q;
    }
}

__global__ __launch_bounds__(COUNT_THREADS) void k_compact_ranges(const int32_t *uniq_recv, const int32_t *uniq_cnt,
                                                                   const int64_t *offsets, const int32_t *range_q,
                                                                   int64_t n_p, int SR, int64_t p0, int64_t nnz,
                                                                   int32_t *origin, int32_t *receiver, int32_t *count)
{
    static_assert(COMPACT_RANGE % COUNT_THREADS == 0, "whole rounds per range");
    for (int64_t r = blockIdx.x; r * COMPACT_RANGE < nnz; r += gridDim.x) {
        int64_t q = range_q[r];
        int64_t base = offsets[q], next = q + 1 < n_p ? offsets[q + 1] : nnz;
#pragma unroll
        for (int i = 0; i < COMPACT_RANGE / COUNT_THREADS; ++i) {
            const int64_t e = r * COMPACT_RANGE + i * COUNT_THREADS + threadIdx.x;
            if (e >= nnz) break;
            while (e >= next) {  // the next patch with edges
                ++q;
                base = next;
                next = q + 1 < n_p ? offsets[q + 1] : nnz;
            }
            const int64_t at = q * SR + (e - base);
            origin[e] = (int32_t)(p0 + q);
            receiver[e] = uniq_recv[at];
            count[e] = uniq_cnt[at];
        }
    }
}

__global__ __launch_bounds__(COUNT_THREADS) void k_compact_edges(const int32_t *uniq_recv, const int32_t *uniq_cnt,
                                                                 const int32_t *n_unique, const int64_t *offsets,
                                                                 int64_t n_p, int SR, int64_t p0, int32_t *origin,
                                                                 int32_t *receiver, int32_t *count, int64_t cap)
{
    for (int64_t q = blockIdx.x; q < n_p; q += gridDim.x) {
        const int64_t off = offsets[q];
        const int nu = n_unique[q];
        // edges past the capacity are dropped; the caller reports the overflow
        const int64_t room = cap - off;
        const int n_write = room <= 0 ? 0 : (room < nu ? (int)room : nu);
        for (int u = threadIdx.x; u < n_write; u += COUNT_THREADS) {
            origin[off + u] = (int32_t)(p0 + q);
            receiver[off + u] = uniq_recv[q * SR + u];
            count[off + u] = uniq_cnt[q * SR + u];
        }
    }
}

// ------------------------------------------------------------------ host side
bool g_tables_ready[64] = {};

int ensure_tables()
{
    int dev = 0;
    PRH_HIP(hipGetDevice(&dev));
    if (dev >= 0 && dev < 64 && g_tables_ready[dev]) return PRH_OK;
    JumpTables t;
    host_jump_tables(&t);
    PRH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_jump), &t, sizeof(t)));
    if (dev >= 0 && dev < 64) g_tables_ready[dev] = true;
    return PRH_OK;
}

struct Buf {
    void *p = nullptr;
    size_t cap = 0;
    template <typename T>
    T *as() const { return static_cast<T *>(p); }
};

// Grow-only per-device workspaces of the tracer (one library stream per
// device, so calls on one device are serialised).
struct Workspace {
    Buf seeds, origins, directions, dda, walls, fin, receiver, uniq_recv, uniq_cnt, n_unique, offsets, counters;
    hipEvent_t ev[5] = {};
    bool events = false;
};

Workspace g_ws[64];
double g_last_ms[4] = {0, 0, 0, 0};

int grow(Buf &b, size_t bytes) { return ensure_capacity(&b.p, &b.cap, bytes); }

// Device bytes of the tracer's workspace per ray of one pass (start record,
// wall record, final state, unique receivers and counts) and per patch
// (unique count and edge offset); prh_trace_rays holds origins, directions,
// records, final states and receivers per ray.
constexpr int64_t PASS_BYTES_PER_RAY = sizeof(DdaRec) + sizeof(WallRec) + sizeof(uint2) + 2 * sizeof(int32_t);
constexpr int64_t PASS_BYTES_PER_PATCH = sizeof(int32_t) + sizeof(int64_t);
constexpr int64_t GIVEN_BYTES_PER_RAY = 6 * sizeof(double) + sizeof(DdaRec) + sizeof(WallRec) + sizeof(uint2) +
                                        sizeof(int32_t);

// Upper bound of the workspace (prh_tracer_set_workspace_limit); a call
// also keeps it within half of the device memory that is free or already
// held by the workspace.
int64_t g_ws_limit = int64_t(32) << 30;

int64_t workspace_held(Workspace &W)
{
    int64_t held = 0;
    for (Buf *b : {&W.seeds, &W.origins, &W.directions, &W.dda, &W.walls, &W.fin, &W.receiver, &W.uniq_recv,
                   &W.uniq_cnt, &W.n_unique, &W.offsets, &W.counters})
        held += (int64_t)b->cap;
    return held;
}

int workspace_free(Workspace &W)
{
    for (Buf *b : {&W.seeds, &W.origins, &W.directions, &W.dda, &W.walls, &W.fin, &W.receiver, &W.uniq_recv,
                   &W.uniq_cnt, &W.n_unique, &W.offsets, &W.counters}) {
        if (b->p) PRH_HIP(hipFree(b->p));
        b->p = nullptr;
        b->cap = 0;
    }
    return PRH_OK;
}

// Edge tables of the counting hold their three arrays in one allocation.
// A destroyed table's block is kept, one per device, for the next table that
// fits in it (at most twice its size), so a traced step neither allocates
// nor frees device memory; prh_tracer_trim releases it and
// prh_tracer_workspace counts it as held.
struct EdgeBlockCache {
    std::mutex m;
    void *p = nullptr;
    size_t bytes = 0;
};
EdgeBlockCache g_edge_cache[64];

constexpr size_t edge_array_bytes(int64_t nnz) { return ((size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t) + 255) & ~size_t(255); }

// A new table of nnz edges (arrays uninitialised) on device dev, current.
int new_edge_table(int dev, int64_t nnz, prh_edges **out)
{
    const size_t a = edge_array_bytes(nnz), need = 3 * a;
    void *p = nullptr;
    size_t bytes = 0;
    {
        EdgeBlockCache &c = g_edge_cache[dev];
        std::lock_guard<std::mutex> lock(c.m);
        if (c.p && c.bytes >= need && c.bytes <= 2 * need) {
            p = c.p;
            bytes = c.bytes;
            c.p = nullptr;
            c.bytes = 0;
        }
    }
    if (!p) {
        if (hipMalloc(&p, need) != hipSuccess) {
            (void)hipGetLastError();
            // the kept block may be what stands in the way
            {
                EdgeBlockCache &c = g_edge_cache[dev];
                std::lock_guard<std::mutex> lock(c.m);
                if (c.p) (void)hipFree(c.p);
                c.p = nullptr;
                c.bytes = 0;
            }
            if (hipMalloc(&p, need) != hipSuccess) {
                (void)hipGetLastError();
                set_error("prh_trace_counts: out of device memory for %lld edges", (long long)nnz);
                return PRH_ERR_HIP;
            }
        }
        bytes = need;
    }
    char *b = static_cast<char *>(p);
    *out = new prh_edges{dev, nnz, reinterpret_cast<int32_t *>(b), reinterpret_cast<int32_t *>(b + a),
                         reinterpret_cast<int32_t *>(b + 2 * a), p, bytes};
    return PRH_OK;
}

// Keeps or frees the block of a destroyed table (device dev current).
void release_edge_block(int dev, void *p, size_t bytes)
{
    // work that may still read the table (on any stream) ends first, as
    // hipFree would wait for it
    (void)hipDeviceSynchronize();
    if (dev < 0 || dev >= 64) {
        (void)hipFree(p);
        return;
    }
    EdgeBlockCache &c = g_edge_cache[dev];
    std::lock_guard<std::mutex> lock(c.m);
    // kept only within the workspace bound, with the workspace
    if ((c.p && c.bytes >= bytes) || (int64_t)bytes + workspace_held(g_ws[dev]) > g_ws_limit) {
        (void)hipFree(p);
        return;
    }
    if (c.p) (void)hipFree(c.p);
    c.p = p;
    c.bytes = bytes;
}

int64_t edge_cache_held(int dev)
{
    EdgeBlockCache &c = g_edge_cache[dev];
    std::lock_guard<std::mutex> lock(c.m);
    return (int64_t)c.bytes;
}

void edge_cache_free(int dev)
{
    EdgeBlockCache &c = g_edge_cache[dev];
    std::lock_guard<std::mutex> lock(c.m);
    if (c.p) (void)hipFree(c.p);
    c.p = nullptr;
    c.bytes = 0;
}

// Workspace bytes a call may use on the current device.
int workspace_budget(Workspace &W, int64_t *budget)
{
    size_t free_b = 0, total_b = 0;
    PRH_HIP(hipMemGetInfo(&free_b, &total_b));
    const int64_t avail = ((int64_t)free_b + workspace_held(W)) / 2;
    *budget = std::min(g_ws_limit, avail);
    return PRH_OK;
}

int resident_waves(const void *kernel, int threads, size_t lds, int64_t *waves)
{
    int dev = 0, cus = 0, per_cu = 0;
    PRH_HIP(hipGetDevice(&dev));
    PRH_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    PRH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds));
    if (per_cu < 1) per_cu = 1;
    *waves = (int64_t)cus * per_cu * (threads / 64);
    return PRH_OK;
}

// Traversal kernel variants: refill threshold x occupancy test mode
// (OccMode).  The variant is a tuning knob (prh_tracer_set_variant, chosen
// by porerad/autotune.py); every variant produces the same receivers.
using TraceKernel = void (*)(TraceArgs);
struct TraceVariant {
    TraceKernel fn;
    int refill_min, occ;
};
const TraceVariant TRACE_VARIANTS[] = {
    {k_trace_stream<32, OCC_BRICK>, 32, OCC_BRICK},   {k_trace_stream<16, OCC_BRICK>, 16, OCC_BRICK},
    {k_trace_stream<24, OCC_BRICK>, 24, OCC_BRICK},   {k_trace_stream<40, OCC_BRICK>, 40, OCC_BRICK},
    {k_trace_stream<48, OCC_BRICK>, 48, OCC_BRICK},   {k_trace_stream<32, OCC_GLOBAL>, 32, OCC_GLOBAL},
    {k_trace_stream<16, OCC_GLOBAL>, 16, OCC_GLOBAL}, {k_trace_stream<24, OCC_GLOBAL>, 24, OCC_GLOBAL},
    {k_trace_stream<16, OCC_AHEAD>, 16, OCC_AHEAD},   {k_trace_stream<24, OCC_AHEAD>, 24, OCC_AHEAD},
    {k_trace_stream<32, OCC_AHEAD>, 32, OCC_AHEAD},   {k_trace_lean<16, false>, 16, OCC_LEAN},
    {k_trace_lean<24, false>, 24, OCC_LEAN},          {k_trace_lean<32, false>, 32, OCC_LEAN},
    {k_trace_lean<8, false>, 8, OCC_LEAN},            {k_trace_lean<12, false>, 12, OCC_LEAN},
};
constexpr int N_TRACE_VARIANTS = sizeof(TRACE_VARIANTS) / sizeof(TRACE_VARIANTS[0]);
// The same variants for geometries whose ny and nz are powers of two.
const TraceVariant TRACE_VARIANTS_POW2[] = {
    {k_trace_stream<32, OCC_BRICK | OCC_POW2>, 32, OCC_BRICK},
    {k_trace_stream<16, OCC_BRICK | OCC_POW2>, 16, OCC_BRICK},
    {k_trace_stream<24, OCC_BRICK | OCC_POW2>, 24, OCC_BRICK},
    {k_trace_stream<40, OCC_BRICK | OCC_POW2>, 40, OCC_BRICK},
    {k_trace_stream<48, OCC_BRICK | OCC_POW2>, 48, OCC_BRICK},
    {k_trace_stream<32, OCC_GLOBAL | OCC_POW2>, 32, OCC_GLOBAL},
    {k_trace_stream<16, OCC_GLOBAL | OCC_POW2>, 16, OCC_GLOBAL},
    {k_trace_stream<24, OCC_GLOBAL | OCC_POW2>, 24, OCC_GLOBAL},
    {k_trace_stream<16, OCC_AHEAD | OCC_POW2>, 16, OCC_AHEAD},
    {k_trace_stream<24, OCC_AHEAD | OCC_POW2>, 24, OCC_AHEAD},
    {k_trace_stream<32, OCC_AHEAD | OCC_POW2>, 32, OCC_AHEAD},
    {k_trace_lean<16, true>, 16, OCC_LEAN},
    {k_trace_lean<24, true>, 24, OCC_LEAN},
    {k_trace_lean<32, true>, 32, OCC_LEAN},
    {k_trace_lean<8, true>, 8, OCC_LEAN},
    {k_trace_lean<12, true>, 12, OCC_LEAN},
};
static_assert(sizeof(TRACE_VARIANTS_POW2) == sizeof(TRACE_VARIANTS), "one power-of-two kernel per variant");

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// The traversal kernel of a variant for geometry g.
// A lean variant needs the walled bitset; without it (an index wider than 31
// bits, or ny or nz below 3) the brick variant 1 runs instead.
constexpr int BRICK_VARIANT = 1;

int usable_variant(const Geometry &g, int variant)
{
    return TRACE_VARIANTS[variant].occ == OCC_LEAN && !g.wocc ? BRICK_VARIANT : variant;
}

const TraceVariant &trace_variant_for(const Geometry &g, int variant)
{
    variant = usable_variant(g, variant);
    return is_pow2(g.ny) && is_pow2(g.nz) ? TRACE_VARIANTS_POW2[variant] : TRACE_VARIANTS[variant];
}
// Default: the lean step refilling at 12 idle lanes, the fastest on the foam
// measured so far (traversal 15.8 ms against 20.5 ms for the brick test).
int g_trace_variant = 15;

// Batch traversal kernels, one per entry of TRACE_VARIANTS (same index).
using BatchKernel = void (*)(const Segment *, int, const DdaRec *, const WallRec *, uint2 *);
struct BatchVariant {
    BatchKernel fn;
    int refill_min, occ;
};
const BatchVariant BATCH_VARIANTS[] = {
    {k_trace_batch<32, OCC_BRICK>, 32, OCC_BRICK},   {k_trace_batch<16, OCC_BRICK>, 16, OCC_BRICK},
    {k_trace_batch<24, OCC_BRICK>, 24, OCC_BRICK},   {k_trace_batch<40, OCC_BRICK>, 40, OCC_BRICK},
    {k_trace_batch<48, OCC_BRICK>, 48, OCC_BRICK},   {k_trace_batch<32, OCC_GLOBAL>, 32, OCC_GLOBAL},
    {k_trace_batch<16, OCC_GLOBAL>, 16, OCC_GLOBAL}, {k_trace_batch<24, OCC_GLOBAL>, 24, OCC_GLOBAL},
    {k_trace_batch<16, OCC_AHEAD>, 16, OCC_AHEAD},   {k_trace_batch<24, OCC_AHEAD>, 24, OCC_AHEAD},
    {k_trace_batch<32, OCC_AHEAD>, 32, OCC_AHEAD},   {k_trace_batch_lean<16, false>, 16, OCC_LEAN},
    {k_trace_batch_lean<24, false>, 24, OCC_LEAN},   {k_trace_batch_lean<32, false>, 32, OCC_LEAN},
    {k_trace_batch_lean<8, false>, 8, OCC_LEAN},     {k_trace_batch_lean<12, false>, 12, OCC_LEAN},
};
static_assert(sizeof(BATCH_VARIANTS) / sizeof(BATCH_VARIANTS[0]) == N_TRACE_VARIANTS,
              "one batch kernel per traversal variant");
const BatchVariant BATCH_VARIANTS_POW2[] = {
    {k_trace_batch<32, OCC_BRICK | OCC_POW2>, 32, OCC_BRICK},
    {k_trace_batch<16, OCC_BRICK | OCC_POW2>, 16, OCC_BRICK},
    {k_trace_batch<24, OCC_BRICK | OCC_POW2>, 24, OCC_BRICK},
    {k_trace_batch<40, OCC_BRICK | OCC_POW2>, 40, OCC_BRICK},
    {k_trace_batch<48, OCC_BRICK | OCC_POW2>, 48, OCC_BRICK},
    {k_trace_batch<32, OCC_GLOBAL | OCC_POW2>, 32, OCC_GLOBAL},
    {k_trace_batch<16, OCC_GLOBAL | OCC_POW2>, 16, OCC_GLOBAL},
    {k_trace_batch<24, OCC_GLOBAL | OCC_POW2>, 24, OCC_GLOBAL},
    {k_trace_batch<16, OCC_AHEAD | OCC_POW2>, 16, OCC_AHEAD},
    {k_trace_batch<24, OCC_AHEAD | OCC_POW2>, 24, OCC_AHEAD},
    {k_trace_batch<32, OCC_AHEAD | OCC_POW2>, 32, OCC_AHEAD},
    {k_trace_batch_lean<16, true>, 16, OCC_LEAN},
    {k_trace_batch_lean<24, true>, 24, OCC_LEAN},
    {k_trace_batch_lean<32, true>, 32, OCC_LEAN},
    {k_trace_batch_lean<8, true>, 8, OCC_LEAN},
    {k_trace_batch_lean<12, true>, 12, OCC_LEAN},
};

// pow2: every geometry of the launch has power-of-two ny and nz.
const BatchVariant &batch_variant_for(int variant, bool pow2 = false)
{
    return pow2 ? BATCH_VARIANTS_POW2[variant] : BATCH_VARIANTS[variant];
}

// Records per queue grab of the traversal (prh_tracer_set_grain).
int g_trace_grain = 512;

// Traversal workgroups for `total` records: the resident ones, capped at one
// grain per wavefront.
unsigned trace_blocks(int64_t resident_waves, int64_t total)
{
    const int64_t wpb = STREAM_THREADS / 64;
    const int64_t need = (total + (int64_t)g_trace_grain * wpb - 1) / ((int64_t)g_trace_grain * wpb);
    const int64_t have = resident_waves / wpb;
    return (unsigned)std::max<int64_t>(1, std::min(need, have));
}

unsigned gen_grid(uint32_t total)
{
    const uint64_t waves = (total + 64ull * GEN_ROUNDS - 1) / (64ull * GEN_ROUNDS);
    return (unsigned)((waves * 64 + TRACE_THREADS - 1) / TRACE_THREADS);
}

// Origins and directions of the rays of patches [p0, p0 + n_p).
int launch_gen_rays(const Geometry &g, uint32_t p0, uint32_t n_p, uint32_t R, const uint64_t *d_seeds, uint32_t total,
                    double *d_orig, double *d_dir, hipStream_t st)
{
    if (total == 0) return PRH_OK;
    GenArgs a{view_of(g), traversal_const(g), p0, n_p, R, R, 0u, d_seeds, total, 0u, d_orig, d_dir, nullptr, nullptr};
    k_gen_rays<false><<<gen_grid(total), TRACE_THREADS, 0, st>>>(a);
    PRH_HIP(hipGetLastError());
    return PRH_OK;
}

// Traversal records of the rays of patches [p0, p0 + n_p).
// R rays per patch: rays r0 .. r0 + R - 1 of blocks of R_stream rays per patch.
int launch_gen_records(const Geometry &g, uint32_t p0, uint32_t n_p, uint32_t R, uint32_t R_stream, uint32_t r0,
                       const uint64_t *d_seeds, uint32_t total, DdaRec *d_dda, WallRec *d_walls, hipStream_t st)
{
    if (total == 0) return PRH_OK;
    GenArgs a{view_of(g), traversal_const(g), p0, n_p, R, R_stream, r0, d_seeds, total, 0u,
              nullptr,    nullptr,          d_dda, d_walls};
    k_gen_rays<true><<<gen_grid(total), TRACE_THREADS, 0, st>>>(a);
    PRH_HIP(hipGetLastError());
    return PRH_OK;
}

int launch_trace(const Geometry &g, const DdaRec *d_dda, const WallRec *d_walls, uint32_t total, uint2 *d_fin,
                 unsigned long long *d_reads, unsigned long long *d_queue, hipStream_t st)
{
    if (total == 0) return PRH_OK;
    const TraceVariant &V = trace_variant_for(g, g_trace_variant);
    const size_t lds = V.occ == OCC_BRICK ? (size_t)g.coarse_words * sizeof(uint32_t) : 0;
    int64_t waves = 0;
    int rc = resident_waves((const void *)V.fn, STREAM_THREADS, lds, &waves);
    if (rc) return rc;
    // the resident workgroups, or fewer when there is not a grain per wavefront
    const unsigned grid = trace_blocks(waves, total);
    TraceArgs a{view_of(g), total, (uint32_t)g_trace_grain, d_dda, d_walls, d_fin, d_reads, d_queue};
    V.fn<<<grid, STREAM_THREADS, lds, st>>>(a);
    PRH_HIP(hipGetLastError());
    return PRH_OK;
}

// Owner of edge tables made along the way; frees whatever it still holds.
struct EdgeParts {
    std::vector<prh_edges *> parts;
    ~EdgeParts()
    {
        for (prh_edges *e : parts) prh_edges_destroy(e);
    }
};

// One traced pass: patches [p0, p0 + n_p) with S seeds (d_seeds, 4 words
// each), R rays per patch (rays r0 .. r0 + R - 1 of each seed's blocks of
// R_stream rays per patch), S R <= MAX_RAYS_PER_PATCH_CALL and
// n_p S R < 2^32: generation, traversal, per-patch counting and compaction
// into a new device edge table.  Adds the occupancy reads and stage times.
int trace_pass(const Geometry &g, Workspace &W, int64_t p0, int64_t n_p, int R, int R_stream, int r0, int S,
               const uint64_t *d_seeds, hipStream_t st, prh_edges **out, unsigned long long *reads, double ms[4])
{
    const int dev = g.device;
    const int SR = R * S;
    const int64_t total = n_p * SR;
    int SRp2 = 1;
    while (SRp2 < SR) SRp2 <<= 1;
    const size_t t1 = total > 0 ? (size_t)total : 1, n1 = n_p > 0 ? (size_t)n_p : 1;
    int rc;
    if ((rc = grow(W.dda, t1 * sizeof(DdaRec))) || (rc = grow(W.walls, t1 * sizeof(WallRec))) ||
        (rc = grow(W.fin, t1 * sizeof(uint2))) || (rc = grow(W.uniq_recv, t1 * sizeof(int32_t))) ||
        (rc = grow(W.uniq_cnt, t1 * sizeof(int32_t))) || (rc = grow(W.n_unique, n1 * sizeof(int32_t))) ||
        (rc = grow(W.offsets, n1 * sizeof(int64_t))) || (rc = grow(W.counters, 3 * sizeof(unsigned long long))))
        return rc;
    // counters: [0] occupancy reads, [1] failures, [2] traversal queue
    PRH_HIP(hipMemsetAsync(W.counters.p, 0, 3 * sizeof(unsigned long long), st));
    *out = nullptr;
    if (total == 0) {
        *out = new prh_edges{dev, 0, nullptr, nullptr, nullptr};
        return PRH_OK;
    }
    PRH_HIP(hipEventRecord(W.ev[0], st));
    unsigned long long *ctr = W.counters.as<unsigned long long>();
    if ((rc = launch_gen_records(g, (uint32_t)p0, (uint32_t)n_p, (uint32_t)R, (uint32_t)R_stream, (uint32_t)r0,
                                 d_seeds, (uint32_t)total, W.dda.as<DdaRec>(), W.walls.as<WallRec>(), st)))
        return rc;
    PRH_HIP(hipEventRecord(W.ev[1], st));
    if ((rc = launch_trace(g, W.dda.as<DdaRec>(), W.walls.as<WallRec>(), (uint32_t)total, W.fin.as<uint2>(), ctr,
                           ctr + 2, st)))
        return rc;
    PRH_HIP(hipEventRecord(W.ev[2], st));
    const unsigned cgrid = (unsigned)(n_p < 65536 ? n_p : 65536);
    if (SRp2 <= COUNT_WAVE_MAX_KEYS) {
        // one wavefront per patch, keys in registers
        const int epl = SRp2 <= 64 ? 1 : SRp2 / 64;
        const int64_t blocks = (n_p + COUNT_THREADS / 64 - 1) / (COUNT_THREADS / 64);
        const unsigned wgrid = (unsigned)(blocks < 65536 ? blocks : 65536);
        auto kern = count_wave_kernel(epl, S);
        kern<<<wgrid, COUNT_THREADS, 0, st>>>(W.fin.as<uint2>(), W.walls.as<WallRec>(), view_of(g), n_p, R, S,
                                              W.uniq_recv.as<int32_t>(),
                                              W.uniq_cnt.as<int32_t>(), W.n_unique.as<int32_t>(), ctr + 1);
    } else {
        const size_t lds = sizeof(int32_t) * (SRp2 + COUNT_THREADS);
        k_count_patches<<<cgrid, COUNT_THREADS, lds, st>>>(W.fin.as<uint2>(), W.walls.as<WallRec>(), view_of(g), n_p,
                                                           R, S, SRp2,
                                                           W.uniq_recv.as<int32_t>(), W.uniq_cnt.as<int32_t>(),
                                                           W.n_unique.as<int32_t>(), ctr + 1);
    }
    PRH_HIP(hipGetLastError());
    PRH_HIP(hipEventRecord(W.ev[3], st));
    // the counters come back with the edge total (one wait for both)
    unsigned long long counters[2] = {0, 0};
    PRH_HIP(hipMemcpyAsync(counters, W.counters.p, sizeof(counters), hipMemcpyDeviceToHost, st));
    int64_t nnz = 0;
    if ((rc = exclusive_scan_i32_to_i64(W.n_unique.as<int32_t>(), W.offsets.as<int64_t>(), n_p, &nnz, st)))
        return rc;
    *reads += counters[0];
    if (counters[1]) {
        set_error("%llu rays did not resolve a receiving patch", counters[1]);
        return PRH_ERR_UNRESOLVED;
    }
    prh_edges *e = nullptr;
    if ((rc = new_edge_table(dev, nnz, &e))) return rc;
    if (nnz > 0) {
        // the wall records are not read after counting: their space holds the range starts
        const int64_t n_ranges = (nnz + COMPACT_RANGE - 1) / COMPACT_RANGE;
        int32_t *range_q = W.walls.as<int32_t>();
        k_range_starts<<<(unsigned)std::min<int64_t>((n_p + 255) / 256, 65536), 256, 0, st>>>(
            W.offsets.as<int64_t>(), n_p, nnz, range_q);
        k_compact_ranges<<<(unsigned)std::min<int64_t>(n_ranges, 65536), COUNT_THREADS, 0, st>>>(
            W.uniq_recv.as<int32_t>(), W.uniq_cnt.as<int32_t>(), W.offsets.as<int64_t>(), range_q, n_p, SR, p0, nnz,
            e->origin, e->receiver, e->count);
    }
    hipError_t err = hipGetLastError();
    if (err == hipSuccess) err = hipEventRecord(W.ev[4], st);
    if (err == hipSuccess) err = hipStreamSynchronize(st);
    if (err != hipSuccess) {
        prh_edges_destroy(e);
        set_error("prh_trace_counts: %s", hipGetErrorString(err));
        return PRH_ERR_HIP;
    }
    float t = 0.f;
    for (int k = 0; k < 3; ++k) {
        PRH_HIP(hipEventElapsedTime(&t, W.ev[k], W.ev[k + 1]));
        ms[k] += t;
    }
    PRH_HIP(hipEventElapsedTime(&t, W.ev[3], W.ev[4]));
    ms[3] += t;
    *out = e;
    return PRH_OK;
}

// The tables of consecutive patch ranges, in order, as one table.
int concatenate_edges(const std::vector<prh_edges *> &parts, int dev, hipStream_t st, prh_edges **out)
{
    int64_t nnz = 0;
    for (const prh_edges *e : parts) nnz += e->nnz;
    prh_edges *res = nullptr;
    int rc = new_edge_table(dev, nnz, &res);
    if (rc) return rc;
    int64_t at = 0;
    hipError_t err = hipSuccess;
    for (const prh_edges *e : parts) {
        if (e->nnz == 0) continue;
        const size_t b = (size_t)e->nnz * sizeof(int32_t);
        if (err == hipSuccess) err = hipMemcpyAsync(res->origin + at, e->origin, b, hipMemcpyDeviceToDevice, st);
        if (err == hipSuccess) err = hipMemcpyAsync(res->receiver + at, e->receiver, b, hipMemcpyDeviceToDevice, st);
        if (err == hipSuccess) err = hipMemcpyAsync(res->count + at, e->count, b, hipMemcpyDeviceToDevice, st);
        at += e->nnz;
    }
    if (err == hipSuccess) err = hipStreamSynchronize(st);
    if (err != hipSuccess) {
        prh_edges_destroy(res);
        set_error("prh_trace_counts: %s", hipGetErrorString(err));
        return PRH_ERR_HIP;
    }
    *out = res;
    return PRH_OK;
}

// ------------------------------------------------------------------ allocation-free counting
// prh_trace_counts_into works in this fixed device array instead of the
// grow-only workspace: the ray records of one pass, the per-patch counts and
// offsets, the scan's tile sums, the seeds and the counters.  Each device
// holds its own copy until the process ends; its address comes from
// hipGetSymbolAddress once per device (the array's name is no device address
// in host code).
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(256) << 20;
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

unsigned char *g_scratch[64] = {};
std::mutex g_scratch_lock[64];

int scratch_address(int dev, unsigned char **out)
{
    if (!g_scratch[dev]) {
        void *p = nullptr;
        PRH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch[dev];
    return PRH_OK;
}

// Bump allocation inside the scratch array, 256-byte aligned.
struct ScratchCarver {
    unsigned char *base;
    size_t used = 0;
    template <typename T>
    T *take(size_t count)
    {
        T *p = reinterpret_cast<T *>(base + used);
        used += (count * sizeof(T) + 255) & ~size_t(255);
        return p;
    }
};

// Scratch bytes of a pass of n_p patches with SR rays each (aligned pieces).
size_t noalloc_pass_bytes(int64_t n_p, int64_t SR)
{
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t rays = (size_t)(n_p * SR);
    return al(rays * sizeof(DdaRec)) + al(rays * sizeof(WallRec)) + al(rays * sizeof(uint2)) +
           2 * al(rays * sizeof(int32_t)) + al((size_t)n_p * sizeof(int32_t)) + al((size_t)n_p * sizeof(int64_t)) +
           al((size_t)scan_scratch_elems(n_p) * sizeof(int64_t));
}

// Element-wise comparison of two device edge tables of n entries (sets *diff).
__global__ __launch_bounds__(256) void k_edges_differ(const int32_t *ao, const int32_t *ar, const int32_t *ac,
                                                      const int32_t *bo, const int32_t *br, const int32_t *bc,
                                                      int64_t n, unsigned int *diff)
{
    bool d = false;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256)
        d |= ao[e] != bo[e] || ar[e] != br[e] || ac[e] != bc[e];
    if (__ballot(d) != 0 && (threadIdx.x & 63u) == 0) atomicOr(diff, 1u);
}

}  // namespace

}  // namespace prh


using namespace prh;

extern "C" {

int prh_last_timings(double *ms, int32_t n)
{
    PRH_CHECK_ARG(ms && n >= 0, "prh_last_timings: null output");
    for (int i = 0; i < n && i < 4; ++i) ms[i] = g_last_ms[i];
    return PRH_OK;
}

int prh_tracer_set_variant(int32_t variant)
{
    PRH_CHECK_ARG(variant >= 0 && variant < N_TRACE_VARIANTS, "prh_tracer_set_variant: variant %d out of range [0, %d)",
                  variant, N_TRACE_VARIANTS);
    g_trace_variant = variant;
    return PRH_OK;
}

int prh_tracer_trim(void)
{
    int dev = 0;
    PRH_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return PRH_OK;
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    PRH_HIP(hipStreamSynchronize(st));
    edge_cache_free(dev);
    return workspace_free(g_ws[dev]);
}

int prh_tracer_set_grain(int32_t records)
{
    PRH_CHECK_ARG(records >= 64 && records <= (1 << 20), "prh_tracer_set_grain: %d records out of range [64, 2^20]",
                  records);
    g_trace_grain = records;
    return PRH_OK;
}

int prh_tracer_grain(int32_t *records)
{
    PRH_CHECK_ARG(records, "prh_tracer_grain: null pointer");
    *records = g_trace_grain;
    return PRH_OK;
}

int prh_tracer_stats(int64_t *wave_steps, int64_t *refills, int32_t reset)
{
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    PRH_HIP(hipStreamSynchronize(st));
    unsigned long long v[2] = {0, 0};
    PRH_HIP(hipMemcpyFromSymbol(v, HIP_SYMBOL(g_trace_stats), sizeof(v)));
    if (wave_steps) *wave_steps = (int64_t)v[0];
    if (refills) *refills = (int64_t)v[1];
    if (reset) {
        const unsigned long long z[2] = {0, 0};
        PRH_HIP(hipMemcpyToSymbol(HIP_SYMBOL(g_trace_stats), z, sizeof(z)));
    }
    return PRH_OK;
}

int prh_tracer_variants(int32_t *n_variants, int32_t *current, int32_t *refill_min, int32_t *occupancy_mode)
{
    if (n_variants) *n_variants = N_TRACE_VARIANTS;
    if (current) *current = g_trace_variant;
    for (int v = 0; v < N_TRACE_VARIANTS; ++v) {
        if (refill_min) refill_min[v] = TRACE_VARIANTS[v].refill_min;
        if (occupancy_mode) occupancy_mode[v] = TRACE_VARIANTS[v].occ;
    }
    return PRH_OK;
}

int prh_generate_rays(const prh_geometry *geo, int64_t p0, int64_t p1, int32_t rays_per_patch,
                      const uint64_t state_inc[4], double *origins, double *directions)
{
    PRH_CHECK_ARG(geo && state_inc && origins && directions, "prh_generate_rays: null pointer");
    const Geometry &g = geo->g;
    PRH_CHECK_ARG(0 <= p0 && p0 <= p1 && p1 <= g.n_patches, "prh_generate_rays: patch range out of bounds");
    PRH_CHECK_ARG(rays_per_patch > 0, "prh_generate_rays: rays_per_patch must be positive");
    const int64_t m = (p1 - p0) * (int64_t)rays_per_patch;
    if (m == 0) return PRH_OK;
    PRH_CHECK_ARG(m < (int64_t(1) << 32), "prh_generate_rays: more than 2^32 rays in one call");
    PRH_HIP(hipSetDevice(g.device));
    int rc = ensure_tables();
    if (rc) return rc;
    hipStream_t st;
    if ((rc = current_stream(&st))) return rc;
    Workspace &W = g_ws[g.device];
    if ((rc = grow(W.seeds, 4 * sizeof(uint64_t)))) return rc;
    if ((rc = grow(W.origins, m * 3 * sizeof(double)))) return rc;
    if ((rc = grow(W.directions, m * 3 * sizeof(double)))) return rc;
    PRH_HIP(hipMemcpyAsync(W.seeds.p, state_inc, 4 * sizeof(uint64_t), hipMemcpyHostToDevice, st));
    if ((rc = launch_gen_rays(g, (uint32_t)p0, (uint32_t)(p1 - p0), (uint32_t)rays_per_patch,
                              W.seeds.as<uint64_t>(), (uint32_t)m, W.origins.as<double>(), W.directions.as<double>(),
                              st)))
        return rc;
    PRH_HIP(hipMemcpyAsync(origins, W.origins.p, m * 3 * sizeof(double), hipMemcpyDeviceToHost, st));
    PRH_HIP(hipMemcpyAsync(directions, W.directions.p, m * 3 * sizeof(double), hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));
    return PRH_OK;
}

int prh_trace_rays(const prh_geometry *geo, const double *origins, const double *directions, int64_t n_rays,
                   int32_t *receiver, int64_t *n_failed)
{
    PRH_CHECK_ARG(geo && receiver && (n_rays == 0 || (origins && directions)), "prh_trace_rays: null pointer");
    PRH_CHECK_ARG(n_rays >= 0, "prh_trace_rays: negative ray count");
    const Geometry &g = geo->g;
    PRH_CHECK_ARG(g.n_patches > 0 && g.cell_code && g.hot_map && g.cold_map,
                  "prh_trace_rays: the geometry has no patch tables (an occupancy-only handle)");
    if (n_failed) *n_failed = 0;
    if (n_rays == 0) return PRH_OK;
    // The traversal's cell indices are 32-bit: origins within |o| < 2^31 (any
    // other finite origin lies far outside every grid), finite directions.
    for (int64_t r = 0; r < 3 * n_rays; ++r)
        PRH_CHECK_ARG(std::isfinite(directions[r]) && std::isfinite(origins[r]) && fabs(origins[r]) < 2147483648.0,
                      "prh_trace_rays: ray %lld has a non-finite direction or an origin outside |o| < 2^31",
                      (long long)(r / 3));
    PRH_HIP(hipSetDevice(g.device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    Workspace &W = g_ws[g.device];
    int64_t budget = 0;
    if ((rc = workspace_budget(W, &budget))) return rc;
    if (workspace_held(W) > budget && (rc = workspace_free(W))) return rc;
    if (workspace_held(W) + edge_cache_held(g.device) > budget) edge_cache_free(g.device);
    // rays in chunks that fit the workspace budget (and 2^31 records)
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(budget / GIVEN_BYTES_PER_RAY, int64_t(1) << 31));
    const int64_t m = std::min(chunk, n_rays);
    if ((rc = grow(W.origins, m * 3 * sizeof(double))) || (rc = grow(W.directions, m * 3 * sizeof(double))) ||
        (rc = grow(W.dda, m * sizeof(DdaRec))) || (rc = grow(W.walls, m * sizeof(WallRec))) ||
        (rc = grow(W.fin, m * sizeof(uint2))) || (rc = grow(W.receiver, m * sizeof(int32_t))) ||
        (rc = grow(W.counters, 3 * sizeof(unsigned long long))))
        return rc;
    unsigned long long failed_total = 0;
    for (int64_t r0 = 0; r0 < n_rays; r0 += m) {
        const int64_t k = std::min(m, n_rays - r0);
        PRH_HIP(hipMemcpyAsync(W.origins.p, origins + 3 * r0, k * 3 * sizeof(double), hipMemcpyHostToDevice, st));
        PRH_HIP(hipMemcpyAsync(W.directions.p, directions + 3 * r0, k * 3 * sizeof(double), hipMemcpyHostToDevice,
                               st));
        PRH_HIP(hipMemsetAsync(W.counters.p, 0, 3 * sizeof(unsigned long long), st));
        const uint32_t total = (uint32_t)k;
        const unsigned grid = (unsigned)((k + TRACE_THREADS - 1) / TRACE_THREADS);
        k_setup_rays<<<grid, TRACE_THREADS, 0, st>>>(view_of(g), traversal_const(g), W.origins.as<double>(),
                                                     W.directions.as<double>(), total, W.dda.as<DdaRec>(),
                                                     W.walls.as<WallRec>());
        PRH_HIP(hipGetLastError());
        unsigned long long *ctr = W.counters.as<unsigned long long>();
        if ((rc = launch_trace(g, W.dda.as<DdaRec>(), W.walls.as<WallRec>(), total, W.fin.as<uint2>(), ctr, ctr + 2,
                               st)))
            return rc;
        k_resolve<<<grid, TRACE_THREADS, 0, st>>>(W.fin.as<uint2>(), W.walls.as<WallRec>(), view_of(g), total,
                                                  W.receiver.as<int32_t>(), ctr + 1);
        PRH_HIP(hipGetLastError());
        unsigned long long counters[2] = {0, 0};
        PRH_HIP(hipMemcpyAsync(receiver + r0, W.receiver.p, k * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        PRH_HIP(hipMemcpyAsync(counters, W.counters.p, sizeof(counters), hipMemcpyDeviceToHost, st));
        PRH_HIP(hipStreamSynchronize(st));
        failed_total += counters[1];
    }
    if (n_failed) *n_failed = (int64_t)failed_total;
    if (failed_total) {
        set_error("%llu rays did not resolve a receiving patch", failed_total);
        return PRH_ERR_UNRESOLVED;
    }
    return PRH_OK;
}

int prh_trace_counts(const prh_geometry *geo, int64_t p0, int64_t p1, int32_t rays_per_patch, int32_t n_seeds,
                     const uint64_t *state_inc, prh_edges **out, int64_t *nnz_out, int64_t *occupancy_reads)
{
    PRH_CHECK_ARG(geo && state_inc && out, "prh_trace_counts: null pointer");
    const Geometry &g = geo->g;
    PRH_CHECK_ARG(0 <= p0 && p0 <= p1 && p1 <= g.n_patches, "prh_trace_counts: patch range out of bounds");
    PRH_CHECK_ARG(rays_per_patch > 0 && n_seeds > 0, "prh_trace_counts: rays_per_patch and n_seeds must be positive");
    const int64_t R = rays_per_patch, S = n_seeds;
    if (R * S > INT32_MAX) {
        // a pooled count must fit the int32 counts of the device edge table
        set_error("prh_trace_counts: rays_per_patch * n_seeds = %lld exceeds 2^31 - 1 rays per patch",
                  (long long)(R * S));
        return PRH_ERR_CAPACITY;
    }
    *out = nullptr;
    PRH_HIP(hipSetDevice(g.device));
    const int dev = g.device;
    int rc = ensure_tables();
    if (rc) return rc;
    hipStream_t st;
    if ((rc = current_stream(&st))) return rc;
    Workspace &W = g_ws[dev];
    if (!W.events) {
        for (auto &e : W.ev) PRH_HIP(hipEventCreate(&e));
        W.events = true;
    }
    int64_t budget = 0;
    if ((rc = workspace_budget(W, &budget))) return rc;
    if (workspace_held(W) > budget && (rc = workspace_free(W))) return rc;
    if (workspace_held(W) + edge_cache_held(g.device) > budget) edge_cache_free(g.device);
    if ((rc = grow(W.seeds, sizeof(uint64_t) * 4 * S))) return rc;
    PRH_HIP(hipMemcpyAsync(W.seeds.p, state_inc, sizeof(uint64_t) * 4 * S, hipMemcpyHostToDevice, st));

    // Seed blocks whose rays a per-patch count can sort in one pass (at most
    // MAX_RAYS_PER_PATCH_CALL per patch): all seeds at once, groups of seeds,
    // or each seed's rays in parts.  Their tables are pooled at the end
    // (pool_counts, view_factors.py:290-300).
    struct Group {
        int s0, S, R, r0;
    };
    std::vector<Group> groups;
    const int64_t CAP = MAX_RAYS_PER_PATCH_CALL;
    if (R * S <= CAP) {
        groups.push_back({0, (int)S, (int)R, 0});
    } else if (R <= CAP) {
        const int64_t per = CAP / R;
        for (int64_t s0 = 0; s0 < S; s0 += per) groups.push_back({(int)s0, (int)std::min(per, S - s0), (int)R, 0});
    } else {
        for (int64_t s = 0; s < S; ++s)
            for (int64_t r0 = 0; r0 < R; r0 += CAP) groups.push_back({(int)s, 1, (int)std::min(CAP, R - r0), (int)r0});
    }

    // Each group's patch range in chunks whose workspace fits the budget.
    double ms[4] = {0, 0, 0, 0};
    unsigned long long reads = 0;
    EdgeParts parts;
    const int64_t n_p = p1 - p0;
    for (const Group &gr : groups) {
        const int64_t SR = (int64_t)gr.S * gr.R;
        const int64_t per_patch = SR * PASS_BYTES_PER_RAY + PASS_BYTES_PER_PATCH;
        int64_t chunk = std::max<int64_t>(1, budget / per_patch);
        chunk = std::min<int64_t>(chunk, ((int64_t(1) << 32) - 1) / SR);
        const uint64_t *d_seeds = W.seeds.as<uint64_t>() + 4 * gr.s0;
        for (int64_t c0 = 0; c0 < std::max<int64_t>(n_p, 1); c0 += chunk) {
            const int64_t c1 = std::min(n_p, c0 + chunk);
            prh_edges *e = nullptr;
            if ((rc = trace_pass(g, W, p0 + c0, c1 - c0, gr.R, (int)R, gr.r0, gr.S, d_seeds, st, &e, &reads, ms)))
                return rc;
            parts.parts.push_back(e);
        }
    }
    prh_edges *result = nullptr;
    if (parts.parts.size() == 1) {
        result = parts.parts[0];
        parts.parts.clear();
    } else if (groups.size() == 1) {
        // consecutive patch ranges of one group: the tables follow each other
        if ((rc = concatenate_edges(parts.parts, dev, st, &result))) return rc;
    } else {
        if ((rc = pool_edge_tables(parts.parts.data(), (int)parts.parts.size(), g.n_patches, st, &result))) return rc;
    }
    for (int k = 0; k < 4; ++k) g_last_ms[k] = ms[k];
    *out = result;
    if (nnz_out) *nnz_out = result->nnz;
    if (occupancy_reads) *occupancy_reads = (int64_t)reads;
    return PRH_OK;
}

int prh_trace_counts_into(const prh_geometry *geo, int64_t p0, int64_t p1, int32_t rays_per_patch, int32_t n_seeds,
                          const uint64_t *state_inc, int32_t *origin, int32_t *receiver, int32_t *count,
                          int64_t capacity, int64_t *nnz_out, int64_t *occupancy_reads)
{
    PRH_CHECK_ARG(geo && state_inc && nnz_out, "prh_trace_counts_into: null pointer");
    const Geometry &g = geo->g;
    PRH_CHECK_ARG(0 <= p0 && p0 <= p1 && p1 <= g.n_patches, "prh_trace_counts_into: patch range out of bounds");
    PRH_CHECK_ARG(rays_per_patch > 0 && n_seeds > 0,
                  "prh_trace_counts_into: rays_per_patch and n_seeds must be positive");
    PRH_CHECK_ARG(capacity >= 0 && (capacity == 0 || (origin && receiver && count)),
                  "prh_trace_counts_into: null output arrays or negative capacity");
    const int64_t SR = (int64_t)rays_per_patch * n_seeds;
    if (SR > MAX_RAYS_PER_PATCH_CALL) {
        set_error("prh_trace_counts_into: rays_per_patch * n_seeds = %lld exceeds %d (one pass per patch)",
                  (long long)SR, MAX_RAYS_PER_PATCH_CALL);
        return PRH_ERR_CAPACITY;
    }
    // fixed part: seeds and five counters; then at least one patch per pass
    const size_t fixed = ((4 * (size_t)n_seeds * sizeof(uint64_t) + 255) & ~size_t(255)) + 256;
    if (fixed + noalloc_pass_bytes(1, SR) > NOALLOC_SCRATCH_BYTES) {
        set_error("prh_trace_counts_into: one patch of %lld rays needs more than the %zu-byte scratch",
                  (long long)SR, NOALLOC_SCRATCH_BYTES);
        return PRH_ERR_CAPACITY;
    }
    PRH_HIP(hipSetDevice(g.device));
    const int dev = g.device;
    PRH_CHECK_ARG(dev >= 0 && dev < 64, "prh_trace_counts_into: device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_lock[dev]);
    int rc = ensure_tables();
    if (rc) return rc;
    hipStream_t st;
    if ((rc = current_stream(&st))) return rc;
    unsigned char *base = nullptr;
    if ((rc = scratch_address(dev, &base))) return rc;

    ScratchCarver sc{base};
    uint64_t *d_seeds = sc.take<uint64_t>(4 * (size_t)n_seeds);
    unsigned long long *ctr = sc.take<unsigned long long>(5);  // reads, failures, queue, carry, start
    int64_t *carry = reinterpret_cast<int64_t *>(ctr + 3), *start = reinterpret_cast<int64_t *>(ctr + 4);
    const size_t room = NOALLOC_SCRATCH_BYTES - sc.used;
    int64_t chunk = 1;
    while (chunk < p1 - p0 && noalloc_pass_bytes(2 * chunk, SR) <= room) chunk *= 2;
    if (noalloc_pass_bytes(chunk, SR) > room) chunk /= 2;
    chunk = std::max<int64_t>(chunk, 1);
    const size_t rays = (size_t)(chunk * SR);
    DdaRec *dda = sc.take<DdaRec>(rays);
    WallRec *walls = sc.take<WallRec>(rays);
    uint2 *fin = sc.take<uint2>(rays);
    int32_t *uniq_recv = sc.take<int32_t>(rays), *uniq_cnt = sc.take<int32_t>(rays);
    int32_t *n_unique = sc.take<int32_t>((size_t)chunk);
    int64_t *offsets = sc.take<int64_t>((size_t)chunk);
    int64_t *scan_tiles = sc.take<int64_t>((size_t)scan_scratch_elems(chunk));

    // the scratch keeps the last call's contents: every counter starts at zero
    PRH_HIP(hipMemcpyAsync(d_seeds, state_inc, 4 * sizeof(uint64_t) * n_seeds, hipMemcpyHostToDevice, st));
    k_zero_u64<<<1, 64, 0, st>>>(ctr, 5);
    PRH_HIP(hipGetLastError());
    const int R = rays_per_patch, S = n_seeds;
    int SRp2 = 1;
    while (SRp2 < SR) SRp2 <<= 1;
    for (int64_t c0 = p0; c0 < p1; c0 += chunk) {
        const int64_t n_p = std::min(chunk, p1 - c0);
        const uint32_t total = (uint32_t)(n_p * SR);
        k_zero_u64<<<1, 64, 0, st>>>(ctr + 2, 1);  // the traversal's queue starts at zero
        if ((rc = launch_gen_records(g, (uint32_t)c0, (uint32_t)n_p, (uint32_t)R, (uint32_t)R, 0u, d_seeds, total,
                                     dda, walls, st)) ||
            (rc = launch_trace(g, dda, walls, total, fin, ctr, ctr + 2, st)))
            return rc;
        const int64_t blocks = (n_p + COUNT_THREADS / 64 - 1) / (COUNT_THREADS / 64);
        if (SRp2 <= COUNT_WAVE_MAX_KEYS) {
            const int epl = SRp2 <= 64 ? 1 : SRp2 / 64;
            auto kern = count_wave_kernel(epl, S);
            kern<<<(unsigned)std::min<int64_t>(blocks, 65536), COUNT_THREADS, 0, st>>>(
                fin, walls, view_of(g), n_p, R, S, uniq_recv, uniq_cnt, n_unique, ctr + 1);
        } else {
            k_count_patches<<<(unsigned)std::min<int64_t>(n_p, 65536), COUNT_THREADS,
                              sizeof(int32_t) * (SRp2 + COUNT_THREADS), st>>>(fin, walls, view_of(g), n_p, R, S, SRp2,
                                                                               uniq_recv, uniq_cnt, n_unique, ctr + 1);
        }
        PRH_HIP(hipGetLastError());
        // edge offsets continue the running total kept on the device
        if ((rc = exclusive_scan_i32_to_i64_carry_async(n_unique, offsets, n_p, scan_tiles, carry, start, st)))
            return rc;
        k_compact_edges<<<(unsigned)std::min<int64_t>(n_p, 65536), COUNT_THREADS, 0, st>>>(
            uniq_recv, uniq_cnt, n_unique, offsets, n_p, (int)SR, c0, origin, receiver, count, capacity);
        PRH_HIP(hipGetLastError());
    }
    unsigned long long host[4] = {0, 0, 0, 0};
    PRH_HIP(hipMemcpyAsync(host, ctr, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    PRH_HIP(hipStreamSynchronize(st));
    const int64_t nnz = (int64_t)host[3];
    *nnz_out = nnz;
    if (occupancy_reads) *occupancy_reads = (int64_t)host[0];
    if (host[1]) {
        set_error("%llu rays did not resolve a receiving patch", host[1]);
        return PRH_ERR_UNRESOLVED;
    }
    if (nnz > capacity) {
        set_error("prh_trace_counts_into: %lld edges exceed the output capacity %lld", (long long)nnz,
                  (long long)capacity);
        return PRH_ERR_CAPACITY;
    }
    return PRH_OK;
}

int prh_tracer_set_workspace_limit(int64_t bytes)
{
    PRH_CHECK_ARG(bytes >= (int64_t(1) << 20), "prh_tracer_set_workspace_limit: %lld bytes is below 1 MiB",
                  (long long)bytes);
    g_ws_limit = bytes;
    return PRH_OK;
}

int prh_tracer_workspace(int64_t *held_bytes, int64_t *limit_bytes)
{
    int dev = 0;
    PRH_HIP(hipGetDevice(&dev));
    PRH_CHECK_ARG(dev >= 0 && dev < 64, "prh_tracer_workspace: device %d out of range", dev);
    if (held_bytes) *held_bytes = workspace_held(g_ws[dev]) + edge_cache_held(dev);
    if (limit_bytes) *limit_bytes = g_ws_limit;
    return PRH_OK;
}

}  // extern "C"

// A multi-geometry batch: segments (geometry, patch range, rays per patch,
// seeds) cut at patch boundaries into pieces that fill fixed-size chunks of
// ray records; per chunk one generation launch and one traversal launch over
// all its pieces, then per piece counting, a scan that carries the running
// edge total on the device, and compaction into one edge arena.  Chunks take
// turns over a few streams, each with its own record buffers, so the grazing
// rays that end one chunk's traversal run beside the next chunk's work; only
// the scans are chained from chunk to chunk (the arena's running total).  Each
// chunk is captured once as two graphs (generation + traversal + counting;
// scans + compaction) that every run replays on the chunk's stream, ordered by
// events: one graph over all streams would be replayed on one queue here and
// lose the overlap.
struct BatchPiece {
    int seg;
    int64_t a, b;      // patch range [a, b) of the segment's geometry
    int64_t ray_off;   // first record within the chunk
    int64_t patch_off; // first per-patch slot within the chunk
};

// Records and per-patch scratch of the chunks of one stream.
struct BatchBuffers {
    hipStream_t stream = nullptr;
    DdaRec *dda = nullptr;
    WallRec *walls = nullptr;
    uint2 *fin = nullptr;
    int32_t *uniq_recv = nullptr, *uniq_cnt = nullptr, *n_unique = nullptr;
    int64_t *offsets = nullptr, *scan_scratch = nullptr;
    hipEvent_t join = nullptr;
};

constexpr int MAX_BATCH_STREAMS = 4;

struct prh_trace_batch {
    int device = 0;
    int n_seg = 0;
    int n_streams = 1;
    bool use_graph = true;
    std::vector<const prh_geometry *> geos;
    std::vector<int> R, S;
    std::vector<int64_t> seed_off;  // first seed word of each segment
    int64_t seed_words = 0;
    std::vector<BatchPiece> pieces;
    std::vector<int> chunk_first;  // first piece of each chunk, then the piece count
    std::vector<int64_t> chunk_waves, chunk_blocks;
    std::vector<char> chunk_pow2;  // every geometry of the chunk has power-of-two ny and nz
    int64_t ray_cap = 0, patch_cap = 0, scratch_elems = 0, edge_cap = 0, rays = 0;
    size_t lds = 0;
    int variant = 0;
    BatchBuffers buf[MAX_BATCH_STREAMS];
    std::vector<hipEvent_t> scanned;  // per chunk: its scans are done (the carry is final)
    hipEvent_t fork = nullptr;
    Segment *segs_dev = nullptr;
    uint64_t *seeds = nullptr;
    unsigned long long *counters = nullptr;  // reads of each segment, then failures
    unsigned long long *queues = nullptr;    // traversal queue of each piece
    int64_t *marks = nullptr;                // running edge total, then each piece's first edge
    int32_t *e_origin = nullptr, *e_receiver = nullptr, *e_count = nullptr;
    std::vector<int64_t> seg_nnz, piece_start;
    int64_t edges = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    // captured graphs: [0] clears the counters, then per chunk front and back
    std::vector<hipGraph_t> graphs;
    std::vector<hipGraphExec_t> execs;
    int64_t bytes = 0;
};

namespace {

constexpr int64_t GEN_UNIT = 64 * GEN_ROUNDS;  // records of one generation wavefront
constexpr int64_t DEFAULT_BATCH_RAYS = int64_t(1) << 28;

int64_t align_gen(int64_t rays) { return (rays + GEN_UNIT - 1) / GEN_UNIT * GEN_UNIT; }

void free_trace_batch(prh_trace_batch *b)
{
    for (hipGraphExec_t e : b->execs)
        if (e) (void)hipGraphExecDestroy(e);
    for (hipGraph_t g : b->graphs)
        if (g) (void)hipGraphDestroy(g);
    for (hipEvent_t e : b->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->scanned)
        if (e) (void)hipEventDestroy(e);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (BatchBuffers &B : b->buf) {
        void *ptrs[] = {B.dda, B.walls, B.fin, B.uniq_recv, B.uniq_cnt, B.n_unique, B.offsets, B.scan_scratch};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
        if (B.join) (void)hipEventDestroy(B.join);
        if (B.stream) (void)hipStreamDestroy(B.stream);
    }
    void *ptrs[] = {b->segs_dev, b->seeds, b->counters, b->queues, b->marks, b->e_origin, b->e_receiver, b->e_count};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
}

void count_launch(const prh_trace_batch *b, const BatchBuffers &B, const BatchPiece &P, hipStream_t st)
{
    const Geometry &g = b->geos[P.seg]->g;
    const int R = b->R[P.seg], S = b->S[P.seg];
    const int64_t np = P.b - P.a;
    int SRp2 = 1;
    while (SRp2 < R * S) SRp2 <<= 1;
    const uint2 *fin = B.fin + P.ray_off;
    const WallRec *walls = B.walls + P.ray_off;
    int32_t *ur = B.uniq_recv + P.ray_off, *uc = B.uniq_cnt + P.ray_off, *nu = B.n_unique + P.patch_off;
    unsigned long long *failed = b->counters + b->n_seg;
    if (SRp2 <= COUNT_WAVE_MAX_KEYS) {
        const int epl = SRp2 <= 64 ? 1 : SRp2 / 64;
        const int64_t blocks = (np + COUNT_THREADS / 64 - 1) / (COUNT_THREADS / 64);
        const unsigned wgrid = (unsigned)(blocks < 65536 ? blocks : 65536);
        auto kern = count_wave_kernel(epl, S);
        kern<<<wgrid, COUNT_THREADS, 0, st>>>(fin, walls, view_of(g), np, R, S, ur, uc, nu, failed);
    } else {
        const unsigned cgrid = (unsigned)(np < 65536 ? np : 65536);
        const size_t lds = sizeof(int32_t) * (SRp2 + COUNT_THREADS);
        k_count_patches<<<cgrid, COUNT_THREADS, lds, st>>>(fin, walls, view_of(g), np, R, S, SRp2, ur, uc, nu,
                                                           failed);
    }
}

// Pieces of a pass, each enqueued on one stream (and captured as one graph).
int enqueue_clear(prh_trace_batch *b, hipStream_t st)
{
    const int n_pieces = (int)b->pieces.size();
    k_zero_u64<<<(b->n_seg + 256) / 256, 256, 0, st>>>(b->counters, b->n_seg + 1);
    k_zero_u64<<<(n_pieces + 255) / 256 + 1, 256, 0, st>>>(b->queues, n_pieces);
    k_zero_u64<<<(n_pieces + 256) / 256, 256, 0, st>>>(reinterpret_cast<unsigned long long *>(b->marks),
                                                       n_pieces + 1);
    PRH_HIP(hipGetLastError());
    return PRH_OK;
}

// Chunk c, first half: generation, traversal, per-piece counting.
int enqueue_front(prh_trace_batch *b, int c, hipStream_t st)
{
    const BatchVariant &V = batch_variant_for(b->variant, b->chunk_pow2[c] != 0);
    const BatchBuffers &B = b->buf[c % b->n_streams];
    const int j0 = b->chunk_first[c], j1 = b->chunk_first[c + 1];
    const unsigned gen_blocks = (unsigned)((b->chunk_waves[c] * 64 + TRACE_THREADS - 1) / TRACE_THREADS);
    k_gen_rays_batch<<<gen_blocks, TRACE_THREADS, 0, st>>>(b->segs_dev + j0, j1 - j0);
    V.fn<<<(unsigned)b->chunk_blocks[c], STREAM_THREADS, b->lds, st>>>(b->segs_dev + j0, j1 - j0, B.dda, B.walls,
                                                                        B.fin);
    PRH_HIP(hipGetLastError());
    for (int j = j0; j < j1; ++j) count_launch(b, B, b->pieces[j], st);
    PRH_HIP(hipGetLastError());
    return PRH_OK;
}

// Chunk c, second half: the scans (continuing the arena's running total) and
// the compaction of each piece.
int enqueue_back(prh_trace_batch *b, int c, hipStream_t st)
{
    const BatchBuffers &B = b->buf[c % b->n_streams];
    for (int j = b->chunk_first[c]; j < b->chunk_first[c + 1]; ++j) {
        const BatchPiece &P = b->pieces[j];
        const int64_t np = P.b - P.a;
        int rc = exclusive_scan_i32_to_i64_carry_async(B.n_unique + P.patch_off, B.offsets + P.patch_off, np,
                                                       B.scan_scratch, b->marks, b->marks + 1 + j, st);
        if (rc) return rc;
        const unsigned cgrid = (unsigned)(np < 65536 ? np : 65536);
        k_compact_edges<<<cgrid, COUNT_THREADS, 0, st>>>(
            B.uniq_recv + P.ray_off, B.uniq_cnt + P.ray_off, B.n_unique + P.patch_off, B.offsets + P.patch_off, np,
            b->R[P.seg] * b->S[P.seg], P.a, b->e_origin, b->e_receiver, b->e_count, b->edge_cap);
        PRH_HIP(hipGetLastError());
    }
    return PRH_OK;
}

// Piece `part` of the pass (0: clear; 1 + 2c: front of chunk c; 2 + 2c: back),
// replayed from its graph or launched directly.
int run_part(prh_trace_batch *b, int part, hipStream_t st)
{
    if (b->use_graph) {
        PRH_HIP(hipGraphLaunch(b->execs[part], st));
        return PRH_OK;
    }
    if (part == 0) return enqueue_clear(b, st);
    return (part & 1) ? enqueue_front(b, (part - 1) / 2, st) : enqueue_back(b, (part - 2) / 2, st);
}

// The pass: stream 0 clears the counters and forks the others; chunk c runs
// on stream c mod n_streams in that stream's buffers (so after the chunk that
// last used them); its scans wait for the previous chunk's.
int enqueue_pass(prh_trace_batch *b)
{
    hipStream_t s0 = b->buf[0].stream;
    int rc = run_part(b, 0, s0);
    if (rc) return rc;
    PRH_HIP(hipEventRecord(b->fork, s0));
    for (int i = 1; i < b->n_streams; ++i) PRH_HIP(hipStreamWaitEvent(b->buf[i].stream, b->fork, 0));
    const int n_chunks = (int)b->chunk_waves.size();
    for (int c = 0; c < n_chunks; ++c) {
        hipStream_t st = b->buf[c % b->n_streams].stream;
        if ((rc = run_part(b, 1 + 2 * c, st))) return rc;
        if (c > 0) PRH_HIP(hipStreamWaitEvent(st, b->scanned[c - 1], 0));
        if ((rc = run_part(b, 2 + 2 * c, st))) return rc;
        PRH_HIP(hipEventRecord(b->scanned[c], st));
    }
    for (int i = 1; i < b->n_streams; ++i) {
        PRH_HIP(hipEventRecord(b->buf[i].join, b->buf[i].stream));
        PRH_HIP(hipStreamWaitEvent(s0, b->buf[i].join, 0));
    }
    return PRH_OK;
}

// Capture part `part` of the pass on st into a graph.
int capture_part(prh_trace_batch *b, int part, hipStream_t st)
{
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        set_error("prh_trace_batch_create: stream capture did not start");
        return PRH_ERR_HIP;
    }
    const int rc = part == 0 ? enqueue_clear(b, st)
                   : (part & 1) ? enqueue_front(b, (part - 1) / 2, st)
                                : enqueue_back(b, (part - 2) / 2, st);
    hipGraph_t graph = nullptr;
    const hipError_t end = hipStreamEndCapture(st, &graph);
    b->graphs.push_back(graph);
    if (rc) return rc;
    if (end != hipSuccess || !graph) {
        set_error("prh_trace_batch_create: graph capture failed: %s", hipGetErrorString(end));
        return PRH_ERR_HIP;
    }
    hipGraphExec_t exec = nullptr;
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        set_error("prh_trace_batch_create: graph instantiation failed");
        return PRH_ERR_HIP;
    }
    b->execs.push_back(exec);
    return PRH_OK;
}

int fail_batch(prh_trace_batch *b, int rc)
{
    free_trace_batch(b);
    delete b;
    return rc;
}

}  // namespace

extern "C" {

int prh_trace_batch_create(int32_t n_segments, const prh_geometry *const *geometries, const int64_t *patch_begin,
                           const int64_t *patch_end, const int32_t *rays_per_patch, const int32_t *n_seeds,
                           int64_t ray_capacity, int64_t edge_capacity, int32_t n_streams, int32_t flags,
                           prh_trace_batch **out)
{
    PRH_CHECK_ARG(out && geometries && patch_begin && patch_end && rays_per_patch && n_seeds,
                  "prh_trace_batch_create: null pointer");
    *out = nullptr;
    PRH_CHECK_ARG(n_segments >= 1, "prh_trace_batch_create: no segments");
    PRH_CHECK_ARG(ray_capacity >= 0 && edge_capacity >= 0, "prh_trace_batch_create: negative capacity");
    PRH_CHECK_ARG(n_streams >= 0 && n_streams <= MAX_BATCH_STREAMS, "prh_trace_batch_create: 0 .. %d streams",
                  MAX_BATCH_STREAMS);
    PRH_CHECK_ARG((flags & ~PRH_BATCH_DIRECT) == 0, "prh_trace_batch_create: unknown flags 0x%x", flags);
    const int device = geometries[0] ? geometries[0]->g.device : 0;
    int64_t total = 0, widest = 0, worst_edges = 0;
    for (int k = 0; k < n_segments; ++k) {
        PRH_CHECK_ARG(geometries[k], "prh_trace_batch_create: null geometry of segment %d", k);
        const Geometry &g = geometries[k]->g;
        PRH_CHECK_ARG(g.device == device, "prh_trace_batch_create: segments on different devices");
        PRH_CHECK_ARG(0 <= patch_begin[k] && patch_begin[k] <= patch_end[k] && patch_end[k] <= g.n_patches,
                      "prh_trace_batch_create: patch range of segment %d out of bounds", k);
        PRH_CHECK_ARG(rays_per_patch[k] >= 1 && n_seeds[k] >= 1, "prh_trace_batch_create: segment %d has no rays", k);
        const int64_t SR = (int64_t)rays_per_patch[k] * n_seeds[k];
        if (SR > MAX_RAYS_PER_PATCH_CALL) {
            set_error("prh_trace_batch_create: segment %d has %lld rays per patch (at most %d)", k, (long long)SR,
                      MAX_RAYS_PER_PATCH_CALL);
            return PRH_ERR_CAPACITY;
        }
        const int64_t np = patch_end[k] - patch_begin[k];
        total += np * SR;
        widest = std::max(widest, align_gen(SR));
        worst_edges += np * std::min<int64_t>(SR, g.n_patches);
    }
    int64_t cap = ray_capacity > 0 ? ray_capacity : std::min(DEFAULT_BATCH_RAYS, align_gen(std::max<int64_t>(total, 1)));
    cap = std::max(cap / GEN_UNIT * GEN_UNIT, widest);
    if (cap >= (int64_t(1) << 32)) {
        set_error("prh_trace_batch_create: a chunk of %lld records exceeds 2^32", (long long)cap);
        return PRH_ERR_CAPACITY;
    }
    PRH_HIP(hipSetDevice(device));
    int rc = ensure_tables();
    if (rc) return rc;
    prh_trace_batch *b = new prh_trace_batch();
    b->device = device;
    b->n_seg = n_segments;
    b->ray_cap = cap;
    b->rays = total;
    b->use_graph = (flags & PRH_BATCH_DIRECT) == 0;
    b->variant = g_trace_variant;
    for (int k = 0; k < n_segments; ++k) b->variant = usable_variant(geometries[k]->g, b->variant);
    b->edge_cap = edge_capacity > 0 ? edge_capacity : std::max<int64_t>(worst_edges, 1);
    for (int k = 0; k < n_segments; ++k) {
        b->geos.push_back(geometries[k]);
        b->R.push_back(rays_per_patch[k]);
        b->S.push_back(n_seeds[k]);
        b->seed_off.push_back(b->seed_words);
        b->seed_words += 4 * (int64_t)n_seeds[k];
    }
    // pieces: whole patches into chunks of at most cap records
    int64_t used = 0, patches = 0;
    b->chunk_first.push_back(0);
    auto close_chunk = [&]() {
        if ((int)b->pieces.size() == b->chunk_first.back()) return;
        b->chunk_first.push_back((int)b->pieces.size());
        b->patch_cap = std::max(b->patch_cap, patches);
        used = patches = 0;
    };
    for (int k = 0; k < n_segments; ++k) {
        const int64_t SR = (int64_t)rays_per_patch[k] * n_seeds[k];
        for (int64_t p = patch_begin[k]; p < patch_end[k];) {
            const int64_t fit = std::min(patch_end[k] - p, (cap - used) / SR);
            if (fit == 0) {
                close_chunk();
                continue;
            }
            b->pieces.push_back(BatchPiece{k, p, p + fit, used, patches});
            used += align_gen(fit * SR);
            patches += fit;
            p += fit;
            b->scratch_elems = std::max(b->scratch_elems, scan_scratch_elems(fit));
        }
    }
    close_chunk();
    const int n_pieces = (int)b->pieces.size();
    const int n_chunks = (int)b->chunk_first.size() - 1;
    // streams (and record buffers): as asked, at most one per chunk, at least one
    b->n_streams = std::max(1, std::min(n_streams > 0 ? (int)n_streams : 2, std::max(n_chunks, 1)));
    bool ok = hipEventCreateWithFlags(&b->ev[0], hipEventDefault) == hipSuccess &&
              hipEventCreateWithFlags(&b->ev[1], hipEventDefault) == hipSuccess &&
              hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess;
    b->scanned.assign(n_chunks, nullptr);
    for (int c = 0; c < n_chunks && ok; ++c)
        ok = hipEventCreateWithFlags(&b->scanned[c], hipEventDisableTiming) == hipSuccess;
    auto alloc = [&](void **ptr, int64_t bytes) {
        bytes = std::max<int64_t>(bytes, 8);
        if (ok && hipMalloc(ptr, (size_t)bytes) != hipSuccess) ok = false;
        b->bytes += bytes;
    };
    const int64_t rec = n_chunks > 0 ? cap : 1;
    for (int i = 0; i < b->n_streams && ok; ++i) {
        BatchBuffers &B = b->buf[i];
        ok = hipStreamCreateWithFlags(&B.stream, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&B.join, hipEventDisableTiming) == hipSuccess;
        alloc((void **)&B.dda, (int64_t)sizeof(DdaRec) * rec);
        alloc((void **)&B.walls, (int64_t)sizeof(WallRec) * rec);
        alloc((void **)&B.fin, (int64_t)sizeof(uint2) * rec);
        alloc((void **)&B.uniq_recv, (int64_t)sizeof(int32_t) * rec);
        alloc((void **)&B.uniq_cnt, (int64_t)sizeof(int32_t) * rec);
        alloc((void **)&B.n_unique, (int64_t)sizeof(int32_t) * std::max<int64_t>(b->patch_cap, 1));
        alloc((void **)&B.offsets, (int64_t)sizeof(int64_t) * std::max<int64_t>(b->patch_cap, 1));
        alloc((void **)&B.scan_scratch, (int64_t)sizeof(int64_t) * std::max<int64_t>(b->scratch_elems, 1));
    }
    alloc((void **)&b->segs_dev, (int64_t)sizeof(Segment) * std::max(n_pieces, 1));
    alloc((void **)&b->seeds, (int64_t)sizeof(uint64_t) * b->seed_words);
    alloc((void **)&b->counters, (int64_t)sizeof(unsigned long long) * (n_segments + 1));
    alloc((void **)&b->queues, (int64_t)sizeof(unsigned long long) * std::max(n_pieces, 1));
    alloc((void **)&b->marks, (int64_t)sizeof(int64_t) * (n_pieces + 1));
    alloc((void **)&b->e_origin, (int64_t)sizeof(int32_t) * b->edge_cap);
    alloc((void **)&b->e_receiver, (int64_t)sizeof(int32_t) * b->edge_cap);
    alloc((void **)&b->e_count, (int64_t)sizeof(int32_t) * b->edge_cap);
    if (!ok) {
        set_error("prh_trace_batch_create: out of device memory or streams (%lld bytes)", (long long)b->bytes);
        return fail_batch(b, PRH_ERR_HIP);
    }
    // launch shapes: generation wavefronts per piece, and up to the resident
    // traversal workgroups per piece (each piece's own queue balances them)
    const BatchVariant &V = batch_variant_for(b->variant);
    uint32_t coarse_words = 0;
    if (V.occ == OCC_BRICK)
        for (const prh_geometry *g : b->geos) coarse_words = std::max<uint32_t>(coarse_words, (uint32_t)g->g.coarse_words);
    b->lds = (size_t)coarse_words * sizeof(uint32_t);
    int64_t waves = 0;
    if ((rc = resident_waves((const void *)V.fn, STREAM_THREADS, b->lds, &waves))) return fail_batch(b, rc);
    std::vector<Segment> segs(n_pieces);
    for (int c = 0; c < n_chunks; ++c) {
        const BatchBuffers &B = b->buf[c % b->n_streams];
        const int j0 = b->chunk_first[c], j1 = b->chunk_first[c + 1];
        int64_t wave_at = 0, block_at = 0;
        bool pow2 = true;
        for (int j = j0; j < j1; ++j) {
            const Geometry &gj = b->geos[b->pieces[j].seg]->g;
            pow2 = pow2 && is_pow2(gj.ny) && is_pow2(gj.nz);
        }
        b->chunk_pow2.push_back(pow2 ? 1 : 0);
        for (int j = j0; j < j1; ++j) {
            const BatchPiece &P = b->pieces[j];
            const Geometry &g = b->geos[P.seg]->g;
            const int64_t rays = (P.b - P.a) * b->R[P.seg] * b->S[P.seg];
            Segment &sg = segs[j];
            sg.gen = GenArgs{view_of(g), traversal_const(g), (uint32_t)P.a, (uint32_t)(P.b - P.a),
                             (uint32_t)b->R[P.seg], (uint32_t)b->R[P.seg], 0u, b->seeds + b->seed_off[P.seg],
                             (uint32_t)rays,
                             (uint32_t)P.ray_off, nullptr, nullptr, B.dda, B.walls};
            sg.gen_wave_begin = (uint32_t)wave_at;
            sg.reads = b->counters + P.seg;
            sg.queue = b->queues + j;
            sg.grain = (uint32_t)g_trace_grain;
            wave_at += (rays + GEN_UNIT - 1) / GEN_UNIT;
            sg.trace_block_begin = (uint32_t)block_at;
            block_at += trace_blocks(waves, rays);
        }
        b->chunk_waves.push_back(wave_at);
        b->chunk_blocks.push_back(block_at);
    }
    if (n_pieces > 0)
        PRH_HIP(hipMemcpy(b->segs_dev, segs.data(), sizeof(Segment) * n_pieces, hipMemcpyHostToDevice));
    b->piece_start.assign(n_pieces + 1, 0);
    b->seg_nnz.assign(n_segments, 0);
    if (b->use_graph) {
        // capture every part once, each on the stream it will be replayed on
        for (int part = 0; part < 1 + 2 * n_chunks; ++part) {
            hipStream_t st = part == 0 ? b->buf[0].stream : b->buf[((part - 1) / 2) % b->n_streams].stream;
            if ((rc = capture_part(b, part, st))) return fail_batch(b, rc);
        }
    }
    *out = b;
    return PRH_OK;
}

int prh_trace_batch_run(prh_trace_batch *b, const uint64_t *state_inc, int64_t *nnz, int64_t *occupancy_reads,
                        double *elapsed_ms)
{
    PRH_CHECK_ARG(b && state_inc, "prh_trace_batch_run: null pointer");
    PRH_HIP(hipSetDevice(b->device));
    const int n = b->n_seg;
    const int n_pieces = (int)b->pieces.size();
    hipStream_t s0 = b->buf[0].stream;
    std::vector<unsigned long long> ctr(n + 1);
    PRH_HIP(hipMemcpyAsync(b->seeds, state_inc, sizeof(uint64_t) * b->seed_words, hipMemcpyHostToDevice, s0));
    PRH_HIP(hipEventRecord(b->ev[0], s0));
    int rc = enqueue_pass(b);
    if (rc) return rc;
    PRH_HIP(hipEventRecord(b->ev[1], s0));
    PRH_HIP(hipMemcpyAsync(ctr.data(), b->counters, sizeof(unsigned long long) * (n + 1), hipMemcpyDeviceToHost, s0));
    PRH_HIP(hipMemcpyAsync(b->piece_start.data(), b->marks, sizeof(int64_t) * (n_pieces + 1), hipMemcpyDeviceToHost,
                           s0));
    PRH_HIP(hipStreamSynchronize(s0));
    if (elapsed_ms) {
        float ms = 0.f;
        PRH_HIP(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
        *elapsed_ms = ms;
    }
    // marks: [0] edges written, [1 + j] first edge of piece j
    b->edges = b->piece_start[0];
    std::fill(b->seg_nnz.begin(), b->seg_nnz.end(), 0);
    for (int j = 0; j < n_pieces; ++j) {
        const int64_t next = j + 1 < n_pieces ? b->piece_start[j + 2] : b->edges;
        b->seg_nnz[b->pieces[j].seg] += next - b->piece_start[j + 1];
    }
    for (int k = 0; k < n; ++k) {
        if (nnz) nnz[k] = b->seg_nnz[k];
        if (occupancy_reads) occupancy_reads[k] = (int64_t)ctr[k];
    }
    if (ctr[n]) {
        set_error("%llu rays did not resolve a receiving patch", ctr[n]);
        return PRH_ERR_UNRESOLVED;
    }
    if (b->edges > b->edge_cap) {
        set_error("prh_trace_batch_run: %lld edges exceed the edge capacity %lld", (long long)b->edges,
                  (long long)b->edge_cap);
        return PRH_ERR_CAPACITY;
    }
    return PRH_OK;
}

int prh_trace_batch_edges(const prh_trace_batch *b, int32_t segment, int32_t *origin, int32_t *receiver,
                          int64_t *count)
{
    PRH_CHECK_ARG(b && segment >= 0 && segment < b->n_seg, "prh_trace_batch_edges: bad segment");
    PRH_CHECK_ARG(b->edges <= b->edge_cap, "prh_trace_batch_edges: the last run overflowed the edge arena");
    PRH_HIP(hipSetDevice(b->device));
    const int n_pieces = (int)b->pieces.size();
    int64_t at = 0;
    std::vector<int32_t> tmp;
    for (int j = 0; j < n_pieces; ++j) {
        if (b->pieces[j].seg != segment) continue;
        const int64_t first = b->piece_start[j + 1];
        const int64_t m = (j + 1 < n_pieces ? b->piece_start[j + 2] : b->edges) - first;
        if (m <= 0) continue;
        if (origin) PRH_HIP(hipMemcpy(origin + at, b->e_origin + first, m * sizeof(int32_t), hipMemcpyDeviceToHost));
        if (receiver)
            PRH_HIP(hipMemcpy(receiver + at, b->e_receiver + first, m * sizeof(int32_t), hipMemcpyDeviceToHost));
        if (count) {
            tmp.resize(m);
            PRH_HIP(hipMemcpy(tmp.data(), b->e_count + first, m * sizeof(int32_t), hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < m; ++i) count[at + i] = tmp[i];
        }
        at += m;
    }
    return PRH_OK;
}

int prh_trace_batch_edges_device(const prh_trace_batch *b, int32_t segment, prh_edges **out)
{
    PRH_CHECK_ARG(b && out && segment >= 0 && segment < b->n_seg, "prh_trace_batch_edges_device: bad argument");
    PRH_CHECK_ARG(b->edges <= b->edge_cap, "prh_trace_batch_edges_device: the last run overflowed the edge arena");
    *out = nullptr;
    PRH_HIP(hipSetDevice(b->device));
    const int n_pieces = (int)b->pieces.size();
    const int64_t nnz = b->seg_nnz[segment];
    const int64_t alloc = nnz > 0 ? nnz : 1;
    int32_t *o = nullptr, *r = nullptr, *c = nullptr;
    PRH_HIP(hipMalloc((void **)&o, alloc * sizeof(int32_t)));
    if (hipMalloc((void **)&r, alloc * sizeof(int32_t)) != hipSuccess ||
        hipMalloc((void **)&c, alloc * sizeof(int32_t)) != hipSuccess) {
        (void)hipFree(o);
        if (r) (void)hipFree(r);
        set_error("prh_trace_batch_edges_device: out of device memory");
        return PRH_ERR_HIP;
    }
    hipStream_t st = b->buf[0].stream;
    int64_t at = 0;
    for (int j = 0; j < n_pieces; ++j) {
        if (b->pieces[j].seg != segment) continue;
        const int64_t first = b->piece_start[j + 1];
        const int64_t m = (j + 1 < n_pieces ? b->piece_start[j + 2] : b->edges) - first;
        if (m <= 0) continue;
        PRH_HIP(hipMemcpyAsync(o + at, b->e_origin + first, m * 4, hipMemcpyDeviceToDevice, st));
        PRH_HIP(hipMemcpyAsync(r + at, b->e_receiver + first, m * 4, hipMemcpyDeviceToDevice, st));
        PRH_HIP(hipMemcpyAsync(c + at, b->e_count + first, m * 4, hipMemcpyDeviceToDevice, st));
        at += m;
    }
    PRH_HIP(hipStreamSynchronize(st));
    *out = new prh_edges{b->device, nnz, o, r, c};
    return PRH_OK;
}

int prh_trace_batch_info(const prh_trace_batch *b, int64_t *info, int32_t n)
{
    PRH_CHECK_ARG(b && info && n >= 0, "prh_trace_batch_info: null pointer");
    const int64_t v[] = {b->bytes,    (int64_t)b->chunk_waves.size(), (int64_t)b->pieces.size(),
                         b->ray_cap,  b->edge_cap,                    b->rays,
                         b->n_streams, b->use_graph ? 1 : 0,          b->variant};
    const int m = (int)(sizeof(v) / sizeof(v[0]));
    for (int i = 0; i < n && i < m; ++i) info[i] = v[i];
    return PRH_OK;
}

void prh_trace_batch_destroy(prh_trace_batch *b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    for (const BatchBuffers &B : b->buf)
        if (B.stream) (void)hipStreamSynchronize(B.stream);
    free_trace_batch(b);
    delete b;
}

int prh_edges_copy(const prh_edges *e, int32_t *origin, int32_t *receiver, int64_t *count)
{
    PRH_CHECK_ARG(e, "prh_edges_copy: null edge table");
    if (e->nnz == 0) return PRH_OK;
    PRH_HIP(hipSetDevice(e->device));
    if (origin) PRH_HIP(hipMemcpy(origin, e->origin, e->nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (receiver) PRH_HIP(hipMemcpy(receiver, e->receiver, e->nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (count) {
        // Widen the int32 device counts into the int64 host array in place:
        // copy into the upper half of the output buffer, then expand upwards
        // (element k is read before the write of element k can reach it).
        int32_t *tmp = reinterpret_cast<int32_t *>(count) + e->nnz;
        PRH_HIP(hipMemcpy(tmp, e->count, e->nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
        for (int64_t k = 0; k < e->nnz; ++k) count[k] = (int64_t)tmp[k];
    }
    return PRH_OK;
}

int prh_edges_equal(const prh_edges *a, const prh_edges *b, int32_t *equal)
{
    PRH_CHECK_ARG(a && b && equal, "prh_edges_equal: null pointer");
    PRH_CHECK_ARG(a->device == b->device, "prh_edges_equal: tables on different devices");
    *equal = 0;
    if (a->nnz != b->nnz) return PRH_OK;
    if (a->nnz == 0) {
        *equal = 1;
        return PRH_OK;
    }
    PRH_HIP(hipSetDevice(a->device));
    hipStream_t st;
    int rc = current_stream(&st);
    if (rc) return rc;
    unsigned int *flag = nullptr;
    PRH_HIP(hipMalloc((void **)&flag, sizeof(unsigned int)));
    unsigned int host = 1;
    hipError_t e = hipMemsetAsync(flag, 0, sizeof(unsigned int), st);
    if (e == hipSuccess) {
        const int64_t blocks = std::min<int64_t>((a->nnz + 255) / 256, 8192);
        k_edges_differ<<<(unsigned)blocks, 256, 0, st>>>(a->origin, a->receiver, a->count, b->origin, b->receiver,
                                                         b->count, a->nnz, flag);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&host, flag, sizeof(unsigned int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(flag);
    PRH_HIP(e);
    *equal = host == 0 ? 1 : 0;
    return PRH_OK;
}

void prh_edges_destroy(prh_edges *e)
{
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->block) {
        release_edge_block(e->device, e->block, e->block_bytes);
    } else {
        if (e->origin) (void)hipFree(e->origin);
        if (e->receiver) (void)hipFree(e->receiver);
        if (e->count) (void)hipFree(e->count);
    }
    delete e;
}

}  // extern "C"
