// Internal definitions shared by the porerad HIP library sources.
// Target: gfx950 (MI355X), wave64.  All FP64 arithmetic on the tracer path is
// compiled with -ffp-contract=off so every product and sum rounds on its own,
// exactly as NumPy and Numba evaluate porerad/view_factors.py.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "porerad_hip.h"

// A device-resident edge table (origin, receiver, count), sorted by
// (origin, receiver) with unique pairs: the output of the tracer's counting.
struct prh_edges {
    int device;
    int64_t nnz;
    int32_t *origin;
    int32_t *receiver;
    int32_t *count;
    // the one allocation holding the three arrays (edge tables of the
    // tracer's counting), or null when each array is an allocation of its own
    void *block;
    size_t block_bytes;
};

namespace prh {

// ------------------------------------------------------------------ errors
void set_error(const char *fmt, ...);

#define PRH_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            ::prh::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),        \
                             __FILE__, __LINE__);                                          \
            return PRH_ERR_HIP;                                                            \
        }                                                                                  \
    } while (0)

#define PRH_CHECK_ARG(cond, ...)                                                           \
    do {                                                                                   \
        if (!(cond)) {                                                                     \
            ::prh::set_error(__VA_ARGS__);                                                 \
            return PRH_ERR_ARG;                                                            \
        }                                                                                  \
    } while (0)

// ------------------------------------------------------------------ 128-bit LCG
// PCG64 (XSL-RR 128/64) state arithmetic on two 64-bit halves.
struct u128 {
    uint64_t lo, hi;
};

__host__ __device__ inline u128 make_u128(uint64_t hi, uint64_t lo) { return u128{lo, hi}; }

// a * b mod 2^128.  The full product of the low halves is formed from four
// 32 x 32 -> 64 products (one v_mad_u64_u32 each, every partial sum below
// 2^64), the cross terms only mod 2^64.
__device__ inline u128 mul128(u128 a, u128 b)
{
    const uint32_t a0 = (uint32_t)a.lo, a1 = (uint32_t)(a.lo >> 32);
    const uint32_t b0 = (uint32_t)b.lo, b1 = (uint32_t)(b.lo >> 32);
    const uint64_t p00 = (uint64_t)a0 * b0;
    const uint64_t t = (uint64_t)a0 * b1 + (p00 >> 32);
    const uint64_t u = (uint64_t)a1 * b0 + (uint32_t)t;
    u128 r;
    r.lo = u << 32 | (uint32_t)p00;
    r.hi = (uint64_t)a1 * b1 + (t >> 32) + (u >> 32) + a.lo * b.hi + a.hi * b.lo;
    return r;
}

__device__ inline u128 add128(u128 a, u128 b)
{
    u128 r;
    const bool carry = __builtin_add_overflow(a.lo, b.lo, &r.lo);
    r.hi = a.hi + b.hi + (carry ? 1u : 0u);
    return r;
}

// Multiplier of the 128-bit LCG under PCG64.
constexpr uint64_t PCG_MULT_HI = 0x2360ED051FC65DA4ULL;
constexpr uint64_t PCG_MULT_LO = 0x4385DF649FCCF645ULL;

// Jump tables: advancing the LCG by 2^b steps maps s -> A[b] s + G[b] inc.
struct JumpTables {
    u128 A[64];
    u128 G[64];
};

// Host-side 128-bit helpers (built with the host compiler's __int128).
void host_jump_tables(JumpTables *t);
void host_seed_words(const uint32_t *words, int n_words, uint64_t out[4]);

// ------------------------------------------------------------------ geometry
// Device-resident patch table of one x-oriented microstructure.
//   occ        occupancy bitset: bit c = cell (i*ny + j)*nz + k (word c >> 5, bit c & 31)
//   coarse     brick bitset: bit b = brick ((i>>s)*cny + (j>>s))*cnz + (k>>s) is set when
//              the 2^s-cell brick holds any solid cell; small enough to sit in LDS
//   cell_code  per cell: (first patch id of the cell << 6) | exposed-face mask
//   patch_*    per patch: origin (x,y,z) FP64, face 0..5 (walls: 0 hot, 1 cold)
//   hot_map / cold_map  wall patch id per (j, k), -1 where the boundary cell is solid
//   wocc       walled occupancy bitset in 4x4x4 blocks, one 64-bit word per block (word
//              P >> 9, at byte offset P >> 6) holding the cell (i + 1, j, k), i in [-1, nx],
//              at bit 63 - (P & 63), where P is the Morton index of (i + 1, j, k)
//              (lean_layout; bits 0-5 of P are the cell within the block, bits 6-8 are
//              zero); the planes i = -1 and i = nx (the hot and cold walls) are set, so
//              one bit test ends a ray at a solid cell or a wall (null when P would need
//              more than 31 bits or ny, nz < 3)
struct Geometry {
    int32_t nx, ny, nz;
    int64_t occ_words;
    int32_t coarse_shift, cnx, cny, cnz;
    int64_t coarse_words;
    uint32_t *coarse;
    double H_m;
    int64_t n_patches, n_wall;
    int device;
    uint32_t *occ;
    uint32_t *cell_code;
    double *patch_origin;  // n_patches * 3
    int8_t *patch_face;    // n_patches
    int32_t *patch_cell;   // n_patches (-1 for walls)
    int8_t *patch_class;   // n_patches
    int32_t *hot_map, *cold_map;
    uint64_t *wocc;
    int64_t wocc_words;
    int32_t wlx, wly, wlz;  // bit widths of nx + 2, ny, nz (each at least 2)
    // cell_code in walled-bitset order: entry (P >> 9) << 6 | (P & 63) is the
    // code of the cell whose index is P (0 for wall and padding cells), so a
    // final state holding P finds its patch with one load; present exactly
    // when wocc is (at most 2^28 entries, 1 GiB)
    uint32_t *wcode;
};

// Bit layout of the walled bitset's index P (Geometry::wocc): a Morton
// order.  The low m bits of each axis's value (m = the smallest of the three
// widths, at least 2 and at most 10) are interleaved as z, y, x, z, y, x, ...
// so bits 0-5 (the bit within a block's word) hold the block's 4x4x4 cells;
// the interleaved bits above those start at bit 9, leaving bits 6-8 zero, so
// P >> 6 is the byte offset of the block's word and the word index is P >> 9;
// a 128 B line holds 8x8x16 cells.  The bits above m of each axis follow in
// fields (z, then y, then x).  The set of an axis's bits is its mask; P is
// the OR of the three dilated values.  The gap bits belong to no axis, so
// adding or taking one from an axis's value (subtracting its mask or its
// mask's lowest bit, then masking) carries or borrows across them.
constexpr uint32_t LEAN_GAP = 3;  // zero bits between the cell within the block and the word index
struct LeanLayout {
    uint32_t mask[3];   // x, y, z
    uint32_t field[3];  // first bit of each axis's field
    uint32_t m;         // interleaved bits per axis
};

__host__ __device__ inline uint32_t morton_spread(uint32_t v)  // bit b -> bit 3b, b < 10
{
    v &= 0x000003ffu;
    v = (v ^ (v << 16)) & 0xff0000ffu;
    v = (v ^ (v << 8)) & 0x0300f00fu;
    v = (v ^ (v << 4)) & 0x030c30c3u;
    v = (v ^ (v << 2)) & 0x09249249u;
    return v;
}

__host__ __device__ inline uint32_t morton_gather(uint32_t v)  // bit 3b -> bit b
{
    v &= 0x09249249u;
    v = (v ^ (v >> 2)) & 0x030c30c3u;
    v = (v ^ (v >> 4)) & 0x0300f00fu;
    v = (v ^ (v >> 8)) & 0xff0000ffu;
    v = (v ^ (v >> 16)) & 0x000003ffu;
    return v;
}

// Interleaved bits 0-5 stay; the ones above move up past the gap.
__host__ __device__ inline uint32_t lean_open_gap(uint32_t v) { return (v & 63u) | (v & ~63u) << LEAN_GAP; }

__host__ __device__ inline LeanLayout lean_layout(int lx, int ly, int lz)
{
    LeanLayout L;
    int m = lx < ly ? lx : ly;
    m = m < lz ? m : lz;
    m = m < 10 ? m : 10;
    L.m = (uint32_t)m;
    L.field[2] = 3u * L.m + LEAN_GAP;
    L.field[1] = L.field[2] + (uint32_t)(lz - m);
    L.field[0] = L.field[1] + (uint32_t)(ly - m);
    const int bits[3] = {lx, ly, lz};
    for (int a = 0; a < 3; ++a)
        L.mask[a] = lean_open_gap(morton_spread((1u << L.m) - 1u) << (2 - a)) |
                    (((1u << (bits[a] - m)) - 1u) << L.field[a]);
    return L;
}

// Value v of axis a (0 x, 1 y, 2 z) spread to its bits of P, and back.
__host__ __device__ inline uint32_t lean_dilate(const LeanLayout &L, int a, uint32_t v)
{
    return lean_open_gap(morton_spread(v & ((1u << L.m) - 1u)) << (2 - a)) | (v >> L.m) << L.field[a];
}

// The same with the interleaved part read from a table of
// lean_open_gap(morton_spread(v)), v < 1024 (LEAN_SPREAD_ENTRIES): opening
// the gap commutes with the axis's shift of 0-2 bits.
constexpr int LEAN_SPREAD_ENTRIES = 1024;
__host__ __device__ inline uint32_t lean_dilate_table(const LeanLayout &L, int a, uint32_t v, const uint32_t *spread)
{
    return spread[v & ((1u << L.m) - 1u)] << (2 - a) | (v >> L.m) << L.field[a];
}

__host__ __device__ inline uint32_t lean_compact(const LeanLayout &L, int a, uint32_t d)
{
    const uint32_t inter = lean_open_gap(morton_spread((1u << L.m) - 1u) << (2 - a));
    const uint32_t low = d & inter;  // the interleaved bits, gap opened
    return morton_gather(((low & 63u) | (low & ~63u) >> LEAN_GAP) >> (2 - a)) |
           ((d & L.mask[a] & ~inter) >> L.field[a]) << L.m;
}

// Device exclusive scan helpers (prh_scan.hip).
int exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, int64_t *total, hipStream_t stream);
int exclusive_scan_i32_to_i64(const int32_t *in, int64_t *out, int64_t n, int64_t *total, hipStream_t stream);
// Stream-ordered form without host synchronisation or allocation (usable
// under graph capture) that continues a running device total: out[i] =
// *carry + sum(in[0..i)), then *start = the old *carry and *carry += sum(in).
// scratch holds scan_scratch_elems(n) values.
int64_t scan_scratch_elems(int64_t n);
int exclusive_scan_i32_to_i64_carry_async(const int32_t *in, int64_t *out, int64_t n, int64_t *scratch,
                                          int64_t *carry, int64_t *start, hipStream_t stream);

// pool_counts of device edge tables (each sorted, unique pairs) into a new
// table: keys sorted, equal pairs' counts summed (prh_assemble.hip).
int pool_edge_tables(const prh_edges *const *parts, int n_parts, int64_t n_patches, hipStream_t st, prh_edges **out);

// Persistent workspace helpers.
int ensure_capacity(void **buf, size_t *cap, size_t bytes);

int current_stream(hipStream_t *s);

}  // namespace prh

struct prh_geometry {
    prh::Geometry g;
};

// A device copy of a uint8 phase array (nx, ny, nz), x-major, 1 = solid, 0 = pore.
struct prh_phase_field {
    int device;
    int32_t nx, ny, nz;
    int64_t cells;
    uint8_t *phases;
};
