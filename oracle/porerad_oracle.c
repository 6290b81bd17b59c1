/*
 * CPU oracle of the porerad radiation hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C99 + OpenMP, the reference algorithm of
 * porerad/view_factors.py (patch table, seeded Lambertian rays, FP64
 * Amanatides-Woo traversal, sorted integer first-hit counts) and the FP64
 * reflection series of porerad/exchange.py.  It is the checker for the HIP
 * path: only tests, the smoke check and the CPU-baseline leg of the benchmark
 * load it.  The product library never links or calls it.
 *
 * Third-party arithmetic restated here (published algorithms, numpy 2.2.6):
 *   - numpy.random.SeedSequence (bit_generator.pyx: hashmix/mix, pool of 4
 *     uint32 words, generate_state) -> PCG64 setseq-128 seeding
 *     (pcg64.h: pcg_setseq_128_srandom_r), XSL-RR 128/64 output, LCG
 *     jump-ahead (pcg_advance_lcg_128) and Generator.random()
 *     = (next64 >> 11) * 2^-53.
 *   - np.remainder for float64 (npy_divmod: fmod, then add the divisor when
 *     the signs differ, +0.0 for an exact zero).
 *   - cos/sin/sqrt from glibc libm (what np.cos/np.sin/np.sqrt call on the
 *     reference's platform); the build uses -ffp-contract=off so every
 *     product and sum is rounded separately, as in NumPy and Numba.
 *
 * Pinning: the 24 archived raw_count_sha256_A/B hashes of
 * data/cases_coupled.csv (tests/test_oracle_hashes.py), numpy's own PCG64 /
 * SeedSequence outputs (tests/test_oracle_rng.py), and brute-force exact
 * ray/plane intersection on tiny grids (tests/test_oracle_tracer.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------ */
/* numpy SeedSequence -> PCG64 (numpy/random/bit_generator.pyx,       */
/* numpy/random/src/pcg64/pcg64.h)                                     */
/* ------------------------------------------------------------------ */
#define SS_INIT_A 0x43b0d7e5u
#define SS_MULT_A 0x931e8875u
#define SS_INIT_B 0x8b51f9ddu
#define SS_MULT_B 0x58f38dedu
#define SS_MIX_MULT_L 0xca01f9ddu
#define SS_MIX_MULT_R 0x4973f715u
#define SS_XSHIFT 16
#define SS_POOL 4

static uint32_t ss_hashmix(uint32_t value, uint32_t *hash_const)
{
    value ^= *hash_const;
    *hash_const *= SS_MULT_A;
    value *= *hash_const;
    value ^= value >> SS_XSHIFT;
    return value;
}

static uint32_t ss_mix(uint32_t x, uint32_t y)
{
    uint32_t r = SS_MIX_MULT_L * x - SS_MIX_MULT_R * y;
    r ^= r >> SS_XSHIFT;
    return r;
}

static const u128 PCG_MULT =
    (((u128)0x2360ED051FC65DA4ULL) << 64) | (u128)0x4385DF649FCCF645ULL;

static u128 u128_of(uint64_t hi, uint64_t lo) { return (((u128)hi) << 64) | (u128)lo; }

/* SeedSequence(entropy words).generate_state(4, uint64) -> PCG64 seeding.
 * out = {state_hi, state_lo, inc_hi, inc_lo}. */
int orc_seed_words(const uint32_t *words, int n_words, uint64_t out[4])
{
    uint32_t pool[SS_POOL];
    uint32_t hc = SS_INIT_A;
    for (int i = 0; i < SS_POOL; ++i)
        pool[i] = ss_hashmix(i < n_words ? words[i] : 0u, &hc);
    for (int s = 0; s < SS_POOL; ++s)
        for (int d = 0; d < SS_POOL; ++d)
            if (s != d) pool[d] = ss_mix(pool[d], ss_hashmix(pool[s], &hc));
    for (int s = SS_POOL; s < n_words; ++s)
        for (int d = 0; d < SS_POOL; ++d)
            pool[d] = ss_mix(pool[d], ss_hashmix(words[s], &hc));
    uint32_t state32[8];
    uint32_t hb = SS_INIT_B;
    for (int i = 0; i < 8; ++i) {
        uint32_t v = pool[i % SS_POOL];
        v ^= hb;
        hb *= SS_MULT_B;
        v *= hb;
        v ^= v >> SS_XSHIFT;
        state32[i] = v;
    }
    uint64_t w[4];
    for (int i = 0; i < 4; ++i) w[i] = (uint64_t)state32[2 * i] | ((uint64_t)state32[2 * i + 1] << 32);
    u128 initstate = u128_of(w[0], w[1]);
    u128 initseq = u128_of(w[2], w[3]);
    u128 inc = (initseq << 1) | 1u;
    u128 state = 0;
    state = state * PCG_MULT + inc;
    state += initstate;
    state = state * PCG_MULT + inc;
    out[0] = (uint64_t)(state >> 64);
    out[1] = (uint64_t)state;
    out[2] = (uint64_t)(inc >> 64);
    out[3] = (uint64_t)inc;
    return 0;
}

static u128 pcg_advance(u128 state, u128 delta, u128 inc)
{
    u128 acc_mult = 1, acc_plus = 0, cur_mult = PCG_MULT, cur_plus = inc;
    while (delta > 0) {
        if (delta & 1) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1) * cur_plus;
        cur_mult *= cur_mult;
        delta >>= 1;
    }
    return acc_mult * state + acc_plus;
}

static inline uint64_t pcg_next64(u128 *state, u128 inc)
{
    *state = *state * PCG_MULT + inc;
    uint64_t hi = (uint64_t)(*state >> 64), lo = (uint64_t)*state;
    unsigned rot = (unsigned)(*state >> 122);
    uint64_t x = hi ^ lo;
    return (x >> rot) | (x << ((64u - rot) & 63u));
}

static inline double pcg_uniform(u128 *state, u128 inc)
{
    return (double)(pcg_next64(state, inc) >> 11) * (1.0 / 9007199254740992.0);
}

void orc_pcg64_advance(uint64_t st[4], uint64_t delta_hi, uint64_t delta_lo)
{
    u128 s = pcg_advance(u128_of(st[0], st[1]), u128_of(delta_hi, delta_lo), u128_of(st[2], st[3]));
    st[0] = (uint64_t)(s >> 64);
    st[1] = (uint64_t)s;
}

void orc_pcg64_random(uint64_t st[4], int64_t count, double *out)
{
    u128 s = u128_of(st[0], st[1]), inc = u128_of(st[2], st[3]);
    for (int64_t i = 0; i < count; ++i) out[i] = pcg_uniform(&s, inc);
    st[0] = (uint64_t)(s >> 64);
    st[1] = (uint64_t)s;
}

void orc_pcg64_raw(uint64_t st[4], int64_t count, uint64_t *out)
{
    u128 s = u128_of(st[0], st[1]), inc = u128_of(st[2], st[3]);
    for (int64_t i = 0; i < count; ++i) out[i] = pcg_next64(&s, inc);
    st[0] = (uint64_t)(s >> 64);
    st[1] = (uint64_t)s;
}

/* ------------------------------------------------------------------ */
/* Patch table (porerad/view_factors.py:70-132)                        */
/* ------------------------------------------------------------------ */

/* np.remainder for float64 */
static inline double py_fmod(double a, double b)
{
    double m = fmod(a, b);
    if (m != 0.0) {
        if ((b < 0.0) != (m < 0.0)) m += b;
    } else {
        m = copysign(0.0, b);
    }
    return m;
}

static inline int64_t py_imod(int64_t a, int64_t b)
{
    int64_t m = a % b;
    if (m != 0 && ((m < 0) != (b < 0))) m += b;
    return m;
}

static inline int exposed_face(const uint8_t *s, int64_t nx, int64_t ny, int64_t nz,
                               int64_t i, int64_t j, int64_t k, int f)
{
    if (!s[(i * ny + j) * nz + k]) return 0;
    switch (f) {
    case 0: return i < nx - 1 && !s[((i + 1) * ny + j) * nz + k];
    case 1: return i > 0 && !s[((i - 1) * ny + j) * nz + k];
    case 2: return !s[(i * ny + (j + 1) % ny) * nz + k];
    case 3: return !s[(i * ny + (j + ny - 1) % ny) * nz + k];
    case 4: return !s[(i * ny + j) * nz + (k + 1) % nz];
    default: return !s[(i * ny + j) * nz + (k + nz - 1) % nz];
    }
}

/* Count patches: returns n_wall in *n_wall_out and the total. */
int64_t orc_count_patches(const uint8_t *solid, int64_t nx, int64_t ny, int64_t nz, int64_t *n_wall_out)
{
    int64_t n_wall = 0, n_int = 0;
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t k = 0; k < nz; ++k) {
            n_wall += !solid[(0 * ny + j) * nz + k];
            n_wall += !solid[((nx - 1) * ny + j) * nz + k];
        }
    for (int64_t i = 0; i < nx; ++i)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t k = 0; k < nz; ++k)
                for (int f = 0; f < 6; ++f) n_int += exposed_face(solid, nx, ny, nz, i, j, k, f);
    *n_wall_out = n_wall;
    return n_wall + n_int;
}

static const double FACE_NORMALS[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

/* Fill the patch table; arrays sized from orc_count_patches. */
void orc_build_patches(const uint8_t *solid, int64_t nx, int64_t ny, int64_t nz,
                       int8_t *patch_class, int32_t *patch_cell, int8_t *patch_face,
                       double *patch_origin, double *patch_normal, int32_t *cell_face_patch,
                       int32_t *hot_wall_patch, int32_t *cold_wall_patch)
{
    int64_t p = 0;
    for (int64_t c = 0; c < nx * ny * nz * 6; ++c) cell_face_patch[c] = -1;
    for (int64_t c = 0; c < ny * nz; ++c) hot_wall_patch[c] = cold_wall_patch[c] = -1;
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t k = 0; k < nz; ++k)
            for (int cls = 0; cls < 2; ++cls) {
                int64_t i = cls == 0 ? 0 : nx - 1;
                if (solid[(i * ny + j) * nz + k]) continue;
                patch_class[p] = (int8_t)cls;
                patch_cell[p] = -1;
                patch_face[p] = -1;
                patch_origin[3 * p + 0] = cls == 0 ? 1.0e-9 : (double)nx - 1.0e-9;
                patch_origin[3 * p + 1] = (double)j + 0.5;
                patch_origin[3 * p + 2] = (double)k + 0.5;
                patch_normal[3 * p + 0] = cls == 0 ? 1.0 : -1.0;
                patch_normal[3 * p + 1] = 0.0;
                patch_normal[3 * p + 2] = 0.0;
                if (cls == 0) hot_wall_patch[j * nz + k] = (int32_t)p;
                else cold_wall_patch[j * nz + k] = (int32_t)p;
                ++p;
            }
    const double offset = 0.5 + 1.0e-9;
    for (int64_t i = 0; i < nx; ++i)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t k = 0; k < nz; ++k)
                for (int f = 0; f < 6; ++f) {
                    if (!exposed_face(solid, nx, ny, nz, i, j, k, f)) continue;
                    int64_t cell = (i * ny + j) * nz + k;
                    patch_class[p] = 2;
                    patch_cell[p] = (int32_t)cell;
                    patch_face[p] = (int8_t)f;
                    double o[3] = {(double)i + 0.5, (double)j + 0.5, (double)k + 0.5};
                    for (int a = 0; a < 3; ++a) {
                        if (FACE_NORMALS[f][a] > 0.0) o[a] += offset;
                        else if (FACE_NORMALS[f][a] < 0.0) o[a] -= offset;
                        patch_normal[3 * p + a] = FACE_NORMALS[f][a];
                    }
                    o[1] = py_fmod(o[1], (double)ny);
                    o[2] = py_fmod(o[2], (double)nz);
                    patch_origin[3 * p + 0] = o[0];
                    patch_origin[3 * p + 1] = o[1];
                    patch_origin[3 * p + 2] = o[2];
                    cell_face_patch[cell * 6 + f] = (int32_t)p;
                    ++p;
                }
}

/* ------------------------------------------------------------------ */
/* Ray generation (porerad/view_factors.py:135-182)                    */
/* ------------------------------------------------------------------ */
#define FACE_EDGE_INSET 1.0e-12
#define TWO_PI (2.0 * 3.141592653589793)

static inline double clip01(double u)
{
    const double lo = FACE_EDGE_INSET, hi = 1.0 - FACE_EDGE_INSET;
    return u < lo ? lo : (u > hi ? hi : u);
}

/* One ray from four uniforms.  normal: the patch normal (one nonzero +-1). */
static inline void make_ray(const double po[3], const double pn[3], const double u[4],
                            double ny, double nz, double o[3], double d[3])
{
    double ta = clip01(u[0]) - 0.5;
    double tb = clip01(u[1]) - 0.5;
    o[0] = po[0];
    o[1] = po[1];
    o[2] = po[2];
    int axis;
    if (pn[0] != 0.0) {
        axis = 0;
        o[1] += ta;
        o[2] += tb;
    } else if (pn[1] != 0.0) {
        axis = 1;
        o[0] += ta;
        o[2] += tb;
    } else {
        axis = 2;
        o[0] += ta;
        o[1] += tb;
    }
    o[1] = py_fmod(o[1], ny);
    o[2] = py_fmod(o[2], nz);
    double radius = sqrt(u[2]);
    double phi = TWO_PI * u[3];
    double la = radius * cos(phi);
    double lb = radius * sin(phi);
    double ct = sqrt(1.0 - u[2]);
    double sign = pn[axis];
    static const int FIRST[3] = {1, 0, 0}, SECOND[3] = {2, 2, 1};
    d[axis] = sign * ct;
    d[FIRST[axis]] = la;
    d[SECOND[axis]] = sign * lb;
    double norm = sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    d[0] /= norm;
    d[1] /= norm;
    d[2] /= norm;
}

/* Origins/directions (p1-p0)*R rays for the patch range, as generate_rays. */
void orc_generate_rays(const double *patch_origin, const double *patch_normal, int64_t ny, int64_t nz,
                       int64_t p0, int64_t p1, int64_t rays, const uint64_t st0[4],
                       double *origins, double *directions)
{
    u128 inc = u128_of(st0[2], st0[3]);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t p = p0; p < p1; ++p) {
        u128 s = pcg_advance(u128_of(st0[0], st0[1]), (u128)p * (u128)rays * 4u, inc);
        for (int64_t r = 0; r < rays; ++r) {
            double u[4];
            for (int c = 0; c < 4; ++c) u[c] = pcg_uniform(&s, inc);
            int64_t m = (p - p0) * rays + r;
            make_ray(patch_origin + 3 * p, patch_normal + 3 * p, u, (double)ny, (double)nz,
                     origins + 3 * m, directions + 3 * m);
        }
    }
}

/* ------------------------------------------------------------------ */
/* FP64 Amanatides-Woo traversal (porerad/view_factors.py:185-258)     */
/* ------------------------------------------------------------------ */
#define TIE_TOLERANCE 1.0e-12
#define INF_T 1.0e300

static inline int64_t initial_cell(double position, double direction)
{
    int64_t index = (int64_t)floor(position);
    if (direction < 0.0 && fabs(position - nearbyint(position)) <= TIE_TOLERANCE) index -= 1;
    return index;
}

static inline void axis_state(double position, double direction, int64_t index, int64_t *step,
                              double *t_max, double *t_delta)
{
    if (direction > 0.0) {
        *step = 1;
        *t_max = ((double)index + 1.0 - position) / direction;
        *t_delta = 1.0 / direction;
    } else if (direction < 0.0) {
        *step = -1;
        *t_max = ((double)index - position) / direction;
        *t_delta = -1.0 / direction;
    } else {
        *step = 0;
        *t_max = INF_T;
        *t_delta = INF_T;
    }
}

/* Returns the receiver (>=0) or -1 on failure; *reads = occupancy reads. */
static int32_t trace_one(const uint8_t *solid, int64_t nx, int64_t ny, int64_t nz,
                         const int32_t *hot_map, const int32_t *cold_map, const int32_t *cell_face_map,
                         const double o[3], const double d[3], int64_t *reads)
{
    double ox = o[0], oy = o[1], oz = o[2];
    double dx = d[0], dy = d[1], dz = d[2];
    *reads = 0;
    if (dx == 0.0) return -1;
    int64_t i = initial_cell(ox, dx), j = initial_cell(oy, dy), k = initial_cell(oz, dz);
    int64_t sx, sy, sz;
    double tx, ty, tz, ddx, ddy, ddz;
    axis_state(ox, dx, i, &sx, &tx, &ddx);
    axis_state(oy, dy, j, &sy, &ty, &ddy);
    axis_state(oz, dz, k, &sz, &tz, &ddz);
    double time = 0.0;
    int face = -1;
    double crossing_scale = 8.0 * (double)nx * (1.0 + (fabs(dy) + fabs(dz)) / fabs(dx)) + 100.0;
    int64_t cap = (int64_t)(crossing_scale < 100000000.0 ? crossing_scale : 100000000.0);
    int64_t max_steps = 10 * nx * ny * nz + 10;
    if (cap > max_steps) max_steps = cap;
    for (int64_t step = 0; step < max_steps; ++step) {
        if (i < 0 || i >= nx) {
            int64_t jj = py_imod((int64_t)floor(oy + dy * time), ny);
            int64_t kk = py_imod((int64_t)floor(oz + dz * time), nz);
            return i < 0 ? hot_map[jj * nz + kk] : cold_map[jj * nz + kk];
        }
        int64_t jj = py_imod(j, ny), kk = py_imod(k, nz);
        int64_t cell = (i * ny + jj) * nz + kk;
        *reads += 1;
        if (solid[cell] != 0) return face >= 0 ? cell_face_map[cell * 6 + face] : -1;
        double next_time = tx;
        if (ty < next_time) next_time = ty;
        if (tz < next_time) next_time = tz;
        int cross_x = fabs(tx - next_time) <= TIE_TOLERANCE;
        int cross_y = fabs(ty - next_time) <= TIE_TOLERANCE;
        int cross_z = fabs(tz - next_time) <= TIE_TOLERANCE;
        if (cross_x) {
            i += sx;
            tx += ddx;
            face = sx > 0 ? 1 : 0;
        }
        if (cross_y) {
            j += sy;
            ty += ddy;
            if (!cross_x) face = sy > 0 ? 3 : 2;
        }
        if (cross_z) {
            k += sz;
            tz += ddz;
            if (!cross_x && !cross_y) face = sz > 0 ? 5 : 4;
        }
        time = next_time;
    }
    return -1;
}

/* trace_first_hit: receiver per ray (-1 = failed).  Returns the failed count;
 * *total_reads = sum of occupancy reads. */
int64_t orc_trace(const uint8_t *solid, int64_t nx, int64_t ny, int64_t nz, const int32_t *hot_map,
                  const int32_t *cold_map, const int32_t *cell_face_map, const double *origins,
                  const double *directions, int64_t n_rays, int32_t *receiver, int64_t *total_reads)
{
    int64_t failed = 0, reads_sum = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : failed, reads_sum)
    for (int64_t r = 0; r < n_rays; ++r) {
        int64_t reads;
        int32_t rec = trace_one(solid, nx, ny, nz, hot_map, cold_map, cell_face_map, origins + 3 * r,
                                directions + 3 * r, &reads);
        receiver[r] = rec;
        failed += rec < 0;
        reads_sum += reads;
    }
    if (total_reads) *total_reads = reads_sum;
    return failed;
}

static int cmp_i32(const void *a, const void *b)
{
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* first_hit_counts for patches [p0, p1): generate + trace + per-patch sorted
 * unique receivers.  The per-patch receivers are written to work[(p-p0)*R ..]
 * sorted, per-patch unique counts to n_unique[p-p0].  Returns the number of
 * failed rays; *total_reads as in orc_trace. */
int64_t orc_first_hit_receivers(const uint8_t *solid, int64_t nx, int64_t ny, int64_t nz,
                                 const int32_t *hot_map, const int32_t *cold_map,
                                 const int32_t *cell_face_map, const double *patch_origin,
                                 const double *patch_normal, int64_t p0, int64_t p1, int64_t rays,
                                 const uint64_t st0[4], int32_t *work, int64_t *n_unique,
                                 int64_t *total_reads)
{
    u128 inc = u128_of(st0[2], st0[3]);
    int64_t failed = 0, reads_sum = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : failed, reads_sum)
    for (int64_t p = p0; p < p1; ++p) {
        u128 s = pcg_advance(u128_of(st0[0], st0[1]), (u128)p * (u128)rays * 4u, inc);
        int32_t *rec = work + (p - p0) * rays;
        for (int64_t r = 0; r < rays; ++r) {
            double u[4], o[3], d[3];
            for (int c = 0; c < 4; ++c) u[c] = pcg_uniform(&s, inc);
            make_ray(patch_origin + 3 * p, patch_normal + 3 * p, u, (double)ny, (double)nz, o, d);
            int64_t reads;
            rec[r] = trace_one(solid, nx, ny, nz, hot_map, cold_map, cell_face_map, o, d, &reads);
            failed += rec[r] < 0;
            reads_sum += reads;
        }
        qsort(rec, (size_t)rays, sizeof(int32_t), cmp_i32);
        int64_t u = 0;
        for (int64_t r = 0; r < rays; ++r)
            if (r == 0 || rec[r] != rec[r - 1]) ++u;
        n_unique[p - p0] = u;
    }
    if (total_reads) *total_reads = reads_sum;
    return failed;
}

/* Compact sorted per-patch receivers into the (origin, receiver, count)
 * edge table; offsets[p-p0] = exclusive prefix of n_unique. */
void orc_compact_edges(const int32_t *work, const int64_t *offsets, int64_t p0, int64_t p1, int64_t rays,
                       int32_t *origin, int32_t *receiver, int64_t *count)
{
#pragma omp parallel for schedule(static)
    for (int64_t p = p0; p < p1; ++p) {
        const int32_t *rec = work + (p - p0) * rays;
        int64_t e = offsets[p - p0] - 1;
        for (int64_t r = 0; r < rays; ++r) {
            if (r == 0 || rec[r] != rec[r - 1]) {
                ++e;
                origin[e] = (int32_t)p;
                receiver[e] = rec[r];
                count[e] = 0;
            }
            count[e] += 1;
        }
    }
}

/* ------------------------------------------------------------------ */
/* FP64 exchange operator application and reflection series            */
/* (porerad/exchange.py:82-122, 217-245).  Edge list sorted by origin;  */
/* the raw products are summed in edge order as np.bincount does.       */
/* ------------------------------------------------------------------ */

/* y = F x (raw_matvec): rows by origin. */
void orc_raw_matvec(const int32_t *origin, const int32_t *receiver, const double *weight, int64_t nnz,
                    int64_t n, const double *x, double *y)
{
    for (int64_t i = 0; i < n; ++i) y[i] = 0.0;
    for (int64_t e = 0; e < nnz; ++e) y[origin[e]] += weight[e] * x[receiver[e]];
}

/* y = F^T x (raw_rmatvec). */
void orc_raw_rmatvec(const int32_t *origin, const int32_t *receiver, const double *weight, int64_t nnz,
                     int64_t n, const double *x, double *y)
{
    for (int64_t i = 0; i < n; ++i) y[i] = 0.0;
    for (int64_t e = 0; e < nnz; ++e) y[receiver[e]] += weight[e] * x[origin[e]];
}

/* Reflection series on a symmetric CSR matrix G = diag(d) S diag(d) with
 * incident = G (flight / area).  This is the same linear map as
 * ExchangeOperator.rmatvec; summation order is row order.  Returns the
 * number of reflections, or -1 if max_reflections were used up.
 * OpenMP over rows, deterministic (each row summed by one thread). */
int64_t orc_reflection_series_csr(const int64_t *rowptr, const int32_t *col, const double *val, int64_t n,
                                  const double *area, const double *emitted, const double *emissivity,
                                  double relative_tolerance, int64_t max_reflections, double *absorbed,
                                  double *flight, double *work, double *residual)
{
    double initial = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        initial += fabs(emitted[i]);
        absorbed[i] = 0.0;
        flight[i] = emitted[i];
    }
    if (initial < 2.2250738585072014e-308) initial = 2.2250738585072014e-308;
    for (int64_t it = 1; it <= max_reflections; ++it) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) work[i] = flight[i] / area[i];
        double l1 = 0.0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : l1)
        for (int64_t i = 0; i < n; ++i) {
            double acc = 0.0;
            for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) acc += val[e] * work[col[e]];
            absorbed[i] += emissivity[i] * acc;
            flight[i] = (1.0 - emissivity[i]) * acc;
            l1 += fabs(flight[i]);
        }
        if (l1 <= relative_tolerance * initial) {
            double s = 0.0;
            for (int64_t i = 0; i < n; ++i) s += flight[i];
            *residual = s;
            return it;
        }
    }
    return -1;
}

/* ------------------------------------------------------------------------ *
 * Multi-bounce Monte Carlo (porerad/properties.py:54-307 trace_properties,
 * porerad/penetration.py:54-305 trace_penetration): the Taichi ray kernels
 * restated in C, one ray at a time, with glibc cos/sin/sqrt.
 * ------------------------------------------------------------------------ */

/* _hash_u32, properties.py:81-88 / penetration.py:58-65 */
static inline uint32_t mc_hash(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

/* _rand01, properties.py:90-93 */
static inline double mc_rand01(uint32_t *state)
{
    *state = mc_hash(*state + 0x9E3779B9u);
    return (double)(*state & 0x00FFFFFFu) / (double)0x01000000u;
}

typedef struct {
    int fate; /* 0 solid, 1 exit x = 0, 2 exit x = nx, 3 step cap */
    double t, n[3];
} mc_hit;

/* _trace, properties.py:96-194 / penetration.py:73-171 */
static mc_hit mc_trace(const uint8_t *solid, int64_t nx, int64_t ny, int64_t nz, double ox, double oy, double oz,
                       double dx0, double dy0, double dz0)
{
    const double norm = sqrt(dx0 * dx0 + dy0 * dy0 + dz0 * dz0);
    const double d[3] = {dx0 / norm, dy0 / norm, dz0 / norm};
    const double o[3] = {ox, oy, oz};
    int64_t idx[3], step[3] = {0, 0, 0};
    double tmax[3] = {1.0e300, 1.0e300, 1.0e300}, tdelta[3] = {1.0e300, 1.0e300, 1.0e300};
    for (int a = 0; a < 3; ++a) {
        idx[a] = (int64_t)floor(o[a]);
        if (d[a] < 0.0 && fabs(o[a] - floor(o[a] + 0.5)) <= 1.0e-12) idx[a] -= 1;
        const double safe = fabs(d[a]) > 0.0 ? d[a] : 1.0;
        if (d[a] > 0.0) {
            step[a] = 1;
            tmax[a] = ((double)(idx[a] + 1) - o[a]) / safe;
            tdelta[a] = 1.0 / safe;
        } else if (d[a] < 0.0) {
            step[a] = -1;
            tmax[a] = ((double)idx[a] - o[a]) / safe;
            tdelta[a] = -1.0 / safe;
        }
    }
    mc_hit h = {-1, 0.0, {0.0, 0.0, 0.0}};
    const int64_t max_steps = 10 * nx * ny * nz + 10;
    int64_t steps = 0;
    while (h.fate == -1 && steps < max_steps) {
        steps += 1;
        if (idx[0] < 0) {
            h.fate = 1;
        } else if (idx[0] >= nx) {
            h.fate = 2;
        } else if (solid[(idx[0] * ny + py_imod(idx[1], ny)) * nz + py_imod(idx[2], nz)] != 0) {
            h.fate = 0;
        } else {
            const double t_next = fmin(tmax[0], fmin(tmax[1], tmax[2]));
            int crossed[3];
            for (int a = 0; a < 3; ++a) crossed[a] = fabs(tmax[a] - t_next) <= 1.0e-12;
            for (int a = 0; a < 3; ++a) {
                if (!crossed[a]) continue;
                idx[a] += step[a];
                tmax[a] += tdelta[a];
                int earlier = 0;
                for (int b = 0; b < a; ++b) earlier |= crossed[b];
                if (!earlier) {
                    h.n[0] = h.n[1] = h.n[2] = 0.0;
                    h.n[a] = -(double)step[a];
                }
            }
            h.t = t_next;
        }
    }
    if (h.fate == -1) h.fate = 3;
    return h;
}

/* _lambertian, properties.py:196-221 / penetration.py:173-198 */
static void mc_lambertian(uint32_t *state, const double n[3], double d[3])
{
    const double xi1 = mc_rand01(state);
    const double xi2 = mc_rand01(state);
    double hx = 1.0, hy = 0.0, hz = 0.0;
    if (fabs(n[0]) > 0.9) {
        hx = 0.0;
        hy = 1.0;
    }
    double u[3] = {hy * n[2] - hz * n[1], hz * n[0] - hx * n[2], hx * n[1] - hy * n[0]};
    const double un = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    for (int a = 0; a < 3; ++a) u[a] /= un;
    const double v[3] = {n[1] * u[2] - n[2] * u[1], n[2] * u[0] - n[0] * u[2], n[0] * u[1] - n[1] * u[0]};
    const double radius = sqrt(xi1);
    const double phi = 2.0 * 3.141592653589793 * xi2;
    const double cx = radius * cos(phi), cy = radius * sin(phi), cz = sqrt(fmax(0.0, 1.0 - xi1));
    for (int a = 0; a < 3; ++a) d[a] = cx * u[a] + cy * v[a] + cz * n[a];
}

/* The reflection loop of both kernels from a solid hit; returns 0 absorbed,
 * 1/2 exit through x = 0 / x = nx, 3 stopped. */
static int mc_reflect(const uint8_t *solid, int64_t nx, int64_t ny, int64_t nz, uint32_t *state, double eps,
                      int max_bounces, int hit_point, double o[3], double d[3], mc_hit h, int64_t *interactions,
                      int64_t *scattered, double *s1, double *s2)
{
    double n[3] = {h.n[0], h.n[1], h.n[2]}, dist = h.t;
    int bounces = 0;
    for (;;) {
        *interactions += 1;
        if (mc_rand01(state) < eps) return 0;
        const double p[3] = {d[0], d[1], d[2]};
        mc_lambertian(state, n, d);
        const double mu = p[0] * d[0] + p[1] * d[1] + p[2] * d[2];
        *scattered += 1;
        *s1 += mu;
        *s2 += mu * mu;
        bounces += 1;
        if (bounces >= max_bounces) return 3;
        for (int a = 0; a < 3; ++a) o[a] = hit_point ? o[a] + n[a] * 1.0e-9 : o[a] + p[a] * dist + n[a] * 1.0e-9;
        const mc_hit g = mc_trace(solid, nx, ny, nz, o[0], o[1], o[2], d[0], d[1], d[2]);
        if (g.fate != 0) return g.fate;
        dist = g.t;
        for (int a = 0; a < 3; ++a) n[a] = g.n[a];
        if (n[0] * n[0] + n[1] * n[1] + n[2] * n[2] <= 1.0e-24)
            for (int a = 0; a < 3; ++a) n[a] = -d[a];
        for (int a = 0; a < 3; ++a) o[a] = o[a] + d[a] * dist;
    }
}

static uint32_t mc_ray_state(uint32_t seed, int64_t ray) { return mc_hash(seed + (uint32_t)(ray + 1) * 747796405u); }

/* trace_properties's kernel (properties.py:223-287): per ray the event or
 * censored first path length (-1 in the other), interactions, reflections
 * and the sums of mu, mu^2. */
void orc_mc_properties(const uint8_t *solid, int64_t nx, int64_t ny, int64_t nz, int64_t n_rays, uint32_t seed,
                       double eps, int max_bounces, int hit_point, double *event_len, double *censored_len,
                       int64_t *interactions, int64_t *scattered, double *mu1, double *mu2)
{
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ray = 0; ray < n_rays; ++ray) {
        uint32_t state = mc_ray_state(seed, ray);
        const double ry = mc_rand01(&state), rz = mc_rand01(&state);
        double o[3] = {0.0, ry * (double)ny, rz * (double)nz}, d[3] = {1.0, 0.0, 0.0};
        mc_hit h = mc_trace(solid, nx, ny, nz, o[0], o[1], o[2], d[0], d[1], d[2]);
        int64_t in = 0, sc = 0;
        double s1 = 0.0, s2 = 0.0;
        if (h.fate == 0) {
            if (h.n[0] * h.n[0] + h.n[1] * h.n[1] + h.n[2] * h.n[2] <= 1.0e-24)
                for (int a = 0; a < 3; ++a) h.n[a] = -d[a];
            event_len[ray] = h.t;
            censored_len[ray] = -1.0;
            if (hit_point)
                for (int a = 0; a < 3; ++a) o[a] = o[a] + d[a] * h.t;
            (void)mc_reflect(solid, nx, ny, nz, &state, eps, max_bounces, hit_point, o, d, h, &in, &sc, &s1, &s2);
        } else {
            event_len[ray] = -1.0;
            censored_len[ray] = h.t;
        }
        interactions[ray] = in;
        scattered[ray] = sc;
        mu1[ray] = s1;
        mu2[ray] = s2;
    }
}

/* trace_penetration's kernel (penetration.py:200-272): per ray 0 reflected,
 * 1 transmitted, 2 absorbed, 3 stopped. */
void orc_mc_penetration(const uint8_t *solid, int64_t nx, int64_t ny, int64_t nz, int64_t n_rays, uint32_t seed,
                        double eps, int max_bounces, int hit_point, uint8_t *fate)
{
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ray = 0; ray < n_rays; ++ray) {
        uint32_t state = mc_ray_state(seed, ray);
        const double ry = mc_rand01(&state), rz = mc_rand01(&state);
        double o[3] = {0.0, ry * (double)ny, rz * (double)nz}, d[3];
        const double normal[3] = {1.0, 0.0, 0.0};
        mc_lambertian(&state, normal, d);
        mc_hit h = mc_trace(solid, nx, ny, nz, o[0], o[1], o[2], d[0], d[1], d[2]);
        int code;
        if (h.fate != 0) {
            code = h.fate == 1 ? 0 : (h.fate == 2 ? 1 : 3);
        } else {
            if (h.n[0] * h.n[0] + h.n[1] * h.n[1] + h.n[2] * h.n[2] <= 1.0e-24)
                for (int a = 0; a < 3; ++a) h.n[a] = -d[a];
            if (hit_point)
                for (int a = 0; a < 3; ++a) o[a] = o[a] + d[a] * h.t;
            int64_t in = 0, sc = 0;
            double s1 = 0.0, s2 = 0.0;
            const int f = mc_reflect(solid, nx, ny, nz, &state, eps, max_bounces, hit_point, o, d, h, &in, &sc, &s1,
                                     &s2);
            code = f == 0 ? 2 : (f == 1 ? 0 : (f == 2 ? 1 : 3));
        }
        fate[ray] = (uint8_t)code;
    }
}

/* ------------------------------------------------------------------ geometry stamping
 * Short shapes of shape_mask (porerad/geometry.py:106-116): per cell the
 * minimum-image offsets ((i - c) + N/2) mod N - N/2 with NumPy's float
 * remainder (fmod, plus N when the result is negative), the local
 * coordinates R^T x as OpenBLAS's dgemm forms them (an FMA chain over
 * k = 0, 1, 2 starting from a rounded product) and the membership test with
 * every division, square and sum rounded on its own.
 * params per shape: cx, cy, cz, rx, ry, rz, then R^T row-major (Rt[r][k] =
 * R[k][r]); kinds: 0 ellipsoid, 1 cylinder. */
static inline double orc_min_image(int64_t i, double c, double N)
{
    const double half = N / 2.0;
    double m = fmod(((double)i - c) + half, N);
    if (m != 0.0) {
        if (m < 0.0) m += N;
    } else {
        m = 0.0;
    }
    return m - half;
}

void orc_stamp_shapes(uint8_t *phases, int64_t N, int64_t n_shapes, const double *params, const int32_t *kinds)
{
    const double Nf = (double)N;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i)
        for (int64_t s = 0; s < n_shapes; ++s) {
            const double *p = params + 15 * s;
            const double *Rt = p + 6;
            const double x = orc_min_image(i, p[0], Nf);
            for (int64_t j = 0; j < N; ++j) {
                const double y = orc_min_image(j, p[1], Nf);
                for (int64_t k = 0; k < N; ++k) {
                    const double z = orc_min_image(k, p[2], Nf);
                    double l[3];
                    for (int r = 0; r < 3; ++r) l[r] = fma(Rt[3 * r + 2], z, fma(Rt[3 * r + 1], y, Rt[3 * r] * x));
                    const double a = l[0] / p[3], b = l[1] / p[4];
                    const double ab = a * a + b * b;
                    int inside;
                    if (kinds[s] == 0) {
                        const double c = l[2] / p[5];
                        inside = ab + c * c <= 1.0;
                    } else {
                        inside = ab <= 1.0 && fabs(l[2]) <= p[5];
                    }
                    if (inside) phases[(i * N + j) * N + k] = 1;
                }
            }
        }
}

/* _long_cylinder's stamping (porerad/geometry.py:63-86) followed by the
 * whole-cell roll of shape_mask (:100-105): cell (node + offset + shift)
 * mod N for every node and disk offset. */
void orc_stamp_offsets(uint8_t *phases, int64_t N, int64_t n_nodes, const int64_t *nodes, int64_t n_off,
                       const int64_t *offsets, const int64_t shift[3])
{
    for (int64_t a = 0; a < n_nodes; ++a)
        for (int64_t b = 0; b < n_off; ++b) {
            int64_t c[3];
            for (int d = 0; d < 3; ++d) c[d] = ((nodes[3 * a + d] + offsets[3 * b + d] + shift[d]) % N + N) % N;
            phases[(c[0] * N + c[1]) * N + c[2]] = 1;
        }
}
