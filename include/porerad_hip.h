/*
 * porerad_hip.h -- C ABI of the MI355X (gfx950) radiation hot path of porerad.
 *
 * Plain C types only (pointers, sizes, status codes); no C++ or torch types
 * cross this boundary.  Every entry point returns PRH_OK (0) or a status code
 * and leaves a message for prh_last_error().  Host pointers are caller-owned
 * and never retained; device memory is owned by the library behind the opaque
 * handles and freed by the matching *_destroy call.  Calls are synchronous
 * with respect to the host unless stated otherwise.
 *
 * Each entry point replaces one function of the reference Python package
 * (paths relative to the reference repository root):
 *
 *   prh_pcg64_seed            numpy PCG64(seed) seeding used by
 *                             generate_rays, porerad/view_factors.py:149-150
 *   prh_libm_sincos           np.cos / np.sin of the azimuth in generate_rays,
 *                             porerad/view_factors.py:168-176
 *   prh_geometry_create       build_surface_patches, porerad/view_factors.py:70-132
 *   prh_geometry_patch_table  the SurfacePatches arrays, porerad/view_factors.py:43-67
 *   prh_generate_rays         generate_rays, porerad/view_factors.py:135-182
 *   prh_trace_rays            trace_first_hit / _trace_kernel,
 *                             porerad/view_factors.py:185-270
 *   prh_trace_counts          first_hit_counts (+ pool_counts over several
 *                             seeds), porerad/view_factors.py:273-300
 *   prh_operator_create       ExchangeOperator.__init__ assembly,
 *                             porerad/exchange.py:47-76
 *   prh_operator_apply        raw_matvec / raw_rmatvec / symmetric_matvec /
 *                             matvec / rmatvec, porerad/exchange.py:82-122
 *   prh_radiosity             reflection_series, porerad/exchange.py:217-245
 *   prh_scalar_counts         scalar_descriptors / interface_area /
 *                             mean_pore_chords / pore_connectivity /
 *                             interface_face_counts, porerad/descriptors.py:38-79,120-123
 *   prh_periodic_chords       periodic_chords, porerad/descriptors.py:82-109
 *   prh_stamp_shapes          shape_mask (short shapes), porerad/geometry.py:106-116
 *   prh_stamp_offsets,        _long_cylinder, porerad/geometry.py:63-86
 *   prh_stamp_long_cylinders
 *   prh_threshold_field       lamellar_field smoothing + threshold,
 *                             porerad/geometry.py:229-238
 */
#ifndef PORERAD_HIP_H
#define PORERAD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define PRH_OK 0
#define PRH_ERR_ARG 1        /* invalid argument or size */
#define PRH_ERR_HIP 2        /* HIP runtime failure */
#define PRH_ERR_UNRESOLVED 3 /* rays did not resolve a receiving patch
                                (RuntimeError, porerad/view_factors.py:268-269) */
#define PRH_ERR_COUNTS 4     /* first-hit counts do not account for every ray
                                (ValueError, porerad/exchange.py:70-71) */
#define PRH_ERR_SERIES 5     /* reflection series did not converge
                                (RuntimeError, porerad/exchange.py:243-244) */
#define PRH_ERR_CAPACITY 6   /* request exceeds a documented capacity */
#define PRH_ERR_BALANCE 7    /* diagonal balancing did not converge
                                (RuntimeError, porerad/exchange.py:197) */

/* ---- operator application modes (prh_operator_apply) ----------------- */
#define PRH_APPLY_SYMMETRIC 0 /* S x              (symmetric_matvec) */
#define PRH_APPLY_MATVEC 1    /* F_p x            (matvec)           */
#define PRH_APPLY_RMATVEC 2   /* F_p^T x          (rmatvec)          */
#define PRH_APPLY_RAW 3       /* F x              (raw_matvec)       */
#define PRH_APPLY_RAW_T 4     /* F^T x            (raw_rmatvec)      */

/* ---- radiosity precisions (prh_radiosity) ---------------------------- */
#define PRH_FP64 0  /* plain FP64 reflection series */
#define PRH_MIXED 1 /* batched Chebyshev sweeps (FP16 gathered directions) + FP64 refinement */
#define PRH_FP64_BATCHED 2 /* the FP64 reflection series of up to 32 emissions per pass over the
                              operator (FP64 slot vectors interleaved per patch: 16 slots = one
                              128 B line per gathered column, 32 slots = two); sums in another
                              order than PRH_FP64, equal to rounding */

typedef struct prh_geometry prh_geometry;
typedef struct prh_edges prh_edges;
typedef struct prh_operator prh_operator;
typedef struct prh_batch prh_batch;
typedef struct prh_coupled prh_coupled;
typedef struct prh_trace_batch prh_trace_batch;
typedef struct prh_phase_field prh_phase_field;

/* Message of the last failing call on this thread ("" if none). */
const char *prh_last_error(void);

/* Library version string. */
const char *prh_version(void);

/* Kernel durations (ms, HIP events on the library stream) of the last
 * prh_trace_counts call: [0] ray generation, [1] traversal, [2] per-patch
 * counting, [3] scan + edge compaction.  n = number of slots to fill. */
int prh_last_timings(double *ms, int32_t n);

/* Tuning knob of the traversal kernel: variant index in [0, n_variants).
 * Each variant fixes the refill threshold (idle lanes of a wavefront before
 * it loads new rays) and the occupancy test: 0 the global bitset, 1 the LDS
 * brick bitset first, 2 the global bitset word loaded ahead of the step, 3 the
 * lean step (walled bitset in 4x4x4 blocks, the current block's word held per
 * lane; geometries without a walled bitset run the brick test instead).  No
 * counterpart in the reference (a launch parameter of this implementation);
 * results do not depend on it.  prh_tracer_variants fills the tables when
 * the pointers are non-NULL (arrays of n_variants entries). */
int prh_tracer_set_variant(int32_t variant);
int prh_tracer_variants(int32_t *n_variants, int32_t *current, int32_t *refill_min, int32_t *occupancy_mode);
/* Wave-steps and refills of the lean traversal on the current device since
 * the last reset (after the library stream has drained); reset != 0 clears
 * them.  Diagnostic with no counterpart in the reference: with the occupancy
 * reads, occupancy_reads / (64 wave_steps) is the share of lanes that step. */
int prh_tracer_stats(int64_t *wave_steps, int64_t *refills, int32_t reset);
/* Tuning knob of the traversal kernel: records a wavefront takes from the
 * launch's shared work queue at a time (64 .. 2^20, default 512).  No
 * counterpart in the reference; results do not depend on it. */
int prh_tracer_set_grain(int32_t records);
/* Free the current device's tracer workspace (ray records of the largest
 * prh_trace_counts / prh_trace_rays pass so far; it grows on demand and is
 * otherwise kept) and the kept block of the last destroyed edge table (see
 * prh_edges_destroy).  No counterpart in the reference. */
int prh_tracer_trim(void);
/* Bound of the tracer workspace (default 32 GiB, at least 1 MiB).  A call
 * also keeps the workspace within half of the device memory that is free or
 * already held by it; larger inputs are traced in several passes (patch
 * ranges, seed blocks), with the same results.  prh_tracer_workspace reports
 * the bytes the current device holds (the workspace and the kept edge-table
 * block) and the bound (either pointer may be NULL), e.g. for an HBM budget
 * that admits other work.  The workspace is
 * the bounded counterpart of the reference's patch_chunk
 * (porerad/view_factors.py:273-287), which bounds its host arrays. */
int prh_tracer_set_workspace_limit(int64_t bytes);
int prh_tracer_workspace(int64_t *held_bytes, int64_t *limit_bytes);
int prh_tracer_grain(int32_t *records);

/* numpy.random.PCG64(seed) seeding: SeedSequence over the little-endian
 * uint32 words of the seed.  state_inc = {state_hi, state_lo, inc_hi, inc_lo}.
 * Replaces np.random.PCG64(int(seed)), porerad/view_factors.py:149. */
int prh_pcg64_seed(const uint32_t *words, int32_t n_words, uint64_t state_inc[4]);

/* sin(x[i]) and cos(x[i]) with the host C library's bits (glibc's dbl-64
 * sin/cos as its FMA build on x86-64 rounds them), the routine the ray
 * generation uses for the azimuth; on the host (on_device 0, no GPU needed)
 * or on the current device.  |x| < 105414350, finite.  Replaces NumPy's
 * cos/sin of 2 pi u, porerad/view_factors.py:168-176. */
int prh_libm_sincos(const double *x, int64_t n, int32_t on_device, double *sin_x, double *cos_x);

/* Upload an x-oriented solid/pore grid (C order, nonzero = solid, i.e. the
 * phases == 1 mask) and build the patch table on the device: wall patches
 * first in (j, k, hot/cold) order, then exposed solid faces in (i, j, k, face)
 * order with periodic y/z.  Replaces build_surface_patches,
 * porerad/view_factors.py:70-132. */
int prh_geometry_create(const uint8_t *solid_x, int32_t nx, int32_t ny, int32_t nz, double H_m,
                        prh_geometry **out);
/* A geometry with the occupancy bitsets only (no patch tables): enough for
 * prh_mc_properties / prh_mc_penetration, which read nothing else.
 * prh_trace_rays rejects it (PRH_ERR_ARG); calls that take a patch range see
 * zero patches.  No counterpart in the reference (its Monte Carlo reads the
 * phase array directly). */
int prh_geometry_create_occupancy(const uint8_t *solid_x, int32_t nx, int32_t ny, int32_t nz, prh_geometry **out);
int prh_geometry_sizes(const prh_geometry *g, int64_t *n_patches, int64_t *n_wall);
/* Download the patch table (arrays of n_patches; origin is n_patches x 3).
 * Any pointer may be NULL.  Replaces the SurfacePatches fields,
 * porerad/view_factors.py:43-67. */
int prh_geometry_patch_table(const prh_geometry *g, int8_t *patch_class, int32_t *patch_cell,
                             int8_t *patch_face, double *patch_origin);
void prh_geometry_destroy(prh_geometry *g);

/* ---- multi-bounce Monte Carlo ray statistics -------------------------- */
/* trace_properties's ray kernel (porerad/properties.py:54-307) for one seed
 * on the x-oriented solid of geometry g: n_rays beam rays along +x enter at
 * (0, u ny, v nz); at a solid surface a ray is absorbed with probability
 * epsilon, else reflected into a Lambertian direction (at most max_bounces
 * reflections; restart_hit_point 0 = the "article" rule, 1 = "hit_point").
 * Per ray the first path length: event_len (solid hit) or censored_len
 * (else), -1 in the other array (host arrays of n_rays).  counters =
 * {surface interactions, reflections}; sums = {sum mu, sum mu^2} of the
 * cosines between incoming and reflected directions. */
int prh_mc_properties(const prh_geometry *g, int64_t n_rays, uint32_t seed, double epsilon, int32_t max_bounces,
                      int32_t restart_hit_point, double *event_len, double *censored_len, int64_t counters[2],
                      double sums[2]);
/* trace_penetration's ray kernel (porerad/penetration.py:54-305): Lambertian
 * rays about +x through x = 0; counts = {reflected (left through x = 0),
 * transmitted (through x = nx), absorbed, stopped at the step or reflection
 * limit}. */
int prh_mc_penetration(const prh_geometry *g, int64_t n_rays, uint32_t seed, double epsilon, int32_t max_bounces,
                       int32_t restart_hit_point, int64_t counts[4]);

/* ---- morphology descriptors (porerad/descriptors.py) ----------------- */
/* A phase field is a copy of a uint8 array (nx, ny, nz), x-major, 1 = solid,
 * 0 = pore (the reference's phase arrays), at most 2^31 - 1 cells, on the
 * current device.  The library returns the integer quantities the descriptors are
 * made of; ratios, means and quantiles are formed from them by the caller. */
int prh_phase_field_create(const uint8_t *phases, int32_t nx, int32_t ny, int32_t nz, prh_phase_field **out);
int prh_phase_field_read(const prh_phase_field *f, uint8_t *phases);
void prh_phase_field_destroy(prh_phase_field *f);

/* Indices of prh_scalar_counts' output. */
#define PRH_DESC_CELLS 0
#define PRH_DESC_SOLID 1              /* cells of phase 1 */
#define PRH_DESC_PORE 2               /* cells of phase 0 */
#define PRH_DESC_BOX_FACES_X 3        /* +0/1/2: faces inside the box between cells whose
                                         solid indicators differ (interface_area,
                                         porerad/descriptors.py:38-43) */
#define PRH_DESC_PORE_RUNS_X 6        /* +0/1/2: pore runs ended by the box faces
                                         (mean_pore_chords, :46-48) */
#define PRH_DESC_PERIODIC_FACES_X 9   /* +0/1/2: pore indicator differs from the next cell
                                         with wrap-around (interface_face_counts, :120-123) */
#define PRH_DESC_LARGEST_CLUSTER 12   /* largest 26-connected pore cluster, cells */
#define PRH_DESC_CLUSTERS 13          /* number of 26-connected pore clusters */
#define PRH_DESC_PERCOLATES 14        /* bit a: one cluster touches both faces of axis a */
#define PRH_DESC_COUNT 15

/* Counts of scalar_descriptors (porerad/descriptors.py:63-79) and
 * pore_connectivity (:51-60, scipy.ndimage.label with a full 3x3x3
 * structure).  connectivity 0 skips the cluster labelling (its three entries
 * are then -1). */
int prh_scalar_counts(const prh_phase_field *f, int32_t connectivity, int64_t counts[PRH_DESC_COUNT]);

/* periodic_chords along axis 0/1/2 (porerad/descriptors.py:82-109): the chord
 * lengths in the reference's order (lines in row-major order of the other two
 * axes, runs along each line; a line whose first and last cells are pore and
 * that is not all pore joins its end runs into one chord placed after all
 * others, in line order), their count, the number of all-pore lines, and the
 * chord-count histogram over lengths 0..hist_len-1 (longer chords are not
 * binned; the reference bins to N + 1 with N = nx).  chords NULL returns the
 * counts only; more chords than capacity is PRH_ERR_CAPACITY with *n_chords
 * set.  Capacity lines * ceil(length / 2) always suffices. */
int prh_periodic_chords(const prh_phase_field *f, int32_t axis, int32_t *chords, int64_t capacity, int64_t *n_chords,
                        int64_t *n_all_pore, int64_t *histogram, int32_t hist_len);

/* ---- microstructure generators (porerad/geometry.py) ----------------- */
/* prh_phase_field_create with phases NULL gives a field of zeros.  Fill,
 * copy (same shape and device) and count the cells of one value (the solid
 * fraction of _bisect_radius, porerad/geometry.py:120-140). */
int prh_phase_field_fill(prh_phase_field *f, uint8_t value);
int prh_phase_field_copy(prh_phase_field *dst, const prh_phase_field *src);
int prh_phase_field_count(const prh_phase_field *f, uint8_t value, int64_t *count);

/* shape_mask's short shapes (porerad/geometry.py:106-116) OR-ed into a cubic
 * field (cells set to 1).  shapes: n_shapes x 16 doubles = centre (3), radii
 * (3; r_z the half length of a cylinder), R^T row-major (9; R's columns are
 * the local axes), reach (offsets beyond it on any axis are outside the
 * shape); kinds: 0 ellipsoid, 1 cylinder.  Membership is evaluated with the
 * reference's FP64 operations (R^T x as an FMA chain, as OpenBLAS forms it). */
int prh_stamp_shapes(prh_phase_field *f, int32_t n_shapes, const double *shapes, const int32_t *kinds);

/* _long_cylinder's stamping (porerad/geometry.py:63-86) and the roll of
 * shape_mask (:100-105): cells (node + offset + shift) mod N set to 1.
 * nodes: n_nodes x 3 in [0, N); offsets: n_offsets x 3 in (-N, N); shift
 * in [0, N). */
int prh_stamp_offsets(prh_phase_field *f, int32_t n_nodes, const int32_t *nodes, int32_t n_offsets,
                      const int32_t *offsets, const int32_t shift[3]);

/* The long cylinders of one render in one call: _long_cylinder for each
 * (porerad/geometry.py:63-86) with the roll of integer centres (:100-105).
 * Cylinder c has the nodes [node_start[c], node_start[c+1]) of nodes (x3, in
 * [0, N)), radius[c] in (0, N/4], the transposed disk frame rq_t[9c..9c+9)
 * (Rq^T row-major) and shift[3c..3c+3) in [0, N); its disk is the offsets o
 * of the (2 pad + 1)^3 box, pad = ceil(r) + 2, with (l0/r)^2 + (l1/r)^2 <=
 * 1 + 1e-12 and |l2| <= 0.5 + 1e-12 for l = Rq^T o (an FMA chain, as NumPy's
 * product rounds), and cells (node + o + shift) mod N are set to 1. */
int prh_stamp_long_cylinders(prh_phase_field *f, int32_t n_cyl, const int32_t *node_start, const int32_t *nodes,
                             const double *rq_t, const double *radius, const int32_t *shift);

/* threshold (porerad/geometry.py:235-238): the field (one FP64 value per cell
 * of f, host) plus, if noise is not NULL (f cubic), lamellar_field's
 * seven-point smoothing of noise (:229-232), thresholded at
 * np.quantile(values, q) (linear rule; the value comes back in *quantile):
 * cells >= it become 1, the others 0. */
int prh_threshold_field(prh_phase_field *f, const double *field, const double *noise, double q, double *quantile);

/* Origins and directions ((p1-p0)*rays_per_patch x 3, FP64) of the seeded
 * Lambertian rays of patches [p0, p1).  Replaces generate_rays,
 * porerad/view_factors.py:135-182. */
int prh_generate_rays(const prh_geometry *g, int64_t p0, int64_t p1, int32_t rays_per_patch,
                      const uint64_t state_inc[4], double *origins, double *directions);

/* First patch hit by each given ray (receiver = -1 where unresolved).
 * Origins may lie anywhere within |o| < 2^31 (lateral indices wrap as the
 * reference's j % ny, k % nz; an x index outside [0, nx) is the wall of that
 * side at once); non-finite values or larger origins are PRH_ERR_ARG.
 * Returns PRH_ERR_UNRESOLVED with *n_failed > 0 if any ray failed.
 * Replaces trace_first_hit, porerad/view_factors.py:261-270. */
int prh_trace_rays(const prh_geometry *g, const double *origins, const double *directions, int64_t n_rays,
                   int32_t *receiver, int64_t *n_failed);

/* Fused ray generation + FP64 DDA + per-patch counting for patches
 * [p0, p1) and n_seeds seeded ray blocks (state_inc: n_seeds x 4 words), all
 * pooled into one device-resident edge table sorted by (origin, receiver).
 * Any patch range and rays_per_patch * n_seeds < 2^31: a pass counts at most
 * 16384 rays per patch, so more seeds are traced in groups (each seed's rays
 * in parts of 16384 when rays_per_patch is larger) whose tables are pooled on
 * the device, and the patch range is cut into passes that fit the workspace
 * bound (prh_tracer_set_workspace_limit).  *occupancy_reads (may be NULL) receives
 * the number of occupancy reads of the traversal (the exact count of
 * solid[i, j, k] loads the reference kernel makes).
 * Replaces first_hit_counts + pool_counts, porerad/view_factors.py:273-300. */
int prh_trace_counts(const prh_geometry *g, int64_t p0, int64_t p1, int32_t rays_per_patch, int32_t n_seeds,
                     const uint64_t *state_inc, prh_edges **out, int64_t *nnz, int64_t *occupancy_reads);
/* prh_trace_counts without any device allocation: the edges go into
 * caller-owned device arrays origin, receiver, count (int32, capacity
 * entries) and *nnz receives their number (PRH_ERR_CAPACITY, nothing beyond
 * capacity written, when they do not fit: *nnz is then the size needed).
 * The work runs in a fixed 256 MiB device array of the library (one per
 * device, held until the process ends) in passes of as many patches as it
 * holds; rays_per_patch * n_seeds <= 16384 (PRH_ERR_CAPACITY above).  Calls on
 * one device share that array and are serialised by the library; the array
 * keeps the last call's contents.  Same edges, bit for bit, as
 * prh_trace_counts.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops
 * with a fixed memory plan). */
int prh_trace_counts_into(const prh_geometry *g, int64_t p0, int64_t p1, int32_t rays_per_patch, int32_t n_seeds,
                          const uint64_t *state_inc, int32_t *origin, int32_t *receiver, int32_t *count,
                          int64_t capacity, int64_t *nnz, int64_t *occupancy_reads);
/* Many first_hit_counts calls of uneven geometries in one graph.  Segment k
 * traces patches [patch_begin[k], patch_end[k]) of geometries[k] with
 * rays_per_patch[k] rays from each of n_seeds[k] seeded blocks, pooled as
 * prh_trace_counts pools them (the same edges, bit for bit).  The segments
 * are cut at patch boundaries into chunks of at most ray_capacity ray records
 * (0: min(2^28, all rays); 112 bytes of HBM per record and stream); each
 * chunk is one ray-generation launch and one traversal launch over all its
 * pieces, followed by per-piece counting and compaction into one device edge
 * arena of edge_capacity entries (0: the worst case, rays per patch bounded by
 * the patch count).  Chunks take turns over n_streams streams (0: 2; at most
 * 4), each with its own record buffers, so consecutive chunks overlap.  The
 * pass is captured once into a HIP graph that prh_trace_batch_run replays
 * (flags PRH_BATCH_DIRECT: launched directly on each run instead).  The
 * geometries must outlive the batch.  Replaces the per-case loop of
 * first_hit_counts + pool_counts calls, porerad/view_factors.py:273-300, as
 * scripts/run_prescribed_field.py:81-85 drives it per case. */
#define PRH_BATCH_DIRECT 1
int prh_trace_batch_create(int32_t n_segments, const prh_geometry *const *geometries, const int64_t *patch_begin,
                           const int64_t *patch_end, const int32_t *rays_per_patch, const int32_t *n_seeds,
                           int64_t ray_capacity, int64_t edge_capacity, int32_t n_streams, int32_t flags,
                           prh_trace_batch **out);
/* Replay the pass with seeds state_inc (each segment's n_seeds x 4 words, in
 * segment order).  nnz[k] / occupancy_reads[k] (either may be NULL) receive
 * segment k's edge count and occupancy reads; *elapsed_ms (may be NULL) the
 * device time of the pass.  PRH_ERR_UNRESOLVED if a ray failed,
 * PRH_ERR_CAPACITY if the edges overflowed the arena. */
int prh_trace_batch_run(prh_trace_batch *b, const uint64_t *state_inc, int64_t *nnz, int64_t *occupancy_reads,
                        double *elapsed_ms);
/* Copy segment k's edges of the last run to host arrays of nnz[k] entries. */
int prh_trace_batch_edges(const prh_trace_batch *b, int32_t segment, int32_t *origin, int32_t *receiver,
                          int64_t *count);
/* Segment k's edges of the last run as a device edge table of their own
 * (for prh_operator_create_device); free it with prh_edges_destroy. */
int prh_trace_batch_edges_device(const prh_trace_batch *b, int32_t segment, prh_edges **out);
/* info[0..n): device bytes, chunks, pieces, ray capacity, edge capacity, rays,
 * streams, graph (1) or direct (0), traversal variant. */
int prh_trace_batch_info(const prh_trace_batch *b, int64_t *info, int32_t n);
void prh_trace_batch_destroy(prh_trace_batch *b);

/* *equal = 1 when two device edge tables (same device) hold the same
 * entries, else 0; compared on the device, nothing is copied to the host.
 * No counterpart in the reference (checks tables too large to copy). */
int prh_edges_equal(const prh_edges *a, const prh_edges *b, int32_t *equal);
/* Copy a device edge table to host arrays of nnz entries. */
int prh_edges_copy(const prh_edges *e, int32_t *origin, int32_t *receiver, int64_t *count);
/* Free an edge table.  The single block that holds a counted table's three
 * arrays is kept (one per device, the largest) for the next counted table
 * that fits in it, so repeated counting neither allocates nor frees device
 * memory; prh_tracer_trim releases it.  Waits for the device first, as
 * freeing device memory does. */
void prh_edges_destroy(prh_edges *e);

/* Assemble the symmetric exchange matrix
 *     S = 0.5 (diag(A) F + F^T diag(A)) + delta diag(A),  F_ij = n_ij / R_i
 * on the device from an edge table sorted by (origin, receiver) with unique
 * pairs whose rows account for every ray.  Replaces the assembly in
 * ExchangeOperator.__init__, porerad/exchange.py:59-76. */
int prh_operator_create(int64_t n, const int32_t *origin, const int32_t *receiver, const int64_t *count,
                        int64_t nnz, const int64_t *rays_per_origin, const double *area,
                        double support_diagonal_fraction, prh_operator **out);
/* The same operator as prh_operator_create, assembled on the device from a
 * device edge table (prh_trace_counts' output, sorted by (origin, receiver),
 * unique pairs): no edge leaves the device.  Row totals are checked against
 * rays_per_origin (PRH_ERR_COUNTS, exchange.py:68-71); n, rays_per_origin and
 * area (host arrays of n) as in prh_operator_create.  The FP64 values of S
 * are bitwise those of prh_operator_create.  Replaces the assembly in
 * ExchangeOperator.__init__, porerad/exchange.py:59-76. */
int prh_operator_create_device(const prh_edges *edges, int64_t n, const int64_t *rays_per_origin, const double *area,
                               double support_diagonal_fraction, prh_operator **out);
/* The operator's count table (origin, receiver, count), sorted, nnz_raw
 * entries (prh_operator_nnz): the edges it was built from. */
int prh_operator_edges(const prh_operator *op, int32_t *origin, int32_t *receiver, int64_t *count);
/* Diagonal balancing factor d (n), used by modes MATVEC/RMATVEC and by
 * prh_radiosity.  Defaults to ones. */
int prh_operator_set_scale(prh_operator *op, const double *scale);
/* Order in which the batched engine (prh_batch_*) sweeps the rows: a
 * permutation of the n patches (order[p] = the patch swept p-th), or NULL
 * for the stored order.  Each XCD sweeps a contiguous run of positions, so a
 * spatially coherent order (patches sorted along a space-filling curve)
 * lets the rows an XCD sweeps together share the L2 lines they gather.
 * Results change only in the rounding of the engine's norm sums.  No
 * counterpart in the reference (performance). */
int prh_operator_set_row_order(prh_operator *op, const int32_t *order);
/* Balance the operator on the device and set d: solve d (S d) = A, as
 * ExchangeOperator._balance, porerad/exchange.py:153-197 (with
 * _closure_error, :142-151): no balancing when d = 1 already closes the rows
 * within tolerance (method 0); else SciPy's Anderson mixing (M 10, w0 0.01,
 * at most 500 iterations, Armijo line search; method 1) when it reaches
 * max |F| <= tolerance and the closure min(max(10 tol, 10 eps), 1e-10); else
 * the damped multiplicative iteration (method 2, at most 20000 steps).
 * PRH_ERR_BALANCE when none converges.  *iterations as balance_iterations. */
int prh_operator_balance(prh_operator *op, double tolerance, int32_t *method, int32_t *iterations);
/* Copy the balancing factor d (n doubles) to the host. */
int prh_operator_get_scale(const prh_operator *op, double *scale);
int prh_operator_nnz(const prh_operator *op, int64_t *nnz_symmetric, int64_t *nnz_raw);
/* The symmetric FP64 matrix S (CSR: rowptr n + 1, col and val nnz_symmetric,
 * columns sorted) to host arrays; any pointer may be NULL.  The matrix of
 * symmetric_matvec, porerad/exchange.py:108-112, as stored on the device. */
int prh_operator_csr(const prh_operator *op, int64_t *rowptr, int32_t *col, double *val);
/* Largest relative defect of A_i F_p,ij = A_j F_p,ji over the count table's
 * pairs, computed on the device from the stored symmetric matrix and the
 * balancing factor.  Replaces ExchangeOperator.reciprocity_error,
 * porerad/exchange.py:124-140. */
int prh_operator_reciprocity(prh_operator *op, double *error);
/* y = op(x) for one of the PRH_APPLY_* modes; x, y host arrays of n. */
int prh_operator_apply(prh_operator *op, int32_t mode, const double *x, double *y);

/* Device time of the FP64 reflection series run on this operator (HIP events
 * around each series): out[0..n) = series, reflections, milliseconds,
 * nonzeros of S, patches, 1 if the integer form K is held, then for the
 * batched FP64 series (PRH_FP64_BATCHED) passes, slot-weighted passes (the
 * slot width summed over passes) and milliseconds.  No counterpart in the
 * reference (measurement). */
int prh_operator_stats(const prh_operator *op, double *out, int32_t n);
/* Measurement probes of the batched radiosity sweep (no counterpart in the
 * reference; they solve nothing).  Both need the integer form K (equal areas
 * and ray counts).
 *
 * prh_probe_gather: device milliseconds of one sweep z' = F_p^T z over every
 * slot of an interleaved slot vector in one of the PRH_PROBE_* formats (the
 * batched engine's tick without its per-slot update), averaged over reps
 * sweeps after one warm-up sweep. */
#define PRH_PROBE_F32X32 0 /* 32 FP32 slots per patch: 128 B per gathered column */
#define PRH_PROBE_F16X64 1 /* 64 FP16 slots: 128 B */
#define PRH_PROBE_F16X32 2 /* 32 FP16 slots: 64 B */
#define PRH_PROBE_F64X16 3 /* 16 FP64 slots: 128 B */
#define PRH_PROBE_F64X32 4 /* 32 FP64 slots: 256 B */
int prh_probe_gather(const prh_operator *op, int32_t format, int32_t reps, double *ms_per_sweep);
/* prh_probe_tiles: the 16 x 4 blocks of K holding >= min_fill nonzeros as
 * dense tiles multiplied by v_mfma_f32_16x16x4f32 into 32 FP32 slots, the
 * rest as a CSR gathered like the engine's tick.  out[0..n_out) = ms of the
 * plain gather of all of K, ms of the tiles, ms of the remainder gather
 * (each averaged over reps), the largest difference between the plain
 * product and tiles + remainder relative to the largest product, tiles,
 * fraction of the nonzeros in tiles, remainder nonzeros. */
int prh_probe_tiles(const prh_operator *op, int32_t min_fill, int32_t reps, double *out, int32_t n_out);
/* prh_probe_order: the FP16 x 64 sweep of prh_probe_gather over S in its
 * stored patch order and with patches renumbered by `order` (n patch
 * indices, a permutation: position p takes patch order[p]), each with plain
 * and with non-temporal (streamed) loads of the matrix.  out[0..n_out) = ms
 * stored/plain, stored/streamed, reordered/plain, reordered/streamed, and the
 * largest difference between the reordered product and the stored one
 * reordered, relative to the largest product (FP16 outputs). */
int prh_probe_order(const prh_operator *op, const int32_t *order, int32_t reps, double *out, int32_t n_out);
/* prh_probe_panels: the FP16 x 64 sweep in n_panels phases, phase k
 * gathering only the columns [k W, (k + 1) W) (W = ceil(n / n_panels)) with
 * FP64 accumulators kept in HBM between phases.  out[0..n_out) = ms of the
 * plain sweep, of the panel sweep in the stored row order and in the
 * operator's row order (prh_operator_set_row_order; the stored order when
 * none is set), and the largest difference between the panel and the plain
 * product relative to the largest (FP16 outputs). */
int prh_probe_panels(const prh_operator *op, int32_t n_panels, int32_t reps, double *out, int32_t n_out);

/* Reflection series J = E + (1 - eps) F_p^T J for n_rhs emissions.
 * emitted: n_rhs x n (row-major, one emission per row); emissivity:
 * n_rhs x n if emissivity_per_rhs else n.  Outputs absorbed (n_rhs x n),
 * residual[n_rhs] (power still in flight), iterations[n_rhs].  Stops each
 * emission when ||flight||_1 <= relative_tolerance * ||E||_1.  Returns
 * PRH_ERR_SERIES if an emission needs more than max_reflections.
 * Replaces reflection_series, porerad/exchange.py:217-245. */
int prh_radiosity(prh_operator *op, int32_t n_rhs, const double *emitted, const double *emissivity,
                  int32_t emissivity_per_rhs, double relative_tolerance, int32_t max_reflections,
                  int32_t precision, double *absorbed, double *residual, int32_t *iterations);
void prh_operator_destroy(prh_operator *op);

/* ---- batched radiosity engine (continuous batching) -------------------
 * Up to 64 emissions ("slots") are solved together on one operator.  Each
 * slot alternates epochs of Chebyshev sweeps (FP16 gathered directions,
 * FP32 recursions, FP64 products) with one FP64 refinement sweep (iterative
 * refinement) and finishes when
 * ||E - (I - diag(1-eps) F_p^T) J||_1 <= relative_tolerance ||E||_1; its
 * result is the absorbed power eps * F_p^T J.  Every sweep serves all
 * running slots; slots are admitted and retired at any sweep while the
 * others run, and a solve returns the same bits alone or in any mix.  The
 * slots' schedules live on the device (a schedule kernel before every
 * sweep) and sweeps are replayed from HIP graphs captured per batch: one
 * sweep per replay, eight when prh_batch_step may run eight or more.  Same
 * contract as reflection_series (porerad/exchange.py:217-245); the serving
 * loop is built on it.  The batch is bound to its operator and must be
 * destroyed before it. */
int prh_batch_create(prh_operator *op, prh_batch **out);
/* Device bytes held by the batch (for HBM-budget admission). */
int prh_batch_bytes(const prh_batch *b, int64_t *bytes);
/* Device time of the engine's ticks (HIP events around every tick, read as
 * the pipeline retires it): out[0..n) = ticks timed, their milliseconds,
 * slot-ticks (running slots summed over the ticks), nonzeros of S, patches,
 * bytes per matrix value read by a tick (2: integer K, 8: FP64 S).  No
 * counterpart in the reference (measurement). */
int prh_batch_stats(const prh_batch *b, double *out, int32_t n);
/* Bit q set = slot q free (64 slots). */
int prh_batch_free_slots(const prh_batch *b, uint64_t *free_mask);
/* Describe the patches for linear-field requests: class per patch (0 hot
 * wall, 1 cold wall, 2 internal), the x index of each internal patch's cell
 * (ignored for walls), the number of cells along x and the cross-section
 * ny nz (H/nx)^2 that turns powers into fluxes (porerad/coupled.py:129-194). */
int prh_batch_set_field(prh_batch *b, const int8_t *patch_class, const int32_t *cell_x, int32_t nx,
                        double cross_section_m2);
/* Admit one emission (n values) with emissivity per patch (eps_per_patch)
 * or one value; returns the slot, PRH_ERR_CAPACITY when all are busy. */
int prh_batch_admit(prh_batch *b, const double *emitted, const double *emissivity, int32_t eps_per_patch,
                    double relative_tolerance, int32_t *slot);
/* Admit the prescribed linear field between T_hot and T_cold with uniform
 * emissivity; the emission eps sigma A (T^4 - T_m^4) is formed on the device
 * exactly as prescribed_field_channels does (porerad/coupled.py:170-194).
 * Needs prh_batch_set_field. */
int prh_batch_admit_linear(prh_batch *b, double T_hot_K, double T_cold_K, double emissivity,
                           double relative_tolerance, int32_t *slot);
/* Queue up to max_ticks sweeps, returning early once a finished slot's
 * results have arrived; replays are pipelined, so the last one may still
 * run on the device when this returns (later calls read it).  done_mask:
 * finished slots waiting for prh_batch_channels/result/release. */
int prh_batch_step(prh_batch *b, int32_t max_ticks, uint64_t *done_mask, uint64_t *active_mask);
/* Channels q_h, q_c, q_i (W m^-2, as prescribed_field_channels), power still
 * in flight and sweep count of a finished slot; keeps the slot.
 * PRH_ERR_SERIES if it did not converge. */
int prh_batch_channels(prh_batch *b, int32_t slot, double *q_W_m2, double *residual, int32_t *iterations);
/* Absorbed power (n), power still in flight and sweep count of a finished
 * slot; frees the slot.  PRH_ERR_SERIES if it did not converge. */
int prh_batch_result(prh_batch *b, int32_t slot, double *absorbed, double *residual, int32_t *iterations);
/* Free a finished slot without reading its absorbed power. */
int prh_batch_release(prh_batch *b, int32_t slot);
void prh_batch_destroy(prh_batch *b);


/* ---- strictly coupled radiation-conduction solve ----------------------
 * Device state of the Newton-Krylov iteration of porerad/coupled.py:232-447:
 * a pool of n_vectors cell vectors (n_cells = nx ny nz doubles each,
 * addressed by slot index), the conduction matrix A (CSR of
 * porerad/conduction.py:57-97 with sorted columns) and the wall response
 * vectors b = hot T_hot + cold T_cold.  With a geometry and its exchange
 * operator (both or neither) the handle also evaluates radiation; without,
 * it serves the pure-conduction solve (conduction.py:100-120).  The caller
 * keeps the control flow (Newton steps, line search, FGMRES least squares).
 * All reductions have a fixed order.  The handle works on a stream of its
 * own, so it runs beside batch engines on the same operator; it uses the
 * operator's series buffers, so FP64 prh_radiosity calls on that operator
 * must not overlap with it.  The handle is bound to its geometry and
 * operator and must be destroyed before them. */
int prh_coupled_create(const prh_geometry *geometry, prh_operator *op, int64_t n_cells, const int64_t *A_rowptr,
                       const int32_t *A_col, const double *A_val, const double *hot_rhs, const double *cold_rhs,
                       int32_t n_vectors, prh_coupled **out);
void prh_coupled_destroy(prh_coupled *c);
/* Copy a host vector into slot v, or slot v to the host; fill with a value. */
int prh_coupled_set(prh_coupled *c, int32_t v, const double *host);
int prh_coupled_get(prh_coupled *c, int32_t v, double *host);
int prh_coupled_fill(prh_coupled *c, int32_t v, double value);
/* y = a x + b y (b == 0: y = a x); y = x / d. */
int prh_coupled_axpby(prh_coupled *c, int32_t y, double a, int32_t x, double b);
int prh_coupled_divide(prh_coupled *c, int32_t y, int32_t x, double d);
/* x . y; statistics of x: out[0] ||x||_2, [1] ||x||_inf, [2] min, [3] max,
 * [4] number of non-finite values (excluded from the others). */
int prh_coupled_dot(prh_coupled *c, int32_t x, int32_t y, double *out);
int prh_coupled_stats(prh_coupled *c, int32_t x, double *out);
/* y = A x. */
int prh_coupled_matvec(prh_coupled *c, int32_t x, int32_t y);
/* Net radiative cell power Q of the field in slot T: walls at T_hot/T_cold,
 * uniform emissivity, FP64 reflection series to rtol (coupled.py:86-105,
 * 45-59).  The patch state stays in the handle for prh_coupled_state. */
int prh_coupled_radiation(prh_coupled *c, int32_t T, double T_hot_K, double T_cold_K, double emissivity,
                          double relative_tolerance, int32_t Q, double *series_residual, int32_t *series_iterations);
/* R = A T - b - Q (coupled.py:362-366). */
int prh_coupled_residual(prh_coupled *c, int32_t T, int32_t Q, double T_hot_K, double T_cold_K, int32_t R);
/* Fix the patch temperatures of the field in slot T as the linearisation
 * point of prh_coupled_jacobian. */
int prh_coupled_linearize(prh_coupled *c, int32_t T, double T_hot_K, double T_cold_K);
/* out = A v - dQ/dT v with the walls fixed (coupled.py:108-126, 380-383). */
int prh_coupled_jacobian(prh_coupled *c, int32_t v, double emissivity, double relative_tolerance, int32_t out,
                         int32_t *series_iterations);
/* Precision of the Jacobian-vector products' radiosity solves: PRH_FP64 (the
 * reference's reflection series, default) or PRH_MIXED (FP32 Chebyshev
 * sweeps over a contiguous FP32 vector with FP64 refinement to the same
 * relative tolerance; the radiation of the Newton residual stays an FP64
 * series).  Replaces the
 * inner solves of radiation_jvp inside newton_krylov, porerad/coupled.py:108-126,
 * 362-425, with their precision as a choice. */
int prh_coupled_set_precision(prh_coupled *c, int32_t precision);
/* Fixed-step Jacobi-PCG on A + diag(emission derivative at the field in
 * slot T), started from zero (coupled.py:292-315): set, then apply. */
int prh_coupled_set_preconditioner(prh_coupled *c, int32_t T, double emissivity, int32_t steps);
int prh_coupled_precondition(prh_coupled *c, int32_t rhs, int32_t out);
/* Two modified Gram-Schmidt passes of w against slots basis .. basis+count-1
 * (count <= 60): h[i] = sum of both passes' coefficients, h[count] = ||w||_2
 * (coupled.py:262-268). */
int prh_coupled_orthogonalize(prh_coupled *c, int32_t basis, int32_t count, int32_t w, double *h);
/* out = x + sum_i y[i] v_{basis+i} (count <= 256). */
int prh_coupled_combine(prh_coupled *c, int32_t out, int32_t x, int32_t basis, int32_t count, const double *y);
/* Jacobi-preconditioned CG for A x = b from the x given, until
 * ||b - A x||_2 <= rtol ||b||_2 (conduction.py:100-120); the iterations and
 * the true relative residual come back. */
int prh_coupled_cg(prh_coupled *c, int32_t b, int32_t x, double relative_tolerance, int32_t max_iterations,
                   int32_t *iterations, double *relative_residual);
/* Patch temperatures, emission and absorbed power of the last
 * prh_coupled_radiation (n_patches each; NULL skips). */
int prh_coupled_state(prh_coupled *c, double *patch_temperature, double *emitted, double *absorbed);

#ifdef __cplusplus
}
#endif

#endif /* PORERAD_HIP_H */
