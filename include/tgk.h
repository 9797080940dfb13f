/*
 * tgk.h -- C ABI of the MI355X tracking engine (libtgk.so).
 *
 * The library replaces the OpenCL backend of the reference's tracking path.
 * Each entry point below names the reference interface it stands in for
 * (paths relative to the reference repository's tractography/ package).
 * Python binds these with ctypes (tractography_amd/_native.py); any FFI that
 * can pass plain pointers and sizes can do the same.
 *
 * Conventions
 *   - Every function returns an int status: TGK_OK (0) or a negative
 *     TGK_ERR_* category; tgk_last_error() then describes the failure
 *     (thread-local, valid until the next call on the same thread).
 *   - Host arrays are C-contiguous and owned by the caller. Device memory is
 *     owned by the runtime/field/domain/tracks handles and freed by their
 *     destroy functions.
 *   - Functions taking `d_` pointers operate on device memory and enqueue on
 *     the runtime's stream without synchronising; all others are blocking.
 *   - One runtime is one GPU and one in-order HIP stream. A handle is not
 *     re-entrant; distinct runtimes may be used from distinct threads.
 */
#ifndef TGK_H
#define TGK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TGK_OK = 0,
    TGK_ERR_INVALID = -1,     /* bad argument, shape or size */
    TGK_ERR_NO_MEMORY = -2,   /* device or host allocation failed */
    TGK_ERR_DEVICE = -3,      /* HIP runtime or kernel failure */
    TGK_ERR_UNSUPPORTED = -4  /* valid request this build cannot serve */
};

enum { TGK_MODEL_SH = 0, TGK_MODEL_DTI = 1 };
enum { TGK_ALGO_TRANSPORT = 0, TGK_ALGO_DIFFUSION = 1, TGK_ALGO_PROBABILISTIC = 2, TGK_ALGO_DETERMINISTIC = 3 };

/* Tracking paths. PLAIN is the reference's integrator, bit-exact with the
 * FP32 host restatement. FAST is the persistent, lane-refilling integrator
 * with a transcendental-free Cartesian model evaluation and FP64 arc length;
 * it is judged against the FP64 oracle. */
enum { TGK_PATH_PLAIN = 0, TGK_PATH_FAST = 1 };

/* Storage of the coefficients read by the FAST path. */
enum { TGK_PRECISION_FP32 = 0, TGK_PRECISION_BF16X2 = 1, TGK_PRECISION_FP16 = 2, TGK_PRECISION_BF16 = 3 };

/* Seeding-from-FOD methods: VALU reproduces seeds.cl bit for bit; MFMA
 * evaluates all 1000 directions as one BF16 (hi + lo split) GEMM. */
enum { TGK_SEEDS_VALU = 0, TGK_SEEDS_MFMA = 1 };

typedef struct tgk_runtime tgk_runtime;
typedef struct tgk_field tgk_field;
typedef struct tgk_domain tgk_domain;
typedef struct tgk_tracks tgk_tracks;

/* Orientation-update algorithm and its local model: the kernel arguments of
 * Transport (algorithm.py:115-117), Diffusion (algorithm.py:231-237),
 * Probabilistic (algorithm.py:378-380) and Deterministic
 * (algorithm.py:619-626) plus the model selection of local.from_shape
 * (local.py:265-282). Fields an algorithm does not use are ignored. */
typedef struct {
    int32_t model;          /* TGK_MODEL_* */
    int32_t algorithm;      /* TGK_ALGO_* */
    int32_t n_parameters;   /* K: 1..45 for SH, 6 for DTI */
    float step_size;
    float gamma;            /* Transport, Diffusion */
    float noise_variance;   /* Diffusion */
    int32_t n_samples;      /* Probabilistic: candidates per update, >= 1 */
    float max_angle;        /* Probabilistic: cap radius in [0, pi] */
    int32_t max_iterations; /* Deterministic: >= 0 */
    float tolerance;        /* Deterministic: gradient-norm threshold */
    float max_step_size;    /* Deterministic: largest update angle */
} tgk_algorithm;

const char *tgk_last_error(void);
int tgk_version(void);

/* ------------------------------------------------------------ runtime --- */

/* Replaces OpenCLRuntime.__init__ / finish (opencl.py:30-84, 163). */
int tgk_runtime_create(int32_t device, tgk_runtime **out);
int tgk_runtime_destroy(tgk_runtime *rt);
int tgk_runtime_synchronize(tgk_runtime *rt);
/* Device name (NUL-terminated, truncated to name_len), compute units, HBM. */
int tgk_runtime_info(tgk_runtime *rt, char *name, int32_t name_len, int32_t *compute_units, int64_t *total_bytes);
/* The runtime's HIP stream (hipStream_t) for callers sharing device buffers. */
int tgk_runtime_stream(tgk_runtime *rt, void **stream);
/* Grid of the FAST path's persistent kernel in blocks (of the launch
 * shape's lanes, 256 by default); <= 0 (the default) sizes it to the
 * device's occupancy. No counterpart in the reference, whose kernel is
 * launched one work-item per seed; set by tractography_amd.autotune. */
int tgk_runtime_set_fast_grid(tgk_runtime *rt, int32_t blocks);
/* No counterpart in the reference (one work-item per seed, tracking.py:174).
 * Launch shape of the FAST path's order-8 SH Transport and Diffusion kernels:
 * 0..TGK_FAST_SHAPES - 1, each a number of lanes per block (streamlines per
 * workgroup) and of a row's 45 coefficients kept in LDS between steps (the
 * rest in registers); tgk_fast_shape_info gives both. Other kernels run
 * shape 0. Results do not depend on it; set by tractography_amd.autotune. */
#define TGK_FAST_SHAPES 5
int tgk_runtime_set_fast_shape(tgk_runtime *rt, int32_t shape);
int tgk_fast_shape_info(int32_t shape, int32_t *lanes_per_block, int32_t *lds_coefficients);
/* No counterpart in the reference. The FAST path runs a tracking call as up
 * to four launches ("generations"): once a generation's streamlines are all
 * started and fewer than half of them are still in flight, it saves those
 * and the next generation resumes them packed into full wavefronts. A
 * generation is cut short only while half its streamlines number at least
 * min_streamlines (default 16384 for SH models, 65536 on mesh domains; DTI
 * calls run in one generation); a large value runs every call in one
 * launch. Results do not depend on it. */
int tgk_runtime_set_fast_repack(tgk_runtime *rt, int32_t min_streamlines);
/* No counterpart in the reference (its kernel reads a voxel row every step):
 * how many coefficient rows and mask values the last FAST-path tracking
 * launch on this runtime read from memory, i.e. how often a lane's cached
 * cell changed. Synchronises the runtime's stream. Used to price the
 * kernel's memory traffic. */
int tgk_runtime_fast_reads(tgk_runtime *rt, uint64_t *rows, uint64_t *mask_values);
/* No counterpart in the reference. Kernel-only timing for launch tuning:
 * uploads `seeds` (n, 8) once, tracks them once untimed and then `repeats`
 * times on the runtime stream (one batch, noise-stream seed 0), and returns
 * the fastest launch in milliseconds (HIP events around the tracking
 * launches only) and its orientation-update count. Allocates, so not while
 * a graph is being recorded. Used by tractography_amd.autotune. */
int tgk_track_timed(tgk_runtime *rt, const tgk_domain *domain, const tgk_field *field, const tgk_algorithm *algorithm,
                    int32_t path, const float *seeds, int64_t n_seeds, uint32_t n_points, float save_at,
                    int32_t repeats, float *milliseconds, uint64_t *steps);

/* ------------------------------------------------------ field, domain --- */

/* Replaces field.Image.get_args (field.py:117-122): uploads the (X, Y, Z, K)
 * float32 image once; `inverse_affine` is the row-major float32 4x4. */
int tgk_field_create(tgk_runtime *rt, const float *data, const uint32_t dims[4], const float inverse_affine[16],
                     tgk_field **out);
int tgk_field_destroy(tgk_field *field);

/* Replaces domain.Image.get_args (domain.py:385-398): mask (X, Y, Z), any
 * nonzero byte is inside. */
int tgk_mask_domain_create(tgk_runtime *rt, const uint8_t *mask, const uint32_t dims[3],
                           const float inverse_affine[16], tgk_domain **out);

/* Replaces BaseDomain.get_args's boundary and BVH buffers
 * (domain.py:170-199) for a domain bounded by triangle meshes: tracking stops
 * when a step's segment hits a boundary triangle (bvh.cl:83-131).
 * vertices (V, 4) f32, triangles (T, 4) u32, cumulative (T,) f32 normalised
 * areas, BVH node bounds (N, 4) f32, nodes (N, 4) i32 (left, right, first,
 * count), order (M,) u32. The culling grid (near, dims, inverse affine) flags
 * cells within `reach` of a triangle bounding box; it may be NULL. */
int tgk_mesh_domain_create(tgk_runtime *rt, const float *vertices, uint32_t n_vertices, const uint32_t *triangles,
                           const float *cumulative, uint32_t n_triangles, const float *bvh_min, const float *bvh_max,
                           const int32_t *bvh_nodes, uint32_t n_nodes, const uint32_t *bvh_order, uint32_t n_order,
                           const uint8_t *near, const uint32_t near_dims[3], const float near_inverse_affine[16],
                           float reach, tgk_domain **out);

/* BVH over triangle bounding boxes (domain.py:72-116), built on the host:
 * leaves hold <= leaf_size triangles; an inner node splits on the longest
 * axis of its box (first axis on ties) at the median of a stable centroid
 * sort; nodes are numbered in pre-order, left first. tgk_bvh_node_count
 * gives N for n_triangles (it depends on the count alone). Inputs are the
 * per-triangle box corners (n_triangles, 3) f64; outputs as taken by
 * tgk_mesh_domain_create: bounds (N, 4) f32 (w = 0), nodes (N, 4) i32
 * (left, right, first, count; left = right = -1 for leaves), ordered
 * triangle slots (n_triangles,) u32. Host only: no device is touched. */
int tgk_bvh_node_count(uint32_t n_triangles, uint32_t leaf_size, uint32_t *n_nodes);
int tgk_bvh_build(const double *triangle_min, const double *triangle_max, uint32_t n_triangles, uint32_t leaf_size,
                  float *bounds_min, float *bounds_max, int32_t *nodes, uint32_t n_nodes, uint32_t *ordered);

/* Per-cell candidate lists of a mesh domain, built on the host for
 * tgk_domain_set_segment_grid (no counterpart in the reference, whose
 * tracking always traverses the BVH, bvh.cl:83-131). Triangle slot t is a
 * candidate of every cell of its inclusive box lo[t]..hi[t] (n_triangles, 3)
 * i64 on a grid of dims cells, numbered (x * dims[1] + y) * dims[2] + z.
 * _count fills cell_start (n_cells + 1) u32 CSR offsets and the entry count;
 * _fill writes entries (n_entries, 12) f32 rows (a.xyz, slot bits,
 * (b - a).xyz, 0, (c - a).xyz, 0), in BVH leaf order (leaf_order, the BVH's
 * ordered slots) within each cell. vertices (V, 4) f32, triangles (T, 4)
 * u32 as for tgk_mesh_domain_create. Host only. */
int tgk_segment_grid_count(const int64_t dims[3], const int64_t *lo, const int64_t *hi, const uint32_t *leaf_order,
                           uint32_t n_triangles, uint32_t *cell_start, uint64_t *n_entries);
int tgk_segment_grid_fill(const int64_t dims[3], const int64_t *lo, const int64_t *hi, const uint32_t *leaf_order,
                          uint32_t n_triangles, const float *vertices, uint32_t n_vertices, const uint32_t *triangles,
                          const uint32_t *cell_start, float *entries, uint64_t n_entries);

/* Attaches the boundary surface used by boundary seeding (domain.py:48-70)
 * to a mask domain; mesh domains already carry theirs. */
int tgk_domain_set_boundary(tgk_domain *domain, const float *vertices, uint32_t n_vertices,
                            const uint32_t *triangles, const float *cumulative, uint32_t n_triangles);
/* Per-cell candidate triangles on the culling grid's cells (the grid given
 * to tgk_mesh_domain_create): cell_start (n_cells + 1) u32 CSR offsets,
 * entries (n_entries, 12) f32 rows (a.xyz, slot bits, (b - a).xyz, 0,
 * (c - a).xyz, 0) per candidate, in BVH leaf order within a cell. The FAST
 * path answers segments up to `reach` long from these lists instead of a
 * traversal. The grid must hold fewer than 2^32 - 2 cells, each listing
 * fewer than 2^26 candidates (TGK_ERR_INVALID otherwise). No counterpart in the reference, whose kernel traverses the
 * BVH for every step (bvh.cl:83-131). */
int tgk_domain_set_segment_grid(tgk_domain *domain, const uint32_t *cell_start, uint64_t n_cells,
                                const float *entries, uint64_t n_entries);
int tgk_domain_destroy(tgk_domain *domain);

/* ------------------------------------------------------------ tracking --- */

/* Replaces tracking.tractogram (tracking.py:22-178) and the batch loop of
 * tractography.tractogram (__init__.py:524-539): seeds (n, 8) f32 rows
 * (x, y, z, 1, dx, dy, dz, 0); seed g draws its noise stream from
 * rng_init(rng_seeds[g / batch_size], g % batch_size) like the reference's
 * per-batch get_rng_seed(); n_points = floor(max_length / save_at) is the
 * point capacity including the seed. Results stay in the returned handle:
 * the per-seed counts and the ragged points in host memory (page-locked,
 * filled by DMA as each chunk is compacted; device memory only when
 * page-locked memory cannot be had). The runtime keeps its working device
 * buffers for later calls. */
int tgk_track(tgk_runtime *rt, const tgk_domain *domain, const tgk_field *field, const tgk_algorithm *algorithm,
              int32_t path, const float *seeds, int64_t n_seeds, const uint32_t *rng_seeds, int64_t batch_size,
              uint32_t n_points, float save_at, tgk_tracks **out);
/* Number of streamlines and of saved points over all of them. */
int tgk_tracks_size(const tgk_tracks *tracks, int64_t *n_seeds, int64_t *total_points);
/* Copies per-seed saved-point counts, orientation-update counts, crossed
 * boundary triangle (0xffffffff when none) and the ragged points
 * (total_points, 3) in seed order. Any output pointer may be NULL. The
 * points come straight from device memory: the call blocks until they are
 * in `points`. */
int tgk_tracks_copy(const tgk_tracks *tracks, uint32_t *lengths, uint32_t *steps, uint32_t *hits, float *points);
/* No counterpart in the reference (its points come back in a dense array):
 * the ragged points (total_points, 3) in place, when the handle holds them
 * in one host buffer (a call tracked in one chunk), else NULL (use
 * tgk_tracks_copy). Valid until tgk_tracks_destroy. */
int tgk_tracks_points(const tgk_tracks *tracks, const float **points);
int tgk_tracks_destroy(tgk_tracks *tracks);

/* Device-resident tracking of one chunk, asynchronous on the runtime stream.
 * d_seeds (n, 8); first_seed is the global index of d_seeds[0] (noise
 * streams); d_dense holds n * n_points * 3 floats (layout private to the
 * path); then tgk_compact_device turns it into ragged points. */
int tgk_track_device(tgk_runtime *rt, const tgk_domain *domain, const tgk_field *field,
                     const tgk_algorithm *algorithm, int32_t path, const float *d_seeds, int64_t first_seed,
                     int64_t n_seeds, const uint32_t *d_rng_seeds, int64_t batch_size, uint32_t n_points,
                     float save_at, uint32_t *d_lengths, uint32_t *d_steps, uint32_t *d_hits, float *d_dense);
/* d_offsets (n + 1) receives the exclusive prefix sum of d_lengths (the last
 * entry is the total); d_points receives the ragged (total, 3) points and
 * must hold up to n * n_points * 3 floats. */
int tgk_compact_device(tgk_runtime *rt, int32_t path, const float *d_dense, int64_t n_seeds, uint32_t n_points,
                       const uint32_t *d_lengths, int64_t *d_offsets, float *d_points);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): tgk_compact_device with the offsets scan's workspace
 * in a fixed 64 KiB device array of the library instead of the runtime's
 * growable buffer, so the call makes no allocation of any kind. Same
 * arguments and results as tgk_compact_device; n_seeds above
 * TGK_COMPACT_NOALLOC_MAX_SEEDS fails with TGK_ERR_INVALID. The array is
 * one per GPU, shared by every runtime on it: calls that overlap on one GPU
 * (on different streams or runtimes) race and must be serialised by the
 * caller. */
#define TGK_COMPACT_NOALLOC_MAX_SEEDS 16775168LL
int tgk_compact_device_noalloc(tgk_runtime *rt, int32_t path, const float *d_dense, int64_t n_seeds,
                               uint32_t n_points, const uint32_t *d_lengths, int64_t *d_offsets, float *d_points);

/* Storage precision of the coefficients the FAST path reads from `field`.
 * Only TGK_PRECISION_FP32 (the default) is served: a FAST-path call on a
 * field set to a 16-bit storage fails with TGK_ERR_UNSUPPORTED. On the
 * 145x174x145 benchmark field, coefficients rounded to FP16 or BF16 move
 * FP64 trajectories by 1.2e-2 and 2.4e-2 voxel (the budget is 1e-3); split
 * hi + lo BF16 moves them by 9.7e-4, leaving no room for the kernel's own
 * rounding, while storing as many bytes as FP32. */
int tgk_field_set_fast_precision(tgk_field *field, int32_t precision);

/* -------------------------------------------------------------- seeds --- */

/* Replaces the kernel of tracking.connectome (tracking.py:388-602):
 * n_seeds boundary-seeded tracks of at most n_steps steps in the mask
 * `domain`, seeded on and ended against `boundary` (a mesh domain over the
 * domain's boundary surface, with cumulative areas and BVH); seed g uses
 * rng_init(rng_seed, g) for its boundary sample and its noise. rows, cols,
 * weights (n_seeds * 9): boundary-vertex pairs and barycentric weight
 * products, row -1 where a track contributes nothing. `path` as for
 * tgk_track: TGK_PATH_PLAIN is the reference's FP32 arithmetic,
 * TGK_PATH_FAST the polynomial model (SH up to 91 coefficients). */
int tgk_connectome(tgk_runtime *rt, const tgk_domain *domain, const tgk_domain *boundary, const tgk_field *field,
                   const tgk_algorithm *algorithm, int32_t path, uint32_t rng_seed, int64_t n_seeds,
                   uint32_t n_steps, float minimum_length, int32_t *rows, int32_t *cols, float *weights);

/* Replaces the kernel of tracking.tod (tracking.py:181-385): n_seeds
 * boundary-seeded tracks of at most n_steps steps in the mask `domain`,
 * seeded on `boundary` (any domain handle that carries the boundary
 * triangles and cumulative areas, e.g. the mask domain itself); seed g uses
 * rng_init(rng_seed, g). At every step inside the domain and the field, the
 * 45-term SH basis of the current orientation is added to the output cell
 * floor(output_inverse_affine * p + 0.5) when that cell is on the
 * (X, Y, Z) = output_dims grid. coefficients: (X, Y, Z, 45) float32, C order;
 * with normalize != 0 each cell is scaled as utils.normalize_odf
 * (utils.py:6-22) does. The sums are float atomics, so their rounding depends
 * on the order in which tracks reach a cell, as the reference's do. `path`
 * as for tgk_connectome; on TGK_PATH_FAST the basis is evaluated without
 * angles (a normalised Legendre recurrence in cos(theta) times powers of
 * x + iy). */
int tgk_tod(tgk_runtime *rt, const tgk_domain *domain, const tgk_domain *boundary, const tgk_field *field,
            const tgk_algorithm *algorithm, int32_t path, uint32_t rng_seed, int64_t n_seeds, uint32_t n_steps,
            const uint32_t output_dims[3], const float output_inverse_affine[16], int32_t normalize,
            float *coefficients);
/* The same on the runtime stream, into device memory the caller owns
 * (X * Y * Z * 45 floats from hipMalloc, cleared by the call: the sums are
 * device float atomics). No allocation, so it may be recorded into a graph. */
int tgk_tod_device(tgk_runtime *rt, const tgk_domain *domain, const tgk_domain *boundary, const tgk_field *field,
                   const tgk_algorithm *algorithm, int32_t path, uint32_t rng_seed, int64_t n_seeds,
                   uint32_t n_steps, const uint32_t output_dims[3], const float output_inverse_affine[16],
                   int32_t normalize, float *d_coefficients);
/* Replaces utils.normalize_odf (utils.py:6-22) on n_cells rows of 45 SH
 * coefficients, in place: row / sqrt(4 pi) / row[0] (float64 arithmetic,
 * one rounding to float32), rows with row[0] == 0 set to 0. */
int tgk_normalize_odf(tgk_runtime *rt, int64_t n_cells, float *coefficients);

/* Replaces domain.seed (domain.py:494-575): n boundary seeds (n, 8). */
int tgk_domain_seed(tgk_runtime *rt, const tgk_domain *domain, uint32_t rng_seed, int64_t n, float *out);

/* Replaces the OpenCL branch of seeds.from_fod (seeds.py:174-216): sparse
 * FOD (nnz, K), voxel indices (nnz, 4) with w = 1, forward affine 4x4,
 * states (n, 2) u32 (updated in place), out (n, 8). method TGK_SEEDS_*:
 * VALU takes SH K <= 45 or DTI; MFMA takes SH K <= 96 (order 12 = 91
 * included) or DTI. The voxel draw is clamped to nnz - 1, which changes no
 * result for nnz <= 2^24. */
int tgk_seeds_from_fod(tgk_runtime *rt, int32_t model, int32_t n_parameters, const float *fod, const float *voxels,
                       uint32_t nnz, const float affine[16], uint32_t *states, int64_t n, int32_t method, float *out);

/* Device-pointer form of tgk_seeds_from_fod, enqueued on the runtime stream
 * without synchronising: d_fod (nnz, K), d_voxels (nnz, 4), d_affine 4x4,
 * d_states (n, 2) advanced in place, d_out (n, 8). The seeding table for
 * (model, K, method) is built on first use or by tgk_seeds_prepare. */
int tgk_seeds_from_fod_device(tgk_runtime *rt, int32_t model, int32_t n_parameters, const float *d_fod,
                              const float *d_voxels, uint32_t nnz, const float *d_affine, uint32_t *d_states,
                              int64_t n, int32_t method, float *d_out);
int tgk_seeds_prepare(tgk_runtime *rt, int32_t model, int32_t n_parameters, int32_t method);

/* ------------------------------------------------------------ HIP graphs --- */

/* No counterpart in the reference (its OpenCL queue replays nothing): for
 * callers that run the same device pipeline many times, e.g. seeding,
 * tracking and compaction over chunks of a large seed set. Between
 * tgk_graph_begin and tgk_graph_end the runtime's stream is recorded, not
 * run; only the *_device entry points may be called, and none of them may
 * allocate: tgk_track_prepare (field transform, work queue, scan workspace
 * for up to max_seeds seeds, kernel occupancy) and tgk_seeds_prepare must
 * come first, or the call fails with TGK_ERR_INVALID. A graph holds the
 * device pointers it was recorded with; tgk_graph_launch replays it on the
 * runtime's stream. */
typedef struct tgk_graph tgk_graph;
int tgk_track_prepare(tgk_runtime *rt, const tgk_domain *domain, const tgk_field *field,
                      const tgk_algorithm *algorithm, int32_t path, int64_t max_seeds);
int tgk_graph_begin(tgk_runtime *rt);
int tgk_graph_end(tgk_runtime *rt, tgk_graph **out);
int tgk_graph_launch(tgk_graph *graph);
int tgk_graph_destroy(tgk_graph *graph);

/* ------------------------------------------------------ one-op drivers --- */

/* Replaces local.evaluate (local.py:285-362): params (n, K), orientations
 * (n, 3), out (n, 3) value, d/dtheta, (d/dphi)/sin(theta). */
int tgk_evaluate(tgk_runtime *rt, int32_t model, int32_t n_parameters, const float *params, const float *orientations,
                 int64_t n, float *out);
/* Replaces algorithm.update_orientation (algorithm.py:701-790). */
int tgk_update_orientation(tgk_runtime *rt, const tgk_algorithm *algorithm, uint32_t rng_seed, const float *params,
                           const float *orientations, int64_t n, float *out);
/* Replaces field.evaluate (field.py:144-197): out (n, K). */
int tgk_field_evaluate(tgk_runtime *rt, const tgk_field *field, const float *points, int64_t n, float *out);
/* Replaces domain.contains (domain.py:433-491) for mask domains: out (n,). */
int tgk_domain_contains(tgk_runtime *rt, const tgk_domain *domain, const float *points, int64_t n, uint8_t *out);
/* Closest boundary hit per segment (bvh.cl:83-131) of a mesh domain. */
int tgk_bvh_segment_hit(tgk_runtime *rt, const tgk_domain *domain, const float *starts, const float *ends, int64_t n,
                        uint8_t *found, uint32_t *hit, float *t);
/* The same query answered from the candidate lists where they decide (the
 * FAST path's method); for the parity tests against the traversal. */
int tgk_grid_segment_hit(tgk_runtime *rt, const tgk_domain *domain, const float *starts, const float *ends,
                         int64_t n, uint8_t *found, uint32_t *hit, float *t);
/* Evaluates the device elementary functions of the numeric contract:
 * fn 0 sin, 1 cos, 2 atan2(x, y), 3 acos, 4 exp, 5 log, 6 sqrt, 7 x / y. */
int tgk_math(tgk_runtime *rt, int32_t fn, const float *x, const float *y, int64_t n, float *out);

#ifdef __cplusplus
}
#endif

#endif /* TGK_H */
