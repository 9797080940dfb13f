/*
 * tgo.h -- host restatement ("oracle") of the reference tracking path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the HIP kernels are
 * compared against; the product library never links, loads or calls it.
 *
 * Two builds of one source (tgo_impl.h):
 *   tgo32_*  FP32, same operation order as the reference OpenCL kernels, with
 *            the elementary functions of tgk_math.h (the numeric contract the
 *            kernels share). The plain HIP path must match it bit for bit.
 *   tgo64_*  FP64 with the platform libm: the accuracy oracle.
 * Each function cites the reference file:line it restates (paths relative to
 * the reference repository's tractography/ package).
 */
#ifndef TGO_H
#define TGO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { TGO_MODEL_SH = 0, TGO_MODEL_DTI = 1 };
enum { TGO_ALGO_TRANSPORT = 0, TGO_ALGO_DIFFUSION = 1, TGO_ALGO_PROBABILISTIC = 2, TGO_ALGO_DETERMINISTIC = 3 };
enum { TGO_DOMAIN_MASK = 0, TGO_DOMAIN_MESH = 1 };

typedef struct {
    int32_t model;          /* TGO_MODEL_* */
    int32_t algorithm;      /* TGO_ALGO_* */
    int32_t n_parameters;   /* K */
    float step_size;
    float gamma;
    float noise_variance;
    int32_t n_samples;      /* Probabilistic */
    float max_angle;        /* Probabilistic */
    int32_t max_iterations; /* Deterministic */
    float tolerance;        /* Deterministic */
    float max_step_size;    /* Deterministic */
} tgo_algorithm;

typedef struct {
    int32_t domain_kind;            /* TGO_DOMAIN_* */
    /* mask domain (domain.py:358-382) */
    const uint8_t *mask;            /* (X, Y, Z) C order, nonzero = inside */
    uint32_t mask_dims[3];
    float mask_inverse_affine[16];  /* row-major float32 */
    /* image field (field.py:88-114) */
    const float *field;             /* (X, Y, Z, K) C order */
    uint32_t field_dims[4];
    float field_inverse_affine[16];
    /* boundary mesh and BVH (domain.py:44-116, bvh.cl) for mesh domains */
    const float *vertices;          /* (V, 4) */
    const uint32_t *triangles;      /* (T, 4) */
    const float *bvh_min;           /* (N, 4) */
    const float *bvh_max;           /* (N, 4) */
    const int32_t *bvh_nodes;       /* (N, 4) left, right, first, count */
    const uint32_t *bvh_order;      /* leaf triangle slots */
    uint32_t n_bvh_nodes;
} tgo_scene;

/* Elementary functions of the FP32 contract; fn: 0 sin, 1 cos, 2 atan2(x, y),
 * 3 acos, 4 exp, 5 log, 6 sqrt. */
void tgo_math32(int fn, const float *x, const float *y, float *out, int64_t n);

/* RNG (core.cl:180-245). kind: 0 raw MWC64X u32, 1 randu f32, 2 randn f32,
 * 3 randn f64, 4 randi(max=aux) u32. */
void tgo_rng_init(uint32_t seed, uint32_t stream_id, uint32_t state[2]);
void tgo_rng_draws(uint32_t state[2], int kind, uint32_t aux, int64_t n, void *out);

/* SH basis l<=8 (spharm.cl:216-413): outputs (n, 45) each. */
void tgo32_sh45(const float *phi, const float *theta, int64_t n, float *ylm, float *dp, float *dt);
void tgo64_sh45(const double *phi, const double *theta, int64_t n, double *ylm, double *dp, double *dt);
/* FP64 real SH of any even order lmax <= 16 (core.py:36-69 convention), K =
 * (lmax + 1)(lmax + 2) / 2 values per point; extends the FP64 models past
 * the reference's l = 8 (tgo64_evaluate/update/track take K up to 153). */
void tgo64_sh_basis(int lmax, const double *phi, const double *theta, int64_t n, double *ylm, double *dp, double *dt);

/* cart2sph (core.cl:142-153): out (n, 2) azimuth, colatitude. */
void tgo32_cart2sph(const float *xyz, int64_t n, float *out);

/* Local model evaluation (local.py:91-120, 212-230): params (n, K),
 * orientations (n, 3), out (n, 3) value, dtheta, dphi/sin(theta). */
void tgo32_evaluate(int model, int K, const float *params, const float *orientations, int64_t n, float *out);
void tgo64_evaluate(int model, int K, const double *params, const double *orientations, int64_t n, double *out);

/* One orientation update per row (algorithm.py:739-758): stream id = row. */
void tgo32_update(const tgo_algorithm *algorithm, uint32_t rng_seed, const float *params,
                  const float *orientations, int64_t n, float *out);
void tgo64_update(const tgo_algorithm *algorithm, uint32_t rng_seed, const double *params,
                  const double *orientations, int64_t n, double *out);

/* Field and domain one-op drivers (field.py:158-175, domain.py:457-468). */
void tgo_field_evaluate(const tgo_scene *scene, const float *points, int64_t n, float *out);
void tgo_domain_contains(const tgo_scene *scene, const float *points, int64_t n, uint8_t *out);

/* Closest boundary hit on a segment (bvh.cl:83-131). Returns found flag per
 * row; hit (n) u32 triangle slot, t (n) f32. */
void tgo_bvh_segment_hit(const tgo_scene *scene, const float *starts, const float *ends, int64_t n,
                         uint8_t *found, uint32_t *hit, float *t);

/* Tracking (tracking.py:94-149). Seed g draws its noise from
 * rng_init(seed_rng[g], seed_stream[g]); the reference's batch loop
 * (__init__.py:524-539) gives global seed i the pair
 * (draw[i / batch_size], i % batch_size). points: (n, n_points, 3). hits:
 * per seed the boundary triangle slot crossed (mesh domains), 0xffffffff when
 * none. FP64: margins (n, n_points) = running minimum over steps of the
 * distance (voxel units) from the field/domain voxel coordinate to the
 * nearest rounding boundary, up to each saved point. */
void tgo32_track(const tgo_scene *scene, const tgo_algorithm *algorithm, const float *seeds, int64_t n,
                 const uint32_t *seed_rng, const uint32_t *seed_stream, uint32_t n_points, float save_at,
                 uint32_t *lengths, uint32_t *steps, uint32_t *hits, float *points, int n_threads);
void tgo64_track(const tgo_scene *scene, const tgo_algorithm *algorithm, const float *seeds, int64_t n,
                 const uint32_t *seed_rng, const uint32_t *seed_stream, uint32_t n_points, float save_at,
                 uint32_t *lengths, uint32_t *steps, uint32_t *hits, double *points, double *margins,
                 int n_threads);
/* tracking.py:457-558 (kernel connectome): scene = mask domain with the
 * boundary's vertices, triangles and BVH; b* = the boundary surface for
 * sampling. rows, cols, weights: 9 per seed (row -1: no contribution). */
void tgo32_connectome(const tgo_scene *scene, const tgo_algorithm *algorithm, const float *bvertices,
                      const uint32_t *btriangles, const float *bcumulative, uint32_t bn, uint32_t rng_seed, int64_t n,
                      uint32_t n_steps, float min_length, int32_t *rows, int32_t *cols, float *weights, int n_threads);
void tgo64_connectome(const tgo_scene *scene, const tgo_algorithm *algorithm, const float *bvertices,
                      const uint32_t *btriangles, const float *bcumulative, uint32_t bn, uint32_t rng_seed, int64_t n,
                      uint32_t n_steps, float min_length, int32_t *rows, int32_t *cols, float *weights, int n_threads);

/* tracking.py:262-357 (kernel tod), serially in seed order: scene = mask
 * domain and field; b* = the boundary surface for sampling; raw (X, Y, Z, 45)
 * zero on entry. abs_sum (X, Y, Z, 45) and flushes (X, Y, Z), optional, zero
 * on entry: per cell the sum of |flushed values| and the flush count. */
void tgo32_tod(const tgo_scene *scene, const tgo_algorithm *algorithm, const float *bvertices,
               const uint32_t *btriangles, const float *bcumulative, uint32_t bn, uint32_t rng_seed, int64_t n,
               uint32_t n_steps, const uint32_t out_dims[3], const float out_inverse_affine[16], float *raw,
               double *abs_sum, uint32_t *flushes);
void tgo64_tod(const tgo_scene *scene, const tgo_algorithm *algorithm, const float *bvertices,
               const uint32_t *btriangles, const float *bcumulative, uint32_t bn, uint32_t rng_seed, int64_t n,
               uint32_t n_steps, const uint32_t out_dims[3], const float out_inverse_affine[16], double *raw,
               double *abs_sum, uint32_t *flushes);

/* One orientation update from an explicit MWC64X state, advanced in place. */
void tgo32_update_state(const tgo_algorithm *algorithm, const float *params, const float orientation[3],
                        uint32_t state[2], float out[3]);
void tgo64_update_state(const tgo_algorithm *algorithm, const double *params, const double orientation[3],
                        uint32_t state[2], double out[3]);

/* Boundary seeding (domain.py:131-166, 530-556): out (n, 8). */
void tgo_domain_seed(const float *vertices, const uint32_t *triangles, const float *cumulative,
                     uint32_t n_triangles, uint32_t rng_seed, int64_t n, float *out);

/* FOD seeding (seeds.cl:12-118): fod (nnz, K), voxels (nnz, 4), affine 4x4,
 * states (n, 2) updated in place, out (n, 8). */
void tgo_seeds_from_fod(int model, int K, const float *fod, const float *voxels, uint32_t nnz,
                        const float *affine, uint32_t *states, int64_t n, float *out, int n_threads);

#ifdef __cplusplus
}
#endif

#endif /* TGO_H */
