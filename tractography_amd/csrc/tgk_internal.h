// tgk_internal.h -- kernel launchers shared between the kernel translation
// units and the C-ABI layer. Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tgk_device.h"

namespace tgk {

// Sets the thread's tgk_last_error() message and returns `code`.
int report_error(int code, const char *fmt, ...);

// Plain FP32 path (tgk_plain.hip): persistent lanes claiming seeds through
// p.queue (zeroed by the launch), seed-major dense output (compacted by
// launch_compact_seed_major, as the fast path's).
hipError_t launch_track_plain(const TrackLaunch &p, hipStream_t stream);

// Boundary connectome (tgk_plain.hip, plain arithmetic): a mask domain, its
// boundary surface with BVH, 9 (row, col, weight) entries per seed.
struct ConnectomeLaunch {
    MaskView mask;
    FieldView field;
    MeshView boundary;      // boundary triangles and their BVH
    const float *cumulative;
    uint32_t n_triangles;
    AlgoParams algo;
    uint32_t rng_seed;
    int64_t n_seeds;
    uint32_t n_steps;
    float min_length;
    int32_t *rows;
    int32_t *cols;
    float *weights;
    uint32_t *queue;         // seed counter of the persistent lanes (zeroed by the launch)
};
hipError_t launch_connectome(const ConnectomeLaunch &c, hipStream_t stream);
// The last-segment hits and the 9 entries per seed from the records a
// connectome tracking kernel left (tgk_plain.hip).
hipError_t launch_connectome_hits(const ConnectomeLaunch &c, hipStream_t stream);
// The connectome on the fast path (tgk_fast.hip): `coeffs` as in
// FastLaunch; the hits pass is launch_connectome_hits.
struct FastConnectomeLaunch {
    ConnectomeLaunch connectome;
    const void *coeffs;
    uint32_t coeff_stride;
    int32_t sh_order;
};
hipError_t launch_connectome_fast(const FastConnectomeLaunch &f, hipStream_t stream);

// Track orientation distribution (tgk_plain.hip): boundary-seeded tracks
// deposit the SH basis of their orientation on an output grid.
struct TodLaunch {
    MaskView mask;
    FieldView field;
    const float4 *vertices;  // boundary triangles for seeding
    const uint4 *triangles;
    const float *cumulative;
    uint32_t n_triangles;
    AlgoParams algo;
    uint32_t rng_seed;
    int64_t n_seeds;
    uint32_t n_steps;
    MaskView grid;           // output grid: dims and inverse affine (mask unused)
    float *coefficients;     // (X, Y, Z, 45) sums, zeroed before the launch
    uint32_t *queue;         // seed counter of the persistent lanes (zeroed by the launch)
};
hipError_t launch_tod(const TodLaunch &t, hipStream_t stream);
// The TOD on the fast path (tgk_fast.hip): `coeffs` as in FastLaunch.
struct FastTodLaunch {
    TodLaunch tod;
    const void *coeffs;
    uint32_t coeff_stride;
    int32_t sh_order;
};
hipError_t launch_tod_fast(const FastTodLaunch &f, hipStream_t stream);
// Fill n floats with +0 (a kernel, so recorded graphs hold no memset node).
hipError_t launch_zero(float *out, int64_t n, hipStream_t stream);
// Rows of K packed floats (src) -> rows of `stride` floats (dst), pad = +0.
hipError_t launch_pad_rows(const float *src, int64_t n_rows, uint32_t K, float *dst, uint32_t stride,
                           hipStream_t stream);
// reference utils.py:6-22 (normalize_odf) on n_cells rows of 45, in place.
hipError_t launch_normalize_odf(float *coefficients, int64_t n_cells, hipStream_t stream);

// Fast path (tgk_fast.hip): persistent lanes with refill, seed-major output,
// run as FAST_GENERATIONS launches (generation 0 claims seeds, each later
// one resumes the streamlines the previous one saved; see track_fast_kernel).
// Each generation holds at most half the previous one's work, so generation
// g runs on grid / 2^g blocks, and the last runs its work to completion.
constexpr int FAST_GENERATIONS = 4;
// By default a generation stops early only while it has at least twice
// this many streamlines; below that, repacking them gains too little.
constexpr uint32_t FAST_MIN_REPACK = 16384;
// The same on mesh domains, whose tail generations gained from fuller ones.
constexpr uint32_t FAST_MIN_REPACK_MESH = 65536;
// float4s per saved streamline.
constexpr int FAST_CONTINUATION_VEC4 = 4;
// Counters of streamlines that left a generation: FAST_END_SHARDS u32, one
// per FAST_END_STRIDE u32 (128 bytes), after a first 128-byte line holding
// the claim counter.
constexpr int FAST_END_SHARDS = 32;
constexpr int FAST_END_STRIDE = 32;
constexpr size_t FAST_COUNTER_BYTES = 128 + (size_t)FAST_END_SHARDS * FAST_END_STRIDE * 4;
// Launch shapes of the order-8 SH Transport and Diffusion tracking kernels
// (lanes per block, coefficients in LDS), indexed 0..FAST_SHAPE_COUNT - 1:
// (256, 40), (128, 40), (512, 40), (256, 32), (256, 44). Other kernels run
// shape 0.
constexpr int FAST_SHAPE_COUNT = 5;
struct FastLaunch {
    TrackLaunch track;
    uint32_t *claims;           // a generation's claim counter (zeroed per generation)
    uint32_t *ended;            // its ended-streamline shards (zeroed per generation)
    unsigned long long *stats;  // [0] coefficient rows, [1] mask values read from memory (zeroed before generation 0)
    float4 *cont[2];            // saved streamlines, FAST_CONTINUATION_VEC4 float4 each; room for the grid's lanes
    uint32_t *cont_count[2];    // entries in cont[i]
    // Set per generation by launch_track_fast:
    const float4 *cont_in;      // the previous generation's saved streamlines (nullptr: claim seeds)
    const uint32_t *cont_in_count;
    float4 *cont_out;
    uint32_t *cont_out_count;
    int32_t last;               // the last generation runs every streamline to its end
    uint32_t min_repack;        // a generation stops early only while half its work is at least this
    int32_t shared_affine;      // mask domain whose inverse affine equals the field's (rows 0..2)
    int32_t precision;          // TGK_PRECISION_*
    const void *coeffs;         // transformed field (layout per precision)
    uint32_t coeff_stride;      // elements per voxel in `coeffs`
    int32_t sh_order;           // 8 (K <= 45) or 12 (K <= 91): polynomial layout of `coeffs`
    int32_t shape;              // launch shape, 0..FAST_SHAPE_COUNT - 1
};
// grid_blocks <= 0 sizes the persistent grid to the device's occupancy
// (computed once per kernel and device, then cached). n_seeds must be below
// 2^31.
hipError_t launch_track_fast(const FastLaunch &p, int grid_blocks, hipStream_t stream);
// Computes and caches that occupancy for the kernel `p` selects, launching
// nothing (so later launches make no device queries, e.g. while a graph is
// being recorded); *lanes receives the lanes of the grid a launch with
// grid_blocks would use (they bound the saved streamlines of a generation).
hipError_t prepare_track_fast(const FastLaunch &p, int grid_blocks, int64_t *lanes);
// Dense seed-major rows (both paths) -> ragged points.
hipError_t launch_compact_seed_major(const float *dense, int64_t n_seeds, uint32_t n_points,
                                     const uint32_t *lengths, const int64_t *offsets, float *out,
                                     hipStream_t stream);
// Per-voxel coefficients the fast path reads: the SH -> polynomial map of
// order sh_order (8: out_stride >= 45, 12: out_stride >= 91) for SH fields of
// K coefficients, a padded copy for DTI (out_stride 8).
hipError_t launch_field_transform(const float *field, int64_t n_voxels, uint32_t stride, int32_t model, int32_t K,
                                  int32_t sh_order, void *out, uint32_t out_stride, hipStream_t stream);

// One-op drivers (tgk_plain.hip).
hipError_t launch_evaluate(int model, int K, const float *params, const float *orientations, int64_t n, float *out,
                           hipStream_t stream);
hipError_t launch_update(const AlgoParams &a, uint32_t rng_seed, const float *params, const float *orientations,
                         int64_t n, float *out, hipStream_t stream);
hipError_t launch_field_evaluate(const FieldView &f, const float *points, int64_t n, float *out, hipStream_t stream);
hipError_t launch_contains(const MaskView &m, const float *points, int64_t n, uint8_t *out, hipStream_t stream);
// method 0: BVH traversal; 1: per-cell candidate lists where they decide.
hipError_t launch_bvh_hit(const MeshView &m, int method, const float *starts, const float *ends, int64_t n,
                          uint8_t *found, uint32_t *hit, float *t, hipStream_t stream);
hipError_t launch_math(int fn, const float *x, const float *y, int64_t n, float *out, hipStream_t stream);
hipError_t launch_domain_seed(const float4 *vertices, const uint4 *triangles, const float *cumulative,
                              uint32_t n_triangles, uint32_t rng_seed, int64_t n, float *out, hipStream_t stream);

// Seeding from FOD (tgk_seeds.hip). VALU: K <= 45, bit-exact; MFMA: K <= 96.
enum SeedsMethod : int32_t { SEEDS_VALU = 0, SEEDS_MFMA = 1 };
hipError_t launch_seeds_from_fod(int model, int K, const float *fod, const float *voxels, uint32_t nnz,
                                 const float *affine, uint32_t *states, int64_t n, int32_t method, float *out,
                                 void *workspace, size_t workspace_bytes, hipStream_t stream);
size_t seeds_from_fod_workspace_bytes(int model, int K, int64_t n, int32_t method);
// The two halves of launch_seeds_from_fod: the per-(model, K, method) table
// (built and uploaded once; synchronises the stream for the MFMA table), and
// the seeding kernel over a prepared table (enqueue only, graph-safe).
hipError_t prepare_seeds_from_fod(int model, int K, int32_t method, void *table, size_t table_bytes,
                                  hipStream_t stream);
hipError_t launch_seeds_from_fod_prepared(int model, int K, const float *fod, const float *voxels, uint32_t nnz,
                                          const float *affine, uint32_t *states, int64_t n, int32_t method,
                                          float *out, const void *table, hipStream_t stream);

// Exclusive scan of u32 lengths into i64 offsets (offsets[n] = total).
hipError_t launch_offsets(const uint32_t *lengths, int64_t n, int64_t *offsets, void *workspace,
                          size_t workspace_bytes, hipStream_t stream);
size_t offsets_workspace_bytes(int64_t n);

}  // namespace tgk
