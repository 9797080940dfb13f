// tgk_device.h -- device-side building blocks shared by the tracking kernels:
// launch parameter blocks, the MWC64X stream, world->voxel lookups of the mask
// domain and the image field, and the boundary BVH traversal.
//
// Everything here follows the reference semantics operation by operation
// (no contraction: translation units including this header for the plain
// path are compiled with -ffp-contract=off).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tgk_math.h"

namespace tgk {

enum Model : int32_t { MODEL_SH = 0, MODEL_DTI = 1 };
enum Algo : int32_t { ALGO_TRANSPORT = 0, ALGO_DIFFUSION = 1, ALGO_PROBABILISTIC = 2, ALGO_DETERMINISTIC = 3 };
enum DomainKind : int32_t { DOMAIN_MASK = 0, DOMAIN_MESH = 1 };

constexpr uint32_t NO_HIT = 0xffffffffu;

// Field as laid out in HBM: (X, Y, Z) voxels of `stride` floats, the first K
// of which are the model coefficients (stride = K rounded up to 4 so every
// voxel starts 16-byte aligned).
struct FieldView {
    const float *data;
    uint32_t dims[3];
    uint32_t n_parameters;
    uint32_t stride;
    float inv[12];  // rows 0..2 of the float32 inverse affine
};

struct MaskView {
    const uint8_t *mask;
    uint32_t dims[3];
    float inv[12];
};

struct MeshView {
    const float4 *vertices;
    const uint4 *triangles;
    const float4 *bvh_min;
    const float4 *bvh_max;
    const int4 *bvh_nodes;
    const uint32_t *bvh_order;
    uint32_t n_nodes;
    // Culling grid: near[cell] = 0 guarantees that no triangle's bounding box
    // comes within `reach` (world units) of the cell, so a segment that starts
    // in the cell and is at most `reach` long cannot hit the boundary and its
    // BVH query is skipped. Cells are floor(inv * p + 0.5) as for masks.
    const uint8_t *near;
    uint32_t near_dims[3];
    float near_inv[12];
    float reach;
    // Candidate triangles per culling-grid cell (CSR, leaf order): entries
    // are 3 float4 rows (a, slot bits), (b - a), (c - a). Null when absent.
    const uint32_t *cell_start;
    const float4 *cell_entries;
};

struct AlgoParams {
    int32_t model;
    int32_t algorithm;
    int32_t n_parameters;
    float step_size;
    float gamma;
    float noise_variance;
    int32_t n_samples;       // Probabilistic: candidates per update
    float max_angle;         // Probabilistic: cap radius (rad)
    int32_t max_iterations;  // Deterministic: Newton iterations
    float tolerance;         // Deterministic: gradient-norm threshold
    float max_step_size;     // Deterministic: largest update angle (rad)
};

struct TrackLaunch {
    int32_t domain_kind;
    MaskView mask;
    MeshView mesh;
    FieldView field;
    AlgoParams algo;
    const float *seeds;          // (n, 8) device
    int64_t first_seed;          // global index of seeds[0] (RNG stream ids)
    int64_t n_seeds;
    const uint32_t *rng_seeds;   // one per batch of `batch_size` global seeds
    int64_t batch_size;
    uint32_t n_points;           // point capacity per streamline, seed included
    float save_at;
    // Orientation updates after which a streamline is stopped regardless:
    // 2 n_points ceil(save_at / step) + 16, which only a `time` accumulator
    // that stopped advancing (step below one ulp of save_at) can reach.
    uint32_t max_steps;
    uint32_t *lengths;
    uint32_t *steps;
    uint32_t *hits;
    float *points;               // dense output, layout chosen by the kernel
    uint32_t *queue;             // seed counter of the plain kernel's persistent lanes (zeroed by its launch)
};

struct V3 {
    float x, y, z;
};

__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }

// OpenCL length() over xyz: sqrt((x*x + y*y) + z*z).
__device__ __forceinline__ float length3(V3 v) { return tgk_sqrtf((v.x * v.x + v.y * v.y) + v.z * v.z); }

// ------------------------------------------------------------- RNG ------
// MWC64X with A = 4294883355 and the reference's seeding (core.cl:180-214).

__device__ __forceinline__ uint32_t mwc64x(uint2 &s)
{
    const uint32_t A = 4294883355u;
    const uint32_t res = s.x ^ s.y;
    // x A + c as one 64-bit multiply-add (v_mad_u64_u32): low half the new
    // x, high half the new carry.
    const uint64_t next = (uint64_t)s.x * A + s.y;
    s.x = (uint32_t)next;
    s.y = (uint32_t)(next >> 32);
    return res;
}

__device__ __forceinline__ uint32_t rng_mix32(uint32_t v)
{
    v ^= v >> 16;
    v *= 0x7feb352du;
    v ^= v >> 15;
    v *= 0x846ca68bu;
    v ^= v >> 16;
    return v;
}

__device__ __forceinline__ uint2 rng_init(uint32_t seed, uint32_t stream_id)
{
    uint32_t x = rng_mix32(seed ^ (stream_id + 0x9e3779b9u));
    x = x < 1u ? 1u : x;
    const uint32_t c = rng_mix32(seed + stream_id * 0x85ebca6bu) % 4294883355u;
    return make_uint2(x, c);
}

__device__ __forceinline__ float randu(uint2 &s) { return (float)(mwc64x(s) >> 8) * 0x1p-24f; }

// Weighted choice among n candidates (algorithm.py:345-375): the candidates
// come from the noise stream (`candidate(state)` draws and returns one),
// weights are `weight(c)` (non-negative), target = randu * total, and the
// choice is the first candidate whose running weight exceeds the target. The
// reference evaluates every candidate twice (a summing pass, then a replay
// of the stream from its start). Here the summing pass also keeps, for 8
// equal blocks of candidates, the running sum and the noise state at the
// block's start; the second pass replays only the block where the sum
// crosses the target, from its saved state and sum. Every sum is the same
// float addition in the same order as the reference's, so the choice, the
// returned candidate and the final noise state (after the chosen
// candidate's draws) are the same bits, for about n + n/16 evaluations
// instead of up to 2n. Returns false, with the state after the n
// candidates' draws, when the total is not positive or no running sum
// exceeds the target.
template <class Candidate, class Weight>
__device__ __forceinline__ bool choose_candidate(int n, uint2 &rng, Candidate &&candidate, Weight &&weight, V3 &chosen)
{
    constexpr int BLOCKS = 8;
    const int per_block = (n + BLOCKS - 1) / BLOCKS;
    float start_sum[BLOCKS];
    uint2 start_state[BLOCKS];
    uint2 state = rng;
    float total = 0.0f;
#pragma unroll
    for (int b = 0; b < BLOCKS; b++) {
        start_sum[b] = total;
        start_state[b] = state;
        const int end = min(n, (b + 1) * per_block);
        for (int i = b * per_block; i < end; i++)
            total += weight(candidate(state));
    }
    rng = state;
    if (!(total > 0.0f))
        return false;
    uint2 target_state = state;
    const float target = randu(target_state) * total;
    // The first block whose running sum at its end exceeds the target.
    int block = BLOCKS;
#pragma unroll
    for (int b = BLOCKS - 1; b >= 0; b--) {
        const float end_sum = b + 1 < BLOCKS ? start_sum[b + 1] : total;
        if (end_sum > target)
            block = b;
    }
    if (block == BLOCKS)
        return false;
    float cumulative = 0.0f;
    uint2 replay = state;
#pragma unroll
    for (int b = 0; b < BLOCKS; b++)
        if (b == block) {
            cumulative = start_sum[b];
            replay = start_state[b];
        }
    const int end = min(n, (block + 1) * per_block);
    for (int i = block * per_block; i < end; i++) {
        const V3 c = candidate(replay);
        cumulative += weight(c);
        if (cumulative > target) {
            rng = replay;
            chosen = c;
            return true;
        }
    }
    return false;  // unreachable: the block's end sum exceeds the target
}

__device__ __forceinline__ uint32_t randi(uint2 &s, uint32_t max) { return (uint32_t)(randu(s) * (float)max); }

// The reference seeds stream (batch k, index within batch) from the per-batch
// host draw; a global seed index maps to it through the batch size.
__device__ __forceinline__ uint2 seed_stream(const TrackLaunch &p, int64_t global_index)
{
    // A 32-bit divide when both operands fit (a 64-bit one expands to about
    // a hundred instructions, paid by the whole wavefront at each refill).
    if ((((uint64_t)global_index | (uint64_t)p.batch_size) >> 32) == 0) {
        const uint32_t index = (uint32_t)global_index, size = (uint32_t)p.batch_size;
        const uint32_t batch = index / size;
        return rng_init(p.rng_seeds[batch], index - batch * size);
    }
    const int64_t batch = global_index / p.batch_size;
    const uint32_t stream = (uint32_t)(global_index - batch * p.batch_size);
    return rng_init(p.rng_seeds[batch], stream);
}

// ------------------------------------------------- world -> voxel -------

// Row dots of the inverse affine with (p, 1), left to right (field.py:91-96).
__device__ __forceinline__ float affine_row(const float *inv, int row, V3 p)
{
    return ((inv[4 * row] * p.x + inv[4 * row + 1] * p.y) + inv[4 * row + 2] * p.z) + inv[4 * row + 3];
}

// Mask domain membership (domain.py:358-382): cells are half-open,
// index = floor(v + 0.5).
__device__ __forceinline__ bool mask_contains(const MaskView &m, V3 p)
{
    const float rx = floorf(affine_row(m.inv, 0, p) + 0.5f);
    const float ry = floorf(affine_row(m.inv, 1, p) + 0.5f);
    const float rz = floorf(affine_row(m.inv, 2, p) + 0.5f);
    // Bitwise &: all three coordinates are computed anyway, and a branch per
    // test costs more than the tests.
    if (!((rx >= 0.0f) & (rx < (float)m.dims[0]) & (ry >= 0.0f) & (ry < (float)m.dims[1]) & (rz >= 0.0f) &
          (rz < (float)m.dims[2])))
        return false;
    const uint64_t idx = ((uint64_t)(uint32_t)rx * m.dims[1] + (uint32_t)ry) * m.dims[2] + (uint32_t)rz;
    return m.mask[idx] != 0;
}

// Field voxel (field.py:88-114): nearest voxel by rint (ties to even);
// returns the voxel's linear index or -1 outside the image.
__device__ __forceinline__ int64_t field_voxel(const FieldView &f, V3 p)
{
    const float rx = rintf(affine_row(f.inv, 0, p));
    const float ry = rintf(affine_row(f.inv, 1, p));
    const float rz = rintf(affine_row(f.inv, 2, p));
    // Bitwise &: all three coordinates are computed anyway, and a branch per
    // test costs more than the tests.
    if (!((rx >= 0.0f) & (rx < (float)f.dims[0]) & (ry >= 0.0f) & (ry < (float)f.dims[1]) & (rz >= 0.0f) &
          (rz < (float)f.dims[2])))
        return -1;
    return ((int64_t)(uint32_t)rx * f.dims[1] + (uint32_t)ry) * f.dims[2] + (uint32_t)rz;
}

// ------------------------------------------------------- boundary -------
// domain_sample_boundary (domain.py:131-166): an area-weighted triangle by
// binary search of the cumulative areas, a uniform point on it from
// sqrt(randu) and randu, the inward normal as orientation, the point moved
// 1e-4 inward. normalize(v) is v / sqrt(dot(v, v)) per component.
struct BoundarySample {
    V3 point, orientation;
    uint32_t triangle;
    float bary[3];
};

__device__ __forceinline__ BoundarySample sample_boundary(const float4 *__restrict__ vertices,
                                                          const uint4 *__restrict__ triangles,
                                                          const float *__restrict__ cumulative, uint32_t n_triangles,
                                                          uint2 &rng)
{
    const float target = randu(rng);
    uint32_t lower = 0, upper = n_triangles;
    while (lower < upper) {
        const uint32_t middle = lower + (upper - lower) / 2;
        if (cumulative[middle] <= target)
            lower = middle + 1;
        else
            upper = middle;
    }
    BoundarySample s;
    s.triangle = lower < n_triangles - 1 ? lower : n_triangles - 1;
    const uint4 ix = triangles[s.triangle];
    const float4 a = vertices[ix.x], b = vertices[ix.y], c = vertices[ix.z];
    const float root_u = tgk_sqrtf(randu(rng));
    const float v = randu(rng);
    const float wa = 1.0f - root_u, wb = root_u * (1.0f - v), wc = root_u * v;
    s.bary[0] = wa;
    s.bary[1] = wb;
    s.bary[2] = wc;
    const float sx = (wa * a.x + wb * b.x) + wc * c.x;
    const float sy = (wa * a.y + wb * b.y) + wc * c.y;
    const float sz = (wa * a.z + wb * b.z) + wc * c.z;
    const float e1x = b.x - a.x, e1y = b.y - a.y, e1z = b.z - a.z;
    const float e2x = c.x - a.x, e2y = c.y - a.y, e2z = c.z - a.z;
    const float nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
    const float len = tgk_sqrtf((nx * nx + ny * ny) + nz * nz);
    s.orientation = v3(-(nx / len), -(ny / len), -(nz / len));
    s.point = v3(sx + 1e-4f * s.orientation.x, sy + 1e-4f * s.orientation.y, sz + 1e-4f * s.orientation.z);
    return s;
}

// ------------------------------------------------------------ BVH -------
// Closest triangle hit on the segment start -> end with t in (1e-6, 1]
// (bvh.cl:16-131): slab test per node, Moller-Trumbore per triangle, stack of
// 64, right child pushed first so the left subtree is visited first, and a
// triangle at an equal t replaces the current hit.

__device__ __forceinline__ bool bvh_node_hit(const MeshView &m, uint32_t node, const float st[3], const float dir[3],
                                             float max_t)
{
    const float4 lo4 = m.bvh_min[node];
    const float4 hi4 = m.bvh_max[node];
    const float lo[3] = {lo4.x, lo4.y, lo4.z};
    const float hi[3] = {hi4.x, hi4.y, hi4.z};
    float t_min = 0.0f, t_max = max_t;
#pragma unroll
    for (int axis = 0; axis < 3; axis++) {
        if (fabsf(dir[axis]) < 1e-12f) {
            if (st[axis] < lo[axis] || st[axis] > hi[axis])
                return false;
        } else {
            const float inv = 1.0f / dir[axis];
            float near_t = (lo[axis] - st[axis]) * inv;
            float far_t = (hi[axis] - st[axis]) * inv;
            if (near_t > far_t) {
                const float tmp = near_t;
                near_t = far_t;
                far_t = tmp;
            }
            t_min = fmaxf(t_min, near_t);
            t_max = fminf(t_max, far_t);
            if (t_min > t_max)
                return false;
        }
    }
    return true;
}

__device__ __forceinline__ bool bvh_triangle_hit(const MeshView &m, uint32_t tri, const float st[3],
                                                 const float dir[3], float max_t, float &hit_t, float &hit_u,
                                                 float &hit_v)
{
    const uint4 ix = m.triangles[tri];
    const float4 a4 = m.vertices[ix.x];
    const float4 b4 = m.vertices[ix.y];
    const float4 c4 = m.vertices[ix.z];
    const float e1[3] = {b4.x - a4.x, b4.y - a4.y, b4.z - a4.z};
    const float e2[3] = {c4.x - a4.x, c4.y - a4.y, c4.z - a4.z};
    const float p[3] = {dir[1] * e2[2] - dir[2] * e2[1], dir[2] * e2[0] - dir[0] * e2[2],
                        dir[0] * e2[1] - dir[1] * e2[0]};
    const float det = (e1[0] * p[0] + e1[1] * p[1]) + e1[2] * p[2];
    if (fabsf(det) < 1e-10f)
        return false;
    const float inv = 1.0f / det;
    const float off[3] = {st[0] - a4.x, st[1] - a4.y, st[2] - a4.z};
    const float u = ((off[0] * p[0] + off[1] * p[1]) + off[2] * p[2]) * inv;
    if (u < -1e-6f || u > 1.0f + 1e-6f)
        return false;
    const float q[3] = {off[1] * e1[2] - off[2] * e1[1], off[2] * e1[0] - off[0] * e1[2],
                        off[0] * e1[1] - off[1] * e1[0]};
    const float v = ((dir[0] * q[0] + dir[1] * q[1]) + dir[2] * q[2]) * inv;
    if (v < -1e-6f || u + v > 1.0f + 1e-6f)
        return false;
    const float t = ((e2[0] * q[0] + e2[1] * q[1]) + e2[2] * q[2]) * inv;
    if (t <= 1e-6f || t > max_t)
        return false;
    hit_t = t;
    hit_u = u;
    hit_v = v;
    return true;
}

// Closest hit with the barycentric coordinates (1 - u - v, u, v) of the hit
// point on the hit triangle (bvh.cl:75-78).
__device__ inline bool bvh_segment_hit_bary(const MeshView &m, V3 start, V3 end, uint32_t &hit, float &hit_t,
                                            float bary[3])
{
    const float st[3] = {start.x, start.y, start.z};
    const float dir[3] = {end.x - start.x, end.y - start.y, end.z - start.z};
    int32_t stack[64];
    uint32_t size = 1;
    stack[0] = 0;
    hit_t = 1.0f;
    hit = 0;
    bool found = false;
    while (size > 0) {
        const uint32_t node = (uint32_t)stack[--size];
        if (!bvh_node_hit(m, node, st, dir, hit_t))
            continue;
        const int4 nd = m.bvh_nodes[node];
        if (nd.w > 0) {
            for (int i = 0; i < nd.w; i++) {
                const uint32_t tri = m.bvh_order[nd.z + i];
                float t, u, v;
                if (bvh_triangle_hit(m, tri, st, dir, hit_t, t, u, v)) {
                    hit_t = t;
                    hit = tri;
                    bary[0] = 1.0f - u - v;
                    bary[1] = u;
                    bary[2] = v;
                    found = true;
                }
            }
        } else {
            if (size + 2 > 64)
                return found;
            stack[size++] = nd.y;
            stack[size++] = nd.x;
        }
    }
    return found;
}

__device__ inline bool bvh_segment_hit(const MeshView &m, V3 start, V3 end, uint32_t &hit, float &hit_t)
{
    float bary[3];
    return bvh_segment_hit_bary(m, start, end, hit, hit_t, bary);
}

// The culling-grid cell of p, or -1 outside the grid.
__device__ __forceinline__ int64_t mesh_grid_cell(const MeshView &m, V3 p)
{
    const float rx = floorf(affine_row(m.near_inv, 0, p) + 0.5f);
    const float ry = floorf(affine_row(m.near_inv, 1, p) + 0.5f);
    const float rz = floorf(affine_row(m.near_inv, 2, p) + 0.5f);
    if (!((rx >= 0.0f) & (rx < (float)m.near_dims[0]) & (ry >= 0.0f) & (ry < (float)m.near_dims[1]) & (rz >= 0.0f) &
          (rz < (float)m.near_dims[2])))
        return -1;
    return ((int64_t)(uint32_t)rx * m.near_dims[1] + (uint32_t)ry) * m.near_dims[2] + (uint32_t)rz;
}

// Moller-Trumbore on a triangle given as a, b - a, c - a: the arithmetic of
// bvh_triangle_hit, with the two edges precomputed (same float32 values).
__device__ __forceinline__ bool grid_triangle_hit(float4 a4, float4 e14, float4 e24, const float st[3],
                                                  const float dir[3], float max_t, float &hit_t)
{
    const float e1[3] = {e14.x, e14.y, e14.z};
    const float e2[3] = {e24.x, e24.y, e24.z};
    const float p[3] = {dir[1] * e2[2] - dir[2] * e2[1], dir[2] * e2[0] - dir[0] * e2[2],
                        dir[0] * e2[1] - dir[1] * e2[0]};
    const float det = (e1[0] * p[0] + e1[1] * p[1]) + e1[2] * p[2];
    if (fabsf(det) < 1e-10f)
        return false;
    const float inv = 1.0f / det;
    const float off[3] = {st[0] - a4.x, st[1] - a4.y, st[2] - a4.z};
    const float u = ((off[0] * p[0] + off[1] * p[1]) + off[2] * p[2]) * inv;
    if (u < -1e-6f || u > 1.0f + 1e-6f)
        return false;
    const float q[3] = {off[1] * e1[2] - off[2] * e1[1], off[2] * e1[0] - off[0] * e1[2],
                        off[0] * e1[1] - off[1] * e1[0]};
    const float v = ((dir[0] * q[0] + dir[1] * q[1]) + dir[2] * q[2]) * inv;
    if (v < -1e-6f || u + v > 1.0f + 1e-6f)
        return false;
    const float t = ((e2[0] * q[0] + e2[1] * q[1]) + e2[2] * q[2]) * inv;
    if (t <= 1e-6f || t > max_t)
        return false;
    hit_t = t;
    return true;
}

// Closest hit of the segment start + t dir (t in (0, 1]) among candidates
// [first, last) of the culling grid's lists, tested in list order with the
// traversal's replacement rule (t <= best replaces).
__device__ __forceinline__ bool grid_candidates_hit(const MeshView &m, uint32_t first, uint32_t last,
                                                    const float st[3], const float dir[3], uint32_t &hit,
                                                    float &hit_t)
{
    bool found = false;
    for (uint32_t i = first; i < last; i++) {
        const float4 a = m.cell_entries[3 * (size_t)i];
        const float4 e1 = m.cell_entries[3 * (size_t)i + 1];
        const float4 e2 = m.cell_entries[3 * (size_t)i + 2];
        float t;
        if (grid_triangle_hit(a, e1, e2, st, dir, hit_t, t)) {
            hit_t = t;
            hit = __float_as_uint(a.w);
            found = true;
        }
    }
    return found;
}

// Closest hit of a segment no longer than `reach` through the candidate
// lists: the candidates of the start cell hold every triangle whose bounding
// box the segment can touch, in the traversal's leaf order, so the same
// replacement rule (t <= best replaces) gives bvh_segment_hit's answer.
// Returns false with `decided` unset when the lists cannot decide (no lists,
// segment too long, start outside the grid).
__device__ inline bool grid_segment_hit(const MeshView &m, V3 start, V3 end, bool &decided, uint32_t &hit,
                                        float &hit_t)
{
    decided = false;
    hit_t = 1.0f;
    hit = 0;
    if (m.cell_start == nullptr || !(m.reach > 0.0f))
        return false;
    const float st[3] = {start.x, start.y, start.z};
    const float dir[3] = {end.x - start.x, end.y - start.y, end.z - start.z};
    const float len2 = (dir[0] * dir[0] + dir[1] * dir[1]) + dir[2] * dir[2];
    if (!(len2 <= m.reach * m.reach))
        return false;
    const int64_t cell = mesh_grid_cell(m, start);
    if (cell < 0)
        return false;
    decided = true;
    return grid_candidates_hit(m, m.cell_start[cell], m.cell_start[cell + 1], st, dir, hit, hit_t);
}

// True when the culling grid proves that a segment from `start` of squared
// length `len2` cannot reach any boundary triangle.
__device__ __forceinline__ bool mesh_far(const MeshView &m, V3 start, float len2)
{
    if (!(m.reach > 0.0f) || !(len2 <= m.reach * m.reach))
        return false;
    const float rx = floorf(affine_row(m.near_inv, 0, start) + 0.5f);
    const float ry = floorf(affine_row(m.near_inv, 1, start) + 0.5f);
    const float rz = floorf(affine_row(m.near_inv, 2, start) + 0.5f);
    // Bitwise &: all three coordinates are computed anyway, and a branch per
    // test costs more than the tests.
    if (!((rx >= 0.0f) & (rx < (float)m.near_dims[0]) & (ry >= 0.0f) & (ry < (float)m.near_dims[1]) & (rz >= 0.0f) &
          (rz < (float)m.near_dims[2])))
        return false;
    const uint64_t idx =
        ((uint64_t)(uint32_t)rx * m.near_dims[1] + (uint32_t)ry) * m.near_dims[2] + (uint32_t)rz;
    return m.near[idx] == 0;
}

// TOD output-grid cell of p (tracking.py:290-300): floor(inverse_affine p + 0.5)
// per axis, C-order linear index, -1 off the grid.
__device__ __forceinline__ int64_t output_cell(const MaskView &grid, V3 p)
{
    const float rx = floorf(affine_row(grid.inv, 0, p) + 0.5f);
    const float ry = floorf(affine_row(grid.inv, 1, p) + 0.5f);
    const float rz = floorf(affine_row(grid.inv, 2, p) + 0.5f);
    if (!((rx >= 0.0f) & (rx < (float)grid.dims[0]) & (ry >= 0.0f) & (ry < (float)grid.dims[1]) & (rz >= 0.0f) &
          (rz < (float)grid.dims[2])))
        return -1;
    return ((int64_t)(uint32_t)rx * grid.dims[1] + (uint32_t)ry) * grid.dims[2] + (uint32_t)rz;
}

// Run sums staged in LDS for one wavefront: up to one record (45 sums and
// the cell) per lane. Stride 45 (odd) keeps the lanes' writes on distinct
// banks; the drain reads the records as one linear array.
struct RunStage {
    float vals[64 * 45];
    int64_t cell[64];
};

// Flushes the runs of the lanes with `need` (wavefront-converged call): each
// such lane writes its record to the stage, then the whole wavefront adds
// the records to the grid with hardware float atomics (global_atomic_add_f32,
// no return), 64 consecutive (record, coefficient) pairs per instruction, so
// one record's 45 adds touch two or three cache lines in one or two
// instructions instead of 45 instructions of scattered lanes. The reference
// adds the same values one at a time (compare-and-swap loop, core.cl:80-99);
// either way the order in which runs of different tracks land is the
// device's.
__device__ __forceinline__ void flush_runs(bool need, int64_t cell, const float acc[45], float *__restrict__ grid,
                                           RunStage &stage)
{
    const uint64_t mask = __ballot(need);
    if (mask == 0)
        return;
    const int lane = __lane_id();
    if (need) {
        const int slot = __popcll(mask & ((1ull << lane) - 1ull));
#pragma unroll
        for (int i = 0; i < 45; i++)
            stage.vals[slot * 45 + i] = acc[i];
        stage.cell[slot] = cell;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int total = __popcll(mask) * 45;
    for (int p = lane; p < total; p += 64) {
        const int r = p / 45;
        unsafeAtomicAdd(grid + stage.cell[r] * 45 + (p - r * 45), stage.vals[p]);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace tgk
