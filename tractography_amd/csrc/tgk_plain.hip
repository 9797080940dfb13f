// tgk_plain.hip -- the plain FP32 tracking path, the boundary-seeded
// products (connectome, TOD) and the one-op drivers.
//
// One integrator per seed, exactly the reference's loop (tracking.py:112-148),
// on persistent lanes that claim the next seed when their streamline ends.
// Compiled with -ffp-contract=off: every result is bit identical to the FP32
// host restatement. The connectome and TOD kernels run the same arithmetic
// on persistent lanes that claim boundary seeds.
//
// Output layout: seed-major, dense[(g * n_points + n) * 3 + c], the fast
// path's layout, so one compaction kernel serves both paths. (A point-major
// layout made each save of a wavefront one contiguous store, but its
// compaction had to gather rows of every point index: 9.2 ms at 1M seeds
// against 2.5 ms, while the saves cost this arithmetic-heavy kernel nothing
// measurable either way.)
#include <hip/hip_runtime.h>

#include "tgk_internal.h"
#include "tgk_plain_model.h"

namespace tgk {

namespace {

__device__ __forceinline__ void store_point(float *dense, uint32_t n_points, uint32_t n, int64_t g, V3 p)
{
    float *dst = dense + (g * (int64_t)n_points + n) * 3;
    dst[0] = p.x;
    dst[1] = p.y;
    dst[2] = p.z;
}

// Waves per SIMD the compiler must fit in the SH tracking kernels, which
// otherwise take 146-148 VGPRs (Transport, Diffusion; three waves), 164-170
// (Probabilistic; three) or 291-295 (Deterministic; one). The Deterministic
// kernel is held to two waves (75% faster at 145x174x145 SH45, spilling
// some 30-50 VGPRs; 3.52 against 5.95 s at one wave on persistent lanes).
// The others: held to four, they spill 10-54 VGPRs. With one lane per seed
// for the whole launch that was still 5% (Transport, Diffusion) and 12%
// (Probabilistic) faster than three waves. On persistent lanes, mask
// domains run faster at three waves (SH45 Diffusion 74.4 against 91.5 ms
// per launch) and mesh domains, whose boundary tests wait on dependent
// node loads, at four (the two-mesh shell 326 against 361 ms). DTI
// kernels keep the compiler's allocation (already below 128 VGPRs, they
// lost 11% when held to exactly four). The connectome and TOD kernels are
// held to two waves for Deterministic only (296 and 342 VGPRs, one wave
// otherwise: 1.52x faster each); the SH Diffusion connectome held to four
// waves ran 11% slower.
template <int MODEL, int ALGO, int DOMAIN> constexpr int plain_min_waves()
{
    if constexpr (MODEL != MODEL_SH)
        return 1;
    else if constexpr (ALGO == ALGO_DETERMINISTIC)
        return 2;
    else
        return DOMAIN == DOMAIN_MESH ? 4 : 3;
}

// ------------------------------------------------------ one-op drivers ---

// Loads K (<= 45) coefficients from an unpadded row.
__device__ __forceinline__ void load_row_scalar(const float *__restrict__ row, int K, float c[45])
{
#pragma unroll
    for (int i = 0; i < 45; i++)
        c[i] = i < K ? row[i] : 0.0f;
}

template <int MODEL>
__device__ __forceinline__ plain::ModelValue evaluate_row(const float *__restrict__ row, int K, V3 o)
{
    if constexpr (MODEL == MODEL_DTI) {
        alignas(16) const float m[8] = {row[0], row[1], row[2], row[3], row[4], row[5], 0.0f, 0.0f};
        return plain::dti_model(m, o);
    } else {
        float c[45];
        load_row_scalar(row, K, c);
        float az, col;
        plain::cart2sph(o, az, col);
        float v, th, ph;
        plain::sh45_contract(c, az, col, v, th, ph);
        const float d = plain::dsoftplus(v);
        plain::ModelValue r;
        r.value = plain::softplus(v);
        r.dtheta = th * d;
        r.dphi = ph * d;
        return r;
    }
}

// local.py:317-334
template <int MODEL>
__global__ void evaluate_kernel(const float *__restrict__ params, int K, const float *__restrict__ orientations,
                                int64_t n, float *__restrict__ out)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n)
        return;
    const V3 o = v3(orientations[3 * g], orientations[3 * g + 1], orientations[3 * g + 2]);
    const plain::ModelValue mv = evaluate_row<MODEL>(params + (int64_t)K * g, K, o);
    out[3 * g] = mv.value;
    out[3 * g + 1] = mv.dtheta;
    out[3 * g + 2] = mv.dphi;
}

// algorithm.py:739-758: stream id = row index.
template <int MODEL, int ALGO>
__global__ void update_kernel(AlgoParams a, uint32_t rng_seed, const float *__restrict__ params,
                              const float *__restrict__ orientations, int64_t n, float *__restrict__ out)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n)
        return;
    const int K = a.n_parameters;
    const float *row = params + (int64_t)K * g;
    // Stage the row in the padded layout update() reads.
    alignas(16) float padded[48];
#pragma unroll
    for (int i = 0; i < 48; i++)
        padded[i] = i < K ? row[i] : 0.0f;
    uint2 rng = rng_init(rng_seed, (uint32_t)g);
    const V3 o = v3(orientations[3 * g], orientations[3 * g + 1], orientations[3 * g + 2]);
    const V3 r = plain::update<MODEL, ALGO, 0>(a, padded, o, rng);
    out[3 * g] = r.x;
    out[3 * g + 1] = r.y;
    out[3 * g + 2] = r.z;
}

// field.py:158-175
__global__ void field_evaluate_kernel(FieldView f, const float *__restrict__ points, int64_t n,
                                      float *__restrict__ out)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n)
        return;
    const int64_t voxel = field_voxel(f, v3(points[3 * g], points[3 * g + 1], points[3 * g + 2]));
    for (uint32_t i = 0; i < f.n_parameters; i++)
        out[(int64_t)f.n_parameters * g + i] = voxel >= 0 ? f.data[voxel * f.stride + i] : 0.0f;
}

// domain.py:457-468
__global__ void contains_kernel(MaskView m, const float *__restrict__ points, int64_t n, uint8_t *__restrict__ out)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n)
        return;
    out[g] = mask_contains(m, v3(points[3 * g], points[3 * g + 1], points[3 * g + 2])) ? 1 : 0;
}

// method 0: BVH traversal; 1: the start cell's candidate list where it
// decides, else the traversal.
__global__ void bvh_hit_kernel(MeshView m, int method, const float *__restrict__ starts,
                               const float *__restrict__ ends, int64_t n, uint8_t *__restrict__ found,
                               uint32_t *__restrict__ hit, float *__restrict__ t)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n)
        return;
    uint32_t tri;
    float ht;
    const V3 a = v3(starts[3 * g], starts[3 * g + 1], starts[3 * g + 2]);
    const V3 b = v3(ends[3 * g], ends[3 * g + 1], ends[3 * g + 2]);
    bool decided = false;
    bool f = method == 1 ? grid_segment_hit(m, a, b, decided, tri, ht) : false;
    if (!decided)
        f = bvh_segment_hit(m, a, b, tri, ht);
    found[g] = f ? 1 : 0;
    hit[g] = tri;
    t[g] = ht;
}

__global__ void math_kernel(int fn, const float *__restrict__ x, const float *__restrict__ y, int64_t n,
                            float *__restrict__ out)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n)
        return;
    const float a = x[g];
    float r;
    switch (fn) {
    case 0: r = tgk_sinf(a); break;
    case 1: r = tgk_cosf(a); break;
    case 2: r = tgk_atan2f(a, y[g]); break;
    case 3: r = tgk_acosf(a); break;
    case 4: r = tgk_expf(a); break;
    case 5: r = tgk_logf(a); break;
    case 6: r = tgk_sqrtf(a); break;
    default: r = a / y[g]; break;
    }
    out[g] = r;
}

// domain.py:131-166 driven as domain.py:530-556; normalize(v) = v / |v|.
__global__ void domain_seed_kernel(const float4 *__restrict__ vertices, const uint4 *__restrict__ triangles,
                                   const float *__restrict__ cumulative, uint32_t n_triangles, uint32_t rng_seed,
                                   int64_t n, float *__restrict__ out)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n)
        return;
    uint2 rng = rng_init(rng_seed, (uint32_t)g);
    const BoundarySample s = sample_boundary(vertices, triangles, cumulative, n_triangles, rng);
    float *row = out + 8 * g;
    row[0] = s.point.x;
    row[1] = s.point.y;
    row[2] = s.point.z;
    row[3] = 1.0f;
    row[4] = s.orientation.x;
    row[5] = s.orientation.y;
    row[6] = s.orientation.z;
    row[7] = 0.0f;
}

// Claims seeds for the idle lanes of the wavefront from a device queue, one
// atomic per refill; returns the lane's new seed index, or -1 when it gets
// none. Sets `drained` (wavefront-uniform) once the queue has passed n.
__device__ __forceinline__ int64_t claim_seed(bool idle, uint32_t *queue, int64_t n, bool &drained)
{
    const uint64_t need = __ballot(idle);
    if (need == 0 || drained)
        return -1;
    const int lane = __lane_id();
    const int leader = __ffsll((unsigned long long)need) - 1;
    const uint32_t count = (uint32_t)__popcll(need);
    uint32_t base = 0;
    if (lane == leader)
        base = atomicAdd(queue, count);
    base = __shfl(base, leader);
    if ((uint64_t)base + count >= (uint64_t)n)
        drained = true;
    if (!idle)
        return -1;
    const uint64_t g = (uint64_t)base + (uint64_t)__popcll(need & ((1ull << lane) - 1ull));
    return g < (uint64_t)n ? (int64_t)g : -1;
}

// tracking.py:112-148 (kernel tractogram), one seed per lane at a time:
// seed g's start from seeds[g], its noise stream from seed_stream, then the
// reference loop (domain test, field voxel, update, move, save every save_at
// of arc length in f32) until a test breaks it or the row or the step count
// is full; its length, step count and boundary hit at the end.
//
// Persistent lanes: a grid that fills the GPU once, whose lanes claim the
// next seed when their streamline ends (claim_seed). With one lane per
// seed for the whole launch, a wavefront issued its longest streamline's
// steps for all 64 lanes, and on a workload of 5 to 3000 steps that left
// about half of them masked off. Each streamline's arithmetic and noise
// stream are those of its seed index, whichever lane runs it, so the
// results do not change.
template <int MODEL, int ALGO, int DOMAIN, int KT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(plain_min_waves<MODEL, ALGO, DOMAIN>())))
void track_plain_kernel(TrackLaunch p)
{
    const float step = p.algo.step_size;
    const float save = p.save_at;
    const uint32_t stride = p.field.stride;
    int64_t g = -1;
    bool drained = false;
    V3 point = v3(0.0f, 0.0f, 0.0f), o = point, previous = point;
    uint2 rng = make_uint2(0u, 0u);
    uint32_t n = 0, steps = 0, hit = NO_HIT;
    float time = 0.0f;
    bool first = true;
    while (true) {
        const int64_t claimed = claim_seed(g < 0, p.queue, p.n_seeds, drained);
        if (claimed >= 0) {
            g = claimed;
            const float4 s0 = reinterpret_cast<const float4 *>(p.seeds)[2 * g];
            const float4 s1 = reinterpret_cast<const float4 *>(p.seeds)[2 * g + 1];
            point = v3(s0.x, s0.y, s0.z);
            o = v3(s1.x, s1.y, s1.z);
            if constexpr (ALGO == ALGO_DIFFUSION || ALGO == ALGO_PROBABILISTIC)
                rng = seed_stream(p, p.first_seed + g);
            store_point(p.points, p.n_points, 0, g, point);
            n = 1;
            steps = 0;
            hit = NO_HIT;
            time = 0.0f;
            previous = point;
            first = true;
        }
        if (__ballot(g >= 0) == 0)
            break;
        if (g < 0)
            continue;
        // One pass of the reference loop; `stop` is its break.
        bool stop = !(n < p.n_points && steps < p.max_steps);
        if (!stop) {
            if constexpr (DOMAIN == DOMAIN_MASK) {
                stop = !mask_contains(p.mask, point);
            } else if (!first) {
                const float dx = point.x - previous.x, dy = point.y - previous.y, dz = point.z - previous.z;
                if (!mesh_far(p.mesh, previous, (dx * dx + dy * dy) + dz * dz)) {
                    uint32_t tri;
                    float t;
                    if (bvh_segment_hit(p.mesh, previous, point, tri, t)) {
                        hit = tri;
                        stop = true;
                    }
                }
            }
        }
        int64_t voxel = -1;
        if (!stop) {
            voxel = field_voxel(p.field, point);
            stop = voxel < 0;
        }
        if (!stop) {
            o = plain::update<MODEL, ALGO, KT>(p.algo, p.field.data + voxel * stride, o, rng);
            steps++;
            previous = point;
            point.x += step * o.x;
            point.y += step * o.y;
            point.z += step * o.z;
            first = false;
            time += step;
            if (time >= save) {
                time -= save;
                store_point(p.points, p.n_points, n, g,
                            v3(point.x - time * o.x, point.y - time * o.y, point.z - time * o.z));
                n++;
            }
            continue;
        }
        p.lengths[g] = n;
        p.steps[g] = steps;
        p.hits[g] = hit;
        g = -1;
    }
}

// tracking.py:457-558 (kernel connectome): boundary seed g from
// rng_init(rng_seed, g), whose state then drives the noise; the loop checks
// the mask, then the field, updates, rejects a non-finite or short
// orientation, and moves; afterwards the last segment is tested against the
// boundary BVH and a hit deposits the outer product of the start and hit
// barycentric weights on the two triangles' vertices (9 entries per seed,
// row -1 when none).
//
// Persistent lanes: a grid that fills the GPU once; a lane whose track ends
// claims the next seed, so tracks of 1 to n_steps steps do not leave lanes
// idle until the longest of their wavefront ends. Every track is computed
// exactly as in a one-lane-per-seed launch; only which lane runs it changes.
template <int MODEL, int ALGO, int KT>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(MODEL == MODEL_SH && ALGO == ALGO_DETERMINISTIC ? 2 : 1)))
void connectome_kernel(ConnectomeLaunch c)
{
    const float step = c.algo.step_size;
    const uint32_t stride = c.field.stride;
    int64_t g = -1;
    bool drained = false;
    uint2 rng = make_uint2(0u, 0u);
    V3 point = v3(0.0f, 0.0f, 0.0f), o = point, previous = point;
    uint32_t tracked = 0, start_triangle = 0;
    float start_bary[3] = {0.0f, 0.0f, 0.0f};
    while (true) {
        const int64_t claimed = claim_seed(g < 0, c.queue, c.n_seeds, drained);
        if (claimed >= 0) {
            g = claimed;
            rng = rng_init(c.rng_seed, (uint32_t)g);
            const BoundarySample s = sample_boundary(c.boundary.vertices, c.boundary.triangles, c.cumulative,
                                                     c.n_triangles, rng);
            point = s.point;
            o = s.orientation;
            previous = point;
            tracked = 0;
            start_triangle = s.triangle;
            for (int i = 0; i < 3; i++)
                start_bary[i] = s.bary[i];
        }
        if (__ballot(g >= 0) == 0)
            break;
        if (g < 0)
            continue;
        bool stop = tracked >= c.n_steps || !mask_contains(c.mask, point);
        int64_t voxel = -1;
        if (!stop) {
            voxel = field_voxel(c.field, point);
            stop = voxel < 0;
        }
        if (!stop) {
            o = plain::update<MODEL, ALGO, KT>(c.algo, c.field.data + voxel * stride, o, rng);
            stop = !(__builtin_isfinite(o.x) && __builtin_isfinite(o.y) && __builtin_isfinite(o.z)) ||
                   length3(o) < 0.5f;
        }
        if (!stop) {
            previous = point;
            point.x += step * o.x;
            point.y += step * o.y;
            point.z += step * o.z;
            tracked++;
            continue;
        }
        // The track has ended: its last segment and start go into the
        // seed's own output slots (weights: previous, point, start
        // barycentrics; rows: start triangle, length test) for
        // connectome_hits_kernel, which tests every segment with all lanes
        // busy instead of one ending lane at a time.
        float *record = c.weights + 9 * g;
        record[0] = previous.x;
        record[1] = previous.y;
        record[2] = previous.z;
        record[3] = point.x;
        record[4] = point.y;
        record[5] = point.z;
        record[6] = start_bary[0];
        record[7] = start_bary[1];
        record[8] = start_bary[2];
        c.rows[9 * g] = (int32_t)start_triangle;
        c.rows[9 * g + 1] = (float)tracked * step < c.min_length ? 0 : 1;
        g = -1;
    }
}

// tracking.py:526-558: the last segment of seed g against the boundary BVH
// (closest hit with barycentrics) and the 9 (row, col, weight) entries, from
// the record connectome_kernel left in the seed's slots.
__global__ __launch_bounds__(256) void connectome_hits_kernel(ConnectomeLaunch c)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= c.n_seeds)
        return;
    int32_t *rows = c.rows + 9 * g;
    int32_t *cols = c.cols + 9 * g;
    float *weights = c.weights + 9 * g;
    const V3 previous = v3(weights[0], weights[1], weights[2]);
    const V3 point = v3(weights[3], weights[4], weights[5]);
    const float start_bary[3] = {weights[6], weights[7], weights[8]};
    const uint32_t start_triangle = (uint32_t)rows[0];
    // Every seed's record is written before this kernel runs; the index is
    // still checked so that no record can send a read out of bounds.
    const bool long_enough = rows[1] != 0 && start_triangle < c.n_triangles;
    uint32_t hit = 0;
    float t, bary[3];
    if (long_enough && bvh_segment_hit_bary(c.boundary, previous, point, hit, t, bary)) {
        const uint4 si = c.boundary.triangles[start_triangle];
        const uint4 ei = c.boundary.triangles[hit];
        const uint32_t sv[3] = {si.x, si.y, si.z};
        const uint32_t ev[3] = {ei.x, ei.y, ei.z};
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                rows[3 * i + j] = (int32_t)sv[i];
                cols[3 * i + j] = (int32_t)ev[j];
                weights[3 * i + j] = start_bary[i] * bary[j];
            }
    } else {
        for (int i = 0; i < 9; i++) {
            rows[i] = -1;
            cols[i] = -1;
            weights[i] = 0.0f;
        }
    }
}


// tracking.py:262-357 (kernel tod): boundary seed g from rng_init(rng_seed,
// g), whose state then drives the noise; each step checks the mask and the
// field, then deposits the SH basis of the current orientation
// (spharm.cl:10-207 at cart2sph(orientation)) on the output cell of the
// point, then updates, rejects a non-finite or short orientation, and moves.
// Consecutive steps in one cell are summed in registers (acc = basis on
// entering a cell, acc += basis after) and flushed when the track leaves the
// cell, leaves the grid, or ends. For the SH model the deposited basis is the
// one the update's model evaluation computes at the same orientation (same
// angles, same values), so it is not evaluated twice. Persistent lanes with
// seed claiming, as the connectome kernel; each iteration has two flush
// points (leaving a cell before the update, ending after it) where the
// wavefront is converged.
template <int MODEL, int ALGO, int KT>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(MODEL == MODEL_SH && ALGO == ALGO_DETERMINISTIC ? 2 : 1)))
void tod_kernel(TodLaunch t)
{
    __shared__ RunStage stages[4];
    RunStage &stage = stages[threadIdx.x >> 6];
    const float step = t.algo.step_size;
    const uint32_t stride = t.field.stride;
    int64_t g = -1;
    bool drained = false;
    uint2 rng = make_uint2(0u, 0u);
    V3 point = v3(0.0f, 0.0f, 0.0f), o = point;
    uint32_t k = 0;
    int64_t previous = -1;
    float acc[45];
#pragma unroll
    for (int i = 0; i < 45; i++)
        acc[i] = 0.0f;
    while (true) {
        const int64_t claimed = claim_seed(g < 0, t.queue, t.n_seeds, drained);
        if (claimed >= 0) {
            g = claimed;
            rng = rng_init(t.rng_seed, (uint32_t)g);
            const BoundarySample s = sample_boundary(t.vertices, t.triangles, t.cumulative, t.n_triangles, rng);
            point = s.point;
            o = s.orientation;
            k = 0;
            previous = -1;
        }
        if (__ballot(g >= 0) == 0)
            break;
        const bool live = g >= 0;
        bool stop = false;
        int64_t voxel = -1, cell = -1;
        if (live) {
            stop = k >= t.n_steps || !mask_contains(t.mask, point);
            if (!stop) {
                voxel = field_voxel(t.field, point);
                stop = voxel < 0;
            }
            if (!stop)
                cell = output_cell(t.grid, point);  // -1 off the grid
        }
        // Leaving a cell (for another or for no cell) ends its run.
        flush_runs(live && !stop && previous >= 0 && cell != previous, previous, acc, t.coefficients, stage);
        if (live && !stop) {
            const bool in_grid = cell >= 0;
            const bool same = cell == previous;
            previous = cell;  // -1: no run
            // A select keeps the reference's `= ylm` on entering a cell (not
            // `0 + ylm`, which differs for -0).
            auto add = [&](int i, float y) {
                if (in_grid)
                    acc[i] = same ? acc[i] + y : y;
            };
            // The update's own model evaluation at o supplies the basis for
            // SH under Transport and Diffusion; otherwise evaluate it here.
            constexpr bool fused = MODEL == MODEL_SH && (ALGO == ALGO_TRANSPORT || ALGO == ALGO_DIFFUSION);
            if constexpr (!fused) {
                if (in_grid) {
                    float az, col;
                    plain::cart2sph(o, az, col);
                    plain::sh45_values(az, col, add);
                }
            }
            o = plain::update<MODEL, ALGO, KT>(t.algo, t.field.data + voxel * stride, o, rng, add);
            stop = !(__builtin_isfinite(o.x) && __builtin_isfinite(o.y) && __builtin_isfinite(o.z)) ||
                   length3(o) < 0.5f;
            if (!stop) {
                point.x += step * o.x;
                point.y += step * o.y;
                point.z += step * o.z;
                k++;
            }
        }
        // The end of a track ends its run.
        const bool ended = live && stop;
        flush_runs(ended && previous >= 0, previous, acc, t.coefficients, stage);
        if (ended)
            g = -1;
    }
}

__global__ void zero_u32_kernel(uint32_t *counter)
{
    if (threadIdx.x == 0)
        *counter = 0u;
}

__global__ __launch_bounds__(256) void zero_kernel(float *__restrict__ out, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = 0.0f;
}

// Rows of K packed floats -> rows of `stride` floats (stride >= K), the pad
// written as +0. One thread per output float, so the stores are coalesced.
__global__ __launch_bounds__(256) void pad_rows_kernel(const float *__restrict__ src, int64_t n_rows, uint32_t K,
                                                       float *__restrict__ dst, uint32_t stride)
{
    const int64_t n = n_rows * stride;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = i / stride;
        const uint32_t k = (uint32_t)(i - row * stride);
        dst[i] = k < K ? src[row * K + k] : 0.0f;
    }
}

// utils.py:6-22 (normalize_odf): numpy forms odf / sqrt(4 pi) in float64
// (sqrt(4 pi) is a float64 scalar), divides by the float32 c0 promoted to
// float64 and rounds once to float32 on the store; cells with c0 == 0 become
// 0. A NaN c0 compares unequal to 0 and propagates, as in numpy.
__global__ __launch_bounds__(256) void normalize_odf_kernel(float *__restrict__ c, int64_t n_cells)
{
    const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= n_cells)
        return;
    float *row = c + cell * 45;
    const float c0 = row[0];
    if (c0 != 0.0f) {
        const double sqrt_4pi = 0x1.c5bf891b4ef6ap+1;  // float64(sqrt(4 pi)) = 3.5449077018110318
        for (int i = 0; i < 45; i++)
            row[i] = (float)(((double)row[i] / sqrt_4pi) / (double)c0);
    } else {
        for (int i = 0; i < 45; i++)
            row[i] = 0.0f;
    }
}

inline unsigned blocks_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

}  // namespace

hipError_t launch_evaluate(int model, int K, const float *params, const float *orientations, int64_t n, float *out,
                           hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    if (model == MODEL_DTI)
        evaluate_kernel<MODEL_DTI><<<blocks_for(n, 128), 128, 0, stream>>>(params, K, orientations, n, out);
    else
        evaluate_kernel<MODEL_SH><<<blocks_for(n, 128), 128, 0, stream>>>(params, K, orientations, n, out);
    return hipGetLastError();
}

template <int MODEL>
hipError_t launch_update_model(const AlgoParams &a, uint32_t rng_seed, const float *params, const float *orientations,
                               int64_t n, float *out, hipStream_t stream)
{
    const unsigned blocks = blocks_for(n, 128);
    switch (a.algorithm) {
    case ALGO_DIFFUSION:
        update_kernel<MODEL, ALGO_DIFFUSION><<<blocks, 128, 0, stream>>>(a, rng_seed, params, orientations, n, out);
        break;
    case ALGO_PROBABILISTIC:
        update_kernel<MODEL, ALGO_PROBABILISTIC><<<blocks, 128, 0, stream>>>(a, rng_seed, params, orientations, n,
                                                                            out);
        break;
    case ALGO_DETERMINISTIC:
        update_kernel<MODEL, ALGO_DETERMINISTIC><<<blocks, 128, 0, stream>>>(a, rng_seed, params, orientations, n,
                                                                            out);
        break;
    default:
        update_kernel<MODEL, ALGO_TRANSPORT><<<blocks, 128, 0, stream>>>(a, rng_seed, params, orientations, n, out);
    }
    return hipGetLastError();
}

hipError_t launch_update(const AlgoParams &a, uint32_t rng_seed, const float *params, const float *orientations,
                         int64_t n, float *out, hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    if (a.model == MODEL_DTI)
        return launch_update_model<MODEL_DTI>(a, rng_seed, params, orientations, n, out, stream);
    return launch_update_model<MODEL_SH>(a, rng_seed, params, orientations, n, out, stream);
}

hipError_t launch_field_evaluate(const FieldView &f, const float *points, int64_t n, float *out, hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    field_evaluate_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(f, points, n, out);
    return hipGetLastError();
}

hipError_t launch_contains(const MaskView &m, const float *points, int64_t n, uint8_t *out, hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    contains_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(m, points, n, out);
    return hipGetLastError();
}

hipError_t launch_bvh_hit(const MeshView &m, int method, const float *starts, const float *ends, int64_t n,
                          uint8_t *found, uint32_t *hit, float *t, hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    bvh_hit_kernel<<<blocks_for(n, 128), 128, 0, stream>>>(m, method, starts, ends, n, found, hit, t);
    return hipGetLastError();
}

hipError_t launch_math(int fn, const float *x, const float *y, int64_t n, float *out, hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    math_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(fn, x, y, n, out);
    return hipGetLastError();
}

// Blocks of `kernel` (256 threads) that fill the current device once; never
// more than the blocks `n` seeds need.
template <class Kernel> hipError_t persistent_blocks(Kernel kernel, int64_t n, unsigned *blocks)
{
    int device = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&device);
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(kernel), 256, 0);
    if (e != hipSuccess)
        return e;
    const int64_t fill = (int64_t)(cus > 0 ? cus : 1) * (per_cu > 0 ? per_cu : 1);
    *blocks = (unsigned)std::min<int64_t>(fill, (n + 255) / 256);
    return hipSuccess;
}

template <class Kernel, class Launch>
hipError_t launch_persistent(Kernel kernel, const Launch &args, uint32_t *queue, int64_t n, hipStream_t stream)
{
    unsigned blocks = 0;
    const hipError_t e = persistent_blocks(kernel, n, &blocks);
    if (e != hipSuccess)
        return e;
    zero_u32_kernel<<<1, 64, 0, stream>>>(queue);
    kernel<<<blocks, 256, 0, stream>>>(args);
    return hipGetLastError();
}

template <int MODEL, int ALGO, int DOMAIN>
hipError_t launch_plain_kt(const TrackLaunch &p, hipStream_t stream)
{
    if (MODEL == MODEL_SH && p.algo.n_parameters == 45)
        return launch_persistent(track_plain_kernel<MODEL, ALGO, DOMAIN, 45>, p, p.queue, p.n_seeds, stream);
    return launch_persistent(track_plain_kernel<MODEL, ALGO, DOMAIN, 0>, p, p.queue, p.n_seeds, stream);
}

template <int MODEL, int ALGO>
hipError_t launch_plain_domain(const TrackLaunch &p, hipStream_t stream)
{
    if (p.domain_kind == DOMAIN_MESH)
        return launch_plain_kt<MODEL, ALGO, DOMAIN_MESH>(p, stream);
    return launch_plain_kt<MODEL, ALGO, DOMAIN_MASK>(p, stream);
}

template <int MODEL>
hipError_t launch_plain_algo(const TrackLaunch &p, hipStream_t stream)
{
    switch (p.algo.algorithm) {
    case ALGO_DIFFUSION: return launch_plain_domain<MODEL, ALGO_DIFFUSION>(p, stream);
    case ALGO_PROBABILISTIC: return launch_plain_domain<MODEL, ALGO_PROBABILISTIC>(p, stream);
    case ALGO_DETERMINISTIC: return launch_plain_domain<MODEL, ALGO_DETERMINISTIC>(p, stream);
    default: return launch_plain_domain<MODEL, ALGO_TRANSPORT>(p, stream);
    }
}

hipError_t launch_track_plain(const TrackLaunch &p, hipStream_t stream)
{
    if (p.n_seeds == 0)
        return hipSuccess;
    if (!p.queue)
        return hipErrorInvalidValue;
    if (p.algo.model == MODEL_DTI)
        return launch_plain_algo<MODEL_DTI>(p, stream);
    return launch_plain_algo<MODEL_SH>(p, stream);
}

template <int MODEL, int ALGO>
hipError_t launch_connectome_kt(const ConnectomeLaunch &c, hipStream_t stream)
{
    const hipError_t e = MODEL == MODEL_SH && c.algo.n_parameters == 45
                             ? launch_persistent(connectome_kernel<MODEL, ALGO, 45>, c, c.queue, c.n_seeds, stream)
                             : launch_persistent(connectome_kernel<MODEL, ALGO, 0>, c, c.queue, c.n_seeds, stream);
    if (e != hipSuccess)
        return e;
    return launch_connectome_hits(c, stream);
}

template <int MODEL>
hipError_t launch_connectome_model(const ConnectomeLaunch &c, hipStream_t stream)
{
    switch (c.algo.algorithm) {
    case ALGO_DIFFUSION: return launch_connectome_kt<MODEL, ALGO_DIFFUSION>(c, stream);
    case ALGO_PROBABILISTIC: return launch_connectome_kt<MODEL, ALGO_PROBABILISTIC>(c, stream);
    case ALGO_DETERMINISTIC: return launch_connectome_kt<MODEL, ALGO_DETERMINISTIC>(c, stream);
    default: return launch_connectome_kt<MODEL, ALGO_TRANSPORT>(c, stream);
    }
}

hipError_t launch_connectome_hits(const ConnectomeLaunch &c, hipStream_t stream)
{
    if (c.n_seeds == 0)
        return hipSuccess;
    connectome_hits_kernel<<<blocks_for(c.n_seeds, 256), 256, 0, stream>>>(c);
    return hipGetLastError();
}

hipError_t launch_connectome(const ConnectomeLaunch &c, hipStream_t stream)
{
    if (c.n_seeds == 0)
        return hipSuccess;
    if (c.algo.model == MODEL_DTI)
        return launch_connectome_model<MODEL_DTI>(c, stream);
    return launch_connectome_model<MODEL_SH>(c, stream);
}

template <int MODEL, int ALGO>
hipError_t launch_tod_kt(const TodLaunch &t, hipStream_t stream)
{
    if (MODEL == MODEL_SH && t.algo.n_parameters == 45)
        return launch_persistent(tod_kernel<MODEL, ALGO, 45>, t, t.queue, t.n_seeds, stream);
    return launch_persistent(tod_kernel<MODEL, ALGO, 0>, t, t.queue, t.n_seeds, stream);
}

template <int MODEL>
hipError_t launch_tod_model(const TodLaunch &t, hipStream_t stream)
{
    switch (t.algo.algorithm) {
    case ALGO_DIFFUSION: return launch_tod_kt<MODEL, ALGO_DIFFUSION>(t, stream);
    case ALGO_PROBABILISTIC: return launch_tod_kt<MODEL, ALGO_PROBABILISTIC>(t, stream);
    case ALGO_DETERMINISTIC: return launch_tod_kt<MODEL, ALGO_DETERMINISTIC>(t, stream);
    default: return launch_tod_kt<MODEL, ALGO_TRANSPORT>(t, stream);
    }
}

hipError_t launch_tod(const TodLaunch &t, hipStream_t stream)
{
    if (t.n_seeds == 0 || t.n_steps == 0)
        return hipSuccess;
    if (t.algo.model == MODEL_DTI)
        return launch_tod_model<MODEL_DTI>(t, stream);
    return launch_tod_model<MODEL_SH>(t, stream);
}

hipError_t launch_zero(float *out, int64_t n, hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
    zero_kernel<<<(unsigned)blocks, 256, 0, stream>>>(out, n);
    return hipGetLastError();
}

hipError_t launch_pad_rows(const float *src, int64_t n_rows, uint32_t K, float *dst, uint32_t stride,
                           hipStream_t stream)
{
    const int64_t n = n_rows * stride;
    if (n == 0)
        return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 16384);
    pad_rows_kernel<<<(unsigned)blocks, 256, 0, stream>>>(src, n_rows, K, dst, stride);
    return hipGetLastError();
}

hipError_t launch_normalize_odf(float *coefficients, int64_t n_cells, hipStream_t stream)
{
    if (n_cells == 0)
        return hipSuccess;
    normalize_odf_kernel<<<blocks_for(n_cells, 256), 256, 0, stream>>>(coefficients, n_cells);
    return hipGetLastError();
}

hipError_t launch_domain_seed(const float4 *vertices, const uint4 *triangles, const float *cumulative,
                              uint32_t n_triangles, uint32_t rng_seed, int64_t n, float *out, hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    domain_seed_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(vertices, triangles, cumulative, n_triangles,
                                                               rng_seed, n, out);
    return hipGetLastError();
}

}  // namespace tgk
