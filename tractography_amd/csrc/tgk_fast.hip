// tgk_fast.hip -- the fast tracking path.
//
// * Persistent lanes: a grid sized to fill the GPU once; every wavefront
//   claims seeds from a device queue (one atomic per wavefront refill) and
//   each lane starts a new streamline as soon as its previous one stops, so
//   streamlines of very different lengths (5 to 3000 steps) do not leave
//   lanes idle until the longest of a wavefront ends.
// * Per-lane cell cache: mask value and the voxel's model coefficients stay
//   on chip (LDS at order 8, registers at order 12 and for DTI) until the
//   point crosses into another cell (about every ten steps at step 0.1).
// * Transcendental-free SH model (tgk_fast_model.h), hardware exp2/log2/cos
//   where the reference calls exp/log/cos, FP64 arc length; the four
//   orientation updates (Transport, Diffusion, Probabilistic, Deterministic).
// * Seed-major dense output rows, compacted into ragged points afterwards.
// * The boundary connectome's tracking loop on the same machinery
//   (connectome_fast_kernel); its hits pass is the plain path's.
//
// Compiled with -ffp-contract=on: a product and a sum in one expression
// become an fma as the source writes them. With =fast the compiler also
// fused across statements, and chose differently in kernel variants that
// differ only in where a row waits (LDS or registers), which changed a few
// trajectories in the last bit; with =on every variant rounds alike.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <utility>
#include <vector>

#include "tgk_fast_model.h"
#include "tgk_internal.h"
#include "tgk_sh_basis.h"

namespace tgk {

namespace {

using fast::Order;

__device__ __forceinline__ void store_row(float *dense, int64_t g, uint32_t n_points, uint32_t n, V3 p)
{
    float *dst = dense + (g * (int64_t)n_points + n) * 3;
    dst[0] = p.x;
    dst[1] = p.y;
    dst[2] = p.z;
}

__device__ __forceinline__ int64_t mask_cell(const MaskView &m, V3 p)
{
    const float rx = floorf(affine_row(m.inv, 0, p) + 0.5f);
    const float ry = floorf(affine_row(m.inv, 1, p) + 0.5f);
    const float rz = floorf(affine_row(m.inv, 2, p) + 0.5f);
    // Bitwise &: all three coordinates are computed anyway, and a branch per
    // test costs more than the tests.
    if (!((rx >= 0.0f) & (rx < (float)m.dims[0]) & (ry >= 0.0f) & (ry < (float)m.dims[1]) & (rz >= 0.0f) &
          (rz < (float)m.dims[2])))
        return -1;
    return ((int64_t)(uint32_t)rx * m.dims[1] + (uint32_t)ry) * m.dims[2] + (uint32_t)rz;
}

// Cell ids of the tracking kernel are 32-bit: the fast path takes mask and
// field grids of fewer than 2^32 - 2 voxels.
constexpr uint32_t NO_CELL = 0xffffffffu;     // outside the grid
constexpr uint32_t UNSET_CELL = 0xfffffffeu;  // nothing cached yet

// The cell of rounded voxel coordinates (rx, ry, rz) on a grid, C order.
__device__ __forceinline__ uint32_t grid_cell(float rx, float ry, float rz, const uint32_t dims[3])
{
    // Bitwise &: all three coordinates are computed anyway, and a branch per
    // test costs more than the tests.
    if (!((rx >= 0.0f) & (rx < (float)dims[0]) & (ry >= 0.0f) & (ry < (float)dims[1]) & (rz >= 0.0f) &
          (rz < (float)dims[2])))
        return NO_CELL;
    return ((uint32_t)rx * dims[1] + (uint32_t)ry) * dims[2] + (uint32_t)rz;
}

// Saved points are stored in groups. A lane appends a 12-byte point about
// every 2.5 steps to its own row, and the grid's 262,144 lanes keep more
// partly written lines open than L2 holds, so single-point stores were
// written back to memory in partial lines several times each (20.9 GB
// written for 6 GB of points at 1M seeds). A group of four is four
// consecutive points starting at a multiple of four, 48 bytes, stored as
// three 16-byte stores when its address is 16-byte aligned (always, for a
// 16-byte aligned buffer and n_points a multiple of 4): 8.2 GB written, and
// 24.2 -> 21.1 ms per launch. The order-12 kernels and the Probabilistic
// and Deterministic ones, whose registers the group would push past a wave
// boundary, store single points.
template <int MODEL, int ALGO, int LMAX> constexpr int point_group()
{
    return (MODEL == MODEL_SH && LMAX == 12) || ALGO == ALGO_PROBABILISTIC || ALGO == ALGO_DETERMINISTIC ? 1 : 4;
}

template <int G> struct Lane {
    int32_t g;          // seed being tracked (-1: idle); calls take fewer than 2^31 seeds
    V3 point, o;
    V3 previous;        // the last step's start (mesh domains: the segment tested against the boundary)
    uint2 rng;
    double time;
    uint32_t n, steps;
    uint32_t field_cell; // cell ids of `point` (UNSET_CELL: none yet); their data is requested by request_cells
    uint32_t mask_cell;
    uint32_t mask_value; // the mask at mask_cell (0 outside the mask image)
    // Mesh domains: culling-grid cell of the last tested segment's start
    // (UNSET_CELL: none yet) and its candidate range [cand_first, cand_last).
    uint32_t mesh_cell, cand_first, cand_last;
    bool field_ok;       // `point` is inside the field image
    bool row_read;       // request_cells read a row / a mask value this iteration
    bool mask_read;
    // The last `held` saved points not yet stored, oldest first, in
    // group[G - held ..]; their indices are n - held .. n - 1.
    // `store_due`: n is a multiple of G, so they close a group.
    V3 group[G];
    uint32_t held;
    bool store_due;
};

template <int G> __device__ __forceinline__ void hold_point(Lane<G> &L, V3 q)
{
#pragma unroll
    for (int i = 0; i + 1 < G; i++)
        L.group[i] = L.group[i + 1];
    L.group[G - 1] = q;
    L.held = min(L.held + 1u, (uint32_t)G);
    L.n++;
    L.store_due = L.n % G == 0;
}

// Stores the held points: a whole aligned group of four as three float4
// stores, otherwise one point at a time (a single-point group, the first
// group after a resumption, or what a stopping or saved streamline holds).
template <int G> __device__ __forceinline__ void store_held(const TrackLaunch &p, Lane<G> &L)
{
    float *row = p.points + ((int64_t)L.g * p.n_points + (L.n - L.held)) * 3;
    if (G == 4 && L.held == G && (reinterpret_cast<uintptr_t>(row) & 15u) == 0) {
        float4 *dst = reinterpret_cast<float4 *>(row);
        const V3 *q = L.group;
        dst[0] = make_float4(q[0].x, q[0].y, q[0].z, q[G > 1 ? 1 : 0].x);
        dst[1] = make_float4(q[G > 1 ? 1 : 0].y, q[G > 1 ? 1 : 0].z, q[G > 2 ? 2 : 0].x, q[G > 2 ? 2 : 0].y);
        dst[2] = make_float4(q[G > 2 ? 2 : 0].z, q[G - 1].x, q[G - 1].y, q[G - 1].z);
    } else {
#pragma unroll
        for (int i = 0; i < G; i++) {
            if ((uint32_t)i >= G - L.held) {
                float *dst = row + 3 * (i - (G - (int)L.held));
                dst[0] = L.group[i].x;
                dst[1] = L.group[i].y;
                dst[2] = L.group[i].z;
            }
        }
    }
    L.held = 0;
    L.store_due = false;
}

// A closed group is stored in the iteration after its last point is saved,
// right after that iteration's wait for its cell data: every iteration
// waits (s_waitcnt vmcnt(0)) for the data requested in the previous one,
// and vector stores count under the same counter, so a store issued at the
// end of an iteration would hold the next one at that wait.
template <int G> __device__ __forceinline__ void store_due_group(const TrackLaunch &p, Lane<G> &L)
{
    if (L.store_due)
        store_held(p, L);
}

// A streamline in flight when a generation of the persistent launch ends
// (see track_fast_kernel): everything its lane needs to resume it, in four
// float4 (the previous point only for mesh domains). Its held points are
// stored first; the cell data are requested again on resumption.
template <int DOMAIN, int G> __device__ __forceinline__ void save_lane(float4 *slot, const Lane<G> &L)
{
    const uint64_t tb = (uint64_t)__double_as_longlong(L.time);
    const V3 prev = DOMAIN == DOMAIN_MESH ? L.previous : v3(0.0f, 0.0f, 0.0f);
    slot[0] = make_float4(L.point.x, L.point.y, L.point.z, L.o.x);
    slot[1] = make_float4(L.o.y, L.o.z, prev.x, prev.y);
    slot[2] = make_float4(prev.z, __uint_as_float((uint32_t)tb), __uint_as_float((uint32_t)(tb >> 32)),
                          __uint_as_float((uint32_t)L.g));
    slot[3] = make_float4(__uint_as_float(L.n), __uint_as_float(L.steps), __uint_as_float(L.rng.x),
                          __uint_as_float(L.rng.y));
}

template <int DOMAIN, int G> __device__ __forceinline__ void resume_lane(const float4 *slot, Lane<G> &L)
{
    const float4 a = slot[0], b = slot[1], c = slot[2], d = slot[3];
    L.point = v3(a.x, a.y, a.z);
    L.o = v3(a.w, b.x, b.y);
    if constexpr (DOMAIN == DOMAIN_MESH)
        L.previous = v3(b.z, b.w, c.x);
    L.time = __longlong_as_double((long long)(((uint64_t)__float_as_uint(c.z) << 32) | __float_as_uint(c.y)));
    L.g = (int32_t)__float_as_uint(c.w);
    L.n = __float_as_uint(d.x);
    L.steps = __float_as_uint(d.y);
    L.rng = make_uint2(__float_as_uint(d.z), __float_as_uint(d.w));
    L.held = 0;
    L.store_due = false;
}

// One iteration of the reference loop (tracking.py:121-147) for a lane with
// a streamline in flight; on termination the lane's results are written and
// the lane becomes idle (L.g = -1).
// Where a lane keeps its cell's coefficients between steps: SH rows of
// order 8 mostly in LDS (float4 columns, one per lane, read back every
// step), so no large register array is carried around the loop; DTI rows
// and order-12 rows in registers.
template <int MODEL, int LMAX> constexpr bool rows_in_lds() { return MODEL == MODEL_SH && LMAX == 8; }
// Threads per block of the fast kernels.
template <int LMAX> constexpr int fast_block() { return 256; }
// float4 columns of a row kept in LDS. At order 8, 10 of the 12 (the first
// 40 of 45 coefficients): 40 KiB per 256-lane block, so four blocks share a
// CU's 160 KiB and each SIMD runs four waves (at 109 VGPRs the registers
// would allow it; the full 48-float row would allow three). The other 5
// coefficients stay in registers.
template <int MODEL, int LMAX> constexpr int lds_vectors() { return rows_in_lds<MODEL, LMAX>() ? 10 : 0; }

// Launch shapes of the order-8 SH Transport and Diffusion tracking kernels,
// chosen per runtime (tgk_runtime_set_fast_shape, the autotuner): lanes per
// block and float4 columns of the row in LDS. Shape 0 is the default above;
// the others trade LDS per block against blocks per CU, or coefficients in
// LDS against registers. Which lane tracks which streamline, and where its
// coefficients wait between steps, changes no result.
struct FastShape {
    int block;
    int lds_vectors;
};
constexpr FastShape FAST_SHAPES[FAST_SHAPE_COUNT] = {{256, 10}, {128, 10}, {512, 10}, {256, 8}, {256, 11}};
template <int MODEL, int ALGO, int LMAX> constexpr bool shaped()
{
    return MODEL == MODEL_SH && LMAX == 8 && (ALGO == ALGO_TRANSPORT || ALGO == ALGO_DIFFUSION);
}
template <int MODEL, int ALGO, int LMAX, int S> constexpr int shape_block()
{
    return shaped<MODEL, ALGO, LMAX>() ? FAST_SHAPES[S].block : fast_block<LMAX>();
}
template <int MODEL, int ALGO, int LMAX, int S> constexpr int shape_lds_vectors()
{
    return shaped<MODEL, ALGO, LMAX>() ? FAST_SHAPES[S].lds_vectors : lds_vectors<MODEL, LMAX>();
}

// Column j of a row from memory into LDS for every lane of the wavefront
// that executes it. The instruction's immediate offset is added to both the
// memory address and the LDS base, so all columns share one vector address
// (the row) and the LDS base is set 16 * j below column j's slot.
template <int BLOCK, int... J>
__device__ __forceinline__ void dma_columns(const void *row, uint32_t wave_base, std::integer_sequence<int, J...>)
{
    (__builtin_amdgcn_global_load_lds(
         row, (__attribute__((address_space(3))) void *)(uintptr_t)(wave_base + (uint32_t)(J * BLOCK * 16 - J * 16)),
         16, J * 16, 0),
     ...);
}

// A cell's row into the lane's cache on a cell change: the first LDS_VEC
// float4 columns of an SH row to LDS (column j of a lane at j * BLOCK +
// lane), the rest to registers; DTI rows to registers.
template <int MODEL, int LMAX, int LDS_VEC = lds_vectors<MODEL, LMAX>(), int BLOCK = fast_block<LMAX>()>
__device__ __forceinline__ void load_row(const float *__restrict__ coeffs, int64_t cell, uint32_t cstride, float *q,
                                         float4 *__restrict__ lds_rows)
{
    constexpr int TERMS = Order<LMAX>::TERMS;
    constexpr int VEC = Order<LMAX>::STRIDE / 4;
    const float4 *row = reinterpret_cast<const float4 *>(coeffs + cell * (int64_t)cstride);
    if constexpr (MODEL == MODEL_SH) {
        // Straight from memory into LDS (no registers held while the loads
        // are in flight): each lane's 16 bytes land at the instruction's
        // wave-uniform LDS base + 16 * lane, i.e. column j of the lane's
        // slot. Lanes that do not reload are masked off and keep theirs.
        // The wavefront's slot base is a scalar: its first lane's byte
        // offset in the block's rows plus the rows' LDS address.
        typedef __attribute__((address_space(3))) char lds_char;
        const uint32_t wave_base =
            (uint32_t)__builtin_amdgcn_readfirstlane((int)((threadIdx.x & ~63u) * 16u)) +
            (uint32_t)(uintptr_t)(lds_char *)lds_rows;
        dma_columns<BLOCK>(static_cast<const void *>(row), wave_base, std::make_integer_sequence<int, LDS_VEC>{});
#pragma unroll
        for (int j = LDS_VEC; j < VEC; j++) {
            const float4 v = row[j];
            if (4 * j < TERMS)
                q[4 * j] = v.x;
            if (4 * j + 1 < TERMS)
                q[4 * j + 1] = v.y;
            if (4 * j + 2 < TERMS)
                q[4 * j + 2] = v.z;
            if (4 * j + 3 < TERMS)
                q[4 * j + 3] = v.w;
        }
    } else {
        const float4 a = row[0], b = row[1];
        q[0] = a.x;
        q[1] = a.y;
        q[2] = a.z;
        q[3] = a.w;
        q[4] = b.x;
        q[5] = b.y;
    }
}

// The LDS part of the cached row back into q, every step.
template <int MODEL, int LMAX, int LDS_VEC = lds_vectors<MODEL, LMAX>(), int BLOCK = fast_block<LMAX>()>
__device__ __forceinline__ void read_row(float *q, const float4 *__restrict__ lds_rows)
{
    constexpr int TERMS = Order<LMAX>::TERMS;
    if constexpr (LDS_VEC > 0) {
        // Loads into LDS (load_row) complete under vmcnt, which the compiler
        // does not wait on before these reads of the same LDS: wait here
        // (s_waitcnt vmcnt(0); expcnt and lgkmcnt left at their maxima).
        __builtin_amdgcn_s_waitcnt(0x0F70);
#pragma unroll
        for (int j = 0; j < LDS_VEC; j++) {
            const float4 v = lds_rows[j * BLOCK + threadIdx.x];
            if (4 * j < TERMS)
                q[4 * j] = v.x;
            if (4 * j + 1 < TERMS)
                q[4 * j + 1] = v.y;
            if (4 * j + 2 < TERMS)
                q[4 * j + 2] = v.z;
            if (4 * j + 3 < TERMS)
                q[4 * j + 3] = v.w;
        }
    }
}

// One orientation update on the fast model. Transport and Diffusion take
// their normal draws (fast::draw_noise) from the caller, already drawn;
// Probabilistic draws from `rng` as it goes.
template <int MODEL, int ALGO, int LMAX>
__device__ __forceinline__ V3 update_orientation(const AlgoParams &a, const float *q, V3 o, uint2 &rng,
                                                 const fast::Noise &nz)
{
    if constexpr (ALGO == ALGO_PROBABILISTIC) {
        return fast::probabilistic<MODEL, LMAX>(a, q, o, rng);
    } else if constexpr (ALGO == ALGO_DETERMINISTIC) {
        return fast::deterministic<MODEL, LMAX>(a, q, o);
    } else if constexpr (MODEL == MODEL_SH) {
        return fast::sh_update<ALGO, LMAX>(a, q, o, nz);
    } else {
        const fast::Direction dir = fast::direction_of(o);
        return fast::update<ALGO>(a, fast::dti_model(q, o, dir), dir, o, nz);
    }
}

template <int MODEL, int ALGO, int LMAX>
__device__ __forceinline__ V3 update_orientation(const AlgoParams &a, const float *q, V3 o, uint2 &rng)
{
    const fast::Noise nz = fast::draw_noise<ALGO>(a, rng);
    return update_orientation<MODEL, ALGO, LMAX>(a, q, o, rng, nz);
}

// Requests the data of L.point's cells one iteration before the loop needs
// it: the mask value when the mask cell changes, the coefficient row when
// the field cell changes (into LDS or registers, see load_row). Their
// memory latency then overlaps the rest of the step, and of the next one
// up to the model evaluation, instead of stalling the wavefront, which
// nearly every step holds at least one lane entering a new cell.
// The voxel coordinates come from one evaluation of the field's inverse
// affine; the mask's, when it is the same matrix (shared_affine: every
// image-domain workload that tracks in its field's grid), reuses it. The mask
// rounds floor(v + 0.5) (domain.py:369), the field rint(v) (field.py:97).
template <int MODEL, int DOMAIN, int LMAX, int G, int LDS_VEC, int BLOCK>
__device__ __forceinline__ void request_cells(const TrackLaunch &p, bool shared_affine,
                                              const float *__restrict__ coeffs, uint32_t cstride, Lane<G> &L, float *q,
                                              float4 *__restrict__ lds_rows)
{
    const float fx = affine_row(p.field.inv, 0, L.point);
    const float fy = affine_row(p.field.inv, 1, L.point);
    const float fz = affine_row(p.field.inv, 2, L.point);
    if constexpr (DOMAIN == DOMAIN_MASK) {
        float mx = fx, my = fy, mz = fz;
        if (!shared_affine) {
            mx = affine_row(p.mask.inv, 0, L.point);
            my = affine_row(p.mask.inv, 1, L.point);
            mz = affine_row(p.mask.inv, 2, L.point);
        }
        const uint32_t cell = grid_cell(floorf(mx + 0.5f), floorf(my + 0.5f), floorf(mz + 0.5f), p.mask.dims);
        if (cell != L.mask_cell) {
            L.mask_cell = cell;
            L.mask_value = cell != NO_CELL ? (uint32_t)p.mask.mask[cell] : 0u;
            L.mask_read = cell != NO_CELL;
        }
    }
    const uint32_t cell = grid_cell(rintf(fx), rintf(fy), rintf(fz), p.field.dims);
    L.field_ok = cell != NO_CELL;
    if (cell != NO_CELL && cell != L.field_cell) {
        L.field_cell = cell;
        load_row<MODEL, LMAX, LDS_VEC, BLOCK>(coeffs, (int64_t)cell, cstride, q, lds_rows);
        L.row_read = true;
    }
}

// Boundary test of the wavefront's last segments against their start cells'
// candidate lists, the (lane, candidate) pairs spread over the wavefront's
// lanes. Lanes near a surface are few (about one in twenty per step), so
// looping per lane would run the longest list for all 64 lanes; here a step
// costs one round of 64 pairs while the lists sum to 64 or fewer. Each pair
// is tested with the bound 1 (grid_triangle_hit), and a lane keeps its
// smallest t, the last in list order among equal t: the candidate that
// testing its list in order with "t <= best replaces" ends on
// (grid_segment_hit), so the hit is the same. Called by every lane of the
// wavefront; `decided` is false where the lists cannot decide (the caller
// then asks the BVH), `found`/`tri` give the hit.
template <int G>
__device__ __forceinline__ void wave_segment_test(const MeshView &m, Lane<G> &L, bool live, bool &decided,
                                                  bool &found, uint32_t &tri)
{
    decided = false;
    found = false;
    tri = 0;
    const int lane = __lane_id();
    const float sx = L.previous.x, sy = L.previous.y, sz = L.previous.z;
    const float dx = L.point.x - sx, dy = L.point.y - sy, dz = L.point.z - sz;
    uint32_t first = 0, count = 0;
    if (live && L.steps > 0 && m.cell_start != nullptr && m.reach > 0.0f &&
        (dx * dx + dy * dy) + dz * dz <= m.reach * m.reach) {
        const int64_t cell = mesh_grid_cell(m, L.previous);
        if (cell >= 0) {
            decided = true;
            if ((uint32_t)cell != L.mesh_cell) {
                L.mesh_cell = (uint32_t)cell;
                L.cand_first = m.cell_start[cell];
                L.cand_last = m.cell_start[cell + 1];
            }
            first = L.cand_first;
            count = L.cand_last - L.cand_first;
        }
    }
    if (__ballot(count != 0) == 0)
        return;
    // Pairs [begin, end) of the flattened list belong to this lane. A list
    // holds fewer than 2^26 candidates (tgk_domain_set_segment_grid), so the
    // sum of 64 fits 32 bits with room for the last round's base + 64.
    uint32_t end = count;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t v = __shfl_up(end, d);
        if (lane >= d)
            end += v;
    }
    const uint32_t total = __shfl(end, 63);
    const uint32_t begin = end - count;
    float best_t = 1.0f;
    uint32_t best_index = 0;
    for (uint32_t base = 0; base < total; base += 64) {
        const uint32_t k = base + (uint32_t)lane;
        // Owner of pair k: the last lane whose begin is at most k (begins
        // never decrease along the lanes; lanes after the owner begin past k).
        int owner = 0;
#pragma unroll
        for (int step = 32; step >= 1; step >>= 1) {
            if (__shfl(begin, owner + step) <= k)
                owner += step;
        }
        const uint32_t index = __shfl(first, owner) + (k - __shfl(begin, owner));
        const float st[3] = {__shfl(sx, owner), __shfl(sy, owner), __shfl(sz, owner)};
        const float dir[3] = {__shfl(dx, owner), __shfl(dy, owner), __shfl(dz, owner)};
        bool h = false;
        float t = 0.0f;
        uint32_t slot = 0;
        if (k < total) {
            const float4 a = m.cell_entries[3 * (size_t)index];
            const float4 e1 = m.cell_entries[3 * (size_t)index + 1];
            const float4 e2 = m.cell_entries[3 * (size_t)index + 2];
            h = grid_triangle_hit(a, e1, e2, st, dir, 1.0f, t);
            slot = __float_as_uint(a.w);
        }
        // Hits are rare: hand each to its owner in turn.
        uint64_t hits = __ballot(h);
        while (hits != 0) {
            const int j = __ffsll((unsigned long long)hits) - 1;
            hits &= hits - 1;
            const int oj = __shfl(owner, j);
            const float tj = __shfl(t, j);
            const uint32_t ij = __shfl(index, j);
            const uint32_t sj = __shfl(slot, j);
            if (lane == oj && (!found || tj < best_t || (tj == best_t && ij > best_index))) {
                found = true;
                best_t = tj;
                best_index = ij;
                tri = sj;
            }
        }
    }
}

// nz: this step's Diffusion draws, made by the caller before anything that
// waits on memory, so they cover the latency of the cell data
// request_cells asked for.
template <int MODEL, int ALGO, int DOMAIN, int LMAX, int G, int LDS_VEC, int BLOCK>
__device__ __forceinline__ void advance(const TrackLaunch &p, bool shared_affine, const float *__restrict__ coeffs,
                                        uint32_t cstride, Lane<G> &L, float *q, float4 *__restrict__ lds_rows,
                                        const fast::Noise &nz, bool seg_decided, bool seg_found, uint32_t seg_tri)
{
    bool stop = !(L.n < p.n_points && L.steps < p.max_steps);
    uint32_t hit = NO_HIT;
    if (!stop) {
        if constexpr (DOMAIN == DOMAIN_MASK) {
            stop = L.mask_value == 0u;
        } else if (L.steps > 0) {
            // The start cell's candidate list (wave_segment_test) answers
            // segments up to `reach` long; longer segments, or starts off the
            // grid, traverse the BVH.
            if (seg_found) {
                hit = seg_tri;
                stop = true;
            } else if (!seg_decided) {
                uint32_t tri;
                float t;
                const float dx = L.point.x - L.previous.x, dy = L.point.y - L.previous.y,
                            dz = L.point.z - L.previous.z;
                if (!mesh_far(p.mesh, L.previous, dx * dx + dy * dy + dz * dz) &&
                    bvh_segment_hit(p.mesh, L.previous, L.point, tri, t)) {
                    hit = tri;
                    stop = true;
                }
            }
        }
    }
    if (!stop)
        stop = !L.field_ok;
    if (stop) {
        store_held(p, L);
        p.lengths[L.g] = L.n;
        p.steps[L.g] = L.steps;
        p.hits[L.g] = hit;
        L.g = -1;
        return;
    }
    read_row<MODEL, LMAX, LDS_VEC, BLOCK>(q, lds_rows);
    store_due_group(p, L);
    L.o = update_orientation<MODEL, ALGO, LMAX>(p.algo, q, L.o, L.rng, nz);
    L.steps++;
    if constexpr (DOMAIN == DOMAIN_MESH)
        L.previous = L.point;
    const float step = p.algo.step_size;
    L.point = v3(L.point.x + step * L.o.x, L.point.y + step * L.o.y, L.point.z + step * L.o.z);
    L.time += (double)step;
    if (L.time >= (double)p.save_at) {
        L.time -= (double)p.save_at;
        const float t = (float)L.time;
        hold_point(L, v3(L.point.x - t * L.o.x, L.point.y - t * L.o.y, L.point.z - t * L.o.z));
    }
    request_cells<MODEL, DOMAIN, LMAX, G, LDS_VEC, BLOCK>(p, shared_affine, coeffs, cstride, L, q, lds_rows);
}

// Waves per SIMD the compiler must fit: the order-8 SH Transport and
// Diffusion kernels fit 128 VGPRs without spilling when held to four waves,
// the most their 40 KiB of LDS per block allows (134-140 otherwise, three
// waves). Held to three, the order-12 Transport and Diffusion kernels take
// 168 VGPRs (170-185 otherwise, two waves) and the order-8 Probabilistic
// one 151-155 (168-181, two waves); held to four, Probabilistic ran 3%
// slower. The order-8 Deterministic kernel, held to four waves, made this
// compiler emit an odd-aligned 64-bit register spill that gfx950 rejects,
// so it keeps its 159-162 VGPRs (three waves). The DTI kernels (90-115
// VGPRs, four or five waves) ran 4% slower held to seven waves and 19%
// slower held to eight (spills).
// A launch shape that puts more of the row in LDS allows fewer blocks per
// CU (160 KiB of LDS each): shape 4 (44 coefficients) three waves per SIMD.
template <int MODEL, int ALGO, int DOMAIN, int LMAX, int S = 0> constexpr int fast_min_waves()
{
    if constexpr (MODEL == MODEL_SH && LMAX == 12 && (ALGO == ALGO_TRANSPORT || ALGO == ALGO_DIFFUSION))
        return 3;
    if constexpr (MODEL == MODEL_SH && LMAX == 8 && (ALGO == ALGO_TRANSPORT || ALGO == ALGO_DIFFUSION)) {
        constexpr int block = FAST_SHAPES[S].block;
        constexpr int per_cu = 163840 / (FAST_SHAPES[S].lds_vectors * block * 16);
        constexpr int waves = per_cu * (block / 64) / 4;
        return waves < 4 ? waves : 4;
    }
    if constexpr (MODEL == MODEL_SH && LMAX == 8 && ALGO == ALGO_PROBABILISTIC)
        return 3;
    return 1;
}

// The persistent launch runs in generations (launch_fast): generation 0
// claims seeds; each later one claims the streamlines the previous one left
// in flight. Once all of a generation's work is claimed, lanes go idle as
// their streamlines end, while a wavefront keeps issuing for its longest
// one: on the 1M-seed benchmark workload half the launch would run with
// its lanes ever sparser. So when the generation's live lanes fall below
// `stop_below` (half its work), every wavefront saves its live lanes and
// ends, and the next generation resumes them packed into full wavefronts,
// a quarter or half as many per SIMD, each running correspondingly faster.
// The last generation, or one with little work, runs to completion. Which
// lane tracks which streamline, and when, changes no result.
template <int MODEL, int ALGO, int DOMAIN, int LMAX, int S>
__global__ __launch_bounds__((shape_block<MODEL, ALGO, LMAX, S>()))
__attribute__((amdgpu_waves_per_eu(fast_min_waves<MODEL, ALGO, DOMAIN, LMAX, S>())))
void track_fast_kernel(FastLaunch f)
{
    constexpr int BLOCK = shape_block<MODEL, ALGO, LMAX, S>();
    constexpr int LDS_VEC = shape_lds_vectors<MODEL, ALGO, LMAX, S>();
    __shared__ float4 lds_rows[LDS_VEC > 0 ? LDS_VEC * BLOCK : 1];
    const TrackLaunch &p = f.track;
    const int lane_id = __lane_id();
    const float *coeffs = static_cast<const float *>(f.coeffs);
    const uint32_t cstride = f.coeff_stride;

    // This generation's work: seeds, or the previous generation's saved
    // streamlines, claimed through f.claims. Streamlines that end or are
    // saved are counted in f.ended: FAST_END_SHARDS counters a cache line
    // apart, each wavefront adding to one of them in batches, so that no
    // single word takes every wavefront's updates.
    const bool resume = f.cont_in != nullptr;
    const uint32_t n_work = resume ? *f.cont_in_count : (uint32_t)p.n_seeds;
    uint32_t stop_below = 0;
    if (!f.last) {
        const uint64_t lanes = (uint64_t)gridDim.x * blockDim.x;
        const uint32_t half = (uint32_t)((resume ? (uint64_t)n_work : min((uint64_t)n_work, lanes)) / 2);
        stop_below = half >= f.min_repack ? half : 0u;
    }
    uint32_t *const my_shard =
        f.ended + ((blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % FAST_END_SHARDS) * FAST_END_STRIDE;

    constexpr int G = point_group<MODEL, ALGO, LMAX>();
    Lane<G> L;
    L.g = -1;
    // An idle lane's position is read (never used) by the wavefront-wide
    // boundary test before its first seed.
    L.point = L.previous = v3(0.0f, 0.0f, 0.0f);
    L.steps = 0;
    L.held = 0;
    L.store_due = false;
    L.row_read = false;
    L.mask_read = false;
    // Rows and mask values this wavefront read (wave-uniform sums).
    uint64_t rows_read = 0, masks_read = 0;
    float q[MODEL == MODEL_SH ? Order<LMAX>::TERMS : 6];
    bool drained = false;
    uint32_t ended_pending = 0;     // ended here and not yet added to the shard
    uint32_t iteration = 0;
    uint32_t live_seen = 0xffffffffu;  // the generation's streamlines in flight at the last count
    bool counting = false;          // shard reads issued, to be summed next iteration
    uint32_t shard_value = 0;

    while (true) {
        // Refill idle lanes, one atomic per wavefront.
        const uint64_t need = __ballot(L.g < 0);
        if (need != 0 && !drained) {
            const int leader = __ffsll((unsigned long long)need) - 1;
            const uint32_t count = (uint32_t)__popcll(need);
            uint32_t base = 0;
            if (lane_id == leader)
                base = atomicAdd(f.claims, count);
            base = __shfl(base, leader);
            if ((uint64_t)base + count >= (uint64_t)n_work)
                drained = true;
            if (L.g < 0) {
                const uint32_t rank = (uint32_t)__popcll(need & ((1ull << lane_id) - 1ull));
                const uint64_t j = (uint64_t)base + rank;
                if (j < (uint64_t)n_work) {
                    if (resume) {
                        resume_lane<DOMAIN, G>(f.cont_in + FAST_CONTINUATION_VEC4 * j, L);
                    } else {
                        const uint64_t g = j;
                        L.g = (int32_t)g;
                        const float4 s0 = reinterpret_cast<const float4 *>(p.seeds)[2 * g];
                        const float4 s1 = reinterpret_cast<const float4 *>(p.seeds)[2 * g + 1];
                        L.point = v3(s0.x, s0.y, s0.z);
                        L.o = v3(s1.x, s1.y, s1.z);
                        L.previous = L.point;
                        if constexpr (ALGO == ALGO_DIFFUSION || ALGO == ALGO_PROBABILISTIC)
                            L.rng = seed_stream(p, p.first_seed + L.g);
                        L.time = 0.0;
                        L.n = 0;
                        L.steps = 0;
                        L.held = 0;
                        hold_point(L, L.point);  // the seed is the streamline's first point
                    }
                    L.field_cell = UNSET_CELL;
                    L.mask_cell = UNSET_CELL;
                    L.mask_value = 0;
                    L.mesh_cell = UNSET_CELL;
                    request_cells<MODEL, DOMAIN, LMAX, G, LDS_VEC, BLOCK>(p, f.shared_affine != 0, coeffs, cstride,
                                                                          L, q, lds_rows);
                }
            }
        }
        const uint64_t live = __ballot(L.g >= 0);
        if (live == 0)
            break;
        // The draws need no memory: made first, they cover the latency of
        // the cell data and shard counts requested last iteration. The
        // stream before them is what a saved streamline resumes from.
        // Only live lanes draw: an idle lane's stream is not initialised (an
        // all-zero MWC64X state never leaves zero, and randn would spin).
        const uint2 rng_before = L.rng;
        fast::Noise nz{0.0f, 0.0f};
        if (L.g >= 0)
            nz = fast::draw_noise<ALGO>(p.algo, L.rng);
        if (counting) {
            // Every claim was made before the shards were read, so the
            // streamlines in flight are the work minus those that left.
            uint32_t left = shard_value;
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1)
                left += __shfl_xor(left, d);
            live_seen = n_work - min(left, n_work);
            counting = false;
        }
        if (live_seen < stop_below) {
            // Few streamlines left in flight: end the generation, handing
            // this wavefront's to the next one.
            const int leader = __ffsll((unsigned long long)live) - 1;
            const uint32_t count = (uint32_t)__popcll(live);
            uint32_t base = 0;
            if (lane_id == leader) {
                base = atomicAdd(f.cont_out_count, count);
                atomicAdd(my_shard, ended_pending + count);
            }
            base = __shfl(base, leader);
            if (L.g >= 0) {
                store_held(p, L);
                L.rng = rng_before;
                save_lane<DOMAIN, G>(f.cont_out +
                              FAST_CONTINUATION_VEC4 *
                                  ((uint64_t)base + (uint32_t)__popcll(live & ((1ull << lane_id) - 1ull))),
                          L);
            }
            ended_pending = 0;
            break;
        }
        bool seg_decided = false, seg_found = false;
        uint32_t seg_tri = 0;
        if constexpr (DOMAIN == DOMAIN_MESH)
            wave_segment_test<G>(p.mesh, L, L.g >= 0, seg_decided, seg_found, seg_tri);
        if (L.g >= 0)
            advance<MODEL, ALGO, DOMAIN, LMAX, G, LDS_VEC, BLOCK>(p, f.shared_affine != 0, coeffs, cstride, L, q,
                                                                 lds_rows, nz, seg_decided, seg_found, seg_tri);
        rows_read += (uint64_t)__popcll(__ballot(L.row_read));
        masks_read += (uint64_t)__popcll(__ballot(L.mask_read));
        L.row_read = false;
        L.mask_read = false;
        ended_pending += (uint32_t)__popcll(live & ~__ballot(L.g >= 0));
        // Once every claim is made, every eighth iteration: add this
        // wavefront's ended streamlines to its shard and read all shards
        // (summed next iteration, after the draws).
        if (stop_below != 0 && drained && (++iteration & 7u) == 0u) {
            if (ended_pending != 0 && lane_id == 0)
                atomicAdd(my_shard, ended_pending);
            ended_pending = 0;
            shard_value = lane_id < FAST_END_SHARDS
                              ? __hip_atomic_load(f.ended + lane_id * FAST_END_STRIDE, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT)
                              : 0u;
            counting = true;
        }
    }
    if (stop_below != 0 && ended_pending != 0 && lane_id == 0)
        atomicAdd(my_shard, ended_pending);
    // A lane that stopped may have requested a row for its last point; the
    // wavefront waits for every load into LDS before it ends, so none lands
    // in LDS another workgroup has been given.
    __builtin_amdgcn_s_waitcnt(0);
    // The wavefront's memory reads of rows and mask values, one atomic each
    // (the launch's totals price the bytes the cell cache had to fetch).
    if (lane_id == 0) {
        atomicAdd(&f.stats[0], (unsigned long long)rows_read);
        atomicAdd(&f.stats[1], (unsigned long long)masks_read);
    }
}

// tracking.py:457-558 (kernel connectome) on the fast model: the plain
// connectome kernel's loop (mask, field, update, finite and length test,
// move) with the tracking kernel's persistent lanes, cell cache and
// polynomial model. A finished track leaves the same record in its seed's
// output slots as on the plain path (last segment, start barycentrics,
// start triangle, length test) for the shared hits kernel.
template <int MODEL, int ALGO, int LMAX>
__global__ __launch_bounds__(fast_block<LMAX>()) void connectome_fast_kernel(FastConnectomeLaunch f)
{
    __shared__ float4 lds_rows[rows_in_lds<MODEL, LMAX>() ? lds_vectors<MODEL, LMAX>() * fast_block<LMAX>() : 1];
    const ConnectomeLaunch &c = f.connectome;
    const float *coeffs = static_cast<const float *>(f.coeffs);
    const uint32_t cstride = f.coeff_stride;
    const float step = c.algo.step_size;
    const int lane_id = __lane_id();

    int64_t g = -1;
    bool drained = false;
    uint2 rng = make_uint2(0u, 0u);
    V3 point = v3(0.0f, 0.0f, 0.0f), o = point, previous = point;
    uint32_t tracked = 0, start_triangle = 0;
    float b0 = 0.0f, b1 = 0.0f, b2 = 0.0f;
    int64_t field_cell = -2, cached_mask_cell = -2;
    bool mask_inside = false;
    float q[MODEL == MODEL_SH ? Order<LMAX>::TERMS : 6];

    while (true) {
        // Refill idle lanes from the queue, one atomic per wavefront.
        const uint64_t need = __ballot(g < 0);
        if (need != 0 && !drained) {
            const int leader = __ffsll((unsigned long long)need) - 1;
            const uint32_t count = (uint32_t)__popcll(need);
            uint32_t base = 0;
            if (lane_id == leader)
                base = atomicAdd(c.queue, count);
            base = __shfl(base, leader);
            if ((uint64_t)base + count >= (uint64_t)c.n_seeds)
                drained = true;
            if (g < 0) {
                const uint64_t claimed = (uint64_t)base + (uint64_t)__popcll(need & ((1ull << lane_id) - 1ull));
                if (claimed < (uint64_t)c.n_seeds) {
                    g = (int64_t)claimed;
                    rng = rng_init(c.rng_seed, (uint32_t)g);
                    const BoundarySample sample = sample_boundary(c.boundary.vertices, c.boundary.triangles,
                                                                  c.cumulative, c.n_triangles, rng);
                    point = sample.point;
                    o = sample.orientation;
                    previous = point;
                    tracked = 0;
                    start_triangle = sample.triangle;
                    b0 = sample.bary[0];
                    b1 = sample.bary[1];
                    b2 = sample.bary[2];
                    field_cell = -2;
                    cached_mask_cell = -2;
                    mask_inside = false;
                }
            }
        }
        if (__ballot(g >= 0) == 0)
            break;
        if (g < 0)
            continue;
        bool stop = tracked >= c.n_steps;
        if (!stop) {
            const int64_t cell = mask_cell(c.mask, point);
            if (cell != cached_mask_cell) {
                cached_mask_cell = cell;
                mask_inside = cell >= 0 && c.mask.mask[cell] != 0;
            }
            stop = !mask_inside;
        }
        int64_t cell = -1;
        if (!stop) {
            cell = field_voxel(c.field, point);
            stop = cell < 0;
        }
        if (!stop) {
            if (cell != field_cell) {
                field_cell = cell;
                load_row<MODEL, LMAX>(coeffs, cell, cstride, q, lds_rows);
            }
            read_row<MODEL, LMAX>(q, lds_rows);
            o = update_orientation<MODEL, ALGO, LMAX>(c.algo, q, o, rng);
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
        float *record = c.weights + 9 * g;
        record[0] = previous.x;
        record[1] = previous.y;
        record[2] = previous.z;
        record[3] = point.x;
        record[4] = point.y;
        record[5] = point.z;
        record[6] = b0;
        record[7] = b1;
        record[8] = b2;
        c.rows[9 * g] = (int32_t)start_triangle;
        c.rows[9 * g + 1] = (float)tracked * step < c.min_length ? 0 : 1;
        g = -1;
    }
}

// tracking.py:262-357 (kernel tod) on the fast model: the plain TOD
// kernel's loop (tod_kernel: mask, field and output cell, flush on leaving a
// cell, deposit, update, finite and length test, move, flush at the end)
// with persistent lanes and the polynomial model. The deposit is the
// reference's 45-term SH basis at the step's orientation (fast::sh45_basis,
// from z and powers of x + iy instead of angles), added to the lane's run
// (acc = Y on entering a cell, acc += Y after); runs are flushed through the
// LDS stages of flush_runs. The row stays in registers: the LDS holds the
// stages.
template <int MODEL, int ALGO, int LMAX>
__global__ __launch_bounds__(256) void tod_fast_kernel(FastTodLaunch f)
{
    __shared__ RunStage stages[4];
    RunStage &stage = stages[threadIdx.x >> 6];
    const TodLaunch &t = f.tod;
    const float *coeffs = static_cast<const float *>(f.coeffs);
    const uint32_t cstride = f.coeff_stride;
    const float step = t.algo.step_size;
    const int lane_id = __lane_id();

    int64_t g = -1;
    bool drained = false;
    uint2 rng = make_uint2(0u, 0u);
    V3 point = v3(0.0f, 0.0f, 0.0f), o = point;
    uint32_t k = 0;
    int64_t previous = -1;
    int64_t field_cell = -2, cached_mask_cell = -2;
    bool mask_inside = false;
    float q[MODEL == MODEL_SH ? Order<LMAX>::TERMS : 6];
    float acc[45];
#pragma unroll
    for (int i = 0; i < 45; i++)
        acc[i] = 0.0f;

    while (true) {
        // Refill idle lanes from the queue, one atomic per wavefront.
        const uint64_t need = __ballot(g < 0);
        if (need != 0 && !drained) {
            const int leader = __ffsll((unsigned long long)need) - 1;
            const uint32_t count = (uint32_t)__popcll(need);
            uint32_t base = 0;
            if (lane_id == leader)
                base = atomicAdd(t.queue, count);
            base = __shfl(base, leader);
            if ((uint64_t)base + count >= (uint64_t)t.n_seeds)
                drained = true;
            if (g < 0) {
                const uint64_t claimed = (uint64_t)base + (uint64_t)__popcll(need & ((1ull << lane_id) - 1ull));
                if (claimed < (uint64_t)t.n_seeds) {
                    g = (int64_t)claimed;
                    rng = rng_init(t.rng_seed, (uint32_t)g);
                    const BoundarySample sample =
                        sample_boundary(t.vertices, t.triangles, t.cumulative, t.n_triangles, rng);
                    point = sample.point;
                    o = sample.orientation;
                    k = 0;
                    previous = -1;
                    field_cell = -2;
                    cached_mask_cell = -2;
                    mask_inside = false;
                }
            }
        }
        if (__ballot(g >= 0) == 0)
            break;
        const bool live = g >= 0;
        bool stop = false;
        int64_t voxel = -1, cell = -1;
        if (live) {
            stop = k >= t.n_steps;
            if (!stop) {
                const int64_t mc = mask_cell(t.mask, point);
                if (mc != cached_mask_cell) {
                    cached_mask_cell = mc;
                    mask_inside = mc >= 0 && t.mask.mask[mc] != 0;
                }
                stop = !mask_inside;
            }
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
            // On the grid: a new run starts from the basis (keep 0) or the
            // run goes on (keep 1); off it, the run is left as it is.
            const float keep = in_grid && cell != previous ? 0.0f : 1.0f;
            previous = cell;
            fast::sh45_basis(o, in_grid ? 1.0f : 0.0f, [&](int i, float y) { acc[i] = fmaf(acc[i], keep, y); });
            if (voxel != field_cell) {
                field_cell = voxel;
                load_row<MODEL, LMAX, 0>(coeffs, voxel, cstride, q, nullptr);
            }
            o = update_orientation<MODEL, ALGO, LMAX>(t.algo, q, o, rng);
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

__global__ void zero_counter(uint32_t *counter) { *counter = 0u; }

// A generation's claim counter, ended-streamline shards and saved-streamline
// count; before generation 0 also the launch's row and mask read totals.
// One block of 64 threads.
__global__ void zero_generation(uint32_t *claims, uint32_t *ended, uint32_t *cont_count, unsigned long long *stats)
{
    const int t = threadIdx.x;
    if (t < FAST_END_SHARDS)
        ended[t * FAST_END_STRIDE] = 0u;
    if (t == 0) {
        *claims = 0u;
        *cont_count = 0u;
        if (stats) {
            stats[0] = 0ull;
            stats[1] = 0ull;
        }
    }
}

// Dense seed-major rows -> ragged points. One wavefront per streamline copies
// its lengths[g] * 3 floats, coalesced on both sides, eight loads in flight
// per lane before their stores (a load-store pair per iteration left the
// copy latency-bound).
__global__ __launch_bounds__(256) void compact_seed_major_kernel(const float *__restrict__ dense, int64_t n_seeds,
                                                                 uint32_t n_points,
                                                                 const uint32_t *__restrict__ lengths,
                                                                 const int64_t *__restrict__ offsets,
                                                                 float *__restrict__ out)
{
    constexpr int UNROLL = 8;
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const int64_t capacity = n_seeds * (int64_t)n_points * 3;
    for (int64_t g = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); g < n_seeds; g += waves) {
        // Lengths are clamped to the row and writes to the output capacity,
        // whatever the caller's lengths and offsets hold: only i in [lo, hi)
        // is copied, for which 0 <= first + i < capacity.
        const int64_t count = (int64_t)min(lengths[g], n_points) * 3;
        const float *src = dense + g * (int64_t)n_points * 3;
        const int64_t first = offsets[g] * 3;
        const int64_t lo = first < 0 ? -first : 0;
        const int64_t hi = min(count, capacity - first);
        for (int64_t base = lo + lane; base < hi; base += 64 * UNROLL) {
            float v[UNROLL];
#pragma unroll
            for (int k = 0; k < UNROLL; k++) {
                const int64_t i = base + 64 * k;
                v[k] = i < hi ? src[i] : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < UNROLL; k++) {
                const int64_t i = base + 64 * k;
                if (i < hi)
                    out[first + i] = v[k];
            }
        }
    }
}

// out[v, r] = sum_i T[r, i] c[v, i] in FP64 for the SH -> polynomial map of
// order L (T row-major TERMS x TERMS in global memory, wave-uniform reads);
// coefficients past the field's K are zero, rows past TERMS of the padded
// stride are zero.
template <int L>
__global__ __launch_bounds__(256) void sh_to_poly_kernel(const float *__restrict__ field, int64_t n_voxels,
                                                         uint32_t stride, uint32_t K, const double *__restrict__ T,
                                                         float *__restrict__ out, uint32_t out_stride)
{
    constexpr int TERMS = Order<L>::TERMS;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n_voxels)
        return;
    double c[TERMS];
#pragma unroll
    for (int i = 0; i < TERMS; i++)
        c[i] = i < (int)K ? (double)field[v * stride + i] : 0.0;
    float *dst = out + v * out_stride;
    for (int r = 0; r < TERMS; r++) {
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < TERMS; i++)
            acc += T[r * TERMS + i] * c[i];
        dst[r] = (float)acc;
    }
    for (uint32_t r = TERMS; r < out_stride; r++)
        dst[r] = 0.0f;
}

__global__ void copy_rows_kernel(const float *__restrict__ field, int64_t n_voxels, uint32_t stride,
                                 float *__restrict__ out, uint32_t out_stride)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n_voxels)
        return;
    for (uint32_t r = 0; r < out_stride; r++)
        out[v * out_stride + r] = r < stride ? field[v * stride + r] : 0.0f;
}

// T[r, i]: coefficient r of the polynomial layout of order L
// (tgk_fast_model.h) per unit SH coefficient i (MRtrix order). The real
// harmonic of order m > 0 is sqrt(2) N_lm P_l^m(z) cos(m phi) (sin for -m)
// with P_l^m = (-1)^m s^m d^m P_l / dz^m and
// N_lm = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!).
template <int L> std::vector<double> sh_to_poly_matrix()
{
    using O = Order<L>;
    constexpr int TERMS = O::TERMS;
    std::vector<double> T((size_t)TERMS * TERMS, 0.0);
    for (int l = 0; l <= L; l += 2) {
        const std::vector<double> P = shb::legendre(l);
        for (int m = 0; m <= l; m++) {
            const double N = std::sqrt((2.0 * l + 1.0) / (4.0 * M_PI) * shb::factorial(l - m) / shb::factorial(l + m));
            const double F = N * (m == 0 ? 1.0 : std::sqrt(2.0)) * ((m & 1) ? -1.0 : 1.0);
            const int parity = m & 1;
            for (int side = 0; side < (m == 0 ? 1 : 2); side++) {
                const int column = l * (l + 1) / 2 + (side == 0 ? m : -m);
                const int base = side == 0 ? O::a_offset(m) : O::b_offset(m);
                for (int j = 0; j < O::poly_terms(m); j++) {
                    const int power = parity + 2 * j;
                    if (power + m > l)
                        continue;
                    const double coefficient = P[power + m] * shb::factorial(power + m) / shb::factorial(power);
                    T[(size_t)(base + j) * TERMS + column] += F * coefficient;
                }
            }
        }
    }
    return T;
}

// The map of order L on the current device, uploaded on first use and kept
// for the life of the process. Runtimes on different threads may ask for it
// at once: the first upload per device happens under a lock.
template <int L> hipError_t device_matrix(const double **out)
{
    static const double *matrices[64] = {};
    static std::mutex lock;
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess)
        return e;
    if (device < 0 || device >= 64)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> guard(lock);
    if (!matrices[device]) {
        const std::vector<double> T = sh_to_poly_matrix<L>();
        void *ptr = nullptr;
        e = hipMalloc(&ptr, T.size() * sizeof(double));
        if (e != hipSuccess)
            return e;
        e = hipMemcpy(ptr, T.data(), T.size() * sizeof(double), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(ptr);
            return e;
        }
        matrices[device] = static_cast<const double *>(ptr);
    }
    *out = matrices[device];
    return hipSuccess;
}

// Blocks that fill the current device once, computed on first use per
// device (atomically published: runtimes on other threads may read it).
template <int MODEL, int ALGO, int DOMAIN, int LMAX, int S> hipError_t occupancy_grid(int *blocks)
{
    static std::atomic<int> cache[64] = {};
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess)
        return e;
    if (device < 0 || device >= 64)
        return hipErrorInvalidDevice;
    if (cache[device].load(std::memory_order_acquire) == 0) {
        int per_cu = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void *>(track_fast_kernel<MODEL, ALGO, DOMAIN, LMAX, S>),
            shape_block<MODEL, ALGO, LMAX, S>(), 0);
        if (e != hipSuccess)
            return e;
        hipDeviceProp_t prop;
        e = hipGetDeviceProperties(&prop, device);
        if (e != hipSuccess)
            return e;
        cache[device].store(prop.multiProcessorCount * (per_cu > 0 ? per_cu : 1), std::memory_order_release);
    }
    *blocks = cache[device].load(std::memory_order_acquire);
    return hipSuccess;
}

// query: only report the lanes of the grid a launch would use.
template <int MODEL, int ALGO, int DOMAIN, int LMAX, int S>
hipError_t launch_fast(const FastLaunch &f, int grid_blocks, bool query, int64_t *lanes, hipStream_t stream)
{
    constexpr int BLOCK = shape_block<MODEL, ALGO, LMAX, S>();
    if (grid_blocks <= 0) {
        const hipError_t e = occupancy_grid<MODEL, ALGO, DOMAIN, LMAX, S>(&grid_blocks);
        if (e != hipSuccess)
            return e;
    }
    if (lanes)
        *lanes = (int64_t)grid_blocks * BLOCK;
    if (query)
        return hipSuccess;
    // Generation g resumes what g - 1 saved in cont[(g - 1) % 2] and saves
    // into cont[g % 2], whose count (generation g - 2's, consumed by g - 1)
    // it clears first.
    // A first generation that is never cut short (half its resident work
    // below the repack threshold) is the only one.
    const uint64_t first_work = std::min<uint64_t>((uint64_t)f.track.n_seeds, (uint64_t)grid_blocks * BLOCK);
    const int generations = first_work / 2 >= f.min_repack ? FAST_GENERATIONS : 1;
    FastLaunch gen = f;
    for (int g = 0; g < generations; g++) {
        gen.cont_in = g == 0 ? nullptr : f.cont[(g - 1) % 2];
        gen.cont_in_count = g == 0 ? nullptr : f.cont_count[(g - 1) % 2];
        gen.cont_out = f.cont[g % 2];
        gen.cont_out_count = f.cont_count[g % 2];
        gen.last = g == generations - 1;
        zero_generation<<<1, 64, 0, stream>>>(f.claims, f.ended, gen.cont_out_count, g == 0 ? f.stats : nullptr);
        const int blocks_g = std::max(1, grid_blocks >> g);
        track_fast_kernel<MODEL, ALGO, DOMAIN, LMAX, S><<<blocks_g, BLOCK, 0, stream>>>(gen);
    }
    return hipGetLastError();
}

template <int MODEL, int ALGO, int DOMAIN>
hipError_t launch_fast_order(const FastLaunch &f, int grid, bool query, int64_t *lanes, hipStream_t stream)
{
    if constexpr (MODEL == MODEL_SH) {
        if (f.sh_order == 12)
            return launch_fast<MODEL, ALGO, DOMAIN, 12, 0>(f, grid, query, lanes, stream);
    }
    if constexpr (shaped<MODEL, ALGO, 8>()) {
        switch (f.shape) {
        case 1: return launch_fast<MODEL, ALGO, DOMAIN, 8, 1>(f, grid, query, lanes, stream);
        case 2: return launch_fast<MODEL, ALGO, DOMAIN, 8, 2>(f, grid, query, lanes, stream);
        case 3: return launch_fast<MODEL, ALGO, DOMAIN, 8, 3>(f, grid, query, lanes, stream);
        case 4: return launch_fast<MODEL, ALGO, DOMAIN, 8, 4>(f, grid, query, lanes, stream);
        default: break;
        }
    }
    return launch_fast<MODEL, ALGO, DOMAIN, 8, 0>(f, grid, query, lanes, stream);
}

template <int MODEL, int ALGO>
hipError_t launch_fast_domain(const FastLaunch &f, int grid, bool query, int64_t *blocks, hipStream_t stream)
{
    if (f.track.domain_kind == DOMAIN_MESH)
        return launch_fast_order<MODEL, ALGO, DOMAIN_MESH>(f, grid, query, blocks, stream);
    return launch_fast_order<MODEL, ALGO, DOMAIN_MASK>(f, grid, query, blocks, stream);
}

template <int MODEL>
hipError_t launch_fast_algo(const FastLaunch &f, int grid, bool query, int64_t *blocks, hipStream_t stream)
{
    switch (f.track.algo.algorithm) {
    case ALGO_DIFFUSION: return launch_fast_domain<MODEL, ALGO_DIFFUSION>(f, grid, query, blocks, stream);
    case ALGO_PROBABILISTIC: return launch_fast_domain<MODEL, ALGO_PROBABILISTIC>(f, grid, query, blocks, stream);
    case ALGO_DETERMINISTIC: return launch_fast_domain<MODEL, ALGO_DETERMINISTIC>(f, grid, query, blocks, stream);
    default: return launch_fast_domain<MODEL, ALGO_TRANSPORT>(f, grid, query, blocks, stream);
    }
}

hipError_t dispatch_fast(const FastLaunch &f, int grid, bool query, int64_t *blocks, hipStream_t stream)
{
    if (f.track.algo.model == MODEL_DTI)
        return launch_fast_algo<MODEL_DTI>(f, grid, query, blocks, stream);
    return launch_fast_algo<MODEL_SH>(f, grid, query, blocks, stream);
}

// Connectome launches: the persistent grid of the kernel's occupancy
// (cached per device), at most one lane per seed, then the hits pass.
template <int MODEL, int ALGO, int LMAX> hipError_t launch_connectome_fast_kernel(const FastConnectomeLaunch &f,
                                                                                  hipStream_t stream)
{
    static std::atomic<int> cache[64] = {};
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess)
        return e;
    if (device < 0 || device >= 64)
        return hipErrorInvalidDevice;
    if (cache[device].load(std::memory_order_acquire) == 0) {
        int per_cu = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void *>(connectome_fast_kernel<MODEL, ALGO, LMAX>), fast_block<LMAX>(),
            0);
        if (e != hipSuccess)
            return e;
        int cus = 0;
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
        if (e != hipSuccess)
            return e;
        cache[device].store((cus > 0 ? cus : 1) * (per_cu > 0 ? per_cu : 1), std::memory_order_release);
    }
    const int64_t fill = cache[device].load(std::memory_order_acquire);
    const int64_t needed = (f.connectome.n_seeds + fast_block<LMAX>() - 1) / fast_block<LMAX>();
    zero_counter<<<1, 1, 0, stream>>>(f.connectome.queue);
    connectome_fast_kernel<MODEL, ALGO, LMAX><<<(unsigned)std::min(fill, needed), fast_block<LMAX>(), 0, stream>>>(f);
    e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    return launch_connectome_hits(f.connectome, stream);
}

template <int MODEL, int ALGO>
hipError_t launch_connectome_fast_order(const FastConnectomeLaunch &f, hipStream_t stream)
{
    if constexpr (MODEL == MODEL_SH) {
        if (f.sh_order == 12)
            return launch_connectome_fast_kernel<MODEL, ALGO, 12>(f, stream);
    }
    return launch_connectome_fast_kernel<MODEL, ALGO, 8>(f, stream);
}

template <int MODEL>
hipError_t launch_connectome_fast_algo(const FastConnectomeLaunch &f, hipStream_t stream)
{
    switch (f.connectome.algo.algorithm) {
    case ALGO_DIFFUSION: return launch_connectome_fast_order<MODEL, ALGO_DIFFUSION>(f, stream);
    case ALGO_PROBABILISTIC: return launch_connectome_fast_order<MODEL, ALGO_PROBABILISTIC>(f, stream);
    case ALGO_DETERMINISTIC: return launch_connectome_fast_order<MODEL, ALGO_DETERMINISTIC>(f, stream);
    default: return launch_connectome_fast_order<MODEL, ALGO_TRANSPORT>(f, stream);
    }
}

// TOD launches: the persistent grid of the kernel's occupancy (cached per
// device), at most one lane per seed.
template <int MODEL, int ALGO, int LMAX> hipError_t launch_tod_fast_kernel(const FastTodLaunch &f, hipStream_t stream)
{
    static std::atomic<int> cache[64] = {};
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess)
        return e;
    if (device < 0 || device >= 64)
        return hipErrorInvalidDevice;
    if (cache[device].load(std::memory_order_acquire) == 0) {
        int per_cu = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void *>(tod_fast_kernel<MODEL, ALGO, LMAX>), 256, 0);
        if (e != hipSuccess)
            return e;
        int cus = 0;
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
        if (e != hipSuccess)
            return e;
        cache[device].store((cus > 0 ? cus : 1) * (per_cu > 0 ? per_cu : 1), std::memory_order_release);
    }
    const int64_t fill = cache[device].load(std::memory_order_acquire);
    const int64_t needed = (f.tod.n_seeds + 255) / 256;
    zero_counter<<<1, 1, 0, stream>>>(f.tod.queue);
    tod_fast_kernel<MODEL, ALGO, LMAX><<<(unsigned)std::min(fill, needed), 256, 0, stream>>>(f);
    return hipGetLastError();
}

template <int MODEL, int ALGO>
hipError_t launch_tod_fast_order(const FastTodLaunch &f, hipStream_t stream)
{
    if constexpr (MODEL == MODEL_SH) {
        if (f.sh_order == 12)
            return launch_tod_fast_kernel<MODEL, ALGO, 12>(f, stream);
    }
    return launch_tod_fast_kernel<MODEL, ALGO, 8>(f, stream);
}

template <int MODEL>
hipError_t launch_tod_fast_algo(const FastTodLaunch &f, hipStream_t stream)
{
    switch (f.tod.algo.algorithm) {
    case ALGO_DIFFUSION: return launch_tod_fast_order<MODEL, ALGO_DIFFUSION>(f, stream);
    case ALGO_PROBABILISTIC: return launch_tod_fast_order<MODEL, ALGO_PROBABILISTIC>(f, stream);
    case ALGO_DETERMINISTIC: return launch_tod_fast_order<MODEL, ALGO_DETERMINISTIC>(f, stream);
    default: return launch_tod_fast_order<MODEL, ALGO_TRANSPORT>(f, stream);
    }
}

}  // namespace

hipError_t launch_track_fast(const FastLaunch &f, int grid_blocks, hipStream_t stream)
{
    if (f.track.n_seeds == 0)
        return hipSuccess;
    if (f.track.n_seeds >= (int64_t)1 << 31)
        return hipErrorInvalidValue;
    return dispatch_fast(f, grid_blocks, false, nullptr, stream);
}

hipError_t prepare_track_fast(const FastLaunch &f, int grid_blocks, int64_t *lanes)
{
    return dispatch_fast(f, grid_blocks, true, lanes, nullptr);
}

hipError_t launch_connectome_fast(const FastConnectomeLaunch &f, hipStream_t stream)
{
    if (f.connectome.n_seeds == 0)
        return hipSuccess;
    if (f.connectome.algo.model == MODEL_DTI)
        return launch_connectome_fast_algo<MODEL_DTI>(f, stream);
    return launch_connectome_fast_algo<MODEL_SH>(f, stream);
}

hipError_t launch_tod_fast(const FastTodLaunch &f, hipStream_t stream)
{
    if (f.tod.n_seeds == 0)
        return hipSuccess;
    if (f.tod.algo.model == MODEL_DTI)
        return launch_tod_fast_algo<MODEL_DTI>(f, stream);
    return launch_tod_fast_algo<MODEL_SH>(f, stream);
}

hipError_t launch_compact_seed_major(const float *dense, int64_t n_seeds, uint32_t n_points,
                                     const uint32_t *lengths, const int64_t *offsets, float *out,
                                     hipStream_t stream)
{
    if (n_seeds == 0)
        return hipSuccess;
    const int64_t blocks = (n_seeds + 3) / 4;
    compact_seed_major_kernel<<<(unsigned)(blocks < 65536 ? blocks : 65536), 256, 0, stream>>>(
        dense, n_seeds, n_points, lengths, offsets, out);
    return hipGetLastError();
}

hipError_t launch_field_transform(const float *field, int64_t n_voxels, uint32_t stride, int32_t model, int32_t K,
                                  int32_t sh_order, void *out, uint32_t out_stride, hipStream_t stream)
{
    if (n_voxels == 0)
        return hipSuccess;
    const unsigned blocks = (unsigned)((n_voxels + 255) / 256);
    if (model == MODEL_SH) {
        const double *T = nullptr;
        hipError_t e = sh_order == 12 ? device_matrix<12>(&T) : device_matrix<8>(&T);
        if (e != hipSuccess)
            return e;
        if (sh_order == 12)
            sh_to_poly_kernel<12><<<blocks, 256, 0, stream>>>(field, n_voxels, stride, (uint32_t)K, T,
                                                              static_cast<float *>(out), out_stride);
        else
            sh_to_poly_kernel<8><<<blocks, 256, 0, stream>>>(field, n_voxels, stride, (uint32_t)K, T,
                                                             static_cast<float *>(out), out_stride);
    } else {
        copy_rows_kernel<<<blocks, 256, 0, stream>>>(field, n_voxels, stride, static_cast<float *>(out), out_stride);
    }
    return hipGetLastError();
}

}  // namespace tgk
