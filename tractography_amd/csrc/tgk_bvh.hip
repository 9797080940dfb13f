// Host-side BVH construction over triangle bounding boxes (reference
// domain.py:72-116), the build half of the mesh-domain termination test.
// Host code only: the tree is built once per mesh and uploaded with
// tgk_mesh_domain_create.
//
// The rule fixes the hit-triangle ids the traversal reports, so it is kept
// exactly:
//   - a node over <= leaf_size triangles is a leaf;
//   - otherwise split on the longest axis of the node's box (the first axis
//     on ties), ordering the node's triangles by box centroid with a stable
//     sort, the left child taking the first n/2;
//   - nodes are numbered in pre-order, left subtree first, and leaves claim
//     consecutive slots of the ordered-triangle list in the same order.
// Centroids and boxes are formed in double from double inputs; bounds are
// rounded to float once at the end. NaN follows NumPy: it propagates through
// min/max, wins argmax, and sorts after every number.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <utility>
#include <vector>

#include "../../include/tgk.h"
#include "tgk_internal.h"

namespace {

uint64_t node_count(uint64_t n, uint32_t leaf_size, std::map<uint64_t, uint64_t> &memo)
{
    if (n <= leaf_size)
        return 1;
    auto it = memo.find(n);
    if (it != memo.end())
        return it->second;
    const uint64_t c = 1 + node_count(n / 2, leaf_size, memo) + node_count(n - n / 2, leaf_size, memo);
    memo.emplace(n, c);
    return c;
}

inline double nan_min(double a, double b) { return (std::isnan(a) || std::isnan(b)) ? NAN : (b < a ? b : a); }
inline double nan_max(double a, double b) { return (std::isnan(a) || std::isnan(b)) ? NAN : (b > a ? b : a); }

struct Builder {
    const double *tmin;  // (n, 3)
    const double *tmax;  // (n, 3)
    uint32_t leaf_size;
    float *bounds_min;   // (N, 4)
    float *bounds_max;   // (N, 4)
    int32_t *nodes;      // (N, 4)
    uint32_t *ordered;   // (n,)
    std::vector<uint32_t> ids;
    std::vector<std::pair<double, uint32_t>> keyed;  // sort scratch
    uint32_t next_node = 0;
    uint32_t next_slot = 0;

    int32_t build(size_t lo, size_t hi)
    {
        const uint32_t id = next_node++;
        double bmin[3], bmax[3];
        for (int k = 0; k < 3; ++k) {
            bmin[k] = tmin[(size_t)ids[lo] * 3 + k];
            bmax[k] = tmax[(size_t)ids[lo] * 3 + k];
        }
        for (size_t i = lo + 1; i < hi; ++i)
            for (int k = 0; k < 3; ++k) {
                bmin[k] = nan_min(bmin[k], tmin[(size_t)ids[i] * 3 + k]);
                bmax[k] = nan_max(bmax[k], tmax[(size_t)ids[i] * 3 + k]);
            }
        for (int k = 0; k < 3; ++k) {
            bounds_min[(size_t)id * 4 + k] = (float)bmin[k];
            bounds_max[(size_t)id * 4 + k] = (float)bmax[k];
        }
        bounds_min[(size_t)id * 4 + 3] = 0.0f;
        bounds_max[(size_t)id * 4 + 3] = 0.0f;
        int32_t *node = nodes + (size_t)id * 4;
        const size_t n = hi - lo;
        if (n <= leaf_size) {
            node[0] = -1;
            node[1] = -1;
            node[2] = (int32_t)next_slot;
            node[3] = (int32_t)n;
            for (size_t i = lo; i < hi; ++i)
                ordered[next_slot++] = ids[i];
            return (int32_t)id;
        }
        // numpy argmax: the first NaN, else the first maximum.
        int axis = 0;
        double best = bmax[0] - bmin[0];
        if (!std::isnan(best))
            for (int k = 1; k < 3; ++k) {
                const double e = bmax[k] - bmin[k];
                if (std::isnan(e)) {
                    axis = k;
                    break;
                }
                if (e > best) {
                    best = e;
                    axis = k;
                }
            }
        keyed.resize(n);
        for (size_t i = 0; i < n; ++i) {
            const uint32_t t = ids[lo + i];
            keyed[i] = {(tmin[(size_t)t * 3 + axis] + tmax[(size_t)t * 3 + axis]) * 0.5, t};
        }
        std::stable_sort(keyed.begin(), keyed.end(), [](const auto &a, const auto &b) {
            return a.first < b.first || (!std::isnan(a.first) && std::isnan(b.first));
        });
        for (size_t i = 0; i < n; ++i)
            ids[lo + i] = keyed[i].second;
        const size_t middle = lo + n / 2;
        node[2] = 0;
        node[3] = 0;
        const int32_t left = build(lo, middle);
        const int32_t right = build(middle, hi);
        node = nodes + (size_t)id * 4;
        node[0] = left;
        node[1] = right;
        return (int32_t)id;
    }
};

}  // namespace

extern "C" {

int tgk_bvh_node_count(uint32_t n_triangles, uint32_t leaf_size, uint32_t *n_nodes)
{
    if (!n_nodes)
        return tgk::report_error(TGK_ERR_INVALID, "NULL argument");
    if (n_triangles == 0 || leaf_size == 0)
        return tgk::report_error(TGK_ERR_INVALID, "empty BVH (%u triangles, leaf size %u)", n_triangles, leaf_size);
    std::map<uint64_t, uint64_t> memo;
    const uint64_t c = node_count(n_triangles, leaf_size, memo);
    if (c > INT32_MAX)
        return tgk::report_error(TGK_ERR_INVALID, "%u triangles need more than 2^31 BVH nodes", n_triangles);
    *n_nodes = (uint32_t)c;
    return TGK_OK;
}

int tgk_bvh_build(const double *triangle_min, const double *triangle_max, uint32_t n_triangles, uint32_t leaf_size,
                  float *bounds_min, float *bounds_max, int32_t *nodes, uint32_t n_nodes, uint32_t *ordered)
{
    if (!triangle_min || !triangle_max || !bounds_min || !bounds_max || !nodes || !ordered)
        return tgk::report_error(TGK_ERR_INVALID, "NULL argument");
    uint32_t expected = 0;
    if (int r = tgk_bvh_node_count(n_triangles, leaf_size, &expected))
        return r;
    if (n_nodes != expected)
        return tgk::report_error(TGK_ERR_INVALID, "BVH over %u triangles has %u nodes, not %u", n_triangles, expected,
                         n_nodes);
    try {
        Builder b{triangle_min, triangle_max, leaf_size, bounds_min, bounds_max, nodes, ordered, {}, {}};
        b.ids.resize(n_triangles);
        for (uint32_t i = 0; i < n_triangles; ++i)
            b.ids[i] = i;
        b.keyed.reserve(n_triangles);
        b.build(0, n_triangles);
    } catch (const std::bad_alloc &) {
        return tgk::report_error(TGK_ERR_NO_MEMORY, "BVH build scratch");
    }
    return TGK_OK;
}

}  // extern "C"

// Per-cell candidate lists of a mesh domain (the fast path's answer for
// short segments): triangle t is a candidate of every cell in its inclusive
// cell box [lo_t, hi_t] (a box with lo > hi on an axis covers no cell).
// Within a cell, candidates follow BVH leaf order, the order in which a
// traversal reaches them, so that testing them in turn with the traversal's
// replacement rule reproduces its closest hit. Cells are numbered
// (x * dims[1] + y) * dims[2] + z.
namespace {

int check_grid_args(const int64_t dims[3], const int64_t *lo, const int64_t *hi, const uint32_t *leaf_order,
                    uint32_t n_triangles, size_t *n_cells)
{
    if (!dims || !lo || !hi || !leaf_order)
        return tgk::report_error(TGK_ERR_INVALID, "NULL argument");
    if (dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0 || dims[0] > (1 << 20) || dims[1] > (1 << 20) ||
        dims[2] > (1 << 20))
        return tgk::report_error(TGK_ERR_INVALID, "grid dims (%lld, %lld, %lld) out of range", (long long)dims[0],
                                 (long long)dims[1], (long long)dims[2]);
    const double cells = (double)dims[0] * (double)dims[1] * (double)dims[2];
    if (cells >= 4294967295.0)
        return tgk::report_error(TGK_ERR_INVALID, "grid of %.0f cells exceeds 32-bit cell ids", cells);
    *n_cells = (size_t)dims[0] * dims[1] * dims[2];
    for (uint32_t t = 0; t < n_triangles; ++t) {
        if (leaf_order[t] >= n_triangles)
            return tgk::report_error(TGK_ERR_INVALID, "leaf order entry %u = %u out of range", t, leaf_order[t]);
        for (int k = 0; k < 3; ++k)
            if (lo[(size_t)t * 3 + k] < 0 || hi[(size_t)t * 3 + k] >= dims[k])
                return tgk::report_error(TGK_ERR_INVALID, "cell box of triangle %u leaves the grid", t);
    }
    return TGK_OK;
}

// Calls visit(cell) for every cell of triangle t's box.
template <class F> inline void for_each_cell(const int64_t dims[3], const int64_t *lo, const int64_t *hi, uint32_t t,
                                             F &&visit)
{
    const int64_t *l = lo + (size_t)t * 3, *h = hi + (size_t)t * 3;
    for (int64_t x = l[0]; x <= h[0]; ++x)
        for (int64_t y = l[1]; y <= h[1]; ++y)
            for (int64_t z = l[2]; z <= h[2]; ++z)
                visit((size_t)((x * dims[1] + y) * dims[2] + z));
}

}  // namespace

extern "C" {

int tgk_segment_grid_count(const int64_t dims[3], const int64_t *lo, const int64_t *hi, const uint32_t *leaf_order,
                           uint32_t n_triangles, uint32_t *cell_start, uint64_t *n_entries)
{
    size_t n_cells = 0;
    if (int r = check_grid_args(dims, lo, hi, leaf_order, n_triangles, &n_cells))
        return r;
    if (!cell_start || !n_entries)
        return tgk::report_error(TGK_ERR_INVALID, "NULL argument");
    try {
        std::vector<uint64_t> count(n_cells, 0);
        for (uint32_t t = 0; t < n_triangles; ++t)
            for_each_cell(dims, lo, hi, t, [&](size_t c) { ++count[c]; });
        uint64_t total = 0;
        cell_start[0] = 0;
        for (size_t c = 0; c < n_cells; ++c) {
            total += count[c];
            if (total > UINT32_MAX)
                return tgk::report_error(TGK_ERR_INVALID, "more than 2^32 - 1 candidate entries");
            cell_start[c + 1] = (uint32_t)total;
        }
        *n_entries = total;
    } catch (const std::bad_alloc &) {
        return tgk::report_error(TGK_ERR_NO_MEMORY, "segment grid counts");
    }
    return TGK_OK;
}

int tgk_segment_grid_fill(const int64_t dims[3], const int64_t *lo, const int64_t *hi, const uint32_t *leaf_order,
                          uint32_t n_triangles, const float *vertices, uint32_t n_vertices, const uint32_t *triangles,
                          const uint32_t *cell_start, float *entries, uint64_t n_entries)
{
    size_t n_cells = 0;
    if (int r = check_grid_args(dims, lo, hi, leaf_order, n_triangles, &n_cells))
        return r;
    if (!vertices || !triangles || !cell_start || !entries)
        return tgk::report_error(TGK_ERR_INVALID, "NULL argument");
    if (cell_start[n_cells] != n_entries)
        return tgk::report_error(TGK_ERR_INVALID, "cell_start ends at %u, not at %llu entries", cell_start[n_cells],
                                 (unsigned long long)n_entries);
    for (size_t i = 0; i < (size_t)n_triangles * 4; ++i)
        if (i % 4 != 3 && triangles[i] >= n_vertices)
            return tgk::report_error(TGK_ERR_INVALID, "triangle %zu names vertex %u of %u", i / 4, triangles[i],
                                     n_vertices);
    try {
        std::vector<uint32_t> cursor(cell_start, cell_start + n_cells);
        for (uint32_t p = 0; p < n_triangles; ++p) {
            const uint32_t slot = leaf_order[p];
            const uint32_t *tri = triangles + (size_t)slot * 4;
            const float *a = vertices + (size_t)tri[0] * 4;
            const float *b = vertices + (size_t)tri[1] * 4;
            const float *c = vertices + (size_t)tri[2] * 4;
            float row[12];
            for (int k = 0; k < 3; ++k) {
                row[k] = a[k];
                row[4 + k] = b[k] - a[k];
                row[8 + k] = c[k] - a[k];
            }
            memcpy(&row[3], &slot, sizeof slot);
            row[7] = 0.0f;
            row[11] = 0.0f;
            bool overflow = false;
            for_each_cell(dims, lo, hi, slot, [&](size_t cell) {
                const uint32_t e = cursor[cell]++;
                if (e >= cell_start[cell + 1]) {
                    overflow = true;
                    return;
                }
                memcpy(entries + (size_t)e * 12, row, sizeof row);
            });
            if (overflow)
                return tgk::report_error(TGK_ERR_INVALID, "cell_start does not match the triangles' cell boxes");
        }
    } catch (const std::bad_alloc &) {
        return tgk::report_error(TGK_ERR_NO_MEMORY, "segment grid cursors");
    }
    return TGK_OK;
}

}  // extern "C"
