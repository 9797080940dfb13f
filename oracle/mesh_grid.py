"""NumPy restatement of a mesh domain's culling grid and per-cell candidate
lists (the host-built tables behind the fast path's short-segment test).

TEST INFRASTRUCTURE ONLY: the checker the native builders
(tgk_segment_grid_count / tgk_segment_grid_fill in libtgk.so) are compared
against. The product never imports it.

The reference has no such tables: its traversal (bvh.cl:83-131) always
walks the BVH. A triangle is a candidate of every cell of its inclusive cell
box [lo, hi]; a cell lists its candidates in BVH leaf order, the order in
which the traversal reaches them.
"""

from __future__ import annotations

import numpy as np


def culling_grid(dims, lo, hi) -> np.ndarray:
    """near (X, Y, Z) u8: 1 on every cell some triangle's box covers."""
    near = np.zeros(tuple(dims), dtype=np.uint8)
    span = (hi - lo).max(axis=0)
    for ox in range(span[0] + 1):
        for oy in range(span[1] + 1):
            for oz in range(span[2] + 1):
                cell = lo + np.array([ox, oy, oz])
                ok = np.all(cell <= hi, axis=1)
                near[cell[ok, 0], cell[ok, 1], cell[ok, 2]] = 1
    return near


def segment_grid(dims, lo, hi, leaf_order, vertices4, indices4):
    """cell_start (n_cells + 1,) u32 and entries (n, 12) f32 rows
    (a, slot bits, b - a, 0, c - a, 0), candidates in leaf order per cell."""
    n_tri = len(indices4)
    position = np.empty(n_tri, dtype=np.int64)
    position[np.asarray(leaf_order, dtype=np.int64)] = np.arange(n_tri)
    cells, slots = [], []
    span = (hi - lo).max(axis=0)
    slot_ids = np.arange(n_tri, dtype=np.int64)
    for ox in range(span[0] + 1):
        for oy in range(span[1] + 1):
            for oz in range(span[2] + 1):
                cell = lo + np.array([ox, oy, oz])
                ok = np.all(cell <= hi, axis=1)
                c = cell[ok]
                cells.append((c[:, 0] * dims[1] + c[:, 1]) * dims[2] + c[:, 2])
                slots.append(slot_ids[ok])
    cells = np.concatenate(cells)
    slots = np.concatenate(slots)
    order = np.lexsort((position[slots], cells))
    cells, slots = cells[order], slots[order]
    n_cells = int(np.prod(dims))
    cell_start = np.zeros(n_cells + 1, dtype=np.uint32)
    np.cumsum(np.bincount(cells, minlength=n_cells), out=cell_start[1:])
    vertices = vertices4[:, :3]
    idx = indices4[slots, :3].astype(np.int64)
    a = vertices[idx[:, 0]]
    entries = np.zeros((len(slots), 12), dtype=np.float32)
    entries[:, 0:3] = a
    entries[:, 3] = slots.astype(np.uint32).view(np.float32)
    entries[:, 4:7] = vertices[idx[:, 1]] - a
    entries[:, 8:11] = vertices[idx[:, 2]] - a
    return cell_start, np.ascontiguousarray(entries)
