"""NumPy restatement of the reference's BVH build (domain.py:72-116).

TEST INFRASTRUCTURE ONLY: the checker the native builder (tgk_bvh_build in
libtgk.so) is compared against. The product never imports it.

The reference recurses (domain.py:85-106); this walks an explicit stack in
the same pre-order (node, left subtree, right subtree). The reference sorts
centroids with NumPy's default argsort, whose order among equal keys is not
specified; this restatement, like the product, fixes it with a stable sort.
"""

from __future__ import annotations

import numpy as np


def build_bvh(triangle_min: np.ndarray, triangle_max: np.ndarray, leaf_size: int = 4):
    """BVH over triangle bounding boxes (domain.py:72-116).

    Returns (bounds_min (N, 4) f32, bounds_max (N, 4) f32, nodes (N, 4) i32
    as (left, right, first, count) with left = right = -1 for leaves, ordered
    triangle slots (M,) u32). The centroid sort is stable, so equal centroids
    keep their slot order.
    """
    triangle_min = np.asarray(triangle_min, dtype=np.float64)
    triangle_max = np.asarray(triangle_max, dtype=np.float64)
    centroids = (triangle_min + triangle_max) * 0.5
    node_min: list[np.ndarray] = []
    node_max: list[np.ndarray] = []
    nodes: list[list[int]] = []
    ordered: list[np.ndarray] = []
    n_ordered = 0
    stack = [(np.arange(len(triangle_min), dtype=np.int64), -1, 0)]
    while stack:
        ids, parent, side = stack.pop()
        node_id = len(nodes)
        if parent >= 0:
            nodes[parent][side] = node_id
        bmin = triangle_min[ids].min(axis=0)
        bmax = triangle_max[ids].max(axis=0)
        node_min.append(bmin)
        node_max.append(bmax)
        if len(ids) <= leaf_size:
            nodes.append([-1, -1, n_ordered, len(ids)])
            ordered.append(ids)
            n_ordered += len(ids)
        else:
            nodes.append([0, 0, 0, 0])
            axis = int(np.argmax(bmax - bmin))
            sorted_ids = ids[np.argsort(centroids[ids, axis], kind="stable")]
            middle = len(sorted_ids) // 2
            stack.append((sorted_ids[middle:], node_id, 1))
            stack.append((sorted_ids[:middle], node_id, 0))
    bounds_min = np.zeros((len(nodes), 4), dtype=np.float32)
    bounds_max = np.zeros((len(nodes), 4), dtype=np.float32)
    bounds_min[:, :3] = np.asarray(node_min)
    bounds_max[:, :3] = np.asarray(node_max)
    return (
        bounds_min,
        bounds_max,
        np.ascontiguousarray(nodes, dtype=np.int32),
        np.ascontiguousarray(np.concatenate(ordered), dtype=np.uint32),
    )
