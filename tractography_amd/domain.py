"""Tracking domains (domain.py of the reference).

* ``Image``: the nonzero voxels of a 3-D mask, cells half-open
  ([i - 0.5, i + 0.5) on each axis). Tracking stops when a point leaves it.
* ``Mesh``: a region bounded by closed triangle surfaces (for example the
  shell between two nested surfaces). Tracking stops when a step's segment
  hits a boundary triangle; the triangle's slot is reported per streamline.

Every domain carries a boundary surface with area-weighted sampling (for
integer seed counts) and a BVH over the same triangles, built with the
reference's rule: leaves of at most 4 triangles, split on the longest axis of
the node bounds at the median of the triangle-bound centroids, nodes numbered
in depth-first preorder, left child first.
"""

from __future__ import annotations

import ctypes
from abc import ABC, abstractmethod

import numpy as np

from . import _native
from .image import image_data
from .runtime import DeviceHandle, HipRuntime
from .utils import apply_affine

Points = np.ndarray
ContainsResult = np.ndarray


class BoundaryMesh:
    """Triangle surface with the attributes the tracking path reads."""

    def __init__(self, vertices, faces):
        self.vertices = np.asarray(vertices, dtype=np.float64).reshape(-1, 3)
        self.faces = np.asarray(faces, dtype=np.int64).reshape(-1, 3)

    @property
    def triangles(self) -> np.ndarray:
        return self.vertices[self.faces]

    @property
    def _cross(self) -> np.ndarray:
        t = self.triangles
        return np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])

    @property
    def area_faces(self) -> np.ndarray:
        return 0.5 * np.linalg.norm(self._cross, axis=1)

    @property
    def area(self) -> float:
        return float(np.sum(self.area_faces))

    @property
    def face_normals(self) -> np.ndarray:
        c = self._cross
        return c / np.linalg.norm(c, axis=1, keepdims=True)

    @property
    def triangles_center(self) -> np.ndarray:
        return self.triangles.mean(axis=1)

    @property
    def bounds(self) -> np.ndarray:
        return np.stack([self.vertices.min(axis=0), self.vertices.max(axis=0)])

    @property
    def is_watertight(self) -> bool:
        """Every directed edge appears once and its reverse once."""
        f = self.faces
        edges = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        forward = {tuple(e) for e in edges.tolist()}
        if len(forward) != len(edges):
            return False
        return all((b, a) in forward for a, b in forward)


def build_bvh(triangle_min: np.ndarray, triangle_max: np.ndarray, leaf_size: int = 4):
    """BVH over triangle bounding boxes (domain.py:72-116), built natively.

    Leaves hold at most `leaf_size` triangles; an inner node splits on the
    longest axis of its box at the median of a stable centroid sort. Returns
    (bounds_min (N, 4) f32, bounds_max (N, 4) f32, nodes (N, 4) i32 as
    (left, right, first, count) with left = right = -1 for leaves, ordered
    triangle slots (M,) u32), nodes in pre-order with the left subtree first.
    """
    triangle_min = np.ascontiguousarray(triangle_min, dtype=np.float64)
    triangle_max = np.ascontiguousarray(triangle_max, dtype=np.float64)
    if triangle_min.ndim != 2 or triangle_min.shape[1] != 3 or triangle_max.shape != triangle_min.shape:
        raise ValueError("triangle bounds must be two (n, 3) arrays")
    n = len(triangle_min)
    if n >= 2**32:
        raise ValueError(f"{n} triangles exceed the BVH's 32-bit slots")
    lib = _native.lib()
    count = ctypes.c_uint32()
    _native.check(lib.tgk_bvh_node_count(n, leaf_size, ctypes.byref(count)))
    n_nodes = count.value
    bounds_min = np.empty((n_nodes, 4), dtype=np.float32)
    bounds_max = np.empty((n_nodes, 4), dtype=np.float32)
    nodes = np.empty((n_nodes, 4), dtype=np.int32)
    ordered = np.empty(n, dtype=np.uint32)
    _native.check(
        lib.tgk_bvh_build(_native.ptr(triangle_min), _native.ptr(triangle_max), n, leaf_size,
                          _native.ptr(bounds_min), _native.ptr(bounds_max), _native.ptr(nodes), n_nodes,
                          _native.ptr(ordered))
    )
    return bounds_min, bounds_max, nodes, ordered


class BaseDomain(ABC):
    """A region constraining tracking, with a sampled boundary and its BVH.

    The boundary arrays (``boundary_vertices``, ``boundary_indices``,
    ``boundary_cumulative_areas``, ``boundary_bvh_*``) are built on first use
    (domain.py:44-116 builds them eagerly).
    """

    def __init__(self, boundary: BoundaryMesh | None = None) -> None:
        self._boundary = boundary
        self._boundary_data = None
        self._device: dict[int, DeviceHandle] = {}

    def _make_boundary(self) -> BoundaryMesh:
        raise NotImplementedError

    @property
    def boundary(self) -> BoundaryMesh:
        if self._boundary is None:
            self._boundary = self._make_boundary()
        return self._boundary

    def _ensure_boundary_data(self):
        if self._boundary_data is not None:
            return self._boundary_data
        boundary = self.boundary
        if len(boundary.faces) == 0:
            raise ValueError("The boundary mesh must contain at least one triangle.")
        vertices = np.asarray(boundary.vertices, dtype=np.float64)
        faces = np.asarray(boundary.faces, dtype=np.uint32)
        areas = np.asarray(boundary.area_faces, dtype=np.float64)
        valid = np.isfinite(areas) & (areas > 0)
        faces = faces[valid]
        areas = areas[valid]
        if len(faces) == 0:
            raise ValueError("The boundary mesh must contain a non-degenerate triangle.")
        vertices4 = np.ones((len(vertices), 4), dtype=np.float32)
        vertices4[:, :3] = vertices
        indices4 = np.zeros((len(faces), 4), dtype=np.uint32)
        indices4[:, :3] = faces
        total = float(np.sum(areas, dtype=np.float64))
        if not np.isfinite(total) or total <= 0:
            raise ValueError("The boundary mesh must have a finite, positive area.")
        cumulative = np.cumsum(areas / total).astype(np.float32)
        cumulative[-1] = 1.0
        triangle_vertices = vertices[faces.astype(np.int64)]
        bvh = build_bvh(triangle_vertices.min(axis=1), triangle_vertices.max(axis=1))
        self._boundary_data = {
            "vertices": np.ascontiguousarray(vertices4),
            "indices": np.ascontiguousarray(indices4),
            "cumulative": np.ascontiguousarray(cumulative),
            "bvh_min": bvh[0],
            "bvh_max": bvh[1],
            "bvh_nodes": bvh[2],
            "bvh_triangles": bvh[3],
        }
        return self._boundary_data

    boundary_vertices = property(lambda self: self._ensure_boundary_data()["vertices"])
    boundary_indices = property(lambda self: self._ensure_boundary_data()["indices"])
    boundary_cumulative_areas = property(lambda self: self._ensure_boundary_data()["cumulative"])
    n_boundary_triangles = property(lambda self: len(self._ensure_boundary_data()["indices"]))
    boundary_bvh_bounds_min = property(lambda self: self._ensure_boundary_data()["bvh_min"])
    boundary_bvh_bounds_max = property(lambda self: self._ensure_boundary_data()["bvh_max"])
    boundary_bvh_nodes = property(lambda self: self._ensure_boundary_data()["bvh_nodes"])
    boundary_bvh_triangles = property(lambda self: self._ensure_boundary_data()["bvh_triangles"])
    n_boundary_bvh_nodes = property(lambda self: len(self._ensure_boundary_data()["bvh_nodes"]))

    @abstractmethod
    def contains(self, points: Points) -> ContainsResult:
        """Authoritative membership of world points (NumPy)."""

    @abstractmethod
    def upload(self, runtime: HipRuntime, with_boundary: bool = False) -> DeviceHandle:
        """Device handle of this domain on ``runtime`` (uploaded once)."""

    def _attach_boundary(self, device: DeviceHandle) -> None:
        if getattr(device, "has_boundary", False):
            return
        data = self._ensure_boundary_data()
        _native.check(
            _native.lib().tgk_domain_set_boundary(
                device.handle, _native.ptr(data["vertices"]), len(data["vertices"]), _native.ptr(data["indices"]),
                _native.ptr(data["cumulative"]), len(data["indices"]),
            )
        )
        device.has_boundary = True


def _voxel_boundary_mesh(mask: np.ndarray, affine: np.ndarray) -> BoundaryMesh:
    """Exposed faces of the union of occupied half-open voxel cells.

    A face is emitted where an occupied voxel borders an empty one or the
    array edge, as two triangles wound so normals point out of the occupied
    region (domain.py:219-264).
    """
    face_definitions = (
        ((-1, 0, 0), ((-1, -1, -1), (-1, -1, 1), (-1, 1, 1), (-1, 1, -1))),
        ((1, 0, 0), ((1, -1, -1), (1, 1, -1), (1, 1, 1), (1, -1, 1))),
        ((0, -1, 0), ((-1, -1, -1), (1, -1, -1), (1, -1, 1), (-1, -1, 1))),
        ((0, 1, 0), ((-1, 1, -1), (-1, 1, 1), (1, 1, 1), (1, 1, -1))),
        ((0, 0, -1), ((-1, -1, -1), (-1, 1, -1), (1, 1, -1), (1, -1, -1))),
        ((0, 0, 1), ((-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1))),
    )
    padded = np.pad(mask, 1, mode="constant", constant_values=False)
    quads = []
    for direction, corner_offsets in face_definitions:
        neighbor = tuple(slice(1 + o, 1 + o + size) for o, size in zip(direction, mask.shape))
        exposed = np.argwhere(mask & ~padded[neighbor])
        if len(exposed) == 0:
            continue
        quads.append(2 * exposed[:, np.newaxis, :] + np.asarray(corner_offsets, dtype=np.int64))
    doubled = np.concatenate(quads, axis=0).reshape(-1, 3)
    corners, inverse = np.unique(doubled, axis=0, return_inverse=True)
    quad_faces = inverse.reshape(-1, 4)
    triangles = np.concatenate((quad_faces[:, (0, 1, 2)], quad_faces[:, (0, 2, 3)]), axis=0)
    world = apply_affine(affine, corners.astype(np.float64) * 0.5)
    if np.linalg.det(affine[:3, :3]) < 0:
        triangles = triangles[:, [0, 2, 1]]
    return BoundaryMesh(world, triangles)


class Image(BaseDomain):
    """Domain of the nonzero voxels of a 3-D image (domain.py:267-430)."""

    def __init__(self, image_data_in):
        if image_data_in.ndim != 3:
            raise ValueError("Only 3D images are supported for image domains.")
        self.values = np.ascontiguousarray(image_data(image_data_in) != 0, dtype=np.int32)
        self.dims = np.ascontiguousarray((*image_data_in.shape, 0), dtype=np.uint32)
        affine = image_data_in.affine
        self.affine = np.asarray(affine if affine is not None else np.eye(4), dtype=np.float64)
        self.inverse_affine = np.ascontiguousarray(np.linalg.inv(self.affine), dtype=np.float32)
        if not np.any(self.values):
            raise ValueError("Image domain must contain at least one nonzero voxel.")
        super().__init__()

    def _make_boundary(self) -> BoundaryMesh:
        return _voxel_boundary_mesh(self.values.astype(bool), self.affine)

    def contains(self, points: Points) -> ContainsResult:
        """NumPy twin (domain.py:400-430): floor(v + 0.5) voxel, mask value."""
        voxel = np.floor(apply_affine(self.inverse_affine, np.asarray(points)) + 0.5).astype(int)
        x, y, z = voxel[:, :3].T
        width, height, depth = self.values.shape
        in_bounds = (0 <= x) & (x < width) & (0 <= y) & (y < height) & (0 <= z) & (z < depth)
        result = np.zeros(x.shape, dtype=np.bool_)
        result[in_bounds] = self.values[tuple(voxel[in_bounds].T)] != 0
        return result

    def upload(self, runtime: HipRuntime, with_boundary: bool = False) -> DeviceHandle:
        key = id(runtime)
        device = self._device.get(key)
        if device is None or device.runtime is not runtime:
            lib = _native.lib()
            handle = ctypes.c_void_p()
            mask = np.ascontiguousarray(self.values != 0, dtype=np.uint8)
            dims = np.ascontiguousarray(self.values.shape, dtype=np.uint32)
            _native.check(
                lib.tgk_mask_domain_create(
                    runtime.handle, _native.ptr(mask), _native.ptr(dims), _native.ptr(self.inverse_affine),
                    ctypes.byref(handle),
                )
            )
            device = DeviceHandle(runtime, handle, lib.tgk_domain_destroy)
            self._device[key] = device
        if with_boundary:
            self._attach_boundary(device)
        return device


class Mesh(BaseDomain):
    """Region bounded by closed triangle surfaces, normals pointing outward.

    Membership (``contains``) is the parity of boundary crossings along +x.
    During tracking the domain ends a streamline when the segment of a step
    hits a boundary triangle (the first test of each iteration, on the
    segment just traversed, as the mask test is on the point just reached);
    the hit triangle's slot is returned per streamline.

    Args:
        vertices, faces: the boundary surfaces as one triangle list.
        cell_size: edge of the cells of the culling grid (world units).
        reach: segments up to this length starting in a cell with no
            triangle bounding box within reach skip the BVH query.
    """

    def __init__(self, vertices, faces, *, cell_size: float = 1.0, reach: float = 0.5):
        super().__init__(BoundaryMesh(vertices, faces))
        self.cell_size = float(cell_size)
        self.reach = float(reach)

    def contains(self, points: Points) -> ContainsResult:
        points = np.asarray(points, dtype=np.float64)
        tri = self.boundary.triangles
        a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
        crossings = np.zeros(len(points), dtype=np.int64)
        for start in range(0, len(points), 256):
            p = points[start : start + 256, None, :]
            # Ray p + s (1, 0, 0), s > 0, against each triangle (barycentric
            # in the y-z projection).
            v0, v1, v2 = a[None] - p, b[None] - p, c[None] - p
            d = (v1[..., 1] - v0[..., 1]) * (v2[..., 2] - v0[..., 2]) - (v1[..., 2] - v0[..., 2]) * (
                v2[..., 1] - v0[..., 1]
            )
            with np.errstate(divide="ignore", invalid="ignore"):
                w1 = ((0 - v0[..., 1]) * (v2[..., 2] - v0[..., 2]) - (0 - v0[..., 2]) * (v2[..., 1] - v0[..., 1])) / d
                w2 = ((v1[..., 1] - v0[..., 1]) * (0 - v0[..., 2]) - (v1[..., 2] - v0[..., 2]) * (0 - v0[..., 1])) / d
                w0 = 1 - w1 - w2
                x = w0 * v0[..., 0] + w1 * v1[..., 0] + w2 * v2[..., 0]
            hit = (d != 0) & (w0 >= 0) & (w1 >= 0) & (w2 >= 0) & (x > 0)
            crossings[start : start + 256] = hit.sum(axis=1)
        return crossings % 2 == 1

    def _grid_cells(self):
        """Geometry of the culling grid and the cells each triangle's bounding
        box reaches: (dims, origin, lo, hi) with per-triangle cell ranges."""
        data = self._ensure_boundary_data()
        tri = self.boundary.vertices[data["indices"][:, :3].astype(np.int64)]
        tmin, tmax = tri.min(axis=1), tri.max(axis=1)
        h = self.cell_size
        margin = self.reach + 0.01 * h
        origin = tmin.min(axis=0) - margin - h
        dims = np.ceil((tmax.max(axis=0) + margin + h - origin) / h).astype(np.int64) + 1
        lo = np.floor((tmin - margin - origin) / h + 0.5).astype(np.int64)
        hi = np.floor((tmax + margin - origin) / h + 0.5).astype(np.int64)
        return dims, origin, np.clip(lo, 0, dims - 1), np.clip(hi, 0, dims - 1)

    def _grid_inverse_affine(self, origin) -> np.ndarray:
        affine = np.eye(4)
        affine[:3, :3] *= self.cell_size
        affine[:3, 3] = origin
        return np.ascontiguousarray(np.linalg.inv(affine), dtype=np.float32)

    def _cell_lists(self):
        """CSR cell -> candidate triangles of the grid (built natively once):
        (dims, origin, lo, hi, cell_start (n_cells + 1,) u32, n_entries)."""
        cached = getattr(self, "_cell_lists_cache", None)
        if cached is not None:
            return cached
        data = self._ensure_boundary_data()
        dims, origin, lo, hi = self._grid_cells()
        dims = np.ascontiguousarray(dims, dtype=np.int64)
        lo = np.ascontiguousarray(lo, dtype=np.int64)
        hi = np.ascontiguousarray(hi, dtype=np.int64)
        cell_start = np.empty(int(np.prod(dims)) + 1, dtype=np.uint32)
        n_entries = ctypes.c_uint64()
        _native.check(
            _native.lib().tgk_segment_grid_count(
                _native.ptr(dims), _native.ptr(lo), _native.ptr(hi), _native.ptr(data["bvh_triangles"]),
                len(data["indices"]), _native.ptr(cell_start), ctypes.byref(n_entries),
            )
        )
        self._cell_lists_cache = (dims, origin, lo, hi, cell_start, n_entries.value)
        return self._cell_lists_cache

    def culling_grid(self):
        """near (X, Y, Z) u8 (1 where a triangle bounding box comes within
        `reach` of the cell), its inverse affine, and the reach it certifies."""
        dims, origin, _, _, cell_start, _ = self._cell_lists()
        near = (np.diff(cell_start) > 0).astype(np.uint8).reshape(tuple(dims))
        return near, self._grid_inverse_affine(origin), np.float32(self.reach)

    def segment_grid(self):
        """Per-cell candidate triangles on the culling grid's cells: every
        triangle whose bounding box comes within `reach` of the cell, in the
        order a BVH traversal reaches them (leaf order), as CSR.

        Returns cell_start (n_cells + 1,) u32 and entries (n, 12) f32 rows
        (a, slot bits, b - a, 0, c - a, 0) in float32 arithmetic. A segment of
        length <= reach starting in a cell can only hit that cell's
        candidates; testing them in order with the traversal's replacement
        rule gives the traversal's closest hit.
        """
        data = self._ensure_boundary_data()
        dims, _, lo, hi, cell_start, n_entries = self._cell_lists()
        entries = np.empty((n_entries, 12), dtype=np.float32)
        _native.check(
            _native.lib().tgk_segment_grid_fill(
                _native.ptr(dims), _native.ptr(lo), _native.ptr(hi), _native.ptr(data["bvh_triangles"]),
                len(data["indices"]), _native.ptr(data["vertices"]), len(data["vertices"]),
                _native.ptr(data["indices"]), _native.ptr(cell_start), _native.ptr(entries), n_entries,
            )
        )
        return cell_start, entries

    def upload(self, runtime: HipRuntime, with_boundary: bool = False) -> DeviceHandle:
        key = id(runtime)
        device = self._device.get(key)
        if device is not None and device.runtime is runtime:
            return device
        data = self._ensure_boundary_data()
        near, near_inverse, reach = self.culling_grid()
        near_dims = np.ascontiguousarray(near.shape, dtype=np.uint32)
        lib = _native.lib()
        handle = ctypes.c_void_p()
        _native.check(
            lib.tgk_mesh_domain_create(
                runtime.handle, _native.ptr(data["vertices"]), len(data["vertices"]), _native.ptr(data["indices"]),
                _native.ptr(data["cumulative"]), len(data["indices"]), _native.ptr(data["bvh_min"]),
                _native.ptr(data["bvh_max"]), _native.ptr(data["bvh_nodes"]), len(data["bvh_nodes"]),
                _native.ptr(data["bvh_triangles"]), len(data["bvh_triangles"]), _native.ptr(np.ascontiguousarray(near)),
                _native.ptr(near_dims), _native.ptr(near_inverse), float(reach), ctypes.byref(handle),
            )
        )
        device = DeviceHandle(runtime, handle, lib.tgk_domain_destroy)
        cell_start, entries = self.segment_grid()
        _native.check(
            lib.tgk_domain_set_segment_grid(handle, _native.ptr(cell_start), len(cell_start) - 1,
                                            _native.ptr(entries), len(entries))
        )
        device.has_boundary = True
        self._device[key] = device
        return device


def contains(domain: BaseDomain, points: Points, runtime: HipRuntime | None = None) -> ContainsResult:
    """Domain membership on the GPU (domain.py:433-491)."""
    points = np.asarray(points)
    n_points = points.shape[0]
    if n_points == 0:
        return np.array([], dtype=np.bool_)
    if not isinstance(domain, Image):
        return domain.contains(points)
    if runtime is None:
        runtime = HipRuntime()
    device = domain.upload(runtime)
    pts = np.ascontiguousarray(points[:, :3], dtype=np.float32)
    out = np.empty(n_points, dtype=np.uint8)
    _native.check(
        _native.lib().tgk_domain_contains(runtime.handle, device.handle, _native.ptr(pts), n_points, _native.ptr(out))
    )
    return out.astype(np.bool_)


def seed(domain: BaseDomain, n_seeds: int, rng_seed: int | None = None, runtime: HipRuntime | None = None):
    """Seeds on the domain boundary, pointing inward (domain.py:494-575).

    Triangles are drawn by area, points uniformly inside them, each nudged
    1e-4 along the inward normal; (n_seeds, 8) float32.
    """
    if not isinstance(n_seeds, (int, np.integer)) or n_seeds < 0:
        raise ValueError("n_seeds must be a non-negative integer.")
    if n_seeds == 0:
        return np.empty((0, 8), dtype=np.float32)
    if runtime is None:
        runtime = HipRuntime()
    if rng_seed is None:
        rng_seed = int(np.random.default_rng().integers(0, 2**32, dtype=np.uint32))
    rng_seed = int(rng_seed) % 2**32
    device = domain.upload(runtime, with_boundary=True)
    out = np.empty((int(n_seeds), 8), dtype=np.float32)
    _native.check(_native.lib().tgk_domain_seed(runtime.handle, device.handle, rng_seed, int(n_seeds), _native.ptr(out)))
    return out


def bvh_segment_hit(domain: BaseDomain, starts, ends, runtime: HipRuntime | None = None):
    """Closest boundary triangle hit per segment on the GPU (bvh.cl:83-131).

    Returns (found (n,) bool, triangle slot (n,) u32, t (n,) f32).
    """
    starts = np.ascontiguousarray(np.asarray(starts)[:, :3], dtype=np.float32)
    ends = np.ascontiguousarray(np.asarray(ends)[:, :3], dtype=np.float32)
    n = len(starts)
    found = np.zeros(n, dtype=np.uint8)
    hit = np.zeros(n, dtype=np.uint32)
    t = np.zeros(n, dtype=np.float32)
    if n == 0:
        return found.astype(bool), hit, t
    if runtime is None:
        runtime = HipRuntime()
    if isinstance(domain, Mesh):
        device = domain.upload(runtime)
    else:
        device = _mesh_handle_for(domain, runtime)
    _native.check(
        _native.lib().tgk_bvh_segment_hit(
            runtime.handle, device.handle, _native.ptr(starts), _native.ptr(ends), n, _native.ptr(found),
            _native.ptr(hit), _native.ptr(t),
        )
    )
    return found.astype(bool), hit, t


def _mesh_handle_for(domain: BaseDomain, runtime: HipRuntime) -> DeviceHandle:
    """A mesh handle over any domain's boundary (for BVH queries)."""
    cache = domain.__dict__.setdefault("_bvh_device", {})
    device = cache.get(id(runtime))
    if device is not None and device.runtime is runtime:
        return device
    data = domain._ensure_boundary_data()
    lib = _native.lib()
    handle = ctypes.c_void_p()
    _native.check(
        lib.tgk_mesh_domain_create(
            runtime.handle, _native.ptr(data["vertices"]), len(data["vertices"]), _native.ptr(data["indices"]),
            _native.ptr(data["cumulative"]), len(data["indices"]), _native.ptr(data["bvh_min"]),
            _native.ptr(data["bvh_max"]), _native.ptr(data["bvh_nodes"]), len(data["bvh_nodes"]),
            _native.ptr(data["bvh_triangles"]), len(data["bvh_triangles"]), None, None, None, 0.0,
            ctypes.byref(handle),
        )
    )
    device = DeviceHandle(runtime, handle, lib.tgk_domain_destroy)
    cache[id(runtime)] = device
    return device
