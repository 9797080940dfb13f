"""Field lookup, mask membership, boundary mesh, BVH and boundary/FOD seeding
of the restatement against the reference's known answers, brute force and
distributional checks."""

import numpy as np
import pytest

import fixtures
import oracle
import tractography_amd as tg


def test_field_lookup_known_answers():
    # test/test_field.py:8-46
    data = np.arange(2 * 3 * 4 * 2, dtype=np.float32).reshape(2, 3, 4, 2)
    field = tg.field.Image(tg.Nifti1Image(data, np.eye(4)))
    points = np.array([[0, 0, 0], [1, 2, 3], [1, 0, 2]], np.float32)
    expected = data[points[:, 0].astype(int), points[:, 1].astype(int), points[:, 2].astype(int)]
    np.testing.assert_array_equal(oracle.field_evaluate(field, points), expected)
    np.testing.assert_array_equal(field.evaluate(points), expected)
    data = np.arange(64, dtype=np.float32).reshape(4, 4, 4, 1)
    field = tg.field.Image(tg.Nifti1Image(data, np.eye(4)))
    points = np.array([[0.5, 1.5, 2.5], [1.5, 0.5, 0.5], [-0.5, 2.5, 2.5]], np.float32)
    expected = data[[0, 2, 0], [2, 0, 2], [2, 0, 2]]
    np.testing.assert_array_equal(oracle.field_evaluate(field, points), expected)
    np.testing.assert_array_equal(field.evaluate(points), expected)
    field = tg.field.Image(tg.Nifti1Image(np.ones((2, 3, 4, 2), np.float32), np.eye(4)))
    points = np.array([[-0.51, 0, 0], [0, 3, 0], [0, 0, 4]], np.float32)
    np.testing.assert_array_equal(oracle.field_evaluate(field, points), np.zeros((3, 2)))


def test_mask_membership_known_answers():
    # test/test_domain.py:12-65, 133-144
    image = np.zeros((3, 3, 3))
    image[1, 1, 1] = 1
    domain = tg.domain.Image(tg.Nifti1Image(image, np.eye(4)))
    rng = np.random.default_rng(0)
    inside = rng.random((50, 3)).astype(np.float32) + 0.5
    assert np.all(oracle.domain_contains(domain, inside))
    assert not oracle.domain_contains(domain, np.zeros((1, 3)))[0]
    affine = np.eye(4)
    affine[:3, 3] = (10, 20, 30)
    shifted = tg.domain.Image(tg.Nifti1Image(image, affine))
    pts = np.array([[11, 21, 31], [10, 20, 30]], np.float32)
    np.testing.assert_array_equal(oracle.domain_contains(shifted, pts), [True, False])
    np.testing.assert_array_equal(shifted.contains(pts), [True, False])
    cube = tg.domain.Image(tg.Nifti1Image(np.ones((1, 1, 1), np.uint8), np.eye(4)))
    pts = np.array([[-0.5, 0, 0], [0.49, 0.49, 0.49], [0.5, 0, 0]], np.float32)
    np.testing.assert_array_equal(oracle.domain_contains(cube, pts), [True, True, False])
    random_points = rng.random((200, 3)) * 5 - 1
    np.testing.assert_array_equal(oracle.domain_contains(domain, random_points), domain.contains(random_points))


def test_voxel_boundary_mesh_topology():
    # test/test_domain.py:68-131
    affine = np.eye(4)
    affine[:3, :3] = np.diag([2.0, 3.0, 4.0])
    affine[:3, 3] = [10.0, 20.0, 30.0]
    domain = tg.domain.Image(tg.Nifti1Image(np.ones((1, 1, 1), np.uint8), affine))
    b = domain.boundary
    assert b.is_watertight and len(b.vertices) == 8 and domain.n_boundary_triangles == 12
    assert domain.boundary_cumulative_areas[-1] == 1.0
    np.testing.assert_allclose(b.bounds, [[9.0, 18.5, 28.0], [11.0, 21.5, 32.0]])
    assert np.all(np.einsum("ij,ij->i", b.triangles_center - [10, 20, 30], b.face_normals) > 0)
    affine[0, 0] = -2.0
    reflected = tg.domain.Image(tg.Nifti1Image(np.ones((1, 1, 1), np.uint8), affine))
    rb = reflected.boundary
    assert np.all(np.einsum("ij,ij->i", rb.triangles_center - affine[:3, 3], rb.face_normals) > 0)
    image = np.ones((3, 3, 3), np.uint8)
    image[1, 1, 1] = 0
    cavity = tg.domain.Image(tg.Nifti1Image(image, np.eye(4))).boundary
    centers = cavity.triangles_center
    inside = np.all((centers >= 0.5) & (centers <= 1.5), axis=1)
    assert inside.sum() == 12
    assert np.all(np.einsum("ij,ij->i", cavity.face_normals[inside], 1.0 - centers[inside]) > 0)
    pair = tg.domain.Image(tg.Nifti1Image(np.ones((2, 1, 1), np.uint8), np.eye(4)))
    assert pair.boundary.is_watertight and len(pair.boundary.vertices) == 12
    assert pair.n_boundary_triangles == 20 and pair.boundary.area == pytest.approx(10.0)
    with pytest.raises(ValueError, match="at least one nonzero voxel"):
        tg.domain.Image(tg.Nifti1Image(np.zeros((2, 2, 2)), np.eye(4)))


def _bvh_invariants(domain):
    nodes = domain.boundary_bvh_nodes
    order = domain.boundary_bvh_triangles
    assert sorted(order.tolist()) == list(range(domain.n_boundary_triangles))
    lo, hi = domain.boundary_bvh_bounds_min, domain.boundary_bvh_bounds_max
    verts = domain.boundary_vertices[:, :3]
    tris = verts[domain.boundary_indices[:, :3].astype(int)]
    for i, (left, right, first, count) in enumerate(nodes):
        if count > 0:
            assert left == -1 and right == -1 and count <= 4
            t = tris[order[first : first + count]]
            assert np.all(t.min(axis=(0, 1)) >= lo[i, :3]) and np.all(t.max(axis=(0, 1)) <= hi[i, :3])
        else:
            assert left == i + 1  # preorder: left child follows its parent
            for child in (left, right):
                assert np.all(lo[child] >= lo[i]) and np.all(hi[child] <= hi[i])


def _brute_force_hit(domain, starts, ends):
    verts = domain.boundary_vertices[:, :3].astype(np.float64)
    tris = verts[domain.boundary_indices[:, :3].astype(int)]
    found, best = [], []
    for s, e in zip(starts.astype(np.float64), ends.astype(np.float64)):
        d = e - s
        a, b, c = tris[:, 0], tris[:, 1], tris[:, 2]
        e1, e2 = b - a, c - a
        p = np.cross(d, e2)
        det = np.einsum("ij,ij->i", e1, p)
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = 1 / det
            off = s - a
            u = np.einsum("ij,ij->i", off, p) * inv
            q = np.cross(off, e1)
            v = (q @ d) * inv
            t = np.einsum("ij,ij->i", e2, q) * inv
        ok = (np.abs(det) > 1e-10) & (u >= 0) & (v >= 0) & (u + v <= 1) & (t > 1e-6) & (t <= 1)
        found.append(bool(ok.any()))
        best.append(np.where(ok, t, np.inf).min())
    return np.array(found), np.array(best)


def test_bvh_closest_hit_matches_brute_force():
    # Segments against a voxel-boundary surface with 2k+ triangles: the BVH
    # traversal finds the same closest hit distance as testing every triangle.
    mask = np.zeros((12, 12, 12), bool)
    idx = np.indices(mask.shape).reshape(3, -1).T
    mask.reshape(-1)[np.linalg.norm(idx - 5.5, axis=1) < 5] = True
    domain = tg.domain.Image(tg.Nifti1Image(mask.astype(np.uint8), np.eye(4)))
    _bvh_invariants(domain)
    rng = np.random.default_rng(5)
    starts = rng.uniform(-1, 12, (3000, 3)).astype(np.float32)
    ends = (starts + rng.standard_normal((3000, 3)) * 2).astype(np.float32)
    found, hit, t = oracle.bvh_hit(domain, starts, ends)
    bf_found, bf_t = _brute_force_hit(domain, starts, ends)
    # Edge grazes within the 1e-6 barycentric tolerance may differ.
    assert np.mean(found == bf_found) > 0.999
    both = found & bf_found
    np.testing.assert_allclose(t[both], bf_t[both], rtol=1e-5, atol=1e-6)
    assert found.sum() > 500


def test_bvh_reference_probes():
    # test/test_domain.py:183-260
    domain = tg.domain.Image(tg.Nifti1Image(np.ones((1, 1, 1), np.uint8), np.eye(4)))
    found, _, _ = oracle.bvh_hit(domain, np.zeros((1, 3)), np.array([[-1.0, 0, 0]]))
    assert found[0]
    seeds = oracle.domain_seed(domain, 1234, 32)
    pts = seeds[:, :3]
    ends = pts + 2.0 * seeds[:, 4:7]
    found, _, _ = oracle.bvh_hit(domain, pts, ends)
    assert np.all(found)


def test_boundary_seeds_are_inside_and_point_inward():
    # test/test_domain.py:153-170
    domain = tg.domain.Image(tg.Nifti1Image(np.ones((5, 5, 5), np.uint8), np.eye(4)))
    seeds = oracle.domain_seed(domain, 123, 64)
    assert seeds.shape == (64, 8)
    np.testing.assert_array_equal(seeds[:, 3], 1.0)
    np.testing.assert_array_equal(seeds[:, 7], 0.0)
    assert np.all(domain.contains(seeds[:, :3]))
    np.testing.assert_allclose(np.linalg.norm(seeds[:, 4:7], axis=1), 1.0, atol=1e-6)
    assert np.all(np.einsum("ij,ij->i", seeds[:, 4:7], 2.0 - seeds[:, :3]) > 0)
    # Triangles are drawn in proportion to area: each of the 6 faces gets 1/6.
    many = oracle.domain_seed(domain, 7, 60000)
    faces = np.argmax(np.abs(many[:, 4:7]) * 10 + (many[:, 4:7] > 0), axis=1)
    counts = np.bincount(faces, minlength=3) / len(many)
    np.testing.assert_allclose(counts, 1 / 3, atol=0.01)


def _second_moment(orientations, weights=None):
    if weights is None:
        weights = np.full(len(orientations), 1.0 / len(orientations))
    return np.einsum("n,ni,nj->ij", weights, orientations, orientations)


def test_fod_seeds_follow_the_sh_density(oracle_threads):
    # test/test_seeds.py:163-187
    directions = tg.core.fibonacci_sphere(1000)
    sh_fod, _, _ = fixtures.cross()
    coefficients = sh_fod.get_fdata()[5, 1, 0]
    fod = np.zeros((1, 1, 1, 45), np.float32)
    fod[0, 0, 0] = coefficients
    inline, voxels, affine = tg.seeds.sparse_fod(tg.Nifti1Image(fod, np.eye(4)))
    states = np.random.default_rng(1).integers(0, 2**32 - 1, (20000, 2), dtype=np.uint32)
    seeds, _ = oracle.seeds_from_fod(0, 45, inline, voxels, affine, states, threads=oracle_threads)
    a, c, _ = tg.core.cart2sph(*directions.T)
    values = tg.core.ishtmtx(a, c, 45)[0] @ coefficients
    w = np.where(100 * values < 30, np.log1p(np.exp(np.clip(100 * values, -700, 30))) / 100, values)
    w = np.maximum(w, 0) / np.maximum(w, 0).sum()
    np.testing.assert_allclose(_second_moment(seeds[:, 4:7]), _second_moment(directions, w), atol=2e-2)
    np.testing.assert_allclose(np.linalg.norm(seeds[:, 4:7], axis=1), 1, atol=1e-5)
    assert np.all(np.abs(seeds[:, :3]) <= 0.5)


def test_fod_seeds_follow_the_dti_density(oracle_threads):
    # test/test_seeds.py:189-213
    directions = tg.core.fibonacci_sphere(1000)
    dti, _, _ = fixtures.tensor_cross()
    coefficients = dti.get_fdata()[10, 1, 0]
    fod = np.zeros((1, 1, 1, 6), np.float32)
    fod[0, 0, 0] = coefficients
    inline, voxels, affine = tg.seeds.sparse_fod(tg.Nifti1Image(fod, dti.affine))
    states = np.random.default_rng(2).integers(0, 2**32 - 1, (20000, 2), dtype=np.uint32)
    seeds, _ = oracle.seeds_from_fod(1, 6, inline, voxels, affine, states, threads=oracle_threads)
    t = np.array([[coefficients[0], coefficients[3], coefficients[4]],
                  [coefficients[3], coefficients[1], coefficients[5]],
                  [coefficients[4], coefficients[5], coefficients[2]]])
    p = np.maximum(np.einsum("ni,ij,nj->n", directions, t, directions), 0)
    np.testing.assert_allclose(_second_moment(seeds[:, 4:7]), _second_moment(directions, p / p.sum()), atol=2e-2)


def test_fod_seeds_stay_in_their_voxels(oracle_threads):
    # test/test_seeds.py:111-127
    fod, mask, _ = fixtures.cross()
    masked = tg.Nifti1Image(fod.get_fdata() * mask.get_fdata()[..., None], fod.affine)
    inline, voxels, affine = tg.seeds.sparse_fod(masked)
    states = np.random.default_rng(3).integers(0, 2**32 - 1, (3000, 2), dtype=np.uint32)
    seeds, _ = oracle.seeds_from_fod(0, 45, inline, voxels, affine, states, threads=oracle_threads)
    assert np.all(seeds[:, :3] >= -0.5) and np.all(seeds[:, :3] <= 9.5)
    np.testing.assert_allclose(np.linalg.norm(seeds[:, 4:7], axis=1), 1, atol=1e-5)
    assert np.all(mask.get_fdata()[tuple(np.round(seeds[:, :3]).astype(int).T)] == 1)


def _stream_uniforms(rng_seed, n_seeds, per_seed):
    """The first `per_seed` randu draws of each boundary seed's stream
    rng_init(rng_seed, g) (core.cl:196-214), as float64."""
    out = np.zeros((n_seeds, per_seed))
    for g in range(n_seeds):
        draws, _ = oracle.rng_draws(oracle.rng_init(rng_seed, g), 1, per_seed)
        out[g] = draws
    return out


def test_boundary_seeds_follow_the_noise_stream_in_float64():
    # domain.py:131-166 restated in float64 from each seed's first three
    # draws: the triangle whose cumulative area first exceeds the first
    # draw, the point (1 - sqrt(u), sqrt(u)(1 - v), sqrt(u) v) on (a, b, c),
    # the inward normal, the 1e-4 nudge. Swapped barycentric roles (sqrt of
    # v instead of u, weights on the wrong vertices) or a wrong normal fail.
    mask = np.zeros((6, 5, 4), np.uint8)
    mask[1:5, 1:4, 1:3] = 1
    mask[2, 2, 3] = 1
    affine = np.array([[0.9, 0.2, 0.0, 3.0], [-0.1, 1.1, 0.3, -2.0], [0.05, 0.0, 1.3, 0.5], [0, 0, 0, 1]])
    domain = tg.domain.Image(tg.Nifti1Image(mask, affine))
    data = domain._ensure_boundary_data()
    vertices = data["vertices"][:, :3].astype(np.float64)
    triangles = data["indices"][:, :3].astype(np.int64)
    cumulative = data["cumulative"].astype(np.float64)
    n = 3000
    seeds = oracle.domain_seed(domain, 99, n)
    u = _stream_uniforms(99, n, 3)
    tri = np.minimum(np.searchsorted(cumulative, u[:, 0], side="right"), len(triangles) - 1)
    a, b, c = (vertices[triangles[tri, k]] for k in range(3))
    root = np.sqrt(u[:, 1])
    sampled = (1 - root)[:, None] * a + (root * (1 - u[:, 2]))[:, None] * b + (root * u[:, 2])[:, None] * c
    outward = np.cross(b - a, c - a)
    inward = -outward / np.linalg.norm(outward, axis=1, keepdims=True)
    np.testing.assert_allclose(seeds[:, 4:7], inward, atol=2e-6)
    np.testing.assert_allclose(seeds[:, :3], sampled + 1e-4 * inward, atol=2e-5)


def test_boundary_points_are_uniform_on_each_face():
    # Points of a one-voxel cube's boundary (two triangles per face) must be
    # uniform on each face square: both in-face coordinates uniform on
    # [-0.5, 0.5] (Kolmogorov-Smirnov at 1e-3). sqrt taken of the wrong
    # draw concentrates points near a vertex and fails.
    from scipy import stats

    domain = tg.domain.Image(tg.Nifti1Image(np.ones((1, 1, 1), np.uint8), np.eye(4)))
    seeds = oracle.domain_seed(domain, 5, 60000)
    normal_axis = np.argmax(np.abs(seeds[:, 4:7]), axis=1)
    for axis in range(3):
        on_face = seeds[normal_axis == axis, :3]
        for k in range(3):
            if k == axis:
                continue
            p = stats.kstest(on_face[:, k], stats.uniform(loc=-0.5, scale=1.0).cdf).pvalue
            assert p > 1e-3, (axis, k, p)


def test_fod_seed_positions_under_a_general_affine(oracle_threads):
    # seeds.cl:68-93: voxel randi(nnz), three jitters randu - 0.5, then the
    # forward affine, restated in float64 from the states' draws. The affine
    # rotates, shears and translates, so a transposed matrix, a dropped
    # translation or jitters on the wrong axes fail.
    fod, _, _ = fixtures.cross(shape=(7, 6, 3), n_coefficients=15)
    affine = np.array([[0.0, 1.2, 0.1, -4.0], [0.9, 0.0, -0.2, 7.5], [0.1, 0.3, 1.1, 2.25], [0, 0, 0, 1]])
    image = tg.Nifti1Image(fod.get_fdata().astype(np.float32), affine)
    inline, voxels, affine32 = tg.seeds.sparse_fod(image)
    states = np.random.default_rng(11).integers(0, 2**32 - 1, (4000, 2), dtype=np.uint32)
    seeds, _ = oracle.seeds_from_fod(0, 15, inline, voxels, affine32, states.copy(), threads=oracle_threads)
    nnz = len(inline)
    expected = np.zeros((len(states), 3))
    for g, state in enumerate(states):
        draws, _ = oracle.rng_draws(state, 1, 4)
        index = int(np.float32(draws[0]) * np.float32(nnz))
        voxel = voxels[index, :3].astype(np.float64) + (draws[1:4].astype(np.float64) - 0.5)
        expected[g] = affine[:3, :3] @ voxel + affine[:3, 3]
    np.testing.assert_allclose(seeds[:, :3], expected, atol=5e-6 * np.abs(expected).max())
    np.testing.assert_array_equal(seeds[:, 3], 1.0)


def test_exact_t_ties_are_rare_on_tracking_segments():
    # The BVH orders triangles with equal centroid coordinates by a stable
    # sort, the reference by numpy's default argsort (domain.py:99), whose
    # order among equal keys is not specified. The two can only return
    # different triangles when two triangles are hit at the same t, i.e. a
    # segment crosses the surface within rounding of a shared edge. On
    # tracking-like segments (0.1 voxel, random positions and directions)
    # across a voxel-boundary surface, count the hits whose second-closest
    # triangle is within a relative 1e-6 of the closest (a superset of
    # float32 exact ties).
    mask = np.zeros((16, 16, 16), bool)
    idx = np.indices(mask.shape).reshape(3, -1).T
    mask.reshape(-1)[np.linalg.norm(idx - 7.5, axis=1) < 6.5] = True
    domain = tg.domain.Image(tg.Nifti1Image(mask.astype(np.uint8), np.eye(4)))
    verts = domain.boundary_vertices[:, :3].astype(np.float64)
    tris = verts[domain.boundary_indices[:, :3].astype(int)]
    rng = np.random.default_rng(8)
    n = 400_000
    starts = rng.uniform(0.5, 14.5, (n, 3))
    d = rng.standard_normal((n, 3))
    d *= 0.1 / np.linalg.norm(d, axis=1, keepdims=True)
    found, _, _ = oracle.bvh_hit(domain, starts.astype(np.float32), (starts + d).astype(np.float32))
    hits = np.flatnonzero(found)
    assert len(hits) > 5000
    a, e1, e2 = tris[:, 0], tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0]
    near_ties = 0
    for i in hits:
        s, dd = starts[i], d[i]
        p = np.cross(dd, e2)
        det = np.einsum("ij,ij->i", e1, p)
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = 1 / det
            off = s - a
            u = np.einsum("ij,ij->i", off, p) * inv
            q = np.cross(off, e1)
            v = (q @ dd) * inv
            t = np.einsum("ij,ij->i", e2, q) * inv
        ok = (np.abs(det) > 1e-10) & (u >= -1e-6) & (v >= -1e-6) & (u + v <= 1 + 1e-6) & (t > 1e-6) & (t <= 1)
        ts = np.sort(t[ok])
        if len(ts) >= 2 and ts[1] - ts[0] <= 1e-6 * ts[0]:
            near_ties += 1
    print(f"\n[ties] {near_ties} near-ties in {len(hits)} boundary hits of {n} segments, {len(tris)} triangles")
    assert near_ties <= 2
