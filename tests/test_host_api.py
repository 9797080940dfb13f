"""The Python surface's validation and early-exit behaviour (no GPU needed):
the reference's error types and messages, and empty inputs returning without
creating a runtime."""

from unittest.mock import patch

import numpy as np
import pytest

import fixtures
import tractography_amd as tg


def test_validate_seed_array_returns_contiguous_float32():
    # test/test_seeds.py:13-22
    seeds = np.zeros((3, 16), dtype=np.float64)[:, ::2]
    seeds[:, 3] = 1.0
    validated = tg.seeds.validate_seed_array(seeds)
    assert validated.shape == (3, 8) and validated.dtype == np.float32 and validated.flags.c_contiguous


def test_from_shape():
    # test/test_local.py:11-24
    assert isinstance(tg.local.from_shape((10, 10, 1, 6)), tg.local.DiffusionTensorImaging)
    sh = tg.local.from_shape((10, 10, 1, 45))
    assert isinstance(sh, tg.local.RealSymmetricSphericalHarmonics) and sh.n_parameters == 45
    assert sh.opencl_model == "MODEL_SYMMETRIC_REAL_SPHERICAL_HARMONICS"
    with pytest.raises(ValueError, match="only DTI and fODF"):
        tg.local.from_shape((10, 10, 1, 12))
    with pytest.raises(ValueError, match="coefficient dimension"):
        tg.local.from_shape(())


def _tiny():
    domain = tg.domain.Image(tg.Nifti1Image(np.ones((1, 1, 1)), np.eye(4)))
    field = tg.field.Image(tg.Nifti1Image(np.ones((1, 1, 1, 1)), np.eye(4)))
    return domain, field


def test_low_level_tracker_validation_and_empty_input():
    # test/test_tractography.py:52-65, test/test_tracking.py:10-25, 56-68
    domain, field = _tiny()
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(1))
    with patch("tractography_amd.tracking.HipRuntime") as factory:
        with pytest.raises(ValueError, match=r"shape \(n_seeds, 8\)"):
            tg.tracking.tractogram(algorithm, domain, field, np.empty((1, 6), np.float32))
        with pytest.raises(ValueError, match="homogeneous values"):
            tg.tracking.tractogram(algorithm, domain, field, np.zeros((1, 8), np.float32))
        result = tg.tracking.tractogram(algorithm, domain, field, np.empty((0, 8), np.float32))
        assert len(result.streamlines) == 0
        np.testing.assert_array_equal(result.affine_to_rasmm, np.eye(4))
        wrong = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(15))
        with pytest.raises(ValueError, match="15 != 1"):
            tg.tracking.tractogram(wrong, domain, field, np.empty((0, 8), np.float32))
        factory.assert_not_called()


def test_public_tractogram_validation():
    # __init__.py:433-500
    fod, _, _ = fixtures.cross()
    with patch("tractography_amd.HipRuntime") as factory:
        with pytest.raises(ValueError, match="save_at"):
            tg.tractogram(fod, 1, save_at=0.0)
        with pytest.raises(ValueError, match="maximum_streamline_length"):
            tg.tractogram(fod, 1, maximum_streamline_length=-1.0)
        with pytest.raises(ValueError, match="must not exceed"):
            tg.tractogram(fod, 1, minimum_streamline_length=5.0, maximum_streamline_length=2.0)
        with pytest.raises(ValueError, match="batch_size"):
            tg.tractogram(fod, 1, batch_size=0)
        with pytest.raises(TypeError, match="field must be"):
            tg.tractogram(np.zeros((2, 2, 2, 45)), 1)
        with pytest.raises(TypeError, match="seeds must be"):
            tg.tractogram(fod, [1, 2])
        with pytest.raises(TypeError, match="algorithm must be"):
            tg.tractogram(fod, 1, algorithm="transport")
        with pytest.raises(ValueError, match="6 != 45"):
            tg.tractogram(fod, 1, algorithm=tg.algorithm.Transport(tg.local.DiffusionTensorImaging()))
        with pytest.raises(TypeError, match="domain must be"):
            tg.tractogram(fod, 1, domain="mask")
        assert len(tg.tractogram(fod, 0)) == 0
        assert len(tg.tractogram(fod, np.empty((0, 8), np.float32))) == 0
        factory.assert_not_called()


def test_empty_one_op_drivers_return_without_runtime():
    with patch("tractography_amd.runtime.HipRuntime.__init__", side_effect=AssertionError("runtime created")):
        model = tg.local.RealSymmetricSphericalHarmonics(15)
        assert tg.local.evaluate(model, np.empty((0, 15)), np.empty((0, 3))).shape == (0, 3)
        algorithm = tg.algorithm.Transport(model)
        out = tg.algorithm.update_orientation(algorithm, np.empty((0, 15), np.float32), np.empty((0, 3), np.float32))
        assert out.shape == (0, 3) and out.dtype == np.float32
        field = tg.field.Image(tg.Nifti1Image(np.zeros((2, 3, 4, 2)), np.eye(4)))
        result = tg.field.evaluate(field, np.empty((0, 3)))
        assert result.shape == (0, 2) and result.dtype == np.float32
        domain, _ = _tiny()
        assert tg.domain.seed(domain, 0).shape == (0, 8)
        with pytest.raises(ValueError, match="non-negative integer"):
            tg.domain.seed(domain, -1)
        fod, _, _ = fixtures.cross()
        assert tg.seeds.from_fod(fod, 0).shape == (0, 8)


def test_point_capacity_matches_the_reference_buffer_rule():
    assert tg.tracking.point_capacity(300.0, 0.25) == 1200
    assert tg.tracking.point_capacity(1.0, 0.001) == 1000
    with pytest.raises(ValueError, match="nbytes must be positive"):
        tg.tracking.point_capacity(0.1, 0.25)


def test_numpy_twins_agree_with_reference_algebra():
    # Transport twin against the analytic quadratic step (test_algorithm.py:54-74).
    params = fixtures.quadratic_fod_coefficients(0.1, 0.3)[None, :]
    o = np.array([tg.core.sph2cart(0.5, 0.8, 1.0)]).reshape(1, 3)
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(15), step_size=0.04, gamma=1.7)
    np.testing.assert_allclose(
        algorithm.update_orientation(params, o)[0],
        fixtures.expected_quadratic_update(o[0], 0.1, 0.3, 0.04, 1.7),
        rtol=2e-4,
        atol=2e-4,
    )


def test_streamlines_container():
    s = tg.Streamlines(np.arange(18, dtype=np.float32).reshape(6, 3), np.array([1, 3, 2]))
    assert len(s) == 3
    np.testing.assert_array_equal(s[1], np.arange(3, 12).reshape(3, 3))
    assert [len(x) for x in s] == [1, 3, 2]
    assert len(tg.Tractogram([np.zeros((2, 3))])) == 1


def test_connectome_zero_seeds_returns_empty_csr_without_a_device():
    # test/test_tracking.py:212-224
    from scipy.sparse import csr_matrix

    domain = tg.domain.Image(tg.Nifti1Image(np.ones((1, 1, 1), dtype=np.uint8), np.eye(4)))
    field = tg.field.Image(tg.Nifti1Image(np.zeros((1, 1, 1, 15)), np.eye(4)))
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(15))
    matrix = tg.tracking.connectome(algorithm, domain, field, n_seeds=0)
    assert isinstance(matrix, csr_matrix)
    assert matrix.shape == (len(domain.boundary.vertices),) * 2
    assert matrix.nnz == 0


def test_connectome_argument_errors():
    domain = tg.domain.Image(tg.Nifti1Image(np.ones((1, 1, 1), dtype=np.uint8), np.eye(4)))
    field = tg.field.Image(tg.Nifti1Image(np.zeros((1, 1, 1, 15)), np.eye(4)))
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(15))
    with pytest.raises(ValueError, match="n_seeds"):
        tg.tracking.connectome(algorithm, domain, field, n_seeds=-1)
    with pytest.raises(ValueError, match="must not exceed"):
        tg.tracking.connectome(algorithm, domain, field, n_seeds=1, maximum_streamline_length=1.0,
                               minimum_streamline_length=2.0)
    wrong = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(45))
    with pytest.raises(ValueError, match="15 != 45|45 != 15"):
        tg.tracking.connectome(wrong, domain, field, n_seeds=1)
    with pytest.raises(TypeError):
        tg.connectome(np.zeros((2, 2, 2, 15)), n_seeds=1)


def test_tod_zero_seeds_returns_empty_image_without_a_device():
    # test/test_tracking.py:164-180
    description = tg.Nifti1Image(np.zeros((2, 1, 1), dtype=np.float32), np.eye(4))
    domain = tg.domain.Image(tg.Nifti1Image(np.ones((1, 1, 1), dtype=np.uint8), np.eye(4)))
    field = tg.field.Image(tg.Nifti1Image(np.ones((1, 1, 1, 1), dtype=np.float32), np.eye(4)))
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(1))
    result = tg.tracking.tod(algorithm, domain, field, description, n_seeds=0)
    assert result.shape == (2, 1, 1, 45)
    assert result.get_data_dtype() == np.float32
    np.testing.assert_array_equal(result.get_fdata(), 0.0)
    # Zero steps (maximum length below one step) also return without a device.
    result = tg.tracking.tod(algorithm, domain, field, description, n_seeds=5, maximum_streamline_length=0.0)
    np.testing.assert_array_equal(result.get_fdata(), 0.0)


def test_tod_argument_errors():
    description = tg.Nifti1Image(np.zeros((2, 1, 1), dtype=np.float32), np.eye(4))
    domain = tg.domain.Image(tg.Nifti1Image(np.ones((1, 1, 1), dtype=np.uint8), np.eye(4)))
    field = tg.field.Image(tg.Nifti1Image(np.ones((1, 1, 1, 15), dtype=np.float32), np.eye(4)))
    algorithm = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(15))
    with pytest.raises(ValueError, match="n_seeds"):
        tg.tracking.tod(algorithm, domain, field, description, n_seeds=-1)
    with pytest.raises(ValueError, match="n_seeds"):
        tg.tracking.tod(algorithm, domain, field, description, n_seeds=1.5)
    with pytest.raises(ValueError, match="maximum_streamline_length"):
        tg.tracking.tod(algorithm, domain, field, description, n_seeds=1, maximum_streamline_length=np.inf)
    with pytest.raises(ValueError, match="3D"):
        tg.tracking.tod(algorithm, domain, field, tg.Nifti1Image(np.zeros((2, 2)), np.eye(4)), n_seeds=1)
    wrong = tg.algorithm.Transport(tg.local.RealSymmetricSphericalHarmonics(45))
    with pytest.raises(ValueError, match="15 != 45|45 != 15"):
        tg.tracking.tod(wrong, domain, field, description, n_seeds=1)
    # The public entry point (__init__.py:257-336).
    with pytest.raises(ValueError, match="n_seeds"):
        tg.tod(np.zeros((2, 2, 2, 15)), n_seeds=-1)
    with pytest.raises(TypeError):
        tg.tod(np.zeros((2, 2, 2, 15)), n_seeds=1)
    fod = tg.Nifti1Image(np.ones((2, 2, 2, 45), dtype=np.float32), np.eye(4))
    with pytest.raises(ValueError, match="four-dimensional"):
        tg.tod(tg.Nifti1Image(np.ones((2, 2, 45)), np.eye(4)), n_seeds=1)
    with pytest.raises(ValueError, match="only DTI"):
        tg.tod(tg.Nifti1Image(np.ones((2, 2, 2, 15)), np.eye(4)), n_seeds=1)
    with pytest.raises(ValueError, match="three-dimensional"):
        tg.tod(fod, n_seeds=1, domain=tg.Nifti1Image(np.ones((2, 2)), np.eye(4)))
    with pytest.raises(TypeError, match="domain"):
        tg.tod(fod, n_seeds=1, domain="mask")
    # An all-zero domain image gives the zero TOD on the field's grid.
    empty = tg.tod(fod, n_seeds=3, domain=tg.Nifti1Image(np.zeros((2, 2, 2), dtype=np.uint8), np.eye(4)))
    assert empty.shape == (2, 2, 2, 45)
    np.testing.assert_array_equal(empty.get_fdata(), 0.0)


def test_opencl_module_names_the_hip_runtime():
    # docs/user-guide/tracking.md:12, 36 build `tg.opencl.OpenCLRuntime()`
    import tractography_amd as tg

    assert tg.opencl.OpenCLRuntime is tg.HipRuntime
    assert tg.OpenCLRuntime is tg.HipRuntime
