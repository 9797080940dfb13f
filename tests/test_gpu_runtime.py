"""Runtime lifetime: closing a runtime frees the fields and domains made on
it first, and later use of either raises instead of touching freed memory."""

import numpy as np
import pytest

import fixtures
import tractography_amd as tg

pytestmark = pytest.mark.gpu


def test_close_releases_children_and_refuses_later_use():
    rng = np.random.default_rng(0)
    field = tg.field.Image(tg.Nifti1Image(rng.standard_normal((4, 5, 6, 15)).astype(np.float32), np.eye(4)))
    domain = tg.domain.Image(tg.Nifti1Image(np.ones((4, 5, 6), dtype=np.int32), np.eye(4)))
    points = np.array([[1.0, 2.0, 3.0], [0.2, 4.4, 5.0]])
    rt = tg.HipRuntime()
    before = tg.field.evaluate(field, points, rt)
    assert tg.domain.contains(domain, points, rt).all()
    handles = [field.upload(rt), domain.upload(rt)]
    rt.close()
    for h in handles:
        with pytest.raises(RuntimeError, match="released"):
            h.handle
    with pytest.raises(RuntimeError, match="closed"):
        tg.field.evaluate(field, points, rt)
    fresh = tg.HipRuntime()
    np.testing.assert_array_equal(tg.field.evaluate(field, points, fresh), before)
    assert tg.domain.contains(domain, points, fresh).all()


def test_calls_refuse_handles_of_another_runtime():
    # A field or domain is uploaded on its own runtime's stream; a call on
    # another runtime is refused rather than racing that upload.
    import ctypes

    from tractography_amd import _native

    rng = np.random.default_rng(1)
    image = tg.Nifti1Image(rng.standard_normal((4, 5, 6, 15)).astype(np.float32), np.eye(4))
    a = tg.HipRuntime()
    b = tg.HipRuntime()
    field = tg.field.Image(image).upload(a)
    points = np.ones((4, 3), dtype=np.float32)
    out = np.zeros((4, 15), dtype=np.float32)
    lib = _native.lib()
    status = lib.tgk_field_evaluate(b.handle, field.handle, _native.ptr(points), 4, _native.ptr(out))
    assert status == _native.ERR_INVALID
    assert b"belong to the runtime" in lib.tgk_last_error()
    _native.check(lib.tgk_field_evaluate(a.handle, field.handle, _native.ptr(points), 4, _native.ptr(out)))
