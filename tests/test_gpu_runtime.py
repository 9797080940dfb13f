"""Runtime lifetime: closing a runtime frees the fields and domains made on
it first, and later use of either raises instead of touching freed memory."""

import numpy as np
import pytest

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
