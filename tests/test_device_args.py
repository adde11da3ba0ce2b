"""Device-tensor arguments are checked on the host before the library sees
them: a tensor on the wrong device, of the wrong dtype, non-contiguous or of
the wrong size raises ValueError, so no kernel reads host memory or past the
end of a buffer. These run without a GPU: the objects are built without
library handles, and every case must fail before any library call."""
import numpy as np
import pytest
import torch

from raptor_amd import rt


def test_dev_checks_dtype_layout_size_and_device():
    with pytest.raises(ValueError, match="dtype"):
        rt._dev(torch.zeros(30, dtype=torch.float64), torch.float32, 30, "x")
    with pytest.raises(ValueError, match="contiguous"):
        rt._dev(torch.zeros((10, 4), dtype=torch.float32)[:, :3], torch.float32, 30, "x")
    with pytest.raises(ValueError, match="elements"):
        rt._dev(torch.zeros(29, dtype=torch.float32), torch.float32, 30, "x")
    with pytest.raises(ValueError, match="at least"):
        rt._dev(torch.zeros(5, dtype=torch.float32), torch.float32, -6, "x")
    with pytest.raises(ValueError, match="GPU tensor"):
        rt._dev(torch.zeros(30, dtype=torch.float32), torch.float32, 30, "x")
    with pytest.raises(ValueError, match="torch tensor"):
        rt._dev(np.zeros(30, np.float32), torch.float32, 30, "x")


def test_host_arrays_must_be_contiguous():
    with pytest.raises(ValueError, match="contiguous"):
        rt._ptr(np.zeros((4, 4), np.float32)[:, :2])


class _NoLib:
    """Stands in for a context: any library call through it is a failure."""

    def check(self, *a, **k):
        raise AssertionError("the library was called")


def _scene(n):
    s = object.__new__(rt.Scene)
    s.ctx, s.n, s.h = _NoLib(), n, None
    return s


def _accel(n):
    a = object.__new__(rt.Accel)
    a.scene, a.ctx, a.h = _scene(n), _NoLib(), None
    return a


@pytest.mark.parametrize("bad", [torch.zeros((10, 3), dtype=torch.float64), torch.zeros((9, 3), dtype=torch.float32),
                                 torch.zeros((10, 4), dtype=torch.float32)[:, :3], torch.zeros((10, 3))])
def test_set_albedo_rejects_bad_tensors(bad):
    with pytest.raises(ValueError):
        _scene(10).set_albedo(bad)


@pytest.mark.parametrize("bad", [torch.zeros((10, 9), dtype=torch.float64), torch.zeros((11, 9)), torch.zeros((10, 9))])
def test_update_rejects_bad_tensors(bad):
    with pytest.raises(ValueError):
        _scene(10).update(bad)


@pytest.mark.parametrize("bad", [torch.zeros((16, 8), dtype=torch.float64), torch.zeros((16, 9))[:, :8],
                                 torch.zeros((16, 7)), torch.zeros((16, 8))])
def test_trace_rejects_bad_ray_tensors(bad):
    a = _accel(10)
    with pytest.raises(ValueError):
        a.trace_closest(bad)
    with pytest.raises(ValueError):
        a.trace_occluded(bad)


def test_batch_rejects_bad_hit_tensors():
    rays = [torch.zeros((64, 8))]
    with pytest.raises(ValueError):
        rt._batch_items([_accel(10)], rays, [torch.zeros((63, 4))])
    with pytest.raises(ValueError):
        rt._batch_items([_accel(10)], rays, [torch.zeros((64, 4), dtype=torch.float64)])
