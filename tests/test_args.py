"""Argument checks of the Python batch calls: a tensor of the wrong dtype,
layout or size is refused on the host before anything reaches the device."""
import pytest
import torch

import fcl_amd as F


class _NoLib:
    """Stands in for a Context: the checks run before any library call, so
    touching the library here would be a test failure."""

    def __getattr__(self, name):
        raise AssertionError("library reached before the argument checks: " + name)


def _collide_args(n=4):
    return dict(query_meshes=torch.zeros((n, 2), dtype=torch.int32),
                tf1=torch.zeros((n, 12), dtype=torch.float64),
                tf2=torch.zeros((n, 12), dtype=torch.float64),
                num_max_contacts=1, counts=torch.zeros(n, dtype=torch.int32),
                pairs=torch.zeros((n, 3), dtype=torch.int32),
                num_pairs=torch.zeros(1, dtype=torch.int64))


def _call(fn, **kw):
    return fn(_NoLib(), **kw)


def test_float32_pose_is_refused():
    kw = _collide_args()
    kw["tf1"] = kw["tf1"].float()
    with pytest.raises(ValueError, match="tf1 must be float64"):
        _call(F.Context.collide_batch_async, **kw)


def test_short_counts_is_refused():
    kw = _collide_args()
    kw["counts"] = kw["counts"][:2]
    with pytest.raises(ValueError, match="counts holds 2 elements, at least 4"):
        _call(F.Context.collide_batch_async, **kw)


def test_int64_query_ids_are_refused():
    kw = _collide_args()
    kw["query_meshes"] = kw["query_meshes"].long()
    with pytest.raises(ValueError, match="query_meshes must be int32"):
        _call(F.Context.collide_batch_async, **kw)


def test_non_contiguous_pose_is_refused():
    kw = _collide_args()
    kw["tf2"] = torch.zeros((12, 4), dtype=torch.float64).t()
    with pytest.raises(ValueError, match="tf2 must be contiguous"):
        _call(F.Context.collide_batch_async, **kw)


def test_host_tensor_is_refused():
    with pytest.raises(ValueError, match="take device tensors"):
        _call(F.Context.collide_batch_async, **_collide_args())


def test_distance_points_too_short():
    n = 3
    kw = dict(query_meshes=torch.zeros((n, 2), dtype=torch.int32),
              tf1=torch.zeros((n, 12), dtype=torch.float64),
              tf2=torch.zeros((n, 12), dtype=torch.float64),
              distance=torch.zeros(n, dtype=torch.float64),
              points=torch.zeros((n, 3), dtype=torch.float64))
    with pytest.raises(ValueError, match="points holds 9 elements, at least 18"):
        _call(F.Context.distance_batch_async, **kw)


def test_mesh_shape_short_primitives():
    n, cap = 4, 3
    rec, _ = F.shape_table([F.BOX], [(1, 1, 1)])
    kw = dict(query_meshes=torch.zeros(n, dtype=torch.int32),
              shapes=torch.from_numpy(rec.view("uint8")), vertices=None,
              query_shapes=torch.zeros(n, dtype=torch.int32),
              tf1=torch.zeros((n, 12), dtype=torch.float64),
              tf2=torch.zeros((n, 12), dtype=torch.float64), num_max_contacts=cap,
              counts=torch.zeros(n, dtype=torch.int32),
              primitives=torch.zeros(n, dtype=torch.int32))
    with pytest.raises(ValueError, match="primitives holds 4 elements, at least 12"):
        _call(F.Context.mesh_shape_collide_batch_async, **kw)


def test_shape_contact_buffers_sized_per_pair_and_cap():
    n, cap = 2, 4
    rec, _ = F.shape_table([F.SPHERE], [(1, 0, 0)])
    kw = dict(shapes=torch.from_numpy(rec.view("uint8")), vertices=None,
              pair_shapes=torch.zeros((n, 2), dtype=torch.int32),
              tf1=torch.zeros((n, 12), dtype=torch.float64),
              tf2=torch.zeros((n, 12), dtype=torch.float64), num_max_contacts=cap,
              counts=torch.zeros(n, dtype=torch.int32),
              contact=(torch.zeros((n * cap, 3), dtype=torch.float64),
                       torch.zeros((n * cap, 3), dtype=torch.float64),
                       torch.zeros(n, dtype=torch.float64)))
    with pytest.raises(ValueError, match="contact depth holds 2 elements, at least 8"):
        _call(F.Context.shape_collide_batch_async, **kw)
