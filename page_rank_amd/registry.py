"""Device residency of PBR matrices: ``get_pbr_gpu_meta`` / ``set_pbr_gpu_meta`` /
``clear_pbr_gpu_meta``, the names the reference's callers use
(Backend/graph_link/graph_link/pbr_registry.py, used by __init__.py:17-56, :66-101).

The reference's PBR objects are C++ classes bound with pybind11, which cannot carry the
uploaded tensors, so it keeps them in a module dictionary keyed by ``id()`` and has to guard
against ids being reused. Here a PBR matrix is a Python object, so the uploaded arrays are
an attribute of the matrix itself: they are freed with it by ordinary reference counting, a
copy never inherits them, and there is no table to go stale.

The arrays are a dict of device tensors with the keys codes (int64 bit patterns of the
uint64 codes), coords, offsets, rem_indptr, rem_col_ind (int32), data, rem_vals (the value
type) and stripe_ptr (int32, first tile of each block row).
"""
from __future__ import annotations

_REQUIRED = ("codes", "coords", "offsets", "data", "rem_indptr", "rem_col_ind", "rem_vals")


def get_pbr_gpu_meta(pbr_mat):
    """The device arrays of ``pbr_mat``, or None while it lives on the host only."""
    return pbr_mat.device_arrays()


def set_pbr_gpu_meta(pbr_mat, meta_dict):
    """Make ``meta_dict`` the device arrays of ``pbr_mat`` (replacing any earlier ones)."""
    missing = [k for k in _REQUIRED if k not in meta_dict]
    if missing:
        raise KeyError(f"device arrays lack {missing}")
    n = int(meta_dict["rem_indptr"].numel())
    if n != pbr_mat.rows + 1:
        raise ValueError(f"rem_indptr has {n} entries for a matrix of {pbr_mat.rows} rows")
    pbr_mat.attach_device_arrays(meta_dict)


def clear_pbr_gpu_meta(pbr_mat):
    """Drop the device arrays of ``pbr_mat``; the matrix keeps (or first fetches) its host
    arrays, so it stays usable."""
    pbr_mat.detach_device_arrays()
