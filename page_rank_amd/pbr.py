"""PBR matrix objects (host view + device residency), mirroring the reference's
PBRMatrix{Int64,Int32}{Float,Double} classes (Backend/graph_link/bindings/bindings.cpp:9-50,
Backend/graph_link/pbr_matrix/pbr_matrix.hpp:98-167) and the injected ``.to()``
(Backend/graph_link/graph_link/__init__.py:17-56).

Differences from the reference, all deliberate:
  * array properties return read-only numpy views instead of fresh copies;
  * a matrix built by the GPU builder is already resident on the device that built it
    (``.to()`` to that device returns ``self``), and its host arrays are fetched lazily;
  * ``.to('cuda')`` and ``.to('cuda:0')`` name the same device (the reference compares
    device strings and re-uploads);
  * float64 matrices upload as float64 and run the FP64 kernels (SpMM and the reference's
    PPR loop); the batched seed-set PPR, the warm start and the tuned SpMM plans are FP32.
"""
from __future__ import annotations

import copy
import enum

import numpy as np

from . import _native
from .registry import clear_pbr_gpu_meta, get_pbr_gpu_meta, set_pbr_gpu_meta


class DataOrder(enum.IntEnum):
    """pbr_matrix.hpp:91-96; only BLOCK_ROW_MAJOR is ever produced."""
    BLOCK_ROW_MAJOR = 0
    BLOCK_COLUMN_MAJOR = 1
    INTERLEAVED_ROW_MAJOR = 2
    INTERLEAVED_COLUMN_MAJOR = 3


_HOST_FIELDS = ("block_codes", "block_coords", "block_offsets", "block_data",
                "remainder_indptr", "remainder_col_ind", "remainder_data")


def _ro(a):
    a = np.asarray(a)
    v = a.view()
    v.flags.writeable = False
    return v


class PBRMatrix:
    """Packed block representation: tiles of block_rows x block_cols with a 64-bit occupancy
    code and packed values, plus a CSR remainder for tiles below min_nnz_per_block."""

    index_dtype = np.dtype(np.int32)
    scalar_dtype = np.dtype(np.float32)

    def __init__(self, rows, cols, block_rows, block_cols, total_nnz, *, host=None,
                 device_meta=None, data_order=DataOrder.BLOCK_ROW_MAJOR):
        self._rows = int(rows)
        self._cols = int(cols)
        self._block_rows = int(block_rows)
        self._block_cols = int(block_cols)
        self._total_nnz = int(total_nnz)
        self._data_order = DataOrder(data_order)
        self._host = None
        if host is not None:
            self._host = {k: _ro(host[k]) for k in _HOST_FIELDS}
        self._dev = None  # device arrays (dict of tensors) while resident on a GPU
        if device_meta is not None:
            set_pbr_gpu_meta(self, device_meta)

    # ---- device residency (the registry functions read and write these) ----
    def device_arrays(self):
        return self._dev

    def attach_device_arrays(self, meta):
        self._dev = meta

    def detach_device_arrays(self):
        if self._dev is not None and self._host is None:
            self._h()  # a GPU-built matrix keeps its arrays on the host
        self._dev = None

    # ---- read-only scalars (bindings.cpp:11-16) ----
    rows = property(lambda self: self._rows)
    cols = property(lambda self: self._cols)
    block_rows = property(lambda self: self._block_rows)
    block_cols = property(lambda self: self._block_cols)
    total_nnz = property(lambda self: self._total_nnz)
    data_order = property(lambda self: self._data_order)

    # ---- arrays (bindings.cpp:17-46) ----
    def _h(self):
        if self._host is None:
            meta = self._dev
            if meta is None:
                raise RuntimeError("PBR matrix has neither host nor device arrays")
            it, st = self.index_dtype, self.scalar_dtype
            self._host = {
                "block_codes": _ro(meta["codes"].cpu().numpy().view(np.uint64)),
                "block_coords": _ro(meta["coords"].cpu().numpy().astype(it, copy=False)),
                "block_offsets": _ro(meta["offsets"].cpu().numpy().astype(it, copy=False)),
                "block_data": _ro(meta["data"].cpu().numpy().view(st)),
                "remainder_indptr": _ro(meta["rem_indptr"].cpu().numpy().astype(it, copy=False)),
                "remainder_col_ind": _ro(meta["rem_col_ind"].cpu().numpy().astype(it, copy=False)),
                "remainder_data": _ro(meta["rem_vals"].cpu().numpy().view(st)),
            }
        return self._host

    block_codes = property(lambda self: self._h()["block_codes"])
    block_coords = property(lambda self: self._h()["block_coords"])
    block_offsets = property(lambda self: self._h()["block_offsets"])
    block_data = property(lambda self: self._h()["block_data"])
    remainder_indptr = property(lambda self: self._h()["remainder_indptr"])
    remainder_col_ind = property(lambda self: self._h()["remainder_col_ind"])
    remainder_data = property(lambda self: self._h()["remainder_data"])

    # ---- counts (pbr_matrix.hpp:156-166) ----
    def accounted_blocks(self) -> int:
        if self._host is None and self._dev is not None:
            return int(self._dev["codes"].numel())
        return int(self._h()["block_codes"].size)

    def compressed_nnz(self) -> int:
        if self._host is None and self._dev is not None:
            return int(self._dev["data"].numel())
        return int(self._h()["block_data"].size)

    def remainder_nnz(self) -> int:
        if self._host is None and self._dev is not None:
            return int(self._dev["rem_col_ind"].numel())
        return int(self._h()["remainder_col_ind"].size)

    def __copy__(self):
        """A new matrix with the same host arrays and no device arrays (the reference's copy
        constructs a new C++ object, which its registry does not know)."""
        new = object.__new__(type(self))
        new.__dict__.update(self.__dict__)
        new._host = self._h()
        new._dev = None
        return new

    def __repr__(self):
        return (f"{type(self).__name__}(rows={self.rows}, cols={self.cols}, "
                f"block={self.block_rows}x{self.block_cols}, tiles={self.accounted_blocks()}, "
                f"tile_nnz={self.compressed_nnz()}, remainder_nnz={self.remainder_nnz()})")

    # ---- device residency (__init__.py:17-56) ----
    def to(self, device="cuda"):
        import torch
        target = None if device is None else torch.device(device)
        if target is not None and target.type == "cuda" and target.index is None:
            target = torch.device("cuda", torch.cuda.current_device())
        meta = get_pbr_gpu_meta(self)
        current = meta["codes"].device if meta is not None else torch.device("cpu")
        tgt = target if target is not None else torch.device("cpu")
        if current == tgt:
            return self
        new = copy.copy(self)
        clear_pbr_gpu_meta(new)
        if tgt.type != "cpu":
            set_pbr_gpu_meta(new, upload(self, tgt))
        return new


class PBRMatrixInt32Float(PBRMatrix):
    index_dtype = np.dtype(np.int32)
    scalar_dtype = np.dtype(np.float32)


class PBRMatrixInt64Float(PBRMatrix):
    index_dtype = np.dtype(np.int64)
    scalar_dtype = np.dtype(np.float32)


class PBRMatrixInt32Double(PBRMatrix):
    index_dtype = np.dtype(np.int32)
    scalar_dtype = np.dtype(np.float64)


class PBRMatrixInt64Double(PBRMatrix):
    index_dtype = np.dtype(np.int64)
    scalar_dtype = np.dtype(np.float64)


_CLASSES = {
    (np.dtype(np.int32), np.dtype(np.float32)): PBRMatrixInt32Float,
    (np.dtype(np.int64), np.dtype(np.float32)): PBRMatrixInt64Float,
    (np.dtype(np.int32), np.dtype(np.float64)): PBRMatrixInt32Double,
    (np.dtype(np.int64), np.dtype(np.float64)): PBRMatrixInt64Double,
}


def pbr_class(index_dtype, scalar_dtype):
    key = (np.dtype(index_dtype), np.dtype(scalar_dtype))
    if key not in _CLASSES:
        raise TypeError(f"Unsupported dtype combination: index={key[0]}, data={key[1]}")
    return _CLASSES[key]


def from_arrays(rows, cols, block_rows, block_cols, block_codes, block_coords, block_offsets,
                block_data, remainder_indptr, remainder_col_ind, remainder_data,
                total_nnz=None):
    """Wrap existing PBR host arrays (for example a saved matrix) in a PBR object."""
    it = np.asarray(block_coords).dtype
    st = np.asarray(block_data).dtype
    cls = pbr_class(it, st)
    host = dict(block_codes=np.asarray(block_codes, dtype=np.uint64),
                block_coords=np.asarray(block_coords, dtype=it),
                block_offsets=np.asarray(block_offsets, dtype=it),
                block_data=np.asarray(block_data, dtype=st),
                remainder_indptr=np.asarray(remainder_indptr, dtype=it),
                remainder_col_ind=np.asarray(remainder_col_ind, dtype=it),
                remainder_data=np.asarray(remainder_data, dtype=st))
    if total_nnz is None:
        total_nnz = host["block_data"].size + host["remainder_data"].size
    return cls(rows, cols, block_rows, block_cols, total_nnz, host=host)


def stripe_index(coords_dev, rows, block_rows, num_tiles, device):
    """Device stripe_ptr[num_stripes + 1] of stripe-ascending tiles."""
    import torch
    ns = (rows + block_rows - 1) // block_rows
    sp = torch.empty(ns + 1, dtype=torch.int32, device=device)
    L = _native.lib()
    with torch.cuda.device(device):
        _native.check(L.gl_pbr_stripe_index(rows, block_rows, num_tiles,
                                            _native.ptr(coords_dev), _native.ptr(sp),
                                            _native.stream_handle(device)),
                      "gl_pbr_stripe_index")
    return sp


def exec_layout(meta, block_rows):
    """The matrix-core SpMM's execution layout of an 8x8 float32 matrix (tile records with
    transposed codes, values in column-major bit order), built on the device at the first
    product and kept in the device arrays; None for other matrices."""
    import torch
    if block_rows != 8 or meta["data"].dtype != torch.float32:
        return None
    if "tile_records" not in meta:
        L = _native.lib()
        dev = meta["codes"].device
        T, V = meta["codes"].numel(), meta["data"].numel()
        recs = torch.empty(L.gl_pbr_exec_layout_bytes(T), dtype=torch.uint8, device=dev)
        vals = torch.empty(max(V, 1), dtype=torch.float32, device=dev)
        with torch.cuda.device(dev):
            _native.check(L.gl_pbr_exec_layout(T, V, _native.ptr(meta["codes"]),
                                               _native.ptr(meta["coords"]),
                                               _native.ptr(meta["offsets"]),
                                               _native.ptr(meta["data"]), _native.ptr(recs),
                                               _native.ptr(vals), _native.stream_handle(dev)),
                          "gl_pbr_exec_layout")
        meta["values_cm"] = vals
        meta["tile_records"] = recs
    return meta["tile_records"], meta["values_cm"]


def upload(p: PBRMatrix, device):
    """Copy host arrays to the device: codes as int64 bit patterns, all indices int32
    (as __init__.py:43-51), values as stored, plus the stripe index."""
    import torch
    _native.lib()  # fail loudly before any copy if the HIP library is missing
    if p.scalar_dtype not in (np.float32, np.float64):
        raise TypeError(f"Unsupported value dtype {p.scalar_dtype}")
    vdt = torch.float64 if p.scalar_dtype == np.float64 else torch.float32
    if p.block_rows not in (2, 4, 8) or p.block_rows != p.block_cols:
        raise ValueError("Unsupported block size for PBR SpMM. Only 2x2, 4x4, and 8x8 are "
                         "supported.")
    if max(p.rows, p.cols, p.total_nnz) >= 2**31:
        raise ValueError("matrix too large for int32 device indices")
    h = p._h()

    def dev(a, dt):
        return torch.as_tensor(np.ascontiguousarray(a)).to(device=device, dtype=dt)

    meta = {
        "codes": torch.as_tensor(np.ascontiguousarray(h["block_codes"]).view(np.int64)).to(device),
        "coords": dev(h["block_coords"], torch.int32),
        "offsets": dev(h["block_offsets"], torch.int32),
        "data": dev(h["block_data"], vdt),
        "rem_indptr": dev(h["remainder_indptr"], torch.int32),
        "rem_col_ind": dev(h["remainder_col_ind"], torch.int32),
        "rem_vals": dev(h["remainder_data"], vdt),
    }
    meta["stripe_ptr"] = stripe_index(meta["coords"], p.rows, p.block_rows,
                                      meta["codes"].numel(), device)
    return meta
