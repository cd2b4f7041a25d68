"""ORACLE — test infrastructure only. ctypes wrapper of oracle/pbr_oracle.cpp.

Follows Backend/graph_link/pbr_matrix/pbr_matrix.cpp:340-477 (builder), :17-68 (CPU SpMM),
:289-338 (PBR -> CSR).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libpbr_oracle.so")
_lib = None

_ITYPES = {np.dtype(np.int32): ("i32", ctypes.c_int32), np.dtype(np.int64): ("i64", ctypes.c_int64)}
_STYPES = {np.dtype(np.float32): "f32", np.dtype(np.float64): "f64"}


def build() -> str:
    """Compile the oracle library (g++) if it is missing or older than its source."""
    src = os.path.join(_HERE, "pbr_oracle.cpp")
    if not os.path.exists(_SO) or os.path.getmtime(_SO) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        _lib = ctypes.CDLL(_SO)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class OraclePbr:
    rows: int
    cols: int
    block_rows: int
    block_cols: int
    total_nnz: int
    block_codes: np.ndarray   # uint64 [T]
    block_coords: np.ndarray  # index [2T]
    block_offsets: np.ndarray # index [T+1]
    block_data: np.ndarray    # scalar [V]
    remainder_indptr: np.ndarray
    remainder_col_ind: np.ndarray
    remainder_data: np.ndarray

    @property
    def num_tiles(self) -> int:
        return int(self.block_codes.size)


def csr_to_pbr(indptr, indices, data, rows, cols, block_rows=2, block_cols=2,
               min_nnz_per_block=2, order="reference") -> OraclePbr:
    """CPU builder. order='reference' (std::unordered_map iteration, as the reference) or
    'canonical' (block columns ascending inside each stripe)."""
    indptr = np.ascontiguousarray(indptr)
    it = indptr.dtype
    indices = np.ascontiguousarray(indices, dtype=it)
    data = np.ascontiguousarray(data)
    isuf, ictype = _ITYPES[np.dtype(it)]
    ssuf = _STYPES[np.dtype(data.dtype)]
    suf = f"{isuf}_{ssuf}"
    L = lib()
    fn = getattr(L, f"orc_csr_to_pbr_{suf}")
    fn.restype = ctypes.c_void_p
    fn.argtypes = [ictype, ictype, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ictype,
                   ictype, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p]
    sizes = np.zeros(3, dtype=np.int64)
    h = fn(rows, cols, _p(indptr), _p(indices), _p(data), block_rows, block_cols,
           int(min_nnz_per_block), 0 if order == "reference" else 1, _p(sizes))
    T, V, R = (int(s) for s in sizes)
    out = OraclePbr(
        rows, cols, block_rows, block_cols, int(indptr[rows]) if rows > 0 else 0,
        np.zeros(T, np.uint64), np.zeros(2 * T, it), np.zeros(T + 1, it),
        np.zeros(V, data.dtype), np.zeros(rows + 1, it), np.zeros(R, it), np.zeros(R, data.dtype))
    get = getattr(L, f"orc_pbr_get_{suf}")
    get.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 7
    get(h, _p(out.block_codes), _p(out.block_coords), _p(out.block_offsets), _p(out.block_data),
        _p(out.remainder_indptr), _p(out.remainder_col_ind), _p(out.remainder_data))
    free = getattr(L, f"orc_pbr_free_{suf}")
    free.argtypes = [ctypes.c_void_p]
    free(h)
    return out


def csr_matrix_to_pbr(mat, block_rows=2, block_cols=2, min_nnz_per_block=2, order="reference"):
    return csr_to_pbr(mat.indptr, mat.indices, mat.data, mat.shape[0], mat.shape[1],
                      block_rows, block_cols, min_nnz_per_block, order)


def spmm(p, X, wide=False):
    """Y = A X on the CPU, tiles in stored order then the remainder (reference order of
    accumulation); wide=True accumulates in float64."""
    it = p.block_coords.dtype
    isuf, ictype = _ITYPES[np.dtype(it)]
    ssuf = _STYPES[np.dtype(p.block_data.dtype)]
    X = np.ascontiguousarray(X, dtype=p.block_data.dtype)
    if X.ndim == 1:
        X = X[:, None]
    K = X.shape[1]
    Y = np.zeros((p.rows, K), dtype=p.block_data.dtype)
    fn = getattr(lib(), f"orc_pbr_spmm_{isuf}_{ssuf}")
    fn.argtypes = [ictype] * 4 + [ctypes.c_int64] + [ctypes.c_void_p] * 7 + [
        ctypes.c_int64, ctypes.c_void_p, ctypes.c_int32]
    fn(p.rows, p.cols, p.block_rows, p.block_cols, p.num_tiles, _p(p.block_codes),
       _p(p.block_coords), _p(p.block_data), _p(p.remainder_indptr), _p(p.remainder_col_ind),
       _p(p.remainder_data), _p(X), K, _p(Y), 1 if wide else 0)
    return Y


def pbr_to_csr(p, strict=False):
    """(indptr, indices, data) of the PBR matrix; strict=True refuses tiles of a trailing
    partial stripe like the reference (raises RuntimeError)."""
    it = p.block_coords.dtype
    isuf, ictype = _ITYPES[np.dtype(it)]
    ssuf = _STYPES[np.dtype(p.block_data.dtype)]
    nnz = int(p.block_data.size + p.remainder_data.size)
    indptr = np.zeros(p.rows + 1, it)
    indices = np.zeros(nnz, it)
    data = np.zeros(nnz, p.block_data.dtype)
    fn = getattr(lib(), f"orc_pbr_to_csr_{isuf}_{ssuf}")
    fn.restype = ctypes.c_int
    fn.argtypes = [ictype] * 4 + [ctypes.c_int64] + [ctypes.c_void_p] * 9 + [ctypes.c_int32]
    rc = fn(p.rows, p.cols, p.block_rows, p.block_cols, p.num_tiles, _p(p.block_codes),
            _p(p.block_coords), _p(p.block_data), _p(p.remainder_indptr),
            _p(p.remainder_col_ind), _p(p.remainder_data), _p(indptr), _p(indices), _p(data),
            1 if strict else 0)
    if rc != 0:
        raise RuntimeError("Block coordinate exceeds matrix dimensions. Check block size and "
                           "matrix dimensions for compatibility.")
    return indptr, indices, data


def bucket_counts(hint, keys):
    """Bucket count of std::unordered_map<int32_t,int>(hint) after each emplace of keys."""
    keys = np.ascontiguousarray(keys, dtype=np.int32)
    counts = np.zeros(keys.size, dtype=np.int64)
    fn = lib().orc_bucket_count_after
    fn.restype = ctypes.c_int64
    fn.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    final = fn(int(hint), _p(keys), keys.size, _p(counts))
    return int(final), counts


def bucket_trace(hint, keys, stripe_ptr, key_dtype=np.int32):
    """Bucket count after each emplace, stripes separated by clear()."""
    keys = np.ascontiguousarray(keys, dtype=key_dtype)
    sptr = np.ascontiguousarray(stripe_ptr, dtype=np.int64)
    counts = np.zeros(keys.size, dtype=np.int64)
    suf = "i32" if np.dtype(key_dtype) == np.int32 else "i64"
    fn = getattr(lib(), f"orc_bucket_trace_{suf}")
    fn.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                   ctypes.c_void_p]
    fn(int(hint), _p(keys), _p(sptr), sptr.size - 1, _p(counts))
    return counts
