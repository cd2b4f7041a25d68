"""Device CSR helpers around the PPR path: GPU column normalization with implicit dangling
handling (O(nnz), no dense 1/N block), PBR -> CSR on the GPU, and an .npz cache of PBR
matrices.

normalize_transition_matrix_gpu  the reference's normalize_transition_matrix
                                 (Backend/graph_link/graph_link/__init__.py:184-207) on the
                                 device; identical float32 values for every non-dangling
                                 column; dangling columns come back as a mask for
                                 run_ppr(..., dangling_mode="uniform")
pbr_to_csr                       Backend/graph_link/graph_link/__init__.py:159-160 ->
                                 pbr_matrix.cpp:289-338, filling caller-allocated arrays
save_pbr / load_pbr              host arrays of a PBR matrix in one .npz (no pickle)
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp
import torch

from . import _native
from .registry import get_pbr_gpu_meta


@dataclass
class DeviceCSR:
    """A CSR matrix whose arrays live on a GPU (int32 indices, float32 values)."""
    rows: int
    cols: int
    indptr: torch.Tensor
    indices: torch.Tensor
    data: torch.Tensor
    canonical: bool = True   # rows sorted by column, no duplicates

    @property
    def shape(self):
        return (self.rows, self.cols)

    @property
    def nnz(self):
        return int(self.indices.numel())

    @property
    def device(self):
        return self.indices.device

    def to_scipy(self) -> sp.csr_matrix:
        return sp.csr_matrix((self.data.cpu().numpy(), self.indices.cpu().numpy(),
                              self.indptr.cpu().numpy()), shape=self.shape)


def _device(device):
    if device is None:
        return torch.device("cuda", torch.cuda.current_device())
    d = torch.device(device)
    if d.type == "cuda" and d.index is None:
        d = torch.device("cuda", torch.cuda.current_device())
    return d


def upload_csr(A, device=None) -> DeviceCSR:
    """Host CSR (scipy) -> DeviceCSR with float32 values and int32 indices."""
    if isinstance(A, DeviceCSR):
        return A
    A = sp.csr_matrix(A)
    if max(A.shape[0], A.shape[1], A.nnz) >= 2**31 - 1:
        raise ValueError("matrix too large for int32 device indices")
    dev = _device(device)
    return DeviceCSR(
        A.shape[0], A.shape[1],
        torch.as_tensor(A.indptr.astype(np.int32, copy=False)).to(dev),
        torch.as_tensor(A.indices.astype(np.int32, copy=False)).to(dev),
        torch.as_tensor(np.ascontiguousarray(A.data.astype(np.float32, copy=False))).to(dev),
        canonical=bool(A.has_canonical_format))


def normalize_columns(m: DeviceCSR):
    """Divide every column of a device CSR by its float32 sum. Returns
    (normalized DeviceCSR, col_sums [cols] float32, dangling [cols] uint8)."""
    if not m.canonical:
        raise ValueError("normalize_columns needs canonical CSR (sorted, no duplicates)")
    dev = m.device
    L = _native.lib()
    wsb = L.gl_csr_normalize_workspace_bytes(m.cols, m.nnz) if m.cols else 0
    if wsb < 0:
        raise ValueError("invalid sizes for normalization")
    ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=dev)
    out = torch.empty_like(m.data)
    sums = torch.zeros(m.cols, dtype=torch.float32, device=dev)
    dang = torch.zeros(m.cols, dtype=torch.uint8, device=dev)
    with torch.cuda.device(dev):
        _native.check(L.gl_csr_normalize_columns_f32(
            m.cols, m.nnz, m.indices.data_ptr(), m.data.data_ptr(), out.data_ptr(),
            sums.data_ptr(), dang.data_ptr(), ws.data_ptr(), wsb,
            _native.stream_handle(dev)), "gl_csr_normalize_columns_f32")
    return DeviceCSR(m.rows, m.cols, m.indptr, m.indices, out, True), sums, dang


def normalize_transition_matrix_gpu(A_csr, device=None):
    """Column-stochastic transition matrix on the device and the dangling mask.

    Every column with a positive sum holds exactly the float32 values of
    normalize_transition_matrix (the reference's host function); a dangling column stays
    empty and is reported in the mask, to be handled by run_ppr(..., dangling=mask,
    dangling_mode="uniform"). That equals the reference's dense 1/N column without
    materializing N entries per dangling node. Returns (DeviceCSR, numpy bool mask)."""
    m = upload_csr(A_csr, device)
    if not m.canonical:
        A = sp.csr_matrix(A_csr).copy()
        A.sum_duplicates()
        m = upload_csr(A, device)
    norm, _, dang = normalize_columns(m)
    return norm, dang.cpu().numpy().astype(bool)


def pbr_to_csr(pbr_mat, indptr, indices, data):
    """Fill indptr[rows+1], indices[nnz], data[nnz] with the canonical CSR of the PBR matrix
    (the reference's signature: caller-allocated arrays, numpy or torch). The conversion runs
    on the GPU; a host-only matrix is uploaded for it. Unlike the reference, a last block row
    that extends past the matrix is accepted."""
    meta = get_pbr_gpu_meta(pbr_mat)
    tmp = None
    if meta is None:
        tmp = pbr_mat.to(_device(None))
        meta = get_pbr_gpu_meta(tmp)
    vdt = meta["data"].dtype
    if vdt not in (torch.float32, torch.float64):
        raise TypeError(f"Unsupported value dtype {vdt}")
    f64 = vdt == torch.float64
    dev = meta["codes"].device
    n = int(np.asarray(indices).shape[0]) if not isinstance(indices, torch.Tensor) \
        else indices.numel()
    if (data.shape[0] if not isinstance(data, torch.Tensor) else data.numel()) != n:
        raise ValueError("indices and data must have the same length")
    rows = pbr_mat.rows
    ip = torch.empty(rows + 1, dtype=torch.int32, device=dev)
    ix = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    vx = torch.empty(max(n, 1), dtype=vdt, device=dev)
    L = _native.lib()
    wsb = (L.gl_pbr_to_csr_f64_workspace_bytes if f64 else L.gl_pbr_to_csr_workspace_bytes)(
        rows, n)
    if wsb < 0:
        raise ValueError("invalid sizes for pbr_to_csr")
    ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=dev)
    fn = L.gl_pbr_to_csr_i32_f64 if f64 else L.gl_pbr_to_csr_i32_f32
    with torch.cuda.device(dev):
        _native.check(fn(
            rows, pbr_mat.cols, pbr_mat.block_rows, pbr_mat.block_cols,
            meta["codes"].numel(), meta["data"].numel(), meta["rem_col_ind"].numel(),
            meta["codes"].data_ptr(), meta["coords"].data_ptr(), meta["offsets"].data_ptr(),
            meta["data"].data_ptr(), meta["rem_indptr"].data_ptr(),
            meta["rem_col_ind"].data_ptr(), meta["rem_vals"].data_ptr(), ip.data_ptr(),
            ix.data_ptr(), vx.data_ptr(), n, ws.data_ptr(), wsb,
            _native.stream_handle(dev)), "gl_pbr_to_csr")
    for src, dst in ((ip, indptr), (ix[:n], indices), (vx[:n], data)):
        if isinstance(dst, torch.Tensor):
            dst.copy_(src.to(dst.dtype))
        else:
            dst[...] = src.cpu().numpy().astype(dst.dtype, copy=False)
    del tmp
    return indptr, indices, data


def pbr_to_scipy(pbr_mat) -> sp.csr_matrix:
    """The PBR matrix as a canonical scipy CSR of its value type."""
    n = pbr_mat.total_nnz
    indptr = np.empty(pbr_mat.rows + 1, np.int32)
    indices = np.empty(n, np.int32)
    data = np.empty(n, np.dtype(pbr_mat.scalar_dtype))
    pbr_to_csr(pbr_mat, indptr, indices, data)
    return sp.csr_matrix((data, indices, indptr), shape=(pbr_mat.rows, pbr_mat.cols))


_NPZ_FIELDS = ("block_codes", "block_coords", "block_offsets", "block_data",
               "remainder_indptr", "remainder_col_ind", "remainder_data")


def save_pbr(pbr_mat, path):
    """Write the host arrays and shape of a PBR matrix to one .npz file."""
    arrays = {k: np.asarray(getattr(pbr_mat, k)) for k in _NPZ_FIELDS}
    np.savez(path, shape=np.array([pbr_mat.rows, pbr_mat.cols, pbr_mat.block_rows,
                                   pbr_mat.block_cols, pbr_mat.total_nnz], dtype=np.int64),
             **arrays)


def load_pbr(path, device=None):
    """Read a .npz written by save_pbr (no pickle); device=None keeps it on the host."""
    from .pbr import from_arrays
    with np.load(path, allow_pickle=False) as z:
        rows, cols, br, bc, _ = (int(v) for v in z["shape"])
        m = from_arrays(rows, cols, br, bc, *(z[k] for k in _NPZ_FIELDS))
    return m if device is None else m.to(device)
