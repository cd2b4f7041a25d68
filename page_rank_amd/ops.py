"""SpMM entry points: pbr_batched_matmul_cuda / pbr_matmul
(Backend/graph_link/graph_link/__init__.py:66-134)."""
from __future__ import annotations

import numpy as np
import torch

from . import _native
from .pbr import exec_layout
from .registry import get_pbr_gpu_meta


def _meta_or_raise(pbr_mat):
    meta = get_pbr_gpu_meta(pbr_mat)
    if meta is None:
        raise RuntimeError("PBRMatrix is not on GPU. Call .to('cuda') first.")
    return meta


def spmm_into(pbr_mat, x: torch.Tensor, y: torch.Tensor, *, accumulate=False, mode="fused",
              kw=0):
    """y (+)= A @ x for one [rows, K] right-hand side, both on the matrix's device.
    kw: RHS columns per launch (0 = library rule, "auto" = measured by the autotuner, which
    then also picks the mode)."""
    meta = _meta_or_raise(pbr_mat)
    vdt = meta["data"].dtype
    if vdt not in (torch.float32, torch.float64) or x.dtype != vdt or y.dtype != vdt:
        raise TypeError(f"Unsupported dtype combination: index={meta['coords'].dtype}, "
                        f"data={vdt}, x={x.dtype}, y={y.dtype}")
    if x.dim() != 2 or x.shape[0] != pbr_mat.cols or y.shape != (pbr_mat.rows, x.shape[1]):
        raise ValueError(f"shape mismatch: A is {pbr_mat.rows}x{pbr_mat.cols}, x {tuple(x.shape)}"
                         f", y {tuple(y.shape)}")
    if x.stride(1) != 1 or y.stride(1) != 1:
        raise ValueError("x and y must have unit stride along K")
    K = x.shape[1]
    if vdt == torch.float64:
        return _spmm_into_f64(pbr_mat, meta, x, y, accumulate)
    if kw == "auto":
        from .autotune import tune_spmm
        t = tune_spmm(pbr_mat, K)
        kw, mode = t["kw"], t["mode"]
    L = _native.lib()
    m = {"fused": _native.GL_SPMM_FUSED, "split": _native.GL_SPMM_SPLIT}[mode]
    with torch.cuda.device(x.device):
        lay = exec_layout(meta, pbr_mat.block_rows)
        recs, vals = (lay[0].data_ptr(), lay[1].data_ptr()) if lay else (None, None)
        rc = L.gl_pbr_spmm_exec_i32_f32(
            pbr_mat.rows, pbr_mat.cols, pbr_mat.block_rows, pbr_mat.block_cols,
            meta["codes"].numel(), meta["data"].numel(), meta["rem_col_ind"].numel(),
            meta["codes"].data_ptr(), meta["coords"].data_ptr(),
            meta["offsets"].data_ptr(), meta["data"].data_ptr(), meta["stripe_ptr"].data_ptr(),
            meta["rem_indptr"].data_ptr(), meta["rem_col_ind"].data_ptr(),
            meta["rem_vals"].data_ptr(), recs, vals, x.data_ptr(), x.stride(0), K,
            y.data_ptr(), y.stride(0), 1 if accumulate else 0, m, kw,
            _native.stream_handle(x.device))
    _native.check(rc, "gl_pbr_spmm_exec_i32_f32")
    return y


def _spmm_into_f64(pbr_mat, meta, x, y, accumulate):
    """FP64 product (bindings_kernels.cu:144, pbr_spmm_cuda_i32_f64)."""
    L = _native.lib()
    with torch.cuda.device(x.device):
        rc = L.gl_pbr_spmm_i32_f64(
            pbr_mat.rows, pbr_mat.cols, pbr_mat.block_rows, pbr_mat.block_cols,
            meta["codes"].numel(), meta["codes"].data_ptr(), meta["coords"].data_ptr(),
            meta["offsets"].data_ptr(), meta["data"].data_ptr(), meta["stripe_ptr"].data_ptr(),
            meta["rem_indptr"].data_ptr(), meta["rem_col_ind"].data_ptr(),
            meta["rem_vals"].data_ptr(), x.data_ptr(), x.stride(0), x.shape[1], y.data_ptr(),
            y.stride(0), 1 if accumulate else 0, _native.stream_handle(x.device))
    _native.check(rc, "gl_pbr_spmm_i32_f64")
    return y


def pbr_batched_matmul_cuda(pbr_mat, x: torch.Tensor, *, mode="fused", kw=0) -> torch.Tensor:
    """Y = A @ x with shape [batch, rows, features] (x is [rows, K] or [batch, rows, K]).
    A float64 matrix takes float64 x (the reference's i32_f64 variant); `mode` and `kw`
    apply to float32."""
    meta = _meta_or_raise(pbr_mat)
    if meta["data"].dtype not in (torch.float32, torch.float64) or x.dtype != meta["data"].dtype:
        raise TypeError(f"Unsupported dtype combination: index={meta['coords'].dtype}, "
                        f"data={meta['data'].dtype}, x={x.dtype}")
    x3 = x.unsqueeze(0) if x.dim() == 2 else x
    x3 = x3.contiguous()
    y = torch.empty((x3.shape[0], pbr_mat.rows, x3.shape[2]), dtype=x3.dtype, device=x3.device)
    for b in range(x3.shape[0]):
        spmm_into(pbr_mat, x3[b], y[b], mode=mode, kw=kw)
    return y


def pbr_matmul(pbr_mat, x: torch.Tensor, *, mode="fused", kw=0):
    """Unified entry point (__init__.py:104-134): x on the GPU -> HIP SpMM (uploading the
    matrix on first use); x on the CPU -> the host SpMM of the library."""
    if x.dim() not in (2, 3):
        raise ValueError(f"Expected 2D or 3D input, got {x.dim()}D")
    if not x.is_cuda:
        y = pbr_batched_matmul_cpu(pbr_mat, x)
        return y.squeeze(0) if x.dim() == 2 else y
    if get_pbr_gpu_meta(pbr_mat) is None:
        pbr_mat = pbr_mat.to(x.device)
    y = pbr_batched_matmul_cuda(pbr_mat, x, mode=mode, kw=kw)
    return y.squeeze(0) if x.dim() == 2 else y


def block_spmm_into(pbr_mat, x: torch.Tensor, y: torch.Tensor, *, accumulate=True, kw=0):
    """y (+)= A_tiles @ x: the block tiles only, the CSR remainder left out."""
    meta = _meta_or_raise(pbr_mat)
    if x.dtype != torch.float32 or y.dtype != torch.float32:
        raise TypeError("block SpMM needs float32 x and y")
    if x.dim() != 2 or x.shape[0] != pbr_mat.cols or y.shape != (pbr_mat.rows, x.shape[1]):
        raise ValueError(f"shape mismatch: A is {pbr_mat.rows}x{pbr_mat.cols}, x {tuple(x.shape)}"
                         f", y {tuple(y.shape)}")
    if x.stride(1) != 1 or y.stride(1) != 1:
        raise ValueError("x and y must have unit stride along K")
    L = _native.lib()
    with torch.cuda.device(x.device):
        rc = L.gl_pbr_block_spmm_i32_f32(
            pbr_mat.rows, pbr_mat.cols, pbr_mat.block_rows, pbr_mat.block_cols,
            meta["codes"].numel(), meta["data"].numel(), meta["codes"].data_ptr(),
            meta["coords"].data_ptr(), meta["offsets"].data_ptr(), meta["data"].data_ptr(),
            meta["stripe_ptr"].data_ptr(), x.data_ptr(), x.stride(0), x.shape[1], y.data_ptr(),
            y.stride(0), 1 if accumulate else 0, kw, _native.stream_handle(x.device))
    _native.check(rc, "gl_pbr_block_spmm_i32_f32")
    return y


def pbr_batched_matmul_triton_gpu(pbr_mat, x: torch.Tensor, y: torch.Tensor, batch_size: int,
                                  width: int):
    """Name and semantics of the reference's block-only product
    (Backend/graph_link/graph_link/pbr_matrix_triton.py:77-110): y += A_tiles @ x over the
    block tiles (the CSR remainder is not included), x and y [rows, width] or
    [batch, rows, width]. Runs the HIP block kernel; there is no Triton in this build."""
    if x.dim() == 2:
        pairs = [(x, y)]
    else:
        pairs = [(x[b], y[b]) for b in range(batch_size)]
    for xb, yb in pairs:
        if xb.shape[-1] != width:
            raise ValueError(f"width {width} does not match x {tuple(xb.shape)}")
        block_spmm_into(pbr_mat, xb, yb, accumulate=True)
    return y


def pbr_batched_matmul_cpu(pbr_mat, x: torch.Tensor, *, threads: int = 0) -> torch.Tensor:
    """Y = A @ x on the host, shape [batch, rows, features] (x is a CPU tensor [cols, K] or
    [batch, cols, K] of the matrix's value type), as the reference's pbr_batched_matmul_cpu
    (Backend/graph_link/pbr_matrix/pbr_matrix.cpp:17-68, bound at bindings.cpp:65-68).
    Runs gl_pbr_spmm_host_* of the native library (bit-identical to the reference's loop)
    on `threads` host threads (0 = all cores)."""
    if x.dim() not in (2, 3):
        raise ValueError(f"Expected 2D or 3D input, got {x.dim()}D")
    if x.is_cuda:
        raise ValueError("pbr_batched_matmul_cpu needs a CPU tensor")
    st = np.dtype(pbr_mat.scalar_dtype)
    it = np.dtype(pbr_mat.index_dtype)
    tdt = torch.float64 if st == np.float64 else torch.float32
    if x.dtype != tdt:
        raise TypeError(f"x is {x.dtype}, the matrix holds {st}")
    x3 = (x.unsqueeze(0) if x.dim() == 2 else x).contiguous()
    if x3.shape[1] != pbr_mat.cols:
        raise ValueError(f"shape mismatch: A is {pbr_mat.rows}x{pbr_mat.cols}, x "
                         f"{tuple(x.shape)}")
    B, _, K = x3.shape
    y = torch.empty((B, pbr_mat.rows, K), dtype=tdt)
    if K == 0 or B == 0:
        return y.zero_()
    L = _native.lib()
    fn = getattr(L, f"gl_pbr_spmm_host_{'i64' if it == np.int64 else 'i32'}_"
                    f"{'f64' if st == np.float64 else 'f32'}")
    arrs = [np.ascontiguousarray(a) for a in (
        pbr_mat.block_codes, pbr_mat.block_coords, pbr_mat.block_offsets, pbr_mat.block_data,
        pbr_mat.remainder_indptr, pbr_mat.remainder_col_ind, pbr_mat.remainder_data)]
    for b in range(B):
        xb, yb = x3[b], y[b]
        rc = fn(pbr_mat.rows, pbr_mat.cols, pbr_mat.block_rows, pbr_mat.block_cols,
                arrs[0].size, *(a.ctypes.data for a in arrs), xb.data_ptr(), K, K,
                yb.data_ptr(), K, int(threads))
        _native.check(rc, "gl_pbr_spmm_host")
    return y


class SpmmPlan:
    """A bound SpMM call (matrix, x, y, mode, kw fixed): ``run()`` costs one C call.
    For launch-bound loops (benchmarks, repeated products on the same buffers)."""

    def __init__(self, pbr_mat, x: torch.Tensor, y: torch.Tensor, *, accumulate=False,
                 mode="fused", kw=0):
        meta = _meta_or_raise(pbr_mat)
        if x.dtype != torch.float32 or y.dtype != torch.float32:
            raise TypeError("SpmmPlan needs float32 x and y")
        if x.dim() != 2 or x.shape[0] != pbr_mat.cols or y.shape != (pbr_mat.rows, x.shape[1]):
            raise ValueError("shape mismatch")
        if x.stride(1) != 1 or y.stride(1) != 1:
            raise ValueError("x and y must have unit stride along K")
        if kw == "auto":
            from .autotune import tune_spmm
            t = tune_spmm(pbr_mat, x.shape[1])
            kw, mode = t["kw"], t["mode"]
        with torch.cuda.device(x.device):
            lay = exec_layout(meta, pbr_mat.block_rows)
        recs, vals = (lay[0].data_ptr(), lay[1].data_ptr()) if lay else (None, None)
        self._keep = (meta, x, y, lay)
        self._fn = _native.lib().gl_pbr_spmm_exec_i32_f32
        m = {"fused": _native.GL_SPMM_FUSED, "split": _native.GL_SPMM_SPLIT}[mode]
        self._args = (
            pbr_mat.rows, pbr_mat.cols, pbr_mat.block_rows, pbr_mat.block_cols,
            meta["codes"].numel(), meta["data"].numel(), meta["rem_col_ind"].numel(),
            meta["codes"].data_ptr(), meta["coords"].data_ptr(),
            meta["offsets"].data_ptr(), meta["data"].data_ptr(), meta["stripe_ptr"].data_ptr(),
            meta["rem_indptr"].data_ptr(), meta["rem_col_ind"].data_ptr(),
            meta["rem_vals"].data_ptr(), recs, vals, x.data_ptr(), x.stride(0), x.shape[1],
            y.data_ptr(), y.stride(0), 1 if accumulate else 0, m, kw,
            _native.stream_handle(x.device))

    def run(self):
        rc = self._fn(*self._args)
        if rc:
            _native.check(rc, "gl_pbr_spmm_exec_i32_f32")
