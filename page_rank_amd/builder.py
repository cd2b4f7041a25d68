"""csr_to_pbr on the GPU (Backend/graph_link/graph_link/__init__.py:137-156 ->
Backend/graph_link/pbr_matrix/pbr_matrix.cpp:340-477), and pbr_to_csr.

The matrix is built by the HIP builder of libgraphlink_hip.so on the current device and stays
resident there (the returned object is already "on cuda"); host arrays are fetched lazily.
order="reference" (default) reproduces the reference's tile order inside each stripe
bit-for-bit; order="canonical" keeps block columns ascending (no host replay).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import torch

from . import _native
from .pbr import pbr_class


def _as_csr_arrays(mat):
    if isinstance(mat, torch.Tensor):
        if mat.device.type != "cpu":
            raise ValueError("Input tensor must be on CPU")
        mat = sp.csr_matrix((mat.values().numpy(), mat.col_indices().numpy(),
                             mat.crow_indices().numpy()), shape=tuple(mat.shape))
    elif not isinstance(mat, sp.csr_matrix):
        raise TypeError(f"Unsupported type {type(mat)}")
    return mat


def _index_dtype(mat):
    it = np.dtype(mat.indices.dtype)
    return it if it in (np.dtype(np.int32), np.dtype(np.int64)) else np.dtype(np.int64)


def _scalar_dtype(mat):
    st = np.dtype(mat.data.dtype)
    return st if st in (np.dtype(np.float32), np.dtype(np.float64)) else np.dtype(np.float32)


def csr_to_pbr(mat, block_rows: int = 2, block_cols: int = 2, min_nnz_per_block: int = 2, *,
               order: str = "reference", device=None):
    """Build a PBR matrix from a canonical CSR matrix on the GPU. `mat` is a scipy CSR, a
    CPU torch sparse CSR tensor, or a DeviceCSR already on the GPU (no upload)."""
    from .csr import DeviceCSR
    if isinstance(mat, DeviceCSR):
        return _build(mat.shape, mat.nnz, mat.indptr, mat.indices, mat.data,
                      np.dtype(np.int32), np.dtype(np.float32), block_rows, block_cols,
                      min_nnz_per_block, order, mat.device, mat.canonical)
    mat = _as_csr_arrays(mat)
    if order not in ("reference", "canonical"):
        raise ValueError("order must be 'reference' or 'canonical'")
    it, st = _index_dtype(mat), _scalar_dtype(mat)
    if max(mat.shape[0], mat.shape[1], mat.nnz) >= 2**31 - 1:
        raise ValueError("matrix too large for int32 device indices")
    if not mat.has_canonical_format:
        # the builder needs sorted, duplicate-free rows (the reference silently produces
        # misaligned values otherwise)
        mat = mat.copy()
        mat.sum_duplicates()
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    if dev.type == "cuda" and dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    with torch.cuda.device(dev):
        indptr = torch.as_tensor(mat.indptr.astype(np.int32, copy=False)).to(dev)
        indices = torch.as_tensor(mat.indices.astype(np.int32, copy=False)).to(dev)
        vals = torch.as_tensor(np.ascontiguousarray(mat.data.astype(st, copy=False))).to(dev)
    return _build(mat.shape, mat.nnz, indptr, indices, vals, it, st, block_rows, block_cols,
                  min_nnz_per_block, order, dev, True)


def _build(shape, nnz, indptr, indices, vals, it, st, block_rows, block_cols,
           min_nnz_per_block, order, dev, canonical):
    if not canonical:
        raise ValueError("csr_to_pbr needs canonical CSR (sorted, no duplicates)")
    if order not in ("reference", "canonical"):
        raise ValueError("order must be 'reference' or 'canonical'")
    cls = pbr_class(it, st)
    rows, cols = shape
    if block_rows < 1 or block_cols < 1 or block_rows * block_cols > 64:
        raise ValueError(f"block {block_rows}x{block_cols} does not fit a 64-bit code")
    if not 0 <= min_nnz_per_block <= 255:
        raise ValueError("min_nnz_per_block must be in [0, 255]")
    L = _native.lib()
    vb = 8 if st == np.float64 else 4
    with torch.cuda.device(dev):
        stream = _native.stream_handle(dev)
        wsb = L.gl_pbr_build_workspace_bytes(rows, block_rows, nnz)
        ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=dev)
        rem_indptr = torch.empty(rows + 1, dtype=torch.int32, device=dev)
        counts = np.zeros(5, dtype=np.int64)
        _native.check(L.gl_pbr_build_count(rows, cols, nnz, block_rows, block_cols,
                                           int(min_nnz_per_block), indptr.data_ptr(),
                                           indices.data_ptr(), rem_indptr.data_ptr(),
                                           ws.data_ptr(), wsb, counts.ctypes.data, stream),
                      "gl_pbr_build_count")
        T, V, R, D, B = (int(c) for c in counts)
        ns = (rows + block_rows - 1) // block_rows
        vdt = torch.float64 if vb == 8 else torch.float32
        want_blocks = order == "reference" and T > 1
        # canonical order: every output in one device buffer, whose views are cut after the
        # fill pass is queued (while it runs); the reference order replaces the tile arrays
        # afterwards, so it allocates them one by one
        shapes = [(T, torch.int64), (2 * T, torch.int32), (T + 1, torch.int32), (V, vdt),
                  (R, torch.int32), (R, vdt), (ns + 1, torch.int32)]
        if want_blocks:
            outs = [torch.empty(n, dtype=dt, device=dev) for n, dt in shapes]
            ptrs = [o.data_ptr() for o in outs]
            blk_ptr = torch.empty(ns + 1, dtype=torch.int32, device=dev)
            blk_bc = torch.empty(max(D, 1), dtype=torch.int32, device=dev)
            blk_rmin = torch.empty(max(D, 1), dtype=torch.int8, device=dev)
            blk_tile = torch.empty(max(D, 1), dtype=torch.int32, device=dev)
        else:
            starts, total = [], 0
            for n, dt in shapes:
                starts.append(total)
                total += (n * dt.itemsize + 255) & ~255
            buf = torch.empty(max(total, 1), dtype=torch.uint8, device=dev)
            base = buf.data_ptr()
            ptrs = [base + o for o in starts]
            blk_ptr = blk_bc = blk_rmin = blk_tile = None
        p = _native.ptr
        _native.check(L.gl_pbr_build_fill(
            rows, cols, nnz, block_rows, block_cols, int(min_nnz_per_block), T, B,
            indptr.data_ptr(), indices.data_ptr(), vals.data_ptr(), vb, ws.data_ptr(), wsb,
            ptrs[0], ptrs[1], ptrs[2], ptrs[3], rem_indptr.data_ptr(), ptrs[4], ptrs[5],
            ptrs[6], p(blk_ptr), p(blk_bc), p(blk_rmin), p(blk_tile), stream),
            "gl_pbr_build_fill")
        if not want_blocks:
            outs = [buf[o:o + n * dt.itemsize].view(dt) for o, (n, dt) in zip(starts, shapes)]
        codes, coords, offsets, data, rem_col, rem_vals, stripe_ptr = outs
        if want_blocks:
            # block lists to pinned host memory (one stream synchronization), the replay of
            # the reference's container on the host, the permutation back the same way
            hosts = [torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
                     for t in (blk_ptr, blk_bc, blk_rmin, blk_tile)]
            for h, t in zip(hosts, (blk_ptr, blk_bc, blk_rmin, blk_tile)):
                h.copy_(t, non_blocking=True)
            torch.cuda.current_stream(dev).synchronize()
            h_ptr, h_bc, h_rmin, h_tile = (h.numpy() for h in hosts)
            perm_h = torch.empty(T, dtype=torch.int32, pin_memory=True)
            _native.check(L.gl_pbr_reference_perm(rows, cols, block_rows, block_cols,
                                                  it.itemsize, h_ptr.ctypes.data,
                                                  h_bc.ctypes.data, h_rmin.ctypes.data,
                                                  h_tile.ctypes.data, perm_h.data_ptr()),
                          "gl_pbr_reference_perm")
            perm_d = perm_h.to(dev, non_blocking=True)
            codes2 = torch.empty_like(codes)
            coords2 = torch.empty_like(coords)
            offsets2 = torch.empty_like(offsets)
            data2 = torch.empty_like(data)
            pwb = L.gl_pbr_permute_workspace_bytes(T)
            pws = torch.empty(max(pwb, 1), dtype=torch.uint8, device=dev)
            _native.check(L.gl_pbr_permute_tiles(T, perm_d.data_ptr(), codes.data_ptr(),
                                                 coords.data_ptr(), offsets.data_ptr(),
                                                 p(data), vb, codes2.data_ptr(),
                                                 coords2.data_ptr(), offsets2.data_ptr(),
                                                 p(data2), pws.data_ptr(), pwb, stream),
                          "gl_pbr_permute_tiles")
            codes, coords, offsets, data = codes2, coords2, offsets2, data2
        meta = {"codes": codes, "coords": coords, "offsets": offsets, "data": data,
                "rem_indptr": rem_indptr, "rem_col_ind": rem_col, "rem_vals": rem_vals}
        meta["stripe_ptr"] = stripe_ptr  # tiles stay in their stripe in either order
    return cls(rows, cols, block_rows, block_cols, nnz, device_meta=meta)
