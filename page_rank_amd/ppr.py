"""Personalized PageRank drivers.

run_personalized_pagerank  drop-in for Backend/graph_link/graph_link/__init__.py:210-271
run_ppr                    batched seed sets, dangling handling, per-column retirement,
                           warm start, hipGraph-captured loop (one C-ABI call)
normalize_transition_matrix  Backend/graph_link/graph_link/__init__.py:184-207
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp
import torch

from . import _native
from .registry import get_pbr_gpu_meta

_DANGLING = {"none": _native.GL_DANGLING_NONE, "uniform": _native.GL_DANGLING_UNIFORM,
             "personalization": _native.GL_DANGLING_PERSONALIZATION}
_RETIRE = {"together": _native.GL_RETIRE_TOGETHER, "per_column": _native.GL_RETIRE_PER_COLUMN}


def normalize_transition_matrix(A_csr: sp.csr_matrix) -> sp.csr_matrix:
    """Column-stochastic transition matrix: each column divided by its sum (float32);
    dangling columns get a dense uniform 1/N column, as the reference does."""
    N = A_csr.shape[0]
    A = A_csr.tocsc().astype(np.float32)
    col_sums = np.asarray(A.sum(axis=0), dtype=np.float32).ravel()
    inv = np.zeros_like(col_sums)
    np.divide(1.0, col_sums, out=inv, where=col_sums > 0)
    A_norm = A.multiply(inv).tocsr()
    dangling = np.where(col_sums == 0)[0]
    if dangling.size > 0:
        r = np.tile(np.arange(N, dtype=np.int32), dangling.size)
        c = np.repeat(dangling.astype(np.int32), N)
        v = np.full(N * dangling.size, 1.0 / N, dtype=np.float32)
        A_norm = A_norm + sp.csr_matrix((v, (r, c)), shape=(N, N))
    return A_norm


@dataclass
class PPRResult:
    X: torch.Tensor              # [N, K] float32 on the device
    iterations: np.ndarray       # [K] int32: iteration at which each column stopped
    converged: np.ndarray        # [K] bool
    history: np.ndarray          # [total_iterations, K] L1 residual (0 once a column stopped)
    total_iterations: int
    feature_passes: int = 0      # passes done before the run (a warm start's feature products)


def _seed_arrays(seed_sets, weights=None):
    ptr = [0]
    idx = []
    w = []
    for k, s in enumerate(seed_sets):
        s = np.atleast_1d(np.asarray(s, dtype=np.int64))
        if s.size == 0:
            raise ValueError(f"seed set {k} is empty")
        idx.append(s)
        if weights is None:
            w.append(np.full(s.size, 1.0 / s.size, dtype=np.float32))
        else:
            ws = np.asarray(weights[k], dtype=np.float32)
            if ws.shape != s.shape:
                raise ValueError("weights must match seed sets")
            w.append(ws)
        ptr.append(ptr[-1] + s.size)
    return (np.asarray(ptr, dtype=np.int32), np.concatenate(idx).astype(np.int32),
            np.concatenate(w).astype(np.float32))


def run_ppr(pbr_mat, seed_sets, *, alpha=0.85, tol=1e-6, max_iter=100, dangling=None,
            dangling_mode="none", retire="per_column", x0=None, weights=None,
            graph_iters=8, kw=0, grid_mult=0, extrapolate_every=0,
            return_history=True, x0_in_place=False, single_workgroup=None) -> PPRResult:
    """Batched PPR, one column per seed set (a list of node-index lists).
    kw / grid_mult: RHS columns per launch and workgroups per CU (0 = library rule;
    kw="auto" = measured by the autotuner, cached per shape).
    extrapolate_every: 0 = plain power iteration; P >= 2 = after every P passes, columns whose
    residual shrinks by a steady ratio rho in [0.5, 0.97] jump ahead along their last update,
    X += rho / (1 - rho) (X_t - X_{t-1}) (Aitken extrapolation of the dominant error mode).
    The stopping rule is unchanged, so the result meets the same residual bound.
    x0_in_place: x0 (a contiguous float32 [N][K] tensor on the matrix's device) becomes the
    result's buffer, with no copy; the caller's start is overwritten.
    single_workgroup: None = library choice (the whole run in one launch of one workgroup for
    small graphs: K <= 64 and nnz * K <= 2^13), True = always (K <= 64), False = never."""
    meta = get_pbr_gpu_meta(pbr_mat)
    if meta is None:
        raise RuntimeError("PBRMatrix is not on GPU. Call .to('cuda') first.")
    if meta["data"].dtype != torch.float32:
        raise TypeError("run_ppr computes in float32; a float64 matrix runs through "
                        "run_personalized_pagerank (the reference loop in FP64) or pbr_matmul")
    if pbr_mat.rows != pbr_mat.cols:
        raise ValueError("PPR needs a square matrix")
    device = meta["codes"].device
    N = pbr_mat.rows
    ptr, idx, w = _seed_arrays(seed_sets, weights)
    if idx.size and (idx.min() < 0 or idx.max() >= N):
        raise ValueError("seed node out of range")
    K = len(seed_sets)
    if kw == "auto":
        from .autotune import tune_ppr
        t = tune_ppr(pbr_mat, K)
        kw = t["kw"]
        if not grid_mult:
            grid_mult = t["grid_mult"]
    dmode = _DANGLING[dangling_mode]
    dmask = None
    if dmode != _native.GL_DANGLING_NONE:
        if dangling is None:
            raise ValueError(f"dangling_mode={dangling_mode!r} needs the dangling mask")
        dmask = torch.as_tensor(np.asarray(dangling, dtype=np.uint8)).to(device)
    if x0 is not None:
        given = x0
        x0 = torch.as_tensor(x0, dtype=torch.float32, device=device).contiguous()
        if x0.shape != (N, K):
            raise ValueError(f"x0 must be [{N}, {K}]")
        if x0_in_place and x0 is not given:
            raise ValueError("x0_in_place needs a contiguous float32 tensor on the device")
    if x0_in_place:
        if x0 is None:
            raise ValueError("x0_in_place needs x0")
        X = x0  # the library runs in place when the start is the output buffer
    else:
        X = torch.empty((N, K), dtype=torch.float32, device=device)
    iters = torch.zeros(K, dtype=torch.int32, device=device)
    conv = torch.zeros(K, dtype=torch.int32, device=device)
    hist = (torch.zeros((max_iter, K), dtype=torch.float32, device=device)
            if return_history else None)
    L = _native.lib()
    wsb = L.gl_ppr_workspace_bytes(N, pbr_mat.block_rows, K, int(idx.size))
    if wsb < 0:
        raise ValueError("invalid PPR sizes")
    ws = torch.empty(wsb, dtype=torch.uint8, device=device)
    opts = _native.GlOpts(alpha=alpha, tol=tol, max_iter=int(max_iter), dangling_mode=dmode,
                          retire_mode=_RETIRE[retire], graph_iters=int(graph_iters), kw=int(kw),
                          grid_mult=int(grid_mult), extrapolate_every=int(extrapolate_every),
                          single_wg=0 if single_workgroup is None else
                          (1 if single_workgroup else -1))
    total = np.zeros(1, dtype=np.int32)
    with torch.cuda.device(device):
        rc = L.gl_ppr_run_i32_f32(
            N, pbr_mat.cols, pbr_mat.block_rows, pbr_mat.block_cols, meta["codes"].numel(),
            meta["data"].numel(), meta["rem_col_ind"].numel(), meta["codes"].data_ptr(), meta["coords"].data_ptr(), meta["offsets"].data_ptr(),
            meta["data"].data_ptr(), meta["stripe_ptr"].data_ptr(),
            meta["rem_indptr"].data_ptr(), meta["rem_col_ind"].data_ptr(),
            meta["rem_vals"].data_ptr(), K, ptr.ctypes.data, idx.ctypes.data, w.ctypes.data,
            _native.ptr(dmask), _native.ptr(x0), opts, X.data_ptr(), iters.data_ptr(),
            conv.data_ptr(), _native.ptr(hist), ws.data_ptr(), wsb, total.ctypes.data,
            _native.stream_handle(device))
    _native.check(rc, "gl_ppr_run_i32_f32")
    T = int(total[0])
    return PPRResult(X=X, iterations=iters.cpu().numpy(), converged=conv.cpu().numpy() != 0,
                     history=(hist[:T].cpu().numpy() if hist is not None
                              else np.zeros((0, K), np.float32)),
                     total_iterations=T)


def run_personalized_pagerank(pbr_mat, source_nodes: torch.Tensor, alpha: float = 0.85,
                              max_iterations: int = 100, tolerance: float = 1e-6,
                              return_history: bool = False):
    """X_{t+1} = alpha * A @ X_t + (1 - alpha) * e_s for one seed per column, all columns
    stopping together when the max L1 update is below tolerance. Returns
    (X, iterations_run, converged[, history]) like the reference."""
    meta = get_pbr_gpu_meta(pbr_mat)
    if meta is None:
        raise RuntimeError("PBRMatrix is not on GPU. Call .to('cuda') first.")
    if meta["data"].dtype == torch.float64:
        return _reference_loop_f64(pbr_mat, meta, source_nodes, alpha, max_iterations,
                                   tolerance, return_history)
    src = torch.as_tensor(source_nodes).to(torch.int32).cpu().numpy().astype(np.int64)
    res = run_ppr(pbr_mat, [[int(s)] for s in src], alpha=alpha, tol=tolerance,
                  max_iter=max_iterations, dangling_mode="none", retire="together",
                  return_history=True)
    iterations = int(res.iterations.max()) if res.iterations.size else 0
    converged = bool(res.converged.all())
    if return_history:
        history = [float(h) for h in res.history.max(axis=1)]
        return res.X, iterations, converged, history
    return res.X, iterations, converged


def _reference_loop_f64(pbr_mat, meta, source_nodes, alpha, max_iterations, tolerance,
                        return_history):
    """The reference loop for a float64 matrix (__init__.py:210-271 with its double
    kernels): one-hot init, then per iteration Y = A X, X = alpha Y + (1 - alpha) e_s with
    per-column L1 residuals, and a host check of their maximum."""
    from .ops import spmm_into
    dev = meta["codes"].device
    src = torch.as_tensor(source_nodes)
    if src.dim() != 1 or src.numel() == 0:
        raise ValueError("source_nodes must be a non-empty 1-D tensor")
    src_h = src.to(torch.int64).cpu()
    N, K = pbr_mat.rows, src.numel()
    if pbr_mat.rows != pbr_mat.cols:
        raise ValueError("PPR needs a square matrix")
    if int(src_h.min()) < 0 or int(src_h.max()) >= N:
        raise ValueError(f"source node out of range [0, {N})")
    src_d = src_h.to(torch.int32).to(dev)
    L = _native.lib()
    X = torch.empty((N, K), dtype=torch.float64, device=dev)
    Y = torch.empty_like(X)
    errors = torch.empty(K, dtype=torch.float64, device=dev)
    wsb = L.gl_ppr_update_f64_workspace_bytes(N, K)
    ws = torch.empty(max(wsb, 8), dtype=torch.uint8, device=dev)
    st = _native.stream_handle(dev)
    with torch.cuda.device(dev):
        _native.check(L.gl_ppr_init_f64(X.data_ptr(), src_d.data_ptr(), N, K, st),
                      "gl_ppr_init_f64")
        history = []
        converged = False
        iterations = max_iterations
        for t in range(max_iterations):
            spmm_into(pbr_mat, X, Y)
            _native.check(L.gl_ppr_update_normalized_f64(
                Y.data_ptr(), X.data_ptr(), src_d.data_ptr(), float(alpha), N, K,
                errors.data_ptr(), ws.data_ptr(), wsb, st), "gl_ppr_update_normalized_f64")
            err = float(errors.max().item())
            history.append(err)
            if err < tolerance:
                iterations = t + 1
                converged = True
                break
    if return_history:
        return X, iterations, converged, history
    return X, iterations, converged
