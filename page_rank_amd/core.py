"""The reference's native per-kernel PPR functions under their own names and signatures
(the graph_link_core names bound at Backend/graph_link/bindings/bindings_kernels.cu:148-161),
over the C ABI of libgraphlink_hip.so. A caller that drives its own host loop the way
Backend/graph_link/graph_link/__init__.py:240-260 does can switch by replacing
``graph_link_core`` with this module:

    init_ppr_cuda_float(X, source_nodes, N, K)
    ppr_update_normalized_cuda_float(Y, X, source_nodes, alpha, N, K, errors)
    missing_mass_cuda_float(Y, col_sums, N, K)
    ppr_update_cuda_float(Y, curr_X, next_X, source_nodes, col_sums, damping, N, K, errors)

and the ``_double`` variants. Tensors are device tensors, contiguous [N, K] (X, Y) or [K]
(errors, col_sums, source_nodes int32). The residuals and column sums are summed in a fixed
order, so repeated calls agree bit for bit (the reference adds float atomics). Work runs on
torch's current stream of the tensors' device.
"""
from __future__ import annotations

import torch

from . import _native

_TYPES = {torch.float32: "f32", torch.float64: "f64"}


def _check_dense(name, t, N, K, dtype):
    if not t.is_cuda or t.dtype != dtype or not t.is_contiguous() or t.numel() != N * K:
        raise ValueError(f"{name} must be a contiguous {dtype} device tensor of {N}x{K} "
                         f"elements, got {tuple(t.shape)} {t.dtype} on {t.device}")


def _check_src(src, K):
    if src.dtype != torch.int32 or not src.is_cuda or src.numel() != K:
        raise ValueError(f"source_nodes must be an int32 device tensor of {K} entries")
    if not src.is_contiguous():
        raise ValueError("source_nodes must be contiguous")


def _workspace(suffix, N, K, device):
    L = _native.lib()
    wsb = getattr(L, f"gl_ppr_update_{suffix}_workspace_bytes")(N, K)
    if wsb < 0:
        raise ValueError(f"invalid sizes N={N} K={K}")
    return torch.empty(max(wsb, 1), dtype=torch.uint8, device=device), wsb


def _init(X, source_nodes, N, K, dtype):
    _check_dense("X", X, N, K, dtype)
    _check_src(source_nodes, K)
    sfx = _TYPES[dtype]
    with torch.cuda.device(X.device):
        rc = getattr(_native.lib(), f"gl_ppr_init_{sfx}")(
            X.data_ptr(), source_nodes.data_ptr(), N, K, _native.stream_handle(X.device))
    _native.check(rc, f"gl_ppr_init_{sfx}")


def _update_normalized(Y, X, source_nodes, alpha, N, K, errors, dtype):
    _check_dense("Y", Y, N, K, dtype)
    _check_dense("X", X, N, K, dtype)
    _check_dense("errors", errors, 1, K, dtype)
    _check_src(source_nodes, K)
    sfx = _TYPES[dtype]
    ws, wsb = _workspace(sfx, N, K, X.device)
    with torch.cuda.device(X.device):
        rc = getattr(_native.lib(), f"gl_ppr_update_normalized_{sfx}")(
            Y.data_ptr(), X.data_ptr(), source_nodes.data_ptr(), float(alpha), N, K,
            errors.data_ptr(), ws.data_ptr(), wsb, _native.stream_handle(X.device))
    _native.check(rc, f"gl_ppr_update_normalized_{sfx}")


def _missing_mass(Y, col_sums, N, K, dtype):
    _check_dense("Y", Y, N, K, dtype)
    _check_dense("col_sums", col_sums, 1, K, dtype)
    sfx = _TYPES[dtype]
    ws, wsb = _workspace(sfx, N, K, Y.device)
    with torch.cuda.device(Y.device):
        rc = getattr(_native.lib(), f"gl_ppr_missing_mass_{sfx}")(
            Y.data_ptr(), col_sums.data_ptr(), N, K, ws.data_ptr(), wsb,
            _native.stream_handle(Y.device))
    _native.check(rc, f"gl_ppr_missing_mass_{sfx}")


def _update(Y, curr_X, next_X, source_nodes, col_sums, damping, N, K, errors, dtype):
    for name, t in (("Y", Y), ("curr_X", curr_X), ("next_X", next_X)):
        _check_dense(name, t, N, K, dtype)
    _check_dense("col_sums", col_sums, 1, K, dtype)
    _check_dense("errors", errors, 1, K, dtype)
    _check_src(source_nodes, K)
    sfx = _TYPES[dtype]
    ws, wsb = _workspace(sfx, N, K, Y.device)
    with torch.cuda.device(Y.device):
        rc = getattr(_native.lib(), f"gl_ppr_update_{sfx}")(
            Y.data_ptr(), curr_X.data_ptr(), next_X.data_ptr(), source_nodes.data_ptr(),
            col_sums.data_ptr(), float(damping), N, K, errors.data_ptr(), ws.data_ptr(), wsb,
            _native.stream_handle(Y.device))
    _native.check(rc, f"gl_ppr_update_{sfx}")


def init_ppr_cuda_float(X, source_nodes, N, K):
    """X = one-hot columns (X[source_nodes[k], k] = 1, every other element 0)."""
    _init(X, source_nodes, N, K, torch.float32)


def init_ppr_cuda_double(X, source_nodes, N, K):
    _init(X, source_nodes, N, K, torch.float64)


def ppr_update_normalized_cuda_float(Y, X, source_nodes, alpha, N, K, errors):
    """X <- alpha Y + (1 - alpha) e_src in place; errors[k] = L1 norm of the update."""
    _update_normalized(Y, X, source_nodes, alpha, N, K, errors, torch.float32)


def ppr_update_normalized_cuda_double(Y, X, source_nodes, alpha, N, K, errors):
    _update_normalized(Y, X, source_nodes, alpha, N, K, errors, torch.float64)


def missing_mass_cuda_float(Y, col_sums, N, K):
    """col_sums[k] = sum_i Y[i, k]."""
    _missing_mass(Y, col_sums, N, K, torch.float32)


def missing_mass_cuda_double(Y, col_sums, N, K):
    _missing_mass(Y, col_sums, N, K, torch.float64)


def ppr_update_cuda_float(Y, curr_X, next_X, source_nodes, col_sums, damping, N, K, errors):
    """next_X = damping Y + [i == src] (damping (1 - col_sums) + 1 - damping);
    errors[k] += L1 norm of next_X - curr_X."""
    _update(Y, curr_X, next_X, source_nodes, col_sums, damping, N, K, errors, torch.float32)


def ppr_update_cuda_double(Y, curr_X, next_X, source_nodes, col_sums, damping, N, K, errors):
    _update(Y, curr_X, next_X, source_nodes, col_sums, damping, N, K, errors, torch.float64)
