"""ORACLE — test infrastructure only. FP64 Personalized-PageRank restatements.

reference_loop   Backend/graph_link/graph_link/__init__.py:210-271 with the kernels it calls,
                 init_ppr_sources_kernel / ppr_update_normalized_kernel
                 (Backend/graph_link/kernels/ppr_kernel.cu:130-216): one-hot start,
                 X <- alpha*A X + (1-alpha) e_s, all columns stop together when
                 max_k L1 residual < tol.
power            the batched generalisation the GPU path implements: seed-set columns,
                 dangling mass (none / uniform / back to the personalization, the latter being
                 batched_missing_mass + batched_ppr_update, ppr_kernel.cu:13-117), per-column
                 retirement, optional warm start.
nx_pagerank      networkx's _pagerank_scipy semantics (networkx 3.4.2
                 algorithms/link_analysis/pagerank_alg.py:459-500): uniform start, stop at
                 L1 < N*tol, result normalised to sum 1. Used by the reference's accuracy test
                 (Tests/test_ppr_accuracy.py:28-40).
solve            exact fixed point of the same equation by sparse LU (small N only).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def personalization_matrix(n, seed_sets, weights=None):
    """Dense [n, K] float64; column k is uniform over seed_sets[k] (duplicates add up)."""
    K = len(seed_sets)
    P = np.zeros((n, K), dtype=np.float64)
    for k, s in enumerate(seed_sets):
        s = np.asarray(s, dtype=np.int64)
        w = (np.full(s.size, 1.0 / s.size) if weights is None
             else np.asarray(weights[k], dtype=np.float64))
        np.add.at(P[:, k], s, w)
    return P


def reference_loop(A, sources, alpha=0.85, max_iterations=100, tolerance=1e-6,
                   dtype=np.float64):
    """Restatement of run_personalized_pagerank (column-stochastic A, one seed per column).
    Returns (X, iterations, converged, history)."""
    A = sp.csr_matrix(A, dtype=dtype)
    n = A.shape[0]
    sources = np.asarray(sources, dtype=np.int64)
    K = sources.size
    X = np.zeros((n, K), dtype=dtype)
    X[sources, np.arange(K)] = 1.0
    E = np.zeros_like(X)
    E[sources, np.arange(K)] = 1.0 - alpha
    history = []
    iterations, converged = max_iterations, False
    for t in range(max_iterations):
        Xn = alpha * (A @ X) + E
        err = np.abs(Xn - X).sum(axis=0)
        X = Xn
        history.append(float(err.max()))
        if err.max() < tolerance:
            iterations, converged = t + 1, True
            break
    return X, iterations, converged, history


def power(A, P, alpha=0.85, tol=1e-6, max_iter=100, dangling=None, dangling_mode="none",
          retire="together", x0=None, progress=None):
    """FP64 batched PPR: X_{t+1} = alpha*A X_t + alpha*d_k*w + (1-alpha)*P, d_k = sum of
    X_t over dangling rows, w = 1/N (uniform) or P[:,k] (personalization).
    `progress(t, err)` (optional) is called after every iteration with its residuals.
    Returns (X, iterations[K], converged[K], history[T][K])."""
    A = sp.csr_matrix(A, dtype=np.float64)
    n = A.shape[0]
    P = np.asarray(P, dtype=np.float64)
    K = P.shape[1]
    X = P.copy() if x0 is None else np.asarray(x0, dtype=np.float64).copy()
    dmask = None if dangling is None else np.asarray(dangling, dtype=bool)
    active = np.ones(K, dtype=bool)
    iters = np.zeros(K, dtype=np.int64)
    conv = np.zeros(K, dtype=bool)
    hist = []
    for t in range(max_iter):
        Xn = alpha * (A @ X) + (1.0 - alpha) * P
        if dangling_mode != "none":
            d = X[dmask].sum(axis=0)
            if dangling_mode == "uniform":
                Xn += alpha * d[None, :] / n
            elif dangling_mode == "personalization":
                Xn += alpha * d[None, :] * P
            else:
                raise ValueError(dangling_mode)
        err = np.abs(Xn - X).sum(axis=0)
        err[~active] = 0.0
        hist.append(err.copy())
        if progress is not None:
            progress(t + 1, err)
        X[:, active] = Xn[:, active]
        if retire == "together":
            done = np.full(K, err[active].max() < tol) & active
        else:
            done = (err < tol) & active
        if t + 1 >= max_iter:
            stop = active.copy()
        else:
            stop = done
        iters[stop] = t + 1
        conv[done] = True
        active &= ~stop
        if not active.any():
            break
    return X, iters, conv, np.array(hist)


def nx_pagerank(A_col, p, alpha=0.85, tol=1e-6, max_iter=100, dangling_weights=None,
                return_iterations=False):
    """networkx _pagerank_scipy semantics on a column-oriented weighted adjacency A_col
    (A_col[target, source] = weight, any non-negative weights). Returns x normalised to 1
    (and the iteration count with return_iterations). Raises RuntimeError (networkx:
    PowerIterationFailedConvergence) after max_iter."""
    A_col = sp.csc_matrix(A_col, dtype=np.float64)
    n = A_col.shape[0]
    S = np.asarray(A_col.sum(axis=0)).ravel()         # out-weight of each source
    Q = np.divide(1.0, S, out=np.zeros_like(S), where=S != 0)
    M = A_col @ sp.diags(Q)                           # column-stochastic except dangling
    x = np.full(n, 1.0 / n)
    p = np.asarray(p, dtype=np.float64)
    p = p / p.sum()
    dw = p if dangling_weights is None else np.asarray(dangling_weights, np.float64)
    dw = dw / dw.sum()
    is_dangling = S == 0
    for it in range(max_iter):
        xlast = x
        x = alpha * (M @ x + x[is_dangling].sum() * dw) + (1 - alpha) * p
        err = np.abs(x - xlast).sum()
        if err < n * tol:
            return (x / x.sum(), it + 1) if return_iterations else x / x.sum()
    raise RuntimeError(f"power iteration failed to converge within {max_iter} iterations")


def solve(A, p, alpha=0.85, dangling=None, dangling_mode="none"):
    """Exact solution of (I - alpha*A - alpha*w d^T) x = (1-alpha) p (small N)."""
    A = sp.csc_matrix(A, dtype=np.float64)
    n = A.shape[0]
    p = np.asarray(p, dtype=np.float64)
    M = sp.identity(n, format="csc") - alpha * A
    if dangling_mode != "none":
        d = np.asarray(dangling, dtype=np.float64)
        w = np.full(n, 1.0 / n) if dangling_mode == "uniform" else p
        M = M - alpha * sp.csc_matrix(np.outer(w, d))
    return spla.spsolve(M.tocsc(), (1 - alpha) * p)


def column_stochastic(A, dangling_mode="none"):
    """Divide each column by its sum (float64), leaving dangling columns empty.
    Returns (A_norm csr, dangling bool mask)."""
    A = sp.csc_matrix(A, dtype=np.float64)
    s = np.asarray(A.sum(axis=0)).ravel()
    inv = np.divide(1.0, s, out=np.zeros_like(s), where=s != 0)
    return sp.csr_matrix(A @ sp.diags(inv)), s == 0
