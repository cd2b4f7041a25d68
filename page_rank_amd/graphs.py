"""Seeded synthetic workloads of the PPR hot path (host generators, numpy).

sbm_edges            seeded stochastic block model (10,000 nodes in 50 communities of 200,
                     p_in 0.04, p_out 0.0002, directed, no self loops)
spmm_matrix          the 1M x 1M block-structured SpMM matrix: every block row holds
                     tiles_per_row distinct 8x8 tiles whose block columns are uniform within
                     +-band blocks of the diagonal (clipped to the matrix), U{fill} nonzeros at
                     distinct uniform positions per tile, plus `scatter` uniformly scattered
                     nonzeros; values U(0,1), float32, canonical CSR
clustered_edges      community-clustered power-law digraph: out-degree Pareto(shape, min)
                     rescaled so the degrees sum to the requested edge count, `intra` of the
                     edges inside the source's community (communities are contiguous node
                     ranges, i.e. already relabelled by community), the rest uniform
transition_matrix    column-stochastic A[target, source] = multiplicity / out-degree (float32
                     CSR, dangling columns left empty) and the dangling mask
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def sbm_edges(n=10_000, community=200, p_in=0.04, p_out=0.0002, seed=0):
    rng = np.random.default_rng(seed)
    nc = n // community
    src, dst = [], []
    for c in range(nc):
        lo = c * community
        m = rng.random((community, community)) < p_in
        np.fill_diagonal(m, False)
        s, d = np.nonzero(m)
        src.append(s + lo)
        dst.append(d + lo)
    # inter-community pairs: Binomial count, then uniform pairs rejecting intra ones
    n_pairs = n * (n - community)
    k = rng.binomial(n_pairs, p_out)
    s = rng.integers(0, n, 2 * k + 100)
    d = rng.integers(0, n, 2 * k + 100)
    ok = (s // community) != (d // community)
    s, d = s[ok][:k], d[ok][:k]
    key = np.unique(s.astype(np.int64) * n + d)
    src.append(key // n)
    dst.append(key % n)
    return np.concatenate(src).astype(np.int64), np.concatenate(dst).astype(np.int64)


def _distinct_offsets(rng, lo, hi, k):
    """k distinct integers per row drawn uniformly from [lo[i], hi[i]] (inclusive)."""
    nrow = lo.size
    out = np.empty((nrow, k), dtype=np.int64)
    todo = np.arange(nrow)
    while todo.size:
        span = (hi[todo] - lo[todo] + 1)[:, None]
        cand = lo[todo, None] + (rng.random((todo.size, 2 * k)) * span).astype(np.int64)
        cand.sort(axis=1)
        dup = np.zeros_like(cand, dtype=bool)
        dup[:, 1:] = cand[:, 1:] == cand[:, :-1]
        cand = np.where(dup, np.iinfo(np.int64).max, cand)
        cand.sort(axis=1)
        good = cand[:, k - 1] != np.iinfo(np.int64).max
        # keep a uniform k-subset of the distinct candidates: shuffle then take k smallest
        sel = todo[good]
        c = cand[good]
        nd = (c != np.iinfo(np.int64).max).sum(axis=1)
        keys = rng.random(c.shape)
        keys[np.arange(c.shape[1])[None, :] >= nd[:, None]] = 2.0
        idx = np.argsort(keys, axis=1)[:, :k]
        out[sel] = np.take_along_axis(c, idx, axis=1)
        todo = todo[~good]
    return out


def spmm_matrix(n=1_000_000, block=8, tiles_per_row=8, band=512, fill=(4, 40),
                scatter=5_000_000, seed=0):
    rng = np.random.default_rng(seed)
    nb = (n + block - 1) // block
    brow = np.arange(nb)
    lo = np.maximum(0, brow - band)
    hi = np.minimum(nb - 1, brow + band)
    bcols = _distinct_offsets(rng, lo, hi, tiles_per_row)          # [nb, tiles]
    T = nb * tiles_per_row
    t_br = np.repeat(brow, tiles_per_row)
    t_bc = bcols.ravel()
    cnt = rng.integers(fill[0], fill[1] + 1, T)
    P = block * block
    rows_l, cols_l = [], []
    chunk = 1 << 17
    for a in range(0, T, chunk):
        b = min(T, a + chunk)
        keys = rng.random((b - a, P), dtype=np.float32)
        order = np.argsort(keys, axis=1)
        take = np.arange(P)[None, :] < cnt[a:b, None]
        pos = order[take]
        tid = np.repeat(np.arange(a, b), cnt[a:b])
        rows_l.append(t_br[tid] * block + pos // block)
        cols_l.append(t_bc[tid] * block + pos % block)
    rows_l.append(rng.integers(0, n, scatter))
    cols_l.append(rng.integers(0, n, scatter))
    r = np.concatenate(rows_l)
    c = np.concatenate(cols_l)
    ok = (r < n) & (c < n)
    key = np.unique(r[ok].astype(np.int64) * n + c[ok])
    rows = key // n
    cols = (key % n).astype(np.int32)
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=n), out=indptr[1:])
    vals = rng.random(key.size, dtype=np.float32)
    A = sp.csr_matrix((vals, cols, indptr.astype(np.int32)), shape=(n, n))
    A.has_sorted_indices = True
    return A


def clustered_edges(n=4_000_000, m=64_000_000, communities=4_000, intra=0.8,
                    pareto_shape=2.1, min_degree=4, seed=0):
    rng = np.random.default_rng(seed)
    d = (rng.pareto(pareto_shape, n) + 1.0) * min_degree
    d *= m / d.sum()  # rescale the Pareto draw to the requested edge count
    deg = np.floor(d + rng.random(n)).astype(np.int64)
    diff = m - int(deg.sum())
    if diff:
        fix = rng.choice(n, abs(diff), replace=True)
        np.add.at(deg, fix, 1 if diff > 0 else -1)
        deg = np.maximum(deg, 0)
    deg = np.minimum(deg, n - 1)
    src = np.repeat(np.arange(n, dtype=np.int64), deg)
    size = n // communities
    comm = np.minimum(src // size, communities - 1)
    lo = comm * size
    hi = np.where(comm == communities - 1, n, lo + size)
    inside = rng.random(src.size, dtype=np.float32) < intra
    u = rng.random(src.size)
    dst = np.where(inside, lo + (u * (hi - lo)).astype(np.int64), (u * n).astype(np.int64))
    return src, dst


def transition_matrix(src, dst, n):
    """Column-stochastic float32 CSR A[dst, src] (duplicate edges add up) + dangling mask."""
    outdeg = np.bincount(src, minlength=n).astype(np.float64)
    key = np.asarray(dst, np.int64) * n + np.asarray(src, np.int64)
    key, mult = np.unique(key, return_counts=True)
    rows = key // n
    cols = key % n
    vals = (mult / outdeg[cols]).astype(np.float32)
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=n), out=indptr[1:])
    A = sp.csr_matrix((vals, cols.astype(np.int32), indptr.astype(np.int32)), shape=(n, n))
    A.has_sorted_indices = True
    return A, outdeg == 0


def seed_sets(n, k, lo, hi, seed=0):
    """k seed sets of sizes uniform in [lo, hi], distinct nodes inside each set."""
    rng = np.random.default_rng(seed)
    return [np.sort(rng.choice(n, int(rng.integers(lo, hi + 1)), replace=False))
            for _ in range(k)]
