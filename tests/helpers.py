"""Shared test helpers: seeded inputs in the style of the reference's tests."""
import numpy as np
import scipy.sparse as sp


def random_csr(n, density, dtype=np.float32, seed=42, m=None):
    """Tests/test_spmm.py:13-15 generator (scipy.sparse.random, seeded)."""
    m = n if m is None else m
    return sp.random(n, m, density=density, format="csr", dtype=dtype,
                     random_state=np.random.default_rng(seed).integers(2**31))


def random_x(n, k, seed=0):
    """Tests/test_spmm.py:18-19."""
    return np.random.default_rng(seed).random((n, k)).astype(np.float32)


def blocky_csr(n, block, tiles_per_row, fill_lo, fill_hi, scatter, seed=0, band=4):
    """Small block-structured matrix: dense-ish tiles near the diagonal plus scattered nnz
    (a miniature of the 1M x 1M SpMM workload)."""
    rng = np.random.default_rng(seed)
    nb = (n + block - 1) // block
    rows, cols = [], []
    for br in range(nb):
        bcs = np.unique(np.clip(br + rng.integers(-band, band + 1, tiles_per_row), 0, nb - 1))
        for bc in bcs:
            k = int(rng.integers(fill_lo, fill_hi + 1))
            pos = rng.choice(block * block, size=min(k, block * block), replace=False)
            r = br * block + pos // block
            c = bc * block + pos % block
            ok = (r < n) & (c < n)
            rows.append(r[ok])
            cols.append(c[ok])
    rows.append(rng.integers(0, n, scatter))
    cols.append(rng.integers(0, n, scatter))
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    key = np.unique(r.astype(np.int64) * n + c)
    r, c = key // n, key % n
    v = rng.random(r.size).astype(np.float32)
    return sp.csr_matrix((v, (r, c)), shape=(n, n))
