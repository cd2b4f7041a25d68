"""CPU checks of gl_pbr_reference_perm, the host replay of the reference builder's
std::unordered_map emission order (Backend/graph_link/pbr_matrix/pbr_matrix.cpp:374-428),
against the oracle's builder, which runs the container itself (oracle/pbr_oracle.cpp).

The block lists are built here from the CSR in numpy (every distinct block of a stripe in
ascending block column, its first row, its canonical tile index or -1), as the GPU builder's
fill pass writes them; the expected permutation maps the oracle's reference-order tiles to
its canonical-order tiles. Cases cover stripes that grow the table (rehashes inside a
stripe, and later stripes with fewer blocks), empty stripes, a partial last stripe, int64
keys, and int32 keys past 2^31 (they wrap, as in the reference).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import pbr as opbr
from page_rank_amd import _native
from tests.helpers import blocky_csr, random_csr


def block_lists(A, br, bc, min_nnz):
    n_rows = A.shape[0]
    ns = (n_rows + br - 1) // br
    ptr, bcs, rmins, tiles = [0], [], [], []
    t = 0
    for s in range(ns):
        first, cnt = {}, {}
        for r in range(s * br, min((s + 1) * br, n_rows)):
            for c in A.indices[A.indptr[r]:A.indptr[r + 1]]:
                b = int(c) // bc
                first.setdefault(b, r - s * br)
                cnt[b] = cnt.get(b, 0) + 1
        for b in sorted(first):
            bcs.append(b)
            rmins.append(first[b])
            if cnt[b] >= min_nnz:
                tiles.append(t)
                t += 1
            else:
                tiles.append(-1)
        ptr.append(len(bcs))
    return (np.array(ptr, np.int32), np.array(bcs, np.int32), np.array(rmins, np.int8),
            np.array(tiles, np.int32), t)


def expected_perm(A, br, bc, min_nnz):
    ref = opbr.csr_matrix_to_pbr(A, br, bc, min_nnz, "reference")
    can = opbr.csr_matrix_to_pbr(A, br, bc, min_nnz, "canonical")
    pos = {(int(r), int(c)): i for i, (r, c) in
           enumerate(zip(can.block_coords[0::2], can.block_coords[1::2]))}
    return np.array([pos[(int(r), int(c))] for r, c in
                     zip(ref.block_coords[0::2], ref.block_coords[1::2])], np.int32)


def run_perm(A, br, bc, min_nnz, index_bytes):
    ptr, bcs, rmins, tiles, T = block_lists(A, br, bc, min_nnz)
    perm = np.full(T, -7, np.int32)
    L = _native.lib()
    rc = L.gl_pbr_reference_perm(A.shape[0], A.shape[1], br, bc, index_bytes,
                                 ptr.ctypes.data, bcs.ctypes.data, rmins.ctypes.data,
                                 tiles.ctypes.data, perm.ctypes.data)
    _native.check(rc, "gl_pbr_reference_perm")
    return perm


def growing_stripes(n_cols, seed):
    """Stripes whose distinct-block counts rise and fall (1, 3, 9, 27, ..., then small and
    large again), so the table rehashes inside stripes and is reused by smaller ones."""
    rng = np.random.default_rng(seed)
    sizes = [1, 3, 9, 27, 5, 81, 2, 0, 150, 40, 243, 7, 300, 12]
    rows, cols = [], []
    for s, k in enumerate(sizes):
        bcs = rng.choice(n_cols // 4, size=k, replace=False)
        for b in bcs:
            r = s * 4 + int(rng.integers(0, 4))
            rows.append(r)
            cols.append(int(b) * 4 + int(rng.integers(0, 4)))
        if k:  # a second entry in some blocks so that min_nnz 2 keeps a few tiles
            rows.append(s * 4)
            cols.append(int(bcs[0]) * 4 + 3 if rows[-1] != s * 4 else int(bcs[0]) * 4 + 3)
    A = sp.csr_matrix((np.ones(len(rows), np.float32), (rows, cols)),
                      shape=(len(sizes) * 4 + 3, n_cols))
    A.sum_duplicates()
    A.sort_indices()
    return A


@pytest.mark.parametrize("case", [
    dict(kind="blocky", n=2000, block=8, mn=1, seed=1, band=40),
    dict(kind="blocky", n=4000, block=4, mn=2, seed=2, band=300),
    dict(kind="blocky", n=3001, block=2, mn=1, seed=3, band=900),
    dict(kind="blocky", n=1500, block=8, mn=4, seed=4, band=180),
    dict(kind="random", n=3000, block=8, mn=2, seed=5, density=0.01),
    dict(kind="random", n=1000, block=3, mn=1, seed=6, density=0.03),
    dict(kind="growing", n_cols=40000, block=4, mn=1, seed=7),
    dict(kind="growing", n_cols=9000, block=4, mn=2, seed=8),
])
@pytest.mark.parametrize("index_bytes", [4, 8])
def test_reference_perm_matches_container(case, index_bytes):
    if case["kind"] == "blocky":
        A = blocky_csr(case["n"], case["block"], 7, 1, 20, scatter=2 * case["n"],
                       seed=case["seed"], band=case["band"])
    elif case["kind"] == "random":
        A = random_csr(case["n"], case["density"], seed=case["seed"])
    else:
        A = growing_stripes(case["n_cols"], case["seed"])
    if index_bytes == 8:
        A = A.copy()
        A.indices = A.indices.astype(np.int64)
        A.indptr = A.indptr.astype(np.int64)
    br = bc = case["block"]
    got = run_perm(A, br, bc, case["mn"], index_bytes)
    want = expected_perm(A, br, bc, case["mn"])
    assert np.array_equal(got, want)


def test_reference_perm_wrapped_int32_keys():
    """Block keys above 2^31 wrap to negative int32 keys; the replay hashes them as the
    reference's container does (sign-extended to size_t)."""
    n_cols = 2**31 - 8
    rows = np.array([0, 0, 1, 80, 80, 80, 81, 88, 89, 89, 89])
    cols = np.array([5, 2**30, 2**31 - 20, 3, 2**30 + 100, 2**31 - 9, 77, 8, 2**29, 2**29 + 8,
                     2**29 + 800])
    A = sp.csr_matrix((np.ones(len(rows), np.float32), (rows, cols)), shape=(90, n_cols))
    A.indices = A.indices.astype(np.int32)
    A.indptr = A.indptr.astype(np.int32)
    got = run_perm(A, 8, 8, 1, 4)
    assert np.array_equal(got, expected_perm(A, 8, 8, 1))
