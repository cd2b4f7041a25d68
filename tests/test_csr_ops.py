"""GPU column normalization (bit-identical to the reference's float32 host arithmetic),
PBR -> CSR on the GPU, the .npz cache, and building from a device CSR."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import page_rank_amd as graph_link
from oracle import pbr as opbr
from oracle import ppr as oppr
from tests.helpers import blocky_csr, random_csr

f32 = np.float32


def _pairwise(a):
    n = len(a)
    if n < 8:
        r = f32(-0.0)
        for v in a:
            r = f32(r + v)
        return r
    if n <= 128:
        r = [a[j] for j in range(8)]
        i = 8
        while i < n - n % 8:
            for j in range(8):
                r[j] = f32(r[j] + a[i + j])
            i += 8
        res = f32(f32(f32(r[0] + r[1]) + f32(r[2] + r[3])) + f32(f32(r[4] + r[5]) + f32(r[6] + r[7])))
        for v in a[i:]:
            res = f32(res + v)
        return res
    h = n // 2
    h -= h % 8
    return f32(_pairwise(a[:h]) + _pairwise(a[h:]))


def test_host_column_sum_order():
    """The summation order the GPU kernel reproduces: numpy.add.reduceat over a float32
    segment is seg[0] + pairwise(seg[1:]). If numpy changed it, this fails first."""
    rng = np.random.default_rng(0)
    for L in list(range(1, 140)) + [255, 256, 257, 1000, 1025, 4099]:
        a = (rng.random(L) * rng.choice([1e-3, 1.0, 1e3], L)).astype(f32)
        want = np.add.reduceat(a, [0])[0]
        got = a[0] if L == 1 else f32(a[0] + _pairwise(a[1:]))
        assert want == got, L


def _hubby(n, seed, dangling=0):
    """Random digraph adjacency A[target, source] with a few hub sources (columns of
    130..3000 entries) and `dangling` empty columns."""
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, n * 6)
    dst = rng.integers(0, n, n * 6)
    hubs = rng.choice(n, 6, replace=False)
    for h, k in zip(hubs, (130, 136, 257, 1000, 2049, 3000)):
        k = min(k, n - 1)
        src = np.r_[src, np.full(k, h)]
        dst = np.r_[dst, rng.choice(n, k, replace=False)]
    w = (rng.random(src.size) * 3 + 0.01).astype(np.float32)
    A = sp.csr_matrix((w, (dst, src)), shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    if dangling:
        cols = rng.choice(n, dangling, replace=False)
        keep = ~np.isin(A.tocoo().col, cols)
        C = A.tocoo()
        A = sp.csr_matrix((C.data[keep], (C.row[keep], C.col[keep])), shape=(n, n))
        A.sum_duplicates()
        A.sort_indices()
    else:
        # every column non-empty
        A = (A + sp.csr_matrix((np.ones(n, np.float32), (np.roll(np.arange(n), 1), np.arange(n))),
                               shape=(n, n))).tocsr()
        A.sum_duplicates()
        A.sort_indices()
    return A.astype(np.float32)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_normalize_is_bit_identical(seed, device):
    A = _hubby(4000, seed)
    want = graph_link.normalize_transition_matrix(A)
    want.sort_indices()
    got, dang = graph_link.normalize_transition_matrix_gpu(A, device)
    assert not dang.any()
    g = got.to_scipy()
    assert np.array_equal(g.indptr, want.indptr) and np.array_equal(g.indices, want.indices)
    assert np.array_equal(g.data.view(np.uint32), want.data.astype(np.float32).view(np.uint32))


@pytest.mark.gpu
def test_normalize_dangling_mask_and_ppr(device):
    A = _hubby(1500, 7, dangling=40)
    want = graph_link.normalize_transition_matrix(A)     # dense 1/N columns for dangling
    got, dang = graph_link.normalize_transition_matrix_gpu(A, device)
    col_sums = np.asarray(A.sum(axis=0)).ravel()
    assert np.array_equal(dang, col_sums == 0) and dang.sum() >= 40
    g = got.to_scipy().tocsc()
    w = want.tocsc()
    for c in np.flatnonzero(~dang)[:300]:
        a, b = g[:, c], w[:, c]
        assert np.array_equal(a.indices, b.indices)
        assert np.array_equal(a.data.view(np.uint32), b.data.astype(np.float32).view(np.uint32))
    # implicit uniform dangling in the PPR run == the reference's densified matrix
    pbr_dense = graph_link.csr_to_pbr(want, 8, 8, 2).to(device)
    pbr_gpu = graph_link.csr_to_pbr(got, 8, 8, 2)
    srcs = [3, 77, 500, 1499]
    X, it, conv = graph_link.run_personalized_pagerank(
        pbr_dense, torch.tensor(srcs, dtype=torch.int32, device=device), tolerance=1e-7,
        max_iterations=300)
    r = graph_link.run_ppr(pbr_gpu, [[s] for s in srcs], tol=1e-7, max_iter=300,
                           dangling=dang, dangling_mode="uniform", retire="together")
    assert conv and r.converged.all()
    assert np.abs(X.cpu().numpy() - r.X.cpu().numpy()).sum(0).max() < 2e-6


@pytest.mark.gpu
@pytest.mark.parametrize("block", [2, 4, 8])
@pytest.mark.parametrize("order", ["reference", "canonical"])
def test_pbr_to_csr_round_trip(block, order, device):
    for A in (blocky_csr(1003, block, 3, 1, block * block, scatter=2000, seed=block),
              random_csr(517, 0.02, seed=9), sp.csr_matrix((37, 41), dtype=np.float32),
              sp.csr_matrix(np.ones((13, 13), np.float32))):
        A = sp.csr_matrix(A, dtype=np.float32)
        A.sort_indices()
        for mn in (0, 2, 65):
            pbr = graph_link.csr_to_pbr(A, block, block, mn, order=order)
            B = graph_link.pbr_to_scipy(pbr)
            assert B.shape == A.shape
            assert np.array_equal(B.indptr, A.indptr) and np.array_equal(B.indices, A.indices)
            assert np.array_equal(B.data.view(np.uint32), A.data.view(np.uint32))


@pytest.mark.gpu
def test_pbr_to_csr_of_oracle_built_matrix_and_reference_signature(device):
    A = blocky_csr(2001, 8, 5, 2, 40, scatter=3000, seed=4)   # 2001 rows: partial last stripe
    o = opbr.csr_matrix_to_pbr(A, 8, 8, 4, order="reference")
    host = graph_link.from_arrays(o.rows, o.cols, 8, 8, o.block_codes, o.block_coords,
                                  o.block_offsets, o.block_data, o.remainder_indptr,
                                  o.remainder_col_ind, o.remainder_data)
    indptr = np.zeros(A.shape[0] + 1, np.int32)
    indices = np.zeros(A.nnz, np.int32)
    data = np.zeros(A.nnz, np.float32)
    graph_link.pbr_to_csr(host, indptr, indices, data)       # host-only matrix: uploaded
    assert np.array_equal(indptr, A.indptr) and np.array_equal(indices, A.indices)
    assert np.array_equal(data, A.data)
    # device outputs
    ti = torch.zeros(A.nnz, dtype=torch.int64, device=device)
    tp = torch.zeros(A.shape[0] + 1, dtype=torch.int64, device=device)
    td = torch.zeros(A.nnz, dtype=torch.float32, device=device)
    graph_link.pbr_to_csr(host.to(device), tp, ti, td)
    assert np.array_equal(ti.cpu().numpy(), A.indices)
    with pytest.raises(ValueError, match="room for"):
        graph_link.pbr_to_csr(host, indptr, indices[:-1], data[:-1])


@pytest.mark.gpu
def test_npz_cache_and_device_csr_build(tmp_path, device):
    A = blocky_csr(600, 4, 3, 2, 12, scatter=500, seed=2)
    pbr = graph_link.csr_to_pbr(A, 4, 4, 2)
    path = tmp_path / "m.npz"
    graph_link.save_pbr(pbr, path)
    back = graph_link.load_pbr(path)
    for k in ("block_codes", "block_coords", "block_offsets", "block_data",
              "remainder_indptr", "remainder_col_ind", "remainder_data"):
        assert np.array_equal(getattr(back, k), getattr(pbr, k)), k
    X = torch.rand((600, 5), device=device)
    y1 = graph_link.pbr_matmul(pbr, X)
    y2 = graph_link.pbr_matmul(graph_link.load_pbr(path, device), X)
    assert torch.equal(y1, y2)
    d = graph_link.upload_csr(A, device)
    p2 = graph_link.csr_to_pbr(d, 4, 4, 2)
    for k in ("block_codes", "block_coords", "block_data", "remainder_col_ind"):
        assert np.array_equal(getattr(p2, k), getattr(pbr, k)), k
