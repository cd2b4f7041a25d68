"""Parity at the benchmark sizes, through checks that do not need a full CPU product.

SpMM on the 1M x 1M block-banded matrix (8 tiles per block row within +-512 blocks, 5M
scattered remainder entries, ~27M nnz) at K = 1, 8, 32: sampled rows against the FP64 scipy
product, the column checksum 1^T (A X) = (1^T A) X in FP64, and linearity. Batched PPR on the
4M-node / 64M-edge block-clustered graph with 64 seed sets, and on the 16M-node / 256M-edge
one with 128: every column a probability vector and a fixed point of the PPR map to the
stopping tolerance.
"""
import numpy as np
import pytest
import torch

import page_rank_amd as graph_link
from page_rank_amd import graphs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def spmm_case(device):
    A = graphs.spmm_matrix(seed=0)
    pbr = graph_link.csr_to_pbr(A, 8, 8, 4, order="canonical").to(device)
    return A, pbr


@pytest.mark.parametrize("K", [1, 8, 32])
def test_spmm_1m_sampled_rows_checksum_linearity(spmm_case, K, device):
    A, pbr = spmm_case
    n = A.shape[0]
    g = torch.Generator(device="cpu").manual_seed(K)
    X1 = torch.rand((n, K), generator=g).to(device)
    X2 = torch.rand((n, K), generator=g).to(device)
    Y1 = graph_link.pbr_matmul(pbr, X1)
    # sampled rows against the FP64 product: FP32 error bound nnz_row * eps * sum|a||x|
    rows = np.arange(0, n, 997)
    x1 = X1.cpu().numpy().astype(np.float64)
    As = A[rows].astype(np.float64)
    ref = As @ x1
    scale = abs(As) @ np.abs(x1) + 1e-30
    got = Y1.cpu().numpy()[rows]
    assert np.max(np.abs(got - ref) / scale) < 1e-5
    # checksum of checksums: column sums of A X equal (column sums of A) . X
    colsum_a = np.asarray(A.astype(np.float64).sum(axis=0)).ravel()
    want = colsum_a @ x1
    have = Y1.double().sum(dim=0).cpu().numpy()
    assert np.allclose(have, want, rtol=1e-5, atol=0)
    # linearity: A (2 X1 - 3 X2) = 2 A X1 - 3 A X2 to FP32 rounding
    Y2 = graph_link.pbr_matmul(pbr, X2)
    Yc = graph_link.pbr_matmul(pbr, 2 * X1 - 3 * X2)
    diff = (Yc - (2 * Y1 - 3 * Y2)).abs().max().item()
    assert diff < 1e-4 * max(1.0, (2 * Y1).abs().max().item())


def test_ppr_4m_fixed_point_and_mass(device):
    n, m, K = 4_000_000, 64_000_000, 64
    src, dst = graphs.clustered_edges(n, m, communities=n // 1000, seed=3)
    A, dang = graphs.transition_matrix(src, dst, n)
    del src, dst
    assert not dang.any()
    pbr = graph_link.csr_to_pbr(A, 8, 8, 4, order="canonical").to(device)
    sets = graphs.seed_sets(n, K, 1, 100, seed=4)
    alpha, tol = 0.85, 1e-6
    res = graph_link.run_ppr(pbr, sets, alpha=alpha, tol=tol, max_iter=300,
                             retire="per_column", return_history=False)
    X = res.X
    assert res.converged.all()
    assert (X >= 0).all().item()
    mass = X.double().sum(dim=0).cpu().numpy()
    assert np.allclose(mass, 1.0, atol=1e-4)
    # fixed point: |alpha A x + (1 - alpha) p - x|_1 per column, p uniform over the seed set
    R = alpha * graph_link.pbr_matmul(pbr, X) - X
    for k, s in enumerate(sets):
        idx = torch.as_tensor(np.asarray(s, dtype=np.int64), device=device)
        R[idx, k] += (1.0 - alpha) / len(s)
    resid = R.abs().double().sum(dim=0).cpu().numpy()
    # the run stops once an update is below tol; the next update is alpha times smaller in
    # norm at most, plus FP32 rounding of the 4M-term sums
    assert resid.max() < 2 * tol


def test_ppr_16m_fixed_point_and_mass(device):
    """The full-scale run: 16M nodes, 256M edges, 16,000 communities, 128 seed sets of
    U{1..200}. X is 8.2 GB, so the step takes its plain-pointer form."""
    n, m, K = 16_000_000, 256_000_000, 128
    src, dst = graphs.clustered_edges(n, m, communities=n // 1000, seed=3)
    A, dang = graphs.transition_matrix(src, dst, n)
    del src, dst
    assert not dang.any()
    pbr = graph_link.csr_to_pbr(A, 8, 8, 4, order="canonical").to(device)
    del A
    sets = graphs.seed_sets(n, K, 1, 200, seed=4)
    alpha, tol = 0.85, 1e-6
    res = graph_link.run_ppr(pbr, sets, alpha=alpha, tol=tol, max_iter=300,
                             retire="per_column", return_history=False)
    X = res.X
    assert res.converged.all()
    assert (X >= 0).all().item()
    mass = torch.sum(X, dim=0, dtype=torch.float64).cpu().numpy()
    assert np.allclose(mass, 1.0, atol=1e-4)
    R = graph_link.pbr_matmul(pbr, X)
    R.mul_(alpha).sub_(X)
    for k, s in enumerate(sets):
        idx = torch.as_tensor(np.asarray(s, dtype=np.int64), device=device)
        R[idx, k] += (1.0 - alpha) / len(s)
    resid = torch.sum(R.abs_(), dim=0, dtype=torch.float64).cpu().numpy()
    assert resid.max() < 2 * tol
