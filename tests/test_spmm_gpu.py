"""GPU SpMM parity (block tiles and CSR remainder) against scipy and the CPU oracle.

Grid of Tests/test_spmm.py:37-88 (N x K x density x block x min_nnz, plus non-power-of-two K),
tolerance atol=1e-4 as the reference (Tests/test_spmm.py:56), plus a tighter relative check
against the oracle's float64-accumulated product.
"""
import numpy as np
import scipy.sparse as sp
import pytest
import torch

import page_rank_amd as graph_link
from oracle import pbr as opbr
from tests.helpers import blocky_csr, random_csr, random_x

pytestmark = pytest.mark.gpu


def _check(A_csr, X, Y_gpu, atol=1e-4):
    Y_ref = A_csr.dot(X)
    assert np.allclose(Y_gpu, Y_ref, atol=atol), f"Max diff: {np.max(np.abs(Y_gpu - Y_ref)):.2e}"
    # against float64 accumulation: FP32 error bound ~ nnz_row * eps * sum|a||x|
    Y64 = A_csr.astype(np.float64).dot(X.astype(np.float64))
    scale = abs(A_csr).astype(np.float64).dot(np.abs(X).astype(np.float64)) + 1e-30
    rel = np.max(np.abs(Y_gpu - Y64) / scale)
    assert rel < 1e-5, f"relative error {rel:.2e}"


@pytest.mark.parametrize("N", [64, 128, 1024])
@pytest.mark.parametrize("num_features", [8, 64, 256])
@pytest.mark.parametrize("density", [0.05, 0.15])
@pytest.mark.parametrize("block_size", [2, 4, 8])
@pytest.mark.parametrize("min_nnz", [1, 4])
def test_spmm_accuracy(N, num_features, density, block_size, min_nnz, device):
    A_csr = random_csr(N, density)
    X = random_x(N, num_features)
    pbr = graph_link.csr_to_pbr(A_csr, block_rows=block_size, block_cols=block_size,
                                min_nnz_per_block=min_nnz).to(device)
    Y = graph_link.pbr_matmul(pbr, torch.from_numpy(X).to(device)).cpu().numpy()
    _check(A_csr, X, Y)


@pytest.mark.parametrize("num_features", [1, 2, 3, 4, 5, 6, 7, 9, 12, 16, 17, 31, 32, 33, 65])
@pytest.mark.parametrize("block_size", [2, 4, 8])
@pytest.mark.parametrize("min_nnz", [1, 4])
def test_spmm_non_pow2_features(num_features, block_size, min_nnz, device):
    N = 128
    A_csr = random_csr(N, density=0.1)
    X = random_x(N, num_features)
    pbr = graph_link.csr_to_pbr(A_csr, block_rows=block_size, block_cols=block_size,
                                min_nnz_per_block=min_nnz).to(device)
    Y = graph_link.pbr_matmul(pbr, torch.from_numpy(X).to(device)).cpu().numpy()
    _check(A_csr, X, Y)


@pytest.mark.parametrize("kw", [1, 2, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("mode", ["fused", "split"])
@pytest.mark.parametrize("block_size", [2, 4, 8])
def test_spmm_every_kernel_width(kw, mode, block_size, device):
    """Each lane layout (bit-lane kw<=8, feature-lane kw>=16) and both launch modes, on a
    matrix with dense tiles (bit 63 set), partial last stripe and a scattered remainder."""
    n = 8 * 37 + 5
    A = blocky_csr(n, 8, 3, 4, 64, scatter=400, seed=kw)
    for K in (1, kw, 3 * kw + 1):
        X = random_x(n, K, seed=K)
        pbr = graph_link.csr_to_pbr(A, block_rows=block_size, block_cols=block_size,
                                    min_nnz_per_block=4).to(device)
        Y = graph_link.pbr_matmul(pbr, torch.from_numpy(X).to(device), mode=mode,
                                  kw=kw).cpu().numpy()
        _check(A, X, Y)
        # the oracle's CPU SpMM over the same PBR arrays agrees as well
        Yo = opbr.spmm(opbr.csr_matrix_to_pbr(A, block_size, block_size, 4), X, wide=True)
        assert np.allclose(Y, Yo, rtol=1e-5, atol=1e-5)


def test_identity_8x8(device):
    """Frontend/test_pipeline.py:6-16: I8 with an 8x8 block and X = 1 gives Y = 1."""
    import scipy.sparse as sp
    A = sp.eye(8, format="csr").astype(np.float32)
    X = torch.ones((8, 1), device=device, dtype=torch.float32)
    pbr = graph_link.csr_to_pbr(A, block_rows=8, block_cols=8).to(device)
    Y = graph_link.pbr_matmul(pbr, X)
    assert torch.equal(Y.cpu(), torch.ones((8, 1)))


def test_edge_cases(device):
    import scipy.sparse as sp
    # empty matrix, no tiles, only remainder, only tiles, single row
    for A in (sp.csr_matrix((16, 16), dtype=np.float32),
              random_csr(16, 0.02, seed=3),
              sp.csr_matrix(np.ones((16, 16), np.float32)),
              sp.csr_matrix(np.ones((1, 1), np.float32))):
        for min_nnz in (0, 1, 65):
            X = random_x(A.shape[1], 5)
            pbr = graph_link.csr_to_pbr(A, 8, 8, min_nnz).to(device)
            Y = graph_link.pbr_matmul(pbr, torch.from_numpy(X).to(device)).cpu().numpy()
            assert np.allclose(Y, A.dot(X), atol=1e-5)


def test_batched_3d_and_accumulate(device):
    A = random_csr(200, 0.05, seed=7)
    X = np.stack([random_x(200, 6, seed=s) for s in range(3)])
    pbr = graph_link.csr_to_pbr(A, 4, 4, 2).to(device)
    Y = graph_link.pbr_matmul(pbr, torch.from_numpy(X).to(device)).cpu().numpy()
    assert Y.shape == (3, 200, 6)
    for b in range(3):
        assert np.allclose(Y[b], A.dot(X[b]), atol=1e-5)
    x = torch.from_numpy(X[0]).to(device)
    y = torch.ones((200, 6), device=device)
    graph_link.spmm_into(pbr, x, y, accumulate=True)
    assert np.allclose(y.cpu().numpy(), A.dot(X[0]) + 1.0, atol=1e-5)


def test_oracle_built_matrix_on_gpu(device):
    """A PBR built by the CPU oracle (reference order) and wrapped with from_arrays runs
    through the same kernels: the SpMM does not depend on tile order within a stripe."""
    A = blocky_csr(1000, 8, 4, 2, 30, scatter=3000, seed=11)
    o = opbr.csr_matrix_to_pbr(A, 8, 8, 4, order="reference")
    pbr = graph_link.from_arrays(o.rows, o.cols, 8, 8, o.block_codes, o.block_coords,
                                 o.block_offsets, o.block_data, o.remainder_indptr,
                                 o.remainder_col_ind, o.remainder_data).to(device)
    X = random_x(1000, 32)
    Y = graph_link.pbr_matmul(pbr, torch.from_numpy(X).to(device)).cpu().numpy()
    _check(A, X, Y)


def test_errors(device):
    A = random_csr(64, 0.1)
    pbr = graph_link.csr_to_pbr(A, 2, 2, 2)
    with pytest.raises(ValueError):
        graph_link.pbr_matmul(pbr, torch.zeros((2, 2, 2, 2), device=device))
    cpu = graph_link.csr_to_pbr(A, 2, 2, 2).to("cpu")
    with pytest.raises(RuntimeError):
        graph_link.pbr_batched_matmul_cuda(cpu, torch.zeros((64, 2), device=device))
    odd = graph_link.csr_to_pbr(A, 3, 3, 1)
    with pytest.raises(ValueError):
        graph_link.pbr_matmul(odd, torch.zeros((64, 2), device=device))


def test_to_semantics(device):
    """Backend/graph_link/graph_link/__init__.py:17-56: same device returns self; another
    device returns a new object registered by id(); the registry entry dies with it."""
    A = random_csr(100, 0.05, seed=2)
    built = graph_link.csr_to_pbr(A, 4, 4, 2)          # built on the GPU
    assert built.to(device) is built
    host = built.to("cpu")
    assert host is not built and graph_link.get_pbr_gpu_meta(host) is None
    assert host.to("cpu") is host
    gpu = host.to(device)
    assert gpu is not host and graph_link.get_pbr_gpu_meta(gpu) is not None
    assert graph_link.get_pbr_gpu_meta(host) is None
    assert gpu.to(device) is gpu
    assert gpu.to("cuda") is gpu
    meta = graph_link.get_pbr_gpu_meta(gpu)
    assert meta["codes"].dtype == torch.int64 and meta["coords"].dtype == torch.int32
    assert np.array_equal(meta["codes"].cpu().numpy().view(np.uint64), host.block_codes)
    graph_link.clear_pbr_gpu_meta(gpu)
    assert graph_link.get_pbr_gpu_meta(gpu) is None


def test_block_only_product_named_like_the_reference(device):
    """pbr_batched_matmul_triton_gpu (pbr_matrix_triton.py:77-110): y += tiles @ x."""
    A = blocky_csr(256, 8, 3, 2, 40, scatter=500, seed=5)
    pbr = graph_link.csr_to_pbr(A, 8, 8, 4).to(device)
    o = opbr.csr_matrix_to_pbr(A, 8, 8, 4, order="canonical")
    tiles_only = A - sp.csr_matrix((o.remainder_data, o.remainder_col_ind, o.remainder_indptr),
                                   shape=A.shape)
    X = random_x(256, 12)
    y = torch.ones((256, 12), device=device)
    graph_link.pbr_batched_matmul_triton_gpu(pbr, torch.from_numpy(X).to(device), y, 1, 12)
    assert np.allclose(y.cpu().numpy(), tiles_only.dot(X) + 1.0, atol=1e-4)
    X3 = np.stack([random_x(256, 12, seed=s) for s in range(2)])
    y3 = torch.zeros((2, 256, 12), device=device)
    graph_link.pbr_batched_matmul_triton_gpu(pbr, torch.from_numpy(X3).to(device), y3, 2, 12)
    for b in range(2):
        assert np.allclose(y3[b].cpu().numpy(), tiles_only.dot(X3[b]), atol=1e-4)


@pytest.mark.parametrize("kw", [0, 16, 32, 64])
@pytest.mark.parametrize("mode", ["fused", "split"])
def test_spmm_densely_tiled_8x8(kw, mode, device):
    """8x8 tilings with about ten tiles per stripe (the 1M x 1M workload's shape, which takes
    the whole-stripe tile batches, and the row-lane tiles at one column): odd
    stripe count, a partial last stripe, an empty stripe beside a full one, unequal tile
    counts in adjacent stripes, a scattered remainder, strided X and Y, accumulate."""
    n = 8 * 41 + 3
    A = blocky_csr(n, 8, 12, 20, 64, scatter=900, seed=13, band=8).tolil()
    A[16:24, :] = 0                       # stripe 2 empty, stripe 3 full
    A[40:48, :] = 0
    A[40:48, 0:8] = 1.0                   # stripe 5: one tile
    A = A.tocsr()
    A.eliminate_zeros()
    pbr = graph_link.csr_to_pbr(A, 8, 8, 4).to(device)
    ns = (n + 7) // 8
    assert pbr.accounted_blocks() >= 6 * ns  # the dense-tiling launch
    for K in (1, 3, 8, 32, 64, 20, 97, 128):
        X = random_x(n, K, seed=K)
        xd = torch.from_numpy(X).to(device)
        Y = graph_link.pbr_matmul(pbr, xd, mode=mode, kw=kw).cpu().numpy()
        _check(A, X, Y)
        # strided operands (row pitch not a multiple of 4) and accumulate
        xs = torch.zeros((n, K + 3), device=device)
        xs[:, :K] = xd
        ys = torch.full((n, K + 5), 2.0, device=device)
        graph_link.spmm_into(pbr, xs[:, :K], ys[:, :K], accumulate=True, mode=mode, kw=kw)
        out = ys.cpu().numpy()
        _check(A, X, out[:, :K] - 2.0)
        assert np.all(out[:, K:] == 2.0)


@pytest.mark.parametrize("mode", ["fused", "split"])
def test_plain_pointer_kernels(mode, device):
    """A matrix with more than 2^24 columns takes the plain-pointer form of the stripe
    kernels (64-bit element offsets; the buffer-descriptor form needs columns and the X row
    pitch below 2^24 and every array below 2 GB). Densely tiled 8x8 stripes with their
    columns spread over the whole range, so the whole-stripe batches run, plus a scattered
    remainder; K covers the bit-lane and both wide feature-lane kernels (X reaches 4.3 GB)."""
    rng = np.random.default_rng(5)
    n, cols = 8 * 40 + 5, (1 << 24) + 40
    rows_l, cols_l = [], []
    for s in range((n + 7) // 8):
        for bc in rng.choice(cols // 8, 9, replace=False):
            pos = rng.choice(64, int(rng.integers(4, 41)), replace=False)
            rows_l.append(s * 8 + pos // 8)
            cols_l.append(bc * 8 + pos % 8)
    rows_l.append(rng.integers(0, n, 600))
    cols_l.append(rng.integers(0, cols, 600))
    r, c = np.concatenate(rows_l), np.concatenate(cols_l)
    keep = r < n
    A = sp.csr_matrix((rng.random(int(keep.sum())).astype(np.float32) + 0.1,
                       (r[keep], c[keep])), shape=(n, cols))
    A.sum_duplicates()
    pbr = graph_link.csr_to_pbr(A, 8, 8, 4).to(device)
    assert pbr.accounted_blocks() >= 6 * ((n + 7) // 8)
    used = np.unique(A.indices)
    A_used = A[:, used]
    idx = torch.from_numpy(used).to(device)
    for K in (1, 8, 32, 64):
        X = torch.rand((cols, K), device=device)
        Y = graph_link.pbr_matmul(pbr, X, mode=mode).cpu().numpy()
        _check(A_used, X[idx].cpu().numpy(), Y)
        del X


@pytest.mark.parametrize("K", [1, 8, 32, 64])
def test_repeat_runs_are_bitwise_identical(K, device):
    """Every output element is owned by one wave and summed in a fixed order (no atomics),
    so repeated products are bitwise identical whatever order the workgroups run in."""
    n = 8 * 300 + 5
    A = blocky_csr(n, 8, 10, 16, 64, scatter=3000, seed=21, band=40)
    pbr = graph_link.csr_to_pbr(A, 8, 8, 4).to(device)
    X = torch.from_numpy(random_x(n, K, seed=K)).to(device)
    ref = graph_link.pbr_matmul(pbr, X)
    for _ in range(3):
        assert torch.equal(graph_link.pbr_matmul(pbr, X), ref)
