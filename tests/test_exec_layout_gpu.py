"""GPU checks of the 8x8 execution layout (gl_pbr_exec_layout) and of the matrix-core tile
kernel that reads it (8 and 32 RHS columns per launch, fused mode).

The layout is checked against its definition computed here with numpy (transposed codes,
column-major values, zero padding records); the product against scipy in FP64 on matrices
that exercise every branch of the kernel: stripes with no tile, with more than 8 tiles (more
than one batch) and more than 64 tiles, dense tiles (batches of more than 256 values), a
last block row and column past the matrix edge, row pitches that are not multiples of 4
(the dword-load form), partial column chunks and accumulation into Y.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import page_rank_amd as gl
from page_rank_amd import _native
from page_rank_amd.pbr import exec_layout

pytestmark = pytest.mark.gpu


def _transpose8(code):
    """Bit r*8+c of code -> bit c*8+r."""
    out = 0
    for r in range(8):
        for c in range(8):
            if (code >> (8 * r + c)) & 1:
                out |= 1 << (8 * c + r)
    return out


def _matrix(n, seed, dense_stripes=0, wide_stripe=False, density=0.04):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=density, format="lil", dtype=np.float32, random_state=rng)
    for s in range(dense_stripes):          # fully dense 8x8 tiles: 64 values each
        r0 = 8 * (2 * s + 1)
        for t in range(10):
            c0 = 8 * ((3 * s + 5 * t) % (n // 8))
            A[r0:r0 + 8, c0:c0 + 8] = rng.random((8, 8)).astype(np.float32) + 0.1
    if wide_stripe:                         # one stripe with more than 64 tiles
        r0 = 8 * (n // 16)
        for t in range(min(n // 8, 80)):
            A[r0:r0 + 8, 8 * t:8 * t + 8] = rng.random((8, 8)).astype(np.float32) * (
                rng.random((8, 8)) < 0.6)
    A[8 * 3:8 * 4, :] = 0                   # a stripe with no tile and no remainder
    A = sp.csr_matrix(A)
    A.eliminate_zeros()
    A.sort_indices()
    return A


def test_layout_matches_definition(device):
    A = _matrix(200, 1, dense_stripes=2)
    pbr = gl.csr_to_pbr(A, 8, 8, 3).to(device)
    meta = gl.get_pbr_gpu_meta(pbr)
    recs, vals = exec_layout(meta, 8)
    torch.cuda.synchronize()
    T = meta["codes"].numel()
    rec = recs.cpu().numpy().view(np.uint32).reshape(-1, 4)
    codes = meta["codes"].cpu().numpy().view(np.uint64)
    coords = meta["coords"].cpu().numpy().reshape(-1, 2)
    offs = meta["offsets"].cpu().numpy()
    data = meta["data"].cpu().numpy()
    v = vals.cpu().numpy()
    assert rec.shape[0] == T + 8
    assert recs.numel() == _native.lib().gl_pbr_exec_layout_bytes(T)
    for t in range(T):
        code = int(codes[t])
        tc = _transpose8(code)
        assert rec[t, 0] == (tc & 0xFFFFFFFF) and rec[t, 1] == (tc >> 32), t
        assert rec[t, 2] == coords[t, 1] and rec[t, 3] == offs[t]
        # values in row-major bit order -> column-major bit order, same range
        rm = [p for p in range(64) if (code >> p) & 1]
        cm = sorted(rm, key=lambda p: (p % 8) * 8 + p // 8)
        got = v[offs[t]:offs[t + 1]]
        want = np.array([data[offs[t] + rm.index(p)] for p in cm], dtype=np.float32)
        assert np.array_equal(got, want), t
    assert np.all(rec[T:, :3] == 0) and np.all(rec[T:, 3] == offs[T])


def _check(A, X, Y, Y0=None):
    ref = A.astype(np.float64) @ X.astype(np.float64)
    if Y0 is not None:
        ref += Y0
    scale = abs(A).astype(np.float64) @ np.abs(X).astype(np.float64) + 1e-30
    err = np.abs(Y - ref)
    assert np.max(err / (scale + (np.abs(Y0) if Y0 is not None else 0))) < 1e-5, np.max(err)


@pytest.mark.parametrize("n,K,kw,ldx_pad,seed", [
    (203, 32, 32, 0, 2),     # last block row and column past the edge, aligned rows
    (256, 32, 32, 1, 3),     # row pitch 33: the dword-load form
    (256, 45, 32, 0, 4),     # a 32-column launch, then a partial chunk in another kernel
    (520, 64, 32, 0, 5),     # two 32-column launches
    (640, 32, 32, 3, 6),
    (203, 8, 8, 0, 7),       # 8 columns per launch
    (256, 5, 8, 2, 8),       # a partial 8-column launch, unaligned pitch
    (520, 24, 8, 1, 9),      # three 8-column launches
])
def test_product_matches_fp64(n, K, kw, ldx_pad, seed, device):
    A = _matrix(n, seed, dense_stripes=3, wide_stripe=True)
    pbr = gl.csr_to_pbr(A, 8, 8, 4).to(device)
    rng = np.random.default_rng(seed)
    Xfull = rng.random((n, K + ldx_pad), dtype=np.float32)
    X = torch.from_numpy(Xfull).to(device)[:, :K]
    Y = torch.empty((n, K), dtype=torch.float32, device=device)
    gl.spmm_into(pbr, X, Y, kw=kw)
    _check(A, Xfull[:, :K], Y.cpu().numpy())
    assert "tile_records" in gl.get_pbr_gpu_meta(pbr)
    # accumulate into Y
    Y0 = rng.random((n, K), dtype=np.float32)
    Y = torch.from_numpy(Y0.copy()).to(device)
    gl.spmm_into(pbr, X, Y, kw=kw, accumulate=True)
    _check(A, Xfull[:, :K], Y.cpu().numpy(), Y0.astype(np.float64))


@pytest.mark.parametrize("K", [32, 8])
def test_exec_path_equals_layout_free_path(K, device):
    """The matrix-core kernel against the lane-per-column / bit-lane kernel on the same
    product."""
    A = _matrix(1024, 7, dense_stripes=4, wide_stripe=True, density=0.02)
    pbr = gl.csr_to_pbr(A, 8, 8, 4).to(device)
    X = torch.rand((1024, K), dtype=torch.float32, device=device)
    Y1 = torch.empty((1024, K), dtype=torch.float32, device=device)
    Y2 = torch.empty_like(Y1)
    gl.spmm_into(pbr, X, Y1, kw=K)                  # execution layout
    meta = gl.get_pbr_gpu_meta(pbr)
    L = _native.lib()
    rc = L.gl_pbr_spmm_i32_f32(1024, 1024, 8, 8, meta["codes"].numel(), meta["data"].numel(),
                               meta["rem_col_ind"].numel(), meta["codes"].data_ptr(),
                               meta["coords"].data_ptr(), meta["offsets"].data_ptr(),
                               meta["data"].data_ptr(), meta["stripe_ptr"].data_ptr(),
                               meta["rem_indptr"].data_ptr(), meta["rem_col_ind"].data_ptr(),
                               meta["rem_vals"].data_ptr(), X.data_ptr(), K, K, Y2.data_ptr(),
                               K, 0, 0, K, _native.stream_handle(device))
    _native.check(rc, "gl_pbr_spmm_i32_f32")
    d = (Y1 - Y2).abs().max().item()
    scale = Y2.abs().max().item()
    assert d <= 1e-6 * scale, d
    # two runs of the matrix-core kernel agree bit for bit
    Y3 = torch.empty_like(Y1)
    gl.spmm_into(pbr, X, Y3, kw=K)
    assert torch.equal(Y1, Y3)
