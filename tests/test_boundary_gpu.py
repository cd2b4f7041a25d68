"""GPU checks of the drop-in boundary beyond the batched loop: the reference's per-kernel PPR
functions in FP32 (page_rank_amd.core, replacing bindings_kernels.cu:148-161), any number of
source columns, max_iterations=0, an in-place warm start through the C ABI, the float64
PBR -> CSR round trip, and concurrent callers on several host threads."""
import threading

import numpy as np
import pytest
import scipy.sparse as sp
import torch

import page_rank_amd as graph_link
from oracle import ppr as oppr
from page_rank_amd import _native, core

pytestmark = pytest.mark.gpu


def _er_graph(n, deg, seed):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, n * deg)
    dst = rng.integers(0, n, n * deg)
    keep = src != dst
    A = sp.csr_matrix((np.ones(keep.sum(), np.float32), (dst[keep], src[keep])), shape=(n, n))
    A.sum_duplicates()
    return A


def test_fp32_host_loop_matches_reference_loop(device):
    """init + (SpMM, update) per iteration, the reference's __init__.py:240-260 loop on the
    FP32 per-kernel entry points, against the FP64 restatement of that loop."""
    n = 3000
    A = graph_link.normalize_transition_matrix(_er_graph(n, 6, 1))
    pbr = graph_link.csr_to_pbr(A, 8, 8, 2).to(device)
    sources = [0, 17, 1234, n - 1]
    K = len(sources)
    src = torch.tensor(sources, dtype=torch.int32, device=device)
    X = torch.full((n, K), 7.0, device=device)  # init writes every element
    errors = torch.zeros(K, device=device)
    core.init_ppr_cuda_float(X, src, n, K)
    assert torch.equal(X.sum(0).cpu(), torch.ones(K))
    hist = []
    for t in range(200):
        Y = graph_link.pbr_matmul(pbr, X)
        core.ppr_update_normalized_cuda_float(Y, X, src, 0.85, n, K, errors)
        hist.append(float(errors.max()))
        if hist[-1] < 1e-6:
            break
    Xr, it_r, conv_r, hist_r = oppr.reference_loop(A, sources, 0.85, 200, 1e-6)
    assert conv_r and abs(len(hist) - it_r) <= 1
    err = np.abs(X.cpu().numpy() - Xr).sum(axis=0).max()
    print(f"host loop: {len(hist)} iterations (oracle {it_r}), max L1 vs FP64 {err:.2e}")
    assert err < 5e-6
    np.testing.assert_allclose(hist[:5], hist_r[:5], rtol=1e-4)


def test_fp32_update_is_deterministic_and_exact_residual(device):
    n, K = 20000, 5
    g = torch.Generator(device="cpu").manual_seed(3)
    Y = torch.rand((n, K), generator=g).to(device)
    X0 = torch.rand((n, K), generator=g).to(device)
    src = torch.tensor([0, 5, 19999, 7, 7], dtype=torch.int32, device=device)
    outs = []
    for _ in range(2):
        X = X0.clone()
        e = torch.empty(K, device=device)
        core.ppr_update_normalized_cuda_float(Y, X, src, 0.85, n, K, e)
        outs.append((X.cpu(), e.cpu()))
    assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])
    Yh, X0h = Y.cpu().double().numpy(), X0.cpu().double().numpy()
    want = 0.85 * Yh
    want[src.cpu().numpy(), np.arange(K)] += 0.15
    np.testing.assert_allclose(outs[0][0].numpy(), want, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(outs[0][1].numpy(), np.abs(want - X0h).sum(0), rtol=1e-5)


def test_fp32_dangling_to_seed_step(device):
    """missing_mass + ppr_update (ppr_kernel.cu:13-117): a matrix with empty dangling
    columns, the lost mass returned to the seed each step, against oracle.power's
    dangling_mode='personalization' (networkx's default)."""
    n = 2000
    Araw = _er_graph(n, 4, 7).tolil()
    Araw[:, :40] = 0  # 40 dangling sources
    An, dang = oppr.column_stochastic(sp.csr_matrix(Araw))
    A32 = sp.csr_matrix(An, dtype=np.float32)
    pbr = graph_link.csr_to_pbr(A32, 8, 8, 2).to(device)
    sources = [50, 100, 1999]
    K = len(sources)
    src = torch.tensor(sources, dtype=torch.int32, device=device)
    X = torch.empty((n, K), device=device)
    core.init_ppr_cuda_float(X, src, n, K)
    Xn = torch.empty_like(X)
    cs = torch.empty(K, device=device)
    for t in range(200):
        Y = graph_link.pbr_matmul(pbr, X)
        core.missing_mass_cuda_float(Y, cs, n, K)
        errors = torch.zeros(K, device=device)  # the reference's step adds into errors
        core.ppr_update_cuda_float(Y, X, Xn, src, cs, 0.85, n, K, errors)
        X, Xn = Xn, X
        if float(errors.max()) < 1e-7:
            break
    P = oppr.personalization_matrix(n, [[s] for s in sources])
    Xo, _, _, _ = oppr.power(An, P, 0.85, 1e-12, 500, dang, "personalization")
    err = np.abs(X.cpu().numpy() - Xo).sum(axis=0).max()
    print(f"dangling-to-seed loop: {t + 1} iterations, max L1 vs FP64 {err:.2e}")
    assert err < 2e-6
    assert np.allclose(X.cpu().numpy().sum(0), 1.0, atol=1e-5)


def test_more_than_1024_sources(device):
    n = 4096
    A = graph_link.normalize_transition_matrix(_er_graph(n, 5, 11))
    pbr = graph_link.csr_to_pbr(A, 8, 8, 2).to(device)
    sources = np.arange(0, n, 3)[:1025]
    X, iters, conv = graph_link.run_personalized_pagerank(
        pbr, torch.tensor(sources, dtype=torch.int32, device=device), tolerance=1e-6,
        max_iterations=200)
    Xr, it_r, conv_r, _ = oppr.reference_loop(A, sources, 0.85, 200, 1e-6)
    assert conv and conv_r and abs(iters - it_r) <= 1 and X.shape == (n, 1025)
    assert np.abs(X.cpu().numpy() - Xr).sum(axis=0).max() < 5e-6


def test_max_iterations_zero_returns_the_start(device):
    A = graph_link.normalize_transition_matrix(_er_graph(500, 4, 2))
    pbr = graph_link.csr_to_pbr(A, 4, 4, 2).to(device)
    src = torch.tensor([3, 9], dtype=torch.int32, device=device)
    X, iters, conv, hist = graph_link.run_personalized_pagerank(pbr, src, max_iterations=0,
                                                                return_history=True)
    assert iters == 0 and conv is False and hist == []
    want = np.zeros((500, 2), np.float32)
    want[[3, 9], [0, 1]] = 1
    assert np.array_equal(X.cpu().numpy(), want)


def test_in_place_warm_start_through_the_c_abi(device):
    """X0 == X_out: the run starts from the caller's X_out (a converged vector converges in
    one pass and stays put) instead of the seed start."""
    n = 3000
    A32, dang = graph_link.graphs.transition_matrix(*graph_link.graphs.sbm_edges(
        n=n, community=100, p_in=0.05, p_out=0.001, seed=4), n)
    pbr = graph_link.csr_to_pbr(A32, 8, 8, 4).to(device)
    sets = [[0], [10, 20, 30]]
    first = graph_link.run_ppr(pbr, sets, tol=1e-7, max_iter=300, dangling=dang,
                               dangling_mode="uniform")
    meta = graph_link.get_pbr_gpu_meta(pbr)
    L = _native.lib()
    X = first.X.clone()
    ptr, idx, w = graph_link.ppr._seed_arrays(sets)
    K = len(sets)
    wsb = L.gl_ppr_workspace_bytes(n, 8, K, int(idx.size))
    ws = torch.empty(wsb, dtype=torch.uint8, device=device)
    iters = torch.zeros(K, dtype=torch.int32, device=device)
    conv = torch.zeros(K, dtype=torch.int32, device=device)
    total = np.zeros(1, np.int32)
    dm = torch.as_tensor(dang.astype(np.uint8)).to(device)
    opts = _native.GlOpts(alpha=0.85, tol=1e-6, max_iter=50,
                          dangling_mode=_native.GL_DANGLING_UNIFORM,
                          retire_mode=_native.GL_RETIRE_PER_COLUMN, graph_iters=0, kw=0,
                          grid_mult=0)
    with torch.cuda.device(device):
        rc = L.gl_ppr_run_i32_f32(
            n, n, 8, 8, meta["codes"].numel(), meta["data"].numel(),
            meta["rem_col_ind"].numel(), meta["codes"].data_ptr(), meta["coords"].data_ptr(),
            meta["offsets"].data_ptr(), meta["data"].data_ptr(), meta["stripe_ptr"].data_ptr(),
            meta["rem_indptr"].data_ptr(), meta["rem_col_ind"].data_ptr(),
            meta["rem_vals"].data_ptr(), K, ptr.ctypes.data, idx.ctypes.data, w.ctypes.data,
            dm.data_ptr(), X.data_ptr(), opts, X.data_ptr(), iters.data_ptr(),
            conv.data_ptr(), None, ws.data_ptr(), wsb, total.ctypes.data,
            _native.stream_handle(device))
    _native.check(rc, "gl_ppr_run_i32_f32")
    assert total[0] == 1 and conv.cpu().numpy().all()
    assert np.abs(X.cpu().numpy() - first.X.cpu().numpy()).sum(axis=0).max() < 1e-6


def test_pbr_to_csr_float64_round_trip(device):
    rng = np.random.default_rng(5)
    A = sp.random(700, 900, density=0.02, format="csr", random_state=rng, dtype=np.float64)
    A.data = rng.random(A.nnz) * 1e-3 + np.pi  # values with all 53 bits in use
    for br in (2, 4, 8):
        p = graph_link.csr_to_pbr(A, br, br, 3).to(device)
        B = graph_link.pbr_to_scipy(p)
        assert B.dtype == np.float64
        assert np.array_equal(B.indptr, A.indptr) and np.array_equal(B.indices, A.indices)
        assert np.array_equal(B.data, A.data)


def test_concurrent_callers(device):
    """Two host threads run the graph-captured PPR loop and split-mode SpMMs at the same
    time (FastAPI runs /api/ppr in a thread pool, as the reference's server does); every
    result must equal the one computed alone."""
    n = 20000
    A32, dang = graph_link.graphs.transition_matrix(*graph_link.graphs.sbm_edges(
        n=n, community=200, p_in=0.03, p_out=0.0005, seed=9), n)
    pbr = graph_link.csr_to_pbr(A32, 8, 8, 4).to(device)
    sets = [[1], [2, 3], [400, 500, 600], [19999]]
    X = torch.rand((n, 24), generator=torch.Generator().manual_seed(1)).to(device)
    # single_workgroup=False: the graph-captured multi-workgroup loop, whatever the size
    ref_ppr = graph_link.run_ppr(pbr, sets, tol=1e-6, max_iter=200, dangling=dang,
                                 dangling_mode="uniform", single_workgroup=False).X.cpu()
    ref_y = graph_link.pbr_matmul(pbr, X, mode="split").cpu()
    fails = []

    def ppr_worker():
        try:
            with torch.cuda.device(device):
                s = torch.cuda.Stream(device)
                with torch.cuda.stream(s):
                    for _ in range(6):
                        r = graph_link.run_ppr(pbr, sets, tol=1e-6, max_iter=200,
                                               dangling=dang, dangling_mode="uniform",
                                               single_workgroup=False)
                        s.synchronize()
                        if not torch.equal(r.X.cpu(), ref_ppr):
                            fails.append("ppr")
        except Exception as e:  # noqa: BLE001
            fails.append(repr(e))

    def spmm_worker():
        try:
            with torch.cuda.device(device):
                s = torch.cuda.Stream(device)
                with torch.cuda.stream(s):
                    for _ in range(40):
                        y = graph_link.pbr_matmul(pbr, X, mode="split")
                        s.synchronize()
                        if not torch.equal(y.cpu(), ref_y):
                            fails.append("spmm")
        except Exception as e:  # noqa: BLE001
            fails.append(repr(e))

    threads = [threading.Thread(target=ppr_worker), threading.Thread(target=ppr_worker),
               threading.Thread(target=spmm_worker)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    assert not any(t.is_alive() for t in threads), "a worker did not finish"
    assert not fails, fails
    Xo, _, _, _ = oppr.power(A32, oppr.personalization_matrix(n, sets), 0.85, 1e-12, 500, dang,
                             "uniform", "per_column")
    assert np.abs(ref_ppr.numpy() - Xo).sum(axis=0).max() < 5e-6
