"""The workloads' own graphs on the HIP path, against FP64 references:

config 1  the seeded 10,000-node SBM (50 communities of 200, p_in 0.04, p_out 0.0002), one
          seed, alpha 0.85, dangling mass returned to the seed (networkx's default), as the
          reference's accuracy test runs networkx (Tests/test_ppr_accuracy.py:28-40);
config 4  a held-out 1M-node / 16M-edge graph of the clustered generator with the committed
          warm-start weights: plain and learned start both reach L1 < 1e-6 against FP64 on 8
          sampled columns (the learned start is judged by the plain method's own bar).
Errors are L1 and max-abs per column, printed with iterations and times."""
import time

import numpy as np
import pytest
import scipy.sparse as sp
import torch

import page_rank_amd as graph_link
from oracle import ppr as oppr
from page_rank_amd import graphs
from page_rank_amd import warmstart as ws

pytestmark = pytest.mark.gpu


def test_config1_sbm_against_fp64(device):
    n = 10_000
    src, dst = graphs.sbm_edges(n=n, community=200, p_in=0.04, p_out=0.0002, seed=0)
    assert 95_000 < src.size < 103_000
    A_col = sp.csr_matrix((np.ones(src.size), (dst, src)), shape=(n, n))
    A64, dang = oppr.column_stochastic(A_col)
    A32, dang32 = graphs.transition_matrix(src, dst, n)
    assert np.array_equal(dang, dang32)
    seed = 0
    p = np.zeros(n)
    p[seed] = 1.0
    exact = oppr.solve(A64, p, 0.85, dang, "personalization")
    x_nx = oppr.nx_pagerank(A_col, p, 0.85, tol=1e-16, max_iter=1000)  # stop at N*tol
    assert np.abs(x_nx - exact).sum() < 1e-9  # the oracle pair agrees first
    pbr = graph_link.csr_to_pbr(A32, 8, 8, 4).to(device)
    for tol in (1.5e-7, 1e-8):
        r = graph_link.run_ppr(pbr, [[seed]], tol=tol, max_iter=100, dangling=dang,
                               dangling_mode="personalization")
        x = r.X[:, 0].cpu().numpy().astype(np.float64)
        l1, mx = np.abs(x - exact).sum(), np.abs(x - exact).max()
        print(f"\nconfig 1, tol {tol:g}: {r.total_iterations} iterations, converged "
              f"{bool(r.converged[0])}, L1 {l1:.2e}, max abs {mx:.2e}")
        assert l1 < 1e-6 and mx < 1e-7
        assert abs(x.sum() - 1) < 1e-5


def test_config4_heldout_graph_warm_and_plain(device):
    t0 = time.perf_counter()
    pbr, sets, dang, dmode = ws.clustered_sample(1_000_000, 16_000_000, 100, seed=5000)
    print(f"\nheld-out graph built in {time.perf_counter() - t0:.1f} s: "
          f"T={pbr.accounted_blocks()} R={pbr.remainder_nnz()}", flush=True)
    params = ws.load(ws.WEIGHTS, device)
    kw = dict(tol=1.5e-7, max_iter=300, dangling=dang, dangling_mode=dmode,
              return_history=False)
    graph = ws.graph_features(pbr)
    for _ in range(2):  # the second pass is timed (graphs captured, caches warm)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        plain = graph_link.run_ppr(pbr, sets, **kw)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        warm = ws.run_warm(params, pbr, sets, graph=graph, **kw)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
    assert plain.converged.all() and warm.converged.all()
    cols = np.random.default_rng(11).choice(len(sets), 8, replace=False)
    A = graph_link.pbr_to_scipy(pbr).astype(np.float64)
    P = oppr.personalization_matrix(pbr.rows, [sets[c] for c in cols])
    ref, _, _, _ = oppr.power(A, P, 0.85, 1e-11, 400, dang, dmode, "per_column")
    sel = torch.as_tensor(cols, device=device)
    for name, r in (("plain", plain), ("warm", warm)):
        d = np.abs(r.X[:, sel].cpu().numpy().astype(np.float64) - ref)
        l1, mx = d.sum(0), d.max(0)
        print(f"{name}: iterations max {r.total_iterations}+{r.feature_passes} mean "
              f"{r.iterations.mean():.1f}+{r.feature_passes}, "
              f"L1 vs FP64 max {l1.max():.2e}, max abs {mx.max():.2e}")
        assert l1.max() < 1e-6
    print(f"wall: plain {(t2 - t1) * 1e3:.1f} ms, warm {(t3 - t2) * 1e3:.1f} ms "
          f"(features + network + iterations)")
