"""CPU pins of the FP64 PPR oracle (oracle/ppr.py) against networkx's pagerank (a library,
the reference's own oracle in Tests/test_ppr_accuracy.py:28-40), sparse LU solves of the
fixed-point equation, and closed forms."""
import networkx as nx
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import ppr as oppr


def random_digraph(n, p, seed, drop_out=()):
    G = nx.gnp_random_graph(n, p, directed=True, seed=seed)
    for u in drop_out:
        G.remove_edges_from(list(G.out_edges(u)))
    return G


def col_adjacency(G, n):
    A = nx.to_scipy_sparse_array(G, nodelist=range(n), format="csr", dtype=np.float64)
    return sp.csr_matrix(A.T)  # A_col[target, source]


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("alpha", [0.85, 0.9])
def test_nx_restatement_matches_networkx(seed, alpha):
    n = 80
    G = random_digraph(n, 0.06, seed, drop_out=(3, 17))
    p = np.zeros(n)
    p[[1, 5]] = [0.25, 0.75]
    want = nx.pagerank(G, alpha=alpha, personalization={i: p[i] for i in range(n)}, tol=1e-10,
                       max_iter=1000)
    got = oppr.nx_pagerank(col_adjacency(G, n), p, alpha, tol=1e-10, max_iter=1000)
    assert np.allclose(got, [want[i] for i in range(n)], atol=1e-9)


@pytest.mark.parametrize("mode", ["uniform", "personalization"])
def test_power_matches_networkx_dangling_conventions(mode):
    n = 120
    G = random_digraph(n, 0.05, 7, drop_out=(0, 9, 50))
    A, dangling = oppr.column_stochastic(col_adjacency(G, n))
    assert dangling[[0, 9, 50]].all()
    seeds = [[4], [10, 11, 12]]
    P = oppr.personalization_matrix(n, seeds)
    X, it, conv, _ = oppr.power(A, P, 0.85, 1e-12, 2000, dangling, mode, "per_column")
    assert conv.all()
    for k in range(2):
        pers = {i: P[i, k] for i in range(n)}
        dw = {i: 1.0 / n for i in range(n)} if mode == "uniform" else None
        want = nx.pagerank(G, alpha=0.85, personalization=pers, dangling=dw, tol=1e-13,
                           max_iter=5000)
        assert np.allclose(X[:, k], [want[i] for i in range(n)], atol=1e-10)
        assert abs(X[:, k].sum() - 1) < 1e-10


@pytest.mark.parametrize("mode", ["none", "uniform", "personalization"])
def test_power_matches_exact_solve(mode):
    n = 200
    G = random_digraph(n, 0.03, 11, drop_out=(2, 3))
    A, dangling = oppr.column_stochastic(col_adjacency(G, n))
    P = oppr.personalization_matrix(n, [[7], [8, 9]])
    X, _, _, hist = oppr.power(A, P, 0.85, 1e-9, 1000, dangling, mode, "together")
    for k in range(2):
        exact = oppr.solve(A, P[:, k], 0.85, dangling, mode)
        # contraction bound: ||x_t - x*||_1 <= alpha/(1-alpha) * last residual
        bound = 0.85 / 0.15 * hist[-1][k] + 1e-12
        assert np.abs(X[:, k] - exact).sum() <= bound


def test_reference_loop_is_power_without_dangling():
    n = 100
    G = random_digraph(n, 0.05, 5)
    A, dangling = oppr.column_stochastic(col_adjacency(G, n))
    src = [0, 25, 50, 99]
    X1, it1, c1, h1 = oppr.reference_loop(A, src, 0.85, 300, 1e-8)
    P = oppr.personalization_matrix(n, [[s] for s in src])
    X2, it2, c2, h2 = oppr.power(A, P, 0.85, 1e-8, 300, None, "none", "together")
    assert c1 and c2.all() and it1 == it2.max()
    assert np.allclose(X1, X2, atol=1e-15)
    assert np.allclose(h1, np.max(h2, axis=1))


def test_reference_loop_against_networkx_uniform_dangling():
    """The reference's test: normalize with dense 1/N dangling columns, then compare the
    loop with networkx using dangling=uniform (Tests/test_ppr_accuracy.py:58-96)."""
    import page_rank_amd as graph_link
    n = 50
    G = nx.erdos_renyi_graph(n, 0.1, directed=True, seed=42)
    A_nx = nx.to_scipy_sparse_array(G, nodelist=range(n), format="csr", dtype=np.float32)
    A_norm = graph_link.normalize_transition_matrix(sp.csr_matrix(A_nx.T))
    X, it, conv, _ = oppr.reference_loop(A_norm, [0, 12], 0.85, 500, 1e-10)
    assert conv
    for k, s in enumerate([0, 12]):
        want = nx.pagerank(G, alpha=0.85, personalization={i: float(i == s) for i in range(n)},
                           dangling={i: 1.0 / n for i in range(n)}, tol=1e-12, max_iter=1000)
        assert np.allclose(X[:, k], [want[i] for i in range(n)], atol=1e-6)


def test_closed_forms():
    # alpha = 0 -> personalization
    A = sp.csr_matrix(np.full((4, 4), 0.25))
    P = oppr.personalization_matrix(4, [[2]])
    X, it, _, _ = oppr.power(A, P, 0.0, 1e-12, 10)
    assert np.allclose(X[:, 0], [0, 0, 1, 0]) and it[0] == 1
    # K5 with uniform personalization -> 1/5
    K5 = sp.csr_matrix((np.ones((5, 5)) - np.eye(5)) / 4)
    X, _, _, _ = oppr.power(K5, oppr.personalization_matrix(5, [[0, 1, 2, 3, 4]]), 0.85,
                            1e-14, 100)
    assert np.allclose(X[:, 0], 0.2)
    # isolated seeded node, dangling mass back to the personalization -> 1 on the seed
    Z = sp.csr_matrix((3, 3))
    X, _, _, _ = oppr.power(Z, oppr.personalization_matrix(3, [[1]]), 0.85, 1e-14, 100,
                            np.ones(3, bool), "personalization")
    assert np.allclose(X[:, 0], [0, 1, 0])


def test_planted_mistake_is_caught():
    """A dropped teleport term in the oracle would fail the networkx comparison."""
    n = 60
    G = random_digraph(n, 0.08, 3)
    A, dangling = oppr.column_stochastic(col_adjacency(G, n))
    P = oppr.personalization_matrix(n, [[1]])
    X, _, _, _ = oppr.power(A, P, 0.85, 1e-12, 2000, dangling, "uniform")
    bad = 0.85 * (A @ X[:, 0])  # one step of the update without (1-alpha) p
    want = nx.pagerank(G, alpha=0.85, personalization={i: P[i, 0] for i in range(n)},
                       dangling={i: 1 / n for i in range(n)}, tol=1e-13, max_iter=5000)
    want = np.array([want[i] for i in range(n)])
    assert np.allclose(X[:, 0], want, atol=1e-10)
    assert not np.allclose(bad, want, atol=1e-3)
