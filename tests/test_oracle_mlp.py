"""CPU pins of the warm-start network oracle (oracle/mlp.py): the analytic FP64 gradient
against central finite differences, and feature values against hand-computed ones."""
import numpy as np

from oracle import mlp as omlp


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, omlp.NUM_PARAMS)
    F = rng.normal(0, 1, (40, omlp.IN))
    t = rng.random(40) * 3
    w = rng.random(40) + 0.5
    g = omlp.grad(p, F, t, w)
    idx = np.r_[0:5, omlp.OFF_B1:omlp.OFF_B1 + 3, omlp.OFF_W2 + 100:omlp.OFF_W2 + 105,
                omlp.OFF_B2, omlp.OFF_W3 + 7, omlp.OFF_B3]
    fd = omlp.finite_difference(p, F, t, w, idx)
    assert np.allclose(g[idx], fd, rtol=1e-5, atol=1e-7)


def test_planted_sign_error_is_caught():
    rng = np.random.default_rng(1)
    p = rng.normal(0, 0.3, omlp.NUM_PARAMS)
    F = rng.normal(0, 1, (20, omlp.IN))
    t, w = rng.random(20), np.ones(20)
    g = omlp.grad(p, F, t, w)
    fd = omlp.finite_difference(p, F, t, w, [omlp.OFF_B3])
    assert np.isclose(g[omlp.OFF_B3], fd[0], rtol=1e-5)
    assert not np.isclose(-g[omlp.OFF_B3], fd[0], rtol=1e-2)


def test_features_by_hand():
    # path 0 -> 1 -> 2 -> 3 -> 4 -> 5 -> 6 as column-stochastic A[target, source]
    n = 7
    A = np.zeros((n, n))
    for i in range(n - 1):
        A[i + 1, i] = 1.0
    P = np.zeros((n, 1))
    P[0, 0] = 1
    v = [P]
    for _ in range(4):
        v.append(A @ v[-1])
    indeg = np.array([0, 1, 1, 1, 1, 1, 1])
    outdeg = np.array([1, 1, 1, 1, 1, 1, 0])
    F, hop = omlp.features(indeg, outdeg, v, 0.85, np.arange(n), np.zeros(n, int))
    assert hop.tolist() == [0, 1, 2, 3, 4, 5, 5]
    assert F[0, 0] == 0 and F[1, 0] == 1.0          # bucket(1) = round(8*log2 2)/8 = 1
    assert F[6, 1] == 0
    assert np.isclose(F[0, 8], np.log1p(n * 0.15) / 8)   # q1 = (1-a) at the seed
    assert np.isclose(F[1, 9], np.log1p(n * 0.15 * 0.85) / 8)   # q2 at hop 1
    assert np.isclose(F[2, 10], np.log1p(n * 0.15 * 0.85 ** 2) / 8)
    assert F[3, 8:11].tolist() == [0, 0, 0]
    assert (F[:, 11] == 1).all()


def test_degree_bucket_is_the_exact_rounding():
    """b(d) = round(8 log2(1+d)) counts the b with (1+d)^16 >= 2^(2b+1); checked in exact
    integer arithmetic, including the degrees closest to a rounding tie."""
    ds = list(range(0, 2000)) + [16065916, 2 ** 24 - 1, 2 ** 24, 123456, 999999]
    rng = np.random.default_rng(0)
    ds += rng.integers(0, 2 ** 24, 500).tolist()
    exact = [min(192, sum(1 for b in range(193) if (1 + d) ** 16 >= 2 ** (2 * b + 1)))
             for d in ds]
    assert omlp.deg_bucket(np.array(ds)).astype(int).tolist() == exact
