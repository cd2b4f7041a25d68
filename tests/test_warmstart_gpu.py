"""GPU checks of the learned warm start: forward, features and the hand-written backward
kernel against the FP64 oracle (oracle/mlp.py), one training step lowering the loss, and the
warm-started power iteration reaching the seed-start fixed point."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import page_rank_amd as graph_link
from page_rank_amd import graphs
from page_rank_amd import warmstart as ws
from oracle import mlp as omlp
from oracle import ppr as oppr

pytestmark = pytest.mark.gpu


def _params(seed=0):
    rng = np.random.default_rng(seed)
    return rng.normal(0, 0.25, omlp.NUM_PARAMS).astype(np.float32)


def test_forward_matches_oracle(device):
    p = _params()
    F = np.random.default_rng(1).normal(0, 1, (1000, 12)).astype(np.float32)
    s = ws.forward(torch.as_tensor(p, device=device), torch.as_tensor(F, device=device))
    so, _ = omlp.forward(p, F)
    assert np.allclose(s.cpu().numpy(), so, rtol=1e-4, atol=1e-5)


@pytest.mark.parametrize("M", [1, 63, 64, 1000, 70_001])
def test_gradient_matches_oracle(M, device):
    p = _params(2)
    rng = np.random.default_rng(M)
    F = rng.normal(0, 1, (M, 12)).astype(np.float32)
    t = (rng.random(M) * 2).astype(np.float32)
    w = (rng.random(M) + 0.5).astype(np.float32)
    g, loss = ws.grad(torch.as_tensor(p, device=device), torch.as_tensor(F, device=device),
                      torch.as_tensor(t, device=device), torch.as_tensor(w, device=device))
    go = omlp.grad(p, F, t, w)
    lo = omlp.loss(p, F, t, w)
    assert abs(loss - lo) <= 1e-4 * abs(lo) + 1e-5
    g = g.cpu().numpy()
    scale = np.abs(go).max()
    assert np.abs(g - go).max() <= 2e-4 * scale + 1e-6, np.abs(g - go).max() / scale


def test_train_step_lowers_loss(device):
    rng = np.random.default_rng(5)
    F = rng.normal(0, 1, (20000, 12)).astype(np.float32)
    t = (np.abs(F[:, 0]) * 2 + 0.1).astype(np.float32)
    w = np.ones(20000, np.float32)
    p = ws.init_params(0, device)
    Ft, tt, wt = (torch.as_tensor(a, device=device) for a in (F, t, w))
    l0 = omlp.loss(p.cpu().numpy(), F, t, w)
    log = ws.train(p, Ft, tt, wt, epochs=20, batch=4096, lr=3e-3)
    l1 = omlp.loss(p.cpu().numpy(), F, t, w)
    assert l1 < 0.5 * l0, (l0, l1, log.losses)


def _graph(n=2000, seed=0):
    rng = np.random.default_rng(seed)
    src = np.repeat(np.arange(n), 6)
    comm = src // 100
    dst = np.where(rng.random(src.size) < 0.8, comm * 100 + rng.integers(0, 100, src.size),
                   rng.integers(0, n, src.size))
    A, dang = graphs.transition_matrix(src, dst, n)
    return A, dang


def test_features_match_oracle(device):
    A, _ = _graph()
    pbr = graph_link.csr_to_pbr(A, 8, 8, 2).to(device)
    sets = [[1, 5], [100], [7, 8, 9]]
    wi = ws.inputs(pbr, sets, 0.85)
    indeg = np.diff(A.indptr)
    outdeg = np.bincount(A.indices, minlength=A.shape[0])
    assert np.array_equal(wi.indeg.cpu().numpy(), indeg)
    assert np.array_equal(wi.outdeg.cpu().numpy(), outdeg)
    nb_want = omlp.deg_bucket(indeg) * 193 + omlp.deg_bucket(outdeg)
    assert np.array_equal(wi.nb.cpu().numpy(), nb_want.astype(np.int64))
    P = oppr.personalization_matrix(A.shape[0], sets)
    v = [P]
    for _ in range(4):
        v.append(A.astype(np.float64) @ v[-1])
    pairs = np.arange(0, A.shape[0] * 3, 7)
    F = ws.pair_features(wi, torch.as_tensor(pairs, device=device)).cpu().numpy()
    Fo, hop = omlp.features(indeg, outdeg, v, 0.85, pairs // 3, pairs % 3)
    assert np.array_equal(F[:, 2:8], Fo[:, 2:8])       # hop one-hot exact
    assert np.allclose(F, Fo, rtol=1e-4, atol=1e-5)


def test_warm_start_reaches_the_same_fixed_point(device):
    A, dang = _graph(3000, 1)
    pbr = graph_link.csr_to_pbr(A, 8, 8, 2).to(device)
    sets = [[3], [10, 20, 30], list(range(500, 540))]
    p = ws.init_params(1, device)
    wi = ws.inputs(pbr, sets, 0.85)
    x0, n_near = ws.predict(p, wi)
    x = x0.cpu().numpy()
    assert n_near > 0 and np.allclose(x.sum(0), 1, atol=1e-5) and (x >= 0).all()
    cold = graph_link.run_ppr(pbr, sets, tol=1e-7, max_iter=300, dangling=dang,
                              dangling_mode="uniform")
    warm = graph_link.run_ppr(pbr, sets, tol=1e-7, max_iter=300, dangling=dang,
                              dangling_mode="uniform", x0=x0)
    assert cold.converged.all() and warm.converged.all()
    assert np.abs(cold.X.cpu().numpy() - warm.X.cpu().numpy()).sum(0).max() < 5e-6


def test_trained_network_saves_iterations(device):
    """Train on four graphs of a family, evaluate on a fifth: the learned start is closer to
    x* than the seed start and converges in fewer iterations to the same result."""
    samples = []
    for g in range(5):
        A, dang = _graph(20_000, 10 + g)
        pbr = graph_link.csr_to_pbr(A, 8, 8, 2).to(device)
        sets = graphs.seed_sets(20_000, 64, 1, 5, seed=g)
        samples.append((pbr, sets, dang, "uniform"))
    params, log = ws.fit(samples[:4], far_samples=1 << 16, epochs=6, batch=1 << 14)
    assert log.losses[-1] < log.losses[0]
    pbr, sets, dang, dmode = samples[4]
    ev = ws.evaluate(params, pbr, sets, dang, dmode, tol=1e-6)
    print(f"\nplain iters {ev.plain_iters.mean():.1f}  warm iters {ev.warm_iters.mean():.1f}  "
          f"start L1 plain {ev.start_l1_plain.mean():.3f} warm {ev.start_l1_warm.mean():.3f}  "
          f"ms plain {ev.plain_ms:.2f} warm {ev.warm_ms:.2f} (predict {ev.predict_ms:.2f})  "
          f"losses {log.losses}")
    assert ev.max_l1_diff < 1e-5
    assert ev.start_l1_warm.mean() < ev.start_l1_plain.mean()
    assert ev.warm_iters.mean() < ev.plain_iters.mean()


def test_assembly_lists_every_near_pair_once(device):
    """1M nodes, 130 seed sets of 20,000 nodes: most (node, set) pairs lie within two hops,
    so every wave of the assembly kernel flushes its near-pair buffer several times. The
    near list must hold each pair with hop <= 2 exactly once (a lost pair would keep x0 = 0,
    a repeated one would inflate the count), and every entry of X0 must be the network's
    output for that pair's features divided by its column's sum, the same bits on every call."""
    n, K = 1_000_000, 130
    rng = np.random.default_rng(7)
    src = np.repeat(np.arange(n), 6)
    dst = rng.integers(0, n, src.size)
    A, _ = graphs.transition_matrix(src, dst, n)
    pbr = graph_link.csr_to_pbr(A, 8, 8, 4).to(device)
    sets = [rng.choice(n, 20_000, replace=False) for _ in range(K)]
    p = ws.init_params(3, device)
    wi = ws.inputs(pbr, sets, 0.85)
    want = int(((wi.v[0] > 0) | (wi.v[1] > 0) | (wi.v[2] > 0)).sum().item())
    x0, n_near = ws.predict(p, wi, blend=None)
    assert n_near == want and n_near > n * K // 4
    assert bool((x0 > 0).all())
    # a sample of pairs through the per-pair path: ratio to X0 is 1 / column sum
    pairs = torch.as_tensor(np.sort(rng.choice(n * K, 200_000, replace=False)), device=device)
    ref = ws.forward(p, ws.pair_features(wi, pairs))
    got = x0.view(-1)[pairs]
    ratio = (got / ref).cpu().numpy()
    cols = (pairs % K).cpu().numpy()
    for k in range(K):
        r = ratio[cols == k]
        assert np.allclose(r, r[0], rtol=2e-5), k
    assert torch.allclose(x0.sum(0), torch.ones(K, device=device), atol=1e-4)
    # a near list smaller than the pair count is refused and predict retries with the count
    x1, n1 = ws.predict(p, wi, near_cap=1 << 16, blend=None)
    assert n1 == want
    # every column sum is taken in a fixed order: the start vector repeats bit for bit
    assert torch.equal(x1, x0)
    # the default start: the 4th power iterate moved BLEND of the way toward x0, renormalized
    xb, _ = ws.predict(p, wi)
    v = [t.double() for t in wi.v]
    a = 0.85
    x4 = (1 - a) * (v[0] + a * v[1] + a * a * v[2] + a ** 3 * v[3]) + a ** 4 * v[4]
    want_b = (x4 + ws.BLEND * (x0.double() - x4)).clamp_min(0)
    want_b = want_b / want_b.sum(0, keepdim=True)
    assert torch.allclose(xb.double(), want_b, rtol=1e-4, atol=1e-9)
    assert torch.allclose(xb.sum(0), torch.ones(K, device=device), atol=1e-4)
