"""Learned warm start for batched PPR.

A 12 -> 64 -> 64 -> 1 MLP (ReLU, ReLU, softplus) maps per-(node, seed set) features to an
unnormalized score; each seed set's column is normalized to sum 1 and used as the start
vector of the power iteration (run_ppr(..., x0=...)). The plain seed-start iteration is both
the fallback and the yardstick: the warm start must reach the same residual, and what it buys
is iterations and wall time. Every kernel (degrees, SpMM for the feature terms, features, forward, backward, Adam,
assembly) is in libgraphlink_hip.so.

Features: degree buckets (in, out), one-hot hop distance to the seed set (0..4, more), the
first three power-iteration terms q_1..q_3 as log1p(N q)/8, and a constant. Pairs beyond
two hops depend only on (degree buckets, hop class), so they read a table evaluated once;
pairs within two hops are evaluated one by one.
"""
from __future__ import annotations

import ctypes
import os
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _native
from .ops import spmm_into
from .registry import get_pbr_gpu_meta


def _lib():
    return _native.lib()


def _ptrs(ts):
    return (ctypes.c_void_p * len(ts))(*[t.data_ptr() for t in ts])


def init_params(seed=0, device="cuda"):
    """He-initialized parameters (bias of the output set so softplus starts near 1)."""
    L = _lib()
    n_in, hid = L.gl_mlp_num_features(), 64
    g = np.random.default_rng(seed)
    W1 = g.normal(0, np.sqrt(2 / n_in), (hid, n_in))
    W2 = g.normal(0, np.sqrt(2 / hid), (hid, hid))
    w3 = g.normal(0, np.sqrt(1 / hid), hid)
    p = np.concatenate([W1.ravel(), np.zeros(hid), W2.ravel(), np.zeros(hid), w3, [0.5]])
    assert p.size == L.gl_mlp_num_params()
    return torch.as_tensor(p.astype(np.float32), device=device)


@dataclass
class WsInputs:
    """Device inputs of the feature kernels for one graph and one batch of seed sets."""
    N: int
    K: int
    alpha: float
    indeg: torch.Tensor
    outdeg: torch.Tensor
    nb: torch.Tensor                        # table row of every node
    v: list = field(default_factory=list)  # A^j P, j = 0..4, each [N][K]

    def vptrs(self):
        return _ptrs(self.v)


def degrees(pbr):
    meta = get_pbr_gpu_meta(pbr)
    if meta is None:
        raise RuntimeError("PBRMatrix is not on GPU. Call .to('cuda') first.")
    if pbr.block_rows != pbr.block_cols or pbr.rows != pbr.cols:
        raise ValueError("degrees need a square matrix with square blocks")
    dev = meta["codes"].device
    n = pbr.rows
    indeg = torch.empty(n, dtype=torch.int32, device=dev)
    outdeg = torch.empty(n, dtype=torch.int32, device=dev)
    _native.check(_lib().gl_ws_degrees(n, pbr.block_rows, meta["codes"].numel(),
                                       meta["codes"].data_ptr(), meta["coords"].data_ptr(),
                                       meta["rem_indptr"].data_ptr(),
                                       meta["rem_col_ind"].data_ptr(), indeg.data_ptr(),
                                       outdeg.data_ptr(), _native.stream_handle(dev)),
                  "gl_ws_degrees")
    return indeg, outdeg


def graph_features(pbr):
    """Per-graph inputs of the feature kernels, computed once per matrix: in- and
    out-degree and the table row (degree buckets) of every node."""
    indeg, outdeg = degrees(pbr)
    nb = torch.empty_like(indeg)
    _native.check(_lib().gl_ws_node_buckets(indeg.numel(), indeg.data_ptr(), outdeg.data_ptr(),
                                            nb.data_ptr(), _native.stream_handle(indeg.device)),
                  "gl_ws_node_buckets")
    return indeg, outdeg, nb


def seed_matrix(n, seed_sets, device):
    """P [n][K]: column k uniform over seed set k (duplicates add up); one upload."""
    K = len(seed_sets)
    sets = [np.atleast_1d(np.asarray(s, dtype=np.int64)) for s in seed_sets]
    flat = np.concatenate([s * K + k for k, s in enumerate(sets)]) if K else np.zeros(0, np.int64)
    w = np.concatenate([np.full(s.size, 1.0 / s.size, np.float32) for s in sets]) if K else \
        np.zeros(0, np.float32)
    P = torch.zeros((n, K), dtype=torch.float32, device=device)
    P.view(-1).index_add_(0, torch.as_tensor(flat).to(device), torch.as_tensor(w).to(device))
    return P


def inputs(pbr, seed_sets, alpha=0.85, graph=None) -> WsInputs:
    """Feature inputs of one batch of seed sets; `graph` = graph_features(pbr), computed
    here when not given."""
    meta = get_pbr_gpu_meta(pbr)
    if meta is None:
        raise RuntimeError("PBRMatrix is not on GPU. Call .to('cuda') first.")
    dev = meta["codes"].device
    indeg, outdeg, nb = graph_features(pbr) if graph is None else graph
    v = [seed_matrix(pbr.rows, seed_sets, dev)]
    for _ in range(4):
        y = torch.empty_like(v[0])
        spmm_into(pbr, v[-1], y)
        v.append(y)
    return WsInputs(pbr.rows, len(seed_sets), alpha, indeg, outdeg, nb, v)


def forward(params, feat):
    out = torch.empty(feat.shape[0], dtype=torch.float32, device=feat.device)
    _native.check(_lib().gl_mlp_forward(params.data_ptr(), feat.data_ptr(), feat.shape[0],
                                        out.data_ptr(), _native.stream_handle(feat.device)),
                  "gl_mlp_forward")
    return out


def pair_features(wi: WsInputs, pairs):
    dev = wi.v[0].device
    feat = torch.empty((pairs.numel(), 12), dtype=torch.float32, device=dev)
    _native.check(_lib().gl_ws_pair_features(wi.N, wi.K, wi.alpha, wi.indeg.data_ptr(),
                                             wi.outdeg.data_ptr(), wi.vptrs(),
                                             pairs.data_ptr(), pairs.numel(), feat.data_ptr(),
                                             _native.stream_handle(dev)),
                  "gl_ws_pair_features")
    return feat


# Weight of the network's start against the 4th power iterate that its feature products
# already hold (gl_ws_blend_start). Measured on the 4M-node clustered graph (64 seed sets,
# tol 1.5e-7): lam 0 (the iterate alone) 41 passes after the 4 products, 0.25 37, 0.5 42,
# 0.75 44, 1 (the network alone, no products counted) 45.
BLEND = 0.25


def predict(params, wi: WsInputs, near_cap=None, blend=BLEND):
    """Normalized warm-start matrix X0 [N][K] and the number of per-pair evaluations.
    blend: None = the network's start alone; lam = the 4th power iterate moved lam of the way
    toward it (the feature products then count as the run's first four passes)."""
    L = _lib()
    dev = wi.v[0].device
    st = _native.stream_handle(dev)
    tsz = L.gl_ws_table_size()
    tfeat = torch.empty((tsz, 12), dtype=torch.float32, device=dev)
    _native.check(L.gl_ws_table_features(tfeat.data_ptr(), st), "gl_ws_table_features")
    table = forward(params, tfeat)
    x0 = torch.empty((wi.N, wi.K), dtype=torch.float32, device=dev)
    # with a blend, the 4th iterate comes out of the same pass over the feature products
    x4 = torch.empty_like(x0) if blend is not None else None
    cap = near_cap if near_cap is not None else max(1 << 16, wi.N * wi.K // 4)
    near = torch.empty(cap, dtype=torch.int64, device=dev)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    _native.check(L.gl_ws_assemble(wi.N, wi.K, wi.alpha, wi.indeg.data_ptr(),
                                   wi.outdeg.data_ptr(), wi.vptrs(), wi.nb.data_ptr(),
                                   table.data_ptr(), x0.data_ptr(), near.data_ptr(), cap,
                                   cnt.data_ptr(), None if x4 is None else x4.data_ptr(), st),
                  "gl_ws_assemble")
    n_near = int(cnt.item())
    if n_near > cap:
        return predict(params, wi, near_cap=n_near, blend=blend)
    pairs = near[:n_near]
    if n_near:
        vals = forward(params, pair_features(wi, pairs))
        _native.check(L.gl_ws_scatter(pairs.data_ptr(), vals.data_ptr(), n_near, x0.data_ptr(), st),
                      "gl_ws_scatter")
    if blend is None:
        # the column sums, then the workgroups' partials (at most 1024 of K floats)
        sums = torch.empty(wi.K * 1025, dtype=torch.float32, device=dev)
        _native.check(L.gl_ws_normalize(x0.data_ptr(), wi.N, wi.K, sums.data_ptr(), st),
                      "gl_ws_normalize")
    else:
        sums = torch.empty(wi.K * 1026, dtype=torch.float32, device=dev)
        _native.check(L.gl_ws_blend_start(x0.data_ptr(), x4.data_ptr(), wi.N, wi.K,
                                          float(blend), sums.data_ptr(), st),
                      "gl_ws_blend_start")
    return x0, n_near


def run_warm(params, pbr, seed_sets, *, graph=None, alpha=0.85, blend=BLEND, **run_kw):
    """PPR from the learned start: features, network, then run_ppr from X0 (run_kw as for
    run_ppr). Returns run_ppr's result; with a blend, the four feature products are passes of
    the same power iteration, and `feature_passes` on the result says so (4, else 0)."""
    from .ppr import run_ppr
    wi = inputs(pbr, seed_sets, alpha, graph=graph)
    x0, _ = predict(params, wi, blend=blend)
    del wi
    res = run_ppr(pbr, seed_sets, alpha=alpha, x0=x0, x0_in_place=True, **run_kw)
    res.feature_passes = 4 if blend is not None else 0
    return res


def training_pairs(wi: WsInputs, x_star: torch.Tensor, far_samples: int, seed=0,
                   near_samples=None):
    """Training set of one graph: the pairs within two hops (all of them, or
    `near_samples` drawn uniformly with weight n_near / near_samples) plus `far_samples`
    uniformly drawn pairs beyond (weight N*K / far_samples), which makes the weighted loss
    an unbiased estimate of the full-column L1. Targets are N * x*."""
    L = _lib()
    dev = wi.v[0].device
    st = _native.stream_handle(dev)
    cap = max(1 << 16, wi.N * wi.K // 4)
    near = torch.empty(cap, dtype=torch.int64, device=dev)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    dummy_table = torch.zeros(L.gl_ws_table_size(), dtype=torch.float32, device=dev)
    x0 = torch.empty((wi.N, wi.K), dtype=torch.float32, device=dev)
    _native.check(L.gl_ws_assemble(wi.N, wi.K, wi.alpha, wi.indeg.data_ptr(),
                                   wi.outdeg.data_ptr(), wi.vptrs(), wi.nb.data_ptr(),
                                   dummy_table.data_ptr(), x0.data_ptr(), near.data_ptr(), cap,
                                   cnt.data_ptr(), None, st),
                  "gl_ws_assemble")
    n_near = min(int(cnt.item()), cap)
    del x0
    g = torch.Generator(device="cpu").manual_seed(seed)
    near = near[:n_near]
    w_near = 1.0
    if near_samples is not None and n_near > near_samples:
        pick = torch.randperm(n_near, generator=g)[:near_samples].to(dev)
        near = near[pick]
        w_near = n_near / near_samples
        n_near = near_samples
    far = torch.randint(0, wi.N * wi.K, (far_samples,), generator=g).to(dev)
    pairs = torch.cat([near, far])
    feat = pair_features(wi, pairs)
    # drop far draws that landed within two hops (they are already in the near list)
    hop_far = feat[n_near:, 2:5].sum(1) == 0
    keep = torch.cat([torch.ones(n_near, dtype=torch.bool, device=dev), hop_far])
    pairs, feat = pairs[keep], feat[keep]
    n_far = int(hop_far.sum().item())
    w = torch.full((pairs.numel(),), float(w_near), dtype=torch.float32, device=dev)
    w[n_near:] = float(wi.N * wi.K) / far_samples
    target = x_star.reshape(-1)[pairs] * wi.N
    return feat, target, w, n_near, n_far


@dataclass
class TrainLog:
    losses: list
    seconds: float


def train(params, feat, target, weight, *, epochs=8, batch=1 << 18, lr=2e-3, seed=0):
    """Adam on the weighted L1 loss with minibatches; all work in HIP kernels."""
    L = _lib()
    dev = feat.device
    st = _native.stream_handle(dev)
    m1 = torch.zeros_like(params)
    m2 = torch.zeros_like(params)
    partial = torch.empty(L.gl_mlp_train_partials(), dtype=torch.float32, device=dev)
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    M = feat.shape[0]
    g = torch.Generator(device="cpu").manual_seed(seed)
    step = 0
    losses = []
    t0 = time.perf_counter()
    total_w = float(weight.sum().item())
    for ep in range(epochs):
        perm = torch.randperm(M, generator=g).to(dev)
        loss.zero_()
        for b in range(0, M, batch):
            idx = perm[b:b + batch]
            f, t, w = feat[idx].contiguous(), target[idx].contiguous(), weight[idx].contiguous()
            step += 1
            scale = 1.0 / max(float(idx.numel()), 1.0) * (M / total_w)
            _native.check(L.gl_mlp_train_step(params.data_ptr(), f.data_ptr(), t.data_ptr(),
                                              w.data_ptr(), idx.numel(), scale, lr,
                                              m1.data_ptr(), m2.data_ptr(), step,
                                              partial.data_ptr(), loss.data_ptr(), st),
                          "gl_mlp_train_step")
        losses.append(float(loss.item()) / total_w)
    torch.cuda.synchronize(dev)
    return TrainLog(losses, time.perf_counter() - t0)


def grad(params, feat, target, weight):
    L = _lib()
    dev = feat.device
    partial = torch.empty(L.gl_mlp_train_partials(), dtype=torch.float32, device=dev)
    g = torch.empty(L.gl_mlp_num_params(), dtype=torch.float32, device=dev)
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    _native.check(L.gl_mlp_grad(params.data_ptr(), feat.data_ptr(), target.data_ptr(),
                                weight.data_ptr(), feat.shape[0], partial.data_ptr(),
                                g.data_ptr(), loss.data_ptr(), _native.stream_handle(dev)),
                  "gl_mlp_grad")
    return g, float(loss.item())


# the committed weights, trained on the clustered generator's 1M-node / 16M-edge graphs
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                       "warmstart_clustered.npz")


def save(params, path):
    np.savez(path, params=params.detach().cpu().numpy())


def load(path, device="cuda"):
    with np.load(path) as z:
        return torch.as_tensor(z["params"], device=device)


def fit(samples, params=None, *, alpha=0.85, far_samples=1 << 18, near_samples=None, epochs=8,
        batch=1 << 16, lr=2e-3, tol=1e-7, max_iter=400, seed=0):
    """Train on a list of (pbr, seed_sets, dangling_mask, dangling_mode) graphs: the
    exact solution of each is computed by run_ppr, then pairs are drawn by
    training_pairs and the network is fitted with train. Returns (params, TrainLog)."""
    from .ppr import run_ppr
    feats, targets, weights = [], [], []
    dev = None
    for j, (pbr, sets, dang, dmode) in enumerate(samples):
        wi = inputs(pbr, sets, alpha)
        dev = wi.v[0].device
        ref = run_ppr(pbr, sets, alpha=alpha, tol=tol, max_iter=max_iter, dangling=dang,
                      dangling_mode=dmode, return_history=False)
        f, t, w, _, _ = training_pairs(wi, ref.X, far_samples, seed=seed + j,
                                       near_samples=near_samples)
        del wi, ref
        feats.append(f), targets.append(t), weights.append(w)
    if params is None:
        params = init_params(seed, dev)
    log = train(params, torch.cat(feats), torch.cat(targets), torch.cat(weights),
                epochs=epochs, batch=batch, lr=lr, seed=seed)
    return params, log


@dataclass
class WarmStartEval:
    """Plain (seed start) against warm start at the same tolerance."""
    plain_iters: np.ndarray
    warm_iters: np.ndarray
    plain_ms: float
    warm_ms: float          # features + network + power iteration
    predict_ms: float       # features + network alone
    start_l1_plain: np.ndarray   # ||x0 - x*||_1 per column
    start_l1_warm: np.ndarray
    max_l1_diff: float      # between the two converged results


def evaluate(params, pbr, seed_sets, dangling=None, dangling_mode="none", *, alpha=0.85,
             tol=1e-6, max_iter=400, x_star=None, reps=3):
    """Plain and warm-started runs to the same tol; times are medians over `reps` runs, the
    warm time including features and network."""
    from .ppr import run_ppr
    dev = get_pbr_gpu_meta(pbr)["codes"].device
    kw = dict(alpha=alpha, tol=tol, max_iter=max_iter, dangling=dangling,
              dangling_mode=dangling_mode, return_history=False)
    run_ppr(pbr, seed_sets, **kw)                       # warm up graph capture / caches
    graph = graph_features(pbr)       # per graph, like the matrix itself
    t_plain, t_warm, t_pred = [], [], []
    for _ in range(max(1, reps)):
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        plain = run_ppr(pbr, seed_sets, **kw)
        torch.cuda.synchronize(dev)
        t1 = time.perf_counter()
        wi = inputs(pbr, seed_sets, alpha, graph=graph)
        x0, _ = predict(params, wi)
        del wi
        torch.cuda.synchronize(dev)
        t2 = time.perf_counter()
        warm = run_ppr(pbr, seed_sets, x0=x0, **kw)
        torch.cuda.synchronize(dev)
        t3 = time.perf_counter()
        t_plain.append(t1 - t0)
        t_warm.append(t3 - t1)
        t_pred.append(t2 - t1)
    if x_star is None:
        x_star = run_ppr(pbr, seed_sets, **{**kw, "tol": min(tol, 1e-7)}).X
    P = seed_matrix(pbr.rows, seed_sets, dev)
    return WarmStartEval(
        plain.iterations, warm.iterations, float(np.median(t_plain)) * 1e3,
        float(np.median(t_warm)) * 1e3, float(np.median(t_pred)) * 1e3,
        (P - x_star).abs().sum(0).cpu().numpy(), (x0 - x_star).abs().sum(0).cpu().numpy(),
        float((plain.X - warm.X).abs().sum(0).max().item()))


def clustered_sample(n, m, sets, seed, block=8, min_nnz=4):
    """One graph of the block-clustered generator (communities of 1,000 nodes) with `sets`
    seed sets of U{1..100} nodes, built on the GPU: (pbr, seed_sets, dangling, mode)."""
    from . import graphs
    from .builder import csr_to_pbr
    src, dst = graphs.clustered_edges(n, m, communities=max(1, n // 1000), seed=seed)
    A, dang = graphs.transition_matrix(src, dst, n)
    del src, dst
    pbr = csr_to_pbr(A, block, block, min_nnz, order="canonical")
    return (pbr, graphs.seed_sets(n, sets, 1, 100, seed=seed + 100000),
            dang if dang.any() else None, "uniform" if dang.any() else "none")


def main(argv=None):
    """Train on seeded clustered graphs and evaluate on held-out ones:

        python -m page_rank_amd.warmstart --graphs 20 --nodes 1000000 --edges 16000000 \
            --sets 100 --out page_rank_amd/data/warmstart_clustered.npz

    or evaluate saved weights only:

        python -m page_rank_amd.warmstart --load page_rank_amd/data/warmstart_clustered.npz
    """
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("--graphs", type=int, default=20)
    ap.add_argument("--heldout", type=int, default=2)
    ap.add_argument("--nodes", type=int, default=1_000_000)
    ap.add_argument("--edges", type=int, default=16_000_000)
    ap.add_argument("--sets", type=int, default=100)
    ap.add_argument("--far-samples", type=int, default=1 << 18)
    ap.add_argument("--near-samples", type=int, default=1 << 18)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--out", default=None)
    ap.add_argument("--load", default=None, help="evaluate these weights; no training")
    a = ap.parse_args(argv)
    if a.load:
        params = load(a.load, torch.device("cuda", torch.cuda.current_device()))
        _evaluate_heldout(params, a, train_seconds=None, losses=None)
        return
    t0 = time.perf_counter()
    feats, targets, weights = [], [], []
    dev = None
    for g in range(a.graphs):
        pbr, sets, dang, dmode = clustered_sample(a.nodes, a.edges, a.sets, seed=1000 + g)
        wi = inputs(pbr, sets)
        dev = wi.v[0].device
        from .ppr import run_ppr
        ref = run_ppr(pbr, sets, tol=1e-7, max_iter=400, dangling=dang, dangling_mode=dmode,
                      return_history=False)
        f, t, w, n_near, n_far = training_pairs(wi, ref.X, a.far_samples, seed=g,
                                                near_samples=a.near_samples)
        feats.append(f), targets.append(t), weights.append(w)
        print(f"graph {g}: {n_near} near + {n_far} far pairs, "
              f"{time.perf_counter() - t0:.1f} s", flush=True)
        del wi, ref, pbr
    params = init_params(0, dev)
    log = train(params, torch.cat(feats), torch.cat(targets), torch.cat(weights),
                epochs=a.epochs, batch=1 << 16, lr=2e-3)
    print(f"trained in {log.seconds:.1f} s, losses {[round(v, 4) for v in log.losses]}",
          flush=True)
    del feats, targets, weights
    if a.out:
        save(params, a.out)
    _evaluate_heldout(params, a, round(log.seconds, 2), log.losses)


def _evaluate_heldout(params, a, train_seconds, losses):
    import json
    evals = []
    for h in range(a.heldout):
        pbr, sets, dang, dmode = clustered_sample(a.nodes, a.edges, a.sets, seed=5000 + h)
        ev = evaluate(params, pbr, sets, dang, dmode, tol=1e-6)
        evals.append({"plain_iters_mean": float(ev.plain_iters.mean()),
                      "warm_iters_mean": float(ev.warm_iters.mean()),
                      "plain_ms": round(ev.plain_ms, 3), "warm_ms": round(ev.warm_ms, 3),
                      "predict_ms": round(ev.predict_ms, 3),
                      "start_l1_plain": float(ev.start_l1_plain.mean()),
                      "start_l1_warm": float(ev.start_l1_warm.mean()),
                      "max_l1_diff": ev.max_l1_diff})
        print(json.dumps(evals[-1]), flush=True)
    print(json.dumps({"graphs": a.graphs if losses is not None else 0, "nodes": a.nodes,
                      "edges": a.edges, "sets": a.sets, "train_seconds": train_seconds,
                      "losses": losses, "heldout": evals}), flush=True)


if __name__ == "__main__":
    main()
