#!/usr/bin/env python3
"""PBR SpMM benchmark on one MI355X (gfx950).

Workload (a "step"): one SpMM Y = A X on the 1M x 1M block-structured matrix — 125k block
rows x 8 occupied 8x8 tiles within +-512 blocks of the diagonal, U{4..40} nonzeros per tile,
5M scattered nonzeros in the CSR remainder (~27M nnz), min_nnz_per_block=4, FP32 values and
RHS U(0,1) — built by the GPU builder, with X resident in HBM. The headline is RHS width
K=32; K=1 and K=8 are reported beside it. Three copies of the matrix rotate between steps so
that no step finds its matrix in the 256 MB Infinity Cache left behind by the previous one.

Prints one JSON line: metric/value (GNZ/s = nnz*K/t/1e9 over the timed steps), the roofline
of the SpMM kernel (algorithmic bytes 20T + 4V + 4(N+1) + 8R + 8NK per launch over the mean
launch time from HIP events on the launch stream, against the 8 TB/s HBM peak) and the CPU
baseline (the CPU restatement of the reference's pbr_batched_matmul_cpu, one thread, on a
row sample of the same matrix).

    python bench.py [--steps 20] [--warmup 5] [--k 32] [--widths 1,8,32] [--no-cpu]
    python bench.py --full-scale      # the 16M-node / 256M-edge PPR run only
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--k", type=int, default=32, help="headline RHS width")
    ap.add_argument("--widths", default="1,8,32", help="RHS widths measured beside it")
    ap.add_argument("--rotate", type=int, default=3, help="matrix copies rotated per step")
    ap.add_argument("--kw", type=int, default=0, help="columns per launch (0 = library)")
    ap.add_argument("--mode", default="fused", choices=["fused", "split"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--only-headline", action="store_true",
                    help="time only the headline width (for profiler runs)")
    ap.add_argument("--no-ppr", action="store_true", help="skip the batched-PPR leg")
    ap.add_argument("--ppr-only", action="store_true", help="run only the batched-PPR leg")
    ap.add_argument("--ppr-nodes", type=int, default=4_000_000)
    ap.add_argument("--ppr-edges", type=int, default=64_000_000)
    ap.add_argument("--ppr-sets", type=int, default=64)
    ap.add_argument("--ppr-set-max", type=int, default=100,
                    help="seed-set sizes are uniform in 1..this")
    ap.add_argument("--ppr-intra", type=float, default=0.8,
                    help="fraction of the PPR graph's edges inside their community")
    ap.add_argument("--ppr-community", type=int, default=1000,
                    help="nodes per community of the PPR graph")
    ap.add_argument("--ppr-profile-passes", type=int, default=0,
                    help="profiling aid: build the PPR graph, run this many plain passes of "
                         "the step (eager launches, no stop) twice, print their time, exit")
    ap.add_argument("--ppr-check-cols", type=int, default=8,
                    help="columns checked against the FP64 CPU power iteration (and timed "
                         "as the CPU baseline of the PPR leg)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first run 3 device-to-device copies of this many MiB (a kernel of "
                         "known traffic that scales the profiler's FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--full-scale", action="store_true",
                    help="only the full-scale PPR run: 16M nodes, 256M edges, 16,000 "
                         "communities, 128 seed sets of U{1..200} nodes, learned warm start "
                         "beside plain, FP64 errors on 8 sampled columns")
    args = ap.parse_args()
    if args.full_scale:
        args.ppr_only = True
        args.ppr_nodes, args.ppr_edges = 16_000_000, 256_000_000
        args.ppr_sets, args.ppr_set_max, args.ppr_check_cols = 128, 200, 8
    return args


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmm_bytes(T, V, R, N, K):
    return 20 * T + 4 * V + 4 * (N + 1) + 8 * R + 8 * N * K


def clone_matrix(pbr, dev):
    import torch
    import page_rank_amd as gl
    meta = gl.get_pbr_gpu_meta(pbr)
    m2 = {k: v.clone() for k, v in meta.items()}
    return type(pbr)(pbr.rows, pbr.cols, pbr.block_rows, pbr.block_cols, pbr.total_nnz,
                     device_meta=m2)


def time_width(mats, K, steps, warmup, kw, mode, dev):
    import torch
    import page_rank_amd as gl
    N = mats[0].rows
    g = torch.Generator(device="cpu").manual_seed(K)
    X = torch.rand((N, K), generator=g, dtype=torch.float32).to(dev)
    Y = torch.empty((N, K), dtype=torch.float32, device=dev)
    plans = [gl.SpmmPlan(m, X, Y, mode=mode, kw=kw) for m in mats]
    for i in range(warmup):
        plans[i % len(plans)].run()
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        ev[i][0].record(st)
        plans[i % len(plans)].run()
        ev[i][1].record(st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = [a.elapsed_time(b) for a, b in ev]
    return wall, launch_ms, X, Y


def cpu_baseline(A, K, seconds):
    """The reference's own CPU SpMM, pbr_batched_matmul_cpu (pbr_matrix.cpp:17-68, compiled
    from its sources by `make -C oracle ref`, one thread), on a row sample of the matrix; the
    oracle's restatement of it when that module is not built."""
    import numpy as np
    rows = 125_000
    S = A[:rows]
    X = np.random.default_rng(1).random((A.shape[1], K), dtype=np.float32)
    ref = None
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle", "_ref"))
        import torch
        import graph_link_core as core
        import graph_link_ref_spmm as rs
        p = core.csr_to_pbr(S.indptr, S.indices, S.data, S.shape[0], S.shape[1], 8, 8, 4)
        Xt = torch.from_numpy(X)
        torch.set_num_threads(1)  # its loop is scalar; only the zero fill could thread
        ref = lambda: rs.pbr_batched_matmul_cpu(p, Xt)  # noqa: E731
        kind, what = "reference", ("the reference's pbr_batched_matmul_cpu "
                                   "(Backend/graph_link/pbr_matrix/pbr_matrix.cpp:17-68), "
                                   "compiled by make -C oracle ref")
    except ImportError as e:
        log(f"  reference CPU module not built ({e}); timing the oracle's restatement")
    if ref is None:
        from oracle import pbr as opbr
        p = opbr.csr_matrix_to_pbr(S, 8, 8, 4)
        ref = lambda: opbr.spmm(p, X)  # noqa: E731
        kind, what = "port", "oracle/pbr_oracle.cpp spmm (float accumulation)"
    ref()
    reps, t = 0, 0.0
    t0 = time.perf_counter()
    while t < seconds:
        ref()
        reps += 1
        t = time.perf_counter() - t0
    gnz = S.nnz * K * reps / t / 1e9
    out = {"value": round(gnz, 4), "unit": "GNZ/s", "cores": 1, "kind": kind,
           "sample": f"rows 0..{rows} of the matrix ({S.nnz} nnz), K={K}, {reps} products "
                     f"in {t:.1f} s, {what}"}
    out.update(host_info())
    out["scipy_csr_f32"] = scipy_spmm(A, K)
    return out


def host_info():
    """Core count and CPU model of the host the CPU baselines ran on."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"host_cores": os.cpu_count(), "host_cpu": model,
            "cores_used": f"1 of {os.cpu_count()}"}


def scipy_spmm(A, K, reps=3):
    """scipy's FP32 CSR product A @ X on the whole matrix (single-threaded), median of
    `reps` after one warm-up: the plain-CSR CPU path beside the PBR one."""
    import numpy as np
    X = np.random.default_rng(2).random((A.shape[1], K), dtype=np.float32)
    A.dot(X)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        A.dot(X)
        ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    return {"gnz_s": round(A.nnz * K / t / 1e9, 4), "ms": round(t * 1e3, 1), "K": K,
            "cores": 1, "sample": f"whole matrix ({A.nnz} nnz), median of {reps}"}


def builder_leg(A, T, V, R, dev):
    """GPU builder on a device-resident CSR (canonical tile order, no host transfer): wall
    time of count + scans + fill, against input CSR + output PBR bytes."""
    import torch
    import page_rank_amd as gl
    d = gl.upload_csr(A, dev)
    gl.csr_to_pbr(d, 8, 8, 4, order="canonical")
    ts = []
    for _ in range(5):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        gl.csr_to_pbr(d, 8, 8, 4, order="canonical")
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    t = min(ts)
    N = A.shape[0]
    nbytes = (4 * (N + 1) + 8 * A.nnz) + (20 * T + 4 * V + 4 * (N + 1) + 8 * R)
    gbs = nbytes / t / 1e9
    log(f"  builder on device CSR: {t*1e3:.3f} ms, {gbs:.0f} GB/s algorithmic")
    return {"ms": round(t * 1e3, 3), "bytes": nbytes, "achieved_gbs": round(gbs, 1),
            "frac": round(gbs / HBM_PEAK_GBS, 4),
            "note": "device CSR in, canonical tile order; wall time of count, scans, one "
                    "host sync and fill (the reference order adds a host replay)"}


def ppr_matrix_bytes(T, V, R, N):
    return 20 * T + 4 * V + 4 * (N + 1) + 8 * R


def _timed_runs(fn, reps=3):
    import numpy as np
    import torch
    fn()
    ts, out = [], None
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), out


def ppr_leg(args, dev):
    """Batched multi-seed PPR on the block-clustered power-law graph (by default 4M nodes,
    64M edges, 4,000 communities, 80 % intra, out-degree Pareto(2.1, min 4)), 64 seed sets
    of U{1..100} nodes, alpha 0.85, per-column retirement, hipGraph-captured loop, launch
    width from the autotuner; with --full-scale 16M nodes, 256M edges, 16,000 communities and
    128 sets of U{1..200}.

    PPR/s counts seed sets that reach L1 < 1e-6 against FP64 per second. The update-norm stop
    leaves up to alpha / (1 - alpha) times tol of error, so the headline runs stop at
    tol 1.5e-7, plain and learned warm start alike, and both are checked against the FP64 CPU
    power iteration on sampled columns (which is also the CPU baseline). The plain run at
    tol 1e-6 is reported beside them."""
    import numpy as np
    import torch
    import page_rank_amd as gl
    from page_rank_amd import graphs
    n, m, K = args.ppr_nodes, args.ppr_edges, args.ppr_sets
    log(f"generating the {n} node / {m} edge clustered graph ...")
    src, dst = graphs.clustered_edges(n, m, communities=max(1, n // args.ppr_community),
                                      intra=args.ppr_intra, seed=3)
    log("  edges drawn; building the column-stochastic CSR ...")
    A, dang = graphs.transition_matrix(src, dst, n)
    del src, dst
    log("  CSR built; building the PBR on the GPU ...")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pbr = gl.csr_to_pbr(A, 8, 8, 4, order="canonical")
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    T, V, R = pbr.accounted_blocks(), pbr.compressed_nnz(), pbr.remainder_nnz()
    log(f"  nnz={A.nnz} T={T} V={V} R={R} dangling={int(dang.sum())}, built in "
        f"{t_build*1e3:.0f} ms (incl. upload)")
    smax = args.ppr_set_max
    sets = graphs.seed_sets(n, K, 1, smax, seed=4)
    dmode = "uniform" if dang.any() else "none"
    if args.ppr_profile_passes:
        np_ = args.ppr_profile_passes
        pk = dict(tol=0.0, max_iter=np_, kw=64 if K >= 64 else 0, graph_iters=0,
                  return_history=False, dangling=dang if dang.any() else None,
                  dangling_mode=dmode)
        tp, _ = _timed_runs(lambda: gl.run_ppr(pbr, sets, **pk), reps=1)
        log(f"  {np_} plain passes: {tp / np_ * 1e3:.3f} ms per pass")
        return {"profile_passes": np_, "ms_per_pass": round(tp / np_ * 1e3, 4),
                "intra": args.ppr_intra, "community": args.ppr_community, "nodes": n,
                "edges_nnz": int(A.nnz), "tiles": T, "remainder_nnz": R}
    # autotuned tiles (the 8x8 threshold) and launch (columns per launch, grid)
    mn_best, pbr, tiles_tab = gl.autotune.tune_ppr_tiles(A, K, built={4: pbr})
    tuned = {k: v for k, v in tiles_tab[mn_best].items() if k != "table"}
    tuned["min_nnz_per_block"] = mn_best
    tuned["ms_per_iter_by_min_nnz"] = {str(k): v["ms_per_iter"] for k, v in tiles_tab.items()}
    T, V, R = pbr.accounted_blocks(), pbr.compressed_nnz(), pbr.remainder_nnz()
    log(f"  autotuned launch: {tuned}; PBR 8x8 min {mn_best}: T={T} V={V} R={R}")
    kw = dict(alpha=0.85, max_iter=300, dangling=dang if dang.any() else None,
              dangling_mode=dmode, retire="per_column", return_history=False,
              kw=tuned["kw"], grid_mult=tuned["grid_mult"])
    mbytes = ppr_matrix_bytes(T, V, R, n)

    def algorithmic(res):
        return res.total_iterations * mbytes + 8 * n * int(res.iterations.sum())

    tol_acc = 1.5e-7
    t, res = _timed_runs(lambda: gl.run_ppr(pbr, sets, tol=tol_acc, **kw))
    gbs = algorithmic(res) / t / 1e9
    log(f"  plain tol {tol_acc}: {K / t:.1f} PPR/s, {t*1e3:.1f} ms, iterations max "
        f"{res.total_iterations} mean {res.iterations.mean():.1f}, {gbs:.0f} GB/s algorithmic")
    out = {"workload": f"ppr-{n // 1_000_000}M-{m // 1_000_000}M-clustered-{K}sets",
           "nodes": n, "edges_nnz": int(A.nnz), "tiles": T, "tile_nnz": V, "remainder_nnz": R,
           "seed_sets": K, "seed_set_size": f"U{{1..{smax}}}", "alpha": 0.85, "tol": tol_acc,
           "dangling_mode": dmode, "launch": tuned, "ppr_per_s": round(K / t, 3),
           "ms": round(t * 1e3, 3), "iterations_max": res.total_iterations,
           "iterations_mean": round(float(res.iterations.mean()), 2),
           "converged": int(res.converged.sum()),
           "algorithmic_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
           "build_ms": round(t_build * 1e3, 1)}
    t6, res6 = _timed_runs(lambda: gl.run_ppr(pbr, sets, tol=1e-6, **kw))
    out["tol_1e-6"] = {"ppr_per_s": round(K / t6, 3), "ms": round(t6 * 1e3, 3),
                       "iterations_max": res6.total_iterations,
                       "iterations_mean": round(float(res6.iterations.mean()), 2)}
    log(f"  plain tol 1e-6: {K / t6:.1f} PPR/s, iterations max {res6.total_iterations}")
    # dominant-mode extrapolation every EXTRAP passes (run_ppr extrapolate_every)
    EXTRAP = 5
    te, eres = _timed_runs(lambda: gl.run_ppr(pbr, sets, tol=tol_acc, extrapolate_every=EXTRAP,
                                              **kw))
    out["extrapolated"] = {"tol": tol_acc, "extrapolate_every": EXTRAP,
                           "ppr_per_s": round(K / te, 3), "ms": round(te * 1e3, 3),
                           "iterations_max": eres.total_iterations,
                           "iterations_mean": round(float(eres.iterations.mean()), 2),
                           "converged": int(eres.converged.sum()),
                           "speedup_vs_plain": round(t / te, 3)}
    log(f"  extrapolated (every {EXTRAP}) tol {tol_acc}: {K / te:.1f} PPR/s, {te*1e3:.1f} ms, "
        f"iterations max {eres.total_iterations} mean {eres.iterations.mean():.1f} "
        f"({t / te:.2f}x plain)")
    wres = wres_x = None
    wpath = os.path.join(ROOT, "page_rank_amd", "data", "warmstart_clustered.npz")
    if os.path.exists(wpath):
        from page_rank_amd import warmstart as ws
        params = ws.load(wpath, dev)
        graph = ws.graph_features(pbr)   # degrees and buckets: once per graph, like the PBR
        tw, wres = _timed_runs(lambda: ws.run_warm(params, pbr, sets, graph=graph,
                                                   tol=tol_acc, **kw))
        out["learned"] = {"tol": tol_acc, "ppr_per_s": round(K / tw, 3),
                          "ms": round(tw * 1e3, 3),
                          "iterations_max": wres.total_iterations + wres.feature_passes,
                          "iterations_mean": round(float(wres.iterations.mean())
                                                   + wres.feature_passes, 2),
                          "feature_passes": wres.feature_passes, "blend": ws.BLEND,
                          "converged": int(wres.converged.sum()),
                          "speedup_vs_plain": round(t / tw, 3),
                          "model": "page_rank_amd/data/warmstart_clustered.npz"}
        log(f"  learned warm start tol {tol_acc}: {K / tw:.1f} PPR/s, {tw*1e3:.1f} ms, "
            f"iterations max {wres.total_iterations}+{wres.feature_passes} mean "
            f"{wres.iterations.mean():.1f}+{wres.feature_passes} ({t / tw:.2f}x plain)")
        twx, wres_x = _timed_runs(lambda: ws.run_warm(params, pbr, sets, graph=graph,
                                                      tol=tol_acc, extrapolate_every=EXTRAP,
                                                      **kw))
        out["learned_extrapolated"] = {
            "tol": tol_acc, "extrapolate_every": EXTRAP, "ppr_per_s": round(K / twx, 3),
            "ms": round(twx * 1e3, 3),
            "iterations_max": wres_x.total_iterations + wres_x.feature_passes,
            "iterations_mean": round(float(wres_x.iterations.mean()) + wres_x.feature_passes, 2),
            "converged": int(wres_x.converged.sum()),
            "speedup_vs_plain": round(t / twx, 3),
            "speedup_vs_extrapolated": round(te / twx, 3)}
        log(f"  learned + extrapolated: {K / twx:.1f} PPR/s, {twx*1e3:.1f} ms "
            f"({t / twx:.2f}x plain, {te / twx:.2f}x extrapolated)")
    if args.ppr_check_cols and not args.no_cpu:
        from oracle import ppr as oppr
        cols = np.random.default_rng(5).choice(K, min(K, args.ppr_check_cols), replace=False)
        P = oppr.personalization_matrix(n, [sets[c] for c in cols])

        def progress(it, err):
            if it % 10 == 0:
                log(f"    FP64 iteration {it}: max residual {err.max():.2e}")
        log(f"  FP64 CPU power iteration on {len(cols)} columns ...")
        t0 = time.perf_counter()
        X64, it64, _, _ = oppr.power(A, P, alpha=0.85, tol=1e-6, max_iter=300,
                                     dangling=dang if dang.any() else None,
                                     dangling_mode=dmode, retire="per_column",
                                     progress=progress)
        t_cpu = time.perf_counter() - t0
        # reference: the same iteration continued to tol 1e-11 (error below 6e-11 per
        # column, four orders under the errors reported)
        log(f"  ... {t_cpu:.1f} s; continuing to tol 1e-11 for the reference ...")
        Xref, _, _, _ = oppr.power(A, P, alpha=0.85, tol=1e-11, max_iter=600,
                                   dangling=dang if dang.any() else None,
                                   dangling_mode=dmode, retire="per_column", x0=X64,
                                   progress=progress)
        sel = torch.as_tensor(cols, device=res.X.device)

        def errs(r):
            Xg = r.X[:, sel].cpu().numpy().astype(np.float64)
            d = np.abs(Xg - Xref)
            return ([float(f"{v:.3e}") for v in d.sum(0)], [float(f"{v:.3e}") for v in d.max(0)])
        out["l1_vs_fp64"], out["max_abs_vs_fp64"] = errs(res)
        out["all_checked_below_1e-6"] = bool(max(out["l1_vs_fp64"]) < 1e-6)
        out["tol_1e-6"]["l1_vs_fp64"], out["tol_1e-6"]["max_abs_vs_fp64"] = errs(res6)
        for key, r in (("extrapolated", eres), ("learned", wres),
                       ("learned_extrapolated", wres_x)):
            if r is None:
                continue
            out[key]["l1_vs_fp64"], out[key]["max_abs_vs_fp64"] = errs(r)
            out[key]["all_checked_below_1e-6"] = bool(max(out[key]["l1_vs_fp64"]) < 1e-6)
        out["checked_columns"] = cols.tolist()
        out["cpu_baseline"] = {"value": round(len(cols) / t_cpu, 4), "unit": "PPR/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{len(cols)} of the {K} seed sets, FP64 scipy power "
                                         f"iteration to tol 1e-6 (oracle/ppr.py power), "
                                         f"{int(it64.max())} iterations in {t_cpu:.1f} s",
                               **host_info()}
        log(f"  FP64 check: plain L1 {out['l1_vs_fp64']}, learned "
            f"{out.get('learned', {}).get('l1_vs_fp64')}; CPU {len(cols)/t_cpu:.3f} PPR/s")
    return out


def config1_leg(dev, reps=7):
    """The reference's CPU path on its CPU-runnable case: networkx's _pagerank_scipy semantics
    (the oracle's restatement, FP64 scipy, as Tests/test_ppr_accuracy.py:28-40 uses it) on the
    seeded 10,000-node SBM (50 communities of 200, p_in 0.04, p_out 0.0002), 1 seed,
    alpha 0.85, dangling mass to the personalization (networkx's default), tol 1e-8, at most
    100 iterations; median of `reps` timed runs after one warm-up. networkx stops when the L1
    update is below N * tol (1e-4 here); the absolute 1e-8 stop is reported beside it. The
    HIP path runs the same graph and is checked against the exact FP64 solution."""
    import numpy as np
    import scipy.sparse as sp
    import torch
    import page_rank_amd as gl
    from oracle import ppr as oppr
    from page_rank_amd import graphs
    n, seed = 10_000, 0
    src, dst = graphs.sbm_edges(n=n, community=200, p_in=0.04, p_out=0.0002, seed=0)
    A_col = sp.csr_matrix((np.ones(src.size), (dst, src)), shape=(n, n))
    p = np.zeros(n)
    p[seed] = 1.0
    ts, it_nx = [], 0
    x_nx, it_nx = oppr.nx_pagerank(A_col, p, 0.85, tol=1e-8, max_iter=100,
                                   return_iterations=True)
    for _ in range(reps):
        t0 = time.perf_counter()
        oppr.nx_pagerank(A_col, p, 0.85, tol=1e-8, max_iter=100)
        ts.append(time.perf_counter() - t0)
    t_nx = float(np.median(ts))
    A64, dang = oppr.column_stochastic(A_col)
    x_abs, it_abs, conv_abs, _ = oppr.power(A64, p[:, None], 0.85, 1e-8, 100, dang,
                                            "personalization")
    exact = oppr.solve(A64, p, 0.85, dang, "personalization")
    A32, _ = graphs.transition_matrix(src, dst, n)
    pbr = gl.csr_to_pbr(A32, 8, 8, 4, order="canonical")
    out = {"workload": "ppr-sbm-10k-50x200", "nodes": n, "edges": int(src.size),
           "dangling": int(dang.sum()), "seed": seed, "alpha": 0.85,
           "cpu_reference": {"value_ms": round(t_nx * 1e3, 3), "iterations": int(it_nx),
                             "stop": "networkx: L1 update < N * tol = 1e-4",
                             "l1_vs_exact": float(f"{np.abs(x_nx - exact).sum():.3e}"),
                             "reps": reps, "cores": 1, "kind": "port",
                             "what": "oracle/ppr.py nx_pagerank (networkx 3.4.2 "
                                     "_pagerank_scipy semantics), FP64",
                             **host_info()},
           "cpu_absolute_tol": {"iterations": int(it_abs[0]), "converged": bool(conv_abs[0]),
                                "l1_vs_exact": float(
                                    f"{np.abs(x_abs[:, 0] - exact).sum():.3e}")}}
    # the library's choice (the multi-workgroup loop at this size), then the whole run in
    # one workgroup on the same graph
    for name, tol, one in (("hip_tol_1e-08", 1e-8, None), ("hip_tol_1.5e-07", 1.5e-7, None),
                           ("hip_one_wg_tol_1e-08", 1e-8, True)):
        tg, r = _timed_runs(lambda: gl.run_ppr(pbr, [[seed]], tol=tol, max_iter=100,
                                               dangling=dang, dangling_mode="personalization",
                                               single_workgroup=one),
                            reps=5)
        x = r.X[:, 0].cpu().numpy().astype(np.float64)
        out[name] = {"ms": round(tg * 1e3, 3), "iterations": r.total_iterations,
                     "converged": bool(r.converged[0]),
                     "l1_vs_fp64": float(f"{np.abs(x - exact).sum():.3e}"),
                     "max_abs_vs_fp64": float(f"{np.abs(x - exact).max():.3e}")}
    log(f"  config 1: CPU networkx path {t_nx*1e3:.2f} ms ({it_nx} it), HIP "
        f"{out['hip_tol_1.5e-07']}, one workgroup {out['hip_one_wg_tol_1e-08']}")
    # request-sized graphs (the /api/ppr case): both paths, one seed, tol 1e-8
    small = {}
    for name, (ns, comm) in (("sbm_1k", (1000, 100)), ("sbm_100", (100, 25))):
        s_src, s_dst = graphs.sbm_edges(n=ns, community=comm, p_in=0.04 * 200 / comm,
                                        p_out=0.0002, seed=seed)
        A_s, dang_s = graphs.transition_matrix(s_src, s_dst, ns)
        pbr_s = gl.csr_to_pbr(A_s, 8, 8, 4, order="canonical")
        row = {"nodes": ns, "nnz": int(A_s.nnz)}
        for leg, one in (("one_wg", True), ("multi_wg", False)):
            tg, r = _timed_runs(lambda: gl.run_ppr(pbr_s, [[0]], tol=1e-8, max_iter=100,
                                                   dangling=dang_s,
                                                   dangling_mode="personalization",
                                                   single_workgroup=one), reps=7)
            row[leg] = {"ms": round(tg * 1e3, 3), "iterations": r.total_iterations}
        small[name] = row
    out["request_sized"] = small
    log(f"  request-sized graphs: {small}")
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested). Seed sets would shard across GPUs with no "
              f"per-iteration communication.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import page_rank_amd as gl
    from page_rank_amd import graphs

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.calibrate_copy_mib:
        src = torch.ones(args.calibrate_copy_mib << 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log(f"calibration copies of {args.calibrate_copy_mib} MiB done")
    if args.ppr_only:
        print(json.dumps({"ppr": ppr_leg(args, dev)}), flush=True)
        return
    log("generating the 1M x 1M SpMM matrix ...")
    A = graphs.spmm_matrix(seed=0)
    log(f"  nnz={A.nnz}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pbr = gl.csr_to_pbr(A, 8, 8, 4, order="canonical")
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    T, V, R, N = pbr.accounted_blocks(), pbr.compressed_nnz(), pbr.remainder_nnz(), pbr.rows
    log(f"  built PBR: T={T} V={V} R={R} in {t_build*1e3:.1f} ms (incl. upload)")
    builder = builder_leg(A, T, V, R, dev)
    mats = [pbr] + [clone_matrix(pbr, dev) for _ in range(max(0, args.rotate - 1))]

    widths = [int(w) for w in args.widths.split(",") if w]
    if args.k not in widths:
        widths.append(args.k)
    if args.only_headline:
        widths = [args.k]
    per_k = {}
    headline = None
    for K in widths:
        wall, launch_ms, X, Y = time_width(mats, K, args.steps, args.warmup, args.kw,
                                           args.mode, dev)
        mean_launch = float(np.mean(launch_ms))
        nbytes = spmm_bytes(T, V, R, N, K)
        gnz = A.nnz * K * args.steps / wall / 1e9
        gbs = nbytes / (mean_launch * 1e-3) / 1e9
        per_k[str(K)] = {"gnz_s": round(gnz, 2), "ms_per_step": round(wall / args.steps * 1e3, 4),
                         "kernel_ms": round(mean_launch, 4), "bytes": nbytes,
                         "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
        log(f"  K={K}: {gnz:.1f} GNZ/s, kernel {mean_launch*1e3:.1f} us, {gbs:.0f} GB/s "
            f"({gbs/HBM_PEAK_GBS*100:.1f}% of HBM peak)")
        # errors against a float64 product on 256 sampled rows (scipy, FP64)
        rows = np.random.default_rng(K).choice(N, 256, replace=False)
        ref = A[rows].astype(np.float64) @ X.cpu().numpy().astype(np.float64)
        got = Y.cpu().numpy()[rows].astype(np.float64)
        assert np.allclose(got, ref, rtol=1e-4, atol=1e-4), f"SpMM check failed at K={K}"
        per_k[str(K)]["max_abs_vs_fp64"] = float(f"{np.abs(got - ref).max():.3e}")
        per_k[str(K)]["l1_rel_vs_fp64"] = float(
            f"{np.abs(got - ref).sum() / np.abs(ref).sum():.3e}")
        per_k[str(K)]["checked_rows"] = 256
        if K == args.k:
            headline = (wall, mean_launch, nbytes, gnz, gbs)
        del X, Y

    wall, mean_launch, nbytes, gnz, gbs = headline
    # HBM bytes per launch: not measurable inside this process (it needs two rocprofv3 --pmc
    # runs); read from the committed summary of those runs of this kernel and workload
    traffic, traffic_source = None, None
    tpath = os.path.join(ROOT, "profiles", f"spmm_1m_k{args.k}_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
        traffic_source = (f"read from profiles/spmm_1m_k{args.k}_traffic.json (rocprofv3 "
                          "FETCH_SIZE / WRITE_SIZE passes, scaled by a copy of known size)")
    cpu = None
    if not args.no_cpu:
        log("timing the CPU baseline ...")
        cpu = cpu_baseline(A, args.k, args.cpu_seconds)
    nnz = int(A.nnz)
    n_mats = len(mats)
    del mats, pbr, A
    torch.cuda.empty_cache()
    ppr = None if (args.no_ppr or args.only_headline) else ppr_leg(args, dev)
    cfg1 = None if (args.no_ppr or args.only_headline or args.no_cpu) else config1_leg(dev)
    line = {
        "metric": "SpMM GNZ/s per RHS (1M x 1M PBR 8x8, ~27M nnz)",
        "value": round(gnz, 2),
        "unit": "GNZ/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "pbr-spmm-1M-8x8-blockband512-scatter5M", "rows": N,
                   "nnz": nnz, "tiles": T, "tile_nnz": V, "remainder_nnz": R,
                   "block": 8, "min_nnz_per_block": 4, "K": args.k, "mode": args.mode,
                   "rotating_matrix_copies": n_mats},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": traffic_source},
        "cpu_baseline": cpu,
        "per_rhs_width": per_k,
        "builder": builder,
        "builder_with_upload_ms": round(t_build * 1e3, 2),
        "ppr": ppr,
        "config1_sbm_10k": cfg1,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
