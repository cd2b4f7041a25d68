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
    ap.add_argument("--ppr-check-cols", type=int, default=2,
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
    """Oracle CPU SpMM (restated pbr_batched_matmul_cpu, one thread) on a row sample."""
    import numpy as np
    from oracle import pbr as opbr
    rows = 125_000
    S = A[:rows]
    p = opbr.csr_matrix_to_pbr(S, 8, 8, 4)
    X = np.random.default_rng(1).random((A.shape[1], K), dtype=np.float32)
    reps, t = 0, 0.0
    t0 = time.perf_counter()
    while t < seconds:
        opbr.spmm(p, X)
        reps += 1
        t = time.perf_counter() - t0
    gnz = S.nnz * K * reps / t / 1e9
    out = {"value": round(gnz, 4), "unit": "GNZ/s", "cores": 1, "kind": "port",
           "sample": f"rows 0..{rows} of the matrix ({S.nnz} nnz), K={K}, {reps} products "
                     f"in {t:.1f} s, oracle/pbr_oracle.cpp spmm (float accumulation)"}
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


def ppr_leg(args, dev):
    """Batched multi-seed PPR on the block-clustered power-law graph (by default 4M nodes,
    64M edges, 4,000 communities, 80 % intra, out-degree Pareto(2.1, min 4)), 64 seed sets
    of U{1..100} nodes, alpha 0.85, per-column L1 tol 1e-6, hipGraph-captured loop; with
    --full-scale 16M nodes, 256M edges, 16,000 communities and 128 sets of U{1..200}.
    Errors against the FP64 CPU power iteration on sampled columns, which is also the CPU
    baseline."""
    import numpy as np
    import torch
    import page_rank_amd as gl
    from page_rank_amd import graphs
    n, m, K = args.ppr_nodes, args.ppr_edges, args.ppr_sets
    log(f"generating the {n} node / {m} edge clustered graph ...")
    src, dst = graphs.clustered_edges(n, m, communities=max(1, n // 1000), seed=3)
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
    kw = dict(alpha=0.85, tol=1e-6, max_iter=300, dangling=dang if dang.any() else None,
              dangling_mode=dmode, retire="per_column", return_history=False)
    gl.run_ppr(pbr, sets, **kw)
    times = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = gl.run_ppr(pbr, sets, **kw)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    col_iters = int(res.iterations.sum())
    nbytes = res.total_iterations * ppr_matrix_bytes(T, V, R, n) + 8 * n * col_iters
    gbs = nbytes / t / 1e9
    log(f"  plain: {K / t:.1f} PPR/s, {t*1e3:.1f} ms, iterations max {res.total_iterations} "
        f"mean {res.iterations.mean():.1f}, {gbs:.0f} GB/s algorithmic")
    out = {"workload": f"ppr-{n // 1_000_000}M-{m // 1_000_000}M-clustered-{K}sets",
           "nodes": n, "edges_nnz": int(A.nnz), "tiles": T, "tile_nnz": V, "remainder_nnz": R,
           "seed_sets": K, "seed_set_size": f"U{{1..{smax}}}", "alpha": 0.85, "tol": 1e-6,
           "dangling_mode": dmode, "ppr_per_s": round(K / t, 3), "ms": round(t * 1e3, 3),
           "iterations_max": res.total_iterations,
           "iterations_mean": round(float(res.iterations.mean()), 2),
           "converged": int(res.converged.sum()),
           "algorithmic_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
           "build_ms": round(t_build * 1e3, 1)}
    wpath = os.path.join(ROOT, "page_rank_amd", "data", "warmstart_clustered.npz")
    if os.path.exists(wpath):
        from page_rank_amd import warmstart as ws
        params = ws.load(wpath, dev)
        graph = ws.graph_features(pbr)   # degrees and buckets: once per graph, like the PBR
        wtimes = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            wi = ws.inputs(pbr, sets, 0.85, graph=graph)
            x0, _ = ws.predict(params, wi)
            del wi
            wres = gl.run_ppr(pbr, sets, x0=x0, **kw)
            torch.cuda.synchronize()
            wtimes.append(time.perf_counter() - t0)
        tw = float(np.median(wtimes))
        out["learned"] = {"ppr_per_s": round(K / tw, 3), "ms": round(tw * 1e3, 3),
                          "iterations_max": wres.total_iterations,
                          "iterations_mean": round(float(wres.iterations.mean()), 2),
                          "converged": int(wres.converged.sum()),
                          "speedup_vs_plain": round(t / tw, 3),
                          "l1_vs_plain_max": float(f"{(wres.X - res.X).abs().sum(0).max().item():.3e}"),
                          "model": "page_rank_amd/data/warmstart_clustered.npz"}
        log(f"  learned warm start: {K / tw:.1f} PPR/s, {tw*1e3:.1f} ms, iterations mean "
            f"{wres.iterations.mean():.1f} ({t / tw:.2f}x plain)")
        del x0, wres
    # tolerance that meets L1 < 1e-6 against FP64 (the update-norm stop leaves up to
    # alpha / (1 - alpha) times tol of error)
    kw_acc = dict(kw, tol=1.5e-7)
    gl.run_ppr(pbr, sets, **kw_acc)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res_acc = gl.run_ppr(pbr, sets, **kw_acc)
    torch.cuda.synchronize()
    t_acc = time.perf_counter() - t0
    out["accurate"] = {"tol": 1.5e-7, "ppr_per_s": round(K / t_acc, 3),
                       "ms": round(t_acc * 1e3, 3), "iterations_max": res_acc.total_iterations}
    log(f"  tol 1.5e-7: {K / t_acc:.1f} PPR/s, iterations max {res_acc.total_iterations}")
    if args.ppr_check_cols and not args.no_cpu:
        from oracle import ppr as oppr
        cols = np.random.default_rng(5).choice(K, min(K, args.ppr_check_cols), replace=False)
        P = oppr.personalization_matrix(n, [sets[c] for c in cols])

        def progress(t, err):
            if t % 10 == 0:
                log(f"    FP64 iteration {t}: max residual {err.max():.2e}")
        log(f"  FP64 CPU power iteration on {len(cols)} columns ...")
        t0 = time.perf_counter()
        X64, it64, _, _ = oppr.power(A, P, alpha=0.85, tol=1e-6, max_iter=300,
                                     dangling=dang if dang.any() else None,
                                     dangling_mode=dmode, retire="per_column",
                                     progress=progress)
        t_cpu = time.perf_counter() - t0
        # reference: the same iteration continued to tol 1e-11 (error below 6e-11 per
        # column, five orders under the errors reported)
        log(f"  ... {t_cpu:.1f} s; continuing to tol 1e-11 for the reference ...")
        Xref, _, _, _ = oppr.power(A, P, alpha=0.85, tol=1e-11, max_iter=600,
                                   dangling=dang if dang.any() else None,
                                   dangling_mode=dmode, retire="per_column", x0=X64,
                                   progress=progress)
        sel = torch.as_tensor(cols, device=res.X.device)
        X = res.X[:, sel].cpu().numpy().astype(np.float64)
        l1 = np.abs(X - Xref).sum(0)
        mx = np.abs(X - Xref).max(0)
        out["l1_vs_fp64"] = [float(f"{v:.3e}") for v in l1]
        out["max_abs_vs_fp64"] = [float(f"{v:.3e}") for v in mx]
        Xa = res_acc.X[:, sel].cpu().numpy().astype(np.float64)
        out["accurate"]["l1_vs_fp64"] = [float(f"{v:.3e}") for v in np.abs(Xa - Xref).sum(0)]
        out["checked_columns"] = cols.tolist()
        out["cpu_baseline"] = {"value": round(len(cols) / t_cpu, 4), "unit": "PPR/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{len(cols)} of the {K} seed sets, FP64 scipy power "
                                         f"iteration to the same tol (oracle/ppr.py power), "
                                         f"{int(it64.max())} iterations in {t_cpu:.1f} s",
                               **host_info()}
        log(f"  FP64 check: L1 {l1}, max abs {mx}; CPU {len(cols)/t_cpu:.3f} PPR/s")
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
        if K == args.k:
            headline = (wall, mean_launch, nbytes, gnz, gbs)
            # spot check against a float64 product on a few rows
            rows = np.random.default_rng(0).choice(N, 64, replace=False)
            ref = A[rows].astype(np.float64) @ X.cpu().numpy().astype(np.float64)
            got = Y.cpu().numpy()[rows]
            assert np.allclose(got, ref, rtol=1e-4, atol=1e-4), "SpMM spot check failed"
        del X, Y

    wall, mean_launch, nbytes, gnz, gbs = headline
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"spmm_1m_k{args.k}_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    cpu = None
    if not args.no_cpu:
        log("timing the CPU baseline ...")
        cpu = cpu_baseline(A, args.k, args.cpu_seconds)
    nnz = int(A.nnz)
    n_mats = len(mats)
    del mats, pbr, A
    torch.cuda.empty_cache()
    ppr = None if (args.no_ppr or args.only_headline) else ppr_leg(args, dev)
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
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
        "per_rhs_width": per_k,
        "builder": builder,
        "builder_with_upload_ms": round(t_build * 1e3, 2),
        "ppr": ppr,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
