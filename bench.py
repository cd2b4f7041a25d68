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
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first run 3 device-to-device copies of this many MiB (a kernel of "
                         "known traffic that scales the profiler's FETCH_SIZE / WRITE_SIZE)")
    return ap.parse_args()


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
    return {"value": round(gnz, 4), "unit": "GNZ/s", "cores": 1, "kind": "port",
            "sample": f"rows 0..{rows} of the matrix ({S.nnz} nnz), K={K}, {reps} products "
                      f"in {t:.1f} s, oracle/pbr_oracle.cpp spmm (float accumulation)"}


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
                   "nnz": int(A.nnz), "tiles": T, "tile_nnz": V, "remainder_nnz": R,
                   "block": 8, "min_nnz_per_block": 4, "K": args.k, "mode": args.mode,
                   "rotating_matrix_copies": len(mats)},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
        "per_rhs_width": per_k,
        "builder": {"ms": round(t_build * 1e3, 2), "note": "includes host->device upload"},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
