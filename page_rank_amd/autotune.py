"""Autotuner of the launch parameters: PBR block size and tile threshold, RHS columns per
launch (kw), SpMM mode (fused / split) and the PPR grid size (workgroups per CU).

Every choice is measured on the device with the real matrix and RHS width. The winner is
cached in memory and in a JSON file, keyed by a shape signature: device name, block size,
rows, tiles, tile values, remainder entries, K, and the operation. Later calls with the same
signature do no timing. ``kw="auto"`` in ``spmm_into`` and ``run_ppr`` routes through here.

Cache file: ``$PAGE_RANK_AMD_TUNE_CACHE`` or ``~/.cache/page_rank_amd/tune.json``.
"""
from __future__ import annotations

import json
import os
import tempfile
import time

import numpy as np
import torch

from .registry import get_pbr_gpu_meta

KW_ALL = (1, 2, 4, 8, 16, 32, 64)
# PPR grid: workgroups per CU; 0 = as many as the device holds of the step kernel at once
GRID_MULTS = (0, 4, 8, 16)


def kw_candidates(K: int):
    """Columns-per-launch values worth timing for K columns: every supported width up to
    the smallest one that covers K in one launch (and at least one feature-lane width)."""
    top = next((w for w in KW_ALL if w >= K), 64)
    out = [w for w in KW_ALL if w <= top]
    if 1 < top < 4:
        out.append(4)  # two columns can run faster in the 4-column kernel (float4 rows)
    if top < 16:
        out.append(16)
    return out[-4:] if len(out) > 4 else out


def cache_path():
    return os.environ.get("PAGE_RANK_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "page_rank_amd",
                                       "tune.json"))


class TuneCache:
    """Signature -> chosen parameters, mirrored to a JSON file (written atomically)."""

    def __init__(self, path=None):
        self.path = path or cache_path()
        self.entries = {}
        try:
            with open(self.path) as f:
                data = json.load(f)
            if isinstance(data, dict):
                self.entries = data
        except (OSError, ValueError):
            pass

    def get(self, key):
        return self.entries.get(key)

    def put(self, key, value):
        self.entries[key] = value
        d = os.path.dirname(self.path)
        try:
            os.makedirs(d, exist_ok=True)
            fd, tmp = tempfile.mkstemp(dir=d, prefix=".tune")
            with os.fdopen(fd, "w") as f:
                json.dump(self.entries, f, indent=1, sort_keys=True)
            os.replace(tmp, self.path)
        except OSError:
            pass  # an unwritable cache only costs a re-tune


_cache = None


def cache() -> TuneCache:
    global _cache
    if _cache is None or _cache.path != cache_path():
        _cache = TuneCache()
    return _cache


def signature(op, pbr, K, device):
    meta = get_pbr_gpu_meta(pbr)
    name = torch.cuda.get_device_name(device).replace("|", "/")
    return (f"{op}|{name}|B{pbr.block_rows}x{pbr.block_cols}|N{pbr.rows}|"
            f"T{meta['codes'].numel()}|V{meta['data'].numel()}|R{meta['rem_col_ind'].numel()}|K{K}")


def _best(candidates, run, reps):
    """Time run(c) for every candidate (1 untimed call, then `reps` timed) and return the
    fastest candidate with its mean milliseconds and the full table."""
    table = []
    for c in candidates:
        run(c)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            run(c)
        torch.cuda.synchronize()
        table.append((c, (time.perf_counter() - t0) / reps * 1e3))
    best = min(table, key=lambda e: e[1])
    return best[0], best[1], table


def tune_spmm(pbr, K, *, reps=5, force=False):
    """Best (kw, mode) for Y = A X with K columns on this matrix."""
    from .ops import spmm_into
    meta = get_pbr_gpu_meta(pbr)
    if meta is None:
        raise RuntimeError("PBRMatrix is not on GPU. Call .to('cuda') first.")
    dev = meta["codes"].device
    key = signature("spmm", pbr, K, dev)
    hit = None if force else cache().get(key)
    if hit is not None:
        return hit
    g = torch.Generator(device="cpu").manual_seed(0)
    x = torch.rand((pbr.cols, K), generator=g).to(dev)
    y = torch.empty((pbr.rows, K), device=dev)
    cands = [(kw, mode) for kw in kw_candidates(K) for mode in ("fused", "split")]
    (kw, mode), ms, table = _best(cands, lambda c: spmm_into(pbr, x, y, kw=c[0], mode=c[1]),
                                  reps)
    out = {"kw": kw, "mode": mode, "ms": round(ms, 4),
           "table": [[c[0], c[1], round(t, 4)] for c, t in table]}
    cache().put(key, out)
    return out


def tune_ppr(pbr, K, *, reps=2, iters=4, force=False):
    """Best (kw, grid_mult) for the fused PPR step with K seed-set columns."""
    from .ppr import run_ppr
    meta = get_pbr_gpu_meta(pbr)
    if meta is None:
        raise RuntimeError("PBRMatrix is not on GPU. Call .to('cuda') first.")
    dev = meta["codes"].device
    key = signature("ppr", pbr, K, dev)
    hit = None if force else cache().get(key)
    if hit is not None:
        return hit
    rng = np.random.default_rng(0)
    sets = [[int(v)] for v in rng.integers(0, pbr.rows, K)]
    cands = [(kw, gm) for kw in kw_candidates(K) if kw >= min(K, 16) or K <= 8
             for gm in GRID_MULTS]
    (kw, gm), ms, table = _best(
        cands, lambda c: run_ppr(pbr, sets, tol=0.0, max_iter=iters, kw=c[0], grid_mult=c[1],
                                 graph_iters=iters, return_history=False), reps)
    out = {"kw": kw, "grid_mult": gm, "ms_per_iter": round(ms / iters, 4),
           "table": [[c[0], c[1], round(t / iters, 4)] for c, t in table]}
    cache().put(key, out)
    return out


# PPR tile thresholds: 4 keeps every 8x8 block with 4+ entries as a tile; 16 and above turn
# the sparse tiles of community graphs into remainder entries (a 64-column tile loads 8 full
# X rows for its few values, two cache lines each, where a remainder entry loads one row)
PPR_MIN_NNZ = (4, 8, 16)


def tune_ppr_tiles(A_csr, K, *, block=8, min_nnz=PPR_MIN_NNZ, reps=2, iters=4, device=None,
                   built=None):
    """Build the matrix with each tile threshold (min_nnz_per_block) on the GPU, tune the PPR
    step on each, and return (best threshold, its PBRMatrix, {threshold: tuned entry}).
    `built` may map a threshold to a matrix already built from A_csr with that threshold and
    block size, which is then used instead of building it again. The other matrices are
    released before returning."""
    from .builder import csr_to_pbr
    built = dict(built or {})
    results, mats = {}, {}
    for mn in min_nnz:
        pbr = built.get(mn)
        if pbr is None:
            pbr = csr_to_pbr(A_csr, block, block, mn, order="canonical", device=device)
        results[mn] = tune_ppr(pbr, K, reps=reps, iters=iters)
        mats[mn] = pbr
    best = min(results, key=lambda mn: results[mn]["ms_per_iter"])
    return best, mats[best], results


def tune_block(A_csr, K, *, blocks=(2, 4, 8), min_nnz_per_block=None, reps=5, device=None):
    """Build the matrix with each block size on the GPU, tune its SpMM, and return
    (best block size, {block: tuned entry}). min_nnz_per_block defaults to half the tile
    area, capped at 4."""
    from .builder import csr_to_pbr
    results = {}
    for b in blocks:
        mn = min_nnz_per_block if min_nnz_per_block is not None else min(4, max(1, b * b // 2))
        pbr = csr_to_pbr(A_csr, b, b, mn, order="canonical", device=device)
        results[b] = tune_spmm(pbr, K, reps=reps)
        del pbr
    best = min(results, key=lambda b: results[b]["ms"])
    return best, results
