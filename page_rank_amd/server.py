"""HTTP service for Personalized PageRank on the HIP path: the reference's contract
(Backend/server/app.py:71-160, schema Frontend/src/lib/ppr/types.ts:5-52) on this build.

    python -m page_rank_amd.server --host 0.0.0.0 --port 8000

GET  /api/health  {"status": "ok", "cuda": bool, "device": name | null}
POST /api/ppr     {"graph": {...}, "params": {...}} ->
                  {"ranks", "iterations", "converged", "convergenceHistory", "degrees"}

The request and response schemas, the 503 (no GPU) and 422 (no nodes / no valid seed)
answers and the degree counts are the reference's, and so is the computation's shape: one
column per seed entry (duplicates included), every column stopping when the largest column
residual falls below the tolerance, ranks = the column mean renormalized to sum 1,
`iterations` = passes run (maxIter when not converged) and `convergenceHistory[t]` = the
largest column residual of pass t. What differs is inside:
- normalization runs on the GPU with dangling columns handled inside the run (their mass
  spread uniformly, as the reference's dense 1/N columns do), so large graphs with many
  dangling nodes fit;
- the matrix is tiled 8x8 (the reference's default is 2x2), which changes no result beyond
  float32 rounding;
- the whole loop is one hipGraph-driven call.
POST /api/ppr/batch takes several seed sets and answers one rank map per set (an addition):
each set is one personalization column (weights 1/|S|) that stops at its own residual.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import torch
from fastapi import FastAPI, HTTPException
from fastapi.middleware.cors import CORSMiddleware
from pydantic import BaseModel, Field

app = FastAPI(title="PPR Analyzer backend (HIP)", version="1.0")
app.add_middleware(CORSMiddleware, allow_origins=["*"], allow_methods=["*"],
                   allow_headers=["*"])


class GraphNode(BaseModel):
    id: str


class GraphEdge(BaseModel):
    source: str
    target: str
    weight: float = 1.0


class Graph(BaseModel):
    nodes: list[GraphNode]
    edges: list[GraphEdge]
    format: str | None = None
    warnings: list[str] = Field(default_factory=list)


class PPRParams(BaseModel):
    alpha: float = 0.85
    maxIter: int = 100
    tolerance: float = 1e-6
    topX: int = 10
    seeds: list[str] = Field(default_factory=list)


class PPRRequest(BaseModel):
    graph: Graph
    params: PPRParams


class PPRBatchRequest(BaseModel):
    graph: Graph
    params: PPRParams
    seedSets: list[list[str]]


def gpu_available() -> bool:
    return torch.cuda.is_available()


@app.get("/api/health")
def health() -> dict:
    cuda = gpu_available()
    return {"status": "ok", "cuda": cuda,
            "device": torch.cuda.get_device_name(0) if cuda else None}


def _prepare(g: Graph):
    """Node index, degree counts (plain edge counts, as the reference) and the adjacency
    A[target, source] = weight over edges with known endpoints and finite positive weight."""
    ids = [n.id for n in g.nodes]
    index_of = {nid: i for i, nid in enumerate(ids)}
    in_deg = {nid: 0.0 for nid in ids}
    out_deg = {nid: 0.0 for nid in ids}
    rows, cols, data = [], [], []
    for e in g.edges:
        if e.source in out_deg:
            out_deg[e.source] += 1.0
        if e.target in in_deg:
            in_deg[e.target] += 1.0
        si, ti = index_of.get(e.source), index_of.get(e.target)
        if si is None or ti is None or not np.isfinite(e.weight) or e.weight <= 0:
            continue
        rows.append(ti)
        cols.append(si)
        data.append(float(e.weight))
    N = len(ids)
    A = sp.csr_matrix((np.asarray(data, dtype=np.float32),
                       (np.asarray(rows, dtype=np.int64), np.asarray(cols, dtype=np.int64))),
                      shape=(N, N))
    A.sum_duplicates()
    return ids, index_of, A, {"in": in_deg, "out": out_deg}


def _solve(A, seed_sets, p: PPRParams, retire):
    import page_rank_amd as gl
    dev = torch.device("cuda", torch.cuda.current_device())
    norm, dang = gl.normalize_transition_matrix_gpu(A, dev)
    pbr = gl.csr_to_pbr(norm, 8, 8, 4, order="canonical")
    return gl.run_ppr(pbr, seed_sets, alpha=p.alpha, tol=p.tolerance, max_iter=max(0, p.maxIter),
                      dangling=dang if dang.any() else None,
                      dangling_mode="uniform" if dang.any() else "none", retire=retire)


def _ranks(x: np.ndarray, ids):
    total = float(x.sum())
    if total > 0:
        x = x / total
    return {ids[i]: float(x[i]) for i in range(len(ids))}


@app.post("/api/ppr")
def run_ppr_endpoint(req: PPRRequest) -> dict:
    if not gpu_available():
        raise HTTPException(status_code=503, detail="GPU device not available on the server.")
    g, p = req.graph, req.params
    if len(g.nodes) == 0:
        raise HTTPException(status_code=422, detail="Graph has no nodes.")
    ids, index_of, A, degrees = _prepare(g)
    seed_idx = [index_of[s] for s in p.seeds if s in index_of]
    if not seed_idx:
        raise HTTPException(status_code=422,
                            detail="Personalized PageRank requires at least one seed node.")
    # one column per seed entry, stopping together (app.py:132-152)
    res = _solve(A, [[i] for i in seed_idx], p, "together")
    r = res.X.mean(dim=1)
    total = float(r.sum().item())
    if total > 0:
        r = r / total
    r = r.detach().cpu().numpy()
    T = res.total_iterations
    history = res.history[:T].max(axis=1) if T > 0 else np.zeros(0, np.float32)
    return {"ranks": {ids[i]: float(r[i]) for i in range(len(ids))},
            "iterations": int(T) if p.maxIter > 0 else int(p.maxIter),
            "converged": bool(res.converged.all()),
            "convergenceHistory": [float(h) for h in history],
            "degrees": degrees}


@app.post("/api/ppr/batch")
def run_ppr_batch(req: PPRBatchRequest) -> dict:
    if not gpu_available():
        raise HTTPException(status_code=503, detail="GPU device not available on the server.")
    g, p = req.graph, req.params
    if len(g.nodes) == 0:
        raise HTTPException(status_code=422, detail="Graph has no nodes.")
    if not req.seedSets:
        raise HTTPException(status_code=422, detail="No seed sets given.")
    ids, index_of, A, degrees = _prepare(g)
    sets = []
    for k, s in enumerate(req.seedSets):
        idx = [index_of[v] for v in s if v in index_of]
        if not idx:
            raise HTTPException(status_code=422, detail=f"Seed set {k} has no valid seed node.")
        sets.append(idx)
    res = _solve(A, sets, p, "per_column")
    X = res.X.detach().cpu().numpy().astype(np.float64)
    return {"results": [{"ranks": _ranks(X[:, k], ids), "iterations": int(res.iterations[k]),
                         "converged": bool(res.converged[k])} for k in range(len(sets))],
            "degrees": degrees}


def main(argv=None):
    import argparse

    import uvicorn
    ap = argparse.ArgumentParser()
    ap.add_argument("--host", default="127.0.0.1")
    ap.add_argument("--port", type=int, default=8000)
    a = ap.parse_args(argv)
    uvicorn.run(app, host=a.host, port=a.port)


if __name__ == "__main__":
    main()
