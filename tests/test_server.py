"""The /api/ppr contract (Backend/server/app.py:71-160): health, 503 without a GPU, 422 for
no nodes / no valid seed / malformed bodies (CPU), and on the GPU the ranks of the karate
club graph against the FP64 oracle with uniform dangling handling."""
import networkx as nx
import numpy as np
import pytest
import scipy.sparse as sp
from fastapi.testclient import TestClient

from page_rank_amd import server
from oracle import ppr as oppr

client = TestClient(server.app)


def _graph(edges, nodes=None):
    nodes = nodes if nodes is not None else sorted({v for e in edges for v in e[:2]}, key=str)
    return {"nodes": [{"id": str(n)} for n in nodes],
            "edges": [{"source": str(a), "target": str(b), "weight": (w[0] if w else 1.0)}
                      for a, b, *w in edges],
            "format": "edge", "warnings": []}


def test_health_reports_device():
    r = client.get("/api/health")
    assert r.status_code == 200
    body = r.json()
    assert body["status"] == "ok" and isinstance(body["cuda"], bool)


def test_503_without_gpu(monkeypatch):
    monkeypatch.setattr(server, "gpu_available", lambda: False)
    r = client.post("/api/ppr", json={"graph": _graph([(0, 1)]), "params": {"seeds": ["0"]}})
    assert r.status_code == 503


def test_422_cases(monkeypatch):
    monkeypatch.setattr(server, "gpu_available", lambda: True)
    r = client.post("/api/ppr", json={"graph": _graph([], nodes=[]), "params": {"seeds": ["a"]}})
    assert r.status_code == 422 and "no nodes" in r.json()["detail"]
    r = client.post("/api/ppr", json={"graph": _graph([(0, 1)]), "params": {"seeds": ["zz"]}})
    assert r.status_code == 422 and "seed" in r.json()["detail"]
    r = client.post("/api/ppr", json={"params": {"seeds": ["0"]}})
    assert r.status_code == 422
    r = client.post("/api/ppr/batch", json={"graph": _graph([(0, 1)]), "params": {},
                                            "seedSets": [["0"], ["nope"]]})
    assert r.status_code == 422


def test_degrees_and_adjacency_follow_the_reference():
    g = server.Graph(**_graph([("a", "b"), ("a", "c", 2.0), ("b", "c"), ("c", "a", -1.0),
                               ("x", "a")], nodes=["a", "b", "c"]))
    ids, index_of, A, deg = server._prepare(g)
    assert deg["out"] == {"a": 2.0, "b": 1.0, "c": 1.0}       # plain counts, bad weight included
    assert deg["in"] == {"a": 2.0, "b": 1.0, "c": 2.0}
    dense = A.toarray()
    assert dense[index_of["c"], index_of["a"]] == 2.0          # A[target, source] = weight
    assert dense[index_of["a"], index_of["c"]] == 0.0          # non-positive weight dropped
    assert A.nnz == 3


@pytest.mark.gpu
def test_karate_ranks_match_fp64(device):
    G = nx.karate_club_graph()
    edges = [(a, b) for a, b in G.edges()] + [(33, 34)]       # node 34: dangling
    nodes = list(range(35))
    body = {"graph": _graph(edges, nodes), "params": {"alpha": 0.85, "maxIter": 200,
                                                      "tolerance": 1e-7, "seeds": ["0", "33", "0"]}}
    r = client.post("/api/ppr", json=body)
    assert r.status_code == 200, r.text
    out = r.json()
    assert out["converged"] and out["iterations"] == len(out["convergenceHistory"])
    A = sp.csr_matrix((np.ones(len(edges)), ([b for a, b in edges], [a for a, b in edges])),
                      shape=(35, 35))
    An, dang = oppr.column_stochastic(A)
    p = np.zeros(35)
    p[0] += 2 / 3
    p[33] += 1 / 3
    x, _, _, _ = oppr.power(An, p[:, None], tol=1e-13, max_iter=1000, dangling=dang,
                            dangling_mode="uniform")
    got = np.array([out["ranks"][str(i)] for i in nodes])
    assert abs(got.sum() - 1) < 1e-6
    assert np.abs(got - x[:, 0]).max() < 5e-6
    # iterations / history as the reference computes them (app.py:132-152): one column per
    # seed entry, all stopping when the largest column residual is below the tolerance
    cols = oppr.personalization_matrix(35, [[0], [33], [0]])
    _, it_o, _, hist_o = oppr.power(An, cols, tol=1e-7, max_iter=200, dangling=dang,
                                    dangling_mode="uniform", retire="together")
    assert abs(out["iterations"] - int(it_o.max())) <= 1
    np.testing.assert_allclose(out["convergenceHistory"][:5], hist_o[:5].max(axis=1), rtol=1e-3)
    assert out["degrees"]["out"]["33"] == 1.0 and out["degrees"]["in"]["34"] == 1.0
    rb = client.post("/api/ppr/batch", json={**body, "seedSets": [["0"], ["33", "1"]]})
    assert rb.status_code == 200 and len(rb.json()["results"]) == 2
