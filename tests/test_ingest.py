"""Edge-list ingest and seed selection (Tests/wiki_ppr.py:29-56, 162) on the CPU, and the
GPU top-k against a numpy ordering (value descending, row ascending on ties)."""
import gzip

import numpy as np
import pytest
import torch

import page_rank_amd as graph_link


def test_load_edge_list_compacts_ids(tmp_path):
    p = tmp_path / "g.csv.gz"
    rows = [(10, "A", 30, "C"), (30, "C", 20, "B"), (10, "A", 20, "B"), (99, "Z", 10, "A"),
            (10, "A", 30, "C")]
    with gzip.open(p, "wt") as f:
        f.write("page_id_from\tpage_title_from\tpage_id_to\tpage_title_to\n")
        for r in rows:
            f.write("\t".join(map(str, r)) + "\n")
    fi, ti, ids, N = graph_link.load_edge_list(p, chunksize=2)
    assert ids.tolist() == [10, 20, 30, 99] and N == 4
    assert fi.tolist() == [0, 2, 0, 3, 0] and ti.tolist() == [2, 1, 1, 0, 2]
    A, out = graph_link.edges_to_adjacency(fi, ti, N)
    assert A[2, 0] == 2.0 and A[1, 2] == 1.0 and A[0, 3] == 1.0   # duplicate edge summed
    assert out.tolist() == [3.0, 0.0, 1.0, 1.0]
    assert graph_link.top_out_degree_seeds(out, 2).tolist()[0] == 0


def _ref_topk(x, k):
    N, K = x.shape
    vals = np.empty((k, K), np.float32)
    rows = np.empty((k, K), np.int64)
    for c in range(K):
        order = np.lexsort((np.arange(N), -x[:, c].astype(np.float64)))[:k]
        vals[:, c] = x[order, c]
        rows[:, c] = order
    return vals, rows


@pytest.mark.gpu
@pytest.mark.parametrize("N,K,k", [(1000, 1, 10), (5000, 64, 100), (3000, 100, 7),
                                   (257, 3, 257), (100, 5, 1)])
def test_topk_matches_numpy(N, K, k, device):
    rng = np.random.default_rng(N + K)
    x = rng.random((N, K)).astype(np.float32)
    x[rng.random((N, K)) < 0.3] = 0.0                 # many ties at zero
    x[:, 0] = np.round(x[:, 0] * 4) / 4               # heavy ties in column 0
    if K > 2:
        x[:, 2] = -x[:, 2]                             # negative values
    v, r = graph_link.topk_columns(torch.from_numpy(x).to(device), k)
    vr, rr = _ref_topk(x, k)
    assert np.array_equal(v.cpu().numpy(), vr)
    assert np.array_equal(r.cpu().numpy(), rr)


@pytest.mark.gpu
def test_topk_of_a_ppr_result(device):
    A, dang = graph_link.graphs.transition_matrix(*graph_link.graphs.sbm_edges(2000, 100, 0.05,
                                                                               0.001, seed=1), 2000)
    pbr = graph_link.csr_to_pbr(A, 8, 8, 2)
    res = graph_link.run_ppr(pbr, [[5], [50, 60]], tol=1e-7, max_iter=200, dangling=dang,
                             dangling_mode="uniform")
    v, r = graph_link.topk_columns(res.X, 20)
    X = res.X.cpu().numpy()
    vr, rr = _ref_topk(X, 20)
    assert np.array_equal(r.cpu().numpy(), rr) and np.array_equal(v.cpu().numpy(), vr)
    v1, r1 = graph_link.topk_columns(res.X[:, 1].contiguous(), 5)
    assert np.array_equal(r1.cpu().numpy(), rr[:5, 1])
