"""Graph ingest, seed selection and top-k around a PPR run (Tests/wiki_ppr.py:29-56, 162,
187): an edge list (CSV/TSV, optionally gzip) to a compact node index and the adjacency
A[target, source], seeds by out-degree, and per-column top-k of the ranks on the GPU."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import torch

from . import _native


def load_edge_list(path, src_col="page_id_from", dst_col="page_id_to", sep=None,
                   chunksize=2_000_000):
    """Read the two id columns of an edge list in chunks and compact the ids to 0..N-1
    (Tests/wiki_ppr.py:29-56). sep=None: tab for .tsv / .gz files, comma otherwise.
    Returns (from_idx, to_idx, node_ids, N)."""
    import pandas as pd
    p = str(path)
    if sep is None:
        sep = "\t" if (p.endswith(".gz") or p.endswith(".tsv")) else ","
    froms, tos = [], []
    for chunk in pd.read_csv(p, sep=sep, usecols=[src_col, dst_col],
                             dtype={src_col: np.int64, dst_col: np.int64},
                             chunksize=chunksize):
        froms.append(chunk[src_col].to_numpy())
        tos.append(chunk[dst_col].to_numpy())
    from_ids = np.concatenate(froms) if froms else np.zeros(0, np.int64)
    to_ids = np.concatenate(tos) if tos else np.zeros(0, np.int64)
    node_ids = np.unique(np.concatenate([from_ids, to_ids]))
    return (np.searchsorted(node_ids, from_ids), np.searchsorted(node_ids, to_ids), node_ids,
            len(node_ids))


def edges_to_adjacency(from_idx, to_idx, N, weights=None):
    """A[target, source] (float32 CSR, duplicate edges summed, canonical) and the
    out-degree (edge weight) of every node."""
    w = np.ones(len(from_idx), np.float32) if weights is None else np.asarray(weights, np.float32)
    A = sp.csr_matrix((w, (np.asarray(to_idx), np.asarray(from_idx))), shape=(N, N))
    A.sum_duplicates()
    A.sort_indices()
    out_deg = np.asarray(A.sum(axis=0), dtype=np.float64).ravel()
    return A, out_deg


def top_out_degree_seeds(out_deg, k):
    """The k nodes of highest out-degree, highest first (Tests/wiki_ppr.py:162)."""
    return np.argsort(np.asarray(out_deg))[::-1][:k].astype(np.int64)


def topk_columns(x: torch.Tensor, k: int):
    """(values [k, K], rows [k, K]) of every column of x [N, K] (float32, on the GPU), the
    largest first and equal values by ascending row; x may be 1-D ([N] -> [k])."""
    squeeze = x.dim() == 1
    x2 = x.unsqueeze(1) if squeeze else x
    if x2.dtype != torch.float32 or x2.stride(1) != 1:
        raise ValueError("topk_columns needs a float32 tensor with unit column stride")
    N, K = x2.shape
    if not 1 <= k <= N:
        raise ValueError(f"k must be in [1, {N}]")
    dev = x2.device
    L = _native.lib()
    wsb = L.gl_topk_workspace_bytes(K, k)
    if wsb < 0:
        raise ValueError("invalid top-k sizes")
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    vals = torch.empty((k, K), dtype=torch.float32, device=dev)
    rows = torch.empty((k, K), dtype=torch.int64, device=dev)
    with torch.cuda.device(dev):
        _native.check(L.gl_topk_columns_f32(x2.data_ptr(), N, K, x2.stride(0), k, vals.data_ptr(),
                                            rows.data_ptr(), ws.data_ptr(), wsb,
                                            _native.stream_handle(dev)), "gl_topk_columns_f32")
    return (vals[:, 0], rows[:, 0]) if squeeze else (vals, rows)
