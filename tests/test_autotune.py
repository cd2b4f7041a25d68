"""Autotuner: candidate sets, the JSON cache (CPU), and on the GPU that tuned launches give
the same results as the default ones and that a cached signature is not timed again."""
import json

import numpy as np
import pytest
import torch

import page_rank_amd as graph_link
from page_rank_amd import autotune
from tests.helpers import blocky_csr, random_x


def test_kw_candidates_cover_every_width():
    assert autotune.kw_candidates(1) == [1, 16]
    assert autotune.kw_candidates(8) == [2, 4, 8, 16]
    assert autotune.kw_candidates(32) == [4, 8, 16, 32]
    assert autotune.kw_candidates(64) == [8, 16, 32, 64]
    assert autotune.kw_candidates(200) == [8, 16, 32, 64]
    for K in range(1, 130):
        assert all(w in autotune.KW_ALL for w in autotune.kw_candidates(K))


def test_cache_round_trip(tmp_path, monkeypatch):
    path = tmp_path / "sub" / "tune.json"
    monkeypatch.setenv("PAGE_RANK_AMD_TUNE_CACHE", str(path))
    c = autotune.cache()
    assert c.get("k") is None
    c.put("k", {"kw": 32, "mode": "fused"})
    assert json.loads(path.read_text())["k"]["kw"] == 32
    monkeypatch.setattr(autotune, "_cache", None)
    assert autotune.cache().get("k") == {"kw": 32, "mode": "fused"}


def test_corrupt_cache_is_ignored(tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    path.write_text("{not json")
    monkeypatch.setenv("PAGE_RANK_AMD_TUNE_CACHE", str(path))
    monkeypatch.setattr(autotune, "_cache", None)
    assert autotune.cache().entries == {}


@pytest.mark.gpu
def test_tuned_spmm_and_ppr_match_defaults(tmp_path, monkeypatch, device):
    monkeypatch.setenv("PAGE_RANK_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(autotune, "_cache", None)
    A = blocky_csr(3000, 8, 4, 2, 30, scatter=4000, seed=3)
    pbr = graph_link.csr_to_pbr(A, 8, 8, 4).to(device)
    X = torch.from_numpy(random_x(3000, 24)).to(device)
    y0 = torch.empty((3000, 24), device=device)
    y1 = torch.empty_like(y0)
    graph_link.spmm_into(pbr, X, y0)
    graph_link.spmm_into(pbr, X, y1, kw="auto")
    t = autotune.tune_spmm(pbr, 24)
    assert t["kw"] in autotune.kw_candidates(24) and t["mode"] in ("fused", "split")
    assert len(t["table"]) == 2 * len(autotune.kw_candidates(24))
    assert np.allclose(y0.cpu().numpy(), y1.cpu().numpy(), rtol=1e-5, atol=1e-6)
    # a cached signature is answered without timing
    monkeypatch.setattr(autotune, "_best", lambda *a, **k: pytest.fail("timed again"))
    assert autotune.tune_spmm(pbr, 24) == t
    monkeypatch.undo()
    monkeypatch.setenv("PAGE_RANK_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    # PPR on a column-stochastic matrix
    Acs = A.multiply(1.0 / np.maximum(np.asarray(A.sum(0)).ravel(), 1e-30)).tocsr().astype(np.float32)
    P = graph_link.csr_to_pbr(Acs, 8, 8, 4).to(device)
    sets = [[i] for i in range(0, 3000, 300)]
    a = graph_link.run_ppr(P, sets, tol=1e-6, max_iter=60)
    b = graph_link.run_ppr(P, sets, tol=1e-6, max_iter=60, kw="auto")
    tp = autotune.tune_ppr(P, len(sets))
    assert tp["grid_mult"] in autotune.GRID_MULTS
    # other launch shapes sum in another order: the stopping step may move by one
    assert np.abs(a.iterations - b.iterations).max() <= 1
    assert np.abs(a.X.cpu().numpy() - b.X.cpu().numpy()).sum(0).max() < 1e-5
    best, res = autotune.tune_block(A, 8, blocks=(4, 8), device=device)
    assert best in (4, 8) and set(res) == {4, 8}


@pytest.mark.gpu
def test_ppr_tile_threshold_tuning(tmp_path, monkeypatch, device):
    monkeypatch.setenv("PAGE_RANK_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(autotune, "_cache", None)
    A = blocky_csr(3000, 8, 4, 2, 30, scatter=4000, seed=5)
    Acs = A.multiply(1.0 / np.maximum(np.asarray(A.sum(0)).ravel(), 1e-30)).tocsr().astype(np.float32)
    P4 = graph_link.csr_to_pbr(Acs, 8, 8, 4).to(device)
    sets = [[i, i + 7] for i in range(0, 3000, 200)]
    best, P, res = autotune.tune_ppr_tiles(Acs, len(sets), device=device, built={4: P4})
    assert best in autotune.PPR_MIN_NNZ and set(res) == set(autotune.PPR_MIN_NNZ)
    # the chosen matrix is the one built with that threshold: more entries leave the tiles
    # as the threshold rises
    tiles = {mn: graph_link.csr_to_pbr(Acs, 8, 8, mn).accounted_blocks()
             for mn in autotune.PPR_MIN_NNZ}
    assert tiles[4] >= tiles[8] >= tiles[16] and tiles[4] > tiles[16]
    assert P.accounted_blocks() == tiles[best] and P.total_nnz == P4.total_nnz
    # same power iteration on another split of the same matrix: sums in another order only
    a = graph_link.run_ppr(P4, sets, tol=1e-7, max_iter=100)
    b = graph_link.run_ppr(P, sets, tol=1e-7, max_iter=100)
    assert np.abs(a.iterations - b.iterations).max() <= 1
    assert np.abs(a.X.cpu().numpy() - b.X.cpu().numpy()).sum(0).max() < 1e-5
