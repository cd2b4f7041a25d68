"""Parity against the reference itself: tests/golden/ref_pbr.npz holds PBR matrices built by
the reference's own C++ builder (Backend/graph_link/pbr_matrix/pbr_matrix.cpp:340-477,
compiled from its sources by `make -C oracle ref`) and products of its own CPU SpMM
(:17-68); tests/golden/make_ref_fixtures.py wrote them.

CPU: the oracle's builder (reference tile order) and the product's host SpMM
(gl_pbr_spmm_host_*) must reproduce them bit for bit. GPU: the product's GPU builder in
reference order must too."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle import pbr as opbr
from page_rank_amd import _native, ops
from page_rank_amd.pbr import from_arrays

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "ref_pbr.npz")
FIELDS = ("block_codes", "block_coords", "block_offsets", "block_data", "remainder_indptr",
          "remainder_col_ind", "remainder_data")


def _cases():
    with np.load(FIXTURE, allow_pickle=False) as z:
        names = [str(n) for n in z["names"]]
        return {n: {k.split("/", 1)[1]: z[k] for k in z.files if k.startswith(n + "/")}
                for n in names}


CASES = _cases()


def test_fixture_covers_the_key_cases():
    names = set(CASES)
    assert {"i32_f32_b8_rehash", "i64_f32_b8_rehash", "i32_f32_b8_wrapkeys"} <= names
    wrap = CASES["i32_f32_b8_wrapkeys"]
    rows, cols, br, bc, _ = wrap["shape"]
    ncb = (cols + bc - 1) // bc
    assert (rows - 1) // br * ncb >= 2**31  # its block indices pass int32


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_builder_matches_reference(name):
    c = CASES[name]
    rows, cols, br, bc, mn = (int(v) for v in c["shape"])
    p = opbr.csr_to_pbr(c["indptr"], c["indices"], c["data"], rows, cols, br, bc, mn,
                        order="reference")
    for f in FIELDS:
        got, want = np.asarray(getattr(p, f)), c[f]
        assert got.dtype == want.dtype, (f, got.dtype, want.dtype)
        assert np.array_equal(got, want), f


@pytest.mark.parametrize("name", sorted(n for n in CASES if "Y" in CASES[n]))
def test_host_spmm_bit_identical_to_reference(name):
    c = CASES[name]
    rows, cols, br, bc, _ = (int(v) for v in c["shape"])
    p = from_arrays(rows, cols, br, bc, *(c[f] for f in FIELDS))
    X = torch.from_numpy(c["X"])
    for threads in (1, 3, 0):
        Y = ops.pbr_batched_matmul_cpu(p, X, threads=threads)[0].numpy()
        assert Y.dtype == c["Y"].dtype
        assert np.array_equal(Y, c["Y"]), (threads, np.abs(Y - c["Y"]).max())
    # and through pbr_matmul with a CPU tensor, 2-D and 3-D
    assert np.array_equal(ops.pbr_matmul(p, X).numpy(), c["Y"])
    X3 = torch.stack([X, X * 0.5])
    Y3 = ops.pbr_matmul(p, X3).numpy()
    assert Y3.shape == (2, rows, X.shape[1]) and np.array_equal(Y3[0], c["Y"])
    # the oracle's restatement of the same loop agrees bit for bit too
    op = opbr.OraclePbr(rows, cols, br, bc, int(c["data"].size), *(c[f] for f in FIELDS))
    assert np.array_equal(opbr.spmm(op, c["X"]), c["Y"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_builder_matches_reference(name, device):
    import page_rank_amd as graph_link
    c = CASES[name]
    rows, cols, br, bc, mn = (int(v) for v in c["shape"])
    A = sp.csr_matrix((c["data"], c["indices"], c["indptr"]), shape=(rows, cols))
    with torch.cuda.device(device):
        p = graph_link.csr_to_pbr(A, br, bc, mn, order="reference")
    for f in FIELDS:
        got, want = np.asarray(getattr(p, f)), c[f]
        assert np.array_equal(got.astype(want.dtype), want), f
