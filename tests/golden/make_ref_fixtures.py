"""Writes tests/golden/ref_pbr.npz: PBR matrices built by the reference's own C++ builder
(Backend/graph_link/pbr_matrix/pbr_matrix.cpp:340-477) and products of its own CPU SpMM
(:17-68), for the tests that pin the oracle's and the product's builders (tile order inside a
stripe included) and the product's host SpMM against the reference itself.

The reference's CPU module is compiled from its sources by `make -C oracle ref` (into
oracle/_ref/, in this container only). Run:
    make -C oracle ref && python tests/golden/make_ref_fixtures.py

Cases: int32 and int64 keys (the builder's hash keys are index_t, so the tile order depends on
them), float32 and float64 values, 8x8 / 4x4 / 2x2 / rectangular blocks, partial last stripes,
stripes with more distinct blocks than the initial bucket count (rehashing), and an int32
matrix whose block index passes 2^31 (truncated keys, pbr_matrix.cpp:386-389).
"""
import os
import sys

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
FIELDS = ("block_codes", "block_coords", "block_offsets", "block_data", "remainder_indptr",
          "remainder_col_ind", "remainder_data")


def load_reference():
    sys.path.insert(0, os.path.join(ROOT, "oracle", "_ref"))
    import torch  # noqa: F401  (the modules link against torch)
    import graph_link_core as core
    import graph_link_ref_spmm as rs
    return core, rs


def clustered(rng, rows, cols, per_row, spread):
    """Rows with entries near a per-stripe centre plus a few far ones (many blocks/stripe)."""
    r, c = [], []
    for i in range(rows):
        centre = (i * cols) // rows
        near = np.clip(centre + rng.integers(-spread, spread + 1, per_row), 0, cols - 1)
        far = rng.integers(0, cols, max(1, per_row // 8))
        cc = np.unique(np.concatenate([near, far]))
        r.append(np.full(cc.size, i))
        c.append(cc)
    r, c = np.concatenate(r), np.concatenate(c)
    v = rng.random(r.size) + 0.25
    return sp.csr_matrix((v, (r, c)), shape=(rows, cols))


def cases(rng):
    yield "i32_f32_b8_rehash", clustered(rng, 96, 4096, 48, 300), np.int32, np.float32, (8, 8), 2
    yield "i64_f32_b8_rehash", clustered(rng, 96, 4096, 48, 300), np.int64, np.float32, (8, 8), 2
    yield "i32_f64_b4_partial", clustered(rng, 250, 600, 12, 40), np.int32, np.float64, (4, 4), 3
    yield "i64_f64_b2_default", sp.random(200, 200, density=0.05, format="csr",
                                          random_state=7), np.int64, np.float64, (2, 2), 2
    yield "i32_f32_b2x3_rect", clustered(rng, 61, 187, 9, 12), np.int32, np.float32, (2, 3), 2
    # int32 keys wrap: block_index = block_row * 131072 + block_col passes 2^31 for
    # block rows >= 16384 (rows >= 131072); entries only in the last stripes keep it small
    rows, cols = 131096, 1 << 20
    tail = clustered(rng, 24, cols, 40, 2000).tocoo()
    big = sp.csr_matrix((tail.data, (tail.row + rows - 24, tail.col)), shape=(rows, cols))
    yield "i32_f32_b8_wrapkeys", big, np.int32, np.float32, (8, 8), 2


def main():
    core, rs = load_reference()
    import torch
    rng = np.random.default_rng(20261204)
    out = {}
    names = []
    for name, A, it, st, (br, bc), mn in cases(rng):
        A = sp.csr_matrix(A, dtype=st)
        A.sum_duplicates()
        A.sort_indices()
        indptr, indices = A.indptr.astype(it), A.indices.astype(it)
        p = core.csr_to_pbr(indptr, indices, A.data, A.shape[0], A.shape[1], br, bc, mn)
        out[f"{name}/shape"] = np.array([A.shape[0], A.shape[1], br, bc, mn], np.int64)
        out[f"{name}/indptr"] = indptr
        out[f"{name}/indices"] = indices
        out[f"{name}/data"] = A.data
        for f in FIELDS:
            out[f"{name}/{f}"] = np.asarray(getattr(p, f))
        if A.shape[0] <= 4096 and not name.startswith("i64_f32"):
            # X in multiples of 1/64 (compresses; the products still round)
            X = torch.from_numpy((rng.integers(1, 65, (A.shape[1], 3)) / 64).astype(st))
            out[f"{name}/X"] = X.numpy()
            out[f"{name}/Y"] = rs.pbr_batched_matmul_cpu(p, X)[0].numpy()
        names.append(name)
        print(f"{name}: T={p.accounted_blocks()} V={p.compressed_nnz()} R={p.remainder_nnz()}")
    out["names"] = np.array(names)
    path = os.path.join(HERE, "ref_pbr.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {len(names)} cases to {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
