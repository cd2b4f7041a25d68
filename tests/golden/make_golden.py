"""Writes tests/golden/pbr_small.json: small CSR matrices with their PBR arrays in canonical
tile order, computed by the naive dictionary builder below (independent of the C++ oracle;
the semantics are those of Backend/graph_link/pbr_matrix/pbr_matrix.cpp:340-477: bit index
r*BC+c, keep a tile iff popcount >= min_nnz_per_block, values in bit order, remainder rows
sorted by column). Run: python tests/golden/make_golden.py"""
import json
import os

import numpy as np
import scipy.sparse as sp


def naive_pbr(A, br, bc, mn):
    A = sp.csr_matrix(A)
    blocks = {}
    for r in range(A.shape[0]):
        for j in range(A.indptr[r], A.indptr[r + 1]):
            c = int(A.indices[j])
            blocks.setdefault((r // br, c // bc), {})[(r % br) * bc + c % bc] = float(A.data[j])
    codes, coords, offs, data = [], [], [0], []
    rem = {r: [] for r in range(A.shape[0])}
    for (sb, cb) in sorted(blocks):
        bits = blocks[(sb, cb)]
        if len(bits) >= mn:
            codes.append(sum(1 << b for b in bits))
            coords += [sb * br, cb * bc]
            data += [bits[b] for b in sorted(bits)]
            offs.append(offs[-1] + len(bits))
        else:
            for b, v in bits.items():
                rem[sb * br + b // bc].append((cb * bc + b % bc, v))
    ptr, cols, vals = [0], [], []
    for r in range(A.shape[0]):
        for c, v in sorted(rem[r]):
            cols.append(c)
            vals.append(v)
        ptr.append(len(cols))
    return dict(block_codes=codes, block_coords=coords, block_offsets=offs, block_data=data,
                remainder_indptr=ptr, remainder_col_ind=cols, remainder_data=vals)


def main():
    rng = np.random.default_rng(2024)
    cases = []
    specs = [("dense_4x4_b2", np.arange(1, 17, dtype=np.float32).reshape(4, 4), (2, 2), 3),
             ("diag_10_b4", np.diag(np.arange(1, 11, dtype=np.float32)), (4, 4), 2),
             ("band_12_b8", np.triu(np.tril(np.ones((12, 12), np.float32), 1), -1), (8, 8), 4),
             ("rect_5x9_b2x3", (rng.random((5, 9)) < 0.4).astype(np.float32), (2, 3), 2),
             ("empty_rows_b8", np.pad(np.eye(3, dtype=np.float32), ((0, 6), (0, 6))), (8, 8), 1),
             ("wide_3x70_b1x64", (rng.random((3, 70)) < 0.3).astype(np.float32) * 2.5, (1, 64), 5)]
    for name, dense, block, mn in specs:
        A = sp.csr_matrix(dense)
        cases.append(dict(name=name, shape=list(A.shape), indptr=A.indptr.tolist(),
                          indices=A.indices.tolist(), data=A.data.tolist(), block=list(block),
                          min_nnz=mn, order="canonical", expect=naive_pbr(A, *block, mn)))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pbr_small.json")
    with open(out, "w") as f:
        json.dump(cases, f, indent=0)
    print(f"wrote {len(cases)} cases to {out}")


if __name__ == "__main__":
    main()
