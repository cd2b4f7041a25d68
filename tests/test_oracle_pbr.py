"""CPU pins of the PBR oracle (oracle/pbr_oracle.cpp): hand-derived fixtures, structural
invariants, an independent Python expansion of tiles + remainder back to the input, the
I8 known answer, and a pure-Python model of the reference's std::unordered_map emission
order (libstdc++ _Hashtable: insert at bucket begin, relink on rehash) driven by the bucket
counts the library reports. The reference ships no golden vectors (its tests recompute
scipy / networkx answers at run time), so these are the pins."""
import json
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import pbr as opbr
from tests.helpers import blocky_csr, random_csr

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------------------------------
# hand-derived fixtures
# ---------------------------------------------------------------------------------------
def test_fixture_4x4_blocks_2x2():
    """A = [[1,0,2,0],[0,3,0,0],[0,0,0,4],[5,0,0,6]], 2x2 blocks, min_nnz 2 (worked by hand:
    block (0,0) bits {0,3} kept, block (0,1) one entry -> remainder row 0, block (1,1) bits
    {1,3} kept, block (1,0) one entry -> remainder row 3)."""
    A = sp.csr_matrix(np.array([[1, 0, 2, 0], [0, 3, 0, 0], [0, 0, 0, 4], [5, 0, 0, 6]],
                               np.float32))
    p = opbr.csr_matrix_to_pbr(A, 2, 2, 2)
    assert p.block_codes.tolist() == [0b1001, 0b1010]
    assert p.block_coords.tolist() == [0, 0, 2, 2]
    assert p.block_offsets.tolist() == [0, 2, 4]
    assert p.block_data.tolist() == [1, 3, 4, 6]
    assert p.remainder_indptr.tolist() == [0, 1, 1, 1, 2]
    assert p.remainder_col_ind.tolist() == [2, 0]
    assert p.remainder_data.tolist() == [2, 5]


def test_identity_8x8_code():
    """I8 in one 8x8 block: bits r*8+r, code = sum 2^(9r) (bit 63 set)."""
    p = opbr.csr_matrix_to_pbr(sp.eye(8, format="csr", dtype=np.float32), 8, 8, 2)
    assert p.block_codes.tolist() == [sum(1 << (9 * r) for r in range(8))]
    assert int(p.block_codes[0]) >> 63 == 1
    assert p.block_data.tolist() == [1.0] * 8
    assert p.remainder_indptr.tolist() == [0] * 9
    Y = opbr.spmm(p, np.ones((8, 1), np.float32))
    assert Y.ravel().tolist() == [1.0] * 8


def test_golden_fixtures():
    """Committed fixtures (tests/golden/make_golden.py): hand-checkable small matrices."""
    with open(os.path.join(GOLDEN, "pbr_small.json")) as f:
        cases = json.load(f)
    for c in cases:
        A = sp.csr_matrix((np.array(c["data"], np.float32), np.array(c["indices"], np.int32),
                           np.array(c["indptr"], np.int32)), shape=tuple(c["shape"]))
        p = opbr.csr_matrix_to_pbr(A, c["block"][0], c["block"][1], c["min_nnz"], c["order"])
        for k in ("block_codes", "block_coords", "block_offsets", "block_data",
                  "remainder_indptr", "remainder_col_ind", "remainder_data"):
            assert np.asarray(getattr(p, k)).tolist() == c["expect"][k], (c["name"], k)


# ---------------------------------------------------------------------------------------
# invariants and round trip
# ---------------------------------------------------------------------------------------
def expand(p):
    """Independent Python expansion of a PBR matrix to {(row, col): value}."""
    out = {}
    for t in range(p.num_tiles):
        code = int(p.block_codes[t])
        r0, c0 = int(p.block_coords[2 * t]), int(p.block_coords[2 * t + 1])
        k = int(p.block_offsets[t])
        for bit in range(p.block_rows * p.block_cols):
            if code >> bit & 1:
                key = (r0 + bit // p.block_cols, c0 + bit % p.block_cols)
                assert key not in out
                out[key] = p.block_data[k]
                k += 1
        assert k == p.block_offsets[t + 1]
    for r in range(p.rows):
        for j in range(p.remainder_indptr[r], p.remainder_indptr[r + 1]):
            key = (r, int(p.remainder_col_ind[j]))
            assert key not in out
            out[key] = p.remainder_data[j]
    return out


def check_invariants(A, p, min_nnz):
    br, bc = p.block_rows, p.block_cols
    pc = [bin(int(c)).count("1") for c in p.block_codes]
    assert p.block_offsets[0] == 0 and np.array_equal(np.diff(p.block_offsets), pc)
    assert p.block_offsets[-1] == p.block_data.size
    assert all(c >= min_nnz for c in pc)
    rows0 = p.block_coords[0::2]
    cols0 = p.block_coords[1::2]
    assert np.all(rows0 % br == 0) and np.all(cols0 % bc == 0)
    assert np.all(np.diff(rows0) >= 0)  # stripes ascend
    assert len(set(zip(rows0.tolist(), cols0.tolist()))) == p.num_tiles
    for r in range(p.rows):
        seg = p.remainder_col_ind[p.remainder_indptr[r]:p.remainder_indptr[r + 1]]
        assert np.all(np.diff(seg) > 0)
    # remainder blocks are below the threshold and never collide with a tile
    tiles = set(zip((rows0 // br).tolist(), (cols0 // bc).tolist()))
    rb = {}
    for r in range(p.rows):
        for j in range(p.remainder_indptr[r], p.remainder_indptr[r + 1]):
            key = (r // br, int(p.remainder_col_ind[j]) // bc)
            rb[key] = rb.get(key, 0) + 1
    assert all(v < min_nnz for v in rb.values())
    assert not (set(rb) & tiles)
    # tiles + remainder reproduce the input bit-for-bit
    coo = A.tocoo()
    want = {(int(r), int(c)): v for r, c, v in zip(coo.row, coo.col, coo.data)}
    got = expand(p)
    assert set(got) == set(want)
    assert all(np.asarray(got[k]).tobytes() == np.asarray(want[k]).tobytes() for k in want)


@pytest.mark.parametrize("block", [(2, 2), (4, 4), (8, 8), (3, 5), (1, 64)])
@pytest.mark.parametrize("min_nnz", [0, 1, 3, 4])
@pytest.mark.parametrize("order", ["reference", "canonical"])
def test_invariants_random(block, min_nnz, order):
    A = random_csr(97, 0.08, seed=block[0] * 10 + min_nnz)
    p = opbr.csr_matrix_to_pbr(A, block[0], block[1], min_nnz, order)
    check_invariants(A, p, min_nnz)
    if order == "canonical":
        for s in np.unique(p.block_coords[0::2]):
            sel = p.block_coords[0::2] == s
            assert np.all(np.diff(p.block_coords[1::2][sel]) > 0)


def test_invariants_blocky():
    A = blocky_csr(803, 8, 4, 1, 40, scatter=900, seed=2)
    for mn in (1, 4, 9):
        check_invariants(A, opbr.csr_matrix_to_pbr(A, 8, 8, mn), mn)


def test_orders_hold_the_same_tiles():
    A = blocky_csr(2000, 8, 6, 1, 30, scatter=3000, seed=4)
    a = opbr.csr_matrix_to_pbr(A, 8, 8, 4, "reference")
    b = opbr.csr_matrix_to_pbr(A, 8, 8, 4, "canonical")
    ka = {(int(a.block_coords[2 * t]), int(a.block_coords[2 * t + 1])): int(a.block_codes[t])
          for t in range(a.num_tiles)}
    kb = {(int(b.block_coords[2 * t]), int(b.block_coords[2 * t + 1])): int(b.block_codes[t])
          for t in range(b.num_tiles)}
    assert ka == kb
    assert np.array_equal(a.remainder_col_ind, b.remainder_col_ind)
    assert a.block_coords.tolist() != b.block_coords.tolist()  # order genuinely differs


def test_round_trip_and_partial_stripe_refusal():
    A = random_csr(50, 0.2, seed=8)  # 50 rows: partial last stripe for 8x8
    p = opbr.csr_matrix_to_pbr(A, 8, 8, 1)
    ip, ix, dv = opbr.pbr_to_csr(p)
    B = sp.csr_matrix((dv, ix, ip), shape=A.shape)
    assert (B != A).nnz == 0
    assert any(p.block_coords[0::2] == 48)
    with pytest.raises(RuntimeError):
        opbr.pbr_to_csr(p, strict=True)  # the reference throws for a trailing partial stripe
    q = opbr.csr_matrix_to_pbr(random_csr(48, 0.2, seed=8), 8, 8, 1)
    opbr.pbr_to_csr(q, strict=True)


def test_cpu_spmm_matches_scipy_and_dense():
    A = blocky_csr(300, 4, 3, 1, 16, scatter=500, seed=6)
    X = np.random.default_rng(1).random((300, 7)).astype(np.float32)
    p = opbr.csr_matrix_to_pbr(A, 4, 4, 3)
    Y = opbr.spmm(p, X, wide=True)
    assert np.allclose(Y, A.astype(np.float64) @ X.astype(np.float64), rtol=1e-6, atol=1e-6)
    assert np.allclose(opbr.spmm(p, X), A.toarray() @ X, rtol=1e-5, atol=1e-5)


# ---------------------------------------------------------------------------------------
# emission order: pure-Python model of libstdc++'s _Hashtable (hashtable.h
# _M_insert_bucket_begin and _M_rehash_aux for unique keys), identity hash on the key
# converted to size_t, bucket counts taken from the library at each insert.
# ---------------------------------------------------------------------------------------
BB = "before_begin"


class HashtableModel:
    def __init__(self, nbkt):
        self.n = nbkt
        self.head = None
        self.nxt = {}
        self.buckets = {}

    def bkt(self, k, n=None):
        return (k % (1 << 64)) % (n or self.n)

    def _after(self, prev):
        return self.head if prev is BB else self.nxt[prev]

    def _link_after(self, prev, node):
        if prev is BB:
            self.head = node
        else:
            self.nxt[prev] = node

    def rehash(self, n):
        p, self.head, bbegin, newb = self.head, None, 0, {}
        while p is not None:
            nx, b = self.nxt[p], self.bkt(p, n)
            if b not in newb:
                self.nxt[p] = self.head
                self.head = p
                newb[b] = BB
                if self.nxt[p] is not None:
                    newb[bbegin] = p
                bbegin = b
            else:
                prev = newb[b]
                self.nxt[p] = self._after(prev)
                self._link_after(prev, p)
            p = nx
        self.buckets, self.n = newb, n

    def insert(self, k, nbkt_after):
        if nbkt_after != self.n:
            self.rehash(nbkt_after)
        b = self.bkt(k)
        if b in self.buckets:
            prev = self.buckets[b]
            self.nxt[k] = self._after(prev)
            self._link_after(prev, k)
        else:
            self.nxt[k] = self.head
            self.head = k
            if self.nxt[k] is not None:
                self.buckets[self.bkt(self.nxt[k])] = k
            self.buckets[b] = BB

    def clear(self):
        self.head, self.nxt, self.buckets = None, {}, {}

    def order(self):
        out, p = [], self.head
        while p is not None:
            out.append(p)
            p = self.nxt[p]
        return out


def model_tile_order(A, br, bc, min_nnz, key_dtype=np.int32):
    """Expected (row0, col0) sequence of kept tiles, derived without the oracle's builder."""
    n_rows, n_cols = A.shape
    ncb = (n_cols + bc - 1) // bc
    hint = int(math.sqrt(ncb))
    stripes = []
    for s in range((n_rows + br - 1) // br):
        seq, seen, cnt = [], set(), {}
        for r in range(s * br, min((s + 1) * br, n_rows)):
            for c in A.indices[A.indptr[r]:A.indptr[r + 1]]:
                b = int(c) // bc
                cnt[b] = cnt.get(b, 0) + 1
                if b not in seen:
                    seen.add(b)
                    seq.append(b)
        stripes.append((s, seq, cnt))
    raw = [s * ncb + b for s, seq, _ in stripes for b in seq]
    keys = np.array(raw, dtype=np.uint64).astype(key_dtype)  # truncation to index_t
    sptr = np.cumsum([0] + [len(seq) for _, seq, _ in stripes])
    counts = opbr.bucket_trace(hint, keys, sptr, key_dtype)
    n0 = opbr.bucket_counts(hint, [])[0]
    model = HashtableModel(n0)
    out, i = [], 0
    for s, seq, cnt in stripes:
        kb = {}
        for b in seq:
            k = int(keys[i])
            model.insert(k, int(counts[i]))
            kb[k] = b
            i += 1
        for k in model.order():
            if cnt[kb[k]] >= min_nnz:
                out.append((s * br, kb[k] * bc))
        model.clear()
    return out


@pytest.mark.parametrize("case", [
    dict(n=2000, block=8, mn=1, seed=1, band=40),
    dict(n=4000, block=4, mn=2, seed=2, band=300),
    dict(n=3001, block=2, mn=1, seed=3, band=900),
    dict(n=1500, block=8, mn=4, seed=4, band=180),
])
def test_reference_order_matches_hashtable_model(case):
    A = blocky_csr(case["n"], case["block"], 7, 1, 20, scatter=2 * case["n"],
                   seed=case["seed"], band=case["band"])
    p = opbr.csr_matrix_to_pbr(A, case["block"], case["block"], case["mn"], "reference")
    got = list(zip(p.block_coords[0::2].tolist(), p.block_coords[1::2].tolist()))
    assert got == model_tile_order(A, case["block"], case["block"], case["mn"])


def test_reference_order_int64_keys():
    A = blocky_csr(1200, 4, 6, 1, 12, scatter=2000, seed=12, band=100)
    A.indices = A.indices.astype(np.int64)
    A.indptr = A.indptr.astype(np.int64)
    p = opbr.csr_matrix_to_pbr(A, 4, 4, 2, "reference")
    got = list(zip(p.block_coords[0::2].tolist(), p.block_coords[1::2].tolist()))
    assert got == model_tile_order(A, 4, 4, 2, key_dtype=np.int64)


def test_key_truncation_is_modelled():
    """Block keys above 2^31 wrap to negative int32 keys (pbr_matrix.cpp:386-389): a wide
    matrix whose stripe-10 and stripe-11 keys exceed int32 still matches the model."""
    n_cols = 2**31 - 8
    rows = np.array([0, 0, 1, 80, 80, 80, 81, 88, 89])
    cols = np.array([5, 2**30, 2**31 - 20, 3, 2**30 + 100, 2**31 - 9, 77, 8, 2**29])
    A = sp.csr_matrix((np.ones(9, np.float32), (rows, cols)), shape=(90, n_cols))
    A.indices = A.indices.astype(np.int32)
    A.indptr = A.indptr.astype(np.int32)
    p = opbr.csr_matrix_to_pbr(A, 8, 8, 1, "reference")
    got = list(zip(p.block_coords[0::2].tolist(), p.block_coords[1::2].tolist()))
    assert got == model_tile_order(A, 8, 8, 1)
    ncb = (n_cols + 7) // 8
    assert 10 * ncb + (2**30 + 100) // 8 > 2**31  # the stripe-10 keys really wrap
