"""ORACLE — test infrastructure only.

CPU restatements of the reference's algorithms for the Personalized-PageRank hot path, used
as the checker by ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg. The product package (``page_rank_amd``) never imports anything from here.

Modules:
  pbr   ctypes wrapper of ``pbr_oracle.cpp`` (builder, CPU SpMM, PBR->CSR round trip).
  ppr   FP64 scipy/numpy power iteration (reference loop, networkx semantics, exact solve).
  mlp   FP64 numpy forward/backward of the warm-start MLP.

Pinning (what each function is checked against, see ``tests/test_oracle_*.py``):
  pbr.csr_to_pbr      structural invariants, tiles+remainder == input bit-for-bit,
                      hand-derived fixtures, and a pure-Python model of the hashtable
                      emission order checked against the library-driven order.
  pbr.spmm            scipy ``csr @ X`` and dense brute force, the I8 known answer.
  pbr.pbr_to_csr      exact round trip, reference refusal of partial stripes.
  ppr.*               networkx ``pagerank`` (library), ``spsolve`` of (I - aA)x = (1-a)p,
                      closed forms (alpha=0, K5, isolated seed), mass conservation.
  mlp.*               central finite differences.
"""
