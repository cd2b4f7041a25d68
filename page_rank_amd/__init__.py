"""page_rank_amd — MI355X (gfx950) Personalized PageRank on the PBR sparse format.

Drop-in for the ``graph_link`` Python API of Dayan135/Page_Rank
(Backend/graph_link/graph_link/__init__.py:274-285): use ``import page_rank_amd as graph_link``.
All compute runs in hand-written HIP kernels of ``lib/libgraphlink_hip.so`` through a C ABI
(include/graph_link_hip.h); torch tensors are only device buffers.
"""
from .pbr import (DataOrder, PBRMatrix, PBRMatrixInt32Double, PBRMatrixInt32Float,
                  PBRMatrixInt64Double, PBRMatrixInt64Float, from_arrays)
from .registry import clear_pbr_gpu_meta, get_pbr_gpu_meta, set_pbr_gpu_meta
from .ops import (SpmmPlan, block_spmm_into, pbr_batched_matmul_cpu, pbr_batched_matmul_cuda,
                  pbr_batched_matmul_triton_gpu, pbr_matmul, spmm_into)
from .ppr import (PPRResult, normalize_transition_matrix, run_personalized_pagerank,
                  run_ppr)
from .builder import csr_to_pbr
from .csr import (DeviceCSR, load_pbr, normalize_transition_matrix_gpu, pbr_to_csr,
                  pbr_to_scipy, save_pbr, upload_csr)
from . import autotune, graphs, ingest, warmstart
from .ingest import edges_to_adjacency, load_edge_list, top_out_degree_seeds, topk_columns

__all__ = [
    "csr_to_pbr",
    "pbr_to_csr",
    "pbr_to_scipy",
    "save_pbr",
    "load_pbr",
    "DeviceCSR",
    "upload_csr",
    "normalize_transition_matrix_gpu",
    "pbr_matmul",
    "pbr_batched_matmul_cpu",
    "pbr_batched_matmul_cuda",
    "pbr_batched_matmul_triton_gpu",
    "get_pbr_gpu_meta",
    "clear_pbr_gpu_meta",
    "normalize_transition_matrix",
    "run_personalized_pagerank",
    "run_ppr",
    "PPRResult",
    "from_arrays",
    "DataOrder",
    "PBRMatrix",
    "PBRMatrixInt32Float",
    "PBRMatrixInt64Float",
    "PBRMatrixInt32Double",
    "PBRMatrixInt64Double",
    "autotune",
    "ingest",
    "load_edge_list",
    "edges_to_adjacency",
    "top_out_degree_seeds",
    "topk_columns",
    "graphs",
    "warmstart",
]
