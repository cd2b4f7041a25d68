"""GPU-metadata registry: device buffers of a PBR matrix keyed by id(matrix).

Same semantics as Backend/graph_link/graph_link/pbr_registry.py:1-31 — entries die with the
matrix (weakref.finalize), and an entry whose remainder indptr does not match the matrix's
row count (a reused id() after garbage collection) is dropped on lookup.
"""
import weakref

_PBR_GPU_CACHE = {}


def get_pbr_gpu_meta(pbr_mat):
    obj_id = id(pbr_mat)
    meta = _PBR_GPU_CACHE.get(obj_id)
    if meta is not None and meta["rem_indptr"].numel() != pbr_mat.rows + 1:
        del _PBR_GPU_CACHE[obj_id]
        return None
    return meta


def set_pbr_gpu_meta(pbr_mat, meta_dict):
    obj_id = id(pbr_mat)
    _PBR_GPU_CACHE[obj_id] = meta_dict
    try:
        weakref.finalize(pbr_mat, _PBR_GPU_CACHE.pop, obj_id, None)
    except TypeError:
        pass


def clear_pbr_gpu_meta(pbr_mat):
    _PBR_GPU_CACHE.pop(id(pbr_mat), None)
