"""ctypes binding of libgraphlink_hip.so (the gfx950 C ABI declared in include/graph_link_hip.h).

The library is built in-tree by ``page_rank_amd.build`` and loaded from
``page_rank_amd/lib/``. There is no fallback: every GPU entry point of the package raises
``NativeLibraryMissing`` when the library is absent.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libgraphlink_hip.so")

GL_OK = 0
GL_ERR_INVALID = 1
GL_ERR_HIP = 2
GL_ERR_WORKSPACE = 3

GL_SPMM_FUSED = 0
GL_SPMM_SPLIT = 1
GL_DANGLING_NONE = 0
GL_DANGLING_UNIFORM = 1
GL_DANGLING_PERSONALIZATION = 2
GL_RETIRE_TOGETHER = 0
GL_RETIRE_PER_COLUMN = 1

c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_f32 = ctypes.c_float
c_p = ctypes.c_void_p


class NativeLibraryMissing(ImportError):
    pass


class NativeError(RuntimeError):
    """A HIP runtime failure reported by the C ABI."""


class GlOpts(ctypes.Structure):
    _fields_ = [("alpha", c_f32), ("tol", c_f32), ("max_iter", c_i32),
                ("dangling_mode", c_i32), ("retire_mode", c_i32), ("graph_iters", c_i32),
                ("kw", c_i32), ("grid_mult", c_i32), ("extrapolate_every", c_i32),
                ("single_wg", c_i32)]


# name -> (restype, argtypes)
_PROTOS = {
    "gl_last_error": (ctypes.c_char_p, []),
    "gl_abi_version": (c_i32, []),
    "gl_device_info": (ctypes.c_int, [c_p, c_p, c_p, c_p]),
    "gl_pbr_stripe_index": (ctypes.c_int, [c_i64, c_i32, c_i64, c_p, c_p, c_p]),
    "gl_pbr_spmm_i32_f32": (ctypes.c_int, [c_i64, c_i64, c_i32, c_i32, c_i64, c_i64, c_i64,
                                           c_p, c_p, c_p, c_p,
                                           c_p, c_p, c_p, c_p, c_p, c_i64, c_i64, c_p, c_i64,
                                           c_i32, c_i32, c_i32, c_p]),
    "gl_pbr_exec_layout_bytes": (c_i64, [c_i64]),
    "gl_pbr_exec_layout": (ctypes.c_int, [c_i64, c_i64, c_p, c_p, c_p, c_p, c_p, c_p, c_p]),
    "gl_pbr_spmm_exec_i32_f32": (ctypes.c_int, [c_i64, c_i64, c_i32, c_i32, c_i64, c_i64,
                                                c_i64, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p,
                                                c_p, c_p, c_p, c_i64, c_i64, c_p, c_i64,
                                                c_i32, c_i32, c_i32, c_p]),
    "gl_pbr_block_spmm_i32_f32": (ctypes.c_int, [c_i64, c_i64, c_i32, c_i32, c_i64, c_i64,
                                                 c_p, c_p,
                                                 c_p, c_p, c_p, c_p, c_i64, c_i64, c_p, c_i64,
                                                 c_i32, c_i32, c_p]),
    "gl_csr_spmm_i32_f32": (ctypes.c_int, [c_i64, c_i64, c_i64, c_p, c_p, c_p, c_p, c_i64, c_i64,
                                           c_p,
                                           c_i64, c_i32, c_i32, c_p]),
    "gl_ppr_workspace_bytes": (c_i64, [c_i64, c_i32, c_i64, c_i64]),
    "gl_ppr_run_i32_f32": (ctypes.c_int, [c_i64, c_i64, c_i32, c_i32, c_i64, c_i64, c_i64,
                                          c_p, c_p, c_p, c_p,
                                          c_p, c_p, c_p, c_p, c_i64, c_p, c_p, c_p, c_p, c_p,
                                          ctypes.POINTER(GlOpts), c_p, c_p, c_p, c_p, c_p,
                                          c_i64, c_p, c_p]),
    "gl_pbr_build_workspace_bytes": (c_i64, [c_i64, c_i32, c_i64]),
    "gl_pbr_build_count": (ctypes.c_int, [c_i64, c_i64, c_i64, c_i32, c_i32, c_i32, c_p, c_p,
                                          c_p, c_p, c_i64, c_p, c_p]),
    "gl_pbr_build_fill": (ctypes.c_int, [c_i64, c_i64, c_i64, c_i32, c_i32, c_i32, c_i64, c_i64,
                                         c_p, c_p, c_p, c_i32, c_p, c_i64, c_p, c_p, c_p, c_p,
                                         c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p]),
    "gl_pbr_reference_perm": (ctypes.c_int, [c_i64, c_i64, c_i32, c_i32, c_i32, c_p, c_p, c_p,
                                             c_p, c_p]),
    "gl_pbr_permute_workspace_bytes": (c_i64, [c_i64]),
    "gl_pbr_permute_tiles": (ctypes.c_int, [c_i64, c_p, c_p, c_p, c_p, c_p, c_i32, c_p, c_p, c_p,
                                            c_p, c_p, c_i64, c_p]),
    "gl_mlp_num_params": (c_i32, []),
    "gl_mlp_num_features": (c_i32, []),
    "gl_ws_num_buckets": (c_i32, []),
    "gl_mlp_forward": (ctypes.c_int, [c_p, c_p, c_i64, c_p, c_p]),
    "gl_mlp_train_partials": (c_i64, []),
    "gl_mlp_train_step": (ctypes.c_int, [c_p, c_p, c_p, c_p, c_i64, c_f32, c_f32, c_p, c_p, c_i32,
                                         c_p, c_p, c_p]),
    "gl_mlp_grad": (ctypes.c_int, [c_p, c_p, c_p, c_p, c_i64, c_p, c_p, c_p, c_p]),
    "gl_ws_degrees": (ctypes.c_int, [c_i64, c_i32, c_i64, c_p, c_p, c_p, c_p, c_p, c_p, c_p]),
    "gl_ws_pair_features": (ctypes.c_int, [c_i64, c_i64, c_f32, c_p, c_p, c_p, c_p, c_i64, c_p,
                                           c_p]),
    "gl_ws_table_size": (c_i64, []),
    "gl_ws_table_features": (ctypes.c_int, [c_p, c_p]),
    "gl_ws_node_buckets": (ctypes.c_int, [c_i64, c_p, c_p, c_p, c_p]),
    "gl_ws_assemble": (ctypes.c_int, [c_i64, c_i64, c_f32, c_p, c_p, c_p, c_p, c_p, c_p, c_p,
                                      c_i64, c_p, c_p, c_p]),
    "gl_ws_scatter": (ctypes.c_int, [c_p, c_p, c_i64, c_p, c_p]),
    "gl_ws_normalize": (ctypes.c_int, [c_p, c_i64, c_i64, c_p, c_p]),
    "gl_ws_blend_start": (ctypes.c_int, [c_p, c_p, c_i64, c_i64, c_f32, c_p, c_p]),
    "gl_csr_normalize_workspace_bytes": (c_i64, [c_i64, c_i64]),
    "gl_csr_normalize_columns_f32": (ctypes.c_int, [c_i64, c_i64, c_p, c_p, c_p, c_p, c_p, c_p,
                                                    c_i64, c_p]),
    "gl_pbr_to_csr_workspace_bytes": (c_i64, [c_i64, c_i64]),
    "gl_pbr_to_csr_f64_workspace_bytes": (c_i64, [c_i64, c_i64]),
    **{f"gl_pbr_to_csr_i32_{ft}": (ctypes.c_int, [c_i64, c_i64, c_i32, c_i32, c_i64, c_i64,
                                                  c_i64, c_p, c_p, c_p, c_p, c_p, c_p, c_p,
                                                  c_p, c_p, c_p, c_i64, c_p, c_i64, c_p])
       for ft in ("f32", "f64")},
    "gl_topk_workspace_bytes": (c_i64, [c_i64, c_i64]),
    "gl_topk_columns_f32": (ctypes.c_int, [c_p, c_i64, c_i64, c_i64, c_i64, c_p, c_p, c_p, c_i64,
                                           c_p]),
    "gl_pbr_spmm_i32_f64": (ctypes.c_int, [c_i64, c_i64, c_i32, c_i32, c_i64, c_p, c_p, c_p,
                                           c_p, c_p, c_p, c_p, c_p, c_p, c_i64, c_i64, c_p,
                                           c_i64, c_i32, c_p]),
    "gl_ppr_init_f32": (ctypes.c_int, [c_p, c_p, c_i64, c_i64, c_p]),
    "gl_ppr_init_f64": (ctypes.c_int, [c_p, c_p, c_i64, c_i64, c_p]),
    "gl_ppr_update_f32_workspace_bytes": (c_i64, [c_i64, c_i64]),
    "gl_ppr_update_f64_workspace_bytes": (c_i64, [c_i64, c_i64]),
    "gl_ppr_update_normalized_f32": (ctypes.c_int, [c_p, c_p, c_p, c_f32, c_i64,
                                                    c_i64, c_p, c_p, c_i64, c_p]),
    "gl_ppr_update_normalized_f64": (ctypes.c_int, [c_p, c_p, c_p, ctypes.c_double, c_i64,
                                                    c_i64, c_p, c_p, c_i64, c_p]),
    "gl_ppr_update_normalized_f32_noalloc": (ctypes.c_int, [c_p, c_p, c_p, c_f32, c_i64, c_i64,
                                                            c_p, c_p]),
    "gl_ppr_missing_mass_f32": (ctypes.c_int, [c_p, c_p, c_i64, c_i64, c_p, c_i64, c_p]),
    "gl_ppr_missing_mass_f64": (ctypes.c_int, [c_p, c_p, c_i64, c_i64, c_p, c_i64, c_p]),
    "gl_ppr_update_f32": (ctypes.c_int, [c_p, c_p, c_p, c_p, c_p, c_f32, c_i64, c_i64, c_p,
                                         c_p, c_i64, c_p]),
    "gl_ppr_update_f64": (ctypes.c_int, [c_p, c_p, c_p, c_p, c_p, ctypes.c_double, c_i64,
                                         c_i64, c_p, c_p, c_i64, c_p]),
    **{f"gl_pbr_spmm_host_{it}_{ft}": (ctypes.c_int, [c_i64, c_i64, c_i32, c_i32, c_i64,
                                                      c_p, c_p, c_p, c_p, c_p, c_p, c_p,
                                                      c_p, c_i64, c_i64, c_p, c_i64, c_i32])
       for it in ("i32", "i64") for ft in ("f32", "f64")},
}

_lib = None


def exported_symbols():
    """Names the C ABI must export (every function declared in include/graph_link_hip.h)."""
    return sorted(_PROTOS)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                "g.build()'` (hipcc --offload-arch=gfx950). The package has no CPU fallback.")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().gl_last_error().decode(errors="replace")


def check(rc: int, what: str = "") -> None:
    if rc == GL_OK:
        return
    msg = f"{what}: {last_error()}" if what else last_error()
    if rc == GL_ERR_INVALID:
        raise ValueError(msg)
    if rc == GL_ERR_WORKSPACE:
        raise MemoryError(msg)
    raise NativeError(msg)


def ptr(t) -> int | None:
    """Device (or host) address of a torch tensor / numpy array, None for None."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data


def stream_handle(device=None) -> int:
    import torch
    return torch.cuda.current_stream(device).cuda_stream
