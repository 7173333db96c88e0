"""ctypes binding of libngspsr.so (the C ABI declared in include/ngspsr.h).

The library is built in-tree (python -m ngspsr._build, or __graft_entry__.build()). Every
product code path goes through it; there is no CPU fallback: loading fails loudly when the
library is absent.
"""
import ctypes
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libngspsr.so")

c_int32, c_int64, c_float, c_double, c_void_p, c_uint64 = (ctypes.c_int32, ctypes.c_int64, ctypes.c_float,
                                                            ctypes.c_double, ctypes.c_void_p, ctypes.c_uint64)


class Grid(ctypes.Structure):
    _fields_ = [("origin", c_int32 * 3), ("dims", c_int32 * 3), ("log2_stride", c_int32), ("pad_", c_int32),
                ("nwords", c_int64), ("bits", c_void_p), ("rank", c_void_p)]


class Level(ctypes.Structure):
    _fields_ = [("grid", Grid), ("coords", c_void_p), ("M", c_int64), ("stride", c_int32), ("pad_", c_int32)]


class PointBins(ctypes.Structure):
    _fields_ = [("cells", Grid), ("cell_start", c_void_p), ("perm", c_void_p), ("b", c_void_p), ("alpha", c_void_p),
                ("alpha_scalar", c_float), ("pad_", c_float), ("n", c_int64)]


class Stencil(ctypes.Structure):
    _fields_ = [("level", Level), ("cells", Grid), ("cell_start", c_void_p), ("point_cell", c_void_p),
                ("basis", c_void_p), ("alpha", c_void_p), ("perm", c_void_p), ("lap", c_void_p), ("u", c_void_p),
                ("alpha_scalar", c_float), ("sf", c_float), ("n", c_int64)]


class HashTable(ctypes.Structure):
    _fields_ = [("table_size", c_int64), ("num_buckets", c_int32), ("evict_bound", c_int32), ("keys", c_void_p),
                ("vals", c_void_p), ("mult", c_uint64 * 3), ("inserted", c_int64), ("attempts", c_int32),
                ("pad_", c_int32)]


P = c_void_p
_SIGNATURES = {
    "ngs_last_error": ([], ctypes.c_char_p),
    "ngs_abi_version": ([], c_int32),
    "ngs_hash_coords": ([P, c_int64, c_int32, P, P], c_int32),
    "ngs_kernel_hash": ([P, c_int64, c_int32, P, c_int32, P, P], c_int32),
    "ngs_hash_sizing": ([c_int64, c_int32, ctypes.POINTER(c_int64), ctypes.POINTER(c_int32),
                         ctypes.POINTER(c_int32)], c_int32),
    "ngs_hash_workspace_bytes": ([c_int64], c_int64),
    "ngs_hash_build": ([ctypes.POINTER(HashTable), P, P, c_int64, c_uint64, P, P], c_int32),
    "ngs_hash_query": ([ctypes.POINTER(HashTable), P, c_int64, P, P], c_int32),
    "ngs_grid_words": ([ctypes.POINTER(c_int32)], c_int64),
    "ngs_scan_workspace_bytes": ([c_int64], c_int64),
    "ngs_grid_clear": ([Grid, P], c_int32),
    "ngs_grid_mark_coords": ([Grid, P, c_int64, P], c_int32),
    "ngs_grid_mark_corners": ([Grid, P, c_int64, c_float, P, c_int32, P], c_int32),
    "ngs_grid_mark_coords_staged": ([Grid, P, c_int64, P, P], c_int32),
    "ngs_grid_mark_corners_staged": ([Grid, P, c_int64, c_float, P, c_int32, P, P], c_int32),
    "ngs_grid_finalize": ([Grid, P, P, P], c_int32),
    "ngs_grid_emit": ([Grid, P, P], c_int32),
    "ngs_grid_query": ([Grid, P, c_int64, P, c_int32, P, P], c_int32),
    "ngs_grid_rank_rows": ([Grid, P, c_int64, P, P], c_int32),
    "ngs_point_bounds": ([P, c_int64, c_float, P, P], c_int32),
    "ngs_point_cells": ([P, c_int64, c_float, c_int32, P, P], c_int32),
    "ngs_voxel_status": ([Grid, Grid, c_int32, P, c_int64, P, P], c_int32),
    "ngs_split_data": ([Grid, P, c_int64, c_float, P, c_int32, P, P], c_int32),
    "ngs_scan_i32": ([P, c_int64, P, P, P], c_int32),
    "ngs_splat": ([Grid, P, c_int64, c_float, P, c_int32, P, P, c_int32, c_double, P, P], c_int32),
    "ngs_splat_gather_workspace_bytes": ([c_int64, c_int64], c_int64),
    "ngs_splat_gather": ([Grid, P, c_int64, P, c_int64, c_float, P, c_int32, P, P, c_int32, c_double, P, P, c_int64, P],
                         c_int32),
    "ngs_eval_field": ([Grid, P, P, c_int64, c_float, c_float, c_int32, P, P, P], c_int32),
    "ngs_eval_field_binned": ([Grid, P, P, P, P, c_int64, c_float, P, P], c_int32),
    "ngs_field_support": ([Grid, P, c_int64, c_float, P, P], c_int32),
    "ngs_grid_mark_dilated": ([Grid, P, c_int64, c_int32, P], c_int32),
    "ngs_test_grid_positions": ([P, c_int64, c_int32, c_int32, c_float, P, P], c_int32),
    "ngs_ddi_forward": ([P, P, P, c_int64, c_int32, P, P, c_int32, P, P, P], c_int32),
    "ngs_ddi_backward": ([P, P, P, c_int64, c_int32, P, P, c_int32, P, P, P, P, P], c_int32),
    "ngs_additive_eval": ([P, P, P, c_int64, P, c_int32, c_float, P, P, P], c_int32),
    "ngs_additive_eval_backward": ([P, P, c_int64, P, c_int32, c_float, P, P, P, P], c_int32),
    "ngs_gather_f32": ([P, P, c_int64, P, P], c_int32),
    "ngs_primal_mc_count": ([P, c_int64, c_int32, P, P], c_int32),
    "ngs_primal_mc_emit": ([P, P, c_int64, c_int32, P, P, P, P, P, P], c_int32),
    "ngs_primal_dedup_workspace_bytes": ([c_int64], c_int64),
    "ngs_primal_dedup": ([P, c_int64, P, P, P, P, P, P, P, P, P, P, P, P, c_int64, P], c_int32),
    "ngs_primal_mc_backward": ([P, P, P, P, c_int64, c_float, P, P], c_int32),
    "ngs_point_levels": ([P, c_int64, c_float, c_float, c_int32, P, P], c_int32),
    "ngs_point_basis": ([P, c_int64, c_float, c_int32, P, P, P], c_int32),
    "ngs_bucket_workspace_bytes": ([c_int64, c_int64], c_int64),
    "ngs_bucket_points": ([P, c_int64, c_int64, P, P, P, c_int64, P], c_int32),
    "ngs_set_tables": ([P, P, c_int32], c_int32),
    "ngs_lhs_row_counts": ([Level, P, P], c_int32),
    "ngs_lhs_assemble": ([Level, PointBins, P, c_float, c_float, P, P, P, P, P, P], c_int32),
    "ngs_assembly_config": ([c_int32], c_int32),
    "ngs_lhs_assemble_runs": ([Level, PointBins, P, c_float, c_float, P, P, P, P, P], c_int32),
    "ngs_assembly_lanes": ([c_int64, c_int64], c_int32),
    "ngs_lhs_assemble_packed": ([Level, PointBins, P, c_float, c_float, P, P, P, P, P, P], c_int32),
    "ngs_pcg_solve_runs": ([P, P, P, P, P, c_int64, c_double, c_int32, c_int32, P, P, ctypes.POINTER(c_int32),
                            ctypes.POINTER(c_double)], c_int32),
    "ngs_pcg_solve_runs_noalloc": ([P, P, P, P, P, c_int64, c_double, c_int32, c_int32, P, ctypes.POINTER(c_int32),
                                    ctypes.POINTER(c_double)], c_int32),
    "ngs_pcg_batch_prepare_runs": ([P, P, P, P, P, c_int64, c_double, c_int32, P, c_int32, c_int32, P, P], c_int32),
    "ngs_spmv_runs": ([P, P, P, P, c_int64, P], c_int32),
    "ngs_residual_f64_runs": ([P, P, P, P, P, c_int64, P, P, P], c_int32),
    "ngs_rhs_cross_stiffness": ([Level, Level, P, P, P], c_int32),
    "ngs_rhs_normals": ([Level, Level, P, P, P], c_int32),
    "ngs_basis_splat": ([Grid, P, P, c_int64, c_float, c_double, P, P], c_int32),
    "ngs_cross_stiffness_transpose": ([Level, Level, P, c_double, P, P], c_int32),
    "ngs_pcg_workspace_bytes": ([c_int64], c_int64),
    "ngs_pcg_config": ([c_int32, c_int32, c_int32, c_int32], c_int32),
    "ngs_pcg_solve": ([P, P, P, P, P, P, c_int64, c_double, c_int32, c_int32, P, P, ctypes.POINTER(c_int32),
                       ctypes.POINTER(c_double)], c_int32),
    "ngs_pcg_batch_entry_bytes": ([], c_int64),
    "ngs_pcg_batch_blocks": ([c_int64], c_int32),
    "ngs_pcg_batch_prepare": ([P, P, P, P, P, P, c_int64, c_double, c_int32, P, c_int32, c_int32, P, P], c_int32),
    "ngs_pcg_batch_prepare_stencil": ([Stencil, P, P, P, c_double, c_int32, P, c_int32, c_int32, P, P], c_int32),
    "ngs_pcg_batch_place": ([P, c_int32], c_int32),
    "ngs_pcg_batch_iterate": ([P, P, c_int32, c_int32, c_int32, c_int32, P], c_int32),
    "ngs_pcg_solve_stencil": ([Stencil, P, P, P, c_double, c_int32, c_int32, P, P, ctypes.POINTER(c_int32),
                               ctypes.POINTER(c_double)], c_int32),
    "ngs_stencil_laplacian": ([c_int32, P, P], c_int32),
    "ngs_bin_points": ([P, c_int64, P, P, P, P, P, P, P], c_int32),
    "ngs_stencil_diag_rhs": ([Stencil, P, c_float, P, P, P], c_int32),
    "ngs_stencil_apply": ([Stencil, P, P, P], c_int32),
    "ngs_runs_slice_counts": ([P, c_int64, P, P], c_int32),
    "ngs_runs_pack": ([P, P, c_int64, P, P, P], c_int32),
    "ngs_spmv_packed": ([P, P, P, P, P, c_int64, P], c_int32),
    "ngs_runs_unpack": ([P, P, c_int64, P, P, P], c_int32),
    "ngs_pcg_solve_packed": ([P, P, P, P, P, P, c_int64, c_double, c_int32, c_int32, P, P, ctypes.POINTER(c_int32),
                              ctypes.POINTER(c_double)], c_int32),
    "ngs_pcg_batch_prepare_packed": ([P, P, P, P, P, P, c_int64, c_double, c_int32, P, c_int32, c_int32, P, P],
                                     c_int32),
    "ngs_residual_f64_packed": ([P, P, P, P, P, P, c_int64, P, P, P], c_int32),
    "ngs_stencil_pass": ([Stencil, P, P, c_int32, P], c_int32),
    "ngs_stencil_row_counts": ([Level, P, P], c_int32),
    "ngs_stencil_residual_f64": ([Stencil, P, P, P, P, P, P, P], c_int32),
    "ngs_pcg_batch_status": ([P, c_int32, P, P, P], c_int32),
    "ngs_spmv": ([P, P, P, P, P, c_int64, P], c_int32),
    "ngs_residual_f64": ([P, P, P, P, P, P, c_int64, P, P, P], c_int32),
    "ngs_conformal_children": ([P, c_int64, c_int32, Grid, c_int32, P, P, P], c_int32),
    "ngs_unrefined_flags": ([P, c_int64, Grid, P, P], c_int32),
    "ngs_compact_rows": ([P, c_int32, P, c_int64, P, P, P, P], c_int32),
    "ngs_expand_primal": ([P, c_int64, c_int32, c_int32, P, P], c_int32),
    "ngs_mark_dual_centres": ([Grid, P, c_int64, c_int32, P], c_int32),
    "ngs_dual_accumulate": ([Grid, P, c_int64, c_int32, c_int32, P, P, P, P], c_int32),
    "ngs_mark_dual_leaves": ([Grid, P, c_int64, c_int32, c_int32, P], c_int32),
    "ngs_mark_dual_leaf_words": ([Grid, Grid, P], c_int32),
    "ngs_dual_gather_cubes": ([P, c_int64, c_int32, P, P, P, P, P, P, P], c_int32),
    "ngs_dual_accumulate_leaves": ([Grid, P, c_int64, c_int32, c_int32, P, P, P, P], c_int32),
    "ngs_dual_compact": ([P, P, c_int32, c_int64, P, P, P, P, P], c_int32),
    "ngs_dual_corner_positions": ([P, c_int64, c_int32, c_float, P, P], c_int32),
    "ngs_leaf_corner_positions": ([P, c_int64, c_int32, c_int32, c_float, P, P], c_int32),
    "ngs_eval_leaf_corners_levels": ([c_int32, P, P, P, c_int64, c_int32, c_int32, c_float, c_float, c_int32, P, P],
                                     c_int32),
    "ngs_eval_leaf_corners": ([Grid, P, P, c_int64, c_int32, c_int32, c_float, c_float, c_int32, P, P], c_int32),
    "ngs_mc_count": ([P, c_int64, P, P, P, P], c_int32),
    "ngs_mc_emit": ([P, c_int64, P, P, P, c_int32, P, P, P, P], c_int32),
    "ngs_mc_remap_faces": ([P, c_int64, P, P], c_int32),
    "ngs_dedup_workspace_bytes": ([c_int64], c_int64),
    "ngs_dedup_vertices": ([P, P, c_int64, c_int32, P, P, P, P, P, P, c_int64, P], c_int32),
    "ngs_f64_to_f32": ([P, P, c_int64, P], c_int32),
    "ngs_chol_analyze": ([P, P, c_int64, P, ctypes.POINTER(c_void_p)], c_int32),
    "ngs_chol_analyze_host": ([P, P, c_int64, ctypes.POINTER(c_void_p)], c_int32),
    "ngs_chol_workspace_bytes": ([c_void_p], c_int64),
    "ngs_chol_info": ([c_void_p, P], c_int32),
    "ngs_chol_permutation": ([c_void_p, P], c_int32),
    "ngs_chol_structure": ([c_void_p, P, P, P, P], c_int32),
    "ngs_chol_factor": ([c_void_p, P, P, P, c_double, P, P, ctypes.POINTER(c_int64)], c_int32),
    "ngs_chol_solve": ([c_void_p, P, P, P, P, P, P], c_int32),
    "ngs_chol_free": ([c_void_p], None),
    "ngs_screen_pairs": ([c_int64, P, P, P, P, P, P, P, c_float, P, P, P, P], c_int32),
    "ngs_weighted_sum": ([P, P, c_int64, P, P], c_int32),
}

STATUS = {0: "ok", 1: "bad argument", 2: "HIP error", 3: "hash bucket overflow", 4: "rehash exhausted",
          5: "not converged", 6: "capacity exceeded"}


class NgsError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


_LIB = None


def lib():
    """The loaded library; raises when it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libngspsr.so not found at {LIB_PATH}; build it with `python -m ngspsr._build` "
                               "(or __graft_entry__.build()). There is no CPU fallback.")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
    return _LIB


def exported_symbols():
    return list(_SIGNATURES)


def call(name, *args):
    st = getattr(lib(), name)(*args)
    if st != 0:
        raise NgsError(st, lib().ngs_last_error().decode())
    return st


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)
_cur_device = getattr(torch._C, "_cuda_getDevice", None)


def stream():
    """The current stream of the current device, as the library's stream argument. Called for every
    launch: the direct lookup costs ~1 us where torch.cuda.current_stream() costs ~8 us."""
    if _raw_stream is not None and _cur_device is not None:
        return ctypes.c_void_p(_raw_stream(_cur_device()))
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class _ThreadBuffers:
    """Per-thread, per-device device buffers reused from call to call in stream order. A buffer
    that this thread last used on another stream is handed over on first use from the new stream:
    the new stream waits for an event recorded on the old one (everything this thread enqueued
    there, which includes every earlier use of the buffer), and the tensor is recorded on the new
    stream, so the caching allocator does not hand its memory out before the new stream's work is
    done when it is replaced."""

    def __init__(self, zero):
        self._tls = threading.local()
        self._zero = zero

    def get(self, key, nbytes, device):
        ents = getattr(self._tls, "ents", None)
        if ents is None:
            ents = self._tls.ents = {}
        cur_ptr = _raw_stream(_cur_device()) if _raw_stream is not None else None
        ent = ents.get(key)
        if ent is not None and (cur_ptr is None or ent[1] != cur_ptr):
            cur = torch.cuda.current_stream()
            if ent[2] != cur:
                ev = torch.cuda.Event()
                ev.record(ent[2])
                cur.wait_event(ev)
                ent[0].record_stream(cur)
            ent[1], ent[2] = cur.cuda_stream, cur
        if ent is None or ent[0].numel() < nbytes:
            # grow geometrically: fewer replacements, so fewer odd-sized blocks left in the cache
            n = max(int(nbytes), 256, 0 if ent is None else int(1.5 * ent[0].numel()))
            buf = torch.zeros(n, dtype=torch.uint8, device=device) if self._zero else \
                torch.empty(n + 256, dtype=torch.uint8, device=device)
            cur = torch.cuda.current_stream()
            ent = ents[key] = [buf, cur.cuda_stream, cur]
        return ent[0]


class Marks:
    """Per-thread, per-device all-zero byte maps for the byte-staged grid marking calls, which leave
    the map all zero again (stream-ordered use, as Scratch)."""
    _bufs = _ThreadBuffers(zero=True)
    CAP = 1 << 31  # bytes; larger boxes mark with atomics

    @classmethod
    def get(cls, nbytes, device):
        return cls._bufs.get(str(device), nbytes, device)


class Scratch:
    """Per-thread, per-device growable scratch buffers for stream-ordered temporary workspaces.

    Buffers are private to the calling thread, so concurrent requests on different streams never
    share one; a thread that moves to another stream hands its buffers over in stream order."""
    _bufs = _ThreadBuffers(zero=False)

    @classmethod
    def get(cls, nbytes, device, slot=0):
        return cls._bufs.get((str(device), slot), nbytes, device)
