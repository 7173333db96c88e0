"""The C ABI library loads and exports every entry point declared in include/ngspsr.h (no GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ngspsr.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ngs_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_declared_symbols():
    from ngspsr import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libngspsr.so is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # every declared entry point has a ctypes signature on the host side
    assert sorted(set(names) - set(_lib.exported_symbols())) == []


def test_host_only_entry_points():
    from ngspsr import _lib
    lib = _lib.lib()
    assert lib.ngs_abi_version() == 1
    ts, nb, eb = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
    assert lib.ngs_hash_sizing(86415, 0, ctypes.byref(ts), ctypes.byref(nb), ctypes.byref(eb)) == 0
    assert (ts.value, nb.value, eb.value) == (262144, 512, 136)
    assert lib.ngs_hash_sizing(0, 0, ctypes.byref(ts), ctypes.byref(nb), ctypes.byref(eb)) == 1
    assert b"positive" in lib.ngs_last_error()
    dims = (ctypes.c_int32 * 3)(10, 10, 33)
    assert lib.ngs_grid_words(dims) == (3300 + 31) // 32
    huge = (ctypes.c_int32 * 3)(100000, 100000, 100000)
    assert lib.ngs_grid_words(huge) == 0


def test_product_fails_loudly_without_library(monkeypatch, tmp_path):
    from ngspsr import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_lib, "_LIB", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.lib()


def test_run_form_rejects_rows_beyond_run_words():
    """Run words index rows with 28 bits: the run-form entry points refuse more than 2^28 rows
    before any device work (argument checks only; the pointers are never dereferenced)."""
    from ngspsr import _lib
    lib = _lib.lib()
    fake = ctypes.c_void_p(256)
    M = (1 << 28) + 1
    assert lib.ngs_spmv_runs(fake, fake, fake, fake, M, None) == 1
    assert b"too many rows" in lib.ngs_last_error()
    assert lib.ngs_residual_f64_runs(fake, fake, fake, fake, fake, M, fake, fake, None) == 1
    assert b"too many rows" in lib.ngs_last_error()


def test_zero_size_calls_need_no_buffers():
    """Zero-size calls return before any argument is dereferenced or any kernel is launched, so empty
    tensors (whose device pointers are null) can be passed through (the mesh path of a cloud whose
    finer levels are empty makes such calls)."""
    from ngspsr import _lib
    lib = _lib.lib()
    assert lib.ngs_dual_compact(None, None, 1, 0, None, None, None, None, None) == 0
    assert lib.ngs_primal_mc_count(None, 0, 3, None, None) == 0
    assert lib.ngs_field_support(_lib.Grid(), None, 0, 1.0, None, None) == 0
    assert lib.ngs_ddi_forward(None, None, None, 0, 4, None, None, 5, None, None, None) == 0


def test_reconstruct_rejects_bad_arguments_before_device_work():
    """Argument errors are raised on the host before any device is selected (no GPU needed)."""
    import numpy as np
    import ngspsr
    xyz = np.random.default_rng(0).uniform(0.2, 0.8, (50, 3)).astype(np.float32)
    nrm = np.tile(np.float32([[0.0, 0.0, 1.0]]), (50, 1))
    bad = xyz.copy()
    bad[3, 1] = np.nan
    with pytest.raises(ValueError, match="non-finite"):
        ngspsr.reconstruct(bad, nrm, 1 / 32)
    with pytest.raises(ValueError, match="non-finite"):
        ngspsr.reconstruct(xyz, np.where(nrm == 1.0, np.inf, nrm), 1 / 32)
    with pytest.raises(ValueError, match="voxel_size"):
        ngspsr.reconstruct(xyz, nrm, 0.0)
    with pytest.raises(ValueError, match="depth"):
        ngspsr.reconstruct(xyz, nrm, 1 / 32, depth=0)
    with pytest.raises(NotImplementedError):
        ngspsr.reconstruct(xyz, nrm, 1 / 32, degree=3)


def test_solver_strings():
    """The reference's solver mini-language (solver.py:180-186) and its modes; values are parsed as
    literals, never evaluated."""
    from ngspsr.solver import parse_solver
    assert parse_solver("pcg@tol=1e-6@maxiter=1000") == ("pcg", {"tol": 1e-6, "maxiter": 1000, "res_fix": False,
                                                                 "verbose": False, "chunk": 32})
    name, conf = parse_solver("mixed")
    assert name == "mixed" and (conf["tol"], conf["fast_tol"], conf["fast_maxiter"]) == (1e-2, 1e-4, 1000)
    assert parse_solver("cholmod")[1]["tol"] == 1e-6  # the singular-pivot tolerance solve_cholmod passes
    name, conf = parse_solver("refine@tol=1e-9@inner_tol=1e-3")
    assert name == "refine" and conf["tol"] == 1e-9 and conf["inner_tol"] == 1e-3
    with pytest.raises(NotImplementedError):
        parse_solver("lu")
    with pytest.raises(ValueError):
        parse_solver("pcg@tol=__import__('os')")
