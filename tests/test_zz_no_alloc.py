"""ngs_pcg_solve_runs_noalloc (the PCG entry point that allocates nothing; its workspace is the
library's own device array) against ngs_pcg_solve_runs with a caller workspace (MI355X).

Both run the same kernels in the same order, so the solutions, iteration counts and residuals are
bit-identical. Two calls in a row, the second on a smaller system, show that the shared array's
contents from the previous call do not leak into the next one. Systems beyond the array's row
limit are rejected before any device work."""
import ctypes

import pytest
import torch

pytestmark = pytest.mark.gpu


def _run_system(cuda, n, vs, seed):
    import ngspsr
    from ngspsr import synthetic
    from ngspsr.reconstructor import PointBinning
    from ngspsr.solver import RunSystem
    xyz, nrm = synthetic.ellipsoids(n, k=3, sigma=0.002, seed=seed)
    res = ngspsr.reconstruct(torch.from_numpy(xyz).to(cuda), torch.from_numpy(nrm).to(cuda), vs, depth=3,
                             min_density=2.0, solver="pcg@tol=1e-6@maxiter=20000", operator="runs",
                             return_result=True)
    recon = res.reconstructor
    bins = PointBinning(res.tree, torch.from_numpy(xyz).to(cuda), 0, 4.0)
    rhs = torch.zeros(res.tree.get_coords_size(None, 0), dtype=torch.float64, device=cuda)
    A = recon.assemble_level(0, bins, None, 1.0, 0.1, rhs, form="runs")
    torch.cuda.synchronize()
    assert isinstance(A, RunSystem)
    g = torch.Generator(device="cpu").manual_seed(seed)
    b = torch.randn(A.M, generator=g).to(cuda)
    return A, b


def _solve(A, b, tol, noalloc):
    from ngspsr._lib import lib, ptr, stream
    x = torch.full((A.M,), float("nan"), dtype=torch.float32, device=A.device)
    it, rr = ctypes.c_int32(-1), ctypes.c_double(-1.0)
    if noalloc:
        st = lib().ngs_pcg_solve_runs_noalloc(ptr(A.runs), ptr(A.rvals), ptr(A.inv_diag), ptr(b), ptr(x), A.M,
                                              float(tol), 20000, 32, stream(), ctypes.byref(it), ctypes.byref(rr))
    else:
        ws = torch.empty(lib().ngs_pcg_workspace_bytes(A.M), dtype=torch.uint8, device=A.device)
        st = lib().ngs_pcg_solve_runs(ptr(A.runs), ptr(A.rvals), ptr(A.inv_diag), ptr(b), ptr(x), A.M, float(tol),
                                      20000, 32, ptr(ws), stream(), ctypes.byref(it), ctypes.byref(rr))
    assert st == 0, lib().ngs_last_error().decode()
    torch.cuda.synchronize()
    return x, it.value, rr.value


def test_noalloc_pcg_equals_workspace_pcg(cuda):
    """A large system, then a smaller one: each no-allocation solve equals the workspace solve."""
    cases = [(60000, 1 / 128, 5), (8000, 1 / 64, 6)]
    systems = []
    for n, vs, seed in cases:
        print(f"[case] noalloc pcg setup n={n} vs={vs} seed={seed}", flush=True)
        systems.append(_run_system(cuda, n, vs, seed))
    assert systems[1][0].M < systems[0][0].M
    for (A, b), (n, vs, seed) in zip(systems, cases):
        print(f"[case] noalloc pcg M={A.M}", flush=True)
        x_ref, it_ref, rr_ref = _solve(A, b, 1e-6, noalloc=False)
        x, it, rr = _solve(A, b, 1e-6, noalloc=True)
        assert it == it_ref and it > 0
        assert rr == rr_ref and rr <= 1e-6
        assert torch.equal(x, x_ref)


def test_noalloc_pcg_rejects_oversize(cuda):
    """Systems above the scratch array's row limit fail with a capacity error before any device
    work (the pointers are never dereferenced)."""
    from ngspsr._lib import lib
    fake = ctypes.c_void_p(256)
    it, rr = ctypes.c_int32(), ctypes.c_double()
    st = lib().ngs_pcg_solve_runs_noalloc(fake, fake, fake, fake, fake, 2790144 + 1, 1e-6, 10, 32, None,
                                          ctypes.byref(it), ctypes.byref(rr))
    assert st == 6
    msg = lib().ngs_last_error().decode()
    assert "2790144 rows" in msg and "scratch" in msg
