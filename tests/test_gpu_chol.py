"""Sparse direct solve ("cholmod", the reference's solve_cusparse) on the MI355X against the oracle.

The supernodal multifrontal Cholesky (FP64) of assembled level systems: solutions equal the
oracle's dense FP64 solve of the same matrix (lower triangle) to 1e-9 relative on small levels and
satisfy the symmetrically completed lower triangle to 1e-10 on every level; the singular Neumann
Laplacian reports the oracle's singular pivot; "cholmod" and the coarse levels of "mixed" run
through it end to end."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _levels(cuda, n=6000, vs=1 / 64, seed=3, depth=4):
    import ngspsr
    from ngspsr import synthetic
    xyz, nrm = synthetic.ellipsoids(n, k=4, sigma=0.002, kappa=200, seed=seed)
    return ngspsr.reconstruct(torch.from_numpy(xyz).to(cuda), torch.from_numpy(nrm).to(cuda), vs, depth=depth,
                              min_density=2.0, solver="pcg@tol=1e-7@maxiter=20000", operator="runs",
                              return_result=True)


def test_factor_solve_matches_dense_oracle(cuda):
    import scipy.sparse as sp
    from ngspsr import solver
    from oracle import direct
    print("[case] chol vs dense oracle on every level", flush=True)
    res = _levels(cuda)
    rng = np.random.default_rng(0)
    for d, (system, rhs) in sorted(res.reconstructor.systems.items()):
        csr = system.to_csr()
        f = solver.CholeskyFactor(csr)
        info = f.info()
        assert f.singularity == -1, d
        a = csr.to_scipy()
        for b in (rhs.double(), torch.from_numpy(rng.standard_normal(csr.M)).to(cuda)):
            x = f.solve(b)
            torch.cuda.synchronize()
            bn = b.cpu().numpy()
            if csr.M <= 4000:
                x_ref, sing = direct.solve_spd(a, bn)
                assert sing == -1
                err = np.linalg.norm(x.cpu().numpy() - x_ref) / np.linalg.norm(x_ref)
                assert err < 1e-9, (d, err)
            # the factor is of the lower triangle completed symmetrically (what cuSOLVER reads): its
            # FP64 residual there is at rounding level; against the full FP32-assembled matrix, whose
            # two triangles agree only to FP32 rounding, the residual is that rounding (~1e-8)
            low = sp.tril(a, format="csr")
            sym = (low + sp.tril(low, -1).T).tocsr()
            xn = x.cpu().numpy()
            r_sym = np.linalg.norm(sym @ xn - bn) / np.linalg.norm(bn)
            assert r_sym < 1e-10, (d, r_sym, info)
            _, rr, bb = system.residual_f64(b.double(), x)
            assert (rr / bb) ** 0.5 < 1e-6, (d, (rr / bb) ** 0.5, info)


def test_singular_neumann_laplacian(cuda):
    """The 1-D Neumann Laplacian (constant null space): the last pivot of the nested-dissection
    order vanishes; the singular row reported is that pivot's row in the original numbering, as the
    oracle's dense factorization of the same order finds it."""
    import ctypes
    import scipy.sparse as sp
    from ngspsr._lib import call, lib, ptr, stream
    from ngspsr.solver import CSRSystem, CholeskyFactor
    from oracle import direct
    print("[case] chol singular matrix", flush=True)
    n = 500
    T = (2 * sp.eye(n) - sp.eye(n, k=1) - sp.eye(n, k=-1)).tolil()
    T[0, 0] = T[n - 1, n - 1] = 1.0
    T = T.tocsr()
    rowptr = torch.from_numpy(T.indptr.astype(np.int32)).to(cuda)
    cols = torch.from_numpy(T.indices.astype(np.int32)).to(cuda)
    vals = torch.from_numpy(T.data.astype(np.float32)).to(cuda)
    inv = torch.from_numpy((1.0 / T.diagonal()).astype(np.float32)).to(cuda)
    f = CholeskyFactor(CSRSystem(rowptr, cols, vals, inv))
    assert f.singularity >= 0
    perm = np.zeros(n, np.int32)
    call("ngs_chol_permutation", f.handle, perm.ctypes.data_as(ctypes.c_void_p))
    P = T[perm][:, perm]
    _, sing_new = direct.solve_spd(P, np.ones(n))
    assert sing_new >= 0 and f.singularity == int(perm[sing_new])


def test_cholmod_and_mixed_end_to_end(cuda):
    """The "cholmod" reconstruction solves every level directly (FP64 solution, residual below
    1e-10) and matches the FP64-refined PCG solution; "mixed" falls back to the direct solve on the
    coarse levels as the reference does."""
    import ngspsr
    from ngspsr import synthetic
    print("[case] cholmod and mixed end to end", flush=True)
    xyz, nrm = synthetic.torus(8000, seed=2)
    xt, nt = torch.from_numpy(xyz).to(cuda), torch.from_numpy(nrm).to(cuda)
    rc = ngspsr.reconstruct(xt, nt, 1 / 64, depth=4, min_density=4.0, solver="cholmod", return_result=True)
    rr = ngspsr.reconstruct(xt, nt, 1 / 64, depth=4, min_density=4.0, solver="refine@tol=1e-11", return_result=True)
    for d, info in rc.solve_info.items():
        assert info.direct and info.singularity == -1 and info.rel_resid < 1e-10, (d, info)
        x, xr = rc.reconstructor.solutions[d], rr.reconstructor.solutions[d]
        assert x.dtype == torch.float64
        assert float((x - xr).norm() / xr.norm()) < 1e-7, d
    assert rc.triangles.shape[0] > 1000
    rm = ngspsr.reconstruct(xt, nt, 1 / 64, depth=4, min_density=4.0, solver="mixed", return_result=True)
    for d, info in rm.solve_info.items():
        # coarse levels direct; the finest PCG unless its residual exceeds 1e-2 (then direct)
        assert info.direct == (d != 0 or info.fallback), d


def test_factor_reuse_for_adjoint(cuda):
    """The factor is kept on the system: a second solve with another right-hand side (the adjoint
    of the backward pass) does not factor again and is as accurate."""
    from ngspsr import solver
    print("[case] chol factor reuse", flush=True)
    res = _levels(cuda, n=4000, vs=1 / 32, depth=3)
    system, rhs = res.reconstructor.systems[0]
    x1, i1 = solver.solve(system, rhs, "cholmod")
    f = system._chol
    g = torch.randn(system.M, device=cuda, dtype=torch.float64)
    x2, i2 = solver.solve(system, g, "cholmod")
    assert system._chol is f
    assert i1.rel_resid < 1e-10 and i2.rel_resid < 1e-10


def test_large_level_factor(cuda):
    """The finest level of the 100k-point torus (128^3, ~86k rows): factored and solved to 1e-10."""
    import time
    import ngspsr
    from ngspsr import solver, synthetic
    print("[case] chol large level", flush=True)
    xyz, nrm = synthetic.torus(100000, seed=0)
    res = ngspsr.reconstruct(torch.from_numpy(xyz).to(cuda), torch.from_numpy(nrm).to(cuda), 1 / 128, depth=4,
                             min_density=8.0, solver="pcg@tol=1e-6@maxiter=10000", return_result=True)
    system, rhs = res.reconstructor.systems[0]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, info = solver.solve(system, rhs, "cholmod")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st = system._chol.info()
    print(f"[result] chol M={system.M} {st} solve+factor {dt:.3f} s rel_resid {info.rel_resid:.2e}", flush=True)
    assert info.rel_resid < 1e-10 and info.singularity == -1
