"""Run form of the assembled level systems against the CSR form (MI355X).

The run form stores the same entries as the CSR of ngs_lhs_assemble, grouped by bitmap column:
the entries, diagonal and RHS are bit-identical; products agree to FP32 reassociation (FP64
reassociation for the residual); solves agree to the solver tolerance; the batched pool runs run
systems beside CSR and stencil ones, each bit-identical to its single solve."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _setup(cuda, per_point_alpha, seed=3, n=20000, vs=1 / 64):
    import ngspsr
    from ngspsr import synthetic
    from ngspsr.reconstructor import PointBinning
    xyz, nrm = synthetic.ellipsoids(n, k=4, sigma=0.002, seed=seed)
    xyz_t, nrm_t = torch.from_numpy(xyz).to(cuda), torch.from_numpy(nrm).to(cuda)
    res = ngspsr.reconstruct(xyz_t, nrm_t, vs, depth=4, min_density=2.0, solver="pcg@tol=1e-7@maxiter=20000",
                             operator="csr", return_result=True)
    tree, recon = res.tree, res.reconstructor
    g = torch.Generator(device="cpu").manual_seed(seed)
    alpha = (torch.rand(xyz.shape[0], generator=g) * 8.0).to(cuda) if per_point_alpha else 4.0
    bins = {d: PointBinning(tree, xyz_t, d, alpha) for d in range(tree.depth)}
    return tree, recon, bins, xyz_t


@pytest.mark.parametrize("per_point_alpha", [False, True])
def test_runs_equal_csr_entries(cuda, per_point_alpha):
    print(f"[case] runs vs csr entries alpha={'point' if per_point_alpha else 'scalar'}", flush=True)
    from ngspsr._lib import call, ptr, stream
    from ngspsr.solver import RunSystem
    tree, recon, bins, xyz_t = _setup(cuda, per_point_alpha)
    g = torch.Generator(device="cpu").manual_seed(11)
    for d in range(tree.depth):
        M = tree.get_coords_size(None, d)
        if M == 0:
            continue
        coarse = (torch.randn(xyz_t.shape[0], generator=g) * 0.1).to(cuda) if d < tree.depth - 1 else None
        base = torch.randn(M, generator=g, dtype=torch.float64).to(cuda)
        rhs_c, rhs_r = base.clone(), base.clone()
        A = recon.assemble_level(d, bins[d], coarse, 0.25 ** d, 0.1, rhs_c, form="csr")
        R = recon.assemble_level(d, bins[d], coarse, 0.25 ** d, 0.1, rhs_r, form="runs")
        torch.cuda.synchronize()
        assert isinstance(R, RunSystem)
        assert torch.equal(A.inv_diag, R.inv_diag), d
        assert torch.equal(rhs_c, rhs_r), d
        B = R.to_csr()
        assert torch.equal(A.rowptr, B.rowptr) and torch.equal(A.cols, B.cols) and torch.equal(A.vals, B.vals), d
        cnt = torch.empty(M, dtype=torch.int32, device=cuda)
        call("ngs_stencil_row_counts", tree.level(d), ptr(cnt), stream())
        assert R.nnz() == int(cnt.sum()), d
        # run words: every run lies inside its bitmap column and holds exactly the column's voxels
        n = R.counts()
        start = R.runs[:, :M] & ((1 << 28) - 1)
        assert R.runs.shape[1] % 64 == 0 and R.runs.shape[1] - M < 64
        assert int(n.max()) <= 5
        rows = torch.arange(M, device=cuda)
        c = tree.grids[d].coords.long()
        s = tree.strides[d]
        for p in range(25):
            dx, dy = p // 5 - 2, p % 5 - 2
            has = n[p] > 0
            if not bool(has.any()):
                continue
            first = c[start[p][has].long()]
            last = c[(start[p][has] + n[p][has] - 1).long()]
            own = c[rows[has]]
            assert torch.equal((first[:, 0] - own[:, 0]) // s, torch.full_like(own[:, 0], dx)), (d, p)
            assert torch.equal((last[:, 1] - own[:, 1]) // s, torch.full_like(own[:, 1], dy)), (d, p)
            assert bool(((first[:, 2] - own[:, 2]) // s >= -2).all()) and bool(((last[:, 2] - own[:, 2]) // s <= 2).all())


def test_runs_products_and_residual(cuda):
    print("[case] runs vs csr products", flush=True)
    from ngspsr import solver
    tree, recon, bins, xyz_t = _setup(cuda, True, seed=7)
    g = torch.Generator(device="cpu").manual_seed(12)
    for d in range(tree.depth):
        M = tree.get_coords_size(None, d)
        if M == 0:
            continue
        rhs = torch.zeros(M, dtype=torch.float64, device=cuda)
        A = recon.assemble_level(d, bins[d], None, 0.25 ** d, 0.1, rhs.clone(), form="csr")
        R = recon.assemble_level(d, bins[d], None, 0.25 ** d, 0.1, rhs.clone(), form="runs")
        x = torch.randn(M, generator=g).to(cuda)
        yc, yr = A.matvec(x), R.matvec(x)
        absA = solver.CSRSystem(A.rowptr, A.cols, A.vals.abs(), A.inv_diag)
        bound = absA.matvec(x.abs())
        # FP32 reassociation of the same terms: a few ulps of sum_j |A_ij x_j|
        assert bool(((yc - yr).abs() <= 1e-5 * bound + 1e-30).all()), (d, float(((yc - yr).abs() / bound).max()))
        # the run form sums in increasing column order: equal to a sequential FP32 sum in that order
        x64 = x.double()
        b64 = torch.randn(M, generator=g, dtype=torch.float64).to(cuda)
        rc, rrc, bbc = A.residual_f64(b64, x64)
        rr_, rrr, bbr = R.residual_f64(b64, x64)
        # ||b||^2 is summed over rows in each kernel's own row order
        assert abs(bbc - bbr) <= 1e-13 * bbc
        assert bool(((rc - rr_).abs() <= 1e-13 * bound.double() + 1e-300).all()), d
        r0, rr0, bb0 = R.residual_f64(b64, None)
        assert torch.equal(r0, b64) and rr0 == bb0


def test_runs_solve_matches_csr_solve(cuda):
    from ngspsr import solver
    print("[case] runs vs csr solve", flush=True)
    tree, recon, bins, xyz_t = _setup(cuda, False, seed=5)
    for d in range(tree.depth):
        M = tree.get_coords_size(None, d)
        if M == 0:
            continue
        rhs_c = torch.zeros(M, dtype=torch.float64, device=cuda)
        rhs_r = torch.zeros(M, dtype=torch.float64, device=cuda)
        A = recon.assemble_level(d, bins[d], None, 0.25 ** d, 0.1, rhs_c, form="csr")
        R = recon.assemble_level(d, bins[d], None, 0.25 ** d, 0.1, rhs_r, form="runs")
        b = (rhs_c + torch.randn(M, dtype=torch.float64, device=cuda)).float()
        xc, ic = solver.pcg(A, b, tol=1e-7, maxiter=20000)
        xr, ir = solver.pcg(R, b, tol=1e-7, maxiter=20000)
        assert abs(ic.iters - ir.iters) <= max(2, ic.iters // 20), (d, ic.iters, ir.iters)
        assert ir.rel_resid <= 1e-7 * 1.0001
        err = float((xr - xc).norm() / xc.norm())
        assert err < 1e-4, (d, err)
        x64, info = solver.refine(R, b, tol=1e-10, inner_tol=1e-4)
        assert info.true_rel_resid <= 1e-10, (d, info)
        # the FP64 residual of the run form equals that of the CSR form (same FP32 entries)
        _, rr, bb = A.residual_f64(b.double(), x64)
        assert (rr / bb) ** 0.5 < 1e-9, d


def test_pool_mixes_runs_csr_and_stencil(cuda):
    """Batched pool with packed run, slot run, CSR and stencil systems together: each result equals
    its single solve."""
    import threading
    from ngspsr import serving, solver
    print("[case] pool runs + csr + stencil", flush=True)
    tree, recon, bins, xyz_t = _setup(cuda, False, seed=13)
    systems = []
    for d in range(2):
        M = tree.get_coords_size(None, d)
        for form in ("runs", "csr", "stencil"):
            rhs = torch.zeros(M, dtype=torch.float64, device=cuda)
            systems.append((recon.assemble_level(d, bins[d], None, 0.25 ** d, 0.1, rhs, form=form), rhs.float()))
        # the same run-form matrix kept in slot form
        packed = systems[-3][0]
        systems.append((solver.RunSystem(packed.runs, packed.rvals.clone(), packed.inv_diag, packed=False),
                        systems[-3][1]))
    torch.cuda.synchronize()
    assert systems[0][0].is_packed and not systems[3][0].is_packed
    assert serving.launch_flags([s for s, _ in systems]) == 1 | 2 | 8
    assert serving.launch_flags([s for s, _ in systems[:1]]) == 8 | 4
    assert serving.launch_flags([systems[3][0]]) == 2 | 4
    pool = serving.BatchedPCG(cuda, chunk=4)
    outs = [None] * len(systems)
    try:
        def work(i):
            s, b = systems[i]
            outs[i] = pool.solve(s, b, tol=1e-6)
        th = [threading.Thread(target=work, args=(i,)) for i in range(len(systems))]
        [t.start() for t in th]
        [t.join() for t in th]
    finally:
        pool.close()
    for (s, b), (xb, ib) in zip(systems, outs):
        xs, info = solver.pcg(s, b, tol=1e-6)
        torch.cuda.synchronize()
        assert ib.iters == info.iters
        assert torch.equal(xb, xs), float((xb - xs).abs().max())


def test_reference_solver_modes(cuda):
    """"mixed" is the reference's: PCG to 1e-4 (at most 1000 iterations), kept when the absolute
    residual is at most 1e-2, otherwise the direct solve ("cholmod": the sparse Cholesky factor with
    refinement against the stored matrix), whose FP32 rounding "mixed" returns."""
    from ngspsr import solver
    print("[case] mixed / cholmod modes", flush=True)
    tree, recon, bins, xyz_t = _setup(cuda, False, seed=17)
    M = tree.get_coords_size(None, 0)
    A = recon.assemble_level(0, bins[0], None, 1.0, 0.1, torch.zeros(M, dtype=torch.float64, device=cuda), form="runs")
    g = torch.Generator(device="cpu").manual_seed(2)
    b = torch.randn(M, generator=g).to(cuda)
    small = b * (1e-3 / float(b.norm()))  # ||b|| = 1e-3: PCG at 1e-4 leaves ||Ax - b|| ~1e-7
    x, info = solver.solve(A, small, "mixed")
    assert not info.fallback and 0 < info.iters <= 1000
    _, rr, bb = A.residual_f64(small.double(), x.double())
    assert rr ** 0.5 <= 1e-2 and (rr / bb) ** 0.5 <= 1e-4 * 1.01
    large = b * (1e6 / float(b.norm()))  # ||b|| = 1e6: 1e-4 relative is far above 1e-2 absolute
    x, info = solver.solve(A, large, "mixed")
    assert info.fallback
    _, rr, bb = A.residual_f64(large.double(), x.double())
    assert (rr / bb) ** 0.5 < 1e-5  # FP32 rounding of the direct solution
    xc, infoc = solver.solve(A, large, "cholmod")
    assert infoc.direct and infoc.true_rel_resid <= 1e-12
    assert torch.equal(xc.float(), x)  # the same factor (kept on the system), deterministic solves


def test_packed_form_equals_slot_form(cuda):
    """The packed run form (values without zero slots, slices of 64 rows staged through LDS), which
    the run-form systems keep by default, gives the slot form's products, FP64 residuals and PCG
    solves bit for bit, on every level of a scene whose level sizes are not multiples of 64; the
    slot values regenerated from it equal the assembled ones."""
    import ngspsr
    from ngspsr import solver, synthetic
    from ngspsr.solver import RunSystem
    print("[case] packed run form", flush=True)
    xyz, nrm = synthetic.ellipsoids(30000, k=5, sigma=0.002, seed=17)
    res = ngspsr.reconstruct(torch.from_numpy(xyz).to(cuda), torch.from_numpy(nrm).to(cuda), 1 / 128, depth=4,
                             min_density=2.0, solver="pcg@tol=1e-6@maxiter=20000", operator="runs",
                             return_result=True)
    g = torch.Generator(device="cpu").manual_seed(3)
    sizes = []
    for d, (system, rhs) in sorted(res.reconstructor.systems.items()):
        assert isinstance(system, RunSystem) and system.is_packed
        sizes.append(system.M)
        pvals, soff = system._packed
        assert int(soff[-1]) == system.nnz()
        slot = RunSystem(system.runs, system.rvals.clone(), system.inv_diag, packed=False)
        assert not slot.is_packed
        # the regenerated slots hold the assembled values: the CSR of both forms is the same
        a, b = system.to_csr(), slot.to_csr()
        assert torch.equal(a.vals, b.vals) and torch.equal(a.cols, b.cols)
        for _ in range(2):
            x = torch.randn(system.M, generator=g).to(cuda)
            assert torch.equal(system.matvec(x), slot.matvec(x)), d
            assert torch.equal(system.matvec(x), system.matvec_slots(x)), d
        b64 = torch.randn(system.M, generator=g, dtype=torch.float64).to(cuda)
        x64 = torch.randn(system.M, generator=g, dtype=torch.float64).to(cuda)
        r1, rr1, bb1 = system.residual_f64(b64, x64)
        r2, rr2, bb2 = slot.residual_f64(b64, x64)
        assert torch.equal(r1, r2), d
        x1, i1 = solver.pcg(system, rhs.float(), tol=1e-6, maxiter=20000)
        x2, i2 = solver.pcg(slot, rhs.float(), tol=1e-6, maxiter=20000)
        assert i1.iters == i2.iters and torch.equal(x1, x2), d
    assert any(m % 64 for m in sizes), sizes


@pytest.mark.parametrize("per_point_alpha", [False, True])
def test_assembly_into_packed_form(cuda, per_point_alpha):
    """Levels that assemble with one lane per row write the packed form directly: run words, slice
    offsets, packed values, inv_diag and RHS equal the slot-form assembly packed afterwards, bit for
    bit (with and without the coarse field); levels with several lanes per row pack afterwards."""
    from ngspsr import solver
    from ngspsr._lib import lib
    print(f"[case] packed assembly alpha={'point' if per_point_alpha else 'scalar'}", flush=True)
    tree, recon, bins, xyz_t = _setup(cuda, per_point_alpha, seed=19)
    g = torch.Generator(device="cpu").manual_seed(4)
    direct = 0
    for d in range(tree.depth):
        M = tree.get_coords_size(None, d)
        if M == 0:
            continue
        coarse = (torch.randn(xyz_t.shape[0], generator=g) * 0.1).to(cuda) if d < tree.depth - 1 else None
        base = torch.randn(M, generator=g, dtype=torch.float64).to(cuda)
        rhs_p, rhs_s = base.clone(), base.clone()
        A = recon.assemble_level(d, bins[d], coarse, 0.25 ** d, 0.1, rhs_p, form="runs")
        solver.RunSystem.PACKED = False
        try:
            S = recon.assemble_level(d, bins[d], coarse, 0.25 ** d, 0.1, rhs_s, form="runs")
        finally:
            solver.RunSystem.PACKED = True
        torch.cuda.synchronize()
        assert A.is_packed and not S.is_packed
        direct += int(lib().ngs_assembly_lanes(int(bins[d].n), M) == 1)
        Sp = solver.RunSystem(S.runs, S.rvals, S.inv_diag, packed=True)
        assert torch.equal(A.runs[:, :M], S.runs[:, :M]), d
        assert torch.equal(A._packed[1], Sp._packed[1]) and torch.equal(A._packed[0], Sp._packed[0]), d
        assert torch.equal(A.inv_diag, S.inv_diag) and torch.equal(rhs_p, rhs_s), d
    assert direct >= 1
