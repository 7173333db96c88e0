"""Multigrid solve with a learned basis (Reconstructor(hash_tree, basis, feat), the generic path of
reconstructor.py:79-195) on the MI355X.

- The B-spline written as a one-component additive basis (AdditiveBasis.get_bezier(), unit features)
  rebuilds the specialised B-spline path's level systems, solutions and field to FP32 rounding.
- A 7-component polynomial basis with per-voxel features (initialisation plus noise): every level's
  matrix and right-hand side equal the FP64 oracle's generic multigrid (fed the GPU's coarser
  solutions), and the field at query points equals the oracle's evaluation."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _tree(cuda, n=5000, seed=21, vs=1 / 32, depth=3, levels=False):
    from ngspsr import synthetic
    from ngspsr.hashtree import HashTree
    xyz, nrm = synthetic.ellipsoids(n, k=3, sigma=0.002, kappa=200, seed=seed)
    xt, nt = torch.from_numpy(xyz).to(cuda), torch.from_numpy(nrm).to(cuda)
    t = HashTree(xt, vs, depth)
    lv = None
    if levels:
        lv = torch.from_numpy(np.random.default_rng(seed).integers(0, depth, n).astype(np.int32)).to(cuda)
    t.build_encoder_hierarchy_adaptive(min_density=4.0, xyz_depth=lv)
    t.reflect_decoder_coords()
    w = 1.0 / t.xyz_density
    nd = {}
    for d in range(depth):
        if t.level_point_counts()[d]:
            nd[d] = t.splat_data(xt, None, d, nt, scale=w, levels=t.xyz_depth, level=d) / float(t.strides[d] ** 3)
    return xyz, nrm, xt, t, nd, w


def test_bspline_as_additive_equals_bspline_path(cuda):
    from ngspsr.additive import AdditiveBasis
    from ngspsr.reconstructor import Reconstructor
    print("[case] learned: bspline as additive vs bspline path", flush=True)
    xyz, nrm, xt, t, nd, w = _tree(cuda, levels=True)
    # FP64-accurate level solutions, so that the right-hand sides (which carry the coarser
    # solutions through the cross-level terms) differ only by the FP32 rounding of the assembly
    solver = "refine@tol=1e-11"
    rb = Reconstructor(t)
    rb.operator = "csr"
    rb.sample_weight = w
    rb.solve_multigrid(t.depth - 1, 0, nd, screen_alpha=4.0, screen_xyz=xt, solver=solver)
    ra = Reconstructor(t, AdditiveBasis.get_bezier())
    ra.sample_weight = w
    ra.solve_multigrid(t.depth - 1, 0, nd, screen_alpha=4.0, screen_xyz=xt, solver=solver)
    assert ra.generic and not rb.generic
    for d in range(t.depth):
        (sb, bb), (sa, ba) = rb.systems[d], ra.systems[d]
        A, B = sb.to_scipy(), sa.to_scipy()
        scale = np.abs(A.data).max()
        assert np.abs((A - B).data).max(initial=0.0) <= 2e-6 * scale, d
        bb_, ba_ = bb.cpu().numpy(), ba.cpu().numpy()
        # the cross-level products cancel against the normal term: FP32 rounding of entries of
        # size sum|a||x| shows at ~1e-5 of the largest entry
        np.testing.assert_allclose(ba_, bb_, rtol=0, atol=3e-5 * np.abs(bb_).max())
        # each FP64 solution satisfies the other path's system to the size of the systems'
        # difference; the solutions themselves differ by the condition number times that
        xa, xb = ra.solutions[d].double(), rb.solutions[d].double()
        _, rr, nb = sb.residual_f64(bb.double(), xa)
        assert (rr / nb) ** 0.5 < 2e-5, (d, (rr / nb) ** 0.5)
        assert float((xa - xb).norm() / xb.norm()) < 1e-4, d
    # the fields carry the solutions' difference (below 1e-4 relative, above)
    q = xt[::5].contiguous()
    ca, cb = ra.evaluate_chi(q), rb.evaluate_chi(q)
    assert float((ca - cb).abs().max()) <= 1e-4 * float(cb.abs().max())
    ga, gb = ra.evaluate_raw_chi(q, compute_grad=True), rb.evaluate_raw_chi(q, compute_grad=True)
    assert float((ga - gb).abs().max()) <= 1e-4 * float(gb.abs().max())
    va, fa = ra.extract_multiscale_mesh(n_upsample=2, build_o3d_mesh=False)
    vb, fb = rb.extract_multiscale_mesh(n_upsample=2, build_o3d_mesh=False)
    assert abs(fa.shape[0] - fb.shape[0]) <= 0.01 * fb.shape[0] and fb.shape[0] > 1000


@pytest.mark.parametrize("alpha_kind", ["scalar", "point"])
def test_polynomial_basis_matches_oracle(cuda, alpha_kind):
    from ngspsr.additive import AdditiveBasis
    from ngspsr.reconstructor import Reconstructor
    from oracle import assembly, learned
    from oracle import hierarchy as H
    print(f"[case] learned: poly-6 basis vs oracle alpha={alpha_kind}", flush=True)
    xyz, nrm, xt, t, nd, w = _tree(cuda, n=4000, seed=22, levels=True)
    basis = AdditiveBasis.get_poly_6()
    g = torch.Generator(device="cpu").manual_seed(5)
    feats = {}
    for d in range(t.depth):
        f = torch.zeros((t.get_coords_size(None, d), basis.get_feature_size()))
        basis.initialize_feature_value(f)
        feats[d] = (f + 0.05 * torch.randn(f.shape, generator=g)).to(cuda)
    rng = np.random.default_rng(6)
    alpha = 4.0 if alpha_kind == "scalar" else rng.uniform(1.0, 8.0, xyz.shape[0]).astype(np.float32)
    alpha_g = alpha if alpha_kind == "scalar" else torch.from_numpy(alpha).to(cuda)
    rec = Reconstructor(t, basis, feats)
    rec.solve_multigrid(t.depth - 1, 0, nd, screen_alpha=alpha_g, screen_xyz=xt, solver="pcg@tol=1e-7@maxiter=20000")
    # the oracle on the same tree, normal data, features and integral tables
    gd = t.xyz_depth.cpu().numpy()
    h = H.Hierarchy(t.depth, t.voxel_size)
    for d in range(t.depth):
        h.set_coords(d, H.quantize_coords(xyz[gd <= d], t.voxel_size, h.strides[d]))
        assert np.array_equal(h.coords[d], t.get_coords(None, d).cpu().numpy())
    comps = [("power", p) for p in range(7)]
    ob = learned.AdditiveBasisOracle(
        comps, [basis.ff_integral[m]["f_f"].data.numpy() for m in range(5)],
        [basis.ff_integral[m]["df_df"].data.numpy() for m in range(5)],
        [[v.data.numpy() for v in basis.f_integral[c]["f"]] for c in range(7)],
        [[v.data.numpy() for v in basis.f_integral[c]["df"]] for c in range(7)],
        [[v.data.numpy() for v in basis.f_integral[c]["f_inv"]] for c in range(7)],
        [[v.data.numpy() for v in basis.f_integral[c]["df_inv"]] for c in range(7)])
    fo = {d: feats[d].cpu().numpy().astype(np.float64) for d in range(t.depth)}
    ndo = {d: v.cpu().numpy().astype(np.float64) for d, v in nd.items()}
    gpu_sol = {d: rec.solutions[d].cpu().numpy().astype(np.float64) for d in range(t.depth)}
    order = list(range(t.depth - 1, -1, -1))
    state = {"i": 0}

    def inject(a, b):
        dd = order[state["i"]]
        state["i"] += 1
        return gpu_sol[dd], 0
    ref = learned.solve_multigrid(h, ob, fo, ndo, xyz, np.asarray(alpha, np.float64) if alpha_kind == "point" else alpha,
                                  0.1, inject)
    for d, (system, rhs) in rec.systems.items():
        A = system.to_scipy()
        R = ref.lhs[d]
        scale = np.abs(R.data).max()
        assert np.abs((A - R).data).max(initial=0.0) <= 2e-5 * scale, d
        want = ref.rhs[d]
        np.testing.assert_allclose(rhs.cpu().numpy(), want, rtol=0, atol=2e-5 * np.abs(want).max())
    q = xyz[::9]
    got = rec.evaluate_raw_chi(torch.from_numpy(q).to(cuda)).cpu().numpy()
    want = learned.evaluate(h, ob, fo, gpu_sol, q)
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-5 * np.abs(want).max())
    # the features matter: the initial features give another system
    rec0 = Reconstructor(t, basis)
    rec0.solve_multigrid(t.depth - 1, 0, nd, screen_alpha=alpha_g, screen_xyz=xt, solver="pcg@tol=1e-7@maxiter=20000")
    A0 = rec0.systems[0][0].to_scipy()
    assert np.abs((A0 - rec.systems[0][0].to_scipy()).data).max() > 1e-3 * np.abs(A0.data).max()
