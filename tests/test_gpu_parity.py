"""GPU parity of every stage of the reconstruct path against the CPU oracle: small scenes in every
level operator form, and the 100k-point torus at 128^3 in full."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _scene(kind, n, seed):
    from ngspsr import synthetic
    if kind == "torus":
        return synthetic.torus(n, seed=seed)
    if kind == "ellipsoids":
        return synthetic.ellipsoids(n, k=4, sigma=0.001, kappa=200, seed=seed)
    return synthetic.sphere(n, seed=seed)


# (scene, points, voxel size, depth, min_density, level operator form); "auto" resolves to the run
# form. The last case is the 100k-point torus at 128^3 with min_density 8 in full.
CASES = [("torus", 4000, 1 / 32, 3, 8.0, "csr"), ("torus", 4000, 1 / 32, 3, 8.0, "stencil"),
         ("torus", 4000, 1 / 32, 3, 8.0, "runs"), ("ellipsoids", 6000, 1 / 64, 4, 8.0, "stencil"),
         ("ellipsoids", 6000, 1 / 64, 4, 8.0, "auto"), ("sphere", 3000, 1 / 32, 3, 32.0, "auto"),
         ("torus", 100000, 1 / 128, 4, 8.0, "auto")]


@pytest.fixture(scope="module", params=CASES, ids=lambda c: f"{c[0]}-{c[1]}-d{c[3]}-{c[5]}")
def pair(request, cuda):
    """GPU reconstruction and oracle on the same input; the oracle's multigrid is fed the GPU
    solutions of the coarser levels so that per-level systems are compared on equal inputs."""
    import ngspsr
    from oracle import assembly, hierarchy, mesh
    kind, n, vs, depth, mind, form = request.param
    print(f"[case] {kind} n={n} vs={vs} depth={depth} min_density={mind} operator={form}", flush=True)
    xyz, nrm = _scene(kind, n, seed=7)
    res = ngspsr.reconstruct(torch.from_numpy(xyz).to(cuda), torch.from_numpy(nrm).to(cuda), vs, depth=depth,
                             min_density=mind, screen_alpha=4.0, solver="pcg@tol=1e-7@maxiter=20000",
                             operator=form, return_result=True)
    torch.cuda.synchronize()
    gpu_depth = res.tree.xyz_depth.cpu().numpy()
    h = hierarchy.Hierarchy(depth, vs)
    for d in range(depth):
        h.set_coords(d, hierarchy.quantize_coords(xyz[gpu_depth <= d], vs, h.strides[d]))
    rho = hierarchy.compute_density(xyz, vs)
    w64 = 1.0 / rho
    nd = assembly.splat_normals(h, xyz, nrm, gpu_depth, w64)
    gpu_sol = {d: res.reconstructor.solutions[d].cpu().numpy().astype(np.float64) for d in range(depth)}
    order = list(range(depth - 1, -1, -1))
    state = {"i": 0}

    def inject(a, b):
        d = order[state["i"]]
        state["i"] += 1
        return gpu_sol[d], 0

    ref = assembly.solve_multigrid(h, nd, xyz, 4.0, 0.1, inject)
    return dict(xyz=xyz, nrm=nrm, vs=vs, depth=depth, min_density=mind, res=res, h=h, rho=rho, nd=nd, ref=ref,
                gpu_depth=gpu_depth, mesh=mesh)


def test_density_and_levels(pair):
    from oracle import hierarchy
    rho_gpu = pair["res"].tree.xyz_density.cpu().numpy()
    np.testing.assert_allclose(rho_gpu, pair["rho"], rtol=2e-5, atol=1e-6)
    lv = hierarchy.level_of_density(pair["rho"], pair["min_density"], 4.0, pair["depth"])
    # levels are integer functions of an FP32 density; only points on a level boundary may differ
    t = np.log(pair["rho"] / pair["min_density"]) / np.log(4.0)
    boundary = np.abs(t - np.round(t)) < 1e-5
    assert np.all((lv == pair["gpu_depth"]) | boundary)


def test_voxel_sets_exact(pair):
    for d in range(pair["depth"]):
        got = pair["res"].tree.get_coords(None, d).cpu().numpy()
        assert np.array_equal(got, pair["h"].coords[d]), f"level {d}"


def test_hash_logical_map(pair):
    from oracle import fnv
    tree = pair["res"].tree
    for d in range(pair["depth"]):
        coords = tree.get_coords(None, d)
        if coords.shape[0] == 0:
            continue
        t = tree.hash_table(d)
        keys = t._sphash(coords).cpu().numpy().astype(np.uint64)
        assert np.array_equal(keys, fnv.fnv_fold60(coords.cpu().numpy()))
        got = t.query(coords).cpu().numpy()
        assert np.array_equal(got, np.arange(coords.shape[0]))
        # absent coordinates miss
        shifted = coords + torch.tensor([[1, 0, 0]], dtype=coords.dtype, device=coords.device) * 10 ** 6
        assert bool((t.query(shifted) == -1).all())
        # the bitmap rank agrees with the cuckoo map
        assert np.array_equal(tree.grids[d].query(coords).cpu().numpy(), got)


@pytest.mark.parametrize("rng", [3, 5])
def test_neighbour_coo_sets(pair, rng):
    h = pair["h"]
    tree = pair["res"].tree
    for d in range(pair["depth"]):
        for dd in range(pair["depth"]):
            if h.num_voxels(d) == 0 or h.num_voxels(dd) == 0:
                continue
            s_ref, t_ref, r_ref, _ = h.self_neighbours(d, dd, rng)
            s, t, r, _ = tree.get_neighbours(d, dd, rng)
            a = np.stack([s_ref, t_ref], 1)
            b = np.stack([s.cpu().numpy(), t.cpu().numpy()], 1)
            ka = np.lexsort(a.T[::-1])
            kb = np.lexsort(b.T[::-1])
            assert np.array_equal(a[ka], b[kb]), (d, dd)
            np.testing.assert_allclose(r.cpu().numpy()[kb], r_ref[ka], atol=1e-5)


def test_splat(pair):
    res = pair["res"]
    tree = res.tree
    w = 1.0 / tree.xyz_density
    for d, ref in pair["nd"].items():
        got = tree.splat_data(torch.from_numpy(pair["xyz"]).cuda(), None, d, torch.from_numpy(pair["nrm"]).cuda(),
                              scale=w, levels=tree.xyz_depth, level=d) / float(tree.strides[d] ** 3)
        np.testing.assert_allclose(got.cpu().numpy(), ref, rtol=1e-4, atol=1e-6 * np.abs(ref).max())


def test_lhs_matrix(pair):
    for d, (system, rhs) in pair["res"].reconstructor.systems.items():
        a = system.to_scipy()
        ref = pair["ref"].lhs[d]
        diff = (a - ref).tocoo()
        scale = np.abs(ref.data).max()
        assert np.abs(diff.data).max() <= 2e-5 * scale if diff.nnz else True, d
        # structure: same stored pattern
        assert (abs(a) > 0).nnz == (abs(ref) > 0).nnz or np.abs(diff.data).max() < 1e-6 * scale


def test_rhs(pair):
    for d, (system, rhs) in pair["res"].reconstructor.systems.items():
        ref = pair["ref"].rhs[d]
        got = rhs.cpu().numpy()
        np.testing.assert_allclose(got, ref, rtol=0, atol=2e-5 * np.abs(ref).max())


def test_solution_residual(pair):
    """The GPU solution of every level satisfies the FP64 oracle system to the solver tolerance."""
    from oracle import pcg
    for d, (system, rhs) in pair["res"].reconstructor.systems.items():
        x = pair["res"].reconstructor.solutions[d].cpu().numpy()
        rel = pcg.relative_residual(pair["ref"].lhs[d], x, pair["ref"].rhs[d])
        assert rel < 5e-6, (d, rel)
        info = pair["res"].reconstructor.solve_info[d]
        assert info.rel_resid <= 1e-7 * 1.0001


def test_mesh_topology_given_same_field(pair):
    """Dual-cube assembly and triangles identical to the oracle's when both use the GPU field
    samples. The GPU keeps only the complete cubes whose corner values change sign (the others
    yield no triangles): the oracle's complete cubes filtered the same way."""
    res = pair["res"]
    v, f, info = res.reconstructor.extract_multiscale_mesh(n_upsample=2, return_info=True)
    vals = info["values"].cpu().numpy()
    om = pair["mesh"]
    v_ref, f_ref, ref = om.extract_multiscale_mesh(pair["h"], None, None, None, n_upsample=2, chi_fn=lambda p: vals)
    np.testing.assert_array_equal(info["pos"].cpu().numpy(), ref["corners"])
    np.testing.assert_array_equal(info["centres"].cpu().numpy(), ref["centres"])
    ref_cubes = ref["acc"][:, :, 0]
    neg = vals[ref_cubes] < 0
    active = neg.any(axis=1) & ~neg.all(axis=1)
    assert 0 < int(active.sum()) < ref_cubes.shape[0]
    np.testing.assert_array_equal(info["cubes"].cpu().numpy(), ref_cubes[active])
    np.testing.assert_array_equal(f.cpu().numpy(), f_ref)
    np.testing.assert_array_equal(v.cpu().numpy(), v_ref)


def test_mesh_close_to_surface(pair):
    res = pair["res"]
    v = res.vertices.cpu().numpy()
    assert v.shape[0] > 100
    xyz = pair["xyz"]
    from scipy.spatial import cKDTree
    d, _ = cKDTree(xyz).query(v)
    assert np.median(d) < pair["vs"], np.median(d)


@pytest.mark.parametrize("nu", [1, 2, 3])
def test_dual_leaf_forms_equal_per_cell(pair, nu):
    """The leaf forms of dual-centre marking and stride-1 dual-cube assembly give the same bitmap,
    counts and complete cubes as the per-sub-cell forms."""
    from ngspsr._lib import call, ptr, stream
    from ngspsr.grid import SparseGrid
    res = pair["res"]
    _, _, info = res.reconstructor.extract_multiscale_mesh(n_upsample=nu, return_info=True)
    tree = res.tree
    cg_leaf = info["centres"]
    leaves, primal = info["leaves"], info["primal"]
    dev = cg_leaf.device
    lo = [int(cg_leaf[:, a].min()) - 1 for a in range(3)]
    hi = [int(cg_leaf[:, a].max()) + 2 for a in range(3)]
    cg = SparseGrid(lo, hi, 1, dev)
    for d, pr in primal.items():
        if pr.shape[0]:
            call("ngs_mark_dual_centres", cg.struct(), ptr(pr), pr.shape[0], tree.strides[d], stream())
    DC = cg.finalize()
    assert torch.equal(cg.coords, cg_leaf)
    acc = torch.zeros((DC, 8), dtype=torch.int32, device=dev)
    cnt = torch.zeros(DC, dtype=torch.int32, device=dev)
    base = 0
    for d, pr in primal.items():
        if pr.shape[0]:
            call("ngs_dual_accumulate", cg.struct(), ptr(pr), pr.shape[0], tree.strides[d], base, None, ptr(acc),
                 ptr(cnt), stream())
        base += pr.shape[0]
    # the mesh's count words also hold the negative-cell counts in their high half
    assert torch.equal(cnt, info["count"] & 0xFFFF)
    full_ = cnt == 8
    ids = info["acc"][full_].long()  # every slot of a complete centre is written
    assert int(ids.min()) >= 0 and int(ids.max()) < info["values"].shape[0]
    neg = (info["values"][ids] < 0).sum(1)
    assert torch.equal(info["count"][full_] >> 16, neg.int())
    full = cnt == 8
    assert torch.equal(acc[full], info["acc"][full])


@pytest.mark.parametrize("nu", [1, 2, 3])
def test_leaf_corner_chi_equals_pointwise(pair, nu):
    """chi at the dual corners, evaluated per leaf (one coefficient window per leaf for nu = 2),
    equals evaluate_chi at the written corner positions bit for bit."""
    print(f"[case] leaf corner chi nu={nu}", flush=True)
    rec = pair["res"].reconstructor
    _, _, info = rec.extract_multiscale_mesh(n_upsample=nu, return_info=True)
    ref = rec.evaluate_chi(info["pos"])
    torch.cuda.synchronize()
    assert info["values"].shape == ref.shape
    assert torch.equal(info["values"], ref), int((info["values"] != ref).sum())


def test_field_queries_match_oracle(pair):
    """evaluate_raw_chi (value, gradient, cover mask, level subset), get_mean_chi and the chunked
    evaluate_chi against the oracle's restatement on the GPU solutions (reconstructor.py:197-255)."""
    print("[case] field queries", flush=True)
    rec = pair["res"].reconstructor
    h, mesh = pair["h"], pair["mesh"]
    sols = {d: rec.solutions[d].cpu().numpy().astype(np.float64) for d in rec.solutions}
    rng = np.random.default_rng(3)
    q = rng.uniform(0.0, 1.0, (6000, 3)).astype(np.float32)  # covered and uncovered queries
    qt = torch.from_numpy(q).cuda()
    v, m = rec.evaluate_raw_chi(qt, compute_mask=True)
    g, m_g = rec.evaluate_raw_chi(qt, compute_mask=True, compute_grad=True)
    g_only = rec.evaluate_raw_chi(qt, compute_grad=True)
    torch.cuda.synchronize()
    ref_v, ref_m = mesh.evaluate_raw_chi(h, sols, q, compute_mask=True)
    ref_g = mesh.evaluate_raw_chi(h, sols, q, compute_grad=True)
    assert m.dtype == torch.bool and np.array_equal(m.cpu().numpy(), ref_m)
    assert 0 < ref_m.sum() < ref_m.size
    assert torch.equal(m, m_g) and torch.equal(g, g_only) and g.shape == (q.shape[0], 3)
    # FP32 sums of at most 27 terms per level against FP64
    np.testing.assert_allclose(v.cpu().numpy(), ref_v, atol=2e-6 * np.abs(ref_v).max())
    np.testing.assert_allclose(g.cpu().numpy(), ref_g, atol=2e-6 * np.abs(ref_g).max())
    assert np.all(ref_v[~ref_m] == 0) and torch.all(v[~m] == 0)
    sub = [d for d in sols if d % 2 == 0]
    v_sub = rec.evaluate_raw_chi(qt, depths=sub).cpu().numpy()
    ref_sub = mesh.evaluate_raw_chi(h, sols, q, depths=sub)
    np.testing.assert_allclose(v_sub, ref_sub, atol=2e-6 * np.abs(ref_sub).max())
    # level set: sample-weighted mean of chi at the input points
    w = rec.sample_weight.cpu().numpy().astype(np.float64)
    ref_mean = mesh.mean_chi(h, sols, pair["xyz"], w)
    chi_pts = mesh.evaluate_raw_chi(h, sols, pair["xyz"])
    mean = rec.get_mean_chi()
    assert abs(mean - ref_mean) <= 2e-6 * np.abs(chi_pts).max()
    # chunked evaluation is pointwise: equal bit for bit to one launch
    full = v - mean
    c, cm = rec.evaluate_chi(qt, compute_mask=True, max_points=777)
    assert torch.equal(c, full) and torch.equal(cm, m)
    assert torch.equal(rec.evaluate_chi(qt, max_points=-1), full)
    assert rec.evaluate_chi(qt[:0], max_points=100).shape == (0,)


def test_points_chi_cache_equals_evaluation(pair):
    """The raw chi at the input points accumulated during the solve (read through the point bins)
    equals evaluate_raw_chi at the same points bit for bit; the level set follows from it."""
    print("[case] points chi cache", flush=True)
    rec = pair["res"].reconstructor
    key, cached = rec._points_chi
    assert rec._key_matches(key)
    direct = rec.evaluate_raw_chi(rec.hash_tree.xyz)
    torch.cuda.synchronize()
    assert torch.equal(cached, direct), int((cached != direct).sum())
