"""The hierarchy API of the learned caller against the oracle (MI355X): dense and subdivided
builds, voxel status, trilinear interpolation (split_data), levels set in the caller's row order
(update_coords), random samples, and the dual mesh over expanded or trimmed leaves."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _scene(n=5000, seed=11):
    from ngspsr import synthetic
    return synthetic.ellipsoids(n, k=3, sigma=0.002, kappa=200, seed=seed)


def _tree(cuda, xyz, vs, depth):
    from ngspsr.hashtree import HashTree
    return HashTree(torch.from_numpy(xyz).to(cuda), vs, depth)


@pytest.mark.parametrize("expand_range,uniform", [(0, True), (2, True), (3, True), (3, False)])
def test_dense_build_matches_oracle(cuda, expand_range, uniform):
    from oracle import hierarchy as H
    print(f"[case] dense build expand_range={expand_range} uniform={uniform}", flush=True)
    xyz, _ = _scene()
    vs, depth = 1 / 64, 4
    t = _tree(cuda, xyz, vs, depth)
    t.build_encoder_hierarchy_dense(expand_range=expand_range, uniform_density=uniform)
    h = H.Hierarchy(depth, vs)
    h.build_dense(xyz, expand_range)
    for d in range(depth):
        np.testing.assert_array_equal(t.get_coords(t.ENCODER, d).cpu().numpy(), h.coords[d])
    assert int(t.xyz_depth.abs().sum()) == 0
    rho = t.xyz_density.cpu().numpy()
    if uniform:
        assert np.all(rho == 1.0)
    else:
        np.testing.assert_allclose(rho, H.compute_density(xyz, vs), rtol=2e-5, atol=1e-6)
    # the decoder branches are separate until reflected (hashtree.py:57-62)
    assert t.get_coords(t.DECODER, 0) is None
    t.reflect_decoder_coords()
    assert torch.equal(t.get_coords(t.DECODER_TMP, 0), t.get_coords(t.ENCODER, 0))


def _count_policy(k):
    """Refine the points of coarser cells holding more than k points (torch or numpy)."""
    def pol(pts, inv):
        lib = torch if isinstance(inv, torch.Tensor) else np
        cnt = lib.bincount(inv)
        return cnt[inv] > k
    return pol


@pytest.mark.parametrize("expand", [False, True])
def test_subdivide_build_matches_oracle(cuda, expand):
    from oracle import hierarchy as H
    print(f"[case] subdivide build expand={expand}", flush=True)
    xyz, _ = _scene(6000, seed=12)
    vs, depth = 1 / 64, 4
    t = _tree(cuda, xyz, vs, depth)
    lv = t.build_hierarchy_subdivide(_count_policy(12), expand=expand, limit_adaptive_depth=4)
    h = H.Hierarchy(depth, vs)
    lv_ref = h.build_subdivide(xyz, _count_policy(12), expand=expand, limit_adaptive_depth=4)
    np.testing.assert_array_equal(lv.cpu().numpy(), lv_ref)
    assert len(set(lv_ref.tolist())) > 1  # the policy splits the points over several levels
    for d in range(depth):
        np.testing.assert_array_equal(t.get_coords(t.ENCODER, d).cpu().numpy(), h.coords[d])
    np.testing.assert_allclose(t.xyz_density.cpu().numpy(), H.compute_density(xyz, vs), rtol=2e-5, atol=1e-6)


def test_subdivide_with_normal_variation_policy(cuda):
    """The learned caller's policy (normal variation per coarser cell, full_net.py:22-36) runs on
    device tensors and its keyword data follows the kept points."""
    print("[case] subdivide normal-variation policy", flush=True)
    xyz, nrm = _scene(6000, seed=13)
    t = _tree(cuda, xyz, 1 / 64, 4)
    calls = []

    def policy(pts, inv, normal):
        calls.append(pts.shape[0])
        assert normal.shape[0] == pts.shape[0] and inv.shape[0] == pts.shape[0]
        n = normal.abs()
        k = int(inv.max()) + 1
        cnt = torch.bincount(inv, minlength=k).clamp(min=1).double()
        var = 0
        for a in range(3):
            m = torch.zeros(k, dtype=torch.float64, device=inv.device).index_add_(0, inv, n[:, a].double()) / cnt
            m2 = torch.zeros(k, dtype=torch.float64, device=inv.device).index_add_(0, inv, n[:, a].double() ** 2) / cnt
            var = var + (m2 - m * m).clamp(min=0).sqrt()
        return (var > 0.05)[inv]
    lv = t.build_hierarchy_subdivide(policy, expand=True, normal=torch.from_numpy(nrm).to(cuda))
    assert len(calls) == 3 and calls[0] == 6000 and calls[1] <= calls[0] and calls[2] <= calls[1]
    assert int(lv.min()) >= 0 and int(lv.max()) <= 3


def test_voxel_status_and_split_data_match_oracle(cuda):
    from oracle import hierarchy as H
    print("[case] voxel status and split_data", flush=True)
    xyz, _ = _scene(5000, seed=14)
    vs, depth = 1 / 64, 4
    t = _tree(cuda, xyz, vs, depth)
    rng = np.random.default_rng(0)
    rho = rng.uniform(0.05, 20.0, xyz.shape[0]).astype(np.float32)
    h = H.Hierarchy(depth, vs)
    lv_ref = h.build_adaptive(xyz, rho, 4.0)
    t.build_encoder_hierarchy_adaptive(xyz_depth=torch.from_numpy(lv_ref).to(cuda))
    for d in range(depth):
        np.testing.assert_array_equal(t.get_coords(t.ENCODER, d).cpu().numpy(), h.coords[d])
        s = 2 ** d
        probe = np.concatenate([h.coords[d], rng.integers(0, 70, size=(3000, 3)) * s,
                                rng.integers(0, 140, size=(500, 3))], 0).astype(np.int32)
        st = t.evaluate_voxel_status(t.ENCODER, torch.from_numpy(probe).to(cuda), d).cpu().numpy()
        np.testing.assert_array_equal(st, h.voxel_status(probe, d))
        assert (d == 0) or np.any(st == 2)
        data = rng.standard_normal((h.coords[d].shape[0], 5)).astype(np.float32)
        out = t.split_data(torch.from_numpy(xyz).to(cuda), t.ENCODER, d, torch.from_numpy(data).to(cuda))
        np.testing.assert_allclose(out.cpu().numpy(), h.split_data(xyz, d, data), rtol=1e-5, atol=1e-5)


def test_update_coords_row_order(cuda):
    """A level set in the caller's order keeps that order through the API: coordinates, hash
    values, neighbour ids, splats, and the solution of a reconstruction, which equals the sorted
    tree's solution permuted."""
    import ngspsr
    from ngspsr.hashtree import HashTree
    from ngspsr.reconstructor import Reconstructor
    print("[case] update_coords row order", flush=True)
    xyz, nrm = _scene(5000, seed=15)
    vs, depth = 1 / 32, 3
    xt, nt = torch.from_numpy(xyz).to(cuda), torch.from_numpy(nrm).to(cuda)
    ref = ngspsr.reconstruct(xt, nt, vs, depth=depth, min_density=4.0, solver="pcg@tol=1e-7@maxiter=20000",
                             return_result=True)
    t = HashTree(xt, vs, depth)
    t.build_encoder_hierarchy_adaptive(min_density=4.0)
    t.reflect_decoder_coords()
    g = torch.Generator(device="cpu").manual_seed(0)
    perms = {}
    for d in range(depth):
        c = t.get_coords(t.DECODER, d)
        p = torch.randperm(c.shape[0], generator=g).to(cuda)
        perms[d] = p
        t.update_coords(t.DECODER, d, c[p])
        assert torch.equal(t.get_coords(t.DECODER, d), c[p])
        assert torch.equal(t.hash_table(d).query(c[p]), torch.arange(c.shape[0], device=cuda))
    # neighbours: ids are rows of the caller's order
    src, tgt, rel, _ = t.get_neighbours(0, 0, 5)
    s_ref, t_ref, r_ref, _ = ref.tree.get_neighbours(0, 0, 5)
    got = {(int(a), int(b)) for a, b in zip(perms[0][src].tolist(), perms[0][tgt].tolist())}
    want = {(int(a), int(b)) for a, b in zip(s_ref.tolist(), t_ref.tolist())}
    assert got == want
    # reconstruction on the permuted tree: the same per-level solutions, permuted
    w = 1.0 / t.xyz_density
    nd = {}
    counts = t.level_point_counts()
    for d in range(depth):
        if counts[d]:
            nd[d] = t.splat_data(xt, None, d, nt, scale=w, levels=t.xyz_depth, level=d) / float(t.strides[d] ** 3)
            nd_ref = ref.tree.splat_data(xt, None, d, nt, scale=w, levels=t.xyz_depth, level=d) / float(t.strides[d] ** 3)
            assert torch.equal(nd[d], nd_ref[perms[d]])
    rec = Reconstructor(t)
    rec.sample_weight = w
    rec.solve_multigrid(depth - 1, 0, nd, screen_alpha=4.0, screen_xyz=xt, solver="pcg@tol=1e-7@maxiter=20000")
    for d in range(depth):
        assert torch.equal(rec.solutions[d], ref.reconstructor.solutions[d][perms[d]]), d
    v, f = rec.extract_multiscale_mesh(n_upsample=2, build_o3d_mesh=False)
    assert torch.equal(v, ref.vertices) and torch.equal(f, ref.triangles)


def test_random_samples_inside_voxels(cuda):
    print("[case] random samples", flush=True)
    xyz, _ = _scene(3000, seed=16)
    vs = 1 / 32
    t = _tree(cuda, xyz, vs, 3)
    t.build_encoder_hierarchy_adaptive(min_density=4.0)
    t.reflect_decoder_coords()
    for d, expand in ((0, 0), (1, 3)):
        q = t.get_random_samples(7, d, expand=expand)
        base = t._coords_ext(d, expand).float()
        assert q.shape == (base.shape[0] * 7, 3)
        loc = q.view(-1, 7, 3) / vs - base[:, None, :]
        s = t.strides[d]
        assert bool((loc >= -1e-3).all()) and bool((loc <= s + 1e-3).all())


@pytest.mark.parametrize("expand,trim", [(3, False), (0, True)])
def test_mesh_expand_trim_matches_oracle(cuda, expand, trim):
    """Leaves, dual cubes and the triangles over expanded or trimmed leaves equal the oracle's
    given the same field samples."""
    import ngspsr
    from oracle import hierarchy as H
    from oracle import mesh as OM
    print(f"[case] mesh expand={expand} trim={trim}", flush=True)
    xyz, nrm = _scene(5000, seed=17)
    vs, depth = 1 / 32, 3
    # points over all levels (imposed: the adaptive rule puts this cloud on one level)
    lv = torch.from_numpy(np.random.default_rng(3).integers(0, depth, xyz.shape[0]).astype(np.int32)).to(cuda)
    res = ngspsr.reconstruct(torch.from_numpy(xyz).to(cuda), torch.from_numpy(nrm).to(cuda), vs, depth=depth,
                             min_density=4.0, solver="pcg@tol=1e-7@maxiter=20000", xyz_depth=lv, return_result=True)
    gd = res.tree.xyz_depth.cpu().numpy()
    assert len(set(gd.tolist())) == depth
    h = H.Hierarchy(depth, vs)
    for d in range(depth):
        h.set_coords(d, H.quantize_coords(xyz[gd <= d], vs, h.strides[d]))
    v, f, info = res.reconstructor.extract_multiscale_mesh(n_upsample=2, expand=expand, trim=trim, return_info=True)
    leaves_ref, md = OM.conformal_leaves(h, expand=expand, trim=trim)
    assert sorted(info["leaves"]) == list(range(md))
    for d in range(md):
        got = info["leaves"][d].cpu().numpy()
        ref = leaves_ref[d]
        if trim:
            np.testing.assert_array_equal(got, ref)
        else:  # conformal children of sorted parents, in the oracle's order
            np.testing.assert_array_equal(got, ref)
    vals = info["values"].cpu().numpy()
    v_ref, f_ref, ref = OM.extract_multiscale_mesh(h, None, None, None, n_upsample=2, chi_fn=lambda p: vals,
                                                   expand=expand, trim=trim)
    np.testing.assert_array_equal(info["pos"].cpu().numpy(), ref["corners"])
    np.testing.assert_array_equal(info["centres"].cpu().numpy(), ref["centres"])
    np.testing.assert_array_equal(f.cpu().numpy(), f_ref)
    np.testing.assert_array_equal(v.cpu().numpy(), v_ref)
    assert f.shape[0] > 100


@pytest.mark.parametrize("alpha_kind", ["scalar", "point"])
def test_per_point_screen_delta_matches_oracle(cuda, alpha_kind):
    """Per-point screening targets (reconstructor.py:143-147; the learned caller passes
    delta * 2 w_p, full_net.py:125): every level's RHS equals the FP64 oracle's, given the same
    coarser solutions."""
    from ngspsr.hashtree import HashTree
    from ngspsr.reconstructor import Reconstructor
    from oracle import assembly, hierarchy as H
    print(f"[case] per-point screen_delta alpha={alpha_kind}", flush=True)
    xyz, nrm = _scene(5000, seed=18)
    vs, depth = 1 / 32, 3
    xt, nt = torch.from_numpy(xyz).to(cuda), torch.from_numpy(nrm).to(cuda)
    t = HashTree(xt, vs, depth)
    t.build_encoder_hierarchy_adaptive(min_density=4.0)
    t.reflect_decoder_coords()
    w = 1.0 / t.xyz_density
    gd = t.xyz_depth.cpu().numpy()
    nd = {}
    for d in range(depth):
        if t.level_point_counts()[d]:
            nd[d] = t.splat_data(xt, None, d, nt, scale=w, levels=t.xyz_depth, level=d) / float(t.strides[d] ** 3)
    rng = np.random.default_rng(5)
    delta = (0.1 * 2 * rng.uniform(0.2, 1.5, xyz.shape[0])).astype(np.float32)
    alpha = 4.0 if alpha_kind == "scalar" else rng.uniform(1.0, 8.0, xyz.shape[0]).astype(np.float32)
    alpha_g = alpha if alpha_kind == "scalar" else torch.from_numpy(alpha).to(cuda)
    rec = Reconstructor(t)
    rec.solve_multigrid(depth - 1, 0, nd, screen_alpha=alpha_g, screen_xyz=xt, screen_delta=torch.from_numpy(delta).to(cuda),
                        solver="pcg@tol=1e-7@maxiter=20000")
    h = H.Hierarchy(depth, vs)
    for d in range(depth):
        h.set_coords(d, H.quantize_coords(xyz[gd <= d], vs, h.strides[d]))
    nd_ref = assembly.splat_normals(h, xyz, nrm, gd, 1.0 / H.compute_density(xyz, vs))
    gpu_sol = {d: rec.solutions[d].cpu().numpy().astype(np.float64) for d in range(depth)}
    order = list(range(depth - 1, -1, -1))
    state = {"i": 0}

    def inject(a, b):
        d = order[state["i"]]
        state["i"] += 1
        return gpu_sol[d], 0
    ref = assembly.solve_multigrid(h, nd_ref, xyz, np.asarray(alpha, np.float64) if alpha_kind == "point" else alpha,
                                   delta.astype(np.float64), inject)
    for d, (system, rhs) in rec.systems.items():
        want = ref.rhs[d]
        got = rhs.cpu().numpy()
        np.testing.assert_allclose(got, want, rtol=0, atol=2e-5 * np.abs(want).max())
        # the targets matter: a uniform target of the mean moves the screening part well past the tolerance
        rhs_s = ref.rhs_parts[d][1]
        pid, vid, val, _ = h.point_basis(xyz, d)
        sf = 0.25 ** d
        a = np.asarray(alpha, np.float64)
        mean_s = sf * np.bincount(vid, weights=val * float(delta.mean()) * (a[pid] if a.ndim else a), minlength=got.shape[0])
        assert np.abs(mean_s - rhs_s).max() > 1e-3 * np.abs(want).max()
