"""Primal sparse marching cubes with the MeshSDF backward, get_test_grid and extract_mesh on the GPU
against the oracle (primal_mc.cu, ops.py:9-101, hashtree.py:209-229, reconstructor.py:257-294)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _field(kind, nl, r, seed):
    """nl distinct leaves in a small box (so that many share faces) and their (r+1)^3 samples."""
    from oracle import primal_mc as pm
    rng = np.random.default_rng(seed)
    cells = np.stack(np.meshgrid(*[np.arange(-2, 4)] * 3, indexing="ij"), -1).reshape(-1, 3)
    base = cells[rng.choice(cells.shape[0], nl, replace=False)].astype(np.float32)
    shape = (nl, r + 1, r + 1, r + 1)
    if kind == "sphere":
        p = pm.sample_positions(base, r)
        sdf = (2.2 - np.linalg.norm(p - 0.7, axis=1)).reshape(shape)
    elif kind == "ties":  # exact zeros and values within the 1e-5 snapping band
        sdf = rng.choice([-1.0, -4e-6, 0.0, 3e-6, 1.0, 1.0 + 5e-6], shape)
    else:
        sdf = rng.standard_normal(shape)
    return base, sdf.astype(np.float32)


CASES = [("random", 30, 1), ("random", 40, 3), ("sphere", 60, 4), ("ties", 25, 2), ("random", 10, 7),
         ("random", 0, 3)]


@pytest.mark.parametrize("kind,nl,r", CASES)
def test_marching_cubes_sparse_exact(cuda, kind, nl, r):
    from ngspsr import ops
    from oracle import primal_mc as pm
    print(f"[case] primal mc {kind} leaves={nl} r={r}", flush=True)
    base, sdf = _field(kind, nl, r, seed=nl + r)
    got = ops.marching_cubes_sparse(torch.from_numpy(base).to(cuda), torch.from_numpy(sdf).to(cuda))
    ref = pm.marching_cubes_sparse(base, sdf)
    assert (nl == 0) == (ref[0].shape[0] == 0)
    for name, g, e in zip(("tris", "vert_ids", "src_ids", "src_weight"), got, ref):
        g = g.cpu().numpy()
        assert g.shape == e.shape and g.dtype == e.dtype, name
        assert np.array_equal(g, e), (name, int((g != e).sum()))


@pytest.mark.parametrize("kind,nl,r", CASES)
def test_marching_cubes_op_and_meshsdf_backward(cuda, kind, nl, r):
    from ngspsr import ops
    from oracle import primal_mc as pm
    print(f"[case] primal mc op {kind} leaves={nl} r={r}", flush=True)
    base, sdf = _field(kind, nl, r, seed=nl + r)
    sdf_t = torch.from_numpy(sdf).to(cuda).requires_grad_()
    v, t, n, unq = ops.marching_cubes_op(torch.from_numpy(base).to(cuda), sdf_t, 2.0)
    rv, rt, rn, runq, rsrc, rw = pm.marching_cubes(base, sdf)
    assert np.array_equal(v.detach().cpu().numpy(), rv) and v.shape == rv.shape
    assert t.dtype == torch.int64 and np.array_equal(t.cpu().numpy(), rt)
    assert np.array_equal(unq.cpu().numpy(), runq)
    np.testing.assert_allclose(n.detach().cpu().numpy(), rn, atol=2e-5)
    gv = np.random.default_rng(1).standard_normal(rv.shape).astype(np.float32)
    (v * torch.from_numpy(gv).to(cuda)).sum().backward()
    ref_g = pm.marching_cubes_backward(gv, rn, rsrc, rw, sdf.size, bandwidth=2.0)
    got_g = sdf_t.grad.cpu().numpy().reshape(-1)
    gmax = float(np.abs(ref_g).max()) if ref_g.size else 0.0
    np.testing.assert_allclose(got_g, ref_g, atol=1e-5 * (1 + gmax))
    assert (gmax > 0) == (rv.shape[0] > 0)


def _tree(cuda):
    import ngspsr
    from ngspsr import synthetic
    xyz, nrm = synthetic.torus(6000, seed=3)
    res = ngspsr.reconstruct(torch.from_numpy(xyz).to(cuda), torch.from_numpy(nrm).to(cuda), 1 / 32, depth=3,
                             min_density=8.0, screen_alpha=4.0, return_result=True)
    return res


@pytest.fixture(scope="module")
def recon(cuda):
    return _tree(cuda)


@pytest.mark.parametrize("expand,conforming", [(0, False), (0, True), (3, False), (5, True)])
def test_test_grid(cuda, recon, expand, conforming):
    """Voxel corners against a numpy restatement of _hash_backend.py:121-147; sample positions
    against the torch expression of hashtree.py:222-228 evaluated by torch on the same GPU."""
    from oracle.hierarchy import CONFORM_OFFSETS
    tree = recon.tree
    print(f"[case] test grid expand={expand} conforming={conforming}", flush=True)
    for depth in (0, 1):
        s = tree.strides[depth]
        c = tree.get_coords(None, depth).cpu().numpy().astype(np.int64)
        if expand >= 3:
            h = expand // 2
            o = np.stack(np.meshgrid(*[np.arange(-h, h + 1)] * 3, indexing="ij"), -1).reshape(-1, 3) * s
            c = np.unique((c[:, None, :] + o[None]).reshape(-1, 3), axis=0)
        if conforming:
            par = np.unique(np.floor(c / s / 2.0).astype(np.int64) * s * 2, axis=0)
            c = (par[:, None, :] + CONFORM_OFFSETS[None] * s).reshape(-1, 3)
        for res in (0, 2, 3, 4, 6, 8):
            q, base = tree.get_test_grid(res, depth, expand=expand, conforming=conforming)
            assert base.dtype == torch.int32 and np.array_equal(base.cpu().numpy(), c)
            if res == 0:
                assert q is None
                continue
            box = torch.linspace(0.0, s, res, device=cuda)
            box = torch.stack(torch.meshgrid(box, box, box, indexing="ij"), dim=3).view(-1, 3)
            ref = (base.unsqueeze(1) + box.unsqueeze(0)).view(-1, 3) * tree.voxel_size
            assert torch.equal(q, ref), (depth, res, int((q != ref).any(1).sum()))


@pytest.mark.parametrize("depth,res", [(0, 4), (1, 6), (0, 8)])
def test_extract_mesh_matches_oracle(cuda, recon, depth, res):
    from oracle import primal_mc as pm
    print(f"[case] extract_mesh depth={depth} res={res}", flush=True)
    rec, tree = recon.reconstructor, recon.tree
    q, base = tree.get_test_grid(res, depth)
    chi = rec.evaluate_chi(q, max_points=2 ** 20)
    v, f, n = rec.extract_mesh(base, chi, depth, build_o3d_mesh=False)
    s = tree.strides[depth]
    rv, rt, rn, _, _, _ = pm.marching_cubes(base.cpu().numpy() / s, chi.cpu().numpy().reshape(-1, res, res, res))
    scale = np.float32(s * tree.voxel_size)
    assert np.array_equal(v.cpu().numpy(), rv * scale) and np.array_equal(f.cpu().numpy(), rt)
    np.testing.assert_allclose(n.cpu().numpy(), rn, atol=2e-5)
    assert f.shape[0] > 100
    # the torus R = 0.35, r = 0.15 about (0.5, 0.5, 0.5): vertices within a voxel of the surface
    p = v.cpu().numpy().astype(np.float64) - 0.5
    d = np.hypot(np.hypot(p[:, 0], p[:, 1]) - 0.35, p[:, 2])
    assert np.abs(d - 0.15).max() < 1 / 32
    m = rec.extract_mesh(base, chi, depth)
    assert np.asarray(m.vertices).shape == tuple(v.shape) and np.asarray(m.triangles).shape == tuple(f.shape)
