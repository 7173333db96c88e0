"""Differentiable solve on the GPU: adjoint gradients vs the FP64 dense-adjoint reference and FP64
central differences (clouds of 10k points at voxel size 1/32, per-voxel weights from a BF16 MLP)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _torus_sdf(p, R=0.35, r=0.15, c=0.5):
    q = p - c
    return np.sqrt((np.sqrt(q[:, 0] ** 2 + q[:, 1] ** 2) - R) ** 2 + q[:, 2] ** 2) - r


@pytest.mark.parametrize("kind", ["torus", "ellipsoids"])
def test_adjoint_gradient(cuda, kind):
    from ngspsr import autograd, synthetic
    from oracle import adjoint, assembly, hierarchy
    print(f"[case] autograd {kind}", flush=True)
    if kind == "torus":
        xyz, nrm = synthetic.torus(10000, seed=11)
    else:
        xyz, nrm = synthetic.ellipsoids(10000, k=3, sigma=0.005, seed=12)
    rng = np.random.default_rng(5)
    vs, L = 1 / 32, 4
    xyz_t, nrm_t = torch.from_numpy(xyz).to(cuda), torch.from_numpy(nrm).to(cuda)
    op, nd = autograd.setup(xyz_t, nrm_t, vs, L, min_density=8.0)
    tree = op.tree
    torch.manual_seed(0)
    mlp = torch.nn.Sequential(torch.nn.Linear(3, 64), torch.nn.ReLU(), torch.nn.Linear(64, 64), torch.nn.ReLU(),
                              torch.nn.Linear(64, 2)).to(cuda, torch.bfloat16)
    centres = torch.cat([tree.get_voxel_centers(d) for d in range(L)])
    out = mlp(centres.to(torch.bfloat16)).float()
    M0 = tree.get_coords_size(None, 0)
    alpha_vox = 4.0 * (1 + 0.5 * torch.tanh(out[:M0, 0]))
    scale = 1 + 0.5 * torch.tanh(out[:, 1])
    cell = torch.floor(xyz_t * tree.inv_vs).int()
    idx = tree.grids[0].query(cell)
    alpha_p = torch.where(idx >= 0, alpha_vox[idx.clamp(min=0)], torch.full_like(idx, 4, dtype=torch.float32))
    q_np = (synthetic.torus(4000, seed=13)[0] if kind == "torus" else xyz[rng.choice(xyz.shape[0], 4000)])
    q_np = (q_np + rng.normal(0, 0.01, q_np.shape)).astype(np.float32)
    target = _torus_sdf(q_np) if kind == "torus" else rng.normal(0, 0.05, 4000)
    q_t = torch.from_numpy(q_np).to(cuda)
    tgt_t = torch.from_numpy(target.astype(np.float32)).to(cuda)

    a = alpha_p.detach().clone().requires_grad_(True)
    s = scale.detach().clone().requires_grad_(True)
    chi = op(a, s, q_t)
    loss = ((chi - tgt_t) ** 2).mean()
    loss.backward()
    ga_gpu = a.grad.double().cpu().numpy()
    gs_gpu = s.grad.double().cpu().numpy()

    # FP64 reference on the same tree
    h = hierarchy.Hierarchy(L, vs)
    for d in range(L):
        h.set_coords(d, tree.get_coords(None, d).cpu().numpy())
    xd = tree.xyz_depth.cpu().numpy()
    rho = tree.xyz_density.double().cpu().numpy()
    nd_ref = assembly.splat_normals(h, xyz, nrm, xd, 1.0 / rho)
    bs = adjoint.BlockSystem(h, nd_ref, xyz, delta=0.1)
    a_np = a.detach().double().cpu().numpy()
    s_np = s.detach().double().cpu().numpy()
    ga, gs, X = bs.gradient(a_np, s_np, q_np, target)
    chi_ref = bs.field(X, q_np)
    assert np.linalg.norm(chi.detach().cpu().numpy() - chi_ref) <= 1e-4 * np.linalg.norm(chi_ref)
    assert np.linalg.norm(ga_gpu - ga) <= 2e-3 * np.linalg.norm(ga), np.linalg.norm(ga_gpu - ga) / np.linalg.norm(ga)
    assert np.linalg.norm(gs_gpu - gs) <= 2e-3 * np.linalg.norm(gs), np.linalg.norm(gs_gpu - gs) / np.linalg.norm(gs)
    # FP64 central differences on the largest entries
    for p in np.argsort(-np.abs(ga))[:3]:
        hstep = 1e-4 * a_np[p]
        ap, am = a_np.copy(), a_np.copy()
        ap[p] += hstep
        am[p] -= hstep
        fd = (bs.loss(ap, s_np, q_np, target)[0] - bs.loss(am, s_np, q_np, target)[0]) / (2 * hstep)
        assert abs(ga_gpu[p] - fd) <= 5e-3 * abs(fd), (p, ga_gpu[p], fd)
    for i in np.argsort(-np.abs(gs))[:3]:
        hstep = 1e-4
        sp_, sm_ = s_np.copy(), s_np.copy()
        sp_[i] += hstep
        sm_[i] -= hstep
        fd = (bs.loss(a_np, sp_, q_np, target)[0] - bs.loss(a_np, sm_, q_np, target)[0]) / (2 * hstep)
        assert abs(gs_gpu[i] - fd) <= 5e-3 * abs(fd), (i, gs_gpu[i], fd)

    # the chain reaches the MLP parameters
    chi2 = op(alpha_p, scale, q_t)
    ((chi2 - tgt_t) ** 2).mean().backward()
    gw = mlp[0].weight.grad
    assert gw is not None and torch.isfinite(gw.float()).all() and gw.float().abs().sum() > 0


def test_config4_batch_of_32_clouds(cuda):
    """Config 4: 32 clouds of 10k points (16 random tori, 16 random 3-ellipsoid scenes, sigma 0.005)
    at voxel size 1/32; per-voxel screening weights and per-basis scales from one random-init BF16
    MLP (3 layers, width 64) on voxel centres; loss = mean over the clouds of the L2 distance of chi
    at 4k seeded query points to the analytic SDF; one backward through all 32 adjoint solves into
    the MLP. Two clouds' parameter gradients are checked against the FP64 dense adjoint."""
    import time
    from ngspsr import autograd, synthetic
    from oracle import adjoint, assembly, hierarchy
    print("[case] config 4: 32 clouds", flush=True)
    vs, L, n = 1 / 32, 4, 10000
    torch.manual_seed(0)
    mlp = torch.nn.Sequential(torch.nn.Linear(3, 64), torch.nn.ReLU(), torch.nn.Linear(64, 64), torch.nn.ReLU(),
                              torch.nn.Linear(64, 2)).to(cuda, torch.bfloat16)
    clouds = []
    for i in range(32):
        rng = np.random.default_rng(100 + i)
        if i % 2 == 0:
            R, r = rng.uniform(0.25, 0.35), rng.uniform(0.08, 0.14)
            c = rng.uniform(0.45, 0.55, 3)
            xyz, nrm = synthetic.torus(n, R=R, r=r, center=c, seed=100 + i)
            sdf = (lambda p, R=R, r=r, c=c: synthetic.torus_sdf(p, R, r, c))
        else:
            xyz, nrm, prm = synthetic.ellipsoids(n, k=3, sigma=0.0, seed=100 + i, return_params=True)
            sdf = (lambda p, prm=prm: synthetic.ellipsoids_sdf(p, prm))
        xyz = (xyz + rng.normal(0, 0.005, xyz.shape)).astype(np.float32)
        q = (xyz[rng.choice(n, 4000)] + rng.normal(0, 0.01, (4000, 3))).astype(np.float32)
        clouds.append((xyz, nrm, q, sdf(q).astype(np.float32)))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses, keep = [], {}
    for i, (xyz, nrm, q, tgt) in enumerate(clouds):
        xyz_t, nrm_t = torch.from_numpy(xyz).to(cuda), torch.from_numpy(nrm).to(cuda)
        op, nd = autograd.setup(xyz_t, nrm_t, vs, L, min_density=8.0)
        tree = op.tree
        centres = torch.cat([tree.get_voxel_centers(d) for d in range(L)])
        out = mlp(centres.to(torch.bfloat16)).float()
        M0 = tree.get_coords_size(None, 0)
        alpha_vox = 4.0 * (1 + 0.5 * torch.tanh(out[:M0, 0]))
        scale = 1 + 0.5 * torch.tanh(out[:, 1])
        idx = tree.grids[0].query(torch.floor(xyz_t * tree.inv_vs).int())
        alpha_p = torch.where(idx >= 0, alpha_vox[idx.clamp(min=0)], torch.full_like(idx, 4, dtype=torch.float32))
        if i in (0, 17):  # leaves for the spot checks
            alpha_p.retain_grad()
            scale.retain_grad()
            keep[i] = (op, nd, alpha_p, scale)
        chi = op(alpha_p, scale, torch.from_numpy(q).to(cuda))
        losses.append(((chi - torch.from_numpy(tgt).to(cuda)) ** 2).mean())
    loss = torch.stack(losses).mean()
    loss.backward()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(f"[result] config 4: 32 clouds forward + adjoint backward {dt:.2f} s, loss {float(loss):.4e}", flush=True)
    for p in mlp.parameters():
        assert p.grad is not None and bool(torch.isfinite(p.grad.float()).all())
    assert float(mlp[0].weight.grad.float().abs().sum()) > 0
    for i, (op, nd, alpha_p, scale) in keep.items():
        xyz, nrm, q, tgt = clouds[i]
        tree = op.tree
        h = hierarchy.Hierarchy(L, vs)
        for d in range(L):
            h.set_coords(d, tree.get_coords(None, d).cpu().numpy())
        nd_ref = assembly.splat_normals(h, xyz, nrm, tree.xyz_depth.cpu().numpy(),
                                        1.0 / tree.xyz_density.double().cpu().numpy())
        bs = adjoint.BlockSystem(h, nd_ref, xyz, delta=0.1)
        a_np = alpha_p.detach().double().cpu().numpy()
        s_np = scale.detach().double().cpu().numpy()
        ga, gs, _ = bs.gradient(a_np, s_np, q, tgt.astype(np.float64))
        # the mean over 32 clouds scales each cloud's gradient by 1/32
        ga_gpu = 32 * alpha_p.grad.double().cpu().numpy()
        gs_gpu = 32 * scale.grad.double().cpu().numpy()
        assert np.linalg.norm(ga_gpu - ga) <= 2e-3 * np.linalg.norm(ga), (i, np.linalg.norm(ga_gpu - ga) / np.linalg.norm(ga))
        assert np.linalg.norm(gs_gpu - gs) <= 2e-3 * np.linalg.norm(gs), (i, np.linalg.norm(gs_gpu - gs) / np.linalg.norm(gs))
