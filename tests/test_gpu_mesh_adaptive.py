"""Dual mesh of strongly adaptive trees (sparse or empty finest level, leaves on every level) against
the oracle's on the same field values, bit for bit: corner positions, faces and vertices. These
exercise the dual-cube gather's coarse-level probes and the per-edge de-duplication of cubes whose
corners repeat leaves."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

# (scene, points, voxel size, depth, min_density, seed): an empty finest level (the torus at
# min_density 32), sparse ones, and leaves spread over all the levels
CASES = [("sphere", 3237, 1 / 64, 3, 4.0, 301276744), ("sphere", 6550, 1 / 64, 4, 8.0, 405750087),
         ("torus", 3177, 1 / 32, 4, 32.0, 69581195), ("sphere", 7656, 1 / 32, 4, 32.0, 628546683)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{c[1]}-d{c[3]}-md{int(c[4])}")
def test_adaptive_mesh_matches_oracle(cuda, case):
    import ngspsr
    from ngspsr import synthetic
    from oracle import hierarchy, mesh as om
    kind, n, vs, depth, mind, seed = case
    print(f"[case] adaptive mesh {kind} n={n} vs={vs} depth={depth} min_density={mind}", flush=True)
    if kind == "torus":
        xyz, nrm = synthetic.torus(n, seed=seed)
    else:
        xyz, nrm = synthetic.sphere(n, seed=seed)
    res = ngspsr.reconstruct(torch.from_numpy(xyz).to(cuda), torch.from_numpy(nrm).to(cuda), vs, depth=depth,
                             min_density=mind, screen_alpha=4.0, solver="pcg@tol=1e-7@maxiter=20000",
                             return_result=True)
    torch.cuda.synchronize()
    gd = res.tree.xyz_depth.cpu().numpy()
    h = hierarchy.Hierarchy(depth, vs)
    for d in range(depth):
        h.set_coords(d, hierarchy.quantize_coords(xyz[gd <= d], vs, h.strides[d]))
    sizes = [int(res.tree.get_coords_size(None, d)) for d in range(depth)]
    assert min(sizes[1:]) > 0 and sizes[0] < sizes[1] * 4, sizes  # adaptive: a sparse or empty finest level
    v, f, info = res.reconstructor.extract_multiscale_mesh(n_upsample=2, return_info=True)
    vals = info["values"].cpu().numpy()
    v_ref, f_ref, ref = om.extract_multiscale_mesh(h, None, None, None, n_upsample=2, chi_fn=lambda p: vals)
    assert f_ref.shape[0] > 1000
    np.testing.assert_array_equal(info["pos"].cpu().numpy(), ref["corners"])
    np.testing.assert_array_equal(f.cpu().numpy(), f_ref)
    np.testing.assert_array_equal(v.cpu().numpy(), v_ref)
