"""reconstruct on degenerate inputs (jspsr/__init__.py:8-61): empty, a handful of points (the finer
levels empty), identical points, a flat cloud; input and output types."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _check_mesh(v, f):
    v, f = np.asarray(v), np.asarray(f)
    assert v.ndim == 2 and v.shape[1] == 3 and f.ndim == 2 and f.shape[1] == 3
    assert np.isfinite(v).all()
    assert f.size == 0 or (f.min() >= 0 and f.max() < v.shape[0])


def test_empty_input(cuda):
    import ngspsr
    print("[case] empty input", flush=True)
    v, f = ngspsr.reconstruct(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32), 1 / 32)
    assert isinstance(v, np.ndarray) and v.shape == (0, 3) and f.shape == (0, 3) and f.dtype == np.int64
    vt, ft = ngspsr.reconstruct(torch.zeros((0, 3), device=cuda), torch.zeros((0, 3), device=cuda), 1 / 32)
    assert vt.is_cuda and vt.shape == (0, 3) and ft.shape == (0, 3)
    with pytest.raises(AssertionError):
        ngspsr.reconstruct(np.zeros((10, 3), np.float32), np.zeros((9, 3), np.float32), 1 / 32)


@pytest.mark.parametrize("kind", ["few", "identical", "flat", "sphere-numpy-float64"])
def test_degenerate_clouds(cuda, kind):
    import ngspsr
    from ngspsr import synthetic
    print(f"[case] degenerate cloud {kind}", flush=True)
    rng = np.random.default_rng(4)
    if kind == "few":  # low density: every point on the coarsest level, the finer levels empty
        xyz, nrm = synthetic.sphere(12, seed=1)
    elif kind == "identical":
        xyz = np.tile(np.float32([[0.5, 0.5, 0.5]]), (500, 1))
        nrm = np.tile(np.float32([[0.0, 0.0, 1.0]]), (500, 1))
    elif kind == "flat":  # a disc in the plane z = 0.5
        r, t = np.sqrt(rng.uniform(0, 0.09, 4000)), rng.uniform(0, 2 * np.pi, 4000)
        xyz = np.stack([0.5 + r * np.cos(t), 0.5 + r * np.sin(t), np.full(4000, 0.5)], 1).astype(np.float32)
        nrm = np.tile(np.float32([[0.0, 0.0, 1.0]]), (4000, 1))
    else:
        xyz, nrm = synthetic.sphere(5000, seed=2)
        xyz, nrm = xyz.astype(np.float64), nrm.astype(np.float64)
    res = ngspsr.reconstruct(xyz, nrm, 1 / 32, depth=3, return_result=True)
    torch.cuda.synchronize()
    _check_mesh(res.vertices.cpu().numpy(), res.triangles.cpu().numpy())
    for d, (coords, x) in res.solution_field.items():
        assert coords.shape[0] == x.shape[0] and bool(torch.isfinite(x).all()), d
    v, f = ngspsr.reconstruct(xyz, nrm, 1 / 32, depth=3)
    assert isinstance(v, np.ndarray) and isinstance(f, np.ndarray)
    _check_mesh(v, f)
    # the learned caller's path on the same field: an empty voxel set gives an empty mesh
    rec = res.reconstructor
    q, base = res.tree.get_test_grid(4, 0)
    vv, ff, nn = rec.extract_mesh(base[:0], rec.evaluate_chi(q[:0]), 0, build_o3d_mesh=False)
    assert vv.shape == (0, 3) and ff.shape == (0, 3) and nn.shape == (0, 3)
    if base.shape[0]:
        vv, ff, nn = rec.extract_mesh(base, rec.evaluate_chi(q), 0, build_o3d_mesh=False)
        _check_mesh(vv.cpu().numpy(), ff.cpu().numpy())
    if kind == "sphere-numpy-float64":
        # a closed surface: the mesh lies near the sphere of radius 0.3 about the centre
        d = np.linalg.norm(v - 0.5, axis=1)
        assert f.shape[0] > 100 and np.abs(d - 0.3).max() < 2 / 32


def test_serving_mixes_degenerate_requests(cuda):
    """An empty request and a low-density one (finer levels empty) in a stream with ordinary ones:
    every request completes, the degenerate ones with the empty mesh reconstruct gives them."""
    import ngspsr
    from ngspsr import serving, synthetic
    print("[case] serving with degenerate requests", flush=True)
    xyz, nrm = synthetic.ellipsoids(20000, k=4, seed=9)
    few_x, few_n = synthetic.sphere(12, seed=1)
    reqs_data = [(xyz, nrm), (np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32)), (few_x, few_n),
                 (xyz[::2].copy(), nrm[::2].copy())]
    srv = serving.ReconstructionServer(cuda, workers=4, tol=1e-6, refine=True)
    try:
        reqs = [serving.Request(i, torch.from_numpy(x).to(cuda), torch.from_numpy(n).to(cuda), 1 / 64, depth=4,
                                min_density=1.0 if i != 2 else 32.0) for i, (x, n) in enumerate(reqs_data)]
        srv.serve(reqs)
    finally:
        srv.close()
    for r in reqs:
        assert r.error is None, (r.rid, r.error)
        v, f = r.result
        _check_mesh(v.cpu().numpy(), f.cpu().numpy())
    assert reqs[1].result[0].shape == (0, 3)
    ref_v, ref_f = ngspsr.reconstruct(few_x, few_n, 1 / 64, depth=4, min_density=32.0)
    assert reqs[2].result[1].shape[0] == ref_f.shape[0]
    assert reqs[0].result[1].shape[0] > 1000 and reqs[3].result[1].shape[0] > 1000
