"""Pins the oracle's learned-basis multigrid (oracle/learned.py) on the CPU.

The degree-2 B-spline is the one-component additive basis with unit features, so the generic path
must rebuild exactly the systems of the B-spline oracle (pinned by quadrature and brute force) when
its tables are the B-spline's exact integrals (f_f = F, df_df = G / M, f = P, df = Q / M, f_inv =
Pinv, df_inv = Qinv: the additive tables carry the 1/M of the chain rule that the B-spline path
applies through its strides). Features enter multiplicatively per axis: doubling every feature
scales each level matrix by 2^6 and the normal part of the right-hand side by 2^3."""
import numpy as np
import pytest

from oracle import assembly, bezier, learned, mesh
from oracle import hierarchy as H


def _bezier_tables():
    ff = [bezier.as_f64(bezier.F_table(2 ** m))[None, None] for m in range(5)]
    dd = [(bezier.as_f64(bezier.G_table(2 ** m)) / 2 ** m)[None, None] for m in range(5)]
    f = [[bezier.as_f64(bezier.P_table(2 ** m)) for m in range(5)]]
    df = [[bezier.as_f64(bezier.Q_table(2 ** m)) / 2 ** m for m in range(5)]]
    f_inv = [[bezier.as_f64(bezier.Pinv_table(2 ** m)) for m in range(1, 5)]]
    df_inv = [[bezier.as_f64(bezier.Qinv_table(2 ** m)) for m in range(1, 5)]]
    return learned.AdditiveBasisOracle([("bezier",)], ff, dd, f, df, f_inv, df_inv)


def _scene():
    rng = np.random.default_rng(4)
    d = rng.standard_normal((700, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    xyz = (0.5 + 0.3 * d).astype(np.float32)
    nrm = d.astype(np.float32)
    h, lv, rho = assembly.build_tree(xyz, 1 / 16, 3, 2.0)
    nd = assembly.splat_normals(h, xyz, nrm, lv, 1.0 / rho)
    return xyz, h, nd


def _dense(a, b):
    return np.linalg.solve(a.toarray(), b), 0


@pytest.mark.parametrize("alpha", [4.0, "point"])
def test_generic_path_with_bspline_tables_equals_bspline_oracle(alpha):
    xyz, h, nd = _scene()
    if alpha == "point":
        alpha = np.random.default_rng(1).uniform(1.0, 6.0, xyz.shape[0])
    ref = assembly.solve_multigrid(h, nd, xyz, alpha, 0.1, _dense)
    basis = _bezier_tables()
    feats = {d: np.ones((h.num_voxels(d), 3)) for d in range(h.depth)}
    got = learned.solve_multigrid(h, basis, feats, nd, xyz, alpha, 0.1, _dense)
    for d in range(h.depth):
        if h.num_voxels(d) == 0:
            continue
        scale = np.abs(ref.lhs[d].data).max()
        assert np.abs((got.lhs[d] - ref.lhs[d]).data).max(initial=0.0) <= 1e-12 * scale, d
        np.testing.assert_allclose(got.rhs[d], ref.rhs[d], rtol=0, atol=1e-12 * np.abs(ref.rhs[d]).max())
        np.testing.assert_allclose(got.solutions[d], ref.solutions[d], rtol=1e-10, atol=1e-12)
    q = xyz[::7]
    np.testing.assert_allclose(learned.evaluate(h, basis, feats, got.solutions, q),
                               mesh.evaluate_raw_chi(h, got.solutions, q), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(learned.evaluate(h, basis, feats, got.solutions, q, compute_grad=True),
                               mesh.evaluate_raw_chi(h, got.solutions, q, compute_grad=True), rtol=1e-9, atol=1e-10)


def test_feature_scaling():
    xyz, h, nd = _scene()
    basis = _bezier_tables()
    one = {d: np.ones((h.num_voxels(d), 3)) for d in range(h.depth)}
    two = {d: 2 * np.ones((h.num_voxels(d), 3)) for d in range(h.depth)}
    d = 0
    # the finest level alone, no coarser solutions: A and the normal RHS only
    r1 = learned.solve_multigrid(h, basis, one, nd, xyz, 4.0, 0.0, _dense, start_depth=0, end_depth=0)
    r2 = learned.solve_multigrid(h, basis, two, nd, xyz, 4.0, 0.0, _dense, start_depth=0, end_depth=0)
    assert np.abs((r2.lhs[d] - 64 * r1.lhs[d]).data).max(initial=0.0) <= 1e-12 * np.abs(r2.lhs[d].data).max()
    np.testing.assert_allclose(r2.rhs[d], 8 * r1.rhs[d], rtol=1e-12, atol=1e-14)
    # a feature of one voxel changes only that voxel's row and column
    f3 = {k: v.copy() for k, v in one.items()}
    f3[0][5] *= 3.0
    r3 = learned.solve_multigrid(h, basis, f3, nd, xyz, 4.0, 0.0, _dense, start_depth=0, end_depth=0)
    diff = (r3.lhs[d] - r1.lhs[d]).tocoo()
    touched = np.unique(np.concatenate([diff.row[np.abs(diff.data) > 0], diff.col[np.abs(diff.data) > 0]]))
    assert 5 in touched
    rows, cols = diff.row[np.abs(diff.data) > 0], diff.col[np.abs(diff.data) > 0]
    assert np.all((rows == 5) | (cols == 5))
