"""Pins the oracle's density, level assignment and level set (mean chi) by closed forms.

compute_density (_hash_backend.py:13-28 with hashtree.py:283-285) is a trilinear splat of ones on
the grid of 4x the voxel size, interpolated back at the points with the degree-2 B-spline basis
B = 2 N2 and divided by 16. The closed forms below are written from the textbook quadratic
B-spline N2 and the tent (trilinear) weights, not from the oracle's code:
- an isolated point at offset u_a (in 4x cells) from the centre of its own cell gets
  prod_a [(1 - u_a) B(u_a) + u_a B(1 - u_a)] / 16 (two splat nodes per axis, at distances u_a and
  1 - u_a);
- a lattice with m points per cell along each axis gives every interior node m^3 splat weight, and
  the basis sums to 2 per axis, so interior points get 8 m^3 / 16.
level_of_density (_hash_backend.py:428-429) puts a point j levels coarser when its density lies in
[min / 4^j, min / 4^(j-1)), clamped to the deepest level. mean_chi (reconstructor.py:228-236) is
the weighted mean of chi at the points; a field whose coefficients are linear in the voxel centre
reproduces the linear function (partition of unity and linear precision of N2), so chi and its
weighted mean are known in closed form."""
import numpy as np
import pytest

from oracle import hierarchy, mesh


def _n2(x):
    """Centred quadratic B-spline (textbook): 3/4 - x^2 on |x| <= 1/2, (3/2 - |x|)^2 / 2 on
    1/2 <= |x| <= 3/2, 0 beyond."""
    a = np.abs(np.asarray(x, np.float64))
    return np.where(a <= 0.5, 0.75 - a * a, np.where(a <= 1.5, 0.5 * (1.5 - a) ** 2, 0.0))


def _b(x):
    return 2.0 * _n2(x)


def test_density_of_isolated_points_closed_form():
    vs = 1.0 / 256
    cell = 4 * vs  # density grid: 4x the voxel size
    rng = np.random.default_rng(0)
    # isolated points: one per 8x8x8 block of density cells, at random offsets inside the block's
    # central cell, so no two points share a splat node or an interpolation neighbourhood
    blocks = np.array([(i, j, k) for i in range(3) for j in range(3) for k in range(3)], np.float64)
    frac = rng.uniform(0.02, 0.98, size=blocks.shape)
    xyz = ((blocks * 8 + 3 + frac) * cell).astype(np.float32)
    rho = hierarchy.compute_density(xyz, vs)
    q = xyz.astype(np.float64) / cell
    u = np.abs(q - np.floor(q) - 0.5)
    per_axis = (1 - u) * _b(u) + u * _b(1 - u)
    expect = per_axis.prod(axis=1) / 16.0
    np.testing.assert_allclose(rho, expect, rtol=2e-6)
    # a point at its cell centre: 1.5 per axis
    assert abs(hierarchy.compute_density(np.array([[3.5, 3.5, 3.5]], np.float32) * cell, vs)[0] - 1.5 ** 3 / 16) < 1e-6


@pytest.mark.parametrize("m", [1, 2, 3])
def test_density_of_interior_lattice_points(m):
    vs = 1.0 / 64
    cell = 4 * vs
    n_cells = 9
    g = (np.arange(n_cells * m) + 0.5) / m * cell
    xyz = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    rho = hierarchy.compute_density(xyz, vs)
    q = xyz.astype(np.float64) / cell
    interior = np.all((q > 3) & (q < n_cells - 3), axis=1)
    assert interior.sum() >= m ** 3
    np.testing.assert_allclose(rho[interior], 8 * m ** 3 / 16.0, rtol=1e-5)
    # no point exceeds the interior value; the outermost ones lose the weight of the missing cells
    assert np.all(rho <= 8 * m ** 3 / 16.0 * (1 + 1e-5))
    outer = np.any((q < 1.0 / m) | (q > n_cells - 1.0 / m), axis=1)
    assert outer.sum() > 0 and np.all(rho[outer] < 0.99 * 8 * m ** 3 / 16.0)


def test_level_of_density_table():
    mn, depth, eps = 8.0, 5, 1e-6
    rows = []  # (j, sign of the perturbation, expected level)
    expected = {(3, 1): 0, (3, -1): 0, (1, 1): 0, (1, -1): 0, (0, 1): 0, (0, -1): 1, (-1, 1): 1, (-1, -1): 2,
                (-2, 1): 2, (-2, -1): 3, (-3, 1): 3, (-3, -1): 4, (-4, 1): 4, (-4, -1): 4, (-6, 1): 4}
    for (j, sg), lv in expected.items():
        rows.append((mn * 4.0 ** j * (1 + sg * eps), lv))
    rho = np.array([r for r, _ in rows])
    want = np.array([lv for _, lv in rows], np.int32)
    got = hierarchy.level_of_density(rho, mn, 4.0, depth)
    assert np.array_equal(got, want), (got, want)
    # a shallower tree clamps to its own deepest level
    got3 = hierarchy.level_of_density(rho, mn, 4.0, 3)
    assert np.array_equal(got3, np.minimum(want, 2))
    # another base: density in [min / 2^j, min / 2^(j-1)) goes j levels coarser
    assert list(hierarchy.level_of_density(np.array([8.0, 7.9, 4.0, 3.9, 1.0]), 8.0, 2.0, 6)) == [0, 1, 1, 2, 3]


@pytest.mark.parametrize("alpha,beta", [(0.7, 0.0), (0.25, -0.5), (-1.0, 2.0)])
def test_mean_chi_of_linear_field(alpha, beta):
    vs = 1.0 / 32
    n = 14
    h = hierarchy.Hierarchy(1, vs)
    ax = np.arange(n)
    coords = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    h.set_coords(0, coords)
    # coefficients linear in the voxel centre's x (finest units): chi = 8 (alpha + beta q_x)
    f = alpha + beta * (h.coords[0][:, 0].astype(np.float64) + 0.5)
    rng = np.random.default_rng(3)
    xyz = (rng.uniform(3.0, n - 3.0, size=(500, 3)) * vs).astype(np.float32)
    w = rng.uniform(0.1, 5.0, size=500)
    q = xyz.astype(np.float64) / vs
    chi = 8.0 * (alpha + beta * q[:, 0])
    np.testing.assert_allclose(mesh.evaluate_raw_chi(h, {0: f}, xyz), chi, rtol=1e-5, atol=1e-5)
    expect = float(np.sum(w * chi) / np.sum(w))
    got = mesh.mean_chi(h, {0: f}, xyz, w)
    assert abs(got - expect) <= 1e-5 * max(1.0, abs(expect))
    if beta != 0.0:
        # the weights matter: the unweighted mean differs
        assert abs(float(np.mean(chi)) - expect) > 1e-3
