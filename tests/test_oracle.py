"""Pins the CPU oracle: known-answer vectors, brute force on tiny inputs, closed forms, invariants."""
import itertools
import json
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import assembly, bezier, fnv, hierarchy, mc_tables, mesh, pcg

KA = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")))


# ------------------------------------------------------------------------------------------- keys
def _fnv_python(words):
    h = 14695981039346656037
    for w in words:
        h ^= w & 0xFFFFFFFF
        h = (h * 1099511628211) % (1 << 64)
    return (h >> 60) ^ (h & ((1 << 60) - 1))


def test_fnv_known_answers():
    for e in KA["fnv_fold60"]:
        assert int(fnv.fnv_fold60(e["coords"])[0]) == int(e["key"], 16)
        assert _fnv_python(e["coords"]) == int(e["key"], 16)


def test_fnv_matches_independent_python_on_random_rows():
    rng = np.random.default_rng(0)
    c = rng.integers(-2 ** 31, 2 ** 31 - 1, size=(200, 3))
    k = fnv.fnv_fold60(c)
    for row, key in zip(c, k):
        assert int(key) == _fnv_python([int(v) for v in row])
    # int64 rows keep their low 32 bits ((unsigned int) cast of hashmap_cuda.cu:255)
    c64 = c.astype(np.int64) + (np.int64(5) << 33)
    assert np.array_equal(fnv.fnv_fold60(c64), k)


def test_kernel_hash_layout():
    c = np.array([[0, 0, 0], [1, 2, 3]])
    off = np.array([[0, 0, 0], [1, 0, 0]])
    k = fnv.kernel_hash(c, off)
    assert k.shape == (2, 2)
    assert int(k[1, 0]) == _fnv_python([1, 0, 0]) and int(k[1, 1]) == _fnv_python([2, 2, 3])


def test_table_sizing_known_answers():
    for e in KA["sizing"]:
        assert fnv.table_sizing(e["n"]) == (e["table_size"], e["num_buckets"], e["evict_bound"])
    assert fnv.table_sizing(3000, enlarge=True)[0] == 4 * 4096


# ---------------------------------------------------------------------------------- basis tables
@pytest.mark.parametrize("name,fn", [("F", bezier.F_table), ("G", bezier.G_table), ("P", bezier.P_table),
                                     ("Q", bezier.Q_table), ("Pinv", bezier.Pinv_table), ("Qinv", bezier.Qinv_table)])
def test_tables_known_answers(name, fn):
    for m, vals in KA["tables"][name].items():
        assert list(fn(int(m))) == [Fraction(v) for v in vals], (name, m)


def test_tables_textbook_stencils_and_sums():
    st = [4 * Fraction(v) for v in KA["stencil_1d"]["stiffness"]]
    ms = [4 * Fraction(v) for v in KA["stencil_1d"]["mass"]]
    assert list(bezier.G_table(1)) == st and list(bezier.F_table(1)) == ms
    for m in (1, 2, 4, 8, 16):
        assert sum(bezier.F_table(m)) == 4 * m and sum(bezier.G_table(m)) == 0
        assert sum(bezier.P_table(m)) == 2 * m and sum(bezier.Q_table(m)) == 0
        assert list(bezier.F_table(m)) == list(reversed(bezier.F_table(m)))  # symmetry


def test_two_scale_relation():
    # B(x/2) = 1/4 sum_j h_j B(x - j), h = (1, 3, 3, 1) at j = -3/2, -1/2, 1/2, 3/2
    x = np.linspace(-3.2, 3.2, 1001)
    rhs = 0.25 * sum(h * bezier.B(x - j) for h, j in zip((1, 3, 3, 1), (-1.5, -0.5, 0.5, 1.5)))
    np.testing.assert_allclose(bezier.B(x / 2), rhs, atol=1e-12)


def _gl(lo, hi, breaks, n=4):
    xs, ws = np.polynomial.legendre.leggauss(n)
    pts = sorted(set([lo, hi] + [b for b in breaks if lo < b < hi]))
    X, W = [], []
    for a, b in zip(pts[:-1], pts[1:]):
        X.append(0.5 * (a + b) + 0.5 * (b - a) * xs)
        W.append(0.5 * (b - a) * ws)
    return np.concatenate(X), np.concatenate(W)


def _basis3(X, Y, Z, c, s):
    """B and grad B (world units) of the voxel with corner c and stride s (finest units = world)."""
    u = [(A - (ci + s / 2)) / s for A, ci in zip((X, Y, Z), c)]
    b = [bezier.B(v) for v in u]
    db = [bezier.dB(v) / s for v in u]
    return b[0] * b[1] * b[2], (db[0] * b[1] * b[2], b[0] * db[1] * b[2], b[0] * b[1] * db[2])


def _breaks(c, s):
    return [ci + s / 2 + k * s for k in (-1.5, -0.5, 0.5, 1.5) for ci in [c]]


@pytest.mark.parametrize("m", [1, 2, 4])
def test_stiffness_entries_against_3d_quadrature(m):
    """integrate_deriv_deriv_product == int grad B_src . grad B_tgt (brute-force 3-D quadrature)."""
    rng = np.random.default_rng(m)
    ss, st = 2, 2 * m
    for _ in range(6):
        ct = rng.integers(-3, 3, 3) * st
        cs = ct + rng.integers(-2 * m - 1, 2 * m + 1, 3) * ss
        rel = ((ct + st / 2) - (cs + ss / 2)) / ss
        val = bezier.integrate_deriv_deriv_product(rel[None], ss, st)[0]
        axes = []
        for a in range(3):
            br = _breaks(cs[a], ss) + _breaks(ct[a], st)
            axes.append(_gl(min(br), max(br), br))
        (X, wx), (Y, wy), (Z, wz) = axes
        XX, YY, ZZ = np.meshgrid(X, Y, Z, indexing="ij")
        W = wx[:, None, None] * wy[None, :, None] * wz[None, None, :]
        _, gs = _basis3(XX, YY, ZZ, cs, ss)
        _, gt = _basis3(XX, YY, ZZ, ct, st)
        brute = np.sum(W * (gs[0] * gt[0] + gs[1] * gt[1] + gs[2] * gt[2]))
        assert abs(val - brute) <= 1e-9 * max(1.0, abs(brute)), (cs, ct, val, brute)


@pytest.mark.parametrize("sd,st", [(2, 2), (2, 4), (2, 8), (4, 2), (8, 2)])
def test_rhs_entries_against_3d_quadrature(sd, st):
    """integrate_const_deriv_product == int_{data voxel} grad B_tgt . V (brute-force quadrature),
    for data finer, equal and (within the reference's |rel| <= 1.5 filter) coarser than target."""
    rng = np.random.default_rng(sd * 10 + st)
    checked = 0
    for _ in range(200):
        ct = rng.integers(-4, 4, 3) * st
        q = np.floor_divide(2 * ct + st, 2 * sd) * sd  # data cell containing the target centre
        cd = q + rng.integers(-1, 2, 3) * sd if sd > st else (ct // sd + rng.integers(-3, 3, 3)) * sd
        rel = ((ct + st / 2) - (cd + sd / 2)) / sd
        if st >= sd:
            m = st // sd
            idx = np.rint(rel + (3 * m - 1) / 2)
            if np.any(idx < 0) or np.any(idx >= 3 * m):
                continue
        else:
            m = sd // st
            if np.any(np.abs(-rel * m) > 1.5 + 1e-6):
                continue
        V = rng.standard_normal(3)
        val = bezier.integrate_const_deriv_product(V[None], rel[None], sd, st)[0]
        axes = []
        for a in range(3):
            axes.append(_gl(cd[a], cd[a] + sd, _breaks(ct[a], st)))
        (X, wx), (Y, wy), (Z, wz) = axes
        XX, YY, ZZ = np.meshgrid(X, Y, Z, indexing="ij")
        W = wx[:, None, None] * wy[None, :, None] * wz[None, None, :]
        _, g = _basis3(XX, YY, ZZ, ct, st)
        brute = np.sum(W * (V[0] * g[0] + V[1] * g[1] + V[2] * g[2]))
        assert abs(val - brute) <= 1e-9 * max(1.0, abs(brute)), (cd, ct, val, brute)
        checked += 1
        if checked >= 8:
            break
    assert checked >= 4


# ---------------------------------------------------------------------------------- hierarchy
def test_trilinear_partition_of_unity_and_corners():
    rng = np.random.default_rng(0)
    xyz = rng.uniform(0.1, 0.9, (500, 3)).astype(np.float32)
    for s in (1, 2, 4):
        c, a = hierarchy.trilinear_weights(xyz, 1 / 16, s)
        np.testing.assert_allclose(a.reshape(8, -1).sum(0), 1.0, atol=1e-6)
        assert np.all(c % s == 0)
        q = xyz.astype(np.float64) * 16
        d = np.abs(c.reshape(8, -1, 3) + s / 2 - q[None])
        assert np.all(d <= s + 1e-6)


def _brute_neighbours(cs, ss, ct, st, rng_):
    """All (source, target) pairs whose target is inside the rng_-cube (target units) around the
    target cell containing the source centre."""
    h = rng_ // 2
    out = []
    for i, c in enumerate(cs):
        q = np.floor((c + ss / 2) / st) * st
        for j, t in enumerate(ct):
            if np.all(np.abs(t - q) <= h * st):
                out.append((i, j))
    return sorted(out)


@pytest.mark.parametrize("d,dd,rng_", [(0, 0, 3), (0, 0, 5), (0, 1, 5), (0, 2, 3), (1, 0, 3), (2, 0, 3)])
def test_neighbour_coo_brute_force(d, dd, rng_):
    rng = np.random.default_rng(d * 7 + dd)
    xyz = rng.uniform(0.3, 0.7, (60, 3)).astype(np.float32)
    h = hierarchy.Hierarchy(3, 1 / 16)
    for k in range(3):
        h.set_coords(k, hierarchy.quantize_coords(xyz, 1 / 16, h.strides[k]))
    s, t, rel, _ = h.self_neighbours(d, dd, rng_)
    got = sorted(zip(s.tolist(), t.tolist()))
    if d <= dd:
        exp = _brute_neighbours(h.coords[d], h.strides[d], h.coords[dd], h.strides[dd], rng_)
    else:  # inverse: pairs of the swapped query, kept when the centres are within 1.5 fine strides
        fine, coarse = dd, d
        exp = []
        for (i, j) in _brute_neighbours(h.coords[fine], h.strides[fine], h.coords[coarse], h.strides[coarse], rng_):
            r = (h.coords[coarse][j] + h.strides[coarse] / 2 - h.coords[fine][i] - h.strides[fine] / 2) / h.strides[fine]
            if np.all(np.abs(r) < rng_ / 2 + 1e-6):
                exp.append((j, i))
        exp = sorted(exp)
    assert got == exp
    # rel = target centre - source centre in source strides (inverse: in coarse-source strides)
    cs = h.coords[d][s] + h.strides[d] / 2
    ct = h.coords[dd][t] + h.strides[dd] / 2
    expect_rel = (ct - cs) / h.strides[d] if d <= dd else (ct - cs) / h.strides[d]
    np.testing.assert_allclose(rel, expect_rel, atol=1e-6)


def test_quantize_matches_brute_force_unique():
    rng = np.random.default_rng(3)
    xyz = rng.uniform(0, 1, (300, 3)).astype(np.float32)
    for s in (1, 2, 4):
        got = hierarchy.quantize_coords(xyz, 1 / 8, s)
        c, _ = hierarchy.trilinear_weights(xyz, 1 / 8, s)
        exp = sorted(set(map(tuple, c.tolist())))
        assert [tuple(r) for r in got.tolist()] == exp


def test_evaluate_interpolated_matches_all_voxel_sum():
    rng = np.random.default_rng(5)
    xyz = rng.uniform(0.3, 0.7, (80, 3)).astype(np.float32)
    h = hierarchy.Hierarchy(2, 1 / 16)
    for k in range(2):
        h.set_coords(k, hierarchy.quantize_coords(xyz, 1 / 16, h.strides[k]))
    q = rng.uniform(0.3, 0.7, (50, 3)).astype(np.float32)
    for d in range(2):
        f = rng.standard_normal(h.num_voxels(d))
        got = h.evaluate_interpolated(q, d, f)
        s = h.strides[d]
        qq = q.astype(np.float64) * 16
        cen = h.coords[d] + s / 2
        u = (qq[:, None, :] - cen[None]) / s
        brute = (bezier.B(u[..., 0]) * bezier.B(u[..., 1]) * bezier.B(u[..., 2])) @ f
        np.testing.assert_allclose(got, brute, atol=1e-9)


def test_evaluate_interpolated_gradient_and_mask_brute_force():
    """Gradient: central differences of the all-voxel sum (per finest-voxel unit of the query).
    Mask: some voxel centre within 1.5 strides of the query's cell centre on every axis."""
    rng = np.random.default_rng(11)
    vs = 1 / 16
    xyz = rng.uniform(0.3, 0.7, (60, 3)).astype(np.float32)
    h = hierarchy.Hierarchy(2, vs)
    for k in range(2):
        h.set_coords(k, hierarchy.quantize_coords(xyz, vs, h.strides[k]))
    q = rng.uniform(0.1, 0.9, (200, 3)).astype(np.float32)
    for d in range(2):
        f = rng.standard_normal(h.num_voxels(d))
        s = h.strides[d]
        cen = h.coords[d] + s / 2

        def brute(qq):
            u = (qq[:, None, :] - cen[None]) / s
            return (bezier.B(u[..., 0]) * bezier.B(u[..., 1]) * bezier.B(u[..., 2])) @ f

        g, m = h.evaluate_interpolated(q, d, f, compute_mask=True, compute_grad=True)
        qq = q.astype(np.float64) / vs
        eps = 1e-5
        for a in range(3):
            e = np.zeros(3)
            e[a] = eps
            fd = (brute(qq + e) - brute(qq - e)) / (2 * eps)
            np.testing.assert_allclose(g[:, a], fd, atol=1e-5)
        cell = np.floor(qq / s) * s + s / 2
        near = np.all(np.abs(cell[:, None, :] - cen[None]) <= 1.01 * s, axis=2).any(1)
        assert np.array_equal(m, near) and 0 < m.sum() < m.size
        v, m2 = h.evaluate_interpolated(q, d, f, compute_mask=True)
        np.testing.assert_allclose(v, brute(qq), atol=1e-9)
        assert np.array_equal(m, m2)


# ---------------------------------------------------------------------------------- assembly
def test_screening_term_duplicate_and_absent_pairs():
    """Hand-computed: point 0 has a(0) = 1, a(1) = 2, b(0) = 5; point 1 has a(1) = 3, b(1) = 7;
    weight 2. Sums: S[0,0] = 10, S[1,0] = 20, S[1,1] = 42, S[0,1] = 0. A pair listed twice gets
    its sum at the first occurrence only; a pair with no points gets 0; a nonzero pair missing
    from the COO is dropped."""
    sda = (np.array([0, 0, 1]), np.array([0, 1, 1]), np.array([1.0, 2.0, 3.0]))
    sdb = (np.array([0, 1]), np.array([0, 1]), np.array([5.0, 7.0]))
    src = np.array([1, 0, 1, 0])
    tgt = np.array([0, 1, 0, 0])
    got = assembly.screening_term(sda, sdb, src, tgt, 2, 2.0)
    np.testing.assert_array_equal(got, [20.0, 0.0, 0.0, 10.0])
    got = assembly.screening_term(sda, sdb, np.array([1, 1]), np.array([1, 0]), 2, np.array([2.0, 0.5]))
    np.testing.assert_array_equal(got, [10.5, 20.0])


def test_screening_term_against_dense_brute_force():
    rng = np.random.default_rng(9)
    xyz = rng.uniform(0.4, 0.6, (30, 3)).astype(np.float32)
    h = hierarchy.Hierarchy(2, 1 / 16)
    for k in range(2):
        h.set_coords(k, hierarchy.quantize_coords(xyz, 1 / 16, h.strides[k]))
    alpha = rng.uniform(1, 2, 30)
    for d, dd in ((0, 0), (0, 1), (1, 1)):
        sda = h.point_basis(xyz, d)[:3]
        sdb = h.point_basis(xyz, dd)[:3]
        s, t, _, _ = h.self_neighbours(d, dd, 5)
        got = assembly.screening_term(sda, sdb, s, t, h.num_voxels(dd), alpha)
        qq = xyz.astype(np.float64) * 16

        def bvals(k):
            u = (qq[:, None, :] - (h.coords[k] + h.strides[k] / 2)[None]) / h.strides[k]
            return bezier.B(u[..., 0]) * bezier.B(u[..., 1]) * bezier.B(u[..., 2])
        dense = (bvals(d) * alpha[:, None]).T @ bvals(dd)
        np.testing.assert_allclose(got, dense[s, t], atol=1e-12)
        # every nonzero of the dense product is inside the COO (no silent misses here)
        assert np.isclose(np.abs(dense).sum(), np.abs(got).sum())


def test_lhs_is_symmetric_positive_definite_small_scene():
    from ngspsr.synthetic import sphere
    xyz, nrm = sphere(400, radius=0.2, seed=1)
    h, xd, rho = assembly.build_tree(xyz, 1 / 16, 2, 8.0)
    nd = assembly.splat_normals(h, xyz, nrm, xd, 1.0 / rho)
    res = assembly.solve_multigrid(h, nd, xyz, 4.0, 0.1)
    for d, a in res.lhs.items():
        dense = a.toarray()
        np.testing.assert_allclose(dense, dense.T, atol=1e-12 * np.abs(dense).max())
        assert np.linalg.eigvalsh(dense).min() > 0
        # the stiffness part of every full interior row sums to zero
        s, t, rel, _ = h.self_neighbours(d, d, 5)
        k = bezier.integrate_deriv_deriv_product(rel, h.strides[d], h.strides[d])
        rows = np.bincount(s, minlength=h.num_voxels(d))
        full = rows == 125
        if full.any():
            np.testing.assert_allclose(np.bincount(s, weights=k, minlength=len(rows))[full], 0, atol=1e-12)


def test_rhs_divergence_identity_constant_field():
    """For a constant field on a data voxel whose target neighbourhood is complete, the RHS
    contributions of that voxel sum to zero (sum of Q_m is zero)."""
    for m in (1, 2, 4):
        rel = np.array([[k * m + o + (m - 1) / 2 for k, o in zip(kk, (0, 0, 0))] for kk in itertools.product((-1, 0, 1), repeat=3)],
                       dtype=np.float64)
        # all m^3 sub-positions of the data voxel inside the target cell
        tot = 0.0
        V = np.array([[0.3, -1.2, 0.7]])
        for off in itertools.product(range(m), repeat=3):
            r = rel - np.array(off)[None]
            tot += bezier.integrate_const_deriv_product(np.repeat(V, r.shape[0], 0), r, 1, m).sum()
        assert abs(tot) < 1e-12


# -------------------------------------------------------------------------------------- solver
def test_pcg_against_scipy_and_dense():
    import scipy.sparse.linalg as spla
    rng = np.random.default_rng(1)
    n = 400
    A = sp.random(n, n, density=0.02, random_state=2)
    A = (A + A.T + sp.eye(n) * 6).tocsr()
    b = rng.standard_normal(n)
    x, it = pcg.solve_pcg(A, b, tol=1e-12, max_iter=1000)
    xd = np.linalg.solve(A.toarray(), b)
    np.testing.assert_allclose(x, xd, atol=1e-9)
    xs, _ = spla.cg(A, b, rtol=1e-12, M=sp.diags(1 / A.diagonal()))
    np.testing.assert_allclose(x, xs, atol=1e-8)
    # stop rule: recurrence residual below tol * ||b|| after `it` iterations, not before
    x1, it1 = pcg.solve_pcg(A, b, tol=1e-6)
    x0, _ = pcg.solve_pcg(A, b, tol=1e-6, max_iter=it1 - 1)
    assert pcg.relative_residual(A, x1, b) < 2e-6 < pcg.relative_residual(A, x0, b) * 10
    # FP32 variant converges to its tolerance
    x32, _ = pcg.solve_pcg(A, b, tol=1e-5, dtype=np.float32)
    assert pcg.relative_residual(A, x32, b) < 2e-5


# ---------------------------------------------------------------------------------------- mesh
def test_mc_table_consistency():
    for c in range(256):
        e = mc_tables.case_edges(c)
        assert len(e) % 3 == 0 and len(e) <= 15
        # every triangle edge crosses a sign change of the case
        for ed in e:
            a, b = mc_tables.EDGE_CORNERS[ed]
            assert ((c >> a) & 1) != ((c >> b) & 1)
        # and every sign-changing edge is used
        for ed, (a, b) in enumerate(mc_tables.EDGE_CORNERS):
            if ((c >> a) & 1) != ((c >> b) & 1):
                assert ed in e
    assert mc_tables.num_verts(0) == 0 and mc_tables.num_verts(255) == 0


def test_dual_slots_match_mc_corner_geometry():
    """On a uniform single-level grid every interior dual centre collects its 8 surrounding cells
    at the MC corner positions (corner k at (k&1 ^ (k>>1)&1, (k>>1)&1, (k>>2)&1))."""
    g = np.stack(np.meshgrid(*([np.arange(4)] * 3), indexing="ij"), -1).reshape(-1, 3)
    h = hierarchy.Hierarchy(1, 1.0)
    h.set_coords(0, g)
    primal = {0: g}
    cen = mesh.dual_centres(h, primal, 1)
    acc, cnt = mesh.accumulate_dual_corners(h, primal, 1, cen)
    full = np.nonzero(cnt == 8)[0]
    assert full.shape[0] == 27  # interior lattice points of a 4^3 block
    for i in full:
        for k in range(8):
            corner = np.array([(k & 1) ^ ((k >> 1) & 1), (k >> 1) & 1, (k >> 2) & 1])
            cell = g[acc[i, k, 0]]
            assert np.array_equal(cell, cen[i] - 1 + corner)


def _sphere_mesh(n_cells=10, r=3.3):
    g = np.stack(np.meshgrid(*([np.arange(n_cells)] * 3), indexing="ij"), -1).reshape(-1, 3)
    h = hierarchy.Hierarchy(1, 1.0)
    h.set_coords(0, g)
    c0 = n_cells / 2.0

    def chi(p):
        return r - np.linalg.norm(p - c0, axis=1)
    v, f, info = mesh.extract_multiscale_mesh(h, None, None, None, n_upsample=1, chi_fn=chi)
    return v, f


def test_dual_mc_closed_surface_is_watertight_sphere():
    v, f = _sphere_mesh()
    assert f.shape[0] > 0
    e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(e, axis=0, return_counts=True)
    assert np.all(cnt == 2)  # manifold and closed
    V, E, F = v.shape[0], cnt.shape[0], f.shape[0]
    assert V - E + F == 2  # sphere
    d = np.abs(np.linalg.norm(v - 5.0, axis=1) - 3.3)
    assert d.max() < 0.2


def test_oracle_end_to_end_small_sphere():
    from ngspsr.synthetic import sphere
    from oracle import reconstruct
    xyz, nrm = sphere(2000, radius=0.3, seed=2)
    out = reconstruct.reconstruct(xyz, nrm, 1 / 32, depth=3, min_density=8.0)
    r = np.linalg.norm(out.v - 0.5, axis=1)
    assert abs(np.median(r) - 0.3) < 1 / 64
    e = np.sort(np.concatenate([out.f[:, [0, 1]], out.f[:, [1, 2]], out.f[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(e, axis=0, return_counts=True)
    # the multi-level dual MC drops incomplete dual cubes, so only most edges are manifold
    assert np.mean(cnt == 2) > 0.95


def test_adjoint_gradient_matches_central_differences():
    """The dense-adjoint reference gradient of the multi-level solve equals FP64 central differences
    in per-point screening weights and per-basis normal-term scales."""
    from ngspsr.synthetic import torus
    from oracle import adjoint
    xyz, nrm = torus(600, seed=3)
    h, xd, rho = assembly.build_tree(xyz, 1 / 16, 3, 8.0)
    nd = assembly.splat_normals(h, xyz, nrm, xd, 1.0 / rho)
    bs = adjoint.BlockSystem(h, nd, xyz, delta=0.1)
    rng = np.random.default_rng(0)
    alpha = 4.0 * (1 + 0.3 * rng.uniform(-1, 1, xyz.shape[0]))
    scale = 1 + 0.3 * rng.uniform(-1, 1, bs.N)
    query = (xyz[rng.choice(xyz.shape[0], 200)] + rng.normal(0, 0.02, (200, 3))).astype(np.float32)
    target = rng.normal(0, 0.1, 200)
    ga, gs, _ = bs.gradient(alpha, scale, query, target)
    # with unit scales the block solve reproduces the coarse-to-fine multigrid of reconstructor.py
    import scipy.sparse.linalg as spla
    X = bs.solve(alpha, np.ones(bs.N))
    ref = assembly.solve_multigrid(h, nd, xyz, alpha, 0.1, lambda a, b: (spla.spsolve(a.tocsc(), b), 0))
    for d in range(h.depth):
        x_ml = ref.solutions[d]
        np.testing.assert_allclose(X[bs.off[d]:bs.off[d + 1]], x_ml, rtol=1e-8, atol=1e-10 * np.abs(x_ml).max())
    for p in rng.choice(xyz.shape[0], 4, replace=False):
        hstep = 1e-4 * alpha[p]
        ap, am = alpha.copy(), alpha.copy()
        ap[p] += hstep
        am[p] -= hstep
        fd = (bs.loss(ap, scale, query, target)[0] - bs.loss(am, scale, query, target)[0]) / (2 * hstep)
        assert abs(fd - ga[p]) <= 1e-6 * max(abs(fd), 1e-12) + 1e-12, (p, fd, ga[p])
    for i in rng.choice(bs.N, 4, replace=False):
        hstep = 1e-4
        sp_, sm_ = scale.copy(), scale.copy()
        sp_[i] += hstep
        sm_[i] -= hstep
        fd = (bs.loss(alpha, sp_, query, target)[0] - bs.loss(alpha, sm_, query, target)[0]) / (2 * hstep)
        assert abs(fd - gs[i]) <= 1e-6 * max(abs(fd), 1e-12) + 1e-13, (i, fd, gs[i])


def test_direct_solve_closed_form_and_singularity():
    """oracle.direct: the inverse of the 1-D Laplacian tridiag(-1, 2, -1) of size n is
    (T^-1)_ij = min(i, j) (n + 1 - max(i, j)) / (n + 1) (1-based); the upper triangle is ignored;
    the Neumann Laplacian (constant null space) reports its last pivot as singular."""
    from oracle import direct
    n = 30
    T = 2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    i = np.arange(1, n + 1)
    inv = np.minimum.outer(i, i) * (n + 1 - np.maximum.outer(i, i)) / (n + 1)
    rng = np.random.default_rng(0)
    b = rng.standard_normal(n)
    x, sing = direct.solve_spd(sp.csr_matrix(T), b)
    assert sing == -1
    np.testing.assert_allclose(x, inv @ b, rtol=1e-12, atol=1e-12)
    garbage_upper = T + np.triu(rng.standard_normal((n, n)), 1)
    x2, _ = direct.solve_spd(garbage_upper, b)
    np.testing.assert_allclose(x2, x, rtol=1e-12, atol=1e-12)
    N = T.copy()
    N[0, 0] = N[-1, -1] = 1.0
    _, sing = direct.solve_spd(N, b)
    assert sing == n - 1
