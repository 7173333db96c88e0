"""Pins the oracle's tree builders, voxel status, trilinear interpolation and mesh leaves.

Each check states the reference's rule another way:
- build_dense (_hash_backend.py:358-378): level 0 is the set of cells floor(q) (dilated by the
  range cube) or the trilinear corners; every coarser level is exactly the set of parents of the
  level below; brute force with Python sets.
- voxel_status (_hash_backend.py:292-318, written there with the 8 conformal children): 2 exactly
  when the voxel is the parent of some voxel one level finer, by a set of parents.
- split_data (_hash_backend.py:320-330): trilinear interpolation reproduces functions that are
  linear in the voxel centre wherever all 8 corners exist.
- build_subdivide (_hash_backend.py:380-412): policies with a known outcome (everything refines,
  nothing refines, refine where x > t) fix every point's level.
- leaves (reconstructor.py:309-336): conformal leaves never overlap and cover the (dilated)
  finest voxels; trimmed leaves are the voxels that are nobody's parent.
"""
import numpy as np
import pytest

from oracle import hierarchy as H
from oracle import mesh as OM


def _cloud(n=600, seed=0):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return (0.5 + 0.3 * d).astype(np.float32)


def _set(c):
    return {tuple(int(v) for v in r) for r in np.asarray(c)}


@pytest.mark.parametrize("expand_range", [0, 1, 2, 3])
def test_build_dense_levels(expand_range):
    vs, depth = 1 / 32, 4
    xyz = _cloud()
    h = H.Hierarchy(depth, vs)
    lv = h.build_dense(xyz, expand_range)
    assert np.all(lv == 0)
    q = xyz.astype(np.float32) * np.float32(32.0)
    if expand_range == 2:
        want = set()
        for p in q.astype(np.float64):
            base = np.floor(p)
            sg = np.sign(p - base - 0.5)
            for o in H.CONFORM_OFFSETS:
                want.add(tuple(int(v) for v in base + o * sg))
    else:
        cells = {tuple(int(v) for v in np.floor(p)) for p in q}
        r = expand_range // 2
        want = {(a + dx, b + dy, c + dz) for a, b, c in cells for dx in range(-r, r + 1) for dy in range(-r, r + 1)
                for dz in range(-r, r + 1)}
    assert _set(h.coords[0]) == want
    for d in range(1, depth):
        s = 2 ** d
        par = {tuple((v // s) * s for v in c) for c in _set(h.coords[d - 1])}
        assert _set(h.coords[d]) == par
        keys = H.pack_keys(h.coords[d])
        assert np.all(np.diff(keys) > 0)


def test_voxel_status_by_parent_sets():
    vs, depth = 1 / 32, 4
    xyz = _cloud(400, seed=1)
    h = H.Hierarchy(depth, vs)
    rng = np.random.default_rng(2)
    lv = h.build_adaptive(xyz, rng.uniform(0.05, 20.0, xyz.shape[0]), 4.0)  # points spread over the levels
    assert len(set(lv.tolist())) == depth
    for d in range(depth):
        s = 2 ** d
        present = _set(h.coords[d])
        parents = {tuple((v // s) * s for v in c) for c in _set(h.coords[d - 1])} if d > 0 else set()
        probe = np.concatenate([h.coords[d], rng.integers(0, 40, size=(200, 3)) * s,
                                rng.integers(0, 80, size=(50, 3))], 0)
        st = h.voxel_status(probe, d)
        assert np.any(st == 2) == (d > 0) and np.any(st == 1) and np.any(st == 0)
        for c, v in zip([tuple(int(x) for x in r) for r in probe], st):
            want = 0 if c not in present else (2 if c in parents else 1)
            assert v == want, (d, c, v, want)


def test_split_data_reproduces_linear_fields():
    vs = 1 / 16
    h = H.Hierarchy(2, vs)
    ax = np.arange(16)
    g0 = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    h.set_coords(0, g0)
    h.set_coords(1, np.unique((g0 // 2) * 2, axis=0))
    rng = np.random.default_rng(4)
    xyz = (rng.uniform(2.0, 14.0, size=(300, 3)) * vs).astype(np.float32)
    q = xyz.astype(np.float64) / vs
    for d in (0, 1):
        s = 2 ** d
        cen = h.coords[d].astype(np.float64) + s / 2.0
        data = np.stack([0.5 + cen[:, 0] - 2 * cen[:, 1], 3.0 * cen[:, 2], np.ones(cen.shape[0])], 1)
        out = h.split_data(xyz, d, data)
        want = np.stack([0.5 + q[:, 0] - 2 * q[:, 1], 3.0 * q[:, 2], np.ones(q.shape[0])], 1)
        np.testing.assert_allclose(out, want, rtol=1e-5, atol=1e-5)
    # a missing corner contributes nothing: one voxel removed, the points around it lose its weight
    h.set_coords(0, g0[~np.all(g0 == 8, axis=1)])
    p = np.array([[8.5, 8.5, 8.5]], np.float32) * np.float32(vs)
    assert abs(h.split_data(p, 0, np.ones((h.coords[0].shape[0], 1)))[0, 0]) < 1e-12


def test_build_subdivide_known_policies():
    vs, depth = 1 / 32, 4
    xyz = _cloud(500, seed=3)
    q = xyz.astype(np.float32) * np.float32(32.0)

    def run(policy, **kw):
        h = H.Hierarchy(depth, vs)
        return h, h.build_subdivide(xyz, policy, **kw)

    h, lv = run(lambda pts, inv: np.ones(pts.shape[0], bool))
    assert np.all(lv == 0)
    for d in range(depth):
        s = 2 ** d
        assert _set(h.coords[d]) == {tuple(int(np.floor(v / s)) * s for v in p) for p in q}
    h, lv = run(lambda pts, inv: np.zeros(pts.shape[0], bool))
    assert np.all(lv == depth - 1)
    assert all(h.coords[d].shape[0] == 0 for d in range(depth - 1))
    # refine where x > 16 voxels: those points reach level 0, the others stay on the coarsest level
    h, lv = run(lambda pts, inv, xs: xs > 16.0, xs=q[:, 0].copy())
    assert np.array_equal(lv, np.where(q[:, 0] > 16.0, 0, depth - 1))
    # the inverse map indexes the coarser level's cells: a policy can read per-cell quantities
    seen = []

    def pol(pts, inv):
        seen.append((pts.shape[0], int(inv.max()) + 1))
        return np.ones(pts.shape[0], bool)
    h, _ = run(pol)
    assert [n for n, _ in seen] == [xyz.shape[0]] * (depth - 1)
    assert [k for _, k in seen] == [h.coords[d].shape[0] for d in range(depth - 1, 0, -1)]
    # expand: every level holds the trilinear corners of the points on it or finer
    h, lv = run(lambda pts, inv, xs: xs > 16.0, expand=True, xs=q[:, 0].copy())
    for d in range(depth):
        assert np.array_equal(h.coords[d], H.quantize_coords(xyz[lv <= d], vs, 2 ** d))
    # limit_adaptive_depth clamps the levels
    _, lv = run(lambda pts, inv: np.zeros(pts.shape[0], bool), limit_adaptive_depth=2)
    assert np.all(lv == 1)


def _cells_of(c, s):
    """Finest-unit cells covered by voxels c of side s."""
    out = set()
    for x, y, z in np.asarray(c).tolist():
        for a in range(s):
            for b in range(s):
                for e in range(s):
                    out.add((x + a, y + b, z + e))
    return out


@pytest.mark.parametrize("expand", [0, 3])
def test_conformal_leaves_partition(expand):
    vs, depth = 1 / 16, 3
    xyz = _cloud(150, seed=5)
    h = H.Hierarchy(depth, vs)
    lv = h.build_adaptive(xyz, np.random.default_rng(5).uniform(0.1, 10.0, xyz.shape[0]), 2.0)
    assert len(set(lv.tolist())) == depth
    leaves, md = OM.conformal_leaves(h, expand=expand)
    covered = set()
    total = 0
    for d in range(md):
        cs = _cells_of(leaves[d], 2 ** d)
        total += len(cs)
        covered |= cs
    assert total == len(covered)  # no two leaves overlap
    fine = h.coords[0].astype(np.int64)
    if expand:
        r = expand // 2
        fine = np.unique((fine[:, None, :] + H.range_kernel(expand)[None].astype(np.int64)).reshape(-1, 3), axis=0)
    assert _set(fine) <= covered


def test_trimmed_leaves_are_voxels_without_children():
    vs, depth = 1 / 16, 3
    xyz = _cloud(150, seed=6)
    h = H.Hierarchy(depth, vs)
    lv = h.build_adaptive(xyz, np.random.default_rng(6).uniform(0.1, 10.0, xyz.shape[0]), 2.0)
    assert len(set(lv.tolist())) == depth
    leaves, md = OM.conformal_leaves(h, trim=True)
    assert md == depth
    for d in range(depth):
        s = 2 ** d
        parents = {tuple((v // s) * s for v in c) for c in _set(h.coords[d - 1])} if d > 0 else set()
        assert _set(leaves[d]) == _set(h.coords[d]) - parents
