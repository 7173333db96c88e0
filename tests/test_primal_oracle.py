"""Primal sparse marching cubes oracle (primal_mc.cu, ops.py:9-101) pinned by closed forms and
mesh invariants; .bnpts / PLY round trips. CPU only."""
import numpy as np
import pytest

from oracle import primal_mc as pm


def _block(n, r):
    base = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    return base, pm.sample_positions(base, r)


@pytest.mark.parametrize("r", [2, 3, 5])
def test_sphere_is_closed_oriented_manifold(r):
    base, p = _block(4, r)
    c, rad = np.array([2.0, 2.1, 1.9]), 1.3
    sdf = (rad - np.linalg.norm(p - c, axis=1)).astype(np.float32).reshape(-1, r + 1, r + 1, r + 1)
    v, t, n, unq, vsrc, vw = pm.marching_cubes(base, sdf)
    # every undirected edge in exactly two triangles, every directed edge once: closed, oriented
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    _, cnt = np.unique(np.sort(e, 1), axis=0, return_counts=True)
    assert set(cnt.tolist()) == {2}
    assert np.unique(e, axis=0).shape[0] == e.shape[0]
    assert v.shape[0] - cnt.shape[0] + t.shape[0] == 2  # Euler characteristic of a sphere
    # vertices near the sphere (linear interpolation of a smooth field), normals outward: toward
    # negative values of a field that is positive inside
    d = np.linalg.norm(v - c, axis=1)
    assert np.abs(d - rad).max() < 0.2 / r ** 2  # chord error of a linear edge fit, about h^2 / (8 R)
    out = (v - c) / d[:, None]
    assert np.sum(n * out, 1).min() > 0.95


def test_plane_vertices_sources_and_edge_ids():
    """Linear field s = 0.37 + x - 0.8 y + 0.3 z: vertices on the plane, each vertex the w-blend of its
    two source samples, which straddle the plane; edge ids follow from the source samples."""
    r = 3
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [3, 2, 1], [-1, 0, 0]], np.float32)
    p = pm.sample_positions(base, r)
    a = np.array([1.0, -0.8, 0.3])
    sdf = (0.37 + p @ a).astype(np.float32).reshape(-1, r + 1, r + 1, r + 1)
    tris, vid, src, w = pm.marching_cubes_sparse(base, sdf)
    assert tris.shape[0] > 0
    q = tris.reshape(-1, 3).astype(np.float64)
    assert np.abs(0.37 + q @ a).max() < 1e-5
    s = src.reshape(-1, 2)
    ww = w.reshape(-1).astype(np.float64)
    blend = ww[:, None] * p[s[:, 0]] + (1 - ww[:, None]) * p[s[:, 1]]
    assert np.abs(blend - q).max() < 1e-5
    f = sdf.reshape(-1)
    assert np.all((f[s[:, 0]] < 0) != (f[s[:, 1]] < 0))
    # sample index -> global sub-lattice point: int(base) * r + local offset
    r1 = r + 1
    loc = np.stack([(s // (r1 * r1)) % r1, (s // r1) % r1, s % r1], -1)
    lat = base[s // r1 ** 3].astype(np.int64) * r + loc
    lo = np.minimum(lat[:, 0], lat[:, 1])
    axis = np.argmax(lat[:, 0] != lat[:, 1], axis=1)
    assert np.all(np.abs(lat[:, 0] - lat[:, 1]).sum(1) == 1)
    assert np.array_equal(vid.reshape(-1, 4), np.concatenate([lo, axis[:, None]], 1))


def test_dedup_rows_and_first_occurrence():
    rng = np.random.default_rng(4)
    r = 2
    base = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [5, 5, 5]], np.float32)
    sdf = rng.standard_normal((4, r + 1, r + 1, r + 1)).astype(np.float32)
    tris, vid, src, w = pm.marching_cubes_sparse(base, sdf)
    v, t, n, unq, vsrc, vw = pm.marching_cubes(base, sdf)
    flat = vid.reshape(-1, 4)
    assert np.array_equal(unq[t.reshape(-1)], flat)
    assert len({tuple(x) for x in unq.tolist()}) == unq.shape[0]
    assert [tuple(x) for x in unq.tolist()] == sorted(tuple(x) for x in unq.tolist())
    first = [int(np.nonzero((flat == u).all(1))[0][0]) for u in unq]
    assert np.array_equal(v, tris.reshape(-1, 3)[first]) and np.array_equal(vw, w.reshape(-1)[first])


def test_meshsdf_gradient_of_a_uniform_shift():
    """For a planar field of unit slope, raising every sample by d moves every vertex by d along the
    outward normal; the MeshSDF gradient of L = sum of vertex x (bandwidth 1) summed over all samples
    equals the finite-difference dL/dd."""
    r = 4
    base, p = _block(2, r)
    sdf0 = (0.83 - p[:, 0]).astype(np.float64)  # positive inside (x < 0.83), normal +x

    def loss(dlt):
        s = (sdf0 + dlt).astype(np.float32).reshape(-1, r + 1, r + 1, r + 1)
        return pm.marching_cubes(base, s)[0][:, 0].astype(np.float64).sum()

    v, t, n, unq, vsrc, vw = pm.marching_cubes(base, sdf0.astype(np.float32).reshape(-1, r + 1, r + 1, r + 1))
    gv = np.zeros_like(v, dtype=np.float64)
    gv[:, 0] = 1.0
    g = pm.marching_cubes_backward(gv, n, vsrc, vw, sdf0.size)
    fd = (loss(1e-3) - loss(-1e-3)) / 2e-3
    assert abs(g.sum() - fd) < 1e-3 * v.shape[0]
    assert np.allclose(n, [1, 0, 0], atol=1e-4)  # unit up to the +1e-6 of the normalisation


def test_io_round_trips(tmp_path):
    from ngspsr import io
    rng = np.random.default_rng(0)
    xyz, nrm = rng.standard_normal((50, 3)), rng.standard_normal((50, 3))
    io.write_bnpts(tmp_path / "a.bnpts", xyz, nrm)
    x2, n2 = io.read_bnpts(tmp_path / "a.bnpts")
    assert np.array_equal(x2, xyz) and np.array_equal(n2, nrm)
    assert (tmp_path / "a.bnpts").stat().st_size == 50 * 6 * 8
    v = rng.standard_normal((20, 3)).astype(np.float32)
    f = rng.integers(0, 20, (30, 3))
    nn = rng.standard_normal((20, 3)).astype(np.float32)
    for normals in (None, nn):
        io.write_ply(tmp_path / "m.ply", v, f, normals)
        v2, f2, n3 = io.read_ply(tmp_path / "m.ply")
        assert np.array_equal(v2, v) and np.array_equal(f2, f)
        assert (n3 is None) if normals is None else np.array_equal(n3, nn)
    m = io.TriangleMesh(v, f, nn)
    assert m.vertices.shape == (20, 3) and m.triangles.dtype == np.int32
