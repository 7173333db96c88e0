"""Primal sparse marching cubes with the MeshSDF backward (oracle; test infrastructure only).

Restates csrc/marching_cubes/primal_mc.cu:7-234 (classification, triangle emission, sdf_interp with
the weight of the first edge end), jspsr/core/ops.py:9-58 (vertex de-duplication and the MeshSDF
backward) and ops.py:85-101 (compute_normal). Emission follows the kernel's float32 operation order
with no fused multiply-add. One choice the reference leaves open is fixed here: a vertex shared by
several triangles takes the position, source samples and weight of its first occurrence in triangle
order (ops.py:20-29 keeps an arbitrary one).
"""
import numpy as np

from .mc_tables import EDGE_CORNERS, case_edges

# corner k of a cube (Bourke numbering), as offsets (x, y, z)
CORNER_OFFSETS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                           [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.int64)
_EA = np.array([a for a, _ in EDGE_CORNERS])
_EB = np.array([b for _, b in EDGE_CORNERS])


def _interp_w(p1, p2, v1, v2):
    """primal_mc.cu:7-19 vectorised in float32: (position, weight of p1)."""
    f = np.float32
    with np.errstate(divide="ignore", invalid="ignore"):
        w2 = (f(0) - v1) / (v2 - v1)
    w1 = (f(1) - w2).astype(np.float32)
    out = (p1 * w1[:, None] + p2 * w2[:, None]).astype(np.float32)
    s1 = np.abs(f(0) - v1) < f(1e-5)
    s2 = ~s1 & (np.abs(f(0) - v2) < f(1e-5))
    s3 = ~s1 & ~s2 & (np.abs(v1 - v2) < f(1e-5))
    out[s1 | s3] = p1[s1 | s3]
    out[s2] = p2[s2]
    w1 = w1.copy()
    w1[s1 | s3] = 1.0
    w1[s2] = 0.0
    return out, w1


def marching_cubes_sparse(base, sdf):
    """marching_cubes_sparse_cuda (primal_mc.cu:176-234).

    :param base: (L, 3) integer-valued leaf corners (leaf units)
    :param sdf: (L, r+1, r+1, r+1) samples
    :return: tris (T, 3, 3) float32, vert_ids (T, 3, 4) int32, src_ids (T, 3, 2) int32,
        src_weight (T, 3, 1) float32; triangles in (leaf, rx, ry, rz) then case-table order.
    """
    base = np.asarray(base, np.float32).reshape(-1, 3)
    sdf = np.asarray(sdf, np.float32)
    nl = base.shape[0]
    r = sdf.shape[1] - 1
    r1 = r + 1
    if nl == 0:
        return (np.zeros((0, 3, 3), np.float32), np.zeros((0, 3, 4), np.int32), np.zeros((0, 3, 2), np.int32),
                np.zeros((0, 3, 1), np.float32))
    g = np.meshgrid(np.arange(nl), np.arange(r), np.arange(r), np.arange(r), indexing="ij")
    li, rx, ry, rz = [a.reshape(-1) for a in g]
    rr = np.stack([rx, ry, rz], 1)
    vals = np.stack([sdf[li, rx + o[0], ry + o[1], rz + o[2]] for o in CORNER_OFFSETS], 1)
    case = ((vals < 0) * (1 << np.arange(8))).sum(1)
    cube_of, edge_of = [], []
    for i in np.nonzero([len(case_edges(int(c))) for c in case])[0]:
        es = case_edges(int(case[i]))
        cube_of += [i] * len(es)
        edge_of += es
    cube_of = np.asarray(cube_of, np.int64)
    edge_of = np.asarray(edge_of, np.int64)
    ca, cb = _EA[edge_of], _EB[edge_of]
    oa, ob = CORNER_OFFSETS[ca], CORNER_OFFSETS[cb]
    b = base[li[cube_of]]
    ri = rr[cube_of]
    sbs = np.float32(1.0) / np.float32(r)
    p1 = (b + (ri + oa).astype(np.float32) * sbs).astype(np.float32)
    p2 = (b + (ri + ob).astype(np.float32) * sbs).astype(np.float32)
    pos, w1 = _interp_w(p1, p2, vals[cube_of, ca], vals[cube_of, cb])
    vid = np.empty((cube_of.shape[0], 4), np.int64)
    vid[:, :3] = np.trunc(b).astype(np.int64) * r + ri + np.minimum(oa, ob)
    vid[:, 3] = np.argmax(oa != ob, axis=1)
    sid0 = r1 ** 3
    lc = li[cube_of].astype(np.int64)
    ea, eb = ri + oa, ri + ob
    src = np.stack([sid0 * lc + (ea[:, 0] * r1 + ea[:, 1]) * r1 + ea[:, 2],
                    sid0 * lc + (eb[:, 0] * r1 + eb[:, 1]) * r1 + eb[:, 2]], 1)
    t = cube_of.shape[0] // 3
    return (pos.reshape(t, 3, 3), vid.astype(np.int32).reshape(t, 3, 4), src.astype(np.int32).reshape(t, 3, 2),
            w1.reshape(t, 3, 1))


def compute_normal(vertices, triangles):
    """ops.py:85-101 in float64: area-weighted vertex normals, divided by length + 1e-6."""
    v = np.asarray(vertices, np.float64)
    t = np.asarray(triangles, np.int64)
    n = np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]])
    acc = np.zeros_like(v)
    for k in range(3):
        np.add.at(acc, t[:, k], n)
    return acc / (np.linalg.norm(acc, axis=1, keepdims=True) + 1e-6)


def marching_cubes(base, sdf):
    """MarchingCubes.forward (ops.py:15-38): vertices (V, 3), triangles (T, 3), normals (V, 3),
    unq (V, 4), plus each vertex's source samples (V, 2) and first-end weight (V,) for the backward."""
    tris, vid, src, w = marching_cubes_sparse(base, sdf)
    flat = vid.reshape(-1, 4)
    if flat.shape[0] == 0:
        return (np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int64), np.zeros((0, 3)), np.zeros((0, 4), np.int32),
                np.zeros((0, 2), np.int32), np.zeros(0, np.float32))
    unq, first, inv = np.unique(flat, axis=0, return_index=True, return_inverse=True)
    vertices = tris.reshape(-1, 3)[first]
    triangles = inv.reshape(-1, 3).astype(np.int64)
    normals = compute_normal(vertices, triangles)
    return vertices, triangles, normals, unq, src.reshape(-1, 2)[first], w.reshape(-1)[first]


def marching_cubes_backward(grad_vert, normals, vsrc, vw, n_samples, bandwidth=1.0):
    """MarchingCubes.backward (ops.py:40-58) in float64: gradient of the flattened samples."""
    g = bandwidth * np.sum(np.asarray(grad_vert, np.float64) * np.asarray(normals, np.float64), axis=1)
    out = np.zeros(n_samples)
    w = np.asarray(vw, np.float64)
    np.add.at(out, np.asarray(vsrc)[:, 0], w * g)
    np.add.at(out, np.asarray(vsrc)[:, 1], (1 - w) * g)
    return out


def sample_positions(base, r):
    """Position (leaf units) of every flat sample index of an (L, r+1, r+1, r+1) block."""
    base = np.asarray(base, np.float64).reshape(-1, 3)
    r1 = r + 1
    g = np.stack(np.meshgrid(np.arange(r1), np.arange(r1), np.arange(r1), indexing="ij"), -1).reshape(-1, 3)
    return (base[:, None, :] + g[None] / r).reshape(-1, 3)
