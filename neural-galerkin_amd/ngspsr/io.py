"""Point-cloud and mesh files of the reconstruct path: `.bnpts` input (examples/main.py:27-28, raw
float64 rows x y z nx ny nz) and binary little-endian PLY output, plus the mesh object
extract_mesh returns when Open3D is absent."""
import numpy as np


def read_bnpts(path):
    """(xyz (N, 3), normal (N, 3)) float64 from a raw float64 N x 6 file."""
    a = np.fromfile(path, dtype="<f8")
    if a.size % 6:
        raise ValueError(f"{path}: size is not a multiple of 6 float64 values")
    a = a.reshape(-1, 6)
    return a[:, :3].copy(), a[:, 3:].copy()


def write_bnpts(path, xyz, normal):
    np.concatenate([np.asarray(xyz, "<f8").reshape(-1, 3), np.asarray(normal, "<f8").reshape(-1, 3)], 1).tofile(path)


def _np(a):
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.asarray(a)


def write_ply(path, vertices, triangles, normals=None):
    """Binary little-endian PLY: float x y z (+ nx ny nz) per vertex, uchar-counted int32 faces."""
    v = _np(vertices).astype("<f4").reshape(-1, 3)
    f = _np(triangles).astype("<i4").reshape(-1, 3)
    n = None if normals is None else _np(normals).astype("<f4").reshape(-1, 3)
    head = ["ply", "format binary_little_endian 1.0", f"element vertex {v.shape[0]}",
            "property float x", "property float y", "property float z"]
    if n is not None:
        head += ["property float nx", "property float ny", "property float nz"]
    head += [f"element face {f.shape[0]}", "property list uchar int vertex_indices", "end_header"]
    vrec = v if n is None else np.concatenate([v, n], 1)
    frec = np.zeros(f.shape[0], dtype=[("n", "u1"), ("i", "<i4", (3,))])
    frec["n"] = 3
    frec["i"] = f
    with open(path, "wb") as fh:
        fh.write(("\n".join(head) + "\n").encode("ascii"))
        fh.write(np.ascontiguousarray(vrec, "<f4").tobytes())
        fh.write(frec.tobytes())


def read_ply(path):
    """(vertices (V, 3), triangles (T, 3), normals (V, 3) or None) from a file write_ply wrote."""
    with open(path, "rb") as fh:
        data = fh.read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    head = data[:end].decode("ascii").split("\n")
    if "format binary_little_endian 1.0" not in head:
        raise ValueError(f"{path}: only binary little-endian PLY is read")
    nv = int(next(l for l in head if l.startswith("element vertex")).split()[-1])
    nf = int(next(l for l in head if l.startswith("element face")).split()[-1])
    has_n = "property float nx" in head
    width = 6 if has_n else 3
    v = np.frombuffer(data, "<f4", nv * width, end).reshape(nv, width)
    frec = np.frombuffer(data, [("n", "u1"), ("i", "<i4", (3,))], nf, end + 4 * nv * width)
    if nf and not np.all(frec["n"] == 3):
        raise ValueError(f"{path}: only triangle faces are read")
    return v[:, :3].copy(), frec["i"].astype(np.int64), (v[:, 3:].copy() if has_n else None)


class TriangleMesh:
    """Mesh with Open3D's TriangleMesh attribute names (vertices, triangles, vertex_normals as numpy
    arrays)."""

    def __init__(self, vertices, triangles, normals=None):
        self.vertices = _np(vertices).astype(np.float64)
        self.triangles = _np(triangles).astype(np.int32)
        self.vertex_normals = None if normals is None else _np(normals).astype(np.float64)

    def write_ply(self, path):
        write_ply(path, self.vertices, self.triangles, self.vertex_normals)


def triangle_mesh(vertices, triangles, normals=None):
    """Open3D TriangleMesh when open3d is importable (reconstructor.py:279-291), else TriangleMesh."""
    try:
        import open3d as o3d
    except ImportError:
        return TriangleMesh(vertices, triangles, normals)
    m = o3d.geometry.TriangleMesh()
    m.vertices = o3d.utility.Vector3dVector(_np(vertices).astype(float))
    m.triangles = o3d.utility.Vector3iVector(_np(triangles).astype(np.int32))
    if normals is not None:
        m.vertex_normals = o3d.utility.Vector3dVector(_np(normals).astype(float))
    return m
