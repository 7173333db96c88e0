"""Sparse voxel hierarchy of the jspsr path (oracle; test infrastructure only).

Restates jspsr/core/_hash_backend.py and jspsr/core/hashtree.py on the CPU with numpy.
Integer outputs (voxel sets, neighbour indices) depend on FP32 arithmetic of the reference; those
steps are mirrored in float32 with the reference's operation order. Accumulated floating-point
quantities (density, splat) are computed in float64.

Division of a float32 tensor by a Python float is computed the way PyTorch's GPU kernel does it for
a CPU-scalar divisor: multiplication by the float32 reciprocal (q = x * fl32(1/fl32(vs))).
For power-of-two voxel sizes both forms are identical.
"""
import numpy as np

from . import bezier

CONFORM_OFFSETS = np.array([(0, 0, 0), (0, 0, 1), (0, 1, 0), (0, 1, 1),
                            (1, 0, 0), (1, 0, 1), (1, 1, 0), (1, 1, 1)], dtype=np.int64)  # _hash_backend.py:74-75

_OFF = 1 << 20


def pack_keys(coords) -> np.ndarray:
    """Order-preserving int64 key of int coords (lexicographic (x, y, z) order, |c| < 2^20)."""
    c = np.asarray(coords, dtype=np.int64) + _OFF
    return (c[:, 0] << 42) | (c[:, 1] << 21) | c[:, 2]


def unique_rows(coords) -> np.ndarray:
    """torch.unique(dim=0) of int rows: lexicographically sorted unique rows (ops.py:103-127)."""
    c = np.asarray(coords, dtype=np.int64).reshape(-1, 3)
    if c.shape[0] == 0:
        return np.zeros((0, 3), dtype=np.int32)
    return np.unique(c, axis=0).astype(np.int32)


class CoordIndex:
    """Exact coord -> row map for a sorted unique coord array (the logical content of the reference's
    cuckoo table, ext.py:14-84: query returns the row or -1)."""

    def __init__(self, coords):
        self.coords = np.asarray(coords, dtype=np.int32).reshape(-1, 3)
        self.keys = pack_keys(self.coords)
        assert np.all(np.diff(self.keys) > 0), "coords must be sorted and unique"

    def query(self, q) -> np.ndarray:
        q = np.asarray(q, dtype=np.int64).reshape(-1, 3)
        if self.keys.shape[0] == 0 or q.shape[0] == 0:
            return np.full(q.shape[0], -1, dtype=np.int64)
        k = pack_keys(q)
        pos = np.searchsorted(self.keys, k)
        pos_c = np.minimum(pos, self.keys.shape[0] - 1)
        return np.where(self.keys[pos_c] == k, pos_c, -1).astype(np.int64)


def f32_inv(v: float) -> np.float32:
    return np.float32(1.0) / np.float32(v)


def scaled_positions(xyz, voxel_size: float) -> np.ndarray:
    """q = xyz / voxel_size in float32 (_hash_backend.py:472, hashtree.py:169)."""
    return np.asarray(xyz, dtype=np.float32) * f32_inv(voxel_size)


def trilinear_weights(xyz, voxel_size: float, stride: int, data=None):
    """_hash_backend.py:469-510. Returns (alpha_coords (8N,3) int32 offset-major, alpha_data).

    alpha_data is float64 (weights, or weights x data rows).
    """
    q = scaled_positions(xyz, voxel_size)
    s = np.float32(stride)
    d_coords = np.floor(q * (np.float32(1) / s)) * s
    rel = (q - d_coords) - np.float32(stride / 2.0)
    sign = np.sign(rel)
    local = (np.abs(rel.astype(np.float64)) / stride)
    coords, alphas = [], []
    for nx, ny, nz in CONFORM_OFFSETS:
        off = np.stack([nx * sign[:, 0], ny * sign[:, 1], nz * sign[:, 2]], axis=1) * s
        coords.append((d_coords + off).astype(np.int32))
        ax = local[:, 0] if nx == 1 else 1 - local[:, 0]
        ay = local[:, 1] if ny == 1 else 1 - local[:, 1]
        az = local[:, 2] if nz == 1 else 1 - local[:, 2]
        a = ax * ay * az
        if data is None:
            alphas.append(a)
        else:
            d = np.asarray(data, dtype=np.float64)
            alphas.append(a * d if d.ndim == 1 else a[:, None] * d)
    return np.concatenate(coords, 0), np.concatenate(alphas, 0)


def quantize_coords(xyz, voxel_size: float, stride: int) -> np.ndarray:
    """_hash_backend.py:457-462: sorted unique trilinear corners at `stride`."""
    c, _ = trilinear_weights(xyz, voxel_size, stride)
    return unique_rows(c)


def coords_neighbours(index: CoordIndex, target_stride: int, source_coords, source_stride,
                      kernel, conv_based=False, transposed=False, float_source=False):
    """_hash_backend.py:176-228.

    `source_coords` are int voxel coords, or float32 positions when float_source is True (the
    evaluate_interpolated / screening use, hashtree.py:169-172, reconstructor.py:113-117).
    Returns (source_ids, target_ids, rel (E,3) float64, nbsizes) in the reference's order:
    kernel-major then source index, or source-major then kernel when transposed.
    """
    kernel = np.asarray(kernel, dtype=np.int64).reshape(-1, 3)
    st = target_stride
    if not conv_based:
        assert source_stride <= st
        if float_source:
            src = np.asarray(source_coords, dtype=np.float32)
            t = src + np.float32(0.5 * source_stride)
            qc = (np.floor(t * (np.float32(1) / np.float32(st))) * np.float32(st)).astype(np.int64)
            c_off = (qc.astype(np.float32) - src).astype(np.float64) / source_stride + (st // source_stride - 1) / 2.0
        else:
            src = np.asarray(source_coords, dtype=np.int64)
            qc = np.floor_divide(2 * src + source_stride, 2 * st) * st
            c_off = (qc - src) / source_stride + (st // source_stride - 1) / 2.0
    else:
        src = np.asarray(source_coords, dtype=np.int64)
        assert source_stride >= st
        qc = src
        c_off = None
    n = qc.shape[0]
    K = kernel.shape[0]
    res = np.empty((K, n), dtype=np.int64)
    for k in range(K):
        res[k] = index.query(qc + kernel[k] * st)
    if transposed:
        res = res.T
    nbsizes = np.sum(res != -1, axis=1)
    if transposed:
        sid, kid = np.nonzero(res != -1)
        tid = res[sid, kid]
    else:
        kid, sid = np.nonzero(res != -1)
        tid = res[kid, sid]
    nt = kernel[kid].astype(np.float64)
    if not conv_based:
        nt = nt * (st / source_stride) + c_off[sid, :3]
    return sid, tid, nt, nbsizes


def range_kernel(n_range: int) -> np.ndarray:
    """Neighbour offsets of an odd range, concentric-cube prefix order.

    The reference loads this order from a data file absent from the repository
    (hashtree.py:103-116); the code requires only that kernel[:r^3] is the r-cube for every odd
    r (_hash_backend.py:56-63,270-274). Fixed here as: shells of increasing Chebyshev radius,
    lexicographic (x, y, z) order inside each shell.
    """
    assert n_range % 2 == 1
    h = n_range // 2
    g = np.stack(np.meshgrid(*([np.arange(-h, h + 1)] * 3), indexing="ij"), -1).reshape(-1, 3)
    shell = np.max(np.abs(g), axis=1)
    order = np.lexsort((g[:, 2], g[:, 1], g[:, 0], shell))
    return g[order].astype(np.int32)


class Hierarchy:
    """SparseFeatureHierarchy (_hash_backend.py:69-510) restricted to the reconstruct path."""

    def __init__(self, depth: int, voxel_size: float):
        self.depth = depth
        self.voxel_size = voxel_size
        self.strides = [2 ** d for d in range(depth)]
        self.coords = [np.zeros((0, 3), np.int32) for _ in range(depth)]
        self.index = [CoordIndex(self.coords[d]) for d in range(depth)]

    def set_coords(self, d, coords):
        self.coords[d] = np.asarray(coords, dtype=np.int32).reshape(-1, 3)
        self.index[d] = CoordIndex(self.coords[d])

    def num_voxels(self, d):
        return self.coords[d].shape[0]

    # _hash_backend.py:414-444
    def build_adaptive(self, xyz, xyz_density, min_density: float, log_base: float = 4.0):
        xyz_depth = level_of_density(xyz_density, min_density, log_base, self.depth)
        for d in range(self.depth):
            self.set_coords(d, quantize_coords(np.asarray(xyz)[xyz_depth <= d], self.voxel_size, self.strides[d]))
        return xyz_depth

    # _hash_backend.py:358-378
    def build_dense(self, xyz, expand_range: int = 0):
        if expand_range == 2:
            c0 = quantize_coords(xyz, self.voxel_size, 1)
        else:
            c0 = unique_rows(np.floor(scaled_positions(xyz, self.voxel_size)).astype(np.int64))
            if expand_range > 0:
                off = range_kernel(expand_range).astype(np.int64)
                c0 = unique_rows((c0.astype(np.int64)[:, None, :] + off[None]).reshape(-1, 3))
        self.set_coords(0, c0)
        for d in range(1, self.depth):
            s = self.strides[d]
            self.set_coords(d, unique_rows(np.floor_divide(self.coords[d - 1].astype(np.int64), s) * s))
        return np.zeros(np.asarray(xyz).shape[0], np.int32)

    # _hash_backend.py:380-412
    def build_subdivide(self, xyz, policy, expand=False, limit_adaptive_depth=100, **kw):
        """policy(current_pts (float32, voxel units), inverse map to the coarser level's unique cells,
        **kw) -> bool mask of the points that go one level finer."""
        xyz = np.asarray(xyz, np.float32)
        cur = scaled_positions(xyz, self.voxel_size)
        n = xyz.shape[0]
        xyz_depth = np.full(n, self.depth - 1, np.int32)
        inds = np.arange(n)
        inv = None
        for d in range(self.depth - 1, -1, -1):
            if d != self.depth - 1:
                m = np.asarray(policy(cur, inv, **kw), bool)
                cur, inds = cur[m], inds[m]
                kw = {k: (v[m] if isinstance(v, np.ndarray) else v) for k, v in kw.items()}
                xyz_depth[inds] -= 1
            s = self.strides[d]
            cells = (np.floor(cur * (np.float32(1) / np.float32(s))) * np.float32(s)).astype(np.int64)
            uc, inv = np.unique(cells, axis=0, return_inverse=True) if cells.shape[0] else (np.zeros((0, 3)), None)
            if inv is not None:
                inv = inv.reshape(-1)
            self.set_coords(d, uc)
        xyz_depth = np.minimum(xyz_depth, limit_adaptive_depth - 1)
        if expand:
            for d in range(self.depth):
                self.set_coords(d, quantize_coords(xyz[xyz_depth <= d], self.voxel_size, self.strides[d]))
        return xyz_depth

    # _hash_backend.py:292-318: 0 absent, 1 present and no child found, 2 some conformal child present
    def voxel_status(self, coords, d):
        coords = np.asarray(coords, np.int64).reshape(-1, 3)
        st = np.zeros(coords.shape[0], np.int64)
        sidx = np.nonzero(self.index[d].query(coords) >= 0)[0]
        st[sidx] = 1
        if d > 0 and sidx.shape[0]:
            ch = (coords[sidx][:, None, :] + CONFORM_OFFSETS[None] * self.strides[d - 1]).reshape(-1, 3)
            q = np.nonzero(self.index[d - 1].query(ch) >= 0)[0] // 8
            st[sidx[q]] = 2
        return st

    # _hash_backend.py:320-330: trilinear interpolation of per-voxel rows at the points
    def split_data(self, xyz, d, data):
        data = np.asarray(data, np.float64)
        ac, aw = trilinear_weights(xyz, self.voxel_size, self.strides[d])
        tid = self.index[d].query(ac)
        n = np.asarray(xyz).shape[0]
        src = np.arange(ac.shape[0]) % n
        ok = tid >= 0
        d2 = data.reshape(data.shape[0], -1)
        out = np.zeros((n, d2.shape[1]))
        for c in range(d2.shape[1]):
            out[:, c] = np.bincount(src[ok], weights=d2[tid[ok], c] * aw[ok], minlength=n)
        return out if data.ndim == 2 else out[:, 0]

    # _hash_backend.py:230-290 (same- and cross-level neighbours, range kernel)
    def self_neighbours(self, source_depth, target_depth, target_range):
        inv = source_depth > target_depth
        if inv:
            source_depth, target_depth = target_depth, source_depth
        sid, tid, nt, nbs = coords_neighbours(self.index[target_depth], self.strides[target_depth],
                                              self.coords[source_depth], self.strides[source_depth],
                                              range_kernel(target_range))
        if not inv:
            return sid, tid, nt, nbs
        near = np.all(np.abs(nt) < target_range / 2.0 + 1.0e-6, axis=1)
        nt = -nt / self.strides[target_depth] * self.strides[source_depth]
        return tid[near], sid[near], nt[near], None

    # _hash_backend.py:332-356
    def splat(self, xyz, d, data=None):
        s = self.strides[d]
        ac, ad = trilinear_weights(xyz, self.voxel_size, s, data)
        tid = self.index[d].query(ac)
        found = tid >= 0
        M = self.num_voxels(d)
        if ad.ndim == 1:
            return np.bincount(tid[found], weights=ad[found], minlength=M)
        return np.stack([np.bincount(tid[found], weights=ad[found, c], minlength=M) for c in range(ad.shape[1])], 1)

    # hashtree.py:157-192 (Bezier basis, no features); the gradient is bezier_tensor.py:62-72 (per
    # finest-voxel unit), the mask marks queries with any voxel in their 27-neighbourhood.
    def evaluate_interpolated(self, query_pos, d, feat, compute_mask=False, compute_grad=False):
        pos = scaled_positions(query_pos, self.voxel_size) - np.float32(0.5)
        s = self.strides[d]
        sid, tid, nt, _ = coords_neighbours(self.index[d], s, pos, 1, range_kernel(3), float_source=True)
        qc = -nt / s
        f = np.asarray(feat, np.float64)[tid]
        n = pos.shape[0]
        b = [bezier.B(qc[:, a]) for a in range(3)]
        if compute_grad:
            db = [bezier.dB(qc[:, a]) / s for a in range(3)]
            terms = (db[0] * b[1] * b[2], b[0] * db[1] * b[2], b[0] * b[1] * db[2])
            val = np.stack([np.bincount(sid, weights=t * f, minlength=n) for t in terms], 1)
        else:
            val = np.bincount(sid, weights=b[0] * b[1] * b[2] * f, minlength=n)
        if compute_mask:
            mask = np.zeros(n, bool)
            mask[sid] = True
            return val, mask
        return val

    def point_basis(self, query_pos, d):
        """Screening pre-evaluation (reconstructor.py:111-120): (pts_ids, vx_ids, values)."""
        pos = scaled_positions(query_pos, self.voxel_size) - np.float32(0.5)
        sid, tid, nt, nbs = coords_neighbours(self.index[d], self.strides[d], pos, 1, range_kernel(3),
                                              transposed=True, float_source=True)
        qc = -nt / self.strides[d]
        return sid, tid, bezier.B(qc[:, 0]) * bezier.B(qc[:, 1]) * bezier.B(qc[:, 2]), nbs


def level_of_density(xyz_density, min_density, log_base, depth):
    """_hash_backend.py:428-429: -clamp(floor(log(rho/min)/log(base)), max=0), clamped to depth-1."""
    t = np.floor(np.log(np.asarray(xyz_density, np.float64) / min_density) / np.log(log_base))
    return np.minimum(-np.minimum(t, 0), depth - 1).astype(np.int32)


def compute_density(xyz, voxel_size: float, density_depth: int = 2) -> np.ndarray:
    """hashtree.py:283-285 + _hash_backend.py:13-28: Bezier-interpolated splat of ones on the
    4x grid, divided by 16."""
    enc_stride = 2 ** density_depth
    vs4 = enc_stride * voxel_size
    h = Hierarchy(1, vs4)
    h.set_coords(0, quantize_coords(xyz, vs4, 1))
    wd = h.splat(xyz, 0)
    return h.evaluate_interpolated(xyz, 0, wd) / (enc_stride ** 2)
