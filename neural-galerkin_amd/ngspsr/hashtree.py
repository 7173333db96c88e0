"""Sparse voxel hierarchy on the GPU (the reconstruct-path part of jspsr/core/hashtree.py and
jspsr/core/_hash_backend.py).

One level per stride 2^d; every level keeps its voxel coordinates (sorted lexicographically,
the reference's torch.unique order) and the ordered bitmap set that maps a coordinate to its row.
The ENCODER/DECODER branches of the reference are reflected on this path, so a single branch is
kept and the branch argument is accepted and ignored.
"""
import numpy as np
import torch

from . import _lib
from .bases import BezierTensorBasis
from .grid import CuckooHashTable, SparseGrid
from ._lib import Level, Scratch, call, lib, ptr, stream

CONFORM_OFFSETS = [(0, 0, 0), (0, 0, 1), (0, 1, 0), (0, 1, 1), (1, 0, 0), (1, 0, 1), (1, 1, 0), (1, 1, 1)]


def f32_reciprocal(v: float) -> float:
    """1/v as PyTorch's GPU kernels apply a division by a CPU scalar: float32(1 / float32(v))."""
    return float(np.float32(1.0) / np.float32(v))


def range_kernel(n_range: int, device="cpu"):
    """Neighbour offsets of an odd range in concentric-cube prefix order (kernel[:r^3] is the
    r-cube for every odd r <= n_range): shells of growing Chebyshev radius, lexicographic inside a
    shell. The reference reads this order from a data file not shipped with it
    (hashtree.py:103-116); only the prefix property is used by its code."""
    assert n_range % 2 == 1, "target_range must be odd."
    h = n_range // 2
    r = np.arange(-h, h + 1)
    g = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    order = np.lexsort((g[:, 2], g[:, 1], g[:, 0], np.abs(g).max(1)))
    return torch.tensor(g[order], dtype=torch.int32, device=device)


class HashTree:
    ENCODER = 0
    DECODER = 1
    DECODER_TMP = 2

    def __init__(self, xyz, voxel_size: float, depth: int):
        if not xyz.is_cuda:
            raise ValueError("HashTree expects a device (cuda) tensor of points")
        self.xyz = xyz.float().contiguous()
        self.device = self.xyz.device
        self._voxel_size = float(voxel_size)
        self._depth = int(depth)
        self.inv_vs = f32_reciprocal(self._voxel_size)
        self.strides = [2 ** d for d in range(self._depth)]
        self.xyz_depth = None
        self.xyz_density = None
        self.is_reflected = False
        self.grids = [None] * self._depth
        self._hash = [None] * self._depth
        self._qmin = self._qmax = None
        self._level_counts = None

    # ------------------------------------------------------------------ basic accessors
    @property
    def voxel_size(self):
        return self._voxel_size

    @property
    def depth(self):
        return self._depth

    def reflect_decoder_coords(self):
        self.is_reflected = True

    def get_stride(self, branch, depth):
        return self.strides[depth]

    def get_coords(self, branch, depth):
        return self.grids[depth].coords

    def get_coords_size(self, branch, depth):
        g = self.grids[depth]
        return 0 if g is None else g.count

    def _coords_ext(self, depth, expand: int = 0, conforming: bool = False):
        """_hash_backend.py:121-147: the level's voxels, dilated by the expand-cube of offsets
        (expand >= 3, odd) and, with conforming, replaced by the 8 children of every distinct parent
        at twice the stride (parents sorted, children in (x, y, z) bit order)."""
        s = self.strides[depth]
        g = self.grids[depth]
        coords = g.coords
        dev = coords.device
        if expand >= 3:
            if expand % 2 == 0:
                raise ValueError("expand must be odd")
            h = expand // 2
            dg = SparseGrid([o - h for o in g.origin], [o + dm - 1 + h for o, dm in zip(g.origin, g.dims)], s, dev)
            if coords.shape[0]:
                call("ngs_grid_mark_dilated", dg.struct(), ptr(coords), coords.shape[0], h, stream())
            dg.finalize()
            coords = dg.coords
            g = dg
        if conforming:
            pg = SparseGrid([(o * s) // (2 * s) - 1 for o in g.origin],
                            [((o + dm) * s) // (2 * s) + 1 for o, dm in zip(g.origin, g.dims)], 2 * s, dev)
            pg.mark_coords(coords)
            pg.finalize()
            npar = pg.count
            out = torch.empty((npar * 8, 3), dtype=torch.int32, device=dev)
            keep = torch.empty(npar * 8, dtype=torch.int32, device=dev)
            if npar:
                call("ngs_conformal_children", ptr(pg.coords), npar, s, SparseGrid.empty(dev).struct(), 0, ptr(out),
                     ptr(keep), stream())
            coords = out
        return coords

    def get_test_grid(self, resolution, depth, expand: int = 0, conforming: bool = False):
        """hashtree.py:209-229: sample positions (N R^3, 3) of an R^3 lattice spanning every voxel of
        the level (corners included, R = resolution) in metric units, and the voxel corners (N, 3)
        int32; (None, corners) for resolution 0."""
        s = self.strides[depth]
        base = self._coords_ext(depth, expand, conforming)
        if resolution == 0:
            return None, base
        n = base.shape[0]
        q = torch.empty((n * resolution ** 3, 3), dtype=torch.float32, device=base.device)
        if n:
            call("ngs_test_grid_positions", ptr(base), n, s, int(resolution), float(self._voxel_size), ptr(q),
                 stream())
        return q, base

    def get_voxel_centers(self, depth, branch=DECODER, normalized=False):
        c = self.get_coords(branch, depth).float() + self.strides[depth] / 2.0
        return c if normalized else c * self._voxel_size

    def level(self, d) -> Level:
        g = self.grids[d]
        lv = Level()
        lv.grid = g.struct()
        lv.coords = g.coords.data_ptr() if g.count else None
        lv.M = g.count
        lv.stride = self.strides[d]
        return lv

    def hash_table(self, d) -> CuckooHashTable:
        """The level's cuckoo table (built on first use): key = FNV fold60 of the coordinate, value =
        row index — the logical content of the reference's _hash_table[d]."""
        if self._hash[d] is None and self.get_coords_size(None, d) > 0:
            self._hash[d] = CuckooHashTable(data=self.grids[d].coords)
        return self._hash[d]

    def get_range_kernel(self, n_range):
        return range_kernel(n_range, self.device)

    # ------------------------------------------------------------------ construction
    def _bounds(self, inv_vs):
        mm = torch.empty(6, dtype=torch.float32, device=self.device)
        call("ngs_point_bounds", ptr(self.xyz), self.xyz.shape[0], inv_vs, ptr(mm), stream())
        v = mm.cpu().numpy()
        return v[:3], v[3:]

    def _voxelise(self, stride, inv_vs, qmin, qmax, levels=None, max_level=0):
        g = SparseGrid.for_points(qmin, qmax, stride, self.device)
        g.mark_corners(self.xyz, inv_vs, levels, max_level)
        g.finalize()
        return g

    def compute_density(self, density_depth: int = 2):
        """hashtree.py:283-285 + _hash_backend.py:13-28: Bezier interpolation of the trilinear
        splat of ones on the (2^density_depth x voxel) grid, divided by 4^density_depth."""
        enc = 2 ** density_depth
        inv4 = f32_reciprocal(enc * self._voxel_size)
        q4min, q4max = self._bounds(inv4)
        g4 = self._voxelise(1, inv4, q4min, q4max)
        w = torch.zeros(g4.count, dtype=torch.float64, device=self.device)
        self._splat(g4, self.xyz, inv4, None, 0, None, None, 0, w)
        wf = w.float()
        rho = torch.empty(self.xyz.shape[0], dtype=torch.float32, device=self.device)
        call("ngs_eval_field", g4.struct(), ptr(wf), ptr(self.xyz), self.xyz.shape[0], inv4, 1.0 / (enc * enc), 0,
             ptr(rho), None, stream())
        return rho

    def build_encoder_hierarchy_adaptive(self, density_depth: int = 2, log_base: float = 4.0, min_density: float = 8.0,
                                         uniform_density: bool = False, limit_adaptive_depth: int = 100,
                                         xyz_depth=None):
        """hashtree.py:266-287 and _hash_backend.py:414-444. `xyz_depth` (optional) imposes the
        per-point levels instead of deriving them from the density."""
        n = self.xyz.shape[0]
        if uniform_density:
            self.xyz_density = torch.full((n,), 1e6, dtype=torch.float32, device=self.device)
        else:
            self.xyz_density = self.compute_density(density_depth)
        if xyz_depth is None:
            lv = torch.empty(n, dtype=torch.int32, device=self.device)
            call("ngs_point_levels", ptr(self.xyz_density), n, float(min_density), float(log_base),
                 min(self._depth, limit_adaptive_depth), ptr(lv), stream())
        else:
            lv = xyz_depth.to(device=self.device, dtype=torch.int32).contiguous()
        self.xyz_depth = lv
        self._level_counts = torch.bincount(lv.long(), minlength=self._depth).cpu().numpy()
        self._qmin, self._qmax = self._bounds(self.inv_vs)
        grids = []
        for d in range(self._depth):
            g = SparseGrid.for_points(self._qmin, self._qmax, self.strides[d], self.device)
            g.mark_corners(self.xyz, self.inv_vs, lv, d)
            grids.append(g)
        for d, g in enumerate(SparseGrid.finalize_all(grids)):
            self.grids[d] = g
        return self.xyz_depth

    def level_point_counts(self):
        return self._level_counts

    # ------------------------------------------------------------------ data on the tree
    def _splat(self, g, xyz, inv_vs, data, nc, scale, levels, level, out):
        """acc += the trilinear splat of the points on g: the binned gather (no FP64 atomics), or the
        atomic scatter where g's members are not emitted or the sizes exceed the gather's."""
        n = xyz.shape[0]
        wsb = lib().ngs_splat_gather_workspace_bytes(n, g.count) if g.coords is not None else -1
        if wsb < 0:
            call("ngs_splat", g.struct(), ptr(xyz), n, inv_vs, ptr(data), nc, ptr(scale), ptr(levels), level, 1.0,
                 ptr(out), stream())
            return
        ws = Scratch.get(wsb, self.device)
        call("ngs_splat_gather", g.struct(), ptr(g.coords), g.count, ptr(xyz), n, inv_vs, ptr(data), nc, ptr(scale),
             ptr(levels), level, 1.0, ptr(out), ptr(ws), wsb, stream())

    def splat_data(self, xyz, branch, data_depth, data=None, check_corr=True, scale=None, levels=None, level=0):
        """_hash_backend.py:332-356 (float64 accumulation, returned as float32)."""
        g = self.grids[data_depth]
        xyz = xyz.float().contiguous()
        if data is None:
            out = torch.zeros(g.count, dtype=torch.float64, device=self.device)
            self._splat(g, xyz, self.inv_vs, None, 0, scale, levels, level, out)
            return out.float()
        data = data.float().contiguous()
        nc = 1 if data.ndim == 1 else data.shape[1]
        out = torch.zeros((g.count, nc), dtype=torch.float64, device=self.device)
        self._splat(g, xyz, self.inv_vs, data, nc, scale, levels, level, out)
        return out.float().squeeze(1) if data.ndim == 1 else out.float()

    def evaluate_interpolated(self, query_pos, basis, basis_feat, feat_depth, feat, compute_mask=False,
                              compute_grad=False):
        """hashtree.py:157-192 for the Bezier basis."""
        q = query_pos.float().contiguous()
        n = q.shape[0]
        f = feat.float().contiguous()
        out = torch.empty(n, dtype=torch.float32, device=self.device)
        grad = torch.empty((n, 3), dtype=torch.float32, device=self.device) if compute_grad else None
        g = self.grids[feat_depth].struct()
        call("ngs_eval_field", g, ptr(f), ptr(q), n, self.inv_vs, 1.0, 0, ptr(out), ptr(grad), stream())
        val = grad if compute_grad else out
        if compute_mask:
            mask = torch.zeros(n, dtype=torch.bool, device=self.device)
            call("ngs_field_support", g, ptr(q), n, self.inv_vs, ptr(mask), stream())
            return val, mask
        return val

    # ------------------------------------------------------------------ neighbour COO (API)
    def get_neighbours_data(self, source_coords, source_stride, target_depth, nn_kernel, branch=None,
                            conv_based=False, transposed=False):
        """_hash_backend.py:176-228: for every source, the target voxels of target_depth around it.
        Returns (source_ids, target_ids, neighbour_types, nbsizes) in the reference's order."""
        st = self.strides[target_depth]
        if not conv_based:
            assert source_stride <= st, "Data must be deeper and has more nodes."
            quantized = torch.div(source_coords + 0.5 * source_stride, st, rounding_mode="floor").int() * st
            c_offset = (quantized - source_coords) / source_stride + (st // source_stride - 1) / 2.0
        else:
            assert source_stride >= st
            quantized = source_coords.int()
        res = self.grids[target_depth].query(quantized, nn_kernel.int() * st)  # (K, N)
        if transposed:
            res = res.T
        nbsizes = torch.sum(res != -1, dim=1)
        if transposed:
            source_ids, kernel_ids = torch.where(res != -1)
            target_ids = res[source_ids, kernel_ids]
        else:
            kernel_ids, source_ids = torch.where(res != -1)
            target_ids = res[kernel_ids, source_ids]
        nt = nn_kernel[kernel_ids]
        if not conv_based:
            nt = nt.float() * (st / source_stride) + c_offset[source_ids, :3]
        return source_ids, target_ids, nt, nbsizes

    def get_neighbours(self, source_depth, target_depth, target_range, branch=None, conv_based=False):
        """_hash_backend.py:230-290 (uncached): (src, tgt, rel, nbsizes)."""
        inv = (not conv_based) and source_depth > target_depth
        if inv:
            source_depth, target_depth = target_depth, source_depth
        src, tgt, nt, nbs = self.get_neighbours_data(self.grids[source_depth].coords, self.strides[source_depth],
                                                     target_depth, self.get_range_kernel(target_range),
                                                     conv_based=conv_based)
        if not inv:
            return src, tgt, nt, nbs
        near = torch.all(nt.abs() < target_range / 2.0 + 1.0e-6, dim=1)
        nt = -nt / self.strides[target_depth] * self.strides[source_depth]
        return tgt[near], src[near], nt[near], None


def basis():
    return BezierTensorBasis()
