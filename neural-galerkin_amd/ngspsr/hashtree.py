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


class _Level:
    """One level of a branch: the ordered bitmap set of its voxels (rank = lexicographic order, the
    order every kernel indexes rows by) and, when the caller set the level's rows in another order
    (update_coords), the coordinates in that order with the maps between row and rank."""
    __slots__ = ("grid", "coords", "rank_of_row", "row_of_rank", "hash")

    def __init__(self, grid, coords=None, rank_of_row=None, row_of_rank=None):
        self.grid = grid
        self.coords = coords
        self.rank_of_row = rank_of_row
        self.row_of_rank = row_of_rank
        self.hash = None

    @property
    def row_coords(self):
        return self.grid.coords if self.coords is None else self.coords


class HashTree:
    """hashtree.py:24-312 on the GPU. Three branches, as the reference keeps (ENCODER, DECODER,
    DECODER_TMP): the build methods fill the encoder, reflect_decoder_coords makes the decoders the
    same object, update_coords replaces one level of one branch. The reconstruction reads the
    decoder branch (reconstructor.py:39)."""
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
        self._enc = [None] * self._depth
        self._dec = [None] * self._depth
        self._tmp = [None] * self._depth
        self._level_counts = None
        self._qbounds = None

    # ------------------------------------------------------------------ basic accessors
    @property
    def voxel_size(self):
        return self._voxel_size

    @property
    def depth(self):
        return self._depth

    def reflect_decoder_coords(self):
        """hashtree.py:57-62: the decoder branches become the encoder branch (one object)."""
        self._dec = self._tmp = self._enc
        self.is_reflected = True

    def _branch(self, branch):
        if branch is None or branch == self.DECODER:
            return self._dec
        if branch == self.ENCODER:
            return self._enc
        if branch == self.DECODER_TMP:
            return self._tmp
        raise NotImplementedError(f"unknown branch {branch}")

    def _level(self, branch, depth) -> _Level:
        lv = self._branch(branch)[depth]
        if lv is None:
            raise ValueError(f"level {depth} of branch {branch} has no voxels set (build the hierarchy, "
                             "reflect_decoder_coords or update_coords first)")
        return lv

    @property
    def grids(self):
        """Bitmap sets of the decoder branch's levels (None where unset), in rank order."""
        return [None if lv is None else lv.grid for lv in self._dec]

    def get_stride(self, branch, depth):
        return self.strides[depth]

    def get_coords(self, branch, depth):
        """(M, 3) int32 voxel coordinates of the level, in the level's row order (None if unset)."""
        lv = self._branch(branch)[depth]
        return None if lv is None else lv.row_coords

    def get_coords_size(self, branch, depth):
        lv = self._branch(branch)[depth]
        return 0 if lv is None else lv.grid.count

    def to_rank(self, depth, data, branch=None):
        """Per-voxel rows given in the level's row order, reordered to rank order (the kernels')."""
        lv = self._branch(branch)[depth]
        if lv is None or lv.row_of_rank is None:
            return data
        return data[lv.row_of_rank]

    def to_rows(self, depth, data, branch=None):
        """Per-voxel rows in rank order, reordered to the level's row order."""
        lv = self._branch(branch)[depth]
        if lv is None or lv.rank_of_row is None:
            return data
        return data[lv.rank_of_row]

    def _grid_of_coords(self, coords, stride):
        """Bitmap set (finalized, emitted) of int32 coordinates that are multiples of stride."""
        n = coords.shape[0]
        if n == 0:
            return SparseGrid.empty(self.device, stride)
        lo = (torch.div(coords.min(0).values, stride, rounding_mode="floor")).tolist()
        hi = (torch.div(coords.max(0).values, stride, rounding_mode="floor")).tolist()
        g = SparseGrid([v - 1 for v in lo], [v + 1 for v in hi], stride, self.device)
        g.mark_coords(coords)
        g.finalize()
        return g

    def update_coords(self, branch, depth, coords):
        """hashtree.py:97-101 / _hash_backend.py:446-452: replace the voxels of one level of one
        branch (None: no voxels). Rows keep the given order: per-voxel data of the level (features,
        splats, solutions) follows it. Coordinates must be distinct multiples of the level's stride."""
        s = self.strides[depth]
        if coords is None:
            coords = torch.zeros((0, 3), dtype=torch.int32, device=self.device)
        assert coords.ndim == 2 and coords.shape[1] == 3, tuple(coords.shape)
        coords = coords.to(device=self.device, dtype=torch.int32).contiguous()
        n = coords.shape[0]
        if n and bool((coords % s != 0).any()):
            raise ValueError(f"update_coords: level {depth} coordinates must be multiples of its stride {s}")
        g = self._grid_of_coords(coords, s)
        if g.count != n:
            raise ValueError(f"update_coords: level {depth} coordinates must be distinct ({n - g.count} repeated)")
        rank = g.query(coords)
        br = self._branch(branch)
        if n == 0 or bool((rank == torch.arange(n, device=self.device)).all()):
            br[depth] = _Level(g)
        else:
            row_of_rank = torch.empty_like(rank)
            row_of_rank[rank] = torch.arange(n, dtype=torch.int64, device=self.device)
            br[depth] = _Level(g, coords, rank, row_of_rank)

    def _coords_ext(self, depth, expand: int = 0, conforming: bool = False):
        """_hash_backend.py:121-147 (decoder branch): the level's voxels in row order, or dilated by
        the expand-cube of offsets (expand >= 3, odd; sorted) and, with conforming, replaced by the 8
        children of every distinct parent at twice the stride (parents sorted, children in (x, y, z)
        bit order)."""
        if expand < 3 and not conforming:
            return self.get_coords(self.DECODER, depth)
        s = self.strides[depth]
        g = self._level(self.DECODER, depth).grid
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

    def get_random_samples(self, sample_per_voxel: int, depth, expand: int = 0):
        """hashtree.py:194-207: sample_per_voxel uniform random positions in every voxel of the level
        (decoder branch, dilated by `expand` as get_test_grid), (N P, 3) in metric units."""
        scale = self.strides[depth]
        base = self._coords_ext(depth, expand)
        local = torch.rand((base.shape[0], sample_per_voxel, 3), device=self.device) * scale
        return ((base.unsqueeze(1) + local).view(-1, 3)) * self._voxel_size

    def get_voxel_centers(self, depth, branch=DECODER, normalized=False):
        c = self.get_coords(branch, depth).float() + self.strides[depth] / 2.0
        return c if normalized else c * self._voxel_size

    def level(self, d) -> Level:
        """Kernel view of decoder level d (rank order)."""
        g = self._level(self.DECODER, d).grid
        lv = Level()
        lv.grid = g.struct()
        lv.coords = g.coords.data_ptr() if g.count else None
        lv.M = g.count
        lv.stride = self.strides[d]
        return lv

    def hash_table(self, d, branch=None) -> CuckooHashTable:
        """The level's cuckoo table (built on first use): key = FNV fold60 of the coordinate, value =
        row index — the logical content of the reference's _hash_table[d]."""
        lv = self._branch(branch)[d]
        if lv is None or lv.grid.count == 0:
            return None
        if lv.hash is None:
            lv.hash = CuckooHashTable(data=lv.row_coords)
        return lv.hash

    def get_range_kernel(self, n_range):
        return range_kernel(n_range, self.device)

    # ------------------------------------------------------------------ construction
    def _bounds(self, inv_vs, xyz=None):
        xyz = self.xyz if xyz is None else xyz
        mm = torch.empty(6, dtype=torch.float32, device=self.device)
        call("ngs_point_bounds", ptr(xyz), xyz.shape[0], inv_vs, ptr(mm), stream())
        v = mm.cpu().numpy()
        return v[:3], v[3:]

    def point_bounds(self, xyz=None):
        """(min, max) per axis of q = xyz / voxel_size: the tree's own points by default (computed
        once), or the given points."""
        own = xyz is None or (xyz.data_ptr() == self.xyz.data_ptr() and xyz.shape == self.xyz.shape
                              and xyz.dtype == self.xyz.dtype)
        if not own:
            return self._bounds(self.inv_vs, xyz.float().contiguous())
        if self._qbounds is None:
            self._qbounds = self._bounds(self.inv_vs)
        return self._qbounds

    def _voxelise(self, stride, inv_vs, qmin, qmax, levels=None, max_level=0):
        g = SparseGrid.for_points(qmin, qmax, stride, self.device)
        g.mark_corners(self.xyz, inv_vs, levels, max_level)
        g.finalize()
        return g

    def _point_cells(self, q, scale, stride):
        """(n, 3) int32 floor(q * scale / stride) * stride (ngs_point_cells)."""
        q = q.float().contiguous()
        out = torch.empty((q.shape[0], 3), dtype=torch.int32, device=self.device)
        if q.shape[0]:
            call("ngs_point_cells", ptr(q), q.shape[0], float(scale), int(stride), ptr(out), stream())
        return out

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

    def _set_encoder_levels(self, grids):
        for d, g in enumerate(grids):
            self._enc[d] = _Level(g)

    def _corner_levels(self, xyz_depth):
        """Level d = trilinear corners (at stride 2^d) of the points whose level is <= d
        (_hash_backend.py:436-441, 457-462), all levels finalized with one count read."""
        qmin, qmax = self.point_bounds()
        grids = []
        for d in range(self._depth):
            g = SparseGrid.for_points(qmin, qmax, self.strides[d], self.device)
            g.mark_corners(self.xyz, self.inv_vs, xyz_depth, d)
            grids.append(g)
        return SparseGrid.finalize_all(grids)

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
        self._set_encoder_levels(self._corner_levels(lv))
        return self.xyz_depth

    def build_encoder_hierarchy_dense(self, expand_range: int = 0, density_depth: int = 2,
                                      uniform_density: bool = False):
        """hashtree.py:240-252 and _hash_backend.py:358-378: every point on the finest level. Level 0
        holds the trilinear corners of the points (expand_range 2) or the cells containing them,
        dilated by the expand_range-cube of offsets (odd expand_range); each coarser level holds the
        parents of the level below. Densities are 1 with uniform_density."""
        n = self.xyz.shape[0]
        if expand_range not in (0, 2) and expand_range % 2 == 0:
            raise ValueError("expand_range must be 0, 2 or odd (the range kernel is odd)")
        if uniform_density:
            self.xyz_density = torch.ones(n, dtype=torch.float32, device=self.device)
        else:
            self.xyz_density = self.compute_density(density_depth)
        self.xyz_depth = torch.zeros(n, dtype=torch.int32, device=self.device)
        self._level_counts = np.array([n] + [0] * (self._depth - 1))
        qmin, qmax = self.point_bounds()
        if expand_range == 2:
            g0 = SparseGrid.for_points(qmin, qmax, 1, self.device)
            g0.mark_corners(self.xyz, self.inv_vs)
        else:
            h = expand_range // 2
            g0 = SparseGrid.for_points(qmin, qmax, 1, self.device, margin=1 + h)
            cells = self._point_cells(self.xyz, self.inv_vs, 1)
            if h == 0:
                g0.mark_coords(cells)
            elif n:
                call("ngs_grid_mark_dilated", g0.struct(), ptr(cells), n, h, stream())
        g0.finalize()
        grids = [g0]
        if self._depth > 1:
            coarse = []
            for d in range(1, self._depth):
                s = self.strides[d]
                g = SparseGrid([(o // s) - 1 for o in g0.origin], [((o + dm - 1) // s) + 1 for o, dm in
                                                                    zip(g0.origin, g0.dims)], s, self.device)
                # floor(c / s_d) s_d of the level-0 voxels: the parents of the level below, level by level
                g.mark_coords(g0.coords)
                coarse.append(g)
            grids += SparseGrid.finalize_all(coarse)
        self._set_encoder_levels(grids)

    def build_hierarchy_subdivide(self, subdivide_policy, expand: bool = False, density_depth: int = 2,
                                  limit_adaptive_depth: int = 100, **policy_kwargs):
        """hashtree.py:254-264 and _hash_backend.py:380-412: top-down subdivision. Every point starts
        on the coarsest level; at each finer level subdivide_policy(points / voxel_size, inverse map
        of the points to the coarser level's voxels, **policy_kwargs) keeps the points that go one
        level finer. A level holds the cells of its points (or, with expand, the trilinear corners of
        the points whose level is <= d). Returns xyz_depth."""
        self.xyz_density = self.compute_density(density_depth)
        n = self.xyz.shape[0]
        qmin, qmax = self.point_bounds()
        current = self.xyz * self.inv_vs  # the reference's xyz / voxel_size in FP32
        xyz_depth = torch.full((n,), self._depth - 1, dtype=torch.int32, device=self.device)
        inds = torch.arange(n, dtype=torch.int64, device=self.device)
        inv_mapping = None
        grids = [None] * self._depth
        for d in range(self._depth - 1, -1, -1):
            if d != self._depth - 1:
                mask = subdivide_policy(current, inv_mapping, **policy_kwargs)
                current = current[mask]
                inds = inds[mask]
                policy_kwargs = {k: v[mask] if isinstance(v, torch.Tensor) else v for k, v in policy_kwargs.items()}
                xyz_depth[inds] -= 1
            s = self.strides[d]
            cells = self._point_cells(current, 1.0, s)
            g = SparseGrid.for_points(qmin, qmax, s, self.device)
            g.mark_coords(cells)
            g.finalize(emit=False, count=False)  # ranks only: the inverse map of the points
            inv_mapping = g.query(cells)
            grids[d] = g
        xyz_depth.clamp_(max=limit_adaptive_depth - 1)
        self.xyz_depth = xyz_depth
        self._level_counts = torch.bincount(xyz_depth.long(), minlength=self._depth).cpu().numpy()
        if expand:
            grids = self._corner_levels(xyz_depth)
        else:
            grids = SparseGrid.finalize_all(grids)
        self._set_encoder_levels(grids)
        return self.xyz_depth

    def level_point_counts(self):
        return self._level_counts

    def evaluate_voxel_status(self, branch, coords, depth):
        """hashtree.py:128-139 / _hash_backend.py:292-318: per coordinate (N, 3), 0 (VS_NON_EXIST) when
        the level has no such voxel, 1 (VS_EXIST_STOP) when it has no children on the next finer level,
        2 (VS_EXIST_CONTINUE) when it has; (N,) int64."""
        lv = self._level(branch, depth)
        coords = coords.to(device=self.device, dtype=torch.int32).contiguous()
        n = coords.shape[0]
        out = torch.empty(n, dtype=torch.int64, device=self.device)
        if n == 0:
            return out
        fp = SparseGrid.empty(self.device, self.strides[depth])
        has_finer = 0
        if depth > 0:
            fine = self._level(branch, depth - 1).grid
            if fine.count:
                s = self.strides[depth]
                fp = SparseGrid([(o * fine.stride) // s - 1 for o in fine.origin],
                                [((o + dm - 1) * fine.stride) // s + 1 for o, dm in zip(fine.origin, fine.dims)], s,
                                self.device)
                fp.mark_coords(fine.coords)  # bits only: the status test reads no ranks
                has_finer = 1
        call("ngs_voxel_status", lv.grid.struct(), fp.struct(), has_finer, ptr(coords), n, ptr(out), stream())
        return out

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
        """_hash_backend.py:332-356 (float64 accumulation, returned as float32), rows in the level's
        row order."""
        g = self._level(branch, data_depth).grid
        xyz = xyz.float().contiguous()
        if data is None:
            out = torch.zeros(g.count, dtype=torch.float64, device=self.device)
            self._splat(g, xyz, self.inv_vs, None, 0, scale, levels, level, out)
            return self.to_rows(data_depth, out.float(), branch)
        data = data.float().contiguous()
        nc = 1 if data.ndim == 1 else data.shape[1]
        out = torch.zeros((g.count, nc), dtype=torch.float64, device=self.device)
        self._splat(g, xyz, self.inv_vs, data, nc, scale, levels, level, out)
        out = out.float().squeeze(1) if data.ndim == 1 else out.float()
        return self.to_rows(data_depth, out, branch)

    def split_data(self, xyz, branch, data_depth, data):
        """hashtree.py:289-298 / _hash_backend.py:320-330: trilinear interpolation of the per-voxel
        rows data (M, K) of a level at the points, (N, K) float32."""
        g = self._level(branch, data_depth).grid
        assert data.shape[0] == g.count, "Tree data does not agree on size."
        xyz = xyz.float().contiguous()
        d2 = data if data.ndim == 2 else data.reshape(-1, 1)
        rows = self.to_rank(data_depth, d2.float(), branch).contiguous()
        nc = rows.shape[1]
        out = torch.empty((xyz.shape[0], nc), dtype=torch.float32, device=self.device)
        if xyz.shape[0] and nc:
            if g.count:
                call("ngs_split_data", g.struct(), ptr(xyz), xyz.shape[0], self.inv_vs, ptr(rows), nc, ptr(out),
                     stream())
            else:
                out.zero_()
        return out if data.ndim == 2 else out.squeeze(1)

    def evaluate_interpolated(self, query_pos, basis, basis_feat, feat_depth, feat, compute_mask=False,
                              compute_grad=False):
        """hashtree.py:157-192 for the Bezier basis (decoder branch)."""
        q = query_pos.float().contiguous()
        n = q.shape[0]
        f = self.to_rank(feat_depth, feat.float()).contiguous()
        out = torch.empty(n, dtype=torch.float32, device=self.device)
        grad = torch.empty((n, 3), dtype=torch.float32, device=self.device) if compute_grad else None
        g = self._level(self.DECODER, feat_depth).grid.struct()
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
        Returns (source_ids, target_ids, neighbour_types, nbsizes) in the reference's order; target
        ids are rows of the target level."""
        st = self.strides[target_depth]
        lv = self._level(branch, target_depth)
        if not conv_based:
            assert source_stride <= st, "Data must be deeper and has more nodes."
            quantized = torch.div(source_coords + 0.5 * source_stride, st, rounding_mode="floor").int() * st
            c_offset = (quantized - source_coords) / source_stride + (st // source_stride - 1) / 2.0
        else:
            assert source_stride >= st
            quantized = source_coords.int()
        res = lv.grid.query(quantized, nn_kernel.int() * st)  # (K, N)
        if transposed:
            res = res.T
        nbsizes = torch.sum(res != -1, dim=1)
        if transposed:
            source_ids, kernel_ids = torch.where(res != -1)
            target_ids = res[source_ids, kernel_ids]
        else:
            kernel_ids, source_ids = torch.where(res != -1)
            target_ids = res[kernel_ids, source_ids]
        if lv.row_of_rank is not None:
            target_ids = lv.row_of_rank[target_ids]
        nt = nn_kernel[kernel_ids]
        if not conv_based:
            nt = nt.float() * (st / source_stride) + c_offset[source_ids, :3]
        return source_ids, target_ids, nt, nbsizes

    def get_neighbours(self, source_depth, target_depth, target_range, branch=None, conv_based=False):
        """_hash_backend.py:230-290 (uncached): (src, tgt, rel, nbsizes), ids in row order."""
        inv = (not conv_based) and source_depth > target_depth
        if inv:
            source_depth, target_depth = target_depth, source_depth
        src, tgt, nt, nbs = self.get_neighbours_data(self.get_coords(branch, source_depth),
                                                     self.strides[source_depth], target_depth,
                                                     self.get_range_kernel(target_range), branch=branch,
                                                     conv_based=conv_based)
        if not inv:
            return src, tgt, nt, nbs
        near = torch.all(nt.abs() < target_range / 2.0 + 1.0e-6, dim=1)
        nt = -nt / self.strides[target_depth] * self.strides[source_depth]
        return tgt[near], src[near], nt[near], None


def basis():
    return BezierTensorBasis()
