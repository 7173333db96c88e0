"""Device sparse sets of lattice points and the cuckoo hash table (host side)."""
import ctypes
import math

import torch

from . import _lib
from ._lib import Grid, HashTable, Marks, NgsError, Scratch, call, lib, ptr, stream


def log2_int(s: int) -> int:
    k = int(round(math.log2(s)))
    if 2 ** k != s:
        raise ValueError("strides must be powers of two")
    return k


class SparseGrid:
    """Ordered bitmap set of lattice points whose coordinates are multiples of `stride`.

    `lo`, `hi` are inclusive bounds in lattice units (coordinate >> log2(stride)). Members are
    ranked in lexicographic (x, y, z) order, which is torch.unique(dim=0) order of their
    coordinates; the rank lookup doubles as a collision-free hash of the set.
    """

    def __init__(self, lo, hi, stride: int, device):
        self.stride = int(stride)
        self.log2 = log2_int(self.stride)
        self.origin = [int(v) for v in lo]
        self.dims = [int(h) - int(l) + 1 for l, h in zip(lo, hi)]
        dims_c = (ctypes.c_int32 * 3)(*self.dims)
        self.nwords = int(lib().ngs_grid_words(dims_c))
        if self.nwords <= 0:
            raise NgsError(6, f"bitmap box {self.dims} exceeds the 2^36-point capacity")
        self.device = torch.device(device)
        self.bits = torch.empty(self.nwords, dtype=torch.int32, device=self.device)
        self.rank = torch.zeros(self.nwords, dtype=torch.int32, device=self.device)
        self.count = None
        self.coords = None
        call("ngs_grid_clear", self.struct(), stream())

    @classmethod
    def empty(cls, device, stride=1):
        g = cls([0, 0, 0], [0, 0, 0], stride, device)
        g.count = 0
        g.coords = torch.zeros((0, 3), dtype=torch.int32, device=g.device)
        return g

    @classmethod
    def for_points(cls, qmin, qmax, stride: int, device, margin: int = 1):
        """Box covering the trilinear corners / containing cells of points with scaled positions
        in [qmin, qmax] at `stride`."""
        lo = [math.floor(float(a) / stride) - margin for a in qmin]
        hi = [math.floor(float(b) / stride) + margin for b in qmax]
        return cls(lo, hi, stride, device)

    def struct(self) -> Grid:
        g = Grid()
        for a in range(3):
            g.origin[a] = self.origin[a]
            g.dims[a] = self.dims[a]
        g.log2_stride = self.log2
        g.nwords = self.nwords
        g.bits = self.bits.data_ptr()
        g.rank = self.rank.data_ptr()
        return g

    def _marks(self, nmarks):
        """The zeroed byte map of the staged marking calls, or None (atomic marking) where the fold,
        which reads one byte per lattice point of the box, would cost more than the atomics it saves
        (more than 64 lattice points per mark) or the box exceeds Marks.CAP lattice points."""
        nb = self.nwords * 32
        return Marks.get(nb, self.device) if nb <= min(Marks.CAP, 64 * nmarks) else None

    def mark_coords(self, coords):
        if coords.shape[0]:
            m = self._marks(coords.shape[0])
            if m is None:
                call("ngs_grid_mark_coords", self.struct(), ptr(coords), coords.shape[0], stream())
            else:
                call("ngs_grid_mark_coords_staged", self.struct(), ptr(coords), coords.shape[0], ptr(m), stream())

    def mark_corners(self, xyz, inv_vs, levels=None, max_level=0):
        """The 8 trilinear corners at the grid stride of the points with level <= max_level."""
        if xyz.shape[0]:
            m = self._marks(8 * xyz.shape[0])
            if m is None:
                call("ngs_grid_mark_corners", self.struct(), ptr(xyz), xyz.shape[0], inv_vs, ptr(levels), max_level,
                     stream())
            else:
                call("ngs_grid_mark_corners_staged", self.struct(), ptr(xyz), xyz.shape[0], inv_vs, ptr(levels),
                     max_level, ptr(m), stream())

    @staticmethod
    def finalize_all(grids, emit=True):
        """finalize() of several grids with one device-to-host read of their member counts."""
        if not grids:
            return grids
        dev = grids[0].device
        cnt = torch.empty(len(grids), dtype=torch.int64, device=dev)
        for i, g in enumerate(grids):
            ws = Scratch.get(lib().ngs_scan_workspace_bytes(g.nwords), dev)
            call("ngs_grid_finalize", g.struct(), ptr(cnt[i:i + 1]), ptr(ws), stream())
        for g, c in zip(grids, cnt.tolist()):
            g.count = int(c)
            if emit:
                g.coords = torch.empty((g.count, 3), dtype=torch.int32, device=g.device)
                if g.count:
                    call("ngs_grid_emit", g.struct(), ptr(g.coords), stream())
        return grids

    def finalize(self, emit=True, count=True) -> int:
        """Ranks, member count (count=False: ranks only, no device-to-host read) and, with emit,
        the members in rank order."""
        ws = Scratch.get(lib().ngs_scan_workspace_bytes(self.nwords), self.device)
        if not count:
            call("ngs_grid_finalize", self.struct(), None, ptr(ws), stream())
            return None
        cnt = torch.empty(1, dtype=torch.int64, device=self.device)
        call("ngs_grid_finalize", self.struct(), ptr(cnt), ptr(ws), stream())
        self.count = int(cnt.item())
        if emit:
            self.coords = torch.empty((self.count, 3), dtype=torch.int32, device=self.device)
            if self.count:
                call("ngs_grid_emit", self.struct(), ptr(self.coords), stream())
        return self.count

    def query(self, coords, offsets=None):
        """rank of coords (+ offsets, added as given), -1 if absent: (N,) or (K, N) int64."""
        coords = coords.to(torch.int32).contiguous()
        n = coords.shape[0]
        if offsets is None:
            out = torch.empty(n, dtype=torch.int64, device=self.device)
            if n:
                call("ngs_grid_query", self.struct(), ptr(coords), n, None, 1, ptr(out), stream())
            return out
        offsets = offsets.to(torch.int32).contiguous()
        k = offsets.shape[0]
        out = torch.empty((k, n), dtype=torch.int64, device=self.device)
        if n and k:
            call("ngs_grid_query", self.struct(), ptr(coords), n, ptr(offsets), k, ptr(out), stream())
        return out


def hash_coords(coords):
    """fold60 FNV-1a keys of int rows (hashmap_cuda.cu:269-305); int64 rows keep their low 32 bits."""
    c = coords.to(torch.int32).contiguous()
    out = torch.empty(c.shape[0], dtype=torch.int64, device=c.device)
    if c.shape[0]:
        call("ngs_hash_coords", ptr(c), c.shape[0], c.shape[1], ptr(out), stream())
    return out


def kernel_hash(coords, offsets):
    """Keys of coords + offsets, (K, N) (hashmap_cuda.cu:307-339)."""
    c = coords.to(torch.int32).contiguous()
    o = offsets.to(torch.int32).contiguous()
    out = torch.empty((o.shape[0], c.shape[0]), dtype=torch.int64, device=c.device)
    if c.shape[0] and o.shape[0]:
        call("ngs_kernel_hash", ptr(c), c.shape[0], c.shape[1], ptr(o), o.shape[0], ptr(out), stream())
    return out


class CuckooHashTable:
    """Cuckoo hash table of voxel keys with the interface of jspsr/ext.py:14-84.

    Built on the GPU: keys are bucketed with wavefront-aggregated atomics, then each 512-slot
    bucket runs a three-function cuckoo insertion in LDS. query() returns the stored value (the
    row index by default) or -1.
    """

    def __init__(self, data=None, hashed_data=None, enlarged: bool = False, values=None, seed: int = 0x5EED):
        if data is not None:
            self.dim = data.size(1)
            assert data.size(0) > 0
            keys = hash_coords(data)
        else:
            self.dim = -1
            keys = hashed_data.contiguous()
        self.device = keys.device
        n = keys.numel()
        ts, nb, eb = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        call("ngs_hash_sizing", n, int(bool(enlarged)), ctypes.byref(ts), ctypes.byref(nb), ctypes.byref(eb))
        self.table_keys = torch.empty(3 * ts.value, dtype=torch.int64, device=self.device)
        self.table_vals = torch.empty(3 * ts.value, dtype=torch.int64, device=self.device)
        t = HashTable()
        t.table_size, t.num_buckets, t.evict_bound = ts.value, nb.value, eb.value
        t.keys, t.vals = self.table_keys.data_ptr(), self.table_vals.data_ptr()
        ws = torch.empty(lib().ngs_hash_workspace_bytes(ts.value), dtype=torch.uint8, device=self.device)
        vals = None if values is None else values.to(torch.int64).contiguous()
        call("ngs_hash_build", ctypes.byref(t), ptr(keys), ptr(vals), n, seed, ptr(ws), stream())
        self.object = t
        self.inserted_size = n

    @classmethod
    def _sphash(cls, coords, offsets=None):
        assert coords.dtype in (torch.int32, torch.int64), coords.dtype
        if offsets is None:
            assert coords.ndim == 2 and coords.shape[1] in (2, 3, 4)
            return hash_coords(coords)
        assert offsets.ndim == 2 and offsets.shape[1] == 3
        assert coords.ndim == 2 and coords.shape[1] in (3, 4)
        return kernel_hash(coords, offsets)

    def query(self, coords, offsets=None):
        assert coords.size(1) == self.dim
        return self.query_hashed(self._sphash(coords, offsets))

    def query_hashed(self, hashed_query):
        sizes = hashed_query.size()
        q = hashed_query.reshape(-1).contiguous()
        out = torch.empty(q.numel(), dtype=torch.int64, device=q.device)
        if q.numel():
            call("ngs_hash_query", ctypes.byref(self.object), ptr(q), q.numel(), ptr(out), stream())
        return out.view(*sizes)
