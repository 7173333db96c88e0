"""Implicit-field evaluation and sparse dual marching cubes (oracle; test infrastructure only).

Restates jspsr/core/reconstructor.py:197-255 (chi), :296-397 (extract_multiscale_mesh with
expand=0, trim=False), csrc/marching_cubes/dual_mc_inds.cu:45-182 (dual-cube corner assembly)
and csrc/marching_cubes/dual_mc.cu:20-189 (cube classification, edge interpolation, vertex keys).
"""
import numpy as np

from .hierarchy import CONFORM_OFFSETS, Hierarchy, range_kernel, unique_rows
from .mc_tables import EDGE_CORNERS, case_edges


class RowLookup:
    """coord -> row for an arbitrary (unsorted, unique) coord table; -1 when absent."""

    def __init__(self, coords):
        c = np.asarray(coords, np.int64).reshape(-1, 3)
        from .hierarchy import pack_keys
        self._pack = pack_keys
        k = pack_keys(c) if c.shape[0] else np.zeros(0, np.int64)
        self.order = np.argsort(k, kind="stable")
        self.keys = k[self.order]

    def query(self, q):
        q = np.asarray(q, np.int64).reshape(-1, 3)
        if self.keys.shape[0] == 0:
            return np.full(q.shape[0], -1, np.int64)
        k = self._pack(q)
        pos = np.minimum(np.searchsorted(self.keys, k), self.keys.shape[0] - 1)
        return np.where(self.keys[pos] == k, self.order[pos], -1)


def evaluate_raw_chi(h: Hierarchy, solutions: dict, xyz, compute_mask=False, compute_grad=False, depths=None):
    """reconstructor.py:197-226: chi (N,) or, with compute_grad, its gradient (N, 3); with
    compute_mask also the mask of queries covered by some evaluated level."""
    n = np.asarray(xyz).shape[0]
    val = np.zeros((n, 3) if compute_grad else n)
    mask = np.zeros(n, bool)
    for d, sol in solutions.items():
        if depths is not None and d not in depths:
            continue
        r = h.evaluate_interpolated(xyz, d, sol, compute_mask, compute_grad)
        if compute_mask:
            val += r[0]
            mask |= r[1]
        else:
            val += r
    return (val, mask) if compute_mask else val


def mean_chi(h, solutions, xyz, sample_weight):
    """reconstructor.py:228-236 (sample-weighted mean of chi at the input points)."""
    chi = evaluate_raw_chi(h, solutions, xyz)
    w = np.asarray(sample_weight, np.float64)
    return float(np.sum(chi * w) / np.sum(w))


def level_coords_ext(h: Hierarchy, d, expand=0, conforming=False):
    """_hash_backend.py:121-147: a level's voxels, dilated by the expand-cube (expand >= 3), and with
    conforming replaced by the 8 children of every distinct parent at twice the stride."""
    s = h.strides[d]
    c = h.coords[d].astype(np.int64)
    if expand >= 3:
        off = range_kernel(expand).astype(np.int64) * s
        c = unique_rows((c[:, None, :] + off[None]).reshape(-1, 3)).astype(np.int64)
    if conforming:
        par = unique_rows(np.floor_divide(c, 2 * s) * 2 * s).astype(np.int64)
        c = (par[:, None, :] + CONFORM_OFFSETS[None] * s).reshape(-1, 3)
    return c


def conformal_leaves(h: Hierarchy, max_depth=100, expand=0, trim=False):
    """reconstructor.py:309-336. Returns ({d: leaf coords}, max_depth)."""
    max_depth = min(max_depth, h.depth)
    base, mask = {}, {}
    if trim:
        for d in range(max_depth):
            st = h.voxel_status(h.coords[d], d)
            base[d] = h.coords[d].astype(np.int64)[st == 1]
        return base, max_depth
    for d in range(max_depth):
        bc = level_coords_ext(h, d, expand, conforming=d < max_depth - 1)
        if d > 0:
            ol = RowLookup(base[d - 1]).query(bc) == -1
            if not np.any(ol):
                max_depth = d
                break
            mask[d] = ol
        base[d] = bc
    for d in range(1, max_depth):
        base[d] = base[d][mask[d]]
    return base, max_depth


def expand_primal(h: Hierarchy, leaves: dict, max_depth: int, n_upsample: int):
    """reconstructor.py:338-348: each leaf split into n^3 sub-cells, upsampled integer units."""
    e = np.arange(n_upsample)
    ec = np.stack(np.meshgrid(e, e, e, indexing="ij"), -1).reshape(-1, 3)
    out = {}
    for d in range(max_depth):
        s = h.strides[d]
        out[d] = (leaves[d][:, None, :] * n_upsample + ec[None] * s).reshape(-1, 3)
    return out


def dual_centres(h, primal, max_depth):
    """reconstructor.py:350-358: sorted unique corners of all primal sub-cells."""
    pts = []
    for d in range(max_depth):
        s = h.strides[d]
        for off in CONFORM_OFFSETS:
            pts.append(primal[d] + off * s)
    return unique_rows(np.concatenate(pts, 0))


def octant_slot(sx, sy, sz):
    """MC corner index of the cell lying in octant (sx, sy, sz) (1 = positive side) around a dual
    centre; the geometric rule behind the corner/edge/face slot tables of dual_mc_inds.cu:16-42."""
    table = {(0, 0, 0): 0, (1, 0, 0): 1, (1, 1, 0): 2, (0, 1, 0): 3,
             (0, 0, 1): 4, (1, 0, 1): 5, (1, 1, 1): 6, (0, 1, 1): 7}
    return table[(sx, sy, sz)]


def boundary_points(s: int):
    """Local lattice offsets on the boundary of a cell of side s (corners, edge and face interiors),
    i.e. the 6 s^2 + 2 points enumerated per cell by dual_mc_inds.cu:45-141."""
    g = np.stack(np.meshgrid(*([np.arange(s + 1)] * 3), indexing="ij"), -1).reshape(-1, 3)
    on = np.any((g == 0) | (g == s), axis=1)
    return g[on]


def _boundary_slots(s: int):
    """For each boundary lattice offset of a side-s cell: the MC slots (octants) the cell covers
    around that lattice point, padded with -1 to 4."""
    bp = boundary_points(s)
    slots = np.full((bp.shape[0], 4), -1, np.int64)
    for k in range(bp.shape[0]):
        sides = [(1,) if v == 0 else (0,) if v == s else (0, 1) for v in bp[k]]
        sl = [octant_slot(x, y, z) for x in sides[0] for y in sides[1] for z in sides[2]]
        slots[k, :len(sl)] = sl
    return bp, slots


def accumulate_dual_corners(h, primal, max_depth, centres):
    """dual_mc_inds.cu:143-182. Returns acc (DC, 8, 2) int32 [primal id, depth], count (DC,).

    Events are processed in (level, primal cell, lattice point) order; an event whose dual centre
    already holds >= 8 registrations is counted but not written (dual_mc_inds.cu:273-277).
    """
    idx = RowLookup(centres)
    DC = centres.shape[0]
    acc = np.zeros((DC, 8, 2), np.int32)
    ev_c, ev_n, ev_slots, ev_pid, ev_d = [], [], [], [], []
    base = 0
    for d in range(max_depth):
        s = h.strides[d]
        P = primal[d]
        bp, slots = _boundary_slots(s)
        q = (P[:, None, :] + bp[None]).reshape(-1, 3)
        dc = idx.query(q)
        hit = np.nonzero(dc >= 0)[0]
        pi, ki = hit // bp.shape[0], hit % bp.shape[0]
        ev_c.append(dc[hit])
        ev_slots.append(slots[ki])
        ev_n.append((slots[ki] >= 0).sum(1))
        ev_pid.append(base + pi)
        ev_d.append(np.full(hit.shape[0], d))
        base += P.shape[0]
    c = np.concatenate(ev_c)
    n = np.concatenate(ev_n)
    sl = np.concatenate(ev_slots)
    pid = np.concatenate(ev_pid)
    dd = np.concatenate(ev_d)
    order = np.argsort(c, kind="stable")
    cs = c[order]
    ns = n[order]
    csum = np.cumsum(ns)
    first = np.searchsorted(cs, cs)
    old = csum - ns - np.where(first > 0, csum[first - 1], 0)
    cnt = np.bincount(c, weights=n, minlength=DC).astype(np.int64)
    ok = order[old < 8]
    for j in range(4):
        e = ok[sl[ok, j] >= 0]
        acc[c[e], sl[e, j], 0] = pid[e]
        acc[c[e], sl[e, j], 1] = dd[e]
    return acc, cnt


def sdf_interp(p1, p2, v1, v2):
    """dual_mc.cu:20-31 in float32."""
    f = np.float32
    if abs(f(0) - v1) < f(1e-5):
        return p1
    if abs(f(0) - v2) < f(1e-5):
        return p2
    if abs(v1 - v2) < f(1e-5):
        return p1
    w2 = (f(0) - v1) / (v2 - v1)
    w1 = f(1) - w2
    return (p1 * w1 + p2 * w2).astype(np.float32)


def _interp_vec(p1, p2, v1, v2):
    """dual_mc.cu:20-31 vectorised in float32."""
    f = np.float32
    with np.errstate(divide="ignore", invalid="ignore"):
        w2 = (f(0) - v1) / (v2 - v1)
    w1 = f(1) - w2
    out = (p1 * w1[:, None] + p2 * w2[:, None]).astype(np.float32)
    snap1 = (np.abs(f(0) - v1) < f(1e-5)) | ((np.abs(f(0) - v2) >= f(1e-5)) & (np.abs(v1 - v2) < f(1e-5)))
    snap2 = (~(np.abs(f(0) - v1) < f(1e-5))) & (np.abs(f(0) - v2) < f(1e-5))
    out[snap1] = p1[snap1]
    out[snap2] = p2[snap2]
    return out


def dual_mc(acc, corner_pos, corner_val):
    """dual_mc.cu:33-189: triangles (T,3,3) float32 and vertex keys (T,3,2) (larger id first), in
    cube order then case-table order."""
    pos = np.asarray(corner_pos, np.float32)
    val = np.asarray(corner_val, np.float32)
    ids = acc[:, :, 0].astype(np.int64)
    v = val[ids]
    case = ((v < 0) * (1 << np.arange(8))).sum(1)
    ntri = np.array([len(case_edges(c)) // 3 for c in range(256)])
    edges_tab = np.full((256, 15), -1, np.int64)
    for c in range(256):
        e = case_edges(c)
        edges_tab[c, :len(e)] = e
    nt = ntri[case]
    cube = np.repeat(np.arange(acc.shape[0]), nt)
    tin = np.arange(cube.shape[0]) - np.repeat(np.cumsum(nt) - nt, nt)
    ec = np.asarray(EDGE_CORNERS, np.int64)
    tris = np.zeros((cube.shape[0], 3, 3), np.float32)
    vids = np.zeros((cube.shape[0], 3, 2), np.int64)
    for k in range(3):
        e = edges_tab[case[cube], 3 * tin + k]
        a, b = ec[e, 0], ec[e, 1]
        ia, ib = ids[cube, a], ids[cube, b]
        tris[:, k] = _interp_vec(pos[ia], pos[ib], v[cube, a], v[cube, b])
        vids[:, k, 0] = np.maximum(ia, ib)
        vids[:, k, 1] = np.minimum(ia, ib)
    return tris, vids


def dedup_vertices(tris, vids):
    """reconstructor.py:379-383: sorted unique vertex keys; the position of a key is taken from
    its first occurrence (the reference's scatter leaves the choice unspecified)."""
    keys = vids.reshape(-1, 2)
    if keys.shape[0] == 0:
        return np.zeros((0, 3), np.float32), np.zeros((0, 3), np.int64)
    unq, first, inv = np.unique(keys, axis=0, return_index=True, return_inverse=True)
    verts = tris.reshape(-1, 3)[first]
    return verts, inv.reshape(-1, 3).astype(np.int64)


def extract_multiscale_mesh(h, solutions, sample_weight, xyz, n_upsample=2, chi_fn=None, expand=0, trim=False):
    """reconstructor.py:296-397 (no Open3D). chi_fn(points) overrides the field (used to feed
    identical samples to two implementations)."""
    leaves, max_depth = conformal_leaves(h, expand=expand, trim=trim)
    primal = expand_primal(h, leaves, max_depth, n_upsample)
    centres = dual_centres(h, primal, max_depth)
    acc, cnt = accumulate_dual_corners(h, primal, max_depth, centres)
    acc = acc[cnt == 8]
    evs = h.voxel_size / n_upsample
    corners = []
    for d in range(max_depth):
        corners.append(((primal[d] + h.strides[d]).astype(np.float32) * np.float32(evs)))
    corners = np.concatenate(corners, 0)
    if chi_fn is None:
        mc = mean_chi(h, solutions, xyz, sample_weight)
        values = evaluate_raw_chi(h, solutions, corners) - mc
    else:
        values = chi_fn(corners)
    tris, vids = dual_mc(acc, corners, values)
    v, f = dedup_vertices(tris, vids)
    return v, f, dict(leaves=leaves, primal=primal, centres=centres, acc=acc, corners=corners, values=values)
