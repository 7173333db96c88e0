"""Primal sparse marching cubes with the MeshSDF backward (jspsr/core/ops.py:9-61,85-101,
csrc/marching_cubes/primal_mc.cu) on the HIP kernels of libngspsr."""
import ctypes

import torch

from ._lib import NgsError, Scratch, call, lib, ptr, stream

_MAX_VERTS_PER_CUBE = 15


def marching_cubes_sparse(base_coords: torch.Tensor, sdf_val: torch.Tensor):
    """marching_cubes.marching_cubes_sparse (primal_mc.cu:176-234).

    :param base_coords: (L, 3) integer-valued leaf corners, in leaf units
    :param sdf_val: (L, r+1, r+1, r+1) samples; sub-cube (rx, ry, rz) spans samples rx..rx+1 etc.
    :return: tris (T, 3, 3) float32 positions (base + offset / r), vert_ids (T, 3, 4) int32 edge ids
        (global sub-lattice corner of the edge's lower end, axis), src_ids (T, 3, 2) int32 flat sample
        indices of the edge's ends, src_weight (T, 3, 1) float32 weight of the first end. Triangles
        come in (leaf, rx, ry, rz) then case-table order.
    """
    dev = sdf_val.device
    base = base_coords.float().contiguous()
    sdf = sdf_val.float().contiguous()
    if sdf.dim() != 4 or not (sdf.shape[1] == sdf.shape[2] == sdf.shape[3]) or sdf.shape[1] < 2:
        raise NgsError("marching_cubes_sparse: sdf_val must be (L, r+1, r+1, r+1) with r >= 1")
    nl, r = sdf.shape[0], sdf.shape[1] - 1
    if base.shape != (nl, 3):
        raise NgsError("marching_cubes_sparse: base_coords must be (L, 3) with L = sdf_val.shape[0]")
    nc = nl * r ** 3
    if nc * _MAX_VERTS_PER_CUBE >= 2 ** 31:
        raise NgsError("marching_cubes_sparse: more than 2^31 / 15 sub-cubes")
    n = 0
    if nc:
        vcount = torch.empty(nc, dtype=torch.int32, device=dev)
        offs = torch.empty(nc + 1, dtype=torch.int32, device=dev)
        call("ngs_primal_mc_count", ptr(sdf), nl, r, ptr(vcount), stream())
        ws = Scratch.get(lib().ngs_scan_workspace_bytes(nc), dev)
        call("ngs_scan_i32", ptr(vcount), nc, ptr(offs), ptr(ws), stream())
        n = int(offs[nc].item())
    tris = torch.empty((n, 3), dtype=torch.float32, device=dev)
    vid = torch.empty((n, 4), dtype=torch.int32, device=dev)
    src = torch.empty((n, 2), dtype=torch.int32, device=dev)
    w = torch.empty(n, dtype=torch.float32, device=dev)
    if n:
        call("ngs_primal_mc_emit", ptr(base), ptr(sdf), nl, r, ptr(offs), ptr(tris), ptr(vid), ptr(src), ptr(w),
             stream())
    t = n // 3
    return tris.view(t, 3, 3), vid.view(t, 3, 4), src.view(t, 3, 2), w.view(t, 3, 1)


def _edge_key_layout(base: torch.Tensor, r: int):
    """Per-axis minimum and bit width of the edge-id coordinates int(b) r + [0, r] of the leaves."""
    mm = torch.empty(6, dtype=torch.float32, device=base.device)
    call("ngs_point_bounds", ptr(base), base.shape[0], 1.0, ptr(mm), stream())
    lo, hi = [int(v) for v in mm[:3].tolist()], [int(v) for v in mm[3:].tolist()]
    org = [a * r for a in lo]
    bits = [max(1, ((b - a) * r + r).bit_length()) for a, b in zip(lo, hi)]
    if sum(bits) + 2 > 64:
        raise NgsError("marching_cubes: the leaf box is too large for 64-bit edge keys")
    return (ctypes.c_int32 * 6)(*org, *bits)


def dedup_vertices(base: torch.Tensor, r: int, tris, vid, src, w):
    """Vertex de-duplication and normals of MarchingCubes.forward (ops.py:15-35): vertices (V, 3),
    triangles (T, 3) int64, normals (V, 3), unq (V, 4) int32, and per vertex its source samples (V, 2)
    and first-end weight (V,). Shared vertices take the record of their first occurrence in triangle
    order."""
    dev = tris.device
    n = vid.shape[0] * 3
    faces = torch.empty(n, dtype=torch.int64, device=dev)
    verts = torch.empty((n, 3), dtype=torch.float32, device=dev)
    normals = torch.empty((n, 3), dtype=torch.float32, device=dev)
    unq = torch.empty((n, 4), dtype=torch.int32, device=dev)
    vsrc = torch.empty((n, 2), dtype=torch.int32, device=dev)
    vw = torch.empty(n, dtype=torch.float32, device=dev)
    nv = torch.zeros(1, dtype=torch.int64, device=dev)
    if n:
        layout = _edge_key_layout(base.float().contiguous(), r)
        wsb = lib().ngs_primal_dedup_workspace_bytes(n)
        ws = Scratch.get(wsb, dev)
        call("ngs_primal_dedup", ptr(vid), n, layout, ptr(tris), ptr(src), ptr(w), ptr(faces), ptr(verts),
             ptr(normals), ptr(unq), ptr(vsrc), ptr(vw), ptr(nv), ptr(ws), wsb, stream())
    v = int(nv.item())
    return verts[:v], faces.view(-1, 3), normals[:v], unq[:v], vsrc[:v], vw[:v]


class MarchingCubes(torch.autograd.Function):
    """Differentiable primal sparse marching cubes (ops.py:9-58, the MeshSDF technique): the
    gradient of a vertex moves the field samples of its edge along the vertex normal."""

    @staticmethod
    def forward(ctx, base_coords, sdf_val, bandwidth=1.0):
        tris, vid, src, w = marching_cubes_sparse(base_coords, sdf_val)
        vertices, triangles, normals, unq, vsrc, vw = dedup_vertices(base_coords, sdf_val.shape[1] - 1, tris, vid,
                                                                     src, w)
        ctx.save_for_backward(normals, vsrc, vw)
        ctx.bandwidth = float(bandwidth)
        ctx.res = sdf_val.size()
        ctx.mark_non_differentiable(triangles, unq)
        return vertices, triangles, normals, unq

    @staticmethod
    def backward(ctx, grad_vert, grad_face, grad_normal, grad_unq):
        normals, vsrc, vw = ctx.saved_tensors
        grad_grid = torch.zeros(ctx.res, dtype=torch.float32, device=normals.device)
        if grad_vert is not None and normals.shape[0]:
            g = grad_vert.float().contiguous()
            call("ngs_primal_mc_backward", ptr(g), ptr(normals), ptr(vsrc), ptr(vw), normals.shape[0], ctx.bandwidth,
                 ptr(grad_grid), stream())
        return None, grad_grid, None


marching_cubes_op = MarchingCubes.apply
