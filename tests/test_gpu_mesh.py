"""Vertex de-duplication of the dual marching cubes (MI355X): the sorted distinct keys, each vertex
interpolated from its first occurrence in that occurrence's corner order, and every triangle
vertex's rank, checked against numpy on repeated and on all-distinct keys."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _expected(keys, tags, pos, val, id_bits):
    from oracle.mesh import _interp_vec
    uniq, first, inv = np.unique(keys, return_index=True, return_inverse=True)
    hi = (uniq >> id_bits).astype(np.int64)
    lo = (uniq & ((1 << id_bits) - 1)).astype(np.int64)
    swap = (tags[first] >> 31) & 1
    ia = np.where(swap == 1, hi, lo)
    ib = np.where(swap == 1, lo, hi)
    return _interp_vec(pos[ia], pos[ib], val[ia], val[ib]), inv.astype(np.int64)


@pytest.mark.parametrize("distinct", [False, True])
def test_dedup_vertices(cuda, distinct):
    from ngspsr._lib import call, lib, ptr, stream
    print(f"[case] dedup distinct={distinct}", flush=True)
    rng = np.random.default_rng(3)
    nprim, id_bits = 5000, 13
    pos = rng.random((nprim, 3), dtype=np.float32)
    val = rng.standard_normal(nprim).astype(np.float32)
    val[:50] = 0.0  # exercise the snapping branches
    npairs = 3000 if distinct else 600
    a = rng.choice(nprim, size=npairs * 2, replace=True).reshape(-1, 2)
    a = a[a[:, 0] != a[:, 1]]
    pairs = np.unique(np.sort(a, axis=1), axis=0)
    if distinct:
        occ = rng.permutation(pairs.shape[0])  # every key once
    else:
        occ = rng.integers(0, pairs.shape[0], size=6000)  # repeated keys
    p = pairs[occ]
    flip = rng.integers(0, 2, size=p.shape[0]).astype(bool)
    ia = np.where(flip, p[:, 1], p[:, 0])
    ib = np.where(flip, p[:, 0], p[:, 1])
    keys = (np.maximum(ia, ib).astype(np.uint64) << np.uint64(id_bits)) | np.minimum(ia, ib).astype(np.uint64)
    n = keys.shape[0]
    tags = (np.arange(n, dtype=np.uint64) | ((ia > ib).astype(np.uint64) << np.uint64(31))).astype(np.uint32)
    k_t = torch.from_numpy(keys.view(np.int64)).to(cuda)
    t_t = torch.from_numpy(tags.view(np.int32)).to(cuda)
    pos_t, val_t = torch.from_numpy(pos).to(cuda), torch.from_numpy(val).to(cuda)
    faces = torch.empty(n, dtype=torch.int64, device=cuda)
    verts = torch.empty((n, 3), dtype=torch.float32, device=cuda)
    nvert = torch.empty(1, dtype=torch.int64, device=cuda)
    wsb = lib().ngs_dedup_workspace_bytes(n)
    ws = torch.empty(wsb, dtype=torch.uint8, device=cuda)
    call("ngs_dedup_vertices", ptr(k_t), ptr(t_t), n, id_bits, ptr(pos_t), ptr(val_t), ptr(faces), ptr(verts),
         ptr(nvert), ptr(ws), wsb, stream())
    V = int(nvert.item())
    ev, ef = _expected(keys, tags, pos, val, id_bits)
    assert V == ev.shape[0]
    np.testing.assert_array_equal(faces.cpu().numpy(), ef)
    np.testing.assert_array_equal(verts[:V].cpu().numpy(), ev)
