"""Packed-column CSR (16-bit codes + plane bases) against the int32-column CSR (MI355X)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _systems(cuda):
    from ngspsr import synthetic
    from ngspsr.hashtree import HashTree
    from ngspsr.reconstructor import PointBinning, Reconstructor
    xyz, nrm = synthetic.ellipsoids(30000, k=4, sigma=0.002, seed=21)
    xyz_t = torch.from_numpy(xyz).to(cuda)
    tree = HashTree(xyz_t, voxel_size=1 / 128, depth=4)
    tree.build_encoder_hierarchy_adaptive(min_density=2.0)
    recon = Reconstructor(tree)
    recon.pack_columns = True
    systems = {}
    for d in range(tree.depth):
        M = tree.get_coords_size(None, d)
        if M:
            rhs = torch.zeros(M, dtype=torch.float64, device=cuda)
            systems[d] = recon.assemble_level(d, PointBinning(tree, xyz_t, d, 4.0), None, 0.25 ** d, 0.1, rhs,
                                              form="csr")
    return (tree, recon), systems


def test_packed_columns_decode_and_products(cuda):
    from ngspsr import solver
    print("[case] packed decode / spmv / solve", flush=True)
    res, systems = _systems(cuda)
    g = torch.Generator(device="cpu").manual_seed(4)
    for d, A in systems.items():
        assert A.packed, d
        rows = torch.repeat_interleave(torch.arange(A.M, device=cuda), (A.rowptr[1:] - A.rowptr[:-1]).long())
        code = A.codes.to(torch.int32) & 0xFFFF
        plane = code >> 13
        cols = A.pbase.view(-1, 8)[rows, plane.long()] + (code & 0x1FFF)
        assert torch.equal(cols, A.cols), d
        plain = solver.CSRSystem(A.rowptr, A.cols, A.vals, A.inv_diag)
        x = torch.randn(A.M, generator=g).to(cuda)
        assert torch.equal(A.matvec(x), plain.matvec(x)), d
        b = torch.randn(A.M, generator=g).to(cuda)
        xp, ip = solver.pcg(A, b, tol=1e-6)
        xc, ic = solver.pcg(plain, b, tol=1e-6)
        assert ip.iters == ic.iters and torch.equal(xp, xc), d


def test_packed_overflow_flag(cuda):
    """A level whose dx = 0 plane spans more than 8191 ranks within five y-rows (five dense
    z-lines of 9000 voxels) sets the overflow flag; the int32 columns stay exact."""
    from ngspsr._lib import Level, Scratch, call, lib, ptr, stream
    from ngspsr.bases import ensure_device_tables
    from ngspsr.grid import SparseGrid
    from ngspsr.reconstructor import PointBinning
    print("[case] packed overflow", flush=True)
    ensure_device_tables(cuda)
    nz = 9000
    z = torch.arange(nz, dtype=torch.int32)
    coords = torch.stack([torch.stack([torch.zeros_like(z), torch.full_like(z, y), z], 1) for y in range(5)])
    coords = coords.reshape(-1, 3).to(cuda)
    grid = SparseGrid([-2, -2, -2], [2, 6, nz + 2], 1, cuda)
    grid.mark_coords(coords)
    M = grid.finalize()
    lv = Level()
    lv.grid = grid.struct()
    lv.coords = grid.coords.data_ptr()
    lv.M = M
    lv.stride = 1

    class _T:
        device = cuda
    bins = PointBinning.empty(_T)
    row_nnz = torch.empty(M, dtype=torch.int32, device=cuda)
    call("ngs_lhs_row_counts", lv, ptr(row_nnz), stream())
    rowptr = torch.empty(M + 1, dtype=torch.int32, device=cuda)
    ws = Scratch.get(lib().ngs_scan_workspace_bytes(M), cuda)
    call("ngs_scan_i32", ptr(row_nnz), M, ptr(rowptr), ptr(ws), stream())
    nnz = int(rowptr[M].item())
    cols = torch.empty(nnz, dtype=torch.int32, device=cuda)
    vals = torch.empty(nnz, dtype=torch.float32, device=cuda)
    codes = torch.empty(nnz, dtype=torch.int16, device=cuda)
    pbase = torch.empty(8 * M, dtype=torch.int32, device=cuda)
    overflow = torch.zeros(1, dtype=torch.int32, device=cuda)
    inv_diag = torch.empty(M, dtype=torch.float32, device=cuda)
    rhs = torch.zeros(M, dtype=torch.float64, device=cuda)
    call("ngs_lhs_assemble_packed", lv, bins.struct(), None, 1.0, 0.1, ptr(rowptr), ptr(cols), ptr(codes), ptr(pbase),
         ptr(overflow), ptr(vals), ptr(inv_diag), ptr(rhs), stream())
    assert int(overflow.item()) == 1
    # int32 columns: every column's voxel is within the 5^3 stencil of its row
    rows = torch.repeat_interleave(torch.arange(M, device=cuda), (rowptr[1:] - rowptr[:-1]).long())
    delta = grid.coords[cols.long()] - grid.coords[rows]
    assert int(delta.abs().max()) <= 2
