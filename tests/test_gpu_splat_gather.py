"""The gather form of the trilinear splat against the atomic scatter: the same FP64 sums (to FP64
rounding: both add FP32 terms, in different orders), with and without point data, with a level
filter, and with points whose stencil's lower node is missing from the grid (the gather's atomic
fallback)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _points(n, seed):
    gen = torch.Generator().manual_seed(seed)
    xyz = torch.rand((n, 3), generator=gen) * 50.0 - 10.0
    xyz[:7] = torch.tensor([[0.5, 0.5, 0.5], [1.0, 2.0, 3.0], [4.5, 4.0, 4.5], [0.0, 0.0, 0.0], [8.0, 8.5, 9.0],
                            [2.5, 2.5, 2.5], [-3.5, 7.0, 1.5]])  # on cell centres and lattice planes (sign 0)
    data = torch.randn((n, 3), generator=gen)
    scale = torch.rand(n, generator=gen) + 0.5
    levels = torch.randint(0, 3, (n,), generator=gen, dtype=torch.int32)
    return xyz.cuda(), data.cuda(), scale.cuda(), levels.cuda()


@pytest.mark.parametrize("stride", [1, 2, 8])
@pytest.mark.parametrize("ncomp", [0, 1, 3])
def test_splat_gather_equals_scatter(stride, ncomp):
    from ngspsr._lib import call, lib, ptr, stream
    from ngspsr.grid import SparseGrid
    n = 30000
    xyz, data, scale, levels = _points(n, 3 + stride)
    dev = xyz.device
    lo, hi = [-12 // stride - 2] * 3, [42 // stride + 2] * 3
    for mark_level, level in ((2, 1), (0, 1)):  # (0, 1): level-1 points' stencils not marked, mostly fallback
        g = SparseGrid(lo, hi, stride, dev)
        g.mark_corners(xyz, 1.0, levels, mark_level)
        g.finalize()
        d = data[:, :ncomp].contiguous() if ncomp else None
        shape = (g.count, ncomp) if ncomp else (g.count,)
        ref = torch.zeros(shape, dtype=torch.float64, device=dev)
        got = torch.zeros(shape, dtype=torch.float64, device=dev)
        call("ngs_splat", g.struct(), ptr(xyz), n, 1.0, ptr(d), ncomp, ptr(scale), ptr(levels), level, 1.0, ptr(ref),
             stream())
        wsb = lib().ngs_splat_gather_workspace_bytes(n, g.count)
        assert wsb > 0
        ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
        call("ngs_splat_gather", g.struct(), ptr(g.coords), g.count, ptr(xyz), n, 1.0, ptr(d), ncomp, ptr(scale),
             ptr(levels), level, 1.0, ptr(got), ptr(ws), wsb, stream())
        r, o = ref.cpu().numpy(), got.cpu().numpy()
        assert np.abs(r).max() > 0
        np.testing.assert_allclose(o, r, rtol=1e-12, atol=1e-14 * np.abs(r).max())


def test_splat_gather_density_all_points():
    """No level filter, no data (the density splat): every point reaches its 8 nodes."""
    from ngspsr._lib import call, lib, ptr, stream
    from ngspsr.grid import SparseGrid
    n = 50000
    xyz, _, _, _ = _points(n, 9)
    dev = xyz.device
    g = SparseGrid([-14] * 3, [44] * 3, 1, dev)
    g.mark_corners(xyz, 1.0)
    g.finalize()
    got = torch.zeros(g.count, dtype=torch.float64, device=dev)
    wsb = lib().ngs_splat_gather_workspace_bytes(n, g.count)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    call("ngs_splat_gather", g.struct(), ptr(g.coords), g.count, ptr(xyz), n, 1.0, None, 0, None, None, 0, 1.0,
         ptr(got), ptr(ws), wsb, stream())
    # the trilinear weights of a point sum to 1
    assert abs(float(got.sum()) - n) < 1e-6 * n


def test_splat_gather_duplicated_cloud():
    """A cloud where most points repeat a few positions: the nodes around those positions gather
    long bins (one thread walks all of their points), and the sums still equal the scatter's."""
    from ngspsr._lib import call, lib, ptr, stream
    from ngspsr.grid import SparseGrid
    n = 40000
    xyz, data, scale, _ = _points(n, 21)
    xyz[: n // 2] = torch.tensor([3.25, 4.5, 5.75], device=xyz.device)   # half the cloud on one point
    xyz[n // 2: 3 * n // 4] = torch.tensor([7.0, 7.0, 7.0], device=xyz.device)  # a quarter on a lattice node
    dev = xyz.device
    g = SparseGrid([-14] * 3, [44] * 3, 1, dev)
    g.mark_corners(xyz, 1.0)
    g.finalize()
    for ncomp in (0, 3):
        d = data[:, :ncomp].contiguous() if ncomp else None
        shape = (g.count, ncomp) if ncomp else (g.count,)
        ref = torch.zeros(shape, dtype=torch.float64, device=dev)
        got = torch.zeros(shape, dtype=torch.float64, device=dev)
        call("ngs_splat", g.struct(), ptr(xyz), n, 1.0, ptr(d), ncomp, ptr(scale), None, 0, 1.0, ptr(ref), stream())
        wsb = lib().ngs_splat_gather_workspace_bytes(n, g.count)
        ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
        call("ngs_splat_gather", g.struct(), ptr(g.coords), g.count, ptr(xyz), n, 1.0, ptr(d), ncomp, ptr(scale),
             None, 0, 1.0, ptr(got), ptr(ws), wsb, stream())
        r, o = ref.cpu().numpy(), got.cpu().numpy()
        np.testing.assert_allclose(o, r, rtol=1e-12, atol=1e-14 * np.abs(r).max())


def test_splat_gather_large_cloud():
    """400k points (their bin sort takes the onesweep radix path): the sums equal the scatter's."""
    from ngspsr._lib import call, lib, ptr, stream
    from ngspsr.grid import SparseGrid
    n = 400_000
    xyz, data, scale, _ = _points(n, 33)
    dev = xyz.device
    g = SparseGrid([-14] * 3, [44] * 3, 1, dev)
    g.mark_corners(xyz, 1.0)
    g.finalize()
    ref = torch.zeros((g.count, 3), dtype=torch.float64, device=dev)
    got = torch.zeros((g.count, 3), dtype=torch.float64, device=dev)
    call("ngs_splat", g.struct(), ptr(xyz), n, 1.0, ptr(data), 3, ptr(scale), None, 0, 1.0, ptr(ref), stream())
    wsb = lib().ngs_splat_gather_workspace_bytes(n, g.count)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    call("ngs_splat_gather", g.struct(), ptr(g.coords), g.count, ptr(xyz), n, 1.0, ptr(data), 3, ptr(scale), None, 0,
         1.0, ptr(got), ptr(ws), wsb, stream())
    r, o = ref.cpu().numpy(), got.cpu().numpy()
    np.testing.assert_allclose(o, r, rtol=1e-12, atol=1e-14 * np.abs(r).max())
