"""Byte-staged grid marking against the atomic marking: the same bitmap, ranks and member count, and
the byte map all zero after every staged call (the state the next staged call relies on)."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _grid(lo, hi, stride):
    from ngspsr.grid import SparseGrid
    return SparseGrid(lo, hi, stride, torch.device("cuda:0"))


def _call_pair(atomic, staged, lo, hi, stride):
    from ngspsr._lib import Marks
    ga, gs = _grid(lo, hi, stride), _grid(lo, hi, stride)
    for g in (ga, gs):
        g.bits.zero_()
    atomic(ga)
    m = Marks.get(gs.nwords * 32, gs.device)
    staged(gs, m)
    torch.cuda.synchronize()
    assert torch.equal(ga.bits, gs.bits)
    assert int(m.count_nonzero()) == 0
    assert ga.finalize() == gs.finalize()
    assert torch.equal(ga.rank, gs.rank) and torch.equal(ga.coords, gs.coords)
    return ga.count


@pytest.mark.parametrize("stride", [1, 4])
def test_staged_coords_equal_atomic(stride):
    from ngspsr._lib import call, ptr, stream
    gen = torch.Generator().manual_seed(11)
    lo, hi = [-5, -7, -3], [40, 33, 70]
    for n in (1, 777, 50000):  # the second call of a size reuses the zeroed byte map
        c = torch.stack([torch.randint(l - 2, h + 3, (n,), generator=gen) for l, h in zip(lo, hi)], 1)
        c = (c * stride).to(torch.int32).cuda()
        cnt = _call_pair(lambda g: call("ngs_grid_mark_coords", g.struct(), ptr(c), n, stream()),
                         lambda g, m: call("ngs_grid_mark_coords_staged", g.struct(), ptr(c), n, ptr(m), stream()),
                         lo, hi, stride)
        inside = [(c[:, a] // stride >= lo[a]) & (c[:, a] // stride <= hi[a]) for a in range(3)]
        expect = torch.unique(c[inside[0] & inside[1] & inside[2]], dim=0).shape[0]
        assert cnt == expect


@pytest.mark.parametrize("stride", [1, 2, 16])
def test_staged_corners_equal_atomic(stride):
    from ngspsr._lib import call, ptr, stream
    gen = torch.Generator().manual_seed(5)
    n = 20000
    xyz = (torch.rand((n, 3), generator=gen) * 60.0 - 20.0).cuda()
    levels = torch.randint(0, 4, (n,), generator=gen, dtype=torch.int32).cuda()
    lo, hi = [-30 // stride - 2] * 3, [50 // stride + 2] * 3
    for lv, maxl in ((None, 0), (levels, 1), (levels, 3)):
        _call_pair(lambda g: call("ngs_grid_mark_corners", g.struct(), ptr(xyz), n, 1.0, ptr(lv), maxl, stream()),
                   lambda g, m: call("ngs_grid_mark_corners_staged", g.struct(), ptr(xyz), n, 1.0, ptr(lv), maxl,
                                     ptr(m), stream()),
                   lo, hi, stride)
