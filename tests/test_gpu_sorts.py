"""Point bucketing (a stable sort of the points by cell rank, then the cell starts) against numpy's
stable argsort, on both sides of the switch from merge sort to the onesweep radix sort (256k items,
keys of at most 24 bits) and past rocPRIM's own merge-sort limit (1M items)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,ncells", [(100_000, 5_000), (262_143, 1 << 17), (262_144, 1 << 17), (300_000, 70_000),
                                      (700_000, (1 << 24) - 3), (1_000_000, 30_000_000), (1_200_000, 400_000)])
def test_bucket_points_stable(n, ncells):
    from ngspsr._lib import call, lib, ptr, stream
    rng = np.random.default_rng(n + ncells)
    rank = rng.integers(0, ncells, n, dtype=np.int64).astype(np.int32)
    rank[: n // 10] = rank[0]  # a long run of one cell: ties keep the input order
    rank[-1] = ncells - 1
    dev = torch.device("cuda:0")
    r = torch.from_numpy(rank).to(dev)
    perm = torch.empty(n, dtype=torch.int32, device=dev)
    cs = torch.empty(ncells + 1, dtype=torch.int32, device=dev)
    wsb = lib().ngs_bucket_workspace_bytes(n, ncells)
    assert wsb > 0
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    print(f"[case] bucket n={n} ncells={ncells}", flush=True)
    call("ngs_bucket_points", ptr(r), n, ncells, ptr(perm), ptr(cs), ptr(ws), wsb, stream())
    ref = np.argsort(rank, kind="stable")
    assert np.array_equal(perm.cpu().numpy(), ref)
    starts = np.searchsorted(rank[ref], np.arange(ncells + 1), side="left")
    assert np.array_equal(cs.cpu().numpy(), starts)
