"""ngspsr: MI355X-native screened-Poisson surface reconstruction (the jspsr `reconstruct` path).

Drop-in for jspsr.reconstruct (jspsr/__init__.py:8-61): points and normals in, triangle mesh out,
with every stage on hand-written gfx950 HIP kernels behind the C ABI of include/ngspsr.h.
"""
import numpy as np
import torch

from ._lib import NgsError, lib  # noqa: F401
from .grid import CuckooHashTable, SparseGrid  # noqa: F401
from .hashtree import HashTree
from .reconstructor import Reconstructor
from .solver import parse_solver  # noqa: F401

__version__ = "1.0.0"
__version_info__ = (1, 0, 0)


def _to_device(a, device):
    if isinstance(a, torch.Tensor):
        return a.detach().to(device=device, dtype=torch.float32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float32))).to(device)


class ReconstructionResult:
    """Sparse solution field (per-level voxel coordinates and basis coefficients) and mesh."""

    def __init__(self, tree, recon, v, f):
        self.tree, self.reconstructor, self.vertices, self.triangles = tree, recon, v, f

    @property
    def solution_field(self):
        return {d: (self.tree.get_coords(None, d), x) for d, x in self.reconstructor.solutions.items()}

    @property
    def solve_info(self):
        return self.reconstructor.solve_info


def reconstruct(xyz, normal, voxel_size: float, depth: int = 4, min_density: float = 32.0, screen_alpha: float = 4.0,
                *, degree: int = 2, solver: str = "pcg@maxiter=100000", n_upsample: int = 2, xyz_depth=None,
                operator: str = "auto", return_result: bool = False, device=None):
    """Screened Poisson reconstruction of oriented points (jspsr/__init__.py:8-61).

    :param xyz: (N, 3) points (numpy array or torch tensor)
    :param normal: (N, 3) normals
    :param voxel_size: width of a finest-level voxel
    :param depth: number of tree levels
    :param min_density: density threshold of the adaptive levels
    :param screen_alpha: screening weight
    :param degree: basis degree (only 2, the reference's BezierTensorBasis, is available)
    :param solver: solver string ("pcg@tol=1e-6@maxiter=10000", "mixed", "cholmod", "refine@tol=1e-8", ...); the default
        keeps the reference's PCG tolerance (1e-5 of ||b||) with an iteration cap
    :param xyz_depth: optional imposed per-point levels (skips the density-derived assignment)
    :param operator: level operator form, "runs" (assembled matrix, one run of consecutive columns
        per bitmap column of the stencil), "csr" (assembled matrix in the reference's CSR layout),
        "stencil" (applied from the voxel bitmap and the binned points) or "auto" (cheapest per
        iteration: the run form at the sizes measured)
    :return: (vertices (V, 3), triangles (T, 3)) as numpy arrays for numpy input, device tensors
        otherwise; with return_result, a ReconstructionResult holding the solution field too.
    """
    if degree != 2:
        raise NotImplementedError("only the degree-2 B-spline basis is available (the reference's BezierTensorBasis)")
    if not (voxel_size > 0 and np.isfinite(voxel_size)):
        raise ValueError(f"reconstruct: voxel_size must be positive and finite, got {voxel_size}")
    if int(depth) != depth or depth < 1:
        raise ValueError(f"reconstruct: depth must be a positive integer, got {depth}")
    if isinstance(xyz, torch.Tensor) and isinstance(normal, torch.Tensor) and xyz.device == normal.device:
        finite = bool(torch.isfinite(xyz).all() & torch.isfinite(normal).all())  # one device read
    else:
        finite = all(bool(torch.isfinite(a).all()) if isinstance(a, torch.Tensor) else bool(np.isfinite(np.asarray(a)).all())
                     for a in (xyz, normal))
    if not finite:
        raise ValueError("reconstruct: xyz or normal holds non-finite values")
    is_numpy = not isinstance(xyz, torch.Tensor)
    if device is None:
        device = xyz.device if (isinstance(xyz, torch.Tensor) and xyz.is_cuda) else torch.device("cuda", torch.cuda.current_device())
    xyz_t = _to_device(xyz, device)
    nrm_t = _to_device(normal, device)
    assert xyz_t.shape[0] == nrm_t.shape[0] and xyz_t.shape[1] == nrm_t.shape[1] == 3, \
        "Input positions and normals should have same size (N, 3)!"
    if xyz_t.shape[0] == 0:
        # no points, no surface: an empty mesh of the usual types (nothing is launched)
        v = torch.zeros((0, 3), dtype=torch.float32, device=device)
        f = torch.zeros((0, 3), dtype=torch.int64, device=device)
        if return_result:
            raise ValueError("reconstruct: no input points, so there is no solution field to return")
        return (v.cpu().numpy(), f.cpu().numpy()) if is_numpy else (v, f)
    with torch.cuda.device(device):
        tree = HashTree(xyz_t, voxel_size=voxel_size, depth=depth)
        tree.build_encoder_hierarchy_adaptive(min_density=min_density, xyz_depth=xyz_depth)
        tree.reflect_decoder_coords()
        sample_weight = 1.0 / tree.xyz_density
        counts = tree.level_point_counts()
        normal_data = {}
        for d in range(depth):
            if counts[d] == 0:
                continue
            nd = tree.splat_data(xyz_t, None, d, nrm_t, scale=sample_weight, levels=tree.xyz_depth, level=d)
            normal_data[d] = nd / float(tree.strides[d] ** 3)
        recon = Reconstructor(tree)
        recon.operator = operator
        recon.sample_weight = sample_weight
        recon.solve_multigrid(depth - 1, 0, normal_data, screen_alpha=screen_alpha, screen_xyz=xyz_t, solver=solver)
        v, f = recon.extract_multiscale_mesh(n_upsample=n_upsample, build_o3d_mesh=False)
    if return_result:
        return ReconstructionResult(tree, recon, v, f)
    if is_numpy:
        return v.cpu().numpy(), f.cpu().numpy()
    return v, f
