"""The plain jspsr.reconstruct path end to end on the CPU (oracle; test infrastructure only).

Follows jspsr/__init__.py:8-61: adaptive hierarchy from density, per-level normal splat weighted by
1/density, coarse-to-fine screened Poisson solve (Jacobi-PCG), dual marching cubes with
n_upsample=2.
"""
import numpy as np

from . import assembly, mesh, pcg
from .hierarchy import Hierarchy, compute_density


class OracleResult:
    pass


def reconstruct(xyz, normal, voxel_size, depth=4, min_density=32.0, screen_alpha=4.0, tol=1e-6,
                max_iter=10000, dtype=np.float64, xyz_depth=None, extract=True):
    xyz = np.asarray(xyz, np.float32)
    normal = np.asarray(normal, np.float32)
    out = OracleResult()
    h = Hierarchy(depth, voxel_size)
    rho = compute_density(xyz, voxel_size)
    if xyz_depth is None:
        xyz_depth = h.build_adaptive(xyz, rho, min_density)
    else:
        from .hierarchy import quantize_coords
        for d in range(depth):
            h.set_coords(d, quantize_coords(xyz[xyz_depth <= d], voxel_size, h.strides[d]))
    w = 1.0 / rho
    nd = assembly.splat_normals(h, xyz, normal, xyz_depth, w)

    def solve(a, b):
        return pcg.solve_pcg(a, b, tol=tol, max_iter=max_iter, dtype=dtype)

    res = assembly.solve_multigrid(h, nd, xyz, screen_alpha, 0.1, solve)
    out.tree, out.xyz_depth, out.density, out.sample_weight = h, xyz_depth, rho, w
    out.normal_data, out.solve = nd, res
    if extract:
        out.v, out.f, out.mesh_info = mesh.extract_multiscale_mesh(h, res.solutions, w, xyz)
    return out
