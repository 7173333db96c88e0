"""CPU oracle for the jspsr screened-Poisson reconstruction path.

TEST INFRASTRUCTURE ONLY. This package restates, on the CPU, the algorithm of the reference
(huangjh-pub/neural-galerkin, package `jspsr`) for the `reconstruct` path. Only `tests/`,
`__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py` may import it, and only as the
checker (or the reported CPU baseline). The product package `ngspsr` never imports it and fails
loudly when its HIP library is missing.

Modules and what pins them (the reference ships no tests or golden vectors):
  fnv          FNV-1a-64 fold60 voxel key, kernel-offset hashing, cuckoo table sizing
               (csrc/common/hashmap_cuda.cu:204-360). Pinned by known-answer vectors and
               by an independent pure-Python FNV implementation.
  bezier       degree-2 B-spline basis and its Galerkin integral tables, exact rationals
               (jspsr/bases/bezier_tensor.py, additive.py:17-52,172-249 conventions). Pinned by
               numerical quadrature, textbook stencils, sum rules and two-scale identities.
  hierarchy    voxelisation, density, adaptive levels, neighbour COO, splat
               (jspsr/core/_hash_backend.py, hashtree.py). Pinned by brute force on tiny inputs
               and by invariants.
  assembly     RHS, stiffness, screening and cross-level correction, multigrid driver
               (jspsr/core/reconstructor.py:59-195). Pinned by brute-force 3-D quadrature and
               dense FP64 assembly on tiny scenes.
  pcg          Jacobi-PCG mirroring csrc/sparse_solve/solve_pcg.cu:15-136 (C, OpenMP).
               Pinned by scipy.sparse.linalg.cg and dense solves.
  mesh         chi evaluation, conformal leaves, dual centres, dual-cube assembly and dual MC
               (reconstructor.py:197-397, csrc/marching_cubes/dual_mc*.cu). Pinned by brute-force
               enumeration, table consistency and closed-surface invariants.
  mc_tables    the standard MC case table (packed form).
"""
