"""Galerkin assembly and the coarse-to-fine solve of the jspsr path (oracle; test infrastructure only).

Restates jspsr/core/reconstructor.py:59-195 (solve_multigrid, _evaluate_screening_term) and
jspsr/__init__.py:38-54 (splat of normals, sample weights) in float64 with numpy/scipy.
Matrices are COO (src, tgt, val) in the reference's neighbour order; the linear solve is
`pcg.solve_pcg` (csrc/sparse_solve/solve_pcg.cu semantics) or a direct solve.
"""
import numpy as np
import scipy.sparse as sp

from . import bezier
from .hierarchy import Hierarchy, compute_density, range_kernel  # noqa: F401


def screening_term(sd_a, sd_b, src, tgt, m_b, pts_weight):
    """reconstructor.py:59-77 + sparse_product.cu:100-137: for each COO pair (src, tgt) the sum over
    points p of w_p * a(p, src) * b(p, tgt); pairs absent from the COO are dropped (silent miss).
    The point sums are the sparse product A^T W B of the (point x voxel) evaluation matrices; a
    pair listed twice in the COO receives the sum once, at its first occurrence (the reference's
    map from pair to COO index keeps one index per pair)."""
    pa, va, xa = sd_a
    pb, vb, xb = sd_b
    E = src.shape[0]
    if E == 0 or pa.shape[0] == 0 or pb.shape[0] == 0:
        return np.zeros(E)
    n = int(max(pa.max(), pb.max())) + 1
    w = np.broadcast_to(np.asarray(pts_weight, np.float64), (n,)) if np.ndim(pts_weight) == 0 else np.asarray(pts_weight, np.float64)
    m_a = int(max(int(np.max(src)), int(va.max()))) + 1
    at = sp.csr_matrix((w[pa] * np.asarray(xa, np.float64), (va, pa)), shape=(m_a, n))
    b = sp.csr_matrix((np.asarray(xb, np.float64), (pb, vb)), shape=(n, m_b))
    s_ = (at @ b).tocsr()
    s_.sum_duplicates()
    s_.sort_indices()
    rows = np.repeat(np.arange(m_a, dtype=np.int64), np.diff(s_.indptr))
    skey = rows * m_b + s_.indices.astype(np.int64)
    key = src.astype(np.int64) * m_b + tgt
    order = np.argsort(key, kind="stable")
    ks = key[order]
    first = np.ones(E, bool)
    first[1:] = ks[1:] != ks[:-1]
    out = np.zeros(E)
    if skey.shape[0] == 0:
        return out
    pos = np.minimum(np.searchsorted(skey, ks), skey.shape[0] - 1)
    hit = (skey[pos] == ks) & first
    out[order[hit]] = s_.data[pos[hit]]
    return out


def coo_matvec(src, tgt, val, x, m_src):
    return np.bincount(src, weights=val * x[tgt], minlength=m_src)


def coo_to_csr(src, tgt, val, m):
    """solver.py:40-48 (row-compressed; column order inside a row is not specified by the reference,
    whose argsort is unstable)."""
    a = sp.csr_matrix((val, (src, tgt)), shape=(m, m))
    a.sum_duplicates()
    return a


class MultigridResult:
    def __init__(self):
        self.solutions = {}
        self.lhs = {}
        self.rhs = {}
        self.rhs_parts = {}
        self.iters = {}


def splat_normals(h: Hierarchy, xyz, normal, xyz_depth, sample_weight):
    """__init__.py:38-48: per level, splat normal * w of the points on that level, / stride^3."""
    data = {}
    for d in range(h.depth):
        mask = xyz_depth == d
        if not np.any(mask):
            continue
        nd = h.splat(np.asarray(xyz)[mask], d, np.asarray(normal, np.float64)[mask] * sample_weight[mask, None])
        data[d] = nd / (h.strides[d] ** 3)
    return data


def solve_multigrid(h: Hierarchy, normal_data: dict, screen_xyz, screen_alpha, screen_delta=0.1,
                    solve=None, start_depth=None, end_depth=0):
    """reconstructor.py:79-195 in float64. `solve(a_csr, b) -> (x, iters)`."""
    from . import pcg
    if solve is None:
        def solve(a, b):
            return pcg.solve_pcg(a, b, tol=1e-6, max_iter=10000, dtype=np.float64)
    if start_depth is None:
        start_depth = h.depth - 1
    res = MultigridResult()
    alpha_arr = screen_alpha
    should_screen = np.ndim(screen_alpha) > 0 or screen_alpha > 0.0
    screen = {}
    if should_screen:
        for d in range(end_depth, start_depth + 1):
            pid, vid, val, _ = h.point_basis(screen_xyz, d)
            screen[d] = (pid, vid, val)
    for d in range(start_depth, end_depth - 1, -1):
        sf = (1 / 4.0) ** d
        M = h.num_voxels(d)
        rhs_n = np.zeros(M)
        for dd, nd in normal_data.items():
            sid, tid, rel, _ = h.self_neighbours(dd, d, 3)
            part = bezier.integrate_const_deriv_product(-nd[sid], rel, h.strides[dd], h.strides[d])
            rhs_n += np.bincount(tid, weights=part, minlength=M)
        rhs_s = np.zeros(M)
        if should_screen:
            pid, vid, val = screen[d]
            mult = (np.asarray(screen_delta) * np.asarray(alpha_arr))
            mult = mult[pid] if np.ndim(mult) > 0 else mult
            rhs_s = sf * np.bincount(vid, weights=val * mult, minlength=M)
        rhs_c = np.zeros(M)
        for dd in range(start_depth, d, -1):
            sid, tid, rel, _ = h.self_neighbours(d, dd, 5)
            a = bezier.integrate_deriv_deriv_product(rel, h.strides[d], h.strides[dd])
            if should_screen:
                a = a + sf * screening_term(screen[d], screen[dd], sid, tid, h.num_voxels(dd), alpha_arr)
            rhs_c -= coo_matvec(sid, tid, a, res.solutions[dd], M)
        rhs = rhs_n + rhs_s + rhs_c
        sid, tid, rel, _ = h.self_neighbours(d, d, 5)
        lhs = bezier.integrate_deriv_deriv_product(rel, h.strides[d], h.strides[d])
        if should_screen:
            lhs = lhs + sf * screening_term(screen[d], screen[d], sid, tid, M, alpha_arr)
        a = coo_to_csr(sid, tid, lhs, M)
        x, it = solve(a, rhs) if M > 0 else (np.zeros(0), 0)
        res.solutions[d] = np.asarray(x, np.float64)
        res.lhs[d] = a
        res.rhs[d] = rhs
        res.rhs_parts[d] = (rhs_n, rhs_s, rhs_c)
        res.iters[d] = it
    return res


def build_tree(xyz, voxel_size, depth, min_density):
    """hashtree.py:266-287 + reflect: density, adaptive levels. Returns (h, xyz_depth, density)."""
    h = Hierarchy(depth, voxel_size)
    rho = compute_density(xyz, voxel_size)
    xyz_depth = h.build_adaptive(xyz, rho, min_density)
    return h, xyz_depth, rho
