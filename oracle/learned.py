"""Multigrid solve with a learned basis (oracle; test infrastructure only).

Restates, in float64 with numpy, the generic path of jspsr/core/reconstructor.py:79-195 (every level
system assembled from COO neighbour pairs with the basis's Galerkin integrals, the screening pairs
and the cross-level correction) for an additive basis with per-voxel features, with the basis
methods of jspsr/bases/additive.py:293-404 (evaluate, integrate_deriv_deriv_product,
integrate_const_deriv_product). The integral tables are inputs (their entries are pinned against
exact symbolic integrals by the additive-basis oracle tests).
"""
import numpy as np

from . import additive
from .assembly import MultigridResult, coo_to_csr, screening_term
from .hierarchy import Hierarchy, coords_neighbours, range_kernel, scaled_positions

MULTS = [1, 2, 4, 8, 16]


class AdditiveBasisOracle:
    """comps: oracle component tuples; ff, dd: per multiple index, (C, C, 3 * 2^m + 2) tables of
    f_f and df_df; f, df: per component, per multiple, the RHS tables; f_inv, df_inv the same for
    coarser data (multiples 2..16)."""

    def __init__(self, comps, ff, dd, f, df, f_inv, df_inv):
        self.comps = comps
        self.ff, self.dd = [np.asarray(a, np.float64) for a in ff], [np.asarray(a, np.float64) for a in dd]
        self.f, self.df = f, df
        self.f_inv, self.df_inv = f_inv, df_inv

    def evaluate(self, feat, xyz, ids):
        fv, _ = additive.additive_eval(self.comps, xyz, feat, ids)
        return fv[:, 0] * fv[:, 1] * fv[:, 2]

    def evaluate_derivative(self, feat, xyz, ids, stride):
        fv, dfv = additive.additive_eval(self.comps, xyz, feat, ids, stride)
        return np.stack([dfv[:, 0] * fv[:, 1] * fv[:, 2], fv[:, 0] * dfv[:, 1] * fv[:, 2],
                         fv[:, 0] * fv[:, 1] * dfv[:, 2]], 1)

    def integrate_dd(self, sfeat, tfeat, rel, ss, ts, sids, tids):
        mult = ts // ss
        idx = np.round(np.asarray(rel, np.float64) + mult * 3 / 2 + 0.5).astype(np.int64)
        m = MULTS.index(mult)
        K = self.ff[m].shape[-1]
        invalid = ~((idx >= 0) & (idx < K))
        idx = np.clip(idx, 0, K - 1)
        p, q = additive.ddi_forward(np.asarray(sfeat, np.float64)[sids], np.asarray(tfeat, np.float64)[tids], idx,
                                    self.ff[m] * ss, self.dd[m] / ss)
        p[invalid] = 0.0
        q[invalid] = 0.0
        return q[:, 0] * p[:, 1] * p[:, 2] + p[:, 0] * q[:, 1] * p[:, 2] + p[:, 0] * p[:, 1] * q[:, 2]

    def integrate_cd(self, data, tfeat, rel, ds, ts, tids):
        rel = np.asarray(rel, np.float64)
        if ts >= ds:
            mult = ts // ds
            idx = np.round(rel + (3 * mult - 1) / 2.0).astype(np.int64)
            m = MULTS.index(mult)
            ft, dft = self.f, self.df
        else:
            mult = ds // ts
            idx = np.round(rel * mult + mult / 2 + 0.5).astype(np.int64)
            m = MULTS.index(mult) - 1
            ft, dft = self.f_inv, self.df_inv
            ds, ts = ts, ds
        t = np.asarray(tfeat, np.float64)[tids]
        p = 0.0
        q = 0.0
        for ci in range(len(self.comps)):
            p = p + t[:, 3 * ci:3 * ci + 3] * (np.asarray(ft[ci][m], np.float64)[idx] * ds)
            q = q + t[:, 3 * ci:3 * ci + 3] * np.asarray(dft[ci][m], np.float64)[idx]
        data = np.asarray(data, np.float64)
        return data[:, 0] * q[:, 0] * p[:, 1] * p[:, 2] + data[:, 1] * p[:, 0] * q[:, 1] * p[:, 2] + \
            data[:, 2] * p[:, 0] * p[:, 1] * q[:, 2]


def screen_data(h: Hierarchy, basis, feat, d, xyz):
    """reconstructor.py:111-120: (point ids, voxel ids, values) of the points around the voxels."""
    pos = scaled_positions(xyz, h.voxel_size) - np.float32(0.5)
    sid, tid, nt, _ = coords_neighbours(h.index[d], h.strides[d], pos, 1, range_kernel(3), transposed=True,
                                        float_source=True)
    qc = -nt / h.strides[d]
    return sid, tid, basis.evaluate(feat, qc, tid)


def solve_multigrid(h: Hierarchy, basis, feats, normal_data, screen_xyz, screen_alpha, screen_delta, solve,
                    start_depth=None, end_depth=0):
    """reconstructor.py:79-195 with a learned basis. `solve(a_csr, b) -> (x, iters)`."""
    if start_depth is None:
        start_depth = h.depth - 1
    res = MultigridResult()
    should_screen = np.ndim(screen_alpha) > 0 or screen_alpha > 0.0
    screen = {}
    if should_screen:
        for d in range(end_depth, start_depth + 1):
            screen[d] = screen_data(h, basis, feats[d], d, screen_xyz)
    for d in range(start_depth, end_depth - 1, -1):
        sf = (1 / 4.0) ** d
        M = h.num_voxels(d)
        rhs = np.zeros(M)
        for dd, nd in normal_data.items():
            sid, tid, rel, _ = h.self_neighbours(dd, d, 3)
            part = basis.integrate_cd(-np.asarray(nd, np.float64)[sid], feats[d], rel, h.strides[dd], h.strides[d], tid)
            rhs += np.bincount(tid, weights=part, minlength=M)
        if should_screen:
            pid, vid, val = screen[d]
            mult = np.asarray(screen_delta) * np.asarray(screen_alpha)
            mult = mult[pid] if np.ndim(mult) > 0 else mult
            rhs += sf * np.bincount(vid, weights=val * mult, minlength=M)
        for dd in range(start_depth, d, -1):
            if h.num_voxels(dd) == 0:
                continue
            sid, tid, rel, _ = h.self_neighbours(d, dd, 5)
            a = basis.integrate_dd(feats[d], feats[dd], rel, h.strides[d], h.strides[dd], sid, tid)
            if should_screen:
                a = a + sf * screening_term(screen[d], screen[dd], sid, tid, h.num_voxels(dd), screen_alpha)
            rhs -= np.bincount(sid, weights=a * res.solutions[dd][tid], minlength=M)
        sid, tid, rel, _ = h.self_neighbours(d, d, 5)
        lhs = basis.integrate_dd(feats[d], feats[d], rel, h.strides[d], h.strides[d], sid, tid)
        if should_screen:
            lhs = lhs + sf * screening_term(screen[d], screen[d], sid, tid, M, screen_alpha)
        a = coo_to_csr(sid, tid, lhs, M)
        x, it = solve(a, rhs) if M > 0 else (np.zeros(0), 0)
        res.solutions[d] = np.asarray(x, np.float64)
        res.lhs[d], res.rhs[d], res.iters[d] = a, rhs, it
    return res


def evaluate(h: Hierarchy, basis, feats, solutions, xyz, compute_grad=False):
    """hashtree.py:157-192 with the basis: chi (or its gradient) summed over the levels."""
    pos = scaled_positions(xyz, h.voxel_size) - np.float32(0.5)
    n = pos.shape[0]
    out = np.zeros((n, 3) if compute_grad else n)
    for d, sol in solutions.items():
        if sol.shape[0] == 0:
            continue
        sid, tid, nt, _ = coords_neighbours(h.index[d], h.strides[d], pos, 1, range_kernel(3), float_source=True)
        qc = -nt / h.strides[d]
        if compute_grad:
            v = basis.evaluate_derivative(feats[d], qc, tid, h.strides[d]) * np.asarray(sol)[tid, None]
            out += np.stack([np.bincount(sid, weights=v[:, a], minlength=n) for a in range(3)], 1)
        else:
            out += np.bincount(sid, weights=basis.evaluate(feats[d], qc, tid) * np.asarray(sol)[tid], minlength=n)
    return out
