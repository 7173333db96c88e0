"""FP64 reference for the differentiable solve (oracle; test infrastructure only).

The coarse-to-fine solve of jspsr/core/reconstructor.py:122-189 is one block-triangular linear
system T X = R over X = (x_0, ..., x_{L-1}): diagonal blocks A_d = K_d + sf_d S_d(alpha),
off-diagonal blocks A_{d,dd} = K_{d,dd} + sf_d S_{d,dd}(alpha) for dd > d (the cascadic RHS
correction moved to the left), R_d = scale_d * n_d + sf_d delta sum_p alpha_p B^d(p).
T and R are affine in alpha: T = T0 + sum_p alpha_p T_p, R = R0 + sum_p alpha_p R_p, with
T_p[d, dd] = sf_d B^d(p) B^dd(p)^T (dd >= d) and R_p[d] = sf_d delta B^d(p). That gives exact
central differences and a dense-adjoint gradient: lambda = T^-T dL/dX,
dL/dalpha_p = lambda^T (R_p - T_p X), dL/dscale_d = lambda_d * n_d.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import bezier
from .hierarchy import Hierarchy


class BlockSystem:
    def __init__(self, h: Hierarchy, normal_data: dict, xyz, delta=0.1):
        self.h = h
        self.L = h.depth
        self.delta = delta
        self.sizes = [h.num_voxels(d) for d in range(self.L)]
        self.off = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.N = self.off[-1]
        n = np.asarray(xyz).shape[0]
        # point basis matrices B^d (n x M_d)
        self.B = {}
        for d in range(self.L):
            pid, vid, val, _ = h.point_basis(xyz, d)
            self.B[d] = sp.csr_matrix((val, (pid, vid)), shape=(n, self.sizes[d]))
        # stiffness part of T and the normal RHS
        rows, cols, vals = [], [], []
        self.n_rhs = np.zeros(self.N)
        for d in range(self.L):
            if self.sizes[d] == 0:
                continue
            for dd, nd in normal_data.items():
                sid, tid, rel, _ = h.self_neighbours(dd, d, 3)
                part = bezier.integrate_const_deriv_product(-nd[sid], rel, h.strides[dd], h.strides[d])
                self.n_rhs[self.off[d]:self.off[d + 1]] += np.bincount(tid, weights=part, minlength=self.sizes[d])
            for dd in range(d, self.L):
                if self.sizes[dd] == 0:
                    continue
                sid, tid, rel, _ = h.self_neighbours(d, dd, 5)
                k = bezier.integrate_deriv_deriv_product(rel, h.strides[d], h.strides[dd])
                rows.append(sid + self.off[d])
                cols.append(tid + self.off[dd])
                vals.append(k)
        self.K = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                               shape=(self.N, self.N))
        self.n = n

    def sf(self, d):
        return 0.25 ** d

    def screening(self, alpha):
        """sum_p alpha_p T_p (all blocks dd >= d) and sum_p alpha_p R_p."""
        blocks = [[None] * self.L for _ in range(self.L)]
        R = np.zeros(self.N)
        Da = sp.diags(np.asarray(alpha, np.float64))
        for d in range(self.L):
            for dd in range(self.L):
                if dd >= d and self.sizes[d] and self.sizes[dd]:
                    blocks[d][dd] = self.sf(d) * (self.B[d].T @ Da @ self.B[dd])
                else:
                    blocks[d][dd] = sp.csr_matrix((self.sizes[d], self.sizes[dd]))
            R[self.off[d]:self.off[d + 1]] = self.sf(d) * self.delta * (self.B[d].T @ np.asarray(alpha, np.float64))
        return sp.bmat(blocks, format="csr"), R

    def system(self, alpha, scale):
        S, Rs = self.screening(alpha)
        return (self.K + S).tocsc(), self.n_rhs * np.asarray(scale, np.float64) + Rs

    def solve(self, alpha, scale):
        T, R = self.system(alpha, scale)
        return spla.spsolve(T, R)

    def field(self, X, query):
        out = np.zeros(np.asarray(query).shape[0])
        for d in range(self.L):
            if self.sizes[d]:
                out += self.h.evaluate_interpolated(query, d, X[self.off[d]:self.off[d + 1]])
        return out

    def query_matrix(self, query):
        """dchi(query)/dX (Q x N), sparse."""
        mats = []
        for d in range(self.L):
            pid, vid, val, _ = self.h.point_basis(query, d)
            mats.append(sp.csr_matrix((val, (pid, vid)), shape=(np.asarray(query).shape[0], self.sizes[d])))
        return sp.hstack(mats, format="csr")

    def loss(self, alpha, scale, query, target):
        X = self.solve(alpha, scale)
        r = self.field(X, query) - target
        return float(np.mean(r * r)), X

    def gradient(self, alpha, scale, query, target):
        """Dense-adjoint gradient of loss() in (alpha, scale)."""
        T, R = self.system(alpha, scale)
        X = spla.spsolve(T, R)
        Qm = self.query_matrix(query)
        r = Qm @ X - target
        dLdX = Qm.T @ (2.0 * r / r.shape[0])
        lam = spla.spsolve(T.T.tocsc(), dLdX)
        g_scale = lam * self.n_rhs
        g_alpha = np.zeros(self.n)
        chi_ge = np.zeros(self.n)
        for d in range(self.L - 1, -1, -1):
            if self.sizes[d] == 0:
                continue
            chi_ge += self.B[d] @ X[self.off[d]:self.off[d + 1]]
            G = self.B[d] @ lam[self.off[d]:self.off[d + 1]]
            g_alpha += self.sf(d) * G * (self.delta - chi_ge)
        return g_alpha, g_scale, X
