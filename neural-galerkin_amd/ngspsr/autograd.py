"""Differentiable screened Poisson solve (adjoint method) on the HIP kernels.

Parameters (what the reference's learned caller trains through the solver, models/full_net.py:
87-133, ops.py:64-82, solver.py:76-98):
  alpha  (N,)        per-point screening weight (ScreenedMultiplication's pts_weight);
  scale  (sum M_d,)  per-basis scale of the normal-field term: b_d = scale_d * n_d + screening
                     - cross-level correction, n_d the divergence RHS of the splatted normals.
Per level d (coarse to fine) the forward solves A_d x_d = b_d with
  A_d     = K_d + sf_d S_d(alpha),                  S_d   = sum_p alpha_p B^d(p) B^d(p)^T
  b_d     = scale_d n_d + sf_d delta sum_p alpha_p B^d(p) - sum_{dd>d} A_{d,dd} x_dd
  A_{d,dd}= K_{d,dd} + sf_d sum_p alpha_p B^d(p) B^dd(p)^T,  sf_d = 4^-d.
The backward runs fine to coarse: g_b_d = A_d^-1 g_x_d (A_d symmetric, the adjoint solve reuses the
assembled matrix); g_scale_d = g_b_d * n_d; with G_d(p) = sum_i g_b_d,i B_i^d(p) and
X_{>=d}(p) = sum_{dd>=d} chi_dd(p):  g_alpha_p += sf_d G_d(p) (delta - X_{>=d}(p));
g_x_dd += -K_{d,dd}^T g_b_d - sf_d sum_p alpha_p G_d(p) B^dd(p) for every coarser dd.
Forward Krylov iterates are not stored: the adjoint is a fresh solve.
"""
import torch

from . import solver as _solver
from ._lib import call, ptr, stream
from .bases import ensure_device_tables
from .hashtree import HashTree
from .reconstructor import PointBinning, Reconstructor


class DifferentiableScreenedPoisson:
    def __init__(self, tree: HashTree, normal_data: dict, xyz, delta: float = 0.1,
                 solver: str = "pcg@tol=1e-7@maxiter=50000", adjoint_tol: float = 1e-7):
        ensure_device_tables(tree.device)
        self.tree = tree
        self.xyz = xyz.float().contiguous()
        self.delta = float(delta)
        self.solver = solver
        self.adjoint_tol = adjoint_tol
        self.L = tree.depth
        self.sizes = [tree.get_coords_size(None, d) for d in range(self.L)]
        self.offsets = [sum(self.sizes[:d]) for d in range(self.L)]
        self.total = sum(self.sizes)
        self.recon = Reconstructor(tree)
        # divergence RHS of the normals per target level (independent of the parameters)
        self.n_rhs = {}
        for d in range(self.L):
            r = torch.zeros(self.sizes[d], dtype=torch.float64, device=tree.device)
            if self.sizes[d]:
                for dd, nd in normal_data.items():
                    call("ngs_rhs_normals", tree.level(d), tree.level(dd), ptr(nd.float().contiguous()), ptr(r), stream())
            self.n_rhs[d] = r

    def split(self, cat):
        return {d: cat[self.offsets[d]:self.offsets[d] + self.sizes[d]] for d in range(self.L)}

    def _field_at_points(self, d, coef, out, accumulate=True):
        if self.sizes[d] == 0:
            return out
        call("ngs_eval_field", self.tree.grids[d].struct(), ptr(coef.float().contiguous()), ptr(self.xyz),
             self.xyz.shape[0], self.tree.inv_vs, 1.0, int(accumulate), ptr(out), None, stream())
        return out

    def forward(self, alpha, scale_cat):
        tree = self.tree
        alpha = alpha.float().contiguous()
        scale = self.split(scale_cat.double())
        n = self.xyz.shape[0]
        self.alpha = alpha
        self.systems, self.x, self.coarse = {}, {}, {}
        self.infos = {}
        for d in range(self.L - 1, -1, -1):
            M = self.sizes[d]
            if M == 0:
                self.x[d] = torch.zeros(0, dtype=torch.float32, device=tree.device)
                continue
            sf = 0.25 ** d
            rhs = self.n_rhs[d] * scale[d]
            coarser = [dd for dd in range(self.L - 1, d, -1) if self.sizes[dd] > 0]
            X = torch.zeros(n, dtype=torch.float32, device=tree.device)
            for dd in coarser:
                self._field_at_points(dd, self.x[dd], X)
                call("ngs_rhs_cross_stiffness", tree.level(d), tree.level(dd), ptr(self.x[dd]), ptr(rhs), stream())
            bins = PointBinning(tree, self.xyz, d, alpha)
            system = self.recon.assemble_level(d, bins, X if coarser else None, sf, self.delta, rhs)
            x, info = _solver.solve(system, rhs, self.solver)
            self.systems[d] = system
            self.x[d] = x.float().contiguous()  # the FP32 kernels below read it ("refine" returns FP64)
            self.coarse[d] = X
            self.infos[d] = info
        return torch.cat([self.x[d] for d in range(self.L)])

    def backward(self, g_x_cat):
        tree = self.tree
        g_x = {d: v.double().clone() for d, v in self.split(g_x_cat).items()}
        n = self.xyz.shape[0]
        g_alpha = torch.zeros(n, dtype=torch.float64, device=tree.device)
        g_scale = {}
        for d in range(self.L):
            M = self.sizes[d]
            if M == 0:
                g_scale[d] = torch.zeros(0, dtype=torch.float64, device=tree.device)
                continue
            sf = 0.25 ** d
            g_b, _ = _solver.pcg(self.systems[d], g_x[d].float(), tol=self.adjoint_tol, maxiter=50000,
                                 raise_on_cap=False)
            g_scale[d] = g_b.double() * self.n_rhs[d]
            G = torch.zeros(n, dtype=torch.float32, device=tree.device)
            self._field_at_points(d, g_b, G, accumulate=False)
            X = self.coarse[d].clone()
            self._field_at_points(d, self.x[d], X)
            g_alpha += sf * G.double() * (self.delta - X.double())
            w = (self.alpha * G).contiguous()
            for dd in range(d + 1, self.L):
                if self.sizes[dd] == 0:
                    continue
                call("ngs_cross_stiffness_transpose", tree.level(d), tree.level(dd), ptr(g_b), -1.0, ptr(g_x[dd]),
                     stream())
                call("ngs_basis_splat", tree.grids[dd].struct(), ptr(w), ptr(self.xyz), n, tree.inv_vs, -sf,
                     ptr(g_x[dd]), stream())
        return g_alpha, torch.cat([g_scale[d] for d in range(self.L)])

    # ----------------------------------------------------------------- field at query points
    def field(self, x_cat, query):
        q = query.float().contiguous()
        out = torch.zeros(q.shape[0], dtype=torch.float32, device=self.tree.device)
        for d, xd in self.split(x_cat).items():
            if self.sizes[d]:
                call("ngs_eval_field", self.tree.grids[d].struct(), ptr(xd.float().contiguous()), ptr(q), q.shape[0],
                     self.tree.inv_vs, 1.0, 1, ptr(out), None, stream())
        return out

    def field_transpose(self, g_chi, query):
        q = query.float().contiguous()
        g = g_chi.float().contiguous()
        outs = []
        for d in range(self.L):
            o = torch.zeros(self.sizes[d], dtype=torch.float64, device=self.tree.device)
            if self.sizes[d]:
                call("ngs_basis_splat", self.tree.grids[d].struct(), ptr(g), ptr(q), q.shape[0], self.tree.inv_vs, 1.0,
                     ptr(o), stream())
            outs.append(o)
        return torch.cat(outs)

    def __call__(self, alpha, scale_cat, query):
        """chi (raw field) at query points, differentiable in alpha and scale_cat."""
        x = _SolveFn.apply(alpha, scale_cat, self)
        return _FieldFn.apply(x, query, self)


class _SolveFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, alpha, scale_cat, op):
        ctx.op = op
        with torch.no_grad():
            return op.forward(alpha.detach(), scale_cat.detach())

    @staticmethod
    def backward(ctx, g_x):
        g_alpha, g_scale = ctx.op.backward(g_x.contiguous())
        return g_alpha.to(torch.float32), g_scale.to(torch.float32), None


class _FieldFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x_cat, query, op):
        ctx.op = op
        ctx.save_for_backward(query)
        with torch.no_grad():
            return op.field(x_cat.detach(), query)

    @staticmethod
    def backward(ctx, g_chi):
        (query,) = ctx.saved_tensors
        return ctx.op.field_transpose(g_chi, query).to(torch.float32), None, None


def setup(xyz, normal, voxel_size, depth, min_density=8.0, xyz_depth=None, **kw):
    """Tree, splatted normals and the differentiable operator for a point cloud on the GPU."""
    tree = HashTree(xyz, voxel_size, depth)
    tree.build_encoder_hierarchy_adaptive(min_density=min_density, xyz_depth=xyz_depth)
    tree.reflect_decoder_coords()
    w = 1.0 / tree.xyz_density
    counts = tree.level_point_counts()
    nd = {}
    for d in range(depth):
        if counts[d]:
            nd[d] = tree.splat_data(xyz, None, d, normal, scale=w, levels=tree.xyz_depth, level=d) / float(
                tree.strides[d] ** 3)
    return DifferentiableScreenedPoisson(tree, nd, xyz, **kw), nd
