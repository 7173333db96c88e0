"""Learned additive bases (jspsr/bases/additive.py, additive_components.py, hermite_ns.py) on the
HIP kernels of libngspsr.

A basis is B(x) = prod_v sum_c m_c,v q_c(x_v): per voxel, 3 coefficients (one per axis) for each
elementary function q_c. The Galerkin integrals of the elementary functions (additive.py:17-52,
172-249) are tabulated here by Gauss-Legendre quadrature on every smooth piece of the integrand
(16 nodes: exact for the polynomial components up to rounding, and to rounding for the sinusoids);
the reference integrates them symbolically with sympy and caches them to files. The derivative
integrals integrate the piecewise derivative, so the jump of a polynomial component at the support
boundary contributes nothing, as in the symbolic integration.
"""
import re

import numpy as np
import torch

from ._lib import NgsError, call, ptr, stream

_GL_X, _GL_W = np.polynomial.legendre.leggauss(16)
N_MULTIPLES = 5  # stride ratios 1, 2, 4, 8, 16 (additive.py:172)

# factor kinds of the kernel's component descriptor
_ONE, _POWER, _BEZIER, _SIN = 0, 1, 2, 3


def _bez(x):
    return np.where((x >= -1.5) & (x < -0.5), (x + 1.5) ** 2,
                    np.where((x >= -0.5) & (x < 0.5), -2 * x * x + 1.5, np.where((x >= 0.5) & (x < 1.5), (x - 1.5) ** 2, 0.0)))


def _dbez(x):
    return np.where((x >= -1.5) & (x < -0.5), 2 * x + 3,
                    np.where((x >= -0.5) & (x < 0.5), -4 * x, np.where((x >= 0.5) & (x < 1.5), 2 * x - 3, 0.0)))


def _support(x):
    return (x >= -1.5) & (x < 1.5)


class AdditiveComponent:
    """Elementary function q(x) on [-1.5, 1.5) (additive_components.py:7-47)."""
    breaks = (-1.5, 1.5)

    def factors(self):
        """Two kernel factors (kind, a0, a1) whose product is q."""
        raise NotImplementedError

    def f(self, x):
        raise NotImplementedError

    def df(self, x):
        raise NotImplementedError


class Power(AdditiveComponent):
    """q(x) = x^p (additive_components.py:74-99)."""

    def __init__(self, power):
        self.power = power

    def __repr__(self):
        return f"Power({self.power:.2f})"

    def factors(self):
        return [(_POWER, float(self.power), 0.0), (_ONE, 0.0, 0.0)]

    def f(self, x):
        return np.where(_support(x), np.ones_like(x) if self.power == 0 else x ** self.power, 0.0)

    def df(self, x):
        if self.power == 0:
            return np.zeros_like(x)
        return np.where(_support(x), self.power * x ** (self.power - 1) if self.power != 1 else np.ones_like(x), 0.0)


class Bezier(AdditiveComponent):
    """q(x) = the degree-2 B-spline B(x) (additive_components.py:102-130)."""
    breaks = (-1.5, -0.5, 0.5, 1.5)

    def __repr__(self):
        return "Bezier()"

    def factors(self):
        return [(_BEZIER, 0.0, 0.0), (_ONE, 0.0, 0.0)]

    def f(self, x):
        return _bez(x)

    def df(self, x):
        return _dbez(x)


class Sinusoid(AdditiveComponent):
    """q(x) = sin(beta x + gamma) (additive_components.py:133-152)."""

    def __init__(self, beta, gamma):
        self.beta, self.gamma = float(beta), float(gamma)

    def __repr__(self):
        return f"Sin({self.beta:.2f}-{self.gamma:.2f})"

    def factors(self):
        return [(_SIN, self.beta, self.gamma), (_ONE, 0.0, 0.0)]

    def f(self, x):
        return np.where(_support(x), np.sin(self.beta * x + self.gamma), 0.0)

    def df(self, x):
        return np.where(_support(x), self.beta * np.cos(self.beta * x + self.gamma), 0.0)


class SinPeaks(Sinusoid):
    """Sinusoid with n peaks on the support (additive_components.py:155-164)."""

    def __init__(self, n_peaks):
        super().__init__(np.pi * n_peaks / 3, -0.5 * np.pi * n_peaks + np.pi)
        self.n_peaks = n_peaks

    def __repr__(self):
        return f"SinPeaks({self.n_peaks})"


class Product(AdditiveComponent):
    """q(x) = ca(x) cb(x) of two elementary components (additive_components.py:50-71)."""

    def __init__(self, ca, cb):
        if isinstance(ca, Product) or isinstance(cb, Product):
            raise NotImplementedError("nested products are not supported")
        self.ca, self.cb = ca, cb
        self.breaks = tuple(sorted(set(ca.breaks) | set(cb.breaks)))

    def __repr__(self):
        return f"Product({self.ca}|{self.cb})"

    def factors(self):
        return [self.ca.factors()[0], self.cb.factors()[0]]

    def f(self, x):
        return self.ca.f(x) * self.cb.f(x)

    def df(self, x):
        return self.ca.f(x) * self.cb.df(x) + self.ca.df(x) * self.cb.f(x)


_NUM = r"\s*(-?[0-9.eE+-]+)\s*"


def parse_component(c):
    """A component or its string form ("Power(2)", "Bezier()", "Sinusoid(b, g)", "SinPeaks(n)"),
    parsed without eval (the reference evals the string, additive.py:150-152)."""
    if isinstance(c, AdditiveComponent):
        return c
    s = c.strip()
    m = re.fullmatch(r"Power\(" + _NUM + r"\)", s)
    if m:
        v = float(m.group(1))
        return Power(int(v) if v.is_integer() else v)
    if re.fullmatch(r"Bezier\(\s*\)", s):
        return Bezier()
    m = re.fullmatch(r"Sinusoid\(" + _NUM + "," + _NUM + r"\)", s)
    if m:
        return Sinusoid(float(m.group(1)), float(m.group(2)))
    m = re.fullmatch(r"SinPeaks\(" + _NUM + r"\)", s)
    if m:
        return SinPeaks(int(float(m.group(1))))
    raise ValueError(f"unknown additive component {c!r}")


def _quad(fun, breaks, lo, hi):
    if not lo < hi:
        return 0.0
    pts = sorted({lo, hi} | {b for b in breaks if lo < b < hi})
    total = 0.0
    for a, b in zip(pts[:-1], pts[1:]):
        xm, xr = 0.5 * (a + b), 0.5 * (b - a)
        total += xr * float(np.sum(_GL_W * fun(xm + xr * _GL_X)))
    return total


_TABLES = {}


def _ff_tables(ca, cb):
    """f_f and df_df of additive.py:17-29 for multiples 2^0..2^4: entry k at dx = k - 1.5 M - 0.5."""
    key = ("ff", repr(ca), repr(cb))
    if key in _TABLES:
        return _TABLES[key]
    f_f, df_df = [], []
    for m in range(N_MULTIPLES):
        M = 2 ** m
        fr, dr = [], []
        for k in range(3 * M + 2):
            dx = k - 1.5 * M - 0.5
            br = [b - dx for b in ca.breaks] + [b * M for b in cb.breaks]
            lo, hi = max(-1.5 - dx, -1.5 * M), min(1.5 - dx, 1.5 * M)
            fr.append(_quad(lambda x: ca.f(x + dx) * cb.f(x / M), br, lo, hi))
            dr.append(_quad(lambda x: ca.df(x + dx) * cb.df(x / M) / M, br, lo, hi))
        f_f.append(np.array(fr))
        df_df.append(np.array(dr))
    _TABLES[key] = (f_f, df_df)
    return f_f, df_df


def _f_tables(c):
    """f, df (additive.py:32-40; entry k at ra = 1.5 M - 1 - k) and f_inv, df_inv (additive.py:43-52;
    multiples 2^1..2^4, entry k at ra = 0.5 - k)."""
    key = ("f", repr(c))
    if key in _TABLES:
        return _TABLES[key]
    f, df, f_inv, df_inv = [], [], [], []
    for m in range(N_MULTIPLES):
        M = 2 ** m
        br = [b * M for b in c.breaks]
        ras = [1.5 * M - 1 - k for k in range(3 * M)]
        f.append(np.array([_quad(lambda x: c.f(x / M), br, ra, ra + 1) for ra in ras]))
        df.append(np.array([_quad(lambda x: c.df(x / M) / M, br, ra, ra + 1) for ra in ras]))
    for m in range(1, N_MULTIPLES):
        M = 2 ** m
        ras = [0.5 - k for k in range(M + 2)]
        f_inv.append(np.array([_quad(c.f, c.breaks, max(ra, -1.5), min(ra + M, 1.5)) for ra in ras]))
        df_inv.append(np.array([_quad(c.df, c.breaks, max(ra, -1.5), min(ra + M, 1.5)) for ra in ras]))
    _TABLES[key] = (f, df, f_inv, df_inv)
    return _TABLES[key]


def _descriptor(components, device):
    d = []
    for c in components:
        for kind, a0, a1 in c.factors():
            d += [float(kind), a0, a1]
    return torch.tensor(d, dtype=torch.float32, device=device)


def _check_ids(ids, m, nc, x):
    if m.dim() != 2 or m.shape[1] != 3 * nc or x.dim() != 2 or x.shape[1] != 3 or ids.shape != (x.shape[0],):
        raise NgsError("additive basis: coefficients must be (M, 3 C), xyz (N, 3), feat_ids (N,)")
    if ids.numel() and not (0 <= int(ids.min()) and int(ids.max()) < m.shape[0]):
        raise NgsError("additive basis: feat_ids outside the feature rows")


class EvalFunc(torch.autograd.Function):
    """Per-axis sums sum_c q_c(x) m_c (additive.py:71-99), on the GPU."""

    @staticmethod
    def forward(ctx, xyz, coeffs, feat_ids, desc, nc):
        x = xyz.float().contiguous()
        ids = feat_ids.long().contiguous()
        m = coeffs.float().contiguous()
        _check_ids(ids, m, nc, x)
        funcs = torch.empty_like(x)
        call("ngs_additive_eval", ptr(x), ptr(m), ptr(ids), x.shape[0], ptr(desc), nc, 1.0, ptr(funcs), None,
             stream())
        ctx.save_for_backward(x, ids, desc)
        ctx.nc, ctx.coeff_shape = nc, m.shape
        return funcs

    @staticmethod
    def backward(ctx, grad_funcs):
        x, ids, desc = ctx.saved_tensors
        g = torch.zeros(ctx.coeff_shape, dtype=torch.float32, device=x.device)
        gf = grad_funcs.float().contiguous()
        call("ngs_additive_eval_backward", ptr(x), ptr(ids), x.shape[0], ptr(desc), ctx.nc, 1.0, ptr(gf), None,
             ptr(g), stream())
        return None, g, None, None, None


class EvalDerivFunc(torch.autograd.Function):
    """The sums and their derivatives / stride (additive.py:102-136), on the GPU."""

    @staticmethod
    def forward(ctx, xyz, coeffs, feat_ids, desc, nc, stride):
        x = xyz.float().contiguous()
        ids = feat_ids.long().contiguous()
        m = coeffs.float().contiguous()
        _check_ids(ids, m, nc, x)
        funcs = torch.empty_like(x)
        dfuncs = torch.empty_like(x)
        call("ngs_additive_eval", ptr(x), ptr(m), ptr(ids), x.shape[0], ptr(desc), nc, float(stride), ptr(funcs),
             ptr(dfuncs), stream())
        ctx.save_for_backward(x, ids, desc)
        ctx.nc, ctx.coeff_shape, ctx.stride = nc, m.shape, float(stride)
        return funcs, dfuncs

    @staticmethod
    def backward(ctx, grad_funcs, grad_dfuncs):
        x, ids, desc = ctx.saved_tensors
        g = torch.zeros(ctx.coeff_shape, dtype=torch.float32, device=x.device)
        gf = (grad_funcs if grad_funcs is not None else torch.zeros_like(x)).float().contiguous()
        gd = (grad_dfuncs if grad_dfuncs is not None else torch.zeros_like(x)).float().contiguous()
        call("ngs_additive_eval_backward", ptr(x), ptr(ids), x.shape[0], ptr(desc), ctx.nc, ctx.stride, ptr(gf),
             ptr(gd), ptr(g), stream())
        return None, g, None, None, None, None


class DDIFunc(torch.autograd.Function):
    """Differentiable nabla B_k . nabla B_l factors (additive.py:55-68, ddi.cu), on the GPU."""

    @staticmethod
    def forward(ctx, source_expanded, target_expanded, rel_pos, f_f_i, df_df_i):
        s = source_expanded.float().contiguous()
        t = target_expanded.float().contiguous()
        r = rel_pos.long().contiguous()
        ff = f_f_i.float().contiguous()
        dd = df_df_i.float().contiguous()
        n, nc, L = s.shape[0], ff.shape[0], ff.shape[2]
        if s.shape != (n, 3 * nc) or t.shape != (n, 3 * nc) or r.shape != (n, 3) or dd.shape != ff.shape:
            raise NgsError("ddi_forward: inconsistent shapes")
        if n and not (0 <= int(r.min()) and int(r.max()) < L):
            raise NgsError("ddi_forward: rel_pos outside the table")
        p = torch.empty((n, 3), dtype=torch.float32, device=s.device)
        q = torch.empty((n, 3), dtype=torch.float32, device=s.device)
        call("ngs_ddi_forward", ptr(s), ptr(t), ptr(r), n, nc, ptr(ff), ptr(dd), L, ptr(p), ptr(q), stream())
        ctx.save_for_backward(r, s, t, ff, dd)
        return p, q

    @staticmethod
    def backward(ctx, grad_p, grad_q):
        r, s, t, ff, dd = ctx.saved_tensors
        gp = (grad_p if grad_p is not None else torch.zeros((s.shape[0], 3), device=s.device)).float().contiguous()
        gq = (grad_q if grad_q is not None else torch.zeros((s.shape[0], 3), device=s.device)).float().contiguous()
        gs = torch.empty_like(s)
        gt = torch.empty_like(t)
        call("ngs_ddi_backward", ptr(s), ptr(t), ptr(r), s.shape[0], ff.shape[0], ptr(ff), ptr(dd), ff.shape[2],
             ptr(gp), ptr(gq), ptr(gs), ptr(gt), stream())
        return gs, gt, None, None, None


def ddi_forward(source_expanded, target_expanded, rel_pos, f_f_i, df_df_i):
    """integration.ddi_forward (ddi.cu:50-76): (p_mats, q_mats) (N, 3)."""
    return DDIFunc.apply(source_expanded, target_expanded, rel_pos, f_f_i, df_df_i)


class AdditiveBasis(torch.nn.Module):
    """B(x) = sum_c m_c q_c (additive.py:139-414)."""

    def __init__(self, components, init_vals):
        super().__init__()
        assert len(components) == len(init_vals)
        self.components = [parse_component(c) for c in components]
        self.init_vals = list(init_vals)
        nc = len(self.components)
        pairs = {(a, b): _ff_tables(self.components[a], self.components[b])
                 for a in range(nc) for b in range(nc)}
        # ff_integral[mult]['f_f' | 'df_df'] (C, C, 3 * 2^mult + 2), additive.py:158-168
        self.ff_integral = []
        for m in range(N_MULTIPLES):
            ent = {}
            for name, j in (("f_f", 0), ("df_df", 1)):
                a = np.array([[pairs[(s, t)][j][m] for t in range(nc)] for s in range(nc)])
                ent[name] = self._param(a, f"ff_integral-{m}-{name}")
            self.ff_integral.append(ent)
        # f_integral[c]['f' | 'df'][mult], ['f_inv' | 'df_inv'][mult - 1]
        self.f_integral = []
        for ci in range(nc):
            f, df, f_inv, df_inv = _f_tables(self.components[ci])
            self.f_integral.append({name: [self._param(v, f"f_integral-{ci}-{name}-{k}") for k, v in enumerate(vals)]
                                    for name, vals in (("f", f), ("df", df), ("f_inv", f_inv), ("df_inv", df_inv))})
        self.register_buffer("desc", _descriptor(self.components, "cpu"), persistent=False)

    def _param(self, a, name):
        p = torch.nn.Parameter(torch.tensor(np.asarray(a), dtype=torch.float32), requires_grad=False)
        self.register_parameter(name.replace(".", "_"), p)
        return p

    @staticmethod
    def get_domain_range():
        return 3

    def feature_to_coefficient(self, feat):
        return feat

    def initialize_feature_value(self, feat):
        assert feat.size(1) == self.get_feature_size()
        for ci in range(len(self.init_vals)):
            feat[:, ci * 3: ci * 3 + 3] = self.init_vals[ci]

    def get_feature_size(self):
        return len(self.components) * 3

    def _desc(self, device):
        if self.desc.device != torch.device(device):
            self.desc = self.desc.to(device)
        return self.desc

    def evaluate(self, feat, xyz, feat_ids):
        """B at local coordinates xyz (N, 3) of voxels feat_ids (additive.py:275-285)."""
        assert feat_ids.size(0) == xyz.size(0), "Input sizes not equal!"
        coeffs = self.feature_to_coefficient(feat)
        funcs = EvalFunc.apply(xyz, coeffs, feat_ids, self._desc(xyz.device), len(self.components))
        return funcs[:, 0] * funcs[:, 1] * funcs[:, 2]

    def evaluate_derivative(self, feat, xyz, feat_ids, stride):
        """nabla B (N, 3), per unit of the stride-scaled coordinate (additive.py:287-300)."""
        coeffs = self.feature_to_coefficient(feat)
        funcs, dfuncs = EvalDerivFunc.apply(xyz, coeffs, feat_ids, self._desc(xyz.device), len(self.components),
                                            stride)
        return torch.stack([dfuncs[:, 0] * funcs[:, 1] * funcs[:, 2],
                            funcs[:, 0] * dfuncs[:, 1] * funcs[:, 2],
                            funcs[:, 0] * funcs[:, 1] * dfuncs[:, 2]], dim=1)

    def integrate_deriv_deriv_product(self, source_feat, target_feat, rel_pos, source_stride, target_stride,
                                      source_ids, target_ids):
        """int nabla B_source . nabla B_target per COO pair (additive.py:302-343)."""
        assert source_stride <= target_stride, "Query must be smaller than reference."
        assert source_ids.size(0) == target_ids.size(0) == rel_pos.size(0), "Input sizes not equal!"
        mult = target_stride // source_stride
        rel = (rel_pos + mult * 3 / 2 + 0.5).round().long()
        mult = [1, 2, 4, 8, 16].index(mult)
        ff = self.ff_integral[mult]
        max_rel = ff["f_f"].size(-1)
        invalid = ~torch.logical_and(rel >= 0, rel < max_rel)
        rel.clamp_(min=0, max=max_rel - 1)
        s = self.feature_to_coefficient(source_feat)[source_ids]
        t = self.feature_to_coefficient(target_feat)[target_ids]
        dev = s.device
        p, q = DDIFunc.apply(s, t, rel, ff["f_f"].data.to(dev) * source_stride, ff["df_df"].data.to(dev) / source_stride)
        p = p.masked_fill(invalid, 0.0)
        q = q.masked_fill(invalid, 0.0)
        return q[:, 0] * p[:, 1] * p[:, 2] + p[:, 0] * q[:, 1] * p[:, 2] + p[:, 0] * p[:, 1] * q[:, 2]

    def integrate_const_deriv_product(self, data, target_feat, rel_pos, data_stride, target_stride, target_ids):
        """int nabla B_target . data per COO pair (additive.py:345-386)."""
        assert data.size(0) == target_ids.size(0) == rel_pos.size(0), "Input sizes not equal!"
        if target_stride >= data_stride:
            mult = target_stride // data_stride
            idx = (rel_pos + (3 * mult - 1) / 2.0).round().long()
            mult = [1, 2, 4, 8, 16].index(mult)
            sfx = ""
        else:
            mult = data_stride // target_stride
            idx = (rel_pos * mult + mult / 2 + 0.5).round().long()
            mult = [2, 4, 8, 16].index(mult)
            sfx = "_inv"
            data_stride, target_stride = target_stride, data_stride
        t = self.feature_to_coefficient(target_feat)[target_ids]
        dev = t.device
        p = 0
        q = 0
        for si in range(len(self.components)):
            tab = self.f_integral[si]
            p = p + t[:, 3 * si:3 * si + 3] * (tab["f" + sfx][mult].data.to(dev)[idx] * data_stride)
            q = q + t[:, 3 * si:3 * si + 3] * tab["df" + sfx][mult].data.to(dev)[idx]
        return data[:, 0] * q[:, 0] * p[:, 1] * p[:, 2] + data[:, 1] * p[:, 0] * q[:, 1] * p[:, 2] + \
            data[:, 2] * p[:, 0] * p[:, 1] * q[:, 2]

    @staticmethod
    def get_bezier():
        return AdditiveBasis([Bezier()], [1.0])

    @staticmethod
    def get_poly_6():
        return AdditiveBasis(["Power(0)", "Power(1)", "Power(2)", "Power(3)", "Power(4)", "Power(5)", "Power(6)"],
                             [0.7394, 0.0, -0.6572, 0.0, 0.1461, 0.0, 0.0])


def hermite_null_space(n_degrees):
    """Basis (rows, ascending degree) of the polynomials of degree <= n_degrees vanishing with their
    derivative at -1.5 and 1.5 (hermite_ns.py:7-21)."""
    from scipy.linalg import null_space
    coef = np.zeros((4, n_degrees + 1))
    for i in range(n_degrees + 1):
        deg = n_degrees - i
        coef[0, i] = 1.5 ** deg
        coef[1, i] = (-1.5) ** deg
        coef[2, i] = deg * 1.5 ** (deg - 1)
        coef[3, i] = deg * (-1.5) ** (deg - 1)
    return null_space(coef)[::-1].T


class AdditiveHermite(AdditiveBasis):
    """Additive basis whose polynomial part lies in the Hermite null space (hermite_ns.py:24-80)."""

    def __init__(self, n_degrees, components, init_vals):
        assert n_degrees > 3, "Degrees <= cubic does not admit two hermite constraints!"
        super().__init__(list(components) + [Power(p) for p in range(n_degrees + 1)],
                         list(init_vals) + [0.0] * (n_degrees + 1))
        self.n_degrees = n_degrees
        ns = torch.from_numpy(hermite_null_space(n_degrees).copy()).float()
        self.ns = torch.nn.Parameter(torch.kron(ns, torch.eye(3)), requires_grad=False)
        self.n_acomp = len(components)

    def feature_to_coefficient(self, feat):
        return torch.cat([feat[:, :self.n_acomp * 3], feat[:, self.n_acomp * 3:] @ self.ns.to(feat.device)], dim=1)

    def initialize_feature_value(self, feat):
        assert feat.size(1) == self.get_feature_size()
        for ci in range(self.n_acomp):
            feat[:, ci * 3: ci * 3 + 3] = self.init_vals[ci]
        ns4 = hermite_null_space(4)
        ns_my = hermite_null_space(self.n_degrees)
        init = np.zeros(self.n_degrees + 1)
        init[0], init[2], init[4] = ns4[0, 0], ns4[0, 2], ns4[0, 4]
        sol = np.linalg.lstsq(ns_my.T, init, rcond=None)[0]
        base = self.n_acomp * 3
        for deg in range(sol.shape[0]):
            feat[:, base + deg * 3: base + deg * 3 + 3] = float(sol[deg])

    def get_feature_size(self):
        return self.n_acomp * 3 + self.ns.size(0)
