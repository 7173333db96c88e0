"""Learned additive bases (oracle; test infrastructure only).

Restates jspsr/bases/additive.py:17-52 (the Galerkin integrals of the elementary functions, here by
sympy's exact symbolic integration piece by piece, as the reference integrates them symbolically), additive.py:55-136 and
csrc/integration/ddi.cu (DDI and the basis evaluation with their gradients) in float64, and
additive_components.py (the elementary functions). Components are tuples: ("power", p),
("bezier",), ("sin", beta, gamma), ("product", c1, c2).
"""
import numpy as np


def q(c, x):
    """Component value as the reference's evaluate computes it (Power and Sinusoid unmasked)."""
    x = np.asarray(x, np.float64)
    k = c[0]
    if k == "power":
        return np.ones_like(x) if c[1] == 0 else x ** c[1]
    if k == "bezier":
        return np.where((x >= -1.5) & (x < -0.5), (x + 1.5) ** 2,
                        np.where((x >= -0.5) & (x < 0.5), -2 * x ** 2 + 1.5, np.where((x >= 0.5) & (x < 1.5), (x - 1.5) ** 2, 0.0)))
    if k == "sin":
        return np.sin(c[1] * x + c[2])
    return q(c[1], x) * q(c[2], x)


def dq(c, x):
    x = np.asarray(x, np.float64)
    k = c[0]
    if k == "power":
        if c[1] == 0:
            return np.zeros_like(x)
        return np.ones_like(x) if c[1] == 1 else c[1] * x ** (c[1] - 1)
    if k == "bezier":
        return np.where((x >= -1.5) & (x < -0.5), 2 * x + 3,
                        np.where((x >= -0.5) & (x < 0.5), -4 * x, np.where((x >= 0.5) & (x < 1.5), 2 * x - 3, 0.0)))
    if k == "sin":
        return c[1] * np.cos(c[1] * x + c[2])
    return q(c[1], x) * dq(c[2], x) + dq(c[1], x) * q(c[2], x)


def pieces(c, x):
    """additive_components.py symbolic forms as (lo, hi, expression) pieces on [-1.5, 1.5), zero
    elsewhere, with rational breakpoints."""
    import sympy as sp
    h = sp.Rational(1, 2)
    k = c[0]
    if k == "power":
        return [(-3 * h, 3 * h, x ** c[1])]
    if k == "bezier":
        return [(-3 * h, -h, (x + 3 * h) ** 2), (-h, h, -2 * x ** 2 + 3 * h), (h, 3 * h, (x - 3 * h) ** 2)]
    if k == "sin":
        return [(-3 * h, 3 * h, sp.sin(sp.nsimplify(c[1]) * x + sp.nsimplify(c[2])))]
    return _product(pieces(c[1], x), pieces(c[2], x))


def _product(pa, pb):
    out = []
    for a0, a1, ea in pa:
        for b0, b1, eb in pb:
            lo, hi = max(a0, b0), min(a1, b1)
            if lo < hi:
                out.append((lo, hi, ea * eb))
    return out


def _integrate(pcs, x, lo=None, hi=None):
    import sympy as sp
    total = 0
    for a, b, e in pcs:
        a = a if lo is None else max(a, lo)
        b = b if hi is None else min(b, hi)
        if a < b:
            total += sp.integrate(e, (x, a, b))
    return float(total)


def ff_entry(ca, cb, m, k):
    """(f_f, df_df) entry k of multiple 2^m (additive.py:17-29): dx = k - 1.5 M - 0.5, integrals of
    f_a(x + dx) f_b(x / M) and of their x-derivatives."""
    import sympy as sp
    x = sp.Symbol("x")
    M = 2 ** m
    dx = sp.Rational(2 * k - 3 * M - 1, 2)
    pa = [(a - dx, b - dx, e.subs(x, x + dx)) for a, b, e in pieces(ca, x)]
    pb = [(a * M, b * M, e.subs(x, x / M)) for a, b, e in pieces(cb, x)]
    ff = _integrate(_product(pa, pb), x)
    dd = _integrate(_product([(a, b, sp.diff(e, x)) for a, b, e in pa], [(a, b, sp.diff(e, x)) for a, b, e in pb]), x)
    return ff, dd


def f_entry(c, m, k):
    """(f, df) entry k of multiple 2^m (additive.py:32-40, inserted in front): ra = 1.5 M - 1 - k."""
    import sympy as sp
    x = sp.Symbol("x")
    M = 2 ** m
    ra = sp.Rational(3 * M - 2 - 2 * k, 2)
    pc = [(a * M, b * M, e.subs(x, x / M)) for a, b, e in pieces(c, x)]
    return _integrate(pc, x, ra, ra + 1), _integrate([(a, b, sp.diff(e, x)) for a, b, e in pc], x, ra, ra + 1)


def f_inv_entry(c, m, k):
    """(f_inv, df_inv) entry k of multiple 2^m, m >= 1 (additive.py:43-52): ra = 0.5 - k."""
    import sympy as sp
    x = sp.Symbol("x")
    M = 2 ** m
    ra = sp.Rational(1 - 2 * k, 2)
    lb, ub = max(ra, -sp.Rational(3, 2)), min(ra + M, sp.Rational(3, 2))
    pc = pieces(c, x)
    return _integrate(pc, x, lb, ub), _integrate([(a, b, sp.diff(e, x)) for a, b, e in pc], x, lb, ub)


def ddi_forward(s, t, rel, ff, dd):
    """ddi.cu:21-76 in float64: p, q (N, 3)."""
    n, nc = s.shape[0], ff.shape[0]
    s3 = np.asarray(s, np.float64).reshape(n, nc, 3)
    t3 = np.asarray(t, np.float64).reshape(n, nc, 3)
    p = np.zeros((n, 3))
    qq = np.zeros((n, 3))
    for v in range(3):
        r = np.asarray(rel)[:, v]
        p[:, v] = np.einsum("ns,nt,nst->n", s3[:, :, v], t3[:, :, v], np.asarray(ff, np.float64)[:, :, r].transpose(2, 0, 1))
        qq[:, v] = np.einsum("ns,nt,nst->n", s3[:, :, v], t3[:, :, v], np.asarray(dd, np.float64)[:, :, r].transpose(2, 0, 1))
    return p, qq


def additive_eval(comps, xyz, coeffs, ids, stride=1.0):
    """EvalDerivFunc.forward (additive.py:106-118) in float64: funcs, dfuncs (N, 3)."""
    x = np.asarray(xyz, np.float64)
    m = np.asarray(coeffs, np.float64)[np.asarray(ids)]
    f = np.zeros_like(x)
    df = np.zeros_like(x)
    for ci, c in enumerate(comps):
        f += q(c, x) * m[:, 3 * ci:3 * ci + 3]
        df += dq(c, x) * m[:, 3 * ci:3 * ci + 3] / stride
    return f, df
