"""Degree-2 B-spline (x2 scaled) basis and its Galerkin integral tables (oracle; test infrastructure).

The reference ships the tables as a pickle that is absent from the repository
(jspsr/bases/bezier_tensor.py:27-34). They are fully determined by closed forms; this module
computes them as exact rationals by piecewise-polynomial integration.

Conventions (jspsr/bases/bezier_tensor.py:75-121 index/scale use; additive.py:17-52,172-249 is the
reference's own generator for the same quantities):
  B(x)        = (x+1.5)^2 | -2x^2+1.5 | (x-1.5)^2 on [-1.5,-.5) | [-.5,.5) | [.5,1.5)   (:17-23,37-44)
  B'(x)       = the derivative function (:47-54), evaluated at the argument, no chain rule.
  m = 2^k, i indexes the table row.
  F_m[i] = int B(x+d) B(x/m) dx,    d = i - 1.5m - 0.5, i in [0, 3m+1]   SHIFTED_SELF_INTEGRAL  (x s_src)
  G_m[i] = int B'(x+d) B'(x/m) dx,  same d                                SHIFTED_DERIVATIVE_INTEGRAL (/ s_tgt)
  P_m[i] = int_{ra}^{ra+1} B(x/m) dx,  ra = 1.5m-1-i, i in [0,3m-1]    DERIVATIVE_PARTIAL_INTEGRAL (x s_data)
  Q_m[i] = int_{ra}^{ra+1} B'(x/m) dx = m (B((ra+1)/m) - B(ra/m))      PARTIAL_INTEGRAL (x s_data/s_tgt)
  Pinv_m[i] = int_{max(ra,-1.5)}^{min(ra+m,1.5)} B(u) du, ra = 0.5-i, i in [0,m+1], m>=2  (x s_tgt)
  Qinv_m[i] = B(min(ra+m,1.5)) - B(max(ra,-1.5))                                           (x 1)
With these, every product in bezier_tensor.py equals the physical integral
int grad B_src . grad B_tgt (resp. int grad B_tgt . V over a data voxel); tests pin this against
brute-force quadrature. The same quantities from additive.py's generator (which differentiates
through the x/m substitution) differ by the factor m that AdditiveBasis applies in its own
scaling (additive.py:337-339 divides df_df by s_src; bezier_tensor.py:90 divides G by s_tgt),
so AdditiveBasis.get_bezier() and BezierTensorBasis agree, as the reference states (:406-409).
"""
from fractions import Fraction as Fr
from functools import lru_cache

import numpy as np

_H = Fr(1, 2)
_BREAKS = [Fr(-3, 2), Fr(-1, 2), Fr(1, 2), Fr(3, 2)]


def _poly_mul(a, b):
    out = [Fr(0)] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            out[i + j] += x * y
    return out


def _poly_eval(p, x):
    r = Fr(0)
    for c in reversed(p):
        r = r * x + c
    return r


def _poly_int(p, lo, hi):
    r = Fr(0)
    for k, c in enumerate(p):
        r += c * (hi ** (k + 1) - lo ** (k + 1)) / (k + 1)
    return r


def _piece(x):
    """Index of the B piece containing x (-1 outside the support)."""
    for k in range(3):
        if _BREAKS[k] <= x < _BREAKS[k + 1]:
            return k
    return -1


# B pieces as polynomials in y (coefficients low->high).
_B_PIECES = [[Fr(9, 4), Fr(3), Fr(1)], [Fr(3, 2), Fr(0), Fr(-2)], [Fr(9, 4), Fr(-3), Fr(1)]]
_DB_PIECES = [[Fr(3), Fr(2)], [Fr(0), Fr(-4)], [Fr(-3), Fr(2)]]


def _affine(p, scale, shift):
    """p(scale*x + shift) as a polynomial in x."""
    out = [Fr(0)]
    lin = [Fr(shift), Fr(scale)]
    powk = [Fr(1)]
    for c in p:
        term = [c * t for t in powk]
        if len(term) > len(out):
            out += [Fr(0)] * (len(term) - len(out))
        for i, t in enumerate(term):
            out[i] += t
        powk = _poly_mul(powk, lin)
    return out


def _pieces_of(fn_pieces, scale, shift):
    """List of (lo, hi, poly in x) for f(scale*x+shift), f piecewise on _BREAKS (scale>0)."""
    res = []
    for k in range(3):
        lo = (_BREAKS[k] - shift) / scale
        hi = (_BREAKS[k + 1] - shift) / scale
        res.append((lo, hi, _affine(fn_pieces[k], scale, shift)))
    return res


def _int_product(pa, pb, lo=None, hi=None):
    total = Fr(0)
    for (alo, ahi, a) in pa:
        for (blo, bhi, b) in pb:
            l, h = max(alo, blo), min(ahi, bhi)
            if lo is not None:
                l = max(l, lo)
            if hi is not None:
                h = min(h, hi)
            if l < h:
                total += _poly_int(_poly_mul(a, b), l, h)
    return total


def B_exact(x):
    k = _piece(Fr(x))
    return Fr(0) if k < 0 else _poly_eval(_B_PIECES[k], Fr(x))


@lru_cache(maxsize=None)
def F_table(m: int):
    d0 = -Fr(3 * m, 2) - _H
    return tuple(_int_product(_pieces_of(_B_PIECES, 1, d0 + i), _pieces_of(_B_PIECES, Fr(1, m), 0))
                 for i in range(3 * m + 2))


@lru_cache(maxsize=None)
def G_table(m: int):
    d0 = -Fr(3 * m, 2) - _H
    return tuple(_int_product(_pieces_of(_DB_PIECES, 1, d0 + i), _pieces_of(_DB_PIECES, Fr(1, m), 0))
                 for i in range(3 * m + 2))


@lru_cache(maxsize=None)
def P_table(m: int):
    out = []
    for i in range(3 * m):
        ra = Fr(3 * m, 2) - 1 - i
        out.append(_int_product(_pieces_of(_B_PIECES, Fr(1, m), 0), [(ra, ra + 1, [Fr(1)])]))
    return tuple(out)


@lru_cache(maxsize=None)
def Q_table(m: int):
    out = []
    for i in range(3 * m):
        ra = Fr(3 * m, 2) - 1 - i
        out.append(m * (B_exact((ra + 1) / m) - B_exact(ra / m)))
    return tuple(out)


@lru_cache(maxsize=None)
def Pinv_table(m: int):
    assert m >= 2
    out = []
    for i in range(m + 2):
        ra = _H - i
        lo, hi = max(ra, -Fr(3, 2)), min(ra + m, Fr(3, 2))
        out.append(_int_product(_pieces_of(_B_PIECES, 1, 0), [(lo, hi, [Fr(1)])]) if lo < hi else Fr(0))
    return tuple(out)


@lru_cache(maxsize=None)
def Qinv_table(m: int):
    assert m >= 2
    out = []
    for i in range(m + 2):
        ra = _H - i
        lo, hi = max(ra, -Fr(3, 2)), min(ra + m, Fr(3, 2))
        out.append(B_exact(hi) - B_exact(lo) if lo < hi else Fr(0))
    return tuple(out)


def as_f64(table):
    return np.array([float(v) for v in table], dtype=np.float64)


# ---------------------------------------------------------------- evaluation (FP64)

def B(x):
    """bezier_tensor.py:37-44 (vectorised, float64)."""
    x = np.asarray(x, dtype=np.float64)
    b1 = (x + 1.5) ** 2
    b2 = -2 * x * x + 1.5
    b3 = (x - 1.5) ** 2
    return np.where((x >= -1.5) & (x < -0.5), b1,
                    np.where((x >= -0.5) & (x < 0.5), b2, np.where((x >= 0.5) & (x < 1.5), b3, 0.0)))


def dB(x):
    """bezier_tensor.py:46-54."""
    x = np.asarray(x, dtype=np.float64)
    return np.where((x >= -1.5) & (x < -0.5), 2 * x + 3,
                    np.where((x >= -0.5) & (x < 0.5), -4 * x, np.where((x >= 0.5) & (x < 1.5), 2 * x - 3, 0.0)))


def integrate_deriv_deriv_product(rel_pos, source_stride: int, target_stride: int):
    """bezier_tensor.py:75-96: int grad B_src . grad B_tgt for COO entries with offset rel_pos."""
    assert source_stride <= target_stride
    m = target_stride // source_stride
    idx = np.rint(np.asarray(rel_pos, dtype=np.float64) + m * 3 / 2 + 0.5).astype(np.int64)
    F, G = as_f64(F_table(m)), as_f64(G_table(m))
    invalid = (idx < 0) | (idx >= F.shape[0])
    idx = np.clip(idx, 0, F.shape[0] - 1)
    c = F[idx] * source_stride
    c[invalid] = 0.0
    g = G[idx] / target_stride
    g[invalid] = 0.0
    return g[:, 0] * c[:, 1] * c[:, 2] + c[:, 0] * g[:, 1] * c[:, 2] + c[:, 0] * c[:, 1] * g[:, 2]


def integrate_const_deriv_product(data, rel_pos, data_stride: int, target_stride: int):
    """bezier_tensor.py:98-121: int grad B_tgt . data over the data voxel."""
    rel_pos = np.asarray(rel_pos, dtype=np.float64)
    if target_stride >= data_stride:
        m = target_stride // data_stride
        idx = np.rint(rel_pos + (3 * m - 1) / 2.0).astype(np.int64)
        dpi, pi = as_f64(P_table(m)), as_f64(Q_table(m))
        i_mult = data_stride / target_stride
        cstride = data_stride
    else:
        m = data_stride // target_stride
        idx = np.rint(rel_pos * m + m / 2 + 0.5).astype(np.int64)
        dpi, pi = as_f64(Pinv_table(m)), as_f64(Qinv_table(m))
        i_mult = 1.0
        cstride = target_stride
    c = dpi[idx] * cstride
    g = pi[idx] * i_mult
    data = np.asarray(data, dtype=np.float64)
    return data[:, 0] * g[:, 0] * c[:, 1] * c[:, 2] + data[:, 1] * c[:, 0] * g[:, 1] * c[:, 2] + \
        data[:, 2] * c[:, 0] * c[:, 1] * g[:, 2]
