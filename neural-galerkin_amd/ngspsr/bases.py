"""Degree-2 tensor B-spline basis of the jspsr path (jspsr/bases/bezier_tensor.py:11-128).

B(x) = 2 N2(x): (x+1.5)^2 on [-1.5,-0.5), -2x^2+1.5 on [-0.5,0.5), (x-1.5)^2 on [0.5,1.5).

The Galerkin integral tables the reference loads from a pickle (bezier_tensor.py:27-34) are
computed here from their definitions by Gauss-Legendre quadrature on every polynomial piece (exact
for these piecewise polynomials up to rounding), for multiples m = 2^k:
  F_m[i] = int B(x+d) B(x/m) dx,  G_m[i] = int B'(x+d) B'(x/m) dx,  d = i - 1.5m - 0.5,  i < 3m+2
  P_m[i] = int_{ra}^{ra+1} B(x/m) dx,  Q_m[i] = int_{ra}^{ra+1} B'(x/m) dx,  ra = 1.5m-1-i,  i < 3m
  Pinv_m[i] = int_{max(ra,-1.5)}^{min(ra+m,1.5)} B,  Qinv_m[i] = B(min(ra+m,1.5)) - B(max(ra,-1.5)),
  ra = 0.5 - i, i < m+2 (m >= 2)
B' is the derivative function evaluated at the argument. With the scalings of bezier_tensor.py
(F x s_src, G / s_tgt, P x s_data, Q x s_data/s_tgt, Pinv x s_tgt) each product is the physical
Galerkin integral.
"""
import math
from fractions import Fraction
from functools import lru_cache

import numpy as np
import torch

_GL_X, _GL_W = np.polynomial.legendre.leggauss(4)
_KNOTS = (-1.5, -0.5, 0.5, 1.5)


def _b(x):
    x = np.asarray(x, np.float64)
    return np.where((x >= -1.5) & (x < -0.5), (x + 1.5) ** 2,
                    np.where((x >= -0.5) & (x < 0.5), -2 * x * x + 1.5, np.where((x >= 0.5) & (x < 1.5), (x - 1.5) ** 2, 0.0)))


def _db(x):
    x = np.asarray(x, np.float64)
    return np.where((x >= -1.5) & (x < -0.5), 2 * x + 3,
                    np.where((x >= -0.5) & (x < 0.5), -4 * x, np.where((x >= 0.5) & (x < 1.5), 2 * x - 3, 0.0)))


def _piecewise_integral(f, breaks, lo, hi):
    pts = sorted(set([lo, hi] + [b for b in breaks if lo < b < hi]))
    total = 0.0
    for a, b in zip(pts[:-1], pts[1:]):
        xm, xr = 0.5 * (a + b), 0.5 * (b - a)
        total += xr * float(np.sum(_GL_W * f(xm + xr * _GL_X)))
    return total


@lru_cache(maxsize=None)
def tables_for_multiple(m: int):
    """(F, G, P, Q, Pinv, Qinv) float64 arrays of multiple m (Pinv/Qinv empty for m = 1)."""
    F, G = [], []
    for i in range(3 * m + 2):
        d = i - 1.5 * m - 0.5
        br = [k - d for k in _KNOTS] + [k * m for k in _KNOTS]
        lo, hi = min(br), max(br)
        F.append(_piecewise_integral(lambda x: _b(x + d) * _b(x / m), br, lo, hi))
        G.append(_piecewise_integral(lambda x: _db(x + d) * _db(x / m), br, lo, hi))
    P, Q = [], []
    for i in range(3 * m):
        ra = 1.5 * m - 1 - i
        br = [k * m for k in _KNOTS]
        P.append(_piecewise_integral(lambda x: _b(x / m), br, ra, ra + 1))
        Q.append(_piecewise_integral(lambda x: _db(x / m), br, ra, ra + 1))
    Pinv, Qinv = [], []
    if m >= 2:
        for i in range(m + 2):
            ra = 0.5 - i
            lo, hi = max(ra, -1.5), min(ra + m, 1.5)
            Pinv.append(_piecewise_integral(_b, list(_KNOTS), lo, hi) if lo < hi else 0.0)
            Qinv.append(float(_b(hi) - _b(lo)) if lo < hi else 0.0)
    # The entries are rationals with small denominators; snap away quadrature rounding.
    snap = [float(Fraction(v).limit_denominator(1 << 20)) for t in (F, G, P, Q, Pinv, Qinv) for v in t]
    out, pos = [], 0
    for t in (F, G, P, Q, Pinv, Qinv):
        out.append(np.array(snap[pos:pos + len(t)], np.float64))
        pos += len(t)
    return tuple(out)


MAX_LEVELS = 8


def packed_tables(levels: int = MAX_LEVELS):
    """Float32 buffer of all tables for m = 1 .. 2^(levels-1) and the offsets ngs_set_tables wants."""
    chunks, offsets, pos = [], [], 0
    for k in range(levels):
        for t in tables_for_multiple(2 ** k):
            offsets.append(pos)
            chunks.append(t.astype(np.float32))
            pos += t.shape[0]
    return np.concatenate(chunks), np.array(offsets, np.int64)


_DEVICE_TABLES = {}


def ensure_device_tables(device):
    """Upload the tables once per device and register them with the library."""
    from . import _lib
    key = str(device)
    if key not in _DEVICE_TABLES:
        buf, off = packed_tables()
        tdev = torch.from_numpy(buf).to(device)
        _lib.call("ngs_set_tables", _lib.ptr(tdev), off.ctypes.data_as(ctypes_p()), MAX_LEVELS)
        _DEVICE_TABLES[key] = (tdev, off)
    return _DEVICE_TABLES[key][0]


def ctypes_p():
    import ctypes
    return ctypes.POINTER(ctypes.c_int64)


class BezierTensorBasis:
    """Evaluation side of BezierTensorBasis (bezier_tensor.py:37-73) on torch tensors."""

    @staticmethod
    def get_domain_range():
        return 3

    @staticmethod
    def evaluate_single(x):
        b1 = (x + 1.5) ** 2
        b2 = -2 * (x ** 2) + 1.5
        b3 = (x - 1.5) ** 2
        m1 = (x >= -1.5) & (x < -0.5)
        m2 = (x >= -0.5) & (x < 0.5)
        m3 = (x >= 0.5) & (x < 1.5)
        return m1 * b1 + m2 * b2 + m3 * b3

    @staticmethod
    def evaluate_derivative_single(x):
        b1 = 2 * x + 3
        b2 = -4 * x
        b3 = 2 * x - 3
        m1 = (x >= -1.5) & (x < -0.5)
        m2 = (x >= -0.5) & (x < 0.5)
        m3 = (x >= 0.5) & (x < 1.5)
        return m1 * b1 + m2 * b2 + m3 * b3

    def evaluate(self, feat, xyz, feat_ids):
        return self.evaluate_single(xyz[..., 0]) * self.evaluate_single(xyz[..., 1]) * self.evaluate_single(xyz[..., 2])

    def evaluate_derivative(self, feat, xyz, feat_ids, stride):
        x, y, z = xyz[..., 0], xyz[..., 1], xyz[..., 2]
        bx, by, bz = self.evaluate_single(x), self.evaluate_single(y), self.evaluate_single(z)
        dx = self.evaluate_derivative_single(x) / stride
        dy = self.evaluate_derivative_single(y) / stride
        dz = self.evaluate_derivative_single(z) / stride
        return torch.stack([dx * by * bz, bx * dy * bz, bx * by * dz], dim=1)

    def integrate_deriv_deriv_product(self, rel_pos, source_stride: int, target_stride: int, **kwargs):
        """bezier_tensor.py:75-96 on torch tensors (COO entries)."""
        assert source_stride <= target_stride
        m = target_stride // source_stride
        F, G, *_ = tables_for_multiple(m)
        idx = (rel_pos + m * 3 / 2 + 0.5).round().long()
        invalid = ~((idx >= 0) & (idx < F.shape[0]))
        idx = idx.clamp(0, F.shape[0] - 1)
        Ft = torch.tensor(F, dtype=torch.float32, device=rel_pos.device)
        Gt = torch.tensor(G, dtype=torch.float32, device=rel_pos.device)
        c = Ft[idx] * source_stride
        c[invalid] = 0.0
        g = Gt[idx] / target_stride
        g[invalid] = 0.0
        return g[:, 0] * c[:, 1] * c[:, 2] + c[:, 0] * g[:, 1] * c[:, 2] + c[:, 0] * c[:, 1] * g[:, 2]

    def integrate_const_deriv_product(self, data, rel_pos, data_stride: int, target_stride: int, **kwargs):
        """bezier_tensor.py:98-121 on torch tensors (COO entries)."""
        if target_stride >= data_stride:
            m = target_stride // data_stride
            idx = (rel_pos + (3 * m - 1) / 2.0).round().long()
            _, _, dpi, pi, _, _ = tables_for_multiple(m)
            i_mult = data_stride / target_stride
            cs = data_stride
        else:
            m = data_stride // target_stride
            idx = (rel_pos * m + m / 2 + 0.5).round().long()
            _, _, _, _, dpi, pi = tables_for_multiple(m)
            i_mult = 1.0
            cs = target_stride
        dev = rel_pos.device
        c = torch.tensor(dpi, dtype=torch.float32, device=dev)[idx] * cs
        g = torch.tensor(pi, dtype=torch.float32, device=dev)[idx] * i_mult
        return data[:, 0] * g[:, 0] * c[:, 1] * c[:, 2] + data[:, 1] * c[:, 0] * g[:, 1] * c[:, 2] + \
            data[:, 2] * c[:, 0] * c[:, 1] * g[:, 2]


def log2_int(s: int) -> int:
    k = int(round(math.log2(s)))
    assert 2 ** k == s, "strides must be powers of two"
    return k
