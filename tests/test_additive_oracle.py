"""Learned additive bases: integral tables, Hermite null space and the DDI restatement pinned by
symbolic integration, closed forms and direct quadrature (additive.py, hermite_ns.py, ddi.cu). CPU."""
import numpy as np
import pytest
import torch

from oracle import additive as OA


def _basis(comps):
    from ngspsr import additive as A
    return A.AdditiveBasis(comps, [1.0] * len(comps))


def test_bezier_tables_equal_the_closed_forms():
    """AdditiveBasis.get_bezier() is by construction the B-spline basis: its tables are the
    rational closed forms (F, G / M, P, Q / M, Pinv, Qinv) up to float32 rounding."""
    from ngspsr import additive as A, bases as B
    b = A.AdditiveBasis.get_bezier()
    for m in range(5):
        M = 2 ** m
        F, G, P, Q, Pi, Qi = B.tables_for_multiple(M)
        np.testing.assert_allclose(b.ff_integral[m]["f_f"].data.numpy()[0, 0], F, atol=2e-7)
        np.testing.assert_allclose(b.ff_integral[m]["df_df"].data.numpy()[0, 0], G / M, atol=2e-7)
        np.testing.assert_allclose(b.f_integral[0]["f"][m].data.numpy(), P, atol=2e-7)
        np.testing.assert_allclose(b.f_integral[0]["df"][m].data.numpy(), Q / M, atol=2e-7)
        if m:
            np.testing.assert_allclose(b.f_integral[0]["f_inv"][m - 1].data.numpy(), Pi, atol=2e-7)
            np.testing.assert_allclose(b.f_integral[0]["df_inv"][m - 1].data.numpy(), Qi, atol=2e-7)


@pytest.mark.parametrize("ca,cb,m,k", [
    (("power", 3), ("bezier",), 1, 2),
    (("sin", 2.0943951023931953, -0.5235987755982991), ("power", 2), 0, 1),
    (("product", ("power", 1), ("bezier",)), ("power", 0), 1, 5),
])
def test_tables_match_symbolic_integration(ca, cb, m, k):
    from ngspsr import additive as A

    def comp(c):
        if c[0] == "power":
            return A.Power(c[1])
        if c[0] == "bezier":
            return A.Bezier()
        if c[0] == "sin":
            return A.Sinusoid(c[1], c[2])
        return A.Product(comp(c[1]), comp(c[2]))

    b = A.AdditiveBasis([comp(ca), comp(cb)], [1.0, 1.0])
    f_f, df_df = OA.ff_entry(ca, cb, m, k)
    assert abs(b.ff_integral[m]["f_f"].data[0, 1, k].item() - f_f) <= 1e-6 * (1 + abs(f_f))
    assert abs(b.ff_integral[m]["df_df"].data[0, 1, k].item() - df_df) <= 1e-6 * (1 + abs(df_df))
    f, df = OA.f_entry(ca, m, k % (3 * 2 ** m))
    assert abs(b.f_integral[0]["f"][m].data[k % (3 * 2 ** m)].item() - f) <= 1e-6 * (1 + abs(f))
    assert abs(b.f_integral[0]["df"][m].data[k % (3 * 2 ** m)].item() - df) <= 1e-6 * (1 + abs(df))
    fi, dfi = OA.f_inv_entry(ca, m + 1, 1)
    assert abs(b.f_integral[0]["f_inv"][m].data[1].item() - fi) <= 1e-6 * (1 + abs(fi))
    assert abs(b.f_integral[0]["df_inv"][m].data[1].item() - dfi) <= 1e-6 * (1 + abs(dfi))


@pytest.mark.parametrize("m", [0, 1, 2])
def test_ddi_is_the_integral_of_the_assembled_axis_functions(m):
    """DDI p and q are the 1-D integrals of the per-axis functions F_s = sum_c s_c q_c and F_t:
    p = s_src int F_s(x + dx) F_t(x / M) dx, q = int F_s'(x + dx) d/dx F_t(x / M) dx / s_src."""
    comps = [("power", 0), ("power", 2), ("bezier",)]
    from ngspsr import additive as A
    b = A.AdditiveBasis([A.Power(0), A.Power(2), A.Bezier()], [1.0, 1.0, 1.0])
    M = 2 ** m
    L = 3 * M + 2
    rng = np.random.default_rng(m)
    n = 12
    s = rng.standard_normal((n, 9))
    t = rng.standard_normal((n, 9))
    rel = rng.integers(0, L, (n, 3))
    ss = 2
    ff = b.ff_integral[m]["f_f"].data.numpy().astype(np.float64) * ss
    dd = b.ff_integral[m]["df_df"].data.numpy().astype(np.float64) / ss
    p, qq = OA.ddi_forward(s, t, rel, ff, dd)
    xs, ws = np.polynomial.legendre.leggauss(40)

    def quad(fun, lo, hi, brk):
        pts = sorted({lo, hi} | {v for v in brk if lo < v < hi})
        return sum(0.5 * (b_ - a) * np.sum(ws * fun(0.5 * (a + b_) + 0.5 * (b_ - a) * xs)) for a, b_ in zip(pts[:-1], pts[1:]))

    def axis_f(coef, x):
        x = np.asarray(x)
        sup = (x >= -1.5) & (x < 1.5)
        return sum(coef[c] * np.where(sup, OA.q(comps[c], x), 0.0) for c in range(3))

    def axis_df(coef, x):
        x = np.asarray(x)
        sup = (x >= -1.5) & (x < 1.5)
        return sum(coef[c] * np.where(sup, OA.dq(comps[c], x), 0.0) for c in range(3))

    for i in range(n):
        for v in range(3):
            dx = rel[i, v] - 1.5 * M - 0.5
            cs, ct = s[i, v::3], t[i, v::3]
            brk = [u - dx for u in (-1.5, -0.5, 0.5, 1.5)] + [u * M for u in (-1.5, -0.5, 0.5, 1.5)]
            lo, hi = max(-1.5 - dx, -1.5 * M), min(1.5 - dx, 1.5 * M)
            ep = ss * quad(lambda x: axis_f(cs, x + dx) * axis_f(ct, x / M), lo, hi, brk)
            eq = quad(lambda x: axis_df(cs, x + dx) * axis_df(ct, x / M) / M, lo, hi, brk) / ss
            assert abs(p[i, v] - ep) < 1e-5 * (1 + abs(ep)) and abs(qq[i, v] - eq) < 1e-5 * (1 + abs(eq))


def test_hermite_null_space_and_initialisation():
    from ngspsr import additive as A
    for n in (4, 6, 8):
        ns = A.hermite_null_space(n)
        assert ns.shape == (n - 3, n + 1)
        for row in ns:
            poly = np.polynomial.Polynomial(row)
            for x in (-1.5, 1.5):
                assert abs(poly(x)) < 1e-10 and abs(poly.deriv()(x)) < 1e-10
    h = A.AdditiveHermite(6, [A.Bezier()], [0.25])
    feat = torch.zeros((5, h.get_feature_size()))
    h.initialize_feature_value(feat)
    coef = h.feature_to_coefficient(feat).numpy().astype(np.float64)
    assert np.all(coef[:, :3] == 0.25)
    # the polynomial part is the degree-4 Hermite polynomial x^4 - 4.5 x^2 + 81/16, up to scale
    ns4 = A.hermite_null_space(4)[0]
    poly = coef[0, 3::3]
    assert np.allclose(poly[:5], ns4, atol=1e-6) and np.allclose(poly[5:], 0, atol=1e-6)
    assert np.allclose(ns4 / ns4[4], [81 / 16, 0, -4.5, 0, 1])
