"""Learned additive bases on the GPU: DDI and basis evaluation (forward and backward) against the
oracle, and the additive form of the B-spline basis against the B-spline basis itself
(additive.py, ddi.cu, bezier_tensor.py)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _dir_check(f, x0, grad, rng, k=4, eps=1e-3):
    """<grad, d> against central differences of the float64 oracle f along random directions d."""
    for _ in range(k):
        d = rng.standard_normal(x0.shape)
        fd = (f(x0 + eps * d) - f(x0 - eps * d)) / (2 * eps)
        got = float(np.sum(grad * d))
        assert abs(got - fd) <= 1e-4 * (1 + abs(fd)), (got, fd)


@pytest.mark.parametrize("m,nc", [(0, 4), (2, 4), (0, 12), (1, 17)])
def test_ddi_forward_backward(cuda, m, nc):
    """nc = 4: tables and coefficient rows in LDS; nc = 12: tables only; nc = 17: tables read from
    global memory."""
    from ngspsr import additive as A
    from oracle import additive as OA
    print(f"[case] ddi m={m} components={nc}", flush=True)
    comps = [A.Power(0), A.Power(2), A.Bezier(), A.SinPeaks(2)] if nc == 4 else \
        [A.Power(p % 7) if p < 7 else A.Product(A.Power(p % 7), A.Bezier()) for p in range(nc - 1)] + [A.Bezier()]
    b = A.AdditiveBasis(comps, [1.0] * nc)
    rng = np.random.default_rng(m)
    n, L = 3000, 3 * 2 ** m + 2
    s = rng.standard_normal((n, 3 * nc)).astype(np.float32)
    t = rng.standard_normal((n, 3 * nc)).astype(np.float32)
    rel = rng.integers(0, L, (n, 3))
    ff = (b.ff_integral[m]["f_f"].data * 2).numpy()
    dd = (b.ff_integral[m]["df_df"].data / 2).numpy()
    st = torch.from_numpy(s).to(cuda).requires_grad_()
    tt = torch.from_numpy(t).to(cuda).requires_grad_()
    p, q = A.ddi_forward(st, tt, torch.from_numpy(rel).to(cuda), torch.from_numpy(ff).to(cuda),
                         torch.from_numpy(dd).to(cuda))
    rp, rq = OA.ddi_forward(s, t, rel, ff, dd)
    np.testing.assert_allclose(p.detach().cpu().numpy(), rp, rtol=1e-5, atol=1e-5 * np.abs(rp).max())
    np.testing.assert_allclose(q.detach().cpu().numpy(), rq, rtol=1e-5, atol=1e-5 * np.abs(rq).max())
    gp = rng.standard_normal((n, 3)).astype(np.float32)
    gq = rng.standard_normal((n, 3)).astype(np.float32)
    ((p * torch.from_numpy(gp).to(cuda)).sum() + (q * torch.from_numpy(gq).to(cuda)).sum()).backward()

    def loss_s(x):
        a, c = OA.ddi_forward(x, t, rel, ff, dd)
        return float(np.sum(a * gp) + np.sum(c * gq))

    def loss_t(x):
        a, c = OA.ddi_forward(s, x, rel, ff, dd)
        return float(np.sum(a * gp) + np.sum(c * gq))

    _dir_check(loss_s, s.astype(np.float64), st.grad.cpu().numpy().astype(np.float64), rng)
    _dir_check(loss_t, t.astype(np.float64), tt.grad.cpu().numpy().astype(np.float64), rng)


def test_additive_evaluation_forward_backward(cuda):
    from ngspsr import additive as A
    from oracle import additive as OA
    print("[case] additive evaluation", flush=True)
    comps = [A.Power(0), A.Power(3), A.Bezier(), A.SinPeaks(2), A.Product(A.Power(1), A.Bezier())]
    ocomps = [("power", 0), ("power", 3), ("bezier",), ("sin", np.pi * 2 / 3, -np.pi + np.pi),
              ("product", ("power", 1), ("bezier",))]
    b = A.AdditiveBasis(comps, [1.0] * 5)
    rng = np.random.default_rng(2)
    n, M, stride = 4000, 60, 4
    xyz = rng.uniform(-1.6, 1.6, (n, 3)).astype(np.float32)
    feat = rng.standard_normal((M, 15)).astype(np.float32)
    ids = rng.integers(0, M, n)
    xt = torch.from_numpy(xyz).to(cuda)
    ft = torch.from_numpy(feat).to(cuda).requires_grad_()
    it = torch.from_numpy(ids).to(cuda)
    val = b.evaluate(ft, xt, it)
    grad = b.evaluate_derivative(ft, xt, it, stride)
    f, df = OA.additive_eval(ocomps, xyz, feat, ids, stride)
    ref_val = f[:, 0] * f[:, 1] * f[:, 2]
    ref_grad = np.stack([df[:, 0] * f[:, 1] * f[:, 2], f[:, 0] * df[:, 1] * f[:, 2], f[:, 0] * f[:, 1] * df[:, 2]], 1)
    np.testing.assert_allclose(val.detach().cpu().numpy(), ref_val, rtol=1e-4, atol=1e-5 * np.abs(ref_val).max())
    np.testing.assert_allclose(grad.detach().cpu().numpy(), ref_grad, rtol=1e-4, atol=1e-5 * np.abs(ref_grad).max())
    w1 = rng.standard_normal(n).astype(np.float32)
    w2 = rng.standard_normal((n, 3)).astype(np.float32)
    ((val * torch.from_numpy(w1).to(cuda)).sum() + (grad * torch.from_numpy(w2).to(cuda)).sum()).backward()

    def loss(fe):
        f_, df_ = OA.additive_eval(ocomps, xyz, fe, ids, stride)
        v = f_[:, 0] * f_[:, 1] * f_[:, 2]
        g = np.stack([df_[:, 0] * f_[:, 1] * f_[:, 2], f_[:, 0] * df_[:, 1] * f_[:, 2], f_[:, 0] * f_[:, 1] * df_[:, 2]], 1)
        return float(np.sum(v * w1) + np.sum(g * w2))

    _dir_check(loss, feat.astype(np.float64), ft.grad.cpu().numpy().astype(np.float64), rng)


@pytest.mark.parametrize("ss,ts", [(1, 1), (1, 2), (2, 8), (1, 16)])
def test_additive_bezier_equals_bspline_basis(cuda, ss, ts):
    """AdditiveBasis.get_bezier() with unit features is the B-spline basis: its integrals equal
    BezierTensorBasis's on the same COO entries (float32 tables, so to rounding)."""
    from ngspsr import additive as A
    from ngspsr.bases import BezierTensorBasis
    print(f"[case] additive bezier strides {ss} {ts}", flush=True)
    b, bz = A.AdditiveBasis.get_bezier(), BezierTensorBasis()
    rng = np.random.default_rng(ss * 31 + ts)
    n, M = 5000, 40
    mult = ts // ss
    # integer offsets, some outside the table (masked to 0 by both)
    rel = torch.from_numpy(rng.integers(-2 * mult - 1, 2 * mult + 2, (n, 3)).astype(np.float32)).to(cuda)
    ones = torch.ones((M, 3), device=cuda)
    sid = torch.from_numpy(rng.integers(0, M, n)).to(cuda)
    tid = torch.from_numpy(rng.integers(0, M, n)).to(cuda)
    got = b.integrate_deriv_deriv_product(ones, ones, rel, ss, ts, sid, tid)
    ref = bz.integrate_deriv_deriv_product(rel, ss, ts)
    torch.testing.assert_close(got, ref, rtol=2e-6, atol=2e-6 * float(ref.abs().max()))
    data = torch.from_numpy(rng.standard_normal((n, 3)).astype(np.float32)).to(cuda)
    for ds, tg in ((ss, ts), (ts, ss)) if ss != ts else ((ss, ts),):
        m = max(ds, tg) // min(ds, tg)
        # offsets whose table index stays in range (neither form masks this lookup)
        if tg >= ds:
            rr = rng.integers(-m, m + 1, (n, 3)).astype(np.float32)
        else:
            rr = rng.integers(-(m // 2), m // 2 + 1, (n, 3)).astype(np.float32) / m
        rr = torch.from_numpy(rr).to(cuda)
        got = b.integrate_const_deriv_product(data, ones, rr, ds, tg, tid)
        ref = bz.integrate_const_deriv_product(data, rr, ds, tg)
        torch.testing.assert_close(got, ref, rtol=2e-6, atol=2e-6 * float(ref.abs().max()))
    xyz = torch.from_numpy(rng.uniform(-1.6, 1.6, (n, 3)).astype(np.float32)).to(cuda)
    torch.testing.assert_close(b.evaluate(ones, xyz, sid), bz.evaluate(None, xyz, None), rtol=1e-6, atol=1e-7)
    torch.testing.assert_close(b.evaluate_derivative(ones, xyz, sid, ts), bz.evaluate_derivative(None, xyz, None, ts),
                               rtol=1e-6, atol=1e-7)
