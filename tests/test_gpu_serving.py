"""Continuous-batching server: results alone vs in the mix, batched vs single-system PCG."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _requests(cuda, k=6):
    from ngspsr import serving, synthetic
    out = []
    for i in range(k):
        rng = np.random.default_rng(300 + i)
        n = int(rng.integers(5000, 40000))
        depth = int(rng.integers(6, 8))
        xyz, nrm = synthetic.ellipsoids(n, k=4, sigma=float(rng.uniform(0, 0.003)), seed=300 + i)
        out.append((torch.from_numpy(xyz).to(cuda), torch.from_numpy(nrm).to(cuda), 2.0 ** -depth))
    return out


def test_batched_pcg_matches_single_solve(cuda):
    import ngspsr
    from ngspsr import serving, solver, synthetic
    print("[case] batched vs single pcg", flush=True)
    systems = []
    for i in range(3):
        xyz, nrm = synthetic.torus(8000 * (i + 1), seed=40 + i)
        res = ngspsr.reconstruct(torch.from_numpy(xyz).to(cuda), torch.from_numpy(nrm).to(cuda), 1 / 64, depth=3,
                                 min_density=4.0, solver="pcg@tol=1e-6", return_result=True)
        systems.append(res.reconstructor.systems[0])
    pool = serving.BatchedPCG(cuda, chunk=4)
    try:
        import threading
        outs = [None] * len(systems)

        def work(i):
            s, b = systems[i]
            outs[i] = pool.solve(s, b, tol=1e-6)
        th = [threading.Thread(target=work, args=(i,)) for i in range(len(systems))]
        [t.start() for t in th]
        [t.join() for t in th]
    finally:
        pool.close()
    for (s, b), (xb, info_b) in zip(systems, outs):
        xs, info_s = solver.pcg(s, b, tol=1e-6)
        torch.cuda.synchronize()
        assert info_b.iters == info_s.iters
        assert torch.equal(xb, xs), float((xb - xs).abs().max())


def test_alone_equals_in_the_mix(cuda):
    from ngspsr import serving
    print("[case] serving alone vs mix", flush=True)
    data = _requests(cuda)
    alone = []
    srv = serving.ReconstructionServer(cuda, workers=1, tol=1e-6, refine=True)
    try:
        for i, (x, n, vs) in enumerate(data):
            r = serving.Request(i, x, n, vs, depth=4, min_density=1.0)
            srv.serve([r])
            assert r.error is None, r.error
            alone.append(r)
    finally:
        srv.close()
    srv = serving.ReconstructionServer(cuda, workers=4, tol=1e-6, refine=True)
    try:
        mix = [serving.Request(i, x, n, vs, depth=4, min_density=1.0) for i, (x, n, vs) in enumerate(data)]
        srv.serve(mix)
    finally:
        srv.close()
    for a, m in zip(alone, mix):
        assert m.error is None, m.error
        assert a.max_rel_residual <= 1e-6 and m.max_rel_residual <= 1e-6
        va, fa = a.result
        vm, fm = m.result
        assert fa.shape == fm.shape and torch.equal(fa, fm)
        scale = float(va.abs().max())
        assert float((va - vm).abs().max()) <= 1e-6 * scale


def test_admission_by_footprint(cuda):
    """A budget of one request's footprint serialises the requests."""
    from ngspsr import serving
    print("[case] admission", flush=True)
    data = _requests(cuda, k=3)
    fp = max(serving.estimate_footprint(x.shape[0], vs) for x, _, vs in data)
    srv = serving.ReconstructionServer(cuda, workers=4, hbm_budget=fp, tol=1e-6, refine=False)
    peak = {"v": 0}
    orig = srv.run_request

    def tracked(req):
        with srv.lock:
            peak["v"] = max(peak["v"], srv.workers_busy)
        return orig(req)
    srv.run_request = tracked
    try:
        reqs = [serving.Request(i, x, n, vs) for i, (x, n, vs) in enumerate(data)]
        srv.serve(reqs)
    finally:
        srv.close()
    assert all(r.error is None for r in reqs)
    assert peak["v"] == 1
