"""Benchmark of the screened-Poisson reconstruction path on one MI355X.

Default workload (config 3, the metric's): a stream of 256 requests, each a seeded scene of 16
random ellipsoids with 1e4-1e6 points (log-uniform), depth 7-9 (voxel size 2^-depth), position
noise 0-0.005, 4 levels, screening 4.0, solved to relative residual 1e-6. It is served plain (one
request at a time, FP32 Jacobi-PCG) and by the continuous-batching server (mixed precision: FP32
inner PCG in one shared batched pool, FP64 residual check and refinement). A step is one pass of
the whole stream through the batched server with every request queued at t = 0; `value` is its
solves/s. Latency is measured under Poisson arrivals at 0.8x the plain capacity. The line also
carries the single-solve leg (config 2: 2M points on 16 noisy ellipsoids at 512^3, refinement to
1e-8) whose finest-level SpMV gives the roofline. With --config c1|c2|c5 the single-solve workload
is the headline (reconstructions/s). Points are in HBM when the timed region starts. Prints one
JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c3|c2|c1|c5] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "neural-galerkin_amd"))

CONFIGS = {
    "c2": dict(workload="ellipsoids16-2M-vs512-L4-mixed1e-8", gen="ellipsoids", n=2_000_000, vs=1 / 512, depth=4,
               min_density=1.0, solver="refine@tol=1e-8@inner_tol=1e-4", sigma=0.002, kappa=200.0),
    "c1": dict(workload="torus-100k-vs128-L4-pcg1e-6", gen="torus", n=100_000, vs=1 / 128, depth=4, min_density=8.0,
               solver="pcg@tol=1e-6@maxiter=10000"),
    "c5": dict(workload="room40boxes-10M-vs1024-L4-mixed1e-7", gen="room", n=10_000_000, vs=1 / 1024, depth=4,
               min_density=1.0, solver="refine@tol=1e-7@inner_tol=1e-4", sigma=0.003),
}
HBM_PEAK_GBS = 8000.0


def make_scene(cfg):
    from ngspsr import synthetic
    if cfg["gen"] == "ellipsoids":
        return synthetic.ellipsoids(cfg["n"], k=16, sigma=cfg["sigma"], kappa=cfg["kappa"], seed=0)
    if cfg["gen"] == "torus":
        return synthetic.torus(cfg["n"], seed=0)
    return synthetic.room(cfg["n"], sigma=cfg["sigma"], seed=0)


def _event_time(fn, reps):
    """Average seconds of fn() over reps launches, HIP events on the current stream (the stream the
    library launches on)."""
    import torch
    for _ in range(3):
        fn()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    e.synchronize()
    return s.elapsed_time(e) / 1e3 / reps


def spmv_roofline(system, reps=50):
    """Average duration of the SpMV kernel of an assembled level system and its bytes per launch.
    CSR (spmv_kernel): algorithmic 8 nnz + 4 (M+1) + 4 M (x) + 4 M (y).
    Run form: algorithmic 4 nnz (values) + 4 R (run words of the R non-empty runs) + 4 M (x) +
    4 M (y), zero slots and empty runs excluded. Packed (spmv_packed_kernel, the solver's form):
    the layout streams 100 M (all run words) + 4 nnz + 8 M + 8 per 64-row slice; the slot form
    (spmv_runs_kernel) streams 600 M + 8 M (all 125 value slots), timed beside it.
    Returns (seconds, algorithmic bytes, nnz, extra dict)."""
    import torch
    from ngspsr._lib import call, ptr, stream
    from ngspsr.solver import RunSystem
    x = torch.randn(system.M, device=system.device)
    y = torch.empty_like(x)
    M = system.M
    nnz = system.nnz()
    if isinstance(system, RunSystem):
        t = _event_time(lambda: call("ngs_spmv_runs", ptr(system.runs), ptr(system.rvals), ptr(x), ptr(y), M,
                                     stream()), reps)
        runs = int((system.counts() > 0).sum().item())
        algo = 4 * nnz + 4 * runs + 8 * M
        extra = {"nonempty_runs": runs, "layout_bytes": 600 * M + 8 * M,
                 "layout_GB/s": round((600 * M + 8 * M) / t / 1e9, 1),
                 "csr_equivalent_bytes": 8 * nnz + 4 * (M + 1) + 8 * M,
                 "csr_equivalent_GB/s": round((8 * nnz + 4 * (M + 1) + 8 * M) / t / 1e9, 1)}
        if not system.is_packed:
            return t, algo, nnz, extra
        # the packed form the solver runs: every run word, the values without the zero slots
        pvals, soff = system._packed
        tp = _event_time(lambda: call("ngs_spmv_packed", ptr(system.runs), ptr(pvals), ptr(soff), ptr(x), ptr(y),
                                      M, stream()), reps)
        equal = bool(torch.equal(system.matvec(x), system.matvec_slots(x)))
        packed_layout = 100 * M + 4 * nnz + 8 * M + 8 * (M // 64 + 1)
        slot = {"kernel": "spmv_runs_kernel (slot form: 5 value slots per run word, zeros included)",
                "launch_us": round(1e6 * t, 2), "achieved": round(algo / t / 1e9, 1),
                "frac": round(algo / t / 1e9 / HBM_PEAK_GBS, 4), **extra}
        extra = {"layout_bytes": int(packed_layout), "layout_GB/s": round(packed_layout / tp / 1e9, 1),
                 "nonempty_runs": runs, "equal_to_slot_form": equal, "slot_form_same_matrix": slot}
        return tp, algo, nnz, extra
    t = _event_time(lambda: call("ngs_spmv", ptr(system.rowptr), ptr(system.cols), ptr(system.vals), ptr(x),
                                 ptr(y), M, stream()), reps)
    return t, 8 * nnz + 4 * (M + 1) + 8 * M, nnz, {}


def stencil_roofline(system, reps=50):
    """Average durations of the two passes of the stencil operator (stencil_points_kernel,
    stencil_rows_kernel) and their algorithmic bytes, each array counted once. With the bitmap
    lookups:
      points: cell coords 12 n + basis 36 n (+ alpha 4 n) + x 4 M + level bitmap 8 W + u 4 n
      rows:   coords 12 M + x 4 M + y 4 M + level bitmap 8 W + cell bitmap 8 Wc + cell_start 4 (C+1)
              + basis 36 n + u 4 n
    (W, Wc: 32-bit words of the level and cell bitmaps, bits + rank)."""
    import torch
    from ngspsr._lib import call, ptr, stream
    x = torch.randn(system.M, device=system.device)
    y = torch.empty_like(x)
    st = system.struct()
    t_pts = _event_time(lambda: call("ngs_stencil_pass", st, ptr(x), ptr(y), 1, stream()), reps)
    t_rows = _event_time(lambda: call("ngs_stencil_pass", st, ptr(x), ptr(y), 2, stream()), reps)
    M, n = system.M, system.n
    W = int(st.level.grid.nwords)
    Wc = int(st.cells.nwords)
    C = system.bins.ncells if n else 0
    b_pts = n * (12 + 36 + 4 + (4 if system.alpha is not None else 0)) + 4 * M + 8 * W
    b_rows = 12 * M + 8 * M + 8 * W + 8 * Wc + 4 * (C + 1) + 40 * n
    return t_pts, b_pts, t_rows, b_rows


def stencil_probe(res, xyz, level=0):
    """The stencil form of one level of a finished reconstruction (rebuilt from its tree and the
    points), timed: the operator form that "auto" weighs against the run form."""
    import torch
    from ngspsr.reconstructor import PointBinning
    rec = res.reconstructor
    tree = res.tree
    M = tree.get_coords_size(None, level)
    rhs = torch.zeros(M, dtype=torch.float64, device=xyz.device)
    bins = PointBinning(tree, xyz, level, 4.0)
    S = rec.assemble_level(level, bins, None, 0.25 ** level, 0.0, rhs, form="stencil")
    t_pts, b_pts, t_rows, b_rows = stencil_roofline(S)
    return {"level": level, "rows": M, "points": int(S.n), "points_us": round(1e6 * t_pts, 2),
            "rows_us": round(1e6 * t_rows, 2), "operator_us": round(1e6 * (t_pts + t_rows), 2),
            "rows_bytes": int(b_rows), "points_bytes": int(b_pts), "rows_GB/s": round(b_rows / t_rows / 1e9, 1)}


def assembly_roofline(res, xyz, name="c2", reps=10):
    """Average duration of the finest-level assembly kernel (ngs_lhs_assemble_runs, screening with a
    coarse field at the points) and its algorithmic bytes, the survey's per-level figure with the run
    form's writes: 4 nnz (values) + 4 R (non-empty run words) + 4 M (inv_diag) + 8 M (rhs read) +
    8 M (rhs write) + 12 M (coords) + per point 12 (cell) + 36 (basis) + 4 (coarse field), bin order.
    The kernel is bound by its point-visit latency (27 rows visit every point), not by HBM."""
    import torch
    from ngspsr._lib import call, ptr, stream
    from ngspsr.reconstructor import PointBinning
    from ngspsr.solver import RunSystem
    rec, tree = res.reconstructor, res.tree
    M = tree.get_coords_size(None, 0)
    bins = PointBinning(tree, xyz, 0, 4.0)
    coarse = torch.randn(xyz.shape[0], device=xyz.device) * 0.1
    P = RunSystem.plane_stride(M)
    runs = torch.empty((25, P), dtype=torch.int32, device=xyz.device)
    rvals = torch.empty((125, P), dtype=torch.float32, device=xyz.device)
    inv_diag = torch.empty(M, dtype=torch.float32, device=xyz.device)
    rhs = torch.zeros(M, dtype=torch.float64, device=xyz.device)
    lv, pb = tree.level(0), bins.struct()
    # the coarse field in bin order, as the reconstruction passes it
    cb = torch.empty_like(coarse)
    call("ngs_gather_f32", ptr(coarse), ptr(bins.perm), bins.n, ptr(cb), stream())
    pb.perm = None
    t_slot = _event_time(lambda: call("ngs_lhs_assemble_runs", lv, pb, ptr(cb), 1.0, 0.1, ptr(runs), ptr(rvals),
                                      ptr(inv_diag), ptr(rhs), stream()), reps)
    R = RunSystem(runs, rvals, inv_diag, packed=False)
    nnz, nruns = R.nnz(), int((R.counts() > 0).sum().item())
    n = xyz.shape[0]
    algo = 4 * nnz + 4 * nruns + (4 + 16 + 12) * M + 52 * n
    out = {"kernel": "lhs_assemble_kernel (finest level, run form, screening with a coarse field)", "bound": "latency",
           "peak": HBM_PEAK_GBS, "unit": "GB/s", "algorithmic_bytes": int(algo), "rows": int(M), "points": int(n)}
    from ngspsr._lib import lib
    if RunSystem.PACKED and lib().ngs_assembly_lanes(int(pb.n), M) == 1:
        # the solver's form: written straight into the packed layout (slice offsets from the row counts)
        nsl = (M + 63) // 64
        row_nnz = torch.zeros(nsl * 64, dtype=torch.int32, device=xyz.device)
        call("ngs_stencil_row_counts", lv, ptr(row_nnz), stream())
        soff = torch.zeros(nsl + 1, dtype=torch.int64, device=xyz.device)
        torch.cumsum(row_nnz.view(nsl, 64).sum(1, dtype=torch.int64), 0, out=soff[1:])
        pvals = torch.empty(int(soff[-1].item()), dtype=torch.float32, device=xyz.device)
        t = _event_time(lambda: call("ngs_lhs_assemble_packed", lv, pb, ptr(cb), 1.0, 0.1, ptr(runs), ptr(pvals),
                                     ptr(soff), ptr(inv_diag), ptr(rhs), stream()), reps)
        out["kernel"] = "lhs_assemble_kernel (finest level, into the packed run form, screening with a coarse field)"
        out["slot_form_launch_us"] = round(1e6 * t_slot, 2)
    else:
        t = t_slot
    out.update({"achieved": round(algo / t / 1e9, 1), "frac": round(algo / t / 1e9 / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic("lhs_assemble_kernel", name), "launch_us": round(1e6 * t, 2)})
    return out


def pcg_iteration_rate(system, iters=200):
    """Seconds per PCG iteration (3 kernels) on the finest system, tolerance 0 so every iteration runs."""
    import torch
    from ngspsr import solver as S
    b = torch.randn(system.M, device=system.device)
    S.pcg(system, b, tol=0.0, maxiter=20, chunk=20, raise_on_cap=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, info = S.pcg(system, b, tol=0.0, maxiter=iters, chunk=iters, raise_on_cap=False)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / max(info.iters, 1)


def cpu_baseline(system, b, budget_s=15.0):
    """The oracle's C/OpenMP Jacobi-PCG (FP32, FP64 dots) on the finest-level system, timed over a
    bounded number of iterations (matrix conversion excluded). Returns seconds per iteration."""
    sys.path.insert(0, ROOT)
    from oracle import pcg as opcg
    prep = opcg.Prepared(arrays=(system.rowptr.cpu().numpy(), system.cols.cpu().numpy(), system.vals.cpu().numpy(),
                                 system.inv_diag.cpu().numpy()))
    bb = b.cpu().numpy()
    t0 = time.perf_counter()
    prep.run(bb, 1)
    t_one = time.perf_counter() - t0
    iters = int(max(2, min(200, budget_s / max(t_one, 1e-3))))
    t0 = time.perf_counter()
    _, it = prep.run(bb, iters)
    per_iter = (time.perf_counter() - t0) / max(it, 1)
    return per_iter, it


def fp64_error(csr, rhs, x, tol=1e-12, max_iter=3000):
    """Relative error of the GPU's finest-level solution against the oracle's all-FP64 Jacobi-PCG
    (C/OpenMP) solve of the same assembled system to relative residual `tol`, on the host, and the
    FP64 relative residual of the stored solution (FP32, as the reference stores it; the refinement's
    FP64 iterate is what max_level_rel_residual reports)."""
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import pcg as opcg
    a = csr.to_scipy()
    b = rhs.double().cpu().numpy()
    t0 = time.perf_counter()
    x64, it = opcg.solve_pcg(a, b, tol=tol, max_iter=max_iter, dtype=np.float64)
    secs = time.perf_counter() - t0
    xg = x.double().cpu().numpy()
    return {"level": 0, "rel_error": float(np.linalg.norm(xg - x64) / np.linalg.norm(x64)),
            "oracle_rel_residual": float(opcg.relative_residual(a, x64, b)), "oracle_iterations": int(it),
            "oracle_tol": tol, "oracle_seconds": round(secs, 1),
            "stored_fp32_solution_rel_residual": float(opcg.relative_residual(a, xg, b))}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_stages(runs=5):
    """CPU baseline by the stated protocol (1 warm-up and `runs` timed runs, median per stage):
    config 1 (100k-point torus, 128^3, 4 levels) through the whole oracle pipeline (numpy stages, the
    C/OpenMP Jacobi-PCG in FP32 to 1e-6 as the reference solves, dual MC), beside the GPU's time for
    the same reconstruction; config 2 (2M points, 512^3) as per-level PCG only: the GPU's assembled
    level systems solved by the oracle's C/OpenMP Jacobi-PCG in FP64 to 1e-8. Prints one JSON line."""
    import numpy as np
    import torch
    import ngspsr
    sys.path.insert(0, ROOT)
    from oracle import assembly, mesh, pcg
    cfg = CONFIGS["c1"]
    xyz, nrm = make_scene(cfg)
    per_run = []
    meta = {}
    for r in range(runs + 1):
        t = {}
        t0 = time.perf_counter()
        h, lv, rho = assembly.build_tree(xyz, cfg["vs"], cfg["depth"], cfg["min_density"])
        t["voxelise_density_levels"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        nd = assembly.splat_normals(h, xyz, nrm, lv, 1.0 / rho)
        t["splat"] = time.perf_counter() - t0
        pcg_s = [0.0]
        iters = {}

        def solve(a, b):
            t1 = time.perf_counter()
            out = pcg.solve_pcg(a, b, tol=1e-6, max_iter=10000, dtype=np.float32)
            pcg_s[0] += time.perf_counter() - t1
            iters[a.shape[0]] = int(out[1])
            return out

        t0 = time.perf_counter()
        res = assembly.solve_multigrid(h, nd, xyz, 4.0, 0.1, solve=solve)
        t["assembly_rhs"] = time.perf_counter() - t0 - pcg_s[0]
        t["pcg"] = pcg_s[0]
        t0 = time.perf_counter()
        v, f, _ = mesh.extract_multiscale_mesh(h, res.solutions, 1.0 / rho, xyz)
        t["mesh"] = time.perf_counter() - t0
        t["total"] = sum(t.values())
        log(f"cpu stages c1 run {r} ({'warm-up' if r == 0 else 'timed'}): " +
            ", ".join(f"{k} {v:.2f} s" for k, v in t.items()))
        if r > 0:
            per_run.append(t)
        meta = {"voxels_by_level": [int(h.num_voxels(d)) for d in range(h.depth)],
                "cpu_pcg_iterations": {int(d): int(i) for d, i in res.iters.items()},
                "cpu_mesh_triangles": int(f.shape[0])}
    med = {k: round(float(np.median([t[k] for t in per_run])), 3) for k in per_run[0]}
    dev = torch.device("cuda:0")
    xt, nt = torch.from_numpy(xyz).to(dev), torch.from_numpy(nrm).to(dev)

    def gpu(c, x, n):
        return ngspsr.reconstruct(x, n, c["vs"], depth=c["depth"], min_density=c["min_density"], screen_alpha=4.0,
                                  solver=c["solver"], return_result=True)
    for _ in range(3):
        gpu(cfg, xt, nt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        g1 = gpu(cfg, xt, nt)
    torch.cuda.synchronize()
    g_s = (time.perf_counter() - t0) / 10
    # config 2: per-level PCG of the GPU's assembled systems, FP64 to 1e-8, on the host
    cfg2 = CONFIGS["c2"]
    x2, n2 = make_scene(cfg2)
    r2 = gpu(cfg2, torch.from_numpy(x2).to(dev), torch.from_numpy(n2).to(dev))
    c2 = {}
    for d, (system, rhs) in sorted(r2.reconstructor.systems.items()):
        a = system.to_scipy()
        b = rhs.double().cpu().numpy()
        ts, its = [], 0
        for r in range(runs + 1):
            t0 = time.perf_counter()
            _, its = pcg.solve_pcg(a, b, tol=1e-8, max_iter=20000, dtype=np.float64)
            if r > 0:
                ts.append(time.perf_counter() - t0)
        c2[int(d)] = {"rows": int(a.shape[0]), "nnz": int(a.nnz), "iterations": int(its),
                      "median_s": round(float(np.median(ts)), 3)}
        log(f"cpu stages c2 level {d}: {c2[int(d)]}")
    c2_total = sum(v["median_s"] for v in c2.values())
    print(json.dumps({
        "protocol": f"1 warm-up + {runs} timed runs, median per stage",
        "cpu_model": _cpu_model(), "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count())),
        "c1": {"workload": cfg["workload"], "cpu_stages_median_s": med, **meta, "gpu_ms": round(1e3 * g_s, 2),
               "gpu_triangles": int(g1.triangles.shape[0]), "cpu_over_gpu": round(med["total"] / g_s, 1),
               "note": "numpy stages single-threaded; the PCG on the OpenMP threads"},
        "c2_pcg_by_level": {"workload": cfg2["workload"], "levels": c2, "cpu_pcg_total_median_s": round(c2_total, 3),
                            "solver": "the oracle's C/OpenMP Jacobi-PCG in FP64 to relative residual 1e-8 on the "
                                      "GPU's assembled level systems (PCG only; assembly and mesh not run on the CPU)"},
    }), flush=True)


COPY_BYTES = 1 << 30


def pmc_probe(res):
    """Kernels for the HBM counter runs: a 1 GiB device copy (its FETCH_SIZE / WRITE_SIZE readings
    calibrate the counters against 1 GiB read + 1 GiB written), then the finest-level SpMV in the
    slot run form, the packed run form and CSR form (the same matrix), 5 launches each."""
    import torch
    from ngspsr._lib import call, ptr, stream
    from ngspsr.solver import RunSystem, StencilSystem
    system = res.reconstructor.systems[0][0]
    csr = system.to_csr() if isinstance(system, (StencilSystem, RunSystem)) else system
    a = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device=csr.device)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    x = torch.randn(csr.M, device=csr.device)
    y = torch.empty_like(x)
    if isinstance(system, RunSystem):
        for _ in range(5):
            call("ngs_spmv_runs", ptr(system.runs), ptr(system.rvals), ptr(x), ptr(y), system.M, stream())
        if system.is_packed:
            pvals, soff = system._packed
            for _ in range(5):
                call("ngs_spmv_packed", ptr(system.runs), ptr(pvals), ptr(soff), ptr(x), ptr(y), system.M, stream())
    for _ in range(5):
        call("ngs_spmv", ptr(csr.rowptr), ptr(csr.cols), ptr(csr.vals), ptr(x), ptr(y), csr.M, stream())
    torch.cuda.synchronize()
    M, nnz = csr.M, csr.nnz()
    algo = {"spmv_kernel": 8 * nnz + 4 * (M + 1) + 8 * M}
    layout = {"spmv_kernel": 8 * csr.vals.shape[0] + 4 * (M + 1) + 8 * M}
    if isinstance(system, RunSystem):
        runs = int((system.counts() > 0).sum().item())
        n = res.tree.xyz.shape[0]
        algo["spmv_runs_kernel"] = algo["spmv_packed_kernel"] = 4 * nnz + 4 * runs + 8 * M
        algo["lhs_assemble_kernel"] = 4 * nnz + 4 * runs + 32 * M + 52 * n
        layout["spmv_runs_kernel"] = 608 * M
        layout["spmv_packed_kernel"] = 100 * M + 4 * nnz + 8 * M + 8 * ((M + 63) // 64 + 1)
    print(json.dumps({"pmc_probe": True, "M": M, "nnz": nnz, "copy_bytes": COPY_BYTES, "algorithmic_bytes": algo,
                      "layout_bytes": layout}), flush=True)


def pmc_summary(probe_json, fetch_csv, write_csv):
    """HBM traffic per launch from two rocprofv3 --pmc runs of --pmc-probe (FETCH_SIZE, then
    WRITE_SIZE; counter_collection CSVs) and the probe's printed line: the 1 GiB copies calibrate
    both counters (bytes = factor x KiB x 1024, factor = copy bytes / the copy's reading); per
    kernel, the readings of its launches (the SpMVs' 5 launches; the finest-level assembly is the
    assembly launch with the largest grid)."""
    import csv
    with open(probe_json) as f:
        probe = json.loads([l for l in f.read().splitlines() if l.startswith("{")][-1])

    def load(path, counter):
        per = {}
        with open(path) as f:
            for r in csv.DictReader(f):
                if r.get("Counter_Name") != counter:
                    continue
                name = r["Kernel_Name"]
                key = next((k for k in ("spmv_packed_kernel", "spmv_runs_kernel", "spmv_kernel", "lhs_assemble_kernel",
                                        "__amd_rocclr_copyBuffer") if k in name), None)
                if key is None or (key == "spmv_kernel" and "bpcg" in name):
                    continue
                dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 if r.get("End_Timestamp") else None
                per.setdefault(key, []).append((int(r.get("Grid_Size") or 0), float(r["Counter_Value"]), dur))
        return per

    fe, wr = load(fetch_csv, "FETCH_SIZE"), load(write_csv, "WRITE_SIZE")
    copy_b = probe["copy_bytes"]
    f_copy = max(v for _, v, _ in fe["__amd_rocclr_copyBuffer"])
    w_copy = max(v for _, v, _ in wr["__amd_rocclr_copyBuffer"])
    ff, wf = copy_b / (f_copy * 1024), copy_b / (w_copy * 1024)
    out = {"calibration": f"device copies of {copy_b} bytes read FETCH_SIZE = {f_copy} KiB and WRITE_SIZE = {w_copy} KiB "
                          f"per dispatch: read bytes = {ff:.3f} x FETCH_SIZE x 1024, written = {wf:.3f} x WRITE_SIZE x 1024",
           "fetch_factor": round(ff, 4), "write_factor": round(wf, 4), "M": probe["M"], "nnz": probe["nnz"],
           "kernels": {}}
    for k in ("spmv_packed_kernel", "spmv_runs_kernel", "spmv_kernel", "lhs_assemble_kernel"):
        if k not in fe or k not in wr:
            continue
        if k == "lhs_assemble_kernel":
            g = max(x[0] for x in fe[k])
            fk, wk = [x for x in fe[k] if x[0] == g][:1], [x for x in wr[k] if x[0] == g][:1]
        else:
            fk, wk = fe[k], wr[k]
        rd = ff * 1024 * sorted(v for _, v, _ in fk)[len(fk) // 2]
        wt = wf * 1024 * sorted(v for _, v, _ in wk)[len(wk) // 2]
        e = {"fetch_size_kib_per_launch": [round(v, 1) for _, v, _ in fk],
             "write_size_kib_per_launch": [round(v, 1) for _, v, _ in wk],
             "hbm_read_bytes_per_launch": int(rd), "hbm_write_bytes_per_launch": int(wt),
             "hbm_bytes_per_launch": int(rd + wt),
             "launch_us_under_counters": [round(d, 1) for _, _, d in fk if d is not None]}
        if k in probe.get("algorithmic_bytes", {}):
            a = probe["algorithmic_bytes"][k]
            e["algorithmic_bytes_per_launch"] = a
            e["traffic_over_algorithmic"] = round((rd + wt) / a, 3)
        if k in probe.get("layout_bytes", {}):
            e["layout_bytes_per_launch"] = probe["layout_bytes"][k]
        out["kernels"][k] = e
    print(json.dumps(out, indent=1))


TRAFFIC_FILES = {"c2": "traffic_ellipsoids2M_512.json", "c5": "traffic_room10M_1024.json"}


def measured_traffic(kernel, name="c2"):
    """HBM bytes per launch of a kernel ("spmv_kernel", "spmv_runs_kernel", or "lhs_assemble_kernel" for
    the finest-level assembly) at a single-solve workload, from the committed counter summary
    (profiles/), if present."""
    if name not in TRAFFIC_FILES:
        return None
    path = os.path.join(ROOT, "profiles", TRAFFIC_FILES[name])
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return t.get("kernels", {}).get(kernel, {}).get("hbm_bytes_per_launch")


def _serving_scene(i, seed0=1000):
    import numpy as np
    from ngspsr import synthetic
    rng = np.random.default_rng(seed0 + i)
    n = synthetic.log_uniform_int(rng, 1e4, 1e6)
    depth = int(rng.integers(7, 10))
    sigma = float(rng.uniform(0, 0.005))
    xyz, nrm = synthetic.ellipsoids(n, k=16, sigma=sigma, kappa=200.0, seed=seed0 + i)
    return xyz, nrm, 2.0 ** -depth


def serving_scenes(n_req, seed0=1000):
    """Requests of the serving stream: random-ellipsoid scenes (16 ellipsoids, kappa 200), points
    log-uniform in [1e4, 1e6], depth uniform in {7, 8, 9} (voxel size 2^-depth), sigma uniform in
    [0, 0.005], seeds seed0 + i. Generated on up to 16 host processes."""
    import concurrent.futures as cf
    import multiprocessing as mp
    workers = max(1, min(16, os.cpu_count() or 1))
    with cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(_serving_scene, range(n_req), [seed0] * n_req, chunksize=4))


def run_serving(n_req=256, steps=20, warmup=5, workers=8, chunk=8, log=None):
    """Config 3: the same stream of n_req requests served plain and batched, at relative residual
    1e-6.
    - plain: one request at a time, the reference's path (`reconstruct`, FP32 Jacobi-PCG to 1e-6);
      one pass, all requests queued at t = 0: its capacity (solves/s) and service latencies.
    - batched: the continuous-batching server with mixed precision (FP32 inner PCG in the shared
      batched pool, FP64 residual check and refinement to 1e-6 per level); `warmup` untimed passes,
      then `steps` timed passes of the whole stream, all requests queued at t = 0 (a step is one pass).
    - latency: Poisson arrivals at 0.8x the plain capacity, served by the batched server and by the
      plain path (first come, first served, one at a time): p50 / p99 of arrival-to-result time.
    Returns (result dict, timed seconds)."""
    import numpy as np
    import torch
    import ngspsr
    from ngspsr import serving
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    scenes = serving_scenes(n_req)
    gen_s = time.perf_counter() - t0
    data = [(torch.from_numpy(x).to(dev), torch.from_numpy(n).to(dev), vs) for x, n, vs in scenes]
    del scenes
    pts = int(sum(x.shape[0] for x, _, _ in data))
    say = log or (lambda m: None)
    say(f"serving: {n_req} requests, {pts} points, generated in {gen_s:.1f} s")

    def plain_one(x, n, vs):
        res = ngspsr.reconstruct(x, n, vs, depth=4, min_density=1.0, solver="pcg@tol=1e-6@maxiter=20000",
                                 return_result=True)
        torch.cuda.synchronize()
        return [res.tree.get_coords_size(None, d) for d in range(4)]

    for x, n, vs in data[:4]:
        plain_one(x, n, vs)
    lat = []
    peaks = []  # (peak bytes allocated by the request, its footprint estimate)
    tp = time.perf_counter()
    for x, n, vs in data:
        t1 = time.perf_counter()
        torch.cuda.reset_peak_memory_stats(dev)
        base = torch.cuda.memory_allocated(dev)
        sizes = plain_one(x, n, vs)
        lat.append(time.perf_counter() - t1)
        peaks.append((torch.cuda.max_memory_allocated(dev) - base, serving.level_footprint(sizes, x.shape[0])))
    plain_s = time.perf_counter() - tp
    ratio = max(p / e for p, e in peaks)
    say(f"serving: largest request peak {max(p for p, _ in peaks) / 2 ** 30:.2f} GiB allocated; "
        f"peak / level footprint at most {ratio:.2f}")
    plain_rate = n_req / plain_s
    plain_peak = torch.cuda.max_memory_reserved(dev)
    say(f"serving: plain {plain_rate:.2f} solves/s, peak HBM reserved {plain_peak / 2 ** 30:.1f} GiB")
    # the plain path's cached blocks live on the default stream, which the server does not use
    torch.cuda.empty_cache()
    torch.cuda.reset_peak_memory_stats(dev)

    srv = serving.ReconstructionServer(dev, workers=workers, tol=1e-6, refine=True, chunk=chunk)

    def hand_off(req):
        # a client takes the mesh as each request finishes; the bench keeps only its size
        if req.result is not None:
            req.mesh_size = (int(req.result[0].shape[0]), int(req.result[1].shape[0]))
            req.result = None

    def one_pass(arrivals=None):
        reqs = [serving.Request(i, x, n, vs, arrival=0.0 if arrivals is None else float(arrivals[i]))
                for i, (x, n, vs) in enumerate(data)]
        ts = time.perf_counter()
        srv.serve(reqs, t0=ts, on_complete=hand_off)
        torch.cuda.synchronize()
        errs = [r for r in reqs if r.error is not None]
        if errs:
            srv.close()
            raise errs[0].error
        return reqs, time.perf_counter() - ts

    for w in range(warmup):
        one_pass()
        say(f"serving: batched warm-up pass {w + 1}/{warmup}, peak HBM reserved "
            f"{torch.cuda.max_memory_reserved(dev) / 2 ** 30:.1f} GiB")
    torch.cuda.synchronize()
    tb = time.perf_counter()
    res_max = 0.0
    for k in range(steps):
        reqs, _ = one_pass()
        res_max = max(res_max, max(r.max_rel_residual for r in reqs))
        say(f"serving: batched pass {k + 1}/{steps}")
    torch.cuda.synchronize()
    timed_s = time.perf_counter() - tb
    batched_rate = steps * n_req / timed_s
    # latency under Poisson arrivals at 0.8x the plain capacity
    lam = 0.8 * plain_rate
    arr = np.cumsum(np.random.default_rng(7).exponential(1.0 / lam, n_req))
    reqs, _ = one_pass(arr)
    lat_b = serving.latency_stats(reqs)
    pool = srv.pool
    srv.close()
    batched_peak = torch.cuda.max_memory_reserved(dev)
    torch.cuda.empty_cache()
    # the plain path under the same arrivals: first come, first served
    tq = time.perf_counter()
    done = []
    for (x, n, vs), a in zip(data, arr):
        now = time.perf_counter() - tq
        if now < a:
            time.sleep(a - now)
        plain_one(x, n, vs)
        done.append(time.perf_counter() - tq - a)
    lat_p = np.array(done)
    say("serving: latency passes done")
    out = {
        "workload": "serving stream: 256 requests, ellipsoid scenes 1e4-1e6 points (log-uniform), depth 7-9, "
                    "sigma 0-0.005, 4 levels, relative residual 1e-6",
        "requests": n_req, "points_total": pts, "passes_timed": steps, "passes_warmup": warmup,
        "batched_mixed_solves_per_s": round(batched_rate, 3),
        "batched_config": {"workers": workers, "pool_chunk": chunk, "precision": "FP32 inner PCG, FP64 residual "
                           "check and refinement to 1e-6"},
        "batched_max_fp64_rel_residual": res_max,
        "plain_fp32_solves_per_s": round(plain_rate, 3),
        "plain_service_latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 2),
                                     "p99": round(1e3 * float(np.percentile(lat, 99)), 2)},
        "north_star_ratio_batched_mixed_vs_plain_fp32": round(batched_rate / plain_rate, 3),
        "poisson_rate_per_s": round(lam, 3), "poisson_rate_basis": "0.8 x plain FP32 capacity",
        "poisson_latency_ms_batched": {"p50": round(lat_b["p50_ms"], 2), "p99": round(lat_b["p99_ms"], 2)},
        "poisson_latency_ms_plain": {"p50": round(1e3 * float(np.percentile(lat_p, 50)), 2),
                                     "p99": round(1e3 * float(np.percentile(lat_p, 99)), 2)},
        "pool_launches": pool.launches,
        "hbm_peak_reserved_gib": {"plain": round(plain_peak / 2 ** 30, 2), "batched": round(batched_peak / 2 ** 30, 2)},
        "request_peak_allocated_gib_max": round(max(p for p, _ in peaks) / 2 ** 30, 3),
        "request_peak_over_level_footprint_max": round(ratio, 3),
    }
    return out, timed_s


def extensions(res, xyz):
    """Timings of the field-query API, the primal marching cubes and the DDI kernel on the
    benchmark's own solution, outside the timed reconstruction loop (HIP events)."""
    import torch
    from ngspsr import additive
    from ngspsr._lib import call, ptr, stream
    rec, tree = res.reconstructor, res.tree
    dev = xyz.device
    gen = torch.Generator(device=dev).manual_seed(0)
    lo, hi = xyz.min(0).values, xyz.max(0).values
    nq = 8_000_000
    q = lo + (hi - lo) * torch.rand((nq, 3), device=dev, generator=gen)
    levels = len(rec.solutions)
    t_val = _event_time(lambda: rec.evaluate_raw_chi(q), 5)
    t_grad = _event_time(lambda: rec.evaluate_raw_chi(q, compute_grad=True), 5)
    out = {"field_query": {"queries": nq, "levels": levels, "chi_ms": round(1e3 * t_val, 3),
                           "chi_Mqueries_per_s": round(nq / t_val / 1e6, 1), "gradient_ms": round(1e3 * t_grad, 3),
                           "gradient_Mqueries_per_s": round(nq / t_grad / 1e6, 1)}}
    del q
    res_mc, depth_mc = 4, 1
    t_grid = _event_time(lambda: tree.get_test_grid(res_mc, depth_mc), 3)
    qg, base = tree.get_test_grid(res_mc, depth_mc)
    t_chi = _event_time(lambda: rec.evaluate_chi(qg, max_points=2 ** 24), 3)
    chi = rec.evaluate_chi(qg, max_points=2 ** 24)
    t_mc = _event_time(lambda: rec.extract_mesh(base, chi, depth_mc, build_o3d_mesh=False), 3)
    v, f, _ = rec.extract_mesh(base, chi, depth_mc, build_o3d_mesh=False)
    out["primal_mesh"] = {"depth": depth_mc, "resolution": res_mc, "leaves": int(base.shape[0]),
                          "samples": int(chi.numel()), "test_grid_ms": round(1e3 * t_grid, 3),
                          "evaluate_chi_ms": round(1e3 * t_chi, 3), "extract_mesh_ms": round(1e3 * t_mc, 3),
                          "vertices": int(v.shape[0]), "triangles": int(f.shape[0])}
    del qg, chi, v, f
    basis = additive.AdditiveBasis.get_poly_6()
    npairs, nc = 4_000_000, len(basis.components)
    s = torch.randn((npairs, 3 * nc), device=dev, generator=gen)
    t = torch.randn((npairs, 3 * nc), device=dev, generator=gen)
    rel = torch.randint(0, 5, (npairs, 3), device=dev, generator=gen)
    ff = basis.ff_integral[0]["f_f"].data.to(dev)
    dd = basis.ff_integral[0]["df_df"].data.to(dev)
    pm = torch.empty((npairs, 3), device=dev)
    qm = torch.empty((npairs, 3), device=dev)
    t_ddi = _event_time(lambda: call("ngs_ddi_forward", ptr(s), ptr(t), ptr(rel), npairs, nc, ptr(ff), ptr(dd), 5,
                                     ptr(pm), ptr(qm), stream()), 10)
    bytes_pair = 2 * 12 * nc + 24 + 24  # s and t rows, rel, p and q
    out["ddi"] = {"kernel": "ddi_forward_kernel", "pairs": npairs, "components": nc, "launch_us": round(1e6 * t_ddi, 2),
                  "Gpairs_per_s": round(npairs / t_ddi / 1e9, 3),
                  "GB/s": round(npairs * bytes_pair / t_ddi / 1e9, 1), "algorithmic_bytes_per_pair": bytes_pair}
    return out


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def single_solve(name, steps, warmup, cpu=True, ext=True, fp64=True):
    """One workload solved end to end `steps` times (after `warmup`), points in HBM: the record of a
    single-solve configuration (c1, c2, c5), with the finest-level SpMV roofline, the assembly, the
    FP64 error, host-I/O latency and, with cpu, the oracle's C/OpenMP PCG on the finest level."""
    import numpy as np
    import torch
    import ngspsr
    cfg = CONFIGS[name]
    dev = torch.device("cuda:0")
    xyz_np, nrm_np = make_scene(cfg)
    xyz = torch.from_numpy(xyz_np).to(dev)
    nrm = torch.from_numpy(nrm_np).to(dev)

    def step():
        return ngspsr.reconstruct(xyz, nrm, cfg["vs"], depth=cfg["depth"], min_density=cfg["min_density"],
                                  screen_alpha=4.0, solver=cfg["solver"], return_result=True)

    for _ in range(warmup):
        res = step()
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats(dev)
    lat = []
    t_start = time.perf_counter()
    for _ in range(steps):
        t0 = time.perf_counter()
        res = step()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
    total = time.perf_counter() - t_start
    peak = torch.cuda.max_memory_allocated(dev)
    lat = np.array(lat)
    log(f"{name}: {steps / total:.3f} reconstructions/s")
    # numpy in, numpy out: the H2D copy of the points and the D2H copy of the mesh included
    io = []
    for _ in range(3):
        t0 = time.perf_counter()
        ngspsr.reconstruct(xyz_np, nrm_np, cfg["vs"], depth=cfg["depth"], min_density=cfg["min_density"],
                           screen_alpha=4.0, solver=cfg["solver"])
        io.append(time.perf_counter() - t0)

    info = res.solve_info
    total_iters = int(sum(i.iters for i in info.values()))
    from ngspsr.solver import RunSystem, StencilSystem
    # the FP64 residual of every level's returned solution, recomputed from the assembled system
    rel_ret = {}
    for d, (s_, b_) in sorted(res.reconstructor.systems.items()):
        _, rr, bb = s_.residual_f64(b_.double(), res.reconstructor.solutions[d].double())
        rel_ret[int(d)] = (rr / bb) ** 0.5 if bb > 0 else 0.0
    system, rhs = res.reconstructor.systems[0]
    forms = {int(d): ("stencil" if isinstance(s_, StencilSystem) else "runs" if isinstance(s_, RunSystem) else "csr")
             for d, (s_, _) in sorted(res.reconstructor.systems.items())}
    csr = system.to_csr() if isinstance(system, (StencilSystem, RunSystem)) else system
    t_spmv, algo, nnz, _ = spmv_roofline(csr)
    t_iter = pcg_iteration_rate(system)
    spmv_obj = {"kernel": "spmv_kernel<8,0,1> (finest-level CSR SpMV, int32 columns, the solve_pcg_diag layout)",
                "bound": "hbm", "achieved": round(algo / t_spmv / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(algo / t_spmv / 1e9 / HBM_PEAK_GBS, 4), "traffic": measured_traffic("spmv_kernel", name),
                "launch_us": round(1e6 * t_spmv, 2), "algorithmic_bytes": int(algo)}
    if isinstance(system, StencilSystem):
        t_pts, b_pts, t_rows, b_rows = stencil_roofline(system)
        iter_bytes = b_pts + b_rows + 44 * system.M
        roof = {"kernel": "stencil_rows_kernel (finest-level row pass of the stencil operator)", "bound": "hbm",
                "achieved": round(b_rows / t_rows / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(b_rows / t_rows / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                "launch_us": round(1e6 * t_rows, 2), "algorithmic_bytes": int(b_rows),
                "point_pass": {"kernel": "stencil_points_kernel", "launch_us": round(1e6 * t_pts, 2),
                               "algorithmic_bytes": int(b_pts), "achieved": round(b_pts / t_pts / 1e9, 1)},
                "operator_us": round(1e6 * (t_pts + t_rows), 2),
                "csr_spmv_us_same_system": round(1e6 * t_spmv, 2)}
    elif isinstance(system, RunSystem):
        t_run, algo_run, _, extra = spmv_roofline(system)
        extra["stencil_form_same_level"] = stencil_probe(res, xyz)
        iter_bytes = algo_run + 44 * system.M
        roof = {"kernel": ("spmv_packed_kernel (finest-level SpMV, packed run form: one run word per bitmap "
                           "column, the values of 64-row slices packed and staged through LDS; the PCG's "
                           "bpcg_packed_dot_kernel runs the same row loop with the p.q partials at 5 waves per SIMD "
                           "instead of 6, about 8% slower per launch)") if system.is_packed
                else "spmv_runs_kernel (finest-level SpMV, run form: one run word per bitmap column, values "
                     "plane-major)", "bound": "hbm", "achieved": round(algo_run / t_run / 1e9, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(algo_run / t_run / 1e9 / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic("spmv_packed_kernel" if system.is_packed else "spmv_runs_kernel", name),
                "launch_us": round(1e6 * t_run, 2),
                "algorithmic_bytes": int(algo_run), **extra, "csr_spmv_us_same_matrix": round(1e6 * t_spmv, 2)}
    else:
        iter_bytes = algo + 44 * system.M
        roof = dict(spmv_obj)
    roof["workload"] = cfg["workload"]
    out = {
        "value": round(steps / total, 4), "unit": "reconstructions/s", "steps": steps, "warmup": warmup,
        "ms_per_step": round(1e3 * total / steps, 3),
        "config": {"workload": cfg["workload"], "points": cfg["n"], "voxel_size": cfg["vs"], "levels": cfg["depth"],
                   "solver": cfg["solver"], "finest_voxels": int(system.M), "finest_nnz": int(nnz),
                   "voxels_by_level": [int(res.tree.get_coords_size(None, d)) for d in range(cfg["depth"])],
                   "operator_forms": forms},
        "timed_region": "points already in HBM; the mesh stays on the device (no host copies)",
        "latency_ms": {"p50": round(1e3 * float(np.percentile(lat, 50)), 3),
                       "p99": round(1e3 * float(np.percentile(lat, 99)), 3)},
        "host_io_latency_ms": {"median": round(1e3 * float(np.median(io)), 3),
                               "note": "numpy points in, numpy mesh out: host-to-device copy of the points and "
                                       "device-to-host copy of the mesh included"},
        "peak_hbm_GiB": round(peak / 2 ** 30, 3),
        "max_level_rel_residual": max(rel_ret.values()),
        "rel_residual_by_level": rel_ret,
        "residual_note": "FP64 relative residual of the solution each level returns (FP64 for refine, FP32 for pcg)",
        "pcg_iterations_total": total_iters,
        "pcg_iterations_by_level": {int(d): int(i.iters) for d, i in sorted(info.items())},
        "pcg_iteration": {"ms": round(1e3 * t_iter, 4), "algorithmic_bytes": int(iter_bytes),
                          "GB/s": round(iter_bytes / t_iter / 1e9, 1)},
        "roofline": roof,
        "csr_spmv": spmv_obj,
        "assembly": assembly_roofline(res, xyz, name),
        "mesh": {"vertices": int(res.vertices.shape[0]), "triangles": int(res.triangles.shape[0])},
    }
    if ext:
        out["extensions"] = extensions(res, xyz)
    if fp64:
        out["solution_error_vs_fp64_oracle"] = fp64_error(csr, rhs, res.reconstructor.solutions[0])
    if cpu:
        per_iter, it = cpu_baseline(csr, rhs)
        out["cpu_pcg_finest"] = {
            "ms_per_iteration": round(per_iter * 1e3, 3), "iterations_timed": it,
            "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count())),
            "note": "the oracle's C/OpenMP FP32 Jacobi-PCG (FP64 dots) on the finest-level system"}
    return out


def cpu_serving_sample(budget_s=30.0):
    """CPU baseline of the serving stream: the oracle pipeline (numpy stages and the C/OpenMP
    Jacobi-PCG in FP32 to 1e-6, the reference's plain method) end to end on the stream's smallest
    requests, smallest first, until budget_s is spent (at least one). Returns (record, the indices
    of the requests run)."""
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import reconstruct as orec
    sizes = []
    for i in range(256):
        rng = np.random.default_rng(1000 + i)
        from ngspsr import synthetic
        n = synthetic.log_uniform_int(rng, 1e4, 1e6)
        depth = int(rng.integers(7, 10))
        sizes.append((n * 2 ** (depth - 7), i, n, depth))
    sizes.sort()
    done, t_all = [], 0.0
    for _, i, n, depth in sizes:
        xyz, nrm, vs = _serving_scene(i)
        t0 = time.perf_counter()
        orec.reconstruct(xyz, nrm, vs, depth=4, min_density=1.0, screen_alpha=4.0, tol=1e-6, dtype=np.float32)
        t_all += time.perf_counter() - t0
        done.append((i, n, depth))
        log(f"cpu sample: request {i} ({n} points, depth {depth}) done, {t_all:.1f} s so far")
        if t_all >= budget_s:
            break
    rec = {"value": round(len(done) / t_all, 5), "unit": "solves/s", "cores": int(os.environ.get("OMP_NUM_THREADS",
                                                                                                   os.cpu_count())),
           "kind": "port",
           "sample": f"the oracle pipeline (numpy stages, C/OpenMP FP32 Jacobi-PCG to 1e-6, dual MC) end to end on the "
                     f"{len(done)} smallest requests of the stream ({', '.join(f'{n} points at depth {d}' for _, n, d in done)}), "
                     f"one run each, {t_all:.1f} s; the stream's requests average 8x more points than these",
           "note": "numpy stages run on one core; the PCG uses the OpenMP threads"}
    return rec, [i for i, _, _ in done]


def plain_pass(n_req=256):
    """The plain path once over the serving stream (after 4 warm-up requests), for a kernel trace:
    prints the wall time of the pass. Run under rocprofv3 --kernel-trace; --trace-busy summarises."""
    import torch
    import ngspsr
    dev = torch.device("cuda:0")
    data = [(torch.from_numpy(x).to(dev), torch.from_numpy(n).to(dev), vs) for x, n, vs in serving_scenes(n_req)]
    for x, n, vs in data[:4]:
        ngspsr.reconstruct(x, n, vs, depth=4, min_density=1.0, solver="pcg@tol=1e-6@maxiter=20000")
    # a short spin kernel on each side of the pass marks its window in a kernel trace
    torch.cuda._sleep(1000)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t0_ns = time.monotonic_ns()
    for x, n, vs in data:
        ngspsr.reconstruct(x, n, vs, depth=4, min_density=1.0, solver="pcg@tol=1e-6@maxiter=20000")
        torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    torch.cuda._sleep(1000)
    torch.cuda.synchronize()
    print(json.dumps({"plain_pass": True, "requests": n_req, "wall_s": round(wall, 4),
                      "solves_per_s": round(n_req / wall, 3), "start_monotonic_ns": t0_ns,
                      "end_monotonic_ns": time.monotonic_ns()}), flush=True)


def batched_pass(n_req=256, workers=8, chunk=8, warmup=2):
    """The batched server once over the serving stream (all requests queued at t = 0, after `warmup`
    passes), for a kernel trace: prints the wall time of the pass. Run under rocprofv3 --kernel-trace;
    --trace-busy summarises the window between the two spin-kernel markers. Under rocprofv3 7.2's
    kernel trace this multi-threaded pass has ended in a host-side SIGSEGV inside the profiler's
    launch hooks in two of three runs (in a request thread's rocPRIM launch, then in the pool's graph
    launch); the unprofiled server runs the same launches in every bench pass."""
    import torch
    from ngspsr import serving
    dev = torch.device("cuda:0")
    data = [(torch.from_numpy(x).to(dev), torch.from_numpy(n).to(dev), vs) for x, n, vs in serving_scenes(n_req)]
    srv = serving.ReconstructionServer(dev, workers=workers, tol=1e-6, refine=True, chunk=chunk)

    def one_pass():
        reqs = [serving.Request(i, x, n, vs) for i, (x, n, vs) in enumerate(data)]
        # as in the timed passes: the mesh is handed off (dropped) as each request finishes
        srv.serve(reqs, t0=time.perf_counter(), on_complete=lambda r: setattr(r, "result", None))
        torch.cuda.synchronize()
        errs = [r for r in reqs if r.error is not None]
        if errs:
            srv.close()
            raise errs[0].error

    for _ in range(warmup):
        one_pass()
    torch.cuda._sleep(1000)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t0_ns = time.monotonic_ns()
    one_pass()
    wall = time.perf_counter() - t0
    torch.cuda._sleep(1000)
    torch.cuda.synchronize()
    launches = srv.pool.launches
    srv.close()
    print(json.dumps({"batched_pass": True, "requests": n_req, "workers": workers, "pool_chunk": chunk,
                      "wall_s": round(wall, 4), "solves_per_s": round(n_req / wall, 3), "pool_launches": launches,
                      "start_monotonic_ns": t0_ns, "end_monotonic_ns": time.monotonic_ns()}), flush=True)


def trace_busy(path, skip_first=0, marker="spin_kernel"):
    """GPU-busy summary of a rocprofv3 kernel trace (CSV): the union of kernel execution intervals
    over the span from the first to the last kernel, the sum of kernel durations, and the largest
    kernels by total time. The window is the kernels between the first and the last `marker` kernel
    when the trace holds two (the plain pass brackets itself with them), else all kernels after the
    first `skip_first`."""
    import csv
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    rows.sort()
    marks = [k for k, r in enumerate(rows) if marker and marker in r[2]]
    if len(marks) >= 2:
        rows = rows[marks[0] + 1:marks[-1]]
    else:
        rows = rows[skip_first:]
    busy, cur_s, cur_e, tot = 0, None, None, 0
    per = {}
    for s_, e_, name in rows:
        tot += e_ - s_
        key = name.split("(")[0]
        per[key] = per.get(key, 0) + (e_ - s_)
        if cur_e is None or s_ > cur_e:
            if cur_e is not None:
                busy += cur_e - cur_s
            cur_s, cur_e = s_, e_
        else:
            cur_e = max(cur_e, e_)
    if cur_e is not None:
        busy += cur_e - cur_s
    span = rows[-1][1] - rows[0][0] if rows else 0
    top = sorted(per.items(), key=lambda kv: -kv[1])[:15]
    print(json.dumps({"trace": os.path.basename(path), "window": "between markers" if len(marks) >= 2 else
                      f"after {skip_first} kernels", "kernels": len(rows), "span_s": span / 1e9,
                      "busy_union_s": busy / 1e9, "busy_fraction": busy / span if span else 0.0,
                      "kernel_time_sum_s": tot / 1e9,
                      "top_kernels_s": {k: round(v / 1e9, 4) for k, v in top}}), flush=True)


def trace_runs(name, runs=3, warmup=2):
    """Reconstructions of a single-solve workload for a kernel trace: `warmup` untimed ones, then
    `runs` with a spin_kernel (torch.cuda._sleep) before and after each, so --trace-stages can cut
    the trace into reconstructions."""
    import torch
    import ngspsr
    cfg = CONFIGS[name]
    dev = torch.device("cuda:0")
    xyz_np, nrm_np = make_scene(cfg)
    xyz, nrm = torch.from_numpy(xyz_np).to(dev), torch.from_numpy(nrm_np).to(dev)
    for i in range(warmup + runs):
        if i >= warmup:
            torch.cuda._sleep(20000)
            torch.cuda.synchronize()
        t0 = time.perf_counter()
        ngspsr.reconstruct(xyz, nrm, cfg["vs"], depth=cfg["depth"], min_density=cfg["min_density"], screen_alpha=4.0,
                           solver=cfg["solver"])
        torch.cuda.synchronize()
        log(f"reconstruction {i}: {1e3 * (time.perf_counter() - t0):.2f} ms")
    torch.cuda._sleep(20000)
    torch.cuda.synchronize()


TRACE_STAGES = [("PCG SpMV with p.q (bpcg_*_dot)", r"bpcg_.*dot"), ("PCG update and direction", r"bpcg_(update|direction)"),
                ("PCG set-up", r"pcg_init|fill_map|write_entry"), ("FP64 residual and norms", r"residual_f64|norms"),
                ("assembly", r"lhs_assemble|lhs_count|runs_pack|runs_slice"),
                ("RHS and cross-level stiffness", r"rhs_normals|cross_stiff"),
                ("field at the points, point bins", r"eval_field|gather_f32|bin_points|point_basis|cell_starts"),
                ("dual mesh", r"dual_|mc_count|mc_emit|remap_slots|dedup|run_heads|eval_leaf|mark_dual|complete_flags|conformal|"
                              r"unrefined|compact_|leaf_corner|weighted_sum"),
                ("voxels, levels, splats", r"stage_corners|fold_marks|emit_kernel|grid_rank|levels_kernel|splat|bounds|"
                                           r"density|mark_coords|clear_words|stage_coords"),
                ("sorts and scans (rocPRIM)", r"rocprim"), ("torch operations and copies", r"at::|rocclr")]


def trace_stages(path):
    """Kernel time of each reconstruction of a --trace-runs kernel trace (CSV), by stage: the windows
    between consecutive spin_kernel markers, with span, kernel time and idle time, and the largest
    kernels of the last window."""
    import csv
    import re
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    rows.sort()
    marks = [i for i, r in enumerate(rows) if "spin_kernel" in r[2]]
    if len(marks) < 2:
        print("no spin_kernel markers: record the trace with bench.py --trace-runs")
        return
    per = {}
    for a, b in zip(marks, marks[1:]):
        w = rows[a + 1:b]
        if not w:
            continue
        span = w[-1][1] - w[0][0]
        busy, cs, ce = 0, None, None
        for s_, e_, _ in w:
            if ce is None or s_ > ce:
                if ce is not None:
                    busy += ce - cs
                cs, ce = s_, e_
            else:
                ce = max(ce, e_)
        busy += ce - cs
        tot, per = {}, {}
        for s_, e_, name in w:
            per[name.split("(")[0][:80]] = per.get(name.split("(")[0][:80], 0) + (e_ - s_)
            lab = next((lb for lb, pat in TRACE_STAGES if re.search(pat, name)), "other")
            tot[lab] = tot.get(lab, 0) + (e_ - s_)
        print(f"reconstruction: {len(w)} kernels, span {span / 1e6:.2f} ms, kernel time "
              f"{sum(e - s for s, e, _ in w) / 1e6:.2f} ms, idle between kernels {(span - busy) / 1e6:.2f} ms")
        for lab in [lb for lb, _ in TRACE_STAGES] + ["other"]:
            print(f"  {lab:34s} {tot.get(lab, 0) / 1e6:7.2f} ms")
    print("largest kernels of the last reconstruction:")
    for k, v in sorted(per.items(), key=lambda kv: -kv[1])[:25]:
        print(f"  {v / 1e6:7.3f} ms  {k}")


def counter_summary(path, kernels=("lhs_assemble", "bpcg_packed_dot", "spmv_packed", "residual_f64_packed")):
    """Per (kernel, grid size) averages of the SQ counters of a rocprofv3 --pmc CSV, per wave of 64
    lanes, with the median dispatch duration and the VALU issue time per SIMD they imply (summed
    instruction counts over the 1024 SIMDs, 4 cycles per wave64 VALU instruction, 2.4 GHz). Dispatches of
    one grid size are split at every duration gap above a factor 1.5 (levels that share the PCG's block
    cap, launches of a converged system)."""
    import csv
    groups = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            name = r["Kernel_Name"]
            if not any(k in name for k in kernels):
                continue
            key = (name.split("(")[0], int(r["Grid_Size"]))
            d = groups.setdefault(key, {}).setdefault(r["Dispatch_Id"], {"dur": 0, "c": {}})
            d["dur"] = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
            d["c"][r["Counter_Name"]] = float(r["Counter_Value"])
    print("| kernel | grid | dispatches | median us | VALU/wave | SALU/wave | VMEM rd/wave | VMEM wr/wave | LDS/wave | "
          "wait/wave cycles | VALU issue us/SIMD |")
    print("|---|---|---|---|---|---|---|---|---|---|---|")
    for (name, grid), ds in sorted(groups.items(), key=lambda kv: (kv[0][0], -kv[0][1])):
        ds = sorted(ds.values(), key=lambda d: d["dur"])
        # split at every duration gap above a factor 1.5 (levels sharing the grid, no-op launches)
        parts, cur = [], [ds[0]]
        for a, b in zip(ds, ds[1:]):
            if b["dur"] / max(a["dur"], 1) > 1.5:
                parts.append(cur)
                cur = []
            cur.append(b)
        parts.append(cur)
        for part in reversed(parts):
            c = {}
            for d in part:
                for k_, v in d["c"].items():
                    c[k_] = c.get(k_, 0.0) + v / len(part)
            waves = grid / 64
            med = part[len(part) // 2]["dur"] / 1e3
            per = lambda k_: f"{c[k_] / waves:.0f}" if k_ in c else "-"
            wait = f"{c['SQ_WAIT_ANY'] / c['SQ_WAVE_CYCLES']:.2f}" if c.get("SQ_WAVE_CYCLES") and "SQ_WAIT_ANY" in c else "-"
            issue = f"{c['SQ_INSTS_VALU'] / 1024 * 4 / 2.4e3:.0f}" if "SQ_INSTS_VALU" in c else "-"
            print(f"| `{name}` | {grid} | {len(part)} | {med:.1f} | {per('SQ_INSTS_VALU')} | {per('SQ_INSTS_SALU')} | "
                  f"{per('SQ_INSTS_VMEM_RD')} | {per('SQ_INSTS_VMEM_WR')} | {per('SQ_INSTS_LDS')} | {wait} | {issue} |")


def trace_levels(path, kernels=("bpcg_packed_dot_kernel", "spmv_packed_kernel", "residual_f64_packed_kernel",
                                  "bpcg_runs_dot_kernel", "spmv_runs_kernel", "residual_f64_runs_kernel",
                                  "bpcg_update_kernel", "bpcg_direction_kernel", "lhs_assemble_kernel",
                                  "runs_pack_kernel", "stencil_rows_kernel", "stencil_points_kernel", "bpcg_stencil")):
    """Per-launch-shape summary of a rocprofv3 kernel trace (CSV): for each kernel whose name holds
    one of `kernels`, launches grouped by grid size (one group per level system, since the grid
    follows the row count), with count, total and average duration. Prints a markdown table."""
    import csv
    groups = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            name = r["Kernel_Name"]
            hit = next((k for k in kernels if k in name), None)
            if hit is None:
                continue
            grid = r.get("Grid_Size") or r.get("Grid_Size_X") or "?"
            key = (hit, name.split("(")[0] if "<" in name else hit, grid)
            groups.setdefault(key, []).append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    print("| kernel | grid size | launches | total ms | average us | working launches | their average us | "
          "min us | max us |")
    print("|---|---|---|---|---|---|---|---|---|")
    for (hit, full, grid), ds in sorted(groups.items(), key=lambda kv: (kv[0][0], -sum(kv[1]))):
        # launches of a converged system return at entry: "working" launches take >= 10% of the longest
        hi = max(ds)
        work = sorted(d for d in ds if d >= 0.1 * hi)
        # systems of different sizes can share a grid (the block count is capped): split the working
        # launches at their largest duration gap when it is a factor above 1.5, one row per cluster
        parts = [work]
        if len(work) > 1:
            k = max(range(1, len(work)), key=lambda j: work[j] / max(work[j - 1], 1))
            if work[k] / max(work[k - 1], 1) > 1.5:
                parts = [work[k:], work[:k]]
        idle = [d for d in ds if d < 0.1 * hi]
        for c, part in enumerate(parts):
            rows = part + (idle if c == len(parts) - 1 else [])
            tot = sum(rows)
            print(f"| `{full}`{' (longer)' if len(parts) > 1 and c == 0 else ' (shorter)' if len(parts) > 1 else ''} | "
                  f"{grid} | {len(rows)} | {tot / 1e6:.3f} | {tot / len(rows) / 1e3:.2f} | {len(part)} | "
                  f"{sum(part) / len(part) / 1e3:.2f} | {min(rows) / 1e3:.2f} | {max(rows) / 1e3:.2f} |")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed steps: passes over the serving stream (c3) or "
                                                          "reconstructions (c1, c2, c5)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS) + ["c3"])
    ap.add_argument("--requests", type=int, default=256, help="requests of the serving stream (c3)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extensions", action="store_true",
                    help="skip the field-query, primal marching cubes and DDI timings")
    ap.add_argument("--cpu-stages", action="store_true",
                    help="per-stage CPU timing of the oracle pipeline (100k-point torus in full, 2M-point ellipsoids "
                         "per-level PCG), 1 warm-up and 5 timed runs (minutes)")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="counter probe: one reconstruction, then 3 device copies of 1 GiB (calibration) and 5 "
                         "launches of the finest-level SpMV; run under rocprofv3 --pmc")
    ap.add_argument("--probe-config", default="c2", choices=sorted(CONFIGS), help="workload of --pmc-probe")
    ap.add_argument("--plain-pass", action="store_true",
                    help="the plain path once over the serving stream, for a kernel trace (rocprofv3)")
    ap.add_argument("--batched-pass", action="store_true",
                    help="one traced pass of the batched server over the serving stream (for --trace-busy)")
    ap.add_argument("--trace-busy", default=None, help="summarise a rocprofv3 kernel-trace CSV (no GPU)")
    ap.add_argument("--trace-skip", type=int, default=0, help="kernels to skip at the start of --trace-busy")
    ap.add_argument("--pmc-summary", nargs=3, default=None, metavar=("PROBE_OUT", "FETCH_CSV", "WRITE_CSV"),
                    help="traffic per launch from the --pmc-probe line and its FETCH_SIZE and WRITE_SIZE counter CSVs "
                         "(no GPU)")
    ap.add_argument("--trace-runs", action="store_true",
                    help="reconstructions of --config (c1, c2, c5) between spin_kernel markers, for a kernel trace")
    ap.add_argument("--trace-stages", default=None,
                    help="per-stage kernel time of each reconstruction of a --trace-runs kernel-trace CSV (no GPU)")
    ap.add_argument("--counter-summary", default=None,
                    help="per-wave SQ instruction counters of the assembly and SpMV kernels in a rocprofv3 --pmc CSV "
                         "(no GPU)")
    ap.add_argument("--trace-levels", default=None,
                    help="per-grid-size table of the PCG, residual and assembly kernels of a rocprofv3 kernel-trace "
                         "CSV (no GPU)")
    args = ap.parse_args()
    if args.trace_busy:
        trace_busy(args.trace_busy, args.trace_skip)
        return
    if args.trace_levels:
        trace_levels(args.trace_levels)
        return
    if args.trace_stages:
        trace_stages(args.trace_stages)
        return
    if args.counter_summary:
        counter_summary(args.counter_summary)
        return
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the path's multi-GPU sharding (by request for serving, "
                                   "by spatial slabs for one large solve) is not launched", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    import ngspsr

    if args.cpu_stages:
        cpu_stages()
        return
    if args.plain_pass:
        plain_pass(args.requests)
        return
    if args.batched_pass:
        batched_pass(args.requests)
        return
    if args.trace_runs:
        trace_runs("c2" if args.config == "c3" else args.config)
        return
    if args.pmc_probe:
        cfg = CONFIGS[args.probe_config]
        xyz_np, nrm_np = make_scene(cfg)
        dev = torch.device("cuda:0")
        res = ngspsr.reconstruct(torch.from_numpy(xyz_np).to(dev), torch.from_numpy(nrm_np).to(dev), cfg["vs"],
                                 depth=cfg["depth"], min_density=cfg["min_density"], screen_alpha=4.0,
                                 solver=cfg["solver"], return_result=True)
        pmc_probe(res)
        return

    common = {"metric": "solves/s + p99 latency at 1e-6 rel. residual, batched vs plain; PCG HBM GB/s",
              "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
              "precision": "FP32 level operators (assembled, run form) and PCG vectors; FP64 dot products, residual "
                           "and refinement", "data": "synthetic"}
    if args.config != "c3":
        out = single_solve(args.config, args.steps, args.warmup, cpu=not args.no_cpu, ext=not args.no_extensions)
        line = {**common, **{k: out[k] for k in ("value", "unit", "steps", "warmup", "ms_per_step")},
                "config": out["config"]}
        line["unit"] = "reconstructions/s"
        line["measure"] = (f"single-solve workload {CONFIGS[args.config]['workload']}: reconstructions/s end to end "
                           "(the serving metric's batched/plain ratio is the c3 workload's)")
        line.update({k: v for k, v in out.items() if k not in line})
        if not args.no_cpu:
            line["cpu_baseline"] = cpu_single_baseline(args.config, out)
        print(json.dumps(line), flush=True)
        return

    # config 3, the metric's workload: the serving stream
    srv, timed_s = run_serving(args.requests, args.steps, args.warmup, log=log)
    log("single-solve leg (2M points, 512^3, refinement to 1e-8)")
    single = single_solve("c2", 10, 2, cpu=False, ext=not args.no_extensions)
    line = {**common,
            "value": srv["batched_mixed_solves_per_s"], "unit": "solves/s",
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * timed_s / args.steps, 3),
            "step": f"one pass over the {args.requests}-request stream, all requests queued at t = 0",
            "config": {"workload": "serving stream of 256 random-ellipsoid requests (1e4-1e6 points, depth 7-9, "
                                   "4 levels) at relative residual 1e-6", "requests": args.requests,
                       "points_total": srv["points_total"], "server": srv["batched_config"]},
            "latency_ms": srv["poisson_latency_ms_batched"],
            "plain_fp32_solves_per_s": srv["plain_fp32_solves_per_s"],
            "batched_vs_plain": srv["north_star_ratio_batched_mixed_vs_plain_fp32"],
            "timed_region": "request points already in HBM; meshes stay on the device",
            "roofline": single["roofline"],
            "pcg_iteration": single["pcg_iteration"],
            "serving": srv,
            "single_solve": {k: v for k, v in single.items() if k not in ("roofline", "pcg_iteration")}}
    if not args.no_cpu:
        rec, _ = cpu_serving_sample()
        line["cpu_baseline"] = rec
    print(json.dumps(line), flush=True)


def cpu_single_baseline(name, out):
    """CPU baseline of a single-solve workload: the oracle's C/OpenMP PCG on the finest level,
    scaled by the GPU run's iterations over all levels (PCG only)."""
    c = out.get("cpu_pcg_finest")
    if c is None:
        return None
    total_iters = out["pcg_iterations_total"]
    rate = 1.0 / (c["ms_per_iteration"] * 1e-3 * total_iters)
    return {"value": round(rate, 5), "unit": "reconstructions/s", "cores": c["cores"], "kind": "port",
            "sample": f"{c['iterations_timed']} iterations of the oracle's C/OpenMP FP32 Jacobi-PCG (FP64 dots) on the "
                      f"finest-level system, {c['ms_per_iteration']} ms/iteration, scaled by the GPU run's "
                      f"{total_iters} PCG iterations over all levels; PCG only (an upper bound on CPU solves/s)"}


if __name__ == "__main__":
    main()
