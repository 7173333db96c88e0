"""Continuous-batching reconstruction server on one GPU.

Requests of uneven size (point count, depth) are admitted while their estimated HBM footprint fits
the budget and a worker slot is free. Each admitted request runs its pipeline (hierarchy, splat,
assembly, mesh) on its own stream in a worker thread and keeps its tree, matrices and Krylov state
on the device. Every linear solve is handed to one shared pool: a pool thread advances all active
level systems of all requests together, one launch per kernel per iteration (batched kernels over
a device work table, replayed as a captured hipGraph), admitting new systems and retiring converged
ones between chunks of iterations. Mixed precision (FP64 refinement around FP32 inner solves)
runs per request around the pool's solves.
"""
import concurrent.futures as cf
import math
import queue
import threading
import time

import numpy as np
import torch

from . import solver as _solver
from ._lib import Marks, NgsError, call, lib, ptr, stream
from .bases import ensure_device_tables
from .hashtree import HashTree
from .reconstructor import Reconstructor


class _Job:
    __slots__ = ("system", "b", "x", "ws", "tol", "max_iter", "done", "iters", "status", "ready", "nblk", "t0")

    def __init__(self, system, b, tol, max_iter):
        self.system = system
        self.b = b.float().contiguous()
        self.x = torch.empty(system.M, dtype=torch.float32, device=system.device)
        self.ws = torch.empty(lib().ngs_pcg_workspace_bytes(system.M), dtype=torch.uint8, device=system.device)
        self.tol = tol
        self.max_iter = max_iter
        self.done = threading.Event()
        self.iters = 0
        self.status = 0
        self.nblk = int(lib().ngs_pcg_batch_blocks(system.M))
        self.ready = torch.cuda.Event()
        self.ready.record(torch.cuda.current_stream())


def launch_flags(systems):
    """ngs_pcg_batch_iterate flags for a set of systems: which SpMV kernels they need."""
    stencil = runs = packed = plain = False
    for s in systems:
        if isinstance(s, _solver.StencilSystem):
            stencil = True
        elif isinstance(s, _solver.RunSystem):
            if s.is_packed:
                packed = True
            else:
                runs = True
        else:
            plain = True
    return (1 if stencil else 0) | (2 if runs else 0) | (0 if plain else 4) | (8 if packed else 0)


class BatchedPCG:
    """Pool of concurrently iterating PCG systems (one batched launch per kernel per iteration)."""

    def __init__(self, device, capacity_blocks=4096, max_systems=256, chunk=8, use_graph=True):
        self.device = torch.device(device)
        self.cap = int(capacity_blocks)
        self.max_sys = int(max_systems)
        self.chunk = int(chunk)
        self.use_graph = bool(use_graph)
        self.eb = int(lib().ngs_pcg_batch_entry_bytes())
        self.host_table = torch.zeros(self.max_sys * self.eb, dtype=torch.uint8).pin_memory()
        self.dev_table = torch.zeros(self.max_sys * self.eb, dtype=torch.uint8, device=self.device)
        self.host_map = torch.full((self.cap,), -1, dtype=torch.int32).pin_memory()
        self.dev_map = torch.full((self.cap,), -1, dtype=torch.int32, device=self.device)
        self.states = torch.zeros(2 * self.max_sys, dtype=torch.int32).pin_memory()
        self.dev_states = torch.zeros(2 * self.max_sys, dtype=torch.int32, device=self.device)
        self.stream = torch.cuda.Stream(self.device)
        self.pending = queue.Queue()
        self.active = []
        self.launches = 0
        self.iterations = 0
        self.busy_s = 0.0
        self._stop = False
        self._err = None
        self._wake = threading.Event()
        self.thread = threading.Thread(target=self._run, name="ngspsr-pcg-pool", daemon=True)
        self.thread.start()

    def solve(self, system, b, tol=1e-6, max_iter=100000):
        """Blocking solve through the pool (callable from any worker thread)."""
        if self._err is not None:
            raise self._err
        job = _Job(system, b, tol, max_iter)
        self.pending.put(job)
        self._wake.set()
        while not job.done.wait(0.5):
            if self._err is not None or not self.thread.is_alive():
                raise self._err or RuntimeError("PCG pool thread stopped")
        if job.status == 2 and self._err is not None:
            raise self._err
        return job.x, _solver.SolveInfo(job.iters, 0.0)

    def close(self):
        self._stop = True
        self._wake.set()
        self.thread.join()

    def _layout(self):
        blk = 0
        hm = self.host_map.numpy()
        hm[:] = -1
        for i, job in enumerate(self.active):
            blk0 = blk
            hm[blk0:blk0 + job.nblk] = i
            call("ngs_pcg_batch_place", self.host_table[i * self.eb:].data_ptr(), blk0)
            blk = blk0 + ((job.nblk + 7) // 8) * 8
        return blk

    def _admit(self):
        used = sum(((j.nblk + 7) // 8) * 8 for j in self.active)
        admitted = []
        while len(self.active) < self.max_sys:
            try:
                job = self.pending.get_nowait()
            except queue.Empty:
                break
            need = ((job.nblk + 7) // 8) * 8
            if used + need > self.cap and self.active:
                self.pending.put(job)  # retry after something retires
                break
            used += need
            self.stream.wait_event(job.ready)
            slot = len(self.active)
            entry = self.host_table[slot * self.eb:(slot + 1) * self.eb]
            s = job.system
            if isinstance(s, _solver.StencilSystem):
                # one system's point scratch (s.u) serves one solve at a time
                call("ngs_pcg_batch_prepare_stencil", s.struct(), ptr(s.inv_diag), ptr(job.b), ptr(job.x),
                     float(job.tol), int(job.max_iter), ptr(job.ws), 0, job.nblk, entry.data_ptr(),
                     self.stream.cuda_stream)
            elif isinstance(s, _solver.RunSystem) and s.is_packed:
                pvals, soff = s._packed
                call("ngs_pcg_batch_prepare_packed", ptr(s.runs), ptr(pvals), ptr(soff), ptr(s.inv_diag), ptr(job.b),
                     ptr(job.x), s.M, float(job.tol), int(job.max_iter), ptr(job.ws), 0, job.nblk, entry.data_ptr(),
                     self.stream.cuda_stream)
            elif isinstance(s, _solver.RunSystem):
                call("ngs_pcg_batch_prepare_runs", ptr(s.runs), ptr(s.rvals), ptr(s.inv_diag), ptr(job.b), ptr(job.x),
                     s.M, float(job.tol), int(job.max_iter), ptr(job.ws), 0, job.nblk, entry.data_ptr(),
                     self.stream.cuda_stream)
            else:
                call("ngs_pcg_batch_prepare", ptr(s.rowptr), ptr(s.cols), ptr(s.vals), ptr(s.inv_diag), ptr(job.b),
                     ptr(job.x), s.M, float(job.tol), int(job.max_iter), ptr(job.ws), 0, job.nblk, entry.data_ptr(),
                     self.stream.cuda_stream)
            self.active.append(job)
            admitted.append(job)
        return admitted

    def _run(self):
        torch.cuda.set_device(self.device)
        try:
            with torch.cuda.stream(self.stream):
                while not self._stop:
                    changed = bool(self._admit())
                    if not self.active:
                        self._wake.wait(0.01)
                        self._wake.clear()
                        continue
                    t0 = time.perf_counter()
                    if changed or self._dirty:
                        self._layout()
                        n = len(self.active)
                        self.dev_table[:n * self.eb].copy_(self.host_table[:n * self.eb], non_blocking=True)
                        self.dev_map.copy_(self.host_map, non_blocking=True)
                        self._dirty = False
                    flags = launch_flags(j.system for j in self.active)
                    call("ngs_pcg_batch_iterate", ptr(self.dev_table), ptr(self.dev_map), self.cap, self.chunk,
                         int(self.use_graph), flags, self.stream.cuda_stream)
                    self.launches += 1
                    self.iterations += self.chunk
                    n = len(self.active)
                    call("ngs_pcg_batch_status", self.host_table.data_ptr(), n, ptr(self.dev_states),
                         self.states.data_ptr(), self.stream.cuda_stream)
                    self.busy_s += time.perf_counter() - t0
                    st = self.states.numpy()[:2 * n].reshape(n, 2)
                    keep = []
                    for i, job in enumerate(self.active):
                        if st[i, 1]:
                            job.iters, job.status = int(st[i, 0]), int(st[i, 1])
                            job.done.set()
                        else:
                            keep.append(job)
                    if len(keep) != len(self.active):
                        # compact the table: entries move with their jobs
                        old = self.host_table.clone()
                        for j, job in enumerate(keep):
                            i = self.active.index(job)
                            self.host_table[j * self.eb:(j + 1) * self.eb] = old[i * self.eb:(i + 1) * self.eb]
                        self.active = keep
                        self._dirty = True
        except Exception as e:  # surface pool failures to waiting requests
            self._err = e
            for job in self.active:
                job.status = 2
                job.done.set()
            while not self.pending.empty():
                j = self.pending.get_nowait()
                j.status = 2
                j.done.set()

    _dirty = False


def estimate_footprint(n_points: int, voxel_size: float = None) -> int:
    """HBM bytes a request reserves at admission, before its tree exists: the tree and splat stages
    (measured under 250 bytes per point at 1e4-1e6 points) plus a margin. The request reserves its
    full footprint (level_footprint) once its tree gives the voxel counts."""
    return int(250.0 * n_points + 64 * 2 ** 20)


def level_footprint(level_sizes, n_points: int) -> int:
    """Peak HBM bytes of a request from its tree's voxel counts: the dual-mesh stage dominates (its
    dual cubes, corner positions and values, vertex keys and their sort), measured at 2.7-5.7 KB
    per finest voxel and up to 3.9 KB per voxel of all levels plus 200 bytes per point on the
    requests of the serving stream (1e4-1e6 points, depth 7-9); solve state and the returned mesh
    are included."""
    return int(4000.0 * sum(int(m) for m in level_sizes) + 200.0 * n_points + 256 * 2 ** 20)


class Request:
    def __init__(self, rid, xyz, normal, voxel_size, depth=4, min_density=1.0, screen_alpha=4.0, arrival=0.0):
        self.rid, self.xyz, self.normal = rid, xyz, normal
        self.voxel_size, self.depth, self.min_density, self.screen_alpha = voxel_size, depth, min_density, screen_alpha
        self.arrival = arrival
        self.footprint = estimate_footprint(xyz.shape[0], voxel_size)
        self.start = self.finish = None
        self.result = None
        self.error = None
        self.max_rel_residual = None


class ReconstructionServer:
    """Continuous batching of reconstructions: admission by HBM footprint, concurrent pipelines,
    one shared batched PCG pool."""

    def __init__(self, device="cuda:0", workers=8, hbm_budget=None, tol=1e-6, refine=True, inner_tol=1e-4,
                 chunk=8, capacity_blocks=4096, use_graph=True, hbm_fraction=0.85):
        """hbm_fraction (None: leave as is) caps this process's caching allocator on the device at
        that fraction of its memory (torch.cuda.set_per_process_memory_fraction, process-wide):
        every worker stream keeps its own cached blocks, and without a cap the cache can fill the
        device and leave the HIP runtime no room for its own allocations (queue scratch); at the
        cap the allocator releases cached blocks and retries instead."""
        self.device = torch.device(device)
        torch.cuda.set_device(self.device)
        if hbm_fraction is not None:
            torch.cuda.set_per_process_memory_fraction(float(hbm_fraction), self.device)
        ensure_device_tables(self.device)
        free, total = torch.cuda.mem_get_info(self.device)
        self.budget = int(hbm_budget if hbm_budget is not None else 0.6 * free)
        self.tol, self.refine, self.inner_tol = tol, refine, inner_tol
        self.pool = BatchedPCG(self.device, capacity_blocks=capacity_blocks, chunk=chunk, use_graph=use_graph)
        self.workers = workers
        # persistent worker threads: each keeps its thread-local scratch and mark buffers (ngspsr._lib
        # Scratch, Marks) from one request to the next, so steady-state requests allocate only their
        # own tree, matrices and Krylov state
        self.executor = cf.ThreadPoolExecutor(max_workers=workers, thread_name_prefix="ngspsr-request")
        # one stream per worker for its lifetime: the caching allocator reuses freed blocks only on
        # the stream that allocated them, so a fresh stream per request would keep every request's
        # peak cached on its own stream
        self._tls = threading.local()
        self.in_use = 0
        self.lock = threading.Condition()
        self.completed = []

    # ---------------------------------------------------------------------- solves via the pool
    def _solve(self, system, rhs):
        if not self.refine:
            x, info = self.pool.solve(system, rhs, tol=self.tol)
            return x, info
        b64 = rhs.double()
        x64 = torch.zeros(system.M, dtype=torch.float64, device=system.device)
        r64, rr, bb = system.residual_f64(b64, None)  # x = 0: r = b without a matrix pass
        nb = math.sqrt(bb)
        rel = math.sqrt(rr) / nb if nb > 0 else 0.0
        info = _solver.SolveInfo()
        while rel > self.tol and info.outer < 20:
            inner_tol = _solver.inner_tolerance(self.tol, self.inner_tol, info.outer)
            e, inner = self.pool.solve(system, r64.float(), tol=inner_tol)
            x64 += e.double()
            info.iters += inner.iters
            info.outer += 1
            r64, rr, _ = system.residual_f64(b64, x64)
            rel = math.sqrt(rr) / nb
        info.rel_resid = info.true_rel_resid = rel
        return x64, info

    def _worker_stream(self):
        s = getattr(self._tls, "stream", None)
        if s is None:
            s = self._tls.stream = torch.cuda.Stream(self.device)
        return s

    def run_request(self, req: Request):
        s = self._worker_stream()
        try:
            with torch.cuda.stream(s):
                req.start = time.perf_counter()
                if req.xyz.shape[0] == 0:
                    # no points, no surface (as reconstruct): nothing is launched
                    req.result = (torch.zeros((0, 3), dtype=torch.float32, device=self.device),
                                  torch.zeros((0, 3), dtype=torch.int64, device=self.device))
                    req.max_rel_residual = 0.0
                    return
                tree = HashTree(req.xyz, req.voxel_size, req.depth)
                tree.build_encoder_hierarchy_adaptive(min_density=req.min_density)
                # the voxel counts are known now: reserve the request's full footprint
                self._reserve(req, level_footprint([tree.get_coords_size(None, d) for d in range(req.depth)],
                                                   req.xyz.shape[0]))
                tree.reflect_decoder_coords()
                w = 1.0 / tree.xyz_density
                counts = tree.level_point_counts()
                nd = {}
                for d in range(req.depth):
                    if counts[d]:
                        nd[d] = tree.splat_data(req.xyz, None, d, req.normal, scale=w, levels=tree.xyz_depth,
                                                level=d) / float(tree.strides[d] ** 3)
                rec = Reconstructor(tree)
                rec.sample_weight = w
                rec.solve_multigrid(req.depth - 1, 0, nd, screen_alpha=req.screen_alpha, screen_xyz=req.xyz,
                                    solve_fn=self._solve)
                v, f = rec.extract_multiscale_mesh(n_upsample=2, build_o3d_mesh=False)
                s.synchronize()
                req.result = (v, f)
                infos = rec.solve_info.values()
                req.max_rel_residual = max((i.rel_resid for i in infos), default=0.0)
        except Exception as e:  # reported per request
            req.error = e
        finally:
            if req.error is not None:
                # let the failed request's launches drain before the worker takes the next one
                try:
                    s.synchronize()
                except Exception:  # the request's error is already recorded
                    pass
            req.finish = time.perf_counter()
            cb = getattr(self, "_on_complete", None)
            if cb is not None:
                try:
                    cb(req)
                except Exception as e:  # reported on the request
                    req.error = req.error or e
            with self.lock:
                self.in_use -= req.footprint
                self.workers_busy -= 1
                self.completed.append(req)
                self.lock.notify_all()

    workers_busy = 0

    def _reserve(self, req, need):
        """Raises the request's reservation to `need` bytes, waiting while it would exceed the
        budget and other requests hold memory (a request alone is always let through)."""
        with self.lock:
            while need > req.footprint and self.in_use - req.footprint + need > self.budget \
                    and self.in_use > req.footprint:
                self.lock.wait()
            if need > req.footprint:
                self.in_use += need - req.footprint
                req.footprint = need

    def serve(self, requests, t0=None, on_complete=None):
        """Runs the requests honouring their arrival offsets (seconds after t0); returns them.
        on_complete(req), when given, is called on the worker thread as each request finishes
        (result or error set, its stream synchronised): the place to hand the mesh on and drop it,
        since results held by the caller are device memory that admission does not count."""
        self._on_complete = on_complete
        reqs = sorted(requests, key=lambda r: r.arrival)
        t0 = time.perf_counter() if t0 is None else t0
        futures = []
        for req in reqs:
            delay = t0 + req.arrival - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            req.arrival_abs = t0 + req.arrival
            with self.lock:
                while self.workers_busy >= self.workers or (self.in_use + req.footprint > self.budget and self.in_use > 0):
                    self.lock.wait()
                self.workers_busy += 1
                self.in_use += req.footprint
            futures.append(self.executor.submit(self.run_request, req))
        for fu in futures:
            fu.result()
        return reqs

    def close(self):
        self.executor.shutdown(wait=True)
        self.pool.close()


def latency_stats(reqs):
    lat = np.array([r.finish - r.arrival_abs for r in reqs])
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3)}
