"""Sparse symmetric solves of the jspsr path (jspsr/core/solver.py), on the HIP PCG.

Solver strings keep the reference's mini-language, "pcg@tol=1e-6@maxiter=1000@res_fix=False"
(solver.py:180-186), parsed without eval. Methods:
  pcg      Jacobi-PCG in FP32 with FP64 reductions; stop when ||r|| <= tol ||b|| (recurrence residual)
           or at maxiter (solve_pcg.cu:15-136 semantics; maxiter < 0 means uncapped in the reference,
           capped here at 10^7 with an error).
  mixed    the reference's "mixed" (solver.py:124-130,160-169): PCG with tol 1e-4 and at most 1000
           iterations; when the absolute residual ||A x - b|| exceeds tol (1e-2) the solve falls
           back to "cholmod".
  cholmod  the reference's direct sparse solve (solve_cusparse.cu:15-42): a supernodal multifrontal
           Cholesky factorization in FP64 (nested-dissection order, fronts factored on the GPU), kept
           on the system for later solves with the same matrix (the adjoint). Returns the FP64
           solution; a pivot at or below tol (1e-6, relative to its diagonal) is reported as the
           singularity, as cuSOLVER reports it.
  refine   FP64 iterative refinement around FP32 inner PCG solves (no reference counterpart):
           r = b - A x in FP64, inner solve A e = r to inner_tol, x += e in FP64, until
           ||b - A x|| / ||b|| <= tol.
"""
import ast
import ctypes

import torch

from . import _lib
from ._lib import NgsError, call, lib, ptr, stream

DEFAULTS = {"pcg": {"tol": 1e-5, "maxiter": -1, "res_fix": False, "verbose": False, "chunk": 32},
            "mixed": {"tol": 1e-2, "fast_tol": 1e-4, "fast_maxiter": 1000, "verbose": False, "chunk": 32},
            "cholmod": {"tol": 1e-6, "verbose": False, "chunk": 32},
            "refine": {"tol": 1e-8, "maxiter": 100000, "inner_tol": 1e-4, "max_outer": 20, "verbose": False,
                       "chunk": 32}}


def parse_solver(spec: str):
    parts = spec.split("@")
    name = parts[0]
    if name not in DEFAULTS:
        raise NotImplementedError(f"solver {name!r} (supported: {sorted(DEFAULTS)})")
    conf = dict(DEFAULTS[name])
    for p in parts[1:]:
        k, v = p.split("=", 1)
        conf[k] = ast.literal_eval(v)
    return name, conf


class CSRSystem:
    """Assembled level system: CSR (rows padded to multiples of 4 entries) and inverse diagonal."""

    def __init__(self, rowptr, cols, vals, inv_diag):
        self.rowptr, self.cols, self.vals, self.inv_diag = rowptr, cols, vals, inv_diag
        self.M = inv_diag.shape[0]
        self.device = vals.device

    @property
    def nnz_stored(self):
        return self.cols.shape[0]

    def nnz(self):
        """Non-padding entries (excludes the explicit zero padding)."""
        return int((self.vals != 0).sum().item())

    def matvec(self, x):
        y = torch.empty(self.M, dtype=torch.float32, device=self.device)
        if self.M:
            call("ngs_spmv", ptr(self.rowptr), ptr(self.cols), ptr(self.vals), ptr(x.float().contiguous()), ptr(y),
                 self.M, stream())
        return y

    def residual_f64(self, b64, x64):
        """(r = b - A x in FP64, ||r||^2, ||b||^2); x64 None means x = 0."""
        r = torch.empty(self.M, dtype=torch.float64, device=self.device)
        norms = torch.empty(2, dtype=torch.float64, device=self.device)
        ws = _lib.Scratch.get(lib().ngs_pcg_workspace_bytes(self.M), self.device, slot=1)
        call("ngs_residual_f64", ptr(self.rowptr), ptr(self.cols), ptr(self.vals), ptr(b64.contiguous()),
             None if x64 is None else ptr(x64.contiguous()), ptr(r), self.M, ptr(norms), ptr(ws), stream())
        nv = norms.cpu()
        return r, float(nv[0]), float(nv[1])

    def to_scipy(self):
        import numpy as np
        import scipy.sparse as sp
        rp = self.rowptr.cpu().numpy()
        a = sp.csr_matrix((self.vals.cpu().numpy().astype(np.float64), self.cols.cpu().numpy(), rp),
                          shape=(self.M, self.M))
        a.eliminate_zeros()
        return a


class RunSystem:
    """Assembled level system in run form (ngs_lhs_assemble_runs): for each row and each of the 25
    bitmap columns (dx, dy) of its 5x5x5 neighbourhood, one run word (first column | count << 28)
    and five value slots, both plane-major with the plane stride P = M rounded up to 64:
    runs (25, P) int32, rvals (125, P) float32. The values are the CSR entries bit for bit; the SpMV
    reads 600 bytes per row with every load an aligned contiguous segment across consecutive rows."""

    COUNT_SHIFT = 28
    # keep the matrix packed (values without the zero slots, ngs_runs_pack): the solver's SpMV and
    # FP64 residual read it; the slot values are regenerated on demand (rvals) for the entry points
    # that take them. False keeps the slot form (same results, bit for bit).
    PACKED = True

    @staticmethod
    def plane_stride(M):
        return (M + 63) // 64 * 64

    def __init__(self, runs, rvals, inv_diag, packed=None, layout=None):
        """rvals may be a caller's temporary when the system packs (it is read once, on the
        current stream). layout: (slice offsets, value count) of the packed form when the caller
        has them (from the level's row counts), else they are counted from the run words."""
        self.runs, self.inv_diag = runs, inv_diag
        self.M = inv_diag.shape[0]
        self.device = inv_diag.device
        assert runs.shape == (25, self.plane_stride(self.M)) and rvals.shape == (125, self.plane_stride(self.M))
        self._rvals = rvals
        self._packed = None
        if (self.PACKED if packed is None else packed) and self.M:
            self._pack(rvals, layout)
            self._rvals = None

    @classmethod
    def from_packed(cls, runs, pvals, soff, inv_diag):
        """A system assembled straight into the packed form (ngs_lhs_assemble_packed)."""
        self = cls.__new__(cls)
        self.runs, self.inv_diag = runs, inv_diag
        self.M = inv_diag.shape[0]
        self.device = inv_diag.device
        assert runs.shape == (25, cls.plane_stride(self.M)) and soff.shape == ((self.M + 63) // 64 + 1,)
        self._rvals = None
        self._packed = (pvals, soff)
        return self

    def _pack(self, rvals, layout=None):
        nsl = (self.M + 63) // 64
        if layout is None:
            counts = torch.empty(nsl, dtype=torch.int64, device=self.device)
            call("ngs_runs_slice_counts", ptr(self.runs), self.M, ptr(counts), stream())
            soff = torch.zeros(nsl + 1, dtype=torch.int64, device=self.device)
            torch.cumsum(counts, 0, out=soff[1:])
            nvals = int(soff[-1].item())
        else:
            soff, nvals = layout
            assert soff.shape == (nsl + 1,)
        pvals = torch.empty(max(nvals, 1), dtype=torch.float32, device=self.device)
        call("ngs_runs_pack", ptr(self.runs), ptr(rvals), self.M, ptr(soff), ptr(pvals), stream())
        self._packed = (pvals, soff)

    @property
    def is_packed(self):
        return self._packed is not None

    @property
    def rvals(self):
        """The slot values (125, P) float32, regenerated from the packed form when the system packs."""
        if self._rvals is None:
            P = self.plane_stride(self.M)
            rv = torch.zeros((125, P), dtype=torch.float32, device=self.device)
            if self.M:
                pvals, soff = self._packed
                call("ngs_runs_unpack", ptr(self.runs), ptr(pvals), self.M, ptr(soff), ptr(rv), stream())
            self._rvals = rv
        return self._rvals

    def counts(self):
        """(25, M) run lengths."""
        return (self.runs[:, :self.M] >> self.COUNT_SHIFT) & 7

    def nnz(self):
        """Stored entries (the CSR entries of the same matrix, without padding)."""
        return int(self.counts().sum(dtype=torch.int64).item())

    def matvec(self, x):
        y = torch.empty(self.M, dtype=torch.float32, device=self.device)
        if self.M:
            xc = x.float().contiguous()
            if self._packed is not None:
                pvals, soff = self._packed
                call("ngs_spmv_packed", ptr(self.runs), ptr(pvals), ptr(soff), ptr(xc), ptr(y), self.M, stream())
            else:
                call("ngs_spmv_runs", ptr(self.runs), ptr(self.rvals), ptr(xc), ptr(y), self.M, stream())
        return y

    def matvec_slots(self, x):
        """y = A x through the slot form (the same sums as matvec, bit for bit)."""
        y = torch.empty(self.M, dtype=torch.float32, device=self.device)
        if self.M:
            call("ngs_spmv_runs", ptr(self.runs), ptr(self.rvals), ptr(x.float().contiguous()), ptr(y), self.M,
                 stream())
        return y

    def residual_f64(self, b64, x64):
        """(r = b - A x in FP64, ||r||^2, ||b||^2); x64 None means x = 0."""
        r = torch.empty(self.M, dtype=torch.float64, device=self.device)
        norms = torch.empty(2, dtype=torch.float64, device=self.device)
        ws = _lib.Scratch.get(lib().ngs_pcg_workspace_bytes(self.M), self.device, slot=1)
        xp = None if x64 is None else ptr(x64.contiguous())
        if self._packed is not None:
            pvals, soff = self._packed
            call("ngs_residual_f64_packed", ptr(self.runs), ptr(pvals), ptr(soff), ptr(b64.contiguous()), xp, ptr(r),
                 self.M, ptr(norms), ptr(ws), stream())
        else:
            call("ngs_residual_f64_runs", ptr(self.runs), ptr(self.rvals), ptr(b64.contiguous()), xp, ptr(r), self.M,
                 ptr(norms), ptr(ws), stream())
        nv = norms.cpu()
        return r, float(nv[0]), float(nv[1])

    def coo(self):
        """(rows, cols, vals) of the stored entries, rows ascending and columns ascending per row."""
        M, dev = self.M, self.device
        n = self.counts()
        start = self.runs[:, :M] & ((1 << self.COUNT_SHIFT) - 1)
        j = torch.arange(5, dtype=torch.int32, device=dev).view(1, 5, 1)
        # (row, plane, slot) order is increasing column order within a row
        present = (j < n.reshape(25, 1, M)).permute(2, 0, 1)
        cols = (start.reshape(25, 1, M) + j).permute(2, 0, 1)[present]
        rows = torch.arange(M, dtype=torch.int32, device=dev).view(1, 1, M).expand(25, 5, M).permute(2, 0, 1)[present]
        vals = self.rvals[:, :M].reshape(25, 5, M).permute(2, 0, 1)[present]
        return rows, cols, vals

    def to_csr(self):
        """The same matrix as the CSR of ngs_lhs_assemble: rows in column order, padded to multiples
        of 4 entries with the row's own column and value 0 (the CSR kernels read 4 entries at a time)."""
        M, dev = self.M, self.device
        r, c, v = self.coo()
        counts = torch.bincount(r, minlength=M)
        padded = (counts + 3) // 4 * 4
        rowptr = torch.zeros(M + 1, dtype=torch.int64, device=dev)
        rowptr[1:] = torch.cumsum(padded, 0)
        first = torch.cumsum(counts, 0) - counts  # index of each row's first entry in (r, c, v)
        rl = r.long()
        pos = rowptr[:-1][rl] + (torch.arange(r.shape[0], device=dev) - first[rl])
        total = int(rowptr[M].item())
        cols = torch.repeat_interleave(torch.arange(M, dtype=torch.int32, device=dev), padded)
        vals = torch.zeros(total, dtype=torch.float32, device=dev)
        cols[pos] = c
        vals[pos] = v
        return CSRSystem(rowptr.to(torch.int32), cols, vals, self.inv_diag)

    def to_scipy(self):
        import numpy as np
        import scipy.sparse as sp
        r, c, v = self.coo()
        a = sp.csr_matrix((v.cpu().numpy().astype(np.float64), (r.cpu().numpy(), c.cpu().numpy())),
                          shape=(self.M, self.M))
        a.eliminate_zeros()
        return a


class StencilSystem:
    """Level system in stencil form: the operator of ngs_lhs_assemble applied from the level bitmap,
    the 125 Laplacian coefficients of the level and the screening points binned by cell, without
    stored entries (per-iteration traffic a few bytes per row instead of 8 per entry). Same
    interface as CSRSystem; to_csr() assembles the identical operator as CSR on demand."""

    def __init__(self, grid, stride: int, bins, sf: float):
        """grid: the level's SparseGrid (voxel bitmap, rank and coordinates), held by the system so
        that the raw pointers of the kernel arguments stay valid for the system's lifetime."""
        from ._lib import Level, Stencil
        self.grid = grid
        level = Level()
        level.grid = grid.struct()
        level.coords = grid.coords.data_ptr() if grid.count else None
        level.M = grid.count
        level.stride = int(stride)
        self.device = grid.device
        self.level = level
        self.bins = bins
        self.sf = float(sf)
        self.M = int(level.M)
        dev = self.device
        self.lap = torch.empty(125, dtype=torch.float32, device=dev)
        call("ngs_stencil_laplacian", int(level.stride), ptr(self.lap), stream())
        n = bins.n
        self.n = n
        self.perm = bins.perm
        if n:
            # the bin-order point arrays of the binning
            self.point_cell, self.basis, self.alpha = bins.point_cell, bins.b, bins.alpha
            self.u = torch.empty(n, dtype=torch.float32, device=dev)
        else:
            self.point_cell = self.basis = self.alpha = self.u = None
        self.inv_diag = torch.empty(self.M, dtype=torch.float32, device=dev)
        st = Stencil()
        st.level = level
        st.cells = bins.cells.struct()
        st.cell_start = bins.cell_start.data_ptr()
        st.point_cell = None if self.point_cell is None else self.point_cell.data_ptr()
        st.basis = None if self.basis is None else self.basis.data_ptr()
        st.alpha = None if self.alpha is None else self.alpha.data_ptr()
        st.perm = None if self.perm is None else self.perm.data_ptr()
        st.lap = self.lap.data_ptr()
        st.u = None if self.u is None else self.u.data_ptr()
        st.alpha_scalar = float(bins.alpha_scalar)
        st.sf = self.sf
        st.n = n
        self._struct = st

    def struct(self):
        return self._struct

    def diag_rhs(self, coarse, delta, rhs):
        """inv_diag and the screening RHS terms (ngs_stencil_diag_rhs)."""
        if self.M:
            call("ngs_stencil_diag_rhs", self._struct, ptr(coarse), float(delta), ptr(self.inv_diag), ptr(rhs),
                 stream())

    def nnz(self):
        """Entries of the equivalent CSR (existing neighbours in the 125-stencil of each row)."""
        if not self.M:
            return 0
        cnt = torch.empty(self.M, dtype=torch.int32, device=self.device)
        call("ngs_stencil_row_counts", self.level, ptr(cnt), stream())
        return int(cnt.sum(dtype=torch.int64).item())

    def matvec(self, x):
        y = torch.empty(self.M, dtype=torch.float32, device=self.device)
        if self.M:
            call("ngs_stencil_apply", self._struct, ptr(x.float().contiguous()), ptr(y), stream())
        return y

    def residual_f64(self, b64, x64):
        """(r = b - A x in FP64, ||r||^2, ||b||^2), the operator evaluated in FP64; x64 None = 0."""
        r = torch.empty(self.M, dtype=torch.float64, device=self.device)
        norms = torch.empty(2, dtype=torch.float64, device=self.device)
        ws = _lib.Scratch.get(lib().ngs_pcg_workspace_bytes(self.M), self.device, slot=1)
        u64 = _lib.Scratch.get(8 * max(self.n, 1), self.device, slot=4)
        call("ngs_stencil_residual_f64", self._struct, ptr(b64.contiguous()),
             None if x64 is None else ptr(x64.contiguous()), ptr(r), ptr(u64), ptr(norms), ptr(ws), stream())
        nv = norms.cpu()
        return r, float(nv[0]), float(nv[1])

    def to_csr(self):
        """The same operator assembled as CSR (ngs_lhs_assemble over the same binned points)."""
        M, dev = self.M, self.device
        row_nnz = torch.empty(M, dtype=torch.int32, device=dev)
        call("ngs_lhs_row_counts", self.level, ptr(row_nnz), stream())
        rowptr = torch.empty(M + 1, dtype=torch.int32, device=dev)
        ws = _lib.Scratch.get(lib().ngs_scan_workspace_bytes(M), dev)
        call("ngs_scan_i32", ptr(row_nnz), M, ptr(rowptr), ptr(ws), stream())
        nnz = int(rowptr[M].item())
        cols = torch.empty(nnz, dtype=torch.int32, device=dev)
        vals = torch.empty(nnz, dtype=torch.float32, device=dev)
        inv_diag = torch.empty(M, dtype=torch.float32, device=dev)
        rhs = torch.zeros(M, dtype=torch.float64, device=dev)
        call("ngs_lhs_assemble", self.level, self.bins.struct(), None, self.sf, 0.0, ptr(rowptr), ptr(cols),
             ptr(vals), ptr(inv_diag), ptr(rhs), stream())
        return CSRSystem(rowptr, cols, vals, inv_diag)

    def to_scipy(self):
        return self.to_csr().to_scipy()


class SolveInfo:
    def __init__(self, iters=0, rel_resid=0.0, outer=0, true_rel_resid=None):
        self.iters, self.rel_resid, self.outer, self.true_rel_resid = iters, rel_resid, outer, true_rel_resid
        self.fallback = False  # "mixed": the direct solve answered
        self.direct = False  # answered by the sparse Cholesky factorization
        self.singularity = -1  # direct solve: row of the first singular pivot (-1: none)

    def __repr__(self):
        return (f"SolveInfo(iters={self.iters}, rel_resid={self.rel_resid:.3e}, outer={self.outer}, "
                f"true_rel_resid={self.true_rel_resid})")


def pcg(system, b, tol=1e-5, maxiter=-1, chunk=32, x_out=None, raise_on_cap=True):
    """FP32 Jacobi-PCG from x = 0. Returns (x, SolveInfo)."""
    M = system.M
    x = x_out if x_out is not None else torch.empty(M, dtype=torch.float32, device=system.device)
    if M == 0:
        return x, SolveInfo()
    ws = _lib.Scratch.get(lib().ngs_pcg_workspace_bytes(M), system.device, slot=2)
    it = ctypes.c_int32()
    rr = ctypes.c_double()
    b = b.float().contiguous()
    if isinstance(system, StencilSystem):
        st = lib().ngs_pcg_solve_stencil(system.struct(), ptr(system.inv_diag), ptr(b), ptr(x), float(tol),
                                         int(maxiter), int(chunk), ptr(ws), stream(), ctypes.byref(it),
                                         ctypes.byref(rr))
    elif isinstance(system, RunSystem):
        if system.is_packed:
            pvals, soff = system._packed
            st = lib().ngs_pcg_solve_packed(ptr(system.runs), ptr(pvals), ptr(soff), ptr(system.inv_diag), ptr(b),
                                            ptr(x), M, float(tol), int(maxiter), int(chunk), ptr(ws), stream(),
                                            ctypes.byref(it), ctypes.byref(rr))
        else:
            st = lib().ngs_pcg_solve_runs(ptr(system.runs), ptr(system.rvals), ptr(system.inv_diag), ptr(b), ptr(x),
                                          M, float(tol), int(maxiter), int(chunk), ptr(ws), stream(),
                                          ctypes.byref(it), ctypes.byref(rr))
    else:
        st = lib().ngs_pcg_solve(ptr(system.rowptr), ptr(system.cols), ptr(system.vals), ptr(system.inv_diag),
                                 ptr(b), ptr(x), M, float(tol), int(maxiter), int(chunk), ptr(ws), stream(),
                                 ctypes.byref(it), ctypes.byref(rr))
    if st != 0 and not (st == 5 and not raise_on_cap):
        raise NgsError(st, lib().ngs_last_error().decode())
    return x, SolveInfo(it.value, rr.value)


FP32_REACH = 1e-6  # relative residual an FP32 Jacobi-PCG reaches on these systems before it stagnates


def inner_tolerance(tol, inner_tol, outer):
    """Relative tolerance of an inner FP32 solve of iterative refinement. Targets the FP32 solver
    reaches on its own (tol >= 1e-6) are aimed at directly by the first inner solve, to half the
    target so that its FP32 rounding still lands below it, and verified in FP64: one FP32 solve and
    one FP64 residual. Tighter targets take inner_tol per step (each step gains that factor)."""
    if outer == 0 and tol >= FP32_REACH:
        return 0.5 * tol
    return inner_tol


def refine(system, b, tol=1e-8, inner_tol=1e-4, max_outer=20, maxiter=100000, chunk=32):
    """FP64 iterative refinement with FP32 PCG inner solves. Returns (x64, SolveInfo)."""
    M = system.M
    b64 = b.double().contiguous()
    x64 = torch.zeros(M, dtype=torch.float64, device=system.device)
    info = SolveInfo()
    if M == 0:
        return x64, info
    r64, rr, bb = system.residual_f64(b64, None)  # x = 0: r = b without a matrix pass
    nb = bb ** 0.5
    rel = (rr ** 0.5) / nb if nb > 0 else 0.0
    while rel > tol and info.outer < max_outer:
        e, inner = pcg(system, r64.float(), tol=inner_tolerance(tol, inner_tol, info.outer), maxiter=maxiter,
                       chunk=chunk, raise_on_cap=False)
        x64 += e.double()
        info.iters += inner.iters
        info.outer += 1
        r64, rr, _ = system.residual_f64(b64, x64)
        rel = (rr ** 0.5) / nb
    info.rel_resid = rel
    info.true_rel_resid = rel
    return x64, info


class CholeskyFactor:
    """Sparse Cholesky factor of a level system (ngs_chol_*): the analysis (order, supernodes, front
    structures) on the host, the FP64 factor in a device workspace. solve(b) -> x (FP64)."""

    def __init__(self, csr, tol=1e-6):
        self.M = csr.M
        self.device = csr.device
        self.csr = csr
        h = ctypes.c_void_p()
        # the pattern read on the host (ordered after the stream that wrote it), analysed there
        rp = csr.rowptr.to(torch.int32).cpu().contiguous()
        cl = csr.cols.to(torch.int32).cpu().contiguous()
        call("ngs_chol_analyze_host", ctypes.c_void_p(rp.data_ptr()), ctypes.c_void_p(cl.data_ptr()), self.M,
             ctypes.byref(h))
        self.handle = h
        self.ws = torch.empty(max(int(lib().ngs_chol_workspace_bytes(h)), 256), dtype=torch.uint8, device=self.device)
        sing = ctypes.c_int64(-1)
        call("ngs_chol_factor", h, ptr(csr.rowptr), ptr(csr.cols), ptr(csr.vals), float(tol), ptr(self.ws), stream(),
             ctypes.byref(sing))
        self.singularity = int(sing.value)

    def info(self):
        st = (ctypes.c_int64 * 8)()
        call("ngs_chol_info", self.handle, st)
        keys = ("supernodes", "heights", "max_front", "factor_entries", "pool_doubles", "index_entries",
                "map_entries", "csr_entries")
        return dict(zip(keys, [int(v) for v in st]))

    def solve(self, b):
        x = torch.empty(self.M, dtype=torch.float64, device=self.device)
        if self.M:
            b = b.contiguous()
            f64 = b.dtype == torch.float64
            call("ngs_chol_solve", self.handle, ptr(self.ws), None if f64 else ptr(b.float()), ptr(b) if f64 else None,
                 None, ptr(x), stream())
        return x

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                lib().ngs_chol_free(h)
            except Exception:  # interpreter shutdown
                pass
            self.handle = None


def cholesky(system, tol=1e-6):
    """The system's Cholesky factor, computed on first use and kept on the system (the adjoint and
    later solves with the same matrix reuse it)."""
    f = getattr(system, "_chol", None)
    if f is None or f.tol != tol:
        csr = system.to_csr() if isinstance(system, (RunSystem, StencilSystem)) else system
        f = CholeskyFactor(csr, tol)
        f.tol = tol
        system._chol = f
    return f


def cholmod(system, b, tol=1e-6, chunk=32, refine_tol=1e-12, max_refine=4):
    """The reference's direct solve (solve_cholmod, solver.py:147-157 -> solve_cusparse.cu:15-42): the
    sparse Cholesky factor of the system (FP64, from its lower triangle), two triangular solves, then
    iterative refinement against the full stored matrix (FP64 residual, correction through the
    factor) until the relative residual is below refine_tol: the assembled FP32 rows are symmetric
    only to rounding, so the factor alone stops near 1e-8. Returns (x64, SolveInfo) with the FP64
    relative residual of x and, in SolveInfo.singularity, the first singular pivot's row (-1: none)."""
    info = SolveInfo()
    if system.M == 0:
        return torch.zeros(0, dtype=torch.float64, device=system.device), info
    f = cholesky(system, tol)
    b64 = b.double().contiguous()
    x = f.solve(b64)
    r, rr, bb = system.residual_f64(b64, x)
    rel = (rr / bb) ** 0.5 if bb > 0 else 0.0
    while rel > refine_tol and info.outer < max_refine:
        x += f.solve(r)
        info.outer += 1
        r, rr, _ = system.residual_f64(b64, x)
        rel_new = (rr / bb) ** 0.5
        if not rel_new < rel:  # no further gain (a singular or indefinite system)
            rel = rel_new
            break
        rel = rel_new
    info.rel_resid = info.true_rel_resid = rel
    info.singularity = f.singularity
    info.direct = True
    return x, info


def mixed(system, b, tol=1e-2, fast_tol=1e-4, fast_maxiter=1000, chunk=32):
    """The reference's "mixed" solve (solve_mixed with solve_pcg(maxiter 1000, tol 1e-4) and the
    cholmod fallback, solver.py:124-130,160-169): PCG first; if ||A x - b|| > tol (absolute), the
    direct-solve stand-in. SolveInfo.fallback tells which one answered. Returns (x float32, SolveInfo)."""
    x, info = pcg(system, b, tol=fast_tol, maxiter=fast_maxiter, chunk=chunk, raise_on_cap=False)
    if system.M == 0:
        return x, info
    _, rr, _ = system.residual_f64(b.double(), x.double())
    if rr ** 0.5 > tol:
        x, info2 = cholmod(system, b, chunk=chunk)
        info2.iters += info.iters
        info2.fallback = True
        return x.float(), info2
    return x, info


def solve(system, b, spec: str = "pcg"):
    """Dispatch on a solver string. Returns (x, SolveInfo): x is float32, or float64 for "refine"
    (the refinement's FP64 iterate, whose residual SolveInfo.rel_resid reports) and "cholmod" (the
    FP64 direct solution)."""
    name, conf = parse_solver(spec)
    if name == "pcg":
        x, info = pcg(system, b, tol=conf["tol"], maxiter=conf["maxiter"], chunk=conf["chunk"])
    elif name == "mixed":
        x, info = mixed(system, b, tol=conf["tol"], fast_tol=conf["fast_tol"], fast_maxiter=conf["fast_maxiter"],
                        chunk=conf["chunk"])
    elif name == "cholmod":
        x, info = cholmod(system, b, tol=conf["tol"], chunk=conf["chunk"])
    else:
        # the FP64 iterate is the solution: its FP64 residual is the one SolveInfo reports (its FP32
        # rounding would sit at the FP32 floor, ~1e-8 relative, above the tighter tolerances)
        x, info = refine(system, b, tol=conf["tol"], inner_tol=conf["inner_tol"], max_outer=conf["max_outer"],
                         maxiter=conf["maxiter"], chunk=conf["chunk"])
    if conf.get("verbose"):
        print("PCG Iteration ended in", info.iters)
    return x, info
