"""Screened Poisson solve and mesh extraction on the GPU (jspsr/core/reconstructor.py)."""
import ctypes

import numpy as np
import torch

from . import solver as _solver
from ._lib import Grid, PointBins, Scratch, call, lib, ptr, stream
from .bases import BezierTensorBasis, ensure_device_tables
from .grid import SparseGrid
from .hashtree import HashTree


class PointBinning:
    """Points of the screening term binned by their containing cell at one level, with the 3x3
    per-axis basis values of the 27 voxels around each point (reconstructor.py:111-120). Per-point
    arrays are stored in bin order: slot k holds point perm[k]."""

    def __init__(self, tree: HashTree, xyz, d: int, alpha, _finalize=True):
        n = xyz.shape[0]
        s = tree.strides[d]
        dev = tree.device
        self.n = n
        cell = torch.empty((n, 3), dtype=torch.int32, device=dev)
        b = torch.empty((n, 9), dtype=torch.float32, device=dev)
        call("ngs_point_basis", ptr(xyz), n, tree.inv_vs, s, ptr(cell), ptr(b), stream())
        qmin, qmax = tree.point_bounds(xyz)
        self.cells = SparseGrid.for_points(qmin, qmax, s, dev)
        self.cells.mark_coords(cell)
        self._pending = (cell, b)
        if _finalize:
            self.cells.finalize(emit=False)
            self._bucket(alpha)

    @classmethod
    def levels(cls, tree: HashTree, xyz, levels, alpha):
        """{d: PointBinning} of several levels, their cell counts read in one device-to-host read."""
        out = {d: cls(tree, xyz, d, alpha, _finalize=False) for d in levels}
        SparseGrid.finalize_all([b.cells for b in out.values()], emit=False)
        for b in out.values():
            b._bucket(alpha)
        return out

    def _bucket(self, alpha):
        """Points sorted by cell rank (the cells grid finalised), per-point arrays in bin order."""
        cell, b = self._pending
        self._pending = None
        n = self.n
        dev = cell.device
        self.ncells = self.cells.count
        rank = torch.empty(n, dtype=torch.int32, device=dev)
        call("ngs_grid_rank_rows", self.cells.struct(), ptr(cell), n, ptr(rank), stream())
        self.perm = torch.empty(n, dtype=torch.int32, device=dev)
        self.cell_start = torch.empty(self.ncells + 1, dtype=torch.int32, device=dev)
        wsb = lib().ngs_bucket_workspace_bytes(n, self.ncells)
        ws = Scratch.get(wsb, dev, slot=3)
        call("ngs_bucket_points", ptr(rank), n, self.ncells, ptr(self.perm), ptr(self.cell_start), ptr(ws), wsb, stream())
        alpha_t = alpha.float().contiguous() if isinstance(alpha, torch.Tensor) else None
        self.point_cell = torch.empty((n, 3), dtype=torch.int32, device=dev)
        self.b = torch.empty((n, 9), dtype=torch.float32, device=dev)
        self.alpha = torch.empty(n, dtype=torch.float32, device=dev) if alpha_t is not None else None
        call("ngs_bin_points", ptr(self.perm), n, ptr(cell), ptr(b), ptr(alpha_t), ptr(self.point_cell), ptr(self.b),
             ptr(self.alpha), stream())
        self.alpha_scalar = 0.0 if alpha_t is not None else float(alpha)

    @classmethod
    def empty(cls, tree):
        obj = cls.__new__(cls)
        obj.n = 0
        obj.cells = SparseGrid.empty(tree.device)
        obj.cell_start = torch.zeros(1, dtype=torch.int32, device=tree.device)
        obj.perm = obj.b = obj.alpha = obj.point_cell = None
        obj.ncells = 0
        obj.alpha_scalar = 0.0
        return obj

    def struct(self) -> PointBins:
        pb = PointBins()
        pb.cells = self.cells.struct()
        pb.cell_start = self.cell_start.data_ptr()
        pb.perm = None if self.perm is None else self.perm.data_ptr()
        pb.b = None if self.b is None else self.b.data_ptr()
        pb.alpha = None if self.alpha is None else self.alpha.data_ptr()
        pb.alpha_scalar = self.alpha_scalar
        pb.n = self.n
        return pb


class Reconstructor:
    def __init__(self, hash_tree: HashTree, basis=None, feat=None):
        """reconstructor.py:26-49. The degree-2 B-spline basis (the default) runs the specialised
        kernels below; any other basis (AdditiveBasis, AdditiveHermite) runs the reference's generic
        path with per-voxel basis features (`feat`: {depth: (M_d, basis.get_feature_size())}, rows in
        the level's row order; initialised by the basis when omitted)."""
        self.hash_tree = hash_tree
        self.basis = basis if basis is not None else BezierTensorBasis()
        self.generic = not isinstance(self.basis, BezierTensorBasis)
        if self.generic and feat is None:
            feat = {}
            for d in range(hash_tree.depth):
                f = torch.zeros((hash_tree.get_coords_size(HashTree.DECODER, d), self.basis.get_feature_size()),
                                device=hash_tree.device)
                self.basis.initialize_feature_value(f)
                feat[d] = f
        self.grid_features = feat if feat is not None else {}
        self.solutions = {}
        self._f32 = {}
        self._points_chi = None
        self.systems = {}
        self.solve_info = {}
        self.sample_weight = None
        self.fixed_level_set = None
        self.branch = HashTree.DECODER
        self.operator = "auto"  # level operator form: "auto", "runs", "csr" or "stencil"
        ensure_device_tables(hash_tree.device)

    def set_fixed_level_set(self, value):
        self.fixed_level_set = value

    def field(self, d):
        """Level d's coefficients as the FP32 vector the field kernels read, in rank order: the
        solution itself, or its FP32 rounding when the solver returned FP64 (iterative refinement),
        reordered when the level's rows are not in rank order (cached per tensor and in-place
        version). The kernels sum in FP32, so the rounding of the coefficients is of the same order as
        the evaluation's own."""
        sol = self.solutions[d]
        lv = self.hash_tree._branch(self.branch)[d]
        if sol.dtype == torch.float32 and (lv is None or lv.row_of_rank is None):
            return sol
        c = self._f32.get(d)
        if c is None or c[0] is not sol or c[1] != sol._version:
            c = (sol, sol._version, self.hash_tree.to_rank(d, sol.float()).contiguous())
            self._f32[d] = c
        return c[2]

    # Per-application time model of the operator forms (picoseconds, fitted to MI355X kernel timings
    # on 0.05M-5.6M-row levels): the run form ~115 ps per row (600 bytes per row at ~5.3 TB/s); CSR
    # ~2.5 ps per stored entry (~90 entries per row); the stencil ~122 ps per row for the Laplacian
    # part, ~3.5 ps per (row, point) visit of the screening part (27 cells around each row) and
    # ~50 ps per point for the point pass. The run form indexes rows with 28 bits.
    RUNS_PS_PER_ROW = 115.0
    CSR_PS_PER_ROW = 2.5 * 90
    STENCIL_PS_PER_ROW = 122.0
    STENCIL_PS_PER_VISIT = 3.5
    STENCIL_PS_PER_POINT = 50.0
    RUNS_MAX_ROWS = 1 << 28

    def operator_form(self, M, n_points, n_cells=0):
        if self.operator in ("runs", "csr", "stencil"):
            form = self.operator
        else:
            visits = 27.0 * n_points / max(n_cells, 1) if n_points else 0.0
            stencil = ((self.STENCIL_PS_PER_ROW + self.STENCIL_PS_PER_VISIT * visits) * M
                       + self.STENCIL_PS_PER_POINT * n_points)
            form = "stencil" if stencil < self.RUNS_PS_PER_ROW * M else "runs"
        if form == "runs" and M > self.RUNS_MAX_ROWS:
            form = "csr"
        return form

    def packed_layouts(self, levels, bins):
        """{d: (slice offsets, value count)} of the packed run form for the levels that take it: the
        level grid alone fixes each row's entry count, so the offsets of all levels are computed
        before any is assembled and their totals come back in one device-to-host read."""
        if not _solver.RunSystem.PACKED:
            return {}
        tree = self.hash_tree
        soffs = {d: self._slice_offsets(d) for d in levels
                 if tree.level(d).M and self.operator_form(tree.level(d).M, bins[d].n,
                                                           getattr(bins[d], "ncells", 0)) == "runs"}
        if not soffs:
            return {}
        totals = torch.cat([s[-1:] for s in soffs.values()]).cpu().tolist()
        return {d: (s, int(t)) for (d, s), t in zip(soffs.items(), totals)}

    def _slice_offsets(self, d):
        """Offsets of level d's 64-row slices in the packed form (slice sums of the row entry
        counts, scanned; nsl + 1 entries)."""
        lv = self.hash_tree.level(d)
        dev = self.hash_tree.device
        nsl = (lv.M + 63) // 64
        row_nnz = torch.zeros(nsl * 64, dtype=torch.int32, device=dev)
        call("ngs_stencil_row_counts", lv, ptr(row_nnz), stream())
        soff = torch.zeros(nsl + 1, dtype=torch.int64, device=dev)
        torch.cumsum(row_nnz.view(nsl, 64).sum(1, dtype=torch.int64), 0, out=soff[1:])
        return soff

    def assemble_level(self, d, bins, coarse, sf, delta, rhs, form=None, layout=None):
        """Level-d system (reconstructor.py:143-183); the screening RHS terms are added to rhs.
        form: "runs" (ngs_lhs_assemble_runs), "csr" (ngs_lhs_assemble), "stencil" (ngs_stencil_*),
        None = operator_form. layout: the level's packed-form offsets from packed_layouts (the run
        form only; counted here when None)."""
        tree = self.hash_tree
        lv = tree.level(d)
        M = lv.M
        dev = tree.device
        if form is None:
            form = self.operator_form(M, bins.n, getattr(bins, "ncells", 0))
        if form == "stencil":
            system = _solver.StencilSystem(tree.grids[d], tree.strides[d], bins, sf)
            system.diag_rhs(coarse, delta, rhs)
            return system
        inv_diag = torch.empty(M, dtype=torch.float32, device=dev)
        pb = bins.struct()
        if coarse is not None and bins.perm is not None and bins.n:
            # the coarse field in bin order: the assembly then streams it with the binned points
            cb = torch.empty(bins.n, dtype=torch.float32, device=dev)
            call("ngs_gather_f32", ptr(coarse), ptr(bins.perm), bins.n, ptr(cb), stream())
            coarse, pb.perm = cb, None
        if form == "runs":
            P = _solver.RunSystem.plane_stride(M)
            runs = torch.empty((25, P), dtype=torch.int32, device=dev)
            if _solver.RunSystem.PACKED and lib().ngs_assembly_lanes(int(pb.n), M) == 1:
                # one lane per row: the assembly writes the packed form itself, at slice offsets from
                # the row counts (64-row slice sums, scanned)
                if layout is None:
                    soff = self._slice_offsets(d)
                    layout = (soff, int(soff[-1].item()))
                soff, nvals = layout
                pvals = torch.empty(max(nvals, 1), dtype=torch.float32, device=dev)
                call("ngs_lhs_assemble_packed", lv, pb, ptr(coarse), float(sf), float(delta), ptr(runs), ptr(pvals),
                     ptr(soff), ptr(inv_diag), ptr(rhs), stream())
                return _solver.RunSystem.from_packed(runs, pvals, soff, inv_diag)
            if _solver.RunSystem.PACKED:
                # the slot values only feed the packing: a thread scratch buffer, reused level to level
                rvals = Scratch.get(125 * P * 4, dev, slot=7)[:125 * P * 4].view(torch.float32).view(125, P)
            else:
                rvals = torch.empty((125, P), dtype=torch.float32, device=dev)
            call("ngs_lhs_assemble_runs", lv, pb, ptr(coarse), float(sf), float(delta), ptr(runs),
                 ptr(rvals), ptr(inv_diag), ptr(rhs), stream())
            return _solver.RunSystem(runs, rvals, inv_diag, layout=layout)
        if form != "csr":
            raise ValueError(f"unknown operator form {form!r}")
        row_nnz = torch.empty(M, dtype=torch.int32, device=dev)
        call("ngs_lhs_row_counts", lv, ptr(row_nnz), stream())
        rowptr = torch.empty(M + 1, dtype=torch.int32, device=dev)
        ws = Scratch.get(lib().ngs_scan_workspace_bytes(M), dev)
        call("ngs_scan_i32", ptr(row_nnz), M, ptr(rowptr), ptr(ws), stream())
        nnz = int(rowptr[M].item())
        cols = torch.empty(nnz, dtype=torch.int32, device=dev)
        vals = torch.empty(nnz, dtype=torch.float32, device=dev)
        call("ngs_lhs_assemble", lv, pb, ptr(coarse), float(sf), float(delta), ptr(rowptr), ptr(cols),
             ptr(vals), ptr(inv_diag), ptr(rhs), stream())
        return _solver.CSRSystem(rowptr, cols, vals, inv_diag)

    def _add_field_at_points(self, d, sol, xyz, bins, out):
        """out += level d's field at the screening points (ngs_eval_field's values bit for bit), read
        through the level's point bins when they hold these points."""
        tree = self.hash_tree
        if bins is not None and bins.perm is not None and bins.n == xyz.shape[0]:
            call("ngs_eval_field_binned", tree.grids[d].struct(), ptr(sol), ptr(bins.point_cell), ptr(bins.b),
                 ptr(bins.perm), bins.n, 1.0, ptr(out), stream())
        else:
            call("ngs_eval_field", tree.grids[d].struct(), ptr(sol), ptr(xyz), xyz.shape[0], tree.inv_vs, 1.0, 1,
                 ptr(out), None, stream())

    def solve_multigrid(self, start_depth, end_depth, normal_data: dict, screen_alpha=0.0, screen_xyz=None,
                        screen_delta=0.1, solver: str = "pcg", verbose: bool = False, solve_fn=None):
        """reconstructor.py:79-195: coarse-to-fine per-level solves; each level's RHS is corrected by
        the coarser solutions (cascadic). solve_fn(system, rhs) -> (x, info) replaces the solver
        (the serving engine routes solves through its batched pool)."""
        if self.generic:
            return self._solve_multigrid_generic(start_depth, end_depth, normal_data, screen_alpha, screen_xyz,
                                                 screen_delta, solver, verbose, solve_fn)
        tree = self.hash_tree
        if isinstance(screen_alpha, torch.Tensor):
            assert screen_xyz is not None and screen_alpha.shape[0] == screen_xyz.shape[0]
            should_screen = True
            alpha = screen_alpha.float().contiguous()
        else:
            should_screen = screen_alpha > 0.0
            alpha = float(screen_alpha)
        delta_t = None
        if isinstance(screen_delta, torch.Tensor):
            # per-point targets (reconstructor.py:143-147): the screening RHS sf sum_p a_p (d_p - X_p) B_i(p)
            # is the kernels' -sf sum_p a_p X'_p B_i(p) with X' = X - d and a scalar target of 0
            assert screen_xyz is not None and screen_delta.shape[0] == screen_xyz.shape[0]
            delta_t = screen_delta.to(device=tree.device, dtype=torch.float32).contiguous()
        self.solutions = {}
        self._points_chi = None
        running, added = None, []
        xyz = screen_xyz.float().contiguous() if should_screen else None
        levels = list(range(end_depth, start_depth + 1))
        bins = PointBinning.levels(tree, xyz, levels, alpha) if should_screen else \
            {d: PointBinning.empty(tree) for d in levels}
        layouts = self.packed_layouts(levels, bins)
        for d in range(start_depth, end_depth - 1, -1):
            M = tree.get_coords_size(None, d)
            if M == 0:
                self.solutions[d] = torch.zeros(0, dtype=torch.float32, device=tree.device)
                continue
            sf = 0.25 ** d
            rhs = torch.zeros(M, dtype=torch.float64, device=tree.device)
            for dd, nd in normal_data.items():
                nd = tree.to_rank(dd, nd.float()).contiguous()
                call("ngs_rhs_normals", tree.level(d), tree.level(dd), ptr(nd), ptr(rhs), stream())
            coarser = [dd for dd in range(start_depth, d, -1) if tree.get_coords_size(None, dd) > 0]
            coarse = None
            if should_screen and coarser:
                # field of the coarser levels at the points, accumulated level by level from the coarsest
                # (the order of evaluate_raw_chi): each level is added once, when it is first needed
                if running is None:
                    running = torch.zeros(xyz.shape[0], dtype=torch.float32, device=tree.device)
                for dd in coarser:
                    if dd not in added:
                        self._add_field_at_points(dd, self.field(dd), xyz, bins[dd], running)
                        added.append(dd)
                coarse = running
            if delta_t is not None and should_screen:
                coarse = (coarse - delta_t) if coarse is not None else -delta_t
            for dd in coarser:
                call("ngs_rhs_cross_stiffness", tree.level(d), tree.level(dd), ptr(self.field(dd)), ptr(rhs),
                     stream())
            system = self.assemble_level(d, bins[d], coarse, sf, 0.0 if delta_t is not None else float(screen_delta),
                                         rhs, layout=layouts.get(d))
            # "mixed" tries PCG only on the finest level; the coarser levels are solved directly
            # (reconstructor.py:186-189)
            level_solver = solver
            if solver.split("@")[0] == "mixed" and d != end_depth:
                level_solver = "cholmod"
            x, info = solve_fn(system, rhs) if solve_fn is not None else _solver.solve(system, rhs, level_solver)
            self.solutions[d] = tree.to_rows(d, x).contiguous()
            self.systems[d] = (system, rhs)
            self.solve_info[d] = info
            if verbose:
                r = system.matvec(x).double() - rhs
                print(f"Solving complete at level {d}, residual = {torch.linalg.norm(r).item()}.")
        if running is not None and self._same_points(xyz):
            # completed with the remaining levels, the running sum is chi at the input points
            for dd, sol in self.solutions.items():
                if sol.numel() and dd not in added:
                    self._add_field_at_points(dd, self.field(dd), xyz, bins[dd], running)
            self._points_chi = (self._solutions_key(), running)
        return self.solutions

    # -------------------------------------------------------------------------- generic bases
    def _screen_data(self, d, base):
        """reconstructor.py:111-120: (point ids, voxel ids, basis values, entries per point) of the
        points around every voxel of level d, point-major."""
        tree = self.hash_tree
        pts, vx, off, nbs = tree.get_neighbours_data(base, 1, d, tree.get_range_kernel(self.basis.get_domain_range()),
                                                     self.branch, transposed=True)
        qc = -off / tree.get_stride(self.branch, d)
        vals = self.basis.evaluate(feat=self.grid_features[d], xyz=qc, feat_ids=vx)
        return pts, vx, vals, nbs

    def _screening_term(self, sa, sb, src, tgt, m_a, m_b, weight):
        """reconstructor.py:59-77 (screened_multiplication): per COO pair (src of level a, tgt of level
        b), sum over the points p around both of w_p B_src(p) B_tgt(p); ngs_screen_pairs locates each
        pair in the row-sorted pair structure instead of the reference's pair hash table."""
        dev = self.hash_tree.device
        E = src.shape[0]
        out = torch.zeros(E, dtype=torch.float32, device=dev)
        if E == 0 or sa[0].numel() == 0 or sb[0].numel() == 0:
            return out
        order = torch.argsort(src.long() * m_b + tgt.long())
        rowptr = torch.zeros(m_a + 1, dtype=torch.int32, device=dev)
        rowptr[1:] = torch.cumsum(torch.bincount(src.long(), minlength=m_a), 0).int()
        cols = tgt[order].int().contiguous()
        b_start = torch.zeros(sb[3].shape[0] + 1, dtype=torch.int32, device=dev)
        b_start[1:] = torch.cumsum(sb[3], 0).int()
        w = weight.float().contiguous() if isinstance(weight, torch.Tensor) else None
        o = torch.zeros(E, dtype=torch.float32, device=dev)
        a_pt, a_ids, a_vals = sa[0].int().contiguous(), sa[1].int().contiguous(), sa[2].float().contiguous()
        b_ids, b_vals = sb[1].int().contiguous(), sb[2].float().contiguous()
        call("ngs_screen_pairs", a_pt.shape[0], ptr(a_pt), ptr(a_ids), ptr(a_vals), ptr(b_start), ptr(b_ids),
             ptr(b_vals), ptr(w), 0.0 if w is not None else float(weight), ptr(rowptr), ptr(cols), ptr(o), stream())
        out[order] = o
        return out

    @staticmethod
    def coo_system(src, tgt, val, M):
        """COO level matrix (unique pairs) as the solver's CSR: rows sorted, columns ascending, rows
        padded to multiples of 4 entries with zero-valued diagonal entries (solver.py:40-68)."""
        dev = val.device
        order = torch.argsort(src.long() * M + tgt.long())
        s, t, v = src[order].long(), tgt[order].int(), val[order].float()
        cnt = torch.bincount(s, minlength=M)
        pcnt = cnt + (-cnt) % 4
        rowptr = torch.zeros(M + 1, dtype=torch.int64, device=dev)
        rowptr[1:] = torch.cumsum(pcnt, 0)
        cols = torch.repeat_interleave(torch.arange(M, dtype=torch.int32, device=dev), pcnt)
        vals = torch.zeros(cols.shape[0], dtype=torch.float32, device=dev)
        first = torch.cumsum(cnt, 0) - cnt
        pos = rowptr[s] + (torch.arange(s.shape[0], device=dev) - first[s])
        cols[pos] = t
        vals[pos] = v
        diag = torch.zeros(M, dtype=torch.float32, device=dev)
        on = s == t.long()
        diag[s[on]] = v[on]
        return _solver.CSRSystem(rowptr.int(), cols, vals, 1.0 / diag)

    def _solve_multigrid_generic(self, start_depth, end_depth, normal_data, screen_alpha, screen_xyz, screen_delta,
                                 solver, verbose, solve_fn):
        """reconstructor.py:79-195 for bases with per-voxel features: COO neighbour pairs of the tree,
        Galerkin integrals of the basis (DDI kernels), the screening pairs (ngs_screen_pairs), the
        level system in CSR and the chosen solver. Every per-voxel array is in the level's row order."""
        tree, basis, br = self.hash_tree, self.basis, self.branch
        dev = tree.device
        feats = self.grid_features
        if isinstance(screen_alpha, torch.Tensor):
            assert screen_xyz is not None and screen_alpha.shape[0] == screen_xyz.shape[0]
            should_screen = True
        else:
            should_screen = screen_alpha > 0.0
        self.solutions = {}
        self._points_chi = None
        nrange = 2 * basis.get_domain_range() - 1
        screen = {}
        if should_screen:
            base = screen_xyz.float() * tree.inv_vs - 0.5  # screen_xyz / voxel_size - 0.5
            for d in range(end_depth, start_depth + 1):
                screen[d] = self._screen_data(d, base)
        for d in range(start_depth, end_depth - 1, -1):
            M = tree.get_coords_size(br, d)
            if M == 0:
                self.solutions[d] = torch.zeros(0, dtype=torch.float32, device=dev)
                continue
            sf = 0.25 ** d
            s_d = tree.get_stride(br, d)
            rhs = torch.zeros(M, dtype=torch.float64, device=dev)
            for dd, nd in normal_data.items():
                nids, tids, off, _ = tree.get_neighbours(dd, d, basis.get_domain_range(), br)
                part = basis.integrate_const_deriv_product(data=-nd.float()[nids], target_feat=feats[d], rel_pos=off,
                                                           data_stride=tree.get_stride(br, dd), target_stride=s_d,
                                                           target_ids=tids)
                rhs.index_add_(0, tids, part.double())
            if should_screen:
                pts, vx, vals, _ = screen[d]
                if isinstance(screen_alpha, torch.Tensor) or isinstance(screen_delta, torch.Tensor):
                    mult = (screen_delta * screen_alpha)[pts]
                else:
                    mult = screen_alpha * screen_delta
                part = torch.zeros(M, dtype=torch.float64, device=dev).index_add_(0, vx, (vals * mult).double())
                rhs += sf * part
            for dd in range(start_depth, d, -1):
                m_dd = tree.get_coords_size(br, dd)
                if m_dd == 0:
                    continue
                src, tgt, rel, _ = tree.get_neighbours(d, dd, nrange, br)
                a = basis.integrate_deriv_deriv_product(source_feat=feats[d], target_feat=feats[dd], rel_pos=rel,
                                                        source_stride=s_d, target_stride=tree.get_stride(br, dd),
                                                        source_ids=src, target_ids=tgt)
                if should_screen:
                    a = a + sf * self._screening_term(screen[d], screen[dd], src, tgt, M, m_dd, screen_alpha)
                rhs.index_add_(0, src, -(self.solutions[dd].double()[tgt] * a.double()))
            src, tgt, rel, _ = tree.get_neighbours(d, d, nrange, br)
            lhs = basis.integrate_deriv_deriv_product(source_feat=feats[d], target_feat=feats[d], rel_pos=rel,
                                                      source_stride=s_d, target_stride=s_d, source_ids=src,
                                                      target_ids=tgt)
            if should_screen:
                lhs = lhs + sf * self._screening_term(screen[d], screen[d], src, tgt, M, M, screen_alpha)
            system = self.coo_system(src, tgt, lhs, M)
            level_solver = "cholmod" if (solver.split("@")[0] == "mixed" and d != end_depth) else solver
            x, info = solve_fn(system, rhs) if solve_fn is not None else _solver.solve(system, rhs, level_solver)
            self.solutions[d] = x.contiguous()
            self.systems[d] = (system, rhs)
            self.solve_info[d] = info
            if verbose:
                r = system.matvec(x).double() - rhs
                print(f"Solving complete at level {d}, residual = {torch.linalg.norm(r).item()}.")
        return self.solutions

    def _evaluate_generic(self, q, compute_mask, compute_grad, depths):
        """hashtree.py:157-192 with the basis and its features, summed over the solved levels."""
        tree, basis = self.hash_tree, self.basis
        n = q.shape[0]
        base = q * tree.inv_vs - 0.5
        out = torch.zeros((n, 3) if compute_grad else n, dtype=torch.float32, device=tree.device)
        mask = torch.zeros(n, dtype=torch.bool, device=tree.device) if compute_mask else None
        for d, sol in self.solutions.items():
            if (depths is not None and d not in depths) or sol.numel() == 0:
                continue
            st = tree.get_stride(self.branch, d)
            bids, tids, off, _ = tree.get_neighbours_data(base, 1, d, tree.get_range_kernel(basis.get_domain_range()),
                                                          self.branch)
            qc = -off / st
            f = self.grid_features[d]
            if compute_grad:
                v = basis.evaluate_derivative(feat=f, xyz=qc, feat_ids=tids, stride=st) * sol.float()[tids, None]
            else:
                v = basis.evaluate(feat=f, xyz=qc, feat_ids=tids) * sol.float()[tids]
            out.index_add_(0, bids, v.float())
            if compute_mask:
                mask[bids] = True
        return (out, mask) if compute_mask else out

    def _same_points(self, xyz):
        q = self.hash_tree.xyz
        return xyz is not None and xyz.data_ptr() == q.data_ptr() and xyz.shape == q.shape and xyz.dtype == q.dtype

    def _solutions_key(self):
        # the tensors themselves (kept alive by the key, so identity cannot be reused) and their
        # in-place version counters
        return [(d, s, s._version) for d, s in self.solutions.items()]

    def _key_matches(self, key):
        cur = list(self.solutions.items())
        return len(cur) == len(key) and all(d == kd and s is ks and s._version == kv
                                            for (d, s), (kd, ks, kv) in zip(cur, key))

    # -------------------------------------------------------------------------- field
    def evaluate_raw_chi(self, xyz, compute_mask: bool = False, compute_grad: bool = False, depths=None):
        """reconstructor.py:197-226: chi at metric positions xyz (N, 3), summed over the solved levels
        in solve order (restricted to `depths` when given). With compute_grad, the gradient (N, 3) per
        finest-voxel unit in place of chi (bezier_tensor.py:62-72). With compute_mask, also the (N,)
        bool mask of queries that some level's basis covers (a voxel among the 27 around the query's
        cell at any evaluated level; hashtree.py:186-190)."""
        assert len(self.solutions) > 0, "Please run solver before evaluation."
        if self.generic:
            return self._evaluate_generic(xyz.float().contiguous(), compute_mask, compute_grad, depths)
        tree = self.hash_tree
        q = xyz.float().contiguous()
        n = q.shape[0]
        out = torch.zeros(n, dtype=torch.float32, device=tree.device)
        grad = torch.zeros((n, 3), dtype=torch.float32, device=tree.device) if compute_grad else None
        mask = torch.zeros(n, dtype=torch.bool, device=tree.device) if compute_mask else None
        for d, sol in self.solutions.items():
            if depths is not None and d not in depths:
                continue
            if sol.numel() == 0:
                continue
            g = tree.grids[d].struct()
            call("ngs_eval_field", g, ptr(self.field(d)), ptr(q), n, tree.inv_vs, 1.0, 1, ptr(out), ptr(grad),
                 stream())
            if compute_mask:
                call("ngs_field_support", g, ptr(q), n, tree.inv_vs, ptr(mask), stream())
        val = grad if compute_grad else out
        return (val, mask) if compute_mask else val

    def get_mean_chi(self):
        """reconstructor.py:228-236: level set = sample-weighted mean of chi at the input points."""
        if self.fixed_level_set is not None:
            return float(self.fixed_level_set)
        cached = getattr(self, "_points_chi", None)
        if cached is not None and self._key_matches(cached[0]):
            chi = cached[1]  # raw chi at the input points, summed during the solve in the same order
        else:
            chi = self.evaluate_raw_chi(self.hash_tree.xyz)
        sums = torch.empty(2, dtype=torch.float64, device=chi.device)
        w = None if self.sample_weight is None else self.sample_weight.float().contiguous()
        call("ngs_weighted_sum", ptr(chi), ptr(w), chi.shape[0], ptr(sums), stream())
        s = sums.cpu()
        return float(s[0] / s[1])

    def evaluate_chi(self, xyz, compute_mask: bool = False, max_points: int = -1, depths=None):
        """reconstructor.py:238-255: chi minus the level set, evaluated in chunks of at most
        max_points queries (all at once when max_points <= 0); with compute_mask also the cover mask."""
        mean = self.get_mean_chi()
        n = xyz.shape[0]
        step = n if max_points <= 0 else max_points
        vals, masks = [], []
        for a in range(0, max(n, 1), max(step, 1)):
            r = self.evaluate_raw_chi(xyz[a:a + step], compute_mask=compute_mask, depths=depths)
            vals.append(r[0] if compute_mask else r)
            if compute_mask:
                masks.append(r[1])
        val = (vals[0] if len(vals) == 1 else torch.cat(vals)) - mean
        if compute_mask:
            return val, (masks[0] if len(masks) == 1 else torch.cat(masks))
        return val

    # -------------------------------------------------------------------------- mesh
    def extract_mesh(self, base_coords, chi_field, chi_depth: int, build_o3d_mesh: bool = True):
        """reconstructor.py:257-294: primal sparse marching cubes over the voxels `base_coords` of
        level chi_depth, each carrying an R^3 block of chi samples (get_test_grid order), with the
        MeshSDF backward to chi_field. Returns (vertices (V, 3), triangles (T, 3) int64, vertex
        normals (V, 3)) in metric units, or with build_o3d_mesh a triangle mesh object (Open3D's when it
        is installed, else ngspsr.io.TriangleMesh with the same attribute names)."""
        from .ops import marching_cubes_op
        scale = self.hash_tree.get_stride(None, chi_depth)
        num_lif = base_coords.shape[0]
        if num_lif == 0:  # no voxels, no surface
            dev = self.hash_tree.device
            vertices = torch.zeros((0, 3), dtype=torch.float32, device=dev)
            triangles = torch.zeros((0, 3), dtype=torch.int64, device=dev)
            if not build_o3d_mesh:
                return vertices, triangles, torch.zeros((0, 3), dtype=torch.float32, device=dev)
            from .io import triangle_mesh
            return triangle_mesh(vertices, triangles, vertices)
        res = round((chi_field.shape[0] // num_lif) ** (1 / 3.0))
        chi = chi_field.reshape(-1, res, res, res)
        vertices, triangles, normals, _ = marching_cubes_op(base_coords.float() / scale, chi)
        vertices = vertices * (scale * self.hash_tree.voxel_size)
        if not build_o3d_mesh:
            return vertices, triangles, normals
        from .io import triangle_mesh
        return triangle_mesh(vertices, triangles, normals)

    def extract_multiscale_mesh(self, n_upsample: int = 1, max_depth: int = 100, expand: int = 0, trim: bool = False,
                                build_o3d_mesh: bool = True, max_points: int = -1, return_info=False):
        """reconstructor.py:296-397: dual marching cubes over the leaves of the tree, each split into
        n_upsample^3 sub-cells. The leaves are those of the conformal tree (every voxel's parent
        replaced by all 8 children, then the cells with finer children dropped), over each level's
        voxels dilated by the expand-cube when expand >= 3; with trim, the voxels without children
        (evaluate_voxel_status == VS_EXIST_STOP). Returns (vertices (V,3) float32, triangles (T,3)
        int64), or a mesh object with build_o3d_mesh."""
        tree = self.hash_tree
        dev = tree.device
        L = min(max_depth, tree.depth)
        # per level: the voxels the leaves come from and their box (lattice units of the level's stride)
        lev_coords, lev_box = {}, {}
        for d in range(L):
            s = tree.strides[d]
            g = tree.grids[d]
            lo, hi = list(g.origin), [o + dm - 1 for o, dm in zip(g.origin, g.dims)]
            coords = g.coords
            if expand >= 3 and not trim:
                if expand % 2 == 0:
                    raise ValueError("expand must be odd")
                h = expand // 2
                lo, hi = [v - h for v in lo], [v + h for v in hi]
                dg = SparseGrid(lo, hi, s, dev)
                if coords.shape[0]:
                    call("ngs_grid_mark_dilated", dg.struct(), ptr(coords), coords.shape[0], h, stream())
                dg.finalize()
                coords = dg.coords
            lev_coords[d], lev_box[d] = coords, (lo, hi)
        # parents (stride 2s) of every level but the last, sorted
        parents = {}
        for d in range(L - 1):
            s = tree.strides[d]
            lo, hi = lev_box[d]
            pg = SparseGrid([(o * s) // (2 * s) - 1 for o in lo], [(o * s) // (2 * s) + 1 for o in hi], 2 * s, dev)
            pg.mark_coords(lev_coords[d])
            parents[d] = pg
        SparseGrid.finalize_all(list(parents.values()))
        leaves = {}
        max_d = L
        for d in range(L):
            s = tree.strides[d]
            if d < L - 1 and not trim:
                npar = parents[d].count
                conf = torch.empty((npar * 8, 3), dtype=torch.int32, device=dev)
                keep = torch.empty(npar * 8, dtype=torch.int32, device=dev)
                fin = parents[d - 1] if d > 0 else SparseGrid.empty(dev)
                call("ngs_conformal_children", ptr(parents[d].coords), npar, s, fin.struct(), int(d > 0), ptr(conf),
                     ptr(keep), stream())
            else:
                conf = lev_coords[d]
                keep = torch.ones(conf.shape[0], dtype=torch.int32, device=dev)
                if d > 0 and conf.shape[0]:
                    call("ngs_unrefined_flags", ptr(conf), conf.shape[0], parents[d - 1].struct(), ptr(keep), stream())
            n = conf.shape[0]
            out = torch.empty((n, 3), dtype=torch.int32, device=dev)
            offs = torch.empty(n + 1, dtype=torch.int32, device=dev)
            ws = Scratch.get(lib().ngs_scan_workspace_bytes(n), dev)
            if n:
                call("ngs_compact_rows", ptr(conf), 3, ptr(keep), n, ptr(out), ptr(offs), ptr(ws), stream())
                k = int(offs[n].item())
            else:
                k = 0
            if d > 0 and k == 0 and not trim:
                max_d = d
                break
            leaves[d] = out[:k]
        L = max_d
        nu = n_upsample
        # n_upsample 2: the dual cubes are gathered per centre from per-level leaf bitmaps
        gather = nu == 2 and L <= 8
        # primal sub-cells (leaf-major, nu^3 per leaf); listed only where read: the per-sub-cell
        # dual-cube assembly of coarser levels (without the gather) and the report
        primal, bases_, nsub, nprim = {}, {}, {}, 0
        for d in range(L):
            s = tree.strides[d]
            lf = leaves[d]
            nsub[d] = lf.shape[0] * nu ** 3
            if (s > 1 and not gather) or return_info:
                pr = torch.empty((nsub[d], 3), dtype=torch.int32, device=dev)
                if lf.shape[0]:
                    call("ngs_expand_primal", ptr(lf), lf.shape[0], nu, s, ptr(pr), stream())
                primal[d] = pr
            bases_[d] = nprim
            nprim += nsub[d]
        # dual centres: sorted unique corners of all primal sub-cells (upsampled units)
        lo = [min((lev_box[d][0][a] - 2) * tree.strides[d] * nu for d in range(L)) for a in range(3)]
        hi = [max((lev_box[d][1][a] + 3) * tree.strides[d] * nu for d in range(L)) for a in range(3)]
        cg = SparseGrid(lo, hi, 1, dev)
        first = 0
        lgrid = None
        if nu == 2 and tree.strides[0] == 1 and leaves[0].shape[0]:
            # finest leaves: centre words computed from a leaf bitmap (no atomics), before the others
            lgrid = self._leaf_grid(0, leaves[0], parents, lev_box, trim)
            call("ngs_mark_dual_leaf_words", cg.struct(), lgrid.struct(), stream())
            first = 1
        for d in range(first, L):
            if leaves[d].shape[0]:
                call("ngs_mark_dual_leaves", cg.struct(), ptr(leaves[d]), leaves[d].shape[0], tree.strides[d], nu,
                     stream())
        DC = cg.finalize(emit=return_info or gather)  # coordinates read by the gathered assembly
        # dual corners = (primal + stride) * voxel_size / n_upsample, and chi there
        pos = torch.empty((nprim, 3), dtype=torch.float32, device=dev)
        scale = float(np.float32(tree.voxel_size / nu))
        for d in range(L):
            if nsub[d]:
                call("ngs_leaf_corner_positions", ptr(leaves[d]), leaves[d].shape[0], tree.strides[d], nu, scale,
                     ptr(pos[bases_[d]:]), stream())
        if self.generic:  # the basis's own evaluation at the corner positions
            values = self.evaluate_chi(pos, max_points=max_points)
        else:
            values = self.leaf_corner_chi(leaves, bases_, nprim, nu, scale) - self.get_mean_chi()
        # primal id per octant, read only for the kept centres (all eight slots written); the count
        # word also counts the registered cells with a negative value
        acc = torch.empty((DC, 8), dtype=torch.int32, device=dev)
        cnt = torch.zeros(DC, dtype=torch.int32, device=dev)
        if gather:
            # n_upsample 2: every level's cubes gathered per centre from per-level leaf bitmaps (no atomics)
            bitmaps, index, base = [], [], []
            for d in range(L):
                if not leaves[d].shape[0]:
                    continue
                if d == 0 and lgrid is not None:
                    lg = lgrid
                else:
                    lg = self._leaf_grid(d, leaves[d], parents, lev_box, trim)
                lg.finalize(emit=False, count=False)  # the leaves are distinct: the rank of each is below n_l
                n_l = leaves[d].shape[0]
                ranks = torch.empty(n_l, dtype=torch.int32, device=dev)
                call("ngs_grid_rank_rows", lg.struct(), ptr(leaves[d]), n_l, ptr(ranks), stream())
                li = torch.empty(max(n_l, 1), dtype=torch.int32, device=dev)
                li[ranks.long()] = torch.arange(n_l, dtype=torch.int32, device=dev)
                bitmaps.append(lg)
                index.append(li)
                base.append(bases_[d])
            nl = len(bitmaps)
            if nl:  # no leaves at all (the finer levels empty, reconstructor.py:322-326): no cubes, empty mesh
                garr = (Grid * nl)(*[g.struct() for g in bitmaps])
                iarr = (ctypes.c_void_p * nl)(*[t.data_ptr() for t in index])
                barr = (ctypes.c_int32 * nl)(*base)
                call("ngs_dual_gather_cubes", ptr(cg.coords), DC, nl, garr, iarr, barr, ptr(values), ptr(acc),
                     ptr(cnt), stream())
        else:
            for d in range(L):
                if not nsub[d]:
                    continue
                if tree.strides[d] == 1:
                    call("ngs_dual_accumulate_leaves", cg.struct(), ptr(leaves[d]), leaves[d].shape[0], nu, bases_[d],
                         ptr(values), ptr(acc), ptr(cnt), stream())
                else:
                    call("ngs_dual_accumulate", cg.struct(), ptr(primal[d]), primal[d].shape[0], tree.strides[d],
                         bases_[d], ptr(values), ptr(acc), ptr(cnt), stream())
        # complete cubes whose corner values change sign (the only ones with triangles), in centre order
        cubes = torch.empty((DC, 8), dtype=torch.int32, device=dev)
        flags = torch.empty(DC, dtype=torch.int32, device=dev)
        offs = torch.empty(DC + 1, dtype=torch.int32, device=dev)
        ws = Scratch.get(lib().ngs_scan_workspace_bytes(DC), dev)
        call("ngs_dual_compact", ptr(acc), ptr(cnt), 1, DC, ptr(cubes), ptr(flags), ptr(offs), ptr(ws), stream())
        K = int(offs[DC].item()) if DC else 0
        cubes = cubes[:K]
        v, f = self.dual_marching_cubes(cubes, pos, values)
        if return_info:
            return v, f, dict(leaves=leaves, primal=primal, centres=cg.coords, acc=acc, count=cnt, cubes=cubes, pos=pos,
                              values=values)
        if build_o3d_mesh:
            from .io import triangle_mesh
            return triangle_mesh(v, f)
        return v, f

    def _leaf_grid(self, d, lf, parents, lev_box, trim):
        """Bitmap of the level-d leaves (stride of level d): the children box of the level's parents
        (conformal leaves), or the level's own box (the last level, or trimmed leaves)."""
        tree = self.hash_tree
        s = tree.strides[d]
        if d in parents and not trim:
            po, pdm = parents[d].origin, parents[d].dims  # lattice units of stride 2s
            lo, hi = [2 * o for o in po], [2 * (o + dm - 1) + 1 for o, dm in zip(po, pdm)]
        else:
            lo, hi = lev_box[d]
        lg = SparseGrid(lo, hi, s, tree.device)
        lg.mark_coords(lf)
        return lg

    def leaf_corner_chi(self, leaves, bases_, nprim, nu, scale):
        """Raw chi at the dual corners of every leaf (the values evaluate_raw_chi gives at the
        positions ngs_dual_corner_positions writes, level by level in the same order)."""
        tree = self.hash_tree
        out = torch.empty(nprim, dtype=torch.float32, device=tree.device)
        levels = [(d, self.field(d)) for d, sol in self.solutions.items() if sol.numel()]
        if not levels:
            return out.zero_()
        garr = (Grid * len(levels))(*[tree.grids[d].struct() for d, _ in levels])
        farr = (ctypes.c_void_p * len(levels))(*[sol.data_ptr() for _, sol in levels])
        for dl, lf in leaves.items():
            if lf.shape[0] == 0:
                continue
            call("ngs_eval_leaf_corners_levels", len(levels), garr, farr, ptr(lf), lf.shape[0], tree.strides[dl], nu,
                 scale, tree.inv_vs, 0, ptr(out[bases_[dl]:]), stream())
        return out

    @staticmethod
    def dual_marching_cubes(cubes, pos, values):
        """dual_mc.cu + reconstructor.py:379-383: triangles of the complete dual cubes, vertices
        de-duplicated by their corner-pair key (sorted)."""
        dev = pos.device
        K = cubes.shape[0]
        if K == 0:
            return torch.zeros((0, 3), dtype=torch.float32, device=dev), torch.zeros((0, 3), dtype=torch.int64, device=dev)
        cubes = cubes.contiguous()
        values = values.float().contiguous()
        # per cube: triangle vertices and the distinct edges they lie on (one key per edge and cube
        # goes to the sort instead of one per triangle vertex)
        counts = torch.empty((2, K), dtype=torch.int32, device=dev)
        call("ngs_mc_count", ptr(cubes), K, ptr(values), ptr(counts[0]), ptr(counts[1]), stream())
        offs = torch.empty((2, K + 1), dtype=torch.int32, device=dev)
        ws = Scratch.get(lib().ngs_scan_workspace_bytes(K), dev)
        call("ngs_scan_i32", ptr(counts[0]), K, ptr(offs[0]), ptr(ws), stream())
        call("ngs_scan_i32", ptr(counts[1]), K, ptr(offs[1]), ptr(ws), stream())
        nv, ne = (int(v) for v in offs[:, K].cpu())  # one read for both totals
        if nv == 0:
            return torch.zeros((0, 3), dtype=torch.float32, device=dev), torch.zeros((0, 3), dtype=torch.int64, device=dev)
        # corner ids fit id_bits bits: keys are sorted on 2 id_bits bits
        id_bits = max(1, (values.shape[0] - 1).bit_length())
        keys = torch.empty(ne, dtype=torch.int64, device=dev)
        tags = torch.empty(ne, dtype=torch.int32, device=dev)
        faces = torch.empty(nv, dtype=torch.int64, device=dev)
        call("ngs_mc_emit", ptr(cubes), K, ptr(values), ptr(offs[0]), ptr(offs[1]), id_bits, ptr(keys), ptr(tags),
             ptr(faces), stream())
        slot_rank = torch.empty(ne, dtype=torch.int64, device=dev)
        verts = torch.empty((ne, 3), dtype=torch.float32, device=dev)
        nvert = torch.empty(1, dtype=torch.int64, device=dev)
        wsb = lib().ngs_dedup_workspace_bytes(ne)
        ws2 = Scratch.get(wsb, dev, slot=4)
        call("ngs_dedup_vertices", ptr(keys), ptr(tags), ne, id_bits, ptr(pos.contiguous()), ptr(values),
             ptr(slot_rank), ptr(verts), ptr(nvert), ptr(ws2), wsb, stream())
        del keys, tags
        call("ngs_mc_remap_faces", ptr(faces), nv, ptr(slot_rank), stream())
        V = int(nvert.item())
        return verts[:V], faces.view(-1, 3)
