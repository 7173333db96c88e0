/*
 * ngspsr C ABI: MI355X (gfx950) kernels of the jspsr screened-Poisson reconstruction path.
 *
 * Conventions
 *  - All pointers are device pointers owned by the caller unless stated otherwise; sizes are
 *    element counts. Every call takes the HIP stream to launch on (a hipStream_t, passed as
 *    ngs_stream_t) and returns an ngs_status; ngs_last_error() gives the text of the last error
 *    on the calling thread.
 *  - Voxel coordinates are int32 triples in finest-voxel units (multiples of the level stride),
 *    the layout of the reference's HashTree coords (jspsr/core/_hash_backend.py:121-147).
 *  - Calls do not synchronise the stream unless documented (table build retries, PCG solve,
 *    count reads). No call allocates device memory except ngs_pcg_solve's graph objects.
 *
 * The reference's operator layer is the pybind11 modules built by setup.py:17-54; each entry
 * point below names the reference function whose role it takes over.
 */
#ifndef NGSPSR_H
#define NGSPSR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *ngs_stream_t; /* hipStream_t */

typedef enum {
    NGS_OK = 0,
    NGS_ERR_ARG = 1,             /* bad argument (shape, size, null pointer) */
    NGS_ERR_HIP = 2,             /* HIP runtime or launch error */
    NGS_ERR_HASH_OVERFLOW = 3,   /* a cuckoo bucket received more keys than it holds */
    NGS_ERR_REHASH = 4,          /* cuckoo insertion failed for every attempted hash family */
    NGS_ERR_NOT_CONVERGED = 5,   /* PCG hit its iteration cap before the tolerance */
    NGS_ERR_CAPACITY = 6         /* a bounded structure (bitmap box, table) would be too large */
} ngs_status;

/* Ordered bitmap set of lattice points ("grid"): a box origin/dims in lattice units
 * (coordinate >> log2_stride), one bit per lattice point in x-major order, and the exclusive
 * popcount prefix per 32-bit word. Member rank = lexicographic (x, y, z) position. */
typedef struct {
    int32_t origin[3];
    int32_t dims[3];
    int32_t log2_stride;
    int32_t pad_;
    int64_t nwords;
    uint32_t *bits;  /* nwords */
    int32_t *rank;   /* nwords */
} ngs_grid;

/* Cuckoo hash table over 60-bit voxel keys: three sub-tables of table_size slots split into
 * 512-slot buckets (bucket = key % num_buckets), multiply-shift slot hash per sub-table. */
typedef struct {
    int64_t table_size;
    int32_t num_buckets;
    int32_t evict_bound;
    uint64_t *keys;      /* 3 * table_size, empty = UINT64_MAX */
    int64_t *vals;       /* 3 * table_size */
    uint64_t mult[3];    /* odd multipliers of the slot hash of each sub-table */
    int64_t inserted;
    int32_t attempts;    /* hash families tried by the last build */
    int32_t pad_;
} ngs_hash_table;

const char *ngs_last_error(void);
int ngs_abi_version(void);

/* ------------------------------------------------------------------------------- keys, hash */

/* Replaces _common.hash_cuda (csrc/common/hashmap_cuda.cu:269-305): keys[i] = fold60 FNV-1a-64
 * of the dim (2..4) uint32 words of row i. */
int ngs_hash_coords(const int32_t *coords, int64_t n, int32_t dim, uint64_t *keys, ngs_stream_t s);

/* Replaces _common.kernel_hash_cuda (hashmap_cuda.cu:307-339): out[k*n + i] = key of
 * coords[i] + offsets[k] (first three words offset, a 4th word copied). */
int ngs_kernel_hash(const int32_t *coords, int64_t n, int32_t dim, const int32_t *offsets, int32_t k,
                    uint64_t *out, ngs_stream_t s);

/* Table sizing of build_hash_table (hashmap_cuda.cu:341-360). */
int ngs_hash_sizing(int64_t n, int32_t enlarge, int64_t *table_size, int32_t *num_buckets,
                    int32_t *evict_bound);
/* Device workspace bytes for ngs_hash_build. */
int64_t ngs_hash_workspace_bytes(int64_t table_size);

/* Replaces _common.build_hash_table (hashmap_cuda.cu:341-379, insert kernels :13-107). Inserts
 * (keys[i], vals ? vals[i] : i). t->table_size/num_buckets/evict_bound/keys/vals must be set by
 * the caller (ngs_hash_sizing). Synchronises the stream to check for bucket overflow and
 * eviction failure; retries with new hash families derived from `seed`. */
int ngs_hash_build(ngs_hash_table *t, const uint64_t *keys, const int64_t *vals, int64_t n,
                   uint64_t seed, void *workspace, ngs_stream_t s);

/* Replaces _common.hash_table_query (hashmap_cuda.cu:381-398) with the CuckooHashTable.query
 * result convention (jspsr/ext.py:74-84): out[i] = value, or -1 when the key is absent. */
int ngs_hash_query(const ngs_hash_table *t, const uint64_t *queries, int64_t n, int64_t *out,
                   ngs_stream_t s);

/* ------------------------------------------------------------------- ordered bitmap set, scan */

/* Words needed by a grid with these dims (0 if the box exceeds the capacity limit). */
int64_t ngs_grid_words(const int32_t dims[3]);
/* Device workspace bytes of the exclusive scans used below (n = words or rows): the single-pass scan's
 * look-back state as the device asks for it, so it needs a current HIP device (-1 without one). */
int64_t ngs_scan_workspace_bytes(int64_t n);

int ngs_grid_clear(ngs_grid g, ngs_stream_t s);
/* Mark int32 coordinate rows (multiples of the grid stride). */
int ngs_grid_mark_coords(ngs_grid g, const int32_t *coords, int64_t n, ngs_stream_t s);
/* Mark c + (dx, dy, dz) * stride for every row c and offset in [-radius, radius]^3: the expansion of
 * get_coords(expand = 2 radius + 1) (_hash_backend.py:132-136) up to the sort, which the bitmap gives. */
int ngs_grid_mark_dilated(ngs_grid g, const int32_t *coords, int64_t n, int32_t radius, ngs_stream_t s);
/* Sample positions of get_test_grid (hashtree.py:209-229): out[(l R^3 + (i R + j) R + k)] =
 * (base[l] + (box_i, box_j, box_k)) * voxel_size with box = torch.linspace(0, stride, R). */
int ngs_test_grid_positions(const int32_t *base, int64_t n, int32_t stride, int32_t resolution, float voxel_size,
                            float *out, ngs_stream_t s);
/* Mark the 8 trilinear corners at the grid stride of every point whose level <= max_level
 * (levels may be NULL: all points). Replaces _quantize_coords (_hash_backend.py:457-462, with
 * _trilinear_weights :469-510) up to the sort, which the bitmap order provides. */
int ngs_grid_mark_corners(ngs_grid g, const float *xyz, int64_t n, float inv_voxel_size,
                          const int32_t *levels, int32_t max_level, ngs_stream_t s);
/* Byte-staged forms of ngs_grid_mark_coords and ngs_grid_mark_corners (same members, same reference
 * lines): each member is recorded by a plain byte store marks[lattice index] = 1, then the bytes are
 * ORed into the bitmap. marks: device buffer of nwords * 32 bytes, 16-byte aligned, all zero on entry
 * and left all zero on return (the calls that share one marks buffer must be stream-ordered). Many
 * points that share a coarse voxel then cost merged stores instead of serialised atomics. */
int ngs_grid_mark_coords_staged(ngs_grid g, const int32_t *coords, int64_t n, uint8_t *marks, ngs_stream_t s);
int ngs_grid_mark_corners_staged(ngs_grid g, const float *xyz, int64_t n, float inv_voxel_size,
                                 const int32_t *levels, int32_t max_level, uint8_t *marks, ngs_stream_t s);
/* Fill rank[] and write the member count to *count (device int64; NULL: ranks only). */
int ngs_grid_finalize(ngs_grid g, int64_t *count, void *workspace, ngs_stream_t s);
/* Write the members in rank order (the reference's torch.unique(dim=0) order). */
int ngs_grid_emit(ngs_grid g, int32_t *coords_out, ngs_stream_t s);
/* out[k*n + i] = rank of coords[i] + offsets[k], or -1 (offsets may be NULL, k = 1). Offsets are
 * added as given (pre-multiplied by the stride, as the reference's callers do).
 * The query of CuckooHashTable.query(coords, offsets) (ext.py:74-84) over a level's voxels. */
int ngs_grid_query(ngs_grid g, const int32_t *coords, int64_t n, const int32_t *offsets, int32_t k,
                   int64_t *out, ngs_stream_t s);
/* min (3) and max (3) of q = xyz * inv_voxel_size over all points, into minmax[6] (device float). */
int ngs_point_bounds(const float *xyz, int64_t n, float inv_voxel_size, float *minmax, ngs_stream_t s);
/* Cells (n, 3) int32 of points at a power-of-two stride: floor(q * scale / stride) * stride, with
 * q * scale in FP32. Replaces the floor divisions of build_hierarchy_dense and
 * build_hierarchy_subdivide (_hash_backend.py:363,396; scale = 1 / voxel_size or 1). */
int ngs_point_cells(const float *q, int64_t n, float scale, int32_t stride, int32_t *cells, ngs_stream_t s);
/* SparseFeatureHierarchy.evaluate_voxel_status (_hash_backend.py:292-318): per coordinate, 0 when
 * absent from grid g, 1 when present without children, 2 when present with children, i.e. when
 * the coordinate is in finer_parents (the parents, at g's stride, of the next finer level; read
 * only when has_finer). Output int64. */
int ngs_voxel_status(ngs_grid g, ngs_grid finer_parents, int32_t has_finer, const int32_t *coords, int64_t n,
                     int64_t *status, ngs_stream_t s);
/* SparseFeatureHierarchy.split_data (_hash_backend.py:320-330): trilinear interpolation of the
 * per-voxel rows data (M, nc) at the points (the splat's 8 corners and weights at g's stride;
 * absent corners add nothing), out (n, nc) float32. */
int ngs_split_data(ngs_grid g, const float *xyz, int64_t n, float inv_voxel_size, const float *data, int32_t nc,
                   float *out, ngs_stream_t s);
/* Exclusive scan of int32 counts into int32 offsets (n + 1 entries; offsets[n] = total). */
int ngs_scan_i32(const int32_t *counts, int64_t n, int32_t *offsets, void *workspace, ngs_stream_t s);

/* --------------------------------------------------------------------- density, levels, splat */

/* Trilinear splat (_hash_backend.py:332-356, splat_data with data = 1 or point rows):
 * acc[id * ncomp + c] += w * (data ? data[p*ncomp + c] * scale[p] : 1) * post, for points with
 * levels[p] == level (levels NULL: all), ids looked up in g. acc is float64. */
int ngs_splat(ngs_grid g, const float *xyz, int64_t n, float inv_voxel_size, const float *data,
              int32_t ncomp, const float *scale, const int32_t *levels, int32_t level, double post,
              double *acc, ngs_stream_t s);
/* The same splat as a gather (no FP64 atomics): points binned by the rank in g of the lower node of
 * their trilinear stencil (a radix sort), then one thread per node of g sums its 8 neighbouring bins
 * in FP64 with the same FP32 weights. node_coords: g's members in rank order (count x 3, as
 * ngs_grid_emit writes them). A point whose lower node is not in g is splatted with atomics. The FP64
 * sums of FP32 terms here are exact, so the result equals ngs_splat's. Workspace from
 * ngs_splat_gather_workspace_bytes(n, count) (-1: sizes beyond 2^31 or no device). */
int64_t ngs_splat_gather_workspace_bytes(int64_t n, int64_t count);
int ngs_splat_gather(ngs_grid g, const int32_t *node_coords, int64_t count, const float *xyz, int64_t n,
                     float inv_voxel_size, const float *data, int32_t ncomp, const float *scale,
                     const int32_t *levels, int32_t level, double post, double *acc, void *workspace,
                     int64_t workspace_bytes, ngs_stream_t s);
/* Basis-interpolated field at points (hashtree.py:157-192, evaluate_interpolated with the Bezier
 * basis): out[p] = (accumulate ? out[p] : 0) + scale * sum_i field[i] B((q_p - c_i)/s) over the 27
 * voxels of g around the cell of q_p = xyz_p * inv_voxel_size. grad_out (nullable, n x 3) receives
 * the derivative with respect to q (finest-voxel units), as evaluate_derivative returns it. */
int ngs_eval_field(ngs_grid g, const float *field, const float *xyz, int64_t n, float inv_voxel_size,
                   float scale, int32_t accumulate, float *out, float *grad_out, ngs_stream_t s);
/* ngs_eval_field with accumulate = 1 and no gradient, for the points binned at the field's level by
 * ngs_bin_points (point_cell, basis in bin order, perm): slot k adds to out[perm[k]]. Bit-identical to
 * ngs_eval_field on the same points; the bitmap and coefficient reads follow the bins. */
int ngs_eval_field_binned(ngs_grid g, const float *field, const int32_t *point_cell, const float *basis,
                          const int32_t *perm, int64_t n, float scale, float *out, ngs_stream_t s);
/* Support of one level's field at the queries (hashtree.py:186-189, compute_mask): mask[p] = 1 (uint8,
 * caller-zeroed; other entries untouched) when any voxel of g is among the 27 that ngs_eval_field sums. */
int ngs_field_support(ngs_grid g, const float *xyz, int64_t n, float inv_voxel_size, uint8_t *mask, ngs_stream_t s);
/* xyz_depth of build_hierarchy_adaptive (_hash_backend.py:428-429) from densities. */
int ngs_point_levels(const float *density, int64_t n, float min_density, float log_base,
                     int32_t depth, int32_t *levels, ngs_stream_t s);

/* ------------------------------------------------------------------------ point binning (screening) */

/* Per point: the cell (at the grid stride) containing it and the 3x3 per-axis basis values of
 * the 27 voxels around it (reconstructor.py:111-120 pre-evaluation): cell[3p..], b[9p..]
 * ordered (x: k=-1,0,1)(y)(z). */
int ngs_point_basis(const float *xyz, int64_t n, float inv_voxel_size, int32_t stride, int32_t *cell,
                    float *b, ngs_stream_t s);
/* Rank of each point's cell in the cell grid (int32). */
int ngs_grid_rank_rows(ngs_grid g, const int32_t *coords, int64_t n, int32_t *out, ngs_stream_t s);
/* Stable bucketing of points by cell rank: perm (points ordered by cell, ties by index) and
 * cell_start (ncells + 1). */
int64_t ngs_bucket_workspace_bytes(int64_t n, int64_t ncells);
int ngs_bucket_points(const int32_t *cell_rank, int64_t n, int64_t ncells, int32_t *perm,
                      int32_t *cell_start, void *workspace, int64_t workspace_bytes, ngs_stream_t s);

/* ----------------------------------------------------------------------------------- assembly */

/* Registers the device copy of the integral tables (jspsr/bases/bezier_tensor.py:27-34, the
 * tables the reference loads from bezier_integrals.pkl): for k = 0..levels-1 (m = 2^k) six float
 * arrays F, G (3m+2), P, Q (3m), Pinv, Qinv (m+2; unused for k = 0) at element offsets
 * offsets[6k + j] of `tables`. The buffer must outlive every assembly call. */
int ngs_set_tables(const float *tables, const int64_t *offsets, int32_t levels);

/* Everything the level-d row kernels need about one tree level. */
typedef struct {
    ngs_grid grid;           /* voxels of the level */
    const int32_t *coords;   /* M x 3 sorted */
    int64_t M;
    int32_t stride;
    int32_t pad_;
} ngs_level;

/* Points binned by their level-d cell, with per-point basis values at level d. Per-point arrays are
 * in bin order (bin k holds point perm[k]; ngs_bin_points writes them). */
typedef struct {
    ngs_grid cells;          /* cells containing >= 1 point */
    const int32_t *cell_start; /* ncells + 1 */
    const int32_t *perm;     /* n: original point id of each bin slot; NULL for the assembly calls
                              * means their per-point coarse field is given in bin order */
    const float *b;          /* n x 9 basis values at level d, bin order */
    const float *alpha;      /* n screening weights in bin order, or NULL with alpha_scalar */
    float alpha_scalar;
    float pad_;
    int64_t n;
} ngs_point_bins;

/* Stencil form of a level system (no reference counterpart: the reference always assembles the
 * matrix, reconstructor.py:170-183 + solver.py:40-48). The operator of ngs_lhs_assemble, applied
 * from the level bitmap, the 125 Laplacian coefficients and the binned screening points:
 *   (A x)_i = sum_o lap[o] x_{i+o} + sf * sum_{p near i} B_i(p) u_p,
 *   u_p     = alpha_p * sum_{j near p} B_j(p) x_j.
 * Point arrays are in bin order (ngs_bin_points). */
typedef struct {
    ngs_level level;           /* rows: the voxels of the level */
    ngs_grid cells;            /* cells holding >= 1 screening point (same stride) */
    const int32_t *cell_start; /* ncells + 1 */
    const int32_t *point_cell; /* n x 3 cell coordinates of the binned points */
    const float *basis;        /* n x 9 per-axis basis values of the binned points */
    const float *alpha;        /* n screening weights of the binned points, or NULL */
    const int32_t *perm;       /* n original point id of each binned point */
    const float *lap;          /* 125 Laplacian coefficients (ngs_stencil_laplacian) */
    float *u;                  /* n FP32 scratch of the point pass (PCG and ngs_stencil_apply) */
    float alpha_scalar;
    float sf;                  /* screening factor of the level (4^-d) */
    int64_t n;                 /* binned points (0: no screening) */
} ngs_stencil;

/* Entries per row of the level's 125-stencil (existing neighbours), rounded up to a multiple of
 * 4 (padding entries are explicit zeros on the diagonal column). */
int ngs_lhs_row_counts(ngs_level lv, int32_t *row_nnz, ngs_stream_t s);

/* Same-level system of level d (reconstructor.py:170-183 LHS; :143-151 screening RHS;
 * :153-168 screening part of the cross-level correction):
 *   A_ij = K(j - i) + sf * sum_p alpha_p B_i(p) B_j(p)  written as CSR in column order,
 *   inv_diag_i = 1 / A_ii,
 *   rhs_i += sf * delta * sum_p alpha_p B_i(p) - sf * sum_p alpha_p B_i(p) coarse_p
 * (coarse may be NULL: the field of coarser levels at the points). K uses the m = 1 tables. */
int ngs_lhs_assemble(ngs_level lv, ngs_point_bins pts, const float *coarse, float sf, float delta,
                     const int32_t *rowptr, int32_t *cols, float *vals, float *inv_diag, double *rhs,
                     ngs_stream_t s);

/* Assembly knob for measurement: lanes per row (1, 4, 16; 0 = automatic: 4 or 16 lanes split the
 * points of a row on levels with few rows and many points per row). Applies to ngs_lhs_assemble,
 * ngs_lhs_assemble_runs and ngs_stencil_diag_rhs alike, so their diagonals stay bit-identical. */
int ngs_assembly_config(int32_t lanes_per_row);

/* ngs_lhs_assemble into the run form of the same matrix instead of CSR (no reference counterpart;
 * the reference stores int32 columns, solver.py:40-48). Row i's neighbours in the bitmap column
 * (dx, dy), p = 5 (dx + 2) + (dy + 2), have consecutive ranks start .. start + n - 1 (n <= 5):
 * runs[p P + i] = start | n << 28 (n = 0: start = i), rvals[(5 p + j) P + i] = A(i, start + j) for
 * j < n, 0 otherwise; 25 int32 and 125 float per row, plane-major with plane stride
 * P = M rounded up to a multiple of 64 (arrays of 25 P and 125 P elements; rows M..P-1 unused).
 * inv_diag and rhs as ngs_lhs_assemble, bit for bit (the values equal the CSR entries). M <= 2^28. */
int ngs_lhs_assemble_runs(ngs_level lv, ngs_point_bins pts, const float *coarse, float sf, float delta,
                          int32_t *runs, float *rvals, float *inv_diag, double *rhs, ngs_stream_t s);

/* ngs_lhs_assemble_runs writing the values straight into the packed run form (ngs_runs_pack's
 * layout; soff from the slice sums of ngs_stencil_row_counts), for levels that assemble with one
 * lane per row: ngs_assembly_lanes(pts.n, M) == 1 (NGS_ERR_ARG otherwise). Same run words,
 * inv_diag, rhs and values, bit for bit. */
int32_t ngs_assembly_lanes(int64_t n_points, int64_t M);
int ngs_lhs_assemble_packed(ngs_level lv, ngs_point_bins pts, const float *coarse, float sf, float delta, int32_t *runs,
                            float *pvals, const int64_t *soff, float *inv_diag, double *rhs, ngs_stream_t s);

/* Stiffness part of the cross-level correction (reconstructor.py:153-168, COO of
 * get_neighbours(d, dd, 5), integrate_deriv_deriv_product bezier_tensor.py:75-96):
 * rhs_i -= sum_j K_{d,dd}(i, j) x_dd[j] over the 125 dd-voxels around the dd-cell of i. */
int ngs_rhs_cross_stiffness(ngs_level fine, ngs_level coarse, const float *x_coarse, double *rhs,
                            ngs_stream_t s);

/* Normal-field RHS (reconstructor.py:125-141, integrate_const_deriv_product
 * bezier_tensor.py:98-121): rhs_i += int grad B_i . (-N) over the data voxels of a data level.
 * Data at the same level is gathered (27 data voxels per target, no atomics); data at a finer
 * level is scattered (27 targets per data voxel, FP64 atomics); data at a coarser
 * level is gathered with the reference's |rel| <= 1.5 filter (_hash_backend.py:254-262). */
int ngs_rhs_normals(ngs_level target, ngs_level data, const float *normals, double *rhs, ngs_stream_t s);

/* The 125 same-level stiffness coefficients K(o), o in [-2,2]^3 lexicographic, of stride `stride`
 * (the K of ngs_lhs_assemble, computed the same way from the m = 1 tables). */
int ngs_stencil_laplacian(int32_t stride, float *lap, ngs_stream_t s);
/* dst[k] = src[idx[k]] (the coarse field in bin order, for the assembly with perm = NULL). */
int ngs_gather_f32(const float *src, const int32_t *idx, int64_t n, float *dst, ngs_stream_t s);
/* Bin-order copies of per-point data: point_cell[k] = cell[perm[k]] (n x 3), basis[k] = b[perm[k]]
 * (n x 9), alpha_out[k] = alpha[perm[k]] when alpha is not NULL. */
int ngs_bin_points(const int32_t *perm, int64_t n, const int32_t *cell, const float *b, const float *alpha,
                   int32_t *point_cell, float *basis, float *alpha_out, ngs_stream_t s);
/* inv_diag and the screening RHS terms of ngs_lhs_assemble for the stencil form (same values,
 * same summation order): inv_diag_i = 1 / A_ii, rhs_i += sf delta sum_p alpha_p B_i(p)
 * - sf sum_p alpha_p B_i(p) coarse_p (coarse in original point order, may be NULL). */
int ngs_stencil_diag_rhs(ngs_stencil op, const float *coarse, float delta, float *inv_diag, double *rhs,
                         ngs_stream_t s);
/* Existing neighbours in the 125-stencil of each row (the entries of the equivalent CSR, unpadded). */
int ngs_stencil_row_counts(ngs_level lv, int32_t *row_nnz, ngs_stream_t s);
/* y = A x (FP32): point pass into op.u, then the row pass. */
int ngs_stencil_apply(ngs_stencil op, const float *x, float *y, ngs_stream_t s);
/* The passes of ngs_stencil_apply separately, for measurement: passes bit 0 = point pass
 * (op.u from x), bit 1 = row pass (y from x and the current op.u). */
int ngs_stencil_pass(ngs_stencil op, const float *x, float *y, int32_t passes, ngs_stream_t s);
/* r = b - A x in FP64 (the operator evaluated in FP64 from the same FP32 data; x NULL = 0),
 * ||r||^2 and ||b||^2 into norms[0..1]; u64: n doubles of scratch; workspace:
 * ngs_pcg_workspace_bytes(M). */
int ngs_stencil_residual_f64(ngs_stencil op, const double *b, const double *x, double *r, double *u64,
                             double *norms, void *workspace, ngs_stream_t s);

/* ----------------------------------------------------------------------- adjoint (backward) */

/* Transpose of ngs_eval_field (the adjoint of evaluate_interpolated, hashtree.py:157-192):
 * out[i] += scale * sum_p w[p] B_i(p) over the 27 voxels of g around every point. */
int ngs_basis_splat(ngs_grid g, const float *w, const float *xyz, int64_t n, float inv_voxel_size, double scale,
                    double *out, ngs_stream_t s);
/* Transpose of the cross-level stiffness of ngs_rhs_cross_stiffness:
 * out_coarse[j] += scale * sum_i K_{d,dd}(i, j) g_fine[i]. The backward of the cascadic RHS
 * correction (reconstructor.py:153-168) through Solver.backward (solver.py:91-98). */
int ngs_cross_stiffness_transpose(ngs_level fine, ngs_level coarse, const float *g_fine, double scale,
                                  double *out_coarse, ngs_stream_t s);

/* ----------------------------------------------------------------------------------- solver */

/* Replaces _sparse_solve.solve_pcg_diag (csrc/sparse_solve/solve_pcg.cu:15-136): Jacobi-PCG from
 * x = 0 with FP32 vectors and FP64 dot products; stops when ||r|| <= tol ||b|| (recurrence
 * residual) or after max_iter iterations (< 0: 10^7). Iterations run in captured hipGraph chunks
 * of `chunk` iterations; the stream is synchronised between chunks. rowptr/cols/vals: CSR as
 * written by ngs_lhs_assemble (rows padded to multiples of 4). On return *iters holds the
 * iteration count and *rel_resid the final recurrence residual ||r||/||b||. */
int64_t ngs_pcg_workspace_bytes(int64_t M);
int ngs_pcg_solve(const int32_t *rowptr, const int32_t *cols, const float *vals, const float *inv_diag,
                  const float *b, float *x, int64_t M, double tol, int32_t max_iter, int32_t chunk,
                  void *workspace, ngs_stream_t s, int32_t *iters, double *rel_resid);

/* ngs_pcg_solve on the run form of the matrix (ngs_lhs_assemble_runs): same iteration and stop rule;
 * the SpMV reads one run word per bitmap column instead of one column index per entry. */
int ngs_pcg_solve_runs(const int32_t *runs, const float *rvals, const float *inv_diag, const float *b, float *x,
                       int64_t M, double tol, int32_t max_iter, int32_t chunk, void *workspace, ngs_stream_t s,
                       int32_t *iters, double *rel_resid);

/* ngs_pcg_solve_runs without a workspace argument, for callers that cannot allocate device memory
 * (code under stream capture, tight serving loops, harnesses that forbid hipMalloc). No reference
 * counterpart. The workspace is a library-owned device array of NGS_PCG_NOALLOC_SCRATCH_BYTES (one
 * copy per GPU, held until the process ends); systems above NGS_PCG_NOALLOC_MAX_ROWS rows are
 * rejected with NGS_ERR_CAPACITY. The call allocates nothing: kernels are launched directly on s
 * (no graph, no private stream or event) and the stream is synchronised between chunks. Same
 * iteration, stop rule and results as ngs_pcg_solve_runs. Every call on one GPU shares that GPU's
 * array, so calls that overlap on one GPU (other threads or streams) race; serialise them. The
 * stream must belong to the current device. */
#define NGS_PCG_NOALLOC_SCRATCH_BYTES (32ll << 20)
#define NGS_PCG_NOALLOC_MAX_ROWS 2790144ll
int ngs_pcg_solve_runs_noalloc(const int32_t *runs, const float *rvals, const float *inv_diag, const float *b,
                               float *x, int64_t M, double tol, int32_t max_iter, int32_t chunk, ngs_stream_t s,
                               int32_t *iters, double *rel_resid);

/* ngs_pcg_solve_runs on the packed run form (ngs_runs_pack): the same iteration and results, bit
 * for bit; the SpMV streams the values without the run form's zero slots. */
int ngs_pcg_solve_packed(const int32_t *runs, const float *pvals, const int64_t *soff, const float *inv_diag,
                         const float *b, float *x, int64_t M, double tol, int32_t max_iter, int32_t chunk,
                         void *workspace, ngs_stream_t s, int32_t *iters, double *rel_resid);

/* ngs_pcg_solve on the stencil form of the system (same iteration, same stop rule; the SpMV is the
 * point pass + row pass of ngs_stencil_apply). */
int ngs_pcg_solve_stencil(ngs_stencil op, const float *inv_diag, const float *b, float *x, double tol,
                          int32_t max_iter, int32_t chunk, void *workspace, ngs_stream_t s, int32_t *iters,
                          double *rel_resid);

/* Solver knobs: use_graph (capture iteration chunks in a hipGraph; default 1), lanes per CSR row
 * in the SpMV (8, 16, 32; 0 = 8), non-temporal loads of the matrix stream (default 0) and
 * interleaved row tiles (blocks of one XCD sweep the XCD's row slab together; default 1). */
int ngs_pcg_config(int32_t use_graph, int32_t lanes_per_row, int32_t nontemporal, int32_t interleave);

/* Batched solves (continuous batching of independent level systems; no reference counterpart —
 * the reference solves one system per call, solve_pcg.cu:15-136). A device table of
 * ngs_pcg_batch_entry_bytes()-sized entries lists the active systems; blk2sys maps every block of
 * the launch to its system (-1: idle). Each system keeps its own Krylov state in its workspace
 * (ngs_pcg_workspace_bytes(M)) and occupies ngs_pcg_batch_blocks(M) consecutive blocks starting
 * at a multiple of 8. */
int64_t ngs_pcg_batch_entry_bytes(void);
int32_t ngs_pcg_batch_blocks(int64_t M);
/* Initialise a system (x = 0, r = b, p = D^-1 b) on stream s and write its host table entry. */
int ngs_pcg_batch_prepare(const int32_t *rowptr, const int32_t *cols, const float *vals, const float *inv_diag,
                          const float *b, float *x, int64_t M, double tol, int32_t max_iter, void *workspace,
                          int32_t blk0, int32_t nblk, void *entry, ngs_stream_t s);
/* Set the first block of a prepared host table entry (a multiple of 8). */
int ngs_pcg_batch_place(void *entry, int32_t blk0);
/* The same as ngs_pcg_batch_prepare for a system in run form. */
int ngs_pcg_batch_prepare_runs(const int32_t *runs, const float *rvals, const float *inv_diag, const float *b,
                               float *x, int64_t M, double tol, int32_t max_iter, void *workspace, int32_t blk0,
                               int32_t nblk, void *entry, ngs_stream_t s);
/* The same as ngs_pcg_batch_prepare for a system in packed run form. */
int ngs_pcg_batch_prepare_packed(const int32_t *runs, const float *pvals, const int64_t *soff, const float *inv_diag,
                                 const float *b, float *x, int64_t M, double tol, int32_t max_iter, void *workspace,
                                 int32_t blk0, int32_t nblk, void *entry, ngs_stream_t s);
/* The same as ngs_pcg_batch_prepare for a system in stencil form. */
int ngs_pcg_batch_prepare_stencil(ngs_stencil op, const float *inv_diag, const float *b, float *x, double tol,
                                  int32_t max_iter, void *workspace, int32_t blk0, int32_t nblk, void *entry,
                                  ngs_stream_t s);
/* `iters` (even) iterations of every active system in one launch per kernel (graph-replayed when
 * use_graph). flags: bit 0 the table may hold stencil systems (adds their point and row passes),
 * bit 1 it may hold run-form systems (adds their SpMV), bit 2 it holds no CSR systems (skips
 * that SpMV), bit 3 it may hold packed-form systems (adds their SpMV). */
int ngs_pcg_batch_iterate(const void *table, const int32_t *blk2sys, int32_t total_blocks, int32_t iters,
                          int32_t use_graph, int32_t flags, ngs_stream_t s);
/* (iter, done) of each system of the host copy of the table into states (2 nsys int32, host),
 * gathered on the device into dev_states (2 nsys int32, device scratch) and copied once;
 * synchronises s. done = 1 converged, 2 iteration cap. */
int ngs_pcg_batch_status(const void *host_table, int32_t nsys, int32_t *dev_states, int32_t *states,
                         ngs_stream_t s);

/* y = A x (FP32) on the run form (terms in increasing column order), exposed for measurement. */
int ngs_spmv_runs(const int32_t *runs, const float *rvals, const float *x, float *y, int64_t M, ngs_stream_t s);
/* Packed run form (no reference counterpart): the run form without its zero value slots. Rows are
 * cut into slices of 64; slice s keeps, plane by plane (p = 0..24), the values of its rows' runs
 * back to back in row order from pvals[soff[s]], with the run words of the run form.
 * ngs_runs_slice_counts writes the number of values of every slice ((M + 63) / 64 int64); soff is
 * its exclusive scan (one more entry, the total); ngs_runs_pack copies the run form's values
 * there and ngs_runs_unpack back (slots past a run get 0; rows M..P-1 are not written);
 * ngs_spmv_packed is y = A x with the same sums as ngs_spmv_runs, bit for bit. */
int ngs_runs_slice_counts(const int32_t *runs, int64_t M, int64_t *counts, ngs_stream_t s);
int ngs_runs_pack(const int32_t *runs, const float *rvals, int64_t M, const int64_t *soff, float *pvals,
                  ngs_stream_t s);
int ngs_runs_unpack(const int32_t *runs, const float *pvals, int64_t M, const int64_t *soff, float *rvals,
                    ngs_stream_t s);
int ngs_spmv_packed(const int32_t *runs, const float *pvals, const int64_t *soff, const float *x, float *y, int64_t M,
                    ngs_stream_t s);

/* y = A x (FP32), the SpMV of the solver, exposed for measurement. */
int ngs_spmv(const int32_t *rowptr, const int32_t *cols, const float *vals, const float *x, float *y,
             int64_t M, ngs_stream_t s);

/* r = b - A x in FP64 (x FP64; NULL = 0, no matrix pass), and ||r||^2, ||b||^2 into norms[0..1]
 * (device double). */
int ngs_residual_f64(const int32_t *rowptr, const int32_t *cols, const float *vals, const double *b,
                     const double *x, double *r, int64_t M, double *norms, void *workspace, ngs_stream_t s);
/* The same on the run form. */
int ngs_residual_f64_runs(const int32_t *runs, const float *rvals, const double *b, const double *x, double *r,
                          int64_t M, double *norms, void *workspace, ngs_stream_t s);
/* The same on the packed run form (the sums of ngs_residual_f64_runs, bit for bit). */
int ngs_residual_f64_packed(const int32_t *runs, const float *pvals, const int64_t *soff, const double *b,
                            const double *x, double *r, int64_t M, double *norms, void *workspace, ngs_stream_t s);

/* Sparse direct solve (replaces _sparse_solve.solve_cusparse, csrc/sparse_solve/solve_cusparse.cu:15-42,
 * the sparse Cholesky behind the "cholmod" solver, jspsr/core/solver.py:147-157): a supernodal
 * multifrontal Cholesky factorization in FP64 of a symmetric positive definite CSR matrix (its
 * lower triangle in the caller's numbering is read, as cuSOLVER reads it, so an upper triangle
 * that differs by rounding is ignored; padded rows repeating their diagonal with value 0 are
 * accepted).
 * ngs_chol_analyze reads the pattern (device arrays, read on stream s, which it synchronises) and
 * computes on the host a
 * nested-dissection order, the supernodes and the front structures; the opaque handle owns only host
 * memory (ngs_chol_free). The device workspace of ngs_chol_workspace_bytes holds the factor and
 * index maps; ngs_chol_factor uploads the maps into it (once per workspace) and factors on the
 * stream. A pivot at or below tol times its column's diagonal marks the matrix singular: the
 * original row index of the first such pivot is stored in *singularity (-1 when there is none, as
 * cuSOLVER reports) and the pivot is replaced by the diagonal so that the factorization completes.
 * ngs_chol_solve solves A x = b with the factor of that workspace (b, x: FP32 or FP64, one of each
 * non-NULL). stats of ngs_chol_info (8 int64): supernodes, tree heights, largest front, entries of
 * L, pool doubles, index-list entries, parent-map entries, CSR entries. ngs_chol_permutation: the
 * order (new -> original row). */
typedef struct ngs_chol_s ngs_chol;
int ngs_chol_analyze(const int32_t *rowptr, const int32_t *cols, int64_t M, ngs_stream_t s, ngs_chol **handle);
/* The same analysis from a pattern in host memory. */
int ngs_chol_analyze_host(const int32_t *rowptr, const int32_t *cols, int64_t M, ngs_chol **handle);
/* The symbolic structure, for inspection: per supernode 8 int64 (first pivot column, pivot
 * columns, front size, parent, pool offset, index-list offset, parent-map offset, height), the
 * front index lists, the child-to-parent position maps and the CSR-entry-to-pool map (sizes from
 * ngs_chol_info stats[0], [5], [6], [7]). */
int ngs_chol_structure(const ngs_chol *handle, int64_t *nodes, int32_t *front_idx, int32_t *cmap, int64_t *amap);
int64_t ngs_chol_workspace_bytes(const ngs_chol *handle);
int ngs_chol_info(const ngs_chol *handle, int64_t *stats);
int ngs_chol_permutation(const ngs_chol *handle, int32_t *perm);
int ngs_chol_factor(ngs_chol *handle, const int32_t *rowptr, const int32_t *cols, const float *vals, double tol,
                    void *workspace, ngs_stream_t s, int64_t *singularity);
int ngs_chol_solve(const ngs_chol *handle, const void *workspace, const float *b, const double *b64, float *x,
                   double *x64, ngs_stream_t s);
void ngs_chol_free(ngs_chol *handle);

/* ------------------------------------------------------------------------------------- mesh */

/* Conformal leaf construction (reconstructor.py:314-336 with get_coords(conforming=True),
 * _hash_backend.py:139-145): the 8 children (CONFORM order) of each sorted parent, and keep[] = 1
 * unless the child's coordinate is a parent of the next finer level (finer_parents, tested only
 * when has_finer). */
int ngs_conformal_children(const int32_t *parents, int64_t np, int32_t stride, ngs_grid finer_parents,
                           int32_t has_finer, int32_t *children, int32_t *keep, ngs_stream_t s);
/* keep[i] = 1 unless coords[i] is a parent of the next finer level (leaf test of the coarsest level). */
int ngs_unrefined_flags(const int32_t *coords, int64_t n, ngs_grid finer_parents, int32_t *keep, ngs_stream_t s);
/* Ordered compaction of int32 rows of `width` columns with keep[] flags; offsets (n + 1) receives
 * the exclusive scan of keep, offsets[n] = kept count. */
int ngs_compact_rows(const int32_t *rows, int32_t width, const int32_t *keep, int64_t n, int32_t *out,
                     int32_t *offsets, void *workspace, ngs_stream_t s);
/* Sub-cell expansion (reconstructor.py:338-348): out[(l*n^3 + e)] = leaf*n + e*stride. */
int ngs_expand_primal(const int32_t *leaves, int64_t n, int32_t n_upsample, int32_t stride, int32_t *out,
                      ngs_stream_t s);
/* Mark the 8 corners of every primal sub-cell (reconstructor.py:350-358). */
int ngs_mark_dual_centres(ngs_grid g, const int32_t *primal, int64_t n, int32_t stride, ngs_stream_t s);
/* Dual-cube assembly (csrc/marching_cubes/dual_mc_inds.cu:45-182): for each primal sub-cell and
 * each lattice point on its boundary that is a dual centre, register id_base + cell in the octant
 * slots the cell covers (the reference also records the level, which nothing downstream reads);
 * acc (DC x 8) int32, count (DC) int32 (caller-zeroed).
 * count bits 0..15: slots registered; bits 16..31: slots whose cell has a negative value, when
 * values (the field at the primal cells' dual corners, indexed by id) is not NULL. The centre
 * grid has stride 1. */
int ngs_dual_accumulate(ngs_grid centres, const int32_t *primal, int64_t n, int32_t stride, int32_t id_base,
                        const float *values, int32_t *acc, int32_t *count, ngs_stream_t s);
/* Keep complete dual cubes (8 slots registered) in order: cubes (K x 8) corner ids; flags (dc)
 * and offsets (dc + 1, offsets[dc] = K) are written on the way. active_only: keep only complete
 * cubes whose corner values change sign (needs counts accumulated with values; the others yield
 * no triangles). acc is read only for the kept centres. */
int ngs_dual_compact(const int32_t *acc, const int32_t *count, int32_t active_only, int64_t dc, int32_t *cubes,
                     int32_t *flags, int32_t *offsets, void *workspace, ngs_stream_t s);
/* Leaf forms of the two calls above (same sets, fewer atomics): the dual centres of all sub-cells
 * of each leaf (lattice leaf*n + o*stride, o in [0, n]^3), and the dual-cube assembly of stride-1
 * leaves (one count update per leaf and lattice point; identical to the per-sub-cell form whenever no
 * dual centre is over-registered). */
int ngs_mark_dual_leaves(ngs_grid g, const int32_t *leaves, int64_t n, int32_t stride, int32_t n_upsample,
                         ngs_stream_t s);
/* The dual centres of stride-1 leaves for n_upsample 2 (the set ngs_mark_dual_leaves marks), written
 * one bitmap word at a time with no atomics: leaf_bitmap holds the leaves (stride-1 grid), every word
 * of centres is overwritten, so call it before marking other levels' centres. */
int ngs_mark_dual_leaf_words(ngs_grid centres, ngs_grid leaf_bitmap, ngs_stream_t s);
/* The dual-cube assembly of all leaf levels for n_upsample 2, gathered per dual centre with no atomics
 * (same slots and counts as the per-sub-cell accumulation when the leaves partition space, which the
 * conformal leaves do): centre_coords (dc x 3, the centres in rank order, upsampled units); per level,
 * fine to coarse (at most 8), a finalized bitmap of its leaves, a device map leaf_index[l][rank in
 * that bitmap] = the leaf's position in the level's leaf list, and id_base[l] (ids id_base + 8
 * position + sub-cell). leaf_bitmaps, leaf_index and id_base are host arrays. count: slots in bits
 * 0..15, slots with a negative value in bits 16..31 (values nullable). The 8 slot ids are written,
 * and the negative slots counted, for complete centres only (8 slots; the only ones
 * ngs_dual_compact reads); the acc rows of the other centres are left as they were. */
int ngs_dual_gather_cubes(const int32_t *centre_coords, int64_t dc, int32_t nlevels, const ngs_grid *leaf_bitmaps,
                          const int32_t *const *leaf_index, const int32_t *id_base, const float *values, int32_t *acc,
                          int32_t *count, ngs_stream_t s);
int ngs_dual_accumulate_leaves(ngs_grid centres, const int32_t *leaves, int64_t n, int32_t n_upsample,
                               int32_t id_base, const float *values, int32_t *acc, int32_t *count, ngs_stream_t s);
/* Dual-corner positions (primal + stride) * voxel_size / n_upsample (reconstructor.py:368-373). */
int ngs_dual_corner_positions(const int32_t *primal, int64_t n, int32_t stride, float scale, float *pos,
                              ngs_stream_t s);
/* The positions of ngs_dual_corner_positions for the sub-cells ngs_expand_primal would list (leaf-major,
 * nu^3 per leaf), computed from the leaves directly, bit for bit. */
int ngs_leaf_corner_positions(const int32_t *leaves, int64_t n, int32_t stride, int32_t n_upsample, float scale,
                              float *pos, ngs_stream_t s);
/* chi of one level at the dual corners of leaves (reconstructor.py:368-375, evaluate_chi at the
 * corner positions): out[l * nu^3 + e] = (accumulate ? out : 0) + the ngs_eval_field value (unit
 * scale) at the position ngs_dual_corner_positions writes for corner e of leaf l, bit for bit; for
 * n_upsample 2 the 27-voxel windows of a leaf's corners are read together. leaf_stride: stride of
 * the leaves' level. */
/* ngs_eval_leaf_corners for up to 8 field levels in one pass (grids, fields: host arrays of nlevels),
 * summed in the given order: out = ((out or 0) + level 0 value) + level 1 value ..., the values of
 * nlevels successive ngs_eval_leaf_corners calls. */
int ngs_eval_leaf_corners_levels(int32_t nlevels, const ngs_grid *grids, const float *const *fields, const int32_t *leaves,
                                 int64_t n, int32_t leaf_stride, int32_t n_upsample, float scale, float inv_voxel_size,
                                 int32_t accumulate, float *out, ngs_stream_t s);
int ngs_eval_leaf_corners(ngs_grid g, const float *field, const int32_t *leaves, int64_t n, int32_t leaf_stride,
                          int32_t n_upsample, float scale, float inv_voxel_size, int32_t accumulate, float *out,
                          ngs_stream_t s);
/* Dual marching cubes (csrc/marching_cubes/dual_mc.cu:33-189): per cube its triangle-vertex count
 * (vcount) and, when ecount is not NULL, the number of distinct cube edges its triangles use; then,
 * from the exclusive scans voffset and eoffset, per distinct edge of each cube (in the order the
 * case's triangles first use them) its key, larger corner id << id_bits | smaller id (ids <
 * 2^id_bits), and a tag, slot index | (the interpolation's first corner is the larger id) << 31,
 * and per triangle vertex (T x 3, cube order then case-table order) the slot of its edge (vslot).
 * ngs_dedup_vertices on the slots gives the rank of each slot; ngs_mc_remap_faces turns vslot
 * into the faces, in place. */
int ngs_mc_count(const int32_t *cubes, int64_t k, const float *values, int32_t *vcount, int32_t *ecount,
                 ngs_stream_t s);
int ngs_mc_emit(const int32_t *cubes, int64_t k, const float *values, const int32_t *voffset, const int32_t *eoffset,
                int32_t id_bits, uint64_t *ekeys, uint32_t *etags, int64_t *vslot, ngs_stream_t s);
/* faces[v] = slot_rank[faces[v]] for v < n (no counterpart in the reference: the second half of its
 * per-vertex de-duplication, reconstructor.py:379-383, once the keys are de-duplicated per slot). */
int ngs_mc_remap_faces(int64_t *faces, int64_t n, const int64_t *slot_rank, ngs_stream_t s);
/* ---- Primal sparse marching cubes (marching_cubes_sparse_cuda, primal_mc.cu:176-234) ----
 * Input: nleaf leaves, each with a base corner base[3l..3l+2] (integer-valued floats, in leaf units)
 * and an (r+1)^3 block of samples sdf[l (r+1)^3 + (x (r+1) + y)(r+1) + z]. Sub-cube (rx, ry, rz) of a
 * leaf spans samples rx..rx+1, ry..ry+1, rz..rz+1; sub-cubes are numbered (l, rx, ry, rz) z-fastest. */
/* vcount[c] = 3 x triangles of sub-cube c (classify_voxels, primal_mc.cu:21-56). */
int ngs_primal_mc_count(const float *sdf, int64_t nleaf, int32_t r, int32_t *vcount, ngs_stream_t s);
/* Triangles of every sub-cube from vertex voffset[c] on (the exclusive scan of vcount), per triangle
 * vertex v: tris[3v..] position (base + offset / r), vert_ids[4v..] = (global sub-lattice corner of
 * the edge's lower end, edge axis), src_ids[2v..] = flat sample indices of the edge's ends in
 * interpolation order, src_weight[v] = weight of the first end (meshing_cube, primal_mc.cu:59-174). */
int ngs_primal_mc_emit(const float *base, const float *sdf, int64_t nleaf, int32_t r, const int32_t *voffset,
                       float *tris, int32_t *vert_ids, int32_t *src_ids, float *src_weight, ngs_stream_t s);
/* Vertex de-duplication of MarchingCubes.forward (ops.py:15-35): the distinct vert_ids rows in
 * lexicographic order (unq, V x 4), faces[v] = row of triangle vertex v, and per distinct vertex the
 * position, the src ids and the weight of its first occurrence in triangle order (the reference keeps
 * an arbitrary one), plus compute_normal (ops.py:85-101) in normals. key_layout (host, 6 int32) =
 * per-axis minimum of the vert_ids coordinates and per-axis bit widths of their span (sum + 2 <= 64).
 * Outputs are sized n (V <= n); *nverts (device int64) = V. */
int64_t ngs_primal_dedup_workspace_bytes(int64_t n);
int ngs_primal_dedup(const int32_t *vert_ids, int64_t n, const int32_t *key_layout, const float *tris,
                     const int32_t *src_ids, const float *src_weight, int64_t *faces, float *verts, float *normals,
                     int32_t *unq, int32_t *vsrc, float *vw, int64_t *nverts, void *workspace, int64_t workspace_bytes,
                     ngs_stream_t s);
/* MarchingCubes.backward (ops.py:40-58): g = bandwidth (grad_vert . normal) per vertex, added as w g to
 * sample vsrc[2v] and (1 - w) g to sample vsrc[2v+1] of grad_grid (caller-zeroed, atomics). */
int ngs_primal_mc_backward(const float *grad_vert, const float *normals, const int32_t *vsrc, const float *vw,
                           int64_t nverts, float bandwidth, float *grad_grid, ngs_stream_t s);
/* ---- Learned additive bases (jspsr/bases/additive.py) ----
 * DDI forward (csrc/integration/ddi.cu:21-76): for pair n and axis v, p_mats[n, v] = sum_si sum_ti
 * s[n, 3 si + v] t[n, 3 ti + v] f_f_i[si][ti][rel_pos[n, v]] and q_mats likewise with df_df_i.
 * s, t are (n, 3 C) float32, rel_pos (n, 3) int64 in [0, table_len), tables (C, C, table_len). */
int ngs_ddi_forward(const float *source_expanded, const float *target_expanded, const int64_t *rel_pos, int64_t n,
                    int32_t n_components, const float *f_f_i, const float *df_df_i, int32_t table_len, float *p_mats,
                    float *q_mats, ngs_stream_t s);
/* DDI backward (ddi.cu:78-141): grad_s, grad_t (n, 3 C) float32, written (not accumulated). */
int ngs_ddi_backward(const float *source_expanded, const float *target_expanded, const int64_t *rel_pos, int64_t n,
                     int32_t n_components, const float *f_f_i, const float *df_df_i, int32_t table_len,
                     const float *grad_p, const float *grad_q, float *grad_s, float *grad_t, ngs_stream_t s);
/* EvalFunc / EvalDerivFunc forward (additive.py:71-122): funcs[n, v] = sum_c q_c(xyz[n, v])
 * coeffs[feat_ids[n], 3 c + v], dfuncs (nullable) the same with q_c' / stride. components: 6 floats
 * per component, two factors (kind, a0, a1): kind 0 = 1, 1 = x^a0, 2 = Bezier, 3 = sin(a0 x + a1). */
int ngs_additive_eval(const float *xyz, const float *coeffs, const int64_t *feat_ids, int64_t n,
                      const float *components, int32_t n_components, float stride, float *funcs, float *dfuncs,
                      ngs_stream_t s);
/* Their backward (additive.py:86-99,124-136): grad_coeffs[feat_ids[n], 3 c + v] += grad_funcs q_c +
 * grad_dfuncs q_c' / stride (grad_dfuncs nullable; grad_coeffs caller-zeroed, atomics). */
int ngs_additive_eval_backward(const float *xyz, const int64_t *feat_ids, int64_t n, const float *components,
                               int32_t n_components, float stride, const float *grad_funcs, const float *grad_dfuncs,
                               float *grad_coeffs, ngs_stream_t s);
/* Replaces _sparse_op.screened_multiplication (csrc/sparse_op/sparse_product.cu:202-258, smult_kernel
 * :100-137) for bases given by per-(point, voxel) values: out[k] += (va * vb) * w[p] (w NULL:
 * w_scalar) for each point p, each entry (a_ids[t], a_vals[t]) with a_pt[t] = p of level a and each
 * entry of p at level b (b_start: n_points + 1 offsets into b_ids, b_vals), k = position of column
 * b_id in row a_id of the pair structure (rowptr, cols ascending within rows); pairs outside the
 * structure are dropped. FP32 atomics into out (caller-zeroed or accumulating). */
int ngs_screen_pairs(int64_t na, const int32_t *a_pt, const int32_t *a_ids, const float *a_vals,
                     const int32_t *b_start, const int32_t *b_ids, const float *b_vals, const float *w,
                     float w_scalar, const int32_t *rowptr, const int32_t *cols, float *out, ngs_stream_t s);
/* Vertex de-duplication (reconstructor.py:379-383): sorted unique keys; faces[tag index] = rank of
 * the element's key (per triangle vertex, or per slot of ngs_mc_emit); the vertex of a key is interpolated (dual_mc.cu:20-31 sdf_interp) on its edge, in the corner
 * order of the key's first occurrence, from the corner positions pos (3 per id) and values — the
 * position the reference's per-triangle vertex has at that occurrence. *nverts (device int64).
 * The keys are radix-sorted on their 2 id_bits low bits. */
int64_t ngs_dedup_workspace_bytes(int64_t nkeys);
int ngs_dedup_vertices(const uint64_t *vkeys, const uint32_t *vtags, int64_t nkeys, int32_t id_bits, const float *pos,
                       const float *values, int64_t *faces, float *verts, int64_t *nverts, void *workspace,
                       int64_t workspace_bytes, ngs_stream_t s);

/* ------------------------------------------------------------------------------------ misc */
int ngs_f64_to_f32(const double *in, float *out, int64_t n, ngs_stream_t s);
/* Weighted sum of field values (FP64) for the level set (reconstructor.py:228-236):
 * sums[0] = sum w_p v_p, sums[1] = sum w_p. */
int ngs_weighted_sum(const float *v, const float *w, int64_t n, double *sums, ngs_stream_t s);

#ifdef __cplusplus
}
#endif
#endif /* NGSPSR_H */
