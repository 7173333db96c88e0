// Galerkin assembly of the screened Poisson system on the sparse voxel hierarchy.
//
// Row-oriented, race-free design: every row of level d is built by one thread from the points in
// the 27 level-d cells that support its basis function (points are pre-binned by cell), with the
// 125-entry stencil row held in registers. This replaces the reference's per-point 27x27 pair
// loop with a hash lookup and an FP32 atomic per matrix entry (csrc/sparse_op/sparse_product.cu:100-137)
// and the enlarged pair hash table (jspsr/core/reconstructor.py:71).
#include "common.h"

namespace ngs {

// Integral tables of one multiple m (device pointers, layout of the reference tables,
// jspsr/bases/bezier_tensor.py:75-121): F/G length 3m+2, P/Q length 3m, Pinv/Qinv length m+2.
struct Tables {
    const float* F;
    const float* G;
    const float* P;
    const float* Q;
    const float* Pinv;
    const float* Qinv;
    int m;
};

template <bool PAD>
__global__ void lhs_count_kernel(ngs_level lv, int32_t* __restrict__ row_nnz) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= lv.M) return;
    const ngs_grid& g = lv.grid;
    const int l2 = g.log2_stride;
    const int lx = (lv.coords[3 * i] >> l2) - g.origin[0], ly = (lv.coords[3 * i + 1] >> l2) - g.origin[1];
    const int lz = (lv.coords[3 * i + 2] >> l2) - g.origin[2];
    int cnt = 0;
    for (int dx = -2; dx <= 2; ++dx)
        for (int dy = -2; dy <= 2; ++dy) {
            int32_t r;
            cnt += __popc(grid_column(g, lx + dx, ly + dy, lz - 2, 5, &r));
        }
    row_nnz[i] = PAD ? (cnt + 3) & ~3 : cnt;
}

// One thread per row, in two passes over the points of the cells around the voxel (points in bin
// order, basis values and weights stored in bin order): the first pass visits all 27 cells and
// accumulates the 75 entries of the x-planes dx = -2..0, the second visits the 18 cells with
// ox >= 0 and accumulates the 50 entries of dx = 1, 2; each pass then writes its planes. Every
// entry receives its point terms in the order of a one-pass loop over all 27 cells (cells in
// (ox, oy, oz) order, points in bin order), so the values do not depend on the split; the split
// keeps at most 75 accumulators live (125 spilled to scratch) for 1.67x the point visits.
// Cells and neighbours are found one bitmap column (3 or 5 consecutive lattice points along z) at
// a time.
// RUNS: the row goes to the run form instead of CSR. The present neighbours of voxel i in the
// bitmap column (dx, dy) (its five z-neighbours) have consecutive ranks start .. start + n - 1;
// runs[p P + i] = start | n << 28 (n = 0: start = i) for p = 5 (dx + 2) + (dy + 2), and
// rvals[(5 p + j) P + i] the value of column start + j (0 for j >= n), P = run_plane_stride(M).
// Both arrays are plane-major, so the run and value loads of consecutive rows are contiguous.
constexpr int RUN_COUNT_SHIFT = 28;

// Lanes per row of the assembly and of stencil_diag_rhs_kernel (the two must agree for their
// diagonals and RHS terms to be bit-identical). A coarse level can have too few rows to fill the
// GPU while each row visits hundreds of points (~27 n / M), so G lanes split a row's points (lane j
// takes points j, j + G, ... of every cell) and their partial sums are combined by an xor
// butterfly. The split costs registers (fewer waves per SIMD), so it is used only when the rows
// alone give fewer than 128k threads. Measured at 2M points: 55k rows, ~1000 visits per row:
// 3.5 ms with 1 lane per row, 2.0 with 4, 1.4 with 16; 261k rows: 1.4, 1.9, 3.2.
static int g_assembly_lanes = 0;  // 0: automatic (ngs_assembly_config)

inline int assembly_lanes(int64_t n_points, int64_t M) {
    if (g_assembly_lanes) return g_assembly_lanes;
    constexpr int64_t FILL = 128 * 1024;
    const double visits = M > 0 ? 27.0 * (double)n_points / (double)M : 0.0;
    return (M >= FILL || visits < 64.0) ? 1 : 16;
}

template <int G>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
    for (int d = G / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Accumulates the NPL x-planes DX0 .. DX0 + NPL - 1 of the row over the cells that reach them
// (|dx - ox| <= 1); the pass holding plane 0 visits all 27 cells and also sums rs and rc.
template <bool HAS_ALPHA, bool HAS_COARSE, int G, int DX0, int NPL>
__device__ __forceinline__ void assemble_pass(const ngs_point_bins& pts, int mx, int my, int mz, int gl,
                                              const float* __restrict__ coarse, float (&acc)[NPL * 25], float& rs,
                                              float& rc) {
    constexpr int OX0 = DX0 - 1 < -1 ? -1 : DX0 - 1;
    constexpr int OX1 = DX0 + NPL > 1 ? 1 : DX0 + NPL;
    constexpr bool CENTRE = DX0 <= 0 && 0 < DX0 + NPL;
    const ngs_grid& cg = pts.cells;
#pragma unroll
    for (int k = 0; k < NPL * 25; ++k) acc[k] = 0.f;
#pragma unroll
    for (int ox = OX0; ox <= OX1; ++ox) {
#pragma unroll
        for (int oy = -1; oy <= 1; ++oy) {
            int32_t cid;
            const uint32_t cm = grid_column(cg, mx + ox, my + oy, mz - 1, 3, &cid);
#pragma unroll
            for (int oz = -1; oz <= 1; ++oz) {
                if (!((cm >> (oz + 1)) & 1u)) continue;
                const int k0 = pts.cell_start[cid], k1 = pts.cell_start[cid + 1];
                ++cid;
                for (int k = k0 + gl; k < k1; k += G) {
                    const float* bp = pts.b + 9 * (int64_t)k;
                    const float bx[3] = {bp[0], bp[1], bp[2]};
                    const float by[3] = {bp[3], bp[4], bp[5]};
                    const float bz[3] = {bp[6], bp[7], bp[8]};
                    const float a = HAS_ALPHA ? pts.alpha[k] : pts.alpha_scalar;
                    // products rounded on their own, sums by explicit FMA: the same arithmetic as
                    // stencil_diag_rhs_kernel, whatever the contraction choices around it
                    float w;
                    {
#pragma clang fp contract(off)
                        w = a * (bx[1 - ox] * by[1 - oy] * bz[1 - oz]);
                    }
                    if (CENTRE) {
                        rs += w;
                        // coarse is in bin order when the bins carry no permutation (a contiguous
                        // stream instead of one random gather per visit; 27 rows visit every point)
                        if (HAS_COARSE) rc = fmaf(w, coarse[pts.perm ? pts.perm[k] : k], rc);
                    }
#pragma unroll
                    for (int pl = 0; pl < NPL; ++pl) {
                        const int vx = DX0 + pl - ox;
                        if (vx < -1 || vx > 1) continue;  // resolved at compile time
#pragma unroll
                        for (int vy = -1; vy <= 1; ++vy) {
                            float wxy;
                            {
#pragma clang fp contract(off)
                                wxy = w * (bx[vx + 1] * by[vy + 1]);
                            }
#pragma unroll
                            for (int vz = -1; vz <= 1; ++vz) {
                                float& e = acc[pl * 25 + (oy + vy + 2) * 5 + (oz + vz + 2)];
                                e = fmaf(wxy, bz[vz + 1], e);
                            }
                        }
                    }
                }
            }
        }
    }
    if (G > 1) {
#pragma unroll
        for (int q = 0; q < NPL * 25; ++q) acc[q] = group_sum<G>(acc[q]);
    }
}

// PK (with RUNS and G = 1): the values go straight to the packed run form (pvals from soff[slice],
// plane by plane, the rows' runs at the wave prefix of their lengths; blocks of 64 threads are
// exactly the 64-row slices) instead of the value slots.
template <bool HAS_ALPHA, bool HAS_COARSE, bool RUNS, int G, bool PK = false>
__global__ void __launch_bounds__(64) lhs_assemble_kernel(ngs_level lv, ngs_point_bins pts,
                                                       const float* __restrict__ coarse, float sf, float sf_delta,
                                                       const float* __restrict__ F1, const float* __restrict__ G1,
                                                       const int32_t* __restrict__ rowptr, int32_t* __restrict__ cols,
                                                       float* __restrict__ vals, float* __restrict__ inv_diag,
                                                       double* __restrict__ rhs, int32_t* __restrict__ runs,
                                                       float* __restrict__ rvals, const int64_t* __restrict__ soff = nullptr) {
    static_assert(!PK || (RUNS && G == 1), "the packed form is written by one lane per row");
    // neighbouring rows share most of their points: each XCD takes a contiguous stretch of rows
    const int64_t t = xcd_contiguous_block() * blockDim.x + threadIdx.x;
    const int64_t i = t / G;  // the G lanes of a row are consecutive lanes of one wave
    const int gl = (int)(t % G);
    if (i >= lv.M) return;
    const int s = lv.stride;
    const int cx = lv.coords[3 * i], cy = lv.coords[3 * i + 1], cz = lv.coords[3 * i + 2];
    const ngs_grid& cg = pts.cells;
    const int c2 = cg.log2_stride;
    const int mx = (cx >> c2) - cg.origin[0], my = (cy >> c2) - cg.origin[1], mz = (cz >> c2) - cg.origin[2];
    float rs = 0.f, rc = 0.f;
    int cnt = 0;
    const float fs = (float)s, inv_s = 1.0f / (float)s;
    const ngs_grid& g = lv.grid;
    const int l2 = g.log2_stride;
    const int lx = (cx >> l2) - g.origin[0], ly = (cy >> l2) - g.origin[1], lz = (cz >> l2) - g.origin[2];
    // CSR: entries go through a 4-deep shift register and leave as aligned 16-byte stores (rows
    // start at multiples of 4): a quarter of the store instructions of per-entry stores.
    typedef float f4 __attribute__((ext_vector_type(4)));
    typedef int i4 __attribute__((ext_vector_type(4)));
    const int32_t base = RUNS ? 0 : rowptr[i];
    float vb0 = 0.f, vb1 = 0.f, vb2 = 0.f, vb3 = 0.f;
    int32_t cb0 = 0, cb1 = 0, cb2 = 0, cb3 = 0;
    auto emit = [&](float v, int32_t col) {
        vb0 = vb1; vb1 = vb2; vb2 = vb3; vb3 = v;
        cb0 = cb1; cb1 = cb2; cb2 = cb3; cb3 = col;
        if ((cnt & 3) == 3) {
            const int64_t k = (int64_t)base + cnt - 3;
            *reinterpret_cast<f4*>(vals + k) = f4{vb0, vb1, vb2, vb3};
            *reinterpret_cast<i4*>(cols + k) = i4{cb0, cb1, cb2, cb3};
        }
        ++cnt;
    };
    const int64_t P = run_plane_stride(lv.M);
    int64_t ppos = PK ? soff[i >> 6] : 0;  // running value position of the slice (wave-uniform)
    // the planes' entries in increasing column order (lexicographic offsets over sorted voxels)
    auto emit_plane = [&](int dx, const float* acc) {
        const float c0 = F1[dx + 2] * fs, g0 = G1[dx + 2] * inv_s;
#pragma unroll
        for (int dy = -2; dy <= 2; ++dy) {
            int32_t j;
            const uint32_t m = grid_column(g, lx + dx, ly + dy, lz - 2, 5, &j);
            const int pl = (dx + 2) * 5 + (dy + 2);
            if (RUNS) runs[pl * P + i] = (m ? j : (int32_t)i) | (__popc(m) << RUN_COUNT_SHIFT);
            int q = 0;
            int pex = 0, ptot = 0;
            if (PK) wave_prefix3(__popc(m), pex, ptot);
#pragma unroll
            for (int dz = -2; dz <= 2; ++dz) {
                if (!((m >> (dz + 2)) & 1u)) continue;
                float v;
                {
#pragma clang fp contract(off)
                    const float c1 = F1[dy + 2] * fs, c2f = F1[dz + 2] * fs;
                    const float g1 = G1[dy + 2] * inv_s, g2 = G1[dz + 2] * inv_s;
                    const float K = g0 * c1 * c2f + c0 * g1 * c2f + c0 * c1 * g2;
                    v = K + sf * acc[(dy + 2) * 5 + (dz + 2)];
                }
                if (dx == 0 && dy == 0 && dz == 0) inv_diag[i] = 1.0f / v;
                if (PK) rvals[ppos + pex + q++] = v;
                else if (RUNS) rvals[(5 * pl + q++) * P + i] = v;
                else emit(v, j++);
            }
            if (PK) ppos += ptot;
            else if (RUNS)
                for (; q < 5; ++q) rvals[(5 * pl + q) * P + i] = 0.f;
        }
    };
    {
        float acc[75];
        assemble_pass<HAS_ALPHA, HAS_COARSE, G, -2, 3>(pts, mx, my, mz, gl, coarse, acc, rs, rc);
        if (G == 1 || gl == 0)
#pragma unroll
            for (int pl = 0; pl < 3; ++pl) emit_plane(pl - 2, acc + 25 * pl);
    }
    {
        float acc[50];
        float unused_s = 0.f, unused_c = 0.f;
        assemble_pass<HAS_ALPHA, HAS_COARSE, G, 1, 2>(pts, mx, my, mz, gl, coarse, acc, unused_s, unused_c);
        if (G == 1 || gl == 0)
#pragma unroll
            for (int pl = 0; pl < 2; ++pl) emit_plane(pl + 1, acc + 25 * pl);
    }
    if (G > 1) {
        rs = group_sum<G>(rs);
        if (HAS_COARSE) rc = group_sum<G>(rc);
        if (gl != 0) return;
    }
    if (!RUNS)
        while (cnt & 3) emit(0.f, (int32_t)i);  // padding: the row itself with a zero value
    rhs[i] += (double)(sf_delta * rs) - (double)(sf * rc);
}

// 125 same-level stiffness coefficients of stride s, computed exactly as in lhs_assemble_kernel.
__global__ void stencil_laplacian_kernel(const float* __restrict__ F1, const float* __restrict__ G1, int s,
                                         float* __restrict__ lap) {
    const int t = threadIdx.x;
    if (t >= 125) return;
    const int dx = t / 25 - 2, dy = (t / 5) % 5 - 2, dz = t % 5 - 2;
    const float fs = (float)s, inv_s = 1.0f / (float)s;
    {
#pragma clang fp contract(off)
        const float c0 = F1[dx + 2] * fs, c1 = F1[dy + 2] * fs, c2 = F1[dz + 2] * fs;
        const float g0 = G1[dx + 2] * inv_s, g1 = G1[dy + 2] * inv_s, g2 = G1[dz + 2] * inv_s;
        lap[t] = g0 * c1 * c2 + c0 * g1 * c2 + c0 * c1 * g2;
    }
}

__global__ void gather_f32_kernel(const float* __restrict__ src, const int32_t* __restrict__ idx, int64_t n,
                                  float* __restrict__ dst) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[idx[k]];
}

__global__ void bin_points_kernel(const int32_t* __restrict__ perm, int64_t n, const int32_t* __restrict__ cell,
                                  const float* __restrict__ b, const float* __restrict__ alpha,
                                  int32_t* __restrict__ pcell, float* __restrict__ basis, float* __restrict__ alpha_out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t p = perm[k];
#pragma unroll
    for (int a = 0; a < 3; ++a) pcell[3 * k + a] = cell[3 * p + a];
#pragma unroll
    for (int j = 0; j < 9; ++j) basis[9 * k + j] = b[9 * p + j];
    if (alpha) alpha_out[k] = alpha[p];
}

// Diagonal and screening RHS of row i for the stencil form: the same products, in the same
// order, as the diagonal entry and rs/rc of lhs_assemble_kernel (cells in (ox, oy, oz) order,
// points in bin order).
template <int G>
__global__ void __launch_bounds__(256) stencil_diag_rhs_kernel(ngs_stencil op, const float* __restrict__ coarse,
                                                               float sf_delta,
                                                               float* __restrict__ inv_diag, double* __restrict__ rhs) {
    const int64_t t = xcd_contiguous_block() * blockDim.x + threadIdx.x;
    const int64_t i = t / G;
    const int gl = (int)(t % G);
    if (i >= op.level.M) return;
    const int s = op.level.stride;
    const int cx = op.level.coords[3 * i], cy = op.level.coords[3 * i + 1], cz = op.level.coords[3 * i + 2];
    float dd = 0.f, rs = 0.f, rc = 0.f;
    if (op.n > 0) {
#pragma unroll
        for (int ox = -1; ox <= 1; ++ox) {
#pragma unroll
            for (int oy = -1; oy <= 1; ++oy) {
#pragma unroll
                for (int oz = -1; oz <= 1; ++oz) {
                    const int cid = grid_find(op.cells, cx + ox * s, cy + oy * s, cz + oz * s);
                    if (cid < 0) continue;
                    const int k0 = op.cell_start[cid], k1 = op.cell_start[cid + 1];
                    for (int k = k0 + gl; k < k1; k += G) {
                        const float* bp = op.basis + 9 * (int64_t)k;
                        const float bx = bp[1 - ox], by = bp[4 - oy], bz = bp[7 - oz];
                        const float a = op.alpha ? op.alpha[k] : op.alpha_scalar;
                        float w, wxy;
                        {
#pragma clang fp contract(off)
                            w = a * (bx * by * bz);
                            wxy = w * (bx * by);
                        }
                        rs += w;
                        if (coarse) rc = fmaf(w, coarse[op.perm[k]], rc);
                        dd = fmaf(wxy, bz, dd);
                    }
                }
            }
        }
    }
    if (G > 1) {
        dd = group_sum<G>(dd);
        rs = group_sum<G>(rs);
        rc = group_sum<G>(rc);
        if (gl != 0) return;
    }
    float v;
    {
#pragma clang fp contract(off)
        v = op.lap[62] + op.sf * dd;  // K(0) + sf * S_ii
    }
    inv_diag[i] = 1.0f / v;
    if (op.n > 0) rhs[i] += (double)(sf_delta * rs) - (double)(op.sf * rc);
}

// rel (in units of the fine stride) and table row for the cross-level pair (fine voxel at c,
// coarse voxel at Q + k s_c): rel = k m + (Q - c)/s_f + (m - 1)/2, row = rel + 1.5 m + 0.5.
__global__ void cross_stiffness_kernel(ngs_level fine, ngs_level coarse, const float* __restrict__ xc, Tables T,
                                       double* __restrict__ rhs) {
    int64_t i = xcd_contiguous_block() * blockDim.x + threadIdx.x;
    if (i >= fine.M) return;
    const int sf_ = fine.stride, sc = coarse.stride, m = T.m;
    int c[3], Q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        c[a] = fine.coords[3 * i + a];
        // floor((c + s_f/2) / s_c) * s_c with exact integer arithmetic
        int num = 2 * c[a] + sf_;
        int den = 2 * sc;
        int q = num >= 0 ? num / den : -((-num + den - 1) / den);
        Q[a] = q * sc;
    }
    const int nrow = 3 * m + 2;
    const float fs = (float)sf_, inv_sc = 1.0f / (float)sc;
    const ngs_grid& g = coarse.grid;
    const int l2 = g.log2_stride;
    const int qx = (Q[0] >> l2) - g.origin[0], qy = (Q[1] >> l2) - g.origin[1], qz = (Q[2] >> l2) - g.origin[2];
    // per-axis table factors, loaded once: z for all five offsets, x and y per column
    float cz[5], gz[5];
    bool okz[5];
#pragma unroll
    for (int t = 0; t < 5; ++t) {
        const int iz = (t - 2) * m + (Q[2] - c[2]) / sf_ + 2 * m;
        okz[t] = iz >= 0 && iz < nrow;
        {
#pragma clang fp contract(off)
            cz[t] = okz[t] ? T.F[iz] * fs : 0.f;
            gz[t] = okz[t] ? T.G[iz] * inv_sc : 0.f;
        }
    }
    float acc = 0.f;
    for (int kx = -2; kx <= 2; ++kx) {
        const int ix = kx * m + (Q[0] - c[0]) / sf_ + 2 * m;
        if (ix < 0 || ix >= nrow) continue;
        float c0, g0;
        {
#pragma clang fp contract(off)
            c0 = T.F[ix] * fs;
            g0 = T.G[ix] * inv_sc;
        }
        for (int ky = -2; ky <= 2; ++ky) {
            const int iy = ky * m + (Q[1] - c[1]) / sf_ + 2 * m;
            if (iy < 0 || iy >= nrow) continue;
            float c1, g1;
            {
#pragma clang fp contract(off)
                c1 = T.F[iy] * fs;
                g1 = T.G[iy] * inv_sc;
            }
            int32_t j;
            const uint32_t msk = grid_column(g, qx + kx, qy + ky, qz - 2, 5, &j);
#pragma unroll
            for (int t = 0; t < 5; ++t) {
                if (!((msk >> t) & 1u)) continue;
                const int32_t jj = j++;
                if (!okz[t]) continue;
                float K;
                {
#pragma clang fp contract(off)
                    K = g0 * c1 * cz[t] + c0 * g1 * cz[t] + c0 * c1 * gz[t];
                }
                acc += K * xc[jj];
            }
        }
    }
    rhs[i] -= (double)acc;
}

// Data at a finer or equal level: scatter each data voxel to the 27 target voxels around the
// target cell containing its centre (get_neighbours(dd, d, 3), _hash_backend.py:176-228).
__global__ void rhs_normals_scatter_kernel(ngs_level tgt, ngs_level dat, const float* __restrict__ nrm, Tables T,
                                           double* __restrict__ rhs) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= dat.M) return;
    const int sd = dat.stride, st = tgt.stride, m = T.m;
    int c[3], Q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        c[a] = dat.coords[3 * v + a];
        int num = 2 * c[a] + sd, den = 2 * st;
        int q = num >= 0 ? num / den : -((-num + den - 1) / den);
        Q[a] = q * st;
    }
    const float n0 = -nrm[3 * v], n1 = -nrm[3 * v + 1], n2 = -nrm[3 * v + 2];
    const float fsd = (float)sd, imult = (float)sd / (float)st;
    for (int kx = -1; kx <= 1; ++kx) {
        int ix = kx * m + (Q[0] - c[0]) / sd + 2 * m - 1;
        for (int ky = -1; ky <= 1; ++ky) {
            int iy = ky * m + (Q[1] - c[1]) / sd + 2 * m - 1;
            for (int kz = -1; kz <= 1; ++kz) {
                int iz = kz * m + (Q[2] - c[2]) / sd + 2 * m - 1;
                if ((unsigned)ix >= (unsigned)(3 * m) || (unsigned)iy >= (unsigned)(3 * m) ||
                    (unsigned)iz >= (unsigned)(3 * m))
                    continue;
                int t = grid_find(tgt.grid, Q[0] + kx * st, Q[1] + ky * st, Q[2] + kz * st);
                if (t < 0) continue;
                float r;
                {
#pragma clang fp contract(off)
                    const float c0 = T.P[ix] * fsd, c1 = T.P[iy] * fsd, c2 = T.P[iz] * fsd;
                    const float g0 = T.Q[ix] * imult, g1 = T.Q[iy] * imult, g2 = T.Q[iz] * imult;
                    r = n0 * g0 * c1 * c2 + n1 * c0 * g1 * c2 + n2 * c0 * c1 * g2;
                }
                atomicAdd(&rhs[t], (double)r);
            }
        }
    }
}

// Data at the same level: the transpose of the scatter above for m = 1, gathered per target voxel t
// from the data voxels t - k s, k in [-1,1]^3, with the same FP32 terms summed in FP64 in a fixed
// order (no atomics).
__global__ void rhs_normals_same_kernel(ngs_level tgt, ngs_level dat, const float* __restrict__ nrm, Tables T,
                                        double* __restrict__ rhs) {
    const int64_t i = xcd_contiguous_block() * blockDim.x + threadIdx.x;
    if (i >= tgt.M) return;
    const int s = tgt.stride;
    const int cx = tgt.coords[3 * i], cy = tgt.coords[3 * i + 1], cz = tgt.coords[3 * i + 2];
    const float fsd = (float)s, imult = 1.0f;
    const ngs_grid& g = dat.grid;
    const int l2 = g.log2_stride;
    const int lx = (cx >> l2) - g.origin[0], ly = (cy >> l2) - g.origin[1], lz = (cz >> l2) - g.origin[2];
    double acc = 0.0;
#pragma unroll
    for (int kx = -1; kx <= 1; ++kx) {
#pragma unroll
        for (int ky = -1; ky <= 1; ++ky) {
            // data voxels t - k s for kz = 1, 0, -1: the column (lz - 1 .. lz + 1), read in kz order
            int32_t r0;
            const uint32_t msk = grid_column(g, lx - kx, ly - ky, lz - 1, 3, &r0);
            if (!msk) continue;
#pragma unroll
            for (int kz = -1; kz <= 1; ++kz) {
                const int bit = 1 - kz;  // lattice offset of t - kz s in the column
                if (!((msk >> bit) & 1u)) continue;
                const int v = r0 + __popc(msk & ((1u << bit) - 1u));
                const float n0 = -nrm[3 * v], n1 = -nrm[3 * v + 1], n2 = -nrm[3 * v + 2];
                const int ix = kx + 1, iy = ky + 1, iz = kz + 1;
                float r;
                {
#pragma clang fp contract(off)
                    const float c0 = T.P[ix] * fsd, c1 = T.P[iy] * fsd, c2 = T.P[iz] * fsd;
                    const float g0 = T.Q[ix] * imult, g1 = T.Q[iy] * imult, g2 = T.Q[iz] * imult;
                    r = n0 * g0 * c1 * c2 + n1 * c0 * g1 * c2 + n2 * c0 * c1 * g2;
                }
                acc += (double)r;
            }
        }
    }
    rhs[i] += acc;
}

// Data at a finer level (m = s_t / s_d >= 2), gathered per target voxel t: the data voxels whose
// target cell is t - k s_t (k in [-1,1]^3) are those of the 3m-wide window starting at t - s_t on
// each axis, and the data voxel at window position j has table row 3m - 1 - j (the row the
// scatter kernel computes for the same pair). Window columns are read from the data bitmap; the
// same FP32 terms are summed in FP64 in a fixed order (no atomics). G lanes share a target (lane j
// takes window x-columns j, j + G, ...) when the targets are few and the windows wide; their FP64
// partials are combined by an xor butterfly.
template <int G>
__global__ void rhs_normals_fine_gather_kernel(ngs_level tgt, ngs_level dat, const float* __restrict__ nrm, Tables T,
                                               double* __restrict__ rhs) {
    const int64_t t = xcd_contiguous_block() * blockDim.x + threadIdx.x;
    const int64_t i = t / G;  // the G lanes of a target are consecutive lanes of one wave
    const int gl = (int)(t % G);
    if (i >= tgt.M) return;
    const int sd = dat.stride, st = tgt.stride, m = T.m, w3 = 3 * m;
    const ngs_grid& g = dat.grid;
    const int l2 = g.log2_stride;
    const int lx0 = ((tgt.coords[3 * i] - st) >> l2) - g.origin[0];
    const int ly0 = ((tgt.coords[3 * i + 1] - st) >> l2) - g.origin[1];
    const int lz0 = ((tgt.coords[3 * i + 2] - st) >> l2) - g.origin[2];
    const float fsd = (float)sd, imult = (float)sd / (float)st;
    double acc = 0.0;
    for (int jx = gl; jx < w3; jx += G) {
        const int ix = w3 - 1 - jx;
        const float px = T.P[ix] * fsd, qx = T.Q[ix] * imult;
        for (int jy = 0; jy < w3; ++jy) {
            const int iy = w3 - 1 - jy;
            const float py = T.P[iy] * fsd, qy = T.Q[iy] * imult;
            for (int z0 = 0; z0 < w3; z0 += 24) {
                const int len = min(24, w3 - z0);
                int32_t r;
                uint32_t msk = grid_column(g, lx0 + jx, ly0 + jy, lz0 + z0, len, &r);
                while (msk) {
                    const int jz = z0 + __ffs(msk) - 1;
                    msk &= msk - 1u;
                    const int iz = w3 - 1 - jz;
                    const float n0 = -nrm[3 * r], n1 = -nrm[3 * r + 1], n2 = -nrm[3 * r + 2];
                    ++r;
                    float v;
                    {
#pragma clang fp contract(off)
                        const float c0 = px, c1 = py, c2 = T.P[iz] * fsd;
                        const float g0 = qx, g1 = qy, g2 = T.Q[iz] * imult;
                        v = n0 * g0 * c1 * c2 + n1 * c0 * g1 * c2 + n2 * c0 * c1 * g2;
                    }
                    acc += (double)v;
                }
            }
        }
    }
    if (G > 1) {
#pragma unroll
        for (int d = G / 2; d > 0; d >>= 1) acc += __shfl_xor(acc, d, 64);
        if (gl != 0) return;
    }
    rhs[i] += acc;
}

// Data at a coarser level: gather, per target voxel, the data voxels around the data cell
// containing its centre, kept when |rel| <= 1.5 target strides (_hash_backend.py:254-262).
__global__ void rhs_normals_gather_kernel(ngs_level tgt, ngs_level dat, const float* __restrict__ nrm, Tables T,
                                          double* __restrict__ rhs) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= tgt.M) return;
    const int st = tgt.stride, sd = dat.stride, m = T.m;  // m = sd / st >= 2
    int c[3], Q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        c[a] = tgt.coords[3 * i + a];
        int num = 2 * c[a] + st, den = 2 * sd;
        int q = num >= 0 ? num / den : -((-num + den - 1) / den);
        Q[a] = q * sd;
    }
    const float fst = (float)st;
    float acc = 0.f;
    for (int kx = -1; kx <= 1; ++kx) {
        // twice the offset of the data centre from the target centre, in target strides
        int r2x = 2 * (kx * m + (Q[0] - c[0]) / st) + (m - 1);
        if (r2x > 3 || r2x < -3) continue;
        for (int ky = -1; ky <= 1; ++ky) {
            int r2y = 2 * (ky * m + (Q[1] - c[1]) / st) + (m - 1);
            if (r2y > 3 || r2y < -3) continue;
            for (int kz = -1; kz <= 1; ++kz) {
                int r2z = 2 * (kz * m + (Q[2] - c[2]) / st) + (m - 1);
                if (r2z > 3 || r2z < -3) continue;
                int v = grid_find(dat.grid, Q[0] + kx * sd, Q[1] + ky * sd, Q[2] + kz * sd);
                if (v < 0) continue;
                // row = round(-rel_inv + m/2 + 1/2) with rel_inv = r2/2
                int ix = (-r2x + m + 1) / 2, iy = (-r2y + m + 1) / 2, iz = (-r2z + m + 1) / 2;
                if ((unsigned)ix > (unsigned)(m + 1) || (unsigned)iy > (unsigned)(m + 1) || (unsigned)iz > (unsigned)(m + 1))
                    continue;
                float r;
                {
#pragma clang fp contract(off)
                    const float c0 = T.Pinv[ix] * fst, c1 = T.Pinv[iy] * fst, c2 = T.Pinv[iz] * fst;
                    const float g0 = T.Qinv[ix], g1 = T.Qinv[iy], g2 = T.Qinv[iz];
                    r = -nrm[3 * v] * g0 * c1 * c2 + -nrm[3 * v + 1] * c0 * g1 * c2 + -nrm[3 * v + 2] * c0 * c1 * g2;
                }
                acc += r;
            }
        }
    }
    rhs[i] += (double)acc;
}

// Transpose of the field evaluation: out[i] += scale * sum_p w_p B_i(p) over the 27 voxels of g
// around each point (the adjoint of hashtree.py:157-192).
__global__ void basis_splat_kernel(ngs_grid g, const float* __restrict__ w, const float* __restrict__ xyz, int64_t n,
                                   float inv_vs, double scale, double* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float wi = w[i];
    if (wi == 0.f) return;
    const int s = 1 << g.log2_stride;
    const float inv_s = 1.0f / (float)s;
    int cell[3];
    float b[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) axis_basis(xyz[3 * i + a] * inv_vs, s, inv_s, &cell[a], b[a]);
    for (int kx = 0; kx < 3; ++kx)
        for (int ky = 0; ky < 3; ++ky)
            for (int kz = 0; kz < 3; ++kz) {
                int id = grid_find(g, cell[0] + (kx - 1) * s, cell[1] + (ky - 1) * s, cell[2] + (kz - 1) * s);
                if (id < 0) continue;
                atomicAdd(&out[id], scale * (double)(wi * (b[0][kx] * b[1][ky] * b[2][kz])));
            }
}

// Transpose of the cross-level stiffness: out_c[j] += scale * sum_i K_{d,dd}(i, j) g_f[i].
__global__ void cross_stiffness_t_kernel(ngs_level fine, ngs_level coarse, const float* __restrict__ gf, Tables T,
                                         double scale, double* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fine.M) return;
    const float gi = gf[i];
    if (gi == 0.f) return;
    const int sf_ = fine.stride, sc = coarse.stride, m = T.m;
    int c[3], Q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        c[a] = fine.coords[3 * i + a];
        int num = 2 * c[a] + sf_, den = 2 * sc;
        int q = num >= 0 ? num / den : -((-num + den - 1) / den);
        Q[a] = q * sc;
    }
    const int nrow = 3 * m + 2;
    const float fs = (float)sf_, inv_sc = 1.0f / (float)sc;
    for (int kx = -2; kx <= 2; ++kx) {
        int ix = kx * m + (Q[0] - c[0]) / sf_ + 2 * m;
        for (int ky = -2; ky <= 2; ++ky) {
            int iy = ky * m + (Q[1] - c[1]) / sf_ + 2 * m;
            for (int kz = -2; kz <= 2; ++kz) {
                int iz = kz * m + (Q[2] - c[2]) / sf_ + 2 * m;
                if (ix < 0 || ix >= nrow || iy < 0 || iy >= nrow || iz < 0 || iz >= nrow) continue;
                int j = grid_find(coarse.grid, Q[0] + kx * sc, Q[1] + ky * sc, Q[2] + kz * sc);
                if (j < 0) continue;
                const float c0 = T.F[ix] * fs, c1 = T.F[iy] * fs, c2 = T.F[iz] * fs;
                const float g0 = T.G[ix] * inv_sc, g1 = T.G[iy] * inv_sc, g2 = T.G[iz] * inv_sc;
                const float K = g0 * c1 * c2 + c0 * g1 * c2 + c0 * c1 * g2;
                atomicAdd(&out[j], scale * (double)(K * gi));
            }
        }
    }
}

}  // namespace ngs

using namespace ngs;

// Tables are uploaded once per process by the host layer.
static const float* g_tab_ptr = nullptr;
static int64_t g_tab_off[8][6];
static int g_tab_levels = 0;

static bool tables_for(int m, Tables* T) {
    int k = 0;
    while ((1 << k) < m) ++k;
    if ((1 << k) != m || k >= g_tab_levels || !g_tab_ptr) return false;
    T->m = m;
    T->F = g_tab_ptr + g_tab_off[k][0];
    T->G = g_tab_ptr + g_tab_off[k][1];
    T->P = g_tab_ptr + g_tab_off[k][2];
    T->Q = g_tab_ptr + g_tab_off[k][3];
    T->Pinv = k > 0 ? g_tab_ptr + g_tab_off[k][4] : nullptr;
    T->Qinv = k > 0 ? g_tab_ptr + g_tab_off[k][5] : nullptr;
    return true;
}

extern "C" {

// Registers the device copy of the integral tables: for k = 0..levels-1 (m = 2^k), six float
// arrays F, G (3m+2), P, Q (3m), Pinv, Qinv (m+2; unused for k = 0) at the given element offsets
// of `tables`. The buffer must stay alive while kernels use it.
int ngs_set_tables(const float* tables, const int64_t* offsets, int32_t levels) {
    NGS_REQUIRE(tables && offsets && levels >= 1 && levels <= 8, "ngs_set_tables: bad arguments");
    g_tab_ptr = tables;
    g_tab_levels = levels;
    for (int k = 0; k < levels; ++k)
        for (int j = 0; j < 6; ++j) g_tab_off[k][j] = offsets[6 * k + j];
    return NGS_OK;
}

int ngs_lhs_row_counts(ngs_level lv, int32_t* row_nnz, ngs_stream_t s) {
    if (lv.M <= 0) return NGS_OK;
    lhs_count_kernel<true><<<grid1d(lv.M, 256), 256, 0, S(s)>>>(lv, row_nnz);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_stencil_row_counts(ngs_level lv, int32_t* row_nnz, ngs_stream_t s) {
    if (lv.M <= 0) return NGS_OK;
    lhs_count_kernel<false><<<grid1d(lv.M, 256), 256, 0, S(s)>>>(lv, row_nnz);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

static int lhs_assemble(ngs_level lv, ngs_point_bins pts, const float* coarse, float sf, float delta,
                        const int32_t* rowptr, int32_t* cols, float* vals, float* inv_diag, double* rhs, int32_t* runs,
                        float* rvals, ngs_stream_t s, const int64_t* soff = nullptr) {
    if (lv.M <= 0) return NGS_OK;
    Tables T;
    NGS_REQUIRE(tables_for(1, &T), "ngs_lhs_assemble: integral tables not registered");
    const float sfd = sf * delta;
    const int lanes = assembly_lanes(pts.n, lv.M);
    dim3 grid(grid1d(lv.M * lanes, 64)), block(64);
    hipStream_t st = S(s);
    if (soff) {
        NGS_REQUIRE(lanes == 1, "ngs_lhs_assemble_packed: this level assembles with several lanes per row "
                                "(ngs_assembly_lanes); assemble the run form and pack it");
#define NGS_ASM_P(A, C)                                                                                          \
    lhs_assemble_kernel<A, C, true, 1, true><<<grid, block, 0, st>>>(lv, pts, coarse, sf, sfd, T.F, T.G, nullptr, \
                                                                     nullptr, nullptr, inv_diag, rhs, runs, rvals, soff)
        if (pts.alpha) {
            if (coarse) NGS_ASM_P(true, true);
            else NGS_ASM_P(true, false);
        } else {
            if (coarse) NGS_ASM_P(false, true);
            else NGS_ASM_P(false, false);
        }
#undef NGS_ASM_P
        NGS_CHECK_LAUNCH();
        return NGS_OK;
    }
#define NGS_ASM_G(A, C, R, LN)                                                                                      \
    lhs_assemble_kernel<A, C, R, LN><<<grid, block, 0, st>>>(lv, pts, coarse, sf, sfd, T.F, T.G, rowptr, cols, vals, \
                                                             inv_diag, rhs, runs, rvals)
#define NGS_ASM(A, C, R)                        \
    do {                                        \
        if (lanes == 16) NGS_ASM_G(A, C, R, 16);  \
        else if (lanes == 4) NGS_ASM_G(A, C, R, 4); \
        else NGS_ASM_G(A, C, R, 1);               \
    } while (0)
#define NGS_ASM_R(A, C)                \
    do {                               \
        if (runs) NGS_ASM(A, C, true);  \
        else NGS_ASM(A, C, false);      \
    } while (0)
    if (pts.alpha) {
        if (coarse) NGS_ASM_R(true, true);
        else NGS_ASM_R(true, false);
    } else {
        if (coarse) NGS_ASM_R(false, true);
        else NGS_ASM_R(false, false);
    }
#undef NGS_ASM_R
#undef NGS_ASM
#undef NGS_ASM_G
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int32_t ngs_assembly_lanes(int64_t n_points, int64_t M) { return assembly_lanes(n_points, M); }

int ngs_lhs_assemble_packed(ngs_level lv, ngs_point_bins pts, const float* coarse, float sf, float delta, int32_t* runs,
                            float* pvals, const int64_t* soff, float* inv_diag, double* rhs, ngs_stream_t s) {
    NGS_REQUIRE(lv.M <= 0 || (runs && pvals && soff && inv_diag && rhs), "ngs_lhs_assemble_packed: null array");
    NGS_REQUIRE(lv.M <= ((int64_t)1 << RUN_COUNT_SHIFT), "ngs_lhs_assemble_packed: too many rows for run words");
    return lhs_assemble(lv, pts, coarse, sf, delta, nullptr, nullptr, nullptr, inv_diag, rhs, runs, pvals, s, soff);
}

int ngs_assembly_config(int32_t lanes_per_row) {
    NGS_REQUIRE(lanes_per_row == 0 || lanes_per_row == 1 || lanes_per_row == 4 || lanes_per_row == 16,
                "ngs_assembly_config: lanes_per_row must be 0, 1, 4 or 16");
    g_assembly_lanes = lanes_per_row;
    return NGS_OK;
}

int ngs_lhs_assemble(ngs_level lv, ngs_point_bins pts, const float* coarse, float sf, float delta,
                     const int32_t* rowptr, int32_t* cols, float* vals, float* inv_diag, double* rhs, ngs_stream_t s) {
    NGS_REQUIRE(lv.M <= 0 || (rowptr && cols && vals && inv_diag && rhs), "ngs_lhs_assemble: null array");
    return lhs_assemble(lv, pts, coarse, sf, delta, rowptr, cols, vals, inv_diag, rhs, nullptr, nullptr, s);
}

int ngs_lhs_assemble_runs(ngs_level lv, ngs_point_bins pts, const float* coarse, float sf, float delta, int32_t* runs,
                          float* rvals, float* inv_diag, double* rhs, ngs_stream_t s) {
    NGS_REQUIRE(lv.M <= 0 || (runs && rvals && inv_diag && rhs), "ngs_lhs_assemble_runs: null array");
    NGS_REQUIRE(lv.M <= ((int64_t)1 << RUN_COUNT_SHIFT), "ngs_lhs_assemble_runs: too many rows for run words");
    return lhs_assemble(lv, pts, coarse, sf, delta, nullptr, nullptr, nullptr, inv_diag, rhs, runs, rvals, s);
}

int ngs_stencil_laplacian(int32_t stride, float* lap, ngs_stream_t s) {
    NGS_REQUIRE(stride >= 1 && (stride & (stride - 1)) == 0 && lap, "ngs_stencil_laplacian: bad arguments");
    Tables T;
    NGS_REQUIRE(tables_for(1, &T), "ngs_stencil_laplacian: integral tables not registered");
    stencil_laplacian_kernel<<<1, 128, 0, S(s)>>>(T.F, T.G, stride, lap);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_gather_f32(const float* src, const int32_t* idx, int64_t n, float* dst, ngs_stream_t s) {
    if (n <= 0) return NGS_OK;
    NGS_REQUIRE(src && idx && dst, "ngs_gather_f32: null argument");
    gather_f32_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(src, idx, n, dst);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_bin_points(const int32_t* perm, int64_t n, const int32_t* cell, const float* b, const float* alpha,
                   int32_t* point_cell, float* basis, float* alpha_out, ngs_stream_t s) {
    if (n <= 0) return NGS_OK;
    NGS_REQUIRE(perm && cell && b && point_cell && basis && (!alpha || alpha_out), "ngs_bin_points: bad arguments");
    bin_points_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(perm, n, cell, b, alpha, point_cell, basis, alpha_out);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_stencil_diag_rhs(ngs_stencil op, const float* coarse, float delta, float* inv_diag, double* rhs, ngs_stream_t s) {
    if (op.level.M <= 0) return NGS_OK;
    NGS_REQUIRE(inv_diag && rhs && op.level.coords && op.lap && (op.n == 0 || (op.basis && op.perm && op.cell_start)),
                "ngs_stencil_diag_rhs: bad arguments");
    const float sfd = op.sf * delta;
    const int lanes = assembly_lanes(op.n, op.level.M);
    const unsigned grid = grid1d(op.level.M * lanes, 256);
    if (lanes == 16) stencil_diag_rhs_kernel<16><<<grid, 256, 0, S(s)>>>(op, coarse, sfd, inv_diag, rhs);
    else if (lanes == 4) stencil_diag_rhs_kernel<4><<<grid, 256, 0, S(s)>>>(op, coarse, sfd, inv_diag, rhs);
    else stencil_diag_rhs_kernel<1><<<grid, 256, 0, S(s)>>>(op, coarse, sfd, inv_diag, rhs);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_rhs_cross_stiffness(ngs_level fine, ngs_level coarse, const float* x_coarse, double* rhs, ngs_stream_t s) {
    if (fine.M <= 0 || coarse.M <= 0) return NGS_OK;
    NGS_REQUIRE(coarse.stride > fine.stride, "ngs_rhs_cross_stiffness: coarse stride must exceed fine stride");
    Tables T;
    NGS_REQUIRE(tables_for(coarse.stride / fine.stride, &T), "ngs_rhs_cross_stiffness: table multiple unavailable");
    cross_stiffness_kernel<<<grid1d(fine.M, 128), 128, 0, S(s)>>>(fine, coarse, x_coarse, T, rhs);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_rhs_normals(ngs_level target, ngs_level data, const float* normals, double* rhs, ngs_stream_t s) {
    if (target.M <= 0 || data.M <= 0) return NGS_OK;
    Tables T;
    if (target.stride == data.stride) {
        NGS_REQUIRE(tables_for(1, &T), "ngs_rhs_normals: table multiple unavailable");
        rhs_normals_same_kernel<<<grid1d(target.M, 256), 256, 0, S(s)>>>(target, data, normals, T, rhs);
    } else if (target.stride > data.stride) {
        const int m = target.stride / data.stride;
        NGS_REQUIRE(tables_for(m, &T), "ngs_rhs_normals: table multiple unavailable");
        // gather (window scan per target, no atomics) unless the data is much sparser than the
        // targets' windows, where the 27-target scatter touches less
        if ((double)target.M * (3.0 * m) * (3.0 * m) < 4.0 * 27.0 * (double)data.M) {
            // lanes per target: enough threads to fill the GPU, at most one per window x-column
            int lanes = 1;
            while (lanes < 16 && lanes * 2 <= 3 * m && target.M * lanes < 512 * 1024) lanes *= 2;
            const unsigned grid = grid1d(target.M * lanes, 128);
            if (lanes == 16) rhs_normals_fine_gather_kernel<16><<<grid, 128, 0, S(s)>>>(target, data, normals, T, rhs);
            else if (lanes == 8) rhs_normals_fine_gather_kernel<8><<<grid, 128, 0, S(s)>>>(target, data, normals, T, rhs);
            else if (lanes == 4) rhs_normals_fine_gather_kernel<4><<<grid, 128, 0, S(s)>>>(target, data, normals, T, rhs);
            else if (lanes == 2) rhs_normals_fine_gather_kernel<2><<<grid, 128, 0, S(s)>>>(target, data, normals, T, rhs);
            else rhs_normals_fine_gather_kernel<1><<<grid, 128, 0, S(s)>>>(target, data, normals, T, rhs);
        } else
            rhs_normals_scatter_kernel<<<grid1d(data.M, 256), 256, 0, S(s)>>>(target, data, normals, T, rhs);
    } else {
        NGS_REQUIRE(tables_for(data.stride / target.stride, &T), "ngs_rhs_normals: table multiple unavailable");
        rhs_normals_gather_kernel<<<grid1d(target.M, 256), 256, 0, S(s)>>>(target, data, normals, T, rhs);
    }
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_basis_splat(ngs_grid g, const float* w, const float* xyz, int64_t n, float inv_voxel_size, double scale, double* out,
                    ngs_stream_t s) {
    if (n <= 0) return NGS_OK;
    basis_splat_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(g, w, xyz, n, inv_voxel_size, scale, out);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_cross_stiffness_transpose(ngs_level fine, ngs_level coarse, const float* g_fine, double scale, double* out_coarse,
                                  ngs_stream_t s) {
    if (fine.M <= 0 || coarse.M <= 0) return NGS_OK;
    NGS_REQUIRE(coarse.stride > fine.stride, "ngs_cross_stiffness_transpose: coarse stride must exceed fine stride");
    Tables T;
    NGS_REQUIRE(tables_for(coarse.stride / fine.stride, &T), "ngs_cross_stiffness_transpose: table multiple unavailable");
    cross_stiffness_t_kernel<<<grid1d(fine.M, 128), 128, 0, S(s)>>>(fine, coarse, g_fine, T, scale, out_coarse);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

}  // extern "C"
