// Stencil form of a level system: the operator that ngs_lhs_assemble writes as CSR, applied
// without stored entries.
//
//   (A x)_i = sum_{o in [-2,2]^3} lap[o] x_{i+o}  +  sf * sum_{p: cell(p) in i + [-1,1]^3} B_i(p) u_p
//   u_p     = alpha_p * sum_{j in cell(p) + [-1,1]^3} B_j(p) x_j
//
// Neighbour ranks come from the level's ordered bitmap: the five z-neighbours of a voxel at one
// (dx, dy) are consecutive lattice bits, and the ranks of the present ones are consecutive, so
// one bitmap window and one rank read give a whole column. Per-row HBM traffic is a few bytes of
// bitmap plus the gathered vector, instead of 8 bytes per stored entry.
#pragma once

#include "common.h"

namespace ngs {

// Row i of A x. lap: the 125 Laplacian coefficients (lexicographic offsets), typically in LDS.
// u: the point pass of the same x.
template <typename T>
__device__ __forceinline__ T stencil_row(const ngs_stencil& op, const float* lap, const T* __restrict__ x_,
                                         const T* __restrict__ u_, int64_t i) {
    const ngs_grid& g = op.level.grid;
    const NGS_GLOBAL T* __restrict__ x = as_global(x_);
    const NGS_GLOBAL T* __restrict__ u = as_global(u_);
    const NGS_GLOBAL int32_t* coords = as_global(op.level.coords);
    const int cx = coords[3 * i], cy = coords[3 * i + 1], cz = coords[3 * i + 2];
    const int l2 = g.log2_stride;
    const int lx = (cx >> l2) - g.origin[0], ly = (cy >> l2) - g.origin[1], lz = (cz >> l2) - g.origin[2];
    T acc = 0;
#pragma unroll 1
    for (int dx = 0; dx < 5; ++dx) {
#pragma unroll
        for (int dy = 0; dy < 5; ++dy) {
            int32_t r;
            const uint32_t m = grid_column(g, lx + dx - 2, ly + dy - 2, lz - 2, 5, &r);
            const float* lp = lap + (dx * 5 + dy) * 5;
#pragma unroll
            for (int dz = 0; dz < 5; ++dz) {
                if ((m >> dz) & 1u) {
                    acc += (T)lp[dz] * x[r];
                    ++r;
                }
            }
        }
    }
    if (op.n == 0) return acc;
    const ngs_grid& cg = op.cells;
    const int c2 = cg.log2_stride;
    const int mx = (cx >> c2) - cg.origin[0], my = (cy >> c2) - cg.origin[1], mz = (cz >> c2) - cg.origin[2];
    T sc = 0;
#pragma unroll
    for (int ox = 0; ox < 3; ++ox) {
#pragma unroll
        for (int oy = 0; oy < 3; ++oy) {
            int32_t c;
            const uint32_t m = grid_column(cg, mx + ox - 1, my + oy - 1, mz - 1, 3, &c);
            if (!m) continue;
#pragma unroll
            for (int oz = 0; oz < 3; ++oz) {
                if (!((m >> oz) & 1u)) continue;
                const int32_t k0 = as_global(op.cell_start)[c], k1 = as_global(op.cell_start)[c + 1];
                ++c;
                // the cell sits at offset (ox-1, oy-1, oz-1) from voxel i: B_i(p) = bx[2-ox] by[2-oy] bz[2-oz]
                for (int32_t k = k0; k < k1; ++k) {
                    const NGS_GLOBAL float* bp = as_global(op.basis) + 9 * (int64_t)k;
                    const T w = (T)bp[2 - ox] * (T)bp[5 - oy] * (T)bp[8 - oz];
                    sc += w * u[k];
                }
            }
        }
    }
    return acc + (T)op.sf * sc;
}

// u_k of binned point k for the vector x: alpha_k sum_j B_j(p_k) x_j over the present voxels j of
// the 27 around the point's cell.
template <typename T>
__device__ __forceinline__ T stencil_point(const ngs_stencil& op, const T* __restrict__ x_, int64_t k) {
    const ngs_grid& g = op.level.grid;
    const NGS_GLOBAL T* __restrict__ x = as_global(x_);
    const int l2 = g.log2_stride;
    const NGS_GLOBAL int32_t* pc = as_global(op.point_cell) + 3 * k;
    const int lx = (pc[0] >> l2) - g.origin[0], ly = (pc[1] >> l2) - g.origin[1], lz = (pc[2] >> l2) - g.origin[2];
    const NGS_GLOBAL float* bp = as_global(op.basis) + 9 * k;
    T sum = 0;
#pragma unroll
    for (int vx = 0; vx < 3; ++vx) {
#pragma unroll
        for (int vy = 0; vy < 3; ++vy) {
            int32_t r;
            const uint32_t m = grid_column(g, lx + vx - 1, ly + vy - 1, lz - 1, 3, &r);
            if (!m) continue;
            const T wxy = (T)bp[vx] * (T)bp[3 + vy];
#pragma unroll
            for (int vz = 0; vz < 3; ++vz) {
                if ((m >> vz) & 1u) {
                    sum += wxy * (T)bp[6 + vz] * x[r];
                    ++r;
                }
            }
        }
    }
    const T a = op.alpha ? (T)as_global(op.alpha)[k] : (T)op.alpha_scalar;
    return a * sum;
}

}  // namespace ngs
