// Sparse dual marching cubes over the conformal leaves of the voxel hierarchy.
#include "common.h"

namespace ngs {

int scan_i32_internal(const int32_t* counts, int64_t n, int32_t* offsets, void* ws, hipStream_t st);
int sort_pairs_u64(const uint64_t* kin, uint64_t* kout, const int32_t* vin, int32_t* vout, int64_t n, int bits,
                   void* tmp, size_t* tmp_bytes, hipStream_t st);

// Standard 256-case marching-cubes triangle table (Lorensen-Cline / Bourke numbering), one case per
// 60-bit word: triangle edge ids as nibbles from the low end, 0xF terminates.
__constant__ uint64_t c_mc_cases[256] = {
    0xfffffffffffffff, 0xffffffffffff380, 0xffffffffffff910, 0xfffffffff189381,
    0xffffffffffffa21, 0xfffffffffa21380, 0xfffffffff920a29, 0xffffff89a8a2382,
    0xffffffffffff2b3, 0xfffffffff0b82b0, 0xfffffffffb32091, 0xffffffb89b912b1,
    0xfffffffff3ab1a3, 0xffffffab8a801a0, 0xffffff9ab9b3093, 0xfffffffffb8aa89,
    0xffffffffffff874, 0xfffffffff437034, 0xfffffffff748910, 0xffffff137174914,
    0xfffffffff748a21, 0xffffffa21403743, 0xffffff748209a29, 0xfff4973727929a2,
    0xfffffffff2b3748, 0xffffff40242b74b, 0xffffffb32748109, 0xfff1292b9b49b74,
    0xffffff487ab31a3, 0xfff4b7401b41ab1, 0xfff30bab9b09874, 0xffffffab99b4b74,
    0xffffffffffff459, 0xfffffffff380459, 0xfffffffff051450, 0xffffff513538458,
    0xfffffffff459a21, 0xffffff594a21803, 0xffffff204245a25, 0xfff8434535235a2,
    0xfffffffffb32459, 0xffffff594b802b0, 0xffffffb32510450, 0xfff584b82852512,
    0xffffff45931ab3a, 0xfffab81a8180594, 0xfff30bab5b05045, 0xffffffb8aa85845,
    0xfffffffff975879, 0xffffff375359039, 0xffffff751710870, 0xfffffffff753351,
    0xffffff21a759879, 0xfff37503505921a, 0xfff25a758528208, 0xffffff7533525a2,
    0xffffff2b3987597, 0xfffb72029279759, 0xfff751871810b32, 0xffffff51771b12b,
    0xfffb3a31a758859, 0x0aba010b7905075, 0x07570805a30b0ab, 0xfffffffff5b75ab,
    0xffffffffffff56a, 0xfffffffff6a5380, 0xfffffffff6a5109, 0xffffff6a5891381,
    0xfffffffff162561, 0xffffff803621561, 0xffffff620609569, 0xfff823625285895,
    0xfffffffff56ab32, 0xffffff56a02b80b, 0xffffff6a5b32910, 0xfffb892b92916a5,
    0xffffff315356b36, 0xfff6b51505b0b80, 0xfff9505606306b3, 0xffffff89bb96956,
    0xfffffffff8746a5, 0xffffffa56374034, 0xffffff7486a5091, 0xfff49737179156a,
    0xffffff874156216, 0xfff743403625521, 0xfff620560509748, 0x962695923497937,
    0xffffff56a4872b3, 0xfffb720242746a5, 0xfff6a5b32874910, 0x6a54b7b492b9129,
    0xfff6b51535b3748, 0xb404b7b016b5b15, 0x74836b630560950, 0xfff9b7974b96956,
    0xfffffffffa4694a, 0xffffff380a946a4, 0xffffff04606a10a, 0xfffa16468618138,
    0xffffff462421941, 0xfff462942921803, 0xfffffffff624420, 0xffffff624428238,
    0xffffff32b46a94a, 0xfff6a4a94b82280, 0xfffa164606102b3, 0x1b8b12184a16146,
    0xfff36b319639469, 0x14641916b0181b8, 0xffffff4600636b3, 0xfffffffff86b846,
    0xffffffa98a876a7, 0xfffa76a907a0370, 0xfff0818717a176a, 0xffffff37117a76a,
    0xfff768981861621, 0x937390976192962, 0xffffff206607087, 0xfffffffff276237,
    0xfff76898a86ab32, 0x7a9a76790b72702, 0xb32a767a1871081, 0xfff17616a71b12b,
    0x63136b619768698, 0xfffffffff76b190, 0xfff06b0b3607087, 0xffffffffffff6b7,
    0xffffffffffffb67, 0xfffffffff67b803, 0xfffffffff67b910, 0xffffff67b138918,
    0xfffffffff7b621a, 0xffffff7b6803a21, 0xffffff7b69a2092, 0xfff89a38a3a27b6,
    0xfffffffff726327, 0xffffff026067807, 0xffffff910732672, 0xfff678891681261,
    0xffffff73171a67a, 0xfff801781a7167a, 0xfff7a69a0a70730, 0xffffff9a88a7a67,
    0xfffffffff68b486, 0xffffff640603b63, 0xffffff109648b68, 0xfff63b139369649,
    0xffffff1a28b6486, 0xfff640b60b03a21, 0xfff9a2920b648b4, 0x36463b34923a39a,
    0xffffff264248328, 0xfffffffff264240, 0xfff834642432091, 0xffffff642241491,
    0xfff1a6648168318, 0xffffff40660a01a, 0x39a9303a6834364, 0xfffffffff4a649a,
    0xfffffffffb67594, 0xffffff67b594380, 0xffffffb67045105, 0xfff51345343867b,
    0xffffffb6721a459, 0xfff594380a217b6, 0xfff204a24a45b67, 0x67b25a523453843,
    0xffffff945267327, 0xfff786260680459, 0xfff045051673263, 0x851584812786826,
    0xfff73167161a459, 0x459078701671a61, 0xa737a6a305a4a04, 0xfffa84a458a7a67,
    0xffffff98b9b6596, 0xfff590650360b63, 0xfffb65510b508b0, 0xffffff1355363b6,
    0xfff65b8b9b59a21, 0xa21965690b603b0, 0x52025a50865b58b, 0xfff35a3a25363b6,
    0xfff283265825985, 0xffffff260069659, 0x826283865081851, 0xfffffffff612651,
    0x698965683a61631, 0xfff06505960a01a, 0xfffffffffa65830, 0xffffffffffff65a,
    0xfffffffffb57a5b, 0xffffff03857ba5b, 0xffffff091ba57b5, 0xfff1381897ba57a,
    0xffffff15717b21b, 0xfffb27571721380, 0xfff7b2209729579, 0x289823295b27257,
    0xffffff573532a52, 0xfff52a578258028, 0xfff2a37353a5109, 0x25752a278129289,
    0xfffffffff573531, 0xffffff571170780, 0xffffff735539309, 0xfffffffff795789,
    0xffffff8ba8a5485, 0xfff03bba50b5405, 0xfff54aba8a48910, 0x41314943b54a4ba,
    0xfff8548b2582152, 0xb151b2b543b0b40, 0x58b8545b2950520, 0xfffffffff3b2549,
    0xfff483543253a52, 0xffffff0244252a5, 0x910854583a532a3, 0xfff2492914252a5,
    0xffffff153358548, 0xfffffffff501540, 0xfff530509358548, 0xffffffffffff549,
    0xffffffba9b947b4, 0xfffba97b9794380, 0xfffb470414b1ba1, 0x4bab474a1843413,
    0xfff219b294b97b4, 0x3801b2b197b9479, 0xffffff04224b47b, 0xfff42343824b47b,
    0xfff947732972a92, 0x70207872a4797a9, 0xa040a1a472a3a73, 0xfffffffff4782a1,
    0xffffff317714194, 0xfff178180714194, 0xfffffffff347304, 0xffffffffffff784,
    0xfffffffff8ba8a9, 0xffffffa9bb93903, 0xffffffba88a0a10, 0xfffffffffa3ba13,
    0xffffff8b99b1b21, 0xfff9b2921b93903, 0xfffffffffb08b20, 0xffffffffffffb23,
    0xffffff98aa82832, 0xfffffffff2902a9, 0xfff8a1810a82832, 0xffffffffffff2a1,
    0xfffffffff819831, 0xffffffffffff190, 0xffffffffffff830, 0xfffffffffffffff,
};

// Edge e interpolates from corner c_edge[e][0] to c_edge[e][1] (argument order of the reference's
// sdf_interp calls); corner k of a dual cube sits at octant (k&1 ^ k>>1&1, k>>1&1, k>>2&1).
__constant__ int8_t c_edge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                     {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

// MC corner index of the cell lying on side (sx, sy, sz) (1 = positive) of a dual centre.
__device__ __forceinline__ int octant_slot(int sx, int sy, int sz) {
    return sz * 4 + (sy ? (sx ? 2 : 3) : (sx ? 1 : 0));
}

__device__ __forceinline__ int case_count(int c) {
    uint64_t w = c_mc_cases[c];
    int n = 0;
    while (n < 15 && ((w >> (4 * n)) & 0xF) != 0xF) ++n;
    return n;
}

__global__ void conformal_children_kernel(const int32_t* __restrict__ par, int64_t np, int s, ngs_grid finer,
                                          int has_finer, int32_t* __restrict__ ch, int32_t* __restrict__ keep) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np * 8) return;
    int64_t p = t >> 3;
    int o = (int)(t & 7);
    int x = par[3 * p] + ((o >> 2) & 1) * s, y = par[3 * p + 1] + ((o >> 1) & 1) * s, z = par[3 * p + 2] + (o & 1) * s;
    ch[3 * t] = x;
    ch[3 * t + 1] = y;
    ch[3 * t + 2] = z;
    keep[t] = (has_finer && grid_test(finer, x, y, z)) ? 0 : 1;
}

__global__ void unrefined_kernel(const int32_t* __restrict__ c, int64_t n, ngs_grid finer, int32_t* __restrict__ keep) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keep[i] = grid_test(finer, c[3 * i], c[3 * i + 1], c[3 * i + 2]) ? 0 : 1;
}

__global__ void compact_rows_kernel(const int32_t* __restrict__ rows, int width, const int32_t* __restrict__ keep,
                                    const int32_t* __restrict__ off, int64_t n, int32_t* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !keep[i]) return;
    int64_t o = off[i];
    for (int j = 0; j < width; ++j) out[o * width + j] = rows[i * width + j];
}

__global__ void expand_kernel(const int32_t* __restrict__ leaves, int64_t n, int nu, int s, int32_t* __restrict__ out) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nu3 = nu * nu * nu;
    if (t >= n * nu3) return;
    int64_t l = t / nu3;
    int e = (int)(t % nu3);
    int ex = e / (nu * nu), ey = (e / nu) % nu, ez = e % nu;
    out[3 * t] = leaves[3 * l] * nu + ex * s;
    out[3 * t + 1] = leaves[3 * l + 1] * nu + ey * s;
    out[3 * t + 2] = leaves[3 * l + 2] * nu + ez * s;
}

__global__ void mark_dual_kernel(ngs_grid g, const int32_t* __restrict__ prim, int64_t n, int s) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * 8) return;
    int64_t p = t >> 3;
    int o = (int)(t & 7);
    grid_mark(g, prim[3 * p] + ((o >> 2) & 1) * s, prim[3 * p + 1] + ((o >> 1) & 1) * s, prim[3 * p + 2] + (o & 1) * s);
}

// Count word of a dual centre: registrations (octant slots filled) in bits 0..15 and, when the
// primal values are known, how many of those slots hold a cell whose value is negative in bits
// 16..31, so completeness and the sign change of a cube are read from one word.
constexpr int COUNT_BITS = 16;
constexpr int32_t COUNT_MASK = (1 << COUNT_BITS) - 1;

// One thread per (primal sub-cell, column (x, y) of its side-s boundary lattice): side columns
// (x or y on the cell's boundary) hold s + 1 boundary points along z, interior columns only z = 0
// and z = s. The column's dual centres come from bitmap windows of up to 31 points.
__global__ void dual_accumulate_kernel(ngs_grid cen, const int32_t* __restrict__ prim, int64_t n, int s,
                                       int id_base, const float* __restrict__ values, int32_t* __restrict__ acc,
                                       int32_t* __restrict__ cnt) {
    const int cols = (s + 1) * (s + 1);
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * cols) return;
    const int64_t p = t / cols;
    const int c = (int)(t % cols);
    const int lx = c / (s + 1), ly = c % (s + 1);
    const bool side_col = lx == 0 || lx == s || ly == 0 || ly == s;
    const int gx = prim[3 * p] + lx - cen.origin[0], gy = prim[3 * p + 1] + ly - cen.origin[1];
    const int gz = prim[3 * p + 2] - cen.origin[2];
    const int pid = id_base + (int)p;
    const int32_t neg = (values && values[pid] < 0.f) ? 1 : 0;
    for (int z0 = 0; z0 <= s; z0 += 31) {
        const int len = min(31, s + 1 - z0);
        int32_t r;
        const uint32_t full = grid_column(cen, gx, gy, gz + z0, len, &r);
        uint32_t m = full;
        if (!side_col) m &= (z0 == 0 ? 1u : 0u) | (z0 + len - 1 == s ? 1u << (len - 1) : 0u);
        while (m) {
            // up to 8 points at a time: their count updates are issued together, then the slots
            constexpr int K = 8;
            int dcs[K], zs[K], old[K];
#pragma unroll
            for (int q = 0; q < K; ++q) {
                const int k = m ? __ffs(m) - 1 : -1;
                zs[q] = k;
                dcs[q] = k >= 0 ? r + __popc(full & ((1u << k) - 1u)) : -1;
                m &= m ? m - 1 : 0u;
            }
#pragma unroll
            for (int q = 0; q < K; ++q) {
                if (dcs[q] < 0) continue;
                const int lz = z0 + zs[q];
                const int nsl = (lx == 0 || lx == s ? 1 : 2) * (ly == 0 || ly == s ? 1 : 2) * (lz == 0 || lz == s ? 1 : 2);
                old[q] = atomicAdd(&cnt[dcs[q]], nsl | ((neg * nsl) << COUNT_BITS));
            }
#pragma unroll
            for (int q = 0; q < K; ++q) {
                if (dcs[q] < 0 || (old[q] & COUNT_MASK) >= 8) continue;
                const int loc[3] = {lx, ly, z0 + zs[q]};
                // sides of the dual centre covered by the cell, per axis: bit0 negative side, bit1 positive
                int side[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) side[a] = loc[a] == 0 ? 2 : (loc[a] == s ? 1 : 3);
                for (int sx = 0; sx < 2; ++sx) {
                    if (!((side[0] >> sx) & 1)) continue;
                    for (int sy = 0; sy < 2; ++sy) {
                        if (!((side[1] >> sy) & 1)) continue;
                        for (int sz = 0; sz < 2; ++sz) {
                            if (!((side[2] >> sz) & 1)) continue;
                            const int sl = octant_slot(sx, sy, sz);
                            acc[(int64_t)dcs[q] * 8 + sl] = pid;
                        }
                    }
                }
            }
        }
    }
}

// Complete cubes (8 registrations); with active_only, only those whose corner values change sign
// (negative-slot count strictly between 0 and 8: the cubes marching cubes emits triangles for).
__global__ void complete_flags_kernel(const int32_t* __restrict__ cnt, int64_t n, int active_only,
                                      int32_t* __restrict__ flag) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t c = cnt[i];
    const int neg = (int)((uint32_t)c >> COUNT_BITS);
    flag[i] = ((c & COUNT_MASK) == 8 && (!active_only || (neg > 0 && neg < 8))) ? 1 : 0;
}

__global__ void compact_cubes_kernel(const int32_t* __restrict__ acc, const int32_t* __restrict__ flag,
                                     const int32_t* __restrict__ off, int64_t n, int32_t* __restrict__ cubes) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    int64_t o = off[i];
#pragma unroll
    for (int k = 0; k < 8; ++k) cubes[o * 8 + k] = acc[i * 8 + k];
}

__global__ void corner_pos_kernel(const int32_t* __restrict__ prim, int64_t n, int s, float scale, float* __restrict__ pos) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n) return;
    pos[i] = (float)(prim[i] + s) * scale;
}

// Dual-corner positions straight from the leaves: corner e of leaf l at (leaf * nu + e * s + s) * scale
// per axis, the same integer and float operations as expand_kernel followed by corner_pos_kernel.
__global__ void leaf_corner_pos_kernel(const int32_t* __restrict__ leaves, int64_t n, int nu, int s, float scale,
                                       float* __restrict__ pos) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int nu3 = nu * nu * nu;
    if (t >= n * nu3) return;
    const int64_t l = t / nu3;
    const int e = (int)(t % nu3);
    const int ee[3] = {e / (nu * nu), (e / nu) % nu, e % nu};
#pragma unroll
    for (int a = 0; a < 3; ++a) pos[3 * t + a] = (float)(leaves[3 * l + a] * nu + ee[a] * s + s) * scale;
}

// Field of one level at the dual corners of leaves: corner e of leaf l (e in [0, nu)^3, leaf-major,
// the order of ngs_expand_primal) sits at ((leaf * nu + e * s_leaf + s_leaf) * scale) / voxel_size
// per axis, computed in the same float operations as corner_pos_kernel followed by eval_field_kernel
// (unit field scale), and summed over the 27 voxels in the same order, so the values equal
// evaluating the written positions. For nu = 2, when on every axis the two corners of a leaf fall in
// the same or adjacent cells of the field level, the leaf's 4x4x4 coefficient window is read once
// (16 bitmap columns) for all eight corners instead of 9 columns per corner; an absent voxel adds
// +0, which leaves the sum unchanged.
__device__ __forceinline__ void eval_leaf_corners_one(const ngs_grid& g, const float* __restrict__ field,
                                                      const int32_t* __restrict__ leaves, int64_t l, int s_leaf, int nu,
                                                      float scale, float inv_vs, int accumulate, float* __restrict__ out) {
#pragma clang fp contract(off)
    const int l2 = g.log2_stride;
    const int sd = 1 << l2;
    const float inv_sd = 1.0f / (float)sd;
    const NGS_GLOBAL int32_t* lf = as_global(leaves) + 3 * l;
    const NGS_GLOBAL float* f = as_global(field);
    const int nu3 = nu * nu * nu;
    const int64_t o0 = l * nu3;
    if (nu == 2) {
        int cell[3][2];
        float b[3][2][3];
        bool window = true;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int p = lf[a] * 2;
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const float pos = (float)(p + e * s_leaf + s_leaf) * scale;
                axis_basis(pos * inv_vs, sd, inv_sd, &cell[a][e], b[a][e]);
            }
            const int dc = cell[a][1] - cell[a][0];
            window = window && (dc == 0 || dc == sd);
        }
        if (window) {
            // lattice points cell[a][0] / sd - 1 + j, j in [0, 4), of the leaf's window; per axis and
            // corner the basis values padded to 4 (the corner's 3 voxels start at j = 0 or j = 1)
            const int wx = (cell[0][0] >> l2) - 1 - g.origin[0], wy = (cell[1][0] >> l2) - 1 - g.origin[1];
            const int wz = (cell[2][0] >> l2) - 1 - g.origin[2];
            float w[3][2][4];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const bool sh = cell[a][1] != cell[a][0];
#pragma unroll
                for (int e = 0; e < 2; ++e) {
                    const bool se = e == 1 && sh;
                    w[a][e][0] = se ? 0.f : b[a][e][0];
                    w[a][e][1] = se ? b[a][e][0] : b[a][e][1];
                    w[a][e][2] = se ? b[a][e][1] : b[a][e][2];
                    w[a][e][3] = se ? b[a][e][2] : 0.f;
                }
            }
            float v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = 0.f;
            // one bitmap column per (x, y) of the window; every corner's terms are added in its own
            // (x, y, z) order, and the padded and absent ones add +0
#pragma unroll
            for (int jx = 0; jx < 4; ++jx)
#pragma unroll
                for (int jy = 0; jy < 4; ++jy) {
                    int32_t r;
                    const uint32_t m = grid_column(g, wx + jx, wy + jy, wz, 4, &r);
                    float c[4];
#pragma unroll
                    for (int jz = 0; jz < 4; ++jz) {
                        const bool present = (m >> jz) & 1u;
                        c[jz] = present ? f[r] : 0.f;
                        r += present ? 1 : 0;
                    }
#pragma unroll
                    for (int ex = 0; ex < 2; ++ex)
#pragma unroll
                        for (int ey = 0; ey < 2; ++ey) {
                            const float wxy = w[0][ex][jx] * w[1][ey][jy];
#pragma unroll
                            for (int ez = 0; ez < 2; ++ez)
#pragma unroll
                                for (int jz = 0; jz < 4; ++jz)
                                    v[(ex * 2 + ey) * 2 + ez] += (wxy * w[2][ez][jz]) * c[jz];
                        }
                }
#pragma unroll
            for (int e = 0; e < 8; ++e) out[o0 + e] = (accumulate ? out[o0 + e] : 0.f) + v[e];
            return;
        }
    }
    // general case: each corner on its own, as eval_field_kernel
    for (int e = 0; e < nu3; ++e) {
        const int ee[3] = {e / (nu * nu), (e / nu) % nu, e % nu};
        int c[3];
        float bb[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float pos = (float)(lf[a] * nu + ee[a] * s_leaf + s_leaf) * scale;
            axis_basis(pos * inv_vs, sd, inv_sd, &c[a], bb[a]);
        }
        const int lx = (c[0] >> l2) - g.origin[0], ly = (c[1] >> l2) - g.origin[1], lz = (c[2] >> l2) - g.origin[2];
        float v = 0.f;
        for (int kx = 0; kx < 3; ++kx)
            for (int ky = 0; ky < 3; ++ky) {
                int32_t id;
                const uint32_t m = grid_column(g, lx + kx - 1, ly + ky - 1, lz - 1, 3, &id);
                if (!m) continue;
                for (int kz = 0; kz < 3; ++kz) {
                    if (!((m >> kz) & 1u)) continue;
                    v += (bb[0][kx] * bb[1][ky] * bb[2][kz]) * f[id++];
                }
            }
        out[o0 + e] = (accumulate ? out[o0 + e] : 0.f) + v;
    }
}

__global__ void eval_leaf_corners_kernel(ngs_grid g, const float* __restrict__ field, const int32_t* __restrict__ leaves,
                                         int64_t n, int s_leaf, int nu, float scale, float inv_vs, int accumulate,
                                         float* __restrict__ out) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n) eval_leaf_corners_one(g, field, leaves, l, s_leaf, nu, scale, inv_vs, accumulate, out);
}

// Up to 8 field levels summed into the corner values in the given order, each exactly as one
// eval_leaf_corners_kernel launch (out = out + level value): the same values as launching per level,
// with the leaf reads and the output kept in cache between levels.
constexpr int MAX_FIELD_LEVELS = 8;
struct FieldLevels {
    ngs_grid g[MAX_FIELD_LEVELS];
    const float* field[MAX_FIELD_LEVELS];
    int32_t n;
};

__global__ void eval_leaf_corners_levels_kernel(FieldLevels fl, const int32_t* __restrict__ leaves, int64_t n,
                                                int s_leaf, int nu, float scale, float inv_vs, int accumulate,
                                                float* __restrict__ out) {
    const int64_t l = xcd_contiguous_block() * blockDim.x + threadIdx.x;
    if (l >= n) return;
    for (int k = 0; k < fl.n; ++k)
        eval_leaf_corners_one(fl.g[k], fl.field[k], leaves, l, s_leaf, nu, scale, inv_vs, accumulate || k > 0, out);
}

__device__ __forceinline__ int cube_case(const int32_t* ids, const float* __restrict__ val) {
    int c = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (val[ids[k]] < 0.f) c |= 1 << k;
    return c;
}

// Distinct cube edges that case c's triangles use.
__device__ __forceinline__ int case_edge_count(int c) {
    const uint64_t w = c_mc_cases[c];
    uint32_t seen = 0;
    for (int n = 0; n < 15; ++n) {
        const int e = (int)((w >> (4 * n)) & 0xF);
        if (e == 0xF) break;
        seen |= 1u << e;
    }
    return __popc(seen);
}

__global__ void mc_count_kernel(const int32_t* __restrict__ cubes, int64_t k, const float* __restrict__ val,
                                int32_t* __restrict__ vc, int32_t* __restrict__ ec) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    const int c = cube_case(cubes + 8 * i, val);
    vc[i] = case_count(c);
    if (ec) ec[i] = case_edge_count(c);
}

// dual_mc.cu:20-31 semantics: snap to an endpoint whose value is within 1e-5 of zero, or to p1
// when the two values are within 1e-5; otherwise linear interpolation of the zero crossing.
__device__ __forceinline__ void sdf_interp(const float* p1, const float* p2, float v1, float v2, float* out) {
#pragma clang fp contract(off)
    if (fabsf(0.0f - v1) < 1.0e-5f) { out[0] = p1[0]; out[1] = p1[1]; out[2] = p1[2]; return; }
    if (fabsf(0.0f - v2) < 1.0e-5f) { out[0] = p2[0]; out[1] = p2[1]; out[2] = p2[2]; return; }
    if (fabsf(v1 - v2) < 1.0e-5f) { out[0] = p1[0]; out[1] = p1[1]; out[2] = p1[2]; return; }
    float w2 = (0.0f - v1) / (v2 - v1);
    float w1 = 1.0f - w2;
    out[0] = p1[0] * w1 + p2[0] * w2;
    out[1] = p1[1] * w1 + p2[1] * w2;
    out[2] = p1[2] * w1 + p2[2] * w2;
}

// Per distinct edge of a cube (slots from eoff[i], in the order the case's triangles first use the
// edges): key (larger id << id_bits | smaller id) and tag = slot | (ia > ib) << 31, where (ia, ib) are
// the edge's corners in sdf_interp argument order; per triangle vertex (from voff[i], cube order then
// case-table order) the slot of its edge in vslot. The first slot of a key in slot order is the edge
// of the key's first triangle vertex in vertex order (same cube, first-used edge), so positions
// computed once per distinct key from its first slot are those of the per-vertex form.
__global__ void mc_emit_kernel(const int32_t* __restrict__ cubes, int64_t k, const float* __restrict__ val,
                               const int32_t* __restrict__ voff, const int32_t* __restrict__ eoff, int id_bits,
                               uint64_t* __restrict__ keys, uint32_t* __restrict__ tags,
                               int64_t* __restrict__ vslot) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    const int32_t* ids = cubes + 8 * i;
    const int c = cube_case(ids, val);
    uint64_t w = c_mc_cases[c];
    int64_t v = voff[i];
    const int64_t e0 = eoff[i];
    uint64_t slot_of = 0;  // 4 bits per cube edge: its slot within the cube
    uint32_t seen = 0;
    int ns = 0;
    for (int n = 0; n < 15; ++n, ++v) {
        int e = (int)((w >> (4 * n)) & 0xF);
        if (e == 0xF) break;
        if (!((seen >> e) & 1u)) {
            seen |= 1u << e;
            slot_of |= (uint64_t)ns << (4 * e);
            const int ia = ids[c_edge[e][0]], ib = ids[c_edge[e][1]];
            const uint32_t hi = (uint32_t)max(ia, ib), lo = (uint32_t)min(ia, ib);
            keys[e0 + ns] = ((uint64_t)hi << id_bits) | lo;
            tags[e0 + ns] = (uint32_t)(e0 + ns) | (ia > ib ? 0x80000000u : 0u);
            ++ns;
        }
        vslot[v] = e0 + (int64_t)((slot_of >> (4 * e)) & 0xFu);
    }
}

// faces[v] <- rank[faces[v]] (slot -> vertex rank), in place.
__global__ void remap_slots_kernel(int64_t* __restrict__ faces, int64_t n, const int64_t* __restrict__ rank) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) faces[v] = rank[faces[v]];
}

__global__ void run_heads_kernel(const uint64_t* __restrict__ k, int64_t n, int32_t* __restrict__ head) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) head[i] = (i == 0 || k[i] != k[i - 1]) ? 1 : 0;
}

__global__ void dedup_scatter_kernel(const int32_t* __restrict__ head, const int32_t* __restrict__ off,
                                     const uint64_t* __restrict__ skeys, const uint32_t* __restrict__ stags, int64_t n,
                                     int id_bits, const float* __restrict__ pos, const float* __restrict__ val,
                                     int64_t* __restrict__ faces, float* __restrict__ verts) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // off[i] = number of run heads strictly before i; the run of element i has rank off[i+1]-1
    const int32_t rank = off[i] + head[i] - 1;
    const uint32_t tag = stags[i];
    faces[tag & 0x7FFFFFFFu] = rank;
    if (head[i]) {
        // the first occurrence (the sort is stable and tags ascend in emission order)
        const uint64_t key = skeys[i];
        const int32_t hi = (int32_t)(key >> id_bits), lo = (int32_t)(key & ((1ull << id_bits) - 1));
        const int32_t ia = (tag >> 31) ? hi : lo, ib = (tag >> 31) ? lo : hi;
        const float p1[3] = {pos[3 * (int64_t)ia], pos[3 * (int64_t)ia + 1], pos[3 * (int64_t)ia + 2]};
        const float p2[3] = {pos[3 * (int64_t)ib], pos[3 * (int64_t)ib + 1], pos[3 * (int64_t)ib + 2]};
        float out[3];
        sdf_interp(p1, p2, val[ia], val[ib], out);
        verts[3 * (int64_t)rank] = out[0];
        verts[3 * (int64_t)rank + 1] = out[1];
        verts[3 * (int64_t)rank + 2] = out[2];
    }
}

// All sub-cell corners of a leaf form the lattice leaf*nu + o*s, o in [0, nu]^3. Along z the lattice
// points of one (x, y) column are bits of the same or the next word, so each column is marked with
// one or two multi-bit atomicOr instead of one atomic per sub-cell corner.
__global__ void mark_dual_leaves_kernel(ngs_grid g, const int32_t* __restrict__ leaves, int64_t n, int s, int nu) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int cols = (nu + 1) * (nu + 1);
    if (t >= n * cols) return;
    int64_t l = t / cols;
    int c = (int)(t % cols);
    int ox = c / (nu + 1), oy = c % (nu + 1);
    int x = leaves[3 * l] * nu + ox * s, y = leaves[3 * l + 1] * nu + oy * s, z0 = leaves[3 * l + 2] * nu;
    int64_t cur_word = -1;
    uint32_t mask = 0;
    for (int oz = 0; oz <= nu; ++oz) {
        int64_t lin;
        if (!grid_cell(g, x, y, z0 + oz * s, &lin)) continue;
        int64_t w = lin >> 5;
        if (w != cur_word) {
            if (mask) grid_mark_word(g, cur_word, mask);
            cur_word = w;
            mask = 0;
        }
        mask |= 1u << (lin & 31);
    }
    if (mask) grid_mark_word(g, cur_word, mask);
}

// Even bits of the result hold the bits of m (bit k -> bit 2k).
__device__ __forceinline__ uint64_t spread_bits(uint32_t m) {
    uint64_t x = m;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}

// Dual centres of stride-1 leaves with n_upsample 2, one output bitmap word per thread and no
// atomics: the centres are {2 L + o, o in [0, 2]^3} over the leaves L, so a centre column (x, y)
// takes the z-bits of the leaf columns X with 2X <= x <= 2X + 2 (the same for y), each leaf bit Z
// setting centre bits 2Z .. 2Z + 2. The word is written whole (it replaces any earlier content);
// coarser leaves are marked afterwards with atomics.
__global__ void mark_dual_leaf_words_kernel(ngs_grid cen, ngs_grid lg) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= cen.nwords) return;
    const int64_t d2 = cen.dims[2], d12 = (int64_t)cen.dims[1] * d2;
    const int64_t nbits = (int64_t)cen.dims[0] * d12;
    int64_t lin = w * 32;
    const int64_t end = min(lin + 32, nbits);
    uint32_t out = 0;
    while (lin < end) {
        // one segment: bits of the word inside one centre column
        const int x = (int)(lin / d12), y = (int)((lin / d2) % cen.dims[1]), z = (int)(lin % d2);
        const int len = (int)min(end - lin, d2 - z);
        const int cx = x + cen.origin[0], cy = y + cen.origin[1], cz = z + cen.origin[2];
        // leaf rows covering centre z in [cz, cz + len): Z from floor((cz - 2) / 2) to floor((cz + len - 1) / 2)
        const int z0 = (cz - 2) >> 1;
        const int nz = ((cz + len - 1) >> 1) - z0 + 1;  // <= 18
        uint64_t seg = 0;
        for (int X = (cx - 1) >> 1; X <= (cx >> 1); ++X) {
            if (2 * X > cx || 2 * X + 2 < cx) continue;
            for (int Y = (cy - 1) >> 1; Y <= (cy >> 1); ++Y) {
                if (2 * Y > cy || 2 * Y + 2 < cy) continue;
                int32_t r;
                const uint32_t m = grid_column(lg, X - lg.origin[0], Y - lg.origin[1], z0 - lg.origin[2], nz, &r);
                if (!m) continue;
                const uint64_t sp = spread_bits(m);
                seg |= sp | (sp << 1) | (sp << 2);  // bit 2k + oz <-> centre z = 2 (z0 + k) + oz
            }
        }
        // centre z = cz + i sits at bit 2 z0 + (bit index) => shift by cz - 2 z0 (2 or 3)
        seg >>= (cz - 2 * z0);
        const uint32_t segm = (uint32_t)seg & (len >= 32 ? 0xFFFFFFFFu : ((1u << len) - 1u));
        out |= segm << (lin - w * 32);
        lin += len;
    }
    cen.bits[w] = out;
}

// Leaf levels of the gathered dual-cube assembly (n_upsample 2): per level the leaf bitmap (stride
// s = 1 << log2_stride), the map from a leaf's bitmap rank to its position in the leaf list, and the
// id of its first sub-cell.
constexpr int MAX_LEAF_LEVELS = 8;
struct LeafLevels {
    ngs_grid g[MAX_LEAF_LEVELS];
    const int32_t* index[MAX_LEAF_LEVELS];
    int32_t base[MAX_LEAF_LEVELS];
    int32_t n;
};

// Dual-cube assembly gathered per dual centre, n_upsample 2, no atomics. Sub-cells of a level-s leaf
// L are [2 L + e s, 2 L + (e + 1) s] per axis in upsampled units (e in {0, 1}), so the octant
// (sx, sy, sz) of centre q lies in the sub-cell with lower corner c = floor((q - 1 + s_a) / s) s per
// axis, of leaf floor(c / 2s) s, sub-cell e = (c / s) mod 2. The leaves partition the covered space,
// so each octant has at most one covering sub-cell, found at the finest level that has its leaf;
// that is the set of registrations the per-sub-cell form makes (every centre lies on the boundary of
// the sub-cells around it). Level 0 (stride 1) is read by bitmap columns. Writes each centre's count
// word (slots | negative-valued slots << 16) and its slots.
__global__ void dual_gather_cubes_kernel(const int32_t* __restrict__ centres, int64_t dcount, LeafLevels lv,
                                         const float* __restrict__ values, int32_t* __restrict__ acc,
                                         int32_t* __restrict__ cnt) {
    const int64_t dc = xcd_contiguous_block() * blockDim.x + threadIdx.x;
    if (dc >= dcount) return;
    const int q[3] = {centres[3 * dc], centres[3 * dc + 1], centres[3 * dc + 2]};
    int32_t ids[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) ids[k] = -1;
    int n = 0;
    // finest level by bitmap columns (when it is a stride-1 level)
    int l0 = 0;
    if (lv.n > 0 && lv.g[0].log2_stride == 0) {
        l0 = 1;
        const ngs_grid& lg = lv.g[0];
        const int lz0 = (q[2] - 1) >> 1, nz = (q[2] >> 1) - lz0 + 1;  // leaf z of sz = 0 and sz = 1
#pragma unroll
        for (int sx = 0; sx < 2; ++sx)
#pragma unroll
            for (int sy = 0; sy < 2; ++sy) {
                const int cx = q[0] - 1 + sx, cy = q[1] - 1 + sy;
                int32_t r;
                const uint32_t m = grid_column(lg, (cx >> 1) - lg.origin[0], (cy >> 1) - lg.origin[1],
                                               lz0 - lg.origin[2], nz, &r);
#pragma unroll
                for (int sz = 0; sz < 2; ++sz) {
                    const int cz = q[2] - 1 + sz;
                    const int k = (cz >> 1) - lz0;  // bit of the leaf in the column window
                    if (!((m >> k) & 1u)) continue;
                    const int32_t rank = r + __popc(m & ((1u << k) - 1u));
                    const int e = ((cx & 1) * 2 + (cy & 1)) * 2 + (cz & 1);
                    ids[octant_slot(sx, sy, sz)] = lv.base[0] + lv.index[0][rank] * 8 + e;
                    ++n;
                }
            }
    }
    // octants left: coarser levels, finest first. Per level the probes of all missing octants go
    // out together (bitmap words, then rank words, then leaf indices, each phase's loads issued
    // before any is used): one chain of three loads per level instead of one per octant and level.
    for (int l = l0; l < lv.n && n < 8; ++l) {
        const ngs_grid& g = lv.g[l];
        const int k = g.log2_stride;
        int64_t lin[8];
        bool ok[8];
        int e[8];
#pragma unroll
        for (int o = 0; o < 8; ++o) {
            const int sx = (o >> 2) & 1, sy = (o >> 1) & 1, sz = o & 1;
            const int c0 = (q[0] - 1 + sx) >> k, c1 = (q[1] - 1 + sy) >> k, c2 = (q[2] - 1 + sz) >> k;  // units of s
            e[o] = ((c0 & 1) * 2 + (c1 & 1)) * 2 + (c2 & 1);
            // the leaf's coordinates ((c >> 1) << k) are multiples of its stride 2^k, as grid_find requires
            lin[o] = 0;
            ok[o] = ids[octant_slot(sx, sy, sz)] < 0 &&
                    grid_cell(g, (c0 >> 1) << k, (c1 >> 1) << k, (c2 >> 1) << k, &lin[o]);
        }
        // loads without branches: a probe that is not needed reads element 0, which every level has
        // (levels without leaves are not passed), and its result is discarded
        const NGS_GLOBAL uint32_t* bits = as_global(g.bits);
        const NGS_GLOBAL int32_t* rnk = as_global(g.rank);
        const NGS_GLOBAL int32_t* index = as_global(lv.index[l]);
        uint32_t w[8];
#pragma unroll
        for (int o = 0; o < 8; ++o) w[o] = bits[ok[o] ? lin[o] >> 5 : 0];
#pragma unroll
        for (int o = 0; o < 8; ++o) ok[o] = ok[o] && ((w[o] >> (lin[o] & 31)) & 1u);
        int32_t rw[8];
#pragma unroll
        for (int o = 0; o < 8; ++o) rw[o] = rnk[ok[o] ? lin[o] >> 5 : 0];
        int32_t li[8];
#pragma unroll
        for (int o = 0; o < 8; ++o)
            li[o] = index[ok[o] ? rw[o] + __popc(w[o] & ((1u << (lin[o] & 31)) - 1u)) : 0];
        const int32_t base = lv.base[l];
#pragma unroll
        for (int o = 0; o < 8; ++o) {  // selects, not branches: the index loads above stay grouped
            int32_t& id = ids[octant_slot((o >> 2) & 1, (o >> 1) & 1, o & 1)];
            id = ok[o] ? base + li[o] * 8 + e[o] : id;
            n += ok[o] ? 1 : 0;
        }
    }
    // only complete centres become cubes: the others write no slots and read no values
    int neg = 0;
    if (n == 8) {
        if (values) {
#pragma unroll
            for (int k = 0; k < 8; ++k) neg += values[ids[k]] < 0.f ? 1 : 0;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[dc * 8 + k] = ids[k];
    }
    cnt[dc] = n | (neg << COUNT_BITS);
}

// Dual-cube assembly for leaves of stride 1 (the finest level): one thread per (leaf, lattice point
// o in [0, nu]^3) registers every sub-cell of the leaf that has the point on its boundary, with one
// aggregated count update. Per axis, o = 0 is covered by sub-cell 0 from the positive side, o = nu
// by sub-cell nu-1 from the negative side, and 0 < o < nu by sub-cells o-1 (negative) and o
// (positive). With values, the number of registered sub-cells with a negative value goes to the
// high half of the count word.
__global__ void dual_accumulate_leaves_kernel(ngs_grid cen, const int32_t* __restrict__ leaves, int64_t n, int nu,
                                              int id_base, const float* __restrict__ values,
                                              int32_t* __restrict__ acc, int32_t* __restrict__ cnt) {
    const int per = (nu + 1) * (nu + 1) * (nu + 1);
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * per) return;
    const int64_t l = t / per;
    const int r = (int)(t % per);
    const int o[3] = {r / ((nu + 1) * (nu + 1)), (r / (nu + 1)) % (nu + 1), r % (nu + 1)};
    const int dc = grid_find(cen, leaves[3 * l] * nu + o[0], leaves[3 * l + 1] * nu + o[1], leaves[3 * l + 2] * nu + o[2]);
    if (dc < 0) return;
    // per axis: up to two (sub-cell index, side) pairs; side 1 = cell on the positive side
    int ce[3][2], sd[3][2], nc[3];
    int nsl = 1;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        nc[a] = 0;
        if (o[a] > 0) { ce[a][nc[a]] = o[a] - 1; sd[a][nc[a]] = 0; ++nc[a]; }
        if (o[a] < nu) { ce[a][nc[a]] = o[a]; sd[a][nc[a]] = 1; ++nc[a]; }
        nsl *= nc[a];
    }
    const int64_t base = (int64_t)id_base + l * (int64_t)(nu * nu * nu);
    int neg = 0;
    if (values)
        for (int i = 0; i < nc[0]; ++i)
            for (int j = 0; j < nc[1]; ++j)
                for (int k = 0; k < nc[2]; ++k)
                    neg += values[base + (ce[0][i] * nu + ce[1][j]) * nu + ce[2][k]] < 0.f ? 1 : 0;
    const int old = atomicAdd(&cnt[dc], nsl | (neg << COUNT_BITS));
    if ((old & COUNT_MASK) >= 8) return;
    for (int i = 0; i < nc[0]; ++i)
        for (int j = 0; j < nc[1]; ++j)
            for (int k = 0; k < nc[2]; ++k) {
                const int sl = octant_slot(sd[0][i], sd[1][j], sd[2][k]);
                const int e = (ce[0][i] * nu + ce[1][j]) * nu + ce[2][k];
                acc[(int64_t)dc * 8 + sl] = (int32_t)(base + e);
            }
}

__global__ void widen_kernel(const int32_t* a, int64_t* b) { *b = *a; }
__global__ void set_i64_kernel(int64_t* p, int64_t v) { *p = v; }

static size_t al256(size_t x) { return (x + 255) / 256 * 256; }

// ------------------------------------------------------------------------ primal sparse MC
// Leaf l carries an (r+1)^3 block of samples, x-major: sdf[l (r+1)^3 + (x (r+1) + y)(r+1) + z], and a
// base corner (integer-valued, in leaf units); sub-cube (rx, ry, rz) of the leaf spans samples
// (rx..rx+1, ry..ry+1, rz..rz+1). One thread per sub-cube, z fastest, so a wave reads consecutive
// samples. Corner k of a sub-cube sits at offset (k&1 ^ k>>1&1, k>>1&1, k>>2&1), as for dual cubes.
__device__ __forceinline__ void corner_offset(int k, int* o) {
    o[0] = (k & 1) ^ ((k >> 1) & 1);
    o[1] = (k >> 1) & 1;
    o[2] = (k >> 2) & 1;
}

__device__ __forceinline__ int primal_case(const float* __restrict__ sdf, int64_t leaf, int r, int rx, int ry, int rz,
                                           float* v) {
    const int r1 = r + 1;
    const int64_t b = leaf * (int64_t)r1 * r1 * r1;
    int c = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        int o[3];
        corner_offset(k, o);
        v[k] = sdf[b + ((int64_t)(rx + o[0]) * r1 + (ry + o[1])) * r1 + (rz + o[2])];
        if (v[k] < 0.f) c |= 1 << k;
    }
    return c;
}

__global__ void primal_count_kernel(const float* __restrict__ sdf, int64_t nleaf, int r, int32_t* __restrict__ vc) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r3 = (int64_t)r * r * r;
    if (t >= nleaf * r3) return;
    const int64_t leaf = t / r3;
    const int sub = (int)(t - leaf * r3);
    float v[8];
    vc[t] = case_count(primal_case(sdf, leaf, r, sub / (r * r), (sub / r) % r, sub % r, v));
}

// sdf_interp of primal_mc.cu:7-19, also returning the weight of p1 (1 when snapped to p1, 0 when
// snapped to p2).
__device__ __forceinline__ float sdf_interp_w(const float* p1, const float* p2, float v1, float v2, float* out) {
#pragma clang fp contract(off)
    if (fabsf(0.0f - v1) < 1.0e-5f) { out[0] = p1[0]; out[1] = p1[1]; out[2] = p1[2]; return 1.0f; }
    if (fabsf(0.0f - v2) < 1.0e-5f) { out[0] = p2[0]; out[1] = p2[1]; out[2] = p2[2]; return 0.0f; }
    if (fabsf(v1 - v2) < 1.0e-5f) { out[0] = p1[0]; out[1] = p1[1]; out[2] = p1[2]; return 1.0f; }
    float w2 = (0.0f - v1) / (v2 - v1);
    float w1 = 1.0f - w2;
    out[0] = p1[0] * w1 + p2[0] * w2;
    out[1] = p1[1] * w1 + p2[1] * w2;
    out[2] = p1[2] * w1 + p2[2] * w2;
    return w1;
}

// Triangles of every sub-cube in (leaf, rx, ry, rz) order at voff[t] / 3 (primal_mc.cu:59-174): per
// triangle vertex the position (base + offset / r), the edge id (global sub-lattice corner of the
// edge's lower end and the edge axis), the flat sample indices of the edge's two ends (sdf_interp
// argument order) and the weight of the first end.
__global__ void primal_emit_kernel(const float* __restrict__ base, const float* __restrict__ sdf, int64_t nleaf, int r,
                                   const int32_t* __restrict__ voff, float* __restrict__ tris,
                                   int32_t* __restrict__ vid, int32_t* __restrict__ src, float* __restrict__ sw) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r3 = (int64_t)r * r * r;
    if (t >= nleaf * r3) return;
    const int64_t leaf = t / r3;
    const int sub = (int)(t - leaf * r3);
    const int rr[3] = {sub / (r * r), (sub / r) % r, sub % r};
    float v[8];
    const int c = primal_case(sdf, leaf, r, rr[0], rr[1], rr[2], v);
    uint64_t w = c_mc_cases[c];
    if ((w & 0xF) == 0xF) return;
    const float sbs = 1.0f / (float)r;
    const float b[3] = {base[3 * leaf], base[3 * leaf + 1], base[3 * leaf + 2]};
    const int r1 = r + 1;
    const int32_t sid0 = r1 * r1 * r1;
    int64_t o = voff[t];
    for (int n = 0; n < 15; ++n, ++o) {
        const int e = (int)((w >> (4 * n)) & 0xF);
        if (e == 0xF) break;
        int oa[3], ob[3];
        corner_offset(c_edge[e][0], oa);
        corner_offset(c_edge[e][1], ob);
        float p1[3], p2[3], q[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            p1[a] = b[a] + (float)(rr[a] + oa[a]) * sbs;
            p2[a] = b[a] + (float)(rr[a] + ob[a]) * sbs;
        }
        const float w1 = sdf_interp_w(p1, p2, v[c_edge[e][0]], v[c_edge[e][1]], q);
        int axis = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            tris[3 * o + a] = q[a];
            vid[4 * o + a] = (int)b[a] * r + rr[a] + min(oa[a], ob[a]);
            if (oa[a] != ob[a]) axis = a;
        }
        vid[4 * o + 3] = axis;
        src[2 * o] = sid0 * (int32_t)leaf + ((rr[0] + oa[0]) * r1 + (rr[1] + oa[1])) * r1 + (rr[2] + oa[2]);
        src[2 * o + 1] = sid0 * (int32_t)leaf + ((rr[0] + ob[0]) * r1 + (rr[1] + ob[1])) * r1 + (rr[2] + ob[2]);
        sw[o] = w1;
    }
}

// Order-preserving 64-bit key of an edge id row (x, y, z, axis): biased coordinates in bx, by, bz
// bits above the 2-bit axis, so key order is the lexicographic order of the signed rows.
struct EdgeKeyLayout {
    int32_t org[3];
    int sh[3];
};

__global__ void primal_pack_kernel(const int32_t* __restrict__ vid, int64_t n, EdgeKeyLayout kl,
                                   uint64_t* __restrict__ keys, uint32_t* __restrict__ tags) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t k = (uint64_t)(uint32_t)vid[4 * i + 3];
#pragma unroll
    for (int a = 0; a < 3; ++a) k |= (uint64_t)(uint32_t)(vid[4 * i + a] - kl.org[a]) << kl.sh[a];
    keys[i] = k;
    tags[i] = (uint32_t)i;
}

// Distinct edges in key order; every record of a vertex is copied from the edge's first occurrence
// in triangle order (the stable sort keeps occurrences in emission order). start[rank] = position
// of the edge's run in the sorted list.
__global__ void primal_scatter_kernel(const int32_t* __restrict__ head, const int32_t* __restrict__ off,
                                      const uint32_t* __restrict__ stags, int64_t n, const float* __restrict__ tris,
                                      const int32_t* __restrict__ vid, const int32_t* __restrict__ src,
                                      const float* __restrict__ sw, int64_t* __restrict__ faces,
                                      float* __restrict__ verts, int32_t* __restrict__ unq, int32_t* __restrict__ vsrc,
                                      float* __restrict__ vw, int32_t* __restrict__ start) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t rank = off[i] + head[i] - 1;
    const int64_t tag = stags[i];
    faces[tag] = rank;
    if (!head[i]) return;
    start[rank] = (int32_t)i;
#pragma unroll
    for (int a = 0; a < 3; ++a) verts[3 * (int64_t)rank + a] = tris[3 * tag + a];
#pragma unroll
    for (int a = 0; a < 4; ++a) unq[4 * (int64_t)rank + a] = vid[4 * tag + a];
    vsrc[2 * (int64_t)rank] = src[2 * tag];
    vsrc[2 * (int64_t)rank + 1] = src[2 * tag + 1];
    vw[rank] = sw[tag];
}

// compute_normal (ops.py:85-101): per vertex the sum of the unnormalised normals (v1 - v0) x (v2 - v0)
// of its triangles, in triangle order (the vertex's run of sorted occurrences), divided by its
// length + 1e-6.
__global__ void primal_normals_kernel(const int32_t* __restrict__ head, const uint32_t* __restrict__ stags,
                                      const int32_t* __restrict__ start, const int64_t* __restrict__ nverts, int64_t n,
                                      const int64_t* __restrict__ faces, const float* __restrict__ verts,
                                      float* __restrict__ normals) {
#pragma clang fp contract(off)
    const int64_t rank = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (rank >= *nverts) return;
    float s[3] = {0.f, 0.f, 0.f};
    int64_t i = start[rank];
    do {
        const int64_t tri = stags[i] / 3;
        const float* v0 = verts + 3 * faces[3 * tri];
        const float* v1 = verts + 3 * faces[3 * tri + 1];
        const float* v2 = verts + 3 * faces[3 * tri + 2];
        const float a[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
        const float b[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
        s[0] += a[1] * b[2] - a[2] * b[1];
        s[1] += a[2] * b[0] - a[0] * b[2];
        s[2] += a[0] * b[1] - a[1] * b[0];
        ++i;
    } while (i < n && !head[i]);
    const float len = sqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]) + 1e-6f;
#pragma unroll
    for (int a = 0; a < 3; ++a) normals[3 * rank + a] = s[a] / len;
}

// MarchingCubes.backward (ops.py:40-58): g = bandwidth (grad_vert . normal); the first sample of the
// vertex's edge receives w g, the second (1 - w) g.
__global__ void primal_backward_kernel(const float* __restrict__ gv, const float* __restrict__ nrm,
                                       const int32_t* __restrict__ vsrc, const float* __restrict__ vw, int64_t nv,
                                       float bandwidth, float* __restrict__ grid) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    const float g = bandwidth * (gv[3 * i] * nrm[3 * i] + gv[3 * i + 1] * nrm[3 * i + 1] + gv[3 * i + 2] * nrm[3 * i + 2]);
    const float w = vw[i];
    atomicAdd(&grid[vsrc[2 * i]], w * g);
    atomicAdd(&grid[vsrc[2 * i + 1]], (1.0f - w) * g);
}

}  // namespace ngs

using namespace ngs;

extern "C" {

int ngs_conformal_children(const int32_t* parents, int64_t np, int32_t stride, ngs_grid finer_parents, int32_t has_finer,
                           int32_t* children, int32_t* keep, ngs_stream_t s) {
    if (np <= 0) return NGS_OK;
    conformal_children_kernel<<<grid1d(np * 8, 256), 256, 0, S(s)>>>(parents, np, stride, finer_parents, has_finer,
                                                                      children, keep);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_unrefined_flags(const int32_t* coords, int64_t n, ngs_grid finer_parents, int32_t* keep, ngs_stream_t s) {
    if (n <= 0) return NGS_OK;
    unrefined_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(coords, n, finer_parents, keep);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_compact_rows(const int32_t* rows, int32_t width, const int32_t* keep, int64_t n, int32_t* out, int32_t* offsets,
                     void* workspace, ngs_stream_t s) {
    NGS_REQUIRE(width >= 1 && workspace && offsets, "ngs_compact_rows: bad arguments");
    if (n <= 0) return NGS_OK;
    int r = scan_i32_internal(keep, n, offsets, workspace, S(s));
    if (r != NGS_OK) return r;
    compact_rows_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(rows, width, keep, offsets, n, out);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_expand_primal(const int32_t* leaves, int64_t n, int32_t n_upsample, int32_t stride, int32_t* out, ngs_stream_t s) {
    NGS_REQUIRE(n_upsample >= 1, "ngs_expand_primal: n_upsample must be >= 1");
    int64_t tot = n * n_upsample * n_upsample * n_upsample;
    if (tot <= 0) return NGS_OK;
    expand_kernel<<<grid1d(tot, 256), 256, 0, S(s)>>>(leaves, n, n_upsample, stride, out);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_mark_dual_centres(ngs_grid g, const int32_t* primal, int64_t n, int32_t stride, ngs_stream_t s) {
    if (n <= 0) return NGS_OK;
    mark_dual_kernel<<<grid1d(n * 8, 256), 256, 0, S(s)>>>(g, primal, n, stride);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_dual_accumulate(ngs_grid centres, const int32_t* primal, int64_t n, int32_t stride, int32_t id_base,
                        const float* values, int32_t* acc, int32_t* count, ngs_stream_t s) {
    NGS_REQUIRE(stride >= 1 && centres.log2_stride == 0, "ngs_dual_accumulate: bad stride");
    if (n <= 0) return NGS_OK;
    const int64_t tot = n * (int64_t)(stride + 1) * (stride + 1);
    dual_accumulate_kernel<<<grid1d(tot, 256), 256, 0, S(s)>>>(centres, primal, n, stride, id_base, values, acc,
                                                                count);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_dual_compact(const int32_t* acc, const int32_t* count, int32_t active_only, int64_t dc, int32_t* cubes,
                     int32_t* flags, int32_t* offsets, void* workspace, ngs_stream_t s) {
    if (dc <= 0) return NGS_OK;
    NGS_REQUIRE(workspace && flags && offsets, "ngs_dual_compact: bad arguments");
    complete_flags_kernel<<<grid1d(dc, 256), 256, 0, S(s)>>>(count, dc, active_only ? 1 : 0, flags);
    NGS_CHECK_LAUNCH();
    int r = scan_i32_internal(flags, dc, offsets, workspace, S(s));
    if (r != NGS_OK) return r;
    compact_cubes_kernel<<<grid1d(dc, 256), 256, 0, S(s)>>>(acc, flags, offsets, dc, cubes);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_dual_corner_positions(const int32_t* primal, int64_t n, int32_t stride, float scale, float* pos, ngs_stream_t s) {
    if (n <= 0) return NGS_OK;
    corner_pos_kernel<<<grid1d(3 * n, 256), 256, 0, S(s)>>>(primal, n, stride, scale, pos);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_mc_count(const int32_t* cubes, int64_t k, const float* values, int32_t* vcount, int32_t* ecount,
                 ngs_stream_t s) {
    if (k <= 0) return NGS_OK;
    NGS_REQUIRE(cubes && values && vcount, "ngs_mc_count: null argument");
    mc_count_kernel<<<grid1d(k, 256), 256, 0, S(s)>>>(cubes, k, values, vcount, ecount);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_mc_emit(const int32_t* cubes, int64_t k, const float* values, const int32_t* voffset, const int32_t* eoffset,
                int32_t id_bits, uint64_t* ekeys, uint32_t* etags, int64_t* vslot, ngs_stream_t s) {
    NGS_REQUIRE(id_bits >= 1 && id_bits <= 31, "ngs_mc_emit: id_bits must be in [1, 31]");
    if (k <= 0) return NGS_OK;
    NGS_REQUIRE(cubes && values && voffset && eoffset && ekeys && etags && vslot, "ngs_mc_emit: null argument");
    mc_emit_kernel<<<grid1d(k, 128), 128, 0, S(s)>>>(cubes, k, values, voffset, eoffset, id_bits, ekeys, etags, vslot);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_mc_remap_faces(int64_t* faces, int64_t n, const int64_t* slot_rank, ngs_stream_t s) {
    if (n <= 0) return NGS_OK;
    NGS_REQUIRE(faces && slot_rank, "ngs_mc_remap_faces: null argument");
    remap_slots_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(faces, n, slot_rank);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

static int dedup_layout(int64_t n, int key_bits, size_t* tmp, size_t off[5], size_t* total) {
    size_t t = 0;
    int r = sort_pairs_u64(nullptr, nullptr, nullptr, nullptr, n, key_bits, nullptr, &t, nullptr);
    if (r != NGS_OK) return r;
    *tmp = t;
    off[0] = al256(t);                          // sorted keys
    off[1] = off[0] + al256(8 * (size_t)n);     // sorted tags
    off[2] = off[1] + al256(4 * (size_t)n);     // heads
    off[3] = off[2] + al256(4 * (size_t)n);     // offsets (n + 1)
    off[4] = off[3] + al256(4 * (size_t)(n + 1));
    const int64_t scan_bytes = ngs_scan_workspace_bytes(n);
    if (scan_bytes < 0) return NGS_ERR_HIP;
    *total = off[4] + al256((size_t)scan_bytes);
    return NGS_OK;
}

static int primal_layout(int64_t n, size_t* tmp, size_t off[8], size_t* total) {
    size_t t = 0;
    int r = sort_pairs_u64(nullptr, nullptr, nullptr, nullptr, n, 64, nullptr, &t, nullptr);
    if (r != NGS_OK) return r;
    *tmp = t;
    off[0] = al256(t);                          // keys
    off[1] = off[0] + al256(8 * (size_t)n);     // tags
    off[2] = off[1] + al256(4 * (size_t)n);     // sorted keys
    off[3] = off[2] + al256(8 * (size_t)n);     // sorted tags
    off[4] = off[3] + al256(4 * (size_t)n);     // heads
    off[5] = off[4] + al256(4 * (size_t)n);     // offsets (n + 1)
    off[6] = off[5] + al256(4 * (size_t)(n + 1));  // run starts
    off[7] = off[6] + al256(4 * (size_t)n);     // scan workspace
    const int64_t scan_bytes = ngs_scan_workspace_bytes(n);
    if (scan_bytes < 0) return NGS_ERR_HIP;
    *total = off[7] + al256((size_t)scan_bytes);
    return NGS_OK;
}

int ngs_primal_mc_count(const float* sdf, int64_t nleaf, int32_t r, int32_t* vcount, ngs_stream_t s) {
    NGS_REQUIRE(r >= 1 && r <= 1024, "ngs_primal_mc_count: resolution must be in [1, 1024]");
    NGS_REQUIRE(nleaf >= 0 && nleaf * (int64_t)(r + 1) * (r + 1) * (r + 1) < ((int64_t)1 << 31),
                "ngs_primal_mc_count: more than 2^31 samples");
    const int64_t t = nleaf * (int64_t)r * r * r;
    if (t <= 0) return NGS_OK;
    NGS_REQUIRE(sdf && vcount, "ngs_primal_mc_count: null argument");
    primal_count_kernel<<<grid1d(t, 256), 256, 0, S(s)>>>(sdf, nleaf, r, vcount);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_primal_mc_emit(const float* base, const float* sdf, int64_t nleaf, int32_t r, const int32_t* voffset,
                       float* tris, int32_t* vert_ids, int32_t* src_ids, float* src_weight, ngs_stream_t s) {
    NGS_REQUIRE(r >= 1 && r <= 1024, "ngs_primal_mc_emit: resolution must be in [1, 1024]");
    NGS_REQUIRE(nleaf >= 0 && nleaf * (int64_t)(r + 1) * (r + 1) * (r + 1) < ((int64_t)1 << 31),
                "ngs_primal_mc_emit: more than 2^31 samples");
    const int64_t t = nleaf * (int64_t)r * r * r;
    if (t <= 0) return NGS_OK;
    NGS_REQUIRE(base && sdf && voffset && tris && vert_ids && src_ids && src_weight, "ngs_primal_mc_emit: null argument");
    primal_emit_kernel<<<grid1d(t, 128), 128, 0, S(s)>>>(base, sdf, nleaf, r, voffset, tris, vert_ids, src_ids,
                                                         src_weight);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int64_t ngs_primal_dedup_workspace_bytes(int64_t n) {
    size_t tmp, off[8], total;
    if (primal_layout(n, &tmp, off, &total) != NGS_OK) return -1;
    return (int64_t)total;
}

int ngs_primal_dedup(const int32_t* vert_ids, int64_t n, const int32_t* key_layout, const float* tris,
                     const int32_t* src_ids, const float* src_weight, int64_t* faces, float* verts, float* normals,
                     int32_t* unq, int32_t* vsrc, float* vw, int64_t* nverts, void* workspace, int64_t workspace_bytes,
                     ngs_stream_t s) {
    NGS_REQUIRE(key_layout && nverts, "ngs_primal_dedup: null key layout or count");
    NGS_REQUIRE(n >= 0 && n < ((int64_t)1 << 31), "ngs_primal_dedup: bad size");
    EdgeKeyLayout kl;
    const int bz = key_layout[5], by = key_layout[4], bx = key_layout[3];
    NGS_REQUIRE(bx >= 0 && by >= 0 && bz >= 0 && bx <= 31 && by <= 31 && bz <= 31 && bx + by + bz + 2 <= 64,
                "ngs_primal_dedup: key layout needs more than 64 bits");
    for (int a = 0; a < 3; ++a) kl.org[a] = key_layout[a];
    kl.sh[2] = 2;
    kl.sh[1] = 2 + bz;
    kl.sh[0] = 2 + bz + by;
    if (n == 0) {
        set_i64_kernel<<<1, 1, 0, S(s)>>>(nverts, 0);
        NGS_CHECK_LAUNCH();
        return NGS_OK;
    }
    NGS_REQUIRE(vert_ids && tris && src_ids && src_weight && faces && verts && normals && unq && vsrc && vw && workspace,
                "ngs_primal_dedup: null argument");
    size_t tmp, off[8], total;
    int r = primal_layout(n, &tmp, off, &total);
    if (r != NGS_OK) return r;
    NGS_REQUIRE((size_t)workspace_bytes >= total, "ngs_primal_dedup: workspace too small");
    char* w = reinterpret_cast<char*>(workspace);
    hipStream_t st = S(s);
    uint64_t* keys = reinterpret_cast<uint64_t*>(w + off[0]);
    uint32_t* tags = reinterpret_cast<uint32_t*>(w + off[1]);
    uint64_t* kout = reinterpret_cast<uint64_t*>(w + off[2]);
    uint32_t* tout = reinterpret_cast<uint32_t*>(w + off[3]);
    int32_t* head = reinterpret_cast<int32_t*>(w + off[4]);
    int32_t* offs = reinterpret_cast<int32_t*>(w + off[5]);
    int32_t* start = reinterpret_cast<int32_t*>(w + off[6]);
    primal_pack_kernel<<<grid1d(n, 256), 256, 0, st>>>(vert_ids, n, kl, keys, tags);
    NGS_CHECK_LAUNCH();
    size_t tb = tmp;
    r = sort_pairs_u64(keys, kout, reinterpret_cast<const int32_t*>(tags), reinterpret_cast<int32_t*>(tout), n,
                       bx + by + bz + 2, w, &tb, st);
    if (r != NGS_OK) return r;
    run_heads_kernel<<<grid1d(n, 256), 256, 0, st>>>(kout, n, head);
    NGS_CHECK_LAUNCH();
    r = scan_i32_internal(head, n, offs, w + off[7], st);
    if (r != NGS_OK) return r;
    primal_scatter_kernel<<<grid1d(n, 256), 256, 0, st>>>(head, offs, tout, n, tris, vert_ids, src_ids, src_weight,
                                                          faces, verts, unq, vsrc, vw, start);
    NGS_CHECK_LAUNCH();
    widen_kernel<<<1, 1, 0, st>>>(&offs[n], nverts);
    NGS_CHECK_LAUNCH();
    // at most n vertices; threads beyond the count read it from the device and return
    primal_normals_kernel<<<grid1d(n, 256), 256, 0, st>>>(head, tout, start, nverts, n, faces, verts, normals);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_primal_mc_backward(const float* grad_vert, const float* normals, const int32_t* vsrc, const float* vw,
                           int64_t nverts, float bandwidth, float* grad_grid, ngs_stream_t s) {
    if (nverts <= 0) return NGS_OK;
    NGS_REQUIRE(grad_vert && normals && vsrc && vw && grad_grid, "ngs_primal_mc_backward: null argument");
    primal_backward_kernel<<<grid1d(nverts, 256), 256, 0, S(s)>>>(grad_vert, normals, vsrc, vw, nverts, bandwidth,
                                                                   grad_grid);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int64_t ngs_dedup_workspace_bytes(int64_t nkeys) {
    size_t tmp, off[5], total;
    if (dedup_layout(nkeys, 64, &tmp, off, &total) != NGS_OK) return -1;
    return (int64_t)total;
}

int ngs_dedup_vertices(const uint64_t* vkeys, const uint32_t* vtags, int64_t nkeys, int32_t id_bits, const float* pos,
                       const float* values, int64_t* faces, float* verts, int64_t* nverts, void* workspace,
                       int64_t workspace_bytes, ngs_stream_t s) {
    NGS_REQUIRE(workspace && pos && values, "ngs_dedup_vertices: null argument");
    NGS_REQUIRE(id_bits >= 1 && id_bits <= 31 && nkeys < ((int64_t)1 << 31), "ngs_dedup_vertices: bad sizes");
    if (nkeys <= 0) return NGS_OK;
    size_t tmp, off[5], total;
    int r = dedup_layout(nkeys, 2 * id_bits, &tmp, off, &total);
    if (r != NGS_OK) return r;
    NGS_REQUIRE((size_t)workspace_bytes >= total, "ngs_dedup_vertices: workspace too small");
    char* w = reinterpret_cast<char*>(workspace);
    hipStream_t st = S(s);
    uint64_t* kout = reinterpret_cast<uint64_t*>(w + off[0]);
    uint32_t* tout = reinterpret_cast<uint32_t*>(w + off[1]);
    int32_t* head = reinterpret_cast<int32_t*>(w + off[2]);
    int32_t* offs = reinterpret_cast<int32_t*>(w + off[3]);
    size_t tb = tmp;
    // keys use 2 id_bits bits: the radix sort skips the empty high digits
    r = sort_pairs_u64(vkeys, kout, reinterpret_cast<const int32_t*>(vtags), reinterpret_cast<int32_t*>(tout), nkeys,
                       2 * id_bits, w, &tb, st);
    if (r != NGS_OK) return r;
    run_heads_kernel<<<grid1d(nkeys, 256), 256, 0, st>>>(kout, nkeys, head);
    NGS_CHECK_LAUNCH();
    r = scan_i32_internal(head, nkeys, offs, w + off[4], st);
    if (r != NGS_OK) return r;
    dedup_scatter_kernel<<<grid1d(nkeys, 256), 256, 0, st>>>(head, offs, kout, tout, nkeys, id_bits, pos, values, faces,
                                                              verts);
    NGS_CHECK_LAUNCH();
    if (nverts) {
        widen_kernel<<<1, 1, 0, st>>>(&offs[nkeys], nverts);
        NGS_CHECK_LAUNCH();
    }
    return NGS_OK;
}

int ngs_mark_dual_leaves(ngs_grid g, const int32_t* leaves, int64_t n, int32_t stride, int32_t n_upsample, ngs_stream_t s) {
    NGS_REQUIRE(n_upsample >= 1 && stride >= 1, "ngs_mark_dual_leaves: bad arguments");
    int64_t tot = n * (n_upsample + 1) * (n_upsample + 1);
    if (tot <= 0) return NGS_OK;
    mark_dual_leaves_kernel<<<grid1d(tot, 256), 256, 0, S(s)>>>(g, leaves, n, stride, n_upsample);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_mark_dual_leaf_words(ngs_grid centres, ngs_grid leaf_bitmap, ngs_stream_t s) {
    NGS_REQUIRE(centres.log2_stride == 0 && leaf_bitmap.log2_stride == 0 && centres.bits && leaf_bitmap.bits &&
                    leaf_bitmap.rank,
                "ngs_mark_dual_leaf_words: stride-1 grids required");
    if (centres.nwords <= 0) return NGS_OK;
    mark_dual_leaf_words_kernel<<<grid1d(centres.nwords, 256), 256, 0, S(s)>>>(centres, leaf_bitmap);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_dual_gather_cubes(const int32_t* centre_coords, int64_t dc, int32_t nlevels, const ngs_grid* leaf_bitmaps,
                          const int32_t* const* leaf_index, const int32_t* id_base, const float* values, int32_t* acc,
                          int32_t* count, ngs_stream_t s) {
    NGS_REQUIRE(nlevels >= 1 && nlevels <= MAX_LEAF_LEVELS && leaf_bitmaps && leaf_index && id_base && acc && count &&
                    (dc == 0 || centre_coords),
                "ngs_dual_gather_cubes: bad arguments");
    if (dc <= 0) return NGS_OK;
    LeafLevels lv{};
    lv.n = nlevels;
    for (int l = 0; l < nlevels; ++l) {
        NGS_REQUIRE(leaf_bitmaps[l].bits && leaf_bitmaps[l].rank && leaf_index[l], "ngs_dual_gather_cubes: null level");
        NGS_REQUIRE(l == 0 || leaf_bitmaps[l].log2_stride > leaf_bitmaps[l - 1].log2_stride,
                    "ngs_dual_gather_cubes: levels must go from fine to coarse");
        lv.g[l] = leaf_bitmaps[l];
        lv.index[l] = leaf_index[l];
        lv.base[l] = id_base[l];
    }
    dual_gather_cubes_kernel<<<grid1d(dc, 256), 256, 0, S(s)>>>(centre_coords, dc, lv, values, acc, count);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_dual_accumulate_leaves(ngs_grid centres, const int32_t* leaves, int64_t n, int32_t n_upsample,
                               int32_t id_base, const float* values, int32_t* acc, int32_t* count, ngs_stream_t s) {
    NGS_REQUIRE(n_upsample >= 1 && centres.log2_stride == 0, "ngs_dual_accumulate_leaves: bad n_upsample or centre grid");
    const int64_t tot = n * (n_upsample + 1) * (n_upsample + 1) * (n_upsample + 1);
    if (tot <= 0) return NGS_OK;
    dual_accumulate_leaves_kernel<<<grid1d(tot, 256), 256, 0, S(s)>>>(centres, leaves, n, n_upsample, id_base,
                                                                      values, acc, count);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_leaf_corner_positions(const int32_t* leaves, int64_t n, int32_t stride, int32_t n_upsample, float scale,
                              float* pos, ngs_stream_t s) {
    NGS_REQUIRE(stride >= 1 && n_upsample >= 1 && pos && (n == 0 || leaves), "ngs_leaf_corner_positions: bad arguments");
    const int64_t tot = n * n_upsample * n_upsample * n_upsample;
    if (tot <= 0) return NGS_OK;
    leaf_corner_pos_kernel<<<grid1d(tot, 256), 256, 0, S(s)>>>(leaves, n, n_upsample, stride, scale, pos);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_eval_leaf_corners_levels(int32_t nlevels, const ngs_grid* grids, const float* const* fields, const int32_t* leaves,
                                 int64_t n, int32_t leaf_stride, int32_t n_upsample, float scale, float inv_voxel_size,
                                 int32_t accumulate, float* out, ngs_stream_t s) {
    NGS_REQUIRE(nlevels >= 1 && nlevels <= MAX_FIELD_LEVELS && grids && fields && n_upsample >= 1 && leaf_stride >= 1 &&
                    out && (n == 0 || leaves),
                "ngs_eval_leaf_corners_levels: bad arguments");
    if (n <= 0) return NGS_OK;
    FieldLevels fl{};
    fl.n = nlevels;
    for (int k = 0; k < nlevels; ++k) {
        NGS_REQUIRE(fields[k], "ngs_eval_leaf_corners_levels: null field");
        fl.g[k] = grids[k];
        fl.field[k] = fields[k];
    }
    eval_leaf_corners_levels_kernel<<<grid1d(n, 128), 128, 0, S(s)>>>(fl, leaves, n, leaf_stride, n_upsample, scale,
                                                                      inv_voxel_size, accumulate, out);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_eval_leaf_corners(ngs_grid g, const float* field, const int32_t* leaves, int64_t n, int32_t leaf_stride,
                          int32_t n_upsample, float scale, float inv_voxel_size, int32_t accumulate, float* out,
                          ngs_stream_t s) {
    NGS_REQUIRE(n_upsample >= 1 && leaf_stride >= 1 && field && out && (n == 0 || leaves),
                "ngs_eval_leaf_corners: bad arguments");
    if (n <= 0) return NGS_OK;
    eval_leaf_corners_kernel<<<grid1d(n, 128), 128, 0, S(s)>>>(g, field, leaves, n, leaf_stride, n_upsample, scale,
                                                               inv_voxel_size, accumulate, out);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

}  // extern "C"
