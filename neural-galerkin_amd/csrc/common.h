// Shared device/host helpers of libngspsr (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/ngspsr.h"

namespace ngs {

// Device buffers reached through struct fields (tables, grids) are global memory; casting them to
// the global address space lets the compiler use global loads with per-counter waits instead of
// flat accesses. Never apply it to LDS or host pointers.
#define NGS_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ NGS_GLOBAL T* as_global(T* p) {
    return (NGS_GLOBAL T*)p;
}

// ----------------------------------------------------------------------------------- errors
void set_error(const std::string& msg);

#define NGS_CHECK_HIP(expr)                                                                   \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess) {                                                               \
            ::ngs::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));              \
            return NGS_ERR_HIP;                                                               \
        }                                                                                     \
    } while (0)

#define NGS_CHECK_LAUNCH()                                                                    \
    do {                                                                                      \
        hipError_t _e = hipGetLastError();                                                    \
        if (_e != hipSuccess) {                                                               \
            ::ngs::set_error(std::string("kernel launch: ") + hipGetErrorString(_e));         \
            return NGS_ERR_HIP;                                                               \
        }                                                                                     \
    } while (0)

#define NGS_REQUIRE(cond, msg)                                                                \
    do {                                                                                      \
        if (!(cond)) {                                                                        \
            ::ngs::set_error(msg);                                                            \
            return NGS_ERR_ARG;                                                               \
        }                                                                                     \
    } while (0)

inline hipStream_t S(ngs_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned grid1d(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    return (unsigned)(g < 1 ? 1 : g);
}

// ------------------------------------------------------------------------- ordered bitmap set
// A set of integer lattice points inside a box, stored as a bitmap in x-major (x, y, z) order
// plus the exclusive popcount prefix of every 32-bit word. The rank of a member is its position
// in lexicographic (x, y, z) order, so emitting members in word order yields the sorted unique
// coordinate list of the reference's torch.unique(dim=0), and the rank lookup is a perfect hash.
__device__ __forceinline__ int64_t grid_linear(const ngs_grid& g, int x, int y, int z) {
    // x, y, z are lattice coordinates relative to origin; caller checked bounds.
    return ((int64_t)x * g.dims[1] + y) * (int64_t)g.dims[2] + z;
}

// coordinate (in finest units, multiple of g.stride) -> lattice position; false if outside.
__device__ __forceinline__ bool grid_cell(const ngs_grid& g, int cx, int cy, int cz, int64_t* lin) {
    // floor division by the (power of two) stride
    int lx = (cx >> g.log2_stride) - g.origin[0];
    int ly = (cy >> g.log2_stride) - g.origin[1];
    int lz = (cz >> g.log2_stride) - g.origin[2];
    if ((unsigned)lx >= (unsigned)g.dims[0] || (unsigned)ly >= (unsigned)g.dims[1] ||
        (unsigned)lz >= (unsigned)g.dims[2])
        return false;
    *lin = grid_linear(g, lx, ly, lz);
    return true;
}

// Rank of coordinate c in the set, or -1. Coordinates that are not multiples of the stride are
// absent.
__device__ __forceinline__ int32_t grid_find(const ngs_grid& g, int cx, int cy, int cz) {
    const int mask = (1 << g.log2_stride) - 1;
    if (((cx | cy | cz) & mask) != 0) return -1;
    int64_t lin;
    if (!grid_cell(g, cx, cy, cz, &lin)) return -1;
    uint32_t w = as_global(g.bits)[lin >> 5];
    uint32_t b = (uint32_t)(lin & 31);
    if (!((w >> b) & 1u)) return -1;
    return as_global(g.rank)[lin >> 5] + __popc(w & ((1u << b) - 1u));
}

__device__ __forceinline__ bool grid_test(const ngs_grid& g, int cx, int cy, int cz) {
    const int mask = (1 << g.log2_stride) - 1;
    if (((cx | cy | cz) & mask) != 0) return false;
    int64_t lin;
    if (!grid_cell(g, cx, cy, cz, &lin)) return false;
    return (as_global(g.bits)[lin >> 5] >> (lin & 31)) & 1u;
}

// Sets the bits `mask` of word w. Bits only go from 0 to 1 while a grid is marked, so a plain read
// that already shows them set makes the atomic unnecessary (a stale read only costs the atomic);
// many points mark the same coarse-level voxels.
__device__ __forceinline__ void grid_mark_word(const ngs_grid& g, int64_t w, uint32_t mask) {
    if ((g.bits[w] & mask) != mask) atomicOr(&g.bits[w], mask);
}

__device__ __forceinline__ void grid_mark(const ngs_grid& g, int cx, int cy, int cz) {
    int64_t lin;
    if (grid_cell(g, cx, cy, cz, &lin)) grid_mark_word(g, lin >> 5, 1u << (lin & 31));
}

// Presence mask of the n (<= 31) consecutive lattice points (lx, ly, lz0 + k), k < n, of g (lattice
// coordinates relative to the origin), and in *r0 the rank of the first present one. Present
// points have consecutive ranks: point k has rank r0 + popc(mask & ((1 << k) - 1)).
__device__ __forceinline__ uint32_t grid_column(const ngs_grid& g, int lx, int ly, int lz0, int n, int32_t* r0) {
    if ((unsigned)lx >= (unsigned)g.dims[0] || (unsigned)ly >= (unsigned)g.dims[1]) return 0u;
    const int zlo = max(lz0, 0), zhi = min(lz0 + n, g.dims[2]);
    if (zlo >= zhi) return 0u;
    const int64_t lin = grid_linear(g, lx, ly, zlo);
    const int64_t w = lin >> 5;
    const int b = (int)(lin & 31);
    const int len = zhi - zlo;
    const NGS_GLOBAL uint32_t* bits = as_global(g.bits);
    const uint32_t w0 = bits[w];
    uint64_t v = w0;
    // the window's last bit is a lattice point of the box, so word w + 1 exists when it is needed
    if (b + len > 32) v |= (uint64_t)bits[w + 1] << 32;
    const uint32_t m = (uint32_t)(v >> b) & ((1u << len) - 1u);
    *r0 = as_global(g.rank)[w] + __popc(w0 & ((1u << b) - 1u));
    return m << (zlo - lz0);
}

__device__ __forceinline__ int lane64() { return (int)(threadIdx.x & 63); }

// Logical index of this block such that the blocks of each XCD form one contiguous range. Blocks are
// dispatched round-robin over the 8 XCDs (block b to XCD b % 8), so consecutive blocks land on
// different XCDs and data shared by neighbouring blocks is fetched into every XCD's L2; with this
// map each XCD sweeps its own stretch of consecutive blocks. A bijection of [0, gridDim.x).
__device__ __forceinline__ int64_t xcd_contiguous_block() {
    const int64_t nb = gridDim.x, b = blockIdx.x, x = b % 8;
    return x * (nb / 8) + min(x, nb % 8) + b / 8;
}

// Exclusive prefix over the wave's active lanes and the wave total of n in [0, 7] (inactive lanes
// count as 0): the positions of the rows' runs in a slice of the packed run form.
__device__ __forceinline__ void wave_prefix3(int n, int& excl, int& total) {
    const uint64_t below = (1ull << lane64()) - 1ull;
    const uint64_t b0 = __ballot(n & 1), b1 = __ballot(n & 2), b2 = __ballot(n & 4);
    excl = __popcll(b0 & below) + 2 * __popcll(b1 & below) + 4 * __popcll(b2 & below);
    total = __popcll(b0) + 2 * __popcll(b1) + 4 * __popcll(b2);
}

// Plane stride of the run form (runs (25, P) int32, rvals (125, P) float32): M rounded up to 64
// rows, so every plane starts on a 256-byte boundary and the loads of a wave's 64 consecutive rows
// are single aligned 256-byte segments.
__host__ __device__ __forceinline__ int64_t run_plane_stride(int64_t M) { return (M + 63) / 64 * 64; }

// ----------------------------------------------------------------------------------- FNV keys
// fold60(FNV-1a-64 over the uint32 words of a coordinate row); csrc/common/hashmap_cuda.cu:248-261.
__device__ __forceinline__ uint64_t fnv_fold60(const int32_t* c, int dim) {
    uint64_t h = 14695981039346656037ull;
    for (int j = 0; j < dim; ++j) {
        h ^= (uint32_t)c[j];
        h *= 1099511628211ull;
    }
    return (h >> 60) ^ (h & 0x0FFFFFFFFFFFFFFFull);
}

// ---------------------------------------------------------------------------- B-spline basis
// B(x) = 2 N2(x): (x+1.5)^2 | -2x^2+1.5 | (x-1.5)^2 on [-1.5,-.5) | [-.5,.5) | [.5,1.5);
// jspsr/bases/bezier_tensor.py:37-44. Evaluated with the reference's operation order.
__device__ __forceinline__ float bspline(float x) {
#pragma clang fp contract(off)
    if (x >= -1.5f && x < -0.5f) { float t = x + 1.5f; return t * t; }
    if (x >= -0.5f && x < 0.5f) return -2.0f * (x * x) + 1.5f;
    if (x >= 0.5f && x < 1.5f) { float t = x - 1.5f; return t * t; }
    return 0.0f;
}

__device__ __forceinline__ float bspline_deriv(float x) {
#pragma clang fp contract(off)
    if (x >= -1.5f && x < -0.5f) return 2.0f * x + 3.0f;
    if (x >= -0.5f && x < 0.5f) return -4.0f * x;
    if (x >= 0.5f && x < 1.5f) return 2.0f * x - 3.0f;
    return 0.0f;
}

// Per-axis basis values of a query position around its containing cell at stride s, following
// the reference's neighbour query for float sources (hashtree.py:169-181,
// _hash_backend.py:196-226): base = q - 0.5; qc = floor((base + 0.5) / s) * s;
// c_off = (qc - base) + (s - 1) / 2; rel_k = k s + c_off; arg = -rel_k / s for k = -1, 0, 1.
// Writes the cell coordinate and b[k+1] = B(arg_k).
__device__ __forceinline__ void axis_basis(float q, int s, float inv_s, int* cell, float b[3]) {
#pragma clang fp contract(off)
    float base = q - 0.5f;
    float t = base + 0.5f;
    int qc = (int)(floorf(t * inv_s) * (float)s);
    float coff = ((float)qc - base) + (float)(s - 1) * 0.5f;
    *cell = qc;
#pragma unroll
    for (int k = -1; k <= 1; ++k) {
        float rel = (float)(k * s) + coff;
        b[k + 1] = bspline(-rel * inv_s);
    }
}

// Same, also returning the derivative of the basis (per unit of q) for gradient queries.
__device__ __forceinline__ void axis_basis_grad(float q, int s, float inv_s, int* cell, float b[3], float db[3]) {
#pragma clang fp contract(off)
    float base = q - 0.5f;
    float t = base + 0.5f;
    int qc = (int)(floorf(t * inv_s) * (float)s);
    float coff = ((float)qc - base) + (float)(s - 1) * 0.5f;
    *cell = qc;
#pragma unroll
    for (int k = -1; k <= 1; ++k) {
        float rel = (float)(k * s) + coff;
        b[k + 1] = bspline(-rel * inv_s);
        db[k + 1] = bspline_deriv(-rel * inv_s) * inv_s;
    }
}

// Trilinear corner selection of a point at stride s (_hash_backend.py:469-510):
// dc = floor(q / s) * s, rel = q - dc - s/2; corner choice per axis is dc or dc + sign(rel) s,
// weight of the moved corner |rel|/s. Returns base corner, per-axis sign, per-axis local weight.
__device__ __forceinline__ void trilinear_axis(float q, int s, float inv_s, int* dc, int* sgn, float* loc) {
#pragma clang fp contract(off)
    float d = floorf(q * inv_s) * (float)s;
    float rel = (q - d) - (float)s * 0.5f;
    *dc = (int)d;
    *sgn = (rel > 0.f) ? 1 : ((rel < 0.f) ? -1 : 0);
    *loc = fabsf(rel) * inv_s;
}

}  // namespace ngs
