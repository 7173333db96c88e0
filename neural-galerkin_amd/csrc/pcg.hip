// Jacobi-preconditioned conjugate gradients on the assembled CSR system (or its stencil form).
//
// One iteration = three kernels, no host round trip (bpcg_* below; a solve of one system is a batch
// of one):
//   K1 spmv_dot:   q = A p; per-block partials of p.q.
//   K2 update:     alpha = rz / sum(p.q); x += alpha p, r -= alpha q; per-block partials of
//                  rz' = r.(D^-1 r) and rr = r.r.
//   K3 direction:  rz', rr summed; convergence (rr <= tol^2 bb) or the iteration cap ends the solve,
//                  otherwise p = D^-1 r + (rz'/rz) p.
// No kernel waits for another block of its own launch: a kernel that needs a global sum reads all
// block partials of the previous kernel and sums them in a fixed order in every block (identical,
// deterministic FP64 results everywhere), and kernel boundaries make the partials visible across
// XCDs. This avoids per-block release fences, which on this GPU write back the XCD's L2. The scalar
// state that K3 advances (rz, iteration count, done flag) is double-buffered by the parity of the
// step within a launch sequence: step u reads slot u&1 and writes slot (u+1)&1, and sequences
// always have an even number of steps, so a new sequence starts on slot 0.
// Row chunks: blocks of one XCD (blockIdx % 8) own one contiguous eighth of the rows (the SpMV
// deals its row tiles round-robin inside that slab), so each XCD's L2 serves the gathers of its
// own slab. Sequences of iterations are captured once into a hipGraph and replayed; kernels of a
// converged system return immediately.
#include "common.h"
#include "stencil.h"

#include <cstddef>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

namespace ngs {

constexpr int PCG_THREADS = 256;
constexpr int NUM_XCD = 8;
constexpr int RUN_COUNT_SHIFT = 28;  // run word: start | n << 28
constexpr int32_t RUN_START_MASK = (1 << RUN_COUNT_SHIFT) - 1;

struct PcgSlot {
    int32_t iter;  // iterations completed
    int32_t done;  // 0 running, 1 converged, 2 iteration cap (set in both slots when the solve ends)
};

struct PcgState {
    double rz[2];   // r.(D^-1 r) of the current iterate, by step parity
    double rr, bb;  // ||r||^2 of the last completed iteration (recurrence), ||b||^2
    double tol2;
    int32_t max_iter, pad;
    PcgSlot slot[2];
};

// Bytes of the state read back after a launch sequence: rr and bb, then slot 0, one contiguous span.
constexpr size_t kStatusBytes = offsetof(PcgState, slot) + sizeof(PcgSlot) - offsetof(PcgState, rr);
static_assert(offsetof(PcgState, bb) == offsetof(PcgState, rr) + sizeof(double), "rr, bb adjacent");
static_assert(kStatusBytes <= 64, "the status span fits the 64-byte pinned staging buffer");

inline void read_status(const void* bytes, double sc[2], int32_t* iter, int32_t* done) {
    const char* c = static_cast<const char*>(bytes);
    std::memcpy(sc, c, sizeof(double) * 2);
    PcgSlot slot;
    std::memcpy(&slot, c + (offsetof(PcgState, slot) - offsetof(PcgState, rr)), sizeof(PcgSlot));
    *iter = slot.iter;
    *done = slot.done;
}

struct PcgArgs {
    const int32_t* rowptr;
    const int32_t* cols;
    const int32_t* runs;    // run form (kind SYS_RUNS): 25 plane-major run words start | n << 28
    const float* rvals;     // run form: 125 plane-major value slots per row; packed form: the values
    const int64_t* soff;    // packed form (kind SYS_PACKED): value offset of every 64-row slice
    const float* vals;
    const float* dinv;
    const float* b;
    float* x;
    float* r;
    float* p;
    float* q;
    double* partials;  // 3 * nblocks: p.q | r.z | r.r (per block)
    PcgState* st;
    int64_t M;
    int64_t chunk;     // rows per block
    int nblocks;
};

__device__ __forceinline__ int logical_block() {
    // Blocks are dispatched round-robin over the XCDs; give each XCD a contiguous range.
    const int nb = gridDim.x, b = blockIdx.x;
    if (nb % NUM_XCD) return b;
    return (b % NUM_XCD) * (nb / NUM_XCD) + b / NUM_XCD;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Block reduce of two doubles (256 threads); the sums are returned in every thread.
__device__ __forceinline__ void block_sum2(double& a, double& b) {
    __shared__ double sa[PCG_THREADS / 64], sb[PCG_THREADS / 64];
    a = wave_sum(a);
    b = wave_sum(b);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();  // the arrays may still be read by a previous call
    if (lane == 0) { sa[w] = a; sb[w] = b; }
    __syncthreads();
    double ta = 0, tb = 0;
#pragma unroll
    for (int i = 0; i < PCG_THREADS / 64; ++i) { ta += sa[i]; tb += sb[i]; }
    a = ta;
    b = tb;
}

// Sums of two partial arrays (n entries each) in a fixed order, returned in every thread of the
// block: every block that calls it computes bit-identical values.
__device__ __forceinline__ void sum_partials2(const double* pa, const double* pb, int n, double& a, double& b) {
    const NGS_GLOBAL double* ga = as_global(pa);
    const NGS_GLOBAL double* gb = as_global(pb);
    double ta = 0, tb = 0;
#pragma unroll 8
    for (int i = threadIdx.x; i < n; i += PCG_THREADS) {
        ta += ga[i];
        tb += gb ? gb[i] : 0.0;
    }
    block_sum2(ta, tb);
    a = ta;
    b = tb;
}

// Block sums of (a, b) written to partials pa[bid], pb[bid] (pb may be null).
__device__ __forceinline__ void store_partials2(double a, double b, double* pa, double* pb, int bid) {
    block_sum2(a, b);
    if (threadIdx.x == 0) {
        as_global(pa)[bid] = a;
        if (pb) as_global(pb)[bid] = b;
    }
}

// Dot product of CSR row `row` with x by G lanes (16-byte loads of 4 values and 4 columns per lane).
// Rows hold multiples of 4 entries. The first 128 entries (every assembled row: at most 125
// neighbours, padded) are read without branches: each lane loads its groups from an index clamped
// to the row's last group (always in bounds; the surplus is masked out of the sum), so the compiler
// issues all matrix loads, then all gathers, with no wait between them. Longer rows of a caller's
// CSR continue in a plain loop. NT: non-temporal loads for the matrix stream.
template <int G, bool NT>
__device__ __forceinline__ float row_dot(int32_t k0, int32_t k1, const NGS_GLOBAL int32_t* __restrict__ cols,
                                         const NGS_GLOBAL float* __restrict__ vals,
                                         const NGS_GLOBAL float* __restrict__ x, int lane) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    typedef int i4 __attribute__((ext_vector_type(4)));
    constexpr int U = 128 / (4 * G);
    float sum = 0.f;
    if (k1 > k0) {
        const int32_t klast = k1 - 4;
        f4 v[U];
        i4 c[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int32_t k = min(k0 + 4 * lane + 4 * G * u, klast);
            if (NT) {
                v[u] = __builtin_nontemporal_load(reinterpret_cast<const NGS_GLOBAL f4*>(vals + k));
                c[u] = __builtin_nontemporal_load(reinterpret_cast<const NGS_GLOBAL i4*>(cols + k));
            } else {
                v[u] = *reinterpret_cast<const NGS_GLOBAL f4*>(vals + k);
                c[u] = *reinterpret_cast<const NGS_GLOBAL i4*>(cols + k);
            }
        }
        float xs[U][4];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            xs[u][0] = x[c[u].x];
            xs[u][1] = x[c[u].y];
            xs[u][2] = x[c[u].z];
            xs[u][3] = x[c[u].w];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const float part = v[u].x * xs[u][0] + v[u].y * xs[u][1] + v[u].z * xs[u][2] + v[u].w * xs[u][3];
            sum += (k0 + 4 * lane + 4 * G * u < k1) ? part : 0.f;
        }
        for (int32_t k = k0 + 128 + 4 * lane; k < k1; k += 4 * G) {
            const f4 vv = *reinterpret_cast<const NGS_GLOBAL f4*>(vals + k);
            const i4 cc = *reinterpret_cast<const NGS_GLOBAL i4*>(cols + k);
            sum += vv.x * x[cc.x] + vv.y * x[cc.y] + vv.z * x[cc.z] + vv.w * x[cc.w];
        }
    }
#pragma unroll
    for (int d = G / 2; d > 0; d >>= 1) sum += __shfl_xor(sum, d, 64);
    return sum;
}

// a * b + c with one rounding, in the sum's type: the row sums of the run and packed forms are
// written with it so that their rounding does not depend on the compiler's contraction choices.
template <typename T>
__device__ __forceinline__ T fma_t(T a, T b, T c);
template <>
__device__ __forceinline__ float fma_t<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <>
__device__ __forceinline__ double fma_t<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Row `row` of the run form times x: the 25 bitmap columns (dx, dy) in order, each a run of n <= 5
// consecutive columns start .. start + n - 1 with its values in slots 0 .. n - 1 (the rest hold 0),
// so the terms are summed in increasing column order. One lane per row: the run words and value
// slots of consecutive rows are contiguous (one 256-byte segment per wave and load), and the x
// reads of a run are consecutive words. The loads of 5 columns are issued before their gathers.
template <typename T, typename XT>
__device__ __forceinline__ T run_row_dot(const NGS_GLOBAL int32_t* __restrict__ runs,
                                         const NGS_GLOBAL float* __restrict__ rvals, const NGS_GLOBAL XT* __restrict__ x,
                                         int64_t P, int64_t row) {
    T sum = 0;
    constexpr int B = 5;
#pragma unroll 1
    for (int p0 = 0; p0 < 25; p0 += B) {
        int32_t w[B];
        float v[B][5];
#pragma unroll
        for (int b = 0; b < B; ++b) {
            w[b] = runs[(p0 + b) * P + row];
#pragma unroll
            for (int j = 0; j < 5; ++j) v[b][j] = rvals[(5 * (p0 + b) + j) * P + row];
        }
        XT xv[B][5];
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int32_t start = w[b] & RUN_START_MASK;
            const int n = (int)((uint32_t)w[b] >> RUN_COUNT_SHIFT);
#pragma unroll
            for (int j = 0; j < 5; ++j) xv[b][j] = j < n ? x[start + j] : (XT)0;
        }
#pragma unroll
        for (int b = 0; b < B; ++b)
#pragma unroll
            for (int j = 0; j < 5; ++j) sum = fma_t<T>((T)v[b][j], (T)xv[b][j], sum);
    }
    return sum;
}

__global__ void __launch_bounds__(PCG_THREADS) pcg_init_kernel(PcgArgs A) {
    const int lb = logical_block();
    const int64_t r0 = (int64_t)lb * A.chunk;
    const int64_t r1 = min(A.M, r0 + A.chunk);
    double bb = 0, rz = 0;
    const NGS_GLOBAL float* __restrict__ b = as_global(A.b);
    const NGS_GLOBAL float* __restrict__ dinv = as_global(A.dinv);
    NGS_GLOBAL float* __restrict__ x = as_global(A.x);
    NGS_GLOBAL float* __restrict__ r = as_global(A.r);
    NGS_GLOBAL float* __restrict__ p = as_global(A.p);
    for (int64_t i = r0 + threadIdx.x; i < r1; i += PCG_THREADS) {
        float bi = b[i];
        float z = dinv[i] * bi;
        x[i] = 0.f;
        r[i] = bi;
        p[i] = z;
        bb += (double)bi * bi;
        rz += (double)bi * z;
    }
    store_partials2(rz, bb, A.partials + A.nblocks, A.partials + 2 * A.nblocks, lb);
}

// One block: initial state from the partials of pcg_init_kernel (r.z in slot 0, ||b||^2).
__global__ void __launch_bounds__(PCG_THREADS) pcg_init_finish_kernel(PcgArgs A, double tol, int max_iter) {
    double rz, bb;
    sum_partials2(A.partials + A.nblocks, A.partials + 2 * A.nblocks, A.nblocks, rz, bb);
    if (threadIdx.x == 0) {
        PcgState* s = A.st;
        s->rz[0] = rz;
        s->rz[1] = rz;
        s->bb = bb;
        s->rr = bb;
        s->tol2 = tol * tol;
        s->max_iter = max_iter;
        const int done = (bb == 0.0) ? 1 : 0;
        s->slot[0].iter = s->slot[1].iter = 0;
        s->slot[0].done = s->slot[1].done = done;
    }
}

// Row tiles of one block in the SpMV kernels. Contiguous: block chunk [lb*chunk, (lb+1)*chunk).
// Interleaved (nb a multiple of 8): the rows are cut into 8 slabs, one per XCD (block index mod 8),
// and each XCD's blocks take tiles of `tile` consecutive rows round-robin, so all blocks of an XCD
// sweep their slab together and the gathered vector window stays small enough for the XCD's L2.
struct RowTiles {
    int64_t first, end, step;  // tile starts: first, first + step, ... < end
    int64_t hi;                // rows of a tile are clipped to hi
};

__device__ __forceinline__ RowTiles tiles_interleaved(int64_t M, int nb, int raw, int tile) {
    RowTiles t;
    const int nbx = nb / NUM_XCD, xcd = raw % NUM_XCD, j = raw / NUM_XCD;
    const int64_t slab = (M + NUM_XCD - 1) / NUM_XCD;
    const int64_t lo = min(M, (int64_t)xcd * slab);
    t.hi = min(M, lo + slab);
    t.first = lo + (int64_t)j * tile;
    t.end = t.hi;
    t.step = (int64_t)nbx * tile;
    return t;
}

__device__ __forceinline__ RowTiles tiles_contiguous(int64_t M, int64_t chunk, int local, int tile) {
    RowTiles t;
    t.first = (int64_t)local * chunk;
    t.hi = min(M, t.first + chunk);
    t.end = t.hi;
    t.step = tile;
    return t;
}

// IL: interleaved tiles (the host launches it only with a grid that is a multiple of 8).
template <int G, bool NT, bool IL>
__global__ void __launch_bounds__(PCG_THREADS) __attribute__((amdgpu_waves_per_eu(8))) spmv_kernel(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
                                                           const float* __restrict__ vals, const float* __restrict__ x,
                                                           float* __restrict__ y, int64_t M, int64_t chunk) {
    const int lane = threadIdx.x % G, grp = threadIdx.x / G;
    constexpr int GROUPS = PCG_THREADS / G;
    const RowTiles t = IL ? tiles_interleaved(M, gridDim.x, blockIdx.x, GROUPS)
                          : tiles_contiguous(M, chunk, logical_block(), GROUPS);
    const NGS_GLOBAL int32_t* rp = as_global(rowptr);
    for (int64_t t0 = t.first; t0 < t.end; t0 += t.step) {
        const int64_t row = t0 + grp;
        if (row >= t.hi) continue;
        float sum = row_dot<G, NT>(rp[row], rp[row + 1], as_global(cols), as_global(vals), as_global(x), lane);
        if (lane == 0) y[row] = sum;
    }
}

// Rows of the run-form kernels: one row per thread, grid-stride over the system's blocks (all blocks
// sweep the rows together, block b taking rows b*256 .. b*256+255 of every 256*nb stretch), which
// measured faster than per-block contiguous ranges for this access pattern.
__device__ __forceinline__ int64_t run_rows_first(int local) { return (int64_t)local * PCG_THREADS + threadIdx.x; }

// y = A x for the run form.
__global__ void __launch_bounds__(PCG_THREADS) spmv_runs_kernel(const int32_t* __restrict__ runs,
                                                                const float* __restrict__ rvals,
                                                                const float* __restrict__ x, float* __restrict__ y,
                                                                int64_t M) {
    const int64_t step = (int64_t)gridDim.x * PCG_THREADS;
    for (int64_t row = run_rows_first(blockIdx.x); row < M; row += step)
        y[row] = run_row_dot<float, float>(as_global(runs), as_global(rvals), as_global(x), run_plane_stride(M), row);
}

// ------------------------------------------------------------------------ packed run form
// The run form without its zero value slots: the rows are cut into slices of 64 (one wave each);
// slice s stores, plane by plane (p = 0..24), the values of its rows' runs back to back in lane
// order (row 64 s + l's run of plane p at the exclusive prefix of the run lengths of lanes < l),
// starting at soff[s]. The run words are those of the run form.
constexpr int PACK_BATCH = 5;                    // planes staged together
constexpr int PACK_STAGE = PACK_BATCH * 5 * 64;  // floats of LDS per wave
static_assert(PACK_BATCH == 5, "batch_prefix5 packs the run lengths of 5 planes");

// Values per slice (the sum of its rows' run lengths). One wave per slice.
__global__ void __launch_bounds__(PCG_THREADS) runs_slice_counts_kernel(const int32_t* __restrict__ runs, int64_t M,
                                                                        int64_t* __restrict__ counts) {
    const int64_t nsl = (M + 63) / 64;
    const int64_t sl = ((int64_t)blockIdx.x * PCG_THREADS + threadIdx.x) >> 6;
    if (sl >= nsl) return;  // uniform over the wave
    const int64_t row = sl * 64 + lane64();
    const int64_t P = run_plane_stride(M);
    int tot = 0;
    for (int p = 0; p < 25; ++p) {
        const int n = row < M ? (int)((uint32_t)runs[p * P + row] >> RUN_COUNT_SHIFT) : 0;
        int ex, t;
        wave_prefix3(n, ex, t);
        tot += t;
    }
    if (lane64() == 0) counts[sl] = tot;
}

// Slot form -> packed form. One wave per slice.
__global__ void __launch_bounds__(PCG_THREADS) runs_pack_kernel(const int32_t* __restrict__ runs,
                                                                const float* __restrict__ rvals, int64_t M,
                                                                const int64_t* __restrict__ soff,
                                                                float* __restrict__ pvals) {
    const int64_t nsl = (M + 63) / 64;
    const int64_t sl = ((int64_t)blockIdx.x * PCG_THREADS + threadIdx.x) >> 6;
    if (sl >= nsl) return;
    const int64_t row = sl * 64 + lane64();
    const int64_t P = run_plane_stride(M);
    int64_t pos = soff[sl];
    for (int p = 0; p < 25; ++p) {
        const int n = row < M ? (int)((uint32_t)runs[p * P + row] >> RUN_COUNT_SHIFT) : 0;
        int ex, t;
        wave_prefix3(n, ex, t);
        for (int j = 0; j < n; ++j) pvals[pos + ex + j] = rvals[(5 * p + j) * P + row];
        pos += t;
    }
}

// Packed form -> slot form (the inverse of runs_pack_kernel; slots past a run hold 0).
__global__ void __launch_bounds__(PCG_THREADS) runs_unpack_kernel(const int32_t* __restrict__ runs,
                                                                  const float* __restrict__ pvals, int64_t M,
                                                                  const int64_t* __restrict__ soff,
                                                                  float* __restrict__ rvals) {
    const int64_t nsl = (M + 63) / 64;
    const int64_t sl = ((int64_t)blockIdx.x * PCG_THREADS + threadIdx.x) >> 6;
    if (sl >= nsl) return;
    const int64_t row = sl * 64 + lane64();
    const int64_t P = run_plane_stride(M);
    int64_t pos = soff[sl];
    for (int p = 0; p < 25; ++p) {
        const int n = row < M ? (int)((uint32_t)runs[p * P + row] >> RUN_COUNT_SHIFT) : 0;
        int ex, t;
        wave_prefix3(n, ex, t);
        if (row < M)
            for (int j = 0; j < 5; ++j) rvals[(5 * p + j) * P + row] = j < n ? pvals[pos + ex + j] : 0.f;
        pos += t;
    }
}

// Inclusive prefix sum over the 64 lanes of a wave (every lane active): shifts within rows of 16,
// then the last lane of rows 0 and 2 into rows 1 and 3, and lane 31 into rows 2 and 3 (DPP).
__device__ __forceinline__ int wave_incl_scan(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);  // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false); // row_bcast:15
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false); // row_bcast:31
    return v;
}

// Positions of a wave's runs in one batch of 5 planes of the packed form: ex[b] = the runs of planes
// before b in the batch plus those of plane b in lanes below this one; returns the batch total. The
// 3-bit run lengths go as 9-bit fields (sums <= 320) into two words, planes 0-2 and 3-4, and two wave
// scans give every field's prefix at once: no ballot masks held in scalar registers, which left the
// row loop short of them (spilled to VGPR lanes) in the table-driven PCG kernel. Every lane of the
// wave must be active.
__device__ __forceinline__ int batch_prefix5(const int32_t (&w)[5], int (&ex)[5]) {
    int n[5];
#pragma unroll
    for (int b = 0; b < 5; ++b) n[b] = (int)((uint32_t)w[b] >> RUN_COUNT_SHIFT);
    const int a0 = n[0] | (n[1] << 9) | (n[2] << 18), a1 = n[3] | (n[4] << 9);
    const int i0 = wave_incl_scan(a0), i1 = wave_incl_scan(a1);
    const int t0 = __builtin_amdgcn_readlane(i0, 63), t1 = __builtin_amdgcn_readlane(i1, 63);
    const int e0 = i0 - a0, e1 = i1 - a1;
    const int T0 = t0 & 511, T1 = (t0 >> 9) & 511, T2 = (t0 >> 18) & 511, T3 = t1 & 511, T4 = (t1 >> 9) & 511;
    ex[0] = e0 & 511;
    ex[1] = T0 + ((e0 >> 9) & 511);
    ex[2] = T0 + T1 + ((e0 >> 18) & 511);
    ex[3] = T0 + T1 + T2 + (e1 & 511);
    ex[4] = T0 + T1 + T2 + T3 + ((e1 >> 9) & 511);
    return T0 + T1 + T2 + T3 + T4;
}

// Orders a wave's LDS writes before its own later reads (each wave stages into its own region, so
// the block's waves need not wait for each other).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Row `row` (lane of the wave of slice sl) of the packed form times x, the terms in the order of
// run_row_dot without its zero slots (so the sums are the same, bit for bit). Per batch of 5 planes
// the wave stages the batch's values (contiguous in pvals) into its LDS region with coalesced loads,
// then every lane reads its runs there and gathers x. Each wave has its own LDS region, ordered by
// wave-scope fences; lanes past M carry empty runs.
template <typename T, typename XT>
__device__ __forceinline__ T packed_row_dot(const NGS_GLOBAL int32_t* __restrict__ runs,
                                            const NGS_GLOBAL float* __restrict__ pvals,
                                            const NGS_GLOBAL XT* __restrict__ x, int64_t P, int64_t row, bool ok,
                                            int64_t pos, float* stage) {
    T sum = 0;
    const int lane = lane64();
#pragma unroll 1
    for (int p0 = 0; p0 < 25; p0 += PACK_BATCH) {
        int32_t w[PACK_BATCH];
        int ex[PACK_BATCH];
#pragma unroll
        for (int b = 0; b < PACK_BATCH; ++b) w[b] = ok ? runs[(p0 + b) * P + row] : 0;
        const int tot = batch_prefix5(w, ex);
        {
            // all of the batch's loads issued together (one memory round trip per batch): lane l takes
            // values l, l + 64, ... of the batch's segment (tot <= PACK_STAGE)
            float t[PACK_STAGE / 64];
#pragma unroll
            for (int u = 0; u < PACK_STAGE / 64; ++u) {
                const int e = u * 64 + lane;
                t[u] = (u * 64 < tot && e < tot) ? pvals[pos + e] : 0.f;
            }
#pragma unroll
            for (int u = 0; u < PACK_STAGE / 64; ++u) {
                const int e = u * 64 + lane;
                if (u * 64 < tot && e < tot) stage[e] = t[u];
            }
        }
        wave_lds_sync();
        float v[PACK_BATCH][5];
        XT xv[PACK_BATCH][5];
#pragma unroll
        for (int b = 0; b < PACK_BATCH; ++b) {
            const int32_t start = w[b] & RUN_START_MASK;
            const int n = (int)((uint32_t)w[b] >> RUN_COUNT_SHIFT);
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                v[b][j] = j < n ? stage[ex[b] + j] : 0.f;
                xv[b][j] = j < n ? x[start + j] : (XT)0;
            }
        }
#pragma unroll
        for (int b = 0; b < PACK_BATCH; ++b)
#pragma unroll
            for (int j = 0; j < 5; ++j) sum = fma_t<T>((T)v[b][j], (T)xv[b][j], sum);
        pos += tot;
        wave_lds_sync();
    }
    return sum;
}

// y = A x for the packed form: blocks of 4 waves take groups of 4 slices, grid-stride. Capped at 5 waves
// per SIMD; with the scan-based run prefixes it needs 80 VGPRs, so the LDS (6 blocks per CU) sets 6.
__global__ void __launch_bounds__(PCG_THREADS) __attribute__((amdgpu_waves_per_eu(5))) spmv_packed_kernel(const int32_t* __restrict__ runs,
                                                                  const float* __restrict__ pvals,
                                                                  const int64_t* __restrict__ soff,
                                                                  const float* __restrict__ x, float* __restrict__ y,
                                                                  int64_t M) {
    __shared__ float stage_all[PCG_THREADS / 64][PACK_STAGE];
    float* stage = stage_all[threadIdx.x >> 6];
    const int64_t nsl = (M + 63) / 64;
    const int64_t groups = (nsl + PCG_THREADS / 64 - 1) / (PCG_THREADS / 64);
    const int64_t P = run_plane_stride(M);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: scalar slice and offset
    for (int64_t g = blockIdx.x; g < groups; g += gridDim.x) {
        const int64_t sl = g * (PCG_THREADS / 64) + wv;
        const int64_t row = sl * 64 + lane64();
        const bool ok = row < M;
        const int64_t pos = sl < nsl ? soff[sl] : 0;
        const float sum = packed_row_dot<float, float>(as_global(runs), as_global(pvals), as_global(x), P, row, ok,
                                                       pos, stage);
        if (ok) y[row] = sum;
    }
}

// r = b - A x in FP64 (FP32 entries, FP64 x and products), 8 lanes per row with 16-byte loads of
// entries and columns. x == nullptr means x = 0 (r = b, no matrix pass).
__global__ void __launch_bounds__(PCG_THREADS) residual_f64_kernel(const int32_t* __restrict__ rowptr,
                                                                   const int32_t* __restrict__ cols,
                                                                   const float* __restrict__ vals, const double* __restrict__ b,
                                                                   const double* __restrict__ x, double* __restrict__ r,
                                                                   int64_t M, int64_t chunk, double* partials) {
    constexpr int G = 8;
    typedef float f4 __attribute__((ext_vector_type(4)));
    typedef int i4 __attribute__((ext_vector_type(4)));
    const int lb = logical_block();
    const int64_t r0 = (int64_t)lb * chunk;
    const int64_t r1 = min(M, r0 + chunk);
    const int lane = threadIdx.x % G, grp = threadIdx.x / G;
    constexpr int GROUPS = PCG_THREADS / G;
    double rr = 0, bb = 0;
    for (int64_t row = r0 + grp; row < r1; row += GROUPS) {
        double sum = 0.0;
        if (x) {
            // same branch-free first 128 entries as row_dot, then a loop for longer rows
            constexpr int U = 128 / (4 * G);
            const int32_t k0 = rowptr[row], k1 = rowptr[row + 1];
            if (k1 > k0) {
                const int32_t klast = k1 - 4;
                f4 v[U];
                i4 c[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int32_t k = min(k0 + 4 * lane + 4 * G * u, klast);
                    v[u] = *reinterpret_cast<const f4*>(vals + k);
                    c[u] = *reinterpret_cast<const i4*>(cols + k);
                }
                double xs[U][4];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    xs[u][0] = x[c[u].x];
                    xs[u][1] = x[c[u].y];
                    xs[u][2] = x[c[u].z];
                    xs[u][3] = x[c[u].w];
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const double part = (double)v[u].x * xs[u][0] + (double)v[u].y * xs[u][1] +
                                        (double)v[u].z * xs[u][2] + (double)v[u].w * xs[u][3];
                    sum += (k0 + 4 * lane + 4 * G * u < k1) ? part : 0.0;
                }
                for (int32_t k = k0 + 128 + 4 * lane; k < k1; k += 4 * G) {
                    const f4 vv = *reinterpret_cast<const f4*>(vals + k);
                    const i4 cc = *reinterpret_cast<const i4*>(cols + k);
                    sum += (double)vv.x * x[cc.x] + (double)vv.y * x[cc.y] + (double)vv.z * x[cc.z] + (double)vv.w * x[cc.w];
                }
            }
#pragma unroll
            for (int d = G / 2; d > 0; d >>= 1) sum += __shfl_xor(sum, d, 64);
        }
        if (lane == 0) {
            const double bi = b[row];
            const double ri = bi - sum;
            r[row] = ri;
            rr += ri * ri;
            bb += bi * bi;
        }
    }
    store_partials2(rr, bb, partials, partials + gridDim.x, lb);
}

// r = b - A x in FP64 for the run form (FP32 values, FP64 x and sums in the same column order).
__global__ void __launch_bounds__(PCG_THREADS) residual_f64_runs_kernel(const int32_t* __restrict__ runs,
                                                                        const float* __restrict__ rvals,
                                                                        const double* __restrict__ b,
                                                                        const double* __restrict__ x,
                                                                        double* __restrict__ r, int64_t M,
                                                                        double* partials) {
    const int64_t step = (int64_t)gridDim.x * PCG_THREADS;
    double rr = 0, bb = 0;
    for (int64_t row = run_rows_first(blockIdx.x); row < M; row += step) {
        const double bi = b[row];
        const double ri = x ? bi - run_row_dot<double, double>(as_global(runs), as_global(rvals), as_global(x),
                                                               run_plane_stride(M), row)
                            : bi;
        r[row] = ri;
        rr += ri * ri;
        bb += bi * bi;
    }
    store_partials2(rr, bb, partials, partials + gridDim.x, blockIdx.x);
}

// r = b - A x in FP64 for the packed form (the sums of residual_f64_runs_kernel, bit for bit): blocks
// take 256-row groups (4 slices), grid-stride. The norms are summed by norms2_kernel afterwards: with
// its FP64 accumulators the row loop here would hold 2 waves per SIMD instead of 3 (172 VGPRs), and a
// separate pass over r and b (16 bytes per row) costs less than the lost occupancy.
__global__ void __launch_bounds__(PCG_THREADS) __attribute__((amdgpu_waves_per_eu(3)))
residual_f64_packed_kernel(const int32_t* __restrict__ runs, const float* __restrict__ pvals,
                           const int64_t* __restrict__ soff, const double* __restrict__ b,
                           const double* __restrict__ x, double* __restrict__ r, int64_t M) {
    __shared__ float stage_all[PCG_THREADS / 64][PACK_STAGE];
    float* stage = stage_all[threadIdx.x >> 6];
    const int64_t step = (int64_t)gridDim.x * PCG_THREADS;
    const int64_t nsl = (M + 63) / 64;
    const int64_t P = run_plane_stride(M);
    for (int64_t base = (int64_t)blockIdx.x * PCG_THREADS; base < M; base += step) {
        const int64_t row = base + threadIdx.x;
        const bool ok = row < M;
        const int64_t sl = row >> 6;
        const int64_t pos = sl < nsl ? soff[sl] : 0;
        const double ax = x ? packed_row_dot<double, double>(as_global(runs), as_global(pvals), as_global(x), P, row,
                                                             ok, pos, stage)
                            : 0.0;
        if (ok) r[row] = x ? b[row] - ax : b[row];
    }
}

// Per-block partials of r.r and b.b over contiguous chunks of rows.
__global__ void __launch_bounds__(PCG_THREADS) norms2_kernel(const double* __restrict__ r, const double* __restrict__ b,
                                                             int64_t M, int64_t chunk, double* partials) {
    const int64_t r0 = (int64_t)blockIdx.x * chunk;
    const int64_t r1 = min(M, r0 + chunk);
    double rr = 0, bb = 0;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += PCG_THREADS) {
        const double ri = r[i], bi = b[i];
        rr += ri * ri;
        bb += bi * bi;
    }
    store_partials2(rr, bb, partials, partials + gridDim.x, blockIdx.x);
}

// One block: norms[0..1] = sums of the two partial arrays of a residual kernel (nb entries each).
__global__ void __launch_bounds__(PCG_THREADS) norms_finish_kernel(const double* partials, int nb, double* norms) {
    double a, b;
    sum_partials2(partials, partials + nb, nb, a, b);
    if (threadIdx.x == 0) {
        norms[0] = a;
        norms[1] = b;
    }
}

template <typename T>
__global__ void __launch_bounds__(PCG_THREADS) stencil_points_kernel(ngs_stencil op, const T* __restrict__ x,
                                                                     T* __restrict__ u) {
    const int64_t k = (int64_t)blockIdx.x * PCG_THREADS + threadIdx.x;
    if (k < op.n) u[k] = stencil_point<T>(op, x, k);
}

__global__ void __launch_bounds__(PCG_THREADS) stencil_rows_kernel(ngs_stencil op, const float* __restrict__ x,
                                                                   float* __restrict__ y, int64_t chunk) {
    __shared__ float lap[125];
    for (int t = threadIdx.x; t < 125; t += PCG_THREADS) lap[t] = op.lap[t];
    __syncthreads();
    const int64_t r0 = (int64_t)logical_block() * chunk;
    const int64_t r1 = min(op.level.M, r0 + chunk);
    for (int64_t row = r0 + threadIdx.x; row < r1; row += PCG_THREADS) y[row] = stencil_row<float>(op, lap, x, op.u, row);
}

__global__ void __launch_bounds__(PCG_THREADS) stencil_residual_kernel(ngs_stencil op, const double* __restrict__ b,
                                                                       const double* __restrict__ x,
                                                                       const double* __restrict__ u,
                                                                       double* __restrict__ r, int64_t chunk,
                                                                       double* partials) {
    __shared__ float lap[125];
    for (int t = threadIdx.x; t < 125; t += PCG_THREADS) lap[t] = op.lap[t];
    __syncthreads();
    const int64_t r0 = (int64_t)logical_block() * chunk;
    const int64_t r1 = min(op.level.M, r0 + chunk);
    double rr = 0, bb = 0;
    for (int64_t row = r0 + threadIdx.x; row < r1; row += PCG_THREADS) {
        const double ri = x ? b[row] - stencil_row<double>(op, lap, x, u, row) : b[row];
        r[row] = ri;
        rr += ri * ri;
        bb += b[row] * b[row];
    }
    store_partials2(rr, bb, partials, partials + gridDim.x, logical_block());
}

static int pcg_blocks(int64_t M, int64_t* chunk) {
    // ~256 rows per block (8 rows per 8-lane group): small levels get enough blocks that each
    // group walks only a few rows, so their iterations are not bound by a long latency chain. At most
    // 2560 = 2 x 256 CUs x 5: the grid-stride SpMV of large systems runs in two full rounds of resident
    // blocks (bpcg_packed_dot_kernel holds 5 blocks per CU)
    int64_t nb = (M + 255) / 256;
    if (nb > 2560) nb = 2560;
    if (nb >= NUM_XCD) nb = (nb / NUM_XCD) * NUM_XCD;
    if (nb < 1) nb = 1;
    // whole 64-row groups per block: the vector loads and stores of each wave are aligned 256-byte
    // segments (the last blocks may get fewer rows, or none)
    *chunk = ((M + nb - 1) / nb + 63) / 64 * 64;
    return (int)nb;
}

static size_t al256(size_t x) { return (x + 255) / 256 * 256; }

static int pcg_layout(int64_t M, PcgArgs* A, char* ws, void** tab = nullptr, int32_t** map = nullptr) {
    int64_t chunk;
    int nb = pcg_blocks(M, &chunk);
    size_t off = 0;
    A->st = reinterpret_cast<PcgState*>(ws + off);
    off += al256(sizeof(PcgState));
    A->partials = reinterpret_cast<double*>(ws + off);
    off += al256(sizeof(double) * 3 * nb);
    A->r = reinterpret_cast<float*>(ws + off);
    off += al256(sizeof(float) * M);
    A->p = reinterpret_cast<float*>(ws + off);
    off += al256(sizeof(float) * M);
    A->q = reinterpret_cast<float*>(ws + off);
    off += al256(sizeof(float) * M);
    if (tab) *tab = ws + off;
    off += 512;  // one batch table entry
    if (map) *map = reinterpret_cast<int32_t*>(ws + off);
    A->nblocks = nb;
    A->chunk = chunk;
    A->M = M;
    return (int)0;
}

static int64_t pcg_ws_bytes(int64_t M) {
    int64_t chunk;
    int nb = pcg_blocks(M, &chunk);
    return (int64_t)(al256(sizeof(PcgState)) + al256(sizeof(double) * 3 * nb) + 3 * al256(sizeof(float) * M) + 512 +
                     al256(sizeof(int32_t) * nb));
}

struct HostState {
    int32_t* pinned = nullptr;
    hipStream_t stream = nullptr;
};

static HostState& host_state() {
    static thread_local HostState hs;
    return hs;
}

// ------------------------------------------------------------------------------ batched solves
// Many independent systems advance together: a device table lists the active systems and the
// contiguous range of blocks each one owns; every kernel looks up its system from its block index,
// so one launch (and one captured graph, whose arguments are the fixed table pointers) serves any
// mix of systems. Systems join and leave between launches by rewriting the table.
enum : int32_t { SYS_CSR = 0, SYS_STENCIL = 1, SYS_RUNS = 2, SYS_PACKED = 3 };

// Flags of a launch sequence (which kernels the table's systems need).
enum : int32_t { FLAG_STENCIL = 1, FLAG_RUNS = 2, FLAG_NO_CSR = 4, FLAG_PACKED = 8 };

struct BatchSys {
    PcgArgs a;
    ngs_stencil op;  // kind == SYS_STENCIL
    int32_t kind;
    int32_t blk0;
    int32_t nblk;
    int32_t pad;
};
static_assert(sizeof(BatchSys) <= 512, "batch table entry must fit its workspace slot");

__device__ __forceinline__ const BatchSys* batch_lookup(const BatchSys* tab, const int32_t* blk2sys, int* local) {
    const int sys = blk2sys[blockIdx.x];
    if (sys < 0) return nullptr;
    const BatchSys* B = tab + sys;
    const int lb = blockIdx.x - B->blk0;
    const int nb = B->nblk;
    *local = (nb % NUM_XCD) ? lb : (lb % NUM_XCD) * (nb / NUM_XCD) + lb / NUM_XCD;
    return B;
}

// (iter, done) of slot 0 of up to kStatusGather systems, whose slot addresses come by value: one
// launch and one copy to the host per status read instead of one copy per system.
constexpr int kStatusGather = 128;
struct StatusSlots {
    const PcgSlot* p[kStatusGather];
};

__global__ void __launch_bounds__(kStatusGather) batch_status_kernel(StatusSlots slots, int32_t n, int32_t* out) {
    const int i = threadIdx.x;
    if (i < n) {
        const PcgSlot sl = *slots.p[i];
        out[2 * i] = sl.iter;
        out[2 * i + 1] = sl.done;
    }
}

// Kernels of one iteration step take the step parity `ph` (0, 1): the state slot they read.

// q = A p and per-block p.q partials for the CSR systems of the table (stencil systems:
// bpcg_stencil_dot_kernel).
template <int G, bool NT, bool IL>
__global__ void __launch_bounds__(PCG_THREADS) __attribute__((amdgpu_waves_per_eu(7)))
bpcg_spmv_dot_kernel(const BatchSys* tab, const int32_t* blk2sys, int ph) {
    int lb;
    const BatchSys* B = batch_lookup(tab, blk2sys, &lb);
    if (!B || B->kind != SYS_CSR) return;
    const PcgArgs& A = B->a;
    if (A.st->slot[ph].done) return;
    const int lane = threadIdx.x % G, grp = threadIdx.x / G;
    constexpr int GROUPS = PCG_THREADS / G;
    const RowTiles t = (IL && B->nblk % NUM_XCD == 0) ? tiles_interleaved(A.M, B->nblk, blockIdx.x - B->blk0, GROUPS)
                                                      : tiles_contiguous(A.M, A.chunk, lb, GROUPS);
    const NGS_GLOBAL int32_t* __restrict__ rowptr = as_global(A.rowptr);
    const NGS_GLOBAL int32_t* __restrict__ cols = as_global(A.cols);
    const NGS_GLOBAL float* __restrict__ vals = as_global(A.vals);
    const NGS_GLOBAL float* __restrict__ p = as_global(A.p);
    NGS_GLOBAL float* __restrict__ q = as_global(A.q);
    double* partials = A.partials;
    double pq = 0;
    for (int64_t t0 = t.first; t0 < t.end; t0 += t.step) {
        const int64_t row = t0 + grp;
        if (row >= t.hi) continue;
        const float prow = p[row];  // read with the row's other loads, not after its store
        const float sum = row_dot<G, NT>(rowptr[row], rowptr[row + 1], cols, vals, p, lane);
        if (lane == 0) {
            q[row] = sum;
            pq += (double)prow * sum;
        }
    }
    store_partials2(pq, 0.0, partials, nullptr, lb);
}

// q = A p and per-block p.q partials for the run-form systems of the table.
__global__ void __launch_bounds__(PCG_THREADS) bpcg_runs_dot_kernel(const BatchSys* tab, const int32_t* blk2sys,
                                                                    int ph) {
    int lb;
    const BatchSys* B = batch_lookup(tab, blk2sys, &lb);
    if (!B || B->kind != SYS_RUNS) return;
    const PcgArgs& A = B->a;
    if (A.st->slot[ph].done) return;
    const int64_t step = (int64_t)B->nblk * PCG_THREADS;
    const NGS_GLOBAL int32_t* __restrict__ runs = as_global(A.runs);
    const NGS_GLOBAL float* __restrict__ rvals = as_global(A.rvals);
    const NGS_GLOBAL float* __restrict__ p = as_global(A.p);
    NGS_GLOBAL float* __restrict__ q = as_global(A.q);
    double pq = 0;
    for (int64_t row = run_rows_first(blockIdx.x - B->blk0); row < A.M; row += step) {
        const float prow = p[row];
        const float sum = run_row_dot<float, float>(runs, rvals, p, run_plane_stride(A.M), row);
        q[row] = sum;
        pq += (double)prow * sum;
    }
    store_partials2(pq, 0.0, A.partials, nullptr, lb);
}

// The same for the packed-form systems: each step of a block takes 4 whole slices (256 rows from a
// multiple of 256), its 4 waves staging their slices' values through LDS in step. Capped at 5 waves per
// SIMD (96 VGPRs, no spills; 118 uncapped). The block count of large systems (pcg_blocks) is two full
// rounds of 5 blocks per CU, so the cap does not leave a partial last round: the finest level runs as at
// 4 waves (574 us), the coarse ones faster (55k rows: 31 -> 27 us per launch). With 2048 blocks the cap
// had cost 606 against 574 us at the finest level.
__global__ void __launch_bounds__(PCG_THREADS) __attribute__((amdgpu_waves_per_eu(5))) bpcg_packed_dot_kernel(const BatchSys* tab, const int32_t* blk2sys,
                                                                      int ph) {
    __shared__ float stage_all[PCG_THREADS / 64][PACK_STAGE];
    int lb;
    const BatchSys* B = batch_lookup(tab, blk2sys, &lb);
    if (!B || B->kind != SYS_PACKED) return;
    const PcgArgs& A = B->a;
    if (A.st->slot[ph].done) return;
    float* stage = stage_all[threadIdx.x >> 6];
    const int64_t step = (int64_t)B->nblk * PCG_THREADS;
    const int64_t nsl = (A.M + 63) / 64;
    const int64_t P = run_plane_stride(A.M);
    const NGS_GLOBAL float* __restrict__ p = as_global(A.p);
    NGS_GLOBAL float* __restrict__ q = as_global(A.q);
    double pq = 0;
    for (int64_t base = (int64_t)(blockIdx.x - B->blk0) * PCG_THREADS; base < A.M; base += step) {
        const int64_t row = base + threadIdx.x;
        const bool ok = row < A.M;
        const int64_t sl = row >> 6;
        const int64_t pos = sl < nsl ? A.soff[sl] : 0;
        const float prow = ok ? p[row] : 0.f;
        const float sum = packed_row_dot<float, float>(as_global(A.runs), as_global(A.rvals), p, P, row, ok, pos,
                                                       stage);
        if (ok) {
            q[row] = sum;
            pq += (double)prow * sum;
        }
    }
    store_partials2(pq, 0.0, A.partials, nullptr, lb);
}

// The same for the stencil systems of the table (row pass; the point pass ran before).
__global__ void __launch_bounds__(PCG_THREADS) bpcg_stencil_dot_kernel(const BatchSys* tab, const int32_t* blk2sys,
                                                                       int ph) {
    int lb;
    const BatchSys* B = batch_lookup(tab, blk2sys, &lb);
    if (!B || B->kind != SYS_STENCIL) return;
    const PcgArgs& A = B->a;
    if (A.st->slot[ph].done) return;
    const int64_t r0 = (int64_t)lb * A.chunk;
    const int64_t r1 = min(A.M, r0 + A.chunk);
    __shared__ float lap[125];
    for (int t = threadIdx.x; t < 125; t += PCG_THREADS) lap[t] = as_global(B->op.lap)[t];
    __syncthreads();
    const ngs_stencil op = B->op;
    const NGS_GLOBAL float* __restrict__ p = as_global(A.p);
    NGS_GLOBAL float* __restrict__ q = as_global(A.q);
    double* partials = A.partials;
    double pq = 0;
    for (int64_t row = r0 + threadIdx.x; row < r1; row += PCG_THREADS) {
        const float sum = stencil_row<float>(op, lap, A.p, op.u, row);
        q[row] = sum;
        pq += (double)p[row] * sum;
    }
    store_partials2(pq, 0.0, partials, nullptr, lb);
}

// Point pass of the stencil systems: u = point form of p, the points split evenly over the
// system's blocks (same XCD-contiguous mapping as its rows).
__global__ void __launch_bounds__(PCG_THREADS) bpcg_points_kernel(const BatchSys* tab, const int32_t* blk2sys, int ph) {
    int lb;
    const BatchSys* B = batch_lookup(tab, blk2sys, &lb);
    if (!B || B->kind != SYS_STENCIL) return;
    const PcgArgs& A = B->a;
    if (A.st->slot[ph].done) return;
    const ngs_stencil op = B->op;
    const int64_t per = (op.n + B->nblk - 1) / B->nblk;
    const int64_t k0 = (int64_t)lb * per;
    const int64_t k1 = min(op.n, k0 + per);
    const float* __restrict__ p = A.p;
    NGS_GLOBAL float* __restrict__ u = as_global(op.u);
    for (int64_t k = k0 + threadIdx.x; k < k1; k += PCG_THREADS) u[k] = stencil_point<float>(op, p, k);
}

// alpha = rz / sum(p.q) (every block sums the partials of the spmv_dot step), then the updates of
// x and r and per-block partials of r.z and r.r.
__global__ void __launch_bounds__(PCG_THREADS) bpcg_update_kernel(const BatchSys* tab, const int32_t* blk2sys, int ph) {
    int lb;
    const BatchSys* B = batch_lookup(tab, blk2sys, &lb);
    if (!B) return;
    const PcgArgs& A = B->a;
    if (A.st->slot[ph].done) return;
    const int nb = B->nblk;
    double* partials = A.partials;
    double pq, unused;
    sum_partials2(partials, nullptr, nb, pq, unused);
    const float alpha = (float)(A.st->rz[ph] / pq);
    const int64_t r0 = (int64_t)lb * A.chunk;
    const int64_t r1 = min(A.M, r0 + A.chunk);
    // local restrict copies: the stores below cannot alias the table or the other vectors
    NGS_GLOBAL float* __restrict__ x = as_global(A.x);
    NGS_GLOBAL float* __restrict__ r = as_global(A.r);
    const NGS_GLOBAL float* __restrict__ p = as_global(A.p);
    const NGS_GLOBAL float* __restrict__ q = as_global(A.q);
    const NGS_GLOBAL float* __restrict__ dinv = as_global(A.dinv);
    double rz = 0, rr = 0;
    // four rows per thread per step: every load of the step is issued before the first store (a
    // store counts against the same wait counter as the loads); single rows for the tail
    constexpr int R = 4;
    int64_t i = r0 + threadIdx.x;
    for (; i + (R - 1) * PCG_THREADS < r1; i += R * PCG_THREADS) {
        float pv[R], xv[R], rv[R], qv[R], dv[R];
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int64_t k = i + u * PCG_THREADS;
            pv[u] = p[k];
            xv[u] = x[k];
            rv[u] = r[k];
            qv[u] = q[k];
            dv[u] = dinv[k];
        }
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int64_t k = i + u * PCG_THREADS;
            const float ri = rv[u] - alpha * qv[u];
            const float z = dv[u] * ri;
            rz += (double)ri * z;
            rr += (double)ri * ri;
            x[k] = xv[u] + alpha * pv[u];
            r[k] = ri;
        }
    }
    for (; i < r1; i += PCG_THREADS) {
        const float ri = r[i] - alpha * q[i];
        const float z = dinv[i] * ri;
        rz += (double)ri * z;
        rr += (double)ri * ri;
        x[i] += alpha * p[i];
        r[i] = ri;
    }
    store_partials2(rz, rr, partials + nb, partials + 2 * nb, lb);
}

// Every block sums rz' and rr of the update step; the iteration ends the solve on convergence
// (rr <= tol^2 bb) or at the cap, otherwise p = D^-1 r + (rz'/rz) p. Block 0 advances the state
// into slot ph^1 (slot ph is still being read by the other blocks).
__global__ void __launch_bounds__(PCG_THREADS) bpcg_direction_kernel(const BatchSys* tab, const int32_t* blk2sys,
                                                                     int ph) {
    int lb;
    const BatchSys* B = batch_lookup(tab, blk2sys, &lb);
    if (!B) return;
    const PcgArgs& A = B->a;
    PcgState* st = A.st;
    if (st->slot[ph].done) return;
    const int nb = B->nblk;
    double rz_new, rr;
    sum_partials2(A.partials + nb, A.partials + 2 * nb, nb, rz_new, rr);
    const int it = st->slot[ph].iter + 1;
    const bool conv = rr <= st->tol2 * st->bb;
    const bool cap = !conv && it >= st->max_iter;
    if (!conv && !cap) {
        const float beta = (float)(rz_new / st->rz[ph]);
        const int64_t r0 = (int64_t)lb * A.chunk;
        const int64_t r1 = min(A.M, r0 + A.chunk);
        NGS_GLOBAL float* __restrict__ p = as_global(A.p);
        const NGS_GLOBAL float* __restrict__ r = as_global(A.r);
        const NGS_GLOBAL float* __restrict__ dinv = as_global(A.dinv);
        constexpr int R = 4;
        int64_t i = r0 + threadIdx.x;
        for (; i + (R - 1) * PCG_THREADS < r1; i += R * PCG_THREADS) {
            float pv[R], rv[R], dv[R];
#pragma unroll
            for (int u = 0; u < R; ++u) {
                const int64_t k = i + u * PCG_THREADS;
                pv[u] = p[k];
                rv[u] = r[k];
                dv[u] = dinv[k];
            }
#pragma unroll
            for (int u = 0; u < R; ++u) p[i + u * PCG_THREADS] = dv[u] * rv[u] + beta * pv[u];
        }
        for (; i < r1; i += PCG_THREADS) p[i] = dinv[i] * r[i] + beta * p[i];
    }
    if (lb == 0 && threadIdx.x == 0) {
        st->rr = rr;
        const int nph = ph ^ 1;
        if (conv || cap) {
            // final in both slots: later steps of either parity see the solve as ended
            const int d = conv ? 1 : 2;
            st->slot[nph].iter = it;
            st->slot[nph].done = d;
            st->slot[ph].iter = it;
            st->slot[ph].done = d;
        } else {
            st->rz[nph] = rz_new;
            st->slot[nph].iter = it;
        }
    }
}

__global__ void write_entry_kernel(BatchSys* dst, BatchSys src) { *dst = src; }

__global__ void fill_map_kernel(int32_t* map, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) map[i] = 0;
}

// SpMV variant: lanes per row (8, 16, 32), non-temporal matrix loads and interleaved row tiles;
// set by ngs_pcg_config.
static int g_spmv_lanes = 8;
static int g_spmv_nt = 0;
static int g_spmv_il = 1;

template <class F, class NT>
static void dispatch_spmv_g(F&& f, NT nt) {
    if (g_spmv_il) {
        if (g_spmv_lanes == 8) f(std::integral_constant<int, 8>(), nt, std::true_type());
        else if (g_spmv_lanes == 32) f(std::integral_constant<int, 32>(), nt, std::true_type());
        else f(std::integral_constant<int, 16>(), nt, std::true_type());
    } else {
        if (g_spmv_lanes == 8) f(std::integral_constant<int, 8>(), nt, std::false_type());
        else if (g_spmv_lanes == 32) f(std::integral_constant<int, 32>(), nt, std::false_type());
        else f(std::integral_constant<int, 16>(), nt, std::false_type());
    }
}

template <class F>
static void dispatch_spmv(F&& f) {
    if (g_spmv_nt) dispatch_spmv_g(f, std::true_type());
    else dispatch_spmv_g(f, std::false_type());
}

static void launch_batch_iteration(const BatchSys* tab, const int32_t* blk2sys, int total_blocks, int flags, int ph,
                                   hipStream_t st) {
    dim3 grid(total_blocks), block(PCG_THREADS);
    if (flags & FLAG_STENCIL) bpcg_points_kernel<<<grid, block, 0, st>>>(tab, blk2sys, ph);
    if (!(flags & FLAG_NO_CSR)) {
        dispatch_spmv([&](auto g, auto nt, auto il) {
            bpcg_spmv_dot_kernel<decltype(g)::value, decltype(nt)::value, decltype(il)::value>
                <<<grid, block, 0, st>>>(tab, blk2sys, ph);
        });
    }
    if (flags & FLAG_RUNS) bpcg_runs_dot_kernel<<<grid, block, 0, st>>>(tab, blk2sys, ph);
    if (flags & FLAG_PACKED) bpcg_packed_dot_kernel<<<grid, block, 0, st>>>(tab, blk2sys, ph);
    if (flags & FLAG_STENCIL) bpcg_stencil_dot_kernel<<<grid, block, 0, st>>>(tab, blk2sys, ph);
    bpcg_update_kernel<<<grid, block, 0, st>>>(tab, blk2sys, ph);
    bpcg_direction_kernel<<<grid, block, 0, st>>>(tab, blk2sys, ph);
}

}  // namespace ngs

using namespace ngs;

extern "C" {

int64_t ngs_pcg_workspace_bytes(int64_t M) { return pcg_ws_bytes(M); }
int ngs_pcg_batch_iterate(const void* table, const int32_t* blk2sys, int32_t total_blocks, int32_t iters,
                          int32_t use_graph, int32_t flags, ngs_stream_t s);

// Environment-free knobs for measurement: graph use and lanes per row (0 = automatic).
static int g_pcg_use_graph = 1;
static int g_pcg_group = 0;
int ngs_pcg_config(int32_t use_graph, int32_t lanes_per_row, int32_t nontemporal, int32_t interleave) {
    NGS_REQUIRE(lanes_per_row == 0 || lanes_per_row == 8 || lanes_per_row == 16 || lanes_per_row == 32,
                "ngs_pcg_config: lanes_per_row must be 0, 8, 16 or 32");
    g_pcg_use_graph = use_graph ? 1 : 0;
    g_pcg_group = lanes_per_row;
    g_spmv_lanes = lanes_per_row ? lanes_per_row : 8;
    g_spmv_nt = nontemporal ? 1 : 0;
    g_spmv_il = interleave ? 1 : 0;
    return NGS_OK;
}

// One system through the table-driven kernels: the single-entry table and its block map live in
// the workspace, so the captured chunk graph depends only on (workspace, blocks, chunk) and is
// replayed by every later solve that reuses the workspace.
static int pcg_solve_entry(BatchSys E, int64_t M, double tol, int32_t max_iter, int32_t chunk, void* workspace,
                           ngs_stream_t s, int32_t* iters, double* rel_resid) {
    if (chunk <= 0) chunk = 32;
    chunk += chunk & 1;  // launch sequences have an even number of steps (state slot parity)
    if (max_iter < 0) max_iter = 10000000;
    hipStream_t st = S(s);
    HostState& hs = host_state();
    if (!hs.pinned) NGS_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&hs.pinned), 64, hipHostMallocDefault));
    if (!hs.stream) NGS_CHECK_HIP(hipStreamCreateWithFlags(&hs.stream, hipStreamNonBlocking));
    hipStream_t work = hs.stream;
    // Order the private capture stream after the caller's stream.
    hipEvent_t ev;
    NGS_CHECK_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    NGS_CHECK_HIP(hipEventRecord(ev, st));
    NGS_CHECK_HIP(hipStreamWaitEvent(work, ev, 0));
    void* tab = nullptr;
    int32_t* map = nullptr;
    PcgArgs A = E.a;
    pcg_layout(M, &A, reinterpret_cast<char*>(workspace), &tab, &map);
    pcg_init_kernel<<<A.nblocks, PCG_THREADS, 0, work>>>(A);
    pcg_init_finish_kernel<<<1, PCG_THREADS, 0, work>>>(A, tol, max_iter);
    E.a = A;
    E.blk0 = 0;
    E.nblk = A.nblocks;
    write_entry_kernel<<<1, 1, 0, work>>>(reinterpret_cast<BatchSys*>(tab), E);
    fill_map_kernel<<<grid1d(A.nblocks, 256), 256, 0, work>>>(map, A.nblocks);
    NGS_CHECK_LAUNCH();
    const int flags = E.kind == SYS_STENCIL ? (FLAG_STENCIL | FLAG_NO_CSR)
                      : E.kind == SYS_RUNS ? (FLAG_RUNS | FLAG_NO_CSR)
                      : E.kind == SYS_PACKED ? (FLAG_PACKED | FLAG_NO_CSR)
                                             : 0;
    int status = NGS_OK;
    int32_t it = 0, done = 0;
    double sc[2];
    while (true) {
        int r = ngs_pcg_batch_iterate(tab, map, A.nblocks, chunk, g_pcg_use_graph, flags,
                                      reinterpret_cast<ngs_stream_t>(work));
        if (r != NGS_OK) return r;
        // one read per chunk: rr, bb and slot 0 (iteration count, done flag)
        NGS_CHECK_HIP(hipMemcpyAsync(hs.pinned, &A.st->rr, kStatusBytes, hipMemcpyDeviceToHost, work));
        NGS_CHECK_HIP(hipStreamSynchronize(work));
        read_status(hs.pinned, sc, &it, &done);
        if (done) break;
    }
    (void)hipEventDestroy(ev);
    if (iters) *iters = it;
    if (rel_resid) *rel_resid = sc[1] > 0 ? std::sqrt(sc[0] / sc[1]) : 0.0;
    if (done == 2) {
        set_error("ngs_pcg_solve: iteration cap reached before the tolerance");
        status = NGS_ERR_NOT_CONVERGED;
    }
    return status;
}

static bool stencil_ok(const ngs_stencil& op) {
    return op.level.coords && op.level.grid.bits && op.level.grid.rank && op.lap &&
           (op.n == 0 || (op.u && op.basis && op.point_cell && op.cell_start && op.cells.bits && op.cells.rank));
}

int ngs_pcg_solve(const int32_t* rowptr, const int32_t* cols, const float* vals, const float* inv_diag, const float* b,
                  float* x, int64_t M, double tol, int32_t max_iter, int32_t chunk, void* workspace, ngs_stream_t s,
                  int32_t* iters, double* rel_resid) {
    NGS_REQUIRE(M >= 0 && workspace, "ngs_pcg_solve: bad arguments");
    NGS_REQUIRE(tol >= 0, "ngs_pcg_solve: negative tolerance");
    if (M == 0) {
        if (iters) *iters = 0;
        if (rel_resid) *rel_resid = 0;
        return NGS_OK;
    }
    BatchSys E{};
    E.kind = SYS_CSR;
    E.a.rowptr = rowptr;
    E.a.cols = cols;
    E.a.vals = vals;
    E.a.dinv = inv_diag;
    E.a.b = b;
    E.a.x = x;
    return pcg_solve_entry(E, M, tol, max_iter, chunk, workspace, s, iters, rel_resid);
}

int ngs_pcg_solve_runs(const int32_t* runs, const float* rvals, const float* inv_diag, const float* b, float* x,
                       int64_t M, double tol, int32_t max_iter, int32_t chunk, void* workspace, ngs_stream_t s,
                       int32_t* iters, double* rel_resid) {
    NGS_REQUIRE(M >= 0 && workspace, "ngs_pcg_solve_runs: bad arguments");
    NGS_REQUIRE(tol >= 0, "ngs_pcg_solve_runs: negative tolerance");
    if (M == 0) {
        if (iters) *iters = 0;
        if (rel_resid) *rel_resid = 0;
        return NGS_OK;
    }
    NGS_REQUIRE(runs && rvals && inv_diag && b && x, "ngs_pcg_solve_runs: null array");
    NGS_REQUIRE(M <= RUN_START_MASK + 1, "ngs_pcg_solve_runs: too many rows for run words");
    BatchSys E{};
    E.kind = SYS_RUNS;
    E.a.runs = runs;
    E.a.rvals = rvals;
    E.a.dinv = inv_diag;
    E.a.b = b;
    E.a.x = x;
    return pcg_solve_entry(E, M, tol, max_iter, chunk, workspace, s, iters, rel_resid);
}

// ------------------------------------------------------------------ solve without allocation
// The workspace of ngs_pcg_solve_runs_noalloc is a file-scope device array, one copy per GPU, held
// until the process ends. On the host the array's name is only a stand-in symbol, never a device
// address: the address of the current device's copy is looked up once per device with
// hipGetSymbolAddress and kept in g_scratch_ptr, and only that address reaches kernels.
static __device__ __align__(256) char scratch[NGS_PCG_NOALLOC_SCRATCH_BYTES];

constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[SCRATCH_MAX_DEVICES];

static int scratch_address(void** out) {
    int dev = -1;
    NGS_CHECK_HIP(hipGetDevice(&dev));
    NGS_REQUIRE(dev >= 0 && dev < SCRATCH_MAX_DEVICES, "ngs_pcg_solve_runs_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("ngs_pcg_solve_runs_noalloc: hipGetSymbolAddress of the scratch array failed: ") +
                      hipGetErrorString(e));
            return NGS_ERR_HIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return NGS_OK;
}

int ngs_pcg_solve_runs_noalloc(const int32_t* runs, const float* rvals, const float* inv_diag, const float* b,
                               float* x, int64_t M, double tol, int32_t max_iter, int32_t chunk, ngs_stream_t s,
                               int32_t* iters, double* rel_resid) {
    NGS_REQUIRE(M >= 0, "ngs_pcg_solve_runs_noalloc: bad arguments");
    NGS_REQUIRE(tol >= 0, "ngs_pcg_solve_runs_noalloc: negative tolerance");
    if (M == 0) {
        if (iters) *iters = 0;
        if (rel_resid) *rel_resid = 0;
        return NGS_OK;
    }
    NGS_REQUIRE(runs && rvals && inv_diag && b && x, "ngs_pcg_solve_runs_noalloc: null array");
    NGS_REQUIRE(M <= RUN_START_MASK + 1, "ngs_pcg_solve_runs_noalloc: too many rows for run words");
    const int64_t need = pcg_ws_bytes(M);
    if (need > (int64_t)NGS_PCG_NOALLOC_SCRATCH_BYTES) {
        set_error("ngs_pcg_solve_runs_noalloc: " + std::to_string(M) + " rows need " + std::to_string(need) +
                  " bytes of workspace, above the " + std::to_string((int64_t)NGS_PCG_NOALLOC_SCRATCH_BYTES) +
                  "-byte scratch array (at most " + std::to_string((int64_t)NGS_PCG_NOALLOC_MAX_ROWS) +
                  " rows); use ngs_pcg_solve_runs with a workspace");
        return NGS_ERR_CAPACITY;
    }
    // the advertised row limit is exactly the largest system whose layout fits the array
    NGS_REQUIRE(pcg_ws_bytes(NGS_PCG_NOALLOC_MAX_ROWS) <= (int64_t)NGS_PCG_NOALLOC_SCRATCH_BYTES &&
                    pcg_ws_bytes(NGS_PCG_NOALLOC_MAX_ROWS + 1) > (int64_t)NGS_PCG_NOALLOC_SCRATCH_BYTES,
                "ngs_pcg_solve_runs_noalloc: row limit does not match the workspace layout");
    void* ws = nullptr;
    if (int r = scratch_address(&ws)) return r;
    if (chunk <= 0) chunk = 32;
    chunk += chunk & 1;  // launch sequences have an even number of steps (state slot parity)
    if (max_iter < 0) max_iter = 10000000;
    hipStream_t st = S(s);
    // Everything runs on the caller's stream, launched directly (no private stream, event or graph).
    // The init kernels write every workspace word the iterations read before they read it: x, r, p,
    // the r.z / b.b partials, the state of both slots, the table entry and the block map.
    BatchSys E{};
    E.kind = SYS_RUNS;
    E.a.runs = runs;
    E.a.rvals = rvals;
    E.a.dinv = inv_diag;
    E.a.b = b;
    E.a.x = x;
    void* tab = nullptr;
    int32_t* map = nullptr;
    PcgArgs A = E.a;
    pcg_layout(M, &A, reinterpret_cast<char*>(ws), &tab, &map);
    pcg_init_kernel<<<A.nblocks, PCG_THREADS, 0, st>>>(A);
    pcg_init_finish_kernel<<<1, PCG_THREADS, 0, st>>>(A, tol, max_iter);
    E.a = A;
    E.blk0 = 0;
    E.nblk = A.nblocks;
    write_entry_kernel<<<1, 1, 0, st>>>(reinterpret_cast<BatchSys*>(tab), E);
    fill_map_kernel<<<grid1d(A.nblocks, 256), 256, 0, st>>>(map, A.nblocks);
    NGS_CHECK_LAUNCH();
    PcgSlot slot{};
    double sc[2];
    alignas(8) char status_bytes[kStatusBytes];
    while (true) {
        if (int r = ngs_pcg_batch_iterate(tab, map, A.nblocks, chunk, 0, FLAG_RUNS | FLAG_NO_CSR, s)) return r;
        // one read per chunk: rr, bb and slot 0 (iteration count, done flag)
        NGS_CHECK_HIP(hipMemcpyAsync(status_bytes, &A.st->rr, kStatusBytes, hipMemcpyDeviceToHost, st));
        NGS_CHECK_HIP(hipStreamSynchronize(st));
        read_status(status_bytes, sc, &slot.iter, &slot.done);
        if (slot.done) break;
    }
    if (iters) *iters = slot.iter;
    if (rel_resid) *rel_resid = sc[1] > 0 ? std::sqrt(sc[0] / sc[1]) : 0.0;
    if (slot.done == 2) {
        set_error("ngs_pcg_solve_runs_noalloc: iteration cap reached before the tolerance");
        return NGS_ERR_NOT_CONVERGED;
    }
    return NGS_OK;
}

int ngs_pcg_solve_stencil(ngs_stencil op, const float* inv_diag, const float* b, float* x, double tol, int32_t max_iter,
                          int32_t chunk, void* workspace, ngs_stream_t s, int32_t* iters, double* rel_resid) {
    const int64_t M = op.level.M;
    NGS_REQUIRE(M >= 0 && workspace && inv_diag && b && x, "ngs_pcg_solve_stencil: bad arguments");
    NGS_REQUIRE(tol >= 0, "ngs_pcg_solve_stencil: negative tolerance");
    if (M == 0) {
        if (iters) *iters = 0;
        if (rel_resid) *rel_resid = 0;
        return NGS_OK;
    }
    NGS_REQUIRE(stencil_ok(op), "ngs_pcg_solve_stencil: incomplete operator");
    BatchSys E{};
    E.kind = SYS_STENCIL;
    E.op = op;
    E.a.dinv = inv_diag;
    E.a.b = b;
    E.a.x = x;
    return pcg_solve_entry(E, M, tol, max_iter, chunk, workspace, s, iters, rel_resid);
}

int ngs_spmv(const int32_t* rowptr, const int32_t* cols, const float* vals, const float* x, float* y, int64_t M,
             ngs_stream_t s) {
    if (M <= 0) return NGS_OK;
    int64_t chunk;
    int nb = pcg_blocks(M, &chunk);
    const int il_saved = g_spmv_il;
    if (nb % NUM_XCD) g_spmv_il = 0;  // interleaved tiles need a grid that is a multiple of 8
    dispatch_spmv([&](auto g, auto nt, auto il) {
        spmv_kernel<decltype(g)::value, decltype(nt)::value, decltype(il)::value>
            <<<nb, PCG_THREADS, 0, S(s)>>>(rowptr, cols, vals, x, y, M, chunk);
    });
    g_spmv_il = il_saved;
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_spmv_runs(const int32_t* runs, const float* rvals, const float* x, float* y, int64_t M, ngs_stream_t s) {
    if (M <= 0) return NGS_OK;
    NGS_REQUIRE(runs && rvals && x && y, "ngs_spmv_runs: null array");
    NGS_REQUIRE(M <= RUN_START_MASK + 1, "ngs_spmv_runs: too many rows for run words");
    int64_t chunk;
    int nb = pcg_blocks(M, &chunk);
    spmv_runs_kernel<<<nb, PCG_THREADS, 0, S(s)>>>(runs, rvals, x, y, M);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_runs_slice_counts(const int32_t* runs, int64_t M, int64_t* counts, ngs_stream_t s) {
    if (M <= 0) return NGS_OK;
    NGS_REQUIRE(runs && counts, "ngs_runs_slice_counts: null array");
    NGS_REQUIRE(M <= RUN_START_MASK + 1, "ngs_runs_slice_counts: too many rows for run words");
    runs_slice_counts_kernel<<<grid1d((M + 63) / 64 * 64, PCG_THREADS), PCG_THREADS, 0, S(s)>>>(runs, M, counts);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_runs_pack(const int32_t* runs, const float* rvals, int64_t M, const int64_t* soff, float* pvals,
                  ngs_stream_t s) {
    if (M <= 0) return NGS_OK;
    NGS_REQUIRE(runs && rvals && soff && pvals, "ngs_runs_pack: null array");
    NGS_REQUIRE(M <= RUN_START_MASK + 1, "ngs_runs_pack: too many rows for run words");
    runs_pack_kernel<<<grid1d((M + 63) / 64 * 64, PCG_THREADS), PCG_THREADS, 0, S(s)>>>(runs, rvals, M, soff, pvals);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_runs_unpack(const int32_t* runs, const float* pvals, int64_t M, const int64_t* soff, float* rvals,
                    ngs_stream_t s) {
    if (M <= 0) return NGS_OK;
    NGS_REQUIRE(runs && pvals && soff && rvals, "ngs_runs_unpack: null array");
    NGS_REQUIRE(M <= RUN_START_MASK + 1, "ngs_runs_unpack: too many rows for run words");
    runs_unpack_kernel<<<grid1d((M + 63) / 64 * 64, PCG_THREADS), PCG_THREADS, 0, S(s)>>>(runs, pvals, M, soff, rvals);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_spmv_packed(const int32_t* runs, const float* pvals, const int64_t* soff, const float* x, float* y, int64_t M,
                    ngs_stream_t s) {
    if (M <= 0) return NGS_OK;
    NGS_REQUIRE(runs && pvals && soff && x && y, "ngs_spmv_packed: null array");
    NGS_REQUIRE(M <= RUN_START_MASK + 1, "ngs_spmv_packed: too many rows for run words");
    const int64_t groups = ((M + 63) / 64 + 3) / 4;
    const int nb = (int)std::min<int64_t>(groups, 8192);
    spmv_packed_kernel<<<nb, PCG_THREADS, 0, S(s)>>>(runs, pvals, soff, x, y, M);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_pcg_solve_packed(const int32_t* runs, const float* pvals, const int64_t* soff, const float* inv_diag,
                         const float* b, float* x, int64_t M, double tol, int32_t max_iter, int32_t chunk,
                         void* workspace, ngs_stream_t s, int32_t* iters, double* rel_resid) {
    NGS_REQUIRE(M >= 0 && workspace, "ngs_pcg_solve_packed: bad arguments");
    NGS_REQUIRE(tol >= 0, "ngs_pcg_solve_packed: negative tolerance");
    if (M == 0) {
        if (iters) *iters = 0;
        if (rel_resid) *rel_resid = 0;
        return NGS_OK;
    }
    NGS_REQUIRE(runs && pvals && soff && inv_diag && b && x, "ngs_pcg_solve_packed: null array");
    NGS_REQUIRE(M <= RUN_START_MASK + 1, "ngs_pcg_solve_packed: too many rows for run words");
    BatchSys E{};
    E.kind = SYS_PACKED;
    E.a.runs = runs;
    E.a.rvals = pvals;
    E.a.soff = soff;
    E.a.dinv = inv_diag;
    E.a.b = b;
    E.a.x = x;
    return pcg_solve_entry(E, M, tol, max_iter, chunk, workspace, s, iters, rel_resid);
}

int ngs_residual_f64_packed(const int32_t* runs, const float* pvals, const int64_t* soff, const double* b,
                            const double* x, double* r, int64_t M, double* norms, void* workspace, ngs_stream_t s) {
    NGS_REQUIRE(workspace && b && r && norms && (!x || (runs && pvals && soff)),
                "ngs_residual_f64_packed: bad arguments");
    NGS_REQUIRE(M <= RUN_START_MASK + 1, "ngs_residual_f64_packed: too many rows for run words");
    if (M <= 0) return NGS_OK;
    int64_t chunk;
    int nb = pcg_blocks(M, &chunk);
    char* w = reinterpret_cast<char*>(workspace);
    double* partials = reinterpret_cast<double*>(w + al256(sizeof(PcgState)));
    const int64_t groups = ((M + 63) / 64 + 3) / 4;
    residual_f64_packed_kernel<<<(int)std::min<int64_t>(groups, 8192), PCG_THREADS, 0, S(s)>>>(runs, pvals, soff, b, x,
                                                                                               r, M);
    norms2_kernel<<<nb, PCG_THREADS, 0, S(s)>>>(r, b, M, chunk, partials);
    norms_finish_kernel<<<1, PCG_THREADS, 0, S(s)>>>(partials, nb, norms);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_residual_f64(const int32_t* rowptr, const int32_t* cols, const float* vals, const double* b, const double* x,
                     double* r, int64_t M, double* norms, void* workspace, ngs_stream_t s) {
    NGS_REQUIRE(workspace && b && r && norms && (!x || (rowptr && cols && vals)), "ngs_residual_f64: bad arguments");
    if (M <= 0) return NGS_OK;
    int64_t chunk;
    int nb = pcg_blocks(M, &chunk);
    char* w = reinterpret_cast<char*>(workspace);
    double* partials = reinterpret_cast<double*>(w + al256(sizeof(PcgState)));
    residual_f64_kernel<<<nb, PCG_THREADS, 0, S(s)>>>(rowptr, cols, vals, b, x, r, M, chunk, partials);
    norms_finish_kernel<<<1, PCG_THREADS, 0, S(s)>>>(partials, nb, norms);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_residual_f64_runs(const int32_t* runs, const float* rvals, const double* b, const double* x, double* r,
                          int64_t M, double* norms, void* workspace, ngs_stream_t s) {
    NGS_REQUIRE(workspace && b && r && norms && (!x || (runs && rvals)), "ngs_residual_f64_runs: bad arguments");
    NGS_REQUIRE(M <= RUN_START_MASK + 1, "ngs_residual_f64_runs: too many rows for run words");
    if (M <= 0) return NGS_OK;
    int64_t chunk;
    int nb = pcg_blocks(M, &chunk);
    char* w = reinterpret_cast<char*>(workspace);
    double* partials = reinterpret_cast<double*>(w + al256(sizeof(PcgState)));
    residual_f64_runs_kernel<<<nb, PCG_THREADS, 0, S(s)>>>(runs, rvals, b, x, r, M, partials);
    norms_finish_kernel<<<1, PCG_THREADS, 0, S(s)>>>(partials, nb, norms);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_stencil_pass(ngs_stencil op, const float* x, float* y, int32_t passes, ngs_stream_t s) {
    const int64_t M = op.level.M;
    if (M <= 0) return NGS_OK;
    NGS_REQUIRE(x && (y || !(passes & 2)) && (passes & 3) && stencil_ok(op), "ngs_stencil_pass: bad arguments");
    if ((passes & 1) && op.n > 0)
        stencil_points_kernel<float><<<grid1d(op.n, PCG_THREADS), PCG_THREADS, 0, S(s)>>>(op, x, op.u);
    if (passes & 2) {
        int64_t chunk;
        int nb = pcg_blocks(M, &chunk);
        stencil_rows_kernel<<<nb, PCG_THREADS, 0, S(s)>>>(op, x, y, chunk);
    }
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_stencil_apply(ngs_stencil op, const float* x, float* y, ngs_stream_t s) { return ngs_stencil_pass(op, x, y, 3, s); }

int ngs_stencil_residual_f64(ngs_stencil op, const double* b, const double* x, double* r, double* u64, double* norms,
                             void* workspace, ngs_stream_t s) {
    const int64_t M = op.level.M;
    if (M <= 0) return NGS_OK;
    NGS_REQUIRE(b && r && norms && workspace && (op.n == 0 || u64) && stencil_ok(op),
                "ngs_stencil_residual_f64: bad arguments");
    if (x && op.n > 0) stencil_points_kernel<double><<<grid1d(op.n, PCG_THREADS), PCG_THREADS, 0, S(s)>>>(op, x, u64);
    int64_t chunk;
    int nb = pcg_blocks(M, &chunk);
    char* w = reinterpret_cast<char*>(workspace);
    double* partials = reinterpret_cast<double*>(w + al256(sizeof(PcgState)));
    stencil_residual_kernel<<<nb, PCG_THREADS, 0, S(s)>>>(op, b, x, u64, r, chunk, partials);
    norms_finish_kernel<<<1, PCG_THREADS, 0, S(s)>>>(partials, nb, norms);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

// ----------------------------------------------------------------------------- batched ABI

int32_t ngs_pcg_batch_blocks(int64_t M);

// Size in bytes of one entry of the batch table (BatchSys).
int64_t ngs_pcg_batch_entry_bytes(void) { return (int64_t)sizeof(BatchSys); }

// Prepare one system for batched iterations: lay out its workspace (as ngs_pcg_solve does),
// run the initialisation (x = 0, r = b, p = D^-1 b, norms) on `s`, and write its table entry
// (host memory, `entry`, ngs_pcg_batch_entry_bytes bytes) with the given block range.
static int batch_prepare(BatchSys* E, int64_t M, double tol, int32_t max_iter, void* workspace, int32_t blk0,
                         int32_t nblk, ngs_stream_t s) {
    PcgArgs A = E->a;
    pcg_layout(M, &A, reinterpret_cast<char*>(workspace));
    if (max_iter < 0) max_iter = 10000000;
    pcg_init_kernel<<<A.nblocks, PCG_THREADS, 0, S(s)>>>(A);
    pcg_init_finish_kernel<<<1, PCG_THREADS, 0, S(s)>>>(A, tol, max_iter);
    NGS_CHECK_LAUNCH();
    E->a = A;
    E->blk0 = blk0;
    E->nblk = nblk;
    return NGS_OK;
}

int ngs_pcg_batch_prepare(const int32_t* rowptr, const int32_t* cols, const float* vals, const float* inv_diag,
                          const float* b, float* x, int64_t M, double tol, int32_t max_iter, void* workspace,
                          int32_t blk0, int32_t nblk, void* entry, ngs_stream_t s) {
    NGS_REQUIRE(M > 0 && workspace && entry, "ngs_pcg_batch_prepare: bad arguments");
    NGS_REQUIRE(nblk == ngs_pcg_batch_blocks(M), "ngs_pcg_batch_prepare: nblk must be ngs_pcg_batch_blocks(M)");
    BatchSys E{};
    E.kind = SYS_CSR;
    E.a.rowptr = rowptr;
    E.a.cols = cols;
    E.a.vals = vals;
    E.a.dinv = inv_diag;
    E.a.b = b;
    E.a.x = x;
    int r = batch_prepare(&E, M, tol, max_iter, workspace, blk0, nblk, s);
    if (r == NGS_OK) *reinterpret_cast<BatchSys*>(entry) = E;
    return r;
}

int ngs_pcg_batch_prepare_runs(const int32_t* runs, const float* rvals, const float* inv_diag, const float* b,
                               float* x, int64_t M, double tol, int32_t max_iter, void* workspace, int32_t blk0,
                               int32_t nblk, void* entry, ngs_stream_t s) {
    NGS_REQUIRE(M > 0 && workspace && entry && runs && rvals && inv_diag && b && x,
                "ngs_pcg_batch_prepare_runs: bad arguments");
    NGS_REQUIRE(M <= RUN_START_MASK + 1, "ngs_pcg_batch_prepare_runs: too many rows for run words");
    NGS_REQUIRE(nblk == ngs_pcg_batch_blocks(M), "ngs_pcg_batch_prepare_runs: nblk must be ngs_pcg_batch_blocks(M)");
    BatchSys E{};
    E.kind = SYS_RUNS;
    E.a.runs = runs;
    E.a.rvals = rvals;
    E.a.dinv = inv_diag;
    E.a.b = b;
    E.a.x = x;
    int r = batch_prepare(&E, M, tol, max_iter, workspace, blk0, nblk, s);
    if (r == NGS_OK) *reinterpret_cast<BatchSys*>(entry) = E;
    return r;
}

int ngs_pcg_batch_prepare_packed(const int32_t* runs, const float* pvals, const int64_t* soff, const float* inv_diag,
                                 const float* b, float* x, int64_t M, double tol, int32_t max_iter, void* workspace,
                                 int32_t blk0, int32_t nblk, void* entry, ngs_stream_t s) {
    NGS_REQUIRE(M > 0 && workspace && entry && runs && pvals && soff && inv_diag && b && x,
                "ngs_pcg_batch_prepare_packed: bad arguments");
    NGS_REQUIRE(M <= RUN_START_MASK + 1, "ngs_pcg_batch_prepare_packed: too many rows for run words");
    NGS_REQUIRE(nblk == ngs_pcg_batch_blocks(M), "ngs_pcg_batch_prepare_packed: nblk must be ngs_pcg_batch_blocks(M)");
    BatchSys E{};
    E.kind = SYS_PACKED;
    E.a.runs = runs;
    E.a.rvals = pvals;
    E.a.soff = soff;
    E.a.dinv = inv_diag;
    E.a.b = b;
    E.a.x = x;
    int r = batch_prepare(&E, M, tol, max_iter, workspace, blk0, nblk, s);
    if (r == NGS_OK) *reinterpret_cast<BatchSys*>(entry) = E;
    return r;
}

int ngs_pcg_batch_prepare_stencil(ngs_stencil op, const float* inv_diag, const float* b, float* x, double tol,
                                  int32_t max_iter, void* workspace, int32_t blk0, int32_t nblk, void* entry,
                                  ngs_stream_t s) {
    const int64_t M = op.level.M;
    NGS_REQUIRE(M > 0 && workspace && entry && inv_diag && b && x, "ngs_pcg_batch_prepare_stencil: bad arguments");
    NGS_REQUIRE(stencil_ok(op), "ngs_pcg_batch_prepare_stencil: incomplete operator");
    NGS_REQUIRE(nblk == ngs_pcg_batch_blocks(M), "ngs_pcg_batch_prepare_stencil: nblk must be ngs_pcg_batch_blocks(M)");
    BatchSys E{};
    E.kind = SYS_STENCIL;
    E.op = op;
    E.a.dinv = inv_diag;
    E.a.b = b;
    E.a.x = x;
    int r = batch_prepare(&E, M, tol, max_iter, workspace, blk0, nblk, s);
    if (r == NGS_OK) *reinterpret_cast<BatchSys*>(entry) = E;
    return r;
}

// Moves a prepared host table entry to the block range starting at blk0.
int ngs_pcg_batch_place(void* entry, int32_t blk0) {
    NGS_REQUIRE(entry && blk0 >= 0 && blk0 % NUM_XCD == 0, "ngs_pcg_batch_place: bad arguments");
    reinterpret_cast<BatchSys*>(entry)->blk0 = blk0;
    return NGS_OK;
}

// Blocks a system of M rows occupies in a batch (its single-system grid).
int32_t ngs_pcg_batch_blocks(int64_t M) {
    int64_t chunk;
    return pcg_blocks(M, &chunk);
}

// Run `iters` batched iterations over the device table (nsys entries) and block map
// (total_blocks entries, -1 = idle block). With use_graph, the iterations are captured once per
// (table, map, total_blocks, iters) and replayed; the table contents may change between calls.
int ngs_pcg_batch_iterate(const void* table, const int32_t* blk2sys, int32_t total_blocks, int32_t iters,
                          int32_t use_graph, int32_t flags, ngs_stream_t s) {
    NGS_REQUIRE(table && blk2sys && total_blocks > 0 && iters > 0 && iters % 2 == 0,
                "ngs_pcg_batch_iterate: bad arguments (iters must be even and positive)");
    const BatchSys* tab = reinterpret_cast<const BatchSys*>(table);
    hipStream_t st = S(s);
    if (!use_graph) {
        for (int i = 0; i < iters; ++i) launch_batch_iteration(tab, blk2sys, total_blocks, flags, i & 1, st);
        NGS_CHECK_LAUNCH();
        return NGS_OK;
    }
    // Graphs depend only on (table, map, blocks, iterations, stream). The cache is per thread (each
    // thread launches on its own streams), bounded, and evicts its oldest entry.
    // the SpMV variant is part of the key: a graph records the kernels of the variant it was captured with
    struct Key { const void* t; const int32_t* m; int nb, it, fl, var; hipStream_t s; hipGraphExec_t e; };
    const int var = g_spmv_lanes * 4 + g_spmv_nt * 2 + g_spmv_il;
    static thread_local std::vector<Key> cache;
    hipGraphExec_t exec = nullptr;
    for (auto& k : cache)
        if (k.t == table && k.m == blk2sys && k.nb == total_blocks && k.it == iters && k.fl == flags && k.var == var &&
            k.s == st)
            exec = k.e;
    if (!exec) {
        hipGraph_t graph;
        NGS_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < iters; ++i) launch_batch_iteration(tab, blk2sys, total_blocks, flags, i & 1, st);
        NGS_CHECK_HIP(hipStreamEndCapture(st, &graph));
        NGS_CHECK_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        (void)hipGraphDestroy(graph);
        if (cache.size() >= 256) {
            (void)hipGraphExecDestroy(cache.front().e);
            cache.erase(cache.begin());
        }
        cache.push_back({table, blk2sys, total_blocks, iters, flags, var, st, exec});
    }
    NGS_CHECK_HIP(hipGraphLaunch(exec, st));
    return NGS_OK;
}

// Copy (iter, done) of each system in the table to host memory states[2*i], states[2*i+1]
// through the device scratch `dev_states` (2*nsys int32): one gather launch per 128 systems and
// one copy; synchronises `s`.
int ngs_pcg_batch_status(const void* host_table, int32_t nsys, int32_t* dev_states, int32_t* states,
                         ngs_stream_t s) {
    NGS_REQUIRE(nsys >= 0, "ngs_pcg_batch_status: negative system count");
    if (nsys == 0) return NGS_OK;
    NGS_REQUIRE(host_table && dev_states && states, "ngs_pcg_batch_status: null argument");
    const BatchSys* tab = reinterpret_cast<const BatchSys*>(host_table);
    hipStream_t st = S(s);
    for (int i0 = 0; i0 < nsys; i0 += ngs::kStatusGather) {
        const int m = std::min(ngs::kStatusGather, nsys - i0);
        ngs::StatusSlots sl{};
        for (int k = 0; k < m; ++k) sl.p[k] = &tab[i0 + k].a.st->slot[0];
        ngs::batch_status_kernel<<<1, ngs::kStatusGather, 0, st>>>(sl, m, dev_states + 2 * (int64_t)i0);
        NGS_CHECK_LAUNCH();
    }
    NGS_CHECK_HIP(hipMemcpyAsync(states, dev_states, 2 * sizeof(int32_t) * (size_t)nsys, hipMemcpyDeviceToHost, st));
    NGS_CHECK_HIP(hipStreamSynchronize(st));
    return NGS_OK;
}

}  // extern "C"
