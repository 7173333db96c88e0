// Keys, cuckoo table, ordered bitmap sets, scans, density/levels, splat and field evaluation.
#include "common.h"

#include <cmath>
#include <mutex>
#include <vector>

namespace ngs {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

// ========================================================================================
// Scan (exclusive prefix sums): single-pass rocPRIM scans.
// ========================================================================================
int scan_counts_device(const int32_t* counts, int64_t n, int32_t* offsets, void* tmp, size_t* tmp_bytes, hipStream_t st);
int scan_popc_device(const uint32_t* words, int64_t n, int32_t* rank, void* tmp, size_t* tmp_bytes, hipStream_t st);

// offsets[0..n] <- exclusive scan of counts[0..n), offsets[n] = total. ws holds ngs_scan_workspace_bytes(n) bytes.
int scan_i32_internal(const int32_t* counts, int64_t n, int32_t* offsets, void* ws, hipStream_t st) {
    size_t need = 0;
    int r = scan_counts_device(nullptr, n, nullptr, nullptr, &need, st);
    if (r != NGS_OK) return r;
    return scan_counts_device(counts, n, offsets, ws, &need, st);
}

__global__ void popc_total_kernel(const uint32_t* words, const int32_t* rank, int64_t n, int64_t* total) {
    *total = n > 0 ? (int64_t)rank[n - 1] + __popc(words[n - 1]) : 0;
}

// ========================================================================================
// FNV keys
// ========================================================================================
__global__ void hash_rows_kernel(const int32_t* __restrict__ c, int64_t n, int dim, uint64_t* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t row[4];
    for (int j = 0; j < dim; ++j) row[j] = c[i * dim + j];
    out[i] = fnv_fold60(row, dim);
}

__global__ void kernel_hash_kernel(const int32_t* __restrict__ c, int64_t n, int dim, const int32_t* __restrict__ off,
                                   int k, uint64_t* __restrict__ out) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * k) return;
    int64_t kk = t / n, i = t % n;
    int32_t row[4];
    for (int j = 0; j < dim; ++j) row[j] = c[i * dim + j];
    for (int j = 0; j < 3; ++j) row[j] += off[kk * 3 + j];
    out[t] = fnv_fold60(row, dim);
}

// ========================================================================================
// Cuckoo table
// ========================================================================================
constexpr int BUCKET = 512;
constexpr uint64_t EMPTY_KEY = ~0ull;

__device__ __forceinline__ int slot_hash(uint64_t key, uint64_t mult) { return (int)((key * mult) >> 55) & (BUCKET - 1); }

struct HashWs {
    int32_t* bucket_count;  // num_buckets
    int32_t* staged;        // num_buckets * BUCKET key indices
    int32_t* flags;         // [0] overflow, [1] eviction failures
};

static HashWs hash_ws(void* ws, int64_t table_size) {
    int64_t nb = (table_size + BUCKET - 1) / BUCKET;
    char* p = reinterpret_cast<char*>(ws);
    HashWs h;
    h.bucket_count = reinterpret_cast<int32_t*>(p);
    h.staged = reinterpret_cast<int32_t*>(p + ((nb * 4 + 255) / 256) * 256);
    h.flags = reinterpret_cast<int32_t*>(p + ((nb * 4 + 255) / 256) * 256 + nb * BUCKET * 4);
    return h;
}

__global__ void zero_i32(int32_t* p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Phase 1: distribute key indices into buckets. Lanes of a wavefront that target the same
// bucket are served by one atomic: the leader reserves popcount(match) slots and every lane
// takes its rank inside the ballot mask.
__global__ void bucket_fill_kernel(const uint64_t* __restrict__ keys, int64_t n, int nb, int32_t* bucket_count,
                                   int32_t* staged, int32_t* flags, int aggregate) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = i < n;
    int b = valid ? (int)(keys[i] % (uint64_t)nb) : -1;
    int slot = -1;
    if (aggregate) {
        const int lane = threadIdx.x & 63;
        uint64_t pending = __ballot(valid);
        while (pending) {
            int leader = __ffsll((unsigned long long)pending) - 1;
            int lb = __shfl(b, leader, 64);
            uint64_t same = __ballot(valid && b == lb) & pending;
            int base = 0;
            if (lane == leader) base = atomicAdd(&bucket_count[lb], __popcll(same));
            base = __shfl(base, leader, 64);
            if (valid && b == lb) slot = base + __popcll(same & ((1ull << lane) - 1ull));
            pending &= ~same;
        }
    } else if (valid) {
        slot = atomicAdd(&bucket_count[b], 1);
    }
    if (!valid) return;
    if (slot < BUCKET)
        staged[(int64_t)b * BUCKET + slot] = (int32_t)i;
    else
        atomicAdd(&flags[0], 1);
}

// Phase 2: one workgroup per bucket runs the three-table cuckoo insertion in LDS.
__global__ void __launch_bounds__(BUCKET) bucket_cuckoo_kernel(const uint64_t* __restrict__ keys,
                                                               const int64_t* __restrict__ vals,
                                                               const int32_t* __restrict__ bucket_count,
                                                               const int32_t* __restrict__ staged, int64_t table_size,
                                                               uint64_t m0, uint64_t m1, uint64_t m2, int evict_bound,
                                                               uint64_t* __restrict__ tkeys, int64_t* __restrict__ tvals,
                                                               int32_t* flags) {
    __shared__ int32_t tab[3][BUCKET];
    const int t = threadIdx.x;
    const int b = blockIdx.x;
    tab[0][t] = -1;
    tab[1][t] = -1;
    tab[2][t] = -1;
    __syncthreads();
    const uint64_t mult[3] = {m0, m1, m2};
    int cnt = bucket_count[b];
    if (cnt > BUCKET) cnt = BUCKET;
    if (t < cnt) {
        int cur = staged[(int64_t)b * BUCKET + t];
        int f = 0;
        int evictions = 0;
        const int limit = 3 * evict_bound;
        while (true) {
            int pos = slot_hash(keys[cur], mult[f]);
            int old = atomicExch(&tab[f][pos], cur);
            if (old < 0) break;
            cur = old;
            f = (f + 1) % 3;
            if (++evictions >= limit) {
                atomicAdd(&flags[1], 1);
                break;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        int idx = tab[f][t];
        int64_t o = (int64_t)f * table_size + (int64_t)b * BUCKET + t;
        tkeys[o] = idx >= 0 ? keys[idx] : EMPTY_KEY;
        tvals[o] = idx >= 0 ? (vals ? vals[idx] : (int64_t)idx) : -1;
    }
}

__global__ void hash_query_kernel(const uint64_t* __restrict__ tkeys, const int64_t* __restrict__ tvals,
                                  int64_t table_size, int nb, uint64_t m0, uint64_t m1, uint64_t m2,
                                  const uint64_t* __restrict__ q, int64_t n, int64_t* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t key = q[i];
    int64_t base = (int64_t)(key % (uint64_t)nb) * BUCKET;
    int64_t res = -1;
    const uint64_t mult[3] = {m0, m1, m2};
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        int64_t o = (int64_t)f * table_size + base + slot_hash(key, mult[f]);
        if (tkeys[o] == key) { res = tvals[o]; break; }
    }
    out[i] = res;
}

static uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// ========================================================================================
// Ordered bitmap sets
// ========================================================================================
__global__ void clear_words(uint32_t* w, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] = 0u;
}

__global__ void mark_coords_kernel(ngs_grid g, const int32_t* __restrict__ c, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    grid_mark(g, c[3 * i], c[3 * i + 1], c[3 * i + 2]);
}

// Byte-staged marking: marks[lattice index] = 1 by a plain store. Stores of the same byte from many
// lanes and waves merge in the caches, where a device-scope atomic OR per word serialises whenever
// many points share a coarse voxel. fold_marks_kernel then ORs each word's 32 bytes into the bitmap
// and clears them, so the byte map is all zero again after every staged call.
__device__ __forceinline__ void stage_mark(const ngs_grid& g, uint8_t* marks, int cx, int cy, int cz) {
    int64_t lin;
    if (grid_cell(g, cx, cy, cz, &lin)) marks[lin] = 1;
}

__global__ void stage_coords_kernel(ngs_grid g, const int32_t* __restrict__ c, int64_t n, uint8_t* marks) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    stage_mark(g, marks, c[3 * i], c[3 * i + 1], c[3 * i + 2]);
}

// Bit k of the result = byte k of x, for bytes that are 0 or 1.
__device__ __forceinline__ uint32_t pack4(uint32_t x) { return (x | (x >> 7) | (x >> 14) | (x >> 21)) & 0xFu; }

__global__ void fold_marks_kernel(uint32_t* __restrict__ bits, uint8_t* __restrict__ marks, int64_t nwords) {
    int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwords) return;
    uint4* p = reinterpret_cast<uint4*>(marks + 32 * w);
    const uint4 a = p[0], b = p[1];
    const uint32_t m = pack4(a.x) | (pack4(a.y) << 4) | (pack4(a.z) << 8) | (pack4(a.w) << 12) | (pack4(b.x) << 16) |
                       (pack4(b.y) << 20) | (pack4(b.z) << 24) | (pack4(b.w) << 28);
    if (m) {
        bits[w] |= m;
        p[0] = make_uint4(0u, 0u, 0u, 0u);
        p[1] = make_uint4(0u, 0u, 0u, 0u);
    }
}

// Marks c + (dx, dy, dz) * stride for every row c and every offset in [-radius, radius]^3; one thread
// per (row, dx, dy), walking dz.
__global__ void mark_dilated_kernel(ngs_grid g, const int32_t* __restrict__ c, int64_t n, int radius) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int w = 2 * radius + 1;
    if (t >= n * w * w) return;
    const int64_t i = t / (w * w);
    const int j = (int)(t - i * w * w);
    const int s = 1 << g.log2_stride;
    const int x = c[3 * i] + (j / w - radius) * s, y = c[3 * i + 1] + (j % w - radius) * s;
    for (int dz = -radius; dz <= radius; ++dz) grid_mark(g, x, y, c[3 * i + 2] + dz * s);
}

// Sample positions of get_test_grid (hashtree.py:209-229): (base + box) * voxel_size per leaf, box =
// torch.linspace(0, stride, R) per axis in ATen's form (start + step i below the halfway index,
// end - step (R - 1 - i) from it, step = (end - start) / (R - 1), each as one fused multiply-add);
// rows (leaf, i, j, k) with k fastest.
__device__ __forceinline__ float linspace_value(float end, int steps, int i) {
    const float step = end / (float)(steps - 1);
    return i < steps / 2 ? fmaf(step, (float)i, 0.0f) : fmaf(-step, (float)(steps - 1 - i), end);
}

__global__ void test_grid_kernel(const int32_t* __restrict__ base, int64_t n, int stride, int res, float vs,
                                 float* __restrict__ out) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r3 = (int64_t)res * res * res;
    if (t >= n * r3) return;
    const int64_t l = t / r3;
    const int sub = (int)(t - l * r3);
    const int idx[3] = {sub / (res * res), (sub / res) % res, sub % res};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float b = (float)base[3 * l + a];
        const float box = res == 1 ? 0.0f : linspace_value((float)stride, res, idx[a]);
        out[3 * t + a] = (b + box) * vs;
    }
}

// The 8 corners of mark_corners_kernel, byte-staged.
__global__ void stage_corners_kernel(ngs_grid g, const float* __restrict__ xyz, int64_t n, float inv_vs,
                                     const int32_t* __restrict__ levels, int max_level, uint8_t* marks) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (levels && levels[i] > max_level) return;
    const int s = 1 << g.log2_stride;
    const float inv_s = 1.0f / (float)s;
    int dc[3], sg[3];
    float loc[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) trilinear_axis(xyz[3 * i + a] * inv_vs, s, inv_s, &dc[a], &sg[a], &loc[a]);
#pragma unroll
    for (int o = 0; o < 8; ++o) {
        const int nx = (o >> 2) & 1, ny = (o >> 1) & 1, nz = o & 1;
        stage_mark(g, marks, dc[0] + nx * sg[0] * s, dc[1] + ny * sg[1] * s, dc[2] + nz * sg[2] * s);
    }
}

__global__ void mark_corners_kernel(ngs_grid g, const float* __restrict__ xyz, int64_t n, float inv_vs,
                                    const int32_t* __restrict__ levels, int max_level) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (levels && levels[i] > max_level) return;
    const int s = 1 << g.log2_stride;
    const float inv_s = 1.0f / (float)s;
    int dc[3], sg[3];
    float loc[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) trilinear_axis(xyz[3 * i + a] * inv_vs, s, inv_s, &dc[a], &sg[a], &loc[a]);
    // the two z-corners of each (x, y) usually share a bitmap word: one atomic for both
#pragma unroll
    for (int o = 0; o < 4; ++o) {
        const int nx = (o >> 1) & 1, ny = o & 1;
        const int cx = dc[0] + nx * sg[0] * s, cy = dc[1] + ny * sg[1] * s;
        int64_t l0, l1;
        const bool v0 = grid_cell(g, cx, cy, dc[2], &l0), v1 = grid_cell(g, cx, cy, dc[2] + sg[2] * s, &l1);
        if (v0 && v1 && (l0 >> 5) == (l1 >> 5)) {
            grid_mark_word(g, l0 >> 5, (1u << (l0 & 31)) | (1u << (l1 & 31)));
        } else {
            if (v0) grid_mark_word(g, l0 >> 5, 1u << (l0 & 31));
            if (v1) grid_mark_word(g, l1 >> 5, 1u << (l1 & 31));
        }
    }
}

__global__ void emit_kernel(ngs_grid g, int32_t* __restrict__ out) {
    int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= g.nwords) return;
    uint32_t bits = g.bits[w];
    if (!bits) return;
    int32_t r = g.rank[w];
    const int dy = g.dims[1], dz = g.dims[2];
    const int64_t dyz = (int64_t)dy * dz;
    // (x, y, z) of the word's bit 0 by division once per word, then advanced bit by bit (z fastest)
    const int64_t lin0 = w * 32;
    int x = (int)(lin0 / dyz);
    const int64_t rem = lin0 - (int64_t)x * dyz;
    int y = (int)(rem / dz);
    int z = (int)(rem - (int64_t)y * dz);
    int cur = 0;
    while (bits) {
        const int b = __ffs(bits) - 1;
        bits &= bits - 1;
        z += b - cur;
        cur = b;
        while (z >= dz) {
            z -= dz;
            if (++y == dy) {
                y = 0;
                ++x;
            }
        }
        out[3 * (int64_t)r] = (x + g.origin[0]) << g.log2_stride;
        out[3 * (int64_t)r + 1] = (y + g.origin[1]) << g.log2_stride;
        out[3 * (int64_t)r + 2] = (z + g.origin[2]) << g.log2_stride;
        ++r;
    }
}

__global__ void grid_query_kernel(ngs_grid g, const int32_t* __restrict__ c, int64_t n, const int32_t* __restrict__ off,
                                  int k, int64_t* __restrict__ out) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * k) return;
    int64_t kk = t / n, i = t % n;
    int x = c[3 * i], y = c[3 * i + 1], z = c[3 * i + 2];
    if (off) { x += off[3 * kk]; y += off[3 * kk + 1]; z += off[3 * kk + 2]; }
    out[t] = grid_find(g, x, y, z);
}

__global__ void grid_rank_rows_kernel(ngs_grid g, const int32_t* __restrict__ c, int64_t n, int32_t* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = grid_find(g, c[3 * i], c[3 * i + 1], c[3 * i + 2]);
}

// Cell of every point at stride s: floor(q scale / s) s per axis, q scale in FP32 (the reference
// divides positions by the voxel size, hashtree.py build paths; _hash_backend.py:363,396). The
// division by the power-of-two stride is exact, so floor(v / s) = floor(v * (1 / s)).
__global__ void point_cells_kernel(const float* __restrict__ q, int64_t n3, float scale, int s, float inv_s,
                                   int32_t* __restrict__ out) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n3) return;
    const float v = q[t] * scale;
    out[t] = (int)floorf(v * inv_s) * s;
}

// Status of coordinates in a level (_hash_backend.py:292-318): 0 absent, 1 present without
// children, 2 present with children. A voxel of the level has children in the next finer level
// exactly when its coordinate is the parent (at the level's stride) of some finer voxel, which is
// what `finer_parents` holds.
__global__ void voxel_status_kernel(ngs_grid g, ngs_grid finer_parents, int has_finer, const int32_t* __restrict__ c,
                                    int64_t n, int64_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int x = c[3 * i], y = c[3 * i + 1], z = c[3 * i + 2];
    int64_t st = 0;
    if (grid_test(g, x, y, z)) st = (has_finer && grid_test(finer_parents, x, y, z)) ? 2 : 1;
    out[i] = st;
}

// Trilinear interpolation of per-voxel rows at the points (_hash_backend.py:320-330): for each
// point the 8 corners of its trilinear stencil at the grid's stride (the splat's corners and
// weights), absent corners skipped; products data * weight in FP32, summed in corner order in FP64.
__global__ void split_data_kernel(ngs_grid g, const float* __restrict__ xyz, int64_t n, float inv_vs,
                                  const float* __restrict__ data, int nc, float* __restrict__ out) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = 1 << g.log2_stride;
    const float inv_s = 1.0f / (float)s;
    int dc[3], sg[3];
    float loc[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) trilinear_axis(xyz[3 * i + a] * inv_vs, s, inv_s, &dc[a], &sg[a], &loc[a]);
    int id[8];
    float w[8];
#pragma unroll
    for (int o = 0; o < 8; ++o) {
        const int nx = (o >> 2) & 1, ny = (o >> 1) & 1, nz = o & 1;
        id[o] = grid_find(g, dc[0] + nx * sg[0] * s, dc[1] + ny * sg[1] * s, dc[2] + nz * sg[2] * s);
        const float ax = nx ? loc[0] : 1.0f - loc[0];
        const float ay = ny ? loc[1] : 1.0f - loc[1];
        const float az = nz ? loc[2] : 1.0f - loc[2];
        w[o] = ax * ay * az;
    }
    for (int c = 0; c < nc; ++c) {
        double acc = 0.0;
#pragma unroll
        for (int o = 0; o < 8; ++o)
            if (id[o] >= 0) acc += (double)(data[(int64_t)id[o] * nc + c] * w[o]);
        out[i * nc + c] = (float)acc;
    }
}

// Float ordering as signed ints for atomic min/max.
__device__ __forceinline__ int f2key(float f) {
    int b = __float_as_int(f);
    return b >= 0 ? b : (b ^ 0x7FFFFFFF);
}
__device__ __forceinline__ float key2f(int k) { return __int_as_float(k >= 0 ? k : (k ^ 0x7FFFFFFF)); }

__global__ void bounds_init(int* mm) {
    if (threadIdx.x < 3) { mm[threadIdx.x] = 0x7FFFFFFF; mm[3 + threadIdx.x] = (int)0x80000000; }
}

__global__ void __launch_bounds__(256) bounds_kernel(const float* __restrict__ xyz, int64_t n, float inv_vs, int* mm) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float q = xyz[3 * i + a] * inv_vs;
            lo[a] = fminf(lo[a], q);
            hi[a] = fmaxf(hi[a], q);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        for (int d = 32; d > 0; d >>= 1) {
            lo[a] = fminf(lo[a], __shfl_xor(lo[a], d, 64));
            hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], d, 64));
        }
    }
    // the block's 4 waves through LDS, then 6 atomics per block: per-wave atomics on the same 6 words
    // serialised (4096 waves cost ~0.25 ms)
    __shared__ float slo[4][3], shi[4][3];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            slo[wv][a] = lo[a];
            shi[wv][a] = hi[a];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float l = slo[0][a], h = shi[0][a];
            for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
                l = fminf(l, slo[k][a]);
                h = fmaxf(h, shi[k][a]);
            }
            atomicMin(&mm[a], f2key(l));
            atomicMax(&mm[3 + a], f2key(h));
        }
    }
}

__global__ void bounds_decode(int* mm) {
    if (threadIdx.x < 6) {
        float f = key2f(mm[threadIdx.x]);
        mm[threadIdx.x] = __float_as_int(f);
    }
}

// ========================================================================================
// Splat, field evaluation, levels, point basis
// ========================================================================================
__global__ void splat_kernel(ngs_grid g, const float* __restrict__ xyz, int64_t n, float inv_vs,
                             const float* __restrict__ data, int ncomp, const float* __restrict__ scale,
                             const int32_t* __restrict__ levels, int level, double post, double* __restrict__ acc) {
#pragma clang fp contract(off)
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (levels && levels[i] != level) return;
    const int s = 1 << g.log2_stride;
    const float inv_s = 1.0f / (float)s;
    int dc[3], sg[3];
    float loc[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) trilinear_axis(xyz[3 * i + a] * inv_vs, s, inv_s, &dc[a], &sg[a], &loc[a]);
    float row[3] = {1.f, 1.f, 1.f};
    if (data) {
        float sc = scale ? scale[i] : 1.0f;
        for (int c = 0; c < ncomp; ++c) row[c] = data[i * ncomp + c] * sc;
    }
#pragma unroll
    for (int o = 0; o < 8; ++o) {
        int nx = (o >> 2) & 1, ny = (o >> 1) & 1, nz = o & 1;
        int id = grid_find(g, dc[0] + nx * sg[0] * s, dc[1] + ny * sg[1] * s, dc[2] + nz * sg[2] * s);
        if (id < 0) continue;
        float ax = nx ? loc[0] : 1.0f - loc[0];
        float ay = ny ? loc[1] : 1.0f - loc[1];
        float az = nz ? loc[2] : 1.0f - loc[2];
        float w = ax * ay * az;
        if (data) {
            for (int c = 0; c < ncomp; ++c) atomicAdd(&acc[(int64_t)id * ncomp + c], (double)(w * row[c]) * post);
        } else {
            atomicAdd(&acc[id], (double)w * post);
        }
    }
}

// Gather form of the splat. Each point is binned by the rank of the lower node of its trilinear
// stencil (min(dc, dc + sg s) per axis); one thread per node then sums, in FP64 registers, the points
// of the 8 bins whose lower node is node - o s, o in {0, 1}^3, with the weights splat_kernel gives
// them (same float operations). Points outside `level` get key `count` (sorted last, never read);
// a point whose lower node is not in g is splatted with atomics here (the gather cannot reach it).
__global__ void splat_key_kernel(ngs_grid g, const float* __restrict__ xyz, int64_t n, float inv_vs,
                                 const float* __restrict__ data, int ncomp, const float* __restrict__ scale,
                                 const int32_t* __restrict__ levels, int level, double post, double* __restrict__ acc,
                                 uint32_t count, uint32_t* __restrict__ keys, int32_t* __restrict__ idx) {
#pragma clang fp contract(off)
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    idx[i] = (int32_t)i;
    keys[i] = count;
    if (levels && levels[i] != level) return;
    const int s = 1 << g.log2_stride;
    const float inv_s = 1.0f / (float)s;
    int dc[3], sg[3];
    float loc[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) trilinear_axis(xyz[3 * i + a] * inv_vs, s, inv_s, &dc[a], &sg[a], &loc[a]);
    const int32_t r = grid_find(g, sg[0] < 0 ? dc[0] - s : dc[0], sg[1] < 0 ? dc[1] - s : dc[1],
                                sg[2] < 0 ? dc[2] - s : dc[2]);
    if (r >= 0) {
        keys[i] = (uint32_t)r;
        return;
    }
    float row[3] = {1.f, 1.f, 1.f};
    if (data) {
        float sc = scale ? scale[i] : 1.0f;
        for (int c = 0; c < ncomp; ++c) row[c] = data[i * ncomp + c] * sc;
    }
    for (int o = 0; o < 8; ++o) {
        int nx = (o >> 2) & 1, ny = (o >> 1) & 1, nz = o & 1;
        int id = grid_find(g, dc[0] + nx * sg[0] * s, dc[1] + ny * sg[1] * s, dc[2] + nz * sg[2] * s);
        if (id < 0) continue;
        float ax = nx ? loc[0] : 1.0f - loc[0];
        float ay = ny ? loc[1] : 1.0f - loc[1];
        float az = nz ? loc[2] : 1.0f - loc[2];
        float w = ax * ay * az;
        if (data) {
            for (int c = 0; c < ncomp; ++c) atomicAdd(&acc[(int64_t)id * ncomp + c], (double)(w * row[c]) * post);
        } else {
            atomicAdd(&acc[id], (double)w * post);
        }
    }
}

// Bin bounds over the sorted keys: [bstart[k], bend[k]) holds the points of key k (both zeroed first).
// The binned points' positions and scaled data rows are copied in bin order, so the gather below reads
// each bin as one contiguous run.
__global__ void splat_bins_kernel(const uint32_t* __restrict__ ks, const int32_t* __restrict__ order, int64_t n,
                                  uint32_t count, const float* __restrict__ xyz, const float* __restrict__ data,
                                  int ncomp, const float* __restrict__ scale, int32_t* __restrict__ bstart,
                                  int32_t* __restrict__ bend, float* __restrict__ bxyz, float* __restrict__ brow) {
#pragma clang fp contract(off)
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = ks[i];
    if (k >= count) return;
    if (i == 0 || ks[i - 1] != k) bstart[k] = (int32_t)i;
    if (i == n - 1 || ks[i + 1] != k) bend[k] = (int32_t)(i + 1);
    const int64_t p = order[i];
#pragma unroll
    for (int a = 0; a < 3; ++a) bxyz[3 * i + a] = xyz[3 * p + a];
    if (data) {
        const float sc = scale ? scale[p] : 1.0f;
        for (int c = 0; c < ncomp; ++c) brow[3 * i + c] = data[p * ncomp + c] * sc;
    }
}

__global__ void splat_gather_kernel(ngs_grid g, const int32_t* __restrict__ nodes, int64_t count, float inv_vs,
                                    int ncomp, const int32_t* __restrict__ bstart, const int32_t* __restrict__ bend,
                                    const float* __restrict__ bxyz, const float* __restrict__ brow, double post,
                                    double* __restrict__ acc) {
#pragma clang fp contract(off)
    const int64_t node = xcd_contiguous_block() * blockDim.x + threadIdx.x;
    if (node >= count) return;
    const int s = 1 << g.log2_stride;
    const float inv_s = 1.0f / (float)s;
    const int N[3] = {nodes[3 * node], nodes[3 * node + 1], nodes[3 * node + 2]};
    const int lx = (N[0] >> g.log2_stride) - g.origin[0], ly = (N[1] >> g.log2_stride) - g.origin[1],
              lz = (N[2] >> g.log2_stride) - g.origin[2];
    double sum[3] = {0.0, 0.0, 0.0};
    // the 8 bins with lower node N - (ox, oy, oz) s: one column query per (ox, oy) covers oz = 1, 0
    for (int oxy = 0; oxy < 4; ++oxy) {
        int32_t r0;
        const uint32_t m = grid_column(g, lx - (oxy >> 1), ly - (oxy & 1), lz - 1, 2, &r0);
        for (int k = 0; k < 2; ++k) {
            if (!((m >> k) & 1u)) continue;
            const int32_t r = r0 + __popc(m & ((1u << k) - 1u));
            const int32_t e = bend[r];
            for (int32_t j = bstart[r]; j < e; ++j) {
                int dc[3], sg[3];
                float loc[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) trilinear_axis(bxyz[3 * j + a] * inv_vs, s, inv_s, &dc[a], &sg[a], &loc[a]);
                // which stencil node of the point this node is, per axis (none: the bin only neighbours it)
                int nn[3];
                bool hit = true;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    if (N[a] == dc[a]) nn[a] = 0;
                    else if (sg[a] != 0 && N[a] == dc[a] + sg[a] * s) nn[a] = 1;
                    else hit = false;
                }
                if (!hit) continue;
                float ax = nn[0] ? loc[0] : 1.0f - loc[0];
                float ay = nn[1] ? loc[1] : 1.0f - loc[1];
                float az = nn[2] ? loc[2] : 1.0f - loc[2];
                float w = ax * ay * az;
                if (ncomp) {
                    for (int c = 0; c < ncomp; ++c) sum[c] += (double)(w * brow[3 * j + c]) * post;
                } else {
                    sum[0] += (double)w * post;
                }
            }
        }
    }
    if (ncomp) {
        for (int c = 0; c < ncomp; ++c) acc[node * ncomp + c] += sum[c];
    } else {
        acc[node] += sum[0];
    }
}

__global__ void eval_field_kernel(ngs_grid g, const float* __restrict__ field, const float* __restrict__ xyz, int64_t n,
                                  float inv_vs, float scale, int accumulate, float* __restrict__ out,
                                  float* __restrict__ grad) {
#pragma clang fp contract(off)
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = 1 << g.log2_stride;
    const float inv_s = 1.0f / (float)s;
    int cell[3];
    float b[3][3], db[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) axis_basis_grad(xyz[3 * i + a] * inv_vs, s, inv_s, &cell[a], b[a], db[a]);
    float v = 0.f, gx = 0.f, gy = 0.f, gz = 0.f;
    const int l2 = g.log2_stride;
    const int lx = (cell[0] >> l2) - g.origin[0], ly = (cell[1] >> l2) - g.origin[1], lz = (cell[2] >> l2) - g.origin[2];
    // the 27 voxels around the cell in (kx, ky, kz) order, one bitmap column per (kx, ky)
    for (int kx = 0; kx < 3; ++kx)
        for (int ky = 0; ky < 3; ++ky) {
            int32_t id;
            const uint32_t m = grid_column(g, lx + kx - 1, ly + ky - 1, lz - 1, 3, &id);
            if (!m) continue;
            for (int kz = 0; kz < 3; ++kz) {
                if (!((m >> kz) & 1u)) continue;
                float f = field[id++];
                v += (b[0][kx] * b[1][ky] * b[2][kz]) * f;
                if (grad) {
                    gx += (db[0][kx] * b[1][ky] * b[2][kz]) * f;
                    gy += (b[0][kx] * db[1][ky] * b[2][kz]) * f;
                    gz += (b[0][kx] * b[1][ky] * db[2][kz]) * f;
                }
            }
        }
    out[i] = (accumulate ? out[i] : 0.f) + v * scale;
    if (grad) {
        float* gp = grad + 3 * i;
        gp[0] = (accumulate ? gp[0] : 0.f) + gx * scale;
        gp[1] = (accumulate ? gp[1] : 0.f) + gy * scale;
        gp[2] = (accumulate ? gp[2] : 0.f) + gz * scale;
    }
}

// eval_field_kernel (accumulate, no gradient) for points binned at the field's level: each point's cell
// and basis values come from the bins (point_basis_kernel's, the values axis_basis gives here too), the
// 27 terms are summed in the same order, and slot k adds to out[perm[k]] (original point order), so
// the result is bit-identical while the bitmap and coefficient reads follow the bins.
__global__ void eval_field_binned_kernel(ngs_grid g, const float* __restrict__ field, const int32_t* __restrict__ cell,
                                         const float* __restrict__ b, const int32_t* __restrict__ perm, int64_t n,
                                         float scale, float* __restrict__ out) {
#pragma clang fp contract(off)
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float* bk = b + 9 * k;
    const int l2 = g.log2_stride;
    const int lx = (cell[3 * k] >> l2) - g.origin[0], ly = (cell[3 * k + 1] >> l2) - g.origin[1],
              lz = (cell[3 * k + 2] >> l2) - g.origin[2];
    float v = 0.f;
    for (int kx = 0; kx < 3; ++kx)
        for (int ky = 0; ky < 3; ++ky) {
            int32_t id;
            const uint32_t m = grid_column(g, lx + kx - 1, ly + ky - 1, lz - 1, 3, &id);
            if (!m) continue;
            for (int kz = 0; kz < 3; ++kz) {
                if (!((m >> kz) & 1u)) continue;
                const float f = field[id++];
                v += (bk[kx] * bk[3 + ky] * bk[6 + kz]) * f;
            }
        }
    const int64_t p = perm[k];
    out[p] = out[p] + v * scale;
}

// Whether any voxel of g lies among the 27 around the cell of each query (the support test of
// eval_field_kernel): mask[i] |= 1.
__global__ void field_support_kernel(ngs_grid g, const float* __restrict__ xyz, int64_t n, float inv_vs,
                                     uint8_t* __restrict__ mask) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = 1 << g.log2_stride;
    const float inv_s = 1.0f / (float)s;
    int cell[3];
    float b[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) axis_basis(xyz[3 * i + a] * inv_vs, s, inv_s, &cell[a], b);
    const int l2 = g.log2_stride;
    const int lx = (cell[0] >> l2) - g.origin[0], ly = (cell[1] >> l2) - g.origin[1], lz = (cell[2] >> l2) - g.origin[2];
    uint32_t any = 0;
    for (int kx = 0; kx < 3; ++kx)
        for (int ky = 0; ky < 3; ++ky) {
            int32_t r;
            any |= grid_column(g, lx + kx - 1, ly + ky - 1, lz - 1, 3, &r);
        }
    if (any) mask[i] = 1;
}

__global__ void levels_kernel(const float* __restrict__ rho, int64_t n, float inv_min, float inv_ln, int depth,
                              int32_t* __restrict__ lv) {
#pragma clang fp contract(off)
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float t = floorf(logf(rho[i] * inv_min) * inv_ln);
    int k = (int)t;
    if (k > 0) k = 0;
    k = -k;
    if (k > depth - 1) k = depth - 1;
    lv[i] = k;
}

__global__ void point_basis_kernel(const float* __restrict__ xyz, int64_t n, float inv_vs, int s,
                                   int32_t* __restrict__ cell, float* __restrict__ b) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float inv_s = 1.0f / (float)s;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int c;
        float bb[3];
        axis_basis(xyz[3 * i + a] * inv_vs, s, inv_s, &c, bb);
        cell[3 * i + a] = c;
        b[9 * i + 3 * a] = bb[0];
        b[9 * i + 3 * a + 1] = bb[1];
        b[9 * i + 3 * a + 2] = bb[2];
    }
}

// cell_start[k], k in [0, ncells], = the first sorted position whose key is >= k (keys at or above
// ncells, the points outside every cell, count as ncells): each position writes the starts of the
// cells between its predecessor's key and its own, so every entry is written once.
__global__ void cell_starts_kernel(const uint32_t* __restrict__ ks, int64_t n, int64_t ncells,
                                   int32_t* __restrict__ cell_start) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    const int64_t k = i < n ? (int64_t)min((uint64_t)ks[i], (uint64_t)ncells) : ncells + 1;
    const int64_t prev = i > 0 ? (int64_t)min((uint64_t)ks[i - 1], (uint64_t)ncells) : -1;
    for (int64_t c = prev + 1; c <= min(k, ncells); ++c) cell_start[c] = (int32_t)i;
}

__global__ void iota_i32(int32_t* p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int32_t)i;
}

__global__ void f64_to_f32_kernel(const double* __restrict__ a, float* __restrict__ b, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

__global__ void __launch_bounds__(256) weighted_sum_kernel(const float* __restrict__ v, const float* __restrict__ w,
                                                           int64_t n, double* sums) {
    double a = 0.0, c = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double wi = w ? (double)w[i] : 1.0;
        a += wi * (double)v[i];
        c += wi;
    }
    for (int d = 32; d > 0; d >>= 1) {
        a += __shfl_xor(a, d, 64);
        c += __shfl_xor(c, d, 64);
    }
    // the block's 4 waves through LDS, then 2 atomics per block (per-wave atomics on the same two
    // words serialise)
    __shared__ double sa[4], sc[4];
    if ((threadIdx.x & 63) == 0) {
        sa[threadIdx.x >> 6] = a;
        sc[threadIdx.x >> 6] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ta = sa[0], tc = sc[0];
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
            ta += sa[k];
            tc += sc[k];
        }
        atomicAdd(&sums[0], ta);
        atomicAdd(&sums[1], tc);
    }
}

__global__ void zero_f64(double* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0.0;
}

// rocPRIM stable radix sort wrapper (sort.hip).
int sort_pairs_u32(const uint32_t* kin, uint32_t* kout, const int32_t* vin, int32_t* vout, int64_t n, int bits,
                   void* tmp, size_t* tmp_bytes, hipStream_t st);

}  // namespace ngs

using namespace ngs;

// ========================================================================================
// C ABI
// ========================================================================================
extern "C" {

const char* ngs_last_error(void) { return g_last_error.c_str(); }
int ngs_abi_version(void) { return 1; }

int ngs_hash_coords(const int32_t* coords, int64_t n, int32_t dim, uint64_t* keys, ngs_stream_t s) {
    NGS_REQUIRE(dim >= 1 && dim <= 4, "ngs_hash_coords: dim must be 1..4");
    NGS_REQUIRE(n >= 0, "ngs_hash_coords: n < 0");
    if (n == 0) return NGS_OK;
    hash_rows_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(coords, n, dim, keys);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_kernel_hash(const int32_t* coords, int64_t n, int32_t dim, const int32_t* offsets, int32_t k, uint64_t* out,
                    ngs_stream_t s) {
    NGS_REQUIRE(dim == 3 || dim == 4, "ngs_kernel_hash: dim must be 3 or 4");
    NGS_REQUIRE(n >= 0 && k >= 0, "ngs_kernel_hash: bad sizes");
    if (n == 0 || k == 0) return NGS_OK;
    kernel_hash_kernel<<<grid1d(n * k, 256), 256, 0, S(s)>>>(coords, n, dim, offsets, k, out);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_hash_sizing(int64_t n, int32_t enlarge, int64_t* table_size, int32_t* num_buckets, int32_t* evict_bound) {
    NGS_REQUIRE(n > 0, "ngs_hash_sizing: n must be positive");
    NGS_REQUIRE(n <= (int64_t)1 << 30, "ngs_hash_sizing: n too large");
    double lg = std::ceil(std::log2((double)n));
    int64_t next_pow2 = (int64_t)std::pow(2.0, lg);
    int64_t ts = (n < 2048) ? 4 * next_pow2 : 2 * next_pow2;
    if (enlarge) ts = 4 * next_pow2;
    if (ts < 512) ts = 512;
    *table_size = ts;
    *num_buckets = (int32_t)((ts + BUCKET - 1) / BUCKET);
    *evict_bound = (int32_t)(8 * lg);
    return NGS_OK;
}

int64_t ngs_hash_workspace_bytes(int64_t table_size) {
    int64_t nb = (table_size + BUCKET - 1) / BUCKET;
    return ((nb * 4 + 255) / 256) * 256 + nb * BUCKET * 4 + 256;
}

int ngs_hash_build(ngs_hash_table* t, const uint64_t* keys, const int64_t* vals, int64_t n, uint64_t seed,
                   void* workspace, ngs_stream_t s) {
    NGS_REQUIRE(t && t->keys && t->vals && workspace, "ngs_hash_build: null table or workspace");
    NGS_REQUIRE(n > 0, "ngs_hash_build: empty key set");
    NGS_REQUIRE(t->num_buckets == (t->table_size + BUCKET - 1) / BUCKET, "ngs_hash_build: inconsistent sizing");
    hipStream_t st = S(s);
    HashWs ws = hash_ws(workspace, t->table_size);
    const int nb = t->num_buckets;
    zero_i32<<<grid1d(nb, 256), 256, 0, st>>>(ws.bucket_count, nb);
    zero_i32<<<1, 64, 0, st>>>(ws.flags, 2);
    NGS_CHECK_LAUNCH();
    bucket_fill_kernel<<<grid1d(n, 256), 256, 0, st>>>(keys, n, nb, ws.bucket_count, ws.staged, ws.flags, nb <= 64 ? 1 : 0);
    NGS_CHECK_LAUNCH();
    int32_t flags[2];
    NGS_CHECK_HIP(hipMemcpyAsync(flags, ws.flags, sizeof(flags), hipMemcpyDeviceToHost, st));
    NGS_CHECK_HIP(hipStreamSynchronize(st));
    if (flags[0] != 0) {
        set_error("ngs_hash_build: bucket overflow (" + std::to_string(flags[0]) + " keys beyond bucket capacity)");
        return NGS_ERR_HASH_OVERFLOW;
    }
    uint64_t state = seed;
    for (int attempt = 0; attempt < 100; ++attempt) {
        for (int f = 0; f < 3; ++f) t->mult[f] = splitmix64(state) | 1ull;
        zero_i32<<<1, 64, 0, st>>>(ws.flags + 1, 1);
        bucket_cuckoo_kernel<<<nb, BUCKET, 0, st>>>(keys, vals, ws.bucket_count, ws.staged, t->table_size, t->mult[0],
                                                    t->mult[1], t->mult[2], t->evict_bound, t->keys, t->vals, ws.flags);
        NGS_CHECK_LAUNCH();
        int32_t fail = 0;
        NGS_CHECK_HIP(hipMemcpyAsync(&fail, ws.flags + 1, sizeof(fail), hipMemcpyDeviceToHost, st));
        NGS_CHECK_HIP(hipStreamSynchronize(st));
        t->attempts = attempt + 1;
        if (fail == 0) {
            t->inserted = n;
            return NGS_OK;
        }
    }
    set_error("ngs_hash_build: cuckoo insertion failed for 100 hash families");
    return NGS_ERR_REHASH;
}

int ngs_hash_query(const ngs_hash_table* t, const uint64_t* queries, int64_t n, int64_t* out, ngs_stream_t s) {
    NGS_REQUIRE(t && t->keys, "ngs_hash_query: null table");
    if (n <= 0) return NGS_OK;
    hash_query_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(t->keys, t->vals, t->table_size, t->num_buckets, t->mult[0],
                                                        t->mult[1], t->mult[2], queries, n, out);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

// Bitmap boxes are limited to 2^36 lattice points (8 GiB of bits).
int64_t ngs_grid_words(const int32_t dims[3]) {
    if (dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) return 0;
    double v = (double)dims[0] * dims[1] * dims[2];
    if (v > 68719476736.0) return 0;
    int64_t bits = (int64_t)dims[0] * dims[1] * dims[2];
    return (bits + 31) / 32;
}

int64_t ngs_scan_workspace_bytes(int64_t n) {
    // the size rocPRIM asks for on the current device (-1 without one): the larger of the two scan forms
    size_t a = 0, b = 0;
    n = n < 0 ? 0 : n;
    if (scan_counts_device(nullptr, n, nullptr, nullptr, &a, nullptr) != NGS_OK) return -1;
    if (scan_popc_device(nullptr, n, nullptr, nullptr, &b, nullptr) != NGS_OK) return -1;
    size_t m = a > b ? a : b;
    return m <= 256 ? 256 : (int64_t)((m + 255) / 256 * 256);  // never 0: callers pass a non-null buffer
}

int ngs_grid_clear(ngs_grid g, ngs_stream_t s) {
    NGS_REQUIRE(g.bits, "ngs_grid_clear: null bits");
    if (g.nwords == 0) return NGS_OK;
    clear_words<<<grid1d(g.nwords, 256), 256, 0, S(s)>>>(g.bits, g.nwords);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_grid_mark_coords(ngs_grid g, const int32_t* coords, int64_t n, ngs_stream_t s) {
    if (n <= 0) return NGS_OK;
    mark_coords_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(g, coords, n);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

static int fold_marks(const ngs_grid& g, uint8_t* marks, hipStream_t st) {
    fold_marks_kernel<<<grid1d(g.nwords, 256), 256, 0, st>>>(g.bits, marks, g.nwords);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_grid_mark_coords_staged(ngs_grid g, const int32_t* coords, int64_t n, uint8_t* marks, ngs_stream_t s) {
    if (n <= 0 || g.nwords <= 0) return NGS_OK;
    NGS_REQUIRE(coords && marks && g.bits, "ngs_grid_mark_coords_staged: null argument");
    NGS_REQUIRE((reinterpret_cast<uintptr_t>(marks) & 15) == 0, "ngs_grid_mark_coords_staged: marks not 16-byte aligned");
    stage_coords_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(g, coords, n, marks);
    NGS_CHECK_LAUNCH();
    return fold_marks(g, marks, S(s));
}

int ngs_grid_mark_dilated(ngs_grid g, const int32_t* coords, int64_t n, int32_t radius, ngs_stream_t s) {
    NGS_REQUIRE(radius >= 0 && radius <= 64, "ngs_grid_mark_dilated: radius must be in [0, 64]");
    const int64_t t = n * (2 * radius + 1) * (2 * radius + 1);
    if (t <= 0) return NGS_OK;
    NGS_REQUIRE(coords, "ngs_grid_mark_dilated: null coordinates");
    mark_dilated_kernel<<<grid1d(t, 256), 256, 0, S(s)>>>(g, coords, n, radius);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_test_grid_positions(const int32_t* base, int64_t n, int32_t stride, int32_t resolution, float voxel_size,
                            float* out, ngs_stream_t s) {
    NGS_REQUIRE(resolution >= 1 && resolution <= 1024 && stride >= 1, "ngs_test_grid_positions: bad resolution");
    const int64_t t = n * (int64_t)resolution * resolution * resolution;
    if (t <= 0) return NGS_OK;
    NGS_REQUIRE(base && out, "ngs_test_grid_positions: null argument");
    test_grid_kernel<<<grid1d(t, 256), 256, 0, S(s)>>>(base, n, stride, resolution, voxel_size, out);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_grid_mark_corners(ngs_grid g, const float* xyz, int64_t n, float inv_voxel_size, const int32_t* levels,
                          int32_t max_level, ngs_stream_t s) {
    if (n <= 0) return NGS_OK;
    mark_corners_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(g, xyz, n, inv_voxel_size, levels, max_level);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_grid_mark_corners_staged(ngs_grid g, const float* xyz, int64_t n, float inv_voxel_size, const int32_t* levels,
                                 int32_t max_level, uint8_t* marks, ngs_stream_t s) {
    if (n <= 0 || g.nwords <= 0) return NGS_OK;
    NGS_REQUIRE(xyz && marks && g.bits, "ngs_grid_mark_corners_staged: null argument");
    NGS_REQUIRE((reinterpret_cast<uintptr_t>(marks) & 15) == 0, "ngs_grid_mark_corners_staged: marks not 16-byte aligned");
    stage_corners_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(g, xyz, n, inv_voxel_size, levels, max_level, marks);
    NGS_CHECK_LAUNCH();
    return fold_marks(g, marks, S(s));
}

int ngs_grid_finalize(ngs_grid g, int64_t* count, void* workspace, ngs_stream_t s) {
    NGS_REQUIRE(g.bits && g.rank && workspace, "ngs_grid_finalize: null pointer");
    hipStream_t st = S(s);
    size_t need = 0;
    int r = scan_popc_device(nullptr, g.nwords, nullptr, nullptr, &need, st);
    if (r != NGS_OK) return r;
    r = scan_popc_device(g.bits, g.nwords, g.rank, workspace, &need, st);
    if (r != NGS_OK) return r;
    if (count) {
        popc_total_kernel<<<1, 1, 0, st>>>(g.bits, g.rank, g.nwords, count);
        NGS_CHECK_LAUNCH();
    }
    return NGS_OK;
}

int ngs_grid_emit(ngs_grid g, int32_t* coords_out, ngs_stream_t s) {
    if (g.nwords == 0) return NGS_OK;
    emit_kernel<<<grid1d(g.nwords, 256), 256, 0, S(s)>>>(g, coords_out);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_grid_query(ngs_grid g, const int32_t* coords, int64_t n, const int32_t* offsets, int32_t k, int64_t* out,
                   ngs_stream_t s) {
    if (!offsets) k = 1;
    if (n <= 0 || k <= 0) return NGS_OK;
    grid_query_kernel<<<grid1d(n * k, 256), 256, 0, S(s)>>>(g, coords, n, offsets, k, out);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_grid_rank_rows(ngs_grid g, const int32_t* coords, int64_t n, int32_t* out, ngs_stream_t s) {
    if (n <= 0) return NGS_OK;
    grid_rank_rows_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(g, coords, n, out);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_point_cells(const float* q, int64_t n, float scale, int32_t stride, int32_t* cells, ngs_stream_t s) {
    NGS_REQUIRE(stride >= 1 && (stride & (stride - 1)) == 0, "ngs_point_cells: stride must be a power of two");
    if (n <= 0) return NGS_OK;
    NGS_REQUIRE(q && cells, "ngs_point_cells: null argument");
    point_cells_kernel<<<grid1d(3 * n, 256), 256, 0, S(s)>>>(q, 3 * n, scale, stride, 1.0f / (float)stride, cells);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_voxel_status(ngs_grid g, ngs_grid finer_parents, int32_t has_finer, const int32_t* coords, int64_t n,
                     int64_t* status, ngs_stream_t s) {
    if (n <= 0) return NGS_OK;
    NGS_REQUIRE(coords && status && g.bits && g.rank, "ngs_voxel_status: null argument");
    NGS_REQUIRE(!has_finer || (finer_parents.bits && finer_parents.log2_stride == g.log2_stride),
                "ngs_voxel_status: the finer level's parents must be a grid at the level's stride");
    voxel_status_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(g, finer_parents, has_finer, coords, n, status);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_split_data(ngs_grid g, const float* xyz, int64_t n, float inv_voxel_size, const float* data, int32_t nc,
                   float* out, ngs_stream_t s) {
    NGS_REQUIRE(nc >= 1, "ngs_split_data: at least one component");
    if (n <= 0) return NGS_OK;
    NGS_REQUIRE(xyz && data && out && g.bits && g.rank, "ngs_split_data: null argument");
    split_data_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(g, xyz, n, inv_voxel_size, data, nc, out);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_point_bounds(const float* xyz, int64_t n, float inv_voxel_size, float* minmax, ngs_stream_t s) {
    NGS_REQUIRE(n > 0, "ngs_point_bounds: no points");
    hipStream_t st = S(s);
    int* mm = reinterpret_cast<int*>(minmax);
    bounds_init<<<1, 64, 0, st>>>(mm);
    unsigned blocks = grid1d(n, 256);
    if (blocks > 256) blocks = 256;
    bounds_kernel<<<blocks, 256, 0, st>>>(xyz, n, inv_voxel_size, mm);
    bounds_decode<<<1, 64, 0, st>>>(mm);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_scan_i32(const int32_t* counts, int64_t n, int32_t* offsets, void* workspace, ngs_stream_t s) {
    NGS_REQUIRE(workspace && offsets, "ngs_scan_i32: null pointer");
    return scan_i32_internal(counts, n, offsets, workspace, S(s));
}

int ngs_splat(ngs_grid g, const float* xyz, int64_t n, float inv_voxel_size, const float* data, int32_t ncomp,
              const float* scale, const int32_t* levels, int32_t level, double post, double* acc, ngs_stream_t s) {
    NGS_REQUIRE(!data || (ncomp >= 1 && ncomp <= 3), "ngs_splat: ncomp must be 1..3");
    if (n <= 0) return NGS_OK;
    splat_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(g, xyz, n, inv_voxel_size, data, ncomp, scale, levels, level, post, acc);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

static int splat_gather_layout(int64_t n, int64_t count, size_t* tmp, size_t off[8], size_t* total) {
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    const int bits = 32 - __builtin_clz((uint32_t)count);  // the keys run up to count
    size_t t = 0;
    int r = sort_pairs_u32(nullptr, nullptr, nullptr, nullptr, n, bits, nullptr, &t, nullptr);
    if (r != NGS_OK) return r;
    *tmp = t;
    off[0] = al(t);                               // keys
    off[1] = off[0] + al(4 * (size_t)n);          // sorted keys
    off[2] = off[1] + al(4 * (size_t)n);          // point ids
    off[3] = off[2] + al(4 * (size_t)n);          // point ids in key order
    off[4] = off[3] + al(4 * (size_t)n);          // bin starts (count)
    off[5] = off[4] + al(4 * (size_t)count);      // bin ends (count)
    off[6] = off[5] + al(4 * (size_t)count);      // positions in bin order (n x 3)
    off[7] = off[6] + al(12 * (size_t)n);         // scaled data rows in bin order (n x 3)
    *total = off[7] + al(12 * (size_t)n);
    return NGS_OK;
}

int64_t ngs_splat_gather_workspace_bytes(int64_t n, int64_t count) {
    if (n <= 0 || count <= 0) return 256;
    if (count >= ((int64_t)1 << 31) || n >= ((int64_t)1 << 31)) return -1;
    size_t tmp, off[8], total;
    if (splat_gather_layout(n, count, &tmp, off, &total) != NGS_OK) return -1;
    return (int64_t)total;
}

int ngs_splat_gather(ngs_grid g, const int32_t* node_coords, int64_t count, const float* xyz, int64_t n,
                     float inv_voxel_size, const float* data, int32_t ncomp, const float* scale, const int32_t* levels,
                     int32_t level, double post, double* acc, void* workspace, int64_t workspace_bytes, ngs_stream_t s) {
    NGS_REQUIRE(!data || (ncomp >= 1 && ncomp <= 3), "ngs_splat_gather: ncomp must be 1..3");
    if (n <= 0 || count <= 0) return NGS_OK;
    NGS_REQUIRE(count < ((int64_t)1 << 31) && n < ((int64_t)1 << 31), "ngs_splat_gather: more than 2^31 points or nodes");
    NGS_REQUIRE(node_coords && xyz && acc && workspace && g.bits && g.rank, "ngs_splat_gather: null argument");
    size_t tmp, off[8], total;
    int r = splat_gather_layout(n, count, &tmp, off, &total);
    if (r != NGS_OK) return r;
    NGS_REQUIRE((size_t)workspace_bytes >= total, "ngs_splat_gather: workspace too small");
    char* w = reinterpret_cast<char*>(workspace);
    hipStream_t st = S(s);
    uint32_t* keys = reinterpret_cast<uint32_t*>(w + off[0]);
    uint32_t* ks = reinterpret_cast<uint32_t*>(w + off[1]);
    int32_t* idx = reinterpret_cast<int32_t*>(w + off[2]);
    int32_t* order = reinterpret_cast<int32_t*>(w + off[3]);
    int32_t* bstart = reinterpret_cast<int32_t*>(w + off[4]);
    int32_t* bend = reinterpret_cast<int32_t*>(w + off[5]);
    float* bxyz = reinterpret_cast<float*>(w + off[6]);
    float* brow = reinterpret_cast<float*>(w + off[7]);
    splat_key_kernel<<<grid1d(n, 256), 256, 0, st>>>(g, xyz, n, inv_voxel_size, data, ncomp, scale, levels, level, post,
                                                     acc, (uint32_t)count, keys, idx);
    NGS_CHECK_LAUNCH();
    const int bits = 32 - __builtin_clz((uint32_t)count);
    size_t tb = tmp;
    r = sort_pairs_u32(keys, ks, idx, order, n, bits, w, &tb, st);
    if (r != NGS_OK) return r;
    clear_words<<<grid1d(count, 256), 256, 0, st>>>(reinterpret_cast<uint32_t*>(bstart), count);
    NGS_CHECK_LAUNCH();
    clear_words<<<grid1d(count, 256), 256, 0, st>>>(reinterpret_cast<uint32_t*>(bend), count);
    NGS_CHECK_LAUNCH();
    splat_bins_kernel<<<grid1d(n, 256), 256, 0, st>>>(ks, order, n, (uint32_t)count, xyz, data, ncomp, scale, bstart,
                                                      bend, bxyz, brow);
    NGS_CHECK_LAUNCH();
    splat_gather_kernel<<<grid1d(count, 256), 256, 0, st>>>(g, node_coords, count, inv_voxel_size, data ? ncomp : 0,
                                                            bstart, bend, bxyz, brow, post, acc);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_eval_field(ngs_grid g, const float* field, const float* xyz, int64_t n, float inv_voxel_size, float scale,
                   int32_t accumulate, float* out, float* grad_out, ngs_stream_t s) {
    if (n <= 0) return NGS_OK;
    eval_field_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(g, field, xyz, n, inv_voxel_size, scale, accumulate, out, grad_out);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_eval_field_binned(ngs_grid g, const float* field, const int32_t* point_cell, const float* basis,
                          const int32_t* perm, int64_t n, float scale, float* out, ngs_stream_t s) {
    if (n <= 0) return NGS_OK;
    NGS_REQUIRE(field && point_cell && basis && perm && out, "ngs_eval_field_binned: null argument");
    eval_field_binned_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(g, field, point_cell, basis, perm, n, scale, out);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_field_support(ngs_grid g, const float* xyz, int64_t n, float inv_voxel_size, uint8_t* mask, ngs_stream_t s) {
    if (n <= 0) return NGS_OK;
    NGS_REQUIRE(mask && xyz, "ngs_field_support: null query or mask pointer");
    field_support_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(g, xyz, n, inv_voxel_size, mask);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_point_levels(const float* density, int64_t n, float min_density, float log_base, int32_t depth, int32_t* levels,
                     ngs_stream_t s) {
    NGS_REQUIRE(min_density > 0 && log_base > 1 && depth >= 1, "ngs_point_levels: bad parameters");
    if (n <= 0) return NGS_OK;
    // PyTorch divides a float tensor by a CPU scalar as a multiplication by the float reciprocal.
    float inv_min = 1.0f / min_density;
    float inv_ln = 1.0f / (float)std::log((double)log_base);
    levels_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(density, n, inv_min, inv_ln, depth, levels);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_point_basis(const float* xyz, int64_t n, float inv_voxel_size, int32_t stride, int32_t* cell, float* b,
                    ngs_stream_t s) {
    NGS_REQUIRE(stride > 0 && (stride & (stride - 1)) == 0, "ngs_point_basis: stride must be a power of two");
    if (n <= 0) return NGS_OK;
    point_basis_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(xyz, n, inv_voxel_size, stride, cell, b);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

static int bucket_layout(int64_t n, int64_t ncells, size_t* sort_tmp, int64_t* off_keys, int64_t* off_vin,
                         int64_t* off_scan, int64_t* off_counts, int64_t* total) {
    size_t tmp = 0;
    int r = sort_pairs_u32(nullptr, nullptr, nullptr, nullptr, n, 32, nullptr, &tmp, nullptr);
    if (r != NGS_OK) return r;
    const int64_t scan_bytes = ngs_scan_workspace_bytes(ncells);
    if (scan_bytes < 0) return NGS_ERR_HIP;
    auto al = [](int64_t x) { return (x + 255) / 256 * 256; };
    *sort_tmp = tmp;
    *off_keys = al((int64_t)tmp);
    *off_vin = *off_keys + al(n * 4);
    *off_scan = *off_vin + al(n * 4);
    *off_counts = *off_scan + al(scan_bytes);
    *total = *off_counts + al(ncells * 4);
    return NGS_OK;
}

int64_t ngs_bucket_workspace_bytes(int64_t n, int64_t ncells) {
    size_t tmp;
    int64_t a, b, c, d, total;
    if (bucket_layout(n, ncells, &tmp, &a, &b, &c, &d, &total) != NGS_OK) return -1;
    return total;
}

int ngs_bucket_points(const int32_t* cell_rank, int64_t n, int64_t ncells, int32_t* perm, int32_t* cell_start,
                      void* workspace, int64_t workspace_bytes, ngs_stream_t s) {
    NGS_REQUIRE(n > 0 && ncells > 0, "ngs_bucket_points: empty input");
    hipStream_t st = S(s);
    size_t tmp;
    int64_t ok, ov, osc, oc, total;
    int r = bucket_layout(n, ncells, &tmp, &ok, &ov, &osc, &oc, &total);
    if (r != NGS_OK) return r;
    NGS_REQUIRE(workspace_bytes >= total, "ngs_bucket_points: workspace too small");
    char* w = reinterpret_cast<char*>(workspace);
    uint32_t* kout = reinterpret_cast<uint32_t*>(w + ok);
    int32_t* vin = reinterpret_cast<int32_t*>(w + ov);
    iota_i32<<<grid1d(n, 256), 256, 0, st>>>(vin, n);
    NGS_CHECK_LAUNCH();
    int bits = 1;
    while (((int64_t)1 << bits) < ncells + 1 && bits < 32) ++bits;
    size_t tb = tmp;
    r = sort_pairs_u32(reinterpret_cast<const uint32_t*>(cell_rank), kout, vin, perm, n, bits, w, &tb, st);
    if (r != NGS_OK) return r;
    // the bounds of each cell's run in the sorted keys (no per-cell atomics, no scan)
    cell_starts_kernel<<<grid1d(n + 1, 256), 256, 0, st>>>(kout, n, ncells, cell_start);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_f64_to_f32(const double* in, float* out, int64_t n, ngs_stream_t s) {
    if (n <= 0) return NGS_OK;
    f64_to_f32_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(in, out, n);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_weighted_sum(const float* v, const float* w, int64_t n, double* sums, ngs_stream_t s) {
    hipStream_t st = S(s);
    zero_f64<<<1, 64, 0, st>>>(sums, 2);
    if (n > 0) {
        unsigned blocks = grid1d(n, 256);
        if (blocks > 256) blocks = 256;
        weighted_sum_kernel<<<blocks, 256, 0, st>>>(v, w, n, sums);
    }
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

}  // extern "C"
