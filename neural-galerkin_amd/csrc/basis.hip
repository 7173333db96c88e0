// Learned additive bases (jspsr/bases/additive.py): the derivative-derivative integral of basis pairs
// (DDI, csrc/integration/ddi.cu) and the evaluation of B(x) = sum_c m_c q_c(x) per axis, with their
// backward passes.
#include "common.h"

namespace ngs {

// p[n, v] = sum_si sum_ti s[n, 3 si + v] t[n, 3 ti + v] ff[si][ti][rel[n, v]], q likewise with dd.
// One thread per pair; the sums run si outer, ti inner, in FP32.
__global__ void ddi_forward_kernel(const float* __restrict__ s, const float* __restrict__ t,
                                   const int64_t* __restrict__ rel, int64_t n, int nc, const float* __restrict__ ff,
                                   const float* __restrict__ dd, int L, float* __restrict__ p, float* __restrict__ q) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* sr = s + i * 3 * nc;
    const float* tr = t + i * 3 * nc;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        const int r = (int)rel[3 * i + v];
        float pa = 0.f, qa = 0.f;
        for (int si = 0; si < nc; ++si) {
            const float sv = sr[3 * si + v];
            for (int ti = 0; ti < nc; ++ti) {
                const float st = sv * tr[3 * ti + v];
                const int64_t k = ((int64_t)si * nc + ti) * L + r;
                pa += st * ff[k];
                qa += st * dd[k];
            }
        }
        p[3 * i + v] = pa;
        q[3 * i + v] = qa;
    }
}

// gs[n, 3 si + v] = sum_ti c(si, ti) t[n, 3 ti + v], gt[n, 3 ti + v] = sum_si c(si, ti) s[n, 3 si + v],
// c = gp[n, v] ff[si][ti][r] + gq[n, v] dd[si][ti][r].
__global__ void ddi_backward_kernel(const float* __restrict__ s, const float* __restrict__ t,
                                    const int64_t* __restrict__ rel, int64_t n, int nc, const float* __restrict__ ff,
                                    const float* __restrict__ dd, int L, const float* __restrict__ gp,
                                    const float* __restrict__ gq, float* __restrict__ gs, float* __restrict__ gt) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* sr = s + i * 3 * nc;
    const float* tr = t + i * 3 * nc;
    float* gsr = gs + i * 3 * nc;
    float* gtr = gt + i * 3 * nc;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        const int r = (int)rel[3 * i + v];
        const float a = gp[3 * i + v], b = gq[3 * i + v];
        for (int si = 0; si < nc; ++si) {
            float acc = 0.f;
            for (int ti = 0; ti < nc; ++ti) {
                const int64_t k = ((int64_t)si * nc + ti) * L + r;
                acc += (a * ff[k] + b * dd[k]) * tr[3 * ti + v];
            }
            gsr[3 * si + v] = acc;
        }
        for (int ti = 0; ti < nc; ++ti) {
            float acc = 0.f;
            for (int si = 0; si < nc; ++si) {
                const int64_t k = ((int64_t)si * nc + ti) * L + r;
                acc += (a * ff[k] + b * dd[k]) * sr[3 * si + v];
            }
            gtr[3 * ti + v] = acc;
        }
    }
}

// The same two kernels with the tables staged in LDS as (ff, dd) pairs in [r][si][ti] order, the
// pair's coefficients of one axis in registers, and a grid-stride loop over pairs (every block loads
// the tables once). Same products and summation order as above, so the results are identical.
constexpr int DDI_MAXC = 16;
constexpr int DDI_LDS_BYTES = 64 * 1024;

__device__ __forceinline__ void ddi_stage_tables(const float* __restrict__ ff, const float* __restrict__ dd, int nc,
                                                 int L, float2* tab) {
    const int nn = nc * nc;
    for (int k = threadIdx.x; k < L * nn; k += blockDim.x) {
        const int r = k / nn, st = k - r * nn;
        tab[k] = make_float2(ff[(int64_t)st * L + r], dd[(int64_t)st * L + r]);
    }
    __syncthreads();
}

__global__ void __launch_bounds__(256) ddi_forward_lds_kernel(const float* __restrict__ s, const float* __restrict__ t,
                                                              const int64_t* __restrict__ rel, int64_t n, int nc,
                                                              const float* __restrict__ ff,
                                                              const float* __restrict__ dd, int L,
                                                              float* __restrict__ p, float* __restrict__ q) {
#pragma clang fp contract(off)
    extern __shared__ float2 tab[];
    ddi_stage_tables(ff, dd, nc, L, tab);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float* sr = s + i * 3 * nc;
        const float* tr = t + i * 3 * nc;
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            const float2* tb = tab + (int64_t)rel[3 * i + v] * nc * nc;
            float tv[DDI_MAXC];
#pragma unroll
            for (int ti = 0; ti < DDI_MAXC; ++ti) tv[ti] = ti < nc ? tr[3 * ti + v] : 0.f;
            float pa = 0.f, qa = 0.f;
            for (int si = 0; si < nc; ++si) {
                const float sv = sr[3 * si + v];
#pragma unroll
                for (int ti = 0; ti < DDI_MAXC; ++ti) {
                    if (ti < nc) {
                        const float st = sv * tv[ti];
                        const float2 w = tb[si * nc + ti];
                        pa += st * w.x;
                        qa += st * w.y;
                    }
                }
            }
            p[3 * i + v] = pa;
            q[3 * i + v] = qa;
        }
    }
}

// Forward with the coefficient rows of 256 consecutive pairs also staged in LDS: the rows are read
// from global memory as one contiguous, coalesced range per tile (a pair's row is 12 C bytes, so
// direct per-thread row reads touch a different cache line per lane). Same arithmetic and order.
constexpr int DDI_TILE = 256;

__global__ void __launch_bounds__(DDI_TILE) ddi_forward_tiled_kernel(
    const float* __restrict__ s, const float* __restrict__ t, const int64_t* __restrict__ rel, int64_t n, int nc,
    const float* __restrict__ ff, const float* __restrict__ dd, int L, float* __restrict__ p, float* __restrict__ q) {
#pragma clang fp contract(off)
    extern __shared__ float2 tab[];
    ddi_stage_tables(ff, dd, nc, L, tab);
    const int w = 3 * nc;
    float* rs = reinterpret_cast<float*>(tab + (int64_t)L * nc * nc);
    float* rt = rs + DDI_TILE * w;
    const int64_t ntiles = (n + DDI_TILE - 1) / DDI_TILE;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t base = tile * DDI_TILE;
        const int rows = (int)min((int64_t)DDI_TILE, n - base);
        __syncthreads();  // the previous tile's rows are no longer read
        for (int k = threadIdx.x; k < rows * w; k += DDI_TILE) {
            rs[k] = s[base * w + k];
            rt[k] = t[base * w + k];
        }
        __syncthreads();
        if ((int)threadIdx.x < rows) {
            const int64_t i = base + threadIdx.x;
            const float* sr = rs + threadIdx.x * w;
            const float* tr = rt + threadIdx.x * w;
#pragma unroll
            for (int v = 0; v < 3; ++v) {
                const float2* tb = tab + (int64_t)rel[3 * i + v] * nc * nc;
                float tv[DDI_MAXC];
#pragma unroll
                for (int ti = 0; ti < DDI_MAXC; ++ti) tv[ti] = ti < nc ? tr[3 * ti + v] : 0.f;
                float pa = 0.f, qa = 0.f;
                for (int si = 0; si < nc; ++si) {
                    const float sv = sr[3 * si + v];
#pragma unroll
                    for (int ti = 0; ti < DDI_MAXC; ++ti) {
                        if (ti < nc) {
                            const float st = sv * tv[ti];
                            const float2 wt = tb[si * nc + ti];
                            pa += st * wt.x;
                            qa += st * wt.y;
                        }
                    }
                }
                p[3 * i + v] = pa;
                q[3 * i + v] = qa;
            }
        }
    }
}

__global__ void __launch_bounds__(256) ddi_backward_lds_kernel(const float* __restrict__ s, const float* __restrict__ t,
                                                               const int64_t* __restrict__ rel, int64_t n, int nc,
                                                               const float* __restrict__ ff,
                                                               const float* __restrict__ dd, int L,
                                                               const float* __restrict__ gp,
                                                               const float* __restrict__ gq, float* __restrict__ gs,
                                                               float* __restrict__ gt) {
#pragma clang fp contract(off)
    extern __shared__ float2 tab[];
    ddi_stage_tables(ff, dd, nc, L, tab);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float* sr = s + i * 3 * nc;
        const float* tr = t + i * 3 * nc;
        float* gsr = gs + i * 3 * nc;
        float* gtr = gt + i * 3 * nc;
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            const float2* tb = tab + (int64_t)rel[3 * i + v] * nc * nc;
            const float a = gp[3 * i + v], b = gq[3 * i + v];
            float tv[DDI_MAXC], gacc[DDI_MAXC];
#pragma unroll
            for (int k = 0; k < DDI_MAXC; ++k) {
                tv[k] = k < nc ? tr[3 * k + v] : 0.f;
                gacc[k] = 0.f;
            }
            for (int si = 0; si < nc; ++si) {
                const float svs = sr[3 * si + v];
                float acc = 0.f;
#pragma unroll
                for (int ti = 0; ti < DDI_MAXC; ++ti) {
                    if (ti < nc) {
                        const float2 w = tb[si * nc + ti];
                        const float c = a * w.x + b * w.y;
                        acc += c * tv[ti];
                        gacc[ti] += c * svs;
                    }
                }
                gsr[3 * si + v] = acc;
            }
#pragma unroll
            for (int ti = 0; ti < DDI_MAXC; ++ti)
                if (ti < nc) gtr[3 * ti + v] = gacc[ti];
        }
    }
}

static bool ddi_use_lds(int nc, int L) { return nc <= DDI_MAXC && (int64_t)nc * nc * L * 8 <= DDI_LDS_BYTES; }
static int64_t ddi_tiled_bytes(int nc, int L) { return (int64_t)nc * nc * L * 8 + 2LL * DDI_TILE * 3 * nc * 4; }

static unsigned ddi_grid(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > 2048 ? 2048 : b));
}

// Elementary functions q(x) of additive_components.py, evaluated as the reference's torch code does
// (no support mask except Bezier's). A component is the product of two factors (kind, a0, a1):
// kind 0 = 1, 1 = x^a0, 2 = Bezier, 3 = sin(a0 x + a1).
constexpr int FACTOR_ONE = 0, FACTOR_POWER = 1, FACTOR_BEZIER = 2, FACTOR_SIN = 3;

__device__ __forceinline__ float ipow(float x, int k) {
    float r = 1.f;
    for (int j = 0; j < k; ++j) r *= x;
    return r;
}

__device__ __forceinline__ void factor_eval(const float* d, float x, float* f, float* df) {
#pragma clang fp contract(off)
    const int kind = (int)d[0];
    if (kind == FACTOR_POWER) {
        const float p = d[1];
        const int ip = (int)p;
        if (p == 0.f) { *f = 1.f; *df = 0.f; return; }
        if ((float)ip == p && ip > 0) {
            *f = ipow(x, ip);
            *df = ip == 1 ? 1.f : p * ipow(x, ip - 1);
        } else {
            *f = powf(x, p);
            *df = p * powf(x, p - 1.f);
        }
    } else if (kind == FACTOR_BEZIER) {
        *f = bspline(x);
        *df = bspline_deriv(x);
    } else if (kind == FACTOR_SIN) {
        *f = sinf(d[1] * x + d[2]);
        *df = d[1] * cosf(d[1] * x + d[2]);
    } else {
        *f = 1.f;
        *df = 0.f;
    }
}

// q(x) and q'(x) of component c (descriptor: two factors of 3 floats each).
__device__ __forceinline__ void component_eval(const float* desc, float x, float* q, float* dq) {
#pragma clang fp contract(off)
    float fa, da, fb, db;
    factor_eval(desc, x, &fa, &da);
    if ((int)desc[3] == FACTOR_ONE) {
        *q = fa;
        *dq = da;
        return;
    }
    factor_eval(desc + 3, x, &fb, &db);
    *q = fa * fb;
    *dq = fa * db + da * fb;  // Product.evaluate_derivative order
}

// funcs[n, v] = sum_c q_c(x[n, v]) m[id[n], 3c + v]; dfuncs[n, v] = sum_c q_c'(x) m / stride (EvalFunc,
// EvalDerivFunc of additive.py:71-136), components summed in order.
__global__ void additive_eval_kernel(const float* __restrict__ xyz, const float* __restrict__ coeffs,
                                     const int64_t* __restrict__ ids, int64_t n, const float* __restrict__ desc, int nc,
                                     float stride, float* __restrict__ funcs, float* __restrict__ dfuncs) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* m = coeffs + ids[i] * 3 * nc;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        const float x = xyz[3 * i + v];
        float f = 0.f, df = 0.f;
        for (int c = 0; c < nc; ++c) {
            float q, dq;
            component_eval(desc + 6 * c, x, &q, &dq);
            const float mc = m[3 * c + v];
            f += q * mc;
            df += dq * mc / stride;
        }
        funcs[3 * i + v] = f;
        if (dfuncs) dfuncs[3 * i + v] = df;
    }
}

// grad_coeffs[id[n], 3c + v] += gf[n, v] q_c(x) + gd[n, v] q_c'(x) / stride (atomics; gd may be NULL).
__global__ void additive_eval_backward_kernel(const float* __restrict__ xyz, const int64_t* __restrict__ ids,
                                              int64_t n, const float* __restrict__ desc, int nc, float stride,
                                              const float* __restrict__ gf, const float* __restrict__ gd,
                                              float* __restrict__ gcoeffs) {
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float* g = gcoeffs + ids[i] * 3 * nc;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        const float x = xyz[3 * i + v];
        const float a = gf[3 * i + v], b = gd ? gd[3 * i + v] : 0.f;
        for (int c = 0; c < nc; ++c) {
            float q, dq;
            component_eval(desc + 6 * c, x, &q, &dq);
            atomicAdd(&g[3 * c + v], gd ? a * q + b * (dq / stride) : a * q);
        }
    }
}

}  // namespace ngs

using namespace ngs;

// Screening term of a level pair for per-(point, voxel) basis values (the generic basis path,
// reconstructor.py:59-77 -> sparse_product.cu:100-137): for every point p, every entry (i, va) of p at
// level a and every entry (j, vb) of p at level b, out[k] += (va vb) w_p where k is the slot of column
// j in row i of the pair structure (rowptr, cols; columns ascending within a row), found by binary
// search; pairs absent from the structure are dropped, as the reference's hash lookup drops them.
// One thread per entry of level a; entries of one point are contiguous (a_start, b_start: n + 1
// offsets).
__global__ void screen_pairs_kernel(int64_t na, const int32_t* __restrict__ a_pt, const int32_t* __restrict__ a_ids,
                                    const float* __restrict__ a_vals, const int32_t* __restrict__ b_start,
                                    const int32_t* __restrict__ b_ids, const float* __restrict__ b_vals,
                                    const float* __restrict__ w, float w_scalar, const int32_t* __restrict__ rowptr,
                                    const int32_t* __restrict__ cols, float* __restrict__ out) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= na) return;
    const int32_t p = a_pt[t], i = a_ids[t];
    const float va = a_vals[t];
    const float wp = w ? w[p] : w_scalar;
    const int32_t r0 = rowptr[i], r1 = rowptr[i + 1];
    for (int32_t k2 = b_start[p]; k2 < b_start[p + 1]; ++k2) {
        const int32_t j = b_ids[k2];
        int32_t lo = r0, hi = r1;
        while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            if (cols[mid] < j) lo = mid + 1;
            else hi = mid;
        }
        if (lo < r1 && cols[lo] == j) atomicAdd(&out[lo], (va * b_vals[k2]) * wp);
    }
}

extern "C" {

int ngs_ddi_forward(const float* source_expanded, const float* target_expanded, const int64_t* rel_pos, int64_t n,
                    int32_t n_components, const float* f_f_i, const float* df_df_i, int32_t table_len, float* p_mats,
                    float* q_mats, ngs_stream_t s) {
    NGS_REQUIRE(n_components >= 1 && n_components <= 64 && table_len >= 1, "ngs_ddi_forward: bad sizes");
    if (n <= 0) return NGS_OK;
    NGS_REQUIRE(source_expanded && target_expanded && rel_pos && f_f_i && df_df_i && p_mats && q_mats,
                "ngs_ddi_forward: null argument");
    if (ddi_use_lds(n_components, table_len) && ddi_tiled_bytes(n_components, table_len) <= DDI_LDS_BYTES)
        ddi_forward_tiled_kernel<<<ddi_grid(n), DDI_TILE, (size_t)ddi_tiled_bytes(n_components, table_len), S(s)>>>(
            source_expanded, target_expanded, rel_pos, n, n_components, f_f_i, df_df_i, table_len, p_mats, q_mats);
    else if (ddi_use_lds(n_components, table_len))
        ddi_forward_lds_kernel<<<ddi_grid(n), 256, (size_t)n_components * n_components * table_len * 8, S(s)>>>(
            source_expanded, target_expanded, rel_pos, n, n_components, f_f_i, df_df_i, table_len, p_mats, q_mats);
    else
        ddi_forward_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(source_expanded, target_expanded, rel_pos, n,
                                                             n_components, f_f_i, df_df_i, table_len, p_mats, q_mats);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_ddi_backward(const float* source_expanded, const float* target_expanded, const int64_t* rel_pos, int64_t n,
                     int32_t n_components, const float* f_f_i, const float* df_df_i, int32_t table_len,
                     const float* grad_p, const float* grad_q, float* grad_s, float* grad_t, ngs_stream_t s) {
    NGS_REQUIRE(n_components >= 1 && n_components <= 64 && table_len >= 1, "ngs_ddi_backward: bad sizes");
    if (n <= 0) return NGS_OK;
    NGS_REQUIRE(source_expanded && target_expanded && rel_pos && f_f_i && df_df_i && grad_p && grad_q && grad_s &&
                    grad_t,
                "ngs_ddi_backward: null argument");
    if (ddi_use_lds(n_components, table_len))
        ddi_backward_lds_kernel<<<ddi_grid(n), 256, (size_t)n_components * n_components * table_len * 8, S(s)>>>(
            source_expanded, target_expanded, rel_pos, n, n_components, f_f_i, df_df_i, table_len, grad_p, grad_q,
            grad_s, grad_t);
    else
        ddi_backward_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(source_expanded, target_expanded, rel_pos, n,
                                                              n_components, f_f_i, df_df_i, table_len, grad_p, grad_q,
                                                              grad_s, grad_t);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_additive_eval(const float* xyz, const float* coeffs, const int64_t* feat_ids, int64_t n,
                      const float* components, int32_t n_components, float stride, float* funcs, float* dfuncs,
                      ngs_stream_t s) {
    NGS_REQUIRE(n_components >= 1 && n_components <= 64 && stride > 0.f, "ngs_additive_eval: bad sizes");
    if (n <= 0) return NGS_OK;
    NGS_REQUIRE(xyz && coeffs && feat_ids && components && funcs, "ngs_additive_eval: null argument");
    additive_eval_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(xyz, coeffs, feat_ids, n, components, n_components, stride,
                                                           funcs, dfuncs);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_additive_eval_backward(const float* xyz, const int64_t* feat_ids, int64_t n, const float* components,
                               int32_t n_components, float stride, const float* grad_funcs, const float* grad_dfuncs,
                               float* grad_coeffs, ngs_stream_t s) {
    NGS_REQUIRE(n_components >= 1 && n_components <= 64 && stride > 0.f, "ngs_additive_eval_backward: bad sizes");
    if (n <= 0) return NGS_OK;
    NGS_REQUIRE(xyz && feat_ids && components && grad_funcs && grad_coeffs, "ngs_additive_eval_backward: null argument");
    additive_eval_backward_kernel<<<grid1d(n, 256), 256, 0, S(s)>>>(xyz, feat_ids, n, components, n_components, stride,
                                                                    grad_funcs, grad_dfuncs, grad_coeffs);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

int ngs_screen_pairs(int64_t na, const int32_t* a_pt, const int32_t* a_ids, const float* a_vals,
                     const int32_t* b_start, const int32_t* b_ids, const float* b_vals, const float* w,
                     float w_scalar, const int32_t* rowptr, const int32_t* cols, float* out, ngs_stream_t s) {
    if (na <= 0) return NGS_OK;
    NGS_REQUIRE(a_pt && a_ids && a_vals && b_start && rowptr && cols && out, "ngs_screen_pairs: null argument");
    screen_pairs_kernel<<<grid1d(na, 256), 256, 0, S(s)>>>(na, a_pt, a_ids, a_vals, b_start, b_ids, b_vals, w,
                                                            w_scalar, rowptr, cols, out);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

}  // extern "C"
