/* Jacobi-PCG restated from csrc/sparse_solve/solve_pcg.cu:15-136 (oracle; test infrastructure only).
 *
 * x0 = 0, r = b; per iteration z = D^-1 r; rho = r.z; p = z (first) or p = (rho/rho1) p + z;
 * q = A p; alpha = rho / p.q; x += alpha p; r -= alpha q, or r = b - A x every ceil(sqrt(n))
 * iterations when res_fix; stop when ||r|| <= tol * ||b|| or after max_iter (< 0: unbounded).
 * Two precisions: all-double (the FP64 oracle) and float vectors/matrix with either float or
 * double dot products (the reference's cuBLAS dots are float).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#define PCG_BODY(real_t, acc_t)                                                             \
    int sqrt_n = (int)ceil(sqrt((double)n));                                                \
    real_t *r = (real_t *)malloc(sizeof(real_t) * (size_t)n);                               \
    real_t *z = (real_t *)malloc(sizeof(real_t) * (size_t)n);                               \
    real_t *p = (real_t *)malloc(sizeof(real_t) * (size_t)n);                               \
    real_t *q = (real_t *)malloc(sizeof(real_t) * (size_t)n);                               \
    acc_t bb = 0;                                                                           \
    _Pragma("omp parallel for reduction(+:bb) schedule(static)")                           \
    for (int i = 0; i < n; ++i) { x[i] = 0; r[i] = b[i]; bb += (acc_t)b[i] * b[i]; }       \
    real_t atol = (real_t)(tol * sqrt((double)bb));                                         \
    int it = 0;                                                                             \
    acc_t rho = 0, rho1;                                                                    \
    double resid = sqrt((double)bb);                                                        \
    while (max_iter < 0 || it < max_iter) {                                                 \
        acc_t rz = 0;                                                                       \
        _Pragma("omp parallel for reduction(+:rz) schedule(static)")                       \
        for (int i = 0; i < n; ++i) { z[i] = inv_diag[i] * r[i]; rz += (acc_t)r[i] * z[i]; } \
        rho1 = rho; rho = rz;                                                               \
        if (it == 0) {                                                                      \
            _Pragma("omp parallel for schedule(static)")                                   \
            for (int i = 0; i < n; ++i) p[i] = z[i];                                        \
        } else {                                                                            \
            real_t beta = (real_t)(rho / rho1);                                             \
            _Pragma("omp parallel for schedule(static)")                                   \
            for (int i = 0; i < n; ++i) p[i] = beta * p[i] + z[i];                          \
        }                                                                                   \
        acc_t pq = 0;                                                                       \
        _Pragma("omp parallel for reduction(+:pq) schedule(dynamic, 1024)")                \
        for (int i = 0; i < n; ++i) {                                                       \
            acc_t s = 0;                                                                    \
            for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) s += (acc_t)val[k] * p[col[k]]; \
            q[i] = (real_t)s; pq += (acc_t)p[i] * q[i];                                      \
        }                                                                                   \
        real_t alpha = (real_t)(rho / pq);                                                  \
        int fix = res_fix && ((it + 1) % sqrt_n == 0);                                      \
        _Pragma("omp parallel for schedule(static)")                                       \
        for (int i = 0; i < n; ++i) x[i] += alpha * p[i];                                   \
        acc_t rr = 0;                                                                       \
        if (fix) {                                                                          \
            _Pragma("omp parallel for reduction(+:rr) schedule(dynamic, 1024)")            \
            for (int i = 0; i < n; ++i) {                                                   \
                acc_t s = 0;                                                                \
                for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) s += (acc_t)val[k] * x[col[k]]; \
                r[i] = b[i] - (real_t)s; rr += (acc_t)r[i] * r[i];                          \
            }                                                                               \
        } else {                                                                            \
            _Pragma("omp parallel for reduction(+:rr) schedule(static)")                   \
            for (int i = 0; i < n; ++i) { r[i] -= alpha * q[i]; rr += (acc_t)r[i] * r[i]; } \
        }                                                                                   \
        ++it;                                                                               \
        resid = sqrt((double)rr);                                                           \
        if ((real_t)resid <= atol) break;                                                   \
    }                                                                                       \
    free(r); free(z); free(p); free(q);                                                     \
    *resid_out = resid;                                                                     \
    return it;

int oracle_pcg_f64(int n, const int32_t *rowptr, const int32_t *col, const double *val, const double *b,
                   const double *inv_diag, double tol, int max_iter, int res_fix, double *x, double *resid_out) {
    PCG_BODY(double, double)
}

int oracle_pcg_f32(int n, const int32_t *rowptr, const int32_t *col, const float *val, const float *b,
                   const float *inv_diag, double tol, int max_iter, int res_fix, float *x, double *resid_out) {
    PCG_BODY(float, float)
}

int oracle_pcg_f32_d(int n, const int32_t *rowptr, const int32_t *col, const float *val, const float *b,
                     const float *inv_diag, double tol, int max_iter, int res_fix, float *x, double *resid_out) {
    PCG_BODY(float, double)
}
