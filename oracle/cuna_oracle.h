/*
 * cuna_oracle.h -- CPU restatement of the CUNA adjoint NFFT, used ONLY as a test
 * checker and as the CPU baseline timed beside the GPU numbers.
 *
 * TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library.  The product (libcunaf.so /
 * libcunad.so) never links, loads or calls it.
 *
 * All arrays are host memory.  Complex arrays are interleaved (re, im) doubles.
 * `dbl` selects which reference build is restated: 0 = libcunaf (float
 * decisions, FILTER_RADIUS 6), 1 = libcunad (double, FILTER_RADIUS 12).
 */
#ifndef CUNA_ORACLE_H
#define CUNA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* bootstrap RNG of src/gridding.cu:67-82 */
uint32_t orc_wang_hash(uint32_t s);
uint32_t orc_xorshift(uint32_t s);
/* d[t] for t in [0, n*nbs): xorshift(wang(seed + t)) % n  (src/gridding.cu:101) */
void orc_bootstrap_indices(uint32_t seed, int n, int nbs, int32_t *d);

/* filter constants (src/gaussian_filter.cu:29-33,125-134) */
void orc_compat_params(int n, int ng, int dbl, double *b, double *normfac, int *radius);
/* per-point decisions (src/gaussian_filter.cu:213-216, src/gridding.cu:43-45) */
void orc_compat_point(double x, int n, int ng, int dbl, int *u, double *eps, int *j);

/* real grid of the compat spread (src/gridding.cu:36-65); grid has ng doubles */
int orc_compat_grid(const double *x, const double *y, int n, int ng, int dbl, double *grid);
/* nbs slabs of ng (src/gridding.cu:86-117) */
int orc_compat_grid_bootstrap(const double *x, const double *y, int n, int ng, int nbs,
                              uint32_t seed, int dbl, double *grid);
/* full compat transform: ng complex outputs (src/cuna_kernel.cu:148-183) */
int orc_compat_adjoint(const double *x, const double *y, int n, int ng, int dbl, double *out);
int orc_compat_adjoint_bootstrap(const double *x, const double *y, int n, int ng, int nbs,
                                 uint32_t seed, int dbl, double *out);

/* unnormalised DFT in place, any n >= 1: z[r] = sum_g z[g] exp(sign*2*pi*i*g*r/n) */
int orc_fft(double *z, int n, int sign);

/* Greengard-Lee Gaussian adjoint NFFT, outputs k = -M/2 .. M/2-1 (M complex) */
int orc_gl_adjoint(const double *x, const double *y, int n, int M, int sigma, int m, double *out);
/* gradient of L with dL/dF_k = g_k (2*M doubles) through the same discrete operator */
int orc_gl_grad(const double *x, const double *y, int n, int M, int sigma, int m,
                const double *g, double *dy, double *dx);

/* direct NDFT F_k = sum_j y_j exp(2 pi i k x_j) for nb bands k0[b] .. k0[b]+len-1;
 * y may be NULL (unit weights).  out has nb*len complex. */
void orc_ndft_bands(const double *x, const double *y, int n, const int64_t *k0, int nb, int len,
                    double *out);
/* analytic NDFT gradient for k = -M/2..M/2-1 */
void orc_ndft_grad(const double *x, const double *y, int n, int M, const double *g,
                   double *dy, double *dx);

/* number of OpenMP threads the oracle uses */
int orc_num_threads(void);

/* forward (type-2) NDFT over k = -M/2..M/2-1 (interleaved complex) */
void orc_ndft_forward(const double *x, int n, const double *fhat, int M, double *f);

/* Lomb-Scargle power at f0 + k df, k < nf (direct FP64) */
void orc_lomb_scargle(const double *t, const double *y, int n, double f0, double df, int nf,
                      double *power);

#ifdef __cplusplus
}
#endif
#endif
