/*
 * cuna.h -- public API of the MI355X-native CUNA adjoint NFFT.
 *
 * Replaces inc/cuna.h of johnh2o2/cunfft_adjoint (source-level drop-in):
 * same names, argument order and meaning, with hipStream_t in place of
 * cudaStream_t and C linkage (extern "C").  Link libcunaf.so (float) or,
 * with -DDOUBLE_PRECISION, libcunad.so (double).
 *
 * All functions compute the reference's transform
 *     f_hat[r] = normfac * exp((r/ng)^2 b/4) * sum_g grid[g] exp(+2 pi i g r/ng),
 *     grid[g]  = sum_j y_j exp(-(eps_j - m)^2/b)/sqrt(pi b),  g = j_j + m,
 * r = 0..ng-1 (see src/gridding.cu:36-65, src/cuna_kernel.cu:148-183,
 * src/gaussian_filter.cu:278-289).  Any ng >= 1 up to 2^23 (libcunaf) / 2^22
 * (libcunad), and powers of two up to 2^25 / 2^24, as cuFFT takes any length.
 * Errors are recorded (cuna_get_last_error) instead of calling exit().
 */
#ifndef CUNA_H
#define CUNA_H
#include "cuna_typedefs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* inc/cuna.h:33-34, src/cuna.cpp:38-49.  Host pointers: x[n], f[n] (f may be
 * NULL for unit weights), fft[ng] receives the transform.  Unlike the
 * reference (whose copy-back never runs), fft is written. */
void cunfft_adjoint(const dTyp *x, const dTyp *f, cTyp *fft, int n, int ng, unsigned int flags);

/* inc/cuna.h:41, src/cuna_kernel.cu:137-145.  Runs the plan built by init_plan;
 * the result is left in p->g_f_hat (and p->g_f_hat_win with
 * CALCULATE_WINDOW_FUNCTION) and copied to p->f_hat unless DONT_TRANSFER_TO_CPU. */
void cunfft_adjoint_from_plan(plan *p);

/* inc/cuna.h:48-49, src/cuna_kernel.cu:148-183.  Device pointers; gpu_fprops
 * from generate_filter_properties for the same x, n, ng.  Runs on the null
 * stream.  f_hat[ng] is overwritten (the reference accumulated into a
 * caller-zeroed f_hat; the results agree for a zeroed f_hat). */
void cunfft_adjoint_raw(const dTyp *x, const dTyp *f_data, Complex *f_hat, const int n,
                        const int ng, const filter_properties *gpu_fprops);

/* inc/cuna.h:51-53, src/cuna_kernel.cu:187-232.  As above, ordered on stream;
 * never blocks the host. */
void cunfft_adjoint_raw_async(const dTyp *x, const dTyp *f_data, Complex *f_hat, const int n,
                              const int ng, const filter_properties *gpu_fprops,
                              hipStream_t stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the transform of cunfft_adjoint_raw_async, with the
 * same arguments and result, that calls no hipMalloc/hipFree.  Its grid
 * lives in a 16 MiB device array of the library (one per GPU, held until the
 * process ends), so ng is at most 2^22 (libcunaf) / 2^21 (libcunad); larger
 * ng is rejected (CUNA_ERROR_UNSUPPORTED).  gpu_fprops must come from
 * generate_filter_properties / generate_pinned_filter_properties for this
 * x, n and ng (they build the plan; this call builds none).  All calls on
 * one GPU share that array: calls that overlap on a GPU (other streams, other
 * threads) race. */
void cunfft_adjoint_raw_async_noalloc(const dTyp *x, const dTyp *f_data, Complex *f_hat,
                                      const int n, const int ng,
                                      const filter_properties *gpu_fprops, hipStream_t stream);

/* inc/cuna.h:55-56 declares this without seed; the reference DEFINES it with
 * seed (src/cuna_kernel.cu:236-289) and that definition is what is exported
 * here.  f_hat holds nbs slabs of ng values: slab s is the transform of
 * bootstrap replicate s (values y[d], d = xorshift(wang(seed + s*n + p)) % n). */
void cunfft_adjoint_raw_async_bootstrap(const dTyp *x, const dTyp *f_data, Complex *f_hat,
                                        const int n, const int ng, const int nbs,
                                        const filter_properties *gpu_fprops, hipStream_t stream,
                                        const unsigned int seed);

#ifdef __cplusplus
}
#endif
#endif /* CUNA_H */
