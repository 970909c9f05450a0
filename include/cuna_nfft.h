/*
 * cuna_nfft.h -- extensions of the CUNA API with no counterpart in the
 * reference library (johnh2o2/cunfft_adjoint exports only inc/cuna.h,
 * inc/cuna_filter.h and init_plan/free_plan).
 *
 *   - accurate mode: textbook Greengard-Lee Gaussian NFFT,
 *       F_k = sum_j y_j exp(2 pi i k x_j),  k = -M/2 .. M/2-1,
 *     on a grid of G = sigma*M cells with 2m taps (rel. error ~9e-9 at
 *     sigma=2, m=8 in exact arithmetic; the FP32 library reaches ~2e-7);
 *   - compat mode: the reference formula (the M outputs are the ng values of
 *     cunfft_adjoint_raw), usable through the same plan objects;
 *   - times input: x_j = frac(df * t_j) reduced in FP64 on the GPU, so long
 *     baselines keep their phase in the FP32 library;
 *   - batches of curves of uneven length (CSR offsets) in one launch;
 *   - bootstrap resampling (accurate mode) and its window transform at 2f;
 *   - backward: gradients of a real loss with respect to y and x;
 *   - the FFT stage on its own.
 *
 * All pointers are device pointers unless stated; every call is ordered on
 * `stream` and returns a CUNA_* status (also kept in cuna_get_last_error).
 * A plan executes on one stream at a time: it owns one workspace.
 */
#ifndef CUNA_NFFT_H
#define CUNA_NFFT_H
#include "cuna_typedefs.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cuna_nfft_plan_s *cuna_nfft_plan;

#define CUNA_MODE_ACCURATE 0
#define CUNA_MODE_COMPAT 1

/* plan flags */
#define CUNA_PLAN_WINDOW2 0x1 /* also bin the points 2x (window transform at 2f) */
#define CUNA_PLAN_UNFUSED 0x2 /* bootstrap through spread + batched FFT kernels instead of
                                 the fused single-launch kernel (grids <= 2^15 cells) */
#define CUNA_PLAN_GATHER 0x4  /* spread in gather form (each cell sums the points that
                                 reach it) instead of the default scatter with fixed-point
                                 integer LDS atomics; both are deterministic */

typedef struct {
    int mode;      /* CUNA_MODE_ACCURATE (default) or CUNA_MODE_COMPAT */
    int sigma;     /* accurate: oversampling, G = sigma*M, power of two >= 2 (default 2) */
    int m;         /* accurate: filter half-width, 2m taps, 2..16 (default 8 / 13) */
    int tile_log2; /* spread tile of 2^tile_log2 cells; 0 = library default */
    int flags;     /* CUNA_PLAN_* */
    int fft_w;     /* FFT pass 1: columns per workgroup (1,2,4,8; 0 = library choice) */
    int fft_h;     /* FFT pass 2: rows (row pairs) per workgroup (1,2,4; 0 = library choice) */
} cuna_nfft_params;

typedef struct {
    int mode, sigma, m, ntaps;
    int n, ncurves, M, G;
    int tile_log2, ntiles;
    double beta;
} cuna_nfft_info;

void cuna_nfft_default_params(cuna_nfft_params *p);

/* n points x (accurate: any real, reduced mod 1; compat: [-1/2, 1/2]).
 * accurate: M modes, M a power of two >= 2; compat: M = ng outputs.
 * Plan creation (every form) is ordered on `stream` and does not wait for
 * it (the first plan of a grid size uploads that size's FFT twiddle tables
 * with a synchronous copy). */
int cuna_nfft_plan_create(cuna_nfft_plan *plan, const dTyp *x, int n, int M,
                          const cuna_nfft_params *params, hipStream_t stream);

/* accurate mode from double times t (device): frequencies k*df, k = -M/2..M/2-1,
 * x_j = frac(df * t_j) evaluated in FP64 */
int cuna_nfft_plan_create_times(cuna_nfft_plan *plan, const double *t, int n, double df, int M,
                                const cuna_nfft_params *params, hipStream_t stream);

/* ncurves curves; curve c holds points offsets[c] .. offsets[c+1]-1 of x.
 * offsets is a HOST array of ncurves+1 ints with offsets[0] = 0. */
int cuna_nfft_plan_create_batch(cuna_nfft_plan *plan, const dTyp *x, const int *offsets,
                                int ncurves, int M, const cuna_nfft_params *params,
                                hipStream_t stream);

/* as above from double times t (device): every curve on the frequency grid
 * k*df, x_j = frac(df * t_j) evaluated in FP64 */
int cuna_nfft_plan_create_batch_times(cuna_nfft_plan *plan, const double *t, const int *offsets,
                                      int ncurves, double df, int M,
                                      const cuna_nfft_params *params, hipStream_t stream);

/* out: ncurves x M complex (curve-major).  y: n values or NULL (unit weights). */
int cuna_nfft_adjoint(cuna_nfft_plan plan, const dTyp *y, Complex *out, hipStream_t stream);

/* accurate mode: nrhs value vectors of one plan in one spread launch and one
 * batched FFT.  y: nrhs x n (row r holds the values of right-hand side r, in
 * the plan's point order); out: ncurves x nrhs x M complex, row c*nrhs + r
 * equal to cuna_nfft_adjoint of row r for curve c.
 * No counterpart in the reference (its batching is over bootstrap replicates,
 * src/cuna_kernel.cu:236-289). */
int cuna_nfft_adjoint_many(cuna_nfft_plan plan, const dTyp *y, int nrhs, Complex *out,
                           hipStream_t stream);

/* accurate mode, per plan: nbs bootstrap replicates.  Replicate s of curve c
 * draws d = offsets[c] + xorshift(wang(seed + s*n_c + p)) % n_c for
 * p = 0 .. n_c-1 (the reference's generator, src/gridding.cu:67-101).
 * out_y[c][s][:]: transform of the values y[d] at points x_d;
 * out_w[c][s][:] (nullable, needs CUNA_PLAN_WINDOW2): unit weights at 2 x_d. */
int cuna_nfft_bootstrap(cuna_nfft_plan plan, const dTyp *y, int nbs, unsigned int seed,
                        Complex *out_y, Complex *out_w, hipStream_t stream);

/* accurate mode, CUNA_PLAN_WINDOW2 plan: Lomb-Scargle periodograms of the
 * same replicates (Press & Rybicki 1989; data centred on the replicate mean,
 * "standard" normalisation, 0 <= P <= 1), computed in the fused bootstrap
 * kernel without writing the spectra.  Frequency k*df (times plans; k/1 for
 * x plans), k = 0 .. M/2-1:
 *   P_k = [YC^2/CC + YS^2/SS] / sum (y - ybar)^2,  tan 2wtau = S2/C2,
 * from F = sum (y_d - ybar) e^{i w t_d} and sum e^{2 i w t_d}.
 * out_power[c*nbs + s][k] (nullable), out_peak[c*nbs + s] = max over
 * k >= 1 and out_kpeak[c*nbs + s] its k (smallest on ties; both nullable).
 * Grids of at most 2^15 cells (FP32) / 2^14 (FP64) and m in {6,8,10,12,13,16};
 * otherwise CUNA_ERROR_UNSUPPORTED.  No counterpart in the reference (its
 * callers compute the periodogram from the two spectra of
 * cunfft_adjoint_raw_async_bootstrap, src/cuna_kernel.cu:236-289). */
int cuna_nfft_bootstrap_ls(cuna_nfft_plan plan, const dTyp *y, int nbs, unsigned int seed,
                           dTyp *out_power, dTyp *out_peak, int *out_kpeak,
                           hipStream_t stream);

/* accurate mode: forward (type-2) NFFT, the adjoint of cuna_nfft_adjoint:
 * f[j] = sum_k fhat[c][M/2 + k] e^{-2 pi i k x_j}, k = -M/2 .. M/2-1, for the
 * points j of curve c (x_j = frac(df t_j) for times plans).  fhat: ncurves x M
 * complex (the adjoint's output layout); f: n complex.  No counterpart in the
 * reference, which has only the adjoint transform. */
int cuna_nfft_forward(cuna_nfft_plan plan, const Complex *fhat, Complex *f, hipStream_t stream);

/* accurate mode: given grad_out = dL/dRe F + i dL/dIm F (ncurves x M complex),
 * writes dL/dy (n) and dL/dx (n); either output may be NULL.  For a times
 * plan dL/dt_j = df * dL/dx_j. */
int cuna_nfft_backward(cuna_nfft_plan plan, const dTyp *y, const Complex *grad_out,
                       dTyp *grad_y, dTyp *grad_x, hipStream_t stream);

int cuna_nfft_plan_info(cuna_nfft_plan plan, cuna_nfft_info *info);

/* Profiling: with events set, cuna_nfft_adjoint records in events[0] and
 * events[1] the start and stop times of its spread kernel (taken by that
 * kernel's dispatch), events[2] between the two FFT passes (large grids
 * only) and events[3] after the last kernel, all on the execution stream.
 * nevents = 0 clears them. */
int cuna_nfft_set_profile_events(cuna_nfft_plan plan, hipEvent_t *events, int nevents);
int cuna_nfft_plan_destroy(cuna_nfft_plan plan);

/* Autotuner: fills *params for accurate-mode plans of these points.  Per
 * sigma in {2, 4} it takes the smallest m whose Gaussian-gridding error bound
 * exp(-pi m (1 - 1/(2 sigma - 1))) plus the arithmetic's rounding floor meets
 * tol (relative L2 error vs the direct NDFT), then times the spread tile
 * (2^8..2^11 cells) and FFT workgroup shapes on the device and keeps the
 * fastest.  Results are cached (appended) in a text file keyed by GPU,
 * precision, n rounded to a power of two, M and tol: cache_path NULL uses
 * $CUNA_TUNE_CACHE or ~/.cache/cuna_tune.txt, "" disables the cache. */
int cuna_nfft_autotune(const dTyp *x, int n, int M, double tol, const char *cache_path,
                       cuna_nfft_params *params, hipStream_t stream);

/* Graphs: every call above made between cuna_stream_capture_begin and
 * cuna_stream_capture_end on `stream` (a created stream, not the null
 * stream) is recorded instead of run; cuna_graph_launch replays the sequence.
 * Run each captured call once beforehand so that its workspaces exist: a
 * capture allocates nothing.  The captured kernels use no memset nodes. */
typedef struct cuna_graph_s *cuna_graph;
int cuna_stream_capture_begin(hipStream_t stream);
int cuna_stream_capture_end(hipStream_t stream, cuna_graph *graph);
int cuna_graph_launch(cuna_graph graph, hipStream_t stream);
int cuna_graph_destroy(cuna_graph graph);

/* Batches of curves through the reference formula: the batched transform and
 * filter generator that the reference keeps commented out
 * (src/cuna_kernel.cu:302-358 cunfft_adjoint_raw_batched_async,
 * src/gaussian_filter.cu:67-99 generate_pinned_filter_properties_batch).
 * Curve i has n[i] points of the concatenated x (and f_data; NULL = unit
 * weights) and ng[i] grid cells (any ng[i] >= 1, n[i] < 2 ng[i]); its
 * ng[i] outputs (exactly those of cunfft_adjoint_raw_async on that curve)
 * go to the start of slab i of f_hat, slab stride idist = max ng, f_hat
 * holding nlc * idist values; the rest of each slab is left as it was.
 * n, ng and gpu_fprops are HOST arrays; gpu_fprops[i] is curve i's device
 * filter_properties from generate_*filter_properties*.  When all of them
 * are such registered structs, the multi-curve plan (curves of equal ng
 * share one spread launch and one batched FFT) is cached until one of them
 * is freed (at most 16 batches; least recently used dropped first); a NULL
 * array or an unregistered struct rebuilds the plan on every call.  As with
 * the reference's E1/E2 arrays, the cached plan belongs to the x it was made
 * from: regenerate the filter properties when x changes.  Like a plan, a
 * cached batch executes on one stream at a time. */
void cunfft_adjoint_raw_batched_async(const dTyp *x, const dTyp *f_data, Complex *f_hat,
                                      const int *n, const int *ng, const int nlc,
                                      const filter_properties *const *gpu_fprops,
                                      hipStream_t stream);
/* generate_pinned_filter_properties for each curve of a batch: h_fprops[i]
 * (host struct with caller-allocated device E1/E2 of n[i] and E3 of R_f
 * entries) and d_fprops[i] (device struct) describe curve i, whose points
 * start at x + n[0] + ... + n[i-1]. */
void generate_pinned_filter_properties_batch(const dTyp *x, const int *n, const int *ng,
                                             const int nlc, filter_properties **h_fprops,
                                             filter_properties **d_fprops, hipStream_t stream);
/* Frees every cached multi-curve plan. */
void cuna_batch_cache_clear(void);

/* The FFT stage alone: in-place unnormalised DFT of `batch` sequences of n
 * values (any n >= 1: powers of two by the Stockham passes, other lengths by
 * Bluestein through them), data[r] = sum_g data[g] exp(sign*2 pi i g r/n).
 * sign = +1 is the reference's CUFFT_INVERSE. */
int cuna_fft(Complex *data, int n, int batch, int sign, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* CUNA_NFFT_H */
