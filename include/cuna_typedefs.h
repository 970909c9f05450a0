/*
 * cuna_typedefs.h -- types of the CUNA C API, HIP-native.
 *
 * Replaces inc/cuna_typedefs.h of johnh2o2/cunfft_adjoint.  Field order and
 * values are kept so that code written against the reference compiles
 * unchanged (a source-level drop-in):
 *   dTyp / Complex / cTyp        inc/cuna_typedefs.h:38-60
 *   filter_properties            inc/cuna_typedefs.h:92-108
 *   FLAGS                        inc/cuna_typedefs.h:110-116
 *   plan                         inc/cuna_typedefs.h:119-148
 *   init_plan / free_plan        inc/cuna_typedefs.h:151-155
 *
 * Define DOUBLE_PRECISION before including (as with the reference) to use
 * libcunad.so; otherwise the header describes libcunaf.so.
 * No CUDA headers are needed; only hipStream_t comes from HIP.
 */
#ifndef CUNA_TYPEDEFS_H
#define CUNA_TYPEDEFS_H

#include <stdio.h>
#include <stdlib.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
#include <complex>
#else
#include <complex.h>
#endif

/* interleaved complex {x, y}: same layout as float2 / double2,
 * cufftComplex / cufftDoubleComplex, std::complex and C99 complex */
typedef struct { float x, y; } cunaFloatComplex;
typedef struct { double x, y; } cunaDoubleComplex;

#ifdef DOUBLE_PRECISION
#define dTyp double
#define Complex cunaDoubleComplex
#ifdef __cplusplus
#define cTyp std::complex<double>
#else
#define cTyp double complex
#endif
#else
#define dTyp float
#define Complex cunaFloatComplex
#ifdef __cplusplus
#define cTyp std::complex<float>
#else
#define cTyp float complex
#endif
#endif

#define PI 3.14159265358979323846

/* Gaussian filter constants and (device) precomputed arrays for one x:
 * E1[i] = exp(-eps_i^2/b)/sqrt(pi b), E2[i] = exp(2 eps_i/b), E3[m] = exp(-m^2/b) */
typedef struct {
    dTyp b, normfac;
    int filter_radius;
    dTyp *E1;
    dTyp *E2;
    dTyp *E3;
} filter_properties;

typedef enum {
    NO_FLAGS = 0x01,
    OUTPUT_INTERMEDIATE = 0x02,
    CALCULATE_WINDOW_FUNCTION = 0x04,
    DONT_TRANSFER_TO_CPU = 0x08,
    PRINT_TIMING = 0x10
} FLAGS;

typedef struct {
    /* host copies */
    dTyp *x_data, *f_data;
    Complex *f_hat, *f_hat_win;
    /* device buffers */
    Complex *g_f_hat, *g_f_hat_win;
    dTyp *g_x_data, *g_f_data;
    unsigned int flags;
    int Ndata, Ngrid, filter_radius;
    /* host and device copies of the filter constants for x */
    filter_properties *fprops_host, *fprops_device;
} plan;

#ifdef __cplusplus
extern "C" {
#endif

/* inc/cuna_typedefs.h:151-152, src/utils.cu:247-278: frees every buffer of p and p itself */
void free_plan(plan *p);

/* inc/cuna_typedefs.h:154-155, src/utils.cu:95-178: copies x, f to host and device
 * buffers, zeroes the device grids and computes the filter properties */
void init_plan(plan *p, const dTyp *x, const dTyp *f, int Ndata, int Ngrid,
               unsigned int plan_flags);

/* ---- error reporting (no counterpart in the reference, which calls exit()) ----
 * Every entry point records its outcome in a thread-local status.
 * cuna_set_abort_on_error(1) restores the reference's behaviour of printing
 * the error and terminating the process. */
enum {
    CUNA_SUCCESS = 0,
    CUNA_ERROR_INVALID_VALUE = 1,
    CUNA_ERROR_UNSUPPORTED = 2,
    CUNA_ERROR_HIP = 3,
    CUNA_ERROR_OUT_OF_MEMORY = 4,
    CUNA_ERROR_NOT_FOUND = 5
};
int cuna_get_last_error(void);
const char *cuna_get_last_error_string(void);
void cuna_clear_last_error(void);
void cuna_set_abort_on_error(int enable);
/* 4 for libcunaf.so, 8 for libcunad.so */
int cuna_real_size(void);

#ifdef __cplusplus
}
#endif

#endif /* CUNA_TYPEDEFS_H */
