/*
 * cuna_filter.h -- filter-property setup of the CUNA API, HIP-native.
 *
 * Replaces the host functions of inc/cuna_filter.h:34-57 of
 * johnh2o2/cunfft_adjoint.  The device kernels declared there
 * (filter, normalize, normalize_bootstrap: inc/cuna_filter.h:64-78) are not
 * host ABI and have no counterpart here.
 *
 * Besides filling filter_properties (b, normfac, filter_radius, E1, E2, E3
 * exactly as src/gaussian_filter.cu:204-228 defines them), each generate_*
 * call builds the library's per-x spread plan (points binned by grid tile) and
 * associates it with the device struct pointer; the raw transforms look it up
 * there.  A struct that was not produced by generate_* still works: the plan
 * is then rebuilt on each call.
 */
#ifndef CUNA_FILTER_H
#define CUNA_FILTER_H
#include "cuna_typedefs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* inc/cuna_filter.h:34-35, src/gaussian_filter.cu:168-177 */
void set_filter_properties(plan *p);

/* inc/cuna_filter.h:37-40, src/gaussian_filter.cu:43-66.  The caller owns
 * h_fprops (host) and d_fprops (device) and must have set h_fprops->E1, E2
 * (device, n values) and E3 (device, filter_radius() values).  Ordered on
 * stream, no host synchronisation. */
void generate_pinned_filter_properties(const dTyp *x, const int n, const int ng,
                                       filter_properties *h_fprops, filter_properties *d_fprops,
                                       hipStream_t stream);

/* inc/cuna_filter.h:47-49, src/gaussian_filter.cu:102-164.  Allocates the
 * host struct (*fprops_host), the device struct (*fprops_device) and E1/E2/E3,
 * computes them from device x and synchronises. */
void generate_filter_properties(const dTyp *x, const int n, const int ng,
                                filter_properties **fprops_host,
                                filter_properties **fprops_device);

/* inc/cuna_filter.h:55-57, src/gaussian_filter.cu:182-190.  Frees E1, E2, E3,
 * the device struct, the host struct and the associated spread plan. */
void free_filter_properties(filter_properties *d_fp, filter_properties *fp);

/* FILTER_RADIUS of this library (6 in libcunaf, 12 in libcunad:
 * src/gaussian_filter.cu:29-33).  Not exported by the reference, which
 * nevertheless requires callers of generate_pinned_filter_properties to
 * allocate E3 with that many values. */
int cuna_filter_radius(void);

#ifdef __cplusplus
}
#endif
#endif /* CUNA_FILTER_H */
