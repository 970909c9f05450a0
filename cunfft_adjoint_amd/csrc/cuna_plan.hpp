// cuna_plan.hpp -- per-x plan: binned points, filter constants, workspaces.
//
// A plan is the MI355X counterpart of the reference's filter_properties +
// plan pair (inc/cuna_typedefs.h:92-148, src/gaussian_filter.cu:102-164): it
// is built once per set of sample positions and reused for every y.
#pragma once

#include <memory>
#include <vector>

#include "cuna_launch.hpp"

namespace cuna {

// Persistent device allocation (hipMalloc / hipFree).
struct DevMem {
    void *ptr = nullptr;
    size_t bytes = 0;
    bool owned = true;  // false: a view into another allocation
    DevMem() = default;
    DevMem(const DevMem &) = delete;
    DevMem &operator=(const DevMem &) = delete;
    void alloc(size_t nbytes) {
        release();
        if (nbytes) CUNA_HIP(hipMalloc(&ptr, nbytes));
        bytes = nbytes;
    }
    void view(void *p, size_t nbytes) {
        release();
        ptr = p;
        bytes = nbytes;
        owned = false;
    }
    void ensure(size_t nbytes) {
        if (nbytes > bytes) alloc(nbytes);
    }
    void release() {
        if (ptr && owned) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        owned = true;
    }
    ~DevMem() { release(); }
    template <typename U>
    U *as() const {
        return static_cast<U *>(ptr);
    }
};

// Points of one or more curves binned by grid tile (stable counting sort).
template <typename T>
struct TiledPoints {
    // host sources of the async uploads (offsets and curve order; the sort's
    // chunk tables), declared first so that they are destroyed after the
    // device buffers (whose hipFree synchronises)
    std::vector<int> plan_h, sort_h;
    int ncurves = 0;
    long n = 0;
    std::vector<int> coff_h;
    int G = 0, tlog = 0, ntiles = 0, Kc = 0, olo = 0, ntaps = 0;
    bool periodic = true;
    DevMem arena;  // one allocation; the buffers below are views into it
    DevMem coff, rec, perm, bucket, eps_orig;
    DevMem rec_orig;  // (anchor, e) by original index
    DevMem order;     // curves by decreasing length (work-queue order)
    DevMem tile_cover;  // per (curve, tile): most points in one window of ntaps anchors
    DevMem tile_desc;   // per curve: the scatter's work order (TileDesc per tile)

    // src: setup arguments with x/t/mode/compat fields filled
    void build(SetupArgs<T> A, const std::vector<int> &offsets, int G_, int tlog_, int ntaps_,
               bool periodic_, bool keep_eps, hipStream_t s);
};

struct ModeSpec {
    int mode = 0;  // CUNA_MODE_*
    int sigma = 2, m = 8, tile_log2 = 0, flags = 0;
    int fft_w = 0, fft_h = 0;  // FFT workgroup shape (0: heuristic)
};

// Bluestein (chirp-z) transform of length ng, for grids whose length is not
// a power of two >= 4 (the reference hands any ng to cuFFT,
// src/cuna_kernel.cu:82,169,208,261):
//   A[r] = sum_g z_g e^{+2 pi i g r/ng} = c_r sum_g (z_g c_g) conj(c_{r-g}),
//   c_k = exp(+i pi k^2/ng),
// a circular convolution of length L = 2^lgL >= 2 ng - 1 done with the
// power-of-two FFT: conv = FFT_+(FFT_-(z c) * FFT_-(b)) / L, with
// b_k = b_{L-k} = conj(c_k) for 0 <= k < ng and zero between.
template <typename T>
struct Bluestein {
    // host sources of the (async) uploads, declared first so that they are
    // destroyed after the device buffers (whose hipFree synchronises)
    std::vector<cplx<T>> chirp_h, b_h;
    int ng = 0, lgL = 0;
    long L = 0;
    DevMem chirp;  // c_k, k < ng
    DevMem bhat;   // FFT_-(b), L values
    DevMem work;   // batch_cap x L complex for the data and as much for the FFT passes
    int batch_cap = 0;
    void init(int ng_, hipStream_t s);
    void ensure(int batch);
    // real grids (batch rows of stride gstride, the first ng cells used) ->
    // the compat deconvolution sink, one row per grid
    void run_real(const T *grid, long gstride, int batch, const SinkCompat<T> &sk,
                  hipStream_t s);
    // in place on batch x ng complex values: data[r] = sum_g data[g] e^{sign 2 pi i g r/ng}
    void run_complex(cplx<T> *data, int batch, int sign, hipStream_t s);
};

// true when a real grid of G cells takes the power-of-two FFT (else Bluestein)
inline bool pow2_grid(long G) { return G >= 4 && is_pow2(G); }

template <typename T>
struct Plan {
    ModeSpec spec;
    int M = 0;  // accurate: modes per curve; compat: ng
    int G = 0, lgN = 0, ntaps = 0, olo = 0;
    T beta = 0, wscale = 1;
    T e3[kMaxTaps] = {};
    // compat constants (src/gaussian_filter.cu:54-57)
    T b_c = 0, normfac_c = 0;
    int R_f = 0;
    // compat plans over several curves with their own b (device arrays, set
    // by the batched transform): per curve (1/b, 1/sqrt(pi b), exp(-1/b)),
    // b/4, and the output offset of the curve's spectrum
    const T *cpar_c = nullptr;
    const T *qb_row_c = nullptr;
    const long *row_off_c = nullptr;
    // kept to identify the x a registered plan belongs to
    const void *x_key = nullptr;
    TiledPoints<T> pts, pts2;  // pts2: points 2x (window at 2f)
    bool has_pts2 = false;
    DevMem grid, work, counts, counter;
    std::unique_ptr<Bluestein<T>> blu;  // compat grids of any other length
    // fused bootstrap: per-curve max |y| and densest tap window of x and 2x
    DevMem ymax, cover;
    DevMem ls_work;  // Lomb-Scargle: window spectra of the resident workgroups
    bool cover_ready = false;
    // optional profiling events, recorded on the execution stream before the
    // first kernel of adjoint() and after each of its kernels
    hipEvent_t prof[4] = {nullptr, nullptr, nullptr, nullptr};

    long n() const { return pts.n; }
    int ncurves() const { return pts.ncurves; }

    // y -> out (ncurves x M complex)
    void adjoint(const T *y, cplx<T> *out, hipStream_t s);
    // the same with the real grid (ncurves x G) in caller-provided memory;
    // allocates nothing
    void adjoint_into(const T *y, cplx<T> *out, T *grid_ptr, hipStream_t s) const;
    // nrhs value vectors (rows of n) -> out (ncurves x nrhs x M), one spread
    // launch and one batched FFT
    void adjoint_many(const T *y, int nrhs, cplx<T> *out, hipStream_t s);
    // compat bootstrap of src/cuna_kernel.cu:236-289 (single curve)
    void bootstrap_compat(const T *y, cplx<T> *out, int nbs, uint32_t seed, hipStream_t s);
    // accurate bootstrap (resampled points)
    void bootstrap(const T *y, int nbs, uint32_t seed, cplx<T> *out_y, cplx<T> *out_w,
                   hipStream_t s);
    // Lomb-Scargle power of each replicate (fused kernel sizes only)
    void bootstrap_ls(const T *y, int nbs, uint32_t seed, T *power, T *peak, int *kpeak,
                      hipStream_t s);
    void backward(const T *y, const cplx<T> *g, T *dy, T *dx, hipStream_t s);
    // forward (type-2) NFFT: fhat (ncurves x M) -> f (n complex)
    void forward(const cplx<T> *fhat, cplx<T> *f, hipStream_t s);
    // allocate every workspace a later call with these sizes needs (so that
    // the call can be captured into a graph without allocating)
    void prepare_bootstrap(int nbs, bool window);
    bool bootstrap_fused_ok(bool window) const;

    SpreadArgs<T> spread_args(const TiledPoints<T> &P) const;
    BootArgs<T> fused_args(const T *y, int nbs, uint32_t seed, hipStream_t s);
    void run_fft(T *grid_ptr, cplx<T> *out, int batch, hipStream_t s,
                 hipEvent_t mid = nullptr) const;
    // twiddle tables (and the Bluestein set-up for other lengths) for this
    // plan's transform; run once at plan creation
    void prepare_fft(hipStream_t s);
};

template <typename T>
std::unique_ptr<Plan<T>> make_plan_accurate(const T *x, const double *t, double df,
                                            const std::vector<int> &offsets, int M,
                                            const ModeSpec &spec, hipStream_t s);
template <typename T>
std::unique_ptr<Plan<T>> make_plan_compat(const T *x, int n, int ng, T *E1, T *E2, T *E3,
                                          const ModeSpec &spec, hipStream_t s);
// b and normfac of the compat filter in the build's precision
template <typename T>
void compat_constants(int n, int ng, T &b, T &normfac, int &R_f);

}  // namespace cuna
