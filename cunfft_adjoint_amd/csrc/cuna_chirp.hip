// cuna_chirp.hip -- Bluestein (chirp-z) transform for lengths that are not
// powers of two, on top of the power-of-two Stockham FFT.
//
// The reference's cuFFT stage takes any ng (cufftPlan1d / cufftPlanMany,
// src/cuna_kernel.cu:82,169,208,261; its test program uses ng = R*N,
// test.cpp:114).  Here a real grid of ng cells is transformed through
//   A[r] = c_r * sum_g (z_g c_g) conj(c_{r-g}),  c_k = exp(+i pi k^2/ng),
// with the convolution done circularly at L = 2^lgL >= 2 ng - 1.  The chirp
// is tabulated on the host in long double from k^2 mod 2 ng (exact integer),
// so its phase carries no error that grows with k.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "cuna_plan.hpp"

namespace cuna {

// a[b][i] = c_i * z[b][i] for i < ng, 0 for ng <= i < L (z real, or complex
// and optionally conjugated)
template <typename T, class In>
__global__ void k_chirp_premul(const In *z, long zstride, int ng, const cplx<T> *chirp, long L,
                               int batch, int conj_in, cplx<T> *a) {
    for (int b = blockIdx.y; b < batch; b += gridDim.y) {
        for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < L;
             i += (long)gridDim.x * blockDim.x) {
            cplx<T> v = cmk(T(0), T(0));
            if (i < ng) {
                if constexpr (std::is_same<In, T>::value) {
                    v = chirp[i] * z[b * zstride + i];
                } else {
                    cplx<T> w = z[b * zstride + i];
                    if (conj_in) w.y = -w.y;
                    v = chirp[i] * w;
                }
            }
            a[(long)b * L + i] = v;
        }
    }
}

// a[b][k] *= bhat[k] * scale
template <typename T>
__global__ void k_chirp_mul(cplx<T> *a, const cplx<T> *bhat, long L, int batch, T scale) {
    for (int b = blockIdx.y; b < batch; b += gridDim.y) {
        for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < L;
             k += (long)gridDim.x * blockDim.x) {
            a[(long)b * L + k] = a[(long)b * L + k] * bhat[k] * scale;
        }
    }
}

// compat sink: out row b, r < ng: c_r conv[b][r] * normfac exp((r/ng)^2 b/4)
template <typename T>
__global__ void k_chirp_post_compat(const cplx<T> *conv, long L, int ng, const cplx<T> *chirp,
                                    int batch, SinkCompat<T> sk) {
    for (int b = blockIdx.y; b < batch; b += gridDim.y) {
        const SinkCompat<T> skb = sk.for_row(b);
        cplx<T> *o = sk.row_out(b);
        for (long r = (long)blockIdx.x * blockDim.x + threadIdx.x; r < ng;
             r += (long)gridDim.x * blockDim.x)
            o[r] = chirp[r] * conv[(long)b * L + r] * skb.fac(r);
    }
}

// plain output, in place: data[b][r] = c_r conv[b][r] (conjugated on request)
template <typename T>
__global__ void k_chirp_post_plain(const cplx<T> *conv, long L, int ng, const cplx<T> *chirp,
                                   int batch, int conj_out, cplx<T> *data) {
    for (int b = blockIdx.y; b < batch; b += gridDim.y) {
        for (long r = (long)blockIdx.x * blockDim.x + threadIdx.x; r < ng;
             r += (long)gridDim.x * blockDim.x) {
            cplx<T> v = chirp[r] * conv[(long)b * L + r];
            if (conj_out) v.y = -v.y;
            data[(long)b * ng + r] = v;
        }
    }
}

static dim3 elem_grid(long n, int batch) {
    const long bx = std::min<long>(ceil_div(std::max(n, 1L), 256), 4096);
    return dim3((unsigned)bx, (unsigned)std::min(std::max(batch, 1), 65535));
}

template <typename T>
void Bluestein<T>::init(int ng_, hipStream_t s) {
    require(ng_ >= 1, "transform length must be positive");
    ng = ng_;
    lgL = std::max(1, ilog2(2L * ng - 1));
    L = 1L << lgL;
    require(lgL <= max_fft_lg<T>(), "transform length too large", CUNA_ERROR_UNSUPPORTED);
    const long double pi = 3.141592653589793238462643383279502884L;
    chirp_h.resize(ng);
    for (long k = 0; k < ng; ++k) {
        const unsigned long long q = ((unsigned long long)k * (unsigned long long)k) %
                                     (2ull * (unsigned long long)ng);
        const long double a = pi * (long double)q / (long double)ng;
        chirp_h[k] = cmk((T)cosl(a), (T)sinl(a));
    }
    b_h.assign(L, cmk(T(0), T(0)));
    for (long k = 0; k < ng; ++k) {
        const cplx<T> v = conjg(chirp_h[k]);
        b_h[k] = v;
        if (k > 0) b_h[L - k] = v;
    }
    chirp.alloc(sizeof(cplx<T>) * ng);
    bhat.alloc(sizeof(cplx<T>) * L);
    CUNA_HIP(hipMemcpyAsync(chirp.ptr, chirp_h.data(), sizeof(cplx<T>) * ng,
                            hipMemcpyHostToDevice, s));
    CUNA_HIP(hipMemcpyAsync(bhat.ptr, b_h.data(), sizeof(cplx<T>) * L, hipMemcpyHostToDevice, s));
    TwiddleCache<T>::get().warm(lgL);
    ensure(1);
    FftShape sh;
    sh.lgN = lgL;
    sh.batch = 1;
    fft_complex<T>(sh, bhat.as<cplx<T>>(), work.as<cplx<T>>() + L, -1, s);
}

template <typename T>
void Bluestein<T>::ensure(int batch) {
    if (batch <= batch_cap) return;
    work.alloc(sizeof(cplx<T>) * (size_t)L * 2 * (size_t)batch);
    batch_cap = batch;
}

template <typename T>
void Bluestein<T>::run_real(const T *grid, long gstride, int batch, const SinkCompat<T> &sk,
                            hipStream_t s) {
    require(batch <= batch_cap, "Bluestein workspace too small for this batch");
    cplx<T> *a = work.as<cplx<T>>();
    cplx<T> *tmp = a + (size_t)L * batch;
    hipLaunchKernelGGL((k_chirp_premul<T, T>), elem_grid(L, batch), dim3(256), 0, s, grid,
                       gstride, ng, chirp.as<cplx<T>>(), L, batch, 0, a);
    CUNA_LAUNCH_CHECK();
    FftShape sh;
    sh.lgN = lgL;
    sh.batch = batch;
    fft_complex<T>(sh, a, tmp, -1, s);
    hipLaunchKernelGGL(k_chirp_mul<T>, elem_grid(L, batch), dim3(256), 0, s, a,
                       bhat.as<cplx<T>>(), L, batch, T(1) / (T)L);
    CUNA_LAUNCH_CHECK();
    fft_complex<T>(sh, a, tmp, 1, s);
    hipLaunchKernelGGL(k_chirp_post_compat<T>, elem_grid(ng, batch), dim3(256), 0, s, a, L, ng,
                       chirp.as<cplx<T>>(), batch, sk);
    CUNA_LAUNCH_CHECK();
}

template <typename T>
void Bluestein<T>::run_complex(cplx<T> *data, int batch, int sign, hipStream_t s) {
    require(batch <= batch_cap, "Bluestein workspace too small for this batch");
    // sign -1 as the conjugate of the sign +1 transform of the conjugate
    const int cj = sign < 0 ? 1 : 0;
    cplx<T> *a = work.as<cplx<T>>();
    cplx<T> *tmp = a + (size_t)L * batch;
    hipLaunchKernelGGL((k_chirp_premul<T, cplx<T>>), elem_grid(L, batch), dim3(256), 0, s, data,
                       (long)ng, ng, chirp.as<cplx<T>>(), L, batch, cj, a);
    CUNA_LAUNCH_CHECK();
    FftShape sh;
    sh.lgN = lgL;
    sh.batch = batch;
    fft_complex<T>(sh, a, tmp, -1, s);
    hipLaunchKernelGGL(k_chirp_mul<T>, elem_grid(L, batch), dim3(256), 0, s, a,
                       bhat.as<cplx<T>>(), L, batch, T(1) / (T)L);
    CUNA_LAUNCH_CHECK();
    fft_complex<T>(sh, a, tmp, 1, s);
    hipLaunchKernelGGL(k_chirp_post_plain<T>, elem_grid(ng, batch), dim3(256), 0, s, a, L, ng,
                       chirp.as<cplx<T>>(), batch, cj, data);
    CUNA_LAUNCH_CHECK();
}

template struct Bluestein<dTyp>;

}  // namespace cuna
