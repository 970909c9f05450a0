// cuna_fft.hpp -- hand-written Stockham FFT for gfx950, staged in LDS.
//
// Replaces the cuFFT stage of the reference (cufftExec{C2C,Z2Z} with
// CUFFT_INVERSE, unnormalised, in place: src/cuna_kernel.cu:43-49,80-99,
// 167-178,206-220,257-276).  The grid of the adjoint NFFT is real, so a real
// grid of G cells is transformed as a complex sequence of N = G/2 values
// z[n] = grid[2n] + i grid[2n+1] (the same memory), and the spectrum is
// recovered pairwise from Z[r], Z[N-r]; the window deconvolution is applied
// in that same epilogue ("fused into the last pass").
//
// Sizes: N <= 2^14 (float) / 2^13 (double) run as one workgroup per
// transform entirely in LDS; larger N = L1*L2 run as two passes (column
// FFTs of length L2 with twiddles, then row FFTs of length L1 whose rows are
// processed in mirror pairs so that the real post-processing stays local).
#pragma once

#include <type_traits>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "cuna_common.hpp"

namespace cuna {

extern __shared__ __attribute__((aligned(16))) unsigned char cuna_smem[];

// ---------------------------------------------------------- butterflies --
// In-register DFTs of radix 2/4/8/16, v[n] -> v[k] = sum_n v[n] w_R^{S n k}.
template <typename T>
__device__ __forceinline__ cplx<T> rot(cplx<T> z, T c, T s) {  // z * (c + i s)
    return cmk(z.x * c - z.y * s, z.x * s + z.y * c);
}
template <typename T, int S>
__device__ __forceinline__ cplx<T> mul_si(cplx<T> z) {  // z * (S i)
    return S > 0 ? cmk(-z.y, z.x) : cmk(z.y, -z.x);
}

template <typename T, int S>
__device__ __forceinline__ void dft2(cplx<T> *v) {
    cplx<T> t = v[0] - v[1];
    v[0] = v[0] + v[1];
    v[1] = t;
}

template <typename T, int S>
__device__ __forceinline__ void dft4(cplx<T> &a0, cplx<T> &a1, cplx<T> &a2, cplx<T> &a3) {
    cplx<T> t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = mul_si<T, S>(a1 - a3);
    a0 = t0 + t2;
    a2 = t0 - t2;
    a1 = t1 + t3;
    a3 = t1 - t3;
}

template <typename T, int S>
__device__ __forceinline__ void dft8(cplx<T> *v) {
    const T h = T(0.70710678118654752440);
    cplx<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    cplx<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<T, S>(e0, e1, e2, e3);
    dft4<T, S>(o0, o1, o2, o3);
    o1 = rot(o1, h, S * h);
    o2 = mul_si<T, S>(o2);
    o3 = rot(o3, -h, S * h);
    v[0] = e0 + o0;
    v[4] = e0 - o0;
    v[1] = e1 + o1;
    v[5] = e1 - o1;
    v[2] = e2 + o2;
    v[6] = e2 - o2;
    v[3] = e3 + o3;
    v[7] = e3 - o3;
}

template <typename T, int S>
__device__ __forceinline__ void dft16(cplx<T> *v) {
    const T c1 = T(0.92387953251128675613), s1 = T(0.38268343236508977173);
    const T h = T(0.70710678118654752440);
    cplx<T> e[8], o[8];
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        e[n] = v[2 * n];
        o[n] = v[2 * n + 1];
    }
    dft8<T, S>(e);
    dft8<T, S>(o);
    o[1] = rot(o[1], c1, S * s1);
    o[2] = rot(o[2], h, S * h);
    o[3] = rot(o[3], s1, S * c1);
    o[4] = mul_si<T, S>(o[4]);
    o[5] = rot(o[5], -s1, S * c1);
    o[6] = rot(o[6], -h, S * h);
    o[7] = rot(o[7], -c1, S * s1);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        v[k] = e[k] + o[k];
        v[k + 8] = e[k] - o[k];
    }
}

template <typename T, int R, int S>
__device__ __forceinline__ void dft(cplx<T> *v) {
    if constexpr (R == 2) dft2<T, S>(v);
    if constexpr (R == 4) dft4<T, S>(v[0], v[1], v[2], v[3]);
    if constexpr (R == 8) dft8<T, S>(v);
    if constexpr (R == 16) dft16<T, S>(v);
}

// ------------------------------------------------------------ radix plan --
// log2 of the radix of pass p for a length-2^lg transform with radices up to
// 2^mr (mr = 4: radix 16, mr = 3: radix 8): full-radix passes, with the
// remainder taken by one smaller pass, or, for lg = k mr + 1, by two
// (radix 8 and 4 after radix-16 passes, 4 and 4 after radix-8 passes).
constexpr int num_passes(int lg, int mr = 4) {
    if (lg <= 0) return 0;
    if (lg <= mr) return 1;
    return lg % mr == 0 ? lg / mr : lg / mr + 1;
}
constexpr int pass_log2(int lg, int p, int mr = 4) {
    if (lg <= mr) return lg;
    const int n = lg / mr, rem = lg % mr;
    if (rem == 0) return mr;
    if (rem == 1) {
        if (p < n - 1) return mr;
        return p == n - 1 ? (mr + 2) / 2 : (mr + 1) / 2;
    }
    return p < n ? mr : rem;
}

// ------------------------------------------------------------ LDS layout --
// Element a of the LDS FFT buffer is stored at lds_sw(a): bits 0-3 are XORed
// with bits 4-7 and bit 4 with bit 8.  The map is a bijection within every
// aligned block of 32 elements, and it makes both the unit-stride reads and
// the stride-16 writes of a radix-16 pass (8- and 16-byte elements, 32 or 64
// banks) conflict-free; unswizzled, those writes hit 2 of the banks per
// 64-byte group.
__device__ __forceinline__ int lds_sw(int a) { return a ^ ((a >> 4) & 31); }

// ------------------------------------------------------ twiddle tables --
// Two-level table for exp(+2 pi i e / Nfull), e < Nfull:
//   lo[e & (B-1)] * hi[e >> log2 B], lo[q] = w^q, hi[q] = w^(q B).
template <typename T>
struct Tw2 {
    const cplx<T> *lo;
    const cplx<T> *hi;
    int lb;  // log2 B
    __device__ __forceinline__ cplx<T> operator()(long e) const {
        return lo[e & ((1L << lb) - 1)] * hi[e >> lb];
    }
    __device__ __forceinline__ cplx<T> operator[](int e) const {
        return lo[e & ((1 << lb) - 1)] * hi[e >> lb];
    }
    // entries of the two tables
    __host__ __device__ int n_lo() const { return 1 << lb; }
};

// v[r] *= w^(r k 2^SHIFT), r = 1 .. R-1 (conjugated for S = -1).  From the
// two-level table (two LDS reads per entry) only w^(k 2^SHIFT) is read; the
// other powers are products of depth <= log2 R: squarings for the powers of
// two, and w^r = w^h w^(r-h) with h the largest power of two below r.  From
// a plain table, or with POWERS off (fewer registers: several butterflies
// per thread), every power is one lookup.
template <typename T, int S, int R, int SHIFT, bool POWERS, class TW>
__device__ __forceinline__ void apply_twiddles(cplx<T> *v, const TW &tw, int k) {
    if constexpr (POWERS && std::is_same<TW, Tw2<T>>::value) {
        cplx<T> w[R];
        w[1] = tw[k << SHIFT];
        if constexpr (S < 0) w[1].y = -w[1].y;
#pragma unroll
        for (int r = 2; r < R; ++r) {
            const int h = 1 << (31 - __builtin_clz(r));
            w[r] = h == r ? w[r / 2] * w[r / 2] : w[h] * w[r - h];
        }
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = v[r] * w[r];
    } else {
#pragma unroll
        for (int r = 1; r < R; ++r) {
            cplx<T> t = tw[(k * r) << SHIFT];
            if constexpr (S < 0) t.y = -t.y;
            v[r] = v[r] * t;
        }
    }
}

// Workgroup barrier that orders LDS only: each wave waits for its own LDS
// operations, not for its outstanding global loads.  The FFT's passes
// exchange data through LDS alone, so global loads issued earlier (the
// epilogue's twiddles) keep flying across the passes.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ----------------------------------------------------------- LDS Stockham --
// W interleaved sequences of length L = 2^LG: element e of sequence w lives at
// lds[lds_sw(e*W + w)].  Pass 0 reads its inputs through `ld(e, w)`; every pass writes
// LDS; the routine returns with natural-order results in LDS after a barrier.
// tw[e] = exp(+2 pi i e / L) (conjugated for S = -1).
struct NoHook {
    __device__ __forceinline__ void operator()() const {}
};

// `hook` runs once, after pass 0 has issued its loads and before it writes
// LDS (used to stage a twiddle table whose loads were issued earlier).
// `tw` is a table pointer or any object with operator[](int) giving w_L^e.
template <typename T, int S, int LG, int W, int NT, int P, int LOGNS, class Loader,
          bool INPLACE = false, class Hook = NoHook, class TW = const cplx<T> *, int MR = 4>
__device__ __forceinline__ void fft_pass(cplx<T> *lds, TW tw, const Loader &ld,
                                         const Hook &hook = Hook()) {
    constexpr int NP = num_passes(LG, MR);
    if constexpr (P < NP) {
        constexpr int LR = pass_log2(LG, P, MR);
        constexpr int R = 1 << LR;
        constexpr int L = 1 << LG;
        constexpr int Q = L >> LR;  // butterflies per sequence
        constexpr int NB = W * Q;
        constexpr int BPT = (NB + NT - 1) / NT;
        constexpr bool EXACT = (NB % NT) == 0;
        cplx<T> v[BPT][R];
        // opaque thread id: in a persistent loop the compiler would otherwise
        // hoist every pass's LDS addresses and twiddles out of the loop and
        // keep them all live (spilling at 1024 threads)
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
#pragma unroll
        for (int q = 0; q < BPT; ++q) {
            const int g = tid + q * NT;
            if (EXACT || g < NB) {
                const int w = g % W, j = g / W;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int e = j + r * Q;
                    if constexpr (P == 0)
                        v[q][r] = ld(e, w);
                    else
                        v[q][r] = lds[lds_sw(e * W + w)];
                }
                // twiddles are global loads: issue them before the barrier
                // (with several butterflies per thread that would hold all
                // their twiddles in registers at once; they follow it then)
                if constexpr (P > 0 && BPT == 1)
                    apply_twiddles<T, S, R, LG - LOGNS - LR, true>(v[q], tw,
                                                                  j & ((1 << LOGNS) - 1));
            }
        }
        if constexpr (P == 0) hook();
        // pass 0 reads through the loader; when that reads this same LDS
        // buffer (INPLACE) all reads must finish before the first write
        if constexpr (P > 0 || INPLACE) lds_barrier();
#pragma unroll
        for (int q = 0; q < BPT; ++q) {
            const int g = tid + q * NT;
            if (EXACT || g < NB) {
                const int w = g % W, j = g / W;
                const int k = j & ((1 << LOGNS) - 1);
                if constexpr (P > 0 && BPT > 1)
                    apply_twiddles<T, S, R, LG - LOGNS - LR, (R <= 4)>(v[q], tw, k);
                dft<T, R, S>(v[q]);
                const int base = ((j >> LOGNS) << (LOGNS + LR)) + k;
#pragma unroll
                for (int r = 0; r < R; ++r) lds[lds_sw((base + (r << LOGNS)) * W + w)] = v[q][r];
            }
        }
        lds_barrier();
        fft_pass<T, S, LG, W, NT, P + 1, LOGNS + LR, Loader, INPLACE, Hook, TW, MR>(lds, tw, ld,
                                                                                   hook);
    }
}

// MR: log2 of the largest radix (4: 16 values per butterfly; 3: 8, so that
// twice the threads share the work)
template <typename T, int S, int LG, int W, int NT, int MR = 4, class Loader, class Hook = NoHook>
__device__ __forceinline__ void lds_fft(cplx<T> *lds, const cplx<T> *__restrict__ tw,
                                        const Loader &ld, const Hook &hook = Hook()) {
    fft_pass<T, S, LG, W, NT, 0, 0, Loader, false, Hook, const cplx<T> *, MR>(lds, tw, ld, hook);
}

// A twiddle table of L entries copied into LDS: load() issues the global
// loads early, store() (run as the pass-0 hook) writes them, so the passes
// read their twiddles from LDS instead of waiting on L2 after each barrier.
template <typename T, int L, int NT>
struct TableStage {
    static constexpr int K = (L + NT - 1) / NT;
    // real and imaginary parts kept as scalars (an array of the complex
    // struct stayed in scratch memory in the FP64 build)
    T vx[K], vy[K];
    __device__ __forceinline__ void load(const cplx<T> *__restrict__ g, int tid) {
        // unconditional (clamped) loads: no branch, so the values stay in
        // registers and the loads issue together
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = tid + k * NT;
            const cplx<T> w = g[(L % NT == 0 || i < L) ? i : L - 1];
            vx[k] = w.x;
            vy[k] = w.y;
        }
    }
    __device__ __forceinline__ void store(cplx<T> *dst, int tid) const {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int i = tid + k * NT;
            if (L % NT == 0 || i < L) dst[i] = cmk(vx[k], vy[k]);
        }
    }
};

// LDS twiddle tables pay off where a pass's table is small next to its data
constexpr bool lds_table(int lg) { return lg <= 12; }

// the data is already in `lds` in natural (unswizzled) order; transformed in
// place, the result is in the swizzled layout
template <typename T, int S, int LG, int NT, class TW>
__device__ __forceinline__ void lds_fft_inplace(cplx<T> *lds, TW tw) {
    auto fetch = [&](int e, int) __attribute__((always_inline)) { return lds[e]; };
    fft_pass<T, S, LG, 1, NT, 0, 0, decltype(fetch), true, NoHook, TW>(lds, tw, fetch);
}


// Device twiddle tables, generated in long double on the host once per size
// and kept for the life of the process (a few KB each).
template <typename T>
class TwiddleCache {
  public:
    static TwiddleCache &get() {
        static TwiddleCache c;
        return c;
    }
    // exp(+2 pi i e / L) for e in [0, L)
    const cplx<T> *fine(int lg) { return table(lg, 0, 1L << lg); }
    Tw2<T> two_level(int lg) {
        int lb = (lg + 1) / 2;
        Tw2<T> t;
        t.lo = table(lg, 1, 1L << lb);          // w_N^q, q < B
        t.hi = table(lg - lb, 0, 1L << (lg - lb));  // w_{N/B}^q
        t.lb = lb;
        return t;
    }

    // Every table an FFT of N = 2^lgN complex values (and its real-input
    // epilogue, G = 2N) may use, on the current device.
    void warm(int lgN) {
        for (int l = 1; l <= std::min(lgN, 14); ++l) fine(l);
        two_level(lgN);
        two_level(lgN + 1);
    }
    // While a NoAlloc guard lives on this thread, a table not yet cached is
    // an error instead of a new allocation.
    struct NoAlloc {
        NoAlloc() { flag() = true; }
        ~NoAlloc() { flag() = false; }
    };

  private:
    static bool &flag() {
        static thread_local bool f = false;
        return f;
    }
    std::mutex mu_;
    std::map<std::tuple<int, int, int>, void *> tabs_;
    const cplx<T> *table(int lg, int kind, long count) {
        int dev = 0;
        CUNA_HIP(hipGetDevice(&dev));
        std::lock_guard<std::mutex> lk(mu_);
        auto key = std::make_tuple(dev, lg, kind);
        auto it = tabs_.find(key);
        if (it != tabs_.end()) return static_cast<const cplx<T> *>(it->second);
        require(!flag(), "twiddle table not prepared for this transform size",
                CUNA_ERROR_INVALID_VALUE);
        std::vector<cplx<T>> h(count);
        const long double two_pi = 6.283185307179586476925286766559005768L;
        long N = 1L << lg;
        for (long e = 0; e < count; ++e) {
            long double a = two_pi * (long double)e / (long double)N;
            h[e].x = (T)cosl(a);
            h[e].y = (T)sinl(a);
        }
        void *d = nullptr;
        CUNA_HIP(hipMalloc(&d, sizeof(cplx<T>) * count));
        CUNA_HIP(hipMemcpy(d, h.data(), sizeof(cplx<T>) * count, hipMemcpyHostToDevice));
        tabs_[key] = d;
        return static_cast<const cplx<T> *>(d);
    }
};

// ------------------------------------------------------------- sinks --
// A sink receives one spectral value A[f] of the real grid (f in [0, N]) and
// writes it, and its conjugate image A[G-f], where the output wants them.

// Accurate mode: out[M/2 + k], k = -M/2 .. M/2-1, deconvolved by
// D(k) = exp(pi^2 beta k^2/G^2)/sqrt(pi beta).
template <typename T>
struct SinkAccurate {
    cplx<T> *out;
    long M;
    long ostride;  // per batch
    T invG, coef, pref;
    // the sink of batch row b and where that row's output starts
    __device__ __forceinline__ const SinkAccurate &for_row(long) const { return *this; }
    __device__ __forceinline__ cplx<T> *row_out(long b) const { return out + b * ostride; }
    __device__ __forceinline__ void operator()(cplx<T> *o, long f, cplx<T> A) const {
        if (f <= M / 2) {
            T kk = (T)f * invG;
            T D = pref * dexp(coef * kk * kk);
            if (f < M / 2) o[M / 2 + f] = A * D;
            if (f > 0) o[M / 2 - f] = conjg(A) * D;
        }
    }
};

// Compat mode (src/gaussian_filter.cu:278-289): out[r] = A[r] * normfac *
// exp((r/ng)^2 * b/4) for every r in [0, ng).
template <typename T>
struct SinkCompat {
    cplx<T> *out;
    long G;
    long ostride;
    T invG, qb, normfac;  // qb = b/4
    // batches of curves with their own b (each nullable): output offset and
    // b/4 of batch row b
    const long *row_off;
    const T *qb_row;
    __device__ __forceinline__ SinkCompat for_row(long b) const {
        SinkCompat r = *this;
        if (qb_row) r.qb = qb_row[b];
        return r;
    }
    __device__ __forceinline__ cplx<T> *row_out(long b) const {
        return out + (row_off ? row_off[b] : b * ostride);
    }
    __device__ __forceinline__ T fac(long r) const {
        T K = (T)r * invG;
        return normfac * dexp(K * K * qb);
    }
    __device__ __forceinline__ void operator()(cplx<T> *o, long f, cplx<T> A) const {
        o[f] = A * fac(f);
        if (f != 0 && 2 * f != G) o[G - f] = conjg(A) * fac(G - f);
    }
};

// A[r] and A[N-r] of the real grid from Z[r], Z[N-r]:
//   E = (Z[r] + conj Z[N-r])/2, O = (Z[r] - conj Z[N-r])/(2i),
//   A[r] = E + w_G^r O,  A[N-r] = conj(E - w_G^r O).
template <typename T, class Sink>
__device__ __forceinline__ void emit_pair(const Sink &sk, cplx<T> *o, long r, long N, cplx<T> Zr,
                                          cplx<T> Znr, cplx<T> wr) {
    cplx<T> c = conjg(Znr);
    cplx<T> E = (Zr + c) * T(0.5);
    cplx<T> D = (Zr - c) * T(0.5);
    cplx<T> O = cmk(D.y, -D.x);
    cplx<T> wO = wr * O;
    sk(o, r, E + wO);
    if (2 * r != N) sk(o, N - r, conjg(E - wO));
}

// ------------------------------------------------------------- loaders --
// Forward: the real grid read as complex z[n] (batch stride N).
template <typename T>
struct LoadGrid {
    // plain loads: the kernel can afford to hold staged twiddles in registers
    static constexpr bool kStage = true;
    const cplx<T> *z;
    long stride;  // per batch
    __device__ __forceinline__ void advance(long b) { z += b * stride; }
    __device__ __forceinline__ cplx<T> at(long n) const { return z[n]; }
};

// Backward (C2R of a Hermitian spectrum): builds Z[q] = E_q + i O_q with
// E_q = H_q + H_{q+N}, O_q = w_G^{-q} (H_q - H_{q+N}),
// H_f = (c_f + conj c_{G-f})/2, c_f = D(k) g[M/2 + k] for f = k mod G in band.
template <typename T>
struct LoadHermitian {
    static constexpr bool kStage = false;  // heavy loader: twiddles stay in global memory
    const cplx<T> *g;  // M values, natural order k = -M/2..M/2-1
    int rot;           // 1: use -i g (the grid then gives the imaginary part)
    long gstride;      // per batch
    long M, G, N;
    T invG, coef, pref;
    Tw2<T> twG;
    __device__ __forceinline__ void advance(long b) { g += b * gstride; }
    __device__ __forceinline__ cplx<T> c(long f) const {
        long k;
        if (f < M / 2)
            k = f;
        else if (f >= G - M / 2 && f < G)
            k = f - G;
        else
            return cmk(T(0), T(0));
        T kk = (T)k * invG;
        T D = pref * dexp(coef * kk * kk);
        const cplx<T> v = g[M / 2 + k] * D;
        return rot ? cmk(v.y, -v.x) : v;  // -i v: imaginary part of the forward transform
    }
    __device__ __forceinline__ cplx<T> H(long f) const {
        long mf = f == 0 ? 0 : G - f;
        return (c(f) + conjg(c(mf))) * T(0.5);
    }
    __device__ __forceinline__ cplx<T> at(long q) const {
        cplx<T> h0 = H(q), h1 = H(q + N);
        cplx<T> E = h0 + h1;
        cplx<T> w = conjg(twG(q));
        cplx<T> O = w * (h0 - h1);
        return E + mul_i(O);
    }
};

// ------------------------------------------------------------- kernels --
// Every kernel issues its twiddle loads together with its data loads: the
// per-pass table goes to LDS (when small, lds_table) through the pass-0 hook,
// and the epilogue's twiddles are fetched into registers up front, so the
// workgroup waits on global memory once instead of once per pass.
// Launch LDS: (sequences + (lds_table(LG) ? 1 : 0)) * 2^LG complex values.

// Two-level twiddle w(e) = lo[e mod B] * hi[e / B], split into its two loads
// (issued early) and the product (formed late).
template <typename T>
struct Tw2Parts {
    cplx<T> lo, hi;
    __device__ __forceinline__ void load(const Tw2<T> &t, long e) {
        lo = t.lo[e & ((1L << t.lb) - 1)];
        hi = t.hi[e >> t.lb];
    }
    __device__ __forceinline__ cplx<T> get() const { return lo * hi; }
};

// One workgroup per transform (batch = blockIdx.x), whole sequence in LDS.
// MODE 0: real post-processing into a sink; MODE 1: plain store of z.
template <typename T, int S, int LG, int NT, class Loader, class Sink, int MODE>
__global__ __launch_bounds__(NT) void k_fft_small(Loader ld, Sink sk, cplx<T> *plain_out,
                                                  const cplx<T> *tw, Tw2<T> twG) {
    constexpr int N = 1 << LG;
    constexpr bool TAB = Loader::kStage && lds_table(LG);
    constexpr int OPT = MODE == 0 ? (N / 2) / NT + 1 : 1;  // epilogue pairs per thread
    cplx<T> *lds = reinterpret_cast<cplx<T> *>(cuna_smem);
    cplx<T> *twl = lds + N;
    const int tid = threadIdx.x;
    const long b = blockIdx.x;
    Loader lb = ld;
    lb.advance(b);
    TableStage<T, (TAB ? N : 1), NT> tab;
    if constexpr (TAB) tab.load(tw, tid);
    Tw2Parts<T> ot[OPT];
    if constexpr (MODE == 0) {
#pragma unroll
        for (int i = 0; i < OPT; ++i) {
            const int r = tid + i * NT;
            if (r <= N / 2) ot[i].load(twG, r);
        }
    }
    auto fetch = [&](int e, int) __attribute__((always_inline)) { return lb.at(e); };
    if constexpr (TAB) {
        auto hook = [&]() __attribute__((always_inline)) { tab.store(twl, tid); };
        lds_fft<T, S, LG, 1, NT>(lds, twl, fetch, hook);
    } else {
        lds_fft<T, S, LG, 1, NT>(lds, tw, fetch);
    }
    if constexpr (MODE == 0) {
        const auto skb = sk.for_row(b);
        cplx<T> *o = sk.row_out(b);
#pragma unroll
        for (int i = 0; i < OPT; ++i) {
            const int r = tid + i * NT;
            if (r <= N / 2) {
                cplx<T> Zr = lds[lds_sw(r)];
                cplx<T> Zn = lds[lds_sw((N - r) & (N - 1))];
                emit_pair(skb, o, (long)r, (long)N, Zr, Zn, ot[i].get());
            }
        }
    } else {
        cplx<T> *o = plain_out + b * N;
        for (int e = tid; e < N; e += NT) o[e] = lds[lds_sw(e)];
    }
}

// Pass 1 of the two-pass FFT: W consecutive columns n1 of the L1 x L2 view
// (n = n1 + L1 n2) per workgroup; length-L2 FFT over n2, multiply by
// w_N^(S n1 r2), store at n1 + L1 r2 of dst (may alias the source).
template <typename T, int S, int LG2, int W, int NT, class Loader, int MR = 4>
__global__ __launch_bounds__(NT) void k_fft_col(Loader ld, cplx<T> *dst, long dst_stride, int L1,
                                                const cplx<T> *tw, Tw2<T> twN, long b0) {
    constexpr int L2 = 1 << LG2;
    constexpr bool TAB = Loader::kStage && lds_table(LG2);
    // FP32: the epilogue twiddles are loaded as the first pass's hook, after
    // its data loads (loads return in order, so the pass waits only for its
    // data), and held through the FFT
    constexpr bool PRE = Loader::kStage && sizeof(T) == 4;
    constexpr int OPT = PRE ? (W * L2 + NT - 1) / NT : 1;
    cplx<T> *lds = reinterpret_cast<cplx<T> *>(cuna_smem);
    cplx<T> *twl = lds + W * L2;
    const int tid = threadIdx.x;
    const int nblk = L1 / W;
    const int blk = xcd_remap(blockIdx.x, nblk);
    const long b = b0 + blockIdx.y;  // launches span at most kMaxGridY batches
    const int n10 = blk * W;
    Loader lb = ld;
    lb.advance(b);
    TableStage<T, (TAB ? L2 : 1), NT> tab;
    if constexpr (TAB) tab.load(tw, tid);
    Tw2Parts<T> ot[OPT];
    auto load_ot = [&]() __attribute__((always_inline)) {
        if constexpr (PRE) {
#pragma unroll
            for (int i = 0; i < OPT; ++i) {
                const int idx = tid + i * NT;
                if ((W * L2) % NT == 0 || idx < W * L2)
                    ot[i].load(twN, (long)(n10 + idx % W) * (idx / W));
            }
        }
    };
    auto fetch = [&](int e, int w) __attribute__((always_inline)) { return lb.at(n10 + w + (long)L1 * e); };
    if constexpr (TAB) {
        auto hook = [&]() __attribute__((always_inline)) {
            tab.store(twl, tid);
            load_ot();
        };
        lds_fft<T, S, LG2, W, NT, MR>(lds, twl, fetch, hook);
    } else {
        lds_fft<T, S, LG2, W, NT, MR>(lds, tw, fetch, load_ot);
    }
    cplx<T> *o = dst + b * dst_stride;
    if constexpr (PRE) {
#pragma unroll
        for (int i = 0; i < OPT; ++i) {
            const int idx = tid + i * NT;
            if ((W * L2) % NT == 0 || idx < W * L2) {
                const int w = idx % W, r2 = idx / W;
                cplx<T> t = ot[i].get();
                if constexpr (S < 0) t.y = -t.y;
                o[n10 + w + (long)L1 * r2] = lds[lds_sw(r2 * W + w)] * t;
            }
        }
    } else {
        for (int idx = tid; idx < W * L2; idx += NT) {
            const int w = idx % W, r2 = idx / W;
            cplx<T> t = twN((long)(n10 + w) * r2);
            if constexpr (S < 0) t.y = -t.y;
            o[n10 + w + (long)L1 * r2] = lds[lds_sw(r2 * W + w)] * t;
        }
    }
}

// Pass 2, real post-processing: workgroup i holds rows rho = i*H + h and their
// mirrors (L2 - rho) mod L2 (row 0's mirror slot holds row L2/2; both are
// self-paired).  Length-L1 FFTs over n1, then pairs (Z[r], Z[N-r]) with
// r = rho + L2 r1 go through emit_pair into the sink.
template <typename T, int LG1, int H, int NT, class Sink, int MR = 4>
__global__ __launch_bounds__(NT) void k_fft_row_real(const cplx<T> *src, long src_stride,
                                                     int L2, Sink sk, const cplx<T> *tw,
                                                     Tw2<T> twG, long b0) {
    constexpr int L1 = 1 << LG1;
    constexpr int W = 2 * H;
    constexpr bool TAB = lds_table(LG1);
    constexpr int OPT = (H * L1 + NT - 1) / NT;
    cplx<T> *lds = reinterpret_cast<cplx<T> *>(cuna_smem);
    cplx<T> *twl = lds + W * L1;
    const int tid = threadIdx.x;
    const int nblk = L2 / (2 * H);
    const int i = xcd_remap(blockIdx.x, nblk);
    const long b = b0 + blockIdx.y;  // launches span at most kMaxGridY batches
    const long N = (long)L1 * L2;
    const cplx<T> *z = src + b * src_stride;
    TableStage<T, (TAB ? L1 : 1), NT> tab;
    if constexpr (TAB) tab.load(tw, tid);
    // epilogue twiddles for r = rho + L2 r1 (row 0 included; its mirror row
    // L2/2 takes its own below): in FP32 loaded as the first pass's hook,
    // after its data loads, so that the pass waits only for its data (loads
    // return in order); in FP64, where that measured slower, before them
    Tw2Parts<T> ot[OPT];
    auto load_ot = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int k = 0; k < OPT; ++k) {
            const int idx = tid + k * NT;
            if ((H * L1) % NT == 0 || idx < H * L1)
                ot[k].load(twG, (long)(i * H + idx % H) + (long)L2 * (idx / H));
        }
    };
    constexpr bool LATE = sizeof(T) == 4;
    if constexpr (!LATE) load_ot();
    auto row_of = [&](int s) -> int {
        int h = s % H;
        int rho = i * H + h;
        if (s < H) return rho;
        return rho == 0 ? L2 / 2 : L2 - rho;
    };
    auto fetch = [&](int e, int s) __attribute__((always_inline)) { return z[(long)row_of(s) * L1 + e]; };
    auto late = [&]() __attribute__((always_inline)) {
        if constexpr (LATE) load_ot();
    };
    if constexpr (TAB) {
        auto hook = [&]() __attribute__((always_inline)) {
            tab.store(twl, tid);
            late();
        };
        lds_fft<T, 1, LG1, W, NT, MR>(lds, twl, fetch, hook);
    } else {
        lds_fft<T, 1, LG1, W, NT, MR>(lds, tw, fetch, late);
    }
    const auto skb = sk.for_row(b);
    cplx<T> *o = sk.row_out(b);
    // pairs: slot h, element r1 of the primary row with its mirror element
#pragma unroll
    for (int k = 0; k < OPT; ++k) {
        const int idx = tid + k * NT;
        if (!((H * L1) % NT == 0 || idx < H * L1)) continue;
        const int h = idx % H, r1 = idx / H;
        const int rho = i * H + h;
        if (rho != 0) {
            long r = rho + (long)L2 * r1;
            cplx<T> Zr = lds[lds_sw(r1 * W + h)];
            cplx<T> Zn = lds[lds_sw((L1 - 1 - r1) * W + (H + h))];
            emit_pair(skb, o, r, N, Zr, Zn, ot[k].get());
        } else {
            // row 0 (slot h=0): element r1 pairs with (L1 - r1) mod L1, r1 <= L1/2
            if (r1 <= L1 / 2) {
                long r = (long)L2 * r1;
                cplx<T> Zr = lds[lds_sw(r1 * W + 0)];
                cplx<T> Zn = lds[lds_sw(((L1 - r1) & (L1 - 1)) * W + 0)];
                emit_pair(skb, o, r, N, Zr, Zn, ot[k].get());
            }
            // row L2/2 (slot H): element r1 pairs with L1-1-r1, r1 < L1/2
            if (r1 < L1 / 2) {
                long r = L2 / 2 + (long)L2 * r1;
                cplx<T> Zr = lds[lds_sw(r1 * W + H)];
                cplx<T> Zn = lds[lds_sw((L1 - 1 - r1) * W + H)];
                emit_pair(skb, o, r, N, Zr, Zn, twG(r));
            }
        }
    }
}

// Pass 2, plain: H consecutive rows per workgroup, length-L1 FFT, result
// X[r2 + L2 r1] stored in natural order.
template <typename T, int S, int LG1, int H, int NT>
__global__ __launch_bounds__(NT) void k_fft_row_plain(const cplx<T> *src, long src_stride,
                                                      int L2, cplx<T> *dst, long dst_stride,
                                                      const cplx<T> *tw, long b0) {
    constexpr int L1 = 1 << LG1;
    constexpr bool TAB = lds_table(LG1);
    cplx<T> *lds = reinterpret_cast<cplx<T> *>(cuna_smem);
    cplx<T> *twl = lds + H * L1;
    const int tid = threadIdx.x;
    const int nblk = L2 / H;
    const int i = xcd_remap(blockIdx.x, nblk);
    const long b = b0 + blockIdx.y;  // launches span at most kMaxGridY batches
    const cplx<T> *z = src + b * src_stride;
    TableStage<T, (TAB ? L1 : 1), NT> tab;
    if constexpr (TAB) tab.load(tw, tid);
    auto fetch = [&](int e, int h) __attribute__((always_inline)) { return z[(long)(i * H + h) * L1 + e]; };
    if constexpr (TAB) {
        auto hook = [&]() __attribute__((always_inline)) { tab.store(twl, tid); };
        lds_fft<T, S, LG1, H, NT>(lds, twl, fetch, hook);
    } else {
        lds_fft<T, S, LG1, H, NT>(lds, tw, fetch);
    }
    cplx<T> *o = dst + b * dst_stride;
    for (int idx = tid; idx < H * L1; idx += NT) {
        const int h = idx % H, r1 = idx / H;
        o[(i * H + h) + (long)L2 * r1] = lds[lds_sw(r1 * H + h)];
    }
}

}  // namespace cuna
