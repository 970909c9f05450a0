// cuna_kernels.hip -- instantiation and launch of the FFT and spread kernels
// for the library's real type (float in libcunaf.so, double in libcunad.so).
#include <hip/hip_ext.h>

#include <algorithm>
#include <cstdlib>

#include "cuna_launch.hpp"

namespace cuna {

// raise the dynamic LDS limit of a kernel once per device (not on every
// launch, so that launches can be captured into a graph).  The record is a
// fixed table: the lookup never allocates host memory (a full table only
// means the attribute is set again on each launch).
template <typename K>
static void set_lds(K kernel, size_t bytes) {
    if (bytes <= 64 * 1024) return;
    struct Entry {
        const void *key;
        int dev;
        size_t bytes;
    };
    constexpr int kEntries = 512;
    static std::mutex mu;
    static Entry done[kEntries];
    static int used = 0;
    const void *key = reinterpret_cast<const void *>(kernel);
    int dev = 0;
    CUNA_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    Entry *e = nullptr;
    for (int i = 0; i < used && !e; ++i)
        if (done[i].key == key && done[i].dev == dev) e = &done[i];
    if (e && e->bytes >= bytes) return;
    CUNA_HIP(hipFuncSetAttribute(key, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    if (!e && used < kEntries) e = &done[used++];
    if (e) *e = Entry{key, dev, bytes};
}

static constexpr int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ----------------------------------------------------------- FFT config --
// threads per workgroup for W sequences of length 2^LG, 2^mr values per thread
constexpr int fft_threads(int lg, int w, int mr = 4) {
    return clampi(((1 << lg) * w) >> mr, 64, 1024);
}
// The two-pass column and row kernels use radix-8 passes (8 values per
// thread, twice the waves of radix 16): measured at N = 2^20, 0.8 and 1.1 us
// faster per pass, the more waves per SIMD overlapping LDS work with loads.
constexpr int kColMr = 3, kRowMr = 3;
// sequences per workgroup so that the LDS tile stays at 64 KB
template <typename T>
constexpr int seq_per_wg(int lg, int cap) {
    return clampi((int)(65536 / ((1L << lg) * (long)sizeof(cplx<T>))), 1, cap);
}
// two-pass split N = L1 * L2, L2 >= L1
constexpr int split_lg1(int lgN) { return lgN / 2; }
// largest log2 N of the single-workgroup inverse (Hermitian) transform
constexpr int kC2rSmallLg = 13;

template <typename T, int S, int LG, class Loader, class Sink, int MODE>
static void run_small(const FftShape &sh, const Loader &ld, const Sink &sk, cplx<T> *plain_out,
                      hipStream_t s) {
    constexpr int NT = fft_threads(LG, 1);
    auto &tc = TwiddleCache<T>::get();
    const cplx<T> *tw = tc.fine(LG);
    Tw2<T> twG = tc.two_level(LG + 1);
    size_t lds = (sizeof(cplx<T>) << LG) * (lds_table(LG) ? 2 : 1);
    auto k = k_fft_small<T, S, LG, NT, Loader, Sink, MODE>;
    set_lds(k, lds);
    hipLaunchKernelGGL(k, dim3(sh.batch), dim3(NT), lds, s, ld, sk, plain_out, tw, twG);
    CUNA_LAUNCH_CHECK();
}

template <typename T, int S, int LG2, int W, class Loader>
static void run_col(const FftShape &sh, const Loader &ld, cplx<T> *dst, int L1, hipStream_t s) {
    constexpr int NT = fft_threads(LG2, W, kColMr);
    auto &tc = TwiddleCache<T>::get();
    const long N = 1L << sh.lgN;
    size_t lds = (sizeof(cplx<T>) << LG2) * (W + (lds_table(LG2) ? 1 : 0));
    auto k = k_fft_col<T, S, LG2, W, NT, Loader, kColMr>;
    set_lds(k, lds);
    require(L1 % W == 0, "fft column split");
    for (long b0 = 0; b0 < sh.batch; b0 += kMaxGridY) {
        const int nb = (int)std::min<long>(kMaxGridY, sh.batch - b0);
        hipLaunchKernelGGL(k, dim3(L1 / W, nb), dim3(NT), lds, s, ld, dst, N, L1, tc.fine(LG2),
                           tc.two_level(sh.lgN), b0);
        CUNA_LAUNCH_CHECK();
    }
}

template <typename T, int LG1, int H, class Sink>
static void run_row_real(const FftShape &sh, const cplx<T> *src, int L2, const Sink &sk,
                         hipStream_t s) {
    constexpr int NT = fft_threads(LG1, 2 * H, kRowMr);
    auto &tc = TwiddleCache<T>::get();
    const long N = 1L << sh.lgN;
    size_t lds = (sizeof(cplx<T>) << LG1) * (2 * H + (lds_table(LG1) ? 1 : 0));
    auto k = k_fft_row_real<T, LG1, H, NT, Sink, kRowMr>;
    set_lds(k, lds);
    require(L2 % (2 * H) == 0, "fft row split");
    for (long b0 = 0; b0 < sh.batch; b0 += kMaxGridY) {
        const int nb = (int)std::min<long>(kMaxGridY, sh.batch - b0);
        hipLaunchKernelGGL(k, dim3(L2 / (2 * H), nb), dim3(NT), lds, s, src, N, L2, sk,
                           tc.fine(LG1), tc.two_level(sh.lgN + 1), b0);
        CUNA_LAUNCH_CHECK();
    }
}

template <typename T, int S, int LG1, int H>
static void run_row_plain(const FftShape &sh, const cplx<T> *src, int L2, cplx<T> *dst,
                          hipStream_t s) {
    constexpr int NT = fft_threads(LG1, H);
    auto &tc = TwiddleCache<T>::get();
    const long N = 1L << sh.lgN;
    size_t lds = (sizeof(cplx<T>) << LG1) * (H + (lds_table(LG1) ? 1 : 0));
    auto k = k_fft_row_plain<T, S, LG1, H, NT>;
    set_lds(k, lds);
    require(L2 % H == 0, "fft row split");
    for (long b0 = 0; b0 < sh.batch; b0 += kMaxGridY) {
        const int nb = (int)std::min<long>(kMaxGridY, sh.batch - b0);
        hipLaunchKernelGGL(k, dim3(L2 / H, nb), dim3(NT), lds, s, src, N, L2, dst, N, tc.fine(LG1),
                           b0);
        CUNA_LAUNCH_CHECK();
    }
}

// Sequences per workgroup of a pass: the largest count whose LDS tile fits in
// 64 KB that still gives >= min_wg workgroups; an environment override
// (CUNA_FFT_W, CUNA_FFT_H) exists for tuning runs.  Measured at N = 2^20
// (1024 x 1024): the column pass is fastest at one 256-thread workgroup per
// CU (W = 4), the row pass with mirror pairs at two or more per CU (H = 1).
template <typename T, int LG, int W>
constexpr bool lds_fits() {
    return (long)W * (1L << LG) * (long)sizeof(cplx<T>) <= 65536;
}

static int env_int(const char *name) {
    const char *v = getenv(name);
    return v ? atoi(v) : 0;
}

// compute units of the current device (cached per device, fixed table)
static int num_cus() {
    constexpr int kDevices = 64;
    static std::mutex mu;
    static int cache[kDevices] = {};
    int dev = 0;
    CUNA_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    if (dev >= 0 && dev < kDevices && cache[dev] > 0) return cache[dev];
    int n = 0;
    CUNA_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    n = std::max(n, 1);
    if (dev >= 0 && dev < kDevices) cache[dev] = n;
    return n;
}

template <typename T>
static int pick_seq(int lg, long nseq, int batch, int per_wg_mult, const char *env,
                    int requested = 0, long min_wg = 512) {
    int forced = requested ? requested : env_int(env);
    auto valid = [&](int w) {
        long lds = (long)w * per_wg_mult * (1L << lg) * (long)sizeof(cplx<T>);
        return w >= 1 && w <= 8 && (w & (w - 1)) == 0 && lds <= 65536 &&
               nseq % (w * per_wg_mult) == 0;
    };
    if (forced && valid(forced)) return forced;
    int best = 1;
    for (int w = 8; w >= 1; w >>= 1) {
        if (!valid(w)) continue;
        if ((nseq / (w * per_wg_mult)) * (long)batch >= min_wg) return w;
        best = std::min(best, w);
    }
    return best;
}

template <typename T, int LG, class F>
static void dispatch_w(int w, F &&f) {
    if constexpr (lds_fits<T, LG, 8>())
        if (w == 8) return f(std::integral_constant<int, 8>{});
    if constexpr (lds_fits<T, LG, 4>())
        if (w == 4) return f(std::integral_constant<int, 4>{});
    if constexpr (lds_fits<T, LG, 2>())
        if (w == 2) return f(std::integral_constant<int, 2>{});
    if (w == 1) return f(std::integral_constant<int, 1>{});
    fail(CUNA_ERROR_UNSUPPORTED, "fft workgroup shape");
}

// dispatch for the row pass with mirror pairs (2H sequences per workgroup)
template <typename T, int LG, class F>
static void dispatch_h(int h, F &&f) {
    if constexpr (lds_fits<T, LG, 8>())
        if (h == 4) return f(std::integral_constant<int, 4>{});
    if constexpr (lds_fits<T, LG, 4>())
        if (h == 2) return f(std::integral_constant<int, 2>{});
    if constexpr (lds_fits<T, LG, 2>())
        if (h == 1) return f(std::integral_constant<int, 1>{});
    fail(CUNA_ERROR_UNSUPPORTED, "fft workgroup shape");
}

// compile-time dispatch on a runtime log2 size
template <int LO, int HI, class F>
static void dispatch_lg(int lg, F &&f) {
    if constexpr (LO <= HI) {
        if (lg == LO) {
            f(std::integral_constant<int, LO>{});
            return;
        }
        dispatch_lg<LO + 1, HI>(lg, std::forward<F>(f));
    } else {
        fail(CUNA_ERROR_UNSUPPORTED, "transform size out of range");
    }
}

template <typename T, class Sink>
static void real_forward_impl(const FftShape &sh, T *grid, const Sink &sk, hipStream_t s,
                              hipEvent_t mid) {
    require(sh.lgN >= 1 && sh.lgN <= max_fft_lg<T>(), "grid size out of range",
            CUNA_ERROR_UNSUPPORTED);
    LoadGrid<T> ld;
    ld.z = reinterpret_cast<const cplx<T> *>(grid);
    ld.stride = 1L << sh.lgN;
    if (sh.lgN <= max_small_lg<T>()) {
        dispatch_lg<1, max_small_lg<T>()>(sh.lgN, [&](auto c) {
            run_small<T, 1, decltype(c)::value, LoadGrid<T>, Sink, 0>(sh, ld, sk, nullptr, s);
        });
        return;
    }
    const int lg1 = split_lg1(sh.lgN), lg2 = sh.lgN - lg1;
    cplx<T> *z = reinterpret_cast<cplx<T> *>(grid);
    const int W = pick_seq<T>(lg2, 1L << lg1, sh.batch, 1, "CUNA_FFT_W", sh.w, num_cus());
    const int H = std::min(4, pick_seq<T>(lg1, 1L << lg2, sh.batch, 2, "CUNA_FFT_H", sh.h));
    dispatch_lg<7, 12>(lg2, [&](auto c) {
        constexpr int LG2 = decltype(c)::value;
        dispatch_w<T, LG2>(W, [&](auto w) {
            run_col<T, 1, LG2, decltype(w)::value, LoadGrid<T>>(sh, ld, z, 1 << lg1, s);
        });
    });
    if (mid) CUNA_HIP(hipEventRecord(mid, s));
    dispatch_lg<7, 12>(lg1, [&](auto c) {
        constexpr int LG1 = decltype(c)::value;
        dispatch_h<T, LG1>(H, [&](auto h) {
            run_row_real<T, LG1, decltype(h)::value, Sink>(sh, z, 1 << lg2, sk, s);
        });
    });
}

template <typename T>
void fft_real_forward(const FftShape &sh, T *grid, const SinkAccurate<T> &sk, hipStream_t s,
                      hipEvent_t mid) {
    real_forward_impl<T, SinkAccurate<T>>(sh, grid, sk, s, mid);
}
template <typename T>
void fft_real_forward(const FftShape &sh, T *grid, const SinkCompat<T> &sk, hipStream_t s,
                      hipEvent_t mid) {
    real_forward_impl<T, SinkCompat<T>>(sh, grid, sk, s, mid);
}

template <typename T>
void fft_c2r_inverse(const FftShape &sh, const LoadHermitian<T> &ld, cplx<T> *work, T *grid,
                     hipStream_t s) {
    require(sh.lgN >= 1 && sh.lgN <= max_fft_lg<T>(), "grid size out of range",
            CUNA_ERROR_UNSUPPORTED);
    cplx<T> *z = reinterpret_cast<cplx<T> *>(grid);
    SinkAccurate<T> dummy{};
    // the Hermitian loader is register-heavy: a 1024-thread workgroup (N =
    // 2^14) would spill, so N >= 2^14 takes the two-pass path in both
    // libraries, whose first split then has L1 >= 2^7
    if (sh.lgN <= kC2rSmallLg) {
        dispatch_lg<1, kC2rSmallLg>(sh.lgN, [&](auto c) {
            run_small<T, -1, decltype(c)::value, LoadHermitian<T>, SinkAccurate<T>, 1>(
                sh, ld, dummy, z, s);
        });
        return;
    }
    const int lg1 = split_lg1(sh.lgN), lg2 = sh.lgN - lg1;
    const int W = pick_seq<T>(lg2, 1L << lg1, sh.batch, 1, "CUNA_FFT_W", sh.w);
    const int H = pick_seq<T>(lg1, 1L << lg2, sh.batch, 1, "CUNA_FFT_H", sh.h);
    dispatch_lg<7, 12>(lg2, [&](auto c) {
        constexpr int LG2 = decltype(c)::value;
        dispatch_w<T, LG2>(W, [&](auto w) {
            run_col<T, -1, LG2, decltype(w)::value, LoadHermitian<T>>(sh, ld, work, 1 << lg1, s);
        });
    });
    dispatch_lg<7, 12>(lg1, [&](auto c) {
        constexpr int LG1 = decltype(c)::value;
        dispatch_w<T, LG1>(H, [&](auto h) {
            run_row_plain<T, -1, LG1, decltype(h)::value>(sh, work, 1 << lg2, z, s);
        });
    });
}

template <typename T>
void fft_complex(const FftShape &sh, cplx<T> *data, cplx<T> *work, int sign, hipStream_t s) {
    require(sh.lgN >= 1 && sh.lgN <= max_fft_lg<T>(), "transform size out of range",
            CUNA_ERROR_UNSUPPORTED);
    require(sign == 1 || sign == -1, "sign must be +1 or -1");
    LoadGrid<T> ld;
    ld.z = data;
    ld.stride = 1L << sh.lgN;
    SinkAccurate<T> dummy{};
    if (sh.lgN <= max_small_lg<T>()) {
        dispatch_lg<1, max_small_lg<T>()>(sh.lgN, [&](auto c) {
            if (sign > 0)
                run_small<T, 1, decltype(c)::value, LoadGrid<T>, SinkAccurate<T>, 1>(sh, ld, dummy,
                                                                                     data, s);
            else
                run_small<T, -1, decltype(c)::value, LoadGrid<T>, SinkAccurate<T>, 1>(
                    sh, ld, dummy, data, s);
        });
        return;
    }
    const int lg1 = split_lg1(sh.lgN), lg2 = sh.lgN - lg1;
    const int W = pick_seq<T>(lg2, 1L << lg1, sh.batch, 1, "CUNA_FFT_W", sh.w);
    const int H = pick_seq<T>(lg1, 1L << lg2, sh.batch, 1, "CUNA_FFT_H", sh.h);
    dispatch_lg<7, 12>(lg2, [&](auto c) {
        constexpr int LG2 = decltype(c)::value;
        dispatch_w<T, LG2>(W, [&](auto w) {
            constexpr int WW = decltype(w)::value;
            if (sign > 0)
                run_col<T, 1, LG2, WW, LoadGrid<T>>(sh, ld, work, 1 << lg1, s);
            else
                run_col<T, -1, LG2, WW, LoadGrid<T>>(sh, ld, work, 1 << lg1, s);
        });
    });
    dispatch_lg<7, 12>(lg1, [&](auto c) {
        constexpr int LG1 = decltype(c)::value;
        dispatch_w<T, LG1>(H, [&](auto h) {
            constexpr int HH = decltype(h)::value;
            if (sign > 0)
                run_row_plain<T, 1, LG1, HH>(sh, work, 1 << lg2, data, s);
            else
                run_row_plain<T, -1, LG1, HH>(sh, work, 1 << lg2, data, s);
        });
    });
}

// ----------------------------------------------------------------- sort --
template <typename T>
void launch_point_setup(const SetupArgs<T> &A, hipStream_t s) {
    const long n = std::max(A.n, (long)A.R_f);
    if (n <= 0) return;
    hipLaunchKernelGGL(k_point_setup<T>, dim3(ceil_div(n, 256)), dim3(256), 0, s, A);
    CUNA_LAUNCH_CHECK();
}

template <typename T>
void launch_digit_keys(const PointRec<T> *rec, long n, int shift, int mask, int *key,
                       hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_digit_keys<T>, dim3(ceil_div(n, 256)), dim3(256), 0, s, rec, n, shift,
                       mask, key);
    CUNA_LAUNCH_CHECK();
}

static SortArgs sort_args(const SortWork &W, int K, const int *keys) {
    SortArgs S;
    S.chunks = W.chunks;
    S.chunk_base = W.chunk_base;
    S.nch = W.nch;
    S.nchunks = W.nchunks;
    S.K = K;
    S.kbits = ilog2(K);
    S.keys = keys;
    S.hist = W.hist;
    return S;
}

template <typename T>
void launch_sort_pass(const SortWork &W, int K, const int *keys, const PointRec<T> *rec_in,
                      const int *perm_in, PointRec<T> *rec_out, int *perm_out, hipStream_t s) {
    if (W.nchunks == 0) return;
    const SortArgs S = sort_args(W, K, keys);
    // waves per block limited by the per-wave LDS counters
    size_t per_wave = sizeof(int) * (size_t)K;
    require(per_wave <= kMaxLds, "too many sort keys", CUNA_ERROR_UNSUPPORTED);
    int wpb = (int)std::max<size_t>(1, std::min<size_t>(4, (64 * 1024) / per_wave));
    size_t lds = per_wave * wpb;
    int nblk = (int)ceil_div(W.nchunks, wpb);
    set_lds(k_hist, lds);
    set_lds(k_scatter<T>, lds);
    hipLaunchKernelGGL(k_hist, dim3(nblk), dim3(64 * wpb), lds, s, S);
    CUNA_LAUNCH_CHECK();
    // exclusive scan of the key-major count matrix
    const long n = (long)K * W.total_chunks;
    const long nb = ceil_div(std::max(n, 1L), kScanTile);
    size_t slds = sizeof(int) * kScanNT;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanNT), slds, s, W.hist, n, W.bsum);
    CUNA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanNT), slds, s, W.bsum, nb);
    CUNA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kScanNT), slds, s, W.hist, n, W.bsum, W.hist);
    CUNA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scatter<T>, dim3(nblk), dim3(64 * wpb), lds, s, S, rec_in, perm_in,
                       rec_out, perm_out);
    CUNA_LAUNCH_CHECK();
}

template <typename T>
void launch_buckets(const int *coff, int ncurves, int Kc, const PointRec<T> *rec, int tlog,
                    int olo, int ohi, int G, int *bucket, hipStream_t s) {
    const long nbk = (long)ncurves * (Kc + 1);
    if (nbk <= 0) return;
    hipLaunchKernelGGL(k_buckets<T>, dim3(ceil_div(nbk, 256)), dim3(256), 0, s, coff, ncurves, Kc,
                       rec, tlog, olo, ohi, G, bucket);
    CUNA_LAUNCH_CHECK();
}

// --------------------------------------------------------------- spread --
// with ev0/ev1 set, the kernel's own start and stop times are recorded in
// them (as its dispatch begins and ends, without separate event packets)
template <typename K, typename T>
static void launch_timed(K k, dim3 grid, dim3 block, size_t lds, hipStream_t s, hipEvent_t ev0,
                         hipEvent_t ev1, const SpreadArgs<T> &A) {
    if (ev0 || ev1)
        hipExtLaunchKernelGGL(k, grid, block, (uint32_t)lds, s, ev0, ev1, 0, A);
    else
        hipLaunchKernelGGL(k, grid, block, lds, s, A);
    CUNA_LAUNCH_CHECK();
}

// grid (ntiles, slabs) in launches of at most kMaxGridY slabs; ev0 records
// the first launch's start, ev1 the last one's stop
template <typename K, typename T>
static void launch_slabs(K k, const SpreadArgs<T> &A, int nslabs, dim3 block, size_t lds,
                         hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    for (int s0 = 0; s0 < nslabs; s0 += kMaxGridY) {
        const int ns = std::min(kMaxGridY, nslabs - s0);
        SpreadArgs<T> B = A;
        B.slab0 = s0;
        launch_timed(k, dim3(A.ntiles, ns), block, lds, s, s0 == 0 ? ev0 : nullptr,
                     s0 + ns == nslabs ? ev1 : nullptr, B);
    }
}

// The scatter form (default): one wave per tile, fixed-point cells per wave
// (32-bit in FP32, 64-bit in FP64), nothing measured at plan time needed to be
// safe.  CUNA_PLAN_GATHER selects the gather form.
template <typename T, int NTAPS, int VM>
static void spread_one(const SpreadArgs<T> &A, int nslabs, hipStream_t s, hipEvent_t ev0,
                       hipEvent_t ev1) {
    if (!A.gather) {
        constexpr int WPT = 1;
        const size_t lds = spread_wv_lds_bytes<T>(A.tlog, NTAPS, WPT);
        require(lds <= kMaxLds, "spread tile too large", CUNA_ERROR_UNSUPPORTED);
        auto k = k_spread_wv<T, NTAPS, VM, WPT>;
        set_lds(k, lds);
        launch_slabs(k, A, nslabs, dim3(64 * WPT), lds, s, ev0, ev1);
        return;
    }
    const size_t lds = spread_lds_bytes<T>(A.tlog, NTAPS, A.cap);
    require(lds <= kMaxLds, "spread tile too large", CUNA_ERROR_UNSUPPORTED);
    auto k = k_spread<T, NTAPS, VM>;
    set_lds(k, lds);
    launch_slabs(k, A, nslabs, dim3(256), lds, s, ev0, ev1);
}

// accurate filters have 2m taps, m in [2, kMaxTaps/2]; compat filters have
// 2 R_f - 1 = 11 (float build) or 23 (double build) taps
template <int LO, int HI, class F>
static void dispatch_m(int m, F &&f) {
    if constexpr (LO <= HI) {
        if (m == LO) {
            f(std::integral_constant<int, 2 * LO>{});
            return;
        }
        dispatch_m<LO + 1, HI>(m, std::forward<F>(f));
    } else {
        fail(CUNA_ERROR_UNSUPPORTED, "unsupported filter width");
    }
}

template <typename T>
void launch_spread(const SpreadArgs<T> &A, int ntaps, int vm, int nslabs, hipStream_t s,
                   hipEvent_t ev0, hipEvent_t ev1) {
    if (ntaps == 11 || ntaps == 23) {
        require(vm != VAL_COUNTS, "counts spread needs an accurate filter");
        if (ntaps == 11) {
            if (vm == VAL_PLAIN)
                spread_one<T, 11, VAL_PLAIN>(A, nslabs, s, ev0, ev1);
            else
                spread_one<T, 11, VAL_COMPAT_BOOT>(A, nslabs, s, ev0, ev1);
        } else {
            if (vm == VAL_PLAIN)
                spread_one<T, 23, VAL_PLAIN>(A, nslabs, s, ev0, ev1);
            else
                spread_one<T, 23, VAL_COMPAT_BOOT>(A, nslabs, s, ev0, ev1);
        }
        return;
    }
    require(ntaps % 2 == 0 && vm != VAL_COMPAT_BOOT, "unsupported filter width",
            CUNA_ERROR_UNSUPPORTED);
    dispatch_m<2, kMaxTaps / 2>(ntaps / 2, [&](auto c) {
        constexpr int NTP = decltype(c)::value;
        if (vm == VAL_PLAIN)
            spread_one<T, NTP, VAL_PLAIN>(A, nslabs, s, ev0, ev1);
        else
            spread_one<T, NTP, VAL_COUNTS>(A, nslabs, s, ev0, ev1);
    });
}

template <typename T>
void launch_tile_cover(const SpreadArgs<T> &A, int ncurves, int *cover, hipStream_t s) {
    if (ncurves <= 0 || A.ntiles <= 0) return;
    const size_t lds = sizeof(int) * (((size_t)1 << A.tlog) + 2 * (size_t)A.ntaps);
    require(lds <= kMaxLds, "spread tile too large", CUNA_ERROR_UNSUPPORTED);
    set_lds(k_tile_cover<T>, lds);
    for (int c0 = 0; c0 < ncurves; c0 += kMaxGridY) {  // blockIdx.y = curve - slab0
        SpreadArgs<T> B = A;
        B.slab0 = c0;
        hipLaunchKernelGGL(k_tile_cover<T>, dim3(A.ntiles, std::min(kMaxGridY, ncurves - c0)),
                           dim3(256), lds, s, B, cover);
        CUNA_LAUNCH_CHECK();
    }
}

void launch_tile_order(const int *bucket, int Kc, int ntiles, int ncurves, bool periodic,
                       const int *cover, TileDesc *desc, hipStream_t s) {
    if (ncurves <= 0 || ntiles <= 0) return;
    require(ntiles <= kOrderMaxTiles, "too many spread tiles for the work order",
            CUNA_ERROR_UNSUPPORTED);
    hipLaunchKernelGGL(k_tile_order, dim3(ncurves), dim3(kOrderThreads), 0, s, bucket, Kc, ntiles,
                       desc);
    CUNA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tile_desc, dim3(ceil_div(ntiles, 256) * ncurves), dim3(256), 0, s, bucket,
                       Kc, ntiles, periodic ? 1 : 0, cover, desc);
    CUNA_LAUNCH_CHECK();
}

template <typename T>
void launch_gather(const GatherArgs<T> &A, int ntaps, int nslabs, hipStream_t s) {
    require(ntaps % 2 == 0, "unsupported filter width", CUNA_ERROR_UNSUPPORTED);
    dispatch_m<2, kMaxTaps / 2>(ntaps / 2, [&](auto c) {
        constexpr int NTP = decltype(c)::value;
        size_t lds = sizeof(T) * (((size_t)1 << A.tlog) + NTP);
        auto k = k_gather<T, NTP>;
        set_lds(k, lds);
        for (int s0 = 0; s0 < nslabs; s0 += kMaxGridY) {
            GatherArgs<T> B = A;
            B.slab0 = s0;
            hipLaunchKernelGGL(k, dim3(A.ntiles, std::min(kMaxGridY, nslabs - s0)), dim3(256),
                               lds, s, B);
            CUNA_LAUNCH_CHECK();
        }
    });
}

// ------------------------------------------------------ fused bootstrap --
// grids up to 2^15 cells (float) / 2^14 (double): 128 KB of LDS, so one
// persistent workgroup per CU
template <typename T>
bool boot_fused_supported(int lgN, int ntaps) {
    return lgN >= 6 && lgN <= max_small_lg<T>() && (ntaps == 12 || ntaps == 16 || ntaps == 20 ||
                                                    ntaps == 24 || ntaps == 26 || ntaps == 32);
}

template <typename T, int LG, int NTAPS, int MODE>
static void boot_fused_one(BootArgs<T> A, hipStream_t s) {
    constexpr int NT = fft_threads(LG, 1) < 256 ? 256 : fft_threads(LG, 1);
    auto &tc = TwiddleCache<T>::get();
    A.twN = tc.two_level(LG);
    A.twG = tc.two_level(LG + 1);
    const size_t lds = boot_lds_bytes<T>(LG);
    static_assert(boot_lds_bytes<T>(LG) <= kMaxLds, "fused bootstrap LDS");
    auto k = k_boot_fused<T, LG, NTAPS, NT, MODE>;
    set_lds(k, lds);
    int dev = 0, ncu = 0;
    CUNA_HIP(hipGetDevice(&dev));
    CUNA_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    int per_cu = 0;
    CUNA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(k),
                                                          NT, lds));
    int nblk = std::max(1, std::min(A.total, ncu * std::max(per_cu, 1)));
    if constexpr (MODE == 1) {
        // one workspace row per workgroup
        require(A.ls_w != nullptr && A.ls_w_slots >= 1, "Lomb-Scargle workspace missing");
        nblk = std::min(nblk, A.ls_w_slots);
    }
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, s, A.counter);
    CUNA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k, dim3(nblk), dim3(NT), lds, s, A);
    CUNA_LAUNCH_CHECK();
}

template <typename T>
void launch_curve_cover(const PointRec<T> *rec, const int *coff, int ncurves, int G, int ntaps,
                        int *cover, hipStream_t s) {
    require(is_pow2(G) && G <= (1 << 15), "cover histogram needs a power-of-two grid <= 2^15");
    if (ncurves <= 0) return;
    const size_t lds = sizeof(int) * (size_t)G;
    set_lds(k_curve_cover<T>, lds);
    hipLaunchKernelGGL(k_curve_cover<T>, dim3(ncurves), dim3(1024), lds, s, rec, coff, G, ntaps,
                       cover);
    CUNA_LAUNCH_CHECK();
}

template <typename T>
void launch_curve_absmax(const T *y, const int *coff, int ncurves, T *out, hipStream_t s) {
    if (ncurves <= 0) return;
    hipLaunchKernelGGL(k_curve_absmax<T>, dim3(ncurves), dim3(256), 0, s, y, coff, out);
    CUNA_LAUNCH_CHECK();
}

template <typename T>
void launch_boot_fused(BootArgs<T> A, int lgN, int ntaps, int mode, hipStream_t s) {
    require(boot_fused_supported<T>(lgN, ntaps), "no fused bootstrap kernel for this size",
            CUNA_ERROR_UNSUPPORTED);
    dispatch_lg<6, max_small_lg<T>()>(lgN, [&](auto c) {
        constexpr int LG = decltype(c)::value;
        auto go = [&](auto m) {
            constexpr int MODE = decltype(m)::value;
            switch (ntaps) {
                case 12: boot_fused_one<T, LG, 12, MODE>(A, s); break;
                case 16: boot_fused_one<T, LG, 16, MODE>(A, s); break;
                case 20: boot_fused_one<T, LG, 20, MODE>(A, s); break;
                case 24: boot_fused_one<T, LG, 24, MODE>(A, s); break;
                case 26: boot_fused_one<T, LG, 26, MODE>(A, s); break;
                default: boot_fused_one<T, LG, 32, MODE>(A, s); break;
            }
        };
        if (mode == 0)
            go(std::integral_constant<int, 0>{});
        else
            go(std::integral_constant<int, 1>{});
    });
}

// ------------------------------------------------------- instantiation --
template void fft_real_forward<dTyp>(const FftShape &, dTyp *, const SinkAccurate<dTyp> &,
                                     hipStream_t, hipEvent_t);
template void fft_real_forward<dTyp>(const FftShape &, dTyp *, const SinkCompat<dTyp> &,
                                     hipStream_t, hipEvent_t);
template void fft_c2r_inverse<dTyp>(const FftShape &, const LoadHermitian<dTyp> &, cplx<dTyp> *,
                                    dTyp *, hipStream_t);
template void fft_complex<dTyp>(const FftShape &, cplx<dTyp> *, cplx<dTyp> *, int, hipStream_t);
template void launch_point_setup<dTyp>(const SetupArgs<dTyp> &, hipStream_t);
template void launch_sort_pass<dTyp>(const SortWork &, int, const int *, const PointRec<dTyp> *,
                                     const int *, PointRec<dTyp> *, int *, hipStream_t);
template void launch_digit_keys<dTyp>(const PointRec<dTyp> *, long, int, int, int *, hipStream_t);
template void launch_buckets<dTyp>(const int *, int, int, const PointRec<dTyp> *, int, int, int, int,
                                   int *, hipStream_t);
template void launch_spread<dTyp>(const SpreadArgs<dTyp> &, int, int, int, hipStream_t, hipEvent_t,
                                  hipEvent_t);
template void launch_gather<dTyp>(const GatherArgs<dTyp> &, int, int, hipStream_t);
template void launch_tile_cover<dTyp>(const SpreadArgs<dTyp> &, int, int *, hipStream_t);
template bool boot_fused_supported<dTyp>(int, int);
template void launch_boot_fused<dTyp>(BootArgs<dTyp>, int, int, int, hipStream_t);
template void launch_curve_cover<dTyp>(const PointRec<dTyp> *, const int *, int, int, int, int *,
                                       hipStream_t);
template void launch_curve_absmax<dTyp>(const dTyp *, const int *, int, dTyp *, hipStream_t);

}  // namespace cuna
