// cuna_api.hip -- plans and the extern "C" API of libcunaf.so / libcunad.so.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <list>
#include <map>
#include <mutex>
#include <unordered_map>

#include "cuna.h"
#include "cuna_filter.h"
#include "cuna_nfft.h"
#include "cuna_plan.hpp"

#define CUNA_API extern "C" __attribute__((visibility("default")))

namespace cuna {

// ----------------------------------------------------------------- errors --
namespace {
thread_local int g_err = CUNA_SUCCESS;
thread_local std::string g_msg;
bool g_abort = false;
}  // namespace

void set_error(int code, const std::string &msg) {
    g_err = code;
    g_msg = msg;
    if (code != CUNA_SUCCESS && g_abort) {
        fprintf(stderr, "CUNA ERROR: %s\n", msg.c_str());
        exit(code);
    }
}
void clear_error() {
    g_err = CUNA_SUCCESS;
    g_msg.clear();
}
bool abort_on_error() { return g_abort; }
[[noreturn]] void fail(int code, const std::string &msg) { throw Error{code, msg}; }

using R = dTyp;

// ------------------------------------------------------------- helpers --
__global__ void k_boot_counts(const int *coff, int ncurves, long n, int nbs, uint32_t seed,
                              int *counts) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nbs) return;
    const int s = (int)(t / n);
    const long i = t % n;
    int lo = 0, hi = ncurves;  // curve c with coff[c] <= i < coff[c+1]
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (coff[mid] <= i)
            lo = mid;
        else
            hi = mid;
    }
    const int c0 = coff[lo];
    const uint32_t nc = (uint32_t)(coff[lo + 1] - c0);
    const uint32_t p = (uint32_t)(i - c0);
    const int d = c0 + (int)(xorshift32(wang_hash(seed + (uint32_t)s * nc + p)) % nc);
    atomicAdd(&counts[(long)s * n + d], 1);
}

template <typename T>
__global__ void k_scale2(const T *in, T *out, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = T(2) * in[i];
}

__global__ void k_fill_int(int *p, long n, int v) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// dst[i] = src[idx[i]]
template <typename T>
__global__ void k_take(const T *src, const int *idx, long n, T *dst) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

// perm[i] = idx[perm[i]]: sorted position -> index into the caller's arrays
__global__ void k_compose(int *perm, const int *idx, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) perm[i] = idx[perm[i]];
}

// ------------------------------------------------------------ TiledPoints --
template <typename T>
void TiledPoints<T>::build(SetupArgs<T> A, const std::vector<int> &offsets, int G_, int tlog_,
                           int ntaps_, bool periodic_, bool keep_eps, hipStream_t s) {
    ncurves = (int)offsets.size() - 1;
    coff_h = offsets;
    n = offsets.back();
    G = G_;
    ntaps = ntaps_;
    olo = -((ntaps - 1) / 2);
    periodic = periodic_;
    // tile: as requested, no larger than the grid, few enough tiles for the
    // sort's per-wave key counters
    const int tmax = 13;
    tlog = std::min(std::min(tlog_, tmax), ilog2(G));
    while ((G >> tlog) > 8192) tlog++;
    require(tlog <= tmax, "grid too large for the spread tiles", CUNA_ERROR_UNSUPPORTED);
    ntiles = (int)ceil_div(G, 1L << tlog);
    if (ntiles > 1 && (1 << tlog) < 2 * ntaps) {
        tlog = ilog2(G);
        ntiles = 1;
    }
    Kc = 3 * ntiles;
    const int K1 = 1 << tlog;

    // chunk table: chunks of at most CH points, never crossing a curve.  One
    // wave ranks a chunk against per-chunk counts of every key: 512 points
    // for the 256-key passes, up to 2048 for larger tiles (first digit)
    const int CH = std::max(512, std::min(2048, 2 * K1));
    std::vector<int4> chunks;
    std::vector<int> base(ncurves), nch(ncurves);
    int total_chunks = 0;
    for (int c = 0; c < ncurves; ++c) {
        int nc = offsets[c + 1] - offsets[c];
        require(nc >= 0, "curve offsets must be non-decreasing");
        int k = (int)ceil_div(nc, CH);
        nch[c] = k;
        base[c] = total_chunks;
        for (int q = 0; q < k; ++q) {
            int st = offsets[c] + q * CH;
            chunks.push_back(make_int4(c, st, std::min(CH, offsets[c + 1] - st), q));
        }
        total_chunks += k;
    }
    const long hist_len = (long)std::max(K1, 256) * std::max(total_chunks, 1);
    // work-queue order of the curves for the fused batched kernels
    std::vector<int> ord(ncurves);
    for (int c = 0; c < ncurves; ++c) ord[c] = c;
    std::stable_sort(ord.begin(), ord.end(), [&](int u, int v) {
        return offsets[u + 1] - offsets[u] > offsets[v + 1] - offsets[v];
    });
    // device memory in two allocations, the plan's buffers and the sort's
    // stream-ordered temporaries, each filled by one upload: allocation and
    // copy calls are most of a plan build's host time
    struct Carve {
        size_t total = 0;
        size_t add(size_t b) {
            const size_t o = total;
            total += (b + 255) & ~size_t(255);
            return o;
        }
    };
    const long n1 = std::max<long>(n, 1);
    const size_t ncov = (size_t)std::max(ncurves, 1) * std::max(ntiles, 1);
    Carve pc;
    const size_t o_up = pc.add(sizeof(int) * (2 * (size_t)ncurves + 1));  // coff, order
    const size_t o_rec = pc.add(sizeof(PointRec<T>) * n1);
    const size_t o_perm = pc.add(sizeof(int) * n1);
    const size_t o_orig = pc.add(sizeof(PointRec<T>) * n1);
    const size_t o_bk = pc.add(sizeof(int) * (size_t)ncurves * (Kc + 1));
    const size_t o_eps = keep_eps ? pc.add(sizeof(T) * n1) : 0;
    const size_t o_cov = pc.add(sizeof(int) * ncov);
    const size_t o_desc = pc.add(sizeof(TileDesc) * ncov);
    arena.alloc(pc.total);
    char *pb = arena.as<char>();
    coff.view(pb + o_up, sizeof(int) * (ncurves + 1));
    order.view(pb + o_up + sizeof(int) * (ncurves + 1), sizeof(int) * ncurves);
    rec.view(pb + o_rec, sizeof(PointRec<T>) * n1);
    perm.view(pb + o_perm, sizeof(int) * n1);
    rec_orig.view(pb + o_orig, sizeof(PointRec<T>) * n1);
    bucket.view(pb + o_bk, sizeof(int) * (size_t)ncurves * (Kc + 1));
    if (keep_eps) {
        eps_orig.view(pb + o_eps, sizeof(T) * n1);
        A.eps_orig = eps_orig.as<T>();
    } else {
        eps_orig.release();
    }
    tile_cover.view(pb + o_cov, sizeof(int) * ncov);
    tile_desc.view(pb + o_desc, sizeof(TileDesc) * ncov);
    plan_h.assign(offsets.begin(), offsets.end());
    plan_h.insert(plan_h.end(), ord.begin(), ord.end());
    CUNA_HIP(hipMemcpyAsync(pb + o_up, plan_h.data(), sizeof(int) * plan_h.size(),
                            hipMemcpyHostToDevice, s));
    Carve tc;
    const size_t o_chunks = tc.add(sizeof(int4) * std::max<size_t>(chunks.size(), 1));
    const size_t o_base = tc.add(sizeof(int) * ncurves);
    const size_t o_nch = tc.add(sizeof(int) * ncurves);
    const size_t o_keys = tc.add(sizeof(int) * n1);
    const size_t o_recA = tc.add(sizeof(PointRec<T>) * n1);
    const size_t o_permA = tc.add(sizeof(int) * n1);
    const size_t o_hist = tc.add(sizeof(int) * hist_len);
    const size_t o_bsum = tc.add(sizeof(int) * ceil_div(hist_len, kScanTile));
    DevBuf tmp;
    tmp.alloc(tc.total, s);
    char *tb = tmp.as<char>();
    // chunk table, chunk bases and counts, laid out as in the temporaries
    sort_h.assign((o_keys - o_chunks) / sizeof(int), 0);
    if (!chunks.empty()) std::memcpy(sort_h.data(), chunks.data(), sizeof(int4) * chunks.size());
    std::memcpy(sort_h.data() + o_base / sizeof(int), base.data(), sizeof(int) * ncurves);
    std::memcpy(sort_h.data() + o_nch / sizeof(int), nch.data(), sizeof(int) * ncurves);
    CUNA_HIP(hipMemcpyAsync(tb, sort_h.data(), sizeof(int) * sort_h.size(), hipMemcpyHostToDevice,
                            s));
    int *const keys = reinterpret_cast<int *>(tb + o_keys);
    PointRec<T> *const recA = reinterpret_cast<PointRec<T> *>(tb + o_recA);
    int *const permA = reinterpret_cast<int *>(tb + o_permA);
    SortWork W;
    W.chunks = reinterpret_cast<const int4 *>(tb + o_chunks);
    W.chunk_base = reinterpret_cast<const int *>(tb + o_base);
    W.nch = reinterpret_cast<const int *>(tb + o_nch);
    W.coff = coff.as<int>();
    W.nchunks = (int)chunks.size();
    W.ncurves = ncurves;
    W.total_chunks = total_chunks;
    W.hist = reinterpret_cast<int *>(tb + o_hist);
    W.bsum = reinterpret_cast<int *>(tb + o_bsum);
    // phase reduction / compat decisions, then an LSD sort by anchor: the cell
    // within the tile, then the tile index in digits of at most 8 bits; the
    // passes alternate between rec/perm and recA/permA, ending in rec/perm
    A.n = n;
    A.G = G;
    A.tlog = tlog;
    A.rec = rec_orig.as<PointRec<T>>();
    A.key = keys;
    launch_point_setup<T>(A, s);
    const int tbits = ilog2(ntiles);
    const int npass = 1 + (tbits + 7) / 8;
    PointRec<T> *const rbuf[2] = {rec.as<PointRec<T>>(), recA};
    int *const pbuf[2] = {perm.as<int>(), permA};
    auto dst = [&](int p) { return (npass - 1 - p) & 1; };
    launch_sort_pass<T>(W, K1, keys, rec_orig.as<PointRec<T>>(), nullptr, rbuf[dst(0)],
                        pbuf[dst(0)], s);
    for (int p = 1; p < npass; ++p) {
        const int shift = 8 * (p - 1), bits = std::min(8, tbits - shift);
        launch_digit_keys<T>(rbuf[dst(p - 1)], n, tlog + shift, (1 << bits) - 1, keys, s);
        launch_sort_pass<T>(W, 1 << bits, keys, rbuf[dst(p - 1)], pbuf[dst(p - 1)],
                            rbuf[dst(p)], pbuf[dst(p)], s);
    }
    launch_buckets<T>(coff.as<int>(), ncurves, Kc, rec.as<PointRec<T>>(), tlog, olo,
                      olo + ntaps - 1, G, bucket.as<int>(), s);
    // the spread's per-tile cover (bounds the sums any cell receives)
    {
        SpreadArgs<T> C{};
        C.rec = rec.as<PointRec<T>>();
        C.bucket = bucket.as<int>();
        C.coff = coff.as<int>();
        C.G = G;
        C.tlog = tlog;
        C.ntiles = ntiles;
        C.Kc = Kc;
        C.olo = olo;
        C.ntaps = ntaps;
        C.periodic = periodic ? 1 : 0;
        C.wrap1 = (periodic && ntiles == 1) ? 1 : 0;
        launch_tile_cover<T>(C, ncurves, tile_cover.as<int>(), s);
        launch_tile_order(bucket.as<int>(), Kc, ntiles, ncurves, periodic, tile_cover.as<int>(),
                          tile_desc.as<TileDesc>(), s);
    }
}

// --------------------------------------------------------------- Plan --
template <typename T>
SpreadArgs<T> Plan<T>::spread_args(const TiledPoints<T> &P) const {
    SpreadArgs<T> A{};
    A.rec = P.rec.template as<PointRec<T>>();
    A.perm = P.perm.template as<int>();
    A.bucket = P.bucket.template as<int>();
    A.coff = P.coff.template as<int>();
    A.G = P.G;
    A.tlog = P.tlog;
    A.ntiles = P.ntiles;
    A.Kc = P.Kc;
    A.olo = P.olo;
    A.ntaps = P.ntaps;
    A.nrep = 1;
    A.periodic = P.periodic ? 1 : 0;
    A.wrap1 = (P.periodic && P.ntiles == 1) ? 1 : 0;
    A.npts = (int)P.n;
    A.cap = sizeof(T) == 4 ? 2048 : 1024;
    A.gather = (spec.flags & CUNA_PLAN_GATHER) ? 1 : 0;
    A.bi = T(1) / beta;
    A.wscale = wscale;
    A.cpar = cpar_c;
    A.tile_cover = P.tile_cover.template as<int>();
    A.tile_desc = P.tile_desc.template as<TileDesc>();
    for (int q = 0; q < kMaxTaps; ++q) A.e3[q] = e3[q];
    return A;
}

template <typename T>
void Plan<T>::prepare_fft(hipStream_t s) {
    if (pow2_grid(G)) {
        TwiddleCache<T>::get().warm(lgN);
        return;
    }
    require(spec.mode == CUNA_MODE_COMPAT, "accurate-mode grids are powers of two");
    blu = std::make_unique<Bluestein<T>>();
    blu->init(G, s);
    blu->ensure(std::max(ncurves(), 1));
}

template <typename T>
void Plan<T>::run_fft(T *grid_ptr, cplx<T> *out, int batch, hipStream_t s, hipEvent_t mid) const {
    FftShape sh;
    sh.lgN = lgN;
    sh.batch = batch;
    sh.w = spec.fft_w;
    sh.h = spec.fft_h;
    if (spec.mode == CUNA_MODE_ACCURATE) {
        SinkAccurate<T> sk;
        sk.out = out;
        sk.M = M;
        sk.ostride = M;
        sk.invG = T(1) / (T)G;
        const double pi = 3.14159265358979323846;
        sk.coef = (T)(pi * pi * (double)beta);
        sk.pref = (T)(1.0 / std::sqrt(pi * (double)beta));
        fft_real_forward<T>(sh, grid_ptr, sk, s, mid);
    } else {
        SinkCompat<T> sk;
        sk.out = out;
        sk.G = G;
        sk.ostride = G;
        sk.invG = T(1) / (T)G;
        sk.qb = b_c * T(0.25);
        sk.normfac = normfac_c;
        sk.row_off = row_off_c;
        sk.qb_row = qb_row_c;
        if (blu)
            blu->run_real(grid_ptr, G, batch, sk, s);
        else
            fft_real_forward<T>(sh, grid_ptr, sk, s, mid);
    }
}

template <typename T>
void Plan<T>::adjoint(const T *y, cplx<T> *out, hipStream_t s) {
    require(out != nullptr, "output pointer is NULL");
    grid.ensure(sizeof(T) * (size_t)G * ncurves());
    adjoint_into(y, out, grid.as<T>(), s);
}

template <typename T>
void Plan<T>::adjoint_into(const T *y, cplx<T> *out, T *grid_ptr, hipStream_t s) const {
    require(out != nullptr, "output pointer is NULL");
    const int nc = ncurves();
    SpreadArgs<T> A = spread_args(pts);
    A.y = y;
    A.grid = grid_ptr;
    A.grid_stride = G;
    // prof[0] and prof[1] take the spread kernel's own start and stop times
    launch_spread<T>(A, ntaps, VAL_PLAIN, nc, s, prof[0], prof[1]);
    run_fft(grid_ptr, out, nc, s, prof[2]);
    if (prof[3]) CUNA_HIP(hipEventRecord(prof[3], s));
}

// Several right-hand sides of one plan: slab c * nrhs + r of the spread takes
// curve c's points with values y[r * n + p] (the spread's per-replicate value
// stride), and the FFT transforms all slabs in one batch, so row c * nrhs + r
// of out equals cuna_nfft_adjoint of row r of y, curve c.
template <typename T>
void Plan<T>::adjoint_many(const T *y, int nrhs, cplx<T> *out, hipStream_t s) {
    require(spec.mode == CUNA_MODE_ACCURATE, "several right-hand sides need an accurate-mode plan");
    require(y != nullptr && out != nullptr, "NULL pointer");
    require(nrhs >= 1, "nrhs must be positive");
    const int nc = ncurves();
    require((long)nc * nrhs <= INT_MAX, "ncurves * nrhs too large", CUNA_ERROR_UNSUPPORTED);
    grid.ensure(sizeof(T) * (size_t)G * nc * nrhs);
    SpreadArgs<T> A = spread_args(pts);
    A.y = y;
    A.y_slab_stride = n();
    A.nrep = nrhs;
    A.grid = grid.as<T>();
    A.grid_stride = G;
    launch_spread<T>(A, ntaps, VAL_PLAIN, nc * nrhs, s, prof[0], prof[1]);
    run_fft(grid.as<T>(), out, nc * nrhs, s, prof[2]);
    if (prof[3]) CUNA_HIP(hipEventRecord(prof[3], s));
}

template <typename T>
void Plan<T>::bootstrap_compat(const T *y, cplx<T> *out, int nbs, uint32_t seed, hipStream_t s) {
    require(spec.mode == CUNA_MODE_COMPAT && ncurves() == 1, "compat bootstrap needs a compat plan");
    require(nbs > 0, "nbs must be positive");
    require(out != nullptr, "output pointer is NULL");
    grid.ensure(sizeof(T) * (size_t)G * nbs);
    if (blu) blu->ensure(nbs);
    SpreadArgs<T> A = spread_args(pts);
    A.y = y;
    A.grid = grid.as<T>();
    A.grid_stride = G;
    A.nrep = nbs;
    A.seed = seed;
    A.eps_orig = pts.eps_orig.template as<T>();
    launch_spread<T>(A, ntaps, VAL_COMPAT_BOOT, nbs, s);
    run_fft(grid.as<T>(), out, nbs, s);
}

template <typename T>
bool Plan<T>::bootstrap_fused_ok(bool window) const {
    return !(spec.flags & CUNA_PLAN_UNFUSED) && boot_fused_supported<T>(lgN, ntaps) &&
           (!window || has_pts2);
}

template <typename T>
void Plan<T>::prepare_bootstrap(int nbs, bool window) {
    if (bootstrap_fused_ok(window)) {
        counter.ensure(sizeof(int));
        ymax.ensure(sizeof(T) * (size_t)ncurves());
        cover.ensure(sizeof(int) * 2 * (size_t)ncurves());
        return;
    }
    counts.ensure(sizeof(int) * (size_t)std::max<long>(n(), 1) * nbs);
    grid.ensure(sizeof(T) * (size_t)G * ncurves() * nbs);
}

// arguments shared by both modes of the fused bootstrap kernel; computes
// the per-curve cover (once per plan) and max |y| (every call) on stream s
template <typename T>
BootArgs<T> Plan<T>::fused_args(const T *y, int nbs, uint32_t seed, hipStream_t s) {
    const int nc = ncurves();
    int *cov = cover.as<int>();
    if (!cover_ready) {
        for (int w = 0; w < (has_pts2 ? 2 : 1); ++w) {
            const TiledPoints<T> &P = w == 0 ? pts : pts2;
            launch_curve_cover<T>(P.rec_orig.template as<PointRec<T>>(), P.coff.template as<int>(),
                                  nc, G, ntaps, cov + w * nc, s);
        }
        cover_ready = true;
    }
    if (y) launch_curve_absmax<T>(y, pts.coff.template as<int>(), nc, ymax.as<T>(), s);
    BootArgs<T> B{};
    B.rec_x = pts.rec_orig.template as<PointRec<T>>();
    B.rec_2x = has_pts2 ? pts2.rec_orig.template as<PointRec<T>>() : nullptr;
    B.y = y;
    B.ymax = ymax.as<T>();
    B.cover_x = cov;
    B.cover_2x = cov + nc;
    B.coff = pts.coff.template as<int>();
    B.order = pts.order.template as<int>();
    B.counter = counter.as<int>();
    B.ncurves = nc;
    B.nbs = nbs;
    B.seed = seed;
    B.bi = T(1) / beta;
    B.wscale = wscale;
    for (int q = 0; q < kMaxTaps; ++q) B.e3[q] = e3[q];
    const double pi = 3.14159265358979323846;
    B.sink.M = M;
    B.sink.ostride = M;
    B.sink.invG = T(1) / (T)G;
    B.sink.coef = (T)(pi * pi * (double)beta);
    B.sink.pref = (T)(1.0 / std::sqrt(pi * (double)beta));
    return B;
}

template <typename T>
void Plan<T>::bootstrap_ls(const T *y, int nbs, uint32_t seed, T *power, T *peak, int *kpeak,
                           hipStream_t s) {
    require(spec.mode == CUNA_MODE_ACCURATE, "Lomb-Scargle bootstrap needs an accurate-mode plan");
    require(has_pts2, "Lomb-Scargle bootstrap needs a plan made with CUNA_PLAN_WINDOW2");
    require(nbs > 0, "nbs must be positive");
    require(y != nullptr, "y is NULL");
    require(power || peak || kpeak, "no output requested");
    require(boot_fused_supported<T>(lgN, ntaps),
            "Lomb-Scargle bootstrap needs a grid of at most 2^15 cells (float) / 2^14 (double) and "
            "m in {6, 8, 10, 12, 13, 16}",
            CUNA_ERROR_UNSUPPORTED);
    prepare_bootstrap(nbs, true);
    // window spectra workspace: one row of M/2 values per resident workgroup
    // (at most one per kMaxLds of LDS on each CU)
    int dev = 0, ncu = 0;
    CUNA_HIP(hipGetDevice(&dev));
    CUNA_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int slots = ncu * std::max(1, (int)(kMaxLds / boot_lds_bytes<T>(lgN)));
    ls_work.ensure(sizeof(cplx<T>) * (size_t)slots * (M / 2));
    BootArgs<T> B = fused_args(y, nbs, seed, s);
    B.ls_power = power;
    B.ls_peak = peak;
    B.ls_kpeak = kpeak;
    B.ls_w = ls_work.as<cplx<T>>();
    B.ls_w_slots = slots;
    B.ntr = 1;
    const long total = (long)ncurves() * nbs;
    require(total < (1L << 31), "too many bootstrap items");
    B.total = (int)total;
    launch_boot_fused<T>(B, lgN, ntaps, 1, s);
}

template <typename T>
void Plan<T>::bootstrap(const T *y, int nbs, uint32_t seed, cplx<T> *out_y, cplx<T> *out_w,
                        hipStream_t s) {
    require(spec.mode == CUNA_MODE_ACCURATE, "bootstrap needs an accurate-mode plan");
    require(nbs > 0, "nbs must be positive");
    require(out_w == nullptr || has_pts2, "window output needs a plan made with CUNA_PLAN_WINDOW2");
    require(out_y != nullptr, "out_y is NULL");
    const int nc = ncurves();
    prepare_bootstrap(nbs, out_w != nullptr);
    if (bootstrap_fused_ok(out_w != nullptr)) {
        BootArgs<T> B = fused_args(y, nbs, seed, s);
        B.out_y = out_y;
        B.out_w = out_w;
        B.ntr = out_w ? 2 : 1;
        const long total = (long)nc * nbs * B.ntr;
        require(total < (1L << 31), "too many bootstrap items");
        B.total = (int)total;
        launch_boot_fused<T>(B, lgN, ntaps, 0, s);
        return;
    }
    const long n_ = n();
    long tot = n_ * nbs;
    hipLaunchKernelGGL(k_fill_int, dim3(ceil_div(std::max(tot, 1L), 256)), dim3(256), 0, s,
                       counts.as<int>(), tot, 0);
    CUNA_LAUNCH_CHECK();
    if (tot > 0) {
        hipLaunchKernelGGL(k_boot_counts, dim3(ceil_div(tot, 256)), dim3(256), 0, s,
                           pts.coff.template as<int>(), nc, n_, nbs, seed, counts.as<int>());
        CUNA_LAUNCH_CHECK();
    }
    for (int pass = 0; pass < 2; ++pass) {
        cplx<T> *out = pass == 0 ? out_y : out_w;
        if (!out) continue;
        const TiledPoints<T> &P = pass == 0 ? pts : pts2;
        SpreadArgs<T> A = spread_args(P);
        A.y = pass == 0 ? y : nullptr;
        A.grid = grid.as<T>();
        A.grid_stride = G;
        A.nrep = nbs;
        A.counts = counts.as<int>();
        A.counts_stride = n_;
        launch_spread<T>(A, ntaps, VAL_COUNTS, nc * nbs, s);
        run_fft(grid.as<T>(), out, nc * nbs, s);
    }
}

template <typename T>
void Plan<T>::backward(const T *y, const cplx<T> *g, T *dy, T *dx, hipStream_t s) {
    require(spec.mode == CUNA_MODE_ACCURATE, "backward needs an accurate-mode plan");
    require(g != nullptr, "gradient pointer is NULL");
    if (!dy && !dx) return;
    const int nc = ncurves();
    const long N = (long)G / 2;
    grid.ensure(sizeof(T) * (size_t)G * nc);
    work.ensure(sizeof(cplx<T>) * (size_t)N * nc);
    const double pi = 3.14159265358979323846;
    LoadHermitian<T> ld;
    ld.rot = 0;
    ld.g = g;
    ld.gstride = M;
    ld.M = M;
    ld.G = G;
    ld.N = N;
    ld.invG = T(1) / (T)G;
    ld.coef = (T)(pi * pi * (double)beta);
    ld.pref = (T)(1.0 / std::sqrt(pi * (double)beta));
    ld.twG = TwiddleCache<T>::get().two_level(lgN + 1);
    FftShape sh;
    sh.lgN = lgN;
    sh.batch = nc;
    sh.w = spec.fft_w;
    fft_c2r_inverse<T>(sh, ld, work.as<cplx<T>>(), grid.as<T>(), s);
    GatherArgs<T> A{};
    A.rec = pts.rec.template as<PointRec<T>>();
    A.perm = pts.perm.template as<int>();
    A.bucket = pts.bucket.template as<int>();
    A.h = grid.as<T>();
    A.h_stride = G;
    A.y = y;
    A.y_stride = 0;
    A.dy = dy;
    A.dx = dx;
    A.out_stride = 0;
    A.dy_elem = 1;
    A.G = G;
    A.tlog = pts.tlog;
    A.ntiles = pts.ntiles;
    A.Kc = pts.Kc;
    A.olo = pts.olo;
    A.nrep = 1;
    A.bi = T(1) / beta;
    A.wscale = wscale;
    A.dxscale = (T)(2.0 * G / (double)beta);
    for (int q = 0; q < kMaxTaps; ++q) A.e3[q] = e3[q];
    launch_gather<T>(A, ntaps, nc, s);
}

// Forward (type-2) NFFT f_j = sum_k fhat_k e^{-2 pi i k x_j}: the real part is
// what backward() computes for dL/dy, Re sum_k fhat_k e^{-2 pi i k x_j}, and
// by linearity the imaginary part is the same with -i fhat; each pass writes
// its part of the interleaved complex output.
template <typename T>
void Plan<T>::forward(const cplx<T> *fhat, cplx<T> *f, hipStream_t s) {
    require(spec.mode == CUNA_MODE_ACCURATE, "forward needs an accurate-mode plan");
    require(fhat != nullptr && f != nullptr, "NULL pointer");
    const int nc = ncurves();
    const long N = (long)G / 2;
    grid.ensure(sizeof(T) * (size_t)G * nc);
    work.ensure(sizeof(cplx<T>) * (size_t)N * nc);
    const double pi = 3.14159265358979323846;
    for (int part = 0; part < 2; ++part) {
        LoadHermitian<T> ld;
        ld.rot = part;
        ld.g = fhat;
        ld.gstride = M;
        ld.M = M;
        ld.G = G;
        ld.N = N;
        ld.invG = T(1) / (T)G;
        ld.coef = (T)(pi * pi * (double)beta);
        ld.pref = (T)(1.0 / std::sqrt(pi * (double)beta));
        ld.twG = TwiddleCache<T>::get().two_level(lgN + 1);
        FftShape sh;
        sh.lgN = lgN;
        sh.batch = nc;
        sh.w = spec.fft_w;
        fft_c2r_inverse<T>(sh, ld, work.as<cplx<T>>(), grid.as<T>(), s);
        GatherArgs<T> A{};
        A.rec = pts.rec.template as<PointRec<T>>();
        A.perm = pts.perm.template as<int>();
        A.bucket = pts.bucket.template as<int>();
        A.h = grid.as<T>();
        A.h_stride = G;
        A.dy = reinterpret_cast<T *>(f) + part;
        A.dy_elem = 2;
        A.G = G;
        A.tlog = pts.tlog;
        A.ntiles = pts.ntiles;
        A.Kc = pts.Kc;
        A.olo = pts.olo;
        A.nrep = 1;
        A.bi = T(1) / beta;
        A.wscale = wscale;
        for (int q = 0; q < kMaxTaps; ++q) A.e3[q] = e3[q];
        launch_gather<T>(A, ntaps, nc, s);
    }
}

// Spread tiles of 256 cells for the one-wave scatter in both libraries (the
// fastest measured at N = 1e6 clustered points, G = 2^21); many small tiles
// keep the chip busy and bound the work of the densest tile
template <typename T>
static int default_tlog() {
    return 8;
}

template <typename T>
std::unique_ptr<Plan<T>> make_plan_accurate(const T *x, const double *t, double df,
                                            const std::vector<int> &offsets, int M,
                                            const ModeSpec &spec, hipStream_t s) {
    require(M >= 2 && is_pow2(M), "M must be a power of two >= 2");
    require(spec.sigma >= 2 && is_pow2(spec.sigma), "sigma must be a power of two >= 2");
    require(spec.m >= 2 && spec.m <= kMaxTaps / 2, "m must be in [2, 16]");
    const long Gl = (long)spec.sigma * M;
    require(ilog2(Gl) - 1 <= max_fft_lg<T>(), "grid too large", CUNA_ERROR_UNSUPPORTED);
    require(offsets.size() >= 2 && offsets.front() == 0, "offsets must start at 0");
    for (size_t c = 1; c < offsets.size(); ++c)
        require(offsets[c] >= offsets[c - 1], "offsets must be non-decreasing");
    const long n = offsets.back();
    require(n == 0 || x || t, "point array is NULL");
    auto P = std::make_unique<Plan<T>>();
    P->spec = spec;
    P->M = M;
    P->G = (int)Gl;
    P->lgN = ilog2(Gl) - 1;
    P->ntaps = 2 * spec.m;
    P->olo = -(spec.m - 1);
    const double pi = 3.14159265358979323846;
    // beta = 2 sigma m / ((2 sigma - 1) pi): the reference's b with R = sigma
    // (src/gaussian_filter.cu:129) and R_f = m
    const double beta = 2.0 * spec.sigma * spec.m / ((2.0 * spec.sigma - 1.0) * pi);
    P->beta = (T)beta;
    P->wscale = T(1);
    for (int q = 0; q < P->ntaps; ++q) {
        double o = (double)(q + P->olo);
        P->e3[q] = (T)std::exp(-o * o / (double)P->beta);
    }
    SetupArgs<T> A{};
    A.x = x;
    A.t = t;
    A.df = df;
    A.mode = t ? SETUP_ACC_T : SETUP_ACC_X;
    int tl = spec.tile_log2 > 0 ? spec.tile_log2 : default_tlog<T>();
    P->pts.build(A, offsets, P->G, tl, P->ntaps, true, false, s);
    if (spec.flags & CUNA_PLAN_WINDOW2) {
        // the points 2x: times with 2 df (FP64), or x doubled (exact in
        // binary floating point)
        SetupArgs<T> B = A;
        if (t) {
            B.df = 2.0 * df;
            P->pts2.build(B, offsets, P->G, tl, P->ntaps, true, false, s);
        } else {
            DevBuf x2;
            x2.alloc(sizeof(T) * std::max<long>(n, 1), s);
            if (n > 0) {
                hipLaunchKernelGGL(k_scale2<T>, dim3(ceil_div(n, 256)), dim3(256), 0, s, x,
                                   x2.as<T>(), n);
                CUNA_LAUNCH_CHECK();
            }
            B.x = x2.as<T>();
            P->pts2.build(B, offsets, P->G, tl, P->ntaps, true, false, s);
        }
        P->has_pts2 = true;
    }
    P->prepare_fft(s);
    return P;
}

template <typename T>
void compat_constants(int n, int ng, T &b, T &normfac, int &R_f) {
    const double pi = 3.14159265358979323846;
    if (sizeof(T) == 4) {
        // src/gaussian_filter.cu:125-134 in float:
        // R = (float)ng / n; b = (float)(((2R)/(2R-1))_float * (6/PI)_double)
        volatile float Rf = (float)ng / (float)n;
        volatile float num = 2.0f * Rf;
        volatile float den = 2.0f * Rf - 1.0f;
        volatile float q = num / den;
        b = (T)(float)((double)q * (6.0 / pi));
        normfac = (T)(float)std::sqrt(2.0 * pi);
        R_f = 6;
    } else {
        double Rd = (double)ng / n;
        b = (T)(2.0 * Rd / (2.0 * Rd - 1.0) * (12.0 / pi));
        normfac = (T)std::sqrt(2.0 * pi);
        R_f = 12;
    }
}

// any ng >= 1 whose transform fits: powers of two >= 4 through the real
// Stockham FFT, other lengths through Bluestein at L >= 2 ng - 1
template <typename T>
static bool compat_ng_ok(long ng) {
    if (ng < 1 || ng > (1L << 26)) return false;
    if (pow2_grid(ng)) return ilog2(ng) - 1 <= max_fft_lg<T>();
    return ilog2(2 * ng - 1) <= max_fft_lg<T>();
}

template <typename T>
std::unique_ptr<Plan<T>> make_plan_compat(const T *x, int n, int ng, T *E1, T *E2, T *E3,
                                          const ModeSpec &spec, hipStream_t s) {
    require(n > 0, "n must be positive");
    require(compat_ng_ok<T>(ng), "ng too large", CUNA_ERROR_UNSUPPORTED);
    // b = 2R/(2R-1) R_f/pi with R = ng/n is negative (NaN weights) for n >= 2 ng
    require((long)n < 2L * ng, "n must be smaller than 2*ng (the filter width b is not positive)");
    require(x != nullptr, "x is NULL");
    auto P = std::make_unique<Plan<T>>();
    P->spec = spec;
    P->spec.mode = CUNA_MODE_COMPAT;
    P->M = ng;
    P->G = ng;
    P->lgN = pow2_grid(ng) ? ilog2(ng) - 1 : 0;
    compat_constants<T>(n, ng, P->b_c, P->normfac_c, P->R_f);
    P->ntaps = 2 * P->R_f - 1;
    P->olo = -(P->R_f - 1);
    P->beta = P->b_c;
    const double pi = 3.14159265358979323846;
    // E1 carries 1/sqrt(pi b) (src/gaussian_filter.cu:219)
    P->wscale = (T)(1.0 / std::sqrt(pi * (double)P->b_c));
    for (int q = 0; q < P->ntaps; ++q) {
        double o = (double)(q + P->olo);
        P->e3[q] = (T)std::exp(-o * o / (double)P->b_c);
    }
    SetupArgs<T> A{};
    A.x = x;
    A.mode = SETUP_COMPAT;
    A.R_f = P->R_f;
    A.b = P->b_c;
    A.e1s = P->wscale;
    A.E1 = E1;
    A.E2 = E2;
    A.E3 = E3;
    int tl = spec.tile_log2 > 0 ? spec.tile_log2 : default_tlog<T>();
    std::vector<int> off = {0, n};
    P->pts.build(A, off, ng, tl, P->ntaps, false, true, s);
    P->prepare_fft(s);
    return P;
}

}  // namespace cuna

namespace cuna {

// the plan machinery is also used by the autotuner's translation unit
template struct TiledPoints<R>;
template struct Plan<R>;
template std::unique_ptr<Plan<R>> make_plan_accurate<R>(const R *, const double *, double,
                                                        const std::vector<int> &, int,
                                                        const ModeSpec &, hipStream_t);

// ------------------------------------------------------------- registry --
// Associates the spread plan with the device filter_properties pointer that
// generate_*filter_properties filled, so that the raw transforms find it.
namespace {
std::mutex g_reg_mu;
std::unordered_map<const void *, std::shared_ptr<Plan<R>>> g_registry;
}  // namespace

static void reg_put(const void *key, std::unique_ptr<Plan<R>> p) {
    std::shared_ptr<Plan<R>> sp(std::move(p));
    std::lock_guard<std::mutex> lk(g_reg_mu);
    g_registry[key] = sp;
}
static std::shared_ptr<Plan<R>> reg_get(const void *key) {
    std::lock_guard<std::mutex> lk(g_reg_mu);
    auto it = g_registry.find(key);
    return it == g_registry.end() ? nullptr : it->second;
}
static void reg_drop(const void *key) {
    std::shared_ptr<Plan<R>> victim;
    {
        std::lock_guard<std::mutex> lk(g_reg_mu);
        auto it = g_registry.find(key);
        if (it != g_registry.end()) {
            victim = it->second;
            g_registry.erase(it);
        }
    }
}

// the plan registered for (fprops, x, n, ng), or a fresh one
static std::shared_ptr<Plan<R>> compat_plan_for(const R *x, int n, int ng,
                                                const filter_properties *fp, hipStream_t s) {
    auto p = fp ? reg_get(fp) : nullptr;
    if (p && p->x_key == x && p->n() == n && p->G == ng) return p;
    ModeSpec spec;
    spec.mode = CUNA_MODE_COMPAT;
    auto fresh = make_plan_compat<R>(x, n, ng, nullptr, nullptr, nullptr, spec, s);
    fresh->x_key = x;
    return std::shared_ptr<Plan<R>>(std::move(fresh));
}

// the reference accepts any ng (cuFFT takes any length), and so does this
// library up to the FFT's largest size
static void check_compat_sizes(int n, int ng) {
    require(n > 0 && ng > 0, "n and ng must be positive");
    require(compat_ng_ok<R>(ng), "ng too large", CUNA_ERROR_UNSUPPORTED);
    require((long)n < 2L * ng, "n must be smaller than 2*ng (the filter width b is not positive)");
}

// ------------------------------------------------- allocation-free entry --
// The entry cunfft_adjoint_raw_async_noalloc (below) keeps the real grid of
// its transform in this file-scope device array instead of a plan workspace,
// so the call allocates nothing.  Every GPU holds its own copy until the
// process ends; all calls on one GPU share it (calls that overlap on a GPU
// race).  The spread writes every cell of the grid, so nothing needs clearing.
constexpr size_t kScratchBytes = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on the current device, resolved once per device
// with hipGetSymbolAddress (in host code the array's name is only a
// stand-in, not a device address).
static void *scratch_ptr() {
    constexpr int kDevices = 64;
    static std::mutex mu;
    static void *ptr[kDevices] = {};
    int dev = 0;
    CUNA_HIP(hipGetDevice(&dev));
    require(dev >= 0 && dev < kDevices, "device index out of range for the scratch table",
            CUNA_ERROR_UNSUPPORTED);
    std::lock_guard<std::mutex> lk(mu);
    if (!ptr[dev]) {
        void *p = nullptr;
        CUNA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        require(p != nullptr, "hipGetSymbolAddress returned no address", CUNA_ERROR_HIP);
        ptr[dev] = p;
    }
    return ptr[dev];
}

static void generate_impl(const R *x, int n, int ng, filter_properties *h, filter_properties *d,
                          hipStream_t s, bool sync) {
    require(h != nullptr && d != nullptr, "filter_properties pointers are NULL");
    require(n > 0 && ng > 0, "n and ng must be positive");
    R b, nf;
    int rf;
    compat_constants<R>(n, ng, b, nf, rf);
    h->b = b;
    h->normfac = nf;
    h->filter_radius = rf;
    CUNA_HIP(hipMemcpyAsync(d, h, sizeof(filter_properties), hipMemcpyHostToDevice, s));
    ModeSpec spec;
    spec.mode = CUNA_MODE_COMPAT;
    auto p = make_plan_compat<R>(x, n, ng, h->E1, h->E2, h->E3, spec, s);
    p->x_key = x;
    // set-up of the allocation-free transform on this device: resolving the
    // scratch array's address loads its code object (the runtime allocates
    // then), so that cunfft_adjoint_raw_async_noalloc itself allocates nothing
    (void)scratch_ptr();
    if (sync) CUNA_HIP(hipStreamSynchronize(s));
    reg_put(d, std::move(p));
}

// ------------------------------------------------ batches of curves --
// The reference's batched compat transform, which it leaves commented out
// (src/cuna_kernel.cu:302-358 with fast_gridding_batched, src/gridding.cu:
// 119-149, and normalize_batch, src/gaussian_filter.cu:303-318): curve i has
// n[i] points of the concatenated x (and f_data) and ng[i] grid cells; its
// spectrum fills the first ng[i] values of slab i of f_hat, slab stride
// idist = max ng.  Curves with the same ng share one plan: one spread launch
// over all of them, each curve with its own filter width b, and one batched
// FFT whose deconvolution also takes b per curve.
struct CompatBatch {
    const void *x_key = nullptr;
    std::vector<int> n, ng;
    std::vector<const void *> fps;  // device filter_properties per curve
    int idist = 0;
    struct Group {
        // host sources of the group's async uploads, declared first so that
        // they outlive the device buffers (whose hipFree synchronises)
        std::vector<int> idx_h;
        std::vector<R> cpar_h, qb_h;
        std::vector<long> ro_h;
        std::unique_ptr<Plan<R>> plan;
        DevMem cpar, qb_row, row_off;
    };
    std::vector<std::unique_ptr<Group>> groups;

    void run(const R *y, cplx<R> *f_hat, hipStream_t s) {
        for (auto &g : groups) g->plan->adjoint(y, f_hat, s);
    }
};

static std::unique_ptr<CompatBatch> make_compat_batch(const R *x, const int *n, const int *ng,
                                                      int nlc, hipStream_t s) {
    require(nlc > 0, "nlc must be positive");
    require(x != nullptr && n != nullptr && ng != nullptr, "x, n and ng must not be NULL");
    auto B = std::make_unique<CompatBatch>();
    B->x_key = x;
    B->n.assign(n, n + nlc);
    B->ng.assign(ng, ng + nlc);
    std::vector<long> off(nlc + 1, 0);
    for (int i = 0; i < nlc; ++i) {
        check_compat_sizes(n[i], ng[i]);
        off[i + 1] = off[i] + n[i];
        B->idist = std::max(B->idist, ng[i]);
    }
    require(off[nlc] <= (long)INT_MAX, "too many points");
    std::map<int, std::vector<int>> by_ng;
    for (int i = 0; i < nlc; ++i) by_ng[ng[i]].push_back(i);
    const double pi = 3.14159265358979323846;
    for (const auto &kv : by_ng) {
        const int g = kv.first;
        const std::vector<int> &cur = kv.second;
        const int nc = (int)cur.size();
        auto G = std::make_unique<CompatBatch::Group>();
        // the group's points in curve order: their indices in x, and offsets
        std::vector<int> &gidx = G->idx_h;
        std::vector<int> goff(1, 0);
        for (int i : cur) {
            for (long j = off[i]; j < off[i + 1]; ++j) gidx.push_back((int)j);
            goff.push_back((int)gidx.size());
        }
        const long gn = (long)gidx.size();
        DevBuf d_idx, d_x;
        d_idx.alloc(sizeof(int) * gn, s);
        d_x.alloc(sizeof(R) * gn, s);
        CUNA_HIP(hipMemcpyAsync(d_idx.ptr, gidx.data(), sizeof(int) * gn, hipMemcpyHostToDevice,
                                s));
        hipLaunchKernelGGL(k_take<R>, dim3(ceil_div(gn, 256)), dim3(256), 0, s, x,
                           d_idx.as<int>(), gn, d_x.as<R>());
        CUNA_LAUNCH_CHECK();
        // one compat plan over the group's curves; the filter constants of
        // its first curve fill the scalar fields, the per-curve arrays below
        // take their place in the kernels
        auto P = std::make_unique<Plan<R>>();
        P->spec.mode = CUNA_MODE_COMPAT;
        P->M = g;
        P->G = g;
        P->lgN = pow2_grid(g) ? ilog2(g) - 1 : 0;
        compat_constants<R>(n[cur[0]], g, P->b_c, P->normfac_c, P->R_f);
        P->ntaps = 2 * P->R_f - 1;
        P->olo = -(P->R_f - 1);
        P->beta = P->b_c;
        P->wscale = (R)(1.0 / std::sqrt(pi * (double)P->b_c));
        for (int q = 0; q < P->ntaps; ++q) {
            const double o = (double)(q + P->olo);
            P->e3[q] = (R)std::exp(-o * o / (double)P->b_c);
        }
        SetupArgs<R> A{};
        A.x = d_x.as<R>();
        A.mode = SETUP_COMPAT;
        A.R_f = P->R_f;
        A.b = P->b_c;
        A.e1s = P->wscale;
        P->pts.build(A, goff, g, default_tlog<R>(), P->ntaps, false, false, s);
        P->prepare_fft(s);
        hipLaunchKernelGGL(k_compose, dim3(ceil_div(std::max(gn, 1L), 256)), dim3(256), 0, s,
                           P->pts.perm.as<int>(), d_idx.as<int>(), gn);
        CUNA_LAUNCH_CHECK();
        // per curve: (1/b, 1/sqrt(pi b), exp(-1/b)) as the single-curve plan
        // forms them, b/4 for the deconvolution, and the output slab
        std::vector<R> &cp = G->cpar_h, &qb = G->qb_h;
        std::vector<long> &ro = G->ro_h;
        cp.resize(3 * (size_t)nc);
        qb.resize(nc);
        ro.resize(nc);
        for (int k = 0; k < nc; ++k) {
            R b, nf;
            int rf;
            compat_constants<R>(n[cur[k]], g, b, nf, rf);
            cp[3 * k] = R(1) / b;
            cp[3 * k + 1] = (R)(1.0 / std::sqrt(pi * (double)b));
            cp[3 * k + 2] = (R)std::exp(-1.0 / (double)b);
            qb[k] = b * R(0.25);
            ro[k] = (long)cur[k] * B->idist;
        }
        G->cpar.alloc(sizeof(R) * cp.size());
        G->qb_row.alloc(sizeof(R) * qb.size());
        G->row_off.alloc(sizeof(long) * ro.size());
        CUNA_HIP(hipMemcpyAsync(G->cpar.ptr, cp.data(), sizeof(R) * cp.size(),
                                hipMemcpyHostToDevice, s));
        CUNA_HIP(hipMemcpyAsync(G->qb_row.ptr, qb.data(), sizeof(R) * qb.size(),
                                hipMemcpyHostToDevice, s));
        CUNA_HIP(hipMemcpyAsync(G->row_off.ptr, ro.data(), sizeof(long) * ro.size(),
                                hipMemcpyHostToDevice, s));
        P->cpar_c = G->cpar.as<R>();
        P->qb_row_c = G->qb_row.as<R>();
        P->row_off_c = G->row_off.as<long>();
        G->plan = std::move(P);
        B->groups.push_back(std::move(G));
    }
    return B;
}

// Multi-curve plan cache: a batch whose curves all have registered filter
// properties (generate_*filter_properties*) is kept, keyed by x, the sizes
// and those pointers, until one of them is freed; at most kBatchCacheMax
// batches, the least recently used dropped first.  Batches with unregistered
// or NULL filter properties are rebuilt on every call.
namespace {
constexpr size_t kBatchCacheMax = 16;
std::mutex g_batch_mu;
std::list<std::shared_ptr<CompatBatch>> g_batches;  // most recently used first
}  // namespace

static std::shared_ptr<CompatBatch> batch_for(const R *x, const int *n, const int *ng, int nlc,
                                              const filter_properties *const *fps,
                                              hipStream_t s) {
    bool cacheable = fps != nullptr && nlc > 0 && n != nullptr && ng != nullptr;
    for (int i = 0; cacheable && i < nlc; ++i) cacheable = fps[i] && reg_get(fps[i]);
    if (cacheable) {
        std::lock_guard<std::mutex> lk(g_batch_mu);
        for (auto it = g_batches.begin(); it != g_batches.end(); ++it) {
            const CompatBatch &b = **it;
            if (b.x_key == x && (int)b.n.size() == nlc &&
                std::equal(b.n.begin(), b.n.end(), n) && std::equal(b.ng.begin(), b.ng.end(), ng) &&
                std::equal(b.fps.begin(), b.fps.end(), fps)) {
                auto hit = *it;
                g_batches.erase(it);
                g_batches.push_front(hit);
                return hit;
            }
        }
    }
    std::shared_ptr<CompatBatch> fresh(make_compat_batch(x, n, ng, nlc, s));
    if (cacheable) {
        fresh->fps.assign(fps, fps + nlc);
        std::shared_ptr<CompatBatch> victim;
        std::lock_guard<std::mutex> lk(g_batch_mu);
        g_batches.push_front(fresh);
        if (g_batches.size() > kBatchCacheMax) {
            victim = g_batches.back();
            g_batches.pop_back();
        }
    }
    return fresh;
}

// drops the cached batches that use the filter properties d (all if NULL)
static void batch_drop(const void *d) {
    std::list<std::shared_ptr<CompatBatch>> victims;
    std::lock_guard<std::mutex> lk(g_batch_mu);
    for (auto it = g_batches.begin(); it != g_batches.end();) {
        const auto &f = (*it)->fps;
        if (!d || std::find(f.begin(), f.end(), d) != f.end()) {
            victims.push_back(*it);
            it = g_batches.erase(it);
        } else {
            ++it;
        }
    }
}

}  // namespace cuna

using namespace cuna;

// ------------------------------------------------------------------ errors --
CUNA_API int cuna_get_last_error(void) { return g_err; }
CUNA_API const char *cuna_get_last_error_string(void) { return g_msg.c_str(); }
CUNA_API void cuna_clear_last_error(void) { clear_error(); }
CUNA_API void cuna_set_abort_on_error(int enable) { g_abort = enable != 0; }
CUNA_API int cuna_real_size(void) { return (int)sizeof(R); }
CUNA_API int cuna_filter_radius(void) { return sizeof(R) == 4 ? 6 : 12; }

// -------------------------------------------------- filter properties --
CUNA_API void generate_filter_properties(const dTyp *x, const int n, const int ng,
                                         filter_properties **fprops_host,
                                         filter_properties **fprops_device) {
    guarded([&] {
        require(fprops_host && fprops_device, "output pointers are NULL");
        check_compat_sizes(n, ng);
        require(x != nullptr, "x is NULL");
        const int rf = cuna_filter_radius();
        filter_properties *h = (filter_properties *)malloc(sizeof(filter_properties));
        require(h != nullptr, "host allocation failed", CUNA_ERROR_OUT_OF_MEMORY);
        memset(h, 0, sizeof *h);
        filter_properties *d = nullptr;
        try {
            CUNA_HIP(hipMalloc(&d, sizeof(filter_properties)));
            CUNA_HIP(hipMalloc(&h->E1, sizeof(R) * n));
            CUNA_HIP(hipMalloc(&h->E2, sizeof(R) * n));
            CUNA_HIP(hipMalloc(&h->E3, sizeof(R) * rf));
            generate_impl(x, n, ng, h, d, nullptr, true);
            CUNA_HIP(hipDeviceSynchronize());
        } catch (...) {
            if (h->E1) (void)hipFree(h->E1);
            if (h->E2) (void)hipFree(h->E2);
            if (h->E3) (void)hipFree(h->E3);
            if (d) (void)hipFree(d);
            free(h);
            throw;
        }
        *fprops_host = h;
        *fprops_device = d;
    });
}

CUNA_API void generate_pinned_filter_properties(const dTyp *x, const int n, const int ng,
                                                filter_properties *h_fprops,
                                                filter_properties *d_fprops, hipStream_t stream) {
    guarded([&] {
        require(h_fprops && h_fprops->E1 && h_fprops->E2 && h_fprops->E3,
                "h_fprops and its E1/E2/E3 arrays must be allocated by the caller");
        generate_impl(x, n, ng, h_fprops, d_fprops, stream, false);
    });
}

CUNA_API void free_filter_properties(filter_properties *d_fp, filter_properties *fp) {
    guarded([&] {
        if (d_fp) batch_drop(d_fp);
        reg_drop(d_fp);
        if (fp) {
            if (fp->E1) CUNA_HIP(hipFree(fp->E1));
            if (fp->E2) CUNA_HIP(hipFree(fp->E2));
            if (fp->E3) CUNA_HIP(hipFree(fp->E3));
        }
        if (d_fp) CUNA_HIP(hipFree(d_fp));
        free(fp);
    });
}

CUNA_API void set_filter_properties(plan *p) {
    guarded([&] {
        require(p != nullptr, "plan is NULL");
        p->filter_radius = cuna_filter_radius();
        filter_properties *h = nullptr, *d = nullptr;
        generate_filter_properties(p->g_x_data, p->Ndata, p->Ngrid, &h, &d);
        if (g_err != CUNA_SUCCESS) fail(g_err, g_msg);
        p->fprops_host = h;
        p->fprops_device = d;
    });
}

// --------------------------------------------------------- raw transforms --
static void raw_impl(const dTyp *x, const dTyp *f_data, Complex *f_hat, int n, int ng,
                     const filter_properties *fp, hipStream_t s) {
    require(f_hat != nullptr, "f_hat is NULL");
    auto p = compat_plan_for(x, n, ng, fp, s);
    p->adjoint(f_data, reinterpret_cast<cplx<R> *>(f_hat), s);
}

CUNA_API void cunfft_adjoint_raw(const dTyp *x, const dTyp *f_data, Complex *f_hat, const int n,
                                 const int ng, const filter_properties *gpu_fprops) {
    guarded([&] { raw_impl(x, f_data, f_hat, n, ng, gpu_fprops, nullptr); });
}

CUNA_API void cunfft_adjoint_raw_async(const dTyp *x, const dTyp *f_data, Complex *f_hat,
                                       const int n, const int ng,
                                       const filter_properties *gpu_fprops, hipStream_t stream) {
    guarded([&] { raw_impl(x, f_data, f_hat, n, ng, gpu_fprops, stream); });
}

CUNA_API void cunfft_adjoint_raw_async_noalloc(const dTyp *x, const dTyp *f_data, Complex *f_hat,
                                               const int n, const int ng,
                                               const filter_properties *gpu_fprops,
                                               hipStream_t stream) {
    guarded([&] {
        require(f_hat != nullptr, "f_hat is NULL");
        require(n > 0 && ng > 0, "n and ng must be positive");
        require((size_t)ng * sizeof(R) <= kScratchBytes,
                "ng exceeds the scratch of the allocation-free transform (2^22 cells in "
                "libcunaf, 2^21 in libcunad)",
                CUNA_ERROR_UNSUPPORTED);
        // the plan must exist already: building one allocates
        auto p = gpu_fprops ? reg_get(gpu_fprops) : nullptr;
        require(p && p->x_key == x && p->n() == n && p->G == ng,
                "gpu_fprops must come from generate_filter_properties or "
                "generate_pinned_filter_properties for this x, n and ng");
        TwiddleCache<R>::NoAlloc guard;
        p->adjoint_into(f_data, reinterpret_cast<cplx<R> *>(f_hat),
                        static_cast<R *>(scratch_ptr()), stream);
    });
}

CUNA_API void cunfft_adjoint_raw_async_bootstrap(const dTyp *x, const dTyp *f_data,
                                                 Complex *f_hat, const int n, const int ng,
                                                 const int nbs,
                                                 const filter_properties *gpu_fprops,
                                                 hipStream_t stream, const unsigned int seed) {
    guarded([&] {
        require(f_hat != nullptr, "f_hat is NULL");
        require(nbs > 0, "nbs must be positive");
        auto p = compat_plan_for(x, n, ng, gpu_fprops, stream);
        p->bootstrap_compat(f_data, reinterpret_cast<cplx<R> *>(f_hat), nbs, seed, stream);
    });
}

CUNA_API void cunfft_adjoint_raw_batched_async(const dTyp *x, const dTyp *f_data,
                                               Complex *f_hat, const int *n, const int *ng,
                                               const int nlc,
                                               const filter_properties *const *gpu_fprops,
                                               hipStream_t stream) {
    guarded([&] {
        require(f_hat != nullptr, "f_hat is NULL");
        auto b = batch_for(x, n, ng, nlc, gpu_fprops, stream);
        b->run(f_data, reinterpret_cast<cplx<R> *>(f_hat), stream);
    });
}

CUNA_API void generate_pinned_filter_properties_batch(const dTyp *x, const int *n, const int *ng,
                                                      const int nlc, filter_properties **h_fprops,
                                                      filter_properties **d_fprops,
                                                      hipStream_t stream) {
    guarded([&] {
        require(nlc > 0, "nlc must be positive");
        require(x && n && ng && h_fprops && d_fprops, "NULL argument");
        long off = 0;
        for (int i = 0; i < nlc; ++i) {
            require(h_fprops[i] && h_fprops[i]->E1 && h_fprops[i]->E2 && h_fprops[i]->E3,
                    "h_fprops[i] and its E1/E2/E3 arrays must be allocated by the caller");
            check_compat_sizes(n[i], ng[i]);
            generate_impl(x + off, n[i], ng[i], h_fprops[i], d_fprops[i], stream, false);
            off += n[i];
        }
    });
}

CUNA_API void cuna_batch_cache_clear(void) {
    guarded([&] { batch_drop(nullptr); });
}

// ------------------------------------------------------------ plan path --
CUNA_API void init_plan(plan *p, const dTyp *x, const dTyp *f, int Ndata, int Ngrid,
                        unsigned int plan_flags) {
    guarded([&] {
        require(p != nullptr, "plan is NULL");
        require(x != nullptr, "x is NULL");
        check_compat_sizes(Ndata, Ngrid);
        memset(p, 0, sizeof *p);
        p->Ndata = Ndata;
        p->Ngrid = Ngrid;
        p->flags = plan_flags;
        p->x_data = (R *)malloc(sizeof(R) * Ndata);
        p->f_data = (R *)malloc(sizeof(R) * Ndata);
        p->f_hat = (Complex *)malloc(sizeof(Complex) * Ngrid);
        if (plan_flags & CALCULATE_WINDOW_FUNCTION)
            p->f_hat_win = (Complex *)malloc(sizeof(Complex) * Ngrid);
        require(p->x_data && p->f_data && p->f_hat, "host allocation failed",
                CUNA_ERROR_OUT_OF_MEMORY);
        memcpy(p->x_data, x, sizeof(R) * Ndata);
        // src/utils.cu:112-116 leaves f_data uninitialised for f == NULL;
        // here NULL means unit weights
        if (f)
            memcpy(p->f_data, f, sizeof(R) * Ndata);
        else
            for (int i = 0; i < Ndata; ++i) p->f_data[i] = R(1);
        CUNA_HIP(hipMalloc(&p->g_f_data, sizeof(R) * Ndata));
        CUNA_HIP(hipMalloc(&p->g_x_data, sizeof(R) * Ndata));
        CUNA_HIP(hipMalloc(&p->g_f_hat, sizeof(Complex) * Ngrid));
        CUNA_HIP(hipMemset(p->g_f_hat, 0, sizeof(Complex) * Ngrid));
        if (plan_flags & CALCULATE_WINDOW_FUNCTION) {
            CUNA_HIP(hipMalloc(&p->g_f_hat_win, sizeof(Complex) * Ngrid));
            CUNA_HIP(hipMemset(p->g_f_hat_win, 0, sizeof(Complex) * Ngrid));
        }
        CUNA_HIP(hipMemcpy(p->g_f_data, p->f_data, sizeof(R) * Ndata, hipMemcpyHostToDevice));
        CUNA_HIP(hipMemcpy(p->g_x_data, p->x_data, sizeof(R) * Ndata, hipMemcpyHostToDevice));
        CUNA_HIP(hipDeviceSynchronize());
        set_filter_properties(p);
        if (g_err != CUNA_SUCCESS) fail(g_err, g_msg);
    });
}

CUNA_API void cunfft_adjoint_from_plan(plan *p) {
    guarded([&] {
        require(p != nullptr && p->g_x_data && p->g_f_hat, "plan is not initialised");
        raw_impl(p->g_x_data, p->g_f_data, p->g_f_hat, p->Ndata, p->Ngrid, p->fprops_device,
                 nullptr);
        if (p->flags & CALCULATE_WINDOW_FUNCTION)
            raw_impl(p->g_x_data, nullptr, p->g_f_hat_win, p->Ndata, p->Ngrid, p->fprops_device,
                     nullptr);
        // the reference meant to copy back unless DONT_TRANSFER_TO_CPU
        // (src/cuna_kernel.cu:142; its test `!(flags | ...)` is always false)
        if (!(p->flags & DONT_TRANSFER_TO_CPU)) {
            CUNA_HIP(hipMemcpy(p->f_hat, p->g_f_hat, sizeof(Complex) * p->Ngrid,
                               hipMemcpyDeviceToHost));
            if ((p->flags & CALCULATE_WINDOW_FUNCTION) && p->f_hat_win)
                CUNA_HIP(hipMemcpy(p->f_hat_win, p->g_f_hat_win, sizeof(Complex) * p->Ngrid,
                                   hipMemcpyDeviceToHost));
        }
    });
}

CUNA_API void free_plan(plan *p) {
    guarded([&] {
        if (!p) return;
        free(p->f_hat);
        free(p->f_hat_win);
        free(p->x_data);
        free(p->f_data);
        if (p->fprops_device || p->fprops_host) free_filter_properties(p->fprops_device, p->fprops_host);
        if (p->g_f_hat) CUNA_HIP(hipFree(p->g_f_hat));
        if (p->g_f_hat_win) CUNA_HIP(hipFree(p->g_f_hat_win));
        if (p->g_f_data) CUNA_HIP(hipFree(p->g_f_data));
        if (p->g_x_data) CUNA_HIP(hipFree(p->g_x_data));
        free(p);
    });
}

CUNA_API void cunfft_adjoint(const dTyp *x, const dTyp *f, cTyp *fft, int n, int ng,
                             unsigned int flags) {
    guarded([&] {
        require(fft != nullptr, "fft is NULL");
        plan *p = (plan *)malloc(sizeof(plan));
        require(p != nullptr, "host allocation failed", CUNA_ERROR_OUT_OF_MEMORY);
        memset(p, 0, sizeof *p);
        init_plan(p, x, f, n, ng, flags & ~DONT_TRANSFER_TO_CPU);
        int err = g_err;
        std::string msg = g_msg;
        if (err == CUNA_SUCCESS) {
            cunfft_adjoint_from_plan(p);
            err = g_err;
            msg = g_msg;
        }
        if (err == CUNA_SUCCESS) memcpy((void *)fft, p->f_hat, sizeof(Complex) * ng);
        free_plan(p);
        if (err != CUNA_SUCCESS) fail(err, msg);
    });
}

// ------------------------------------------------------------- extensions --
struct cuna_nfft_plan_s {
    std::unique_ptr<Plan<R>> p;
};

static ModeSpec spec_from(const cuna_nfft_params *params) {
    cuna_nfft_params d;
    cuna_nfft_default_params(&d);
    const cuna_nfft_params &q = params ? *params : d;
    ModeSpec s;
    s.mode = q.mode;
    s.sigma = q.sigma;
    s.m = q.m;
    s.tile_log2 = q.tile_log2;
    s.flags = q.flags;
    s.fft_w = q.fft_w;
    s.fft_h = q.fft_h;
    require(s.fft_w >= 0 && s.fft_w <= 8 && s.fft_h >= 0 && s.fft_h <= 4,
            "fft_w must be in [0, 8] and fft_h in [0, 4]");
    require(s.mode == CUNA_MODE_ACCURATE || s.mode == CUNA_MODE_COMPAT, "unknown mode");
    require((s.flags & ~(CUNA_PLAN_WINDOW2 | CUNA_PLAN_UNFUSED | CUNA_PLAN_GATHER)) == 0,
            "unknown plan flags");
    require(s.tile_log2 == 0 || (s.tile_log2 >= 6 && s.tile_log2 <= 14), "tile_log2 out of range");
    return s;
}

CUNA_API void cuna_nfft_default_params(cuna_nfft_params *p) {
    if (!p) return;
    p->mode = CUNA_MODE_ACCURATE;
    p->sigma = 2;
    p->m = sizeof(R) == 4 ? 8 : 13;
    p->tile_log2 = 0;
    p->flags = 0;
    p->fft_w = 0;
    p->fft_h = 0;
}

CUNA_API int cuna_nfft_plan_create(cuna_nfft_plan *plan, const dTyp *x, int n, int M,
                                   const cuna_nfft_params *params, hipStream_t stream) {
    return guarded([&] {
        require(plan != nullptr, "plan is NULL");
        require(n >= 0, "n must be non-negative");
        ModeSpec spec = spec_from(params);
        auto h = new cuna_nfft_plan_s;
        try {
            if (spec.mode == CUNA_MODE_COMPAT)
                h->p = make_plan_compat<R>(x, n, M, nullptr, nullptr, nullptr, spec, stream);
            else
                h->p = make_plan_accurate<R>(x, nullptr, 0.0, {0, n}, M, spec, stream);
        } catch (...) {
            delete h;
            throw;
        }
        *plan = h;
    });
}

CUNA_API int cuna_nfft_plan_create_times(cuna_nfft_plan *plan, const double *t, int n, double df,
                                         int M, const cuna_nfft_params *params,
                                         hipStream_t stream) {
    return guarded([&] {
        require(plan != nullptr, "plan is NULL");
        require(n >= 0, "n must be non-negative");
        require(std::isfinite(df) && df > 0, "df must be positive");
        ModeSpec spec = spec_from(params);
        require(spec.mode == CUNA_MODE_ACCURATE, "times plans are accurate-mode plans");
        require(n == 0 || t != nullptr, "t is NULL");
        auto h = new cuna_nfft_plan_s;
        try {
            h->p = make_plan_accurate<R>(nullptr, t, df, {0, n}, M, spec, stream);
        } catch (...) {
            delete h;
            throw;
        }
        *plan = h;
    });
}

CUNA_API int cuna_nfft_plan_create_batch(cuna_nfft_plan *plan, const dTyp *x, const int *offsets,
                                         int ncurves, int M, const cuna_nfft_params *params,
                                         hipStream_t stream) {
    return guarded([&] {
        require(plan != nullptr && offsets != nullptr, "plan or offsets is NULL");
        require(ncurves >= 1, "ncurves must be positive");
        ModeSpec spec = spec_from(params);
        require(spec.mode == CUNA_MODE_ACCURATE, "batched plans are accurate-mode plans");
        std::vector<int> off(offsets, offsets + ncurves + 1);
        auto h = new cuna_nfft_plan_s;
        try {
            h->p = make_plan_accurate<R>(x, nullptr, 0.0, off, M, spec, stream);
        } catch (...) {
            delete h;
            throw;
        }
        *plan = h;
    });
}

CUNA_API int cuna_nfft_plan_create_batch_times(cuna_nfft_plan *plan, const double *t,
                                               const int *offsets, int ncurves, double df, int M,
                                               const cuna_nfft_params *params,
                                               hipStream_t stream) {
    return guarded([&] {
        require(plan != nullptr && offsets != nullptr, "plan or offsets is NULL");
        require(ncurves >= 1, "ncurves must be positive");
        require(std::isfinite(df) && df > 0, "df must be positive");
        ModeSpec spec = spec_from(params);
        require(spec.mode == CUNA_MODE_ACCURATE, "batched plans are accurate-mode plans");
        std::vector<int> off(offsets, offsets + ncurves + 1);
        require(off.back() == 0 || t != nullptr, "t is NULL");
        auto h = new cuna_nfft_plan_s;
        try {
            h->p = make_plan_accurate<R>(nullptr, t, df, off, M, spec, stream);
        } catch (...) {
            delete h;
            throw;
        }
        *plan = h;
    });
}

CUNA_API int cuna_nfft_adjoint(cuna_nfft_plan plan, const dTyp *y, Complex *out,
                               hipStream_t stream) {
    return guarded([&] {
        require(plan && plan->p, "plan is NULL");
        plan->p->adjoint(y, reinterpret_cast<cplx<R> *>(out), stream);
    });
}

CUNA_API int cuna_nfft_adjoint_many(cuna_nfft_plan plan, const dTyp *y, int nrhs, Complex *out,
                                    hipStream_t stream) {
    return guarded([&] {
        require(plan && plan->p, "plan is NULL");
        plan->p->adjoint_many(y, nrhs, reinterpret_cast<cplx<R> *>(out), stream);
    });
}

CUNA_API int cuna_nfft_bootstrap(cuna_nfft_plan plan, const dTyp *y, int nbs, unsigned int seed,
                                 Complex *out_y, Complex *out_w, hipStream_t stream) {
    return guarded([&] {
        require(plan && plan->p, "plan is NULL");
        plan->p->bootstrap(y, nbs, seed, reinterpret_cast<cplx<R> *>(out_y),
                           reinterpret_cast<cplx<R> *>(out_w), stream);
    });
}

CUNA_API int cuna_nfft_bootstrap_ls(cuna_nfft_plan plan, const dTyp *y, int nbs,
                                    unsigned int seed, dTyp *out_power, dTyp *out_peak,
                                    int *out_kpeak, hipStream_t stream) {
    return guarded([&] {
        require(plan && plan->p, "plan is NULL");
        plan->p->bootstrap_ls(y, nbs, seed, out_power, out_peak, out_kpeak, stream);
    });
}

CUNA_API int cuna_nfft_forward(cuna_nfft_plan plan, const Complex *fhat, Complex *f,
                               hipStream_t stream) {
    return guarded([&] {
        require(plan && plan->p, "plan is NULL");
        plan->p->forward(reinterpret_cast<const cplx<R> *>(fhat), reinterpret_cast<cplx<R> *>(f),
                         stream);
    });
}

CUNA_API int cuna_nfft_backward(cuna_nfft_plan plan, const dTyp *y, const Complex *grad_out,
                                dTyp *grad_y, dTyp *grad_x, hipStream_t stream) {
    return guarded([&] {
        require(plan && plan->p, "plan is NULL");
        plan->p->backward(y, reinterpret_cast<const cplx<R> *>(grad_out), grad_y, grad_x, stream);
    });
}

CUNA_API int cuna_nfft_plan_info(cuna_nfft_plan plan, cuna_nfft_info *info) {
    return guarded([&] {
        require(plan && plan->p && info, "plan or info is NULL");
        const Plan<R> &p = *plan->p;
        info->mode = p.spec.mode;
        info->sigma = p.spec.mode == CUNA_MODE_ACCURATE ? p.spec.sigma : 0;
        info->m = p.spec.mode == CUNA_MODE_ACCURATE ? p.spec.m : p.R_f;
        info->ntaps = p.ntaps;
        info->n = (int)p.n();
        info->ncurves = p.ncurves();
        info->M = p.M;
        info->G = p.G;
        info->tile_log2 = p.pts.tlog;
        info->ntiles = p.pts.ntiles;
        info->beta = (double)p.beta;
    });
}

CUNA_API int cuna_nfft_set_profile_events(cuna_nfft_plan plan, hipEvent_t *events, int nevents) {
    return guarded([&] {
        require(plan && plan->p, "plan is NULL");
        require(nevents >= 0 && nevents <= 4 && (nevents == 0 || events), "bad event list");
        for (int i = 0; i < 4; ++i) plan->p->prof[i] = i < nevents ? events[i] : nullptr;
    });
}

CUNA_API int cuna_nfft_plan_destroy(cuna_nfft_plan plan) {
    return guarded([&] { delete plan; });
}

// ------------------------------------------------------------- graphs --
struct cuna_graph_s {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
};

CUNA_API int cuna_stream_capture_begin(hipStream_t stream) {
    return guarded([&] {
        require(stream != nullptr, "capture needs a created stream (not the null stream)");
        CUNA_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    });
}

CUNA_API int cuna_stream_capture_end(hipStream_t stream, cuna_graph *graph) {
    return guarded([&] {
        require(graph != nullptr, "graph is NULL");
        auto G = new cuna_graph_s;
        hipError_t e = hipStreamEndCapture(stream, &G->g);
        if (e != hipSuccess) {
            delete G;
            hip_check(e, "hipStreamEndCapture", __FILE__, __LINE__);
        }
        e = hipGraphInstantiate(&G->e, G->g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(G->g);
            delete G;
            hip_check(e, "hipGraphInstantiate", __FILE__, __LINE__);
        }
        *graph = G;
    });
}

CUNA_API int cuna_graph_launch(cuna_graph graph, hipStream_t stream) {
    return guarded([&] {
        require(graph && graph->e, "graph is NULL");
        CUNA_HIP(hipGraphLaunch(graph->e, stream));
    });
}

CUNA_API int cuna_graph_destroy(cuna_graph graph) {
    return guarded([&] {
        if (!graph) return;
        if (graph->e) CUNA_HIP(hipGraphExecDestroy(graph->e));
        if (graph->g) CUNA_HIP(hipGraphDestroy(graph->g));
        delete graph;
    });
}

CUNA_API int cuna_fft(Complex *data, int n, int batch, int sign, hipStream_t stream) {
    return guarded([&] {
        require(data != nullptr, "data is NULL");
        require(n >= 1, "n must be positive");
        require(batch >= 1, "batch must be positive");
        require(sign == 1 || sign == -1, "sign must be +1 or -1");
        if (n < 2 || !is_pow2(n)) {
            // any other length: Bluestein, set up for this call (as the
            // reference creates its cuFFT plan on every call); the call
            // waits for its stream before the workspace is released
            Bluestein<R> b;
            b.init(n, stream);
            b.ensure(batch);
            b.run_complex(reinterpret_cast<cplx<R> *>(data), batch, sign, stream);
            CUNA_HIP(hipStreamSynchronize(stream));
            return;
        }
        FftShape sh;
        sh.lgN = ilog2(n);
        sh.batch = batch;
        DevBuf work;
        if (sh.lgN > max_small_lg<R>()) work.alloc(sizeof(cplx<R>) * (size_t)n * batch, stream);
        fft_complex<R>(sh, reinterpret_cast<cplx<R> *>(data), work.as<cplx<R>>(), sign, stream);
    });
}
