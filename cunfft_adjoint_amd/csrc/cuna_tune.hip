// cuna_tune.hip -- autotuner for accurate-mode plans.
//
// For a point set and an accuracy target it picks the oversampling sigma and
// filter half-width m (the smallest m per sigma whose Gaussian-gridding error
// bound meets the target), then measures on the device the spread tile size
// and the FFT workgroup shapes, and keeps the fastest combination.  Results
// are cached in a text file keyed by device, precision, n (rounded to a power
// of two), M and the target.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <vector>

#include "cuna_nfft.h"
#include "cuna_plan.hpp"

#define CUNA_API extern "C" __attribute__((visibility("default")))

namespace cuna {
namespace {

using R = dTyp;

// relative L2 error of Gaussian gridding with 2m taps at oversampling sigma
// (Greengard & Lee 2004: exp(-pi m (1 - 1/(2 sigma - 1)))), plus the
// rounding floor of the arithmetic
double error_bound(int sigma, int m) {
    const double pi = 3.14159265358979323846;
    double e = std::exp(-pi * m * (1.0 - 1.0 / (2.0 * sigma - 1.0)));
    double floor = sizeof(R) == 4 ? 4e-7 : 2e-15 * m;
    return e + floor;
}

__global__ void k_fill(R *p, long n, R v) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

std::string cache_path_of(const char *path) {
    if (path) return path;
    if (const char *e = getenv("CUNA_TUNE_CACHE")) return e;
    const char *home = getenv("HOME");
    return std::string(home ? home : "/tmp") + "/.cache/cuna_tune.txt";
}

std::string key_of(int n, int M, double tol) {
    int dev = 0;
    CUNA_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    CUNA_HIP(hipGetDeviceProperties(&prop, dev));
    std::ostringstream k;
    k << prop.gcnArchName << "|cu" << prop.multiProcessorCount << "|" << (sizeof(R) == 4 ? "f" : "d")
      << "|n2^" << ilog2(std::max(n, 1)) << "|M" << M << "|tol";
    char buf[32];
    snprintf(buf, sizeof buf, "%.1e", tol);
    k << buf;
    std::string s = k.str();
    for (auto &ch : s)
        if (ch == ' ') ch = '_';
    return s;
}

std::mutex g_cache_mu;

bool cache_get(const std::string &file, const std::string &key, cuna_nfft_params &p) {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    std::ifstream in(file);
    if (!in) return false;
    std::string line, k;
    bool found = false;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        int sg, m, tl, w, h;
        if (ls >> k >> sg >> m >> tl >> w >> h && k == key) {
            p.sigma = sg;
            p.m = m;
            p.tile_log2 = tl;
            p.fft_w = w;
            p.fft_h = h;
            found = true;  // the last entry wins
        }
    }
    return found;
}

void cache_put(const std::string &file, const std::string &key, const cuna_nfft_params &p) {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    std::ofstream out(file, std::ios::app);
    if (!out) return;  // an unwritable cache is not an error
    out << key << " " << p.sigma << " " << p.m << " " << p.tile_log2 << " " << p.fft_w << " "
        << p.fft_h << "\n";
}

struct Timer {
    hipEvent_t ev[4] = {};
    Timer() {
        for (auto &e : ev) CUNA_HIP(hipEventCreate(&e));
    }
    ~Timer() {
        for (auto &e : ev) (void)hipEventDestroy(e);
    }
    // median over reps of (stage b end - stage a start) of plan.adjoint
    float run(Plan<R> &P, const R *y, cplx<R> *out, int a, int b, hipStream_t s, int reps) {
        for (int i = 0; i < 2; ++i) P.adjoint(y, out, s);
        std::vector<float> t;
        for (int r = 0; r < reps; ++r) {
            for (int i = 0; i < 4; ++i) P.prof[i] = ev[i];
            P.adjoint(y, out, s);
            for (int i = 0; i < 4; ++i) P.prof[i] = nullptr;
            CUNA_HIP(hipStreamSynchronize(s));
            float ms = 0;
            CUNA_HIP(hipEventElapsedTime(&ms, ev[a], ev[b]));
            t.push_back(ms);
        }
        std::sort(t.begin(), t.end());
        return t[t.size() / 2];
    }
};

}  // namespace
}  // namespace cuna

using namespace cuna;

CUNA_API int cuna_nfft_autotune(const dTyp *x, int n, int M, double tol, const char *cache_path,
                                cuna_nfft_params *params, hipStream_t stream) {
    return guarded([&] {
        require(params != nullptr, "params is NULL");
        require(n > 0 && x != nullptr, "x must hold n > 0 points");
        require(M >= 2 && is_pow2(M), "M must be a power of two >= 2");
        require(tol > 0 && tol < 1, "tol must be in (0, 1)");
        cuna_nfft_params best;
        cuna_nfft_default_params(&best);
        const std::string file = cache_path_of(cache_path);
        const bool use_cache = !(cache_path && cache_path[0] == 0);
        const std::string key = key_of(n, M, tol);
        if (use_cache && cache_get(file, key, best)) {
            *params = best;
            return;
        }
        // candidates: smallest m per sigma meeting the bound
        std::vector<std::pair<int, int>> cands;
        for (int sigma : {2, 4}) {
            if ((long)sigma * M / 2 > (1L << max_fft_lg<R>())) continue;
            for (int m = 2; m <= kMaxTaps / 2; ++m)
                if (error_bound(sigma, m) <= tol) {
                    cands.push_back({sigma, m});
                    break;
                }
        }
        require(!cands.empty(), "no filter meets this accuracy target in this precision",
                CUNA_ERROR_UNSUPPORTED);
        DevBuf y, out;
        y.alloc(sizeof(R) * n, stream);
        out.alloc(sizeof(cplx<R>) * M, stream);
        hipLaunchKernelGGL(k_fill, dim3(ceil_div(n, 256)), dim3(256), 0, stream, y.as<R>(), (long)n,
                           R(1));
        CUNA_LAUNCH_CHECK();
        Timer tm;
        double best_ms = 1e30;
        for (auto [sigma, m] : cands) {
            ModeSpec spec;
            spec.sigma = sigma;
            spec.m = m;
            // spread: tile size
            int best_t = 9;
            float best_spread = 1e30f;
            std::unique_ptr<Plan<R>> keep;
            for (int tl : {8, 9, 10, 11}) {
                spec.tile_log2 = tl;
                auto P = make_plan_accurate<R>(x, nullptr, 0.0, {0, n}, M, spec, stream);
                float t = tm.run(*P, y.as<R>(), out.as<cplx<R>>(), 0, 1, stream, 5);
                if (t < best_spread) {
                    best_spread = t;
                    best_t = tl;
                    keep = std::move(P);
                }
            }
            // FFT workgroup shapes (two-pass sizes only)
            int bw = 0, bh = 0;
            float best_fft = tm.run(*keep, y.as<R>(), out.as<cplx<R>>(), 1, 3, stream, 5);
            if (keep->lgN > max_small_lg<R>()) {
                for (int w : {1, 2, 4, 8})
                    for (int h : {1, 2, 4}) {
                        keep->spec.fft_w = w;
                        keep->spec.fft_h = h;
                        float t = tm.run(*keep, y.as<R>(), out.as<cplx<R>>(), 1, 3, stream, 5);
                        if (t < best_fft) {
                            best_fft = t;
                            bw = w;
                            bh = h;
                        }
                    }
            }
            const double total = (double)best_spread + best_fft;
            if (total < best_ms) {
                best_ms = total;
                best.sigma = sigma;
                best.m = m;
                best.tile_log2 = best_t;
                best.fft_w = bw;
                best.fft_h = bh;
            }
        }
        CUNA_HIP(hipStreamSynchronize(stream));
        if (use_cache) cache_put(file, key, best);
        *params = best;
    });
}
