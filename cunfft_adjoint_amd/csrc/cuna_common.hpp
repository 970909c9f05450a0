// cuna_common.hpp -- shared definitions of the HIP library (both precisions).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <new>
#include <string>

#include "cuna_typedefs.h"

namespace cuna {

// ------------------------------------------------------------------ errors --
// Thread-local last error.  The reference exits the process on any CUDA or
// cuFFT error (inc/cuna_typedefs.h:80-89, src/cufft_utils.cu:78-83); here the
// error is recorded and, only when abort-on-error is enabled, fatal.
void set_error(int code, const std::string &msg);
void clear_error();
bool abort_on_error();

struct Error {
    int code;
    std::string msg;
};

// Throws cuna::Error; every extern "C" entry point catches it and records it.
[[noreturn]] void fail(int code, const std::string &msg);

inline void hip_check(hipError_t e, const char *what, const char *file, int line) {
    if (e != hipSuccess) {
        char buf[512];
        snprintf(buf, sizeof buf, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
        fail(e == hipErrorOutOfMemory ? CUNA_ERROR_OUT_OF_MEMORY : CUNA_ERROR_HIP, buf);
    }
}
#define CUNA_HIP(call) ::cuna::hip_check((call), #call, __FILE__, __LINE__)
#define CUNA_LAUNCH_CHECK() ::cuna::hip_check(hipGetLastError(), "kernel launch", __FILE__, __LINE__)

inline void require(bool ok, const std::string &msg, int code = CUNA_ERROR_INVALID_VALUE) {
    if (!ok) fail(code, msg);
}

// Runs f, turning a cuna::Error (or any exception) into the recorded
// thread-local status; returns that status.  Every extern "C" entry uses it.
template <class F>
int guarded(F &&f) {
    try {
        clear_error();
        f();
        return CUNA_SUCCESS;
    } catch (const Error &e) {
        set_error(e.code, e.msg);
        return e.code;
    } catch (const std::bad_alloc &) {
        set_error(CUNA_ERROR_OUT_OF_MEMORY, "host allocation failed");
        return CUNA_ERROR_OUT_OF_MEMORY;
    } catch (...) {
        set_error(CUNA_ERROR_HIP, "unexpected exception");
        return CUNA_ERROR_HIP;
    }
}

// --------------------------------------------------------------- complex --
template <typename T>
struct alignas(2 * sizeof(T)) cplx {
    T x, y;
};

template <typename T>
__host__ __device__ __forceinline__ cplx<T> cmk(T a, T b) {
    cplx<T> r;
    r.x = a;
    r.y = b;
    return r;
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b) {
    return cmk(a.x + b.x, a.y + b.y);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b) {
    return cmk(a.x - b.x, a.y - b.y);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator*(cplx<T> a, cplx<T> b) {
    return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator*(cplx<T> a, T s) {
    return cmk(a.x * s, a.y * s);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> conjg(cplx<T> a) {
    return cmk(a.x, -a.y);
}
// multiply by +i
template <typename T>
__host__ __device__ __forceinline__ cplx<T> mul_i(cplx<T> a) {
    return cmk(-a.y, a.x);
}

// ------------------------------------------------------------ device math --
__device__ __forceinline__ float dexp(float v) { return __expf(v); }
__device__ __forceinline__ double dexp(double v) { return exp(v); }

// Load from memory that no kernel running at the same time writes (plan
// tables), through the constant address space: with a uniform address the
// compiler then issues a scalar load (s_load), which waits on its own counter
// instead of behind the wave's outstanding vector loads.
template <typename U>
__device__ __forceinline__ U ldc(const U *p) {
    using CP = const __attribute__((address_space(4))) U *;
    return *__builtin_bit_cast(CP, p);
}

// ----------------------------------------------------------- launch shape --
// MI355X has 8 XCDs, each with its own L2; consecutive workgroup ids are
// dispatched round-robin over the XCDs.  Remapping makes neighbouring logical
// blocks (which share cache lines) run on the same XCD.
constexpr int kNumXcd = 8;
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
    if (nblocks < kNumXcd * 2) return bid;
    int per = nblocks / kNumXcd;  // blocks per XCD in the regular part
    int regular = per * kNumXcd;
    if (bid >= regular) return bid;
    return (bid % kNumXcd) * per + bid / kNumXcd;
}

inline int ilog2(long v) {
    int r = 0;
    while ((1L << r) < v) r++;
    return r;
}
inline bool is_pow2(long v) { return v > 0 && (v & (v - 1)) == 0; }
inline long ceil_div(long a, long b) { return (a + b - 1) / b; }

// ----------------------------------------------------------- allocation --
// Stream-ordered device buffer (hipMallocAsync / hipFreeAsync).
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    hipStream_t stream = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    void alloc(size_t nbytes, hipStream_t s) {
        release();
        stream = s;
        bytes = nbytes;
        if (nbytes) CUNA_HIP(hipMallocAsync(&ptr, nbytes, s));
    }
    void release() {
        if (ptr) (void)hipFreeAsync(ptr, stream);
        ptr = nullptr;
        bytes = 0;
    }
    ~DevBuf() { release(); }
    template <typename U>
    U *as() const {
        return static_cast<U *>(ptr);
    }
};

}  // namespace cuna
