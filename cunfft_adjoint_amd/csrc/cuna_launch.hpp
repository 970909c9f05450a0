// cuna_launch.hpp -- host launchers for the FFT and spread kernels.
// Defined in cuna_kernels.hip, instantiated for the library's real type.
#pragma once

#include "cuna_common.hpp"
#include "cuna_fft.hpp"
#include "cuna_spread.hpp"

namespace cuna {

// complex length N = 2^lgN of the real grid G = 2N, `batch` transforms
struct FftShape {
    int lgN = 0;
    int batch = 1;
    int w = 0;  // columns per workgroup in pass 1 (0: heuristic)
    int h = 0;  // rows (real: row pairs) per workgroup in pass 2 (0: heuristic)
};

template <typename T>
constexpr int max_small_lg() {
    return sizeof(T) == 4 ? 14 : 13;
}
template <typename T>
constexpr int max_fft_lg() {
    return sizeof(T) == 4 ? 24 : 23;
}

// Real grid (batch x G reals, read as batch x N complex, modified in place
// when two passes are used) -> spectrum through the sink.
// `mid` (optional) is recorded on s between the two passes of a two-pass FFT.
template <typename T>
void fft_real_forward(const FftShape &sh, T *grid, const SinkAccurate<T> &sk, hipStream_t s,
                      hipEvent_t mid = nullptr);
template <typename T>
void fft_real_forward(const FftShape &sh, T *grid, const SinkCompat<T> &sk, hipStream_t s,
                      hipEvent_t mid = nullptr);

// Hermitian spectrum (loader) -> real grid (batch x G reals) by an inverse
// (sign -1) transform; `work` holds batch x N complex.
template <typename T>
void fft_c2r_inverse(const FftShape &sh, const LoadHermitian<T> &ld, cplx<T> *work, T *grid,
                     hipStream_t s);

// Plain complex transform of batch sequences of N = 2^lgN values, in place
// (used by tests of the FFT stage through the library).
template <typename T>
void fft_complex(const FftShape &sh, cplx<T> *data, cplx<T> *work, int sign, hipStream_t s);

// sort-by-tile plan building
struct SortWork {
    const int4 *chunks;
    const int *chunk_base;
    const int *nch;
    const int *coff;
    int nchunks, ncurves, total_chunks;
    int *hist;  // >= K * total_chunks entries for the largest K used
    int *bsum;
};
template <typename T>
void launch_point_setup(const SetupArgs<T> &A, hipStream_t s);
// one stable counting-sort pass over keys in [0, K)
template <typename T>
void launch_sort_pass(const SortWork &W, int K, const int *keys, const PointRec<T> *rec_in,
                      const int *perm_in, PointRec<T> *rec_out, int *perm_out, hipStream_t s);
// key = (anchor >> shift) & mask, the digits of the tile index
template <typename T>
void launch_digit_keys(const PointRec<T> *rec, long n, int shift, int mask, int *key,
                       hipStream_t s);
// first sorted index of every (curve, 3 * tile + edge class) key, from points
// sorted by anchor
template <typename T>
void launch_buckets(const int *coff, int ncurves, int Kc, const PointRec<T> *rec, int tlog,
                    int olo, int ohi, int G, int *bucket, hipStream_t s);

// ev0/ev1 (optional) receive the spread kernel's own start and stop times
template <typename T>
void launch_spread(const SpreadArgs<T> &A, int ntaps, int vm, int nslabs, hipStream_t s,
                   hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
template <typename T>
void launch_gather(const GatherArgs<T> &A, int ntaps, int nslabs, hipStream_t s);
// per (curve, tile) point cover of the spread's scale (A: rec, bucket, tiles)
template <typename T>
void launch_tile_cover(const SpreadArgs<T> &A, int ncurves, int *cover, hipStream_t s);
// per curve, the scatter's work order (tiles by decreasing point count)
void launch_tile_order(const int *bucket, int Kc, int ntiles, int ncurves, bool periodic,
                       const int *cover, TileDesc *desc, hipStream_t s);

// fused batched bootstrap (grid in LDS): false when this size/filter has no
// fused kernel
template <typename T>
bool boot_fused_supported(int lgN, int ntaps);
// per-curve densest tap window (G <= 2^15) and max |y|
template <typename T>
void launch_curve_cover(const PointRec<T> *rec, const int *coff, int ncurves, int G, int ntaps,
                        int *cover, hipStream_t s);
template <typename T>
void launch_curve_absmax(const T *y, const int *coff, int ncurves, T *out, hipStream_t s);
template <typename T>
// mode 0: spectra of y at x and of the window at 2x; 1: Lomb-Scargle power
void launch_boot_fused(BootArgs<T> A, int lgN, int ntaps, int mode, hipStream_t s);

// bytes of dynamic LDS a launch may use on gfx950
constexpr size_t kMaxLds = 160 * 1024;
// the largest grid y dimension; launches over more slabs or batches split
constexpr int kMaxGridY = 65535;

}  // namespace cuna
