// cuna_spread.hpp -- sort-by-tile Gaussian spread (and its transpose, the
// gather of the backward pass) for gfx950.
//
// Replaces fast_gridding / fast_gridding_bootstrap (src/gridding.cu:36-117),
// which scatter every tap of every point with a global atomicAdd.  Here the
// points of one x are binned once per plan by a stable two-pass counting sort
// into buckets key = 3*tile + class (class 0: taps reach the previous tile,
// 2: the next tile, 1: interior), ordered by anchor within a bucket.  One
// workgroup owns one tile: it takes the points of its own bucket plus the
// boundary classes of its two neighbours and accumulates the taps that land
// in its tile in LDS, then writes the tile once with coalesced stores.  No
// global atomics, and every accumulation form is bitwise reproducible:
//   k_spread_wv (default): one wave per tile, integer atomics in fixed point
//     (32-bit cells in FP32, 64-bit in FP64; integer sums do not depend on
//     their order), tiles taken densest first;
//   k_spread (on request, CUNA_PLAN_GATHER): gather, each cell summed by
//     one thread in a fixed order.
// The fused bootstrap kernel (k_boot_fused) keeps a whole curve's grid in
// LDS; the backward's gather (k_gather) is the transpose of the spread.
#pragma once

#include <type_traits>

#include "cuna_common.hpp"
#include "cuna_fft.hpp"

namespace cuna {

constexpr int kMaxTaps = 32;

template <typename T>
struct PointRec {
    int a;  // anchor cell
    T e;    // weight parameter: taps at a+o weigh exp(-(o-e)^2/beta)
};

// ---------------------------------------------------- bootstrap RNG --
// src/gridding.cu:67-82
__device__ __host__ __forceinline__ uint32_t wang_hash(uint32_t s) {
    s = (s ^ 61u) ^ (s >> 16);
    s *= 9u;
    s = s ^ (s >> 4);
    s *= 0x27d4eb2du;
    s = s ^ (s >> 15);
    return s;
}
__device__ __host__ __forceinline__ uint32_t xorshift32(uint32_t s) {
    s ^= (s << 13);
    s ^= (s >> 17);
    s ^= (s << 5);
    return s;
}

// x mod d for 32-bit unsigned x and a divisor 1 <= d < 2^31 shared by many x,
// exact (Granlund & Montgomery 1994, unsigned division by a run-time
// invariant): l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1,
// q = (t + ((x - t) >> min(l,1))) >> max(l-1,0) with t = mulhi(m, x).
struct FastMod {
    uint32_t d, m;
    int s1, s2;
    __device__ __host__ void init(uint32_t d_) {
        d = d_;
        int l = 0;
        while ((1ull << l) < d) ++l;
        m = (uint32_t)(((1ull << 32) * ((1ull << l) - d)) / d + 1);
        s1 = l < 1 ? l : 1;
        s2 = l > 1 ? l - 1 : 0;
    }
    __device__ __forceinline__ uint32_t operator()(uint32_t x) const {
        const uint32_t t = __umulhi(m, x);
        const uint32_t q = (t + ((x - t) >> s1)) >> s2;
        return x - q * d;
    }
};

// ------------------------------------------------------ point setup --
enum SetupMode { SETUP_COMPAT = 0, SETUP_ACC_X = 1, SETUP_ACC_T = 2 };

template <typename T>
struct SetupArgs {
    const T *x;        // SETUP_COMPAT / SETUP_ACC_X
    const double *t;   // SETUP_ACC_T: x = frac(df * t) computed in FP64
    double df;
    int mode;
    int G;             // grid cells per curve (ng in compat)
    int R_f;           // compat filter radius
    T b, e1s;          // compat b and 1/sqrt(pi b)
    T *E1, *E2, *E3;   // compat filter_properties arrays (nullable)
    T *eps_orig;       // compat: eps by original index (nullable)
    long n;
    int tlog;          // tile = 2^tlog cells
    PointRec<T> *rec;  // out: (anchor, e) by original index
    int *key;          // out: anchor & (tile - 1), the first sort digit
};

// compat decisions in the build's precision, one rounding per operation
// (src/gaussian_filter.cu:213-216): u = (int)(ng*x - R_f), eps = ng*x - u;
// src/gridding.cu:43-45: j = (int)((x + 0.5)*(ng - 1)) in double, clamped.
template <typename T>
__device__ __forceinline__ void compat_point(T x, int ng, int R_f, int &j, T &eps) {
#pragma clang fp contract(off)
    T prod = (T)ng * x;
    T v = prod - (T)R_f;
    int u = (int)v;
    eps = prod - (T)u;
    j = (int)(((double)x + 0.5) * (double)(ng - 1));
    if (j >= ng) j = ng - 1;
    if (j < 0) j = 0;
}

template <typename T>
__device__ __forceinline__ void setup_point(const SetupArgs<T> &A, long i, int &a, T &e) {
    if (A.mode == SETUP_COMPAT) {
        compat_point<T>(A.x[i], A.G, A.R_f, a, e);
        if (A.E1) {
            // src/gaussian_filter.cu:216-220
            T e2 = e / A.b;
            A.E1[i] = (T)exp(-e * e2) * A.e1s;
            A.E2[i] = (T)exp(2 * e2);
        }
        if (A.eps_orig) A.eps_orig[i] = e;
    } else {
        // accurate: t = G*x, anchor floor(t) mod G, e = t - floor(t).  With G
        // a power of two, G*x and t - floor(t) are exact in FP64 and the
        // anchor is reduced in integers: no bit of x is lost (reducing x to
        // [0, 1) first would round x + 1 for negative x).  Times: x = df*t
        // in FP64 (one rounding).
        const double xv = A.mode == SETUP_ACC_T ? A.df * A.t[i] : (double)A.x[i];
        const double tg = xv * (double)A.G;
        const double c = floor(tg);
        T f = (T)(tg - c);
        // (non-finite x: anchor 0 and a NaN e, which makes the output NaN)
        long long ci = fabs(c) < 0x1p62 ? (long long)c : 0;
        if (f >= T(1)) {  // rounding of tg - c to T
            f = T(0);
            ci += 1;
        }
        a = (int)(ci & (long long)(A.G - 1));  // G is a power of two
        e = f;
    }
}

template <typename T>
__global__ void k_point_setup(SetupArgs<T> A) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (A.E3 && i < A.R_f) {
        // src/gaussian_filter.cu:222-226: E3[m] = exp(-m^2/b)
        const int m = (int)i;
        A.E3[m] = (T)exp(-(T)(m * m) / A.b);
    }
    if (i >= A.n) return;
    int a;
    T e;
    setup_point(A, i, a, e);
    PointRec<T> pr;
    pr.a = a;
    pr.e = e;
    A.rec[i] = pr;
    A.key[i] = a & ((1 << A.tlog) - 1);
}

__device__ __forceinline__ int tile_key(int a, int tlog, int olo, int ohi, int G) {
    int tile = a >> tlog;
    int ts = tile << tlog;
    int te = min(ts + (1 << tlog), G);
    int cls = (a + olo < ts) ? 0 : ((a + ohi >= te) ? 2 : 1);
    return 3 * tile + cls;
}

// sort digits after the first: bits [shift, shift + log2(mask + 1)) of the
// tile index a >> tlog
template <typename T>
__global__ void k_digit_keys(const PointRec<T> *rec, long n, int shift, int mask, int *key) {
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) key[s] = (rec[s].a >> shift) & mask;
}

// A stable counting sort pass over K keys.  Points are cut into chunks that
// never cross a curve; the per-chunk counts form a key-major matrix
// hist[curve base + key * nch + chunk], whose exclusive scan gives every
// chunk's first position for every key.  One wave handles one chunk.
struct SortArgs {
    const int4 *chunks;     // {curve, start, len, column}
    const int *chunk_base;  // chunks before each curve
    const int *nch;         // chunks per curve
    int nchunks, K, kbits;
    const int *keys;
    int *hist;
};

__device__ __forceinline__ long hist_index(const SortArgs &S, int4 cd, int k) {
    return (long)S.K * S.chunk_base[cd.x] + (long)k * S.nch[cd.x] + cd.w;
}

static __global__ void k_hist(SortArgs S) {
    int *cnt_all = reinterpret_cast<int *>(cuna_smem);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ch = blockIdx.x * (blockDim.x >> 6) + wave;
    int *cnt = cnt_all + wave * S.K;
    for (int k = lane; k < S.K; k += 64) cnt[k] = 0;
    __syncthreads();
    int4 cd = make_int4(0, 0, 0, 0);
    if (ch < S.nchunks) {
        cd = S.chunks[ch];
        for (int q = lane; q < cd.z; q += 64) atomicAdd(&cnt[S.keys[(long)cd.y + q]], 1);
    }
    __syncthreads();
    if (ch < S.nchunks)
        for (int k = lane; k < S.K; k += 64) S.hist[hist_index(S, cd, k)] = cnt[k];
}

// scatter with ranks that keep the input order among equal keys
template <typename T>
__global__ void k_scatter(SortArgs S, const PointRec<T> *rec_in, const int *perm_in,
                          PointRec<T> *rec_out, int *perm_out) {
    int *cur_all = reinterpret_cast<int *>(cuna_smem);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ch = blockIdx.x * (blockDim.x >> 6) + wave;
    int *cur = cur_all + wave * S.K;
    int4 cd = make_int4(0, 0, 0, 0);
    if (ch < S.nchunks) {
        cd = S.chunks[ch];
        for (int k = lane; k < S.K; k += 64) cur[k] = S.hist[hist_index(S, cd, k)];
    }
    __syncthreads();
    if (ch < S.nchunks) {
        const unsigned long long lt = (1ull << lane) - 1ull;
        for (int r0 = 0; r0 < cd.z; r0 += 64) {
            const int q = r0 + lane;
            const bool act = q < cd.z;
            const long i = (long)cd.y + q;
            const int key = act ? S.keys[i] : 0;
            unsigned long long peers = __ballot(act);
            for (int bit = 0; bit < S.kbits; ++bit) {
                const int bv = (key >> bit) & 1;
                unsigned long long m = __ballot(bv);
                peers &= bv ? m : ~m;
            }
            const int rank = __popcll(peers & lt);
            const int gcount = __popcll(peers);
            int pos = 0;
            if (act) pos = cur[key] + rank;
            __builtin_amdgcn_wave_barrier();
            if (act) {
                rec_out[pos] = rec_in[i];
                perm_out[pos] = perm_in ? perm_in[i] : (int)i;
                if (rank == gcount - 1) cur[key] = pos + 1;
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// bucket[c*(Kc+1) + k] = first sorted index of the curve's points with
// tile_key >= k (k = Kc: the curve's end).  The points are sorted by anchor,
// and tile_key is non-decreasing in the anchor (the class is 0 for a prefix
// of the tile's cells and 2 for a suffix), so a binary search finds it.
template <typename T>
__global__ void k_buckets(const int *coff, int ncurves, int Kc, const PointRec<T> *rec, int tlog,
                          int olo, int ohi, int G, int *bucket) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)ncurves * (Kc + 1)) return;
    const int c = (int)(idx / (Kc + 1)), k = (int)(idx % (Kc + 1));
    int lo = k < Kc ? coff[c] : coff[c + 1], hi = coff[c + 1];
    while (lo < hi) {
        const int mid = lo + ((hi - lo) >> 1);
        if (tile_key(rec[mid].a, tlog, olo, ohi, G) < k)
            lo = mid + 1;
        else
            hi = mid;
    }
    bucket[idx] = lo;
}

// ------------------------------------------------ exclusive int scan --
constexpr int kScanNT = 1024, kScanIPT = 4, kScanTile = kScanNT * kScanIPT;

__device__ __forceinline__ int block_exclusive_scan(int v, int *sh, int &total) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int off = 1; off < kScanNT; off <<= 1) {
        int add = tid >= off ? sh[tid - off] : 0;
        __syncthreads();
        sh[tid] += add;
        __syncthreads();
    }
    total = sh[kScanNT - 1];
    int incl = sh[tid];
    __syncthreads();
    return incl - v;
}

static __global__ __launch_bounds__(kScanNT) void k_scan_reduce(const int *in, long n, int *bsum) {
    int *sh = reinterpret_cast<int *>(cuna_smem);
    long base = (long)blockIdx.x * kScanTile + (long)threadIdx.x * kScanIPT;
    int s = 0;
#pragma unroll
    for (int q = 0; q < kScanIPT; ++q)
        if (base + q < n) s += in[base + q];
    int total;
    block_exclusive_scan(s, sh, total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

static __global__ __launch_bounds__(kScanNT) void k_scan_sums(int *bsum, long nb) {
    int *sh = reinterpret_cast<int *>(cuna_smem);
    int carry = 0;
    for (long c0 = 0; c0 < nb; c0 += kScanNT) {
        long i = c0 + threadIdx.x;
        int v = i < nb ? bsum[i] : 0;
        int total;
        int ex = block_exclusive_scan(v, sh, total);
        if (i < nb) bsum[i] = ex + carry;
        carry += total;
        __syncthreads();
    }
}

static __global__ __launch_bounds__(kScanNT) void k_scan_apply(const int *in, long n, const int *bsum,
                                                        int *out) {
    int *sh = reinterpret_cast<int *>(cuna_smem);
    long base = (long)blockIdx.x * kScanTile + (long)threadIdx.x * kScanIPT;
    int v[kScanIPT];
    int s = 0;
#pragma unroll
    for (int q = 0; q < kScanIPT; ++q) {
        v[q] = base + q < n ? in[base + q] : 0;
        s += v[q];
    }
    int total;
    int ex = block_exclusive_scan(s, sh, total) + bsum[blockIdx.x];
#pragma unroll
    for (int q = 0; q < kScanIPT; ++q) {
        if (base + q < n) out[base + q] = ex;
        ex += v[q];
    }
}

// ----------------------------------------------------------- spread --
enum ValueMode { VAL_PLAIN = 0, VAL_COMPAT_BOOT = 1, VAL_COUNTS = 2 };

// One entry of a curve's spread work order: the tile, its three sorted
// point ranges [b, e) (empty where a neighbour is absent: non-periodic grid
// edges) and its cover.  Built once per plan with the tiles in decreasing
// order of their own point counts (ties by tile index), so that the densest
// tiles start first and the launch ends on light ones.
struct TileDesc {
    int tile, b0, e0, b1, e1, b2, e2, cover;
};
// scalar loads of one entry (merged into one 32-byte s_load)
__device__ __forceinline__ TileDesc ldc_desc(const TileDesc *p) {
    const int *q = reinterpret_cast<const int *>(p);
    TileDesc d;
    d.tile = ldc(q);
    d.b0 = ldc(q + 1);
    d.e0 = ldc(q + 2);
    d.b1 = ldc(q + 3);
    d.e1 = ldc(q + 4);
    d.b2 = ldc(q + 5);
    d.e2 = ldc(q + 6);
    d.cover = ldc(q + 7);
    return d;
}

template <typename T>
struct SpreadArgs {
    const PointRec<T> *rec;  // points sorted by anchor within each curve
    const int *perm;
    const int *bucket;
    const int *coff;       // curve point offsets (ncurves + 1)
    const T *y;            // values by original index, or null (unit weights)
    long y_slab_stride;    // added per replicate/slab for VAL_PLAIN with several rhs
    int npts;              // points of all curves (rec and perm hold max(npts, 1))
    T *grid;
    long grid_stride;      // per slab
    int G, tlog, ntiles, Kc, olo, ntaps, nrep;
    int slab0;  // slab (plan-time kernels: curve) of blockIdx.y = 0
    int periodic;
    int wrap1;  // periodic and a single tile covers the grid
    int cap;    // points staged in LDS per round (gather form)
    int gather; // FP32: use the gather form instead of the fixed-point scatter
    T bi;       // 1/beta
    T wscale;   // weight prefactor
    T e3[kMaxTaps];  // exp(-o^2/beta), o = olo + q
    // nullable: per curve (bi, wscale, exp(-bi)) in place of the three above,
    // for batches of compat curves whose filter widths differ
    const T *cpar;
    // nullable: per (curve, tile), the most points whose anchors fall in one
    // window of ntaps consecutive anchors (no cell receives more weights)
    const int *tile_cover;
    // nullable: per curve, ntiles work-order entries (the scatter kernels
    // take entry blockIdx.x of their curve instead of tile blockIdx.x)
    const TileDesc *tile_desc;
    // VAL_COMPAT_BOOT (src/gridding.cu:86-117): value y[d] and weight
    // parameter eps_orig[d] of the drawn point d, position of point p
    uint32_t seed;
    const T *eps_orig;
    // VAL_COUNTS: value counts[rep][p] * y[p]
    const int *counts;
    long counts_stride;
};

// Tap q (0 <= q < NTAPS) sits at offset o = q - Q0 from the anchor,
// Q0 = (NTAPS-1)/2: compat (NTAPS = 2R_f - 1) has o in [-R_f+1, R_f-1]
// (src/gridding.cu:53-58), accurate (NTAPS = 2m) has o in [-m+1, m].
// Weights by the Greengard-Lee factorisation exp(-(o-e)^2/beta) =
// E1 * E2^o * E3[o]: three exponentials per point, one multiply per tap.
template <typename T, int NTAPS>
__device__ __forceinline__ void tap_weights(T v, T e, T bi, const T *e3, T *w) {
    constexpr int Q0 = (NTAPS - 1) / 2;
    const T E1 = dexp(-e * e * bi);
    const T E2 = dexp(T(2) * e * bi);
    const T E2i = dexp(T(-2) * e * bi);
    T up = v * E1;
#pragma unroll
    for (int q = Q0; q < NTAPS; ++q) {
        w[q] = up * e3[q];
        up *= E2;
    }
    T dn = v * E1 * E2i;
#pragma unroll
    for (int q = Q0 - 1; q >= 0; --q) {
        w[q] = dn * e3[q];
        dn *= E2i;
    }
}

// value and weight parameter of sorted point s (original index p)// value and weight parameter of sorted point s (original index p)
template <typename T, int VM>
__device__ __forceinline__ void point_value(const SpreadArgs<T> &A, int p, int c, int rep, T &v,
                                            T &e) {
    if constexpr (VM == VAL_PLAIN) {
        v = A.y ? A.y[p + rep * A.y_slab_stride] : T(1);
    } else if constexpr (VM == VAL_COMPAT_BOOT) {
        if (A.y) {
            const int c0 = A.coff[c];
            const uint32_t nc = (uint32_t)(A.coff[c + 1] - c0);
            const uint32_t t = (uint32_t)rep * nc + (uint32_t)(p - c0);
            const int d = c0 + (int)(xorshift32(wang_hash(A.seed + t)) % nc);
            v = A.y[d];
            e = A.eps_orig[d];
        } else {
            v = T(1);
        }
    } else {
        v = (T)A.counts[p + rep * A.counts_stride] * (A.y ? A.y[p] : T(1));
    }
}

template <typename T>
struct StagedPoint {
    int a;  // anchor relative to the tile's first cell
    T e, v;
};

// LDS bytes of one spread workgroup
template <typename T>
__host__ __device__ inline size_t spread_lds_bytes(int tlog, int ntaps, int cap) {
    size_t b = sizeof(T) * ((size_t)1 << tlog) + sizeof(int) * (((size_t)1 << tlog) + 2 * ntaps + 1);
    b = (b + 15) & ~(size_t)15;
    return b + sizeof(StagedPoint<T>) * (size_t)cap;
}

// The sorted points a tile's cells depend on, as up to three ranges with the
// offset that maps their anchors to this tile's cells, in increasing order of
// mapped anchor: the previous tile's right-edge class, this tile, the next
// tile's left-edge class.  A tiny periodic grid (taps wrap more than once)
// instead uses 2S+1 shifted copies of all its points.
struct TileRanges {
    int b0, n0, o0, b1, n1, b2, n2, o2, tstart, total, S, G;
    bool tiny;
    template <typename T>
    __device__ __forceinline__ void init(const SpreadArgs<T> &A, int tile, int c, int ntaps) {
        load(A, tile, c, ntaps);
        finish();
    }
    // issues the bucket loads (b* = first sorted index, n* = one past the
    // last until finish()); nothing here waits for them
    template <typename T>
    __device__ __forceinline__ void load(const SpreadArgs<T> &A, int tile, int c, int ntaps) {
        const int nt = A.ntiles;
        const int *bk = A.bucket + (long)c * (A.Kc + 1);
        const int TT = 1 << A.tlog;
        G = A.G;
        tstart = tile << A.tlog;
        tiny = A.wrap1 && A.G < 2 * ntaps;
        S = tiny ? (ntaps + A.G - 1) / A.G + 1 : 0;
        b0 = n0 = o0 = b2 = n2 = o2 = 0;
        b1 = ldc(bk + 3 * tile);
        n1 = ldc(bk + 3 * tile + 3);
        if (!tiny) {
            int prev = tile - 1, next = tile + 1;
            bool hp = true, hn = true;
            if (prev < 0) {
                prev = nt - 1;
                hp = A.periodic;
            }
            if (next >= nt) {
                next = 0;
                hn = A.periodic;
            }
            if (hp) {
                b0 = ldc(bk + 3 * prev + 2);
                n0 = ldc(bk + 3 * prev + 3);
                o0 = (prev << A.tlog) + TT;
            }
            if (hn) {
                b2 = ldc(bk + 3 * next);
                n2 = ldc(bk + 3 * next + 1);
                o2 = (next << A.tlog) - TT;
            }
        }
    }
    // the same ranges from a work-order entry (one scalar load of 32 bytes
    // in place of the bucket loads)
    template <typename T>
    __device__ __forceinline__ void from_desc(const SpreadArgs<T> &A, const TileDesc &d,
                                              int ntaps) {
        const int nt = A.ntiles;
        const int TT = 1 << A.tlog;
        G = A.G;
        tstart = d.tile << A.tlog;
        tiny = A.wrap1 && A.G < 2 * ntaps;
        S = tiny ? (ntaps + A.G - 1) / A.G + 1 : 0;
        b0 = d.b0;
        n0 = d.e0;
        b1 = d.b1;
        n1 = d.e1;
        b2 = d.b2;
        n2 = d.e2;
        const int prev = d.tile > 0 ? d.tile - 1 : nt - 1;
        const int next = d.tile + 1 < nt ? d.tile + 1 : 0;
        o0 = (prev << A.tlog) + TT;
        o2 = (next << A.tlog) - TT;
        finish();
    }
    // counts and total from the loaded bounds
    __device__ __forceinline__ void finish() {
        n0 -= b0;
        n1 -= b1;
        n2 -= b2;
        total = tiny ? n1 * (2 * S + 1) : n0 + n1 + n2;
    }
    // an interior tile: its three ranges are adjacent buckets (keys 3t-1 ..
    // 3t+3) and share the offset tstart, so point j is sorted index b0 + j
    __device__ __forceinline__ bool contiguous() const {
        return !tiny && b0 + n0 == b1 && b1 + n1 == b2 && o0 == tstart && o2 == tstart;
    }
    // j-th point of the tile: sorted index s and the offset ro (mapped anchor
    // = a - ro); selects rather than branches, so that loads indexed by it
    // stay in one basic block and issue together
    __device__ __forceinline__ void get(int j, int &s, int &ro) const {
        if (tiny) {  // uniform
            s = b1 + j % n1;
            ro = -(j / n1 - S) * G;
            return;
        }
        // (selects between computed values only: a select between two
        // fields would become a select of their addresses and keep this
        // struct out of registers)
        const bool in0 = j < n0, in1 = j < n0 + n1;
        const int s0 = b0 + j, s1 = b1 + (j - n0), s2 = b2 + (j - n0 - n1);
        s = in0 ? s0 : (in1 ? s1 : s2);
        ro = o2 + (in1 ? tstart - o2 : 0) + (in0 ? o0 - tstart : 0);
    }
};

// grid: blockIdx.x = tile, blockIdx.y = slab (curve * nrep + replicate).
// Output-driven: every cell of the tile is owned by one thread, which sums
// v_i exp(-(o_i - e_i)^2/beta) over the points i whose taps reach it (o_i =
// cell - anchor_i).  Points are sorted by anchor, so those are a contiguous
// window of the staged list; start[L] (first point with anchor >= L) is built
// in one pass over the points.  Deterministic: each cell sums in a fixed order.
template <typename T, int NTAPS, int VM>
__global__ __launch_bounds__(256) void k_spread(SpreadArgs<T> A) {
    constexpr int NT = 256, Q0 = (NTAPS - 1) / 2, QH = NTAPS - 1 - Q0;
    constexpr int LMIN = -NTAPS;
    const int TT = 1 << A.tlog;
    const int LMAX = TT + NTAPS + 1;  // start[] covers cells LMIN .. LMAX-1
    T *gacc = reinterpret_cast<T *>(cuna_smem);
    int *start = reinterpret_cast<int *>(gacc + TT);
    size_t off_pts = sizeof(T) * TT + sizeof(int) * (TT + 2 * NTAPS + 1);
    off_pts = (off_pts + 15) & ~(size_t)15;
    StagedPoint<T> *pts = reinterpret_cast<StagedPoint<T> *>(cuna_smem + off_pts);

    // consecutive tiles go to different XCDs (hardware round-robin): dense
    // regions of clustered points are spread over all of them
    const int tile = blockIdx.x;
    const int slab = A.slab0 + blockIdx.y;
    const int c = slab / A.nrep, rep = slab % A.nrep;
    TileRanges R;
    R.init(A, tile, c, NTAPS);
    const int tstart = R.tstart, total = R.total;
    const T bi = A.cpar ? A.cpar[3 * c] : A.bi;
    const T wscale = A.cpar ? A.cpar[3 * c + 1] : A.wscale;

    for (int g = threadIdx.x; g < TT; g += NT) gacc[g] = T(0);
    for (int base = 0; base < total; base += A.cap) {
        const int cnt = min(A.cap, total - base);
        __syncthreads();
        // stage this round's points: anchor in tile coordinates, e, value;
        // four per thread at a time with branch-free loads (slots past the
        // end read point 0 of the round and are not stored)
        for (int i0 = 0; i0 < cnt; i0 += 4 * NT) {
            int ak[4], pk[4];
            T ek[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int i = i0 + k * NT + threadIdx.x;
                int s, ro;
                R.get(base + (i < cnt ? i : 0), s, ro);
                const PointRec<T> pr = A.rec[s];
                ak[k] = pr.a - ro;
                ek[k] = pr.e;
                pk[k] = A.perm[s];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int i = i0 + k * NT + threadIdx.x;
                T v, e = ek[k];
                point_value<T, VM>(A, pk[k], c, rep, v, e);
                if (i < cnt) {
                    StagedPoint<T> sp;
                    sp.a = ak[k];
                    sp.e = e;
                    sp.v = v;
                    pts[i] = sp;
                }
            }
        }
        for (int L = threadIdx.x; L < LMAX - LMIN; L += NT) start[L] = cnt;
        __syncthreads();
        // start[L - LMIN] = first staged point with anchor >= L
        for (int i = threadIdx.x; i < cnt; i += NT) {
            const int hi = min(pts[i].a, LMAX - 1);
            const int lo = i == 0 ? LMIN : max(pts[i - 1].a + 1, LMIN);
            for (int L = lo; L <= hi; ++L) start[L - LMIN] = i;
        }
        __syncthreads();
        for (int g = threadIdx.x; g < TT; g += NT) {
            const int lo = start[g - QH - LMIN];
            const int hi = start[g + Q0 + 1 - LMIN];
            T sum = T(0);
            for (int i = lo; i < hi; ++i) {
                const StagedPoint<T> sp = pts[i];
                const T d = (T)(g - sp.a) - sp.e;
                sum += sp.v * dexp(-d * d * bi);
            }
            gacc[g] += sum;
        }
    }
    __syncthreads();
    T *gout = A.grid + (long)slab * A.grid_stride + tstart;
    const int tlen = min(TT, A.G - tstart);
    for (int g = threadIdx.x; g < tlen; g += NT) gout[g] = gacc[g] * wscale;
}

// ---------------------------------------------- fixed-point scatter --
// Scatter form (both libraries): a thread takes a point, forms its NTAPS
// weights with three exponentials, and adds them to the tile's LDS cells
// with integer atomics in fixed point (32-bit cells in FP32, 64-bit in
// FP64).  Integer sums do not depend on the order of the additions, so the
// result is bitwise reproducible; float LDS atomics would be neither
// reproducible nor fast on this hardware, and double LDS atomics are
// compare-and-swap loops.  A round of up to CAP points uses the scale 2^s
// with min(cnt, cover) * wmax * 2^s <= 2^headroom (cover: the most points of
// the tile in one window of ntaps anchors), wmax the largest weight the round adds
// to the tile's own cells (bounded by max|v| * wscale for accurate filters),
// so none of them can overflow (halo cells, which are not written out, may).
// Split of a scaled weight x (|x| < 2^62) into a two's-complement 64-bit
// integer: h = floor(x / 2^32) and l = x - h 2^32 in [0, 2^32).  In float,
// exact for x >= 0; for -2^31 < x < 0, l is rounded to a multiple of 2^8,
// an error far below the float weight's own rounding.  In double, l is exact
// and rounded to the nearest integer (at most 2^32 - 1, so a tiny negative x
// gives -1, one unit of the scale).
__device__ __forceinline__ unsigned long long to_fixed64(float x) {
    const float h = floorf(x * 0x1p-32f);
    const float l = fminf(fmaf(h, -0x1p32f, x), 4294967040.0f);
    return ((unsigned long long)(unsigned)(int)h << 32) | (unsigned)l;
}
__device__ __forceinline__ unsigned long long to_fixed64(double x) {
    const double h = floor(x * 0x1p-32);
    const double l = fmin(rint(fma(h, -0x1p32, x)), 4294967295.0);
    return ((unsigned long long)(unsigned)(int)h << 32) | (unsigned)l;
}

// Plan time, one workgroup per (tile, curve): cover[c * ntiles + tile] = the
// most points (the tile's own and its neighbours' edge classes) whose anchors
// fall in one window of ntaps consecutive anchors that reaches a cell of the
// tile, from a histogram of the anchors in LDS.
template <typename T>
__global__ __launch_bounds__(256) void k_tile_cover(SpreadArgs<T> A, int *cover) {
    constexpr int NT = 256;
    int *hist = reinterpret_cast<int *>(cuna_smem);
    __shared__ int red[NT / 64];
    const int tile = blockIdx.x, c = A.slab0 + blockIdx.y;
    const int TT = 1 << A.tlog, nt = A.ntaps, Q0 = (nt - 1) / 2, NH = TT + 2 * nt;
    TileRanges R;
    R.init(A, tile, c, nt);
    for (int i = threadIdx.x; i < NH; i += NT) hist[i] = 0;
    __syncthreads();
    for (int j = threadIdx.x; j < R.total; j += NT) {
        int s, ro;
        R.get(j, s, ro);
        const int a = A.rec[s].a - ro + nt;
        if (a >= 0 && a < NH) atomicAdd(&hist[a], 1);
    }
    __syncthreads();
    // cell g takes the points with anchors g - (nt - 1 - Q0) .. g + Q0
    int best = 0;
    for (int g = threadIdx.x; g < TT; g += NT) {
        int sum = 0;
        for (int q = 0; q < nt; ++q) sum += hist[g - (nt - 1 - Q0) + nt + q];
        best = max(best, sum);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) best = max(best, __shfl_xor(best, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        int m = red[0];
        for (int w = 1; w < NT / 64; ++w) m = max(m, red[w]);
        cover[(long)c * A.ntiles + tile] = m;
    }
}

// Work order of the scatter kernels, one workgroup per curve: the curve's
// tiles by decreasing own point count, as a stable counting sort over
// kOrderBins bins of the count (exact below 128 points, in steps of 8 up to
// 1143, one bin above), ties in tile order; position i's TileDesc gets its
// tile, and k_tile_desc fills in the rest.
// The order only schedules the tiles: every tile's cells come out the same
// in any order.  Each of the 16 waves ranks a contiguous segment of tiles.
constexpr int kOrderThreads = 1024, kOrderWaves = kOrderThreads / 64, kOrderBins = 256;
constexpr int kOrderMaxTiles = 8192;
static_assert(kOrderThreads == kScanNT, "the work order scans with one value per thread");

__device__ __forceinline__ int order_bin(int npts) {
    const int r = npts < 128 ? npts : 128 + min(127, (npts - 128) >> 3);
    return kOrderBins - 1 - r;  // bin 0: the densest tiles
}

static __global__ __launch_bounds__(kOrderThreads) void k_tile_order(const int *bucket, int Kc,
                                                                     int ntiles, TileDesc *desc) {
    __shared__ int cnt[kOrderBins * kOrderWaves];  // [bin][wave]: counts, then starts
    __shared__ int scan_sh[kOrderThreads];
    __shared__ unsigned char bin_sh[kOrderMaxTiles];
    __shared__ short order[kOrderMaxTiles];
    const int c = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int *bk = bucket + (long)c * (Kc + 1);
    for (int i = tid; i < kOrderBins * kOrderWaves; i += kOrderThreads) cnt[i] = 0;
    // every thread's loads first (clamped to in-bounds tiles), then their
    // uses, so that the loads are in flight together
    constexpr int TPT = kOrderMaxTiles / kOrderThreads;
    int lo[TPT], hi[TPT];
#pragma unroll
    for (int r = 0; r < TPT; ++r) {
        const int tc = min(r * kOrderThreads + tid, ntiles - 1);
        lo[r] = bk[3 * tc];
        hi[r] = bk[3 * tc + 3];
    }
#pragma unroll
    for (int r = 0; r < TPT; ++r) {
        const int t = r * kOrderThreads + tid;
        if (t < ntiles) bin_sh[t] = (unsigned char)order_bin(hi[r] - lo[r]);
    }
    __syncthreads();
    const int seg = (ntiles + kOrderWaves - 1) / kOrderWaves;
    const int s0 = min(ntiles, wave * seg), s1 = min(ntiles, s0 + seg);
    for (int t = s0 + lane; t < s1; t += 64) atomicAdd(&cnt[bin_sh[t] * kOrderWaves + wave], 1);
    __syncthreads();
    // exclusive scan in (bin, wave) order: a wave's tiles of one bin follow
    // the earlier waves' tiles of that bin
    constexpr int IPT = kOrderBins * kOrderWaves / kOrderThreads;
    int v[IPT], sum = 0;
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
        v[q] = cnt[IPT * tid + q];
        sum += v[q];
    }
    int total;
    int ex = block_exclusive_scan(sum, scan_sh, total);
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
        cnt[IPT * tid + q] = ex;
        ex += v[q];
    }
    __syncthreads();
    // rank within the wave's 64 tiles of a step: lanes with the same bin
    const unsigned long long lt = (1ull << lane) - 1ull;
    for (int t0 = s0; t0 < s1; t0 += 64) {
        const int t = t0 + lane;
        const bool act = t < s1;
        const int b = act ? bin_sh[t] : 0;
        unsigned long long peers = __ballot(act);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const unsigned long long m = __ballot((b >> bit) & 1);
            peers &= ((b >> bit) & 1) ? m : ~m;
        }
        const int rank = __popcll(peers & lt);
        int pos = 0;
        if (act) pos = cnt[b * kOrderWaves + wave] + rank;
        __builtin_amdgcn_wave_barrier();
        if (act) {
            order[pos] = (short)t;
            if (rank == __popcll(peers) - 1) cnt[b * kOrderWaves + wave] = pos + 1;
        }
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    for (int i = tid; i < ntiles; i += kOrderThreads) desc[(long)c * ntiles + i].tile = order[i];
}

// The rest of each TileDesc: tile t's own range and its neighbours' edge
// classes (prev = t - 1 and next = t + 1, wrapping when periodic).  Many
// workgroups share the random reads of the bucket table.
static __global__ __launch_bounds__(256) void k_tile_desc(const int *bucket, int Kc, int ntiles,
                                                          int periodic, const int *cover,
                                                          TileDesc *desc) {
    const int nblk = (ntiles + 255) / 256;  // workgroups per curve
    const int c = blockIdx.x / nblk, i = (blockIdx.x % nblk) * 256 + threadIdx.x;
    if (i >= ntiles) return;
    const int *bk = bucket + (long)c * (Kc + 1);
    TileDesc *dp = desc + (long)c * ntiles + i;
    const int t = dp->tile;
    const int prev = t > 0 ? t - 1 : ntiles - 1, next = t + 1 < ntiles ? t + 1 : 0;
    const bool hp = t > 0 || periodic != 0, hn = t + 1 < ntiles || periodic != 0;
    TileDesc d;
    d.tile = t;
    d.b1 = bk[3 * t];
    d.e1 = bk[3 * t + 3];
    d.b0 = hp ? bk[3 * prev + 2] : 0;
    d.e0 = hp ? bk[3 * prev + 3] : 0;
    d.b2 = hn ? bk[3 * next] : 0;
    d.e2 = hn ? bk[3 * next + 1] : 0;
    d.cover = cover ? cover[(long)c * ntiles + t] : 0x7fffffff;
    *dp = d;
}

// Accumulators of the scatter: 32-bit cells (ds_add_u32) hold sums within
// 2^30, 64-bit cells (FP64 library) within 2^62.  In FP64 each weight is also
// kept within 2^`single` = 2^50 (a second bound on the scale) and rounded to
// the nearest integer by adding 1.5 * 2^52: for |x| < 2^51 the sum lies in
// [2^52, 2^53), whose spacing is 1, so its low mantissa bits are round(x) and
//   round(x) = bits(x + 1.5 * 2^52) - bits(1.5 * 2^52)
// as 64-bit integers (one add, or one fma with the weight's product, and one
// 64-bit subtract per tap).
template <typename T>
struct Fixed64 {
    using U = unsigned long long;
    static constexpr int headroom = 62;
    static constexpr int single = sizeof(T) == 8 ? 50 : 0;  // 0: no per-weight bound
    static constexpr double kMagic = 0x1.8p52;
    static constexpr U kMagicBits = 0x4338000000000000ull;
    __device__ static U fixed(T x) {
        if constexpr (sizeof(T) == 8)
            return (U)__double_as_longlong(x + kMagic) - kMagicBits;
        else
            return to_fixed64(x);
    }
    // round(a * b), FP64 only
    __device__ static U fixed_fma(double a, double b) {
        return (U)__double_as_longlong(fma(a, b, kMagic)) - kMagicBits;
    }
};
struct Fixed32 {
    using U = unsigned;
    static constexpr int headroom = 30;
    static constexpr int single = 0;
    __device__ static U fixed(float x) { return (U)__float2int_rn(x); }
};

// The NTAPS fixed-point weights of one point, rounded to nearest:
// n_q = round(v exp(-(q - Q0 - e)^2 bi)), v already carrying the scale, each
// handed to emit(q, n_q) as soon as it is formed (outward from the anchor),
// so that no array of NTAPS values stays live.  In FP32 with the table e3
// (accurate filters), each tap is one multiply (the step E2^{+-1}), one fma
// that also adds the rounding half (with the sign of v: all taps of a point
// share it) and one truncating conversion; FP64 rounds by FX::fixed_fma.
// Otherwise (compat filters) the weights come by ratio recurrences, which
// need only c1 = exp(-bi), and FX::fixed per tap.
template <typename T, int NTAPS, class FX, bool TABLE, class Emit>
__device__ __forceinline__ void tap_fixed(T v, T e, T bi, const T *e3, T c1, Emit &&emit) {
    constexpr int Q0 = (NTAPS - 1) / 2;
    if constexpr (TABLE && sizeof(T) == 4 && std::is_same<FX, Fixed32>::value) {
        const float E1 = dexp(-e * e * bi);
        const float E2 = dexp(2.0f * e * bi);
        const float E2i = dexp(-2.0f * e * bi);
        const float h = copysignf(0.5f, v);
        float up = v * E1;
#pragma unroll
        for (int q = Q0; q < NTAPS; ++q) {
            emit(q, (unsigned)(int)fmaf(up, e3[q], h));
            up *= E2;
        }
        float dn = v * E1 * E2i;
#pragma unroll
        for (int q = Q0 - 1; q >= 0; --q) {
            emit(q, (unsigned)(int)fmaf(dn, e3[q], h));
            dn *= E2i;
        }
    } else if constexpr (TABLE && sizeof(T) == 8 && std::is_same<FX, Fixed64<T>>::value) {
        const double E1 = dexp(-e * e * bi);
        const double E2 = dexp(2.0 * e * bi);
        const double E2i = dexp(-2.0 * e * bi);
        double up = v * E1;
#pragma unroll
        for (int q = Q0; q < NTAPS; ++q) {
            emit(q, FX::fixed_fma(up, e3[q]));
            up *= E2;
        }
        double dn = v * E1 * E2i;
#pragma unroll
        for (int q = Q0 - 1; q >= 0; --q) {
            emit(q, FX::fixed_fma(dn, e3[q]));
            dn *= E2i;
        }
    } else if constexpr (TABLE) {
        T w[NTAPS];
        tap_weights<T, NTAPS>(v, e, bi, e3, w);
#pragma unroll
        for (int q = 0; q < NTAPS; ++q) emit(q, FX::fixed(w[q]));
    } else {
        // two ratio recurrences, which need exp(-bi) instead of the table:
        // with w_o = v exp(-(o-e)^2 bi), w_{o+1} = w_o R_o, R_0 = E2 exp(-bi),
        // R_{o+1} = R_o exp(-2 bi); downwards w_{o-1} = w_o L_o with
        // L_0 = exp(-bi)/E2, L_{o-1} = L_o exp(-2 bi)
        const T D = c1 * c1;
        const T E1 = dexp(-e * e * bi);
        const T E2 = dexp(T(2) * e * bi);
        const T E2i = dexp(T(-2) * e * bi);
        const T w0 = v * E1;
        T up = w0, r = E2 * c1;
#pragma unroll
        for (int q = Q0; q < NTAPS; ++q) {
            emit(q, FX::fixed(up));
            up *= r;
            r *= D;
        }
        T dn = w0, l = E2i * c1;
#pragma unroll
        for (int q = Q0 - 1; q >= 0; --q) {
            dn *= l;
            l *= D;
            emit(q, FX::fixed(dn));
        }
    }
}


// ------------------------------------------- per-wave fixed-point scatter --
// Each wave of the tile's workgroup accumulates its own share of the points
// into its own LDS cells, in fixed point with integer atomics (32-bit cells
// in FP32, 64-bit in FP64), with its own power-of-two scale:
//   min(cnt, cover) * vmax * wscale * 2^s <= 2^headroom,
// cnt = the wave's points this round (<= 64 * kSpreadSlots), vmax their
// largest |value| (compat filters: largest kept weight / wscale), cover = the
// tile's most points in one window of ntaps anchors.  No cell can overflow,
// whatever the density, so no wider fallback and no plan-time read-back is
// needed; a wave's scale needs no workgroup barrier; and the waves never
// contend for one address.  After the last round the waves' cells are
// converted and summed in a fixed wave order, so the grid is bitwise
// reproducible.  Resolution: 2^-headroom * min(cnt, cover) of the round's
// largest weight (FP32: 2^-24 for cover <= 64, 2^-22 at worst).  A wave with
// several rounds folds each into its own float (double) cells.
constexpr int kSpreadSlots = 4;  // point slots per lane in a round

template <typename T>
using SpreadFixed = std::conditional_t<sizeof(T) == 4, Fixed32, Fixed64<T>>;

template <typename T>
__host__ __device__ inline size_t spread_wv_lds_bytes(int tlog, int ntaps, int waves) {
    const size_t tt = (size_t)1 << tlog;
    const size_t acc = sizeof(typename SpreadFixed<T>::U) * (tt + 2 * (size_t)ntaps);
    return (size_t)waves * (acc + sizeof(T) * tt);
}

template <typename T, int NTAPS, int VM, int WPT>
__global__ __launch_bounds__(64 * WPT) void k_spread_wv(SpreadArgs<T> A) {
    using FX = SpreadFixed<T>;
    using Acc = typename FX::U;
    using SAcc = std::make_signed_t<Acc>;
    constexpr int NT = 64 * WPT, PPT = kSpreadSlots, CAP = NT * PPT;
    constexpr int Q0 = (NTAPS - 1) / 2, HALO = NTAPS;
    constexpr T kBig = sizeof(T) == 4 ? T(3.0e38f) : T(1.0e300);
    constexpr int kMaxExp = sizeof(T) == 4 ? 100 : 1000;
    __shared__ int sexp_sh[WPT];
    __shared__ int fin_sh[WPT];
    const int TT = 1 << A.tlog;
    const int NC = TT + 2 * HALO;  // cells -HALO .. TT+HALO-1
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    Acc *acc_all = reinterpret_cast<Acc *>(cuna_smem);
    T *facc_all = reinterpret_cast<T *>(acc_all + WPT * NC);
    Acc *acc = acc_all + wave * NC;
    T *facc = facc_all + wave * TT;

    const int slab = A.slab0 + blockIdx.y;
    const int c = slab / A.nrep, rep = slab % A.nrep;
    TileRanges R;
    int cover = CAP;
    if (A.tile_desc) {
        const TileDesc d = ldc_desc(A.tile_desc + (long)c * A.ntiles + blockIdx.x);
        R.from_desc(A, d, NTAPS);
        cover = d.cover;
    } else {
        R.init(A, blockIdx.x, c, NTAPS);
        if (A.tile_cover) cover = A.tile_cover[(long)c * A.ntiles + blockIdx.x];
    }
    const int total = R.total;
    const bool multi = total > CAP;
    const T c1 = A.cpar ? A.cpar[3 * c + 2] : A.e3[Q0 + 1];  // exp(-bi)
    const T bi = A.cpar ? A.cpar[3 * c] : A.bi;
    const T wscale = A.cpar ? A.cpar[3 * c + 1] : A.wscale;
    const int tlen = min(TT, A.G - R.tstart);  // cells of the tile inside the grid
    int ew = 0;
    (void)frexp(wscale, &ew);

    for (int g = lane; g < NC; g += 64) acc[g] = Acc(0);
    if (multi)
        for (int g = lane; g < TT; g += 64) facc[g] = T(0);
    __builtin_amdgcn_s_waitcnt(0xc07f);  // the wave's clears land before its atomics
    __builtin_amdgcn_wave_barrier();
    bool finite = true;
    int sexp = 0;
    const bool contig = R.contiguous();  // uniform
    for (int base = 0; base < total; base += CAP) {
        const int cnt = min(CAP, total - base);
        // thread t takes the round's points t*spl .. t*spl + spl-1, spl the
        // fewest slots that hold cnt points: a wave's lanes then sit on
        // points spread over the whole tile, so one atomic instruction seldom
        // puts two lanes on one cell (the points are sorted by anchor), and
        // slots past spl are empty for every lane.  Branch-free loads (all
        // slots issue together); a slot past the end reads point 0.
        const int spl = (cnt + NT - 1) / NT;  // uniform
        auto slot = [&](int k) { return k < spl ? tid * spl + k : CAP; };
        int am[PPT], pk[PPT], sk[PPT], rk[PPT];
        T ev[PPT], vv[PPT];
        // this wave's points in the round
        const int wcnt = max(0, min(64 * spl, cnt - wave * 64 * spl));
        // the original indices first, then the records: the values gathered
        // through perm below wait only for perm (loads return in order), and
        // their loads overlap the records'
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const int i = slot(k);
            const int j = base + (i < cnt ? i : 0);
            if (contig) {
                sk[k] = R.b0 + j;
                rk[k] = R.tstart;
            } else {
                R.get(j, sk[k], rk[k]);
            }
            pk[k] = A.perm[sk[k]];
        }
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const PointRec<T> pr = A.rec[sk[k]];
            am[k] = pr.a - rk[k];
            ev[k] = pr.e;
        }
        T vmax = T(0);
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            T v, e = ev[k];
            point_value<T, VM>(A, pk[k], c, rep, v, e);
            const bool ok = slot(k) < cnt;
            vv[k] = ok ? v : T(0);
            ev[k] = e;
            T av = fabs(vv[k]);
            if constexpr (NTAPS % 2 == 1) {
                // compat filter: e lies up to R_f + 1 cells from the taps and
                // taps off the grid are dropped, so the largest weight this
                // tile keeps, at its tap o nearest e, can be far below |v|
                const int lo = max(-Q0, -am[k]), hi = min(NTAPS - 1 - Q0, tlen - 1 - am[k]);
                const T o = fmin(fmax(rint(e), T(lo)), T(hi));
                av = lo <= hi ? av * dexp(-(o - e) * (o - e) * bi) : T(0);
            }
            // a non-finite value or position (NaN x gives NaN e) makes the
            // tile's output NaN
            if (!(av <= kBig) || (ok && !(fabs(e) <= kBig))) av = T(INFINITY);
            vmax = fmax(vmax, av);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) vmax = fmax(vmax, __shfl_xor(vmax, o));
        if (!(vmax <= kBig)) finite = false;
        const T bound = (T)min(wcnt, cover) * vmax * wscale;
        int eb = -64;
        if (bound > T(0) && finite) (void)frexp(bound, &eb);
        sexp = min(FX::headroom - eb, kMaxExp - ew);
        if constexpr (FX::single > 0) {
            int e1 = -64;
            const T b1 = vmax * wscale;
            if (b1 > T(0) && finite) (void)frexp(b1, &e1);
            sexp = min(sexp, FX::single - e1);
        }
        const T mul = ldexp(wscale, sexp);
        if (finite) {
#pragma unroll
            for (int k = 0; k < PPT; ++k) {
                if (slot(k) < cnt) {
                    // accurate filters (even NTAPS) share one table e3; compat
                    // batches carry a filter per curve (cpar), so recurrences
                    const int c0 = am[k] - Q0;  // first tap's cell
                    Acc *dst = acc + HALO + c0;
                    if (R.tiny) {
                        tap_fixed<T, NTAPS, FX, NTAPS % 2 == 0>(
                            vv[k] * mul, ev[k], bi, A.e3, c1, [&](int q, Acc n) {
                                const int cell = c0 + q;
                                if (cell >= -HALO && cell < TT + HALO) atomicAdd(dst + q, n);
                            });
                    } else {
                        tap_fixed<T, NTAPS, FX, NTAPS % 2 == 0>(
                            vv[k] * mul, ev[k], bi, A.e3, c1,
                            [&](int q, Acc n) { atomicAdd(dst + q, n); });
                    }
                }
            }
        }
        if (multi) {
            // fold the round into this wave's float cells; the wave's own
            // atomics must have landed first (LDS counter drained)
            __builtin_amdgcn_s_waitcnt(0xc07f);
            __builtin_amdgcn_wave_barrier();
            const T inv = ldexp(T(1), -sexp);
            for (int g = lane; g < TT; g += 64) {
                facc[g] += (T)(SAcc)acc[HALO + g] * inv;
                acc[HALO + g] = Acc(0);
            }
            for (int h = lane; h < 2 * HALO; h += 64) acc[h < HALO ? h : TT + h] = Acc(0);
            __builtin_amdgcn_s_waitcnt(0xc07f);
            __builtin_amdgcn_wave_barrier();
        }
    }
    if (lane == 0) {
        sexp_sh[wave] = sexp;
        fin_sh[wave] = finite ? 1 : 0;
    }
    // every wave's cells and scale are read below by all threads
    __syncthreads();
    bool all_finite = true;
    T inv[WPT];
#pragma unroll
    for (int w = 0; w < WPT; ++w) {
        all_finite = all_finite && fin_sh[w] != 0;
        inv[w] = ldexp(T(1), -sexp_sh[w]);
    }
    T *gout = A.grid + (long)slab * A.grid_stride + R.tstart;
    for (int g = tid; g < tlen; g += NT) {
        T v = T(0);
#pragma unroll
        for (int w = 0; w < WPT; ++w)
            v += multi ? facc_all[w * TT + g] : (T)(SAcc)acc_all[w * NC + HALO + g] * inv[w];
        gout[g] = all_finite ? v : T(NAN);
    }
}

// ------------------------------------------- fused batched bootstrap --
// One persistent workgroup per CU drains a queue of (curve, replicate,
// transform) items.  An item draws the replicate's points
// d = c0 + xorshift(wang(seed + rep*n_c + p)) % n_c, p < n_c, spreads them
// into a grid of G = 2N cells held entirely in LDS, transforms it in place
// (real input as N complex) and writes the deconvolved spectrum:
// transform 0 is y at the points x_d, transform 1 the unit-weight window at 2 x_d.
//
// LDS float atomics are slow on this hardware, so the grid accumulates in
// fixed point with integer atomics (int32 in the FP32 library, int64 in FP64),
// which are also order-independent.  The power-of-two scale comes from the
// curve's max |y| and its densest tap window; an addition that overflows is
// detected from the returned old value and the item is redone at a smaller
// scale.  The integer grid occupies the same LDS words as the complex
// buffer of the FFT, whose first pass reads and converts it.
template <typename T>
struct BootArgs {
    const PointRec<T> *rec_x;   // (anchor, frac) of x by original index
    const PointRec<T> *rec_2x;  // of 2x (window), nullable
    const T *y;                 // nullable: unit weights
    const T *ymax;              // per curve max |y| (when y is set)
    const int *cover_x;         // per curve: most points within one tap window
    const int *cover_2x;
    const int *coff;
    const int *order;           // curves, longest first
    int *counter;               // work-queue head, zeroed before the launch
    cplx<T> *out_y, *out_w;
    int ncurves, nbs, ntr, total;
    uint32_t seed;
    T bi, wscale;
    T e3[kMaxTaps];
    SinkAccurate<T> sink;       // out pointer set per item
    Tw2<T> twN;                 // w_N^e, e < N (two-level, staged into LDS)
    Tw2<T> twG;                 // w_G^e, e < G (two-level, staged into LDS)
    // Lomb-Scargle mode (each nullable): per replicate, power at k = 0..M/2-1,
    // and its maximum over k >= 1 with the k where it occurs
    T *ls_power;
    T *ls_peak;
    int *ls_kpeak;
    // ... and a workspace of ls_w_slots rows of M/2 values: a workgroup keeps
    // its current window spectrum in row blockIdx.x
    cplx<T> *ls_w;
    int ls_w_slots;
};

// LDS of the fused kernel: the grid (G reals = N complex) and the four
// twiddle sub-tables (Tw2 of N and of G).
template <typename T>
__host__ __device__ constexpr int boot_tw_entries(int lg) {
    return (1 << ((lg + 1) / 2)) + (1 << (lg - (lg + 1) / 2)) + (1 << ((lg + 2) / 2)) +
           (1 << (lg + 1 - (lg + 2) / 2));
}
template <typename T>
__host__ __device__ constexpr size_t boot_lds_bytes(int lg) {
    return sizeof(cplx<T>) * (((size_t)1 << lg) + (size_t)boot_tw_entries<T>(lg));
}

static __global__ void k_reset_counter(int *c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0;
}

template <typename I>
struct alignas(2 * sizeof(I)) FixedPair {
    I x, y;
};

template <typename T>
struct FixedOf;
template <>
struct FixedOf<float> {
    using I = int;
    static constexpr int headroom_bits = 30;
    __device__ static I round(float v) { return __float2int_rn(v); }
};
template <>
struct FixedOf<double> {
    using I = long long;
    static constexpr int headroom_bits = 61;
    __device__ static I round(double v) { return __double2ll_rn(v); }
};

// Spectrum value A[r] of the real grid from the pair (Z[r], Z[N-r]) of its
// complex transform (the first of the two values emit_pair produces).
template <typename T>
__device__ __forceinline__ cplx<T> real_spectrum_at(cplx<T> Zr, cplx<T> Znr, cplx<T> wr) {
    cplx<T> c = conjg(Znr);
    cplx<T> E = (Zr + c) * T(0.5);
    cplx<T> D = (Zr - c) * T(0.5);
    return E + wr * cmk(D.y, -D.x);
}

// Lomb-Scargle power (Press & Rybicki 1989, "standard" normalisation, data
// centred) from F = sum (y_i - ybar) e^{i w t_i} and W = sum e^{2 i w t_i}
// over n points with yy = sum (y_i - ybar)^2:
//   tan 2wtau = Im W / Re W,  YC = Re F cos wtau + Im F sin wtau,
//   YS = Im F cos wtau - Re F sin wtau,  CC = (n + |W|)/2,  SS = (n - |W|)/2,
//   P = (YC^2/CC + YS^2/SS) / yy   (a term whose denominator vanishes is 0).
// Reciprocals: the FP32 library uses the hardware's (v_rcp_f32, v_rsq_f32,
// about 1 ulp; the periodogram's error budget is far larger), FP64 divides.
__device__ __forceinline__ float ls_rcp(float v) { return __builtin_amdgcn_rcpf(v); }
__device__ __forceinline__ double ls_rcp(double v) { return 1.0 / v; }
__device__ __forceinline__ float ls_rsqrt(float v) { return __builtin_amdgcn_rsqf(v); }
__device__ __forceinline__ double ls_rsqrt(double v) { return 1.0 / sqrt(v); }

template <typename T>
__device__ __forceinline__ T ls_power(cplx<T> F, cplx<T> W, T n, T yy) {
    const T h2 = W.x * W.x + W.y * W.y;
    const T ih = h2 > T(0) ? ls_rsqrt(h2) : T(0);
    const T h = h2 * ih;
    const T c2 = h2 > T(0) ? W.x * ih : T(1), s2 = h2 > T(0) ? W.y * ih : T(0);
    const T cw = sqrt(fmax(T(0), (T(1) + c2) * T(0.5)));
    const T sw = copysign(sqrt(fmax(T(0), (T(1) - c2) * T(0.5))), s2);
    const T yc = F.x * cw + F.y * sw, ys = F.y * cw - F.x * sw;
    const T cc = (n + h) * T(0.5), ss = (n - h) * T(0.5);
    const T tiny = n * T(sizeof(T) == 4 ? 1e-6 : 1e-14);
    T p = (cc > tiny ? yc * yc * ls_rcp(cc) : T(0)) + (ss > tiny ? ys * ys * ls_rcp(ss) : T(0));
    return yy > T(0) ? p * ls_rcp(yy) : T(0);
}

// A value every thread of the workgroup holds equally, moved to scalar
// registers (it then takes no vector registers across the transforms).
__device__ __forceinline__ double uniform_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)b);
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// MODE 0: items (curve, replicate, transform) -> deconvolved spectra.
// MODE 1: items (curve, replicate) -> Lomb-Scargle power of the replicate:
// the window transform (unit weights at 2x) goes to the workgroup's row of
// the workspace (each thread writes, and later reads back, its own
// frequencies), then the transform of the mean-subtracted values at x gives
// the power directly; nothing else is written.
template <typename T, int LG, int NTAPS, int NT, int MODE>
__global__ __launch_bounds__(NT) void k_boot_fused(BootArgs<T> A) {
    using I = typename FixedOf<T>::I;
    using U = std::make_unsigned_t<I>;
    static_assert(sizeof(I) == sizeof(T), "fixed-point grid aliases the real grid");
    constexpr int N = 1 << LG, G = 2 * N, Q0 = (NTAPS - 1) / 2;
    constexpr int LBN = (LG + 1) / 2, LBG = (LG + 2) / 2;  // as TwiddleCache::two_level
    constexpr int NLN = 1 << LBN, NHN = 1 << (LG - LBN), NLG = 1 << LBG, NHG = 1 << (LG + 1 - LBG);
    constexpr int PB = 4;  // resampled points per thread whose loads are in flight together
    constexpr int kMaxExp = sizeof(T) == 4 ? 100 : 1000;
    T *grid = reinterpret_cast<T *>(cuna_smem);
    U *gu = reinterpret_cast<U *>(cuna_smem);
    cplx<T> *z = reinterpret_cast<cplx<T> *>(cuna_smem);
    cplx<T> *tl = z + N;  // twiddle sub-tables, staged once per workgroup
    const int tid = threadIdx.x;
    for (int i = tid; i < NLN; i += NT) tl[i] = A.twN.lo[i];
    for (int i = tid; i < NHN; i += NT) tl[NLN + i] = A.twN.hi[i];
    for (int i = tid; i < NLG; i += NT) tl[NLN + NHN + i] = A.twG.lo[i];
    for (int i = tid; i < NHG; i += NT) tl[NLN + NHN + NLG + i] = A.twG.hi[i];
    Tw2<T> twN, twG;
    twN.lo = tl;
    twN.hi = tl + NLN;
    twN.lb = LBN;
    twG.lo = tl + NLN + NHN;
    twG.hi = tl + NLN + NHN + NLG;
    twG.lb = LBG;
    __shared__ int item_sh;
    __shared__ int dmax_sh[NT / 64];
    // draws per block of 32 anchors (at most G/32 = 1024 blocks)
    constexpr int NB = G / 32;
    static_assert(NB >= 2 && NTAPS <= 32, "a tap window spans at most two anchor blocks");
    __shared__ int hist_sh[NB];
    __shared__ double red_sh[2 * (NT / 64)];
    __shared__ T peak_sh[NT / 64];
    __shared__ int kpk_sh[NT / 64];

    // Grid of one transform: the draws of replicate `rep` of the curve at
    // c0 (nc points), positions from P, values y[d] - shift (or 1), |value|
    // <= vm; spread in fixed point, converted, transformed in place into z.
    auto transform = [&](int c0, int nc, int rep, const FastMod &fm, const PointRec<T> *P,
                         const T *y, T shift, double vm, int cover) __attribute__((always_inline)) {
        int sexp = 0;
        if (vm <= 1e300) {  // false for NaN and inf
            // No cell receives more than cnt weights of size <= vm * wscale, so
            // with cnt * vm * wscale * 2^sexp <= 2^headroom no cell can wrap.
            // First attempt: cnt = 4x the densest window of the original
            // points.  The draws' own densest window is counted on the way
            // (draws per block of 32 anchors; a cell's taps come from at most
            // 32 consecutive anchors, so from two adjacent blocks): if it
            // exceeds cnt the item is redone at that exact count.
            int cnt = 4 * max(cover, 1);
            for (int attempt = 0; attempt < 2; ++attempt) {
                const double bound = vm * (double)cnt * (double)A.wscale;
                const int e = bound > 0 ? (int)ceil(log2(bound)) : -kMaxExp;
                sexp = min(max(FixedOf<T>::headroom_bits - e, -kMaxExp), kMaxExp);
                const T mul = (T)ldexp((double)A.wscale, sexp);
                for (int i = tid; i < G; i += NT) gu[i] = U(0);
                for (int i = tid; i < NB; i += NT) hist_sh[i] = 0;
                __syncthreads();
                for (int p0 = 0; p0 < nc; p0 += NT * PB) {
                    PointRec<T> pr[PB];
                    T vv[PB];
                    // branch-free: out-of-range slots draw for p = 0 (and are
                    // skipped below), so all loads issue together
                    int dk[PB];
#pragma unroll
                    for (int k = 0; k < PB; ++k) {
                        int p = p0 + k * NT + tid;
                        p = p < nc ? p : 0;
                        const uint32_t t = (uint32_t)rep * (uint32_t)nc + (uint32_t)p;
                        dk[k] = c0 + (int)fm(xorshift32(wang_hash(A.seed + t)));
                    }
#pragma unroll
                    for (int k = 0; k < PB; ++k) pr[k] = P[dk[k]];
                    if (y) {
#pragma unroll
                        for (int k = 0; k < PB; ++k) vv[k] = y[dk[k]] - shift;
                    } else {
#pragma unroll
                        for (int k = 0; k < PB; ++k) vv[k] = T(1);
                    }
#pragma unroll
                    for (int k = 0; k < PB; ++k) {
                        if (p0 + k * NT + tid < nc) {
                            const int base = pr[k].a - Q0;
                            if constexpr (sizeof(T) == 4) {
                                // each tap's rounded weight to its atomic as
                                // it is formed (weights <= 2^headroom_bits)
                                tap_fixed<T, NTAPS, Fixed32, true>(
                                    vv[k] * mul, pr[k].e, A.bi, A.e3, T(0), [&](int q, unsigned n) {
                                        atomicAdd(&gu[(base + q) & (G - 1)], (U)n);
                                    });
                            } else {
                                T w[NTAPS];
                                tap_weights<T, NTAPS>(vv[k] * mul, pr[k].e, A.bi, A.e3, w);
#pragma unroll
                                for (int q = 0; q < NTAPS; ++q)
                                    atomicAdd(&gu[(base + q) & (G - 1)],
                                              (U)FixedOf<T>::round(w[q]));
                            }
                            atomicAdd(&hist_sh[pr[k].a >> 5], 1);
                        }
                    }
                }
                __syncthreads();
                int dmax = 0;
                for (int b = tid; b < NB; b += NT)
                    dmax = max(dmax, hist_sh[b] + hist_sh[(b + 1) & (NB - 1)]);
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) dmax = max(dmax, __shfl_xor(dmax, o));
                if ((tid & 63) == 0) dmax_sh[tid >> 6] = dmax;
                __syncthreads();
                dmax = dmax_sh[0];
#pragma unroll
                for (int w = 1; w < NT / 64; ++w) dmax = max(dmax, dmax_sh[w]);
                // every thread has read dmax_sh before the next attempt or item
                // writes it (the FFT and the emit below have barriers too)
                __syncthreads();
                if (dmax <= cnt) break;  // uniform
                cnt = dmax;
            }
            const T inv = (T)ldexp(1.0, -sexp);
            if constexpr (MODE == 0) {
                // the FFT's first pass reads the fixed-point grid (complex e
                // = cells 2e, 2e+1) and converts it to reals as it loads
                auto fetch = [&](int e, int) __attribute__((always_inline)) {
                    const FixedPair<I> q = reinterpret_cast<const FixedPair<I> *>(gu)[e];
                    return cmk((T)q.x * inv, (T)q.y * inv);
                };
                fft_pass<T, 1, LG, 1, NT, 0, 0, decltype(fetch), true, NoHook, Tw2<T>>(z, twN,
                                                                                    fetch);
            } else {
                // (Lomb-Scargle keeps the window spectrum in registers; with
                // the transform outside this branch it needs fewer of them)
                for (int i = tid; i < G; i += NT) grid[i] = (T)(I)gu[i] * inv;
            }
        } else {
            // a non-finite value: every spectrum value is NaN
            for (int i = tid; i < G; i += NT) grid[i] = (T)NAN;
            if constexpr (MODE == 0) __syncthreads();
        }
        if constexpr (MODE == 1) {
            __syncthreads();
            lds_fft_inplace<T, 1, LG, NT>(z, twN);
        }
    };
    // deconvolved spectrum at frequency f (0 <= f < M/2) from z
    auto spectrum = [&](int f) __attribute__((always_inline)) -> cplx<T> {
        const SinkAccurate<T> &sk = A.sink;
        const cplx<T> a = real_spectrum_at(z[lds_sw(f)], z[lds_sw((N - f) & (N - 1))], twG(f));
        const T kk = (T)f * sk.invG;
        return a * (sk.pref * dexp(sk.coef * kk * kk));
    };

    for (;;) {
        if (tid == 0) item_sh = atomicAdd(A.counter, 1);
        __syncthreads();
        // uniform over the workgroup: scalar registers (so are the curve's
        // offsets and sizes loaded through it)
        const int item = __builtin_amdgcn_readfirstlane(item_sh);
        if (item >= A.total) break;  // uniform over the workgroup
        const int per_curve = A.nbs * (MODE == 0 ? A.ntr : 1);
        const int ci = A.order[item / per_curve];
        const int rem = item % per_curve;
        const int c0 = A.coff[ci];
        const int nc = A.coff[ci + 1] - c0;
        FastMod fm;
        fm.init((uint32_t)max(nc, 1));
        if constexpr (MODE == 0) {
            const int rep = rem / A.ntr, tr = rem % A.ntr;
            const bool usey = tr == 0 && A.y;
            transform(c0, nc, rep, fm, tr == 0 ? A.rec_x : A.rec_2x, usey ? A.y : nullptr, T(0),
                      usey ? (double)A.ymax[ci] : 1.0, tr == 0 ? A.cover_x[ci] : A.cover_2x[ci]);
            SinkAccurate<T> sk = A.sink;
            cplx<T> *o = (tr == 0 ? A.out_y : A.out_w) + ((long)ci * A.nbs + rep) * sk.M;
            for (int r = tid; r <= N / 2; r += NT)
                emit_pair(sk, o, (long)r, (long)N, z[lds_sw(r)], z[lds_sw((N - r) & (N - 1))],
                          twG(r));
            __syncthreads();
        } else {
            const int rep = rem;
            const long row = (long)ci * A.nbs + rep;
            const int half = (int)(A.sink.M / 2);
            // moments of the drawn values about the first draw (shifted sums
            // keep the variance accurate when |mean| >> spread)
            double s1 = 0.0, s2 = 0.0;
            T shift = T(0);
            if (nc > 0) {
                shift = A.y[c0 + (int)fm(xorshift32(wang_hash(A.seed + (uint32_t)rep * (uint32_t)nc)))];
                for (int p0 = 0; p0 < nc; p0 += NT * PB) {
                    T vv[PB];
#pragma unroll
                    for (int k = 0; k < PB; ++k) {
                        int p = p0 + k * NT + tid;
                        p = p < nc ? p : 0;
                        const uint32_t t = (uint32_t)rep * (uint32_t)nc + (uint32_t)p;
                        vv[k] = A.y[c0 + (int)fm(xorshift32(wang_hash(A.seed + t)))];
                    }
#pragma unroll
                    for (int k = 0; k < PB; ++k) {
                        if (p0 + k * NT + tid < nc) {
                            const double d = (double)vv[k] - (double)shift;
                            s1 += d;
                            s2 += d * d;
                        }
                    }
                }
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                s1 += __shfl_xor(s1, o);
                s2 += __shfl_xor(s2, o);
            }
            if ((tid & 63) == 0) {
                red_sh[2 * (tid >> 6)] = s1;
                red_sh[2 * (tid >> 6) + 1] = s2;
            }
            __syncthreads();
            s1 = 0.0;
            s2 = 0.0;
            for (int w = 0; w < NT / 64; ++w) {
                s1 += red_sh[2 * w];
                s2 += red_sh[2 * w + 1];
            }
            s1 = uniform_f64(s1);
            s2 = uniform_f64(s2);
            const double nd = (double)max(nc, 1);
            const double mean_off = s1 / nd;
            const T ybar = (T)((double)shift + mean_off);
            const T yy = (T)fmax(0.0, s2 - s1 * mean_off);
            // window W(2w), parked in this workgroup's workspace row
            transform(c0, nc, rep, fm, A.rec_2x, nullptr, T(0), 1.0, A.cover_2x[ci]);
            cplx<T> *wrow = A.ls_w + (long)blockIdx.x * half;
            for (int f = tid; f < half; f += NT) wrow[f] = spectrum(f);
            __syncthreads();
            // F(w) of the mean-subtracted values, then the power
            const double vm = (double)A.ymax[ci] + fabs((double)ybar);
            transform(c0, nc, rep, fm, A.rec_x, A.y, ybar, vm, A.cover_x[ci]);
            T best = T(-1);
            int kbest = 0;
            for (int f = tid; f < half; f += NT) {
                const T pw = ls_power(spectrum(f), wrow[f], (T)nc, yy);
                if (A.ls_power) A.ls_power[row * half + f] = pw;
                if (f >= 1 && pw > best) {
                    best = pw;
                    kbest = f;
                }
            }
            if (A.ls_peak || A.ls_kpeak) {
                // block arg-max (ties: smallest k)
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) {
                    const T ob = __shfl_xor(best, o);
                    const int ok = __shfl_xor(kbest, o);
                    if (ob > best || (ob == best && ok < kbest)) {
                        best = ob;
                        kbest = ok;
                    }
                }
                if ((tid & 63) == 0) {
                    peak_sh[tid >> 6] = best;
                    kpk_sh[tid >> 6] = kbest;
                }
                __syncthreads();
                if (tid == 0) {
                    for (int w = 1; w < NT / 64; ++w) {
                        if (peak_sh[w] > best || (peak_sh[w] == best && kpk_sh[w] < kbest)) {
                            best = peak_sh[w];
                            kbest = kpk_sh[w];
                        }
                    }
                    if (A.ls_peak) A.ls_peak[row] = best < T(0) ? T(0) : best;
                    if (A.ls_kpeak) A.ls_kpeak[row] = kbest;
                }
            }
            __syncthreads();
        }
    }
}

// per-curve max |y|
template <typename T>
__global__ void k_curve_absmax(const T *y, const int *coff, T *out) {
    const int c = blockIdx.x;
    const int b = coff[c], e = coff[c + 1];
    T m = T(0);
    for (int i = b + threadIdx.x; i < e; i += blockDim.x) m = fmax(m, fabs(y[i]));
    __shared__ T red[256];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = red[0];
}

// per-curve densest window: the most points whose anchors fall in one
// cyclic run of NTAPS cells of the G-cell grid.  One workgroup per curve,
// anchor histogram in LDS (G ints).
template <typename T>
__global__ void k_curve_cover(const PointRec<T> *rec, const int *coff, int G, int ntaps,
                              int *cover) {
    int *h = reinterpret_cast<int *>(cuna_smem);
    const int c = blockIdx.x;
    for (int i = threadIdx.x; i < G; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (int i = coff[c] + threadIdx.x; i < coff[c + 1]; i += blockDim.x)
        atomicAdd(&h[rec[i].a & (G - 1)], 1);
    __syncthreads();
    int best = 0;
    for (int g = threadIdx.x; g < G; g += blockDim.x) {
        int sum = 0;
        for (int k = 0; k < ntaps; ++k) sum += h[(g + k) & (G - 1)];
        best = max(best, sum);
    }
    __shared__ int red;
    if (threadIdx.x == 0) red = 0;
    __syncthreads();
    atomicMax(&red, best);
    __syncthreads();
    if (threadIdx.x == 0) cover[c] = red;
}

// ----------------------------------------------------------- gather --
// Transpose of the accurate spread, for the backward pass: for each point of
// a tile, sum_o w(o) h[a+o] and sum_o (o - e) w(o) h[a+o] over the real grid h.
template <typename T>
struct GatherArgs {
    const PointRec<T> *rec;
    const int *perm;
    const int *bucket;
    const T *h;         // real grid per slab (slab = curve * nrep + rep)
    long h_stride;
    const T *y;         // values (for d/dx), nullable -> 1
    long y_stride;      // per rep
    T *dy;              // nullable
    T *dx;              // nullable
    long out_stride;    // per rep
    int dy_elem;        // element stride of dy (2: one part of an interleaved complex array)
    int G, tlog, ntiles, Kc, olo, nrep;
    int slab0;  // slab of blockIdx.y = 0
    T bi, wscale, dxscale;  // dxscale = 2 G / beta
    T e3[kMaxTaps];
};

template <typename T, int NTAPS>
__global__ __launch_bounds__(256) void k_gather(GatherArgs<T> A) {
    constexpr int NT = 256;
    constexpr int Q0 = (NTAPS - 1) / 2;
    const int TT = 1 << A.tlog;
    T *hs = reinterpret_cast<T *>(cuna_smem);  // cells tstart-Q0 .. tstart+TT-Q0+NTAPS-1
    const int tile = xcd_remap(blockIdx.x, A.ntiles);
    const int slab = A.slab0 + blockIdx.y;
    const int c = slab / A.nrep, rep = slab % A.nrep;
    const int *bk = A.bucket + (long)c * (A.Kc + 1);
    const int tstart = tile << A.tlog;
    const T *h = A.h + (long)slab * A.h_stride;
    const int span = TT + NTAPS;
    for (int i = threadIdx.x; i < span; i += NT) {
        int cell = (tstart - Q0 + i) % A.G;
        if (cell < 0) cell += A.G;
        hs[i] = h[cell];
    }
    __syncthreads();
    const long send = bk[3 * tile + 3];
    for (long s = bk[3 * tile] + threadIdx.x; s < send; s += NT) {
        const PointRec<T> pr = A.rec[s];
        const int p = A.perm[s];
        const T e = pr.e;
        const int rel = pr.a - tstart;  // hs index of tap q is rel + q
        T w[NTAPS];
        tap_weights<T, NTAPS>(A.wscale, e, A.bi, A.e3, w);
        T sw = T(0), sd = T(0);
#pragma unroll
        for (int q = 0; q < NTAPS; ++q) {
            const T hv = hs[rel + q] * w[q];
            sw += hv;
            sd += ((T)(q - Q0) - e) * hv;
        }
        const long oo = (long)rep * A.out_stride;
        if (A.dy) A.dy[oo + (long)p * A.dy_elem] = sw;
        if (A.dx) {
            const T v = A.y ? A.y[(long)rep * A.y_stride + p] : T(1);
            A.dx[oo + p] = A.dxscale * v * sd;
        }
    }
}

}  // namespace cuna
