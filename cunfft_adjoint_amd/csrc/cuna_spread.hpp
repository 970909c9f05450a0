// cuna_spread.hpp -- sort-by-tile Gaussian spread (and its transpose, the
// gather of the backward pass) for gfx950.
//
// Replaces fast_gridding / fast_gridding_bootstrap (src/gridding.cu:36-117),
// which scatter every tap of every point with a global atomicAdd.  Here the
// points of one x are binned once per plan by a stable counting sort into
// buckets key = 3*tile + class (class 0: taps reach the previous tile, 2: the
// next tile, 1: interior).  One workgroup owns one tile: it walks the points
// of its own bucket plus the boundary classes of its two neighbours,
// accumulates the taps that land in its tile in LDS (one private copy per
// wave, summed in a fixed order, so the result is bitwise reproducible) and
// writes the tile once with coalesced stores.  No global atomics.
#pragma once

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
    int tlog, ntiles, Kc, olo, ohi;
    bool periodic;
    const int4 *chunks;  // {curve, start, len, column}
    const long *mbase;   // matrix base per curve
    const int *nch;      // chunks per curve
    int nchunks;
    int *keys;
    PointRec<T> *rec_tmp;
    int *hist;
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
        // accurate: t = G*x mod G, anchor floor(t), e = t - floor(t)
        double xs;
        if (A.mode == SETUP_ACC_T) {
            double s = A.df * A.t[i];
            xs = s - floor(s);  // [0, 1): FP64 phase reduction
        } else {
            double xv = (double)A.x[i];
            xs = xv - floor(xv);
        }
        double tg = xs * (double)A.G;
        double c = floor(tg);
        int ci = (int)c;
        T f = (T)(tg - c);
        if (f >= T(1)) {  // rounding of tg - c to T
            f = T(0);
            ci += 1;
        }
        if (ci >= A.G) ci -= A.G;
        if (!(ci >= 0 && ci < A.G)) ci = 0;  // non-finite x: keep the index in range
        a = ci;
        e = f;
    }
}

__device__ __forceinline__ int tile_key(int a, int tlog, int olo, int ohi, int G) {
    int tile = a >> tlog;
    int ts = tile << tlog;
    int te = min(ts + (1 << tlog), G);
    int cls = (a + olo < ts) ? 0 : ((a + ohi >= te) ? 2 : 1);
    return 3 * tile + cls;
}

// one wave per chunk, blockDim/64 chunks per block; LDS counters per wave
template <typename T>
__global__ __launch_bounds__(256) void k_setup_hist(SetupArgs<T> A) {
    int *cnt_all = reinterpret_cast<int *>(cuna_smem);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ch = blockIdx.x * (blockDim.x >> 6) + wave;
    int *cnt = cnt_all + wave * A.Kc;
    for (int k = lane; k < A.Kc; k += 64) cnt[k] = 0;
    __syncthreads();
    int4 cd = make_int4(0, 0, 0, 0);
    if (ch < A.nchunks) {
        cd = A.chunks[ch];
        for (int r0 = 0; r0 < cd.z; r0 += 64) {
            int q = r0 + lane;
            if (q < cd.z) {
                long i = (long)cd.y + q;
                int a;
                T e;
                setup_point(A, i, a, e);
                int key = tile_key(a, A.tlog, A.olo, A.ohi, A.G);
                A.keys[i] = key;
                PointRec<T> pr;
                pr.a = a;
                pr.e = e;
                A.rec_tmp[i] = pr;
                atomicAdd(&cnt[key], 1);
            }
        }
    }
    if (ch == 0 && A.E3 && threadIdx.x < A.R_f) {
        // src/gaussian_filter.cu:222-226: E3[m] = exp(-m^2/b)
        int m = threadIdx.x;
        A.E3[m] = (T)exp(-(T)(m * m) / A.b);
    }
    __syncthreads();
    if (ch < A.nchunks) {
        const long base = A.mbase[cd.x];
        const int nc = A.nch[cd.x];
        for (int k = lane; k < A.Kc; k += 64) A.hist[base + (long)k * nc + cd.w] = cnt[k];
    }
}

struct ScatterArgs {
    const int4 *chunks;
    const long *mbase;
    const int *nch;
    int nchunks, Kc, kbits;
    const int *keys;
    const int *offsets;  // scanned matrix
};

template <typename T>
__global__ __launch_bounds__(256) void k_scatter(ScatterArgs S, const PointRec<T> *rec_tmp,
                                                 PointRec<T> *rec, int *perm) {
    int *cur_all = reinterpret_cast<int *>(cuna_smem);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ch = blockIdx.x * (blockDim.x >> 6) + wave;
    int *cur = cur_all + wave * S.Kc;
    int4 cd = make_int4(0, 0, 0, 0);
    if (ch < S.nchunks) {
        cd = S.chunks[ch];
        const long base = S.mbase[cd.x];
        const int nc = S.nch[cd.x];
        for (int k = lane; k < S.Kc; k += 64) cur[k] = S.offsets[base + (long)k * nc + cd.w];
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
                rec[pos] = rec_tmp[i];
                perm[pos] = (int)i;
                if (rank == gcount - 1) cur[key] = pos + 1;
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// bucket[c*(Kc+1) + k] = first sorted index of key k of curve c
static __global__ void k_buckets(const int *offsets, const long *mbase, const int *nch, const int *coff,
                          int ncurves, int Kc, int *bucket) {
    long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long total = (long)ncurves * (Kc + 1);
    if (idx >= total) return;
    int c = (int)(idx / (Kc + 1)), k = (int)(idx % (Kc + 1));
    int v;
    if (k == Kc || nch[c] == 0)
        v = k == Kc ? coff[c + 1] : coff[c];
    else
        v = offsets[mbase[c] + (long)k * nch[c]];
    bucket[idx] = v;
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

template <typename T>
struct SpreadArgs {
    const PointRec<T> *rec;
    const int *perm;
    const int *bucket;
    const int *coff;       // curve point offsets (ncurves + 1)
    const T *y;            // values by original index, or null (unit weights)
    long y_slab_stride;    // added per replicate/slab for VAL_PLAIN with several rhs
    T *grid;
    long grid_stride;      // per slab
    int G, tlog, ntiles, Kc, olo, ntaps, nrep;
    int periodic;
    int wrap1;  // periodic and a single tile covers the grid
    T bi;       // 1/beta
    T wscale;   // weight prefactor
    T e3[kMaxTaps];  // exp(-o^2/beta), o = olo + q
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

// value and weight parameter of sorted point s (original index p)
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

template <typename T, int NTAPS>
__device__ __forceinline__ void deposit(const SpreadArgs<T> &A, T *acc, int TT, int a, T e, T v,
                                        int off) {
    constexpr int Q0 = (NTAPS - 1) / 2;
    T w[NTAPS];
    tap_weights<T, NTAPS>(v * A.wscale, e, A.bi, A.e3, w);
    const int rel = a - off - Q0;  // cell of tap q = 0
    if (rel >= 0 && rel + NTAPS <= TT) {
#pragma unroll
        for (int q = 0; q < NTAPS; ++q) atomicAdd(&acc[rel + q], w[q]);
    } else if (A.wrap1) {
        // single periodic tile (TT = G, a power of two): taps wrap around
#pragma unroll
        for (int q = 0; q < NTAPS; ++q) atomicAdd(&acc[(rel + q) & (TT - 1)], w[q]);
    } else {
#pragma unroll
        for (int q = 0; q < NTAPS; ++q) {
            const int cidx = rel + q;
            if (cidx >= 0 && cidx < TT) atomicAdd(&acc[cidx], w[q]);
        }
    }
}

// Points [begin, end) of the sorted order, seen from the tile whose first
// cell is `off`.  Each thread takes U points per round and issues all their
// loads before using any, so that the dependent gathers overlap.
template <typename T, int NTAPS, int VM>
__device__ __forceinline__ void spread_range(const SpreadArgs<T> &A, T *acc, int TT, int begin,
                                             int end, int off, int c, int rep) {
    constexpr int NT = 256, U = 4;
    for (int s0 = begin + (int)threadIdx.x; s0 < end; s0 += NT * U) {
        PointRec<T> pr[U];
        int p[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int s = s0 + u * NT;
            if (s < end) {
                pr[u] = A.rec[s];
                p[u] = A.perm[s];
            }
        }
        T v[U], e[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            e[u] = pr[u].e;
            if (s0 + u * NT < end) point_value<T, VM>(A, p[u], c, rep, v[u], e[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (s0 + u * NT < end && (VM != VAL_COUNTS || v[u] != T(0)))
                deposit<T, NTAPS>(A, acc, TT, pr[u].a, e[u], v[u], off);
    }
}

// grid: blockIdx.x = tile, blockIdx.y = slab (curve * nrep + replicate)
template <typename T, int NTAPS, int VM>
__global__ __launch_bounds__(256) void k_spread(SpreadArgs<T> A) {
    constexpr int NT = 256, NW = NT / 64;
    const int TT = 1 << A.tlog;
    T *acc_all = reinterpret_cast<T *>(cuna_smem);
    const int nt = A.ntiles;
    const int tile = xcd_remap(blockIdx.x, nt);
    const int slab = blockIdx.y;
    const int c = slab / A.nrep, rep = slab % A.nrep;
    const int *bk = A.bucket + (long)c * (A.Kc + 1);
    for (int i = threadIdx.x; i < NW * TT; i += NT) acc_all[i] = T(0);
    __syncthreads();
    T *acc = acc_all + (threadIdx.x >> 6) * TT;
    const int tstart = tile << A.tlog;
    // own tile, all classes
    spread_range<T, NTAPS, VM>(A, acc, TT, bk[3 * tile], bk[3 * tile + 3], tstart, c, rep);
    if (nt > 1) {
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
        if (hp)  // previous tile's right-edge class, seen from this tile
            spread_range<T, NTAPS, VM>(A, acc, TT, bk[3 * prev + 2], bk[3 * prev + 3],
                                       (prev << A.tlog) + TT, c, rep);
        if (hn)  // next tile's left-edge class
            spread_range<T, NTAPS, VM>(A, acc, TT, bk[3 * next], bk[3 * next + 1],
                                       (next << A.tlog) - TT, c, rep);
    }
    __syncthreads();
    T *g = A.grid + (long)slab * A.grid_stride + tstart;
    const int tlen = min(TT, A.G - tstart);
    for (int i = threadIdx.x; i < tlen; i += NT) {
        T s = acc_all[i];
#pragma unroll
        for (int w = 1; w < NW; ++w) s += acc_all[w * TT + i];
        g[i] = s;
    }
}

// ------------------------------------------- fused batched bootstrap --
// One persistent workgroup per CU drains a queue of (curve, replicate,
// transform) items.  An item draws the replicate's points
// d = c0 + xorshift(wang(seed + rep*n_c + p)) % n_c, p < n_c, spreads them
// into a grid of G = 2N cells held entirely in LDS, transforms it in place
// (real input as N complex) and writes the deconvolved spectrum:
// transform 0 is y at the points x_d, transform 1 the unit-weight window at 2 x_d.
template <typename T>
struct BootArgs {
    const PointRec<T> *rec_x;   // (anchor, frac) of x by original index
    const PointRec<T> *rec_2x;  // of 2x (window), nullable
    const T *y;                 // nullable: unit weights
    const int *coff;
    const int *order;           // curves, longest first
    int *counter;               // work-queue head, zeroed before the launch
    cplx<T> *out_y, *out_w;
    int ncurves, nbs, ntr, total;
    uint32_t seed;
    T bi, wscale;
    T e3[kMaxTaps];
    SinkAccurate<T> sink;       // out pointer set per item
    const cplx<T> *tw;
    Tw2<T> twG;
};

static __global__ void k_reset_counter(int *c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0;
}

template <typename T, int LG, int NTAPS, int NT>
__global__ __launch_bounds__(NT) void k_boot_fused(BootArgs<T> A) {
    constexpr int N = 1 << LG, G = 2 * N, Q0 = (NTAPS - 1) / 2, U = 4;
    T *grid = reinterpret_cast<T *>(cuna_smem);
    cplx<T> *z = reinterpret_cast<cplx<T> *>(cuna_smem);
    __shared__ int item_sh;
    for (;;) {
        if (threadIdx.x == 0) item_sh = atomicAdd(A.counter, 1);
        for (int i = threadIdx.x; i < G; i += NT) grid[i] = T(0);
        __syncthreads();
        const int item = item_sh;
        if (item >= A.total) break;  // uniform over the workgroup
        const int per_curve = A.nbs * A.ntr;
        const int ci = A.order[item / per_curve];
        const int rem = item % per_curve;
        const int rep = rem / A.ntr, tr = rem % A.ntr;
        const int c0 = A.coff[ci];
        const int nc = A.coff[ci + 1] - c0;
        const PointRec<T> *P = tr == 0 ? A.rec_x : A.rec_2x;
        for (int p0 = threadIdx.x; p0 < nc; p0 += NT * U) {
            PointRec<T> pr[U];
            T v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int p = p0 + u * NT;
                if (p < nc) {
                    const uint32_t t = (uint32_t)rep * (uint32_t)nc + (uint32_t)p;
                    const int d = c0 + (int)(xorshift32(wang_hash(A.seed + t)) % (uint32_t)nc);
                    pr[u] = P[d];
                    v[u] = (tr == 0 && A.y) ? A.y[d] : T(1);
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (p0 + u * NT < nc) {
                    T w[NTAPS];
                    tap_weights<T, NTAPS>(v[u] * A.wscale, pr[u].e, A.bi, A.e3, w);
                    const int base = pr[u].a - Q0;
#pragma unroll
                    for (int q = 0; q < NTAPS; ++q) atomicAdd(&grid[(base + q) & (G - 1)], w[q]);
                }
            }
        }
        __syncthreads();
        lds_fft_inplace<T, 1, LG, NT>(z, A.tw);
        SinkAccurate<T> sk = A.sink;
        cplx<T> *o = (tr == 0 ? A.out_y : A.out_w) + ((long)ci * A.nbs + rep) * sk.M;
        for (int r = threadIdx.x; r <= N / 2; r += NT)
            emit_pair(sk, o, (long)r, (long)N, z[r], z[(N - r) & (N - 1)], A.twG(r));
        __syncthreads();
    }
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
    int G, tlog, ntiles, Kc, olo, nrep;
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
    const int slab = blockIdx.y;
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
        if (A.dy) A.dy[oo + p] = sw;
        if (A.dx) {
            const T v = A.y ? A.y[(long)rep * A.y_stride + p] : T(1);
            A.dx[oo + p] = A.dxscale * v * sd;
        }
    }
}

}  // namespace cuna
