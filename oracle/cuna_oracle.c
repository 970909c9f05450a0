/*
 * cuna_oracle.c -- CPU restatement of the CUNA adjoint NFFT.
 *
 * TEST INFRASTRUCTURE ONLY.  It is the checker for the HIP library and the CPU
 * baseline that bench.py times; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it.  The product never calls it.
 *
 * Each function cites the reference file:line whose behaviour it restates.
 * Built with -ffp-contract=off so that the float decisions of the single
 * precision build are evaluated with one rounding per operation.
 */
#include "cuna_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---------------------------------------------------------------- RNG --- */

/* src/gridding.cu:67-74 */
uint32_t orc_wang_hash(uint32_t s) {
    s = (s ^ 61u) ^ (s >> 16);
    s *= 9u;
    s = s ^ (s >> 4);
    s *= 0x27d4eb2du;
    s = s ^ (s >> 15);
    return s;
}

/* src/gridding.cu:76-82 */
uint32_t orc_xorshift(uint32_t s) {
    s ^= (s << 13);
    s ^= (s >> 17);
    s ^= (s << 5);
    return s;
}

/* src/gridding.cu:90-101: thread t draws d = xorshift(wang(seed + t)) % Ndata */
void orc_bootstrap_indices(uint32_t seed, int n, int nbs, int32_t *d) {
    long total = (long)n * nbs;
    for (long t = 0; t < total; t++)
        d[t] = (int32_t)(orc_xorshift(orc_wang_hash(seed + (uint32_t)t)) % (uint32_t)n);
}

/* ----------------------------------------------------- compat formula --- */

/* src/gaussian_filter.cu:29-33 (radius), :125-134 (b, normfac).
 * float build: R = (float)ng / n; b = (float)(((2R)/(2R-1))_float * (6/PI)_double) */
void orc_compat_params(int n, int ng, int dbl, double *b, double *normfac, int *radius) {
    if (!dbl) {
        volatile float R = (float)ng / (float)n;
        volatile float num = 2.0f * R;
        volatile float den = 2.0f * R - 1.0f;
        volatile float q = num / den;
        float bf = (float)((double)q * (6.0 / ORC_PI));
        *b = (double)bf;
        *normfac = (double)(float)sqrt(2.0 * ORC_PI);
        *radius = 6;
    } else {
        double R = (double)ng / n;
        *b = 2.0 * R / (2.0 * R - 1.0) * (12.0 / ORC_PI);
        *normfac = sqrt(2.0 * ORC_PI);
        *radius = 12;
    }
}

/* src/gaussian_filter.cu:213-216: u = (int)(Ngrid*x - R_f), eps = Ngrid*x - u,
 * both in the build's precision; src/gridding.cu:43: j = (int)((x+0.5)*(Ngrid-1))
 * in double arithmetic (0.5 is a double literal) for both builds. */
void orc_compat_point(double x, int n, int ng, int dbl, int *u, double *eps, int *j) {
    (void)n;
    if (!dbl) {
        volatile float xf = (float)x;
        volatile float prod = (float)ng * xf;
        volatile float v = prod - 6.0f;
        int uu = (int)v;
        volatile float e = prod - (float)uu;
        *u = uu;
        *eps = (double)e;
        *j = (int)(((double)xf + 0.5) * (double)(ng - 1));
    } else {
        double prod = (double)ng * x;
        int uu = (int)(prod - 12.0);
        *u = uu;
        *eps = prod - (double)uu;
        *j = (int)((x + 0.5) * (double)(ng - 1));
    }
}

/* weight of tap m: E1*E2^m*E3[|m|] = exp(-(eps-m)^2/b)/sqrt(pi*b)
 * (src/gaussian_filter.cu:219-225 and :264-273) */
static double compat_weight(double eps, int m, double b) {
    double d = eps - (double)m;
    return exp(-d * d / b) / sqrt(ORC_PI * b);
}

/* src/gridding.cu:36-65 (plain) */
int orc_compat_grid(const double *x, const double *y, int n, int ng, int dbl, double *grid) {
    if (n <= 0 || ng <= 0) return -1;
    double b, normfac;
    int R;
    orc_compat_params(n, ng, dbl, &b, &normfac, &R);
    memset(grid, 0, sizeof(double) * (size_t)ng);
    for (int i = 0; i < n; i++) {
        int u, j;
        double eps;
        orc_compat_point(x[i], n, ng, dbl, &u, &eps, &j);
        if (j >= ng) j = ng - 1; /* :44-45 */
        if (j < 0) j = 0;
        double val = y ? (dbl ? y[i] : (double)(float)y[i]) : 1.0;
        int mstart = -R + 1, mend = R; /* :53-58 */
        if (j + mstart < 0) mstart = -j;
        if (j + mend > ng) mend = ng - j;
        for (int m = mstart; m < mend; m++) grid[j + m] += val * compat_weight(eps, m, b);
    }
    return 0;
}

/* src/gridding.cu:86-117 (bootstrap): grid index from point p = t % n without
 * clamp, value and weights from the drawn point d */
int orc_compat_grid_bootstrap(const double *x, const double *y, int n, int ng, int nbs,
                              uint32_t seed, int dbl, double *grid) {
    if (n <= 0 || ng <= 0 || nbs <= 0) return -1;
    double b, normfac;
    int R;
    orc_compat_params(n, ng, dbl, &b, &normfac, &R);
    memset(grid, 0, sizeof(double) * (size_t)ng * (size_t)nbs);
    double *eps = (double *)malloc(sizeof(double) * (size_t)n);
    int *jj = (int *)malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; i++) {
        int u;
        orc_compat_point(x[i], n, ng, dbl, &u, &eps[i], &jj[i]);
    }
    long total = (long)n * nbs;
    for (long t = 0; t < total; t++) {
        int p = (int)(t % n);
        int j = jj[p];
        int d;
        double val;
        if (y) {
            d = (int)(orc_xorshift(orc_wang_hash(seed + (uint32_t)t)) % (uint32_t)n);
            val = dbl ? y[d] : (double)(float)y[d];
        } else {
            d = p;
            val = 1.0;
        }
        int mstart = -R + 1, mend = R;
        if (j + mstart < 0) mstart = -j;
        if (j + mend > ng) mend = ng - j;
        double *slab = grid + (size_t)(t / n) * (size_t)ng;
        for (int m = mstart; m < mend; m++) slab[j + m] += val * compat_weight(eps[d], m, b);
    }
    free(eps);
    free(jj);
    return 0;
}

/* ------------------------------------------------------------------ FFT --- */

static int is_pow2(long v) { return v > 0 && (v & (v - 1)) == 0; }

/* exp(sign * 2 pi i * num / den) with the argument reduced exactly */
static void unit_root(long num, long den, int sign, double *c, double *s) {
    long r = num % den;
    if (r < 0) r += den;
    double a = 2.0 * ORC_PI * (double)r / (double)den;
    *c = cos(a);
    *s = sign * sin(a);
}

/* iterative radix-2, twiddles from direct cos/sin (no recurrence) */
static void fft_pow2(double *z, long n, int sign) {
    for (long i = 1, j = 0; i < n; i++) {
        long bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t0 = z[2 * i], t1 = z[2 * i + 1];
            z[2 * i] = z[2 * j];
            z[2 * i + 1] = z[2 * j + 1];
            z[2 * j] = t0;
            z[2 * j + 1] = t1;
        }
    }
    double *tw = (double *)malloc(sizeof(double) * (size_t)n);
    for (long k = 0; k < n / 2; k++) unit_root(k, n, sign, &tw[2 * k], &tw[2 * k + 1]);
    for (long len = 2; len <= n; len <<= 1) {
        long half = len >> 1, step = n / len;
#pragma omp parallel for schedule(static) if (n >= 65536)
        for (long blk = 0; blk < n; blk += len) {
            for (long k = 0; k < half; k++) {
                double wr = tw[2 * k * step], wi = tw[2 * k * step + 1];
                double *a = z + 2 * (blk + k), *bp = z + 2 * (blk + k + half);
                double br = bp[0] * wr - bp[1] * wi, bi = bp[0] * wi + bp[1] * wr;
                bp[0] = a[0] - br;
                bp[1] = a[1] - bi;
                a[0] += br;
                a[1] += bi;
            }
        }
    }
    free(tw);
}

/* Bluestein chirp-z for any n: g*r = (g^2 + r^2 - (r-g)^2)/2 */
static void fft_bluestein(double *z, long n, int sign) {
    long L = 1;
    while (L < 2 * n - 1) L <<= 1;
    double *a = (double *)calloc((size_t)(2 * L), sizeof(double));
    double *bb = (double *)calloc((size_t)(2 * L), sizeof(double));
    double *ch = (double *)malloc(sizeof(double) * (size_t)(2 * n));
    for (long q = 0; q < n; q++) {
        /* chirp c_q = exp(sign*i*pi*q^2/n) = unit_root(q^2 mod 2n, 2n) */
        long q2 = (long)(((__int128)q * q) % (2 * n));
        unit_root(q2, 2 * n, sign, &ch[2 * q], &ch[2 * q + 1]);
    }
    for (long g = 0; g < n; g++) {
        a[2 * g] = z[2 * g] * ch[2 * g] - z[2 * g + 1] * ch[2 * g + 1];
        a[2 * g + 1] = z[2 * g] * ch[2 * g + 1] + z[2 * g + 1] * ch[2 * g];
    }
    for (long q = 0; q < n; q++) {
        bb[2 * q] = ch[2 * q];
        bb[2 * q + 1] = -ch[2 * q + 1];
        if (q > 0) {
            bb[2 * (L - q)] = ch[2 * q];
            bb[2 * (L - q) + 1] = -ch[2 * q + 1];
        }
    }
    fft_pow2(a, L, -1);
    fft_pow2(bb, L, -1);
    for (long i = 0; i < L; i++) {
        double re = a[2 * i] * bb[2 * i] - a[2 * i + 1] * bb[2 * i + 1];
        double im = a[2 * i] * bb[2 * i + 1] + a[2 * i + 1] * bb[2 * i];
        a[2 * i] = re;
        a[2 * i + 1] = im;
    }
    fft_pow2(a, L, +1);
    for (long r = 0; r < n; r++) {
        double re = a[2 * r] / L, im = a[2 * r + 1] / L;
        z[2 * r] = re * ch[2 * r] - im * ch[2 * r + 1];
        z[2 * r + 1] = re * ch[2 * r + 1] + im * ch[2 * r];
    }
    free(a);
    free(bb);
    free(ch);
}

/* The reference's stage 2 is cufftExec{C2C,Z2Z}(..., CUFFT_INVERSE), unnormalised
 * (src/cuna_kernel.cu:43-49,80-99): sign = +1 restates it. */
int orc_fft(double *z, int n, int sign) {
    if (n <= 0 || (sign != 1 && sign != -1)) return -1;
    if (n == 1) return 0;
    if (is_pow2(n))
        fft_pow2(z, n, sign);
    else
        fft_bluestein(z, n, sign);
    return 0;
}

/* src/gaussian_filter.cu:278-289 (normalize): F[r] *= normfac*exp((r/ng)^2*b/4) */
static void compat_normalize(double *F, int ng, double b, double normfac) {
    for (int r = 0; r < ng; r++) {
        double K = (double)r / (double)ng;
        double fac = normfac * exp(K * K * b * 0.25);
        F[2 * r] *= fac;
        F[2 * r + 1] *= fac;
    }
}

/* src/cuna_kernel.cu:148-183: grid -> in-place inverse FFT -> normalize */
int orc_compat_adjoint(const double *x, const double *y, int n, int ng, int dbl, double *out) {
    double *grid = (double *)malloc(sizeof(double) * (size_t)ng);
    if (orc_compat_grid(x, y, n, ng, dbl, grid)) {
        free(grid);
        return -1;
    }
    for (int i = 0; i < ng; i++) {
        out[2 * i] = grid[i];
        out[2 * i + 1] = 0.0;
    }
    free(grid);
    orc_fft(out, ng, +1);
    double b, normfac;
    int R;
    orc_compat_params(n, ng, dbl, &b, &normfac, &R);
    compat_normalize(out, ng, b, normfac);
    return 0;
}

/* src/cuna_kernel.cu:237-289: bootstrap spread, batched FFT (dist ng),
 * normalize_bootstrap (src/gaussian_filter.cu:291-302, k = i % ng) */
int orc_compat_adjoint_bootstrap(const double *x, const double *y, int n, int ng, int nbs,
                                 uint32_t seed, int dbl, double *out) {
    size_t tot = (size_t)ng * (size_t)nbs;
    double *grid = (double *)malloc(sizeof(double) * tot);
    if (orc_compat_grid_bootstrap(x, y, n, ng, nbs, seed, dbl, grid)) {
        free(grid);
        return -1;
    }
    double b, normfac;
    int R;
    orc_compat_params(n, ng, dbl, &b, &normfac, &R);
    for (size_t i = 0; i < tot; i++) {
        out[2 * i] = grid[i];
        out[2 * i + 1] = 0.0;
    }
    free(grid);
    for (int s = 0; s < nbs; s++) {
        double *slab = out + 2 * (size_t)s * ng;
        orc_fft(slab, ng, +1);
        compat_normalize(slab, ng, b, normfac);
    }
    return 0;
}

/* ------------------------------------------ Greengard-Lee (accurate) --- */
/* Not a reference function: the textbook Gaussian gridding NFFT
 * (Greengard & Lee 2004) on a grid of G = sigma*M cells with 2m taps
 * g = floor(t)-m+1 .. floor(t)+m around t = G*x mod G, weight exp(-(g-t)^2/beta),
 * beta = 2*sigma*m/((2*sigma-1)*pi) (the reference's b formula with R = sigma,
 * src/gaussian_filter.cu:129), and exact deconvolution
 * F_k = exp(pi^2 k^2 beta/G^2)/sqrt(pi beta) * FFT_+(grid)[k mod G]. */

static double gl_beta(int sigma, int m) {
    return 2.0 * sigma * m / ((2.0 * sigma - 1.0) * ORC_PI);
}

static void gl_spread(const double *x, const double *y, int n, long G, int m, double beta,
                      double *grid) {
    int nth = orc_num_threads();
    if (nth < 1) nth = 1;
    double *priv = (double *)calloc((size_t)G * (size_t)nth, sizeof(double));
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *pg = priv + (size_t)tid * (size_t)G;
#pragma omp for schedule(static)
        for (int j = 0; j < n; j++) {
            /* t = G x is exact (G a power of two), and so is t - floor(t);
             * the anchor is reduced mod G in integers.  (Shifting a negative
             * t by G first would round away its last bit.) */
            double t = (double)G * x[j];
            double c = floor(t);
            double d = t - c;
            long ci = (long)c % G;
            double v = y ? y[j] : 1.0;
            for (int o = -m + 1; o <= m; o++) {
                double s = (double)o - d;
                long g = ci + o;
                g %= G;
                if (g < 0) g += G;
                pg[g] += v * exp(-s * s / beta);
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (long g = 0; g < G; g++) {
        double acc = 0.0;
        for (int t = 0; t < nth; t++) acc += priv[(size_t)t * (size_t)G + (size_t)g];
        grid[g] = acc;
    }
    free(priv);
}

int orc_gl_adjoint(const double *x, const double *y, int n, int M, int sigma, int m, double *out) {
    if (n < 0 || M <= 0 || (M & 1) || sigma < 1 || m < 1) return -1;
    long G = (long)sigma * M;
    double beta = gl_beta(sigma, m);
    double *grid = (double *)malloc(sizeof(double) * (size_t)G);
    gl_spread(x, y, n, G, m, beta, grid);
    double *z = (double *)malloc(sizeof(double) * 2 * (size_t)G);
    for (long g = 0; g < G; g++) {
        z[2 * g] = grid[g];
        z[2 * g + 1] = 0.0;
    }
    free(grid);
    orc_fft(z, (int)G, +1);
    double inv = 1.0 / sqrt(ORC_PI * beta);
    for (int i = 0; i < M; i++) {
        long k = (long)i - M / 2;
        long r = ((k % G) + G) % G;
        double kk = (double)k / (double)G;
        double D = exp(ORC_PI * ORC_PI * beta * kk * kk) * inv;
        out[2 * i] = z[2 * r] * D;
        out[2 * i + 1] = z[2 * r + 1] * D;
    }
    free(z);
    return 0;
}

/* Gradient of L through the discrete operator F = D . FFT_+ . S (see above):
 * with h = FFT_-(D g) on the grid, dL/dy_j = sum_g w(g-t_j) Re h_g and
 * dL/dx_j = -G y_j sum_g w'(g-t_j) Re h_g, w'(s) = -2 s/beta w(s). */
int orc_gl_grad(const double *x, const double *y, int n, int M, int sigma, int m,
                const double *g, double *dy, double *dx) {
    if (n < 0 || M <= 0 || (M & 1) || sigma < 1 || m < 1) return -1;
    long G = (long)sigma * M;
    double beta = gl_beta(sigma, m);
    double inv = 1.0 / sqrt(ORC_PI * beta);
    double *h = (double *)calloc(2 * (size_t)G, sizeof(double));
    for (int i = 0; i < M; i++) {
        long k = (long)i - M / 2;
        long r = ((k % G) + G) % G;
        double kk = (double)k / (double)G;
        double D = exp(ORC_PI * ORC_PI * beta * kk * kk) * inv;
        h[2 * r] = g[2 * i] * D;
        h[2 * r + 1] = g[2 * i + 1] * D;
    }
    orc_fft(h, (int)G, -1);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++) {
        double t = (double)G * x[j]; /* exact, as in gl_spread */
        double c = floor(t), d = t - c;
        long ci = (long)c % G;
        double sw = 0.0, sd = 0.0;
        for (int o = -m + 1; o <= m; o++) {
            double s = (double)o - d;
            long gg = ((ci + o) % G + G) % G;
            double w = exp(-s * s / beta);
            double re = h[2 * gg];
            sw += w * re;
            sd += (-2.0 * s / beta) * w * re;
        }
        dy[j] = sw;
        dx[j] = -(double)G * (y ? y[j] : 1.0) * sd;
    }
    free(h);
    return 0;
}

/* ----------------------------------------------------------------- NDFT --- */

/* exp(2 pi i k x) with k*x reduced mod 1 exactly (fma gives the product's
 * rounding error), so the phase stays accurate for |k| up to 2^31 */
static void phase_root(int64_t k, double x, double *c, double *s) {
    double kd = (double)k;
    double hi = kd * x;
    double lo = fma(kd, x, -hi);
    double nn = nearbyint(hi);
    double f = (hi - nn) + lo;
    double a = 2.0 * ORC_PI * f;
    *c = cos(a);
    *s = sin(a);
}

#define ORC_RESEED 1024

void orc_ndft_bands(const double *x, const double *y, int n, const int64_t *k0, int nb, int len,
                    double *out) {
    int chunks = (len + ORC_RESEED - 1) / ORC_RESEED;
    long ntask = (long)nb * chunks;
#pragma omp parallel for schedule(dynamic, 1)
    for (long task = 0; task < ntask; task++) {
        int bnd = (int)(task / chunks), ch = (int)(task % chunks);
        int off = ch * ORC_RESEED;
        int cnt = len - off < ORC_RESEED ? len - off : ORC_RESEED;
        int64_t kstart = k0[bnd] + off;
        double *acc = (double *)calloc(2 * (size_t)cnt, sizeof(double));
        for (int j = 0; j < n; j++) {
            double zr, zi, sr, si;
            phase_root(kstart, x[j], &zr, &zi);
            phase_root(1, x[j], &sr, &si);
            double v = y ? y[j] : 1.0;
            for (int q = 0; q < cnt; q++) {
                acc[2 * q] += v * zr;
                acc[2 * q + 1] += v * zi;
                double tr = zr * sr - zi * si;
                zi = zr * si + zi * sr;
                zr = tr;
            }
        }
        double *o = out + 2 * ((size_t)bnd * len + off);
        memcpy(o, acc, sizeof(double) * 2 * (size_t)cnt);
        free(acc);
    }
}

/* dL/dy_j = Re sum_k conj(g_k) e^{2 pi i k x_j},
 * dL/dx_j = Re sum_k conj(g_k) (2 pi i k) y_j e^{2 pi i k x_j}, k = -M/2..M/2-1 */
void orc_ndft_grad(const double *x, const double *y, int n, int M, const double *g,
                   double *dy, double *dx) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int j = 0; j < n; j++) {
        double sy = 0.0, sx = 0.0;
        for (int c0 = 0; c0 < M; c0 += ORC_RESEED) {
            int cnt = M - c0 < ORC_RESEED ? M - c0 : ORC_RESEED;
            double zr, zi, sr, si;
            phase_root((int64_t)c0 - M / 2, x[j], &zr, &zi);
            phase_root(1, x[j], &sr, &si);
            for (int q = 0; q < cnt; q++) {
                int i = c0 + q;
                double k = (double)(i - M / 2);
                double gr = g[2 * i], gi = g[2 * i + 1];
                /* conj(g) * z = (gr zr + gi zi) + i (gr zi - gi zr) */
                double re = gr * zr + gi * zi;
                double im = gr * zi - gi * zr;
                sy += re;
                /* Re[(2 pi i k)(re + i im)] = -2 pi k im */
                sx += -2.0 * ORC_PI * k * im;
                double tr = zr * sr - zi * si;
                zi = zr * si + zi * sr;
                zr = tr;
            }
        }
        dy[j] = sy;
        dx[j] = sx * (y ? y[j] : 1.0);
    }
}

/* Lomb-Scargle periodogram, direct FP64 (Press & Rybicki 1989, eq. 2-3;
 * Scargle 1982): data centred on its mean, "standard" normalisation.  For
 * w = 2 pi (f0 + k df), tau from tan(2 w tau) = sum sin 2wt / sum cos 2wt,
 *   P_k = [ (sum y' cos w(t-tau))^2 / sum cos^2 w(t-tau)
 *         + (sum y' sin w(t-tau))^2 / sum sin^2 w(t-tau) ] / sum y'^2,
 * a term whose denominator is below 1e-12 n dropped.  No reference
 * counterpart: CUNA's callers compute it from the two bootstrap spectra. */
void orc_lomb_scargle(const double *t, const double *y, int n, double f0, double df, int nf,
                      double *power) {
    double ybar = 0.0, yy = 0.0;
    for (int i = 0; i < n; i++) ybar += y[i];
    ybar = n > 0 ? ybar / n : 0.0;
    for (int i = 0; i < n; i++) yy += (y[i] - ybar) * (y[i] - ybar);
#pragma omp parallel for schedule(dynamic, 8)
    for (int k = 0; k < nf; k++) {
        const double f = f0 + k * df;
        double s2 = 0.0, c2 = 0.0;
        for (int i = 0; i < n; i++) {
            /* phase reduced before the trig call: 2 f t mod 1 */
            double ph = 2.0 * f * t[i];
            ph -= floor(ph);
            s2 += sin(2.0 * ORC_PI * ph);
            c2 += cos(2.0 * ORC_PI * ph);
        }
        const double wtau = 0.5 * atan2(s2, c2); /* w * tau */
        double yc = 0.0, ys = 0.0, cc = 0.0, ss = 0.0;
        for (int i = 0; i < n; i++) {
            double ph = f * t[i];
            ph -= floor(ph);
            const double a = 2.0 * ORC_PI * ph - wtau;
            const double c = cos(a), s = sin(a);
            yc += (y[i] - ybar) * c;
            ys += (y[i] - ybar) * s;
            cc += c * c;
            ss += s * s;
        }
        double p = 0.0;
        if (cc > 1e-12 * n) p += yc * yc / cc;
        if (ss > 1e-12 * n) p += ys * ys / ss;
        power[k] = yy > 0.0 ? p / yy : 0.0;
    }
}

/* Forward (type-2) NDFT, the adjoint of orc_ndft_bands over k = -M/2..M/2-1:
 * f_j = sum_k fhat_k exp(-2 pi i k x_j) (fhat and f interleaved complex),
 * with the same exact phase reduction, re-seeded every ORC_RESEED modes.
 * No reference counterpart (CUNA has only the adjoint transform). */
void orc_ndft_forward(const double *x, int n, const double *fhat, int M, double *f) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int j = 0; j < n; j++) {
        double sr = 0.0, si = 0.0;
        for (int c0 = 0; c0 < M; c0 += ORC_RESEED) {
            int cnt = M - c0 < ORC_RESEED ? M - c0 : ORC_RESEED;
            double zr, zi, wr, wi;
            /* exp(+2 pi i k x) then conjugated */
            phase_root((int64_t)c0 - M / 2, x[j], &zr, &zi);
            phase_root(1, x[j], &wr, &wi);
            for (int q = 0; q < cnt; q++) {
                int i = c0 + q;
                double gr = fhat[2 * i], gi = fhat[2 * i + 1];
                /* g * conj(z) */
                sr += gr * zr + gi * zi;
                si += gi * zr - gr * zi;
                double tr = zr * wr - zi * wi;
                zi = zr * wi + zi * wr;
                zr = tr;
            }
        }
        f[2 * j] = sr;
        f[2 * j + 1] = si;
    }
}
