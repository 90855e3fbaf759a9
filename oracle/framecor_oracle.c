/*
 * framecor_oracle.c -- the CPU oracle of the framecor correlator.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (slipcam_amd/) loads this library. It is loaded
 * by the tests, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py, always as the
 * checker (or the timed CPU baseline), never as the code path that is measured or shipped.
 *
 * It holds two correlators over the same windows, frames and outputs as the reference:
 *
 *  oracle_correlate_fft -- a C restatement of the reference CPU backend:
 *      int64 summed-area tables .................. cxx/cpu.cpp:24-42
 *      zero-mean, zero-padded reference windows .. cxx/common.cpp:30-48
 *      cached spectra of the reference windows ... cxx/cpu.cpp:92-102
 *      per window: search window minus its mean, r2c, times conj(reference spectrum), c2r,
 *      normalization with the tables, first maximum, peak fit .... cxx/cpu.cpp:117-167
 *      per-window math ........................... cxx/surface.h:21-146
 *    FFTW (which the reference links) is not available here, so the transforms are this file's own
 *    mixed-radix Stockham FFT in FP32 arithmetic with twiddles rounded from FP64. The results agree
 *    with FFTW's to FP32 rounding, not bit for bit.
 *
 *  oracle_correlate_direct -- the FP64 direct-NCC oracle: the correlation sums are exact integers
 *    (int64 sums of uint8 products), the NCC is evaluated in FP64 from exact integer moments,
 *        ncc = (n*sum(FT) - sum(T)*sum(F)) / sqrt((n*sum(T^2) - sum(T)^2) * (n*sum(F^2) - sum(F)^2)),
 *    the argmax is taken on the FP64 surface (first maximum in row-major order), and the peak fit,
 *    SNR and covariance of cxx/surface.h:71-146 are evaluated in FP64 on the FP64 surface.
 *
 * Parity pins: tests/test_oracle.py checks the FFT against numpy.fft, the tables against numpy
 * cumsum, the direct sums against brute force and scipy, the peak fit against closed-form
 * paraboloids, and both correlators against the known-answer shift tests of the reference
 * (tests/test_framecor.py:34-95 of the reference) and against each other.
 *
 * Build: oracle/Makefile (gcc -O2 -std=c11 -ffp-contract=off -fopenmp).
 */

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_PI 3.14159265358979323846

/* ------------------------------------------------------------------------------------------------
 * Complex FFT: Stockham autosort, mixed radix (4, 2, 3, 5, then any odd factor), FP32 arithmetic.
 * Data are interleaved (re, im) floats. sign -1: forward; +1: inverse (unnormalized).
 * ---------------------------------------------------------------------------------------------- */

typedef struct {
    int n;
    int nf;
    int fac[40];
    float *tw;      /* n complex: exp(-2 pi i k / n), from FP64 */
} cplan;

static int cplan_init(cplan *p, int n)
{
    p->n = n;
    p->nf = 0;
    int m = n;
    static const int first[4] = {4, 2, 3, 5};
    for (int t = 0; t < 4; t++)
        while (m % first[t] == 0 && p->nf < 40) {
            p->fac[p->nf++] = first[t];
            m /= first[t];
        }
    for (int r = 7; m > 1 && p->nf < 40; r += 2)
        while (m % r == 0) {
            p->fac[p->nf++] = r;
            m /= r;
        }
    if (m > 1) return -1;
    p->tw = (float *)malloc(sizeof(float) * 2 * (size_t)(n > 0 ? n : 1));
    if (!p->tw) return -1;
    for (int k = 0; k < n; k++) {
        const double a = -2.0 * ORACLE_PI * (double)k / (double)n;
        p->tw[2 * k] = (float)cos(a);
        p->tw[2 * k + 1] = (float)sin(a);
    }
    return 0;
}

static void cplan_free(cplan *p)
{
    free(p->tw);
    p->tw = NULL;
}

/* x: n complex in/out; y: n complex workspace */
static void cfft(const cplan *p, float *x, float *y, int sign)
{
    const int n = p->n;
    const float *tw = p->tw;
    float *src = x, *dst = y;
    int Ns = 1;
    for (int s = 0; s < p->nf; s++) {
        const int r = p->fac[s], nr = n / r, step = n / (Ns * r);
        for (int j = 0; j < nr; j++) {
            float vr[r], vi[r];
            const int jm = j % Ns;
            for (int q = 0; q < r; q++) {
                float ar = src[2 * (j + q * nr)], ai = src[2 * (j + q * nr) + 1];
                if (q && jm) {
                    const int t = (jm * q * step) % n;
                    const float wr = tw[2 * t], wi = sign < 0 ? tw[2 * t + 1] : -tw[2 * t + 1];
                    const float br = ar * wr - ai * wi, bi = ar * wi + ai * wr;
                    ar = br;
                    ai = bi;
                }
                vr[q] = ar;
                vi[q] = ai;
            }
            float Vr[r], Vi[r];
            if (r == 2) {
                Vr[0] = vr[0] + vr[1]; Vi[0] = vi[0] + vi[1];
                Vr[1] = vr[0] - vr[1]; Vi[1] = vi[0] - vi[1];
            }
            else if (r == 4) {
                const float s0r = vr[0] + vr[2], s0i = vi[0] + vi[2];
                const float d0r = vr[0] - vr[2], d0i = vi[0] - vi[2];
                const float s1r = vr[1] + vr[3], s1i = vi[1] + vi[3];
                const float d1r = vr[1] - vr[3], d1i = vi[1] - vi[3];
                /* d1 times -i (forward) or +i (inverse) */
                const float er = sign < 0 ? d1i : -d1i, ei = sign < 0 ? -d1r : d1r;
                Vr[0] = s0r + s1r; Vi[0] = s0i + s1i;
                Vr[1] = d0r + er;  Vi[1] = d0i + ei;
                Vr[2] = s0r - s1r; Vi[2] = s0i - s1i;
                Vr[3] = d0r - er;  Vi[3] = d0i - ei;
            }
            else {
                const int rs = n / r;
                for (int c = 0; c < r; c++) {
                    float accr = 0.0f, acci = 0.0f;
                    for (int q = 0; q < r; q++) {
                        const int t = ((q * c) % r) * rs;
                        const float wr = tw[2 * t], wi = sign < 0 ? tw[2 * t + 1] : -tw[2 * t + 1];
                        accr += vr[q] * wr - vi[q] * wi;
                        acci += vr[q] * wi + vi[q] * wr;
                    }
                    Vr[c] = accr;
                    Vi[c] = acci;
                }
            }
            const int base = (j / Ns) * Ns * r + jm;
            for (int q = 0; q < r; q++) {
                dst[2 * (base + q * Ns)] = Vr[q];
                dst[2 * (base + q * Ns) + 1] = Vi[q];
            }
        }
        float *t = src;
        src = dst;
        dst = t;
        Ns *= r;
    }
    if (src != x) memcpy(x, src, sizeof(float) * 2 * (size_t)n);
}

/* the 2-D real transform of an L x L real array (row stride L) into its L x (L/2 + 1) half spectrum
 * (row stride L/2 + 1, interleaved complex); work: 4 L complex */
static void rfft2(const cplan *p, const float *in, float *spec, float *work)
{
    const int L = p->n, H = L / 2 + 1;
    float *z = work, *y = work + 2 * L;
    /* rows, two at a time: z = a + i b, then A, B from Z and conj(Z(L - k)) */
    for (int r = 0; r < L; r += 2) {
        const int pair = r + 1 < L;
        for (int x = 0; x < L; x++) {
            z[2 * x] = in[(size_t)r * L + x];
            z[2 * x + 1] = pair ? in[(size_t)(r + 1) * L + x] : 0.0f;
        }
        cfft(p, z, y, -1);
        for (int k = 0; k < H; k++) {
            const int kk = (L - k) % L;
            const float zr = z[2 * k], zi = z[2 * k + 1], cr = z[2 * kk], ci = -z[2 * kk + 1];
            spec[2 * ((size_t)r * H + k)] = 0.5f * (zr + cr);
            spec[2 * ((size_t)r * H + k) + 1] = 0.5f * (zi + ci);
            if (pair) {
                /* (Z - conj(Z')) / (2 i) */
                spec[2 * ((size_t)(r + 1) * H + k)] = 0.5f * (zi - ci);
                spec[2 * ((size_t)(r + 1) * H + k) + 1] = -0.5f * (zr - cr);
            }
        }
    }
    /* columns */
    for (int k = 0; k < H; k++) {
        for (int r = 0; r < L; r++) {
            z[2 * r] = spec[2 * ((size_t)r * H + k)];
            z[2 * r + 1] = spec[2 * ((size_t)r * H + k) + 1];
        }
        cfft(p, z, y, -1);
        for (int r = 0; r < L; r++) {
            spec[2 * ((size_t)r * H + k)] = z[2 * r];
            spec[2 * ((size_t)r * H + k) + 1] = z[2 * r + 1];
        }
    }
}

/* the inverse of rfft2, unnormalized (L^2 times the inverse), for the Hermitian half spectrum
 * spec (overwritten); the imaginary parts of the columns 0 and L/2 (L even) are ignored in the row
 * pass, as a complex-to-real transform does */
static void irfft2(const cplan *p, float *spec, float *out, float *work)
{
    const int L = p->n, H = L / 2 + 1;
    float *z = work, *y = work + 2 * L;
    for (int k = 0; k < H; k++) {
        for (int r = 0; r < L; r++) {
            z[2 * r] = spec[2 * ((size_t)r * H + k)];
            z[2 * r + 1] = spec[2 * ((size_t)r * H + k) + 1];
        }
        cfft(p, z, y, +1);
        for (int r = 0; r < L; r++) {
            spec[2 * ((size_t)r * H + k)] = z[2 * r];
            spec[2 * ((size_t)r * H + k) + 1] = z[2 * r + 1];
        }
    }
    for (int r = 0; r < L; r += 2) {
        const int pair = r + 1 < L;
        const float *A = spec + 2 * (size_t)r * H, *B = pair ? spec + 2 * (size_t)(r + 1) * H : NULL;
        for (int k = 0; k < L; k++) {
            float ar, ai, br = 0.0f, bi = 0.0f;
            if (k < H) {
                ar = A[2 * k];
                ai = A[2 * k + 1];
                if (B) {
                    br = B[2 * k];
                    bi = B[2 * k + 1];
                }
                if (k == 0 || 2 * k == L) {
                    ai = 0.0f;
                    bi = 0.0f;
                }
            }
            else {
                ar = A[2 * (L - k)];
                ai = -A[2 * (L - k) + 1];
                if (B) {
                    br = B[2 * (L - k)];
                    bi = -B[2 * (L - k) + 1];
                }
            }
            /* z = a + i b */
            z[2 * k] = ar - bi;
            z[2 * k + 1] = ai + br;
        }
        cfft(p, z, y, +1);
        for (int x = 0; x < L; x++) {
            out[(size_t)r * L + x] = z[2 * x];
            if (pair) out[(size_t)(r + 1) * L + x] = z[2 * x + 1];
        }
    }
}

/* ------------------------------------------------------------------------------------------------
 * Summed-area tables and the per-window math
 * ---------------------------------------------------------------------------------------------- */

/* (h + 1) x (w + 1) int64 tables of the pixels and of their squares, first row and column zero
 * (cxx/cpu.cpp:24-42) */
void oracle_summed_area(const uint8_t *frame, int h, int w, int64_t *sat, int64_t *sat2)
{
    const size_t stride = (size_t)w + 1;
    memset(sat, 0, sizeof(int64_t) * stride);
    memset(sat2, 0, sizeof(int64_t) * stride);
    for (int y = 0; y < h; y++) {
        int64_t a = 0, b = 0;
        int64_t *s = sat + (size_t)(y + 1) * stride, *s2 = sat2 + (size_t)(y + 1) * stride;
        s[0] = s2[0] = 0;
        for (int x = 0; x < w; x++) {
            const int64_t px = frame[(size_t)y * w + x];
            a += px;
            b += px * px;
            s[x + 1] = s[x + 1 - stride] + a;
            s2[x + 1] = s2[x + 1 - stride] + b;
        }
    }
}

/* the sum of the n x n box at (y, x) (cxx/surface.h:21-25) */
static int64_t box_sum(const int64_t *sat, int64_t stride, int y, int x, int n)
{
    return sat[(y + n) * stride + x + n] - sat[y * stride + x + n] - sat[(y + n) * stride + x]
        + sat[y * stride + x];
}

/* the normalized correlation in FP32 (cxx/surface.h:30-35) */
static float normalized_f32(float corr, float norm, int64_t s, int64_t s2, int n)
{
    const int64_t v = (int64_t)n * s2 - s * s;
    if (v <= 0 || norm <= 0) return 0.0f;
    return corr / (norm * sqrtf((float)v / (float)n));
}

static void invalidate(int64_t q, float *offset, float *peak, float *snr, float *cov)
{
    offset[2 * q] = offset[2 * q + 1] = NAN;
    peak[q] = snr[q] = 0.0f;
    cov[3 * q] = cov[3 * q + 1] = 99.0f;
    cov[3 * q + 2] = 0.0f;
}

static float parabola_f32(float m, float c, float p)
{
    const float den = m - 2 * c + p;
    return den < 0 ? 0.5f * (m - p) / den : 0.0f;
}

static double parabola_f64(double m, double c, double p)
{
    const double den = m - 2 * c + p;
    return den < 0 ? 0.5 * (m - p) / den : 0.0;
}

/* the SNR and the covariance of the peak (u, v) of a surface whose values have type S (float or
 * double) (cxx/surface.h:107-145): GET(i, j) reads the surface as S; the SNR sums squares in FP64;
 * the sums of neighbours in the curvature are taken in S, then promoted, as the reference does */
#define SNR_AND_COV(S, GET, f0d)                                                                    \
    do {                                                                                            \
        const int nn = stat < M ? stat : M;                                                         \
        double sum = 0.0;                                                                           \
        int count = 0;                                                                              \
        for (int a = 0; a < nn; a++) {                                                              \
            const int i = u + a - nn / 2;                                                           \
            if (i < 0 || i >= M) continue;                                                          \
            for (int b = 0; b < nn; b++) {                                                          \
                const int j = v + b - nn / 2;                                                       \
                if (j < 0 || j >= M) continue;                                                      \
                const double c = GET(i, j);                                                         \
                sum += c * c;                                                                       \
                count++;                                                                            \
            }                                                                                       \
        }                                                                                           \
        const double mean = count > 1 ? (sum - (f0d) * (f0d)) / (count - 1) : 0.0;                 \
        snr[q] = mean > 0 ? (float)((f0d) * (f0d) / mean) : 0.0f;                                   \
        cov[3 * q] = cov[3 * q + 1] = 99.0f;                                                        \
        cov[3 * q + 2] = 0.0f;                                                                      \
        if (inside) {                                                                               \
            const double T = tsize;                                                                 \
            const S sxx = GET(u + 1, v) + GET(u - 1, v);                                            \
            const S syy = GET(u, v + 1) + GET(u, v - 1);                                            \
            const S sxy = GET(u + 1, v + 1) + GET(u - 1, v - 1) - GET(u + 1, v - 1) - GET(u - 1, v + 1); \
            const double dxx = -(sxx - 2.0 * (f0d)) * T;                                            \
            const double dyy = -(syy - 2.0 * (f0d)) * T;                                            \
            const double dxy = sxy * 0.25 * T;                                                      \
            double n2 = 1.0 - (f0d) > 0 ? 1.0 - (f0d) : 0.0;                                        \
            const double n4 = n2 * n2 * 0.5 * T;                                                    \
            n2 *= 2;                                                                                \
            const double wd = dxy * dxy - dxx * dyy;                                                \
            if (fabs(wd) >= 1e-2) {                                                                 \
                const double w2 = wd * wd;                                                          \
                cov[3 * q] = (float)((-n2 * wd * dyy + n4 * (dyy * dyy + dxy * dxy)) / w2);         \
                cov[3 * q + 1] = (float)((-n2 * wd * dxx + n4 * (dxx * dxx + dxy * dxy)) / w2);     \
                cov[3 * q + 2] = (float)((n2 * wd * dxy - n4 * (dxx + dyy) * dxy) / w2);            \
            }                                                                                       \
        }                                                                                           \
    } while (0)

/* the outputs of the window q from its FP32 surface s (M x M, row stride `stride`) and its integer
 * peak (u, v): the FP32 quadratic fit of cxx/surface.h:71-105, the FP64 SNR and covariance of
 * cxx/surface.h:107-145 */
void oracle_write_peak_f32(const float *s, int stride, int M, int u, int v, int R, int gy, int gx,
                           int stat, int tsize, int64_t q, float *offset, float *peak, float *snr,
                           float *cov)
{
    const float f0 = s[u * stride + v];
    const int inside = u > 0 && u < M - 1 && v > 0 && v < M - 1;
    float du = 0.0f, dv = 0.0f;
    if (inside) {
        float f[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                f[i][j] = s[(u + i - 1) * stride + v + j - 1];
        const float c1 = (f[2][0] + f[2][1] + f[2][2] - f[0][0] - f[0][1] - f[0][2]) / 6;
        const float c2 = (f[0][2] + f[1][2] + f[2][2] - f[0][0] - f[1][0] - f[2][0]) / 6;
        const float c3 = (f[0][0] + f[0][1] + f[0][2] + f[2][0] + f[2][1] + f[2][2]
                          - 2 * (f[1][0] + f[1][1] + f[1][2])) / 6;
        const float c5 = (f[0][0] + f[1][0] + f[2][0] + f[0][2] + f[1][2] + f[2][2]
                          - 2 * (f[0][1] + f[1][1] + f[2][1])) / 6;
        const float c4 = (f[2][2] - f[2][0] - f[0][2] + f[0][0]) / 4;
        const float det = 4 * c3 * c5 - c4 * c4;
        int fitted = 0;
        if (c3 < 0 && det > 0) {
            du = (c4 * c2 - 2 * c5 * c1) / det;
            dv = (c4 * c1 - 2 * c3 * c2) / det;
            fitted = fabsf(du) <= 1 && fabsf(dv) <= 1;
        }
        if (!fitted) {
            du = parabola_f32(f[0][1], f[1][1], f[2][1]);
            dv = parabola_f32(f[1][0], f[1][1], f[1][2]);
        }
    }
    offset[2 * q] = (float)(gy + u - R) + du;
    offset[2 * q + 1] = (float)(gx + v - R) + dv;
    peak[q] = f0;
    const double f0d = (double)f0;
#define GET_F32(i, j) (s[(i) * stride + (j)])
    SNR_AND_COV(float, GET_F32, f0d);
#undef GET_F32
}

/* the same outputs from an FP64 surface, with the quadratic fit in FP64; the offsets are rounded
 * to float once, from (gy + u - R) + du in FP64 */
void oracle_write_peak_f64(const double *s, int stride, int M, int u, int v, int R, int gy, int gx,
                           int stat, int tsize, int64_t q, float *offset, float *peak, float *snr,
                           float *cov)
{
    const double f0 = s[u * stride + v];
    const int inside = u > 0 && u < M - 1 && v > 0 && v < M - 1;
    double du = 0.0, dv = 0.0;
    if (inside) {
        double f[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                f[i][j] = s[(u + i - 1) * stride + v + j - 1];
        const double c1 = (f[2][0] + f[2][1] + f[2][2] - f[0][0] - f[0][1] - f[0][2]) / 6;
        const double c2 = (f[0][2] + f[1][2] + f[2][2] - f[0][0] - f[1][0] - f[2][0]) / 6;
        const double c3 = (f[0][0] + f[0][1] + f[0][2] + f[2][0] + f[2][1] + f[2][2]
                           - 2 * (f[1][0] + f[1][1] + f[1][2])) / 6;
        const double c5 = (f[0][0] + f[1][0] + f[2][0] + f[0][2] + f[1][2] + f[2][2]
                           - 2 * (f[0][1] + f[1][1] + f[2][1])) / 6;
        const double c4 = (f[2][2] - f[2][0] - f[0][2] + f[0][0]) / 4;
        const double det = 4 * c3 * c5 - c4 * c4;
        int fitted = 0;
        if (c3 < 0 && det > 0) {
            du = (c4 * c2 - 2 * c5 * c1) / det;
            dv = (c4 * c1 - 2 * c3 * c2) / det;
            fitted = fabs(du) <= 1 && fabs(dv) <= 1;
        }
        if (!fitted) {
            du = parabola_f64(f[0][1], f[1][1], f[2][1]);
            dv = parabola_f64(f[1][0], f[1][1], f[1][2]);
        }
    }
    offset[2 * q] = (float)((double)(gy + u - R) + du);
    offset[2 * q + 1] = (float)((double)(gx + v - R) + dv);
    peak[q] = (float)f0;
#define GET_F64(i, j) (s[(i) * stride + (j)])
    SNR_AND_COV(double, GET_F64, f0);
#undef GET_F64
}

/* ------------------------------------------------------------------------------------------------
 * The correlator
 * ---------------------------------------------------------------------------------------------- */

typedef struct {
    int h, w, W, R, L, H, M, ny, nx, nwin, stat, threads;
    int *down, *across;
    int ready;
    cplan plan;
    float *norm;        /* nwin: the norms of the zero-mean reference windows (FP32, as the reference) */
    float *spec;        /* nwin x L x H complex: the spectra of the reference windows */
    uint8_t *tmpl;      /* nwin x W x W: the reference windows */
    int64_t *tsum;      /* nwin: sum(T) */
    int64_t *tsum2;     /* nwin: sum(T^2) */
} oracle_cor;

static int nthreads_of(const oracle_cor *c)
{
#ifdef _OPENMP
    return c->threads > 0 ? c->threads : omp_get_max_threads();
#else
    (void)c;
    return 1;
#endif
}

void oracle_destroy(oracle_cor *c)
{
    if (!c) return;
    cplan_free(&c->plan);
    free(c->down);
    free(c->across);
    free(c->norm);
    free(c->spec);
    free(c->tmpl);
    free(c->tsum);
    free(c->tsum2);
    free(c);
}

/* the parameters are validated by the caller (the tests use the product's own checks) */
oracle_cor *oracle_create(int h, int w, const int *down, int ny, const int *across, int nx, int window,
                          int search, int stat, int threads)
{
    oracle_cor *c = (oracle_cor *)calloc(1, sizeof(oracle_cor));
    if (!c) return NULL;
    c->h = h;
    c->w = w;
    c->W = window;
    c->R = search;
    c->L = window + 2 * search;
    c->H = c->L / 2 + 1;
    c->M = 2 * search + 1;
    c->ny = ny;
    c->nx = nx;
    c->nwin = ny * nx;
    c->stat = stat;
    c->threads = threads;
    c->down = (int *)malloc(sizeof(int) * (size_t)ny);
    c->across = (int *)malloc(sizeof(int) * (size_t)nx);
    c->norm = (float *)malloc(sizeof(float) * (size_t)c->nwin);
    c->spec = (float *)malloc(sizeof(float) * 2 * (size_t)c->nwin * c->L * c->H);
    c->tmpl = (uint8_t *)malloc((size_t)c->nwin * window * window);
    c->tsum = (int64_t *)malloc(sizeof(int64_t) * (size_t)c->nwin);
    c->tsum2 = (int64_t *)malloc(sizeof(int64_t) * (size_t)c->nwin);
    if (!c->down || !c->across || !c->norm || !c->spec || !c->tmpl || !c->tsum || !c->tsum2
        || cplan_init(&c->plan, c->L)) {
        oracle_destroy(c);
        return NULL;
    }
    memcpy(c->down, down, sizeof(int) * (size_t)ny);
    memcpy(c->across, across, sizeof(int) * (size_t)nx);
    return c;
}

/* the reference window k, zero-mean (the mean in FP64), zero-padded to L x L; its FP32 norm
 * (cxx/common.cpp:30-48) */
static float reference_window(const oracle_cor *c, const uint8_t *frame, int k, float *padded)
{
    const int L = c->L, W = c->W;
    const int y0 = c->down[k / c->nx], x0 = c->across[k % c->nx];
    double sum = 0.0;
    for (int i = 0; i < W; i++)
        for (int j = 0; j < W; j++)
            sum += frame[(size_t)(y0 + i) * c->w + x0 + j];
    const double mean = sum / ((double)W * W);
    memset(padded, 0, sizeof(float) * (size_t)L * L);
    double sq = 0.0;
    for (int i = 0; i < W; i++)
        for (int j = 0; j < W; j++) {
            const float d = (float)(frame[(size_t)(y0 + i) * c->w + x0 + j] - mean);
            padded[i * L + j] = d;
            sq += (double)d * d;
        }
    return (float)sqrt(sq);
}

int oracle_set_reference(oracle_cor *c, const uint8_t *frame)
{
    const int L = c->L, H = c->H, W = c->W, nt = nthreads_of(c);
    const size_t S = (size_t)L * H;
    int failed = 0;
#pragma omp parallel num_threads(nt)
    {
        float *buf = (float *)malloc(sizeof(float) * (size_t)L * L);
        float *work = (float *)malloc(sizeof(float) * 8 * (size_t)L);
        if (!buf || !work) {
#pragma omp atomic write
            failed = 1;
        }
#pragma omp for schedule(dynamic, 8)
        for (int k = 0; k < c->nwin; k++) {
            if (!buf || !work) continue;
            c->norm[k] = reference_window(c, frame, k, buf);
            rfft2(&c->plan, buf, c->spec + 2 * S * (size_t)k, work);
            const int y0 = c->down[k / c->nx], x0 = c->across[k % c->nx];
            int64_t s = 0, s2 = 0;
            for (int i = 0; i < W; i++)
                for (int j = 0; j < W; j++) {
                    const int64_t px = frame[(size_t)(y0 + i) * c->w + x0 + j];
                    c->tmpl[(size_t)k * W * W + (size_t)i * W + j] = (uint8_t)px;
                    s += px;
                    s2 += px * px;
                }
            c->tsum[k] = s;
            c->tsum2[k] = s2;
        }
        free(buf);
        free(work);
    }
    if (failed) return -1;
    c->ready = 1;
    return 0;
}

/* the search window of window k with the gross offset (gy, gx); whether it is in the frame
 * (cxx/surface.h:39-45) */
static int search_window(const oracle_cor *c, int k, int gy, int gx, int *sy, int *sx)
{
    *sy = c->down[k / c->nx] + gy - c->R;
    *sx = c->across[k % c->nx] + gx - c->R;
    return *sy >= 0 && *sx >= 0 && *sy + c->L <= c->h && *sx + c->L <= c->w;
}

/* one window of one frame, the reference algorithm (cxx/cpu.cpp:117-167) */
static void fft_window(const oracle_cor *c, const uint8_t *frame, const int64_t *sat, const int64_t *sat2,
                       int64_t q, int k, const int *gross, float *buf, float *sp, float *work,
                       float *offset, float *peak, float *snr, float *cov)
{
    const int L = c->L, H = c->H, M = c->M, W = c->W;
    const int gy = gross ? gross[2 * q] : 0, gx = gross ? gross[2 * q + 1] : 0;
    const int64_t stride = c->w + 1;
    int sy, sx;
    if (c->norm[k] <= 0 || !search_window(c, k, gy, gx, &sy, &sx)) {
        invalidate(q, offset, peak, snr, cov);
        return;
    }
    const float mean = (float)box_sum(sat, stride, sy, sx, L) / ((float)L * L);
    for (int i = 0; i < L; i++)
        for (int j = 0; j < L; j++)
            buf[i * L + j] = frame[(size_t)(sy + i) * c->w + sx + j] - mean;
    rfft2(&c->plan, buf, sp, work);
    const float *r = c->spec + 2 * (size_t)k * L * H;
    for (size_t i = 0; i < (size_t)L * H; i++) {
        const float ar = sp[2 * i], ai = sp[2 * i + 1], br = r[2 * i], bi = r[2 * i + 1];
        sp[2 * i] = ar * br + ai * bi;
        sp[2 * i + 1] = ai * br - ar * bi;
    }
    irfft2(&c->plan, sp, buf, work);
    const float scale = 1.0f / ((float)L * L);
    const int n = W * W;
    float best = -2.0f;
    int bu = 0, bv = 0;
    for (int u = 0; u < M; u++)
        for (int v = 0; v < M; v++) {
            float *cc = buf + u * L + v;
            *cc = normalized_f32(*cc * scale, c->norm[k], box_sum(sat, stride, sy + u, sx + v, W),
                                 box_sum(sat2, stride, sy + u, sx + v, W), n);
            if (*cc > best) {
                best = *cc;
                bu = u;
                bv = v;
            }
        }
    if (!(best > 0)) {
        invalidate(q, offset, peak, snr, cov);
        return;
    }
    oracle_write_peak_f32(buf, L, M, bu, bv, c->R, gy, gx, c->stat, n, q, offset, peak, snr, cov);
}

/* the reference algorithm over n frames (cxx/cpu.cpp:104-113); gross: n x nwin x 2 or NULL;
 * mask: nwin flags of the windows to correlate (the others are left untouched), or NULL */
int oracle_correlate_fft(oracle_cor *c, const uint8_t *frames, int n, const int *gross,
                         const uint8_t *mask, float *offset, float *peak, float *snr, float *cov)
{
    if (!c->ready) return -2;
    const int L = c->L, H = c->H, nt = nthreads_of(c);
    const size_t pixels = (size_t)c->h * c->w, tsz = (size_t)(c->h + 1) * (c->w + 1);
    int64_t *sat = (int64_t *)malloc(sizeof(int64_t) * tsz), *sat2 = (int64_t *)malloc(sizeof(int64_t) * tsz);
    if (!sat || !sat2) {
        free(sat);
        free(sat2);
        return -1;
    }
    int failed = 0;
    for (int f = 0; f < n && !failed; f++) {
        const uint8_t *frame = frames + (size_t)f * pixels;
        oracle_summed_area(frame, c->h, c->w, sat, sat2);
#pragma omp parallel num_threads(nt)
        {
            float *buf = (float *)malloc(sizeof(float) * (size_t)L * L);
            float *sp = (float *)malloc(sizeof(float) * 2 * (size_t)L * H);
            float *work = (float *)malloc(sizeof(float) * 8 * (size_t)L);
            if (!buf || !sp || !work) {
#pragma omp atomic write
                failed = 1;
            }
#pragma omp for schedule(dynamic, 8)
            for (int k = 0; k < c->nwin; k++) {
                if (!buf || !sp || !work) continue;
                if (mask && !mask[k]) continue;
                fft_window(c, frame, sat, sat2, (int64_t)f * c->nwin + k, k, gross, buf, sp, work,
                           offset, peak, snr, cov);
            }
            free(buf);
            free(sp);
            free(work);
        }
    }
    free(sat);
    free(sat2);
    return failed ? -1 : 0;
}

/* the exact correlation sums of a W x W template with an L x L chip (row strides ldt, ldc):
 * out[u * M + v] = sum_{i,j < W} t[i][j] * chip[u + i][v + j], M = L - W + 1 */
void oracle_raw_direct(const uint8_t *chip, int ldc, int L, const uint8_t *t, int ldt, int W, int64_t *out)
{
    const int M = L - W + 1;
    for (int u = 0; u < M; u++)
        for (int v = 0; v < M; v++) {
            int64_t acc = 0;
            for (int i = 0; i < W; i++) {
                const uint8_t *a = chip + (size_t)(u + i) * ldc + v, *b = t + (size_t)i * ldt;
                int32_t row = 0;
                for (int j = 0; j < W; j++) row += (int32_t)a[j] * b[j];
                acc += row;
            }
            out[(size_t)u * M + v] = acc;
        }
}

/* the FP64 NCC from the exact moments; 0 where the search box is flat (as cxx/surface.h:30-35) */
static double ncc_f64(int64_t raw, int64_t n, int64_t tsum, int64_t vt, int64_t fs, int64_t fs2)
{
    const int64_t vf = n * fs2 - fs * fs;
    if (vf <= 0 || vt <= 0) return 0.0;
    const double num = (double)n * (double)raw - (double)tsum * (double)fs;
    return num / sqrt((double)vt * (double)vf);
}

/* the FP64 direct-NCC oracle over n frames; surfaces: NULL, or n x nwin x M x M doubles for the
 * FP64 surfaces of the valid windows (NaN elsewhere) */
int oracle_correlate_direct(oracle_cor *c, const uint8_t *frames, int n, const int *gross,
                            const uint8_t *mask, float *offset, float *peak, float *snr, float *cov,
                            double *surfaces)
{
    if (!c->ready) return -2;
    const int L = c->L, M = c->M, W = c->W, nt = nthreads_of(c);
    const int64_t nn = (int64_t)W * W, stride = c->w + 1;
    const size_t pixels = (size_t)c->h * c->w, tsz = (size_t)(c->h + 1) * (c->w + 1);
    int64_t *sat = (int64_t *)malloc(sizeof(int64_t) * tsz), *sat2 = (int64_t *)malloc(sizeof(int64_t) * tsz);
    if (!sat || !sat2) {
        free(sat);
        free(sat2);
        return -1;
    }
    int failed = 0;
    for (int f = 0; f < n && !failed; f++) {
        const uint8_t *frame = frames + (size_t)f * pixels;
        oracle_summed_area(frame, c->h, c->w, sat, sat2);
#pragma omp parallel num_threads(nt)
        {
            int64_t *raw = (int64_t *)malloc(sizeof(int64_t) * (size_t)M * M);
            double *s = (double *)malloc(sizeof(double) * (size_t)M * M);
            if (!raw || !s) {
#pragma omp atomic write
                failed = 1;
            }
#pragma omp for schedule(dynamic, 1)
            for (int k = 0; k < c->nwin; k++) {
                if (!raw || !s) continue;
                if (mask && !mask[k]) continue;
                const int64_t q = (int64_t)f * c->nwin + k;
                double *out = surfaces ? surfaces + (size_t)q * M * M : NULL;
                if (out)
                    for (int i = 0; i < M * M; i++) out[i] = NAN;
                const int gy = gross ? gross[2 * q] : 0, gx = gross ? gross[2 * q + 1] : 0;
                int sy, sx;
                if (c->norm[k] <= 0 || !search_window(c, k, gy, gx, &sy, &sx)) {
                    invalidate(q, offset, peak, snr, cov);
                    continue;
                }
                oracle_raw_direct(frame + (size_t)sy * c->w + sx, c->w, L, c->tmpl + (size_t)k * W * W, W, W, raw);
                const int64_t vt = nn * c->tsum2[k] - c->tsum[k] * c->tsum[k];
                double best = -2.0;
                int bu = 0, bv = 0;
                for (int u = 0; u < M; u++)
                    for (int v = 0; v < M; v++) {
                        const double x = ncc_f64(raw[(size_t)u * M + v], nn, c->tsum[k], vt,
                                                 box_sum(sat, stride, sy + u, sx + v, W),
                                                 box_sum(sat2, stride, sy + u, sx + v, W));
                        s[(size_t)u * M + v] = x;
                        if (x > best) {
                            best = x;
                            bu = u;
                            bv = v;
                        }
                    }
                if (out) memcpy(out, s, sizeof(double) * (size_t)M * M);
                if (!(best > 0)) {
                    invalidate(q, offset, peak, snr, cov);
                    continue;
                }
                oracle_write_peak_f64(s, M, M, bu, bv, c->R, gy, gx, c->stat, (int)nn, q, offset, peak, snr, cov);
            }
            free(raw);
            free(s);
        }
    }
    free(sat);
    free(sat2);
    return failed ? -1 : 0;
}

/* a bound of the Python slice a[i:] or a[:i] over a length-M axis (positive step) */
static int py_slice_bound(int64_t i, int M)
{
    if (i < 0) {
        i += M;
        if (i < 0) i = 0;
    }
    return i > M ? M : (int)i;
}

/* The periodicity of python/slipcam/auto_slip.py:105-129: per window, the NCC of the reference window
 * with the frame around it (its search window at its own position, no gross offset), then the largest
 * value at a local maximum (a point equal to the largest value of its 5 x 5 neighbourhood clipped to the
 * surface: scipy.ndimage.maximum_filter, size 5, whose default 'reflect' mode repeats in-surface values)
 * outside the rows and columns of the slice [search - exclude : search + exclude + 1] (:124-125); 0 if
 * there is none (:126); 1 where the template is flat, tn < 1e-6 with tn the FP64 norm of the template
 * minus its FP64 mean (:113-116). The NCC is the exact one (int64 sums, FP64 from exact moments, as
 * ncc_f64); the reference evaluates the same quotient with FP64 FFTs (:117-122).
 * out: nwin doubles. Returns -3 where a search window leaves the frame, -2 before set_reference. */
int oracle_periodicity(oracle_cor *c, const uint8_t *frame, int exclude, double *out)
{
    if (!c->ready) return -2;
    const int L = c->L, M = c->M, W = c->W, nt = nthreads_of(c);
    const int64_t nn = (int64_t)W * W, stride = c->w + 1;
    for (int k = 0; k < c->nwin; k++) {
        int sy, sx;
        if (!search_window(c, k, 0, 0, &sy, &sx)) return -3;
    }
    const int lo = py_slice_bound((int64_t)c->R - exclude, M), hi = py_slice_bound((int64_t)c->R + exclude + 1, M);
    const size_t tsz = (size_t)(c->h + 1) * (c->w + 1);
    int64_t *sat = (int64_t *)malloc(sizeof(int64_t) * tsz), *sat2 = (int64_t *)malloc(sizeof(int64_t) * tsz);
    if (!sat || !sat2) {
        free(sat);
        free(sat2);
        return -1;
    }
    oracle_summed_area(frame, c->h, c->w, sat, sat2);
    int failed = 0;
#pragma omp parallel num_threads(nt)
    {
        int64_t *raw = (int64_t *)malloc(sizeof(int64_t) * (size_t)M * M);
        double *s = (double *)malloc(sizeof(double) * (size_t)M * M);
        if (!raw || !s) {
#pragma omp atomic write
            failed = 1;
        }
#pragma omp for schedule(dynamic, 1)
        for (int k = 0; k < c->nwin; k++) {
            if (!raw || !s) continue;
            const uint8_t *t = c->tmpl + (size_t)k * W * W;
            const double mean = (double)c->tsum[k] / (double)nn;
            double sq = 0.0;
            for (int64_t i = 0; i < nn; i++) sq += ((double)t[i] - mean) * ((double)t[i] - mean);
            if (sqrt(sq) < 1e-6) {
                out[k] = 1.0;
                continue;
            }
            int sy, sx;
            search_window(c, k, 0, 0, &sy, &sx);
            oracle_raw_direct(frame + (size_t)sy * c->w + sx, c->w, L, t, W, W, raw);
            const int64_t vt = nn * c->tsum2[k] - c->tsum[k] * c->tsum[k];
            for (int u = 0; u < M; u++)
                for (int v = 0; v < M; v++)
                    s[(size_t)u * M + v] = ncc_f64(raw[(size_t)u * M + v], nn, c->tsum[k], vt,
                                                   box_sum(sat, stride, sy + u, sx + v, W),
                                                   box_sum(sat2, stride, sy + u, sx + v, W));
            int any = 0;
            double best = 0.0;
            for (int u = 0; u < M; u++)
                for (int v = 0; v < M; v++) {
                    if (u >= lo && u < hi && v >= lo && v < hi) continue;
                    const double x = s[(size_t)u * M + v];
                    double m = x;
                    for (int uu = u - 2; uu <= u + 2; uu++)
                        for (int vv = v - 2; vv <= v + 2; vv++)
                            if (uu >= 0 && uu < M && vv >= 0 && vv < M && s[(size_t)uu * M + vv] > m)
                                m = s[(size_t)uu * M + vv];
                    if (x == m && (!any || x > best)) {
                        best = x;
                        any = 1;
                    }
                }
            out[k] = any ? best : 0.0;
        }
        free(raw);
        free(s);
    }
    free(sat);
    free(sat2);
    return failed ? -1 : 0;
}

/* the parabolic refinement of python/slipcam/stabilize.py:37-39: the vertex of the parabola through
 * (-1, m), (0, c), (1, p), 0 where its curvature is below 1e-12 */
static double parabola_refine(double c, double m, double p)
{
    const double den = m - 2 * c + p;
    return fabs(den) < 1e-12 ? 0.0 : 0.5 * (m - p) / den;
}

/* The peak of python/slipcam/stabilize.py:22-45 ncc_peak for every window of the grid against one
 * frame: the NCC of the reference window with its search window (at the window's position plus its
 * gross offset; the exact NCC of ncc_f64, which the reference evaluates with FP64 FFTs, :29-34), its
 * first maximum in row-major order (np.argmax, :35-36), each coordinate refined by the parabola
 * through it and its two neighbours when it is not on the border (:40-41), relative to the centre
 * (M - 1) / 2 (:42-43). NaN, NaN, 0 for a flat template (tn < 1e-6, :24-28) or a search window leaving
 * the frame (coarse_track skips those, :85-86). gross: nwin x 2 or NULL; offsets: nwin x 2 (down,
 * across), peak: nwin. */
int oracle_ncc_peaks(oracle_cor *c, const uint8_t *frame, const int *gross, double *offsets, double *peak)
{
    if (!c->ready) return -2;
    const int L = c->L, M = c->M, W = c->W, nt = nthreads_of(c);
    const int64_t nn = (int64_t)W * W, stride = c->w + 1;
    const size_t tsz = (size_t)(c->h + 1) * (c->w + 1);
    int64_t *sat = (int64_t *)malloc(sizeof(int64_t) * tsz), *sat2 = (int64_t *)malloc(sizeof(int64_t) * tsz);
    if (!sat || !sat2) {
        free(sat);
        free(sat2);
        return -1;
    }
    oracle_summed_area(frame, c->h, c->w, sat, sat2);
    int failed = 0;
#pragma omp parallel num_threads(nt)
    {
        int64_t *raw = (int64_t *)malloc(sizeof(int64_t) * (size_t)M * M);
        double *s = (double *)malloc(sizeof(double) * (size_t)M * M);
        if (!raw || !s) {
#pragma omp atomic write
            failed = 1;
        }
#pragma omp for schedule(dynamic, 1)
        for (int k = 0; k < c->nwin; k++) {
            if (!raw || !s) continue;
            offsets[2 * k] = offsets[2 * k + 1] = NAN;
            peak[k] = 0.0;
            const uint8_t *t = c->tmpl + (size_t)k * W * W;
            const double mean = (double)c->tsum[k] / (double)nn;
            double sq = 0.0;
            for (int64_t i = 0; i < nn; i++) sq += ((double)t[i] - mean) * ((double)t[i] - mean);
            int sy, sx;
            if (sqrt(sq) < 1e-6 || !search_window(c, k, gross ? gross[2 * k] : 0, gross ? gross[2 * k + 1] : 0, &sy, &sx))
                continue;
            oracle_raw_direct(frame + (size_t)sy * c->w + sx, c->w, L, t, W, W, raw);
            const int64_t vt = nn * c->tsum2[k] - c->tsum[k] * c->tsum[k];
            int bi = 0;
            for (int i = 0; i < M * M; i++) {
                const int u = i / M, v = i % M;
                s[i] = ncc_f64(raw[i], nn, c->tsum[k], vt, box_sum(sat, stride, sy + u, sx + v, W),
                               box_sum(sat2, stride, sy + u, sx + v, W));
                if (s[i] > s[bi]) bi = i;
            }
            const int u = bi / M, v = bi % M;
            const double du = (u > 0 && u < M - 1) ? parabola_refine(s[bi], s[bi - M], s[bi + M]) : 0.0;
            const double dv = (v > 0 && v < M - 1) ? parabola_refine(s[bi], s[bi - 1], s[bi + 1]) : 0.0;
            const double c0 = (M - 1) / 2.0;
            offsets[2 * k] = u + du - c0;
            offsets[2 * k + 1] = v + dv - c0;
            peak[k] = s[bi];
        }
        free(raw);
        free(s);
    }
    free(sat);
    free(sat2);
    return failed ? -1 : 0;
}

/* the norms of the reference windows (FP32, as the reference) and the integer template moments */
int oracle_reference_stats(const oracle_cor *c, float *norm, int64_t *tsum, int64_t *tsum2)
{
    if (!c->ready) return -2;
    memcpy(norm, c->norm, sizeof(float) * (size_t)c->nwin);
    memcpy(tsum, c->tsum, sizeof(int64_t) * (size_t)c->nwin);
    memcpy(tsum2, c->tsum2, sizeof(int64_t) * (size_t)c->nwin);
    return 0;
}

/* the spectra of the reference windows, nwin x L x (L/2 + 1) complex */
const float *oracle_reference_spectra(const oracle_cor *c)
{
    return c->spec;
}

/* standalone transforms, for pinning against numpy.fft */
int oracle_cfft(int n, float *x, int sign)
{
    cplan p;
    if (cplan_init(&p, n)) return -1;
    float *y = (float *)malloc(sizeof(float) * 2 * (size_t)n);
    if (!y) {
        cplan_free(&p);
        return -1;
    }
    cfft(&p, x, y, sign);
    free(y);
    cplan_free(&p);
    return 0;
}

int oracle_rfft2(int L, const float *in, float *spec)
{
    cplan p;
    if (cplan_init(&p, L)) return -1;
    float *work = (float *)malloc(sizeof(float) * 8 * (size_t)L);
    if (!work) {
        cplan_free(&p);
        return -1;
    }
    rfft2(&p, in, spec, work);
    free(work);
    cplan_free(&p);
    return 0;
}

int oracle_irfft2(int L, float *spec, float *out)
{
    cplan p;
    if (cplan_init(&p, L)) return -1;
    float *work = (float *)malloc(sizeof(float) * 8 * (size_t)L);
    if (!work) {
        cplan_free(&p);
        return -1;
    }
    irfft2(&p, spec, out, work);
    free(work);
    cplan_free(&p);
    return 0;
}

int oracle_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
