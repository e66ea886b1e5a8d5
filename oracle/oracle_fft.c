/*
 * Oracle (test infrastructure only): FP64 real FFT used by the CPU
 * restatement of libmadam's noise-prior filter.
 *
 * The reference calls FFTW3 (third-party, not vendored in the reference
 * tree; no version pinned): src/fourier_fftw_2003.f90:69-70 plans
 * fftw_plan_dft_r2c_1d / fftw_plan_dft_c2r_1d, dfft (101-109) is the
 * unnormalised forward r2c transform X_k = sum_n x_n exp(-2 pi i k n / N),
 * k = 0..N/2, and dfftinv (113-123) is the c2r inverse multiplied by 1/N.
 * FFTW's c2r ignores the imaginary parts of X_0 and X_{N/2}.
 *
 * Here: iterative radix-2 complex FFT of length N/2 plus the standard
 * real-to-half-complex split.  Parity against numpy.fft is checked in
 * tests/test_oracle_fft.py.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_fft.h"

struct offt_plan {
    long n;          /* real length, power of two >= 4 */
    long h;          /* n / 2 complex points */
    double *tw;      /* h/2 complex twiddles exp(-2 pi i k / h) */
    double *rtw;     /* n/4+1 complex twiddles exp(-2 pi i k / n) */
    long *rev;       /* bit reversal of 0..h-1 */
};

static const double TWO_PI = 6.283185307179586476925286766559;

offt_plan *offt_create(long n)
{
    if (n < 4 || (n & (n - 1)) != 0) return NULL;
    offt_plan *p = (offt_plan *)calloc(1, sizeof(*p));
    p->n = n;
    p->h = n / 2;
    long h = p->h;
    p->tw = (double *)malloc(sizeof(double) * 2 * (h / 2 + 1));
    for (long k = 0; k <= h / 2; ++k) {
        double a = TWO_PI * (double)k / (double)h;
        p->tw[2 * k] = cos(a);
        p->tw[2 * k + 1] = -sin(a);
    }
    p->rtw = (double *)malloc(sizeof(double) * 2 * (n / 4 + 1));
    for (long k = 0; k <= n / 4; ++k) {
        double a = TWO_PI * (double)k / (double)n;
        p->rtw[2 * k] = cos(a);
        p->rtw[2 * k + 1] = -sin(a);
    }
    p->rev = (long *)malloc(sizeof(long) * h);
    int bits = 0;
    while ((1L << bits) < h) ++bits;
    for (long i = 0; i < h; ++i) {
        long r = 0, x = i;
        for (int b = 0; b < bits; ++b) { r = (r << 1) | (x & 1); x >>= 1; }
        p->rev[i] = r;
    }
    return p;
}

void offt_destroy(offt_plan *p)
{
    if (!p) return;
    free(p->tw); free(p->rtw); free(p->rev); free(p);
}

long offt_size(const offt_plan *p) { return p->n; }

/* twiddle exp(-2 pi i k / h) for 0 <= k < h, from the quarter table */
static inline void tw_get(const offt_plan *p, long k, double *c, double *s)
{
    long h = p->h;
    if (k <= h / 2) { *c = p->tw[2 * k]; *s = p->tw[2 * k + 1]; }
    else { long m = h - k; *c = p->tw[2 * m]; *s = -p->tw[2 * m + 1]; }
}

/* in-place complex FFT of length h, sign -1 (forward) or +1 (inverse, unscaled) */
static void cfft(const offt_plan *p, double *z, int sign)
{
    long h = p->h;
    for (long i = 0; i < h; ++i) {
        long r = p->rev[i];
        if (r > i) {
            double t0 = z[2 * i], t1 = z[2 * i + 1];
            z[2 * i] = z[2 * r]; z[2 * i + 1] = z[2 * r + 1];
            z[2 * r] = t0; z[2 * r + 1] = t1;
        }
    }
    for (long len = 2; len <= h; len <<= 1) {
        long half = len >> 1, step = h / len;
        for (long j = 0; j < half; ++j) {
            double wr, wi;
            tw_get(p, j * step, &wr, &wi);
            if (sign > 0) wi = -wi;
            for (long i = j; i < h; i += len) {
                long a = i, b = i + half;
                double xr = z[2 * b] * wr - z[2 * b + 1] * wi;
                double xi = z[2 * b] * wi + z[2 * b + 1] * wr;
                z[2 * b] = z[2 * a] - xr;
                z[2 * b + 1] = z[2 * a + 1] - xi;
                z[2 * a] += xr;
                z[2 * a + 1] += xi;
            }
        }
    }
}

/* x[n] real -> X[n/2+1] complex (interleaved re,im), unnormalised */
void offt_r2c(const offt_plan *p, const double *x, double *X, double *work)
{
    long n = p->n, h = p->h;
    double *z = work;
    memcpy(z, x, sizeof(double) * n);  /* z_m = x_{2m} + i x_{2m+1} */
    cfft(p, z, -1);
    /* split */
    X[0] = z[0] + z[1];
    X[1] = 0.0;
    X[2 * h] = z[0] - z[1];
    X[2 * h + 1] = 0.0;
    for (long k = 1; k < h; ++k) {
        double zr = z[2 * k], zi = z[2 * k + 1];
        double cr = z[2 * (h - k)], ci = -z[2 * (h - k) + 1];
        double er = 0.5 * (zr + cr), ei = 0.5 * (zi + ci);      /* even part */
        double orr = 0.5 * (zi - ci), oi = -0.5 * (zr - cr);    /* odd part = (z - conj)/(2i) */
        double wr, wi;
        if (k <= n / 4) { wr = p->rtw[2 * k]; wi = p->rtw[2 * k + 1]; }
        else { long m = h - k; wr = -p->rtw[2 * m]; wi = p->rtw[2 * m + 1]; }
        X[2 * k] = er + (orr * wr - oi * wi);
        X[2 * k + 1] = ei + (orr * wi + oi * wr);
    }
}

/* X[n/2+1] complex -> x[n] real, unnormalised (x = N * ifft); imag of X_0, X_{N/2} ignored */
void offt_c2r(const offt_plan *p, const double *X, double *x, double *work)
{
    long n = p->n, h = p->h;
    double *z = work;
    for (long k = 0; k < h; ++k) {
        /* Z_k = E_k + i W^{-k} O_k with E = (X_k + conj X_{h-k})/2 ... scaled by 2 for c2r */
        double ar = X[2 * k], ai = (k == 0) ? 0.0 : X[2 * k + 1];
        double br = X[2 * (h - k)], bi = (h - k == h) ? 0.0 : -X[2 * (h - k) + 1];
        if (k == 0) { br = X[2 * h]; bi = 0.0; }
        double er = ar + br, ei = ai + bi;            /* 2 E_k */
        double dr = ar - br, di = ai - bi;            /* 2 W^k O_k */
        double wr, wi;                                /* W^{-k} = conj(exp(-2 pi i k / n)) */
        if (k <= n / 4) { wr = p->rtw[2 * k]; wi = -p->rtw[2 * k + 1]; }
        else { long m = h - k; wr = -p->rtw[2 * m]; wi = -p->rtw[2 * m + 1]; }
        double or_ = dr * wr - di * wi, oi = dr * wi + di * wr;  /* 2 O_k */
        z[2 * k] = er - oi;        /* 2E + i * 2O */
        z[2 * k + 1] = ei + or_;
    }
    cfft(p, z, +1);
    memcpy(x, z, sizeof(double) * n);
}
