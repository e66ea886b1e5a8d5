/* Oracle (test infrastructure only): FP64 real FFT, FFTW r2c/c2r semantics. */
#ifndef ORACLE_FFT_H
#define ORACLE_FFT_H

typedef struct offt_plan offt_plan;

offt_plan *offt_create(long n);
void offt_destroy(offt_plan *p);
long offt_size(const offt_plan *p);
/* work must hold n doubles */
void offt_r2c(const offt_plan *p, const double *x, double *X, double *work);
void offt_c2r(const offt_plan *p, const double *X, double *x, double *work);

#endif
