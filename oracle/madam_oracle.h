/*
 * Oracle (test infrastructure only — never linked into the product library):
 * FP64 CPU restatement of libmadam's destripe() for a single MPI rank.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library.  Each function in madam_oracle.c cites the reference
 * file:line it follows.
 */
#ifndef MADAM_ORACLE_H
#define MADAM_ORACLE_H
#include <stdint.h>

typedef struct {
    int32_t status;          /* 0 = ok, otherwise error (msg) */
    char msg[512];
    int32_t nmap, ndet, nside_map, nside_cross, nthreads;
    int64_t npix_map, npix_cross, nsamp, nbase;
    /* maps at nside_map, [npix][nmap] (reference map(nmap, 0:npix-1)) */
    double *map;             /* destriped map after makemaps (0 where rejected) */
    double *binmap;          /* binned map after makemaps */
    double *cc;              /* [npix][nmap][nmap] inverse pixel matrix, 0 where rejected */
    double *crit;            /* [npix] pixmode criterion (rcond for pixmode 2) */
    int32_t *hits;           /* [npix] */
    int32_t *mask;           /* [npix] outmask */
    /* baselines [ndet][nbase] */
    double *aa, *yba, *nna;
    int64_t *base_start;     /* [nbase] first sample of each baseline */
    int64_t *base_len;       /* [nbase] */
    /* solver trace */
    int32_t niter;
    double rrinit, rzinit;
    double *rr;              /* [niter] rr/rrinit per iteration */
    int32_t nof, nband_max, nchol, ncgfallback;
    int32_t precond_kind;    /* 0 none, 1 diagonal, 2 fprecond, 3 band */
    /* timings, seconds */
    double t_setup, t_bin, t_init, t_precond, t_iter, t_final, t_total;
} oracle_result;

/*
 * subchunk_pass: which pass of the reference's subchunk loop to compute
 * (0 = full data, k = the k-th subchunk); see src/smadam.F90:194-201.
 */
oracle_result *oracle_destripe(int subchunk_pass, const char *parstring, long ndet,
                               const char *detstring, const double *detweights,
                               long nsamp, long nnz, const double *timestamps,
                               int32_t *pix, const float *pixweights, double *signal,
                               long nperiod, const long *periods, const long *npsd,
                               long npsdtot, const double *psdstarts, long npsdbin,
                               const double *psdfreqs, long npsdval, const double *psdvals);
void oracle_free(oracle_result *r);

/* dense 3x3 helpers exposed for unit tests */
int oracle_invert_eig(double *cc, int n, double *cdet, double *rcond);
int oracle_band_cholesky(double *band, long n, int nband);
void oracle_band_solve(const double *band, long n, int nband, double *z);
/* noise filter for a single PSD (tests): fcov[nof/2+1] */
int oracle_build_filter(double fsample, int nshort, long nof, int radiometers,
                        double sigma, long npsdbin, const double *psdfreqs,
                        const double *psd, double *fcov_out);
void oracle_interpolate_psd(long n_in, const double *freq, const double *psd,
                            long n_out, const double *tf, double *tp);

#endif
