/*
 * Oracle (test infrastructure only): FP64 CPU restatement of libmadam's
 * destripe() on one MPI rank, with OpenMP.
 *
 * This file is the checker for the HIP product path.  It is never linked into
 * libmadam.so; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  It restates the reference's algorithm and order
 * of operations (file:line citations at each function, paths relative to the
 * libmadam source tree), with these deliberate differences:
 *   - maps are indexed by global nested pixel instead of the compacted
 *     hit-submap ("locmap") index of src/pointing.f90:115-191 (a relabelling);
 *   - flagged samples are skipped everywhere.  The reference resets
 *     dummy_pixel to 12*nside_max**2 after compaction (src/smadam.F90:300),
 *     so its `ip == dummy_pixel` tests never match; they are harmless because
 *     the compacted dummy slot holds zeros, except for a thread-count dependent
 *     leak in baseline_to_map, which is not reproduced;
 *   - flag_bad_baselines (src/read_data.f90:30-75) is a no-op for the same
 *     reason (it counts dummy samples as good) and is not reproduced;
 *   - LAPACK dsyev / dpotrf / dpbtrf are replaced by a cyclic Jacobi
 *     eigen-solver, an unblocked Cholesky and an unblocked band Cholesky;
 *     FFTW by oracle_fft.c.  These agree to rounding (tests pin them).
 * Single-precision literals of the reference (e.g. `rcond < 1e-12`) are kept
 * as F32() so thresholds compare identically.
 * Unsupported (error): basis_order > 0, file_inmask, file_spectrum,
 * file_covmat, nmap not in {1, 3}.
 */
#include <math.h>
#include <omp.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "madam_oracle.h"
#include "oracle_fft.h"

#define F32(x) ((double)(float)(x))
#define PI_D 3.1415926535897932384626433832795

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ------------------------------------------------------------------------ */
/* parameters: src/commonparam.f90:200-305 (defaults), src/inputparam.f90   */
/* ------------------------------------------------------------------------ */

typedef struct {
    int info, nthreads, nsubchunk, isubchunk;
    double dnshort;
    int nside_map, nside_cross, nside_submap;
    double good_baseline_fraction;
    int pixmode_map, pixmode_cross;
    double pixlim_map, pixlim_cross;
    int incomplete_matrices, allow_decoupling, kfirst, basis_order;
    int iter_min, iter_max;
    double cglimit, fsample;
    int mode_detweight, temperature_only, noise_weights_from_psd, radiometers;
    int kfilter;
    double diagfilter;
    int precond_width_min, precond_width_max, use_fprecond, use_cgprecond;
    int rm_monopole, write_tod, ndetset, nsurvey, mcmode;
    int do_binmap, do_hits, do_mask;
    int has_inmask, has_spectrum, has_covmat;
    int nshort;
} oprm;

static void set_defaults(oprm *p)
{
    memset(p, 0, sizeof(*p));
    p->info = 2;
    p->nthreads = 1;
    p->nsubchunk = 0;
    p->isubchunk = 0;
    p->dnshort = -1;
    p->nside_map = 512;
    p->nside_cross = -1;
    p->nside_submap = 16;
    p->pixmode_map = 2;
    p->pixmode_cross = 2;
    p->pixlim_map = F32(1e-6);
    p->pixlim_cross = F32(1e-3);
    p->kfirst = 1;
    p->basis_order = 0;
    p->iter_min = 3;
    p->iter_max = 1000;
    p->cglimit = F32(1e-12);
    p->fsample = 1;
    p->mode_detweight = 0;
    p->radiometers = 1;
    p->precond_width_min = 10;
    p->precond_width_max = 100;
}

typedef struct { char *msg; int err; } perr;

static void fail(char *msg, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(msg, 512, fmt, ap);
    va_end(ap);
}

static void trim(char *s)
{
    size_t n = strlen(s);
    while (n > 0 && (s[n - 1] == ' ' || s[n - 1] == '\t' || s[n - 1] == '\n' || s[n - 1] == '\r')) s[--n] = 0;
    size_t i = 0;
    while (s[i] == ' ' || s[i] == '\t') ++i;
    if (i) memmove(s, s + i, n - i + 1);
}

/* first list-directed token */
static void first_token(const char *v, char *tok, size_t cap)
{
    size_t i = 0, j = 0;
    while (v[i] == ' ' || v[i] == '\t') ++i;
    while (v[i] && v[i] != ' ' && v[i] != ',' && v[i] != '\t' && j + 1 < cap) tok[j++] = v[i++];
    tok[j] = 0;
}

static int read_int(const char *v, int *out)
{
    char tok[128], *end;
    first_token(v, tok, sizeof tok);
    if (!tok[0]) return -1;
    long x = strtol(tok, &end, 10);
    if (*end) return -1;
    *out = (int)x;
    return 0;
}

static int read_real(const char *v, double *out)
{
    char tok[128], *end;
    first_token(v, tok, sizeof tok);
    if (!tok[0]) return -1;
    for (char *c = tok; *c; ++c) if (*c == 'd' || *c == 'D') *c = 'e';
    double x = strtod(tok, &end);
    if (*end) return -1;
    *out = x;
    return 0;
}

static int read_logical(const char *v, int *out)
{
    char tok[128];
    first_token(v, tok, sizeof tok);
    const char *c = tok;
    if (*c == '.') ++c;
    if (*c == 'T' || *c == 't') { *out = 1; return 0; }
    if (*c == 'F' || *c == 'f') { *out = 0; return 0; }
    return -1;
}

/* src/inputparam.f90:169-371 (read_line) */
static int read_line(oprm *p, char *line, char *msg)
{
    trim(line);
    if (!line[0]) return 0;
    char *eq = strchr(line, '=');
    if (line[0] == '#') return 0;
    if (!eq) { fail(msg, "Syntax error in parameter definition: %s", line); return -1; }
    *eq = 0;
    char key[1024], value[1024];
    snprintf(key, sizeof key, "%s", line);
    snprintf(value, sizeof value, "%s", eq + 1);
    trim(key);
    trim(value);
    char *hash = strchr(value, '#');
    if (hash) { *hash = 0; trim(value); }
    if (strncmp(key, "detector", 8) == 0) strcpy(key, "detector");
    if (!value[0] && strncmp(key, "file", 4) != 0) return 0;
    int e = 0;
#define KI(name, field) else if (!strcmp(key, name)) e = read_int(value, &p->field);
#define KR(name, field) else if (!strcmp(key, name)) e = read_real(value, &p->field);
#define KL(name, field) else if (!strcmp(key, name)) e = read_logical(value, &p->field);
#define KIGN_I(name) else if (!strcmp(key, name)) { int t_; e = read_int(value, &t_); }
#define KIGN_R(name) else if (!strcmp(key, name)) { double t_; e = read_real(value, &t_); }
#define KIGN_L(name) else if (!strcmp(key, name)) { int t_; e = read_logical(value, &t_); }
#define KIGN_S(name) else if (!strcmp(key, name)) { }
    if (0) {}
    KI("info", info)
    KI("nthreads", nthreads)
    KI("nsubchunk", nsubchunk)
    KI("isubchunk", isubchunk)
    else if (!strcmp(key, "time_unit")) {
        if (value[0] != 'p' && value[0] != 'P') { double t_; e = read_real(value, &t_); }
    }
    KR("base_first", dnshort)
    KR("nshort", dnshort)
    KI("nside_map", nside_map)
    KI("nside_cross", nside_cross)
    KI("nside_submap", nside_submap)
    KR("good_baseline_fraction", good_baseline_fraction)
    KIGN_L("concatenate_messages")
    KIGN_L("allreduce")
    KIGN_L("reassign_submaps")
    KI("pixmode_map", pixmode_map)
    KI("pixmode_cross", pixmode_cross)
    KR("pixlim_map", pixlim_map)
    KR("pixlim_cross", pixlim_cross)
    KL("incomplete_matrices", incomplete_matrices)
    KL("allow_decoupling", allow_decoupling)
    KL("kfirst", kfirst)
    else if (!strcmp(key, "basis_func")) {
        if (strncmp(value, "poly", 4) && strncmp(value, "four", 4) && strncmp(value, "cheb", 4) &&
            strncmp(value, "lege", 4)) {
            fail(msg, "Unknown function basis : %s", value);
            return -1;
        }
    }
    else if (!strcmp(key, "basis_order")) {
        e = read_int(value, &p->basis_order);
        if (!e && p->basis_order < 0) { fail(msg, "basis order must be nonnegative"); return -1; }
    }
    KI("iter_min", iter_min)
    KI("iter_max", iter_max)
    KR("cglimit", cglimit)
    KR("fsample", fsample)
    KI("mode_detweight", mode_detweight)
    KIGN_L("flag_by_horn")
    KIGN_L("write_cut")
    KIGN_L("checknan")
    KIGN_L("sync_output")
    KIGN_L("skip_existing")
    KL("temperature_only", temperature_only)
    KIGN_L("force_pol")
    KL("noise_weights_from_psd", noise_weights_from_psd)
    KL("radiometers", radiometers)
    KIGN_I("psdlen")
    KIGN_I("psd_down")
    KIGN_I("psd_downsample")
    KL("kfilter", kfilter)
    KR("diagfilter", diagfilter)
    KIGN_L("unaligned_fft")
    KI("precond_width_min", precond_width_min)
    KI("precond_width_max", precond_width_max)
    KL("use_fprecond", use_fprecond)
    KL("use_cgprecond", use_cgprecond)
    KL("rm_monopole", rm_monopole)
    KIGN_S("path_output")
    KIGN_S("file_root")
    KIGN_L("write_map")
    KL("write_binmap", do_binmap)
    KL("write_hits", do_hits)
    KIGN_L("write_matrix")
    KIGN_L("write_wcov")
    KIGN_L("write_base")
    KL("write_mask", do_mask)
    KIGN_L("write_leakmatrix")
    KIGN_S("unit_tod")
    KIGN_S("file_gap_out")
    KIGN_S("file_mc")
    KL("write_tod", write_tod)
    else if (!strcmp(key, "file_inmask")) p->has_inmask = value[0] != 0;
    else if (!strcmp(key, "file_spectrum")) p->has_spectrum = value[0] != 0;
    KIGN_S("file_gap")
    KIGN_L("binary_output")
    KIGN_I("nwrite_binary")
    else if (!strcmp(key, "file_covmat")) p->has_covmat = value[0] != 0;
    else if (!strcmp(key, "detset") || !strcmp(key, "detset_nopol")) {
        char *c = strchr(value, ':');
        if (!c) { fail(msg, "Failed to parse: %s for valid detector set name", value); return -1; }
        /* "name : all" names the full set (src/inputparam.f90:430-438) */
        char rest[512];
        snprintf(rest, sizeof rest, "%s", c + 1);
        trim(rest);
        if (!(rest[0] == 0 || !strcmp(rest, "all") || !strcmp(rest, "All") || !strcmp(rest, "ALL"))) p->ndetset++;
    }
    else if (!strcmp(key, "survey")) {
        char *c = strchr(value, ':');
        if (!c) { fail(msg, "Failed to parse: %s for valid survey name", value); return -1; }
        char rest[512];
        snprintf(rest, sizeof rest, "%s", c + 1);
        trim(rest);
        if (!strchr(rest, '-')) {
            if (!(rest[0] == 0 || !strcmp(rest, "all") || !strcmp(rest, "All") || !strcmp(rest, "ALL"))) {
                fail(msg, "Failed to parse: %s for valid survey span marker (-)", rest);
                return -1;
            }
        } else p->nsurvey++;
    }
    KIGN_L("bin_subsets")
    KL("mcmode", mcmode)
    else { fail(msg, "Unknown parameter: %s", key); return -1; }
    if (e) { fail(msg, "Failed to parse %s from %s", key, value); return -1; }
    return 0;
}

/* src/inputparam.f90:24-67 (read_parameters) */
static int read_parameters(oprm *p, const char *parstring, char *msg)
{
    const char *s = parstring;
    while (*s) {
        const char *e = strchr(s, ';');
        size_t n = e ? (size_t)(e - s) : strlen(s);
        char line[1024];
        if (n >= sizeof line) n = sizeof line - 1;
        memcpy(line, s, n);
        line[n] = 0;
        if (read_line(p, line, msg)) return -1;
        if (!e) break;
        s = e + 1;
    }
    /* dnshort = nint(dnshort * fsample): baseline length given in seconds */
    p->dnshort = (double)llround(p->dnshort * p->fsample);
    /* check_missing_parameters, src/inputparam.f90:500-505 */
    if (p->kfirst && p->dnshort < 0) p->dnshort = F32(1e10);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* detectors: src/inputparam.f90:73-163, src/commonparam.f90:23-36           */
/* ------------------------------------------------------------------------ */

typedef struct {
    int npsd;
    int has_psd;
    double *weights, *sigmas, *psdstarts, *plateaus;
    double *psds; /* [npsd][npsdbin] */
} odet;

typedef struct {
    oprm p;
    long ndet, nsamp, nnz;
    int nmap, nside_max;
    long npsdbin;
    const double *psdfreqs;
    odet *det;
    /* intervals (periods) */
    long ninterval;
    int64_t *intervals, *ival_start;  /* samples */
    int64_t *noba_pp, *kstart_pp;     /* baselines */
    int64_t nbase, noba_pp_max;
    int64_t *bstart, *blen;
    double *btime;
    int16_t *subchunk;
    /* per (det, interval) weight used by the projections, 0 = skipped */
    double *dw_ival;                   /* [ndet][ninterval] */
    /* filter */
    long nof;
    int npsdtot;
    double *fcov;                      /* [npsdtot][nof/2+1] (real) */
    int *psddet;
    double *psdstart, *psdstop;
    offt_plan *plan;
} octx;

static int psd_index_det(const octx *c, int idet, double t)
{
    /* src/noise_routines.f90:211-236 */
    const odet *d = &c->det[idet];
    if (!d->has_psd && d->npsd == 1) return 0;
    int ipsd = 0;
    for (;;) {
        if (ipsd == d->npsd - 1) {
            if (d->psdstarts[ipsd] > t) ipsd = -1;
            return ipsd;
        }
        if (d->psdstarts[ipsd + 1] > t) return ipsd;
        ipsd++;
    }
}

static int psd_index(const octx *c, int idet, double t)
{
    /* src/noise_routines.f90:193-208 (global index over all detectors) */
    for (int i = 0; i < c->npsdtot; ++i)
        if (c->psddet[i] == idet && c->psdstart[i] <= t && c->psdstop[i] >= t) return i;
    return -1;
}

/* src/noise_routines.f90:130-190 */
void oracle_interpolate_psd(long n_in, const double *freq, const double *psd,
                            long n_out, const double *tf, double *tp)
{
    double *lf = (double *)malloc(sizeof(double) * n_in);
    double *lp = (double *)malloc(sizeof(double) * n_in);
    long startbin = 0;
    while (startbin < n_in && freq[startbin] == 0) startbin++;
    for (long i = 0; i < n_in; ++i) { lf[i] = log(freq[i]); lp[i] = log(psd[i]); }
    long j = startbin;
    for (long i = 0; i < n_out; ++i) {
        if (tf[i] == 0) { tp[i] = 0; continue; }
        double flog = log(tf[i]);
        if (j < n_in - 2) {
            while (lf[j + 1] < flog) {
                j++;
                if (j == n_in - 2) break;
            }
        }
        if (psd[j] > 0 && psd[j + 1] > 0) {
            double r = (flog - lf[j]) / (lf[j + 1] - lf[j]);
            tp[i] = exp(lp[j] * (1.0 - r) + lp[j + 1] * r);
        } else {
            double r = (tf[i] - freq[j]) / (freq[j + 1] - freq[j]);
            tp[i] = psd[j] * (1.0 - r) + psd[j + 1] * r;
        }
    }
    free(lf);
    free(lp);
}

/* ------------------------------------------------------------------------ */
/* small dense linear algebra: src/matrix.f90                                */
/* ------------------------------------------------------------------------ */

/* cyclic Jacobi eigen-decomposition of a symmetric n x n (n <= 3) matrix a
 * (row-major, full), eigenvalues ascending, eigenvectors in columns of v */
static int jacobi_eig(int n, double *a, double *ev, double *v)
{
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) v[i * n + j] = (i == j);
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0, diag = 0;
        for (int i = 0; i < n; ++i) {
            diag += a[i * n + i] * a[i * n + i];
            for (int j = i + 1; j < n; ++j) off += a[i * n + j] * a[i * n + j];
        }
        if (off == 0.0 || off < 1e-34 * diag) break;
        for (int pp = 0; pp < n; ++pp)
            for (int q = pp + 1; q < n; ++q) {
                double apq = a[pp * n + q];
                if (apq == 0.0) continue;
                double app = a[pp * n + pp], aqq = a[q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
                for (int k = 0; k < n; ++k) {
                    double akp = a[k * n + pp], akq = a[k * n + q];
                    a[k * n + pp] = cs * akp - sn * akq;
                    a[k * n + q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = a[pp * n + k], aqk = a[q * n + k];
                    a[pp * n + k] = cs * apk - sn * aqk;
                    a[q * n + k] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double vkp = v[k * n + pp], vkq = v[k * n + q];
                    v[k * n + pp] = cs * vkp - sn * vkq;
                    v[k * n + q] = sn * vkp + cs * vkq;
                }
            }
    }
    for (int i = 0; i < n; ++i) ev[i] = a[i * n + i];
    /* sort ascending (dsyev order) */
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (ev[j] < ev[i]) {
                double t = ev[i]; ev[i] = ev[j]; ev[j] = t;
                for (int k = 0; k < n; ++k) { t = v[k * n + i]; v[k * n + i] = v[k * n + j]; v[k * n + j] = t; }
            }
    for (int i = 0; i < n; ++i) if (isnan(ev[i])) return 1;
    return 0;
}

/* cc: Fortran column-major n x n (cc(i,j) at [i + j*n]); upper triangle used.
 * src/matrix.f90:53-126 (invert_eig).  Returns sing. */
static int invert_eig_fm(double *cc, int n, int incomplete, double *cdet, double *rcond)
{
    int good[3], nf = 0;
    if (incomplete) {
        for (int i = 0; i < n; ++i) if (fabs(cc[i + i * n]) > F32(1e-30)) good[nf++] = i;
        if (nf == 0) { good[0] = 0; nf = 1; }   /* reference: "nfound = 1", good(1) unset */
    } else {
        for (int i = 0; i < n; ++i) good[i] = i;
        nf = n;
    }
    double a[9], v[9], ev[3];
    for (int i = 0; i < nf; ++i)
        for (int j = 0; j < nf; ++j) {
            int gi = good[i < j ? i : j], gj = good[i < j ? j : i];   /* upper triangle */
            a[i * nf + j] = cc[gi + gj * n];
        }
    int info = jacobi_eig(nf, a, ev, v);
    double det = 1, mn = ev[0], mx = ev[0];
    for (int i = 0; i < nf; ++i) { det *= ev[i]; if (ev[i] < mn) mn = ev[i]; if (ev[i] > mx) mx = ev[i]; }
    *cdet = det;
    *rcond = mn / mx;
    if (info != 0 || mx < F32(1e-30) || *rcond < F32(1e-12)) {
        for (int i = 0; i < n * n; ++i) cc[i] = 0;
        *cdet = 0;
        *rcond = 0;
        return 1;
    }
    for (int i = 0; i < nf; ++i)
        for (int j = 0; j < nf; ++j) {
            double s = 0;
            for (int k = 0; k < nf; ++k) s += (v[i * nf + k] / ev[k]) * v[j * nf + k];
            cc[good[i] + good[j] * n] = s;
        }
    return 0;
}

int oracle_invert_eig(double *cc, int n, double *cdet, double *rcond)
{
    return invert_eig_fm(cc, n, 0, cdet, rcond);
}

/* src/matrix.f90:132-179 (invert_eig_s); returns noeig */
static int invert_eig_s_fm(double *cc, int n, double plim, double *cdet, double *rcond)
{
    double a[9], v[9], ev[3];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) a[i * n + j] = cc[(i < j ? i : j) + (i < j ? j : i) * n];
    int info = jacobi_eig(n, a, ev, v);
    double det = 1, mn = ev[0], mx = ev[0];
    for (int i = 0; i < n; ++i) { det *= ev[i]; if (ev[i] < mn) mn = ev[i]; if (ev[i] > mx) mx = ev[i]; }
    *cdet = det;
    *rcond = mn / mx;
    if (info != 0 || mx < F32(1e-30)) {
        for (int i = 0; i < n * n; ++i) cc[i] = 0;
        *cdet = 0; *rcond = 0;
        return 0;
    }
    int noeig = 0;
    for (int i = 0; i < n; ++i) {
        if (ev[i] < fabs(plim * mx)) ev[i] = 0;
        else { ev[i] = 1.0 / ev[i]; noeig++; }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0;
            for (int k = 0; k < n; ++k) s += (v[i * n + k] * ev[k]) * v[j * n + k];
            cc[i + j * n] = s;
        }
    return noeig;
}

/* src/matrix.f90:287-327 (invert_LU via dpotrf/dpotri on the upper triangle) */
static int invert_lu_fm(double *cc, int n, double *cdet)
{
    double u[9] = {0};
    for (int j = 0; j < n; ++j) {
        double s = cc[j + j * n];
        for (int k = 0; k < j; ++k) s -= u[k + j * n] * u[k + j * n];
        if (!(s > 0)) return 1;
        double ujj = sqrt(s);
        u[j + j * n] = ujj;
        for (int i = j + 1; i < n; ++i) {
            double t = cc[j + i * n];
            for (int k = 0; k < j; ++k) t -= u[k + j * n] * u[k + i * n];
            u[j + i * n] = t / ujj;
        }
    }
    double det = 1;
    for (int i = 0; i < n; ++i) det *= u[i + i * n];
    *cdet = det;
    /* inverse = U^-1 U^-T */
    double ui[9] = {0};
    for (int j = 0; j < n; ++j) {
        ui[j + j * n] = 1.0 / u[j + j * n];
        for (int i = j - 1; i >= 0; --i) {
            double s = 0;
            for (int k = i + 1; k <= j; ++k) s += u[i + k * n] * ui[k + j * n];
            ui[i + j * n] = -s / u[i + i * n];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0;
            for (int k = (i > j ? i : j); k < n; ++k) s += ui[i + k * n] * ui[j + k * n];
            cc[i + j * n] = s;
        }
    return 0;
}

/* src/matrix.f90:15-51 */
static int test_decoupling(double *cc, int n, int allow)
{
    if (!allow) return 1;
    int istart = 2;
    for (int i = 0; i < n; ++i)
        if (cc[i + i * n] > F32(1e-30))
            for (int j = 1; j < n; ++j)
                if (fabs(cc[0 + j * n]) / cc[i + i * n] > F32(1e-10)) istart = 1;
    int any = 0;
    for (int j = 1; j < n; ++j) if (fabs(cc[j + 0 * n]) > F32(1e-10)) any = 1;
    istart = any ? 1 : 2;
    if (istart == 2 && cc[0] > F32(1e-30)) {
        cc[0] = 1.0 / cc[0];
        for (int j = 1; j < n; ++j) { cc[j] = 0; cc[j * n] = 0; }
    }
    return istart;
}

/* apply inversion to the (istart..n) sub-block in place */
static void sub_extract(const double *cc, int n, int s, double *out)
{
    int m = n - s;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) out[i + j * m] = cc[(i + s) + (j + s) * n];
}
static void sub_store(double *cc, int n, int s, const double *in)
{
    int m = n - s;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) cc[(i + s) + (j + s) * n] = in[i + j * m];
}

/* ------------------------------------------------------------------------ */
/* band Cholesky (DPBTRF / DPBTRS, 'L'): src/noise_routines.f90:1259-1301    */
/* band[j*nband + m] = A(j+m, j), m = 0..nband-1                             */
/* ------------------------------------------------------------------------ */

int oracle_band_cholesky(double *b, long n, int nband)
{
    int kd = nband - 1;
    for (long j = 0; j < n; ++j) {
        double ajj = b[j * nband];
        if (!(ajj > 0)) return (int)(j + 1);
        ajj = sqrt(ajj);
        b[j * nband] = ajj;
        long kn = kd < n - 1 - j ? kd : n - 1 - j;
        for (long m = 1; m <= kn; ++m) b[j * nband + m] /= ajj;
        /* trailing symmetric rank-1 update (dsyr 'L') */
        for (long c = 1; c <= kn; ++c) {
            double xc = b[j * nband + c];
            double *col = &b[(j + c) * nband];
            for (long r = c; r <= kn; ++r) col[r - c] -= b[j * nband + r] * xc;
        }
    }
    return 0;
}

void oracle_band_solve(const double *b, long n, int nband, double *z)
{
    int kd = nband - 1;
    /* L y = z */
    for (long j = 0; j < n; ++j) {
        z[j] /= b[j * nband];
        double zj = z[j];
        long kn = kd < n - 1 - j ? kd : n - 1 - j;
        for (long m = 1; m <= kn; ++m) z[j + m] -= zj * b[j * nband + m];
    }
    /* L^T x = y */
    for (long j = n - 1; j >= 0; --j) {
        double s = z[j];
        long kn = kd < n - 1 - j ? kd : n - 1 - j;
        for (long m = 1; m <= kn; ++m) s -= b[j * nband + m] * z[j + m];
        z[j] = s / b[j * nband];
    }
}

/* ------------------------------------------------------------------------ */
/* noise filter: src/noise_routines.f90:242-292 (init_filter),              */
/* 367-739 (build_filter, get_spectrum_interp)                              */
/* ------------------------------------------------------------------------ */

static void get_spectrum_interp(const oprm *p, long npsdbin, const double *psdfreqs,
                                const double *psd, double plateau_in, long nof,
                                const double *f, double *spectrum, double *sigma_out)
{
    oracle_interpolate_psd(npsdbin, psdfreqs, psd, nof, f, spectrum);
    double plateau;
    if (p->noise_weights_from_psd) {
        plateau = plateau_in;
    } else {
        if (p->radiometers) {
            double fr = p->fsample / 2, d;
            oracle_interpolate_psd(npsdbin, psdfreqs, psd, 1, &fr, &d);
            plateau = d;
        } else {
            double fr[1000], d[1000];
            for (int i = 0; i < 1000; ++i) fr[i] = 1 + (double)(i * 10) / 1000;
            oracle_interpolate_psd(npsdbin, psdfreqs, psd, 1000, fr, d);
            plateau = d[0];
            for (int i = 1; i < 1000; ++i) if (d[i] < plateau) plateau = d[i];
        }
        if (sigma_out) *sigma_out = sqrt(plateau * p->fsample);
    }
    double margin = p->radiometers ? F32(1e-6) : F32(1e-3);
    for (long i = 0; i < nof; ++i) {
        if (spectrum[i] > plateau * (1 + margin)) spectrum[i] = spectrum[i] - plateau;
        else spectrum[i] = spectrum[i] * margin;
    }
    for (long i = 0; i < nof; ++i) {
        if (spectrum[i] <= 0) {
            int done = 0;
            for (long j = i + 1; j < nof; ++j)
                if (spectrum[j] > 0) { spectrum[i] = spectrum[j]; done = 1; break; }
            if (!done)
                for (long j = i - 1; j >= 0; --j)
                    if (spectrum[j] > 0) { spectrum[i] = spectrum[j]; break; }
        }
    }
    if (f[0] == 0) spectrum[0] = 0;
}

static void build_filter_one(const oprm *p, long nof, double sigma, long npsdbin,
                             const double *psdfreqs, const double *psd, double plateau,
                             double *fcov, double *sigma_out)
{
    double fa = p->fsample / p->nshort;
    long nh = nof / 2 + 1;
    double *aspec = (double *)calloc(nh, sizeof(double));
    double *f = (double *)malloc(sizeof(double) * nof);
    double *g = (double *)malloc(sizeof(double) * nof);
    double *sp = (double *)malloc(sizeof(double) * nof);
    for (int k = 0; k <= 2; ++k) {
        for (long i = 0; i < nof; ++i) f[i] = k * fa + ((double)i * fa) / nof;
        get_spectrum_interp(p, npsdbin, psdfreqs, psd, plateau, nof, f, sp, k == 0 ? sigma_out : NULL);
        for (long i = 0; i < nof; ++i) {
            double x = PI_D * f[i] / fa;
            if (x < F32(1e-30)) g[i] = 1;
            else { double s = sin(x); g[i] = (s * s) / (x * x); }
        }
        if (k == 0) aspec[0] = sp[0];
        else aspec[0] = aspec[0] + 2 * sp[0] * g[0];
        for (long i = 1; i < nh; ++i) aspec[i] = aspec[i] + sp[i] * g[i] + sp[nof - i] * g[nof - i];
    }
    for (long i = 0; i < nh; ++i) {
        if (fabs(aspec[i]) > sigma * F32(1e-30)) aspec[i] = 1.0 / (fa * aspec[i]);
        else aspec[i] = 0;
    }
    aspec[0] = 0;
    memcpy(fcov, aspec, sizeof(double) * nh);
    free(aspec); free(f); free(g); free(sp);
}

int oracle_build_filter(double fsample, int nshort, long nof, int radiometers, double sigma,
                        long npsdbin, const double *psdfreqs, const double *psd, double *fcov_out)
{
    oprm p;
    set_defaults(&p);
    p.fsample = fsample;
    p.nshort = nshort;
    p.radiometers = radiometers;
    build_filter_one(&p, nof, sigma, npsdbin, psdfreqs, psd, 0.0, fcov_out, NULL);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* trim_interval: src/noise_routines.f90:764-813                            */
/* ------------------------------------------------------------------------ */

static int64_t trim_interval(const octx *c, const double *nna_det, int64_t *kstart, int64_t *kstop)
{
    int64_t ks = *kstart, ke = *kstop;
    int any = 0;
    for (int64_t k = ks; k < ke; ++k) if (nna_det[k] != 0) { any = 1; break; }
    if (!any) return 0;
    while (ks < ke) { if (nna_det[ks] != 0) break; ks++; }
    while (ke > ks) { if (nna_det[ke - 1] != 0) break; ke--; }
    int64_t noba = ke - ks;
    int64_t gapmin = (int64_t)(60 * c->p.fsample / c->p.dnshort + 1);
    int64_t gaplen = 0;
    for (int64_t i = ks; i < ke; ++i) {
        if (nna_det[i] != 0) gaplen = 0;
        else {
            gaplen++;
            if (gaplen == gapmin) {
                noba = (i + 1) - gaplen - ks;
                *kstart = ks; *kstop = ke;
                return noba;
            }
        }
    }
    *kstart = ks; *kstop = ke;
    return noba;
}

/* ------------------------------------------------------------------------ */
/* preconditioner: src/noise_routines.f90:920-1153, 1159-1409               */
/* ------------------------------------------------------------------------ */

typedef struct {
    int64_t idet, kstart, noba;
    int ipsd;          /* global PSD index or -1 */
    int nband;         /* 0 = CG fallback */
    double *band;      /* [noba][nband] */
    int skip;          /* no PSD / zero weight: z = r */
} oseg;

typedef struct {
    int kind;          /* 0 none, 1 diagonal, 2 fprecond, 3 band */
    double *prec_diag; /* [ndet][nbase] */
    double *fprecond;  /* [npsdtot][nh] */
    double *invcov;    /* [npsdtot][nof] */
    oseg *seg;
    long nseg;
    int nband_max, nchol, ncg;
} oprec;

static void convolve_segment(const octx *c, const double *fc, const double *x, double *y,
                             int64_t noba, double *xx, double *fx, double *work)
{
    long nof = c->nof, nh = nof / 2 + 1;
    double s = 0;
    for (int64_t i = 0; i < noba; ++i) s += x[i];
    s = s / noba;
    for (int64_t i = 0; i < noba; ++i) xx[i] = x[i] - s;
    for (long i = noba; i < nof; ++i) xx[i] = 0;
    offt_r2c(c->plan, xx, fx, work);
    for (long k = 0; k < nh; ++k) { fx[2 * k] *= fc[k]; fx[2 * k + 1] *= fc[k]; }
    offt_c2r(c->plan, fx, xx, work);
    double inv = 1 / (double)nof;
    for (long i = 0; i < nof; ++i) xx[i] = xx[i] * inv;
    s = 0;
    for (int64_t i = 0; i < noba; ++i) s += xx[i];
    s = s / noba;
    for (int64_t i = 0; i < noba; ++i) y[i] = xx[i] - s;
}

/* one trimmed segment's convolution on its own, exported for the tests */
void oracle_convolve(long nof, const double *fcov, const double *x, long noba, double *y)
{
    octx c;
    memset(&c, 0, sizeof c);
    c.nof = nof;
    c.plan = offt_create(nof);
    double *xx = (double *)malloc(sizeof(double) * nof);
    double *fx = (double *)malloc(sizeof(double) * 2 * (nof / 2 + 1));
    double *work = (double *)malloc(sizeof(double) * nof);
    convolve_segment(&c, fcov, x, y, noba, xx, fx, work);
    free(xx); free(fx); free(work);
    offt_destroy(c.plan);
}

/* convolve_pp: y = x, then per (det, interval) segments; src/noise_routines.f90:816-913 */
static void convolve_pp(const octx *c, const double *fcall, const double *nna, const double *x, double *y)
{
    int64_t nb = c->nbase;
    memcpy(y, x, sizeof(double) * nb * c->ndet);
    long ntask = c->ndet * c->ninterval;
    long nof = c->nof, nh = nof / 2 + 1;
#pragma omp parallel
    {
        double *xx = (double *)malloc(sizeof(double) * nof);
        double *fx = (double *)malloc(sizeof(double) * 2 * nh);
        double *work = (double *)malloc(sizeof(double) * nof);
#pragma omp for schedule(static, 1)
        for (long t = 0; t < ntask; ++t) {
            int idet = (int)(t / c->ninterval);
            long ival = t % c->ninterval;
            int64_t kstart = c->kstart_pp[ival], kstop = kstart + c->noba_pp[ival];
            int ipsd = psd_index(c, idet, c->btime[kstart]);
            if (ipsd == -1) continue;
            const double *fc = fcall + (size_t)ipsd * nh;
            for (;;) {
                int64_t noba = trim_interval(c, nna + (size_t)idet * nb, &kstart, &kstop);
                if (noba == 0) break;
                convolve_segment(c, fc, x + (size_t)idet * nb + kstart, y + (size_t)idet * nb + kstart,
                                 noba, xx, fx, work);
                kstart += noba;
            }
        }
        free(xx); free(fx); free(work);
    }
}

/* apply_CG_preconditioner: src/noise_routines.f90:1305-1409 */
static void cg_precond(const octx *c, const double *fc, const double *nna, const double *invcov,
                       int64_t noba, double *x, const double *b, double *xx, double *fx, double *work)
{
    double *resid = (double *)malloc(sizeof(double) * noba);
    double *prop = (double *)malloc(sizeof(double) * noba);
    double *aprop = (double *)malloc(sizeof(double) * noba);
    double *zres = (double *)malloc(sizeof(double) * noba);
    double *prec = (double *)malloc(sizeof(double) * noba);
    for (int64_t i = 0; i < noba; ++i) { x[i] = 0; resid[i] = b[i] - x[i]; }
    double rr = 0;
    for (int64_t i = 0; i < noba; ++i) rr += resid[i] * resid[i];
    if (rr == 0) goto done;
    double rr0 = rr;
    for (int64_t i = 0; i < noba; ++i) { prec[i] = nna[i] + invcov[0]; if (prec[i] != 0) prec[i] = 1 / prec[i]; }
    for (int64_t i = 0; i < noba; ++i) zres[i] = prec[i] * resid[i];
    double rz = 0;
    for (int64_t i = 0; i < noba; ++i) rz += resid[i] * zres[i];
    memcpy(prop, zres, sizeof(double) * noba);
    int iter = 1;
    for (;;) {
        convolve_segment(c, fc, prop, aprop, noba, xx, fx, work);
        for (int64_t i = 0; i < noba; ++i) aprop[i] = aprop[i] + nna[i] * prop[i];
        double anorm = 0;
        for (int64_t i = 0; i < noba; ++i) anorm += prop[i] * aprop[i];
        double alpha = rz / anorm;
        for (int64_t i = 0; i < noba; ++i) x[i] = x[i] + alpha * prop[i];
        for (int64_t i = 0; i < noba; ++i) resid[i] = resid[i] - alpha * aprop[i];
        rr = 0;
        for (int64_t i = 0; i < noba; ++i) rr += resid[i] * resid[i];
        if (rr / rr0 < F32(1e-6) || iter == 100) break;
        for (int64_t i = 0; i < noba; ++i) zres[i] = prec[i] * resid[i];
        double rz_old = rz;
        rz = 0;
        for (int64_t i = 0; i < noba; ++i) rz += resid[i] * zres[i];
        double beta = rz / rz_old;
        for (int64_t i = 0; i < noba; ++i) prop[i] = zres[i] + beta * prop[i];
        iter++;
    }
done:
    free(resid); free(prop); free(aprop); free(zres); free(prec);
}

static int construct_preconditioner(octx *c, const double *nna, oprec *pr, char *msg)
{
    const oprm *p = &c->p;
    memset(pr, 0, sizeof(*pr));
    int64_t nb = c->nbase;
    if (p->precond_width_max < 1) { pr->kind = 0; return 0; }
    if (!p->kfilter || p->precond_width_max == 1) {
        pr->kind = 1;
        pr->prec_diag = (double *)calloc((size_t)nb * c->ndet, sizeof(double));
        for (int idet = 0; idet < c->ndet; ++idet)
            for (long iv = 0; iv < c->ninterval; ++iv) {
                int64_t ks = c->kstart_pp[iv], no = c->noba_pp[iv];
                int ip = psd_index_det(c, idet, c->btime[ks]);
                for (int64_t k = ks; k < ks + no; ++k) {
                    double *pd = &pr->prec_diag[(size_t)idet * nb + k];
                    if (ip < 0) *pd = 1.;
                    else if (nna[(size_t)idet * nb + k] == 0) *pd = 1. / c->det[idet].weights[ip];
                    else *pd = 1. / nna[(size_t)idet * nb + k];
                }
            }
        return 0;
    }
    long nh = c->nof / 2 + 1;
    pr->fprecond = (double *)calloc((size_t)c->npsdtot * nh, sizeof(double));
    {
        int ipsd = 0;
        for (int idet = 0; idet < c->ndet; ++idet)
            for (int j = 0; j < c->det[idet].npsd; ++j, ++ipsd)
                for (long k = 0; k < nh; ++k)
                    pr->fprecond[(size_t)ipsd * nh + k] =
                        1.0 / (p->nshort * c->det[idet].weights[j] + c->fcov[(size_t)ipsd * nh + k]);
    }
    if (p->use_fprecond) { pr->kind = 2; return 0; }
    pr->kind = 3;
    long nof = c->nof;
    pr->invcov = (double *)malloc(sizeof(double) * nof * c->npsdtot);
    {
        double *work = (double *)malloc(sizeof(double) * nof);
        for (int ip = 0; ip < c->npsdtot; ++ip) {
            double *fx = (double *)malloc(sizeof(double) * 2 * nh);
            for (long k = 0; k < nh; ++k) { fx[2 * k] = c->fcov[(size_t)ip * nh + k]; fx[2 * k + 1] = 0; }
            double *out = pr->invcov + (size_t)ip * nof;
            offt_c2r(c->plan, fx, out, work);
            double inv = 1 / (double)nof;
            for (long i = 0; i < nof; ++i) out[i] *= inv;
            free(fx);
        }
        free(work);
    }
    /* enumerate segments */
    long cap = 16, ns = 0;
    oseg *seg = (oseg *)malloc(sizeof(oseg) * cap);
    for (int idet = 0; idet < c->ndet; ++idet)
        for (long iv = 0; iv < c->ninterval; ++iv) {
            int64_t ks = c->kstart_pp[iv], ke = ks + c->noba_pp[iv];
            for (;;) {
                int64_t noba = trim_interval(c, nna + (size_t)idet * nb, &ks, &ke);
                if (noba == 0) break;
                if (ns == cap) { cap *= 2; seg = (oseg *)realloc(seg, sizeof(oseg) * cap); }
                oseg *s = &seg[ns++];
                memset(s, 0, sizeof(*s));
                s->idet = idet; s->kstart = ks; s->noba = noba;
                int ipd = psd_index_det(c, idet, c->btime[ks]);
                s->skip = (ipd < 0 || c->det[idet].weights[ipd] == 0);
                s->ipsd = psd_index(c, idet, c->btime[ks]);
                if (s->ipsd < 0) s->skip = 1;
                ks += noba;
            }
        }
    pr->seg = seg;
    pr->nseg = ns;
    if (p->use_cgprecond) { pr->ncg = (int)ns; return 0; }
    const double plim = F32(1 - 1e-5f);
    int nchol = 0, ncg = 0, nbmax = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : nchol, ncg) reduction(max : nbmax)
    for (long is = 0; is < ns; ++is) {
        oseg *s = &seg[is];
        if (s->skip) continue;
        const double *ic = pr->invcov + (size_t)s->ipsd * nof;
        const double *nn = nna + (size_t)s->idet * nb + s->kstart;
        double ptot = 0;
        for (long i = 0; i < nof / 2; ++i) ptot += ic[i] * ic[i];
        int tr = 1, ierr = -1, nband = 0;
        for (;;) {
            nband = (int)(s->noba < tr * p->precond_width_min ? s->noba : tr * p->precond_width_min);
            if (nband > p->precond_width_max) nband = p->precond_width_max;
            double pp = 0;
            for (int i = 0; i < nband; ++i) pp += ic[i] * ic[i];
            if (pp > plim * ptot) { ierr = 0; break; }
            if (nband == s->noba || nband == p->precond_width_max) break;
            if (tr == 1000) break;
            tr++;
        }
        if (ierr == 0) {
            for (;;) {
                double *b = (double *)malloc(sizeof(double) * (size_t)s->noba * nband);
                for (int64_t j = 0; j < s->noba; ++j)
                    for (int m = 0; m < nband; ++m) b[j * nband + m] = ic[m];
                for (int64_t j = 0; j < s->noba; ++j) b[j * nband] += nn[j];
                ierr = oracle_band_cholesky(b, s->noba, nband);
                if (ierr == 0) { s->band = b; s->nband = nband; break; }
                free(b);
                if (tr == 1000) break;
                if (nband == s->noba || nband == p->precond_width_max) break;
                tr++;
                nband = nband + p->precond_width_min < p->precond_width_max ? nband + p->precond_width_min
                                                                           : p->precond_width_max;
            }
        }
        if (ierr != 0) ncg++;
        else { nchol++; if (nband > nbmax) nbmax = nband; }
    }
    pr->nchol = nchol;
    pr->ncg = ncg;
    pr->nband_max = nbmax;
    (void)msg;
    return 0;
}

static void free_prec(oprec *pr)
{
    free(pr->prec_diag);
    free(pr->fprecond);
    free(pr->invcov);
    for (long i = 0; i < pr->nseg; ++i) free(pr->seg[i].band);
    free(pr->seg);
}

/* preconditioning_band: src/noise_routines.f90:1159-1256 */
static void apply_preconditioner(const octx *c, const oprec *pr, const double *nna, const double *r, double *z)
{
    int64_t nb = c->nbase;
    size_t n = (size_t)nb * c->ndet;
    memcpy(z, r, sizeof(double) * n);
    if (pr->kind == 0) return;
    if (pr->kind == 1) {
        for (size_t i = 0; i < n; ++i) z[i] = r[i] * pr->prec_diag[i];
        return;
    }
    if (pr->kind == 2) { convolve_pp(c, pr->fprecond, nna, r, z); return; }
    long nof = c->nof, nh = nof / 2 + 1;
#pragma omp parallel
    {
        double *xx = (double *)malloc(sizeof(double) * nof);
        double *fx = (double *)malloc(sizeof(double) * 2 * nh);
        double *work = (double *)malloc(sizeof(double) * nof);
#pragma omp for schedule(dynamic, 1)
        for (long is = 0; is < pr->nseg; ++is) {
            const oseg *s = &pr->seg[is];
            if (s->skip) continue;
            double *zz = z + (size_t)s->idet * nb + s->kstart;
            const double *rr = r + (size_t)s->idet * nb + s->kstart;
            if (s->band) {
                memcpy(zz, rr, sizeof(double) * s->noba);
                oracle_band_solve(s->band, s->noba, s->nband, zz);
            } else {
                cg_precond(c, c->fcov + (size_t)s->ipsd * nh, nna + (size_t)s->idet * nb + s->kstart,
                           pr->invcov + (size_t)s->ipsd * nof, s->noba, zz, rr, xx, fx, work);
            }
        }
        free(xx); free(fx); free(work);
    }
}

/* ------------------------------------------------------------------------ */
/* main driver                                                              */
/* ------------------------------------------------------------------------ */

static inline int sample_active(const octx *c, int pass, int64_t i)
{
    return pass == 0 || c->subchunk[i] == pass;
}

static void ccmultiply(int nmap, const double *cc, double *m, int64_t npix)
{
    /* src/map_routines.f90:24-77, cc stored [ip][b][a] = cc(a+1, b+1) */
    if (nmap == 1) {
        for (int64_t ip = 0; ip < npix; ++ip) {
            if (cc[ip] == 0) m[ip] = 0;
            else m[ip] = cc[ip] * m[ip];
        }
        return;
    }
#pragma omp parallel for schedule(static)
    for (int64_t ip = 0; ip < npix; ++ip) {
        const double *C = cc + ip * 9;
        double *M = m + ip * 3;
        if (C[0] == 0) { M[0] = M[1] = M[2] = 0; continue; }
        /* C[a + 3b] = cc(a+1, b+1) */
        double tm = C[0] * M[0] + C[1] * M[1] + C[2] * M[2];
        double qm = C[1] * M[0] + C[4] * M[1] + C[5] * M[2];
        double um = C[2] * M[0] + C[5] * M[1] + C[8] * M[2];
        M[0] = tm; M[1] = qm; M[2] = um;
    }
}

oracle_result *oracle_destripe(int pass, const char *parstring, long ndet, const char *detstring,
                               const double *detweights, long nsamp, long nnz, const double *timestamps,
                               int32_t *pix, const float *pixweights, double *signal, long nperiod,
                               const long *periods, const long *npsd, long npsdtot, const double *psdstarts,
                               long npsdbin, const double *psdfreqs, long npsdval, const double *psdvals)
{
    double t_start = now_s();
    oracle_result *R = (oracle_result *)calloc(1, sizeof(oracle_result));
    octx C;
    memset(&C, 0, sizeof C);
    octx *c = &C;
    oprm *p = &c->p;
    char *msg = R->msg;
    oprec pr;
    memset(&pr, 0, sizeof pr);
    double *loc = NULL, *loccc = NULL, *cca = NULL, *cc = NULL, *wamap = NULL, *map = NULL, *binmap = NULL;
    int32_t *lochits = NULL;
    double *aa = NULL, *yba = NULL, *nna = NULL;
    double *r = NULL, *pv = NULL, *z = NULL, *ap = NULL, *ro = NULL, *lm = NULL, *wam = NULL;
    char *rmask = NULL;
    (void)detstring;
    (void)npsdval;

    set_defaults(p);
    if (read_parameters(p, parstring, msg)) goto error;
    /* src/smadam.F90:138-144 */
    if (p->mcmode && p->nsubchunk > 1) { fail(msg, "MCMode is not compatible with nsubchunk > 1."); goto error; }
    if (p->temperature_only && nnz != 1) { fail(msg, "temperature_only=T but pointing weights are polarized"); goto error; }
    if (ndet > 10000) { fail(msg, "ndet > NDETMAX"); goto error; }
    if (nnz != 1 && nnz != 3) { fail(msg, "nnz=%ld not supported (nmap must be 1 or 3)", nnz); goto error; }
    if (p->has_inmask || p->has_spectrum || p->has_covmat) {
        fail(msg, "file_inmask / file_spectrum / file_covmat are not supported"); goto error;
    }
    c->ndet = ndet; c->nsamp = nsamp; c->nnz = nnz; c->nmap = (int)nnz;
    c->npsdbin = npsdbin; c->psdfreqs = psdfreqs;

    /* read_detectors: src/inputparam.f90:73-163 */
    c->det = (odet *)calloc(ndet, sizeof(odet));
    {
        long ipsd = 0;
        for (long id = 0; id < ndet; ++id) {
            odet *d = &c->det[id];
            int nint = (int)npsd[id];
            d->npsd = nint == 0 ? 1 : nint;
            d->weights = (double *)malloc(sizeof(double) * d->npsd);
            d->sigmas = (double *)malloc(sizeof(double) * d->npsd);
            d->psdstarts = (double *)calloc(d->npsd, sizeof(double));
            d->plateaus = (double *)malloc(sizeof(double) * d->npsd);
            double w = detweights[id];
            for (int j = 0; j < d->npsd; ++j) {
                d->weights[j] = w;
                d->sigmas[j] = w > 0 ? 1.0 / sqrt(w) : 0.0;
                d->plateaus[j] = d->sigmas[j] * d->sigmas[j] / p->fsample;
            }
            if (nint > 0) {
                d->has_psd = 1;
                d->psds = (double *)malloc(sizeof(double) * nint * npsdbin);
                for (int j = 0; j < nint; ++j) {
                    d->psdstarts[j] = psdstarts[ipsd];
                    memcpy(d->psds + (size_t)j * npsdbin, psdvals + (size_t)ipsd * npsdbin, sizeof(double) * npsdbin);
                    ipsd++;
                }
            }
        }
    }

    /* pointing range check: src/smadam.F90:158-171 */
    {
        int nside_max = p->nside_map > p->nside_cross ? p->nside_map : p->nside_cross;
        int64_t pmin = INT64_MAX, pmax = INT64_MIN;
        for (long i = 0; i < ndet * nsamp; ++i) { if (pix[i] < pmin) pmin = pix[i]; if (pix[i] > pmax) pmax = pix[i]; }
        if (ndet * nsamp > 0) {
            if (pmin < -1) { fail(msg, "Too small pixel number %lld", (long long)pmin); goto error; }
            if (pmax > 12LL * nside_max * nside_max - 1) { fail(msg, "Too large pixel number %lld", (long long)pmax); goto error; }
        }
    }

    /* check_files: src/read_data.f90:203-323 */
    c->ninterval = nperiod;
    c->intervals = (int64_t *)malloc(sizeof(int64_t) * nperiod);
    for (long ip = 0; ip < nperiod; ++ip) {
        int64_t len = ip < nperiod - 1 ? periods[ip + 1] - periods[ip] : nsamp - periods[ip];
        c->intervals[ip] = len;
    }
    for (long id = 0; id < ndet; ++id)
        for (long i = 0; i < nsamp; ++i) {
            const float *w = pixweights + ((size_t)id * nsamp + i) * nnz;
            int allz = 1;
            for (long j = 0; j < nnz; ++j) if (w[j] != 0) { allz = 0; break; }
            if (allz && !(pix[(size_t)id * nsamp + i] < 0)) pix[(size_t)id * nsamp + i] = -1;
        }
    {
        int64_t istart = 0;
        for (long ip = 0; ip < nperiod; ++ip) {
            int64_t istop = istart + c->intervals[ip];
            for (long id = 0; id < ndet; ++id) {
                int32_t *pp = pix + (size_t)id * nsamp;
                int64_t ngood = 0;
                for (int64_t i = istart; i < istop; ++i) if (pp[i] >= 0) ngood++;
                if (ngood == 0) continue;
                double frac = (double)ngood / c->intervals[ip];
                if (frac < F32(1e-3)) for (int64_t i = istart; i < istop; ++i) pp[i] = -1;
            }
            istart = istop;
        }
    }

    /* init_parameters: src/parameter_control.f90:27-409 */
    if (!p->radiometers) {
        for (long id = 0; id < ndet; ++id) {
            odet *d = &c->det[id];
            if (!d->has_psd) continue;
            for (int j = 0; j < d->npsd; ++j) {
                double *psd = d->psds + (size_t)j * npsdbin;
                long im = 0;
                for (long i = 1; i < npsdbin; ++i) if (psd[i] > psd[im]) im = i;
                for (long i = 0; i < im; ++i) psd[i] = psd[im];
            }
        }
    }
    if (p->noise_weights_from_psd) {
        /* measure_noise_weights: src/noise_routines.f90:65-127 */
        for (long id = 0; id < ndet; ++id) {
            odet *d = &c->det[id];
            if (!d->has_psd) { fail(msg, "noise_weights_from_psd needs PSDs for every detector"); goto error; }
            for (int j = 0; j < d->npsd; ++j) {
                const double *psd = d->psds + (size_t)j * npsdbin;
                if (p->radiometers) {
                    double fr = p->fsample / 2, pl;
                    oracle_interpolate_psd(npsdbin, psdfreqs, psd, 1, &fr, &pl);
                    d->sigmas[j] = sqrt(pl * p->fsample);
                    d->plateaus[j] = pl;
                } else {
                    int nbin = 10000;
                    double fstep = p->fsample * .5 / nbin;
                    double *fr = (double *)malloc(sizeof(double) * nbin), *dd = (double *)malloc(sizeof(double) * nbin);
                    for (int i = 0; i < nbin; ++i) fr[i] = fstep * (i + 1);
                    double fbase = 1 / (p->dnshort / p->fsample);
                    int ibase = (int)(fbase / fstep), ilower = (int)(1 / fstep), iupper = (int)(10 / fstep);
                    oracle_interpolate_psd(npsdbin, psdfreqs, psd, nbin, fr, dd);
                    double pl = dd[ilower - 1];
                    for (int i = ilower; i < iupper && i < nbin; ++i) if (dd[i] < pl) pl = dd[i];
                    double rms = ibase * pl;
                    for (int i = ibase; i < nbin; ++i) rms = rms + dd[i];
                    rms = sqrt(rms / nbin * p->fsample);
                    d->sigmas[j] = rms;
                    d->plateaus[j] = pl;
                    d->weights[j] = 1 / (pl * p->fsample);
                    free(fr); free(dd);
                }
            }
        }
    }
    if (p->nsubchunk < 2) { p->nsubchunk = 0; p->isubchunk = 0; }
    if (p->isubchunk > p->nsubchunk) { fail(msg, "ERROR: isubchunk > nsubchunk"); goto error; }
    if (pass < 0 || pass > p->nsubchunk) { fail(msg, "subchunk pass %d out of range", pass); goto error; }
    if (p->good_baseline_fraction < 0 || p->good_baseline_fraction > 1) {
        fail(msg, "ERROR: good_baseline_fraction must be in 0..1"); goto error;
    }
    if (p->nside_cross < 0) p->nside_cross = p->nside_map;
    {
        int ns[3] = {p->nside_map, p->nside_submap, p->nside_cross};
        for (int i = 0; i < 3; ++i)
            if (ns[i] < 1 || (ns[i] & (ns[i] - 1))) { fail(msg, "nside must be a power of two"); goto error; }
    }
    if (p->nside_map < p->nside_cross) { fail(msg, "nside_cross > nside_map not supported by Madam/TOAST"); goto error; }
    if (p->pixlim_cross < 0) p->pixlim_cross = p->pixlim_map;
    if (p->pixmode_map < 0 || p->pixmode_map > 2) { fail(msg, "Pixmode_map must be in range [0-2]"); goto error; }
    if (p->pixmode_cross < 0 || p->pixmode_cross > 4) { fail(msg, "Pixmode_cross must be in range [0-4]"); goto error; }
    if (p->basis_order != 0) { fail(msg, "basis_order > 0 is not supported"); goto error; }
    if (c->nmap == 1) p->temperature_only = 1;
    for (long id = 0; id < ndet; ++id) {
        odet *d = &c->det[id];
        for (int j = 0; j < d->npsd; ++j) {
            if (p->mode_detweight == 0) d->weights[j] = d->sigmas[j] == 0 ? 0 : 1.0 / d->plateaus[j] / p->fsample;
            else if (p->mode_detweight == 1) d->weights[j] = 1;
            else { fail(msg, "mode_detweight=%d not supported", p->mode_detweight); goto error; }
        }
    }
    c->nside_max = p->nside_map > p->nside_cross ? p->nside_map : p->nside_cross;
    p->nshort = (int)(p->dnshort + .5);
    if (!p->kfirst) {
        p->dnshort = 1e6;
        p->nshort = (int)(p->dnshort + .5);
        p->precond_width_min = 0;
        p->precond_width_max = 0;
        p->kfilter = 0;
    } else if (p->precond_width_max < p->precond_width_min) p->precond_width_max = p->precond_width_min;
    c->noba_pp = (int64_t *)malloc(sizeof(int64_t) * nperiod);
    c->kstart_pp = (int64_t *)malloc(sizeof(int64_t) * nperiod);
    c->ival_start = (int64_t *)malloc(sizeof(int64_t) * nperiod);
    c->nbase = 0;
    c->noba_pp_max = 0;
    for (long ip = 0; ip < nperiod; ++ip) {
        c->noba_pp[ip] = (int64_t)((c->intervals[ip] - 1) / p->dnshort + 1);
        c->kstart_pp[ip] = c->nbase;
        c->nbase += c->noba_pp[ip];
        if (c->noba_pp[ip] > c->noba_pp_max) c->noba_pp_max = c->noba_pp[ip];
    }

    /* start_timeloop: src/parameter_control.f90:752-950; baseline_times 956-977 */
    c->bstart = (int64_t *)malloc(sizeof(int64_t) * (c->nbase + 1));
    c->blen = (int64_t *)malloc(sizeof(int64_t) * (c->nbase + 1));
    c->btime = (double *)malloc(sizeof(double) * (c->nbase + 1));
    {
        int64_t m = 0, s0 = 0;
        for (long ip = 0; ip < nperiod; ++ip) {
            c->ival_start[ip] = s0;
            double dn = 0.0;
            int64_t n = 0;
            for (int64_t k = 0; k < c->noba_pp[ip] - 1; ++k) {
                int64_t n0 = n;
                dn = dn + p->dnshort;
                n = (int64_t)(dn + .5);
                c->blen[m++] = n - n0;
            }
            c->blen[m++] = c->intervals[ip] - n;
            if (p->kfirst && (c->blen[m - 1] < 0 || c->blen[m - 1] > (int64_t)p->dnshort + 1)) {
                fail(msg, "ERROR in start_timeloop: Lengths do not match."); goto error;
            }
            s0 += c->intervals[ip];
        }
        int64_t st = 0;
        for (int64_t k = 0; k < c->nbase; ++k) {
            c->bstart[k] = st;
            st += c->blen[k];
            c->btime[k] = (c->bstart[k] < nsamp && c->bstart[k] >= 0) ? timestamps[c->bstart[k]] : 0.0;
        }
    }
    c->subchunk = (int16_t *)calloc(nsamp > 0 ? nsamp : 1, sizeof(int16_t));
    if (p->nsubchunk > 0) {
        int64_t off = 0;
        for (long ip = 0; ip < nperiod; ++ip) {
            int64_t n = c->intervals[ip], sublen = n / p->nsubchunk;
            for (int isub = 1; isub <= p->nsubchunk; ++isub) {
                int64_t s = off + sublen * (isub - 1), e = s + sublen - 1;
                if (isub == p->nsubchunk) e = off + n - 1;
                for (int64_t i = s; i <= e && i < nsamp; ++i) if (i >= 0) c->subchunk[i] = (int16_t)isub;
            }
            off += n;
        }
    }

    /* per (det, interval) detector weight: 0 when there is no PSD or zero weight */
    c->dw_ival = (double *)calloc((size_t)ndet * nperiod, sizeof(double));
    for (long id = 0; id < ndet; ++id)
        for (long ip = 0; ip < nperiod; ++ip) {
            int ipsd = c->noba_pp[ip] > 0 ? psd_index_det(c, (int)id, c->btime[c->kstart_pp[ip]]) : -1;
            c->dw_ival[id * nperiod + ip] = ipsd < 0 ? 0.0 : c->det[id].weights[ipsd];
        }

    /* filter: init_filter + build_filter */
    if (p->kfilter) {
        int64_t nof_min = 2 * c->noba_pp_max;
        if (p->precond_width_max > nof_min - 1) nof_min = p->precond_width_max + 1;
        long nof = 1;
        while (nof < nof_min) nof *= 2;
        if (nof < 4) nof = 4;
        c->nof = nof;
        c->plan = offt_create(nof);
        long nh = nof / 2 + 1;
        c->npsdtot = 0;
        for (long id = 0; id < ndet; ++id) {
            if (!c->det[id].has_psd) { fail(msg, "All detectors must have at least one PSD to build the filter"); goto error; }
            c->npsdtot += c->det[id].npsd;
        }
        c->fcov = (double *)malloc(sizeof(double) * nh * c->npsdtot);
        c->psddet = (int *)malloc(sizeof(int) * c->npsdtot);
        c->psdstart = (double *)malloc(sizeof(double) * c->npsdtot);
        c->psdstop = (double *)malloc(sizeof(double) * c->npsdtot);
        int ipt = 0;
        for (long id = 0; id < ndet; ++id) {
            odet *d = &c->det[id];
            for (int j = 0; j < d->npsd; ++j, ++ipt) {
                c->psddet[ipt] = (int)id;
                c->psdstart[ipt] = d->psdstarts[j];
                c->psdstop[ipt] = j < d->npsd - 1 ? d->psdstarts[j + 1] : F32(1e30);
                double sig_new = d->sigmas[j];
                build_filter_one(p, nof, d->sigmas[j], npsdbin, psdfreqs, d->psds + (size_t)j * npsdbin,
                                 d->plateaus[j], c->fcov + (size_t)ipt * nh, &sig_new);
                d->sigmas[j] = sig_new;
                for (long k = 0; k < nh; ++k)
                    if (isnan(c->fcov[(size_t)ipt * nh + k])) { fail(msg, "NaN in fcov. Check input PSD and sigma."); goto error; }
            }
        }
    }
    R->t_setup = now_s() - t_start;

    /* ---------------- binning: pixel_matrix, count_hits, bin_tod ---------------- */
    double t0 = now_s();
    int nmap = c->nmap, ncc = nmap * nmap;
    int64_t nside_max = c->nside_max;
    int64_t npix_max = 12 * nside_max * nside_max;
    int64_t npix_map = 12LL * p->nside_map * p->nside_map;
    int64_t npix_cross = 12LL * p->nside_cross * p->nside_cross;
    int64_t deg_map = npix_max / npix_map, deg_cross = npix_max / npix_cross;
    int64_t nb = c->nbase;
    loccc = (double *)calloc((size_t)npix_max * ncc, sizeof(double));
    loc = (double *)calloc((size_t)npix_max * nmap, sizeof(double));
    lochits = (int32_t *)calloc((size_t)npix_max, sizeof(int32_t));
    int nthreads = omp_get_max_threads();
    if (p->nthreads > 1 && p->nthreads < nthreads) nthreads = p->nthreads;
    R->nthreads = nthreads;
#pragma omp parallel num_threads(nthreads)
    {
        int tid = omp_get_thread_num(), nt = omp_get_num_threads();
        int64_t first = tid * npix_max / nt, last = (tid + 1) * npix_max / nt - 1;
        for (long id = 0; id < ndet; ++id)
            for (long iv = 0; iv < nperiod; ++iv) {
                double dwt = c->dw_ival[id * nperiod + iv];
                if (dwt == 0) continue;
                double sq = sqrt(dwt);
                for (int64_t k = c->kstart_pp[iv]; k < c->kstart_pp[iv] + c->noba_pp[iv]; ++k)
                    for (int64_t i = c->bstart[k]; i < c->bstart[k] + c->blen[k]; ++i) {
                        if (!sample_active(c, pass, i)) continue;
                        int64_t ip = pix[(size_t)id * nsamp + i];
                        if (ip < first || ip > last) continue;
                        const float *w = pixweights + ((size_t)id * nsamp + i) * nnz;
                        double d = signal[(size_t)id * nsamp + i];
                        lochits[ip] += 1;
                        if (nmap == 1) {
                            loccc[ip] = loccc[ip] + dwt;
                            loc[ip] = loc[ip] + d * dwt;
                        } else {
                            double w1 = (double)w[0] * sq, w2 = (double)w[1] * sq, w3 = (double)w[2] * sq;
                            double *L = loccc + ip * 9;   /* L[a + 3b] = cc(a+1, b+1) */
                            L[0] = L[0] + w1 * w1;
                            L[3] = L[3] + w1 * w2;
                            L[6] = L[6] + w1 * w3;
                            L[4] = L[4] + w2 * w2;
                            L[7] = L[7] + w2 * w3;
                            L[8] = L[8] + w3 * w3;
                            double *M = loc + ip * 3;
                            M[0] = M[0] + (double)w[0] * d * dwt;
                            M[1] = M[1] + (double)w[1] * d * dwt;
                            M[2] = M[2] + (double)w[2] * d * dwt;
                        }
                    }
            }
    }
    /* collect_cc / collect_map / collect_hits with nside degrade: maptod_transfer.f90:461-936 */
    cc = (double *)calloc((size_t)npix_map * ncc, sizeof(double));
    cca = (double *)calloc((size_t)npix_cross * ncc, sizeof(double));
    map = (double *)calloc((size_t)npix_map * nmap, sizeof(double));
    binmap = (double *)calloc((size_t)npix_map * nmap, sizeof(double));
    wamap = (double *)calloc((size_t)npix_cross * nmap, sizeof(double));
    R->hits = (int32_t *)calloc((size_t)npix_map, sizeof(int32_t));
    for (int64_t q = 0; q < npix_map; ++q)
        for (int64_t ch = q * deg_map; ch < (q + 1) * deg_map; ++ch) {
            for (int a = 0; a < ncc; ++a) cc[q * ncc + a] += loccc[ch * ncc + a];
            for (int a = 0; a < nmap; ++a) {
                if (p->kfirst) map[q * nmap + a] += loc[ch * nmap + a];
                binmap[q * nmap + a] += loc[ch * nmap + a];
            }
            R->hits[q] += lochits[ch];
        }
    if (p->kfirst)
        for (int64_t q = 0; q < npix_cross; ++q)
            for (int64_t ch = q * deg_cross; ch < (q + 1) * deg_cross; ++ch) {
                for (int a = 0; a < ncc; ++a) cca[q * ncc + a] += loccc[ch * ncc + a];
                for (int a = 0; a < nmap; ++a) wamap[q * nmap + a] += loc[ch * nmap + a];
            }
    R->t_bin = now_s() - t0;

    aa = (double *)calloc((size_t)nb * ndet + 1, sizeof(double));
    yba = (double *)calloc((size_t)nb * ndet + 1, sizeof(double));
    nna = (double *)calloc((size_t)nb * ndet + 1, sizeof(double));

    if (p->kfirst) {
        t0 = now_s();
        /* invert_pixelmatrix_cross: src/map_routines.f90:83-212 */
        if (nmap == 1) {
            for (int64_t q = 0; q < npix_cross; ++q) cca[q] = cca[q] != 0 ? 1.0 / cca[q] : 0;
        } else {
#pragma omp parallel for schedule(static)
            for (int64_t q = 0; q < npix_cross; ++q) {
                double *Cq = cca + q * 9;
                int allz = 1;
                for (int a = 0; a < 9; ++a) if (Cq[a] != 0) { allz = 0; break; }
                if (allz) continue;
                int istart = test_decoupling(Cq, 3, p->allow_decoupling) - 1;
                double sub[9], cdet = 0, rc = 0;
                sub_extract(Cq, 3, istart, sub);
                int m = 3 - istart, sing = 0;
                switch (p->pixmode_cross) {
                case 0:
                    sing = invert_lu_fm(sub, m, &cdet);
                    sub_store(Cq, 3, istart, sub);
                    if (cdet < p->pixlim_cross || sing) memset(Cq, 0, sizeof(double) * 9);
                    break;
                case 1: {
                    double tr = 0, trn;
                    for (int a = istart; a < 3; ++a) tr += Cq[a + 3 * a];
                    trn = fabs(tr) < F32(1.e-30) ? 0 : pow(3.0 / tr, 3);
                    sing = invert_lu_fm(sub, m, &cdet);
                    sub_store(Cq, 3, istart, sub);
                    if (cdet * trn < p->pixlim_cross || sing) memset(Cq, 0, sizeof(double) * 9);
                    break;
                }
                case 2:
                    sing = invert_eig_fm(sub, m, p->incomplete_matrices, &cdet, &rc);
                    sub_store(Cq, 3, istart, sub);
                    if (rc < p->pixlim_cross || sing) memset(Cq, 0, sizeof(double) * 9);
                    break;
                case 3:
                    if (Cq[0] > F32(1e-30)) {
                        if (istart == 0) Cq[0] = 1.0 / Cq[0];
                        for (int a = 1; a < 9; ++a) Cq[a] = 0;
                    } else memset(Cq, 0, sizeof(double) * 9);
                    break;
                case 4:
                    invert_eig_s_fm(sub, m, p->pixlim_cross, &cdet, &rc);
                    sub_store(Cq, 3, istart, sub);
                    break;
                }
            }
        }
        /* initialize_a: src/smadam_routines.f90:428-671 */
        ccmultiply(nmap, cca, wamap, npix_cross);
        lm = (double *)malloc(sizeof(double) * (size_t)npix_max * nmap);
        for (int64_t ch = 0; ch < npix_max; ++ch)
            for (int a = 0; a < nmap; ++a) lm[ch * nmap + a] = wamap[(ch / deg_cross) * nmap + a];
#pragma omp parallel for schedule(dynamic, 1)
        for (long id = 0; id < ndet; ++id)
            for (long iv = 0; iv < nperiod; ++iv) {
                double dwt = c->dw_ival[id * nperiod + iv];
                if (dwt == 0) continue;
                for (int64_t k = c->kstart_pp[iv]; k < c->kstart_pp[iv] + c->noba_pp[iv]; ++k) {
                    double y = 0, n = 0;
                    for (int64_t i = c->bstart[k]; i < c->bstart[k] + c->blen[k]; ++i) {
                        if (!sample_active(c, pass, i)) continue;
                        int64_t ip = pix[(size_t)id * nsamp + i];
                        if (ip < 0) continue;
                        const double *L = lm + ip * nmap;
                        int allz = 1;
                        for (int a = 0; a < nmap; ++a) if (L[a] != 0) { allz = 0; break; }
                        if (allz) continue;
                        y = y + signal[(size_t)id * nsamp + i];
                        n = n + 1;
                        if (nmap == 1) y = y - L[0];
                        else {
                            const float *w = pixweights + ((size_t)id * nsamp + i) * nnz;
                            y = y - (L[0] * (double)w[0] + L[1] * (double)w[1] + L[2] * (double)w[2]);
                        }
                    }
                    yba[(size_t)id * nb + k] = y * dwt;
                    nna[(size_t)id * nb + k] = n * dwt;
                }
            }
        /* empty baselines: eigenvalues == 0 -> nna = yba = 0 (src/smadam_routines.f90:604-623) */
        for (size_t i = 0; i < (size_t)nb * ndet; ++i)
            if (fabs(nna[i]) < F32(1e-30)) { nna[i] = 0; yba[i] = 0; }
        R->t_init = now_s() - t0;

        t0 = now_s();
        if (construct_preconditioner(c, nna, &pr, msg)) goto error;
        R->t_precond = now_s() - t0;
        R->precond_kind = pr.kind;
        R->nband_max = pr.nband_max;
        R->nchol = pr.nchol;
        R->ncgfallback = pr.ncg;
        R->nof = (int)c->nof;

        /* iterate_a: src/smadam_routines.f90:677-1483 */
        t0 = now_s();
        size_t nv = (size_t)nb * ndet;
        r = (double *)calloc(nv + 1, sizeof(double));
        pv = (double *)calloc(nv + 1, sizeof(double));
        z = (double *)calloc(nv + 1, sizeof(double));
        ap = (double *)calloc(nv + 1, sizeof(double));
        ro = (double *)calloc(nv + 1, sizeof(double));
        rmask = (char *)calloc(nv + 1, 1);
        for (long id = 0; id < ndet; ++id)
            for (long iv = 0; iv < nperiod; ++iv) {
                int64_t ks = c->kstart_pp[iv], ke = ks + c->noba_pp[iv];
                for (;;) {
                    int64_t noba = trim_interval(c, nna + (size_t)id * nb, &ks, &ke);
                    if (noba == 0) break;
                    for (int64_t k = ks; k < ks + noba; ++k) rmask[(size_t)id * nb + k] = 1;
                    ks += noba;
                }
            }
        memcpy(r, yba, sizeof(double) * nv);
        apply_preconditioner(c, &pr, nna, r, z);
        memcpy(pv, z, sizeof(double) * nv);
        double rz = 0, rr = 0;
        for (size_t i = 0; i < nv; ++i) if (rmask[i]) { rz += r[i] * z[i]; rr += r[i] * r[i]; }
        double rzinit = rz, rrinit = rr;
        R->rzinit = rzinit;
        R->rrinit = rrinit;
        R->rr = (double *)calloc(p->iter_max + 1, sizeof(double));
        int istep = 0;
        wam = (double *)malloc(sizeof(double) * (size_t)npix_cross * nmap);
        if (!(rzinit < F32(1.e-30)) && !isnan(rz)) {
            for (;;) {
                if (istep >= p->iter_max) break;
                istep++;
                /* baseline_to_map: pixel-range ownership, each thread scans all samples */
                memset(lm, 0, sizeof(double) * (size_t)npix_max * nmap);
#pragma omp parallel num_threads(nthreads)
                {
                    int tid = omp_get_thread_num(), nt = omp_get_num_threads();
                    int64_t npt = (npix_max + nt - 1) / nt;
                    int64_t first = tid * npt, last = first + npt - 1;
                    for (long id = 0; id < ndet; ++id)
                        for (long iv = 0; iv < nperiod; ++iv) {
                            double dwt = c->dw_ival[id * nperiod + iv];
                            if (dwt == 0) continue;
                            for (int64_t k = c->kstart_pp[iv]; k < c->kstart_pp[iv] + c->noba_pp[iv]; ++k) {
                                double pw = pv[(size_t)id * nb + k] * dwt;
                                for (int64_t i = c->bstart[k]; i < c->bstart[k] + c->blen[k]; ++i) {
                                    if (!sample_active(c, pass, i)) continue;
                                    int64_t ip = pix[(size_t)id * nsamp + i];
                                    if (ip < first || ip > last) continue;
                                    if (nmap == 1) lm[ip] = lm[ip] + pw;
                                    else {
                                        const float *w = pixweights + ((size_t)id * nsamp + i) * nnz;
                                        double *L = lm + ip * 3;
                                        L[0] = L[0] + (double)w[0] * pw;
                                        L[1] = L[1] + (double)w[1] * pw;
                                        L[2] = L[2] + (double)w[2] * pw;
                                    }
                                }
                            }
                        }
                }
                memset(wam, 0, sizeof(double) * (size_t)npix_cross * nmap);
                for (int64_t q = 0; q < npix_cross; ++q)
                    for (int64_t ch = q * deg_cross; ch < (q + 1) * deg_cross; ++ch)
                        for (int a = 0; a < nmap; ++a) wam[q * nmap + a] += lm[ch * nmap + a];
                ccmultiply(nmap, cca, wam, npix_cross);
                for (int64_t ch = 0; ch < npix_max; ++ch)
                    for (int a = 0; a < nmap; ++a) lm[ch * nmap + a] = wam[(ch / deg_cross) * nmap + a];
                if (p->kfilter) convolve_pp(c, c->fcov, nna, pv, ap);
                else {
                    for (size_t i = 0; i < nv; ++i) ap[i] = 0;
                    if (p->diagfilter != 0)
                        for (size_t i = 0; i < nv; ++i) ap[i] = nna[i] * pv[i] * p->diagfilter;
                }
                /* map_to_baseline: (detector, interval) tasks */
#pragma omp parallel for schedule(static, 1) num_threads(nthreads)
                for (long t = 0; t < ndet * nperiod; ++t) {
                    long id = t / nperiod, iv = t % nperiod;
                    double dwt = c->dw_ival[id * nperiod + iv];
                    if (dwt == 0) continue;
                    for (int64_t k = c->kstart_pp[iv]; k < c->kstart_pp[iv] + c->noba_pp[iv]; ++k) {
                        double apn = 0;
                        for (int64_t i = c->bstart[k]; i < c->bstart[k] + c->blen[k]; ++i) {
                            if (!sample_active(c, pass, i)) continue;
                            int64_t ip = pix[(size_t)id * nsamp + i];
                            if (ip < 0) continue;
                            const double *L = lm + ip * nmap;
                            if (nmap == 1) apn = apn + L[0];
                            else {
                                const float *w = pixweights + ((size_t)id * nsamp + i) * nnz;
                                apn = apn + (double)w[0] * L[0] + (double)w[1] * L[1] + (double)w[2] * L[2];
                            }
                        }
                        size_t kk = (size_t)id * nb + k;
                        apn = nna[kk] * pv[kk] - apn * dwt;
                        ap[kk] = ap[kk] + apn;
                    }
                }
                double pap = 0;
                for (size_t i = 0; i < nv; ++i) if (rmask[i]) pap += pv[i] * ap[i];
                double alpha = rz / pap;
                memcpy(ro, r, sizeof(double) * nv);
                for (size_t i = 0; i < nv; ++i) aa[i] = aa[i] + alpha * pv[i];
                for (size_t i = 0; i < nv; ++i) r[i] = r[i] - alpha * ap[i];
                apply_preconditioner(c, &pr, nna, r, z);
                double rzo = rz, rz2 = 0;
                rz = 0; rr = 0;
                for (size_t i = 0; i < nv; ++i)
                    if (rmask[i]) { rz += r[i] * z[i]; rz2 += ro[i] * z[i]; rr += r[i] * r[i]; }
                double beta = (rz - rz2) / rzo;
                R->rr[istep - 1] = rr / rrinit;
                if (rr / rrinit > 1e3) { fail(msg, "CG is diverging"); R->niter = istep; goto error; }
                if (rr / rrinit < p->cglimit && istep > p->iter_min) break;
                if (rz == 0) break;
                for (size_t i = 0; i < nv; ++i) pv[i] = z[i] + beta * pv[i];
            }
        }
        R->niter = istep;
        R->t_iter = now_s() - t0;

        /* subtract_baselines_a: src/smadam_routines.f90:1489-1539 */
        t0 = now_s();
        memset(lm, 0, sizeof(double) * (size_t)npix_max * nmap);
        for (long id = 0; id < ndet; ++id)
            for (long iv = 0; iv < nperiod; ++iv) {
                double dwt = c->dw_ival[id * nperiod + iv];
                if (dwt == 0) continue;
                for (int64_t k = c->kstart_pp[iv]; k < c->kstart_pp[iv] + c->noba_pp[iv]; ++k)
                    for (int64_t i = c->bstart[k]; i < c->bstart[k] + c->blen[k]; ++i) {
                        if (!sample_active(c, pass, i)) continue;
                        int64_t ip = pix[(size_t)id * nsamp + i];
                        if (ip < 0) continue;
                        double aw = aa[(size_t)id * nb + k] * dwt;
                        if (nmap == 1) lm[ip] = lm[ip] - aw;
                        else {
                            const float *w = pixweights + ((size_t)id * nsamp + i) * nnz;
                            double *L = lm + ip * 3;
                            L[0] = L[0] - aw * (double)w[0];
                            L[1] = L[1] - aw * (double)w[1];
                            L[2] = L[2] - aw * (double)w[2];
                        }
                    }
            }
        for (int64_t q = 0; q < npix_map; ++q)
            for (int64_t ch = q * deg_map; ch < (q + 1) * deg_map; ++ch)
                for (int a = 0; a < nmap; ++a) map[q * nmap + a] += lm[ch * nmap + a];
    } else {
        t0 = now_s();
    }

    /* invert_pixelmatrix_map: src/map_routines.f90:218-321 */
    R->mask = (int32_t *)calloc((size_t)npix_map, sizeof(int32_t));
    R->crit = (double *)calloc((size_t)npix_map, sizeof(double));
    if (nmap == 1) {
        for (int64_t q = 0; q < npix_map; ++q) {
            if (cc[q] != 0) { cc[q] = 1.0 / cc[q]; R->mask[q] = 1; }
            else { cc[q] = 0; R->mask[q] = 0; }
        }
    } else {
#pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < npix_map; ++q) {
            double *Cq = cc + q * 9;
            int allz = 1;
            for (int a = 0; a < 9; ++a) if (Cq[a] != 0) { allz = 0; break; }
            if (allz) continue;
            int istart = test_decoupling(Cq, 3, p->allow_decoupling) - 1;
            double sub[9], cdet = 0, rc = 0, pcrit = 0;
            sub_extract(Cq, 3, istart, sub);
            int m = 3 - istart, sing = 0;
            if (p->pixmode_map == 0) { sing = invert_lu_fm(sub, m, &cdet); pcrit = cdet; }
            else if (p->pixmode_map == 1) {
                double tr = 0, trn;
                for (int a = istart; a < 3; ++a) tr += Cq[a + 3 * a];
                trn = fabs(tr) < F32(1e-30) ? 0 : pow(3.0 / tr, 3);
                sing = invert_lu_fm(sub, m, &cdet);
                pcrit = cdet * trn;
            } else { sing = invert_eig_fm(sub, m, p->incomplete_matrices, &cdet, &rc); pcrit = rc; }
            sub_store(Cq, 3, istart, sub);
            if (pcrit < p->pixlim_map || sing) { memset(Cq, 0, sizeof(double) * 9); R->mask[q] = 0; }
            else R->mask[q] = 1;
            R->crit[q] = (double)(float)pcrit;
        }
    }
    /* makemaps: src/map_routines.f90:327-357 */
    ccmultiply(nmap, cc, binmap, npix_map);
    ccmultiply(nmap, cc, map, npix_map);
    if (p->rm_monopole) {
        int64_t cover = 0;
        for (int64_t q = 0; q < npix_map; ++q) cover += R->mask[q] > 0;
        if (cover > 0) {
            double mono = 0;
            for (int64_t q = 0; q < npix_map; ++q) mono += map[q * nmap];
            mono = mono / cover;
            for (int64_t q = 0; q < npix_map; ++q) if (R->mask[q] > 0) map[q * nmap] -= mono;
            mono = 0;
            for (int64_t q = 0; q < npix_map; ++q) mono += binmap[q * nmap];
            mono = mono / cover;
            for (int64_t q = 0; q < npix_map; ++q) if (R->mask[q] > 0) binmap[q * nmap] -= mono;
        }
    }

    /* clean_tod: src/smadam_routines.f90:1545-1585 (final full-data pass only) */
    if (pass == 0 && p->kfirst && (p->write_tod || p->ndetset > 0 || p->nsurvey > 0)) {
        for (long id = 0; id < ndet; ++id)
            for (int64_t k = 0; k < nb; ++k)
                for (int64_t i = c->bstart[k]; i < c->bstart[k] + c->blen[k]; ++i)
                    signal[(size_t)id * nsamp + i] = signal[(size_t)id * nsamp + i] - 1.0 * aa[(size_t)id * nb + k];
    }
    /* restore_pixels: flagged samples end as 12*nside_max**2 (src/pointing.f90:168-191) */
    for (size_t i = 0; i < (size_t)ndet * nsamp; ++i) if (pix[i] < 0) pix[i] = (int32_t)npix_max;
    R->t_final = now_s() - t0;

    R->nmap = nmap;
    R->ndet = (int)ndet;
    R->nside_map = p->nside_map;
    R->nside_cross = p->nside_cross;
    R->npix_map = npix_map;
    R->npix_cross = npix_cross;
    R->nsamp = nsamp;
    R->nbase = nb;
    R->map = map; map = NULL;
    R->binmap = binmap; binmap = NULL;
    R->cc = cc; cc = NULL;
    R->aa = aa; aa = NULL;
    R->yba = yba; yba = NULL;
    R->nna = nna; nna = NULL;
    R->base_start = c->bstart; c->bstart = NULL;
    R->base_len = c->blen; c->blen = NULL;
    R->status = 0;
    goto cleanup;
error:
    R->status = 1;
    /* restore flagged pixels like the reference would have on a normal exit */
cleanup:
    free(loc); free(loccc); free(lochits); free(cca); free(cc); free(wamap); free(map); free(binmap);
    free(aa); free(yba); free(nna); free(r); free(pv); free(z); free(ap); free(ro); free(rmask); free(lm); free(wam);
    free_prec(&pr);
    if (c->det) {
        for (long id = 0; id < ndet; ++id) {
            free(c->det[id].weights); free(c->det[id].sigmas); free(c->det[id].psdstarts);
            free(c->det[id].plateaus); free(c->det[id].psds);
        }
        free(c->det);
    }
    free(c->intervals); free(c->ival_start); free(c->noba_pp); free(c->kstart_pp);
    free(c->bstart); free(c->blen); free(c->btime); free(c->subchunk); free(c->dw_ival);
    free(c->fcov); free(c->psddet); free(c->psdstart); free(c->psdstop);
    offt_destroy(c->plan);
    R->t_total = now_s() - t_start;
    return R;
}

void oracle_free(oracle_result *r)
{
    if (!r) return;
    free(r->map); free(r->binmap); free(r->cc); free(r->crit); free(r->hits); free(r->mask);
    free(r->aa); free(r->yba); free(r->nna); free(r->base_start); free(r->base_len); free(r->rr);
    free(r);
}
