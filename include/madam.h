/*
 * libmadam.so — MI355X (gfx950) implementation of libmadam's destriping solve.
 *
 * The first three entry points keep the reference's C ABI unchanged; the
 * madam_* entry points are additive (in-memory outputs, device-resident
 * buffers, error codes).  Paths cite the reference source tree.
 *
 * Buffer types behind void* (src/smadam.F90:151-154, src/pointing.f90:13-14,
 * src/tod_storage.f90:14-15): timestamps float64[nsamp], pix int32[ndet][nsamp],
 * pixweights float32[ndet][nsamp][nnz], signal float64[ndet][nsamp].  As in
 * the reference, flagged samples are left as 12*nside_max**2 in the caller's
 * pix array, and the baselines are subtracted from signal when write_tod or a
 * detset / survey is given.  Not re-entrant: one call at a time per process.
 * destripe() writes the reference's FITS products (map, binned map, hits,
 * mask, pixel matrices; subchunk and subset variants) under path_output;
 * madam_get_maps returns the same products in memory.
 */
#ifndef LIBMADAM_AMD_MADAM_H
#define LIBMADAM_AMD_MADAM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reference: src/madam.h:2, src/smadam.F90:64-590.  `comm` is a Fortran MPI
 * handle in the reference; this single-GPU library records and ignores it. */
void destripe(int comm, char *parstring, long ndet, char *detstring, double *detweights, long nsamp, long nnz,
              void *timestamps, void *pix, void *pixweights, void *signal, long nperiod, long *periods, long *npsd,
              long npsdtot, double *psdstarts, long npsdbin, double *psdfreqs, long npsdval, double *psdvals);

/* Reference: src/smadam.F90:593-918 (bind(c) name destripe_with_cache; not in
 * the reference madam.h).  Solves again on the cached configuration of a
 * destripe(mcmode=T) call, computing the right-hand side from a zeroed map. */
void destripe_with_cache(int comm, long ndet, long nsamp, long nnz, void *timestamps, void *pix, void *pixweights,
                         void *signal, char *outpath);

/* Reference: src/smadam.F90:921-951. */
void clear_caches(void);

/* ---------------------------------------------------------------- additive API
 * Replaces the reference's module-global state (src/commonparam.f90:38-194)
 * with an explicit handle.  All functions return 0 on success and 1 on error
 * (message from madam_last_error).  on_device != 0 means the pointers are
 * device pointers on the handle's GPU and are used in place.
 */
void *madam_create(int device);
void madam_destroy(void *ctx);
/* parameters, detectors, intervals, baselines: src/inputparam.f90:24-163,
 * src/read_data.f90:203-323, src/parameter_control.f90:27-409, 752-977.
 * subchunk_pass selects the pass of the subchunk loop (0 = full data). */
int madam_configure(void *ctx, const char *parstring, long ndet, const char *detstring, const double *detweights,
                    long nsamp, long nnz, const double *timestamps, long nperiod, const long *periods,
                    const long *npsd, long npsdtot, const double *psdstarts, long npsdbin, const double *psdfreqs,
                    long npsdval, const double *psdvals, int subchunk_pass);
/* flagging and pointing setup (src/read_data.f90:241-272, src/pointing.f90:115-191) */
int madam_set_pointing(void *ctx, int32_t *pix, float *pixweights, int on_device);
/* binning, inversion, PCG, final maps (src/smadam.F90:316-433) */
int madam_solve(void *ctx, double *signal, int on_device);
/* write the flagged samples back into a host pix array as 12*nside_max**2 */
int madam_restore_pixels(void *ctx, int32_t *host_pix);
/* out6 = {nmap, npix_map, nbase (per detector), ndet, iterations, nside_map} */
int madam_get_sizes(void *ctx, long *out6);
/* maps [npix][nmap], cc [npix][nmap][nmap] inverse pixel matrices; NULL skips */
int madam_get_maps(void *ctx, double *map, double *binmap, int32_t *hits, double *cc, int32_t *mask, double *crit);
/* baselines [ndet][nbase] */
int madam_get_baselines(void *ctx, double *aa, double *yba, double *nna);
/* run statistics, out[0..n): niter, rrinit, rzinit, t_pointing, t_bin, t_init,
 * t_precond, t_iter, t_final, t_scatter, t_gather, t_prior, t_apply_prec,
 * n_scatter, n_gather, precond_kind, nband, nseg, nof, rr_final, npix_hit,
 * nactive, t_total, t_step, ncg, band_chunks, band_chunks_unique, band_blocks,
 * band_blocks_unique, rec_bytes, band_rows_factored, prec_reused,
 * pointing_reused, nwide (the order of STAT_NAMES in libmadam_amd/__init__.py) */
int madam_get_stats(void *ctx, double *out, int n);
int madam_get_rr_history(void *ctx, double *out, int n);
int madam_set_profiling(void *ctx, int on);
/* benchmark / continuation: run exactly niter PCG iterations without the exit
 * tests (restart != 0 starts again from a = 0); products of the last
 * madam_solve are not refreshed */
int madam_iterate(void *ctx, int niter, int restart);
int madam_reset_timers(void *ctx);
/* Batched solve: nrhs (1..64) signals that share this handle's pointing
 * (Monte Carlo realisations of one scan: the reference solves them one
 * destripe_with_cache call at a time, src/smadam.F90:593-918), stored
 * [nrhs][ndet][nsamp] as float64 (tod_f32 = 0) or float32 (tod_f32 = 1).
 * The set-up (pixel matrices, nna, segments, filters, preconditioner) is that
 * of the first signal and serves all of them.  precision 0: one PCG over all
 * right-hand sides in FP64, each with the reference's recurrence and exit
 * tests (src/smadam_routines.f90:677-924); precision 1: FP32 operator and
 * vectors (FP64 dot products and scalars) inside an FP64 iterative refinement
 * to the same cglimit on the FP64 residual.  The signals are not modified. */
int madam_solve_batch(void *ctx, int nrhs, const void *signals, int tod_f32, int on_device, int precision);
/* destriped and binned maps [npix][nmap] of right-hand side irhs (NULL skips) */
int madam_get_batch_maps(void *ctx, int irhs, double *map, double *binmap);
/* baselines and right-hand side [ndet][nbase] of right-hand side irhs */
int madam_get_batch_baselines(void *ctx, int irhs, double *aa, double *yba);
/* out[0..n): last PCG's steps, status (1 converged, 2 diverging, 4 iteration
 * limit, 8 nothing to solve, 16 stalled), its rr / rr0, refinement passes,
 * the FP64 residual's rr / rrinit, t_setup, t_solve, PCG steps over the
 * batch, their time, entries per pixel-side work item (autotuned) */
int madam_get_batch_info(void *ctx, int irhs, double *out, int n);
/* one operator of the last madam_solve on host vectors x, y of ndet * nbase
 * baselines: which 0 = noise prior y = C_a^-1 x (cinvmul, src/noise_routines.f90:745-913),
 * 1 = preconditioner y = M^-1 x (src/noise_routines.f90:1159-1409), 2 = the
 * destriping operator y = A x of one PCG step (src/smadam_routines.f90:808-872) */
int madam_apply(void *ctx, int which, const double *x, double *y);
/* the noise filter of global PSD ipsd, fcov[nof / 2 + 1] (build_filter,
 * src/noise_routines.f90:367-490), after a madam_solve with kfilter = T */
int madam_get_filter(void *ctx, int ipsd, double *fcov);
/* projection kernels (low 4 bits): 0 = segmented 512-entry tiles on the pixel
 * side and LDS tiles on the baseline side (default), 1 = wave per pixel / lane
 * group per baseline, 2 = LDS tiles on both sides, 3 = segmented tiles on both
 * sides; + 16 = band factorization computes every row (no periodic
 * continuation of repeating 64-row blocks) */
int madam_set_kernel_variant(void *ctx, int variant);
/* pointing expansion of the synthetic ring scan: scan5 = {fsample, spin_period,
 * prec_period, opening_turns, lon0_turns}, dets3[ndet][3] = {psi_turns,
 * open_offset_turns, spin_offset_turns}; outputs pix[ndet][n], weights[ndet][n][3] */
int madam_expand_pointing(const double *scan5, const double *dets3, int ndet, long nside, long i0, long n,
                          int32_t *pix, float *weights, int on_device);
const char *madam_last_error(void);
/* status of the last destripe / destripe_with_cache call (0 ok); only
 * meaningful when MADAM_RAISE is set, otherwise errors end the process like
 * the reference's MPI_Abort */
int madam_last_status(void);
int madam_cache_handle(void **out);

#ifdef __cplusplus
}
#endif

#endif
