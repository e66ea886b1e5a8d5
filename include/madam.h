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
/* hit counts [npix] of the pointing set with madam_set_pointing (count_hits,
 * src/smadam_routines.f90:349-422) */
int madam_get_hits(void *ctx, int32_t *hits);
/* The samples grouped by map pixel, which fixes the order of every per-pixel
 * sum (the reference visits a pixel's samples in time order inside each
 * thread's pixel range, src/smadam_routines.f90:111-422; no entry point of its
 * own): off [npix + 1] the first entry of each pixel, src [nactive] the sample
 * index (det * nsamp + t) of each entry, ascending inside a pixel, bord
 * [ndet * nbase] the baseline (det * nbase + k) held by each slot of the
 * projections' baseline order.  NULL buffers are skipped. */
int madam_get_csr(void *ctx, int64_t *off, uint32_t *src, int32_t *bord);
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
/* benchmark: exactly niter steps of the last batch's PCG (the FP32 inner one
 * for precision 1) on every right-hand side, restarted from 0, without the
 * exit tests; the batch's results are not refreshed */
int madam_iterate_batch(void *ctx, int niter);
/* destriped and binned maps [npix][nmap] of right-hand side irhs (NULL skips) */
int madam_get_batch_maps(void *ctx, int irhs, double *map, double *binmap);
/* baselines and right-hand side [ndet][nbase] of right-hand side irhs */
int madam_get_batch_baselines(void *ctx, int irhs, double *aa, double *yba);
/* out[0..n): last PCG's steps, status (1 converged, 2 diverging, 4 iteration
 * limit, 8 nothing to solve, 16 stalled), its rr / rr0, refinement passes,
 * the FP64 residual's rr / rrinit, t_setup, t_solve, PCG steps over the
 * batch, their time, entries per pixel-side work item (autotuned); with
 * madam_set_profiling on, the summed HIP-event times of the steps' pixel
 * side, prior, baseline side, preconditioner and whole step, and the number
 * of steps timed; the run-length records of the FP32 operator on the baseline
 * and pixel sides (0 when it reads the sample records) (n <= 18) */
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
 * continuation of repeating 64-row blocks); + 32 / + 64 = band carry chains
 * always / never several per wave (default: several when there are enough
 * of them); no variant changes a result bit */
int madam_set_kernel_variant(void *ctx, int variant);
/* pointing expansion of the synthetic ring scan: scan5 = {fsample, spin_period,
 * prec_period, opening_turns, lon0_turns}, dets3[ndet][3] = {psi_turns,
 * open_offset_turns, spin_offset_turns}; outputs pix[ndet][n], weights[ndet][n][3] */
int madam_expand_pointing(const double *scan5, const double *dets3, int ndet, long nside, long i0, long n,
                          int32_t *pix, float *weights, int on_device);
/* No counterpart in the reference.  madam_expand_pointing for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): pix / weights are device buffers
 * of the caller on the current device, the detector table is staged in a
 * fixed per-GPU array of the library (ndet <= 4096, 96 KiB per GPU, held
 * until the process ends), the work is queued on `stream` (a hipStream_t of
 * the current device; NULL = the null stream) and nothing is allocated or
 * freed.  Calls that overlap on one GPU share that array and race; under
 * graph capture dets3 must stay valid while the graph is replayed. */
int madam_expand_pointing_noalloc(const double *scan5, const double *dets3, int ndet, long nside, long i0, long n,
                                  int32_t *pix, float *weights, void *stream);
/* ------------------------------------------------------------------ request loop
 * Destripe jobs served on one GPU (no counterpart in the reference, whose
 * destripe() serves one job per call and keeps one cached configuration in
 * Monte Carlo mode, src/smadam.F90:440-448, 593-951).  Jobs are admitted by
 * an estimate of their device memory against hbm_budget_bytes, first-fit in
 * arrival order, up to max_inflight at once, each on its own stream; jobs
 * with equal non-zero scan_key share one resident set-up (pointing, pixel
 * matrices, filters, preconditioner: the key must identify the pointing and
 * the parameters), and those waiting together are solved as one batch of
 * right-hand sides.  The job's buffers are read in place until it is done;
 * on_device: pix, pixweights and signal are device memory of the server's
 * GPU (the other arrays are host memory).  precision as madam_solve_batch;
 * signal float64 or float32 (tod_f32). */
void *madam_server_create(int device, double hbm_budget_bytes, int max_inflight);
void madam_server_destroy(void *srv);
/* returns the job id, or -1 when the job alone exceeds the budget */
long madam_server_submit(void *srv, const char *parstring, long ndet, const double *detweights, long nsamp, long nnz,
                         const double *timestamps, const int32_t *pix, const float *pixweights, const void *signal,
                         int tod_f32, long nperiod, const long *periods, const long *npsd, long npsdtot,
                         const double *psdstarts, long npsdbin, const double *psdfreqs, long npsdval,
                         const double *psdvals, unsigned long long scan_key, int precision, int on_device);
/* 0 done, 1 failed (madam_server_error), 2 not finished within timeout_s (< 0: wait) */
int madam_server_wait(void *srv, long id, double timeout_s);
const char *madam_server_error(void *srv, long id);
/* map, binmap [npix][nmap], hits [npix], aa [ndet][nbase] (NULL skips);
 * sizes6 = {npix, nmap, nbase, ndet, batch size, resident scan};
 * times4 = {submitted, started, done (seconds, steady clock), PCG steps} */
int madam_server_result(void *srv, long id, double *map, double *binmap, int32_t *hits, double *aa, long *sizes6,
                        double *times4);
int madam_server_release(void *srv, long id);
/* out = {done, failed, batches, jobs in batches, jobs on resident scans,
 * evictions, peak estimated bytes, most jobs in flight, resident scans,
 * evictions forced by the measured free device memory, least free device
 * memory seen (bytes)}; n <= 11 */
int madam_server_stats(void *srv, double *out, int n);
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
