// Batched, mixed-precision PCG: interfaces between the host orchestration
// (madam_batch_host.cpp) and the kernels (madam_batch.hip).  R right-hand
// sides share one context's pointing and set-up; baseline vectors are
// RHS-major x[r * nbtot + b] (see madam_batch.hip for the layouts).
#pragma once
#include "madam_internal.hpp"

namespace madam {

// pixel-side work item: entries e0 .. e0 + n - 1 of the cross pixel of
// ordinal k; part >= 0 stores the sums as piece `part` of a long pixel
struct BsItem {
    int64_t e0;
    int32_t n;
    int32_t k;
    int32_t part;
    int32_t pad;
};
// a long pixel: ordinal k, its pieces part0 .. part0 + nparts - 1
struct BsFix {
    int32_t k;
    int32_t part0;
    int32_t nparts;
};
struct BsTables {
    const BsItem *items = nullptr;
    int64_t nitems = 0;
    const BsFix *fix = nullptr;
    int64_t nfix = 0;
    int64_t nparts = 0;
};

// PCG scalars of one right-hand side (device)
struct BatchScal {
    double rz, pap, beta, rr, rr0;
    double lim;     // exit threshold on rr / rr0 (used when the call passes lim < 0)
    double rrmin;   // smallest rr so far (stall test)
    int32_t steps, status;   // status: 1 converged, 2 diverging, 4 iteration limit, 8 nothing to solve, 16 stalled
    int32_t nstall, pad;
};

// the power-of-two kernel width for nrhs right-hand sides (-1 above 64)
int batch_width(int nrhs);

// p = z + beta p and the next operator's pw (batch_pw) in one pass
template <typename T>
void batch_update_p_pw(hipStream_t s, int R, const BatchScal *sc, const int *active, T *p, const T *z, int64_t nbtot,
                       const int32_t *bslot, const T *dwb, T *pw);
template <typename T>
void batch_pw(hipStream_t s, int R, const T *p, int64_t nbtot, const int32_t *bord, const int32_t *bslot,
              const T *dwb_s, const T *dwb, const int *active, T *pw);
// the earliest baseline (index in its detector) among each non-empty pixel's entries
void batch_pixel_keys(hipStream_t s, const int64_t *segstart, int64_t nz, const void *ent, int rec_b,
                      const int32_t *bord, int64_t nbase, int64_t *key);
// run-length records of the two projections (FP32 path): count_only marks the
// run starts and scans them into pos[n + 1]; the second call writes the
// records, and per detector the baselines' first records (cst[ndet][nbase + 1],
// relative to dbase[det]) or the pixels' first entries (cseg[nz + 1])
void batch_rle_gather(hipStream_t s, const PixW *rec, int ndet, int64_t nsamp, const int64_t *bstart, int64_t nbase,
                      int dsh, int temperature, uint32_t *bmask, int32_t *flag, int64_t *pos, int64_t *scan_tmp,
                      PixW *out, int64_t *cst, int64_t *dbase, bool count_only);
void batch_rle_scatter(hipStream_t s, const PixW *ent, int64_t nent, const int64_t *segstart, int64_t nz,
                       int temperature, uint32_t *smask, int32_t *flag, int64_t *pos, int64_t *scan_tmp, PixW *out,
                       int64_t *cseg, bool count_only);
template <typename T>
void batch_scatter(hipStream_t s, int R, int rl, int nmap, const BsTables &bt, const void *ent, const T *pw,
                   const T *cc_c, const int32_t *nzq, T *m, T *part);
template <typename T>
// bstride, dbase: 0, null for the sample records (bstart shared by the
// detectors); nbase + 1 and the per-detector record bases for run-length
// records (bstart[det][nbase + 1] relative to dbase[det]; R >= 8, ndeg 1)
void batch_gather(hipStream_t s, int R, int rl, int nmap, int ndet, int64_t nsamp, int64_t nbase,
                  const int64_t *bstart, const void *rec, const T *m, int64_t ndeg, const T *dwb, const T *nna,
                  const T *p, T *ap, const int *active, int64_t bstride = 0, const int64_t *dbase = nullptr);
template <typename T>
void batch_dots(hipStream_t s, int which, int R, const T *a, const T *b, const T *c, const int8_t *mask, int64_t n,
                double *partials);
// stage 0: pap; 1: rz, rr, beta and the exit tests; 2: start of a solve (rz, rr0, active where allow[r],
// or everywhere when allow is null); 3: test of the refinement residual (rr / rr0 < lim, steps = passes).
// lim < 0 takes each right-hand side's own BatchScal::lim; stall > 0 stops a right-hand side whose rr has
// not fallen below 0.999 of its smallest value for `stall` steps.
void batch_final(hipStream_t s, int R, const double *partials, int stage, BatchScal *sc, int *active, double lim,
                 int iter_min, int iter_max, const int *allow = nullptr, int stall = 0);
// inner[r].lim = max(cglimit * outer[r].rr0 / outer[r].rr, floor): the reduction one refinement pass needs
void batch_inner_limits(hipStream_t s, int R, const BatchScal *outer, BatchScal *inner, double cglimit, double floor);
size_t batch_partials_doubles(int R);
template <typename T>
void batch_update_r(hipStream_t s, int R, const BatchScal *sc, const int *active, T *x, T *r, T *ro, const T *p,
                    const T *ap, int64_t n);
template <typename T>
void batch_update_p(hipStream_t s, int R, const BatchScal *sc, const int *active, T *p, const T *z, int64_t n);
template <typename TI, typename TO>
void batch_cast(hipStream_t s, const TI *in, TO *out, int64_t n);
template <typename T>
void batch_accumulate(hipStream_t s, int R, double *a, const T *x, const int *was_active, int64_t n);
void batch_residual(hipStream_t s, const double *b, const double *ap, double *r, int64_t n);
template <typename T>
void batch_diag(hipStream_t s, int R, const T *r, const double *pd, T *z, int64_t n);
template <typename T>
void batch_copy(hipStream_t s, T *d, const T *a, int64_t n);
template <typename T>
void batch_fill(hipStream_t s, T *d, T v, int64_t n);
// the map of right-hand side r from the RHS-minor layout into double4 per pixel
void batch_map_unpack(hipStream_t s, const void *m, int is_float, int64_t npix, int nmap, int R, int r, double *out4);

}  // namespace madam
