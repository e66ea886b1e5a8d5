// Internal interfaces between the host orchestration (madam_host.cpp) and the
// HIP kernels (madam_kernels.hip, madam_fft.hip, madam_band.hip).
//
// Device data model for one scan (pointing) on one GPU:
//   samples      detector-major, index g = det * nsamp + i (uint32_t; the
//                library rejects ndet * nsamp >= 2^32)
//   pixw         16-byte record per sample {pix, w0, w1, w2}; pix = -1 for
//                every sample that takes no part in this pass (flagged, zero
//                detector weight, no PSD period, other subchunk)
//   csr          samples regrouped by map pixel: off[npix_map + 1], and per
//                entry {bidx, w0, w1, w2} plus the source sample index; the
//                cross pixel of map pixel q is q / ndeg, so a cross pixel's
//                entries are a contiguous range of the map-pixel CSR
//   baselines    global baseline index b = det * nbase + k; vectors of
//                length nbtot = ndet * nbase
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace madam {

struct PixW {
    int32_t pix;
    float w0, w1, w2;
};

// Pointing records of the two per-iteration projections.  REC16 is the full
// record {index, w0, w1, w2}; when every active sample has an intensity weight
// of exactly 1 (the (1, cos 2psi, sin 2psi) convention), REC12 {index, w1, w2}
// and, for temperature-only maps, REC4 {index} carry the same information in
// fewer bytes (1.0f * x == x, so the results are the same bits).
enum RecLayout { REC16 = 0, REC12 = 1, REC4 = 2 };
struct Rec12 {
    int32_t i;
    float w1, w2;
};
inline int rec_bytes(int rl) { return rl == REC16 ? 16 : (rl == REC12 ? 12 : 4); }
// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2).
// Remap so that XCD x owns a contiguous range of logical blocks: consecutive
// work (the same ring of the scan, or neighbouring nested pixels) then shares
// one L2 instead of being spread over eight.
__device__ inline int64_t xcd_block(int64_t bid, int64_t nblk)
{
    const int64_t NX = 8;
    int64_t per = nblk / NX;
    if (bid >= per * NX) return bid;
    return (bid % NX) * per + bid / NX;
}

// zero records after the end of every record array (pixel 0, weights 0)
static constexpr int REC_PAD = 128;

// record i of a record array in layout RL, as a full record (w0 = 1 for the compact layouts)
template <int RL>
__device__ inline PixW load_rec(const void *base, int64_t i)
{
    PixW r;
    if constexpr (RL == REC16) {
        r = ((const PixW *)base)[i];
    } else if constexpr (RL == REC12) {
        Rec12 c = ((const Rec12 *)base)[i];
        r.pix = c.i;
        r.w0 = 1.0f;
        r.w1 = c.w1;
        r.w2 = c.w2;
    } else {
        r.pix = ((const int32_t *)base)[i];
        r.w0 = 1.0f;
        r.w1 = r.w2 = 0.0f;
    }
    return r;
}

// load_rec with the non-temporal hint: the projections stream their records
// once per pass, so the records should not displace the map and baseline
// rows that the same kernel reads many times from L2
typedef int rec_v4i __attribute__((ext_vector_type(4)));
template <int RL>
__device__ inline PixW load_rec_nt(const void *base, int64_t i)
{
    PixW r;
    if constexpr (RL == REC16) {
        rec_v4i v = __builtin_nontemporal_load((const rec_v4i *)base + i);
        r.pix = v.x;
        r.w0 = __int_as_float(v.y);
        r.w1 = __int_as_float(v.z);
        r.w2 = __int_as_float(v.w);
    } else if constexpr (RL == REC12) {
        const int32_t *q = (const int32_t *)base + 3 * i;
        r.pix = __builtin_nontemporal_load(q);
        r.w0 = 1.0f;
        r.w1 = __int_as_float(__builtin_nontemporal_load(q + 1));
        r.w2 = __int_as_float(__builtin_nontemporal_load(q + 2));
    } else {
        r.pix = __builtin_nontemporal_load((const int32_t *)base + i);
        r.w0 = 1.0f;
        r.w1 = r.w2 = 0.0f;
    }
    return r;
}

// two- and four-wide vector types of the projections' element type
template <typename T>
struct VecT;
template <>
struct VecT<double> {
    using v2 = double2;
    using v4 = double4;
};
template <>
struct VecT<float> {
    using v2 = float2;
    using v4 = float4;
};

struct CsrEnt {
    int32_t b;   // global baseline index
    float w0, w1, w2;
};

// Segment of consecutive baselines with nonzero nna inside one (det, interval),
// as produced by the reference's trim_interval (src/noise_routines.f90:764-813).
struct Segment {
    int64_t b0;      // first global baseline index
    int32_t noba;    // number of baselines
    int32_t det;
    int32_t ival;
    int32_t ipsd;    // global PSD index for the filter, -1 = none
    int32_t skip;    // preconditioner leaves z = r on this segment
    int32_t pad;
};

#define MADAM_CHECK(x)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) throw madam::HipError(e_, #x, __FILE__, __LINE__);           \
    } while (0)

struct HipError {
    hipError_t err;
    const char *what;
    const char *file;
    int line;
    HipError(hipError_t e, const char *w, const char *f, int l) : err(e), what(w), file(f), line(l) {}
};

// ------------------------------------------------------------------ kernels.hip

void launch_flag_and_count(hipStream_t s, int32_t *pix, const float *w, int nnz, int ndet, int64_t nsamp,
                           const int64_t *ival_start, int ninterval, unsigned long long *valid_counts,
                           unsigned long long *nzero);
void launch_flag_intervals(hipStream_t s, int32_t *pix, int64_t nsamp, const int64_t *ival_start,
                           int ninterval, int ndet, const int8_t *flag_ival);
// the packed records of the samples (pixel -1 for inactive ones) and, in the
// same pass, the CSR sort keys (the pixel, flag_key for inactive samples), the
// compact records (REC12 for nmap 3, REC4 for nmap 1; those of REC12 are only
// valid when *nonunit stays 0) and *nonunit += the number of active samples
// whose w0 is not exactly 1 (*nonunit must start at 0)
void launch_pack(hipStream_t s, const int32_t *pix, const float *w, int nnz, int nmap, int ndet, int64_t nsamp,
                 const int64_t *ival_start, int ninterval, const double *dw_ival, int nsubchunk, int pass,
                 PixW *pixw, uint32_t *key, uint32_t flag_key, void *pixc, unsigned long long *nonunit);
void launch_exclusive_scan_i32_to_i64(hipStream_t s, const int32_t *in, int64_t n, int64_t *out, int64_t *tmp,
                                      int64_t tmp_len);
// TT: the TOD's type (double, or float for the mixed-precision solves)
template <typename TT>
void launch_bin(hipStream_t s, int nmap, int64_t npix_map, const int64_t *off, const CsrEnt *ent, const uint32_t *src,
                const TT *tod, const double *dwb, const double *sdwb, double *cc6, double *rhs4);
void launch_degrade(hipStream_t s, int nmap, int64_t npix_cross, int64_t ndeg, const double *cc6, const double *rhs4,
                    double *cc6x, double *rhs4x);
void launch_bin_subset(hipStream_t s, int nmap, int64_t npix_map, const int64_t *off, const CsrEnt *ent,
                       const uint32_t *src, const uint8_t *smask, const double *tod, const double *dwb,
                       const double *sdwb, double *cc6, double *rhs4, int32_t *hits);
void launch_invert(hipStream_t s, int nmap, int64_t npix, const double *cc6, double *ccinv6, double *crit,
                   int32_t *mask, int pixmode, double pixlim, int incomplete, int allow_decoupling,
                   unsigned long long *ngood);
void launch_ccmultiply(hipStream_t s, int nmap, int64_t npix, const double *ccinv6, const double *rhs4, double *out4);
template <typename TT>
void launch_init_a(hipStream_t s, int nmap, int ndet, int64_t nsamp, int64_t nbase, const int64_t *bstart,
                   const PixW *pixw, const TT *tod, const double *m4, int64_t ndeg, const double *dwb,
                   double *yba, double *nna, int group);
void launch_scatter_pixels(hipStream_t s, int nmap, int64_t npix_cross, int64_t ndeg, const int64_t *off,
                           const CsrEnt *ent, const double *pw, const double *ccinv6, double *m4);
void launch_gather_baselines(hipStream_t s, int nmap, int ndet, int64_t nsamp, int64_t nbase,
                             const int64_t *bstart, const PixW *pixw, const double *m4, int64_t ndeg,
                             const double *dwb, const double *nna, const double *p, double *ap, int group);
// rec: the pointing records in layout rl (PixW for REC16); T as launch_scatter_seg
template <typename T>
void launch_gather_flat(hipStream_t s, int nmap, int ndet, int64_t nsamp, int64_t nbase, const int64_t *bstart,
                        const void *rec, int rl, const T *m4, int64_t ndeg, const T *dwb, const T *nna, const T *p,
                        T *ap);
void launch_scatter_flat(hipStream_t s, int nmap, int64_t npix_cross, int64_t ndeg, const int64_t *off,
                         const CsrEnt *ent, const double *pw, const double *ccinv6, double *m4);
// entry-balanced pixel-side projection: tinfo from launch_tile_info (once per
// CSR), scratch of scatter_tiles_scratch_bytes for the cross-tile partial sums
void launch_scatter_tiles(hipStream_t s, int nmap, int64_t npix_cross, int64_t ndeg, const int64_t *off, int64_t nent,
                          const CsrEnt *ent, const double *pw, const double *ccinv6, double *m4, const void *tinfo,
                          void *scratch);
size_t scatter_tiles_scratch_bytes(int64_t nent);
size_t scatter_tiles_info_bytes(int64_t nent);
void launch_tile_info(hipStream_t s, int64_t npix_cross, int64_t ndeg, const int64_t *off, int64_t nent, void *tinfo);
// segmented tile reductions (one wave per 512-entry tile, no LDS)
int64_t segred_ntiles(int64_t nent);
size_t segred_heads_bytes(int64_t nent);
void segred_build(hipStream_t s, const int64_t *segstart, int64_t nseg, int64_t nent, uint8_t *heads,
                  int64_t *tile_k0, int64_t *fix_end, int64_t *fix_ord);
void segred_pixel_segments(hipStream_t s, const int64_t *off, int64_t npc, int64_t ndeg, int32_t *flag, int64_t *pos,
                           int64_t *scan_tmp, int64_t nchunks, int32_t *nzq, int64_t *segstart);
void segred_baseline_segments(hipStream_t s, const int64_t *bstart, int64_t nbase, int ndet, int64_t nsamp,
                              int64_t *segstart);
// ent: the CSR entries in layout rl (CsrEnt for REC16); T = double, or float
// for the mixed-precision operator (pw, cc_c, m4 and the partial sums in T,
// part: 2 per tile of VecT<T>::v4, at most the bytes of the double4 ones)
template <typename T>
void launch_scatter_seg(hipStream_t s, int nmap, int64_t nent, const void *ent, int rl, const T *pw,
                        const uint8_t *heads, const int64_t *tile_k0, const int64_t *fix_end, const int64_t *fix_ord,
                        const int32_t *nzq, int64_t nzpix, const T *cc_c, T *m4, void *part);
// cc_c[k] = ccinv6[nzq[k]]: the inverse pixel matrices in the pixel side's ordinal order
void segred_compact_cc(hipStream_t s, const int32_t *nzq, int64_t nseg, const double *ccinv6, double *cc_c);
void launch_gather_seg(hipStream_t s, int nmap, int64_t nent, const PixW *pixw, const double *m4, int64_t ndeg,
                       const uint8_t *heads, const int64_t *tile_k0, const int64_t *fix_end, const int64_t *fix_ord,
                       const double *dwb, const double *nna, const double *p, double *ap, double4 *part);
// out[i] = in[ord[i]]
void launch_gather_order(hipStream_t s, const double *in, const int32_t *ord, double *out, int64_t n);
// pw[i] = p[ord[i]] * dwb_s[i]: the pixel-side input in slot order
void launch_pw_slots(hipStream_t s, const double *p, const int32_t *ord, const double *dwb_s, double *pw, int64_t n);
void launch_subtract_baselines(hipStream_t s, int nmap, int64_t npix_map, const int64_t *off, const CsrEnt *ent,
                               const double *aa, const double *dwb, double *rhs4);
void launch_clean_tod(hipStream_t s, int ndet, int64_t nsamp, int64_t nbase, const int64_t *bstart,
                      const double *aa, double *tod);
void launch_set_segments_mask(hipStream_t s, const Segment *seg, int64_t nseg, int8_t *rmask);
void launch_find_segments(hipStream_t s, const double *nna, int ndet, int64_t nbase, const int64_t *kstart_pp,
                          const int64_t *noba_pp, int ninterval, int64_t gapmin, Segment *out,
                          unsigned long long *nseg, int64_t cap);
void launch_fill(hipStream_t s, double *p, double v, int64_t n);
void launch_bits_differ(hipStream_t s, const double *a, const double *b, int64_t n, int *flag);
void launch_copy(hipStream_t s, double *dst, const double *src, int64_t n);
// elementwise: out = a * x
void launch_diag_precond(hipStream_t s, const double *r, const double *pd, double *z, int64_t n);
void launch_prec_diag_build(hipStream_t s, const double *nna, const double *wb_inv, const int8_t *haspsd,
                            double *pd, int64_t n);
// masked dot products; partial sums per block -> finalize
void launch_dots(hipStream_t s, int which, const double *a, const double *b, const double *c, const double *d,
                 const int8_t *mask, int64_t n, double *partials, int nblocks);
int dots_nblocks(int64_t n);
void launch_finalize(hipStream_t s, const double *partials, int nblocks, int nval, double *out);
// launch_finalize of the three r/z dot products into scal[2..4], then, in the
// same launch, beta = (scal[2] - scal[3]) / scal[0] into scal[5], the old rz
// into scal[6] and the new rz into scal[0]
void launch_finalize_scalars(hipStream_t s, const double *partials, int nblocks, double *scal);
// PCG vector steps, scalars read from device memory
void launch_pcg_update_r(hipStream_t s, const double *scal, double *aa, double *r, double *ro, const double *p,
                         const double *ap, int64_t n);
void launch_pcg_update_p(hipStream_t s, const double *scal, double *p, const double *z, int64_t n);
void launch_axpby_nna(hipStream_t s, double *ap, const double *nna, const double *p, double factor, int64_t n);

void launch_minmax(hipStream_t s, const int32_t *pix, int64_t n, int *mn, int *mx);
void launch_restore_pixels(hipStream_t s, int32_t *pix, int64_t n, int32_t dummy);
// pix == dummy -> -1
void launch_dummy_to_flag(hipStream_t s, int32_t *pix, int64_t n, int32_t dummy);
// *flag = 1 if any a[i] != b[i] (flag must start at 0)
void launch_words_differ(hipStream_t s, const uint32_t *a, const uint32_t *b, int64_t n, int *flag);
void launch_fprecond(hipStream_t s, const double *fcov, int64_t nh, const double *wpsd, int nshort, double *out,
                     int64_t npsd);

// inner-CG preconditioner pieces (segments x per-segment scalars; act[s] != 0 = active)
void launch_cg_init(hipStream_t s, const Segment *seg, int64_t nseg, const int *act, const double *ic0, const double *r,
                    const double *nna, double *x, double *resid, double *prec, double *zr, double *p);
void launch_cg_dot(hipStream_t s, const Segment *seg, int64_t nseg, const int *act, const double *a, const double *b,
                   double *partial, double *out);
void launch_cg_nna(hipStream_t s, const Segment *seg, int64_t nseg, const int *act, const double *nna, const double *p,
                   double *ap);
void launch_cg_axpy(hipStream_t s, const Segment *seg, int64_t nseg, const int *act, const double *alpha, double *x,
                    const double *p, double *resid, const double *ap);
void launch_cg_prec(hipStream_t s, const Segment *seg, int64_t nseg, const int *act, const double *prec,
                    const double *resid, double *zr);
void launch_cg_p(hipStream_t s, const Segment *seg, int64_t nseg, const int *act, const double *beta, const double *zr,
                 double *p);

// ------------------------------------------------------------------ fft.hip

// Twiddle factors, all from one allocation and small enough to stay cached:
//   sub1[j] = W_L1^j (j < L1 / 2), sub2[j] = W_L2^j (j < L2 / 2) for the LDS sub-FFTs;
//   W_N^m = na[m % 1024] * nb[m / 1024] and W_nof^k = fa[k % 1024] * fb[k / 1024]
//   (two-level tables; W_x = exp(-2 pi i / x)).
struct FftPlan {
    int64_t N;       // complex length (= nof / 2)
    int64_t L1, L2;  // N = L1 * L2, column pass length L1, row pass length L2
    double *tw_mem;
    const double2 *sub1, *sub2, *na, *nb, *fa, *fb;
    float *tw32_mem;   // the same tables rounded to FP32 (the FP32-work convolution)
    const float2 *sub1f, *sub2f, *naf, *nbf, *faf, *fbf;
};

void fft_plan_init(FftPlan &p, int64_t nof, hipStream_t s);
void fft_plan_free(FftPlan &p);
// convolve every segment: y = x outside segments is the caller's job (copy first)
// fc: filter spectra, nrow x fc_stride; fc_rows: the same in row order (fft_spectrum_rows)
// T (float or double) is the type of x and y; the transforms run in FP64,
// except for T = float with fc_rows32 (the row-order spectra in FP32) given:
// then the fused passes work in FP32 (the operator of the mixed-precision
// inner solve; segment means stay FP64)
template <typename T>
void fft_convolve_segments(hipStream_t s, const FftPlan &p, const Segment *dseg, int64_t nseg_dev, const T *x, T *y,
                           const double *fc, const double *fc_rows, int64_t fc_stride, double *work, int64_t batch,
                           int64_t nseg_max, const Segment *hseg, const float *fc_rows32 = nullptr);
// fT = fc reordered for the fused row pass: f[r + L1 k2] at [r * L2 + k2], f[N] at [N]
void fft_spectrum_rows(hipStream_t s, const FftPlan &p, const double *fc, int64_t nrow, int64_t stride, double *fT);
// nseg_max: the most segments one fft_convolve_segments call takes
size_t fft_convolve_work_doubles(const FftPlan &p, int64_t batch, int64_t nseg_max);
// inverse real FFT of fcov rows (invcov = c2r(fcov) / nof)
void fft_c2r_rows(hipStream_t s, const FftPlan &p, const double *fc, int64_t nrows, double *out, double *work);
// noise filter (build_filter) on the device
void launch_build_filter(hipStream_t s, int64_t nof, double fsample, int nshort, int radiometers, double sigma,
                         double plateau, int64_t n_in, int64_t startbin, const double *tab, double *fcov, int *bad);

// ------------------------------------------------------------------ band.hip

struct BandPrec {
    int64_t nseg = 0;
    int kd = 0;             // max bandwidth - 1
    int chunk = 0;          // rows per partition
    int64_t nchunk = 0;     // total partitions over all segments
    double *L = nullptr;    // factor rows while building (freed once re-laid out)
    double *Tf = nullptr;   // forward carry response [chunk][kd][kd]
    double *Tb = nullptr;   // backward carry response
    void *impl = nullptr;   // chunk tables, interleaved factor, carry-chain groups
    // partitions (and 64-partition blocks) that are not bitwise repeats of the previous one
    int64_t nchunk_unique = 0, nblock = 0, nblock_unique = 0;
    int64_t nrows = 0;      // baseline rows the factored segments cover
    int64_t nrows_factored = 0;   // of those, rows the column loop computed (the rest repeat a block)
    // carry chains: -1 several chains per wave when there are enough of them,
    // 1 always, 0 never (one chain per wave); the same results either way
    int chain_pack = -1;
};

// kernel_variant bits 32 / 64: carry chains always / never packed several per wave
inline int band_chain_pack(int kernel_variant)
{
    return (kernel_variant & 32) ? 1 : ((kernel_variant & 64) ? 0 : -1);
}

int band_build(hipStream_t s, BandPrec &bp, const Segment *hseg, const Segment *dseg, int64_t nseg,
               const int *nband_seg_host, const double *invcov_dev, int64_t nof, const double *nna, int *info_out,
               bool periodic = true);
// z = M^-1 r for nrhs right-hand sides at a stride of vstride elements (T =
// float or double; the factor and the solves are FP64)
template <typename T>
void band_apply(hipStream_t s, const BandPrec &bp, const T *r, T *z, int nrhs = 1, int64_t vstride = 0);
// make room for nrhs right-hand sides in band_apply (allocates; call outside graph capture)
void band_reserve(const BandPrec &bp, int nrhs);
void band_free(BandPrec &bp);

// ------------------------------------------------------------------ pointing.hip

struct ScanParams {
    double fsample, spin_period, prec_period, opening_turns, lon0_turns;
};
struct ScanDet {
    double psi_turns, open_offset_turns, spin_offset_turns;
};
void launch_expand_pointing(hipStream_t s, ScanParams sp, const ScanDet *dets, int ndet, int64_t nside,
                            int64_t i0, int64_t n, int64_t ld, int32_t *pix, float *w);

// the message madam_last_error() returns (madam_host.cpp)
void set_last_error(const char *msg);

}  // namespace madam
