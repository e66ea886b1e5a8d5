// Host-side state of one destriping configuration on one GPU: the parsed
// parameters (src/inputparam.f90:24-371, defaults src/commonparam.f90:200-305),
// the detectors and their PSDs (src/inputparam.f90:73-163), and every device
// buffer of the set-up and the solves.  Shared by the host orchestration
// (madam_host.cpp) and the batched, mixed-precision solves
// (madam_batch_host.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <string>
#include <vector>

#include "madam_batch.hpp"
#include "madam_internal.hpp"

#define F32(x) ((double)(float)(x))

namespace madam {

struct Error {
    std::string msg;
};

struct Params {
    int info = 2, nthreads = 1, nsubchunk = 0, isubchunk = 0;
    double dnshort = -1;
    int nside_map = 512, nside_cross = -1, nside_submap = 16;
    double good_baseline_fraction = 0;
    int pixmode_map = 2, pixmode_cross = 2;
    double pixlim_map = F32(1e-6), pixlim_cross = F32(1e-3);
    int incomplete_matrices = 0, allow_decoupling = 0, kfirst = 1, basis_order = 0;
    int iter_min = 3, iter_max = 1000;
    double cglimit = F32(1e-12), fsample = 1;
    int mode_detweight = 0, temperature_only = 0, noise_weights_from_psd = 0, radiometers = 1;
    int kfilter = 0;
    double diagfilter = 0;
    int precond_width_min = 10, precond_width_max = 100, use_fprecond = 0, use_cgprecond = 0;
    int rm_monopole = 0, write_tod = 0, ndetset = 0, nsurvey = 0, mcmode = 0;
    bool has_inmask = false, has_spectrum = false, has_covmat = false;
    int nshort = 0;
    // products (src/commonparam.f90:200-305, src/inputparam.f90:310-360)
    std::string path_output, file_root = "madam", unit_tod = "unknown";
    int write_map = 1, write_binmap = 0, write_hits = 0, write_mask = 0, write_matrix = 0, write_wcov = 0;
    int write_cut = 0, bin_subsets = 0, force_pol = 0, skip_existing = 0, write_leakmatrix = 0;
    struct DetSet {
        std::string name;
        std::vector<std::string> dets;
        bool nopol = false;
    };
    struct Survey {
        std::string name;
        std::vector<double> starts, stops;
    };
    std::vector<DetSet> detsets;     // detsets(1..ndetset)
    std::vector<Survey> surveys;     // surveys(1..nsurvey)
    std::string detset0, survey0;    // names of the full detector set / full data span
};

struct Detector {
    int npsd = 1;
    bool has_psd = false;
    std::vector<double> weights, sigmas, psdstarts, plateaus, psds;   // psds [npsd][npsdbin]
};

template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    void alloc(size_t count)
    {
        if (count <= n && p) return;
        release();
        MADAM_CHECK(hipMalloc(&p, sizeof(T) * (count > 0 ? count : 1)));
        n = count;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DBuf() { release(); }
};

// The samples sorted by map pixel (madam_csr.hip): keys (pixel, or npix_map
// for flagged samples, which sort last) and sample indices, ascending in
// sample index inside each pixel; the sorted pairs are k[cur], v[cur].
struct CsrSort {
    DBuf<uint32_t> k[2], v[2];
    int cur = 0;
    int64_t n = 0, npix_map = 0;
};
// buffers for n samples; the caller writes the keys of sample g to k[0][g]
void csr_alloc(CsrSort &cs, int64_t n, int64_t npix_map);
// sorts the samples by key
void csr_sort(hipStream_t s, CsrSort &cs);
// hits[p] for p < npix_map: the number of unflagged samples of pixel p
void csr_hits(hipStream_t s, const CsrSort &cs, int32_t *hits);
// the first `total` (unflagged) sorted samples as CSR entries: the record with
// its baseline's slot (bslot may be null: the baseline index), the sample
// index, and for rl != REC16 the compact copy of the entries in layout rl
void csr_fill(hipStream_t s, const CsrSort &cs, const PixW *pixw, int64_t nsamp, const int64_t *bstart, int64_t nbase,
              int64_t total, const int32_t *bslot, int rl, CsrEnt *ent, uint32_t *src, void *entc);
// locality order of the baselines: bord[slot] = baseline, bslot[baseline] = slot
void baseline_slots(hipStream_t s, const PixW *pixw, int ndet, int64_t nsamp, const int64_t *bstart, int64_t nbase,
                    int64_t npix_map, int32_t *bord, int32_t *bslot);

enum { ST_NITER, ST_RRINIT, ST_RZINIT, ST_T_POINTING, ST_T_BIN, ST_T_INIT, ST_T_PRECOND, ST_T_ITER, ST_T_FINAL,
       ST_T_SCATTER, ST_T_GATHER, ST_T_PRIOR, ST_T_APPLY_PREC, ST_N_SCATTER, ST_N_GATHER, ST_PRECOND_KIND,
       ST_NBAND, ST_NSEG, ST_NOF, ST_RR_FINAL, ST_NPIX_HIT, ST_NACTIVE, ST_T_TOTAL, ST_T_STEP, ST_NCG,
       ST_BAND_CHUNKS, ST_BAND_CHUNKS_UNIQUE, ST_BAND_BLOCKS, ST_BAND_BLOCKS_UNIQUE, ST_REC_BYTES,
       ST_BAND_ROWS_FACTORED, ST_PREC_REUSED, ST_POINTING_REUSED, ST_NWIDE, ST_COUNT };

struct Context;
struct BatchState;
void batch_release(Context &c);
void batch_invalidate(Context &c);

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    bool configured = false, pointed = false, solved = false;
    Params p;
    int ndet = 0, nnz = 0, nmap = 0, nside_max = 0;
    int64_t nsamp = 0, npix_map = 0, npix_cross = 0, ndeg = 1;
    int pass = 0;
    std::vector<Detector> det;
    std::vector<std::string> det_names;
    std::vector<double> timestamps;   // host copy, for the survey spans of the subset products
    bool subsets_enabled = false;     // destripe(): bin the detset / survey subsets inside the solve
    std::string subsets_path;
    std::vector<double> psdfreqs;
    // intervals
    int nival = 0;
    std::vector<int64_t> ivs;          // [nival + 1] starts
    std::vector<int64_t> noba_pp, kstart_pp;
    int64_t nbase = 0, nbtot = 0, noba_pp_max = 0;
    std::vector<int64_t> bstart;       // [nbase + 1]
    std::vector<double> btime;
    std::vector<double> dw_ival;       // [ndet][nival]
    std::vector<int> ipsd_det_ival, ipsd_glob_ival;
    // filter
    int64_t nof = 0;
    int npsdtot = 0;
    std::vector<int> psddet;
    std::vector<double> psdstart, psdstop, psd_w;   // per global PSD
    // device: plan
    DBuf<int64_t> d_ivs, d_bstart, d_kstart_pp, d_noba_pp;
    DBuf<double> d_dw_ival, d_dwb, d_sdwb, d_wbinv;
    DBuf<int8_t> d_haspsd;
    // device: pointing
    DBuf<int32_t> d_pix_own;
    DBuf<float> d_w_own;
    int32_t *d_pix = nullptr;
    float *d_w = nullptr;
    DBuf<PixW> d_pixw;
    DBuf<int32_t> d_hits;
    DBuf<int64_t> d_off, d_scan_tmp;
    DBuf<CsrEnt> d_ent;
    // locality order of the baselines (the CSR entries carry slot numbers):
    // bord[slot] = baseline, bslot[baseline] = slot, and the baseline vectors
    // the entries index, in slot order
    DBuf<int32_t> d_bord, d_bslot;
    // the projections' pointing records in the compact layout rec_layout
    // (REC12 / REC4), or unused (REC16: d_pixw and d_ent themselves)
    int rec_layout = 0;
    DBuf<char> d_pixc, d_entc;
    DBuf<double> d_dwb_s, d_sdwb_s, d_aa_s;
    DBuf<uint32_t> d_src;
    int64_t nactive = 0;
    bool pix_host_dirty = false;
    // device: maps
    DBuf<double> d_cc6, d_rhs4, d_rhsd4, d_cc6x, d_rhs4x, d_ccinv6x, d_m4x, d_ccinv6, d_crit;
    DBuf<int32_t> d_mask;
    DBuf<double> d_map4, d_binmap4;
    // device: baselines
    DBuf<double> d_tod_own;
    double *d_tod = nullptr;
    DBuf<double> d_yba, d_nna, d_aa, d_r, d_ro, d_z, d_p, d_pw, d_ap, d_prec, d_scal, d_part;
    DBuf<int8_t> d_rmask;
    DBuf<char> d_tilescratch, d_tileinfo;
    // segmented-tile projections: pixel side (s) and baseline side (g)
    struct SegTiles {
        DBuf<uint8_t> heads;
        DBuf<int64_t> k0, fix_end, fix_ord;
        int64_t nent = 0;
        bool ok = false;
    } sr_s, sr_g;
    DBuf<int32_t> d_nzq;
    DBuf<int64_t> d_segstart_s;   // CSR start of each non-empty cross pixel (pixel-side ordinal order)
    DBuf<double> d_ccinv_c;   // inverse cross-pixel matrices in the pixel side's ordinal order
    int64_t nzpix = 0;
    bool tiles_built = false;
    DBuf<double4> d_sr_part;
    // segments / prior
    std::vector<Segment> seg;
    DBuf<Segment> d_seg;
    std::vector<Segment> segp;         // seg with the preconditioner's PSD index
    DBuf<Segment> d_segp;
    std::vector<Segment> segconv;      // segments the prior convolves (PSD found)
    DBuf<Segment> d_segconv;
    // the convolved segments (the band and CG preconditioner segments) cover
    // every baseline, so the identity copy outside them is skipped
    bool prior_covers_all = false, prec_covers_all = false;
    DBuf<unsigned long long> d_cnt;
    DBuf<double> d_fcov, d_fprec, d_invcov, d_fftwork, d_tab;
    DBuf<double> d_fcov_rows, d_fprec_rows;   // the spectra in the fused row pass's order
    FftPlan fft{};
    bool fft_ready = false;
    int64_t fft_batch = 0, fft_nseg_max = 0;
    int precond_kind = 0;
    // inner-CG preconditioner segments (band width criterion or Cholesky failed)
    std::vector<Segment> segcg;
    std::vector<double> ic0cg;
    DBuf<Segment> d_segcg, d_segact;
    DBuf<double> d_cg_resid, d_cg_prec, d_cg_zr, d_cg_p, d_cg_ap, d_cg_part, d_cg_ic0, d_cg_sc;
    DBuf<int> d_cg_act;
    BandPrec band;
    bool band_ready = false;
    // segments, filters and preconditioner of the last solve, kept while the
    // configuration, pointing and kernel choice stay and reused when nna is
    // bitwise the same (they are functions of those alone)
    bool prec_cached = false;
    int prec_cached_variant = 0;
    DBuf<double> d_nna_key;
    DBuf<int> d_flag;
    // results
    double stats[ST_COUNT] = {0};
    std::vector<double> rr_hist;
    int profile = 0;
    int pcg_steps_done = 0;
    // projections: 0 segmented pixel side + LDS-tile baseline side, 1 wave per pixel / lane
    // group per baseline, 2 LDS tiles on both sides, 3 segmented tiles on both sides
    int kernel_variant = 0;

    // batched / mixed-precision solves (madam_batch_host.cpp)
    BatchState *batch = nullptr;

    ~Context()
    {
        batch_release(*this);
        if (band_ready) band_free(band);
        if (fft_ready) fft_plan_free(fft);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

inline double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename T>
inline void h2d(T *d, const T *h, size_t n, hipStream_t s)
{
    if (n) MADAM_CHECK(hipMemcpyAsync(d, h, sizeof(T) * n, hipMemcpyHostToDevice, s));
}

template <typename T>
inline void d2h(T *h, const T *d, size_t n, hipStream_t s)
{
    if (n) MADAM_CHECK(hipMemcpyAsync(h, d, sizeof(T) * n, hipMemcpyDeviceToHost, s));
}

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    Timer()
    {
        MADAM_CHECK(hipEventCreate(&a));
        MADAM_CHECK(hipEventCreate(&b));
    }
    ~Timer()
    {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    void start(hipStream_t s) { MADAM_CHECK(hipEventRecord(a, s)); }
    void stop(hipStream_t s) { MADAM_CHECK(hipEventRecord(b, s)); }
    double ms()
    {
        MADAM_CHECK(hipEventSynchronize(b));
        float v = 0;
        MADAM_CHECK(hipEventElapsedTime(&v, a, b));
        return v;
    }
};

// error with a message for madam_last_error(); the C-ABI entries catch it
[[noreturn]] void fail(const char *fmt, ...);

// ------------------------------------------------------------------ host functions shared across files

void ensure_stream(Context &c);
int group_for(const Context &c);
// binning, inversion, projection tables, initialize_a and (unless nna is
// bitwise that of the cached set-up) segments, filters and preconditioner of
// one solve (src/smadam.F90:316-400), for a TOD of type TT on the device
template <typename TT>
void solve_setup(Context &c, const TT *tod);
void cg_apply(Context &c, const double *r, double *z);

// batched and mixed-precision solves (madam_batch_host.cpp)
void solve_batch(Context &c, int nrhs, const void *signals, int tod_f32, int on_device, int precision);
void batch_get_maps(Context &c, int irhs, double *map, double *binmap);
void batch_get_baselines(Context &c, int irhs, double *aa, double *yba);
void batch_get_info(Context &c, int irhs, double *out, int n);
void batch_iterate(Context &c, int niter);
void batch_reset_timers(Context &c);

}  // namespace madam
