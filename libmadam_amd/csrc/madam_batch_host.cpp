// Batched and mixed-precision destriping solves on one context.
//
// A batch is nrhs right-hand sides that share the context's pointing: Monte
// Carlo realisations of one scan, the reference's destripe_with_cache cycle
// (src/smadam.F90:593-918) done for many signals at once.  The set-up of the
// first signal (pixel matrices, projection tables, nna, segments, filters,
// preconditioner) serves all of them; each signal adds its own binned map and
// right-hand side yba (src/smadam_routines.f90:252-671).
//
// precision 0: one PCG over all right-hand sides in FP64, the recurrence of
//   src/smadam_routines.f90:677-924 per right-hand side, each with its own
//   scalars and exit tests (a finished right-hand side is masked out).
// precision 1: the same PCG with an FP32 operator and FP32 vectors (FP64 dot
//   products and scalars) inside an FP64 iterative refinement:
//   r = yba - A64 a; solve A32 d = r to the reduction the tolerance still
//   needs; a += d; until rr / rrinit < cglimit on the FP64 residual.
//
// One PCG step is the same sequence of launches every time, so it is
// captured once into a HIP graph and replayed (unless the inner-CG
// preconditioner fallback, which synchronises with the host, is in use).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>
#include <vector>

#include "madam_context.hpp"

namespace madam {

struct BatchState {
    int nrhs = 0, R = 0, prec = 0;
    // pixel-side work items (rebuilt when the pointing changes)
    DBuf<BsItem> items;
    DBuf<BsFix> fix;
    BsTables bt;
    int piece = 0;            // entries per work item (the autotuner's choice)
    bool tables_ok = false;
    const void *tables_ent = nullptr;   // the CSR the tables were built for
    std::vector<int64_t> item_order;    // pixels in first-hit order (build_items)
    const void *order_ent = nullptr;
    // run-length records of the FP32 operator (build_rle)
    bool rle_ok = false;
    DBuf<PixW> rg_rec, rs_rec;
    DBuf<int64_t> rg_st, rg_base, rs_seg;
    DBuf<BsItem> items_c;
    DBuf<BsFix> fix_c;
    BsTables bt_c;
    int64_t n_rg = 0, n_rs = 0;
    // FP32 copies of the shared per-baseline / per-pixel terms
    DBuf<float> dwb_s32, dwb32, nna32, cc_c32;
    DBuf<float> m4f;   // the one-RHS FP32 map, 4 per pixel (the FP64 one is the context's m4x)
    DBuf<float> fcov_rows32, fprec_rows32;   // the filter spectra in row order, FP32 (empty: FP64 transforms)
    // FP64 per right-hand side (RHS-major)
    DBuf<double> yba, a64, r64, ap64, rhs4, map4, binmap4;
    DBuf<double> pw64, m64, part64;
    // inner solve vectors (float or double by precision)
    DBuf<char> x, r, ro, z, p, ap, pw, m, part;
    DBuf<double> dots;
    DBuf<BatchScal> sc, sc64;
    DBuf<int> active, active64, was_active, ones;
    // the prior's segments repeated for every right-hand side
    std::vector<Segment> segR;
    DBuf<Segment> d_segR;
    DBuf<double> fftwork;
    int64_t fft_nseg = 0;
    int64_t fft_batch32 = 1;   // segments per FFT batch of the FP32 convolution
    // per right-hand side results
    std::vector<BatchScal> hsc, hsc64;
    std::vector<int> inner_steps;
    // the captured PCG step, per vector type
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    hipGraph_t graph[2] = {nullptr, nullptr};
    // the same step without exit tests, for madam_iterate_batch
    hipGraphExec_t gexec_free = nullptr;
    hipGraph_t graph_free = nullptr;
    bool use_graph = true;
    double t_setup = 0, t_solve = 0, t_inner = 0;
    int64_t n_inner = 0, n_outer = 0;
    // profiling (Context::profile): HIP-event times of the parts of each PCG step, summed
    double t_scatter = 0, t_prior = 0, t_gather = 0, t_prec = 0, t_step = 0;
    int64_t n_prof = 0;

    void drop_graphs()
    {
        for (int i = 0; i < 2; ++i) {
            if (gexec[i]) (void)hipGraphExecDestroy(gexec[i]);
            if (graph[i]) (void)hipGraphDestroy(graph[i]);
            gexec[i] = nullptr;
            graph[i] = nullptr;
        }
        if (gexec_free) (void)hipGraphExecDestroy(gexec_free);
        if (graph_free) (void)hipGraphDestroy(graph_free);
        gexec_free = nullptr;
        graph_free = nullptr;
    }
    ~BatchState() { drop_graphs(); }
};

void batch_release(Context &c)
{
    delete c.batch;
    c.batch = nullptr;
}

// a new pointing: the work items and the captured steps refer to the old one
void batch_invalidate(Context &c)
{
    if (!c.batch) return;
    c.batch->tables_ok = false;
    c.batch->rle_ok = false;
    c.batch->item_order.clear();
    c.batch->drop_graphs();
    c.batch->nrhs = 0;
}

template <typename T>
static T *tp(DBuf<char> &b)
{
    return (T *)b.p;
}

// ------------------------------------------------------------------ pixel-side work items

static const void *ent_records(const Context &c);

// Each non-empty cross pixel is one item, or, beyond `piece` entries, pieces
// of `piece` entries whose sums the fix-up pass adds in order.
//
// The items are listed in the order in which the scan first reaches their
// pixels (the earliest baseline among a pixel's entries, ties by pixel): the
// waves that run together then read the baseline slots of one stretch of the
// scan, which the L2 holds.  The order of the items changes no sum.
// item order (first-hit) of the non-empty cross pixels, kept per pointing
static const std::vector<int64_t> &item_order(Context &c, BatchState &B)
{
    int64_t nz = c.nzpix;
    std::vector<int64_t> &order = B.item_order;
    if ((int64_t)order.size() != nz || B.order_ent != ent_records(c)) {
        order.resize(nz);
        for (int64_t k = 0; k < nz; ++k) order[k] = k;
        const char *eo = getenv("MADAM_BS_ORDER");
        if (!(eo && atoi(eo) == 0) && nz > 0) {
            DBuf<int64_t> dkey;
            dkey.alloc(nz);
            batch_pixel_keys(c.stream, c.d_segstart_s.p, nz, ent_records(c), rec_bytes(c.rec_layout), c.d_bord.p,
                             c.nbase, dkey.p);
            MADAM_CHECK(hipGetLastError());
            std::vector<int64_t> key(nz);
            d2h(key.data(), dkey.p, nz, c.stream);
            MADAM_CHECK(hipStreamSynchronize(c.stream));
            std::stable_sort(order.begin(), order.end(), [&key](int64_t a, int64_t b) { return key[a] < key[b]; });
        }
        B.order_ent = ent_records(c);
    }
    return order;
}

// the work items over the pixel segments seg_dev[nz + 1] (entries or run-length records)
static void build_items_into(Context &c, BatchState &B, int piece, const int64_t *seg_dev, DBuf<BsItem> &ditems,
                             DBuf<BsFix> &dfix, BsTables &bt)
{
    int64_t nz = c.nzpix;
    std::vector<int64_t> st(nz + 1);
    d2h(st.data(), seg_dev, nz + 1, c.stream);
    const std::vector<int64_t> &order = item_order(c, B);
    MADAM_CHECK(hipStreamSynchronize(c.stream));
    std::vector<BsItem> items;
    std::vector<BsFix> fix;
    items.reserve(nz + nz / 8);
    int32_t nparts = 0;
    for (int64_t k : order) {
        int64_t e0 = st[k], n = st[k + 1] - st[k];
        if (n <= piece) {
            items.push_back(BsItem{e0, (int32_t)n, (int32_t)k, -1, 0});
            continue;
        }
        BsFix f{(int32_t)k, nparts, 0};
        for (int64_t a = 0; a < n; a += piece) {
            int32_t m = (int32_t)std::min<int64_t>(piece, n - a);
            items.push_back(BsItem{e0 + a, m, (int32_t)k, nparts++, 0});
            f.nparts++;
        }
        fix.push_back(f);
    }
    ditems.alloc(std::max<size_t>(items.size(), 1));
    dfix.alloc(std::max<size_t>(fix.size(), 1));
    h2d(ditems.p, items.data(), items.size(), c.stream);
    h2d(dfix.p, fix.data(), fix.size(), c.stream);
    MADAM_CHECK(hipStreamSynchronize(c.stream));
    bt.items = ditems.p;
    bt.nitems = (int64_t)items.size();
    bt.fix = dfix.p;
    bt.nfix = (int64_t)fix.size();
    bt.nparts = nparts;
}

static void build_items(Context &c, BatchState &B, int piece)
{
    build_items_into(c, B, piece, c.d_segstart_s.p, B.items, B.fix, B.bt);
    B.piece = piece;
    B.rle_ok = false;
}

// Run-length records of the FP32 operator's two projections (R >= 8): one
// record per run of samples of one baseline on one cross pixel, and per run of
// CSR entries of one pixel from one baseline slot, weights summed in stream
// order; the pixel side's items are rebuilt over them with the same cut.
// MADAM_RLE=0 keeps the sample records.
static void build_rle(Context &c, BatchState &B)
{
    hipStream_t s = c.stream;
    B.rle_ok = false;
    const char *er = getenv("MADAM_RLE");
    if (er && atoi(er) == 0) return;
    int dsh = 0;
    while ((1ll << dsh) < c.ndeg) ++dsh;
    int temperature = c.nmap == 1 ? 1 : 0;
    int64_t n = (int64_t)c.ndet * c.nsamp, nent = c.nactive;
    int64_t nmax = std::max(n, nent);
    DBuf<int32_t> flag;
    DBuf<int64_t> pos, tmp;
    DBuf<uint32_t> mask;
    flag.alloc(std::max<int64_t>(nmax, 1));
    pos.alloc(nmax + 1);
    tmp.alloc((nmax + 1 + 4095) / 4096 + 1);
    mask.alloc((std::max<int64_t>(c.nsamp, nent) + 31) / 32 + 1);
    // baseline side
    int64_t tot = 0;
    batch_rle_gather(s, c.d_pixw.p, c.ndet, c.nsamp, c.d_bstart.p, c.nbase, dsh, temperature, mask.p, flag.p, pos.p,
                     tmp.p, nullptr, nullptr, nullptr, true);
    d2h(&tot, pos.p + n, 1, s);
    MADAM_CHECK(hipStreamSynchronize(s));
    B.rg_rec.alloc(tot + REC_PAD);
    MADAM_CHECK(hipMemsetAsync(B.rg_rec.p + tot, 0, sizeof(PixW) * REC_PAD, s));
    B.rg_st.alloc((size_t)c.ndet * (c.nbase + 1));
    B.rg_base.alloc(std::max(c.ndet, 1));
    batch_rle_gather(s, c.d_pixw.p, c.ndet, c.nsamp, c.d_bstart.p, c.nbase, dsh, temperature, mask.p, flag.p, pos.p,
                     tmp.p, B.rg_rec.p, B.rg_st.p, B.rg_base.p, false);
    B.n_rg = tot;
    // pixel side
    batch_rle_scatter(s, (const PixW *)c.d_ent.p, nent, c.d_segstart_s.p, c.nzpix, temperature, mask.p, flag.p, pos.p,
                      tmp.p, nullptr, nullptr, true);
    d2h(&tot, pos.p + nent, 1, s);
    MADAM_CHECK(hipStreamSynchronize(s));
    B.rs_rec.alloc(tot + REC_PAD);
    MADAM_CHECK(hipMemsetAsync(B.rs_rec.p + tot, 0, sizeof(PixW) * REC_PAD, s));
    B.rs_seg.alloc(c.nzpix + 1);
    batch_rle_scatter(s, (const PixW *)c.d_ent.p, nent, c.d_segstart_s.p, c.nzpix, temperature, mask.p, flag.p,
                      pos.p, tmp.p, B.rs_rec.p, B.rs_seg.p, false);
    MADAM_CHECK(hipGetLastError());
    B.n_rs = tot;
    build_items_into(c, B, B.piece, B.rs_seg.p, B.items_c, B.fix_c, B.bt_c);
    B.rle_ok = true;
}

static const void *ent_records(const Context &c)
{
    return c.rec_layout == REC16 ? (const void *)c.d_ent.p : (const void *)c.d_entc.p;
}

static const void *sample_records(const Context &c)
{
    return c.rec_layout == REC16 ? (const void *)c.d_pixw.p : (const void *)c.d_pixc.p;
}

// ------------------------------------------------------------------ operator pieces

template <typename T>
struct Ops;   // the per-type buffers of the shared terms

template <>
struct Ops<double> {
    static const double *dwb_s(Context &c, BatchState &) { return c.d_dwb_s.p; }
    static const double *dwb(Context &c, BatchState &) { return c.d_dwb.p; }
    static const double *nna(Context &c, BatchState &) { return c.d_nna.p; }
    static const double *cc(Context &c, BatchState &) { return c.d_ccinv_c.p; }
    static double *m4(Context &c, BatchState &) { return c.d_m4x.p; }
    static const float *fcov32(BatchState &) { return nullptr; }
    static const float *fprec32(BatchState &) { return nullptr; }
};
template <>
struct Ops<float> {
    static const float *dwb_s(Context &, BatchState &B) { return B.dwb_s32.p; }
    static const float *dwb(Context &, BatchState &B) { return B.dwb32.p; }
    static const float *nna(Context &, BatchState &B) { return B.nna32.p; }
    static const float *cc(Context &, BatchState &B) { return B.cc_c32.p; }
    static float *m4(Context &, BatchState &B) { return B.m4f.p; }
    static const float *fcov32(BatchState &B) { return B.fcov_rows32.p; }
    static const float *fprec32(BatchState &B) { return B.fprec_rows32.p; }
};

// events around the parts of a PCG step when profiling
struct StepTimers {
    Timer sc, pr, ga, pc, st;
};

// y = A x for the R right-hand sides (src/smadam_routines.f90:808-872): pixel
// side with the ccmultiply, prior, baseline side
template <typename T>
static void batch_operator(Context &c, BatchState &B, const T *x, T *y, T *pw, T *m, T *part, const int *active,
                           StepTimers *tm = nullptr, bool pw_ready = false)
{
    hipStream_t s = c.stream;
    int R = B.R;
    int64_t nb = c.nbtot;
    // one right-hand side: the single solve's segmented-tile pixel side and
    // flat baseline side in the operator's type; several: the batched kernels
    bool one = R == 1;
    if (!pw_ready) batch_pw<T>(s, R, x, nb, c.d_bord.p, c.d_bslot.p, Ops<T>::dwb_s(c, B), Ops<T>::dwb(c, B), active, pw);
    if (tm) tm->sc.start(s);
    // the FP32 operator at R >= 8 reads the run-length records
    bool rle = std::is_same<T, float>::value && B.rle_ok && R >= 8;
    if (one)
        launch_scatter_seg<T>(s, c.nmap, c.nactive, ent_records(c), c.rec_layout, pw, c.sr_s.heads.p, c.sr_s.k0.p,
                              c.sr_s.fix_end.p, c.sr_s.fix_ord.p, c.d_nzq.p, c.nzpix, Ops<T>::cc(c, B),
                              Ops<T>::m4(c, B), c.d_sr_part.p);
    else if (rle)
        batch_scatter<T>(s, R, REC16, c.nmap, B.bt_c, B.rs_rec.p, pw, Ops<T>::cc(c, B), c.d_nzq.p, m, part);
    else
        batch_scatter<T>(s, R, c.rec_layout, c.nmap, B.bt, ent_records(c), pw, Ops<T>::cc(c, B), c.d_nzq.p, m, part);
    if (tm) {
        tm->sc.stop(s);
        tm->pr.start(s);
    }
    if (c.p.kfilter) {
        if (!c.prior_covers_all) batch_copy<T>(s, y, x, nb * R);
        if (!B.segR.empty())
            fft_convolve_segments<T>(s, c.fft, B.d_segR.p, (int64_t)B.segR.size(), x, y, c.d_fcov.p, c.d_fcov_rows.p,
                                     c.nof / 2 + 1, B.fftwork.p, Ops<T>::fcov32(B) ? B.fft_batch32 : c.fft_batch,
                                     B.fft_nseg, B.segR.data(), Ops<T>::fcov32(B));
    } else {
        batch_fill<T>(s, y, (T)0, nb * R);
    }
    if (tm) {
        tm->pr.stop(s);
        tm->ga.start(s);
    }
    if (one)
        launch_gather_flat<T>(s, c.nmap, c.ndet, c.nsamp, c.nbase, c.d_bstart.p, sample_records(c), c.rec_layout,
                              Ops<T>::m4(c, B), c.ndeg, Ops<T>::dwb(c, B), Ops<T>::nna(c, B), x, y);
    else if (rle)
        batch_gather<T>(s, R, REC16, c.nmap, c.ndet, c.nsamp, c.nbase, B.rg_st.p, B.rg_rec.p, m, 1,
                        Ops<T>::dwb(c, B), Ops<T>::nna(c, B), x, y, active, c.nbase + 1, B.rg_base.p);
    else
        batch_gather<T>(s, R, c.rec_layout, c.nmap, c.ndet, c.nsamp, c.nbase, c.d_bstart.p, sample_records(c), m,
                        c.ndeg, Ops<T>::dwb(c, B), Ops<T>::nna(c, B), x, y, active);
    if (tm) tm->ga.stop(s);
}

// z = M^-1 r for the R right-hand sides (src/noise_routines.f90:1159-1409)
template <typename T>
static void batch_precond(Context &c, BatchState &B, const T *r, T *z)
{
    hipStream_t s = c.stream;
    int R = B.R;
    int64_t nb = c.nbtot;
    switch (c.precond_kind) {
    case 0: batch_copy<T>(s, z, r, nb * R); break;
    case 1: batch_diag<T>(s, R, r, c.d_prec.p, z, nb); break;
    case 2:
        batch_copy<T>(s, z, r, nb * R);
        if (!B.segR.empty())
            fft_convolve_segments<T>(s, c.fft, B.d_segR.p, (int64_t)B.segR.size(), r, z, c.d_fprec.p,
                                     c.d_fprec_rows.p, c.nof / 2 + 1, B.fftwork.p,
                                     Ops<T>::fprec32(B) ? B.fft_batch32 : c.fft_batch, B.fft_nseg,
                                     B.segR.data(),
                                     Ops<T>::fprec32(B));
        break;
    default:
        if (!c.prec_covers_all) batch_copy<T>(s, z, r, nb * R);
        c.band.chain_pack = band_chain_pack(c.kernel_variant);
        band_apply<T>(s, c.band, r, z, R, nb);
        if (!c.segcg.empty()) {
            // the inner-CG fallback segments, one right-hand side at a time in FP64
            DBuf<double> r1, z1;
            r1.alloc(nb);
            z1.alloc(nb);
            for (int k = 0; k < B.nrhs; ++k) {
                batch_cast<T, double>(s, r + (int64_t)k * nb, r1.p, nb);
                batch_cast<T, double>(s, z + (int64_t)k * nb, z1.p, nb);
                cg_apply(c, r1.p, z1.p);
                batch_cast<double, T>(s, z1.p, z + (int64_t)k * nb, nb);
            }
        }
        break;
    }
}

// ------------------------------------------------------------------ one PCG over the batch

template <typename T>
struct Vec {
    T *x, *r, *ro, *z, *p, *ap, *pw, *m, *part;
};

// one PCG step (src/smadam_routines.f90:785-924) for every active right-hand side
template <typename T>
static void pcg_step(Context &c, BatchState &B, const Vec<T> &v, double lim, int iter_min, int iter_max, int stall,
                     StepTimers *tm = nullptr)
{
    hipStream_t s = c.stream;
    int R = B.R;
    int64_t nb = c.nbtot;
    if (tm) tm->st.start(s);
    // the previous step's p update (or the start) left pw ready; the FP32
    // inner solve keeps no r_old (beta from ap . z, stage 4)
    const bool fuse = true, lean = std::is_same<T, float>::value;
    batch_operator<T>(c, B, v.p, v.ap, v.pw, v.m, v.part, B.active.p, tm, fuse);
    batch_dots<T>(s, 0, R, v.p, v.ap, nullptr, c.d_rmask.p, nb, B.dots.p);
    batch_final(s, R, B.dots.p, 0, B.sc.p, B.active.p, lim, iter_min, iter_max);
    batch_update_r<T>(s, R, B.sc.p, B.active.p, v.x, v.r, lean ? nullptr : v.ro, v.p, v.ap, nb);
    if (tm) tm->pc.start(s);
    batch_precond<T>(c, B, v.r, v.z);
    if (tm) tm->pc.stop(s);
    batch_dots<T>(s, 1, R, v.r, v.z, lean ? v.ap : v.ro, c.d_rmask.p, nb, B.dots.p);
    batch_final(s, R, B.dots.p, lean ? 4 : 1, B.sc.p, B.active.p, lim, iter_min, iter_max, nullptr, stall);
    batch_update_p_pw<T>(s, R, B.sc.p, B.active.p, v.p, v.z, nb, c.d_bslot.p, Ops<T>::dwb(c, B), v.pw);
    if (tm) {
        tm->st.stop(s);
        B.t_scatter += tm->sc.ms() * 1e-3;
        B.t_prior += tm->pr.ms() * 1e-3;
        B.t_gather += tm->ga.ms() * 1e-3;
        B.t_prec += tm->pc.ms() * 1e-3;
        B.t_step += tm->st.ms() * 1e-3;
        B.n_prof++;
    }
}

static int count_active(Context &c, BatchState &B)
{
    std::vector<int> a(B.R);
    d2h(a.data(), B.active.p, B.R, c.stream);
    MADAM_CHECK(hipStreamSynchronize(c.stream));
    int n = 0;
    for (int v : a) n += v != 0;
    return n;
}

// x = A^-1 b on the right-hand sides allowed by `allow` (null = all): PCG from
// x = 0 until each right-hand side's exit test (lim < 0: its own BatchScal::lim)
template <typename T>
static int64_t batch_pcg(Context &c, BatchState &B, const Vec<T> &v, const T *b, const int *allow, double lim,
                         int iter_min, int iter_max, int stall)
{
    hipStream_t s = c.stream;
    int R = B.R;
    int64_t nb = c.nbtot;
    batch_copy<T>(s, v.r, b, nb * R);
    batch_fill<T>(s, v.x, (T)0, nb * R);
    batch_precond<T>(c, B, v.r, v.z);
    batch_copy<T>(s, v.p, v.z, nb * R);
    batch_dots<T>(s, 1, R, v.r, v.z, v.r, c.d_rmask.p, nb, B.dots.p);
    batch_final(s, R, B.dots.p, 2, B.sc.p, B.active.p, lim, iter_min, iter_max, allow);
    // the first step's pixel-side input (the steps leave it for the next one)
    batch_pw<T>(s, R, v.p, nb, c.d_bord.p, c.d_bslot.p, Ops<T>::dwb_s(c, B), Ops<T>::dwb(c, B), B.active.p, v.pw);
    int gi = std::is_same<T, float>::value ? 1 : 0;
    // profiling times every part with events, so the steps run as launched
    bool graph = B.use_graph && c.segcg.empty() && !c.profile;
    StepTimers tmr;
    int64_t steps = 0;
    double t0 = now();
    while (count_active(c, B) > 0) {
        if (c.profile) {
            pcg_step<T>(c, B, v, lim, iter_min, iter_max, stall, &tmr);
        } else if (graph) {
            if (!B.gexec[gi]) {
                MADAM_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                pcg_step<T>(c, B, v, lim, iter_min, iter_max, stall);
                MADAM_CHECK(hipStreamEndCapture(s, &B.graph[gi]));
                MADAM_CHECK(hipGraphInstantiate(&B.gexec[gi], B.graph[gi], nullptr, nullptr, 0));
            }
            MADAM_CHECK(hipGraphLaunch(B.gexec[gi], s));
        } else {
            pcg_step<T>(c, B, v, lim, iter_min, iter_max, stall);
        }
        MADAM_CHECK(hipGetLastError());
        steps++;
    }
    B.t_inner += now() - t0;
    B.n_inner += steps;
    return steps;
}

// ------------------------------------------------------------------ set-up and solve

template <typename T>
static Vec<T> alloc_vectors(Context &c, BatchState &B)
{
    int64_t n = c.nbtot * B.R;
    size_t e = sizeof(T);
    for (DBuf<char> *b : {&B.x, &B.r, &B.ro, &B.z, &B.p, &B.ap}) b->alloc(std::max<int64_t>(n, 1) * e);
    B.pw.alloc(std::max<int64_t>(n, 1) * e);
    size_t mb = (size_t)std::max<int64_t>(c.npix_cross, 1) * c.nmap * B.R * e;
    B.m.alloc(mb);
    // the rows of pixels without hits are never written: zero, so that the
    // baseline side may weight them by 0
    MADAM_CHECK(hipMemsetAsync(B.m.p, 0, mb, c.stream));
    B.part.alloc((size_t)std::max<int64_t>(B.bt.nparts, 1) * c.nmap * B.R * e);
    return Vec<T>{tp<T>(B.x), tp<T>(B.r), tp<T>(B.ro), tp<T>(B.z), tp<T>(B.p), tp<T>(B.ap),
                  tp<T>(B.pw), tp<T>(B.m), tp<T>(B.part)};
}

// Autotuner: entries per pixel-side work item.  Short items balance the
// long pixels (pole and caustic pixels hold thousands of entries) at the cost
// of more fix-up partial sums; the best cut depends on the scan and on R.  The
// candidates are timed on this pointing with HIP events and the fastest kept.
template <typename T>
static void tune_pieces(Context &c, BatchState &B, const Vec<T> &v)
{
    static const int cand[] = {256, 1024, 4096};
    hipStream_t s = c.stream;
    // MADAM_BS_PIECE pins the cut (the tuned cut can differ from run to run,
    // and with it the rounding of the long pixels' sums)
    const char *pe = getenv("MADAM_BS_PIECE");
    if (pe && atoi(pe) > 0) {
        build_items(c, B, atoi(pe));
        return;
    }
    hipEvent_t e0, e1;
    MADAM_CHECK(hipEventCreate(&e0));
    MADAM_CHECK(hipEventCreate(&e1));
    float best = 1e30f;
    int bestp = cand[1];
    for (int piece : cand) {
        build_items(c, B, piece);
        B.part.alloc((size_t)std::max<int64_t>(B.bt.nparts, 1) * c.nmap * B.R * sizeof(T));
        T *part = tp<T>(B.part);
        // a warm-up launch, then two timed
        batch_scatter<T>(s, B.R, c.rec_layout, c.nmap, B.bt, ent_records(c), v.pw, Ops<T>::cc(c, B), c.d_nzq.p, v.m,
                         part);
        MADAM_CHECK(hipEventRecord(e0, s));
        for (int k = 0; k < 2; ++k)
            batch_scatter<T>(s, B.R, c.rec_layout, c.nmap, B.bt, ent_records(c), v.pw, Ops<T>::cc(c, B), c.d_nzq.p,
                             v.m, part);
        MADAM_CHECK(hipEventRecord(e1, s));
        MADAM_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        MADAM_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            bestp = piece;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    build_items(c, B, bestp);
}

static void make_segments(Context &c, BatchState &B)
{
    B.segR.clear();
    int64_t ns = (int64_t)c.segconv.size();
    for (int r = 0; r < B.R; ++r)
        for (int64_t i = 0; i < ns; ++i) {
            Segment g = c.segconv[i];
            g.b0 += (int64_t)r * c.nbtot;
            B.segR.push_back(g);
        }
    if (!B.segR.empty()) {
        B.d_segR.alloc(B.segR.size());
        h2d(B.d_segR.p, B.segR.data(), B.segR.size(), c.stream);
    }
    B.fft_nseg = std::max<int64_t>((int64_t)B.segR.size(), 1);
    // the FP32 work buffer holds twice the segments of the FP64 one in the
    // same bytes (the batch stays resident in the last-level cache)
    B.fft_batch32 = std::min<int64_t>(std::min<int64_t>(2 * c.fft_batch, 64), B.fft_nseg);
    if (const char *e = getenv("MADAM_FFT_BATCH32")) B.fft_batch32 = std::max<int64_t>(std::min<int64_t>(atoi(e), 64), 1);
    if (c.p.kfilter)
        B.fftwork.alloc(fft_convolve_work_doubles(c.fft, std::max<int64_t>(c.fft_batch, B.fft_batch32), B.fft_nseg));
}

// binned map and right-hand side of one signal (src/smadam_routines.f90:252-671),
// with the shared pixel matrices of the set-up
template <typename TT>
static void bin_signal(Context &c, const TT *tod, double *rhs4, double *yba, DBuf<double> &cc6, DBuf<double> &rhs4x,
                       DBuf<double> &m4x, DBuf<double> &nna)
{
    hipStream_t s = c.stream;
    launch_bin<TT>(s, c.nmap, c.npix_map, c.d_off.p, c.d_ent.p, c.d_src.p, tod, c.d_dwb_s.p, c.d_sdwb_s.p, cc6.p,
                   rhs4);
    const double *rx = rhs4;
    if (c.ndeg > 1) {
        launch_degrade(s, c.nmap, c.npix_cross, c.ndeg, cc6.p, rhs4, cc6.p + c.npix_map * 6, rhs4x.p);
        rx = rhs4x.p;
    }
    launch_ccmultiply(s, c.nmap, c.npix_cross, c.d_ccinv6x.p, rx, m4x.p);
    launch_init_a<TT>(s, c.nmap, c.ndet, c.nsamp, c.nbase, c.d_bstart.p, c.d_pixw.p, tod, m4x.p, c.ndeg, c.d_dwb.p,
                      yba, nna.p, group_for(c));
}

template <typename TT>
static void solve_batch_t(Context &c, int nrhs, const TT *signals, int on_device, int precision)
{
    if (!c.pointed) fail("madam_solve_batch called before the pointing was set");
    if (!c.p.kfirst) fail("madam_solve_batch needs destriping (kfirst = T)");
    if (c.p.diagfilter != 0 && !c.p.kfilter) fail("madam_solve_batch does not take diagfilter");
    if (precision != 0 && precision != 1) fail("madam_solve_batch: precision must be 0 (FP64) or 1 (mixed)");
    int R = batch_width(nrhs);
    if (nrhs < 1 || R < 0) fail("madam_solve_batch: nrhs = %d outside 1..64", nrhs);
    hipStream_t s = c.stream;
    double tstart = now();
    int64_t n = (int64_t)c.ndet * c.nsamp, nb = c.nbtot, np = c.npix_map;
    if (!c.batch) c.batch = new BatchState();
    BatchState &B = *c.batch;
    if (B.R != R || B.prec != precision || !B.tables_ok || B.tables_ent != ent_records(c)) B.drop_graphs();
    B.nrhs = nrhs;
    B.prec = precision;
    B.t_inner = 0;
    B.n_inner = 0;
    B.n_outer = 0;
    B.t_scatter = B.t_prior = B.t_gather = B.t_prec = B.t_step = 0;
    B.n_prof = 0;
    const char *ge = getenv("MADAM_BATCH_GRAPH");
    B.use_graph = !(ge && ge[0] == '0');
    // the signals on the device, one at a time when they are in host memory
    DBuf<TT> tod1;
    auto signal = [&](int k) -> const TT * {
        if (on_device) return signals + (int64_t)k * n;
        tod1.alloc(std::max<int64_t>(n, 1));
        h2d(tod1.p, signals + (int64_t)k * n, n, s);
        return tod1.p;
    };
    // shared set-up with the first signal
    solve_setup<TT>(c, signal(0));
    bool rebuilt_prec = c.stats[ST_PREC_REUSED] == 0;
    if (rebuilt_prec || B.R != R) B.drop_graphs();
    B.R = R;
    B.yba.alloc((size_t)std::max<int64_t>(nb, 1) * R);
    B.rhs4.alloc((size_t)np * 4 * R);
    batch_fill<double>(s, B.yba.p, 0.0, nb * R);
    batch_copy<double>(s, B.yba.p, c.d_yba.p, nb);
    batch_copy<double>(s, B.rhs4.p, c.d_rhs4.p, np * 4);
    {
        DBuf<double> cc6, rhs4x, m4x, nna;
        cc6.alloc((size_t)(np + c.npix_cross) * 6);
        rhs4x.alloc((size_t)c.npix_cross * 4);
        m4x.alloc((size_t)c.npix_cross * 4);
        nna.alloc(std::max<int64_t>(nb, 1));
        for (int k = 1; k < nrhs; ++k)
            bin_signal<TT>(c, signal(k), B.rhs4.p + (int64_t)k * np * 4, B.yba.p + (int64_t)k * nb, cc6, rhs4x, m4x,
                           nna);
        MADAM_CHECK(hipStreamSynchronize(s));
    }
    if (c.precond_kind == 3) band_reserve(c.band, R);
    make_segments(c, B);
    // per-RHS state
    B.dots.alloc(batch_partials_doubles(R));
    B.sc.alloc(R);
    B.sc64.alloc(R);
    B.active.alloc(R);
    B.active64.alloc(R);
    B.was_active.alloc(R);
    B.ones.alloc(R);
    {
        std::vector<int> one(R, 0);
        for (int k = 0; k < nrhs; ++k) one[k] = 1;
        h2d(B.ones.p, one.data(), R, s);   // padding right-hand sides beyond nrhs stay inactive
    }
    B.a64.alloc((size_t)std::max<int64_t>(nb, 1) * R);
    const Params &p = c.p;
    if (precision == 0) {
        Vec<double> v = alloc_vectors<double>(c, B);
        if (R > 1 && (!B.tables_ok || B.tables_ent != ent_records(c))) {
            tune_pieces<double>(c, B, v);
            B.part.alloc((size_t)std::max<int64_t>(B.bt.nparts, 1) * c.nmap * R * sizeof(double));
            v.part = tp<double>(B.part);
            B.tables_ok = true;
            B.tables_ent = ent_records(c);
            B.drop_graphs();
        }
        v.x = B.a64.p;
        B.drop_graphs();   // the solution vector is bound into the captured step
        double t0 = now();
        batch_pcg<double>(c, B, v, B.yba.p, B.ones.p, p.cglimit, p.iter_min, p.iter_max, 0);
        B.hsc.resize(R);
        d2h(B.hsc.data(), B.sc.p, R, s);
        MADAM_CHECK(hipStreamSynchronize(s));
        B.n_outer = 1;
        B.t_solve = now() - t0;
        for (int k = 0; k < nrhs; ++k)
            if (B.hsc[k].status == 2) fail("CG is diverging (right-hand side %d)", k);
    } else {
        // FP32 copies of the shared terms
        B.dwb_s32.alloc(std::max<int64_t>(nb, 1));
        B.dwb32.alloc(std::max<int64_t>(nb, 1));
        B.nna32.alloc(std::max<int64_t>(nb, 1));
        B.cc_c32.alloc((size_t)std::max<int64_t>(c.nzpix, 1) * 6);
        batch_cast<double, float>(s, c.d_dwb_s.p, B.dwb_s32.p, nb);
        batch_cast<double, float>(s, c.d_dwb.p, B.dwb32.p, nb);
        batch_cast<double, float>(s, c.d_nna.p, B.nna32.p, nb);
        batch_cast<double, float>(s, c.d_ccinv_c.p, B.cc_c32.p, c.nzpix * 6);
        // the FP32 operator's noise prior and Fourier preconditioner transform in
        // FP32 (MADAM_FFT32=0 keeps them in FP64)
        const char *e32 = getenv("MADAM_FFT32");
        bool fft32 = !(e32 && atoi(e32) == 0);
        int64_t nspec = (int64_t)(c.nof / 2 + 1) * c.npsdtot;
        const float *old_f = B.fcov_rows32.p, *old_p = B.fprec_rows32.p;
        if (fft32 && c.p.kfilter && c.d_fcov_rows.p) {
            B.fcov_rows32.alloc((size_t)nspec);
            batch_cast<double, float>(s, c.d_fcov_rows.p, B.fcov_rows32.p, nspec);
        } else {
            B.fcov_rows32.release();
        }
        if (fft32 && c.precond_kind == 2 && c.d_fprec_rows.p) {
            B.fprec_rows32.alloc((size_t)nspec);
            batch_cast<double, float>(s, c.d_fprec_rows.p, B.fprec_rows32.p, nspec);
        } else {
            B.fprec_rows32.release();
        }
        if (B.fcov_rows32.p != old_f || B.fprec_rows32.p != old_p) B.drop_graphs();   // bound into the captured step
        Vec<float> v = alloc_vectors<float>(c, B);
        B.m4f.alloc((size_t)std::max<int64_t>(c.npix_cross, 1) * 4);
        if (R > 1 && (!B.tables_ok || B.tables_ent != ent_records(c))) {
            tune_pieces<float>(c, B, v);
            B.tables_ok = true;
            B.tables_ent = ent_records(c);
            B.drop_graphs();
        }
        if (R >= 8 && !B.rle_ok) {
            build_rle(c, B);
            B.drop_graphs();
        }
        int64_t nparts = std::max<int64_t>(B.bt.nparts, B.rle_ok ? B.bt_c.nparts : 0);
        B.part.alloc((size_t)std::max<int64_t>(nparts, 1) * c.nmap * R * sizeof(float));
        v.part = tp<float>(B.part);
        // the FP64 operator's own buffers (its pieces in the same tables)
        B.pw64.alloc((size_t)std::max<int64_t>(nb, 1) * R);
        B.m64.alloc((size_t)std::max<int64_t>(c.npix_cross, 1) * c.nmap * R);
        MADAM_CHECK(hipMemsetAsync(B.m64.p, 0, sizeof(double) * (size_t)std::max<int64_t>(c.npix_cross, 1) * c.nmap * R,
                                   s));
        B.part64.alloc((size_t)std::max<int64_t>(B.bt.nparts, 1) * c.nmap * R);
        B.r64.alloc((size_t)std::max<int64_t>(nb, 1) * R);
        B.ap64.alloc((size_t)std::max<int64_t>(nb, 1) * R);
        double t0 = now();
        batch_fill<double>(s, B.a64.p, 0.0, nb * R);
        batch_copy<double>(s, B.r64.p, B.yba.p, nb * R);
        // rrinit per right-hand side: outer stage 2 on r = yba (src/smadam_routines.f90:737-756)
        batch_dots<double>(s, 1, R, B.r64.p, B.r64.p, B.r64.p, c.d_rmask.p, nb, B.dots.p);
        batch_final(s, R, B.dots.p, 2, B.sc64.p, B.active64.p, p.cglimit, 0, 64, B.ones.p);
        const int max_outer = 20;
        float *r32 = tp<float>(B.ro);   // the inner right-hand side lives in the PCG's ro buffer until it starts
        DBuf<float> b32;
        b32.alloc((size_t)std::max<int64_t>(nb, 1) * R);
        (void)r32;
        for (int outer = 0; outer < max_outer; ++outer) {
            std::vector<int> a(R);
            d2h(a.data(), B.active64.p, R, s);
            MADAM_CHECK(hipStreamSynchronize(s));
            int na = 0;
            for (int k = 0; k < R; ++k) na += a[k] != 0;
            if (na == 0) break;
            // inner FP32 solve of A d = r to the reduction still needed
            batch_cast<double, float>(s, B.r64.p, b32.p, nb * R);
            batch_inner_limits(s, R, B.sc64.p, B.sc.p, p.cglimit, 1e-12);
            MADAM_CHECK(hipMemcpyAsync(B.was_active.p, B.active64.p, sizeof(int) * R, hipMemcpyDeviceToDevice, s));
            batch_pcg<float>(c, B, v, b32.p, B.was_active.p, -1.0, p.iter_min, p.iter_max, 50);
            // a += d, then the FP64 residual r = yba - A64 a
            batch_accumulate<float>(s, R, B.a64.p, v.x, B.was_active.p, nb);
            batch_operator<double>(c, B, B.a64.p, B.ap64.p, B.pw64.p, B.m64.p, B.part64.p, B.ones.p);
            batch_residual(s, B.yba.p, B.ap64.p, B.r64.p, nb * R);
            batch_dots<double>(s, 1, R, B.r64.p, B.r64.p, B.r64.p, c.d_rmask.p, nb, B.dots.p);
            batch_final(s, R, B.dots.p, 3, B.sc64.p, B.active64.p, p.cglimit, 0, max_outer);
            B.n_outer++;
        }
        B.hsc.resize(R);
        B.hsc64.resize(R);
        d2h(B.hsc.data(), B.sc.p, R, s);
        d2h(B.hsc64.data(), B.sc64.p, R, s);
        MADAM_CHECK(hipStreamSynchronize(s));
        B.t_solve = now() - t0;
    }
    // final maps per right-hand side (src/smadam_routines.f90:1489-1539, src/map_routines.f90:218-357)
    c.d_ccinv6.alloc(np * 6);
    c.d_mask.alloc(np);
    c.d_crit.alloc(np);
    MADAM_CHECK(hipMemsetAsync(c.d_crit.p, 0, sizeof(double) * np, s));
    launch_invert(s, c.nmap, np, c.d_cc6.p, c.d_ccinv6.p, c.d_crit.p, c.d_mask.p, p.pixmode_map, p.pixlim_map,
                  p.incomplete_matrices, p.allow_decoupling, nullptr);
    B.map4.alloc((size_t)np * 4 * R);
    B.binmap4.alloc((size_t)np * 4 * R);
    {
        DBuf<double> aa_s, rhsd4;
        aa_s.alloc(std::max<int64_t>(nb, 1));
        rhsd4.alloc(np * 4);
        for (int k = 0; k < nrhs; ++k) {
            const double *rhs = B.rhs4.p + (int64_t)k * np * 4;
            MADAM_CHECK(hipMemcpyAsync(rhsd4.p, rhs, sizeof(double) * np * 4, hipMemcpyDeviceToDevice, s));
            launch_gather_order(s, B.a64.p + (int64_t)k * nb, c.d_bord.p, aa_s.p, nb);
            launch_subtract_baselines(s, c.nmap, np, c.d_off.p, c.d_ent.p, aa_s.p, c.d_dwb_s.p, rhsd4.p);
            launch_ccmultiply(s, c.nmap, np, c.d_ccinv6.p, rhsd4.p, B.map4.p + (int64_t)k * np * 4);
            launch_ccmultiply(s, c.nmap, np, c.d_ccinv6.p, rhs, B.binmap4.p + (int64_t)k * np * 4);
        }
    }
    MADAM_CHECK(hipGetLastError());
    MADAM_CHECK(hipStreamSynchronize(s));
    B.t_setup = now() - tstart - B.t_solve;
}

void solve_batch(Context &c, int nrhs, const void *signals, int tod_f32, int on_device, int precision)
{
    if (tod_f32) solve_batch_t<float>(c, nrhs, (const float *)signals, on_device, precision);
    else solve_batch_t<double>(c, nrhs, (const double *)signals, on_device, precision);
}

// Benchmark / continuation: exactly niter PCG steps of the last batch's PCG
// (the inner FP32 one for precision 1) on every right-hand side, restarted
// from x = 0 on its right-hand side, without the exit tests (src/smadam_routines.f90:785-924).
template <typename T>
static void iterate_t(Context &c, BatchState &B, const T *b, int niter)
{
    hipStream_t s = c.stream;
    int R = B.R;
    int64_t nb = c.nbtot;
    Vec<T> v{tp<T>(B.x), tp<T>(B.r), tp<T>(B.ro), tp<T>(B.z), tp<T>(B.p), tp<T>(B.ap), tp<T>(B.pw), tp<T>(B.m),
             tp<T>(B.part)};
    batch_copy<T>(s, v.r, b, nb * R);   // x: the scratch solution vector, the batch's own stays
    batch_fill<T>(s, v.x, (T)0, nb * R);
    batch_precond<T>(c, B, v.r, v.z);
    batch_copy<T>(s, v.p, v.z, nb * R);
    batch_dots<T>(s, 1, R, v.r, v.z, v.r, c.d_rmask.p, nb, B.dots.p);
    batch_final(s, R, B.dots.p, 2, B.sc.p, B.active.p, 0.0, 0, 1 << 30, B.ones.p);
    batch_pw<T>(s, R, v.p, nb, c.d_bord.p, c.d_bslot.p, Ops<T>::dwb_s(c, B), Ops<T>::dwb(c, B), B.active.p, v.pw);
    const int big = 1 << 30;
    StepTimers tmr;
    bool graph = B.use_graph && c.segcg.empty() && !c.profile;
    for (int k = 0; k < niter; ++k) {
        if (c.profile) {
            pcg_step<T>(c, B, v, 0.0, 0, big, 0, &tmr);
        } else if (graph) {
            if (!B.gexec_free) {
                MADAM_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                pcg_step<T>(c, B, v, 0.0, 0, big, 0);
                MADAM_CHECK(hipStreamEndCapture(s, &B.graph_free));
                MADAM_CHECK(hipGraphInstantiate(&B.gexec_free, B.graph_free, nullptr, nullptr, 0));
            }
            MADAM_CHECK(hipGraphLaunch(B.gexec_free, s));
        } else {
            pcg_step<T>(c, B, v, 0.0, 0, big, 0);
        }
    }
    MADAM_CHECK(hipGetLastError());
    MADAM_CHECK(hipStreamSynchronize(s));
}

void batch_iterate(Context &c, int niter)
{
    if (!c.batch || c.batch->nrhs == 0) fail("madam_iterate_batch needs a completed madam_solve_batch");
    BatchState &B = *c.batch;
    if (B.prec == 0) {
        iterate_t<double>(c, B, B.yba.p, niter);
    } else {
        DBuf<float> b32;
        b32.alloc((size_t)std::max<int64_t>(c.nbtot, 1) * B.R);
        batch_cast<double, float>(c.stream, B.yba.p, b32.p, c.nbtot * B.R);
        iterate_t<float>(c, B, b32.p, niter);
    }
}

void batch_reset_timers(Context &c)
{
    if (!c.batch) return;
    BatchState &B = *c.batch;
    B.t_scatter = B.t_prior = B.t_gather = B.t_prec = B.t_step = 0;
    B.n_prof = 0;
}

static BatchState &batch_of(Context &c, int irhs)
{
    if (!c.batch || c.batch->nrhs == 0) fail("no batched solution available");
    if (irhs < 0 || irhs >= c.batch->nrhs) fail("right-hand side %d outside 0..%d", irhs, c.batch->nrhs - 1);
    return *c.batch;
}

void batch_get_maps(Context &c, int irhs, double *map, double *binmap)
{
    BatchState &B = batch_of(c, irhs);
    int64_t np = c.npix_map;
    int nm = c.nmap;
    std::vector<double> m4(np * 4), b4(np * 4);
    d2h(m4.data(), B.map4.p + (int64_t)irhs * np * 4, np * 4, c.stream);
    d2h(b4.data(), B.binmap4.p + (int64_t)irhs * np * 4, np * 4, c.stream);
    MADAM_CHECK(hipStreamSynchronize(c.stream));
    for (int64_t q = 0; q < np; ++q)
        for (int a = 0; a < nm; ++a) {
            if (map) map[q * nm + a] = m4[q * 4 + a];
            if (binmap) binmap[q * nm + a] = b4[q * 4 + a];
        }
}

void batch_get_baselines(Context &c, int irhs, double *aa, double *yba)
{
    BatchState &B = batch_of(c, irhs);
    int64_t nb = c.nbtot;
    if (aa) d2h(aa, B.a64.p + (int64_t)irhs * nb, nb, c.stream);
    if (yba) d2h(yba, B.yba.p + (int64_t)irhs * nb, nb, c.stream);
    MADAM_CHECK(hipStreamSynchronize(c.stream));
}

// out: steps of the (last inner) PCG, its status, its rr / rr0, refinement
// passes, the FP64 residual's rr / rrinit (mixed precision; FP64: the PCG's),
// t_setup, t_solve, inner PCG steps over the batch, t_inner, piece; with
// profiling, the summed event times of the steps' pixel side, prior,
// baseline side, preconditioner and whole step, and the steps timed
void batch_get_info(Context &c, int irhs, double *out, int n)
{
    BatchState &B = batch_of(c, irhs);
    const BatchScal &S = B.hsc[irhs];
    double v[18];
    v[0] = S.steps;
    v[1] = S.status;
    v[2] = S.rr0 > 0 ? S.rr / S.rr0 : 0;
    if (B.prec == 1) {
        const BatchScal &O = B.hsc64[irhs];
        v[3] = O.steps;
        v[4] = O.rr0 > 0 ? O.rr / O.rr0 : 0;
    } else {
        v[3] = 1;
        v[4] = v[2];
    }
    v[5] = B.t_setup;
    v[6] = B.t_solve;
    v[7] = (double)B.n_inner;
    v[8] = B.t_inner;
    v[9] = B.piece;
    v[10] = B.t_scatter;
    v[11] = B.t_prior;
    v[12] = B.t_gather;
    v[13] = B.t_prec;
    v[14] = B.t_step;
    v[15] = (double)B.n_prof;
    // run-length records of the FP32 operator (0: sample records)
    v[16] = B.rle_ok ? (double)B.n_rg : 0;
    v[17] = B.rle_ok ? (double)B.n_rs : 0;
    for (int i = 0; i < n && i < 18; ++i) out[i] = v[i];
}

}  // namespace madam
