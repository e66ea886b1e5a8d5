// Batched PCG kernels: R right-hand sides (Monte Carlo realisations of one
// scan, src/smadam.F90:593-918 solves them one at a time) that share the
// pointing, the pixel matrices, the noise filters and the preconditioner, in
// FP64 or in FP32 (the operator of the mixed-precision solve, which refines
// its FP32 solutions with FP64 residuals of the same kernels in FP64).
//
// Layouts.  Baseline vectors are RHS-major, x[r * nbtot + b], so that the
// prior's FFTs and the band solves see R times the segments of one solve.
// The projections read RHS-minor copies: pw[slot * R + r], the pixel side's
// input in baseline-slot order, and maps m[(q * NMAP + c) * R + r], so the R
// lanes that share one pointing record read R consecutive values.
//
// Pixel side (baseline_to_map + ccmultiply, src/smadam_routines.f90:1038-1148,
// src/map_routines.f90:24-77): one wave per work item, a run of the CSR entries
// of one cross pixel.  Lane (g, r), G = 64 / R entry groups, sums entries
// g, g + G, ... for right-hand side r; the G partial sums are then added over
// the lanes with xor shuffles.  Pixels longer than one item are cut into
// pieces whose sums a fix-up pass adds in piece order.
//
// Baseline side (map_to_baseline, src/smadam_routines.f90:1235-1376): one wave
// per run of consecutive baselines of one detector.  Lane (k, g, r) takes
// baseline k of a step of KB, samples g, g + G, ... of it, right-hand side r;
// the samples' pointing records are read once for all R right-hand sides.
// For R >= 8 (k_bgather_w) the waves of the co-pointed time windows of a few
// detectors share one XCD, so that its L2 holds the map rows they read.
//
// Every sum is taken in a fixed order, so results are bitwise reproducible.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <cstdlib>

#include "madam_batch.hpp"

namespace madam {

static constexpr int BTPB = 256;

static int bgrid(int64_t n, int per_block, int cap = 65535 * 8)
{
    int64_t g = (n + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

// workgroups go round-robin over the 8 XCDs: give each XCD a contiguous range
__device__ inline int64_t bx_block(int64_t bid, int64_t nblk)
{
    const int64_t NX = 8;
    int64_t per = nblk / NX;
    if (bid >= per * NX) return bid;
    return (bid % NX) * per + bid / NX;
}

#define BATCH_DISPATCH_R(RV, STMT)                                             \
    switch (RV) {                                                              \
    case 1: { constexpr int R = 1; STMT; } break;                              \
    case 2: { constexpr int R = 2; STMT; } break;                              \
    case 4: { constexpr int R = 4; STMT; } break;                              \
    case 8: { constexpr int R = 8; STMT; } break;                              \
    case 16: { constexpr int R = 16; STMT; } break;                            \
    case 32: { constexpr int R = 32; STMT; } break;                            \
    case 64: { constexpr int R = 64; STMT; } break;                            \
    default: throw HipError(hipErrorInvalidValue, "batch: unsupported RHS width", __FILE__, __LINE__);   \
    }

int batch_width(int nrhs)
{
    int w = 1;
    while (w < nrhs) w *= 2;
    return w <= 64 ? w : -1;
}

// ---------------------------------------------------------------- pixel side

// pw[bslot[b] * R + r] = p[r * nbtot + b] * dwb[b] (0 for finished right-hand
// sides): a tile of BPW_TB baselines x R is read row by row (RHS-major,
// coalesced), turned over in LDS and written one slot row of R values at a time
static constexpr int BPW_TB = 64;

// one right-hand side: pw[slot] = p[bord[slot]] * dwb_s[slot], written in order
template <typename T>
__global__ void k_bpw1(const T *p, int64_t nbtot, const int32_t *bord, const T *dwb_s, const int *active, T *pw)
{
    bool act = active[0] != 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nbtot; i += (int64_t)gridDim.x * blockDim.x)
        pw[i] = act ? p[bord[i]] * dwb_s[i] : (T)0;
}

template <typename T, int R>
__global__ void __launch_bounds__(BTPB) k_bpw(const T *p, int64_t nbtot, const int32_t *bslot, const T *dwb,
                                              const int *active, T *pw)
{
    __shared__ T tile[R][BPW_TB + 1];
    int64_t b0 = (int64_t)blockIdx.x * BPW_TB;
    int nbt = (int)(b0 + BPW_TB <= nbtot ? BPW_TB : nbtot - b0);
    for (int t = threadIdx.x; t < R * BPW_TB; t += BTPB) {
        int r = t / BPW_TB, j = t - r * BPW_TB;
        if (j < nbt) tile[r][j] = p[(int64_t)r * nbtot + b0 + j];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < R * BPW_TB; t += BTPB) {
        int j = t / R, r = t - j * R;
        if (j < nbt) {
            int64_t b = b0 + j;
            pw[(int64_t)bslot[b] * R + r] = active[r] ? tile[r][j] * dwb[b] : (T)0;
        }
    }
}

// m_r = ccinv . a_r, the inverse in lower-triangle storage [c11, c21, c31, c22, c32, c33]
template <typename T, int R, int NMAP>
__device__ inline void bs_finish(const T *cc, T a0, T a1, T a2, T *mq)
{
    if (NMAP == 1) {
        mq[0] = cc[0] * a0;
    } else {
        mq[0] = cc[0] * a0 + cc[1] * a1 + cc[2] * a2;
        mq[R] = cc[1] * a0 + cc[3] * a1 + cc[4] * a2;
        mq[2 * R] = cc[2] * a0 + cc[4] * a1 + cc[5] * a2;
    }
}

// The entries are taken in chunks of CH = U * G, the next chunk's records in
// flight with the current chunk's baseline values.  R >= 8: the chunk's
// records go through the wave's LDS slots (one record per lane in, broadcast
// reads out), with the weights of entries past the item's end zeroed there,
// so that the per-entry selects and masks stay off the VALU (the kernel is
// bound by VALU issue).  Record reads run past the item's end (the arrays
// carry zero records past their end, and every record holds a real slot).
template <typename T, int R, int RL, int NMAP>
__global__ void __launch_bounds__(BTPB) k_bscatter(const BsItem *__restrict__ items, int64_t nitems,
                                                   const void *__restrict__ ent, const T *__restrict__ pw,
                                                   const T *cc_c, const int32_t *nzq, T *m, T *part)
{
    constexpr int G = 64 / R;
    constexpr int U = R >= 64 ? 8 : (R >= 8 ? 16 / G : 4);
    constexpr int CH = U * G;
    constexpr int RB = RL == REC16 ? 16 : (RL == REC12 ? 12 : 4);
    int lane = threadIdx.x & 63;
    int r = lane % R, g = lane / R;
    const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int64_t it = bx_block(blockIdx.x, gridDim.x) * (BTPB / 64) + wib;
    if (it >= nitems) return;   // uniform over the wave
    BsItem in = items[it];
    T a0 = 0, a1 = 0, a2 = 0;
    if (in.n > 0) {
        const char *ente = (const char *)ent + (size_t)in.e0 * RB;
        const int n = in.n;
        if constexpr (R < 8) {
            // narrow rows: one record per lane per entry, the next chunk's
            // records in flight with the current baseline values
            auto load_lane = [&](PixW *e, int jb) {
#pragma unroll
                for (int u = 0; u < U; ++u) e[u] = load_rec_nt<RL>(ente, (int64_t)(uint32_t)(jb + u * G + g));
            };
            PixW ea[U], eb[U];
            auto step = [&](const PixW *cur, PixW *nxt, int j0) -> bool {
                bool more = j0 + CH < n;
                load_lane(nxt, more ? j0 + CH : j0);
                __builtin_amdgcn_sched_barrier(0);
                T v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) v[u] = pw[(int64_t)cur[u].pix * R + r];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    T fv = (j0 + u * G + g < n) ? v[u] : (T)0;
                    if (NMAP == 1) {
                        a0 += fv;
                    } else {
                        a0 += (T)cur[u].w0 * fv;
                        a1 += (T)cur[u].w1 * fv;
                        a2 += (T)cur[u].w2 * fv;
                    }
                }
                return more;
            };
            load_lane(ea, 0);
            for (int j0 = 0;; j0 += 2 * CH) {
                if (!step(ea, eb, j0)) break;
                if (!step(eb, ea, j0 + CH)) break;
            }
        } else {
            // lane j < CH loads entry record j of the chunk (one coalesced
            // load, one chunk ahead) and writes it to the wave's LDS slot j
            // with the weights zeroed past the item's end; lane (g, r) reads
            // the records of its entries u * G + g (broadcast reads)
            __shared__ int4 srec[BTPB / 64][CH];
            int4 *rl = srec[wib];
            PixW rr{0, 0.0f, 0.0f, 0.0f};
            if (lane < CH) rr = load_rec_nt<RL>(ente, (int64_t)(uint32_t)lane);
            for (int j0 = 0;; j0 += CH) {
                bool more = j0 + CH < n;
                if (lane < CH) {
                    bool ok = j0 + lane < n;
                    rl[lane] = make_int4(ok ? rr.pix : 0, __float_as_int(ok ? rr.w0 : 0.0f),
                                         __float_as_int(ok ? rr.w1 : 0.0f), __float_as_int(ok ? rr.w2 : 0.0f));
                }
                __builtin_amdgcn_wave_barrier();
                PixW e[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    int4 q = rl[u * G + g];
                    e[u].pix = q.x;
                    e[u].w0 = __int_as_float(q.y);
                    e[u].w1 = __int_as_float(q.z);
                    e[u].w2 = __int_as_float(q.w);
                }
                T v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) v[u] = pw[(int64_t)e[u].pix * R + r];
                if (more && lane < CH) rr = load_rec_nt<RL>(ente, (int64_t)(uint32_t)(j0 + CH + lane));   // in flight
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    a0 += (T)e[u].w0 * v[u];
                    if (NMAP == 3) {
                        a1 += (T)e[u].w1 * v[u];
                        a2 += (T)e[u].w2 * v[u];
                    }
                }
                if (!more) break;
            }
        }
    }
#pragma unroll
    for (int o = R; o < 64; o <<= 1) {
        a0 += __shfl_xor(a0, o, 64);
        if (NMAP == 3) {
            a1 += __shfl_xor(a1, o, 64);
            a2 += __shfl_xor(a2, o, 64);
        }
    }
    if (g == 0) {
        if (in.part >= 0) {
            T *P = part + (int64_t)in.part * NMAP * R + r;
            P[0] = a0;
            if (NMAP == 3) {
                P[R] = a1;
                P[2 * R] = a2;
            }
        } else {
            bs_finish<T, R, NMAP>(cc_c + (int64_t)in.k * 6, a0, a1, a2, m + (int64_t)nzq[in.k] * NMAP * R + r);
        }
    }
}

// thread per (long pixel, r): its pieces' sums in piece order
template <typename T, int R, int NMAP>
__global__ void k_bscatter_fix(const BsFix *fix, int64_t nfix, const T *part, const T *cc_c, const int32_t *nzq, T *m)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nfix * R) return;
    int64_t f = t / R;
    int r = (int)(t - f * R);
    BsFix fx = fix[f];
    T a0 = 0, a1 = 0, a2 = 0;
    for (int u = 0; u < fx.nparts; ++u) {
        const T *P = part + (int64_t)(fx.part0 + u) * NMAP * R + r;
        a0 += P[0];
        if (NMAP == 3) {
            a1 += P[R];
            a2 += P[2 * R];
        }
    }
    bs_finish<T, R, NMAP>(cc_c + (int64_t)fx.k * 6, a0, a1, a2, m + (int64_t)nzq[fx.k] * NMAP * R + r);
}

template <typename T>
void batch_scatter(hipStream_t s, int R, int rl, int nmap, const BsTables &bt, const void *ent, const T *pw,
                   const T *cc_c, const int32_t *nzq, T *m, T *part)
{
    if (bt.nitems <= 0) return;
    unsigned g = (unsigned)((bt.nitems + BTPB / 64 - 1) / (BTPB / 64));
    BATCH_DISPATCH_R(R, ({
        if (nmap == 1) {
            if (rl == REC16)
                k_bscatter<T, R, REC16, 1><<<g, BTPB, 0, s>>>(bt.items, bt.nitems, ent, pw, cc_c, nzq, m, part);
            else
                k_bscatter<T, R, REC4, 1><<<g, BTPB, 0, s>>>(bt.items, bt.nitems, ent, pw, cc_c, nzq, m, part);
            if (bt.nfix > 0)
                k_bscatter_fix<T, R, 1><<<bgrid(bt.nfix * R, BTPB), BTPB, 0, s>>>(bt.fix, bt.nfix, part, cc_c, nzq, m);
        } else {
            if (rl == REC12)
                k_bscatter<T, R, REC12, 3><<<g, BTPB, 0, s>>>(bt.items, bt.nitems, ent, pw, cc_c, nzq, m, part);
            else
                k_bscatter<T, R, REC16, 3><<<g, BTPB, 0, s>>>(bt.items, bt.nitems, ent, pw, cc_c, nzq, m, part);
            if (bt.nfix > 0)
                k_bscatter_fix<T, R, 3><<<bgrid(bt.nfix * R, BTPB), BTPB, 0, s>>>(bt.fix, bt.nfix, part, cc_c, nzq, m);
        }
    }));
}

// key[k] = the earliest baseline (index in its detector's timeline) among the
// CSR entries of non-empty cross pixel k: the order in which a scan first
// reaches the pixels, so that the pixel side's concurrent waves read the
// baseline slots of one stretch of the scan
__global__ void k_pixel_first_time(const int64_t *segstart, int64_t nz, const void *ent, int rec_b,
                                   const int32_t *bord, int64_t nbase, int64_t *key)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nz; k += (int64_t)gridDim.x * blockDim.x) {
        int64_t t = INT64_MAX;
        for (int64_t e = segstart[k]; e < segstart[k + 1]; ++e) {
            int32_t slot = *(const int32_t *)((const char *)ent + e * rec_b);
            int64_t b = bord[slot];
            int64_t kk = b % nbase;
            t = kk < t ? kk : t;
        }
        key[k] = t;
    }
}

void batch_pixel_keys(hipStream_t s, const int64_t *segstart, int64_t nz, const void *ent, int rec_b,
                      const int32_t *bord, int64_t nbase, int64_t *key)
{
    if (nz <= 0) return;
    k_pixel_first_time<<<bgrid(nz, BTPB), BTPB, 0, s>>>(segstart, nz, ent, rec_b, bord, nbase, key);
}

// ---------------------------------------------------------------- run-length records
// Consecutive samples of one baseline on one (cross) pixel, and consecutive
// CSR entries of one pixel from one baseline slot, contribute
// (sum of weights) x (map value) and (sum of weights) x (baseline value): one
// record {index, W0, W1, W2} stands for the run (W0 counts the samples of a
// temperature-only map, whose projections take no weight).  On a ring scan at
// Nside 512 a run holds 1.6 samples.  Runs are found and summed in stream
// order (deterministic); positions come from an exclusive scan of the run
// starts.

__global__ void k_rle_mark(const int64_t *starts, int64_t n, uint32_t *mask)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = starts[k];
        atomicOr(mask + (i >> 5), 1u << (i & 31));
    }
}

__device__ inline bool rle_bit(const uint32_t *mask, int64_t i) { return (mask[i >> 5] >> (i & 31)) & 1u; }

// baseline side: rec = the samples' full records, detector-major
__device__ inline int32_t rle_q(const PixW &r, int dsh) { return r.pix >= 0 ? r.pix >> dsh : -1; }

__global__ void k_rle_gather_flags(const PixW *rec, int64_t n, int64_t nsamp, const uint32_t *bmask, int dsh,
                                   int32_t *flag)
{
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = g % nsamp;
        int32_t q = rle_q(rec[g], dsh);
        bool st = q >= 0;
        if (st && i > 0 && !rle_bit(bmask, i)) st = rle_q(rec[g - 1], dsh) != q;
        flag[g] = st ? 1 : 0;
    }
}

__global__ void k_rle_gather_write(const PixW *rec, int64_t n, int64_t nsamp, const uint32_t *bmask, int dsh,
                                   int temperature, const int32_t *flag, const int64_t *pos, PixW *out)
{
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
        if (!flag[g]) continue;
        int64_t i = g % nsamp;
        PixW r = rec[g];
        int32_t q = rle_q(r, dsh);
        float w0 = temperature ? 1.0f : r.w0, w1 = r.w1, w2 = r.w2;
        for (int64_t j = i + 1, gj = g + 1; j < nsamp && !rle_bit(bmask, j); ++j, ++gj) {
            PixW rj = rec[gj];
            if (rle_q(rj, dsh) != q) break;
            w0 += temperature ? 1.0f : rj.w0;
            w1 += rj.w1;
            w2 += rj.w2;
        }
        out[pos[g]] = PixW{q, w0, w1, w2};
    }
}

// per detector: its first record, and each baseline's first record relative to it
__global__ void k_rle_gather_starts(const int64_t *bstart, int64_t nbase, int ndet, int64_t nsamp,
                                    const int64_t *pos, int64_t *cst, int64_t *dbase)
{
    int64_t tot = (int64_t)ndet * (nbase + 1);
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        int64_t det = t / (nbase + 1), k = t - det * (nbase + 1);
        int64_t g0 = det * nsamp;
        cst[t] = pos[g0 + bstart[k]] - pos[g0];
        if (k == 0) dbase[det] = pos[g0];
    }
}

// pixel side: ent = the CSR entries' full records {slot, w0, w1, w2}, pixel segments at smask
__global__ void k_rle_scatter_flags(const PixW *ent, int64_t n, const uint32_t *smask, int32_t *flag)
{
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        flag[e] = (e == 0 || rle_bit(smask, e) || ent[e].pix != ent[e - 1].pix) ? 1 : 0;
}

__global__ void k_rle_scatter_write(const PixW *ent, int64_t n, const uint32_t *smask, int temperature,
                                    const int32_t *flag, const int64_t *pos, PixW *out)
{
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        if (!flag[e]) continue;
        PixW r = ent[e];
        float w0 = temperature ? 1.0f : r.w0, w1 = r.w1, w2 = r.w2;
        for (int64_t j = e + 1; j < n && !rle_bit(smask, j) && ent[j].pix == r.pix; ++j) {
            w0 += temperature ? 1.0f : ent[j].w0;
            w1 += ent[j].w1;
            w2 += ent[j].w2;
        }
        out[pos[e]] = PixW{r.pix, w0, w1, w2};
    }
}

__global__ void k_rle_scatter_segs(const int64_t *segstart, int64_t nz, const int64_t *pos, int64_t *cseg)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k <= nz; k += (int64_t)gridDim.x * blockDim.x)
        cseg[k] = pos[segstart[k]];
}

void batch_rle_gather(hipStream_t s, const PixW *rec, int ndet, int64_t nsamp, const int64_t *bstart, int64_t nbase,
                      int dsh, int temperature, uint32_t *bmask, int32_t *flag, int64_t *pos, int64_t *scan_tmp,
                      PixW *out, int64_t *cst, int64_t *dbase, bool count_only)
{
    int64_t n = (int64_t)ndet * nsamp;
    if (n <= 0) return;
    if (count_only) {
        MADAM_CHECK(hipMemsetAsync(bmask, 0, sizeof(uint32_t) * ((nsamp + 31) / 32 + 1), s));
        k_rle_mark<<<bgrid(nbase, BTPB), BTPB, 0, s>>>(bstart, nbase, bmask);
        k_rle_gather_flags<<<bgrid(n, BTPB), BTPB, 0, s>>>(rec, n, nsamp, bmask, dsh, flag);
        int64_t nchunks = (n + 1 + 4095) / 4096;
        launch_exclusive_scan_i32_to_i64(s, flag, n, pos, scan_tmp, nchunks);
        return;
    }
    k_rle_gather_write<<<bgrid(n, BTPB), BTPB, 0, s>>>(rec, n, nsamp, bmask, dsh, temperature, flag, pos, out);
    k_rle_gather_starts<<<bgrid((int64_t)ndet * (nbase + 1), BTPB), BTPB, 0, s>>>(bstart, nbase, ndet, nsamp, pos, cst,
                                                                                 dbase);
}

void batch_rle_scatter(hipStream_t s, const PixW *ent, int64_t nent, const int64_t *segstart, int64_t nz,
                       int temperature, uint32_t *smask, int32_t *flag, int64_t *pos, int64_t *scan_tmp, PixW *out,
                       int64_t *cseg, bool count_only)
{
    if (nent <= 0) return;
    if (count_only) {
        MADAM_CHECK(hipMemsetAsync(smask, 0, sizeof(uint32_t) * ((nent + 31) / 32 + 1), s));
        k_rle_mark<<<bgrid(nz, BTPB), BTPB, 0, s>>>(segstart, nz, smask);
        k_rle_scatter_flags<<<bgrid(nent, BTPB), BTPB, 0, s>>>(ent, nent, smask, flag);
        int64_t nchunks = (nent + 1 + 4095) / 4096;
        launch_exclusive_scan_i32_to_i64(s, flag, nent, pos, scan_tmp, nchunks);
        return;
    }
    k_rle_scatter_write<<<bgrid(nent, BTPB), BTPB, 0, s>>>(ent, nent, smask, temperature, flag, pos, out);
    k_rle_scatter_segs<<<bgrid(nz + 1, BTPB), BTPB, 0, s>>>(segstart, nz, pos, cseg);
}

template <typename T>
void batch_pw(hipStream_t s, int R, const T *p, int64_t nbtot, const int32_t *bord, const int32_t *bslot,
              const T *dwb_s, const T *dwb, const int *active, T *pw)
{
    if (nbtot <= 0) return;
    if (R == 1) {
        k_bpw1<T><<<bgrid(nbtot, BTPB), BTPB, 0, s>>>(p, nbtot, bord, dwb_s, active, pw);
        return;
    }
    unsigned g = (unsigned)((nbtot + BPW_TB - 1) / BPW_TB);
    BATCH_DISPATCH_R(R, (k_bpw<T, R><<<g, BTPB, 0, s>>>(p, nbtot, bslot, dwb, active, pw)));
}

// ---------------------------------------------------------------- baseline side

static constexpr int BG_RUN = 64;   // baselines per wave of k_bgather
static constexpr int BG_RUN_MAX = 63;   // baselines per wave of k_bgather_w (one lane each, plus the end)

// time-major wave order: the waves of one time window, all detectors, run
// together (co-pointed detectors read the same map pixels)
template <typename T, int R, int RL, int NMAP>
__global__ void __launch_bounds__(BTPB) k_bgather(int ndet, int64_t nsamp, int64_t nbase, const int64_t *bstart,
                                                  const void *rec, const T *m, int dsh, const T *dwb, const T *nna,
                                                  const T *p, T *ap, const int *active)
{
    constexpr int G = R >= 8 ? 64 / R : 8;   // sample groups per baseline
    constexpr int KB = 64 / (G * R);        // baselines per step
    int lane = threadIdx.x & 63;
    int r = lane % R, g = (lane / R) % G, kk = lane / (R * G);
    int64_t nruns = (nbase + BG_RUN - 1) / BG_RUN;
    int64_t wv = bx_block(blockIdx.x, gridDim.x) * (BTPB / 64) + (threadIdx.x >> 6);
    int64_t run = wv / ndet, det = wv - run * ndet;
    if (run >= nruns) return;   // uniform over the wave
    int64_t k0 = run * BG_RUN;
    int64_t k1 = k0 + BG_RUN < nbase ? k0 + BG_RUN : nbase;
    int64_t nbtot = (int64_t)ndet * nbase;
    bool act = active[r] != 0;
    const int64_t rbase = det * nsamp;
    for (int64_t kb = k0; kb < k1; kb += KB) {
        int64_t k = kb + kk;
        bool valid = k < k1;
        int64_t i0 = valid ? bstart[k] : 0, i1 = valid ? bstart[k + 1] : 0;
        T acc = 0;
        int64_t i = i0 + g;
        // four samples per step, their records and then their map values in flight together
        for (; i + 3 * G < i1; i += 4 * G) {
            PixW e[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) e[u] = load_rec_nt<RL>(rec, rbase + i + u * G);
            T mv[4][NMAP];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const T *mq = m + (int64_t)((e[u].pix >= 0 ? e[u].pix : 0) >> dsh) * NMAP * R + r;
#pragma unroll
                for (int c = 0; c < NMAP; ++c) mv[u][c] = mq[c * R];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                T v = NMAP == 1 ? mv[u][0] : (T)e[u].w0 * mv[u][0] + (T)e[u].w1 * mv[u][1] + (T)e[u].w2 * mv[u][2];
                acc += e[u].pix >= 0 ? v : (T)0;
            }
        }
        for (; i < i1; i += G) {
            PixW e = load_rec_nt<RL>(rec, rbase + i);
            if (e.pix >= 0) {
                const T *mq = m + (int64_t)(e.pix >> dsh) * NMAP * R + r;
                acc += NMAP == 1 ? mq[0] : (T)e.w0 * mq[0] + (T)e.w1 * mq[R] + (T)e.w2 * mq[2 * R];
            }
        }
#pragma unroll
        for (int o = R; o < R * G; o <<= 1) acc += __shfl_xor(acc, o, 64);
        if (g == 0 && valid && act) {
            int64_t b = det * nbase + k, ix = (int64_t)r * nbtot + b;
            ap[ix] = ap[ix] + (nna[b] * p[ix] - acc * dwb[b]);
        }
    }
}

// R >= 8: one baseline at a time per wave, G = 64 / R sample groups.  The
// baselines of a run are cut into chunks of CH = U * G samples.  A chunk's
// records are wave-uniform (scalar loads, one copy per wave rather than one
// per lane); each lane takes its samples' values from them, and the next
// chunk's records are loaded while the current chunk's map values are in
// flight.  Record reads run past a baseline's end (up to CH - 1 records, the
// arrays carry zero records past their end); those samples weigh 0, as do
// flagged ones (which read the row of pixel 0: the rows of pixels without
// hits are zero, the others finite).  The run's baseline starts sit one per
// lane.  Sums are taken in chunk order, samples g, g + G, ... per group, then
// over the groups by xor shuffles.
//
// ndg > 0: detector groups per XCD split.  Workgroups go round-robin over the
// 8 XCDs; XCD x takes detector group x % ndg (of ndg, detectors contiguous)
// over time slice x / ndg (of 8 / ndg), time-major inside, so that the L2 of
// each XCD holds the map pixels of a few detectors only.  ndg = 0: every XCD
// takes a contiguous slice of the time-major (run, detector) order.
//
// The baseline sums of a run go to the wave's LDS rows (run_len x (R + 1),
// dynamic shared memory, BTPB / 64 waves), and the run ends with one pass that
// reads p, ap, nna, dwb and writes ap as R contiguous row segments of the run.
template <typename T, int R, int RL, int NMAP, int U>
__global__ void __launch_bounds__(BTPB) k_bgather_w(int ndet, int64_t nsamp, int64_t nbase, int run_len, int ndg,
                                                    const int64_t *__restrict__ bstart, int64_t bstride,
                                                    const int64_t *__restrict__ dbase, const void *__restrict__ rec,
                                                    const T *__restrict__ m, int dsh, const T *dwb, const T *nna,
                                                    const T *p, T *ap, const int *active)
{
    extern __shared__ __align__(16) unsigned char bg_lds[];
    constexpr int G = 64 / R;
    constexpr int CH = U * G;   // samples per chunk
    constexpr int RB = RL == REC16 ? 16 : (RL == REC12 ? 12 : 4);
    int lane = threadIdx.x & 63;
    int r = lane % R, g = lane / R;
    int64_t nruns = (nbase + run_len - 1) / run_len;
    int64_t run, det;
    const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // the wave in the block (uniform)
    T *stage = reinterpret_cast<T *>(bg_lds) + wib * run_len * (R + 1);
    if (ndg > 0) {
        int xcd = blockIdx.x & 7, ntg = 8 / ndg;
        int dg = xcd % ndg, tg = xcd / ndg;
        int d0 = (int)((int64_t)dg * ndet / ndg), d1 = (int)((int64_t)(dg + 1) * ndet / ndg);
        int64_t r0 = (int64_t)tg * nruns / ntg, r1 = (int64_t)(tg + 1) * nruns / ntg;
        int64_t lw = (int64_t)(blockIdx.x >> 3) * (BTPB / 64) + wib;
        int nd = d1 - d0;
        if (nd <= 0) return;
        run = r0 + lw / nd;
        det = d0 + lw % nd;
        if (run >= r1) return;   // uniform over the wave
    } else {
        int64_t wv = bx_block(blockIdx.x, gridDim.x) * (BTPB / 64) + wib;
        run = wv / ndet;
        det = wv - run * ndet;
        if (run >= nruns) return;   // uniform over the wave
    }
    int64_t k0 = run * run_len;
    int nk = (int)(k0 + run_len < nbase ? run_len : nbase - k0);
    // sample indices inside the detector's stream fit 32 bits (nsamp < 2^31)
    // (run-length records: per-detector baseline starts and record bases)
    int32_t bsl = (int32_t)bstart[det * bstride + k0 + (lane < nk ? lane : nk)];   // run_len <= 63: lane nk holds the end
    int64_t nbtot = (int64_t)ndet * nbase;
    const char *recd = (const char *)rec + (size_t)(dbase ? dbase[det] : det * nsamp) * RB;
    // the baseline bounds are read with readlane: the chunk control is scalar
    int kc = 0;
    int32_t ib = __builtin_amdgcn_readlane(bsl, 0), ie = __builtin_amdgcn_readlane(bsl, 1);
    int32_t i = ib;
    T acc = 0;
    // The chunk's records: lane j < CH loads record j (one coalesced load,
    // one chunk ahead) and writes it to the wave's LDS slot j as
    // {pix, w0, w1, w2} with the pixel and weights zeroed for samples past the
    // baseline's end and flagged ones; lane (g, r) then reads the records of
    // its samples u * G + g (broadcast reads).  The per-sample selects and
    // masks stay off the VALU.
    int4 *rl = reinterpret_cast<int4 *>(reinterpret_cast<T *>(bg_lds) + (BTPB / 64) * run_len * (R + 1)) + wib * CH;
    PixW rr{0, 0.0f, 0.0f, 0.0f};
    if (lane < CH) rr = load_rec_nt<RL>(recd, (int64_t)(uint32_t)(i + lane));
    for (;;) {
        bool last = i + CH >= ie;
        int nkc = kc;
        int32_t ni = i + CH, nib = ib, nie = ie;
        if (last) {
            nkc = kc + 1;
            nib = ie;
            nie = __builtin_amdgcn_readlane(bsl, nkc + 1 < 64 ? nkc + 1 : 63);
            ni = nib;
        }
        bool more = nkc < nk;
        if (lane < CH) {
            bool ok = i + lane < ie && rr.pix >= 0;
            rl[lane] = make_int4(ok ? rr.pix : 0, __float_as_int(ok ? rr.w0 : 0.0f), __float_as_int(ok ? rr.w1 : 0.0f),
                                 __float_as_int(ok ? rr.w2 : 0.0f));
        }
        __builtin_amdgcn_wave_barrier();
        PixW e[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            int4 q = rl[u * G + g];
            e[u].pix = q.x;
            e[u].w0 = __int_as_float(q.y);
            e[u].w1 = __int_as_float(q.z);
            e[u].w2 = __int_as_float(q.w);
        }
        T mv[U][NMAP];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const T *mq = m + (int64_t)(e[u].pix >> dsh) * NMAP * R + r;
#pragma unroll
            for (int c = 0; c < NMAP; ++c) mv[u][c] = mq[c * R];
        }
        // the next chunk's records, in flight with the map values
        if (more && lane < CH) rr = load_rec_nt<RL>(recd, (int64_t)(uint32_t)(ni + lane));
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (NMAP == 1) {
                acc += (T)e[u].w0 * mv[u][0];
            } else {
                acc += (T)e[u].w0 * mv[u][0];
                acc += (T)e[u].w1 * mv[u][1];
                acc += (T)e[u].w2 * mv[u][2];
            }
        }
        if (last) {
#pragma unroll
            for (int o = R; o < 64; o <<= 1) acc += __shfl_xor(acc, o, 64);
            if (g == 0) stage[kc * (R + 1) + r] = acc;
            acc = 0;
        }
        if (!more) break;
        i = ni;
        kc = nkc;
        ib = nib;
        ie = nie;
    }
    __builtin_amdgcn_wave_barrier();
    // ap = ap + (nna p - sum dwb) over the run, row segment by row segment
    const int64_t bb0 = det * nbase + k0;
    for (int t = lane; t < nk * R; t += 64) {
        int r2 = t / nk, j = t - r2 * nk;
        if (active[r2]) {
            int64_t b = bb0 + j, ix = (int64_t)r2 * nbtot + b;
            ap[ix] = ap[ix] + (nna[b] * p[ix] - stage[j * (R + 1) + r2] * dwb[b]);
        }
    }
}

static int bg_run_len()
{
    static int v = [] {
        const char *e = getenv("MADAM_BG_RUN");
        int x = e ? atoi(e) : 32;
        return x < 1 ? 1 : (x > BG_RUN_MAX ? BG_RUN_MAX : x);
    }();
    return v;
}

static bool bg_xcd_split()
{
    static bool v = [] {
        const char *e = getenv("MADAM_BG_XCD");
        return e ? atoi(e) != 0 : true;
    }();
    return v;
}

template <typename T>
void batch_gather(hipStream_t s, int R, int rl, int nmap, int ndet, int64_t nsamp, int64_t nbase,
                  const int64_t *bstart, const void *rec, const T *m, int64_t ndeg, const T *dwb, const T *nna,
                  const T *p, T *ap, const int *active, int64_t bstride, const int64_t *dbase)
{
    if (ndet <= 0 || nbase <= 0) return;
    int dsh = 0;
    while ((1ll << dsh) < ndeg) ++dsh;
    if (R >= 8) {
        int run_len = bg_run_len();
        int64_t nruns = (nbase + run_len - 1) / run_len;
        int64_t nw = (int64_t)ndet * nruns;
        unsigned g = (unsigned)((nw + BTPB / 64 - 1) / (BTPB / 64));
        int ndg = 0;
        if (bg_xcd_split()) {
            // about four detectors per group, the rest of the XCDs in time
            // slices: 16 detectors gather 6 % faster in 4 groups x 2 slices
            // than in 8 groups x 1, 8 in 2 x 4 and 4 in 1 x 8 by 15 % and
            // 11 % (MADAM_BG_NDG = 1, 2, 4 or 8 overrides)
            ndg = ndet >= 32 ? 8 : ndet >= 16 ? 4 : ndet >= 8 ? 2 : 1;
            if (const char *e = getenv("MADAM_BG_NDG")) {
                int v = atoi(e);
                if ((v == 1 || v == 2 || v == 4 || v == 8) && v <= ndet) ndg = v;
            }
            int ntg = 8 / ndg;
            // the largest (detector group x time slice) rectangle, in waves, on every XCD
            int64_t w = (int64_t)((ndet + ndg - 1) / ndg) * ((nruns + ntg - 1) / ntg);
            g = (unsigned)(8 * ((w + BTPB / 64 - 1) / (BTPB / 64)));
        }
#define BGW_ARGS ndet, nsamp, nbase, run_len, ndg, bstart, bstride, dbase, rec, m, dsh, dwb, nna, p, ap, active
#define BGW_LAUNCH(RR, UU)                                                                                           \
    do {                                                                                                             \
        size_t lds = sizeof(T) * (BTPB / 64) * run_len * (RR + 1) + (BTPB / 64) * 16 * sizeof(int4);                 \
        if (nmap == 1 && rl == REC16)                                                                                \
            k_bgather_w<T, RR, REC16, 1, UU><<<g, BTPB, lds, s>>>(BGW_ARGS);                                         \
        else if (nmap == 1)                                                                                          \
            k_bgather_w<T, RR, REC4, 1, UU><<<g, BTPB, lds, s>>>(BGW_ARGS);                                          \
        else if (rl == REC12)                                                                                        \
            k_bgather_w<T, RR, REC12, 3, UU><<<g, BTPB, lds, s>>>(BGW_ARGS);                                         \
        else                                                                                                         \
            k_bgather_w<T, RR, REC16, 3, UU><<<g, BTPB, lds, s>>>(BGW_ARGS);                                         \
    } while (0)
        switch (R) {   // 16 records per chunk (8 at R = 64)
        case 8: BGW_LAUNCH(8, 2); break;
        case 16: BGW_LAUNCH(16, 4); break;
        case 32: BGW_LAUNCH(32, 8); break;
        default: BGW_LAUNCH(64, 8); break;
        }
#undef BGW_LAUNCH
#undef BGW_ARGS
        return;
    }
    if (dbase)
        throw HipError(hipErrorInvalidValue, "batch_gather: run-length records need R >= 8", __FILE__, __LINE__);
    int64_t nw = (int64_t)ndet * ((nbase + BG_RUN - 1) / BG_RUN);
    unsigned g = (unsigned)((nw + BTPB / 64 - 1) / (BTPB / 64));
    BATCH_DISPATCH_R(R, ({
        if (nmap == 1)
            k_bgather<T, R, REC4, 1><<<g, BTPB, 0, s>>>(ndet, nsamp, nbase, bstart, rec, m, dsh, dwb, nna, p, ap,
                                                          active);
        else if (rl == REC12)
            k_bgather<T, R, REC12, 3><<<g, BTPB, 0, s>>>(ndet, nsamp, nbase, bstart, rec, m, dsh, dwb, nna, p, ap,
                                                           active);
        else
            k_bgather<T, R, REC16, 3><<<g, BTPB, 0, s>>>(ndet, nsamp, nbase, bstart, rec, m, dsh, dwb, nna, p, ap,
                                                           active);
    }));
}

// ---------------------------------------------------------------- vector steps (RHS-major, grid.y = r)

static constexpr int BDOT_BLK = 512;   // blocks per right-hand side for the dot products

// four consecutive vector elements in one load (rows of length n with
// n % 4 == 0 start on 4-element boundaries)
template <typename T>
struct Quad;
template <>
struct Quad<float> {
    using V = float4;
};
template <>
struct Quad<double> {
    using V = double4;
};

// masked dot products of one right-hand side per grid row, FP64 sums:
// which 0: s0 = a.b;  which 1: s0 = a.b, s1 = c.b, s2 = a.a
template <typename T, int WHICH>
__global__ void __launch_bounds__(BTPB) k_bdots(const T *__restrict__ a, const T *__restrict__ b,
                                                const T *__restrict__ c, const int8_t *__restrict__ mask, int64_t n,
                                                double *partials)
{
    int64_t off = (int64_t)blockIdx.y * n;
    a += off;
    b += off;
    if (WHICH == 1) c += off;
    double s0 = 0, s1 = 0, s2 = 0;
    int64_t i0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    auto acc = [&](int8_t m, T av, T bv, T cv) {
        if (!m) return;
        double ai = (double)av, bi = (double)bv;
        s0 += ai * bi;
        if (WHICH == 1) {
            s1 += (double)cv * bi;
            s2 += ai * ai;
        }
    };
    if ((n & 3) == 0) {
        // four elements per load: loads only in the loop body, so they pipeline
        using V = typename Quad<T>::V;
        const V *a4 = (const V *)a, *b4 = (const V *)b, *c4 = (const V *)c;
        const char4 *m4 = (const char4 *)mask;
        for (int64_t q = i0; q < (n >> 2); q += stride) {
            char4 m = m4[q];
            V av = a4[q], bv = b4[q], cv = av;
            if (WHICH == 1) cv = c4[q];
            acc(m.x, av.x, bv.x, cv.x);
            acc(m.y, av.y, bv.y, cv.y);
            acc(m.z, av.z, bv.z, cv.z);
            acc(m.w, av.w, bv.w, cv.w);
        }
    } else {
        for (int64_t i = i0; i < n; i += stride) acc(mask[i], a[i], b[i], WHICH == 1 ? c[i] : a[i]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        s0 += __shfl_xor(s0, o, 64);
        s1 += __shfl_xor(s1, o, 64);
        s2 += __shfl_xor(s2, o, 64);
    }
    __shared__ double sh[3][BTPB / 64];
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { sh[0][wv] = s0; sh[1][wv] = s1; sh[2][wv] = s2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t0 = 0, t1 = 0, t2 = 0;
        for (int k = 0; k < BTPB / 64; ++k) { t0 += sh[0][k]; t1 += sh[1][k]; t2 += sh[2][k]; }
        double *o = partials + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * 3;
        o[0] = t0;
        o[1] = t1;
        o[2] = t2;
    }
}

// Per-RHS scalars (BatchScal) from the dot partials, one block per right-hand
// side.  stage 0: pap.  stage 1: rz_new, rz2, rr, beta (Polak-Ribiere, as the
// single solve), then the exit tests of src/smadam_routines.f90:905-918 on
// the active right-hand sides: rr / rr0 > 1e3 diverging, rr / rr0 < lim after
// more than iter_min steps or rz = 0 converged, iter_max steps reached.  A
// right-hand side that stops is marked inactive before its p update.
// stage 2 (start of a solve): rz, rr0 = rr; active if rz >= 1e-30 and not NaN.
// stage 3: rr only (the refinement's residual test).  stage 4: stage 1 with
// rz2 = ap . z in place of r_old . z (beta = -(ap . z) / pap, the same
// Polak-Ribiere beta, since r_old = r + alpha ap).
__global__ void __launch_bounds__(BTPB) k_bfinal(const double *partials, int nblk, int stage, BatchScal *sc,
                                                 int *active, double lim, int iter_min, int iter_max,
                                                 const int *allow, int stall)
{
    int r = blockIdx.x;
    __shared__ double sh[3][BTPB];
    double t[3] = {0, 0, 0};
    const double *P = partials + (int64_t)r * nblk * 3;
    for (int i = threadIdx.x; i < nblk; i += BTPB)
        for (int q = 0; q < 3; ++q) t[q] += P[i * 3 + q];
    for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] = t[q];
    __syncthreads();
    for (int o = BTPB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    BatchScal &S = sc[r];
    if (stage == 0) {
        S.pap = sh[0][0];
        return;
    }
    if (stage == 2) {
        S.rz = sh[0][0];
        S.rr = sh[2][0];
        S.rr0 = sh[2][0];
        S.rrmin = sh[2][0];
        S.steps = 0;
        S.nstall = 0;
        S.status = 0;
        bool ok = S.rz >= 1e-30 && S.rz == S.rz && (!allow || allow[r]);
        active[r] = ok ? 1 : 0;
        if (!ok) S.status = 8;
        return;
    }
    if (!active[r]) return;
    double L = lim < 0 ? S.lim : lim;
    if (stage == 3) {
        double rr = sh[2][0];
        S.rr = rr;
        S.steps += 1;
        if (rr / S.rr0 < L) {
            S.status = 1;
            active[r] = 0;
        } else if (S.steps >= iter_max) {
            S.status = 4;
            active[r] = 0;
        }
        return;
    }
    double rzn = sh[0][0], rz2 = sh[1][0], rr = sh[2][0];
    // stage 1: rz2 = r_old . z; stage 4: rz2 = ap . z, and r_old = r + alpha ap
    // gives r_old . z = rzn + alpha (ap . z): the same Polak-Ribiere beta
    // without keeping r_old
    S.beta = stage == 4 ? -rz2 / S.pap : (rzn - rz2) / S.rz;
    S.rz = rzn;
    S.rr = rr;
    S.steps += 1;
    if (rr < 0.999 * S.rrmin) {
        S.rrmin = rr;
        S.nstall = 0;
    } else {
        S.nstall += 1;
    }
    double q = rr / S.rr0;
    if (q > 1e3) {
        S.status = 2;
        active[r] = 0;
    } else if ((q < L && S.steps > iter_min) || rzn == 0) {
        S.status = 1;
        active[r] = 0;
    } else if (S.steps >= iter_max) {
        S.status = 4;
        active[r] = 0;
    } else if (stall > 0 && S.nstall >= stall) {
        S.status = 16;
        active[r] = 0;
    }
}

__global__ void k_binner_lim(int R, const BatchScal *outer, BatchScal *inner, double cglimit, double floor)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    double need = outer[r].rr > 0 ? cglimit * outer[r].rr0 / outer[r].rr : 1.0;
    inner[r].lim = need > floor ? need : floor;
}

void batch_inner_limits(hipStream_t s, int R, const BatchScal *outer, BatchScal *inner, double cglimit, double floor)
{
    k_binner_lim<<<1, 64, 0, s>>>(R, outer, inner, cglimit, floor);
}

// alpha = rz / pap; ro = r; x += alpha p; r -= alpha ap (active right-hand sides)
template <typename T>
__global__ void k_bupdate_r(const BatchScal *sc, const int *active, T *__restrict__ x, T *__restrict__ r,
                            T *__restrict__ ro, const T *__restrict__ p, const T *__restrict__ ap, int64_t n)
{
    int y = blockIdx.y;
    if (!active[y]) return;
    T alpha = (T)(sc[y].rz / sc[y].pap);
    int64_t off = (int64_t)y * n;
    int64_t i0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    if ((n & 3) == 0) {
        using V = typename Quad<T>::V;
        V *x4 = (V *)(x + off), *r4 = (V *)(r + off), *ro4 = ro ? (V *)(ro + off) : nullptr;
        const V *p4 = (const V *)(p + off), *ap4 = (const V *)(ap + off);
        for (int64_t q = i0; q < (n >> 2); q += stride) {
            V ri = r4[q], xi = x4[q], pi = p4[q], ai = ap4[q];
            if (ro4) ro4[q] = ri;
            xi.x += alpha * pi.x;
            xi.y += alpha * pi.y;
            xi.z += alpha * pi.z;
            xi.w += alpha * pi.w;
            ri.x -= alpha * ai.x;
            ri.y -= alpha * ai.y;
            ri.z -= alpha * ai.z;
            ri.w -= alpha * ai.w;
            x4[q] = xi;
            r4[q] = ri;
        }
    } else {
        for (int64_t i = off + i0; i < off + n; i += stride) {
            T ri = r[i];
            if (ro) ro[i] = ri;
            x[i] = x[i] + alpha * p[i];
            r[i] = ri - alpha * ap[i];
        }
    }
}

// p = z + beta p (active right-hand sides)
template <typename T>
__global__ void k_bupdate_p(const BatchScal *sc, const int *active, T *p, const T *z, int64_t n)
{
    int y = blockIdx.y;
    if (!active[y]) return;
    T beta = (T)sc[y].beta;
    int64_t off = (int64_t)y * n;
    for (int64_t i = off + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < off + n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = z[i] + beta * p[i];
}

// p = z + beta p on a tile of BPW_TB baselines x R (RHS-major rows), and the
// next operator's pw[bslot[b] * R + r] = p[r][b] * dwb[b] from the same tile
// (0 for finished right-hand sides): k_bupdate_p and k_bpw in one pass
template <typename T, int R>
__global__ void __launch_bounds__(BTPB) k_bupdate_p_pw(const BatchScal *sc, const int *active, T *__restrict__ p,
                                                       const T *__restrict__ z, int64_t nbtot,
                                                       const int32_t *__restrict__ bslot, const T *__restrict__ dwb,
                                                       T *__restrict__ pw)
{
    __shared__ T tile[R][BPW_TB + 1];
    int64_t b0 = (int64_t)blockIdx.x * BPW_TB;
    int nbt = (int)(b0 + BPW_TB <= nbtot ? BPW_TB : nbtot - b0);
    if ((nbtot & 3) == 0 && nbt == BPW_TB) {
        // whole tile, rows on 4-element boundaries: every load of the thread
        // is issued before the first store
        using V = typename Quad<T>::V;
        constexpr int QR = BPW_TB / 4;                          // quads per row
        constexpr int QPT = (R * QR + BTPB - 1) / BTPB;         // quads per thread
        V pv[QPT], zv[QPT];
#pragma unroll
        for (int k = 0; k < QPT; ++k) {
            int qd = threadIdx.x + k * BTPB;
            if (qd < R * QR) {
                int r = qd / QR, j = (qd - r * QR) * 4;
                int64_t i = (int64_t)r * nbtot + b0 + j;
                pv[k] = *(const V *)(p + i);
                zv[k] = *(const V *)(z + i);
            }
        }
#pragma unroll
        for (int k = 0; k < QPT; ++k) {
            int qd = threadIdx.x + k * BTPB;
            if (qd < R * QR) {
                int r = qd / QR, j = (qd - r * QR) * 4;
                V v = pv[k];
                if (active[r]) {
                    T beta = (T)sc[r].beta;
                    v.x = zv[k].x + beta * v.x;
                    v.y = zv[k].y + beta * v.y;
                    v.z = zv[k].z + beta * v.z;
                    v.w = zv[k].w + beta * v.w;
                    *(V *)(p + (int64_t)r * nbtot + b0 + j) = v;
                }
                tile[r][j] = v.x;
                tile[r][j + 1] = v.y;
                tile[r][j + 2] = v.z;
                tile[r][j + 3] = v.w;
            }
        }
    } else {
        for (int t = threadIdx.x; t < R * BPW_TB; t += BTPB) {
            int r = t / BPW_TB, j = t - r * BPW_TB;
            if (j < nbt) {
                int64_t i = (int64_t)r * nbtot + b0 + j;
                T v = p[i];
                if (active[r]) {
                    v = z[i] + (T)sc[r].beta * v;
                    p[i] = v;
                }
                tile[r][j] = v;
            }
        }
    }
    __syncthreads();
    // pw rows: a thread keeps one right-hand side (BTPB is a multiple of R),
    // its slots and weights loaded before the stores
    constexpr int NP = (BPW_TB * R + BTPB - 1) / BTPB;               // passes
    int32_t sl[NP];
    T dw[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        int t = threadIdx.x + k * BTPB;
        int j = t / R;
        sl[k] = -1;
        if (t < R * BPW_TB && j < nbt) {
            sl[k] = bslot[b0 + j];
            dw[k] = dwb[b0 + j];
        }
    }
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        int t = threadIdx.x + k * BTPB;
        int j = t / R, r = t - j * R;
        if (sl[k] >= 0) pw[(int64_t)sl[k] * R + r] = active[r] ? tile[r][j] * dw[k] : (T)0;
    }
}

// one right-hand side: p = z + beta p in baseline order (coalesced) and pw
// written scattered to the slots, in place of a pass gathering p through the
// slot order
template <typename T>
__global__ void k_bupdate_p_pw1(const BatchScal *sc, const int *active, T *__restrict__ p, const T *__restrict__ z,
                                int64_t n, const int32_t *__restrict__ bslot, const T *__restrict__ dwb,
                                T *__restrict__ pw)
{
    bool on = active[0] != 0;
    T beta = (T)sc[0].beta;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        T v = p[i];
        T w = dwb[i];
        int32_t sl = bslot[i];
        if (on) {
            v = z[i] + beta * v;
            p[i] = v;
        }
        pw[sl] = on ? v * w : (T)0;
    }
}

template <typename T>
void batch_update_p_pw(hipStream_t s, int R, const BatchScal *sc, const int *active, T *p, const T *z, int64_t nbtot,
                       const int32_t *bslot, const T *dwb, T *pw)
{
    if (nbtot <= 0) return;
    if (R == 1) {
        k_bupdate_p_pw1<T><<<(unsigned)bgrid(nbtot, BTPB, 4096 * 4), BTPB, 0, s>>>(sc, active, p, z, nbtot, bslot, dwb,
                                                                                  pw);
        return;
    }
    unsigned g = (unsigned)((nbtot + BPW_TB - 1) / BPW_TB);
    BATCH_DISPATCH_R(R, (k_bupdate_p_pw<T, R><<<g, BTPB, 0, s>>>(sc, active, p, z, nbtot, bslot, dwb, pw)));
}

template <typename T>
void batch_dots(hipStream_t s, int which, int R, const T *a, const T *b, const T *c, const int8_t *mask, int64_t n,
                double *partials)
{
    dim3 g(BDOT_BLK, (unsigned)R);
    if (which == 0) k_bdots<T, 0><<<g, BTPB, 0, s>>>(a, b, c, mask, n, partials);
    else k_bdots<T, 1><<<g, BTPB, 0, s>>>(a, b, c, mask, n, partials);
}

void batch_final(hipStream_t s, int R, const double *partials, int stage, BatchScal *sc, int *active, double lim,
                 int iter_min, int iter_max, const int *allow, int stall)
{
    k_bfinal<<<(unsigned)R, BTPB, 0, s>>>(partials, BDOT_BLK, stage, sc, active, lim, iter_min, iter_max, allow, stall);
}

size_t batch_partials_doubles(int R) { return (size_t)R * BDOT_BLK * 3; }

template <typename T>
void batch_update_r(hipStream_t s, int R, const BatchScal *sc, const int *active, T *x, T *r, T *ro, const T *p,
                    const T *ap, int64_t n)
{
    dim3 g((unsigned)bgrid(n, BTPB, 4096), (unsigned)R);
    k_bupdate_r<T><<<g, BTPB, 0, s>>>(sc, active, x, r, ro, p, ap, n);
}

template <typename T>
void batch_update_p(hipStream_t s, int R, const BatchScal *sc, const int *active, T *p, const T *z, int64_t n)
{
    dim3 g((unsigned)bgrid(n, BTPB, 4096), (unsigned)R);
    k_bupdate_p<T><<<g, BTPB, 0, s>>>(sc, active, p, z, n);
}

// ---------------------------------------------------------------- elementwise helpers

template <typename TI, typename TO>
__global__ void k_bcast(const TI *in, TO *out, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (TO)in[i];
}

template <typename TI, typename TO>
void batch_cast(hipStream_t s, const TI *in, TO *out, int64_t n)
{
    if (n > 0) k_bcast<TI, TO><<<bgrid(n, BTPB, 65535 * 4), BTPB, 0, s>>>(in, out, n);
}

// a64 += x (the refinement step of right-hand sides that were active in the inner solve)
template <typename T>
__global__ void k_baccum(double *a, const T *x, const int *was_active, int64_t n)
{
    int y = blockIdx.y;
    if (!was_active[y]) return;
    int64_t off = (int64_t)y * n;
    for (int64_t i = off + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < off + n;
         i += (int64_t)gridDim.x * blockDim.x)
        a[i] = a[i] + (double)x[i];
}

template <typename T>
void batch_accumulate(hipStream_t s, int R, double *a, const T *x, const int *was_active, int64_t n)
{
    dim3 g((unsigned)bgrid(n, BTPB, 4096), (unsigned)R);
    k_baccum<T><<<g, BTPB, 0, s>>>(a, x, was_active, n);
}

// r = b - ap (FP64 residual of the refinement)
__global__ void k_bresid(const double *b, const double *ap, double *r, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        r[i] = b[i] - ap[i];
}

void batch_residual(hipStream_t s, const double *b, const double *ap, double *r, int64_t n)
{
    if (n > 0) k_bresid<<<bgrid(n, BTPB, 65535 * 4), BTPB, 0, s>>>(b, ap, r, n);
}

// z = r * pd[b] (diagonal preconditioner, shared over the right-hand sides)
template <typename T>
__global__ void k_bdiag(const T *r, const double *pd, T *z, int64_t n)
{
    int64_t off = (int64_t)blockIdx.y * n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        z[off + i] = (T)((double)r[off + i] * pd[i]);
}

template <typename T>
void batch_diag(hipStream_t s, int R, const T *r, const double *pd, T *z, int64_t n)
{
    dim3 g((unsigned)bgrid(n, BTPB, 4096), (unsigned)R);
    k_bdiag<T><<<g, BTPB, 0, s>>>(r, pd, z, n);
}

template <typename T>
__global__ void k_bcopy(T *d, const T *a, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = a[i];
}

template <typename T>
void batch_copy(hipStream_t s, T *d, const T *a, int64_t n)
{
    if (n > 0) k_bcopy<T><<<bgrid(n, BTPB, 65535 * 4), BTPB, 0, s>>>(d, a, n);
}

template <typename T>
__global__ void k_bfill(T *d, T v, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = v;
}

template <typename T>
void batch_fill(hipStream_t s, T *d, T v, int64_t n)
{
    if (n > 0) k_bfill<T><<<bgrid(n, BTPB, 65535 * 4), BTPB, 0, s>>>(d, v, n);
}

// out[(q * nmap + c) * R + r] -> out4[r][q][4] (double) for the final maps
template <typename T>
__global__ void k_bmap_unpack(const T *m, int64_t npix, int nmap, int R, int r, double *out4)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < npix; q += (int64_t)gridDim.x * blockDim.x) {
        double4 v;
        v.x = (double)m[(q * nmap + 0) * R + r];
        v.y = nmap == 3 ? (double)m[(q * nmap + 1) * R + r] : 0.0;
        v.z = nmap == 3 ? (double)m[(q * nmap + 2) * R + r] : 0.0;
        v.w = 0.0;
        ((double4 *)out4)[q] = v;
    }
}

// ---------------------------------------------------------------- explicit instantiations

#define BATCH_INST(T)                                                                                              \
    template void batch_scatter<T>(hipStream_t, int, int, int, const BsTables &, const void *, const T *, const T *, \
                                   const int32_t *, T *, T *);                                                     \
    template void batch_update_p_pw<T>(hipStream_t, int, const BatchScal *, const int *, T *, const T *, int64_t,  \
                                       const int32_t *, const T *, T *);                                          \
    template void batch_pw<T>(hipStream_t, int, const T *, int64_t, const int32_t *, const int32_t *, const T *,   \
                               const T *, const int *, T *); \
    template void batch_gather<T>(hipStream_t, int, int, int, int, int64_t, int64_t, const int64_t *, const void *, \
                                  const T *, int64_t, const T *, const T *, const T *, T *, const int *, int64_t,      \
                                  const int64_t *);          \
    template void batch_dots<T>(hipStream_t, int, int, const T *, const T *, const T *, const int8_t *, int64_t,    \
                                double *);                                                                         \
    template void batch_update_r<T>(hipStream_t, int, const BatchScal *, const int *, T *, T *, T *, const T *,     \
                                    const T *, int64_t);                                                           \
    template void batch_update_p<T>(hipStream_t, int, const BatchScal *, const int *, T *, const T *, int64_t);     \
    template void batch_accumulate<T>(hipStream_t, int, double *, const T *, const int *, int64_t);                \
    template void batch_diag<T>(hipStream_t, int, const T *, const double *, T *, int64_t);                        \
    template void batch_copy<T>(hipStream_t, T *, const T *, int64_t);                                             \
    template void batch_fill<T>(hipStream_t, T *, T, int64_t);

BATCH_INST(float)
BATCH_INST(double)

template void batch_cast<double, float>(hipStream_t, const double *, float *, int64_t);
template void batch_cast<float, double>(hipStream_t, const float *, double *, int64_t);
template void batch_cast<double, double>(hipStream_t, const double *, double *, int64_t);
template void batch_cast<float, float>(hipStream_t, const float *, float *, int64_t);

void batch_map_unpack(hipStream_t s, const void *m, int is_float, int64_t npix, int nmap, int R, int r, double *out4)
{
    if (npix <= 0) return;
    if (is_float)
        k_bmap_unpack<float><<<bgrid(npix, BTPB, 65535 * 4), BTPB, 0, s>>>((const float *)m, npix, nmap, R, r, out4);
    else
        k_bmap_unpack<double><<<bgrid(npix, BTPB, 65535 * 4), BTPB, 0, s>>>((const double *)m, npix, nmap, R, r,
                                                                           out4);
}

}  // namespace madam
