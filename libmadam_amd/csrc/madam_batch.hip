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
// per run of BG_RUN baselines of one detector.  Lane (k, g, r) takes baseline
// k of a step of KB, samples g, g + G, ... of it, right-hand side r; the
// samples' pointing records are read once for all R right-hand sides.
//
// Every sum is taken in a fixed order, so results are bitwise reproducible.
#include <hip/hip_runtime.h>

#include <cstdint>

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

// pw[slot * R + r] = p[r * nbtot + bord[slot]] * dwb_s[slot] (0 for finished right-hand sides)
template <typename T, int R>
__global__ void k_bpw(const T *p, int64_t nbtot, const int32_t *bord, const T *dwb_s, const int *active, T *pw)
{
    int64_t n = nbtot * R;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t slot = i / R;
        int r = (int)(i - slot * R);
        T v = 0;
        if (active[r]) v = p[r * nbtot + bord[slot]] * dwb_s[slot];
        pw[i] = v;
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

template <typename T, int R, int RL, int NMAP>
__global__ void __launch_bounds__(BTPB) k_bscatter(const BsItem *items, int64_t nitems, const void *ent, const T *pw,
                                                   const T *cc_c, const int32_t *nzq, T *m, T *part)
{
    constexpr int G = 64 / R;
    int lane = threadIdx.x & 63;
    int r = lane % R, g = lane / R;
    int64_t it = bx_block(blockIdx.x, gridDim.x) * (BTPB / 64) + (threadIdx.x >> 6);
    if (it >= nitems) return;   // uniform over the wave
    BsItem in = items[it];
    T a0 = 0, a1 = 0, a2 = 0;
    int j = g;
    // four entries per step, every load issued before the arithmetic
    for (; j + 3 * G < in.n; j += 4 * G) {
        PixW e[4];
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) e[u] = load_rec<RL>(ent, in.e0 + j + u * G);
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = pw[(int64_t)e[u].pix * R + r];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (NMAP == 1) {
                a0 += v[u];
            } else {
                a0 += (T)e[u].w0 * v[u];
                a1 += (T)e[u].w1 * v[u];
                a2 += (T)e[u].w2 * v[u];
            }
        }
    }
    for (; j < in.n; j += G) {
        PixW e = load_rec<RL>(ent, in.e0 + j);
        T v = pw[(int64_t)e.pix * R + r];
        if (NMAP == 1) {
            a0 += v;
        } else {
            a0 += (T)e.w0 * v;
            a1 += (T)e.w1 * v;
            a2 += (T)e.w2 * v;
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

template <typename T>
void batch_pw(hipStream_t s, int R, const T *p, int64_t nbtot, const int32_t *bord, const T *dwb_s,
              const int *active, T *pw)
{
    if (nbtot <= 0) return;
    BATCH_DISPATCH_R(R, (k_bpw<T, R><<<bgrid(nbtot * R, BTPB), BTPB, 0, s>>>(p, nbtot, bord, dwb_s, active, pw)));
}

// ---------------------------------------------------------------- baseline side

static constexpr int BG_RUN = 64;   // baselines per wave

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
    // time-major: the waves of one time window, all detectors, run together
    // (co-pointed detectors read the same map pixels)
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
            for (int u = 0; u < 4; ++u) e[u] = load_rec<RL>(rec, rbase + i + u * G);
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
            PixW e = load_rec<RL>(rec, rbase + i);
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

template <typename T>
void batch_gather(hipStream_t s, int R, int rl, int nmap, int ndet, int64_t nsamp, int64_t nbase,
                  const int64_t *bstart, const void *rec, const T *m, int64_t ndeg, const T *dwb, const T *nna,
                  const T *p, T *ap, const int *active)
{
    int64_t nw = (int64_t)ndet * ((nbase + BG_RUN - 1) / BG_RUN);
    if (nw <= 0) return;
    int dsh = 0;
    while ((1ll << dsh) < ndeg) ++dsh;
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

// masked dot products of one right-hand side per grid row, FP64 sums:
// which 0: s0 = a.b;  which 1: s0 = a.b, s1 = c.b, s2 = a.a
template <typename T, int WHICH>
__global__ void __launch_bounds__(BTPB) k_bdots(const T *a, const T *b, const T *c, const int8_t *mask, int64_t n,
                                                double *partials)
{
    int64_t off = (int64_t)blockIdx.y * n;
    a += off;
    b += off;
    if (WHICH == 1) c += off;
    double s0 = 0, s1 = 0, s2 = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (!mask[i]) continue;
        double ai = (double)a[i], bi = (double)b[i];
        s0 += ai * bi;
        if (WHICH == 1) {
            s1 += (double)c[i] * bi;
            s2 += ai * ai;
        }
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
    S.beta = (rzn - rz2) / S.rz;
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
__global__ void k_bupdate_r(const BatchScal *sc, const int *active, T *x, T *r, T *ro, const T *p, const T *ap,
                            int64_t n)
{
    int y = blockIdx.y;
    if (!active[y]) return;
    T alpha = (T)(sc[y].rz / sc[y].pap);
    int64_t off = (int64_t)y * n;
    for (int64_t i = off + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < off + n;
         i += (int64_t)gridDim.x * blockDim.x) {
        T ri = r[i];
        ro[i] = ri;
        x[i] = x[i] + alpha * p[i];
        r[i] = ri - alpha * ap[i];
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
    template void batch_pw<T>(hipStream_t, int, const T *, int64_t, const int32_t *, const T *, const int *, T *); \
    template void batch_gather<T>(hipStream_t, int, int, int, int, int64_t, int64_t, const int64_t *, const void *, \
                                  const T *, int64_t, const T *, const T *, const T *, T *, const int *);          \
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
