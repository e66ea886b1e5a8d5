// Band-Cholesky preconditioner for the PCG baseline solve.
//
// Reference: construct_preconditioner / get_cholesky_decomposition /
// apply_cholesky_decomposition (src/noise_routines.f90:920-1153, 1259-1301):
// per trimmed segment, the band matrix M with M(i, j) = invcov(|i - j|) for
// |i - j| < nband plus nna on the diagonal is factored with DPBTRF('L') and
// applied with DPBTRS.
//
// The factor here is the same global Cholesky factor L (one wave per segment,
// right-looking, rolling (kd+1)^2 window in LDS).  The triangular solves are
// parallelised exactly, without approximation, by partitioning each segment
// into chunks of CH rows: a chunk's sweep is an affine function of the KD-wide
// carry state entering it, s_out = s0_out + T s_in, with T precomputed per
// chunk.  A solve is then: sweep every chunk with zero carry (parallel), chain
// the carries, sweep again with the true carries (parallel).  The chain is
// itself split in three: groups of G chunks chain from zero in parallel, one
// wave per segment chains the groups with the precomputed group products
// T_last ... T_first, and the groups chain again from their true carries, so
// the sequential depth is 2 G + nchunk / G instead of nchunk.  Segments with a
// smaller band are zero-padded to the largest KD.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "madam_internal.hpp"

namespace madam {

static constexpr int BAND_CH = 64;    // rows per chunk (one thread each in the sweeps)
static constexpr int BAND_G = 64;     // chunks per group of the carry chain (at most 64: one per lane)
static_assert(BAND_G <= 64, "");
static_assert(BAND_CH > 0, "");

struct ChunkInfo {
    int64_t b0;    // first global baseline row of the chunk
    int32_t n;     // rows
    int32_t seg;   // segment
    int32_t first, last;   // first / last chunk of its segment
};


// ---------------------------------------------------------------- factorization

// Register-window variant: one wave per segment, lane a holds row j + a of the
// active (K x K) window, W[b] = A(j + a, j + b).  Per column: broadcast the
// pivot, scale the column, rank-1 update, shift the window up by one row.
// broadcast of lane `src` (wave-uniform) through the scalar unit
__device__ inline double readlane_d(double v, int src)
{
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_readlane(lo, src);
    hi = __builtin_amdgcn_readlane(hi, src);
    return __hiloint2double(hi, lo);
}

// 1 / sqrt(a) for a > 0: the hardware estimate refined by two Newton steps
// y <- y (3 - a y^2) / 2 (relative error of the estimate ~2^-23, so the result is
// within about an ulp); shorter than a correctly rounded sqrt followed by a divide
__device__ inline double rsqrt_nr(double a)
{
    double y = __builtin_amdgcn_rsq(a);
    double h = 0.5 * a;
    y = y * fma(-h * y, y, 1.5);
    y = y * fma(-h * y, y, 1.5);
    return y;
}

// lane l takes lane l + 1's value (row_shl:1 inside 16-lane rows; lane 15 of a row gets 0)
__device__ inline double shift_down_row(double v)
{
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_update_dpp(0, lo, 0x101, 0xf, 0xf, false);
    hi = __builtin_amdgcn_update_dpp(0, hi, 0x101, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

// Register-window variant: one wave per segment, lane a holds row j + a of the
// active (K x K) window, W[b] = A(j + a, j + b).  Per column: broadcast the
// pivot, scale the column, rank-1 update, shift the window up by one row.  The
// diagonal terms nna of the rows entering the window are loaded 64 rows ahead.
//
// Periodic continuation.  The window at the start of column j holds rows
// j .. j+K-1 of A (Toeplitz terms, nna on the diagonal) minus the products of
// columns j-K+1 .. j-1 of L, in a fixed order, so it is a function of those
// columns and of those rows' nna.  If the 64 columns of block j0 equal those of
// block j0 - 64 bitwise and nna is one value v on rows j0 .. e-1, the window
// after block j0 equals the one after block j0 - 64, and every block j0 + 64 q'
// whose rows end before e (j0 + 64 q' + 64 + K <= e) repeats block j0 bitwise.
// The wave then records (j0, q) for k_band_fill_periodic, which copies the
// block, and resumes at the first block past them with the window after block
// j0, the window the full column loop would reach there.  nna runs are tracked
// as the rows are loaded: rstart is the first row of the run the last loaded
// row belongs to (rows before K are treated as a run of their own).
constexpr int BAND_MAXJ = 16;   // recorded repeats per segment

struct NnaRun {
    int64_t rstart;
    long long v;
};

// fold rows r0 .. r0+63 (lane t: row r0 + t, value x if r0 + t < n) into the run
__device__ inline void nna_run_add(NnaRun &run, int64_t r0, int64_t n, double x, int lane)
{
    long long xb = __double_as_longlong(x);
    long long prev = __shfl_up(xb, 1, 64);
    if (lane == 0) prev = run.v;
    uint64_t d = __ballot(r0 + lane < n && xb != prev);
    if (d) run.rstart = r0 + 63 - __builtin_clzll(d);
    int64_t last = n - 1 - r0 < 63 ? n - 1 - r0 : 63;
    if (last >= 0) run.v = __shfl(xb, (int)last, 64);
}

// first row >= r (up to n) whose nna differs from v
__device__ inline int64_t nna_run_end(const double *nn, int64_t r, int64_t n, long long v, int lane)
{
    for (; r < n; r += 512) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            int64_t i = r + 64 * u + lane;
            long long x = __double_as_longlong(nn[i < n ? i : n - 1]);
            uint64_t d = __ballot(i < n && x != v);
            if (d) return r + 64 * u + __builtin_ctzll(d);
        }
    }
    return n;
}

template <int KD>
__global__ void __launch_bounds__(64) k_band_factor_w(const Segment *seg, int64_t nseg, const int *nband_seg,
                                                      const double *invcov, int64_t nof, const double *nna,
                                                      double *Lb, int *info, int64_t *per_src, int64_t *per_q)
{
    constexpr int K = KD + 1;
    constexpr bool DPP_SHIFT = K <= 15;   // the window fits one 16-lane DPP row
    int lane = threadIdx.x & 63;
    int64_t s = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    if (s >= nseg) return;
    Segment sg = seg[s];
    int nb = nband_seg[s];
    if (nb <= 0 || sg.skip) return;
    const double *ic = invcov + (int64_t)sg.ipsd * nof;
    const double *nn = nna + sg.b0;
    int64_t n = sg.noba;
    double icv[K];
#pragma unroll
    for (int m = 0; m < K; ++m) icv[m] = m < nb ? ic[m] : 0.0;
    double W[K];
#pragma unroll
    for (int bb = 0; bb < K; ++bb) {
        // A(a, bb) for a = lane, lower triangle
        double v = 0.0;
        if (lane < K && bb <= lane && lane < n) {
            int d = lane - bb;
#pragma unroll
            for (int m = 0; m < K; ++m) if (m == d) v = icv[m];
            if (bb == lane) v += nn[lane];
        }
        W[bb] = v;
    }
    // Columns go in blocks of 64: block j0 needs the nna of rows j0 + K .. j0 + K + 63
    // (lane t holds row j0 + K + t); the next block's values are loaded one block
    // ahead, so the only memory wait is once per block.
    double nnext = K + lane < n ? nn[K + lane] : 0.0;
    NnaRun run{K, 0};
    int njump = 0;
    int64_t jfirst = 0;   // first block computed since the last repeat (the copies land later)
    if (per_src) {
        run.v = __double_as_longlong(__shfl(nnext, 0, 64));
        nna_run_add(run, K, n, nnext, lane);
    }
    for (int64_t j0 = 0; j0 < n; j0 += 64) {
        double nbuf = nnext;
        nnext = j0 + 64 + K + lane < n ? nn[j0 + 64 + K + lane] : 0.0;
        int tn = n - j0 < 64 ? (int)(n - j0) : 64;
        for (int t = 0; t < tn; ++t) {
            int64_t j = j0 + t;
            double ajj = readlane_d(W[0], 0);
            if (!(ajj > 0)) {
                if (lane == 0) info[s] = (int)(j + 1);
                return;
            }
            double dinv = rsqrt_nr(ajj);
            double l = (lane >= 1 && lane < K && j + lane < n) ? W[0] * dinv : 0.0;
            if (lane == 0) Lb[(sg.b0 + j) * K] = dinv;
            else if (lane < K) Lb[(sg.b0 + j) * K + lane] = l;
            // rank-1 update of rows 1..KD: W[bb] -= l_a l_bb for 1 <= bb <= a
#pragma unroll
            for (int bb = 1; bb < K; ++bb) {
                double lb = readlane_d(l, bb);
                if (bb <= lane) W[bb] -= l * lb;
            }
            // the row entering the window: i = j + K
            int64_t i = j + K;
            double nni = readlane_d(nbuf, t);
            // shift: new row a = old row a + 1, columns shifted left by one
            double Wn[K];
#pragma unroll
            for (int bb = 0; bb < K - 1; ++bb)
                Wn[bb] = DPP_SHIFT ? shift_down_row(W[bb + 1]) : __shfl(W[bb + 1], lane + 1, 64);
            Wn[K - 1] = 0.0;
            // new last row: entries A(i, j + 1 + bb) = ic[K - 1 - bb] (+ nna on the diagonal)
            if (lane == K - 1) {
#pragma unroll
                for (int bb = 0; bb < K; ++bb) {
                    double v = 0.0;
                    if (i < n) {
                        v = icv[K - 1 - bb];
                        if (bb == K - 1) v += nni;
                    }
                    Wn[bb] = v;
                }
            }
#pragma unroll
            for (int bb = 0; bb < K; ++bb) W[bb] = Wn[bb];
        }
        if (!per_src || njump == BAND_MAXJ) continue;
        nna_run_add(run, j0 + 64 + K, n, nnext, lane);   // rows up to j0 + 127 + K seen
        if (tn < 64 || j0 < jfirst + 64 || run.rstart > j0 || j0 + 128 + K > n) continue;
        bool eq = true;
        if (lane < K) {
            const double *cur = Lb + (sg.b0 + j0) * K + lane, *prv = cur - 64 * K;
#pragma unroll 8
            for (int t = 0; t < 64; ++t)
                eq = eq && __double_as_longlong(cur[t * K]) == __double_as_longlong(prv[t * K]);
        }
        if (__ballot(!eq) != 0) continue;
        int64_t e = nna_run_end(nn, j0 + 128 + K, n, run.v, lane);
        int64_t q = (e - 64 - K - j0) / 64;   // blocks j0 + 64, ..., j0 + 64 q repeat block j0
        if (q < 1) continue;
        if (lane == 0) {
            per_src[s * BAND_MAXJ + njump] = j0;
            per_q[s * BAND_MAXJ + njump] = q;
        }
        ++njump;
        j0 += 64 * q;
        jfirst = j0 + 64;
        // resume at block jfirst: the window's rows jfirst .. jfirst+K-1 lie in
        // the run; its entering rows go through the run tracker now
        nnext = jfirst + K + lane < n ? nn[jfirst + K + lane] : 0.0;
        nna_run_add(run, jfirst + K, n, nnext, lane);
    }
}

// Lb rows of block src + 64 q' (q' = 1 .. q) = those of block src, for every
// repeat the factorization recorded
__global__ void k_band_fill_periodic(const Segment *seg, int64_t nseg, int K, const int64_t *per_src,
                                     const int64_t *per_q, double *Lb)
{
    int blk = 64 * K;
    for (int64_t s = blockIdx.y; s < nseg; s += gridDim.y)
        for (int k = 0; k < BAND_MAXJ; ++k) {
            int64_t q = per_q[s * BAND_MAXJ + k];
            if (q <= 0) break;
            double *from = Lb + (seg[s].b0 + per_src[s * BAND_MAXJ + k]) * K;
            for (int64_t qq = 1 + blockIdx.x; qq <= q; qq += gridDim.x)
                for (int e = threadIdx.x; e < blk; e += blockDim.x) from[qq * blk + e] = from[e];
        }
}

// ---------------------------------------------------------------- partitioned solves


// The factor is re-laid out per chunk for the sweeps, blocked by wave:
// Li[(((c / 64) * CH + i) * K + m) * 64 + c % 64] is L(b0_c + i + m, b0_c + i)
// (m = 0 holds 1 / L(i, i)).  The lanes of a wave (consecutive chunks c) read
// consecutive addresses, and a wave's whole sweep is one contiguous block.
//
// Where the band matrix is stationary (equal baseline lengths, no flags) the
// factor's rows reach a bitwise fixed point or a short cycle (period 1, 2 or
// 4, all dividing the chunk length) within a few hundred rows, so most chunks
// repeat the previous one exactly.  bmap[w] names a block whose 64 chunks are
// bitwise identical to those of block w (the first block of a run of equal
// blocks), and cmap[c] the first chunk of c's run of equal chunks: the sweeps
// and the carry chains read through them, so a run streams its factor and its
// transfer matrices from cache once instead of from memory per chunk.  The
// values read are the same bits, so the results do not change.  A negative
// map entry -1 - w marks a block whose 64 chunks are all equal to each other:
// every lane then reads lane 0's column of block w, one cache line per load.
__device__ inline const double *li_chunk(const double *Li, const int32_t *bmap, int64_t c, int K)
{
    int64_t w = c >> 6, lane = c & 63;
    if (bmap) {
        int32_t m = bmap[w];
        w = m >= 0 ? m : -1 - m;
        if (m < 0) lane = 0;
    }
    return Li + w * (int64_t)BAND_CH * K * 64 + lane;
}

// The sweeps read r and write z in the solve's own layout (baseline order,
// one right-hand side after another) through one LDS tile per 64-chunk block:
// load k of the wave reads rows 0..63 of chunk k of the block (one contiguous
// piece of the vector); row i of chunk j sits at tile[j * 64 + (i ^ j)], so
// that both the loads (one chunk, all rows) and the sweeps (lane j on chunk j,
// one row) spread over the banks without padding: the FP64 tile is then
// exactly 32 KB, five blocks per CU (four with a padded row).  Each lane keeps
// its own chunk's first row and row count, and the loads take them with lane
// reads.  Rows past a chunk's end read as 0 and are never written back.
static_assert(BAND_CH == 64, "the tile maps one row of a chunk to each lane of the wave");
// Strides of one right-hand side (grid dimension y) in the solve's buffers:
// s = per-chunk carry states, g = per-group states
struct BandRhs {
    int64_t s, g;
};

template <typename V>
__device__ inline void band_tile_in(const V *v, uint32_t b0, int n, V *tile)
{
    const int lane = threadIdx.x;
    constexpr int PB = 128 / sizeof(V);   // loads in flight per lane before their LDS stores
#pragma unroll
    for (int k0 = 0; k0 < 64; k0 += PB) {
        V in[PB];
#pragma unroll
        for (int k = 0; k < PB; ++k) {
            const int nk = __builtin_amdgcn_readlane(n, k0 + k);
            const uint32_t bk = (uint32_t)__builtin_amdgcn_readlane((int)b0, k0 + k);
            in[k] = lane < nk ? v[bk + (uint32_t)lane] : (V)0;
        }
#pragma unroll
        for (int k = 0; k < PB; ++k) tile[(k0 + k) * 64 + (lane ^ (k0 + k))] = in[k];
    }
    __syncthreads();
}

template <typename V>
__device__ inline void band_tile_out(V *v, uint32_t b0, int n, const V *tile)
{
    __syncthreads();
    const int lane = threadIdx.x;
#pragma unroll 16
    for (int k = 0; k < 64; ++k) {
        const int nk = __builtin_amdgcn_readlane(n, k);
        const uint32_t bk = (uint32_t)__builtin_amdgcn_readlane((int)b0, k);
        if (lane < nk) v[bk + (uint32_t)lane] = tile[k * 64 + (lane ^ k)];
    }
}

// compact row-major copy [i][m] of lane 0's column of block reps[u] (a block
// whose chunks are all equal), for the scalar-load sweeps
__global__ void k_band_compact_uniform(const double *Li, int K, const int32_t *reps, int64_t nu, double *Lu)
{
    int64_t tot = nu * BAND_CH * K;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        int64_t u = t / (BAND_CH * K), im = t % (BAND_CH * K);
        Lu[t] = Li[((int64_t)reps[u] * BAND_CH * K + im) * 64];
    }
}

// ceq[c] = 1 if chunk c holds the same row count and bitwise the same factor
// rows as chunk c - 1
__global__ void k_band_chunk_eq(const ChunkInfo *ch, int64_t nch, int K, const double *Li, uint8_t *ceq)
{
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nch; c += (int64_t)gridDim.x * blockDim.x) {
        bool eq = c > 0 && ch[c].n == ch[c - 1].n;
        if (eq) {
            const double *a = li_chunk(Li, nullptr, c, K), *b = li_chunk(Li, nullptr, c - 1, K);
            int n = ch[c].n;
            for (int64_t t = 0; t < (int64_t)n * K && eq; ++t)
                eq = __double_as_longlong(a[t * 64]) == __double_as_longlong(b[t * 64]);
        }
        ceq[c] = eq ? 1 : 0;
    }
}

__global__ void k_band_relayout(const ChunkInfo *ch, int64_t nch, int64_t NCP, int K, int CH, const double *Lb,
                                double *Li)
{
    int64_t tot = NCP * (int64_t)CH * K;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        int64_t lane = t & 63, rest = t >> 6;
        int64_t m = rest % K, rest2 = rest / K;
        int64_t i = rest2 % CH, w = rest2 / CH;
        int64_t c = w * 64 + lane;
        double v = 0.0;
        if (c < nch) {
            ChunkInfo ci = ch[c];
            if (i < ci.n) v = Lb[(ci.b0 + i) * K + m];
        }
        Li[t] = v;
    }
}

// Rows are processed in groups of BAND_PF: the factor rows and right-hand
// sides of a group are loaded first (they do not depend on the carry state),
// so each thread keeps BAND_PF rows of loads in flight instead of one.
// (fewer rows per group for the wide bands, to stay within the register file)
template <int KD>
struct BandPF {
    static constexpr int value = KD <= 9 ? 4 : (KD <= 19 ? 8 : (KD <= 29 ? 4 : 2));
};

// forward sweep of one chunk: L y = r, carry state acc[KD] = pending sums for the
// next KD rows (acc[0] for the first row of the chunk).  If store, writes y.
// lic is the chunk's factor column, li_chunk(...) resolved once by the caller
// (a map load inside the row loop would wait for every load in flight), with
// element m of row i at lic[(i * K + m) * MS]: MS = 64 for the interleaved
// factor, MS = 1 for a block's compact copy read with scalar loads; r and
// y are the chunk's rows in the block's LDS tile (row i at [i ^ key]); y may be
// r (each row is read before it is written).
template <int KD, int MS = 64, typename V = double>
__device__ inline void fwd_sweep(const double *lic, int n, const V *r, double *acc, V *y, bool store,
                                 bool zero_rhs, int key = 0)
{
    constexpr int K = KD + 1, BAND_PF = BandPF<KD>::value;
    for (int i0 = 0; i0 < n; i0 += BAND_PF) {
        double rw[BAND_PF][K], rh[BAND_PF];
#pragma unroll
        for (int u = 0; u < BAND_PF; ++u) {
            int i = i0 + u;
            if (i < n) {
                const double *row = lic + (int64_t)i * K * MS;
#pragma unroll
                for (int m = 0; m < K; ++m) rw[u][m] = row[m * MS];
                rh[u] = zero_rhs ? 0.0 : (double)r[i ^ key];
            }
        }
#pragma unroll
        for (int u = 0; u < BAND_PF; ++u) {
            int i = i0 + u;
            if (i < n) {
                double v = (rh[u] + acc[0]) * rw[u][0];
#pragma unroll
                for (int m = 1; m < KD; ++m) acc[m - 1] = acc[m] - rw[u][m] * v;
                acc[KD - 1] = -rw[u][KD] * v;
                if (store) y[i ^ key] = (V)v;
            }
        }
    }
}

// backward sweep: L^T x = y, state xs[0..KD) = x of the KD rows following the chunk
// (x may be y, as for fwd_sweep)
template <int KD, int MS = 64, typename V = double>
__device__ inline void bwd_sweep(const double *lic, int n, const V *y, double *xs, V *x, bool store,
                                 bool zero_rhs, int key = 0)
{
    constexpr int K = KD + 1, BAND_PF = BandPF<KD>::value;
    for (int i0 = n - 1; i0 >= 0; i0 -= BAND_PF) {
        double rw[BAND_PF][K], rh[BAND_PF];
#pragma unroll
        for (int u = 0; u < BAND_PF; ++u) {
            int i = i0 - u;
            if (i >= 0) {
                const double *row = lic + (int64_t)i * K * MS;
#pragma unroll
                for (int m = 0; m < K; ++m) rw[u][m] = row[m * MS];
                rh[u] = zero_rhs ? 0.0 : (double)y[i ^ key];
            }
        }
#pragma unroll
        for (int u = 0; u < BAND_PF; ++u) {
            int i = i0 - u;
            if (i >= 0) {
                double s = rh[u];
#pragma unroll
                for (int m = 1; m <= KD; ++m) s -= rw[u][m] * xs[m - 1];
                double xi = s * rw[u][0];
#pragma unroll
                for (int m = KD - 1; m > 0; --m) xs[m] = xs[m - 1];
                xs[0] = xi;
                if (store) x[i ^ key] = (V)xi;
            }
        }
    }
}

// T matrices: response of the outgoing state to a unit incoming state
template <int KD>
__global__ void __launch_bounds__(64) k_band_response(const ChunkInfo *ch, int64_t nch, int64_t NCP, const double *Lb, double *Tf,
                                double *Tb)
{
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nch * KD) return;
    int64_t c = t / KD;
    int j = (int)(t % KD);
    ChunkInfo ci = ch[c];
    double st[KD];
#pragma unroll
    for (int m = 0; m < KD; ++m) st[m] = (m == j) ? 1.0 : 0.0;
    fwd_sweep<KD, 64, double>(li_chunk(Lb, nullptr, c, KD + 1), ci.n, nullptr, st, nullptr, false, true);
#pragma unroll
    for (int m = 0; m < KD; ++m) Tf[(c * KD + m) * KD + j] = st[m];
#pragma unroll
    for (int m = 0; m < KD; ++m) st[m] = (m == j) ? 1.0 : 0.0;
    bwd_sweep<KD, 64, double>(li_chunk(Lb, nullptr, c, KD + 1), ci.n, nullptr, st, nullptr, false, true);
#pragma unroll
    for (int m = 0; m < KD; ++m) Tb[(c * KD + m) * KD + j] = st[m];
}

// zero-carry forward sweep of every chunk: its outgoing partial state s0
template <int KD, typename V>
__global__ void __launch_bounds__(64) k_band_fwd0(const ChunkInfo *ch, int64_t nch, const int32_t *bmap,
                                                  const double *Lb, const int32_t *uidx,
                                                  const double *__restrict__ Lu, const V *r, int64_t vstride,
                                                  double *s0, BandRhs st)
{
    __shared__ V tile[64 * 64];
    r += blockIdx.y * vstride;
    s0 += blockIdx.y * st.s;
    int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const uint32_t b0 = c < nch ? (uint32_t)ch[c].b0 : 0u;   // baseline rows < ndet * nsamp < 2^32
    const int n = c < nch ? ch[c].n : 0;
    band_tile_in(r, b0, n, tile);
    if (c >= nch) return;   // past the last barrier
    const int lane = threadIdx.x;
    const V *tl = tile + lane * 64;
    double acc[KD];
#pragma unroll
    for (int m = 0; m < KD; ++m) acc[m] = 0;
    int32_t ui = uidx[blockIdx.x];
    if (ui >= 0)
        fwd_sweep<KD, 1, V>(Lu + (int64_t)ui * BAND_CH * (KD + 1), __builtin_amdgcn_readfirstlane(n), tl, acc,
                            (V *)nullptr, false, false, lane);
    else
        fwd_sweep<KD, 64, V>(li_chunk(Lb, bmap, c, KD + 1), n, tl, acc, (V *)nullptr, false, false, lane);
#pragma unroll
    for (int m = 0; m < KD; ++m) s0[c * KD + m] = acc[m];
}

// chain carries inside each segment: one wave per segment, lane m < KD holds
// state component m; s_next = s0[c] + T[c] s, sequential over the chunks
// Group tables: group g covers chunks gfirst[g] .. gfirst[g] + gn[g] - 1 of one segment;
// the segment's groups are sfirst[s] .. sfirst[s + 1] - 1.
// Wave per group: chain the group's chunks (increasing order for dir 0,
// decreasing for dir 1) starting from zero (cg[g] = outgoing state) or from the
// group's true incoming state Sg[g] (writing every chunk's incoming state sin).
// Both chain levels run one wave per chain with lane m < KD holding state
// component m.  The chain's additive terms are staged in LDS before the loop,
// the state is broadcast through LDS, and the transfer matrix stays in
// registers while consecutive steps share it (runs of equal chunks or groups,
// through cmap / gmap); it is reloaded only where the run changes.  The sums
// are taken in the same order on every path.
// v + tr . st: the KD products are independent, and a pairwise tree adds them
// (depth log2 KD instead of a KD-long dependent chain of FMAs on the path
// from one step's state to the next)
template <int KD>
__device__ inline double chain_step(double v, const double (&tr)[KD], double st, double *sb, int base = 0)
{
    // the state goes to the wave's LDS row and comes back as KD broadcast
    // reads of the lane's chain (lanes base .. base + KD - 1), cheaper per
    // step than 2 KD lane reads through the scalar unit
    sb[threadIdx.x & 63] = st;
    __builtin_amdgcn_wave_barrier();
    double pr[KD];
#pragma unroll
    for (int q = 0; q < KD; ++q) pr[q] = tr[q] * sb[base + q];
#pragma unroll
    for (int w = 1; w < KD; w *= 2)
#pragma unroll
        for (int q = 0; q + w < KD; q += 2 * w) pr[q] += pr[q + w];
    __builtin_amdgcn_wave_barrier();
    return v + pr[0];
}

template <int KD>
__global__ void __launch_bounds__(64) k_band_chain_local(const int64_t *gfirst, const int32_t *gn, int64_t ngrp,
                                                         const double *s0, const double *T, const int32_t *cmap,
                                                         int dir, const double *Sg, double *sin, double *cg,
                                                         BandRhs rs)
{
    __shared__ double s0l[BAND_G * KD];
    __shared__ double stb[64];   // the state, broadcast to the lanes each step
    s0 += blockIdx.y * rs.s;
    if (Sg) Sg += blockIdx.y * rs.g;
    if (sin) sin += blockIdx.y * rs.s;
    if (cg) cg += blockIdx.y * rs.g;
    int lane = threadIdx.x & 63;
    int64_t g = blockIdx.x;
    if (g >= ngrp) return;
    bool on = lane < KD;
    int64_t c0 = gfirst[g];
    int n = gn[g];
    double st = (Sg && on) ? Sg[g * KD + lane] : 0.0;
    // eight loads in flight before their LDS stores (n <= BAND_G)
    for (int k0 = 0; k0 < KD; k0 += 8) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            int t = lane + 64 * (k0 + k);
            v[k] = k0 + k < KD && t < n * KD ? s0[c0 * KD + t] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            int t = lane + 64 * (k0 + k);
            if (k0 + k < KD && t < n * KD) s0l[t] = v[k];
        }
    }
    // lane t holds the transfer-matrix index of the group's chunk t (n <= 64)
    int myct = lane < n ? (cmap ? cmap[c0 + lane] : (int32_t)(c0 + lane)) : 0;
    __syncthreads();
    int64_t cur = -1;
    double tr[KD];
#pragma unroll
    for (int q = 0; q < KD; ++q) tr[q] = 0.0;
    // Each chunk's incoming state replaces its additive term in LDS (the slot
    // is read first) and the states go out in one coalesced pass after the
    // chain: a global store inside the loop would make every step wait for
    // the previous step's store (the wait before the step's arithmetic counts
    // stores and loads alike).
    for (int t = 0; t < n; ++t) {
        int lc = dir == 0 ? t : n - 1 - t;
        int64_t ct = __builtin_amdgcn_readlane(myct, lc);
        if (ct != cur) {   // uniform over the wave
            cur = ct;
#pragma unroll
            for (int q = 0; q < KD; ++q) tr[q] = on ? T[(ct * KD + lane) * KD + q] : 0.0;
        }
        double v = on ? s0l[lc * KD + lane] : 0.0;
        if (sin && on) s0l[lc * KD + lane] = st;
        st = chain_step<KD>(v, tr, st, stb);
    }
    if (cg && on) cg[g * KD + lane] = st;
    if (sin) {
        __syncthreads();
        for (int t = lane; t < n * KD; t += 64) sin[c0 * KD + t] = s0l[t];
    }
}

__device__ inline int64_t readlane64(int64_t v, int src)
{
    int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffff), src);
    int hi = __builtin_amdgcn_readlane((int)(v >> 32), src);
    return ((int64_t)hi << 32) | (uint32_t)lo;
}

// CPW = 64 / KD chains (group, right-hand side) per wave, lanes k * KD + m
// (m < KD) of chain k, so that the chains' steps share each instruction (one
// chain per wave leaves 55 of 64 lanes idle in every step; with few chains,
// too few packed waves to hide the window staging, band_apply keeps
// k_band_chain_local).  Chain id = rhs *
// ngrp + g.  The chains' additive terms and transfer-matrix indices are
// staged in LDS in windows of CW chunks along the chain (a window of one
// chain is one contiguous piece of s0); each chunk's incoming state replaces
// its term in LDS (the slot is read first) and leaves with the window, so that
// no global store sits inside the sequential loop.
template <int KD>
__global__ void __launch_bounds__(64) k_band_chain_pack(const int64_t *gfirst, const int32_t *gn, int64_t ngrp,
                                                        int nrhs, const double *s0, const double *T,
                                                         const int32_t *cmap, int dir, const double *Sg, double *sin,
                                                         double *cg, BandRhs rs)
{
    constexpr int CPW = 64 / KD, CW = 16;
    __shared__ double s0l[CPW * CW * KD];
    __shared__ int ctl[CPW * CW];
    __shared__ double stb[64];   // the states, broadcast to the lanes each step
    const int lane = threadIdx.x & 63;
    const int k = lane / KD, m = lane - k * KD;
    const int64_t id = (int64_t)blockIdx.x * CPW + k;
    const bool on = k < CPW && id < ngrp * nrhs;
    const int64_t rhs = on ? id / ngrp : 0, g = on ? id - rhs * ngrp : 0;
    const int64_t c0 = on ? gfirst[g] : 0;
    const int n = on ? gn[g] : 0;
    double st = (Sg && on) ? Sg[rhs * rs.g + g * KD + m] : 0.0;
    int nmax = n;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) nmax = max(nmax, __shfl_xor(nmax, o, 64));
    const int kb = k < CPW ? k : 0;   // lanes past the last chain use chain 0's slots
    int cur = -1;
    double tr[KD];
#pragma unroll
    for (int q = 0; q < KD; ++q) tr[q] = 0.0;
    for (int w0 = 0; w0 < nmax; w0 += CW) {
        // window w0 .. w0 + CW - 1 of every chain: chunks lo .. lo + cnt - 1
        __syncthreads();
        for (int kk = 0; kk < CPW; ++kk) {
            const int nk = __builtin_amdgcn_readlane(n, kk * KD);
            const int cnt = min(CW, nk - w0);
            if (cnt <= 0) continue;   // uniform
            const int lo = dir == 0 ? w0 : nk - w0 - cnt;
            const int64_t ck = readlane64(c0, kk * KD) + lo, rk = readlane64(rhs, kk * KD);
            const double *sp = s0 + rk * rs.s + ck * KD;
            double v[(CW * KD + 63) / 64];
#pragma unroll
            for (int j = 0; j < (CW * KD + 63) / 64; ++j) {
                int t = lane + 64 * j;
                v[j] = t < cnt * KD ? sp[t] : 0.0;
            }
            if (lane < cnt) ctl[kk * CW + lane] = cmap ? cmap[ck + lane] : (int32_t)(ck + lane);
#pragma unroll
            for (int j = 0; j < (CW * KD + 63) / 64; ++j) {
                int t = lane + 64 * j;
                if (t < cnt * KD) s0l[kk * CW * KD + t] = v[j];
            }
        }
        __syncthreads();
        const int cnt = min(CW, n - w0), lo = dir == 0 ? w0 : n - w0 - cnt;
        const int tend = min(w0 + CW, nmax);
        for (int t = w0; t < tend; ++t) {
            const bool act = on && t < n;
            const int sl = (dir == 0 ? t : n - 1 - t) - lo;   // slot in the window (act)
            double v = 0.0;
            if (act) {
                int ct = ctl[kb * CW + sl];
                if (ct != cur) {
                    cur = ct;
#pragma unroll
                    for (int q = 0; q < KD; ++q) tr[q] = T[((int64_t)ct * KD + m) * KD + q];
                }
                v = s0l[(kb * CW + sl) * KD + m];
                if (sin) s0l[(kb * CW + sl) * KD + m] = st;
            }
            double ns = chain_step<KD>(v, tr, st, stb, kb * KD);
            if (act) st = ns;
        }
        if (sin) {
            __syncthreads();
            for (int kk = 0; kk < CPW; ++kk) {
                const int nk = __builtin_amdgcn_readlane(n, kk * KD);
                const int ck_cnt = min(CW, nk - w0);
                if (ck_cnt <= 0) continue;
                const int klo = dir == 0 ? w0 : nk - w0 - ck_cnt;
                const int64_t ck = readlane64(c0, kk * KD) + klo, rk = readlane64(rhs, kk * KD);
                for (int t = lane; t < ck_cnt * KD; t += 64) sin[rk * rs.s + ck * KD + t] = s0l[kk * CW * KD + t];
            }
        }
    }
    if (cg && on) cg[rhs * rs.g + g * KD + m] = st;
}

// wave per segment: S_{g+1} = cg[g] + Gm[g] S_g over the segment's groups in
// chain order, staging cg in windows of 64 groups
template <int KD>
__global__ void __launch_bounds__(64) k_band_chain_top(const int64_t *sfirst, int64_t nseg, const double *cg,
                                                       const double *Gm, const int32_t *gmap, int dir, double *Sg,
                                                       BandRhs rs)
{
    constexpr int W = 64;
    __shared__ double cgl[W * KD];
    __shared__ double stb[64];
    cg += blockIdx.y * rs.g;
    Sg += blockIdx.y * rs.g;
    int lane = threadIdx.x & 63;
    int64_t s = blockIdx.x;
    if (s >= nseg) return;
    bool on = lane < KD;
    int64_t g0 = sfirst[s], g1 = sfirst[s + 1];
    int64_t n = g1 - g0;
    double st = 0.0;
    int64_t cur = -1;
    double tr[KD];
#pragma unroll
    for (int q = 0; q < KD; ++q) tr[q] = 0.0;
    for (int64_t t0 = 0; t0 < n; t0 += W) {
        int64_t m = n - t0 < W ? n - t0 : W;
        // window of chain positions t0 .. t0 + m - 1: groups lg0 .. lg0 + m - 1
        int64_t lg0 = dir == 0 ? t0 : n - t0 - m;
        __syncthreads();
        // eight loads in flight before their LDS stores (m <= W)
        for (int k0 = 0; k0 < KD; k0 += 8) {
            double v[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                int64_t t = lane + 64 * (k0 + k);
                v[k] = k0 + k < KD && t < m * KD ? cg[(g0 + lg0) * KD + t] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                int64_t t = lane + 64 * (k0 + k);
                if (k0 + k < KD && t < m * KD) cgl[t] = v[k];
            }
        }
        // lane t holds the group-matrix index of the window's group t
        int mygt = lane < m ? (gmap ? gmap[g0 + lg0 + lane] : (int32_t)(g0 + lg0 + lane)) : 0;
        __syncthreads();
        // (the groups' incoming states staged in LDS as in k_band_chain_local)
        for (int64_t t = 0; t < m; ++t) {
            int lw = (int)(dir == 0 ? t : m - 1 - t);   // position in the window
            int64_t gt = __builtin_amdgcn_readlane(mygt, lw);
            if (gt != cur) {
                cur = gt;
#pragma unroll
                for (int q = 0; q < KD; ++q) tr[q] = on ? Gm[(gt * KD + lane) * KD + q] : 0.0;
            }
            double v = on ? cgl[lw * KD + lane] : 0.0;
            if (on) cgl[lw * KD + lane] = st;
            st = chain_step<KD>(v, tr, st, stb);
        }
        __syncthreads();
        for (int64_t t = lane; t < m * KD; t += 64) Sg[(g0 + lg0) * KD + t] = cgl[t];
    }
}

// wave per group: Gm = T_last ... T_first in chain order (lane m holds row m)
template <int KD>
__global__ void __launch_bounds__(64) k_band_group_mats(const int64_t *gfirst, const int32_t *gn, int64_t ngrp,
                                                        const double *T, int dir, double *Gm)
{
    int lane = threadIdx.x & 63;
    int64_t g = blockIdx.x;
    if (g >= ngrp) return;
    bool on = lane < KD;
    double M[KD];
#pragma unroll
    for (int j = 0; j < KD; ++j) M[j] = (on && j == lane) ? 1.0 : 0.0;
    int64_t c0 = gfirst[g];
    int n = gn[g];
    for (int t = 0; t < n; ++t) {
        int64_t c = dir == 0 ? c0 + t : c0 + n - 1 - t;
        double tr[KD];
#pragma unroll
        for (int q = 0; q < KD; ++q) tr[q] = on ? T[(c * KD + lane) * KD + q] : 0.0;
        double Mn[KD];
#pragma unroll
        for (int j = 0; j < KD; ++j) {
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < KD; ++q) v += tr[q] * __shfl(M[j], q, 64);
            Mn[j] = v;
        }
#pragma unroll
        for (int j = 0; j < KD; ++j) M[j] = Mn[j];
    }
    if (on) {
#pragma unroll
        for (int j = 0; j < KD; ++j) Gm[(g * KD + lane) * KD + j] = M[j];
    }
}

// final forward sweep with true carries gives y (written to z, which the last
// backward sweep reads), then the zero-carry backward sweep of the same chunk,
// on y in LDS, gives its backward partial state
template <int KD, typename V>
__global__ void __launch_bounds__(64) k_band_fwd1_bwd0(const ChunkInfo *ch, int64_t nch, const int32_t *bmap,
                                                       const double *Lb, const int32_t *uidx,
                                                       const double *__restrict__ Lu, const V *r, int64_t vstride,
                                                       const double *sinf, V *z, double *s0b, BandRhs st)
{
    __shared__ V tile[64 * 64];
    r += blockIdx.y * vstride;
    z += blockIdx.y * vstride;
    sinf += blockIdx.y * st.s;
    s0b += blockIdx.y * st.s;
    int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const uint32_t b0 = c < nch ? (uint32_t)ch[c].b0 : 0u;
    const int nc = c < nch ? ch[c].n : 0;
    band_tile_in(r, b0, nc, tile);
    if (c < nch) {
        const int lane = threadIdx.x;
        V *tl = tile + lane * 64;
        double acc[KD];
#pragma unroll
        for (int m = 0; m < KD; ++m) acc[m] = sinf[c * KD + m];
        int32_t ui = uidx[blockIdx.x];
        if (ui >= 0) {
            const double *lu = Lu + (int64_t)ui * BAND_CH * (KD + 1);
            int n = __builtin_amdgcn_readfirstlane(nc);
            fwd_sweep<KD, 1, V>(lu, n, tl, acc, tl, true, false, lane);
#pragma unroll
            for (int m = 0; m < KD; ++m) acc[m] = 0;
            bwd_sweep<KD, 1, V>(lu, n, tl, acc, (V *)nullptr, false, false, lane);
        } else {
            const double *lic = li_chunk(Lb, bmap, c, KD + 1);
            fwd_sweep<KD, 64, V>(lic, nc, tl, acc, tl, true, false, lane);
#pragma unroll
            for (int m = 0; m < KD; ++m) acc[m] = 0;
            bwd_sweep<KD, 64, V>(lic, nc, tl, acc, (V *)nullptr, false, false, lane);
        }
#pragma unroll
        for (int m = 0; m < KD; ++m) s0b[c * KD + m] = acc[m];
    }
    band_tile_out(z, b0, nc, tile);
}

// final backward sweep with true carries: z = x
template <int KD, typename V>
__global__ void __launch_bounds__(64) k_band_bwd1(const ChunkInfo *ch, int64_t nch, const int32_t *bmap,
                                                  const double *Lb, const int32_t *uidx,
                                                  const double *__restrict__ Lu, const double *sinb, V *z,
                                                  int64_t vstride, BandRhs st)
{
    __shared__ V tile[64 * 64];
    sinb += blockIdx.y * st.s;
    z += blockIdx.y * vstride;
    int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const uint32_t b0 = c < nch ? (uint32_t)ch[c].b0 : 0u;
    const int nc = c < nch ? ch[c].n : 0;
    band_tile_in((const V *)z, b0, nc, tile);
    if (c < nch) {
        const int lane = threadIdx.x;
        V *tl = tile + lane * 64;
        double xs[KD];
#pragma unroll
        for (int m = 0; m < KD; ++m) xs[m] = sinb[c * KD + m];
        int32_t ui = uidx[blockIdx.x];
        if (ui >= 0)
            bwd_sweep<KD, 1, V>(Lu + (int64_t)ui * BAND_CH * (KD + 1), __builtin_amdgcn_readfirstlane(nc), tl, xs,
                                tl, true, false, lane);
        else
            bwd_sweep<KD, 64, V>(li_chunk(Lb, bmap, c, KD + 1), nc, tl, xs, tl, true, false, lane);
    }
    band_tile_out(z, b0, nc, tile);
}

// ---------------------------------------------------------------- host side

struct BandImpl {
    double *Li = nullptr;   // interleaved factor
    int64_t NCP = 0;
    ChunkInfo *ch = nullptr;
    int64_t *seg_first = nullptr;
    double *s0f = nullptr, *s0b = nullptr, *sinf = nullptr, *sinb = nullptr;
    int64_t nseg_used = 0;
    // carry-chain groups
    int64_t ngrp = 0;
    int64_t *gfirst = nullptr, *sgfirst = nullptr;
    int32_t *gn = nullptr;
    double *Gf = nullptr, *Gb = nullptr, *cg = nullptr, *Sg = nullptr;
    // runs of bitwise-equal chunks and blocks of the factor (see li_chunk)
    int32_t *cmap = nullptr, *bmap = nullptr;
    // runs of groups with equal chunk sequences (equal group matrices)
    int32_t *gmap = nullptr;
    // uniform blocks: uidx[w] = slot of the block's compact factor copy in Lu, or -1
    int32_t *uidx = nullptr;
    double *Lu = nullptr;
    int rcap = 1;   // right-hand sides the carry-state buffers hold
};

static int kd_class(int kd)
{
    static const int cls[] = {9, 19, 29, 39, 49};
    for (int c : cls)
        if (kd <= c) return c;
    return -1;
}

#define BAND_DISPATCH(KDV, STMT)                         \
    switch (KDV) {                                        \
    case 9: { constexpr int KD = 9; STMT; } break;        \
    case 19: { constexpr int KD = 19; STMT; } break;      \
    case 29: { constexpr int KD = 29; STMT; } break;      \
    case 39: { constexpr int KD = 39; STMT; } break;      \
    default: { constexpr int KD = 49; STMT; } break;      \
    }

// Returns 0 on success; fills info_out[s] (0 ok, >0 failing column) for each segment.
int band_build(hipStream_t s, BandPrec &bp, const Segment *hseg, const Segment *dseg, int64_t nseg,
                    const int *nband_seg_host, const double *invcov_dev, int64_t nof, const double *nna,
                    int *info_out, bool periodic)
{
    int kdmax = 0;
    for (int64_t i = 0; i < nseg; ++i)
        if (!hseg[i].skip && nband_seg_host[i] - 1 > kdmax) kdmax = nband_seg_host[i] - 1;
    int KD = kd_class(kdmax < 1 ? 1 : kdmax);
    if (KD < 0) return -1;   // band wider than 50: not supported
    int K = KD + 1;
    bp.kd = KD;
    bp.chunk = BAND_CH;
    bp.nseg = nseg;
    // storage for the factor over every baseline row of every segment
    int64_t rows_max = 0;
    for (int64_t i = 0; i < nseg; ++i) rows_max = std::max<int64_t>(rows_max, hseg[i].b0 + hseg[i].noba);
    MADAM_CHECK(hipMalloc(&bp.L, sizeof(double) * (rows_max > 0 ? rows_max : 1) * K));
    int *dnb = nullptr, *dinfo = nullptr;
    MADAM_CHECK(hipMalloc(&dnb, sizeof(int) * (nseg > 0 ? nseg : 1)));
    MADAM_CHECK(hipMalloc(&dinfo, sizeof(int) * (nseg > 0 ? nseg : 1)));
    MADAM_CHECK(hipMemcpyAsync(dnb, nband_seg_host, sizeof(int) * nseg, hipMemcpyHostToDevice, s));
    MADAM_CHECK(hipMemsetAsync(dinfo, 0, sizeof(int) * (nseg > 0 ? nseg : 1), s));
    int64_t *dper = nullptr;   // per_src[nseg][BAND_MAXJ], per_q[nseg][BAND_MAXJ]
    int64_t ns1 = nseg > 0 ? nseg : 1, nj = ns1 * BAND_MAXJ;
    MADAM_CHECK(hipMalloc(&dper, sizeof(int64_t) * 2 * nj));
    MADAM_CHECK(hipMemsetAsync(dper, 0, sizeof(int64_t) * 2 * nj, s));
    if (nseg > 0) {
        BAND_DISPATCH(KD, (k_band_factor_w<KD><<<(unsigned)nseg, 64, 0, s>>>(dseg, nseg, dnb, invcov_dev, nof, nna,
                                                                            bp.L, dinfo, periodic ? dper : nullptr,
                                                                            dper + nj)));
        if (periodic) {
            unsigned gy = (unsigned)std::min<int64_t>(nseg, 65535);
            unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>(256, 8192 / nseg));
            k_band_fill_periodic<<<dim3(gx, gy), 256, 0, s>>>(dseg, nseg, K, dper, dper + nj, bp.L);
        }
    }
    MADAM_CHECK(hipGetLastError());
    MADAM_CHECK(hipMemcpyAsync(info_out, dinfo, sizeof(int) * nseg, hipMemcpyDeviceToHost, s));
    std::vector<int64_t> perq(nj, 0);
    MADAM_CHECK(hipMemcpyAsync(perq.data(), dper + nj, sizeof(int64_t) * nj, hipMemcpyDeviceToHost, s));
    MADAM_CHECK(hipStreamSynchronize(s));
    (void)hipFree(dnb);
    (void)hipFree(dinfo);
    (void)hipFree(dper);
    bp.nrows_factored = 0;
    for (int64_t i = 0; i < nseg; ++i)
        if (!hseg[i].skip && nband_seg_host[i] > 0) {
            bp.nrows_factored += hseg[i].noba;
            for (int k = 0; k < BAND_MAXJ; ++k) bp.nrows_factored -= 64 * perq[i * BAND_MAXJ + k];
        }
    // chunk table over the segments that factored
    std::vector<ChunkInfo> ch;
    std::vector<int64_t> first;
    for (int64_t i = 0; i < nseg; ++i) {
        if (hseg[i].skip || nband_seg_host[i] <= 0 || info_out[i] != 0) continue;
        first.push_back((int64_t)ch.size());
        int64_t n = hseg[i].noba;
        for (int64_t r = 0; r < n; r += BAND_CH) {
            ChunkInfo ci;
            ci.b0 = hseg[i].b0 + r;
            ci.n = (int32_t)(n - r < BAND_CH ? n - r : BAND_CH);
            ci.seg = (int32_t)i;
            ci.first = r == 0;
            ci.last = r + BAND_CH >= n;
            ch.push_back(ci);
        }
    }
    first.push_back((int64_t)ch.size());
    bp.nchunk = (int64_t)ch.size();
    bp.nrows = 0;
    for (const ChunkInfo &ci : ch) bp.nrows += ci.n;
    BandImpl *im = new BandImpl();
    im->nseg_used = (int64_t)first.size() - 1;
    // groups of BAND_G chunks inside each segment
    std::vector<int64_t> gfirst, sgfirst;
    std::vector<int32_t> gn;
    for (int64_t si = 0; si < im->nseg_used; ++si) {
        sgfirst.push_back((int64_t)gfirst.size());
        for (int64_t c = first[si]; c < first[si + 1]; c += BAND_G) {
            gfirst.push_back(c);
            gn.push_back((int32_t)std::min<int64_t>(BAND_G, first[si + 1] - c));
        }
    }
    sgfirst.push_back((int64_t)gfirst.size());
    im->ngrp = (int64_t)gfirst.size();
    {
        int64_t ng = im->ngrp > 0 ? im->ngrp : 1;
        MADAM_CHECK(hipMalloc(&im->gfirst, sizeof(int64_t) * ng));
        MADAM_CHECK(hipMalloc(&im->gn, sizeof(int32_t) * ng));
        MADAM_CHECK(hipMalloc(&im->sgfirst, sizeof(int64_t) * sgfirst.size()));
        MADAM_CHECK(hipMalloc(&im->Gf, sizeof(double) * ng * KD * KD));
        MADAM_CHECK(hipMalloc(&im->Gb, sizeof(double) * ng * KD * KD));
        MADAM_CHECK(hipMalloc(&im->cg, sizeof(double) * ng * KD));
        MADAM_CHECK(hipMalloc(&im->Sg, sizeof(double) * ng * KD));
        if (im->ngrp > 0) {
            MADAM_CHECK(hipMemcpyAsync(im->gfirst, gfirst.data(), sizeof(int64_t) * gfirst.size(),
                                       hipMemcpyHostToDevice, s));
            MADAM_CHECK(hipMemcpyAsync(im->gn, gn.data(), sizeof(int32_t) * gn.size(), hipMemcpyHostToDevice, s));
        }
        MADAM_CHECK(hipMemcpyAsync(im->sgfirst, sgfirst.data(), sizeof(int64_t) * sgfirst.size(),
                                   hipMemcpyHostToDevice, s));
    }
    int64_t nc = bp.nchunk > 0 ? bp.nchunk : 1;
    MADAM_CHECK(hipMalloc(&im->ch, sizeof(ChunkInfo) * nc));
    MADAM_CHECK(hipMalloc(&im->seg_first, sizeof(int64_t) * first.size()));
    MADAM_CHECK(hipMalloc(&bp.Tf, sizeof(double) * nc * KD * KD));
    MADAM_CHECK(hipMalloc(&bp.Tb, sizeof(double) * nc * KD * KD));
    MADAM_CHECK(hipMalloc(&im->s0f, sizeof(double) * nc * KD));
    MADAM_CHECK(hipMalloc(&im->s0b, sizeof(double) * nc * KD));
    MADAM_CHECK(hipMalloc(&im->sinf, sizeof(double) * nc * KD));
    MADAM_CHECK(hipMalloc(&im->sinb, sizeof(double) * nc * KD));
    if (bp.nchunk > 0)
        MADAM_CHECK(hipMemcpyAsync(im->ch, ch.data(), sizeof(ChunkInfo) * ch.size(), hipMemcpyHostToDevice, s));
    MADAM_CHECK(hipMemcpyAsync(im->seg_first, first.data(), sizeof(int64_t) * first.size(), hipMemcpyHostToDevice, s));
    if (bp.nchunk > 0) {
        im->NCP = (bp.nchunk + 63) / 64 * 64;
        MADAM_CHECK(hipMalloc(&im->Li, sizeof(double) * im->NCP * BAND_CH * K));
        int64_t tot = im->NCP * BAND_CH * K;
        unsigned g = (unsigned)std::min<int64_t>((tot + 255) / 256, 65535 * 16);
        k_band_relayout<<<g, 256, 0, s>>>(im->ch, bp.nchunk, im->NCP, K, BAND_CH, bp.L, im->Li);
        MADAM_CHECK(hipGetLastError());
        MADAM_CHECK(hipStreamSynchronize(s));
        (void)hipFree(bp.L);   // the sweeps only read the interleaved copy
        bp.L = nullptr;
        int64_t nt = bp.nchunk * KD;
        BAND_DISPATCH(KD, (k_band_response<KD><<<(unsigned)((nt + 63) / 64), 64, 0, s>>>(im->ch, bp.nchunk, im->NCP,
                                                                                         im->Li, bp.Tf, bp.Tb)));
        MADAM_CHECK(hipGetLastError());
        if (im->ngrp > 0) {
            BAND_DISPATCH(KD, ({
                k_band_group_mats<KD><<<(unsigned)im->ngrp, 64, 0, s>>>(im->gfirst, im->gn, im->ngrp, bp.Tf, 0, im->Gf);
                k_band_group_mats<KD><<<(unsigned)im->ngrp, 64, 0, s>>>(im->gfirst, im->gn, im->ngrp, bp.Tb, 1, im->Gb);
            }));
            MADAM_CHECK(hipGetLastError());
        }
        // runs of equal chunks and blocks
        uint8_t *dceq = nullptr;
        MADAM_CHECK(hipMalloc(&dceq, bp.nchunk));
        k_band_chunk_eq<<<(unsigned)std::min<int64_t>((bp.nchunk + 255) / 256, 65535), 256, 0, s>>>(
            im->ch, bp.nchunk, K, im->Li, dceq);
        MADAM_CHECK(hipGetLastError());
        std::vector<uint8_t> ceq(bp.nchunk);
        MADAM_CHECK(hipMemcpyAsync(ceq.data(), dceq, bp.nchunk, hipMemcpyDeviceToHost, s));
        MADAM_CHECK(hipStreamSynchronize(s));
        (void)hipFree(dceq);
        std::vector<int32_t> cmap(bp.nchunk), bmap(im->NCP / 64);
        for (int64_t c = 0; c < bp.nchunk; ++c) {
            cmap[c] = ceq[c] ? cmap[c - 1] : (int32_t)c;
            bp.nchunk_unique += ceq[c] ? 0 : 1;
        }
        for (int64_t w = 0; w < (int64_t)bmap.size(); ++w) {
            bool same = w > 0;
            for (int64_t c = w * 64; c < std::min<int64_t>(w * 64 + 64, bp.nchunk) && same; ++c)
                same = cmap[c] == cmap[c - 64];
            bmap[w] = same ? bmap[w - 1] : (int32_t)w;
            bp.nblock_unique += same ? 0 : 1;
        }
        // blocks whose chunks are all one run read a single column (see li_chunk)
        for (int64_t w = 0; w < (int64_t)bmap.size(); ++w) {
            int64_t c0 = w * 64, c1 = std::min<int64_t>(w * 64 + 64, bp.nchunk);
            bool uni = true;
            for (int64_t c = c0 + 1; c < c1 && uni; ++c) uni = cmap[c] == cmap[c0];
            if (uni) bmap[w] = -1 - bmap[w];
        }
        // one compact copy per distinct uniform block
        std::vector<int32_t> uidx(bmap.size(), -1), reps;
        for (int64_t w = 0; w < (int64_t)bmap.size(); ++w) {
            if (bmap[w] >= 0) continue;
            int32_t wr = -1 - bmap[w];
            if (reps.empty() || reps.back() != wr) reps.push_back(wr);   // runs are contiguous
            uidx[w] = (int32_t)reps.size() - 1;
        }
        MADAM_CHECK(hipMalloc(&im->uidx, sizeof(int32_t) * uidx.size()));
        MADAM_CHECK(hipMemcpyAsync(im->uidx, uidx.data(), sizeof(int32_t) * uidx.size(), hipMemcpyHostToDevice, s));
        MADAM_CHECK(hipMalloc(&im->Lu, sizeof(double) * std::max<size_t>(reps.size(), 1) * BAND_CH * K));
        if (!reps.empty()) {
            int32_t *dreps = nullptr;
            MADAM_CHECK(hipMalloc(&dreps, sizeof(int32_t) * reps.size()));
            MADAM_CHECK(hipMemcpyAsync(dreps, reps.data(), sizeof(int32_t) * reps.size(), hipMemcpyHostToDevice, s));
            int64_t tot = (int64_t)reps.size() * BAND_CH * K;
            k_band_compact_uniform<<<(unsigned)std::min<int64_t>((tot + 255) / 256, 65535), 256, 0, s>>>(
                im->Li, K, dreps, (int64_t)reps.size(), im->Lu);
            MADAM_CHECK(hipGetLastError());
            MADAM_CHECK(hipStreamSynchronize(s));
            (void)hipFree(dreps);
        }
        bp.nblock = (int64_t)bmap.size();
        // groups whose chunks repeat the previous group's one for one have the
        // same group matrices (computed from the same bits in the same order)
        std::vector<int32_t> gmap(std::max<int64_t>(im->ngrp, 1));
        for (int64_t g = 0; g < im->ngrp; ++g) {
            bool same = g > 0 && gn[g] == gn[g - 1];
            for (int t = 0; t < gn[g] && same; ++t) same = cmap[gfirst[g] + t] == cmap[gfirst[g - 1] + t];
            gmap[g] = same ? gmap[g - 1] : (int32_t)g;
        }
        MADAM_CHECK(hipMalloc(&im->gmap, sizeof(int32_t) * gmap.size()));
        MADAM_CHECK(hipMemcpyAsync(im->gmap, gmap.data(), sizeof(int32_t) * gmap.size(), hipMemcpyHostToDevice, s));
        MADAM_CHECK(hipMalloc(&im->cmap, sizeof(int32_t) * cmap.size()));
        MADAM_CHECK(hipMalloc(&im->bmap, sizeof(int32_t) * bmap.size()));
        MADAM_CHECK(hipMemcpyAsync(im->cmap, cmap.data(), sizeof(int32_t) * cmap.size(), hipMemcpyHostToDevice, s));
        MADAM_CHECK(hipMemcpyAsync(im->bmap, bmap.data(), sizeof(int32_t) * bmap.size(), hipMemcpyHostToDevice, s));
    }
    bp.impl = im;
    MADAM_CHECK(hipStreamSynchronize(s));
    return 0;
}

// grow the carry-state buffers to nrhs right-hand sides
void band_reserve(const BandPrec &bp, int nrhs)
{
    BandImpl *im = (BandImpl *)bp.impl;
    if (!im || bp.nchunk == 0 || nrhs <= im->rcap) return;
    int KD = bp.kd;
    int64_t nc = bp.nchunk, ng = im->ngrp > 0 ? im->ngrp : 1;
    for (double **b : {&im->s0f, &im->s0b, &im->sinf, &im->sinb}) {
        (void)hipFree(*b);
        MADAM_CHECK(hipMalloc(b, sizeof(double) * nc * KD * nrhs));
    }
    for (double **b : {&im->cg, &im->Sg}) {
        (void)hipFree(*b);
        MADAM_CHECK(hipMalloc(b, sizeof(double) * ng * KD * nrhs));
    }
    im->rcap = nrhs;
}

// z = M^-1 r on every factored segment, for nrhs right-hand sides at a
// stride of vstride elements (grid dimension y of every launch)
template <typename T>
void band_apply(hipStream_t s, const BandPrec &bp, const T *r, T *z, int nrhs, int64_t vstride)
{
    BandImpl *im = (BandImpl *)bp.impl;
    if (!im || bp.nchunk == 0) return;
    if (nrhs > im->rcap)
        throw HipError(hipErrorInvalidValue, "band_apply: more right-hand sides than band_reserve made room for",
                       __FILE__, __LINE__);
    int KD = bp.kd;
    unsigned R = (unsigned)nrhs;
    dim3 g((unsigned)((bp.nchunk + 63) / 64), R);
    dim3 gs((unsigned)im->nseg_used, R);   // one 64-lane wave per segment
    BandRhs st{bp.nchunk * KD, (im->ngrp > 0 ? im->ngrp : 1) * KD};
    // the LDS tiles hold r and z in their own type (FP32 halves their
    // traffic); the sweeps compute in FP64
    using V = T;
    BAND_DISPATCH(KD, ({
        // 64 / KD chunk-group chains per wave where that leaves at least 1024
        // waves (k_band_chain_pack), else one chain per wave
        const int64_t nchain = im->ngrp * (int64_t)R;
        const bool pack = bp.chain_pack < 0 ? nchain >= 1024 * (64 / KD) : bp.chain_pack != 0;
        dim3 gp((unsigned)((nchain + 64 / KD - 1) / (64 / KD))), gg((unsigned)im->ngrp, R);
        auto chain = [&](const double *s0x, const double *Tx, int dir, const double *Sgx, double *sinx, double *cgx) {
            if (pack)
                k_band_chain_pack<KD><<<gp, 64, 0, s>>>(im->gfirst, im->gn, im->ngrp, nrhs, s0x, Tx, im->cmap, dir, Sgx,
                                                        sinx, cgx, st);
            else
                k_band_chain_local<KD><<<gg, 64, 0, s>>>(im->gfirst, im->gn, im->ngrp, s0x, Tx, im->cmap, dir, Sgx,
                                                         sinx, cgx, st);
        };
        k_band_fwd0<KD, V><<<g, 64, 0, s>>>(im->ch, bp.nchunk, im->bmap, im->Li, im->uidx, im->Lu, r, vstride,
                                            im->s0f, st);
        chain(im->s0f, bp.Tf, 0, nullptr, nullptr, im->cg);
        k_band_chain_top<KD><<<gs, 64, 0, s>>>(im->sgfirst, im->nseg_used, im->cg, im->Gf, im->gmap, 0, im->Sg, st);
        chain(im->s0f, bp.Tf, 0, im->Sg, im->sinf, nullptr);
        k_band_fwd1_bwd0<KD, V><<<g, 64, 0, s>>>(im->ch, bp.nchunk, im->bmap, im->Li, im->uidx, im->Lu, r, vstride,
                                                 im->sinf, z, im->s0b, st);
        chain(im->s0b, bp.Tb, 1, nullptr, nullptr, im->cg);
        k_band_chain_top<KD><<<gs, 64, 0, s>>>(im->sgfirst, im->nseg_used, im->cg, im->Gb, im->gmap, 1, im->Sg, st);
        chain(im->s0b, bp.Tb, 1, im->Sg, im->sinb, nullptr);
        k_band_bwd1<KD, V><<<g, 64, 0, s>>>(im->ch, bp.nchunk, im->bmap, im->Li, im->uidx, im->Lu, im->sinb, z,
                                            vstride, st);
    }));
}

template void band_apply<double>(hipStream_t, const BandPrec &, const double *, double *, int, int64_t);
template void band_apply<float>(hipStream_t, const BandPrec &, const float *, float *, int, int64_t);

void band_free(BandPrec &bp)
{
    BandImpl *im = (BandImpl *)bp.impl;
    if (im) {
        if (im->Li) (void)hipFree(im->Li);
        if (im->cmap) (void)hipFree(im->cmap);
        if (im->gmap) (void)hipFree(im->gmap);
        if (im->uidx) (void)hipFree(im->uidx);
        if (im->Lu) (void)hipFree(im->Lu);
        if (im->bmap) (void)hipFree(im->bmap);
        (void)hipFree(im->ch);
        (void)hipFree(im->seg_first);
        (void)hipFree(im->s0f);
        (void)hipFree(im->s0b);
        (void)hipFree(im->sinf);
        (void)hipFree(im->sinb);
        (void)hipFree(im->gfirst);
        (void)hipFree(im->gn);
        (void)hipFree(im->sgfirst);
        (void)hipFree(im->Gf);
        (void)hipFree(im->Gb);
        (void)hipFree(im->cg);
        (void)hipFree(im->Sg);
        delete im;
    }
    if (bp.L) (void)hipFree(bp.L);
    if (bp.Tf) (void)hipFree(bp.Tf);
    if (bp.Tb) (void)hipFree(bp.Tb);
    bp = BandPrec();
}

}  // namespace madam
