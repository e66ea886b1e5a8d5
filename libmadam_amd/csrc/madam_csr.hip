// Deterministic regrouping of the sample stream by map pixel (the CSR that the
// binning and the pixel-side projection read), and the locality order of the
// baselines.
//
// The reference owns pixel ranges per OpenMP thread and visits samples in time
// order inside each pixel (src/smadam_routines.f90:111-422, 1038-1148).  Here a
// stable LSD radix sort of (pixel, sample index) pairs gives the same order:
// inside a pixel the entries are ascending in sample index, so every FP64
// per-pixel sum is taken in one fixed order on every run, and consecutive
// entries of a pixel mostly share a baseline (one ring crossing), which keeps
// the baseline gathers of the pixel-side projection in cache.  The hit counts
// are the run lengths of the sorted keys (no atomics), so they are exact by
// construction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "madam_context.hpp"

namespace madam {

static constexpr int CSR_TPB = 256;

static int grid_of(int64_t n)
{
    int64_t g = (n + CSR_TPB - 1) / CSR_TPB;
    if (g < 1) g = 1;
    if (g > 65535 * 8) g = 65535 * 8;
    return (int)g;
}

// ------------------------------------------------------------ radix sort
//
// Stable LSD radix sort of 32-bit keys with 32-bit values, 8 bits per pass.
// A pass is three launches:
//   * k_rs_count: every block counts the digits of its fixed range of the
//     input in an LDS histogram per wave (LDS atomics, no global ones) and
//     writes one count per (digit, block), digit-major;
//   * an exclusive scan of those counts: the output position of the first
//     element of each digit of each block;
//   * k_rs_scatter: every block walks its range in tiles of RS_TILE elements.
//     A wave ranks its elements among the lanes with the same digit with
//     eight ballots (a wave-level multisplit) and keeps a running count per
//     digit in LDS, so the rank of an element is the number of earlier
//     elements of its tile with the same digit; the tile is reordered by
//     digit in LDS and written out in runs, each run after the previous
//     tiles' elements of that digit.
// Ties keep their input order in every pass, so the result is the stable sort
// and does not depend on the launch configuration.

static constexpr int RS_TPB = 256;                 // threads per block (4 waves)
static constexpr int RS_WAVES = RS_TPB / 64;
static constexpr int RS_IPT = 16;                  // elements per thread per tile
static constexpr int RS_TILE = RS_TPB * RS_IPT;    // 4096
static constexpr int RS_RADIX = 256;

__global__ void __launch_bounds__(RS_TPB) k_rs_count(const uint32_t *keys, int64_t n, int shift, int64_t per_blk,
                                                     int nblk, int32_t *counts)
{
    __shared__ uint32_t h[RS_WAVES][RS_RADIX];
    int tid = threadIdx.x, w = tid >> 6;
    for (int i = tid; i < RS_WAVES * RS_RADIX; i += RS_TPB) (&h[0][0])[i] = 0;
    __syncthreads();
    int64_t b0 = (int64_t)blockIdx.x * per_blk, b1 = std::min(n, b0 + per_blk);
    // per_blk is a multiple of RS_TILE, so b0 is 16-byte aligned
    int64_t nv = (b1 - b0) >> 2;
    const uint4 *k4 = (const uint4 *)(keys + b0);
    for (int64_t v = tid; v < nv; v += RS_TPB) {
        uint4 k = k4[v];
        atomicAdd(&h[w][(k.x >> shift) & 255], 1u);
        atomicAdd(&h[w][(k.y >> shift) & 255], 1u);
        atomicAdd(&h[w][(k.z >> shift) & 255], 1u);
        atomicAdd(&h[w][(k.w >> shift) & 255], 1u);
    }
    for (int64_t e = b0 + nv * 4 + tid; e < b1; e += RS_TPB) atomicAdd(&h[w][(keys[e] >> shift) & 255], 1u);
    __syncthreads();
    uint32_t c = 0;
    for (int ww = 0; ww < RS_WAVES; ++ww) c += h[ww][tid];
    counts[(int64_t)tid * nblk + blockIdx.x] = (int32_t)c;
}

template <bool IOTA>
__global__ void __launch_bounds__(RS_TPB) k_rs_scatter(const uint32_t *kin, const uint32_t *vin, uint32_t *kout,
                                                       uint32_t *vout, int64_t n, int shift, int64_t per_blk,
                                                       int nblk, const int64_t *offs)
{
    __shared__ uint32_t lk[RS_TILE], lv[RS_TILE];
    __shared__ uint32_t wc[RS_WAVES][RS_RADIX];   // per wave: running count, then exclusive prefix over waves
    __shared__ uint32_t lstart[RS_RADIX], tot[RS_RADIX], wsum[RS_WAVES];
    __shared__ int64_t gbase[RS_RADIX];
    int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    uint64_t lt = (1ull << lane) - 1;
    gbase[tid] = offs[(int64_t)tid * nblk + blockIdx.x];
    int64_t b0 = (int64_t)blockIdx.x * per_blk, b1 = std::min(n, b0 + per_blk);
    for (int64_t t0 = b0; t0 < b1; t0 += RS_TILE) {
        for (int i = tid; i < RS_WAVES * RS_RADIX; i += RS_TPB) (&wc[0][0])[i] = 0;
        __syncthreads();
        uint32_t key[RS_IPT], val[RS_IPT], pos[RS_IPT];
        // element (wave, item, lane) = t0 + (w * RS_IPT + i) * 64 + lane: the
        // tile's order is (wave, item, lane)
#pragma unroll
        for (int i = 0; i < RS_IPT; ++i) {
            int64_t e = t0 + (int64_t)(w * RS_IPT + i) * 64 + lane;
            bool ok = e < b1;
            key[i] = ok ? kin[e] : 0u;
            val[i] = ok ? (IOTA ? (uint32_t)e : vin[e]) : 0u;
        }
#pragma unroll
        for (int i = 0; i < RS_IPT; ++i) {
            int64_t e = t0 + (int64_t)(w * RS_IPT + i) * 64 + lane;
            bool ok = e < b1;
            uint32_t d = (key[i] >> shift) & 255;
            uint64_t peers = __ballot(ok);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                uint64_t bb = __ballot((d >> b) & 1);
                peers &= ((d >> b) & 1) ? bb : ~bb;
            }
            uint32_t rank = (uint32_t)__popcll(peers & lt);
            uint32_t before = ok ? wc[w][d] : 0u;
            pos[i] = before + rank;
            // the lowest lane of each digit group advances the wave's count (the
            // group's reads above are one instruction, issued before this write)
            if (ok && rank == 0) wc[w][d] = before + (uint32_t)__popcll(peers);
        }
        __syncthreads();
        {
            uint32_t s = 0;
            for (int ww = 0; ww < RS_WAVES; ++ww) {
                uint32_t c = wc[ww][tid];
                wc[ww][tid] = s;
                s += c;
            }
            tot[tid] = s;
            // exclusive scan of the digit totals: where each digit's run starts in the tile
            uint32_t v = s;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                uint32_t t = __shfl_up(v, o, 64);
                if (lane >= o) v += t;
            }
            if (lane == 63) wsum[w] = v;
            __syncthreads();
            uint32_t add = 0;
            for (int ww = 0; ww < w; ++ww) add += wsum[ww];
            lstart[tid] = add + v - s;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < RS_IPT; ++i) {
            int64_t e = t0 + (int64_t)(w * RS_IPT + i) * 64 + lane;
            if (e < b1) {
                uint32_t d = (key[i] >> shift) & 255;
                uint32_t lp = lstart[d] + wc[w][d] + pos[i];
                lk[lp] = key[i];
                lv[lp] = val[i];
            }
        }
        __syncthreads();
        int cnt = (int)std::min<int64_t>(RS_TILE, b1 - t0);
        for (int k = tid; k < cnt; k += RS_TPB) {
            uint32_t kk = lk[k];
            uint32_t d = (kk >> shift) & 255;
            int64_t g = gbase[d] + (k - (int)lstart[d]);
            kout[g] = kk;
            vout[g] = lv[k];
        }
        __syncthreads();
        gbase[tid] += tot[tid];
        // the next tile resets wc only after every thread passed this barrier
        __syncthreads();
    }
}

// number of passes for keys below 2^bits
static int rs_passes(int bits) { return std::max(1, (bits + 7) / 8); }

// Sorts n (key, value) pairs by the low `bits` bits of the keys (higher bits
// must be zero).  iota: the values are the input positions (vals_in unread).
// The sorted pairs end in (k[0], v[0]) or (k[1], v[1]); the return value says
// which.  Every buffer holds n elements.
static int radix_sort_pairs(hipStream_t s, uint32_t *k[2], uint32_t *v[2], int64_t n, int bits, bool iota)
{
    if (n <= 0) return 0;
    int64_t ntiles = (n + RS_TILE - 1) / RS_TILE;
    // about 16 blocks per CU on large inputs, whole tiles per block
    int64_t tiles_per_blk = std::max<int64_t>(1, (ntiles + 4095) / 4096);
    int64_t per_blk = tiles_per_blk * RS_TILE;
    int nblk = (int)((n + per_blk - 1) / per_blk);
    int64_t ncount = (int64_t)RS_RADIX * nblk;
    DBuf<int32_t> counts;
    DBuf<int64_t> offs, tmp;
    counts.alloc(ncount);
    offs.alloc(ncount + 1);
    tmp.alloc((ncount + 1 + 4095) / 4096);
    int cur = 0;
    int np = rs_passes(bits);
    for (int p = 0; p < np; ++p) {
        int shift = 8 * p;
        k_rs_count<<<nblk, RS_TPB, 0, s>>>(k[cur], n, shift, per_blk, nblk, counts.p);
        launch_exclusive_scan_i32_to_i64(s, counts.p, ncount, offs.p, tmp.p, tmp.n);
        if (p == 0 && iota)
            k_rs_scatter<true><<<nblk, RS_TPB, 0, s>>>(k[cur], nullptr, k[cur ^ 1], v[cur ^ 1], n, shift, per_blk,
                                                       nblk, offs.p);
        else
            k_rs_scatter<false><<<nblk, RS_TPB, 0, s>>>(k[cur], v[cur], k[cur ^ 1], v[cur ^ 1], n, shift, per_blk,
                                                        nblk, offs.p);
        MADAM_CHECK(hipGetLastError());
        cur ^= 1;
    }
    // the count and offset buffers are freed on return: finish with them first
    MADAM_CHECK(hipStreamSynchronize(s));
    return cur;
}

static int key_bits(int64_t max_key)
{
    int b = 1;
    while (b < 32 && ((int64_t)1 << b) <= max_key) ++b;
    return b;
}

// ------------------------------------------------------------ CSR by pixel

// hits[k] = length of the run of key k in the sorted keys: the run's first
// element adds -start and its last element adds end + 1 (two integer atomics
// per pixel, not per sample; the sum is exact in any order, and the modulo 2^32
// arithmetic of the partial sums cancels); hits must start at zero, and
// flagged samples (flag_key) are not counted
__global__ void k_csr_hits(const uint32_t *keys, int64_t n, uint32_t flag_key, uint32_t *hits)
{
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
        uint32_t k = keys[j];
        if (k == flag_key) continue;
        if (j == 0 || keys[j - 1] != k) atomicAdd(&hits[k], 0u - (uint32_t)j);
        if (j + 1 == n || keys[j + 1] != k) atomicAdd(&hits[k], (uint32_t)(j + 1));
    }
}

// btab[t] = the baseline holding sample t (the last baseline that starts at or
// before t; baselines are common to all detectors)
__global__ void k_baseline_of_sample(const int64_t *bstart, int64_t nbase, int64_t nsamp, int32_t *btab)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nsamp; t += (int64_t)gridDim.x * blockDim.x) {
        int64_t lo = 0, hi = nbase - 1;
        while (lo < hi) {
            int64_t mid = (lo + hi + 1) >> 1;
            if (bstart[mid] <= t) lo = mid;
            else hi = mid - 1;
        }
        btab[t] = (int32_t)lo;
    }
}

// entry j of the CSR: the record of sample g = sorted_vals[j] with its baseline
// slot in place of the pixel, its sample index, and the compact copy in layout rl
// (each block fills a contiguous range of entries, and the ranges of one XCD
// are contiguous: neighbouring pixels share ring crossings, so their entries
// gather the same record lines, which then stay in one L2)
__global__ void k_csr_fill(const PixW *pixw, int64_t nsamp, const int32_t *btab, int64_t nbase, int64_t total,
                           const uint32_t *order, const int32_t *bslot, int rl, CsrEnt *ent, uint32_t *src,
                           void *entc)
{
    int64_t nblk = gridDim.x, per = (total + nblk - 1) / nblk;
    int64_t j0 = xcd_block(blockIdx.x, nblk) * per, j1 = std::min(total, j0 + per);
    for (int64_t j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
        uint32_t g = order[j];
        PixW r = pixw[g];
        int64_t det = g / nsamp;
        int64_t b = det * nbase + btab[(int64_t)g - det * nsamp];
        CsrEnt e;
        e.b = bslot ? bslot[b] : (int32_t)b;
        e.w0 = r.w0;
        e.w1 = r.w1;
        e.w2 = r.w2;
        ent[j] = e;
        src[j] = g;
        if (rl == REC12) {
            Rec12 c;
            c.i = e.b;
            c.w1 = e.w1;
            c.w2 = e.w2;
            ((Rec12 *)entc)[j] = c;
        } else if (rl == REC4) {
            ((int32_t *)entc)[j] = e.b;
        }
    }
}

void csr_alloc(CsrSort &cs, int64_t n, int64_t npix_map)
{
    cs.n = n;
    cs.npix_map = npix_map;
    cs.cur = 0;
    for (int i = 0; i < 2; ++i) {
        cs.k[i].alloc(std::max<int64_t>(n, 1));
        cs.v[i].alloc(std::max<int64_t>(n, 1));
    }
}

void csr_sort(hipStream_t s, CsrSort &cs)
{
    uint32_t *k[2] = {cs.k[0].p, cs.k[1].p}, *v[2] = {cs.v[0].p, cs.v[1].p};
    cs.cur = radix_sort_pairs(s, k, v, cs.n, key_bits(cs.npix_map), true);
}

void csr_hits(hipStream_t s, const CsrSort &cs, int32_t *hits)
{
    MADAM_CHECK(hipMemsetAsync(hits, 0, sizeof(int32_t) * std::max<int64_t>(cs.npix_map, 1), s));
    if (cs.n > 0)
        k_csr_hits<<<grid_of(cs.n), CSR_TPB, 0, s>>>(cs.k[cs.cur].p, cs.n, (uint32_t)cs.npix_map, (uint32_t *)hits);
    MADAM_CHECK(hipGetLastError());
}

void csr_fill(hipStream_t s, const CsrSort &cs, const PixW *pixw, int64_t nsamp, const int64_t *bstart, int64_t nbase,
              int64_t total, const int32_t *bslot, int rl, CsrEnt *ent, uint32_t *src, void *entc)
{
    if (total <= 0) return;
    DBuf<int32_t> btab;
    btab.alloc(std::max<int64_t>(nsamp, 1));
    k_baseline_of_sample<<<grid_of(nsamp), CSR_TPB, 0, s>>>(bstart, nbase, nsamp, btab.p);
    k_csr_fill<<<grid_of(total), CSR_TPB, 0, s>>>(pixw, nsamp, btab.p, nbase, total, cs.v[cs.cur].p, bslot, rl, ent,
                                                  src, entc);
    MADAM_CHECK(hipGetLastError());
    MADAM_CHECK(hipStreamSynchronize(s));   // btab is freed on return
}

// ------------------------------------------------------------ baseline slots
//
// Locality order of the baselines for the pixel-side projection.  The entries
// of one pixel read the baseline vector at every baseline that crosses the
// pixel; in time order those lie a scan period apart for one detector and
// nbase apart across detectors, one cache line each.  Sorting the baselines
// by (map pixel of their middle sample, baseline index in the detector's
// timeline, detector) puts the co-pointed detectors at one time, and the
// successive passes over one place on the sky, next to each other, so that a
// pixel's entries read a few lines (on a ring scan at Nside 512 this cuts the
// projection's L1 accesses by about half).  The order is a relabelling only:
// the entries carry slot numbers, the vectors they index are gathered into
// slot order, and every per-pixel sum keeps its fixed order.
//
// The baselines enter the sort in (baseline index, detector) order and are
// sorted stably by the pixel alone, which gives that three-level order.

__global__ void k_bslot_keys(const PixW *pixw, int ndet, int64_t nsamp, const int64_t *bstart, int64_t nbase,
                             uint32_t empty_key, uint32_t *key, uint32_t *val)
{
    int64_t nb = (int64_t)ndet * nbase;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nb; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t k = i / ndet, det = i - k * ndet;
        int64_t a = bstart[k], e = bstart[k + 1];
        uint32_t cell = empty_key;   // empty or flagged middle: after every pixel
        if (e > a) {
            int32_t pix = pixw[det * nsamp + (a + e - 1) / 2].pix;
            if (pix >= 0) cell = (uint32_t)pix;
        }
        key[i] = cell;
        val[i] = (uint32_t)(det * nbase + k);
    }
}

__global__ void k_bslot_invert(const uint32_t *bord, int64_t nb, int32_t *bord_out, int32_t *bslot)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nb; i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t b = bord[i];
        bord_out[i] = (int32_t)b;
        bslot[b] = (int32_t)i;
    }
}

void baseline_slots(hipStream_t s, const PixW *pixw, int ndet, int64_t nsamp, const int64_t *bstart, int64_t nbase,
                    int64_t npix_map, int32_t *bord, int32_t *bslot)
{
    int64_t nb = (int64_t)ndet * nbase;
    if (nb == 0) return;
    DBuf<uint32_t> kb[2], vb[2];
    for (int i = 0; i < 2; ++i) {
        kb[i].alloc(nb);
        vb[i].alloc(nb);
    }
    k_bslot_keys<<<grid_of(nb), CSR_TPB, 0, s>>>(pixw, ndet, nsamp, bstart, nbase, (uint32_t)npix_map, kb[0].p,
                                                 vb[0].p);
    MADAM_CHECK(hipGetLastError());
    uint32_t *k[2] = {kb[0].p, kb[1].p}, *v[2] = {vb[0].p, vb[1].p};
    int cur = radix_sort_pairs(s, k, v, nb, key_bits(npix_map), false);
    k_bslot_invert<<<grid_of(nb), CSR_TPB, 0, s>>>(v[cur], nb, bord, bslot);
    MADAM_CHECK(hipGetLastError());
    MADAM_CHECK(hipStreamSynchronize(s));
}

}  // namespace madam
