// Deterministic regrouping of the sample stream by map pixel (the CSR that the
// binning and the pixel-side projection read).
//
// The reference owns pixel ranges per OpenMP thread and visits samples in time
// order inside each pixel (src/smadam_routines.f90:111-422, 1038-1148).  Here a
// stable LSD radix sort of (pixel, sample index) pairs gives the same order:
// inside a pixel the entries are ascending in sample index, so every FP64
// per-pixel sum is taken in one fixed order on every run, and consecutive
// entries of a pixel mostly share a baseline (one ring crossing), which keeps
// the baseline gathers of the pixel-side projection in cache.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdint>

#include "madam_internal.hpp"

namespace madam {

static constexpr int CSR_TPB = 256;

__global__ void k_csr_keys(const PixW *pixw, int64_t n, uint32_t flagged_key, uint32_t *key, uint32_t *val)
{
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
        int32_t p = pixw[g].pix;
        key[g] = p >= 0 ? (uint32_t)p : flagged_key;
        val[g] = (uint32_t)g;
    }
}

__device__ inline int64_t csr_find_baseline(const int64_t *bstart, int64_t nbase, int64_t i)
{
    int64_t lo = 0, hi = nbase - 1;
    while (lo < hi) {
        int64_t mid = (lo + hi + 1) >> 1;
        if (bstart[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void k_csr_fill(const PixW *pixw, int64_t nsamp, const int64_t *bstart, int64_t nbase, int64_t total,
                           const uint32_t *src, const int32_t *bslot, CsrEnt *ent)
{
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < total; j += (int64_t)gridDim.x * blockDim.x) {
        uint32_t g = src[j];
        PixW r = pixw[g];
        int64_t det = g / nsamp;
        int64_t k = csr_find_baseline(bstart, nbase, (int64_t)g - det * nsamp);
        int64_t b = det * nbase + k;
        CsrEnt e;
        e.b = bslot ? bslot[b] : (int32_t)b;
        e.w0 = r.w0;
        e.w1 = r.w1;
        e.w2 = r.w2;
        ent[j] = e;
    }
}

static int grid_of(int64_t n)
{
    int64_t g = (n + CSR_TPB - 1) / CSR_TPB;
    if (g < 1) g = 1;
    if (g > 65535 * 8) g = 65535 * 8;
    return (int)g;
}

// device scratch released on every exit path (errors are thrown by MADAM_CHECK)
struct ScratchBuf {
    void *p = nullptr;
    explicit ScratchBuf(size_t bytes) { MADAM_CHECK(hipMalloc(&p, bytes > 0 ? bytes : 1)); }
    ~ScratchBuf()
    {
        if (p) (void)hipFree(p);
    }
    ScratchBuf(const ScratchBuf &) = delete;
    ScratchBuf &operator=(const ScratchBuf &) = delete;
    template <typename T>
    T *as() const { return (T *)p; }
};

void csr_build_sorted(hipStream_t s, const PixW *pixw, int ndet, int64_t nsamp, const int64_t *bstart, int64_t nbase,
                      int64_t npix_map, int64_t total, const int32_t *bslot, CsrEnt *ent, uint32_t *src)
{
    int64_t n = (int64_t)ndet * nsamp;
    if (n == 0) return;
    int end_bit = 1;
    while (end_bit < 32 && ((int64_t)1 << end_bit) <= npix_map) ++end_bit;   // npix_map (flag key) fits
    ScratchBuf keys(sizeof(uint32_t) * n), vals(sizeof(uint32_t) * n), keys_out(sizeof(uint32_t) * n);
    // the sort writes n values; the CSR keeps the first `total` (flagged samples sort last)
    ScratchBuf vals_out(total < n ? sizeof(uint32_t) * n : 0);
    uint32_t *vout = total < n ? vals_out.as<uint32_t>() : src;
    k_csr_keys<<<grid_of(n), CSR_TPB, 0, s>>>(pixw, n, (uint32_t)npix_map, keys.as<uint32_t>(), vals.as<uint32_t>());
    MADAM_CHECK(hipGetLastError());
    size_t tmp_bytes = 0;
    MADAM_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys.as<uint32_t>(), keys_out.as<uint32_t>(),
                                                   vals.as<uint32_t>(), vout, n, 0, end_bit, s));
    ScratchBuf tmp(tmp_bytes);
    MADAM_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, keys.as<uint32_t>(), keys_out.as<uint32_t>(),
                                                   vals.as<uint32_t>(), vout, n, 0, end_bit, s));
    if (vout != src && total > 0)
        MADAM_CHECK(hipMemcpyAsync(src, vout, sizeof(uint32_t) * total, hipMemcpyDeviceToDevice, s));
    if (total > 0) k_csr_fill<<<grid_of(total), CSR_TPB, 0, s>>>(pixw, nsamp, bstart, nbase, total, src, bslot, ent);
    MADAM_CHECK(hipGetLastError());
    MADAM_CHECK(hipStreamSynchronize(s));
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

__global__ void k_bslot_keys(const PixW *pixw, int ndet, int64_t nsamp, const int64_t *bstart, int64_t nbase,
                             uint64_t *key, int32_t *val)
{
    int64_t nb = (int64_t)ndet * nbase;
    for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < nb; b += (int64_t)gridDim.x * blockDim.x) {
        int64_t det = b / nbase, k = b - det * nbase;
        int64_t a = bstart[k], e = bstart[k + 1];
        uint64_t cell = 0xffffffffull;   // empty or flagged middle: after every pixel
        if (e > a) {
            int32_t pix = pixw[det * nsamp + (a + e - 1) / 2].pix;
            if (pix >= 0) cell = (uint64_t)(uint32_t)pix;
        }
        key[b] = (cell << 32) | (uint64_t)(uint32_t)(k * ndet + det);
        val[b] = (int32_t)b;
    }
}

__global__ void k_bslot_invert(const int32_t *bord, int64_t nb, int32_t *bslot)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nb; i += (int64_t)gridDim.x * blockDim.x)
        bslot[bord[i]] = (int32_t)i;
}

void baseline_slots(hipStream_t s, const PixW *pixw, int ndet, int64_t nsamp, const int64_t *bstart, int64_t nbase,
                    int32_t *bord, int32_t *bslot)
{
    int64_t nb = (int64_t)ndet * nbase;
    if (nb == 0) return;
    ScratchBuf keys(sizeof(uint64_t) * nb), keys_out(sizeof(uint64_t) * nb), vals(sizeof(int32_t) * nb);
    k_bslot_keys<<<grid_of(nb), CSR_TPB, 0, s>>>(pixw, ndet, nsamp, bstart, nbase, keys.as<uint64_t>(),
                                                 vals.as<int32_t>());
    MADAM_CHECK(hipGetLastError());
    size_t tmp_bytes = 0;
    MADAM_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys.as<uint64_t>(), keys_out.as<uint64_t>(),
                                                   vals.as<int32_t>(), bord, nb, 0, 64, s));
    ScratchBuf tmp(tmp_bytes);
    MADAM_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, keys.as<uint64_t>(), keys_out.as<uint64_t>(),
                                                   vals.as<int32_t>(), bord, nb, 0, 64, s));
    k_bslot_invert<<<grid_of(nb), CSR_TPB, 0, s>>>(bord, nb, bslot);
    MADAM_CHECK(hipGetLastError());
    MADAM_CHECK(hipStreamSynchronize(s));
}

}  // namespace madam
