// Segmented reductions of a sorted entry stream, one wave per 512-entry tile:
// the two per-iteration pointing projections of the PCG.
//
//   pixel side    m_q = ccinv_q . sum_{j in pixel q} w_j p[b_j]
//                 (baseline_to_map_order0 + ccmultiply,
//                  src/smadam_routines.f90:1038-1148, src/map_routines.f90:24-77)
//   baseline side ap_b += nna_b p_b - dwb_b sum_{i in baseline b} w_i . m[pix_i]
//                 (map_to_baseline_order0, src/smadam_routines.f90:1235-1376)
//
// Entries are consumed in stream order; a segment (a cross pixel's CSR range,
// or a baseline's sample range) starts where the head bitmask has a bit.  The
// wave loads its tile coalesced and transposes the per-entry products through
// LDS so that lane l holds entries t0 + 4 l .. t0 + 4 l + 3: it sums its runs
// sequentially, finishes segments that begin and end inside it, and one
// segmented inclusive scan over the 64 lanes joins runs that cross lanes.  A
// segment that crosses a tile edge leaves a partial sum per tile (slot 1 in
// the tile where it starts, slot 0 in every later tile) and is finished by the
// fix-up pass in tile order.  Every sum is taken in one fixed order, so the
// results are bitwise reproducible; every wave does the same work whatever the
// segment lengths are (pole pixels hold tens of thousands of entries).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "madam_internal.hpp"

namespace madam {

static constexpr int SR_PER_LANE = 8;
static constexpr int SR_TILE = 64 * SR_PER_LANE;
static constexpr int SR_TPB = 256;
static constexpr int SR_WAVES = SR_TPB / 64;

static int sr_grid(int64_t n, int cap = 65535 * 8)
{
    int64_t g = (n + SR_TPB - 1) / SR_TPB;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

__device__ inline int64_t sr_xcd_block(int64_t bid, int64_t nblk)
{
    // workgroups go round-robin over the 8 XCDs; give each XCD a contiguous range
    const int64_t NX = 8;
    int64_t per = nblk / NX;
    if (bid >= per * NX) return bid;
    return (bid % NX) * per + bid / NX;
}

// Lane shifts through DPP (VALU, no LDS round trip); a lane whose source is
// out of range, or whose row is masked off, gets 0, the identity of both
// scans below.
template <int CTRL, int ROWMASK>
__device__ inline int sr_dpp(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWMASK, 0xf, true);
}
template <int CTRL, int ROWMASK>
__device__ inline float sr_dpp(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWMASK, 0xf, true));
}
template <int CTRL, int ROWMASK>
__device__ inline double sr_dpp(double v)
{
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, ROWMASK, 0xf, true);
    hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, ROWMASK, 0xf, true);
    return __hiloint2double(hi, lo);
}
static constexpr int DPP_ROW_SHR = 0x110;   // + n: row_shr:n
static constexpr int DPP_WAVE_SHR1 = 0x138, DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143;

// one step of the segmented inclusive scan: (f, v) = (fo, vo) (+) (f, v)
template <typename T, int NC, int CTRL, int ROWMASK>
__device__ inline void sr_seg_step(T (&v)[NC], int &f)
{
    T o[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) o[c] = sr_dpp<CTRL, ROWMASK>(v[c]);
    int of = sr_dpp<CTRL, ROWMASK>(f);
    if (!f) {
#pragma unroll
        for (int c = 0; c < NC; ++c) v[c] = o[c] + v[c];
    }
    f |= of;
}

// Segmented inclusive scan over the 64 lanes with the operator
// (f1, v1) + (f2, v2) = (f1 | f2, f2 ? v2 : v1 + v2): shifts by 1, 2, 4, 8
// inside rows of 16 lanes, then the row totals across rows (row_bcast:15 into
// rows 1 and 3, row_bcast:31 into rows 2 and 3).
template <typename T, int NC>
__device__ inline void sr_seg_scan(T (&v)[NC], bool head, int lane)
{
    (void)lane;
    int f = head ? 1 : 0;
    sr_seg_step<T, NC, DPP_ROW_SHR + 1, 0xf>(v, f);
    sr_seg_step<T, NC, DPP_ROW_SHR + 2, 0xf>(v, f);
    sr_seg_step<T, NC, DPP_ROW_SHR + 4, 0xf>(v, f);
    sr_seg_step<T, NC, DPP_ROW_SHR + 8, 0xf>(v, f);
    sr_seg_step<T, NC, DPP_ROW_BCAST15, 0xa>(v, f);
    sr_seg_step<T, NC, DPP_ROW_BCAST31, 0xc>(v, f);
}

// inclusive sum over the 64 lanes, same pattern
__device__ inline int sr_scan_sum(int h)
{
    h += sr_dpp<DPP_ROW_SHR + 1, 0xf>(h);
    h += sr_dpp<DPP_ROW_SHR + 2, 0xf>(h);
    h += sr_dpp<DPP_ROW_SHR + 4, 0xf>(h);
    h += sr_dpp<DPP_ROW_SHR + 8, 0xf>(h);
    h += sr_dpp<DPP_ROW_BCAST15, 0xa>(h);
    h += sr_dpp<DPP_ROW_BCAST31, 0xc>(h);
    return h;
}

// Reduce one tile.  prod[u] are the lane's entry contributions (zero past the
// end), hb the head bits of its 8 entries, k0 the ordinal of the segment that
// holds the tile's first entry, cont_last whether the segment that starts last
// in the tile continues into the next tile.  fin(ordinal, sum) finishes a
// segment; part(slot, sum) stores a partial sum of this tile.  Returns the
// number of segment heads in the tile.
template <typename T, int NC, typename Fin, typename Part>
__device__ inline int sr_reduce_tile(const T (&prod)[SR_PER_LANE][NC], unsigned hb, int lane, int64_t k0,
                                      bool head0, bool cont_last, Fin fin, Part part)
{
    int nh = __popc(hb);
    int H = sr_scan_sum(nh);
    int Hex = H - nh;                    // heads in the lanes before this one
    int Htot = __builtin_amdgcn_readlane(H, 63);
    int64_t kb = k0 - (head0 ? 1 : 0);   // ordinal of a head h = kb + #heads in [t0, h]
    T pre[NC], cur[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) pre[c] = cur[c] = (T)0;
    int hk = 0;
#pragma unroll
    for (int u = 0; u < SR_PER_LANE; ++u) {
        if ((hb >> u) & 1u) {
            if (hk == 0) {
#pragma unroll
                for (int c = 0; c < NC; ++c) pre[c] = cur[c];
            } else {
                fin(kb + Hex + hk, cur);   // began and ended inside this lane
            }
#pragma unroll
            for (int c = 0; c < NC; ++c) cur[c] = (T)0;
            ++hk;
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) cur[c] += prod[u][c];
    }
    if (hk == 0) {
#pragma unroll
        for (int c = 0; c < NC; ++c) pre[c] = cur[c];
    }
    T v[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = hk ? cur[c] : pre[c];
    sr_seg_scan<T, NC>(v, hk > 0, lane);
    T sprev[NC];   // the previous lane's scan value (0 in lane 0)
#pragma unroll
    for (int c = 0; c < NC; ++c) sprev[c] = sr_dpp<DPP_WAVE_SHR1, 0xf>(v[c]);
    if (hk > 0) {
        int firstu = __ffs(hb) - 1;
        if (lane > 0 || firstu > 0) {   // the segment that ends before this lane's first head
            T tot[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) tot[c] = pre[c] + sprev[c];
            if (Hex == 0) part(0, tot);   // began in an earlier tile
            else fin(kb + Hex, tot);
        }
    }
    if (lane == 63) {
        if (Htot == 0) part(0, v);
        else if (cont_last) part(1, v);
        else fin(kb + Htot, v);
    }
    return Htot;
}

template <typename T>
__device__ inline void sr_part_store(typename VecT<T>::v4 *part, int64_t j, int slot, const T *v, int nc)
{
    typename VecT<T>::v4 r;
    r.x = v[0];
    r.y = nc > 1 ? v[1] : (T)0;
    r.z = nc > 2 ? v[2] : (T)0;
    r.w = (T)0;
    part[j * 2 + slot] = r;
}

template <typename T>
__device__ inline void sr_ccmul_store(int nmap, const T *cc, const T *m, T *out4)
{
    // lower-triangle storage [c11, c21, c31, c22, c32, c33] of the symmetric inverse
    typename VecT<T>::v4 r;
    if (nmap == 1) {
        r.x = cc[0] * m[0];
        r.y = r.z = (T)0;
    } else {
        r.x = cc[0] * m[0] + cc[1] * m[1] + cc[2] * m[2];
        r.y = cc[1] * m[0] + cc[3] * m[1] + cc[4] * m[2];
        r.z = cc[2] * m[0] + cc[4] * m[1] + cc[5] * m[2];
    }
    r.w = (T)0;
    *(typename VecT<T>::v4 *)out4 = r;
}

// ------------------------------------------------------------ pixel side

// Entries are loaded coalesced (lane l takes t0 + 64 u + l) and their products
// are transposed through a per-wave LDS buffer to the lane-contiguous layout
// of the reduction; the pad of one double per 16 keeps both directions free of
// bank conflicts (16 lanes of a 64-bit access cover the 32 banks once).
static constexpr int SR_PAD = SR_TILE + SR_TILE / 16;

// Orders this wave's LDS accesses: the LDS serves one wave's requests in
// issue order, so a wave that only exchanges data through its own LDS slice
// needs no workgroup barrier, only a fence that keeps the compiler from
// moving accesses across this point.
__device__ inline void sr_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline int sr_pad(int i) { return i + (i >> 4); }
__device__ inline int sr_col(int u, int lane) { return sr_pad(u * 64 + lane); }
__device__ inline int sr_row(int u, int lane) { return sr_pad(lane * SR_PER_LANE + u); }

// head bits of the lane's SR_PER_LANE entries starting at e0 (a multiple of SR_PER_LANE)
// Per-tile metadata in one int64: the ordinal of the segment holding the
// tile's first entry (bits 0-60), whether that entry is a segment head (bit
// 61) and whether the segment that starts last in the tile continues into the
// next tile (bit 62).
__host__ __device__ inline int64_t sr_meta_k0(int64_t m) { return m & ((1ll << 61) - 1); }
__host__ __device__ inline bool sr_meta_head0(int64_t m) { return (m >> 61) & 1; }
__host__ __device__ inline bool sr_meta_cont(int64_t m) { return (m >> 62) & 1; }

// (the head bytes are padded to whole tiles, so the load is always in bounds)
__device__ inline unsigned sr_head_bits(const uint8_t *heads, int64_t e0, int64_t nent)
{
    int64_t rem = nent - e0;
    int nv = rem <= 0 ? 0 : (rem >= SR_PER_LANE ? SR_PER_LANE : (int)rem);
    unsigned byte = heads[e0 >> 3];
    return (byte >> (e0 & 7)) & ((1u << nv) - 1u);
}

// The transpose goes one component at a time through a 4.3 KB per-wave
// buffer, the only LDS the kernel uses.  With eight entries per lane the
// lane-crossing scan is shared by twice the entries that four gave (the
// kernel is bound by VALU issue more than by HBM); it takes 94 registers,
// five waves per SIMD, without spills.
// The pixels finished in a tile have consecutive ordinals: the lane that
// completes a sum deposits it in the transpose buffer at slot k - k0, and one
// uniform pass then has lane l finish ordinal k0 + l, with the inverse matrix
// (compact ordinal order) and the pixel index that lane l loaded before the
// reduction.  Only a tile that finishes more than 64 pixels reads the
// finishing data of the rest where they complete.
template <typename T, int NC, int WAVES, int RL>
__global__ void __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(5, 8))) k_sr_scatter(int64_t nent, const void *ent, const T *pw,
                                                           const uint8_t *heads, const int64_t *tile_k0,
                                                           const int64_t *fix_end, const int32_t *nzq,
                                                           int64_t nzpix, const T *cc_c, T *m4,
                                                           typename VecT<T>::v4 *part)
{
    using V2 = typename VecT<T>::v2;
    __shared__ T tr[WAVES][SR_PAD];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int64_t j = sr_xcd_block(blockIdx.x, gridDim.x) * WAVES + w;
    int64_t t0 = j * SR_TILE;
    if (t0 >= nent) return;   // uniform over the wave; waves synchronise only with themselves
    PixW e[SR_PER_LANE];   // the entries as full records: .pix is the baseline slot
    T x[SR_PER_LANE];
    int64_t k0 = 0;
    bool head0 = false, cont_last = false;
    unsigned hb = 0u;
    int64_t e0 = t0 + lane * SR_PER_LANE;
    {
        // tile metadata first: these loads are independent of the entry stream
        hb = sr_head_bits(heads, e0, nent);
        int64_t meta = tile_k0[j];   // one load: ordinal and the two flags (sr_meta_*)
        k0 = sr_meta_k0(meta);
        head0 = sr_meta_head0(meta);
        cont_last = sr_meta_cont(meta);
        // unconditional loads from clamped indices, so that they all issue
        // before the first wait; entries past the end are neutralised after
#pragma unroll
        for (int u = 0; u < SR_PER_LANE; ++u) {
            int64_t idx = t0 + u * 64 + lane;
            e[u] = load_rec_nt<RL>(ent, idx < nent ? idx : nent - 1);
        }
#pragma unroll
        for (int u = 0; u < SR_PER_LANE; ++u) {
            bool ok = t0 + u * 64 + lane < nent && e[u].pix >= 0;
            T v = pw[ok ? e[u].pix : 0];
            x[u] = ok ? v : (T)0;
        }
    }
    T prod[SR_PER_LANE][NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
#pragma unroll
        for (int u = 0; u < SR_PER_LANE; ++u) {
            float wc = NC == 1 ? 1.0f : (c == 0 ? e[u].w0 : (c == 1 ? e[u].w1 : e[u].w2));
            tr[w][sr_col(u, lane)] = NC == 1 ? x[u] : (T)wc * x[u];
        }
        sr_wave_sync();
#pragma unroll
        for (int u = 0; u < SR_PER_LANE; ++u) prod[u][c] = tr[w][sr_row(u, lane)];
        sr_wave_sync();
    }
    // heads in the tile, from one ballot per entry slot (scalar work)
    int htot = 0;
#pragma unroll
    for (int u = 0; u < SR_PER_LANE; ++u) htot += __popcll(__ballot((hb >> u) & 1u));
    // finished ordinals are kb + 1 .. kb + htot - cont_last with kb = k0 - head0,
    // i.e. slots lo .. hi relative to k0
    int lo = head0 ? 0 : 1;
    int hi = htot - (cont_last ? 1 : 0) - (head0 ? 1 : 0);
    const int nmap = NC == 1 ? 1 : 3;
    auto part_fn = [&](int slot, const T *s) { sr_part_store<T>(part, j, slot, s, NC); };
    if (hi < 64) {
        // the common case: every finished pixel has a slot in the transpose buffer
        // only the lanes that finish a pixel load its inverse matrix (three
        // 16-byte loads of 48 contiguous bytes); issued here, the loads
        // overlap the reduction
        bool fin_lane = lane >= lo && lane <= hi;
        V2 cc[3];
        int32_t q = 0;
#pragma unroll
        for (int i = 0; i < 3; ++i) cc[i].x = cc[i].y = (T)0;
        if (fin_lane) {
            q = nzq[k0 + lane];
            const V2 *src = (const V2 *)(cc_c + (k0 + lane) * 6);
#pragma unroll
            for (int i = 0; i < 3; ++i) cc[i] = src[i];
        }
        sr_reduce_tile<T, NC>(
            prod, hb, lane, k0, head0, cont_last,
            [&](int64_t k, const T *s) {
                int loc = (int)(k - k0);
#pragma unroll
                for (int c = 0; c < NC; ++c) tr[w][loc * NC + c] = s[c];
            },
            part_fn);
        sr_wave_sync();
        if (fin_lane) {
            T s[3] = {(T)0, (T)0, (T)0};
#pragma unroll
            for (int c = 0; c < NC; ++c) s[c] = tr[w][lane * NC + c];
            const T ccd[6] = {cc[0].x, cc[0].y, cc[1].x, cc[1].y, cc[2].x, cc[2].y};
            sr_ccmul_store<T>(nmap, ccd, s, m4 + (int64_t)q * 4);
        }
    } else {
        // a tile of many short segments: finish each where it completes
        sr_reduce_tile<T, NC>(
            prod, hb, lane, k0, head0, cont_last,
            [&](int64_t k, const T *s) { sr_ccmul_store<T>(nmap, cc_c + k * 6, s, m4 + (int64_t)nzq[k] * 4); },
            part_fn);
    }
}

// compact copy of the inverse pixel matrices in non-empty-pixel order
__global__ void k_sr_compact_cc(const int32_t *nzq, int64_t nseg, const double *ccinv6, double *cc_c)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nseg * 6; t += (int64_t)gridDim.x * blockDim.x)
        cc_c[t] = ccinv6[(int64_t)nzq[t / 6] * 6 + t % 6];
}

void segred_compact_cc(hipStream_t s, const int32_t *nzq, int64_t nseg, const double *ccinv6, double *cc_c)
{
    if (nseg <= 0) return;
    k_sr_compact_cc<<<sr_grid(nseg * 6), SR_TPB, 0, s>>>(nzq, nseg, ccinv6, cc_c);
    MADAM_CHECK(hipGetLastError());
}

// Finish the segments that start in a tile and continue into later ones.
// One wave per 64 tiles, lane l taking tile j = 64 w + l: a chain of up to
// SR_FIX_SHORT later tiles is summed by its lane in tile order; a longer one
// (the most-hit pixels span hundreds of tiles) by the whole wave, lane l
// adding tiles j + 1 + l, j + 1 + l + 64, ..., then a butterfly over the
// lanes.  Both orders are fixed, so the sums are reproducible.
static constexpr int SR_FIX_SHORT = 16;
template <typename T>
__global__ void __launch_bounds__(SR_TPB) k_sr_scatter_fix(int nmap, int64_t ntiles, const int64_t *fix_end,
                                                           const int64_t *fix_ord, const int32_t *nzq,
                                                           const typename VecT<T>::v4 *part, const T *cc_c, T *m4)
{
    using V4 = typename VecT<T>::v4;
    int lane = threadIdx.x & 63;
    int64_t w = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    int64_t j = w * 64 + lane;
    if (w * 64 >= ntiles) return;   // uniform over the wave
    int64_t jl = j < ntiles ? fix_end[j] : -1;
    bool longc = jl >= 0 && jl - j > SR_FIX_SHORT;
    if (jl >= 0 && !longc) {
        V4 a = part[j * 2 + 1];
        V4 v[SR_FIX_SHORT];
#pragma unroll
        for (int u = 0; u < SR_FIX_SHORT; ++u)
            if (j + 1 + u <= jl) v[u] = part[(j + 1 + u) * 2];
#pragma unroll
        for (int u = 0; u < SR_FIX_SHORT; ++u)
            if (j + 1 + u <= jl) {
                a.x += v[u].x;
                a.y += v[u].y;
                a.z += v[u].z;
            }
        int64_t kk = fix_ord[j];
        T m[3] = {a.x, a.y, a.z};
        sr_ccmul_store<T>(nmap, cc_c + kk * 6, m, m4 + (int64_t)nzq[kk] * 4);
    }
    uint64_t lm = __ballot(longc);
    while (lm) {
        int src = __builtin_ctzll(lm);
        lm &= lm - 1;
        int64_t j2 = __shfl(j, src, 64), jl2 = __shfl(jl, src, 64);
        T sx = (T)0, sy = (T)0, sz = (T)0;
        for (int64_t jj = j2 + 1 + lane; jj <= jl2; jj += 64) {
            V4 v = part[jj * 2];
            sx += v.x;
            sy += v.y;
            sz += v.z;
        }
        for (int o = 32; o > 0; o >>= 1) {
            sx += __shfl_xor(sx, o, 64);
            sy += __shfl_xor(sy, o, 64);
            sz += __shfl_xor(sz, o, 64);
        }
        if (lane == 0) {
            V4 a = part[j2 * 2 + 1];
            int64_t kk = fix_ord[j2];
            T m[3] = {a.x + sx, a.y + sy, a.z + sz};
            sr_ccmul_store<T>(nmap, cc_c + kk * 6, m, m4 + (int64_t)nzq[kk] * 4);
        }
    }
}

// ------------------------------------------------------------ baseline side

template <int NC>
__global__ void __launch_bounds__(SR_TPB) k_sr_gather(int64_t nent, const PixW *pixw, const double *m4, int64_t ndeg,
                                                      const uint8_t *heads, const int64_t *tile_k0,
                                                      const int64_t *fix_end, const double *dwb, const double *nna,
                                                      const double *p, double *ap, double4 *part)
{
    __shared__ double tr[SR_WAVES][SR_PAD];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int64_t j = sr_xcd_block(blockIdx.x, gridDim.x) * SR_WAVES + w;
    int64_t t0 = j * SR_TILE;
    bool active = t0 < nent;
    if (active) {
        PixW r[SR_PER_LANE];
#pragma unroll
        for (int u = 0; u < SR_PER_LANE; ++u) {
            int64_t idx = t0 + u * 64 + lane;
            if (idx < nent) r[u] = pixw[idx];
            else r[u].pix = -1;
        }
        double4 m[SR_PER_LANE];
#pragma unroll
        for (int u = 0; u < SR_PER_LANE; ++u)
            if (r[u].pix >= 0) m[u] = *(const double4 *)(m4 + (r[u].pix / ndeg) * 4);
#pragma unroll
        for (int u = 0; u < SR_PER_LANE; ++u) {
            double v = 0.0;
            if (r[u].pix >= 0)
                v = NC == 1 ? m[u].x
                            : (double)r[u].w0 * m[u].x + (double)r[u].w1 * m[u].y + (double)r[u].w2 * m[u].z;
            tr[w][sr_col(u, lane)] = v;
        }
    }
    __syncthreads();
    if (!active) return;
    int64_t e0 = t0 + lane * SR_PER_LANE;
    unsigned hb = sr_head_bits(heads, e0, nent);
    double prod[SR_PER_LANE][1];
#pragma unroll
    for (int u = 0; u < SR_PER_LANE; ++u) prod[u][0] = tr[w][sr_row(u, lane)];
    int64_t meta = tile_k0[j];
    int64_t k0 = sr_meta_k0(meta);
    bool head0 = sr_meta_head0(meta);
    bool cont_last = sr_meta_cont(meta);
    sr_reduce_tile<double, 1>(
        prod, hb, lane, k0, head0, cont_last,
        [&](int64_t b, const double *s) { ap[b] = ap[b] + (nna[b] * p[b] - s[0] * dwb[b]); },
        [&](int slot, const double *s) { sr_part_store<double>(part, j, slot, s, 1); });
}

__global__ void k_sr_gather_fix(int64_t ntiles, const int64_t *fix_end, const int64_t *fix_ord, const double4 *part,
                                const double *dwb, const double *nna, const double *p, double *ap)
{
    int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= ntiles) return;
    int64_t jl = fix_end[j];
    if (jl < 0) return;
    double a = part[j * 2 + 1].x;
    for (int64_t jj = j + 1; jj <= jl; ++jj) a += part[jj * 2].x;
    int64_t b = fix_ord[j];
    ap[b] = ap[b] + (nna[b] * p[b] - a * dwb[b]);
}

// ------------------------------------------------------------ set-up

// segstart[k] for k < nseg: first entry of segment k (strictly increasing), segstart[nseg] = nent
__global__ void k_sr_heads(const int64_t *segstart, int64_t nseg, uint32_t *heads32)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nseg; k += (int64_t)gridDim.x * blockDim.x) {
        int64_t s = segstart[k];
        atomicOr(&heads32[s >> 5], 1u << (s & 31));
    }
}

__global__ void k_sr_tiles(const int64_t *segstart, int64_t nseg, int64_t *tile_k0, int64_t *fix_end,
                           int64_t *fix_ord)
{
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nseg; k += (int64_t)gridDim.x * blockDim.x) {
        int64_t a = segstart[k], e = segstart[k + 1];
        int64_t ja = a / SR_TILE, jl = (e - 1) / SR_TILE;
        // tiles whose first entry lies in [a, e)
        for (int64_t t = (a + SR_TILE - 1) / SR_TILE; t * SR_TILE < e; ++t) tile_k0[t] = k;
        if (jl > ja) {
            fix_end[ja] = jl;
            fix_ord[ja] = k;
        }
    }
}

__global__ void k_sr_pack_meta(int64_t nt, const uint8_t *heads, const int64_t *fix_end, int64_t *tile_k0)
{
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nt; j += (int64_t)gridDim.x * blockDim.x) {
        int64_t m = tile_k0[j];
        if (heads[j * (SR_TILE / 8)] & 1u) m |= 1ll << 61;
        if (fix_end[j] >= 0) m |= 1ll << 62;
        tile_k0[j] = m;
    }
}

__global__ void k_sr_fill_i64(int64_t *x, int64_t v, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = v;
}


int64_t segred_ntiles(int64_t nent) { return (nent + SR_TILE - 1) / SR_TILE; }

size_t segred_heads_bytes(int64_t nent)
{
    // lanes read the byte of their first entry; pad to whole 32-bit words of whole tiles
    return (size_t)(segred_ntiles(nent) + 1) * (SR_TILE / 8);
}

void segred_build(hipStream_t s, const int64_t *segstart, int64_t nseg, int64_t nent, uint8_t *heads,
                  int64_t *tile_k0, int64_t *fix_end, int64_t *fix_ord)
{
    int64_t nt = segred_ntiles(nent);
    MADAM_CHECK(hipMemsetAsync(heads, 0, segred_heads_bytes(nent), s));
    if (nt == 0) return;
    k_sr_fill_i64<<<sr_grid(nt), SR_TPB, 0, s>>>(tile_k0, 0, nt);
    k_sr_fill_i64<<<sr_grid(nt), SR_TPB, 0, s>>>(fix_end, -1, nt);
    k_sr_fill_i64<<<sr_grid(nt), SR_TPB, 0, s>>>(fix_ord, -1, nt);
    if (nseg > 0) {
        k_sr_heads<<<sr_grid(nseg), SR_TPB, 0, s>>>(segstart, nseg, (uint32_t *)heads);
        k_sr_tiles<<<sr_grid(nseg), SR_TPB, 0, s>>>(segstart, nseg, tile_k0, fix_end, fix_ord);
    }
    k_sr_pack_meta<<<sr_grid(nt), SR_TPB, 0, s>>>(nt, heads, fix_end, tile_k0);
    MADAM_CHECK(hipGetLastError());
}

// segments of the pixel side: the non-empty cross pixels, in order
__global__ void k_sr_nonempty(const int64_t *off, int64_t npc, int64_t ndeg, int32_t *flag)
{
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < npc; c += (int64_t)gridDim.x * blockDim.x)
        flag[c] = off[(c + 1) * ndeg] > off[c * ndeg] ? 1 : 0;
}

__global__ void k_sr_compact(const int64_t *off, int64_t npc, int64_t ndeg, const int32_t *flag, const int64_t *pos,
                             int32_t *nzq, int64_t *segstart)
{
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < npc; c += (int64_t)gridDim.x * blockDim.x) {
        if (flag[c]) {
            nzq[pos[c]] = (int32_t)c;
            segstart[pos[c]] = off[c * ndeg];
        }
        if (c == npc - 1) segstart[pos[c] + flag[c]] = off[npc * ndeg];
    }
}

void segred_pixel_segments(hipStream_t s, const int64_t *off, int64_t npc, int64_t ndeg, int32_t *flag, int64_t *pos,
                           int64_t *scan_tmp, int64_t nchunks, int32_t *nzq, int64_t *segstart)
{
    k_sr_nonempty<<<sr_grid(npc), SR_TPB, 0, s>>>(off, npc, ndeg, flag);
    launch_exclusive_scan_i32_to_i64(s, flag, npc, pos, scan_tmp, nchunks);
    k_sr_compact<<<sr_grid(npc), SR_TPB, 0, s>>>(off, npc, ndeg, flag, pos, nzq, segstart);
    MADAM_CHECK(hipGetLastError());
}

// segments of the baseline side: every baseline of every detector
__global__ void k_sr_base_starts(const int64_t *bstart, int64_t nbase, int ndet, int64_t nsamp, int64_t *segstart)
{
    int64_t nb = (int64_t)ndet * nbase;
    for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b <= nb; b += (int64_t)gridDim.x * blockDim.x) {
        if (b == nb) segstart[b] = (int64_t)ndet * nsamp;
        else segstart[b] = (b / nbase) * nsamp + bstart[b % nbase];
    }
}

void segred_baseline_segments(hipStream_t s, const int64_t *bstart, int64_t nbase, int ndet, int64_t nsamp,
                              int64_t *segstart)
{
    k_sr_base_starts<<<sr_grid((int64_t)ndet * nbase + 1), SR_TPB, 0, s>>>(bstart, nbase, ndet, nsamp, segstart);
    MADAM_CHECK(hipGetLastError());
}

// ------------------------------------------------------------ launchers

template <typename T>
void launch_scatter_seg(hipStream_t s, int nmap, int64_t nent, const void *ent, int rl, const T *pw,
                        const uint8_t *heads, const int64_t *tile_k0, const int64_t *fix_end, const int64_t *fix_ord,
                        const int32_t *nzq, int64_t nzpix, const T *cc_c, T *m4, void *part_v)
{
    using V4 = typename VecT<T>::v4;
    V4 *part = (V4 *)part_v;
    int64_t nt = segred_ntiles(nent);
    if (nt == 0) return;
    unsigned grid = (unsigned)((nt + SR_WAVES - 1) / SR_WAVES);
#define MADAM_SC(NCV, RLV)                                                                                       \
    k_sr_scatter<T, NCV, SR_WAVES, RLV><<<grid, SR_TPB, 0, s>>>(nent, ent, pw, heads, tile_k0, fix_end, nzq, nzpix,   \
                                                                 cc_c, m4, part)
    if (nmap == 1) {
        if (rl == REC4) MADAM_SC(1, REC4);
        else MADAM_SC(1, REC16);
    } else {
        if (rl == REC12) MADAM_SC(3, REC12);
        else MADAM_SC(3, REC16);
    }
#undef MADAM_SC
    k_sr_scatter_fix<T><<<(unsigned)((nt + SR_TPB - 1) / SR_TPB), SR_TPB, 0, s>>>(nmap, nt, fix_end, fix_ord, nzq,
                                                                               part, cc_c, m4);
}

template void launch_scatter_seg<double>(hipStream_t, int, int64_t, const void *, int, const double *,
                                         const uint8_t *, const int64_t *, const int64_t *, const int64_t *,
                                         const int32_t *, int64_t, const double *, double *, void *);
template void launch_scatter_seg<float>(hipStream_t, int, int64_t, const void *, int, const float *, const uint8_t *,
                                        const int64_t *, const int64_t *, const int64_t *, const int32_t *, int64_t,
                                        const float *, float *, void *);

void launch_gather_seg(hipStream_t s, int nmap, int64_t nent, const PixW *pixw, const double *m4, int64_t ndeg,
                       const uint8_t *heads, const int64_t *tile_k0, const int64_t *fix_end, const int64_t *fix_ord,
                       const double *dwb, const double *nna, const double *p, double *ap, double4 *part)
{
    int64_t nt = segred_ntiles(nent);
    if (nt == 0) return;
    unsigned grid = (unsigned)((nt + SR_WAVES - 1) / SR_WAVES);
    if (nmap == 1)
        k_sr_gather<1><<<grid, SR_TPB, 0, s>>>(nent, pixw, m4, ndeg, heads, tile_k0, fix_end, dwb, nna, p, ap, part);
    else
        k_sr_gather<3><<<grid, SR_TPB, 0, s>>>(nent, pixw, m4, ndeg, heads, tile_k0, fix_end, dwb, nna, p, ap, part);
    k_sr_gather_fix<<<sr_grid(nt, 1 << 30), SR_TPB, 0, s>>>(nt, fix_end, fix_ord, part, dwb, nna, p, ap);
}

}  // namespace madam
