// HIP kernels for the FP64 destriping path on gfx950 (MI355X, wave64).
//
// Reference behaviour (paths relative to the libmadam source tree):
//   check_files flagging        src/read_data.f90:203-323
//   pixel_matrix / bin_tod /    src/smadam_routines.f90:111-422
//     count_hits
//   invert_pixelmatrix_*        src/map_routines.f90:83-321, src/matrix.f90:15-327
//   ccmultiply                  src/map_routines.f90:24-77
//   initialize_a                src/smadam_routines.f90:428-671
//   baseline_to_map_order0_*    src/smadam_routines.f90:1038-1148
//   map_to_baseline_order0_*    src/smadam_routines.f90:1235-1376
//   PCG vector steps            src/smadam_routines.f90:732-922
//   subtract_baselines_a        src/smadam_routines.f90:1489-1539
//   clean_tod                   src/smadam_routines.f90:1545-1585
//   trim_interval               src/noise_routines.f90:764-813
//
// Design: the reference scatters every sample into a thread-owned pixel range
// (each OpenMP thread re-reads all samples).  Here the pointing is regrouped
// once per scan into a CSR by map pixel, so P^T W F p becomes a segmented
// gather-reduction per pixel (no atomics, deterministic for a fixed CSR), and
// F^T W P m is a segmented reduction per baseline over the time-ordered stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "madam_internal.hpp"

namespace madam {

#define F32(x) ((double)(float)(x))

static constexpr int TPB = 256;

static inline int grid_for(int64_t n, int per_block, int cap = 65535 * 4)
{
    int64_t g = (n + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

__device__ inline int find_interval(const int64_t *ivs, int nival, int64_t i)
{
    // ivs[0..nival] cumulative interval starts, ivs[0] = 0
    if (i < ivs[0] || i >= ivs[nival]) return -1;
    int lo = 0, hi = nival - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (ivs[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ inline int64_t find_baseline(const int64_t *bstart, int64_t nbase, int64_t i)
{
    int64_t lo = 0, hi = nbase - 1;
    while (lo < hi) {
        int64_t mid = (lo + hi + 1) >> 1;
        if (bstart[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

static constexpr int ITEMS_PER_BLOCK = 64;

template <typename T>
__device__ inline T wave_sum(T v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <int G, typename T>
__device__ inline T group_sum(T v)
{
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
    return v;
}

// ---------------------------------------------------------------- check_files

__global__ void k_flag_and_count(int32_t *pix, const float *w, int nnz, int ndet, int64_t nsamp,
                                 const int64_t *ivs, int nival, unsigned long long *valid,
                                 unsigned long long *nzero)
{
    const int64_t TILE = 4096;
    int64_t tpd = (nsamp + TILE - 1) / TILE;
    __shared__ unsigned long long s_c[TPB / 64], s_z[TPB / 64];
    for (int64_t t = blockIdx.x; t < (int64_t)ndet * tpd; t += gridDim.x) {
        int det = (int)(t / tpd);
        int64_t i0 = (t % tpd) * TILE, i1 = i0 + TILE < nsamp ? i0 + TILE : nsamp;
        int j0 = find_interval(ivs, nival, i0);
        int64_t jend = j0 >= 0 ? ivs[j0 + 1] : -1;
        unsigned long long c = 0, zc = 0;
        for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
            int64_t g = (int64_t)det * nsamp + i;
            int32_t p = pix[g];
            bool allz = true;
            for (int j = 0; j < nnz; ++j) allz = allz && (w[g * nnz + j] == 0.0f);
            if (p >= 0 && allz) {
                pix[g] = -1;
                p = -1;
                zc++;
            }
            if (p >= 0) {
                if (j0 >= 0 && i < jend) c++;
                else {
                    int j = find_interval(ivs, nival, i);
                    if (j >= 0) atomicAdd(&valid[(int64_t)det * nival + j], 1ull);
                }
            }
        }
        c = wave_sum(c);
        zc = wave_sum(zc);
        int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        if (lane == 0) { s_c[wv] = c; s_z[wv] = zc; }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long cs = 0, zs = 0;
            for (int k = 0; k < TPB / 64; ++k) { cs += s_c[k]; zs += s_z[k]; }
            if (cs && j0 >= 0) atomicAdd(&valid[(int64_t)det * nival + j0], cs);
            if (zs) atomicAdd(nzero, zs);
        }
        __syncthreads();
    }
}

void launch_flag_and_count(hipStream_t s, int32_t *pix, const float *w, int nnz, int ndet, int64_t nsamp,
                           const int64_t *ivs, int nival, unsigned long long *valid, unsigned long long *nzero)
{
    int64_t tiles = (int64_t)ndet * ((nsamp + 4095) / 4096);
    k_flag_and_count<<<grid_for(tiles, 1), TPB, 0, s>>>(pix, w, nnz, ndet, nsamp, ivs, nival, valid, nzero);
}

__global__ void k_flag_intervals(int32_t *pix, int64_t nsamp, const int64_t *ivs, int nival, int ndet,
                                 const int8_t *flag)
{
    int64_t n = (int64_t)ndet * nsamp;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
        int det = (int)(g / nsamp);
        int j = find_interval(ivs, nival, g % nsamp);
        if (j >= 0 && flag[(int64_t)det * nival + j]) pix[g] = -1;
    }
}

void launch_flag_intervals(hipStream_t s, int32_t *pix, int64_t nsamp, const int64_t *ivs, int nival, int ndet,
                           const int8_t *flag)
{
    k_flag_intervals<<<grid_for((int64_t)ndet * nsamp, TPB), TPB, 0, s>>>(pix, nsamp, ivs, nival, ndet, flag);
}

// ---------------------------------------------------------------- packing

__global__ void k_pack(const int32_t *pix, const float *w, int nnz, int nmap, int ndet, int64_t nsamp,
                       const int64_t *ivs, int nival, const double *dw_ival, int nsubchunk, int pass, PixW *out,
                       uint32_t *key, uint32_t flag_key, void *pixc, unsigned long long *nonunit)
{
    int64_t n = (int64_t)ndet * nsamp;
    unsigned long long nu = 0;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
        int det = (int)(g / nsamp);
        int64_t i = g % nsamp;
        int32_t p = pix[g];
        int j = find_interval(ivs, nival, i);
        bool act = p >= 0 && j >= 0 && dw_ival[(int64_t)det * nival + j] != 0.0;
        if (act && pass != 0) {
            // subchunk of sample i (src/parameter_control.f90:822-836)
            int64_t len = ivs[j + 1] - ivs[j], sublen = len / nsubchunk, o = i - ivs[j];
            int64_t isub = sublen == 0 ? nsubchunk : o / sublen + 1;
            if (isub > nsubchunk) isub = nsubchunk;
            act = isub == pass;
        }
        PixW r;
        r.pix = act ? p : -1;
        if (nmap == 3) {
            r.w0 = w[g * nnz];
            r.w1 = w[g * nnz + 1];
            r.w2 = w[g * nnz + 2];
        } else {
            r.w0 = w[g * nnz];
            r.w1 = 0.0f;
            r.w2 = 0.0f;
        }
        out[g] = r;
        key[g] = r.pix >= 0 ? (uint32_t)r.pix : flag_key;
        nu += (r.pix >= 0 && r.w0 != 1.0f) ? 1 : 0;
        if (nmap == 3) {
            Rec12 c;
            c.i = r.pix;
            c.w1 = r.w1;
            c.w2 = r.w2;
            ((Rec12 *)pixc)[g] = c;
        } else {
            ((int32_t *)pixc)[g] = r.pix;
        }
    }
    for (int o = 32; o > 0; o >>= 1) nu += __shfl_xor(nu, o, 64);
    if ((threadIdx.x & 63) == 0 && nu) atomicAdd(nonunit, nu);
}

void launch_pack(hipStream_t s, const int32_t *pix, const float *w, int nnz, int nmap, int ndet, int64_t nsamp,
                 const int64_t *ivs, int nival, const double *dw_ival, int nsubchunk, int pass, PixW *pixw,
                 uint32_t *key, uint32_t flag_key, void *pixc, unsigned long long *nonunit)
{
    k_pack<<<grid_for((int64_t)ndet * nsamp, TPB), TPB, 0, s>>>(pix, w, nnz, nmap, ndet, nsamp, ivs, nival,
                                                                dw_ival, nsubchunk, pass, pixw, key, flag_key, pixc,
                                                                nonunit);
}

// ---------------------------------------------------------------- CSR by pixel

static constexpr int SCAN_CHUNK = 4096;

__global__ void k_scan_blocks(const int32_t *in, int64_t n, int64_t *bsum)
{
    int64_t base = (int64_t)blockIdx.x * SCAN_CHUNK;
    int64_t acc = 0;
    for (int k = threadIdx.x; k < SCAN_CHUNK; k += blockDim.x) {
        int64_t i = base + k;
        if (i < n) acc += in[i];
    }
    acc = wave_sum(acc);
    __shared__ int64_t sh[TPB / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t t = 0;
        for (int k = 0; k < TPB / 64; ++k) t += sh[k];
        bsum[blockIdx.x] = t;
    }
}

__global__ void k_scan_top(int64_t *bsum, int64_t nb)
{
    // single block, exclusive scan in place, sequential over chunks of blockDim
    __shared__ int64_t sh[1024];
    int64_t carry = 0;
    for (int64_t base = 0; base < nb; base += blockDim.x) {
        int64_t i = base + threadIdx.x;
        int64_t v = i < nb ? bsum[i] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < (int)blockDim.x; o <<= 1) {
            int64_t t = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) bsum[i] = carry + sh[threadIdx.x] - v;
        int64_t tot = sh[blockDim.x - 1];
        __syncthreads();
        carry += tot;
    }
}

__global__ void k_scan_final(const int32_t *in, int64_t n, const int64_t *bsum, int64_t *out)
{
    // one block per chunk; each thread owns SCAN_CHUNK / blockDim consecutive items
    const int per = SCAN_CHUNK / TPB;
    int64_t base = (int64_t)blockIdx.x * SCAN_CHUNK + threadIdx.x * per;
    int64_t loc[per];
    int64_t s = 0;
    for (int k = 0; k < per; ++k) {
        int64_t i = base + k;
        loc[k] = s;
        s += i < n ? in[i] : 0;
    }
    __shared__ int64_t sh[TPB];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < TPB; o <<= 1) {
        int64_t t = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    int64_t off = bsum[blockIdx.x] + sh[threadIdx.x] - s;
    for (int k = 0; k < per; ++k) {
        int64_t i = base + k;
        if (i <= n) out[i] = off + loc[k];   // out has n + 1 entries, out[n] = total
    }
}

void launch_exclusive_scan_i32_to_i64(hipStream_t s, const int32_t *in, int64_t n, int64_t *out, int64_t *tmp,
                                      int64_t tmp_len)
{
    // out[0..n] (n + 1 entries); in is read up to n-1; tmp holds one int64 per chunk
    int64_t nb = (n + 1 + SCAN_CHUNK - 1) / SCAN_CHUNK;
    (void)tmp_len;
    k_scan_blocks<<<(unsigned)nb, TPB, 0, s>>>(in, n, tmp);
    k_scan_top<<<1, 1024, 0, s>>>(tmp, nb);
    k_scan_final<<<(unsigned)nb, TPB, 0, s>>>(in, n, tmp, out);
}

// ---------------------------------------------------------------- binning

// pixel_matrix + bin_tod, one wave per map pixel over its CSR range.
// cc6 = upper triangle (11,12,13,22,23,33), rhs4 = (T, Q, U, 0).
template <typename TT>
__global__ void k_bin(int nmap, int64_t npix, const int64_t *off, const CsrEnt *ent, const uint32_t *src,
                      const TT *tod, const double *dwb, const double *sdwb, double *cc6, double *rhs4)
{
    int lane = threadIdx.x & 63;
    int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t q = wave; q < npix; q += nwaves) {
        int64_t j0 = off[q], j1 = off[q + 1];
        double c0 = 0, c1 = 0, c2 = 0, c3 = 0, c4 = 0, c5 = 0, r0 = 0, r1 = 0, r2 = 0;
        for (int64_t j = j0 + lane; j < j1; j += 64) {
            CsrEnt e = ent[j];
            double d = (double)tod[src[j]];
            double dw = dwb[e.b];
            if (nmap == 3) {
                double sq = sdwb[e.b];
                double a0 = (double)e.w0 * sq, a1 = (double)e.w1 * sq, a2 = (double)e.w2 * sq;
                c0 += a0 * a0; c1 += a0 * a1; c2 += a0 * a2;
                c3 += a1 * a1; c4 += a1 * a2; c5 += a2 * a2;
                r0 += (double)e.w0 * d * dw;
                r1 += (double)e.w1 * d * dw;
                r2 += (double)e.w2 * d * dw;
            } else {
                c0 += dw;
                r0 += d * dw;
            }
        }
        c0 = wave_sum(c0); r0 = wave_sum(r0);
        if (nmap == 3) {
            c1 = wave_sum(c1); c2 = wave_sum(c2); c3 = wave_sum(c3); c4 = wave_sum(c4); c5 = wave_sum(c5);
            r1 = wave_sum(r1); r2 = wave_sum(r2);
        }
        if (lane == 0) {
            double *C = cc6 + q * 6;
            C[0] = c0; C[1] = c1; C[2] = c2; C[3] = c3; C[4] = c4; C[5] = c5;
            double *R = rhs4 + q * 4;
            R[0] = r0; R[1] = r1; R[2] = r2; R[3] = 0.0;
        }
    }
}

template <typename TT>
void launch_bin(hipStream_t s, int nmap, int64_t npix_map, const int64_t *off, const CsrEnt *ent, const uint32_t *src,
                const TT *tod, const double *dwb, const double *sdwb, double *cc6, double *rhs4)
{
    k_bin<TT><<<grid_for(npix_map, TPB / 64, 1 << 20), TPB, 0, s>>>(nmap, npix_map, off, ent, src, tod, dwb, sdwb,
                                                                   cc6, rhs4);
}
template void launch_bin<double>(hipStream_t, int, int64_t, const int64_t *, const CsrEnt *, const uint32_t *,
                                 const double *, const double *, const double *, double *, double *);
template void launch_bin<float>(hipStream_t, int, int64_t, const int64_t *, const CsrEnt *, const uint32_t *,
                                const float *, const double *, const double *, double *, double *);

// pixel_matrix + count_hits + bin_tod over the samples of one subset (detector
// set x survey spans, src/smadam.F90:954-1257): smask[g] != 0 selects sample g
__global__ void k_bin_subset(int nmap, int64_t npix, const int64_t *off, const CsrEnt *ent, const uint32_t *src,
                             const uint8_t *smask, const double *tod, const double *dwb, const double *sdwb,
                             double *cc6, double *rhs4, int32_t *hits)
{
    int lane = threadIdx.x & 63;
    int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t q = wave; q < npix; q += nwaves) {
        int64_t j0 = off[q], j1 = off[q + 1];
        double c0 = 0, c1 = 0, c2 = 0, c3 = 0, c4 = 0, c5 = 0, r0 = 0, r1 = 0, r2 = 0;
        int nh = 0;
        for (int64_t j = j0 + lane; j < j1; j += 64) {
            uint32_t g = src[j];
            if (!smask[g]) continue;
            CsrEnt e = ent[j];
            double d = tod[g];
            double dw = dwb[e.b];
            ++nh;
            if (nmap == 3) {
                double sq = sdwb[e.b];
                double a0 = (double)e.w0 * sq, a1 = (double)e.w1 * sq, a2 = (double)e.w2 * sq;
                c0 += a0 * a0; c1 += a0 * a1; c2 += a0 * a2;
                c3 += a1 * a1; c4 += a1 * a2; c5 += a2 * a2;
                r0 += (double)e.w0 * d * dw;
                r1 += (double)e.w1 * d * dw;
                r2 += (double)e.w2 * d * dw;
            } else {
                c0 += dw;
                r0 += d * dw;
            }
        }
        c0 = wave_sum(c0); r0 = wave_sum(r0);
        nh = wave_sum(nh);
        if (nmap == 3) {
            c1 = wave_sum(c1); c2 = wave_sum(c2); c3 = wave_sum(c3); c4 = wave_sum(c4); c5 = wave_sum(c5);
            r1 = wave_sum(r1); r2 = wave_sum(r2);
        }
        if (lane == 0) {
            double *C = cc6 + q * 6;
            C[0] = c0; C[1] = c1; C[2] = c2; C[3] = c3; C[4] = c4; C[5] = c5;
            double *R = rhs4 + q * 4;
            R[0] = r0; R[1] = r1; R[2] = r2; R[3] = 0.0;
            hits[q] = nh;
        }
    }
}

void launch_bin_subset(hipStream_t s, int nmap, int64_t npix_map, const int64_t *off, const CsrEnt *ent,
                       const uint32_t *src, const uint8_t *smask, const double *tod, const double *dwb,
                       const double *sdwb, double *cc6, double *rhs4, int32_t *hits)
{
    k_bin_subset<<<grid_for(npix_map, TPB / 64, 1 << 20), TPB, 0, s>>>(nmap, npix_map, off, ent, src, smask, tod,
                                                                      dwb, sdwb, cc6, rhs4, hits);
}

__global__ void k_degrade(int nmap, int64_t npc, int64_t ndeg, const double *cc6, const double *rhs4, double *cc6x,
                          double *rhs4x)
{
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < npc; c += (int64_t)gridDim.x * blockDim.x) {
        double C[6] = {0, 0, 0, 0, 0, 0}, R[4] = {0, 0, 0, 0};
        for (int64_t q = c * ndeg; q < (c + 1) * ndeg; ++q) {
            for (int a = 0; a < 6; ++a) C[a] += cc6[q * 6 + a];
            for (int a = 0; a < 3; ++a) R[a] += rhs4[q * 4 + a];
        }
        for (int a = 0; a < 6; ++a) cc6x[c * 6 + a] = C[a];
        for (int a = 0; a < 4; ++a) rhs4x[c * 4 + a] = R[a];
    }
}

void launch_degrade(hipStream_t s, int nmap, int64_t npix_cross, int64_t ndeg, const double *cc6, const double *rhs4,
                    double *cc6x, double *rhs4x)
{
    k_degrade<<<grid_for(npix_cross, TPB), TPB, 0, s>>>(nmap, npix_cross, ndeg, cc6, rhs4, cc6x, rhs4x);
}

// ---------------------------------------------------------------- 3x3 inversion

// Cyclic Jacobi eigen-decomposition of symmetric n x n (n <= 3), a row-major.
__device__ bool jacobi_eig(int n, double *a, double *ev, double *v)
{
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) v[i * 3 + j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0, diag = 0;
        for (int i = 0; i < n; ++i) {
            diag += a[i * 3 + i] * a[i * 3 + i];
            for (int j = i + 1; j < n; ++j) off += a[i * 3 + j] * a[i * 3 + j];
        }
        if (off == 0.0 || off < 1e-34 * diag) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = a[p * 3 + q];
                if (apq == 0.0) continue;
                double app = a[p * 3 + p], aqq = a[q * 3 + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
                for (int k = 0; k < n; ++k) {
                    double akp = a[k * 3 + p], akq = a[k * 3 + q];
                    a[k * 3 + p] = cs * akp - sn * akq;
                    a[k * 3 + q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = a[p * 3 + k], aqk = a[q * 3 + k];
                    a[p * 3 + k] = cs * apk - sn * aqk;
                    a[q * 3 + k] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double vkp = v[k * 3 + p], vkq = v[k * 3 + q];
                    v[k * 3 + p] = cs * vkp - sn * vkq;
                    v[k * 3 + q] = sn * vkp + cs * vkq;
                }
            }
    }
    for (int i = 0; i < n; ++i) ev[i] = a[i * 3 + i];
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (ev[j] < ev[i]) {
                double t = ev[i]; ev[i] = ev[j]; ev[j] = t;
                for (int k = 0; k < n; ++k) { t = v[k * 3 + i]; v[k * 3 + i] = v[k * 3 + j]; v[k * 3 + j] = t; }
            }
    for (int i = 0; i < n; ++i)
        if (isnan(ev[i])) return false;
    return true;
}

// invert_eig on the sub-block given by index list g[0..nf) of the full
// (col-major, C[a + 3b]) symmetric matrix using its upper triangle.
// Returns sing; writes the inverse of the sub-block into C.
__device__ bool inv_eig_block(double *C, const int *g, int nf, double *cdet, double *rcond)
{
    double a[9], v[9], ev[3];
    for (int i = 0; i < nf; ++i)
        for (int j = 0; j < nf; ++j) {
            int gi = g[i < j ? i : j], gj = g[i < j ? j : i];
            a[i * 3 + j] = C[gi + 3 * gj];
        }
    bool ok = jacobi_eig(nf, a, ev, v);
    double det = 1, mn = ev[0], mx = ev[0];
    for (int i = 0; i < nf; ++i) { det *= ev[i]; mn = fmin(mn, ev[i]); mx = fmax(mx, ev[i]); }
    *cdet = det;
    *rcond = mn / mx;
    if (!ok || mx < F32(1e-30) || *rcond < F32(1e-12)) {
        for (int i = 0; i < 9; ++i) C[i] = 0;
        *cdet = 0;
        *rcond = 0;
        return true;
    }
    for (int i = 0; i < nf; ++i)
        for (int j = 0; j < nf; ++j) {
            double s = 0;
            for (int k = 0; k < nf; ++k) s += (v[i * 3 + k] / ev[k]) * v[j * 3 + k];
            C[g[i] + 3 * g[j]] = s;
        }
    return false;
}

__device__ int inv_eig_s_block(double *C, int s0, double plim, double *cdet, double *rcond)
{
    int n = 3 - s0;
    double a[9], v[9], ev[3];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            int gi = s0 + (i < j ? i : j), gj = s0 + (i < j ? j : i);
            a[i * 3 + j] = C[gi + 3 * gj];
        }
    bool ok = jacobi_eig(n, a, ev, v);
    double det = 1, mn = ev[0], mx = ev[0];
    for (int i = 0; i < n; ++i) { det *= ev[i]; mn = fmin(mn, ev[i]); mx = fmax(mx, ev[i]); }
    *cdet = det;
    *rcond = mn / mx;
    if (!ok || mx < F32(1e-30)) {
        for (int i = 0; i < 9; ++i) C[i] = 0;
        *cdet = 0; *rcond = 0;
        return 0;
    }
    int noeig = 0;
    for (int i = 0; i < n; ++i) {
        if (ev[i] < fabs(plim * mx)) ev[i] = 0;
        else { ev[i] = 1.0 / ev[i]; noeig++; }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0;
            for (int k = 0; k < n; ++k) s += (v[i * 3 + k] * ev[k]) * v[j * 3 + k];
            C[(s0 + i) + 3 * (s0 + j)] = s;
        }
    return noeig;
}

// dpotrf('U') + dpotri on the sub-block s0..2; returns sing
__device__ bool inv_lu_block(double *C, int s0, double *cdet)
{
    int n = 3 - s0;
    double u[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = 0; j < n; ++j) {
        double s = C[(s0 + j) + 3 * (s0 + j)];
        for (int k = 0; k < j; ++k) s -= u[k + 3 * j] * u[k + 3 * j];
        if (!(s > 0)) return true;
        double ujj = sqrt(s);
        u[j + 3 * j] = ujj;
        for (int i = j + 1; i < n; ++i) {
            double t = C[(s0 + j) + 3 * (s0 + i)];
            for (int k = 0; k < j; ++k) t -= u[k + 3 * j] * u[k + 3 * i];
            u[j + 3 * i] = t / ujj;
        }
    }
    double det = 1;
    for (int i = 0; i < n; ++i) det *= u[i + 3 * i];
    *cdet = det;
    double ui[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = 0; j < n; ++j) {
        ui[j + 3 * j] = 1.0 / u[j + 3 * j];
        for (int i = j - 1; i >= 0; --i) {
            double s = 0;
            for (int k = i + 1; k <= j; ++k) s += u[i + 3 * k] * ui[k + 3 * j];
            ui[i + 3 * j] = -s / u[i + 3 * i];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0;
            for (int k = (i > j ? i : j); k < n; ++k) s += ui[i + 3 * k] * ui[j + 3 * k];
            C[(s0 + i) + 3 * (s0 + j)] = s;
        }
    return false;
}

// pixmode semantics of invert_pixelmatrix_cross (map == 0) and _map (map == 1)
__global__ void k_invert(int nmap, int64_t npix, const double *cc6, double *ccinv6, double *crit, int32_t *mask,
                         int pixmode, double pixlim, int incomplete, int allow_dec, int is_map,
                         unsigned long long *ngood)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < npix; q += (int64_t)gridDim.x * blockDim.x) {
        const double *u = cc6 + q * 6;
        double out[6] = {0, 0, 0, 0, 0, 0};
        int good = 0;
        double pc = 0;
        if (nmap == 1) {
            if (u[0] != 0) { out[0] = 1.0 / u[0]; good = 1; }
            pc = 0;
        } else {
            double C[9] = {u[0], 0, 0, u[1], u[3], 0, u[2], u[4], u[5]};   // C[a + 3b] = cc(a+1, b+1), upper
            bool allz = true;
            for (int a = 0; a < 6; ++a) allz = allz && (u[a] == 0.0);
            if (!allz) {
                int s0 = 0;
                if (allow_dec) {
                    // test_decoupling (src/matrix.f90:15-51)
                    int istart = 2;
                    for (int i = 0; i < 3; ++i)
                        if (C[i + 3 * i] > F32(1e-30))
                            for (int j = 1; j < 3; ++j)
                                if (fabs(C[3 * j]) / C[i + 3 * i] > F32(1e-10)) istart = 1;
                    bool any = false;
                    for (int j = 1; j < 3; ++j) any = any || fabs(C[j]) > F32(1e-10);
                    istart = any ? 1 : 2;
                    if (istart == 2 && C[0] > F32(1e-30)) {
                        C[0] = 1.0 / C[0];
                        for (int j = 1; j < 3; ++j) { C[j] = 0; C[3 * j] = 0; }
                    }
                    s0 = istart - 1;
                }
                double cdet = 0, rc = 0;
                bool sing = false;
                int mode = pixmode;
                if (mode == 0 || mode == 1) {
                    double tr = 0;
                    for (int a = s0; a < 3; ++a) tr += C[a + 3 * a];
                    sing = inv_lu_block(C, s0, &cdet);
                    pc = cdet;
                    if (mode == 1) {
                        double trn = fabs(tr) < F32(1e-30) ? 0 : (3.0 / tr) * (3.0 / tr) * (3.0 / tr);
                        pc = cdet * trn;
                    }
                    if (pc < pixlim || sing) { for (int a = 0; a < 9; ++a) C[a] = 0; good = 0; }
                    else good = 1;
                } else if (mode == 2) {
                    int g[3], nf = 0;
                    if (incomplete) {
                        for (int i = s0; i < 3; ++i) if (fabs(C[i + 3 * i]) > F32(1e-30)) g[nf++] = i;
                        if (nf == 0) { g[0] = s0; nf = 1; }
                    } else {
                        for (int i = s0; i < 3; ++i) g[nf++] = i;
                    }
                    sing = inv_eig_block(C, g, nf, &cdet, &rc);
                    pc = rc;
                    if (rc < pixlim || sing) { for (int a = 0; a < 9; ++a) C[a] = 0; good = 0; }
                    else good = 1;
                } else if (mode == 3) {
                    if (C[0] > F32(1e-30)) {
                        if (s0 == 0) C[0] = 1.0 / C[0];
                        for (int a = 1; a < 9; ++a) C[a] = 0;
                        good = 1;
                    } else { for (int a = 0; a < 9; ++a) C[a] = 0; good = 0; }
                } else {
                    int noeig = inv_eig_s_block(C, s0, pixlim, &cdet, &rc);
                    good = noeig > 0;
                }
                // store the lower triangle, which is what ccmultiply reads
                out[0] = C[0]; out[1] = C[1]; out[2] = C[2]; out[3] = C[4]; out[4] = C[5]; out[5] = C[8];
            }
        }
        for (int a = 0; a < 6; ++a) ccinv6[q * 6 + a] = out[a];
        if (is_map) {
            if (mask) mask[q] = good;
            if (crit) crit[q] = (double)(float)pc;
        }
        if (good && ngood) atomicAdd(ngood, 1ull);
    }
}

void launch_invert(hipStream_t s, int nmap, int64_t npix, const double *cc6, double *ccinv6, double *crit,
                   int32_t *mask, int pixmode, double pixlim, int incomplete, int allow_dec,
                   unsigned long long *ngood)
{
    k_invert<<<grid_for(npix, TPB), TPB, 0, s>>>(nmap, npix, cc6, ccinv6, crit, mask, pixmode, pixlim, incomplete,
                                                allow_dec, mask != nullptr, ngood);
}

__device__ inline void ccmul(int nmap, const double *c, const double *m, double *o)
{
    if (c[0] == 0.0) { o[0] = 0; o[1] = 0; o[2] = 0; return; }
    if (nmap == 1) { o[0] = c[0] * m[0]; o[1] = 0; o[2] = 0; return; }
    o[0] = c[0] * m[0] + c[1] * m[1] + c[2] * m[2];
    o[1] = c[1] * m[0] + c[3] * m[1] + c[4] * m[2];
    o[2] = c[2] * m[0] + c[4] * m[1] + c[5] * m[2];
}

__global__ void k_ccmultiply(int nmap, int64_t npix, const double *ccinv6, const double *in4, double *out4)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < npix; q += (int64_t)gridDim.x * blockDim.x) {
        double o[3];
        ccmul(nmap, ccinv6 + q * 6, in4 + q * 4, o);
        double4 r;
        r.x = o[0]; r.y = o[1]; r.z = o[2]; r.w = 0.0;
        *(double4 *)(out4 + q * 4) = r;
    }
}

void launch_ccmultiply(hipStream_t s, int nmap, int64_t npix, const double *ccinv6, const double *rhs4, double *out4)
{
    k_ccmultiply<<<grid_for(npix, TPB), TPB, 0, s>>>(nmap, npix, ccinv6, rhs4, out4);
}

// ---------------------------------------------------------------- time-ordered per-baseline reductions

// initialize_a: yba, nna for order-0 baselines.  G lanes per baseline.
template <int G, typename TT>
__global__ void k_init_a(int nmap, int ndet, int64_t nsamp, int64_t nbase, const int64_t *bstart, const PixW *pixw,
                         const TT *tod, const double *m4, int64_t ndeg, const double *dwb, double *yba,
                         double *nna)
{
    int lane = threadIdx.x & (G - 1);
    int gib = threadIdx.x / G, ngib = blockDim.x / G;
    int64_t nbt = (int64_t)ndet * nbase;
    int64_t lb = xcd_block(blockIdx.x, gridDim.x);
    int64_t b1 = (lb + 1) * ITEMS_PER_BLOCK < nbt ? (lb + 1) * ITEMS_PER_BLOCK : nbt;
    for (int64_t b = lb * ITEMS_PER_BLOCK + gib; b < b1; b += ngib) {
        int det = (int)(b / nbase);
        int64_t k = b - (int64_t)det * nbase;
        int64_t i0 = bstart[k], i1 = bstart[k + 1];
        const PixW *P = pixw + (int64_t)det * nsamp;
        const TT *T = tod + (int64_t)det * nsamp;
        double y = 0, n = 0;
        for (int64_t i = i0 + lane; i < i1; i += G) {
            PixW r = P[i];
            if (r.pix < 0) continue;
            const double *L = m4 + (r.pix / ndeg) * 4;
            double4 m = *(const double4 *)L;
            if (m.x == 0.0 && m.y == 0.0 && m.z == 0.0) continue;
            double dot = nmap == 1 ? m.x : (m.x * (double)r.w0 + m.y * (double)r.w1 + m.z * (double)r.w2);
            y += (double)T[i] - dot;
            n += 1.0;
        }
        y = group_sum<G>(y);
        n = group_sum<G>(n);
        if (lane == 0) {
            double dw = dwb[b];
            double yy = y * dw, nn = n * dw;
            if (fabs(nn) < F32(1e-30)) { yy = 0; nn = 0; }
            yba[b] = yy;
            nna[b] = nn;
        }
    }
}

#define DISPATCH_G(group, KERNEL, ...)                                              \
    switch (group) {                                                                \
    case 4: KERNEL<4><<<grid, TPB, 0, s>>>(__VA_ARGS__); break;                     \
    case 8: KERNEL<8><<<grid, TPB, 0, s>>>(__VA_ARGS__); break;                     \
    case 16: KERNEL<16><<<grid, TPB, 0, s>>>(__VA_ARGS__); break;                   \
    case 32: KERNEL<32><<<grid, TPB, 0, s>>>(__VA_ARGS__); break;                   \
    default: KERNEL<64><<<grid, TPB, 0, s>>>(__VA_ARGS__); break;                   \
    }

template <typename TT>
void launch_init_a(hipStream_t s, int nmap, int ndet, int64_t nsamp, int64_t nbase, const int64_t *bstart,
                   const PixW *pixw, const TT *tod, const double *m4, int64_t ndeg, const double *dwb,
                   double *yba, double *nna, int group)
{
    int grid = (int)(((int64_t)ndet * nbase + ITEMS_PER_BLOCK - 1) / ITEMS_PER_BLOCK);
    switch (group) {
    case 4: k_init_a<4, TT><<<grid, TPB, 0, s>>>(nmap, ndet, nsamp, nbase, bstart, pixw, tod, m4, ndeg, dwb, yba, nna); break;
    case 8: k_init_a<8, TT><<<grid, TPB, 0, s>>>(nmap, ndet, nsamp, nbase, bstart, pixw, tod, m4, ndeg, dwb, yba, nna); break;
    case 16: k_init_a<16, TT><<<grid, TPB, 0, s>>>(nmap, ndet, nsamp, nbase, bstart, pixw, tod, m4, ndeg, dwb, yba, nna); break;
    case 32: k_init_a<32, TT><<<grid, TPB, 0, s>>>(nmap, ndet, nsamp, nbase, bstart, pixw, tod, m4, ndeg, dwb, yba, nna); break;
    default: k_init_a<64, TT><<<grid, TPB, 0, s>>>(nmap, ndet, nsamp, nbase, bstart, pixw, tod, m4, ndeg, dwb, yba, nna); break;
    }
}
template void launch_init_a<double>(hipStream_t, int, int, int64_t, int64_t, const int64_t *, const PixW *,
                                    const double *, const double *, int64_t, const double *, double *, double *, int);
template void launch_init_a<float>(hipStream_t, int, int, int64_t, int64_t, const int64_t *, const PixW *,
                                   const float *, const double *, int64_t, const double *, double *, double *, int);

// map_to_baseline_order0: ap[b] += nna[b] p[b] - dw[b] * sum_i w_i . m[cross(pix_i)]
template <int G>
__global__ void k_gather_baselines(int nmap, int ndet, int64_t nsamp, int64_t nbase, const int64_t *bstart,
                                   const PixW *pixw, const double *m4, int64_t ndeg, const double *dwb,
                                   const double *nna, const double *p, double *ap)
{
    int lane = threadIdx.x & (G - 1);
    int gib = threadIdx.x / G, ngib = blockDim.x / G;
    int64_t nbt = (int64_t)ndet * nbase;
    int64_t lb = xcd_block(blockIdx.x, gridDim.x);
    int64_t b1 = (lb + 1) * ITEMS_PER_BLOCK < nbt ? (lb + 1) * ITEMS_PER_BLOCK : nbt;
    for (int64_t b = lb * ITEMS_PER_BLOCK + gib; b < b1; b += ngib) {
        int det = (int)(b / nbase);
        int64_t k = b - (int64_t)det * nbase;
        int64_t i0 = bstart[k], i1 = bstart[k + 1];
        const PixW *P = pixw + (int64_t)det * nsamp;
        double apn = 0;
        for (int64_t i = i0 + lane; i < i1; i += G) {
            PixW r = P[i];
            if (r.pix < 0) continue;
            double4 m = *(const double4 *)(m4 + (r.pix / ndeg) * 4);
            if (nmap == 1) apn += m.x;
            else apn += (double)r.w0 * m.x + (double)r.w1 * m.y + (double)r.w2 * m.z;
        }
        apn = group_sum<G>(apn);
        if (lane == 0) {
            double v = nna[b] * p[b] - apn * dwb[b];
            ap[b] = ap[b] + v;
        }
    }
}

void launch_gather_baselines(hipStream_t s, int nmap, int ndet, int64_t nsamp, int64_t nbase, const int64_t *bstart,
                             const PixW *pixw, const double *m4, int64_t ndeg, const double *dwb, const double *nna,
                             const double *p, double *ap, int group)
{
    int grid = (int)(((int64_t)ndet * nbase + ITEMS_PER_BLOCK - 1) / ITEMS_PER_BLOCK);
    DISPATCH_G(group, k_gather_baselines, nmap, ndet, nsamp, nbase, bstart, pixw, m4, ndeg, dwb, nna, p, ap);
}

// baseline_to_map_order0 as a per-cross-pixel gather over the CSR, fused with ccmultiply
__global__ void k_scatter_pixels(int nmap, int64_t npc, int64_t ndeg, const int64_t *off, const CsrEnt *ent,
                                 const double *pw, const double *ccinv6, double *m4)
{
    int lane = threadIdx.x & 63;
    int wib = threadIdx.x >> 6, nwib = blockDim.x >> 6;
    int64_t lb = xcd_block(blockIdx.x, gridDim.x);
    int64_t c1 = (lb + 1) * ITEMS_PER_BLOCK < npc ? (lb + 1) * ITEMS_PER_BLOCK : npc;
    for (int64_t c = lb * ITEMS_PER_BLOCK + wib; c < c1; c += nwib) {
        int64_t j0 = off[c * ndeg], j1 = off[(c + 1) * ndeg];
        double a0 = 0, a1 = 0, a2 = 0;
        for (int64_t j = j0 + lane; j < j1; j += 64) {
            CsrEnt e = ent[j];
            double v = pw[e.b];
            if (nmap == 1) a0 += v;
            else {
                a0 += (double)e.w0 * v;
                a1 += (double)e.w1 * v;
                a2 += (double)e.w2 * v;
            }
        }
        a0 = wave_sum(a0);
        if (nmap == 3) { a1 = wave_sum(a1); a2 = wave_sum(a2); }
        if (lane == 0) {
            double m[3] = {a0, a1, a2}, o[3];
            ccmul(nmap, ccinv6 + c * 6, m, o);
            double4 r;
            r.x = o[0]; r.y = o[1]; r.z = o[2]; r.w = 0.0;
            *(double4 *)(m4 + c * 4) = r;
        }
    }
}

void launch_scatter_pixels(hipStream_t s, int nmap, int64_t npix_cross, int64_t ndeg, const int64_t *off,
                           const CsrEnt *ent, const double *pw, const double *ccinv6, double *m4)
{
    int grid = (int)((npix_cross + ITEMS_PER_BLOCK - 1) / ITEMS_PER_BLOCK);
    k_scatter_pixels<<<grid, TPB, 0, s>>>(nmap, npix_cross, ndeg, off, ent, pw, ccinv6, m4);
}

// ---------------------------------------------------------------- flat tiles
// Flat variants of the two per-iteration projections: every thread of a block
// issues coalesced loads for a tile of the stream (maximum memory-level
// parallelism), the per-sample products go to LDS, and a segmented reduction
// from LDS produces one value per baseline (or pixel).  Deterministic for a
// fixed CSR.  A block owns whole baselines (pixels), so no atomics are needed.

static constexpr int FLAT_TILE = 2048;
static constexpr int FLAT_ITEMS = 64;    // baselines / pixels per block
static constexpr int FLAT_LANES = 4;     // lanes per item in the reduction phase

// Baseline side: a block owns FG_K consecutive baselines of FG_D consecutive
// detectors, i.e. one time window of every one of those detectors (the
// baseline boundaries are common to all detectors), and the samples are
// loaded time-major across the detectors: a wave's lanes hold 64 / FG_D
// consecutive samples of each.  Co-pointed detectors (the polarisation pair
// of one horn, detectors that share a scan path) then read each map pixel in
// the same load instruction, one cache-line access for all of them; the loads
// of the pointing stream stay in whole 256-byte runs.
static constexpr int FG_D = 4;                    // detectors per block
static constexpr int FG_ITEMS = 256;              // (detector, baseline) items per block
static constexpr int FG_LANES = FG_ITEMS >= 256 ? 1 : 256 / FG_ITEMS;   // lanes per item (reduction)
static constexpr int FG_IPT = FG_ITEMS >= 256 ? FG_ITEMS / 256 : 1;     // items per thread
static constexpr int FG_K = FG_ITEMS / FG_D;      // baselines per detector per block
static constexpr int FG_T = FLAT_TILE / FG_D;     // samples per detector per LDS tile

template <typename T, int NMAP, int RL>
__global__ void __launch_bounds__(256) k_gather_flat(int ndet, int64_t nsamp, int64_t nbase,
                                                     const int64_t *bstart, const void *pixw, const T *m4,
                                                     int64_t ndeg, const T *dwb, const T *nna,
                                                     const T *p, T *ap)
{
    using V2 = typename VecT<T>::v2;
    __shared__ T vals[FLAT_TILE + 1];    // vals[d * FG_T + t]
    __shared__ int64_t bs[FG_K + 1];      // sample starts of the block's baselines, bs[nk] = end
    int64_t ngrp = (ndet + FG_D - 1) / FG_D;
    int64_t lb = xcd_block(blockIdx.x, gridDim.x);
    int64_t kb = lb / ngrp, g = lb - kb * ngrp;   // time-major: neighbouring blocks share the sky
    int64_t k0 = kb * FG_K;
    if (k0 >= nbase) return;
    int nk = (int)(nbase - k0 < FG_K ? nbase - k0 : FG_K);
    int d0 = (int)g * FG_D, nd = ndet - d0 < FG_D ? ndet - d0 : FG_D;
    for (int t = threadIdx.x; t <= nk; t += blockDim.x) bs[t] = bstart[k0 + t];
    int sub = threadIdx.x % FG_LANES;
    int id[FG_IPT], ik[FG_IPT];   // items (detector, baseline)
    bool own[FG_IPT];
    T acc[FG_IPT];
#pragma unroll
    for (int i = 0; i < FG_IPT; ++i) {
        int item = threadIdx.x / FG_LANES + i * (256 / FG_LANES);
        id[i] = item / FG_K;
        ik[i] = item % FG_K;
        own[i] = id[i] < nd && ik[i] < nk;
        acc[i] = (T)0;
    }
    __syncthreads();
    const int dsh = __ffsll((unsigned long long)ndeg) - 1;   // ndeg is a power of 4
    int64_t S0 = bs[0], S1 = bs[nk];
    for (int64_t t0 = S0; t0 < S1; t0 += FG_T) {
        int len = (int)(S1 - t0 < FG_T ? S1 - t0 : FG_T);
        // one time sample per lane, all FG_D detectors of the block: the
        // pointing records (one contiguous row per detector), then the map
        // entry of each detector, loaded only when its pixel differs from the
        // earlier detectors' (co-pointed detectors share one load).  A wave
        // whose samples all lie past the end issues no loads (the address
        // path costs per instruction, whatever the exec mask).
        for (int tt0 = threadIdx.x; tt0 - (int)(threadIdx.x & 63) < len; tt0 += blockDim.x) {
            const int tb = __builtin_amdgcn_readfirstlane(tt0 - (int)(threadIdx.x & 63));
            if (tb >= len) break;   // uniform over the wave
            int tl = tt0 < len ? tt0 : len - 1;
            PixW r[FG_D];
#pragma unroll
            for (int d = 0; d < FG_D; ++d)
                r[d] = load_rec_nt<RL>(pixw, (int64_t)(d0 + (d < nd ? d : 0)) * nsamp + t0 + tl);
            int q[FG_D];
            bool ld[FG_D];
#pragma unroll
            for (int d = 0; d < FG_D; ++d) {
                q[d] = r[d].pix >= 0 ? (r[d].pix >> dsh) : 0;
                bool seen = false;
#pragma unroll
                for (int e = 0; e < d; ++e) seen = seen || q[e] == q[d];
                ld[d] = !seen;
            }
            V2 mxy[FG_D];   // (m.x, m.y); m.y unused for NMAP = 1
            T mz[FG_D];
#pragma unroll
            for (int d = 0; d < FG_D; ++d) {
                if (ld[d]) {
                    const T *mp = m4 + (int64_t)q[d] * 4;
                    if (NMAP == 1) {
                        mxy[d].x = mp[0];
                    } else {
                        mxy[d] = *(const V2 *)mp;
                        mz[d] = mp[2];
                    }
                }
            }
            // keeps the loads above from being sunk into the per-sample code below
            asm volatile("" ::: "memory");
#pragma unroll
            for (int d = 1; d < FG_D; ++d) {
                if (!ld[d]) {
                    int src = 0;
#pragma unroll
                    for (int e = 1; e < d; ++e)
                        if (q[e] == q[d]) src = e;
                    V2 cxy = mxy[0];
                    T cz = mz[0];
#pragma unroll
                    for (int e = 1; e < d; ++e)
                        if (src == e) {
                            cxy = mxy[e];
                            cz = mz[e];
                        }
                    mxy[d] = cxy;
                    mz[d] = cz;
                }
            }
#pragma unroll
            for (int d = 0; d < FG_D; ++d) {
                T v = NMAP == 1 ? mxy[d].x
                                : (T)r[d].w0 * mxy[d].x + (T)r[d].w1 * mxy[d].y + (T)r[d].w2 * mz[d];
                v = (d < nd && r[d].pix >= 0) ? v : (T)0;
                vals[tt0 < len ? d * FG_T + tt0 : FLAT_TILE] = v;   // past the end: the spare slot
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < FG_IPT; ++i)
            if (own[i]) {
                int64_t t1 = t0 + len;
                int64_t a = bs[ik[i]] > t0 ? bs[ik[i]] : t0, e = bs[ik[i] + 1] < t1 ? bs[ik[i] + 1] : t1;
                for (int64_t t = a + sub; t < e; t += FG_LANES) acc[i] += vals[id[i] * FG_T + (t - t0)];
            }
        __syncthreads();
    }
    // the per-baseline terms, loaded after the sample stream (holding them
    // across it would cost the registers of two waves per SIMD)
#pragma unroll
    for (int i = 0; i < FG_IPT; ++i) {
        T a = group_sum<FG_LANES>(acc[i]);
        if (own[i] && sub == 0) {
            int64_t b = (int64_t)(d0 + id[i]) * nbase + k0 + ik[i];
            ap[b] = ap[b] + (nna[b] * p[b] - a * dwb[b]);
        }
    }
}

__global__ void __launch_bounds__(256) k_scatter_flat(int nmap, int64_t npc, int64_t ndeg, const int64_t *off,
                                                      const CsrEnt *ent, const double *pw, const double *ccinv6,
                                                      double *m4)
{
    __shared__ double vals[3][FLAT_TILE];
    __shared__ int64_t qs[FLAT_ITEMS + 1];
    int64_t lb = xcd_block(blockIdx.x, gridDim.x);
    int64_t c0 = lb * FLAT_ITEMS, c1 = c0 + FLAT_ITEMS < npc ? c0 + FLAT_ITEMS : npc;
    int nc = (int)(c1 - c0);
    if (nc <= 0) return;
    for (int t = threadIdx.x; t <= nc; t += blockDim.x) qs[t] = off[(c0 + t) * ndeg];
    __syncthreads();
    int item = threadIdx.x / FLAT_LANES, sub = threadIdx.x % FLAT_LANES;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    int64_t S0 = qs[0], S1 = qs[nc];
    for (int64_t t0 = S0; t0 < S1; t0 += FLAT_TILE) {
        int64_t t1 = t0 + FLAT_TILE < S1 ? t0 + FLAT_TILE : S1;
        for (int64_t j = t0 + threadIdx.x; j < t1; j += blockDim.x) {
            CsrEnt e = ent[j];
            double v = pw[e.b];
            int64_t o = j - t0;
            if (nmap == 1) vals[0][o] = v;
            else {
                vals[0][o] = (double)e.w0 * v;
                vals[1][o] = (double)e.w1 * v;
                vals[2][o] = (double)e.w2 * v;
            }
        }
        __syncthreads();
        if (item < nc) {
            int64_t a = qs[item] > t0 ? qs[item] : t0, e = qs[item + 1] < t1 ? qs[item + 1] : t1;
            for (int64_t j = a + sub; j < e; j += FLAT_LANES) {
                a0 += vals[0][j - t0];
                if (nmap == 3) { a1 += vals[1][j - t0]; a2 += vals[2][j - t0]; }
            }
        }
        __syncthreads();
    }
    a0 = group_sum<FLAT_LANES>(a0);
    a1 = group_sum<FLAT_LANES>(a1);
    a2 = group_sum<FLAT_LANES>(a2);
    if (item < nc && sub == 0) {
        int64_t c = c0 + item;
        double m[3] = {a0, a1, a2}, o[3];
        ccmul(nmap, ccinv6 + c * 6, m, o);
        double4 r;
        r.x = o[0]; r.y = o[1]; r.z = o[2]; r.w = 0.0;
        *(double4 *)(m4 + c * 4) = r;
    }
}

// Entry-balanced scatter: block j reduces CSR entries [j T, (j+1) T) (T =
// FLAT_TILE).  Pixels wholly inside the tile are finished in place; the
// partial sums of the first / last pixel of the tile, when they continue into a
// neighbouring tile, go to part[j][0] / part[j][1] and are combined in tile
// order by k_scatter_fixup.  Every block does the same amount of work however
// uneven the hit counts are (the scan's pole pixels hold tens of thousands of
// samples).
struct TileInfo {
    int64_t qa, qb;   // first and last cross pixel touching the tile
};

__device__ inline int64_t cross_of_entry(const int64_t *off, int64_t ndeg, int64_t npc, int64_t j)
{
    // largest c with off[c * ndeg] <= j
    int64_t lo = 0, hi = npc - 1;
    while (lo < hi) {
        int64_t mid = (lo + hi + 1) >> 1;
        if (off[mid * ndeg] <= j) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(256) k_scatter_tiles(int nmap, int64_t ndeg, const int64_t *off, int64_t nent,
                                                       const CsrEnt *ent, const double *pw, const double *ccinv6,
                                                       double *m4, double4 *part, const TileInfo *tinfo)
{
    __shared__ double vals[3][FLAT_TILE];
    int64_t j = xcd_block(blockIdx.x, gridDim.x);
    int64_t t0 = j * FLAT_TILE, t1 = t0 + FLAT_TILE < nent ? t0 + FLAT_TILE : nent;
    if (t0 >= nent) return;
    TileInfo ti = tinfo[j];
    for (int64_t jj = t0 + threadIdx.x; jj < t1; jj += blockDim.x) {
        CsrEnt e = ent[jj];
        double v = pw[e.b];
        int64_t o = jj - t0;
        if (nmap == 1) vals[0][o] = v;
        else {
            vals[0][o] = (double)e.w0 * v;
            vals[1][o] = (double)e.w1 * v;
            vals[2][o] = (double)e.w2 * v;
        }
    }
    __syncthreads();
    int64_t qa = ti.qa, qb = ti.qb;
    int item = threadIdx.x / FLAT_LANES, sub = threadIdx.x % FLAT_LANES;
    for (int64_t q = qa + item; q <= qb; q += blockDim.x / FLAT_LANES) {
        int64_t a = off[q * ndeg], e = off[(q + 1) * ndeg];
        int64_t lo = a > t0 ? a : t0, hi = e < t1 ? e : t1;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int64_t jj = lo + sub; jj < hi; jj += FLAT_LANES) {
            a0 += vals[0][jj - t0];
            if (nmap == 3) { a1 += vals[1][jj - t0]; a2 += vals[2][jj - t0]; }
        }
        a0 = group_sum<FLAT_LANES>(a0);
        a1 = group_sum<FLAT_LANES>(a1);
        a2 = group_sum<FLAT_LANES>(a2);
        if (sub == 0) {
            if (a >= t0 && e <= t1) {
                double m[3] = {a0, a1, a2}, o[3];
                ccmul(nmap, ccinv6 + q * 6, m, o);
                double4 r;
                r.x = o[0]; r.y = o[1]; r.z = o[2]; r.w = 0.0;
                *(double4 *)(m4 + q * 4) = r;
            } else {
                double4 r;
                r.x = a0; r.y = a1; r.z = a2; r.w = 0.0;
                part[j * 2 + (q == qa ? 0 : 1)] = r;
            }
        }
    }
}

// one thread per tile: finish the pixel that starts in tile j and continues past it
__global__ void k_scatter_fixup(int nmap, int64_t ntiles, int64_t ndeg, const int64_t *off, const double4 *part,
                                const TileInfo *tinfo, const double *ccinv6, double *m4)
{
    int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= ntiles) return;
    int64_t t0 = j * FLAT_TILE, t1 = t0 + FLAT_TILE;
    TileInfo ti = tinfo[j];
    int64_t q = ti.qb;
    int64_t a = off[q * ndeg], e = off[(q + 1) * ndeg];
    if (!(a >= t0 && e > t1)) return;   // q must start in this tile and continue
    double4 acc = part[j * 2 + (q == ti.qa ? 0 : 1)];
    for (int64_t jj = j + 1; jj < ntiles; ++jj) {
        double4 v = part[jj * 2 + 0];   // q is the first pixel of every later tile it reaches
        acc.x += v.x; acc.y += v.y; acc.z += v.z;
        if (e <= (jj + 1) * FLAT_TILE) break;
    }
    double m[3] = {acc.x, acc.y, acc.z}, o[3];
    ccmul(nmap, ccinv6 + q * 6, m, o);
    double4 r;
    r.x = o[0]; r.y = o[1]; r.z = o[2]; r.w = 0.0;
    *(double4 *)(m4 + q * 4) = r;
}

// tile -> first / last cross pixel; depends on the CSR only, computed once per pointing
__global__ void k_tile_info(int64_t npc, int64_t ndeg, const int64_t *off, int64_t nent, TileInfo *tinfo)
{
    int64_t ntiles = (nent + FLAT_TILE - 1) / FLAT_TILE;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < ntiles; j += (int64_t)gridDim.x * blockDim.x) {
        int64_t t0 = j * FLAT_TILE, t1 = t0 + FLAT_TILE < nent ? t0 + FLAT_TILE : nent;
        tinfo[j].qa = cross_of_entry(off, ndeg, npc, t0);
        tinfo[j].qb = cross_of_entry(off, ndeg, npc, t1 - 1);
    }
}

size_t scatter_tiles_info_bytes(int64_t nent)
{
    int64_t ntiles = (nent + FLAT_TILE - 1) / FLAT_TILE;
    return (size_t)(ntiles > 0 ? ntiles : 1) * sizeof(TileInfo);
}

void launch_tile_info(hipStream_t s, int64_t npix_cross, int64_t ndeg, const int64_t *off, int64_t nent, void *tinfo)
{
    int64_t ntiles = (nent + FLAT_TILE - 1) / FLAT_TILE;
    if (ntiles == 0) return;
    k_tile_info<<<grid_for(ntiles, TPB), TPB, 0, s>>>(npix_cross, ndeg, off, nent, (TileInfo *)tinfo);
}

void launch_scatter_tiles(hipStream_t s, int nmap, int64_t npix_cross, int64_t ndeg, const int64_t *off, int64_t nent,
                          const CsrEnt *ent, const double *pw, const double *ccinv6, double *m4, const void *tinfo_v,
                          void *scratch)
{
    (void)npix_cross;
    int64_t ntiles = (nent + FLAT_TILE - 1) / FLAT_TILE;
    if (ntiles == 0) return;
    double4 *part = (double4 *)scratch;
    const TileInfo *tinfo = (const TileInfo *)tinfo_v;
    k_scatter_tiles<<<(unsigned)ntiles, 256, 0, s>>>(nmap, ndeg, off, nent, ent, pw, ccinv6, m4, part, tinfo);
    k_scatter_fixup<<<(unsigned)((ntiles + 255) / 256), 256, 0, s>>>(nmap, ntiles, ndeg, off, part, tinfo, ccinv6,
                                                                      m4);
}

size_t scatter_tiles_scratch_bytes(int64_t nent)
{
    int64_t ntiles = (nent + FLAT_TILE - 1) / FLAT_TILE;
    return (size_t)ntiles * 2 * sizeof(double4) + 256;
}

template <typename T>
void launch_gather_flat(hipStream_t s, int nmap, int ndet, int64_t nsamp, int64_t nbase, const int64_t *bstart,
                        const void *pixw, int rl, const T *m4, int64_t ndeg, const T *dwb, const T *nna, const T *p,
                        T *ap)
{
    int64_t grid = ((ndet + FG_D - 1) / FG_D) * ((nbase + FG_K - 1) / FG_K);
    if (grid > 0)
    {
#define MADAM_GF(NM, RLV)                                                                                        \
    k_gather_flat<T, NM, RLV><<<(unsigned)grid, 256, 0, s>>>(ndet, nsamp, nbase, bstart, pixw, m4, ndeg, dwb, nna, p, \
                                                             ap)
        if (nmap == 1) {
            if (rl == REC4) MADAM_GF(1, REC4);
            else MADAM_GF(1, REC16);
        } else {
            if (rl == REC12) MADAM_GF(3, REC12);
            else MADAM_GF(3, REC16);
        }
#undef MADAM_GF
    }
}
template void launch_gather_flat<double>(hipStream_t, int, int, int64_t, int64_t, const int64_t *, const void *, int,
                                         const double *, int64_t, const double *, const double *, const double *,
                                         double *);
template void launch_gather_flat<float>(hipStream_t, int, int, int64_t, int64_t, const int64_t *, const void *, int,
                                        const float *, int64_t, const float *, const float *, const float *, float *);

void launch_scatter_flat(hipStream_t s, int nmap, int64_t npix_cross, int64_t ndeg, const int64_t *off,
                         const CsrEnt *ent, const double *pw, const double *ccinv6, double *m4)
{
    int64_t grid = (npix_cross + FLAT_ITEMS - 1) / FLAT_ITEMS;
    k_scatter_flat<<<(unsigned)grid, 256, 0, s>>>(nmap, npix_cross, ndeg, off, ent, pw, ccinv6, m4);
}

// subtract_baselines_a: rhs4 (binned P^T W d at map resolution) += -P^T W F a
__global__ void k_subtract(int nmap, int64_t npix, const int64_t *off, const CsrEnt *ent, const double *aa,
                           const double *dwb, double *rhs4)
{
    int lane = threadIdx.x & 63;
    int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t q = wave; q < npix; q += nwaves) {
        int64_t j0 = off[q], j1 = off[q + 1];
        double a0 = 0, a1 = 0, a2 = 0;
        for (int64_t j = j0 + lane; j < j1; j += 64) {
            CsrEnt e = ent[j];
            double aw = aa[e.b] * dwb[e.b];
            if (nmap == 1) a0 -= aw;
            else {
                a0 -= aw * (double)e.w0;
                a1 -= aw * (double)e.w1;
                a2 -= aw * (double)e.w2;
            }
        }
        a0 = wave_sum(a0);
        if (nmap == 3) { a1 = wave_sum(a1); a2 = wave_sum(a2); }
        if (lane == 0) {
            rhs4[q * 4] += a0;
            rhs4[q * 4 + 1] += a1;
            rhs4[q * 4 + 2] += a2;
        }
    }
}

void launch_subtract_baselines(hipStream_t s, int nmap, int64_t npix_map, const int64_t *off, const CsrEnt *ent,
                               const double *aa, const double *dwb, double *rhs4)
{
    k_subtract<<<grid_for(npix_map, TPB / 64, 1 << 20), TPB, 0, s>>>(nmap, npix_map, off, ent, aa, dwb, rhs4);
}

__global__ void k_clean_tod(int ndet, int64_t nsamp, int64_t nbase, const int64_t *bstart, const double *aa,
                            double *tod)
{
    int64_t n = (int64_t)ndet * nsamp;
    int64_t end = bstart[nbase];
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
        int det = (int)(g / nsamp);
        int64_t i = g % nsamp;
        if (i >= end) continue;
        int64_t k = find_baseline(bstart, nbase, i);
        tod[g] = tod[g] - 1.0 * aa[(int64_t)det * nbase + k];
    }
}

void launch_clean_tod(hipStream_t s, int ndet, int64_t nsamp, int64_t nbase, const int64_t *bstart, const double *aa,
                      double *tod)
{
    k_clean_tod<<<grid_for((int64_t)ndet * nsamp, TPB), TPB, 0, s>>>(ndet, nsamp, nbase, bstart, aa, tod);
}

// ---------------------------------------------------------------- segments (trim_interval)

// One wave per (det, interval).  A segment is a maximal run of baselines from a
// nonzero-nna baseline to the last nonzero-nna baseline before a run of at
// least gapmin zero baselines (or the end of the interval).
__global__ void k_find_segments(const double *nna, int ndet, int64_t nbase, const int64_t *kstart_pp,
                                const int64_t *noba_pp, int nival, int64_t gapmin, Segment *out,
                                unsigned long long *nseg, int64_t cap)
{
    int lane = threadIdx.x & 63;
    int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
    int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t t = wave; t < (int64_t)ndet * nival; t += nwaves) {
        int det = (int)(t / nival), iv = (int)(t % nival);
        int64_t k0 = kstart_pp[iv], k1 = k0 + noba_pp[iv];
        const double *nn = nna + (int64_t)det * nbase;
        int64_t start = -1, last = -1;
        for (int64_t base = k0; base < k1; base += 64) {
            int64_t k = base + lane;
            bool nz = k < k1 && nn[k] != 0.0;
            unsigned long long m = __ballot(nz);
            if (m == 0) continue;
            if (gapmin > 63 && start >= 0) {
                // no gap of gapmin zeros fits inside one 64-wide word: only the
                // run from the previous nonzero to the first set bit can split
                int first = __ffsll((long long)m) - 1;
                int64_t kk = base + first;
                if (kk - last - 1 >= gapmin) {
                    if (lane == 0) {
                        unsigned long long pos = atomicAdd(nseg, 1ull);
                        if ((int64_t)pos < cap) {
                            Segment sg;
                            sg.b0 = (int64_t)det * nbase + start; sg.noba = (int32_t)(last - start + 1);
                            sg.det = det; sg.ival = iv; sg.ipsd = -1; sg.skip = 0; sg.pad = 0;
                            out[pos] = sg;
                        }
                    }
                    start = kk;
                }
                last = base + 63 - __clzll((long long)m);
                continue;
            }
            while (m) {
                int bit = __ffsll((long long)m) - 1;
                m &= m - 1;
                int64_t kk = base + bit;
                if (start < 0) start = kk;
                else if (kk - last - 1 >= gapmin) {
                    if (lane == 0) {
                        unsigned long long pos = atomicAdd(nseg, 1ull);
                        if ((int64_t)pos < cap) {
                            Segment sg;
                            sg.b0 = (int64_t)det * nbase + start; sg.noba = (int32_t)(last - start + 1);
                            sg.det = det; sg.ival = iv; sg.ipsd = -1; sg.skip = 0; sg.pad = 0;
                            out[pos] = sg;
                        }
                    }
                    start = kk;
                }
                last = kk;
            }
        }
        if (start >= 0 && lane == 0) {
            unsigned long long pos = atomicAdd(nseg, 1ull);
            if ((int64_t)pos < cap) {
                Segment sg;
                sg.b0 = (int64_t)det * nbase + start; sg.noba = (int32_t)(last - start + 1);
                sg.det = det; sg.ival = iv; sg.ipsd = -1; sg.skip = 0; sg.pad = 0;
                out[pos] = sg;
            }
        }
    }
}

void launch_find_segments(hipStream_t s, const double *nna, int ndet, int64_t nbase, const int64_t *kstart_pp,
                          const int64_t *noba_pp, int ninterval, int64_t gapmin, Segment *out,
                          unsigned long long *nseg, int64_t cap)
{
    k_find_segments<<<grid_for((int64_t)ndet * ninterval, TPB / 64), TPB, 0, s>>>(nna, ndet, nbase, kstart_pp, noba_pp,
                                                                                 ninterval, gapmin, out, nseg, cap);
}

__global__ void k_set_mask(const Segment *seg, int64_t nseg, int8_t *rmask)
{
    for (int64_t s = blockIdx.x; s < nseg; s += gridDim.x) {
        Segment sg = seg[s];
        for (int64_t k = threadIdx.x; k < sg.noba; k += blockDim.x) rmask[sg.b0 + k] = 1;
    }
}

void launch_set_segments_mask(hipStream_t s, const Segment *seg, int64_t nseg, int8_t *rmask)
{
    if (nseg > 0) k_set_mask<<<grid_for(nseg, 1), TPB, 0, s>>>(seg, nseg, rmask);
}

// ---------------------------------------------------------------- vector helpers

__global__ void k_fill(double *p, double v, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
void launch_fill(hipStream_t s, double *p, double v, int64_t n)
{
    k_fill<<<grid_for(n, TPB), TPB, 0, s>>>(p, v, n);
}

// flag[0] = 1 if a and b differ bitwise anywhere (flag cleared by the caller)
__global__ void k_bits_differ(const double *a, const double *b, int64_t n, int *flag)
{
    bool d = false;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d = d || __double_as_longlong(a[i]) != __double_as_longlong(b[i]);
    if (d) flag[0] = 1;
}
void launch_bits_differ(hipStream_t s, const double *a, const double *b, int64_t n, int *flag)
{
    if (n > 0) k_bits_differ<<<grid_for(n, TPB), TPB, 0, s>>>(a, b, n, flag);
}

__global__ void k_copy(double *d, const double *a, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) d[i] = a[i];
}
void launch_copy(hipStream_t s, double *dst, const double *src, int64_t n)
{
    k_copy<<<grid_for(n, TPB), TPB, 0, s>>>(dst, src, n);
}

__global__ void k_diag(const double *r, const double *pd, double *z, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        z[i] = r[i] * pd[i];
}
void launch_diag_precond(hipStream_t s, const double *r, const double *pd, double *z, int64_t n)
{
    k_diag<<<grid_for(n, TPB), TPB, 0, s>>>(r, pd, z, n);
}

// construct_preconditioner, diagonal branch (src/noise_routines.f90:943-974):
// no PSD -> 1, nna == 0 -> 1 / w, else 1 / nna.  wb_inv holds 1/w per baseline.
__global__ void k_prec_diag(const double *nna, const double *wb_inv, const int8_t *haspsd, double *pd, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double v;
        if (!haspsd[i]) v = 1.0;
        else if (nna[i] == 0.0) v = wb_inv[i];
        else v = 1.0 / nna[i];
        pd[i] = v;
    }
}
void launch_prec_diag_build(hipStream_t s, const double *nna, const double *wb_inv, const int8_t *haspsd, double *pd,
                            int64_t n)
{
    k_prec_diag<<<grid_for(n, TPB), TPB, 0, s>>>(nna, wb_inv, haspsd, pd, n);
}

__global__ void k_nna_term(double *ap, const double *nna, const double *p, double f, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        ap[i] = nna[i] * p[i] * f;
}
void launch_axpby_nna(hipStream_t s, double *ap, const double *nna, const double *p, double factor, int64_t n)
{
    k_nna_term<<<grid_for(n, TPB), TPB, 0, s>>>(ap, nna, p, factor, n);
}

// ---------------------------------------------------------------- deterministic masked dots

static constexpr int DOT_BLOCKS_MAX = 8192;
static constexpr int DOT_CHUNK = 8;   // elements per thread per step (one 8-byte mask load)

int dots_nblocks(int64_t n)
{
    int64_t b = (n + TPB * DOT_CHUNK - 1) / (TPB * DOT_CHUNK);
    if (b < 1) b = 1;
    if (b > DOT_BLOCKS_MAX) b = DOT_BLOCKS_MAX;
    return (int)b;
}

// which = 0: sum a*b (1 value)
// which = 1: sum a*b, sum c*b, sum a*a (3 values: rz, rz2, rr with a=r, b=z, c=ro)
// Only elements with mask != 0 count.  Each thread takes chunks of 8 elements
// and issues all of a chunk's loads (mask as one 8-byte word, the vectors as
// 16-byte pairs) before the arithmetic; the block partials and their final
// sum are in a fixed order, so the result depends on n alone.
__global__ void __launch_bounds__(TPB) k_dots(int which, const double *a, const double *b, const double *c,
                                              const int8_t *mask, int64_t n, double *partials)
{
    double s0 = 0, s1 = 0, s2 = 0;
    int64_t nch = n / DOT_CHUNK;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t gt = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    for (int64_t ch = gt; ch < nch; ch += stride) {
        uint64_t mk = ((const uint64_t *)mask)[ch];
        double2 A[DOT_CHUNK / 2], B[DOT_CHUNK / 2], C[DOT_CHUNK / 2];
#pragma unroll
        for (int u = 0; u < DOT_CHUNK / 2; ++u) {
            A[u] = ((const double2 *)a)[ch * (DOT_CHUNK / 2) + u];
            B[u] = ((const double2 *)b)[ch * (DOT_CHUNK / 2) + u];
            if (which == 1) C[u] = ((const double2 *)c)[ch * (DOT_CHUNK / 2) + u];
        }
#pragma unroll
        for (int u = 0; u < DOT_CHUNK / 2; ++u)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                bool m = ((mk >> (8 * (2 * u + h))) & 0xff) != 0;
                double ai = h ? A[u].y : A[u].x, bi = h ? B[u].y : B[u].x;
                s0 += m ? ai * bi : 0.0;
                if (which == 1) {
                    double ci = h ? C[u].y : C[u].x;
                    s1 += m ? ci * bi : 0.0;
                    s2 += m ? ai * ai : 0.0;
                }
            }
    }
    for (int64_t i = nch * DOT_CHUNK + gt; i < n; i += stride) {
        if (!mask[i]) continue;
        double ai = a[i], bi = b[i];
        s0 += ai * bi;
        if (which == 1) {
            s1 += c[i] * bi;
            s2 += ai * ai;
        }
    }
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    __shared__ double sh[3][TPB / 64];
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { sh[0][wv] = s0; sh[1][wv] = s1; sh[2][wv] = s2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t0 = 0, t1 = 0, t2 = 0;
        for (int k = 0; k < TPB / 64; ++k) { t0 += sh[0][k]; t1 += sh[1][k]; t2 += sh[2][k]; }
        partials[blockIdx.x * 3] = t0;
        partials[blockIdx.x * 3 + 1] = t1;
        partials[blockIdx.x * 3 + 2] = t2;
    }
}

void launch_dots(hipStream_t s, int which, const double *a, const double *b, const double *c, const double *d,
                 const int8_t *mask, int64_t n, double *partials, int nblocks)
{
    (void)d;
    k_dots<<<nblocks, TPB, 0, s>>>(which, a, b, c, mask, n, partials);
}

// sums the block partials in a fixed order into out[0..nval): thread t adds
// blocks t, t + 1024, ... (eight loads in flight per value), then a tree
static constexpr int FIN_TPB = 1024;
// with scal (out = scal + 2): then beta = (scal[2] - scal[3]) / scal[0] into
// scal[5], the old rz into scal[6] and the new rz into scal[0]
__global__ void __launch_bounds__(FIN_TPB) k_finalize(const double *partials, int nblocks, int nval, double *out,
                                                      double *scal)
{
    __shared__ double sh[3][FIN_TPB];
    double t[3] = {0, 0, 0};
    for (int i0 = threadIdx.x; i0 < nblocks; i0 += 8 * FIN_TPB) {
        double v[8][3];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            int i = i0 + u * FIN_TPB;
#pragma unroll
            for (int q = 0; q < 3; ++q) v[u][q] = i < nblocks ? partials[i * 3 + q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int q = 0; q < 3; ++q) t[q] += v[u][q];
    }
    for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] = t[q];
    __syncthreads();
    for (int o = FIN_TPB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int q = 0; q < nval; ++q) out[q] = sh[q][0];
        if (scal) {
            double beta = (scal[2] - scal[3]) / scal[0];
            scal[5] = beta;
            scal[6] = scal[0];
            scal[0] = scal[2];
        }
    }
}

void launch_finalize(hipStream_t s, const double *partials, int nblocks, int nval, double *out)
{
    k_finalize<<<1, FIN_TPB, 0, s>>>(partials, nblocks, nval, out, nullptr);
}

void launch_finalize_scalars(hipStream_t s, const double *partials, int nblocks, double *scal)
{
    k_finalize<<<1, FIN_TPB, 0, s>>>(partials, nblocks, 3, scal + 2, scal);
}

// scal: [0] rz, [1] pap, [2] rz_new, [3] rz2, [4] rr, [5] beta
__global__ void k_update_r(const double *scal, double *aa, double *r, double *ro, const double *p, const double *ap,
                           int64_t n)
{
    double alpha = scal[0] / scal[1];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double ri = r[i];
        ro[i] = ri;
        aa[i] = aa[i] + alpha * p[i];
        r[i] = ri - alpha * ap[i];
    }
}

void launch_pcg_update_r(hipStream_t s, const double *scal, double *aa, double *r, double *ro, const double *p,
                         const double *ap, int64_t n)
{
    k_update_r<<<grid_for(n, TPB), TPB, 0, s>>>(scal, aa, r, ro, p, ap, n);
}

__global__ void k_update_p(const double *scal, double *p, const double *z, int64_t n)
{
    double beta = scal[5];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = z[i] + beta * p[i];
}

void launch_pcg_update_p(hipStream_t s, const double *scal, double *p, const double *z, int64_t n)
{
    k_update_p<<<grid_for(n, TPB), TPB, 0, s>>>(scal, p, z, n);
}

__global__ void k_gather_order(const double *in, const int32_t *ord, double *out, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = in[ord[i]];
}

void launch_gather_order(hipStream_t s, const double *in, const int32_t *ord, double *out, int64_t n)
{
    if (n > 0) k_gather_order<<<grid_for(n, TPB), TPB, 0, s>>>(in, ord, out, n);
}

__global__ void k_pw_slots(const double *p, const int32_t *ord, const double *dwb_s, double *pw, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        pw[i] = p[ord[i]] * dwb_s[i];
}

void launch_pw_slots(hipStream_t s, const double *p, const int32_t *ord, const double *dwb_s, double *pw, int64_t n)
{
    if (n > 0) k_pw_slots<<<grid_for(n, TPB), TPB, 0, s>>>(p, ord, dwb_s, pw, n);
}

}  // namespace madam

namespace madam {

__global__ void k_minmax(const int32_t *pix, int64_t n, int *mn, int *mx)
{
    int lo = INT32_MAX, hi = INT32_MIN;
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
        int v = pix[g];
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    }
    for (int o = 32; o > 0; o >>= 1) {
        int a = __shfl_xor(lo, o, 64), b = __shfl_xor(hi, o, 64);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(mn, lo);
        atomicMax(mx, hi);
    }
}

void launch_minmax(hipStream_t s, const int32_t *pix, int64_t n, int *mn, int *mx)
{
    k_minmax<<<grid_for(n, TPB * 8, 4096), TPB, 0, s>>>(pix, n, mn, mx);
}

__global__ void k_restore(int32_t *pix, int64_t n, int32_t dummy)
{
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x)
        if (pix[g] < 0) pix[g] = dummy;
}

void launch_restore_pixels(hipStream_t s, int32_t *pix, int64_t n, int32_t dummy)
{
    k_restore<<<grid_for(n, TPB), TPB, 0, s>>>(pix, n, dummy);
}

// the dummy pixel a previous call left in the caller's array flags the sample again
__global__ void k_dummy_to_flag(int32_t *pix, int64_t n, int32_t dummy)
{
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x)
        if (pix[g] == dummy) pix[g] = -1;
}

void launch_dummy_to_flag(hipStream_t s, int32_t *pix, int64_t n, int32_t dummy)
{
    if (n > 0) k_dummy_to_flag<<<grid_for(n, TPB), TPB, 0, s>>>(pix, n, dummy);
}

__global__ void k_words_differ(const uint32_t *a, const uint32_t *b, int64_t n, int *flag)
{
    bool d = false;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d = d || a[i] != b[i];
    if (d) flag[0] = 1;
}

void launch_words_differ(hipStream_t s, const uint32_t *a, const uint32_t *b, int64_t n, int *flag)
{
    if (n > 0) k_words_differ<<<grid_for(n, TPB), TPB, 0, s>>>(a, b, n, flag);
}

// beta = (rz_new - rz2) / rz_old; rz_old <- rz_new   (scal layout as k_update_r)
// fprecond rows: 1 / (nshort * w + fcov)  (src/noise_routines.f90:977-993)
__global__ void k_fprecond(const double *fcov, int64_t nh, const double *wpsd, int nshort, double *out, int64_t npsd)
{
    int64_t n = nh * npsd;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t p = i / nh;
        out[i] = 1.0 / (nshort * wpsd[p] + fcov[i]);
    }
}

void launch_fprecond(hipStream_t s, const double *fcov, int64_t nh, const double *wpsd, int nshort, double *out,
                     int64_t npsd)
{
    k_fprecond<<<grid_for(nh * npsd, TPB), TPB, 0, s>>>(fcov, nh, wpsd, nshort, out, npsd);
}

}  // namespace madam

// ---------------------------------------------------------------- inner-CG preconditioner
// apply_CG_preconditioner (src/noise_routines.f90:1305-1409), batched over the
// segments whose band Cholesky was not available; per-segment scalars.
namespace madam {

static constexpr int CG_NB = 32;

// x = 0, resid = r, prec = 1 / (nna + invcov(0)) where nonzero, zr = prec resid, p = zr
__global__ void k_cg_init(const Segment *seg, const int *act, const double *ic0, const double *r, const double *nna,
                          double *x, double *resid, double *prec, double *zr, double *p)
{
    int64_t s = blockIdx.y;
    if (!act[s]) return;
    Segment g = seg[s];
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < g.noba; j += (int64_t)gridDim.x * blockDim.x) {
        int64_t b = g.b0 + j;
        x[b] = 0.0;
        double rs = r[b] - 0.0;
        resid[b] = rs;
        double pc = nna[b] + ic0[s];
        if (pc != 0) pc = 1 / pc;
        prec[b] = pc;
        double z = pc * rs;
        zr[b] = z;
        p[b] = z;
    }
}

// partial[s * CG_NB + blk] = sum a * b over segment s
__global__ void k_cg_dot(const Segment *seg, const int *act, const double *a, const double *bv, double *partial)
{
    int64_t s = blockIdx.y;
    if (!act[s]) return;
    Segment g = seg[s];
    double acc = 0;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < g.noba; j += (int64_t)CG_NB * blockDim.x)
        acc += a[g.b0 + j] * bv[g.b0 + j];
    acc = wave_sum(acc);
    __shared__ double sh[TPB / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int k = 0; k < TPB / 64; ++k) t += sh[k];
        partial[s * CG_NB + blockIdx.x] = t;
    }
}

__global__ void k_cg_dot_final(const double *partial, const int *act, int64_t nseg, double *out)
{
    int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= nseg || !act[s]) return;
    double t = 0;
    for (int k = 0; k < CG_NB; ++k) t += partial[s * CG_NB + k];
    out[s] = t;
}

// ap += nna * p
__global__ void k_cg_nna(const Segment *seg, const int *act, const double *nna, const double *p, double *ap)
{
    int64_t s = blockIdx.y;
    if (!act[s]) return;
    Segment g = seg[s];
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < g.noba; j += (int64_t)gridDim.x * blockDim.x) {
        int64_t b = g.b0 + j;
        ap[b] = ap[b] + nna[b] * p[b];
    }
}

// x += alpha p; resid -= alpha ap
__global__ void k_cg_axpy(const Segment *seg, const int *act, const double *alpha, double *x, const double *p,
                          double *resid, const double *ap)
{
    int64_t s = blockIdx.y;
    if (!act[s]) return;
    Segment g = seg[s];
    double al = alpha[s];
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < g.noba; j += (int64_t)gridDim.x * blockDim.x) {
        int64_t b = g.b0 + j;
        x[b] = x[b] + al * p[b];
        resid[b] = resid[b] - al * ap[b];
    }
}

// zr = prec * resid
__global__ void k_cg_prec(const Segment *seg, const int *act, const double *prec, const double *resid, double *zr)
{
    int64_t s = blockIdx.y;
    if (!act[s]) return;
    Segment g = seg[s];
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < g.noba; j += (int64_t)gridDim.x * blockDim.x) {
        int64_t b = g.b0 + j;
        zr[b] = prec[b] * resid[b];
    }
}

// p = zr + beta p
__global__ void k_cg_p(const Segment *seg, const int *act, const double *beta, const double *zr, double *p)
{
    int64_t s = blockIdx.y;
    if (!act[s]) return;
    Segment g = seg[s];
    double bt = beta[s];
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < g.noba; j += (int64_t)gridDim.x * blockDim.x) {
        int64_t b = g.b0 + j;
        p[b] = zr[b] + bt * p[b];
    }
}

void launch_cg_init(hipStream_t s, const Segment *seg, int64_t nseg, const int *act, const double *ic0, const double *r,
                    const double *nna, double *x, double *resid, double *prec, double *zr, double *p)
{
    k_cg_init<<<dim3(64, (unsigned)nseg), TPB, 0, s>>>(seg, act, ic0, r, nna, x, resid, prec, zr, p);
}
void launch_cg_dot(hipStream_t s, const Segment *seg, int64_t nseg, const int *act, const double *a, const double *b,
                   double *partial, double *out)
{
    k_cg_dot<<<dim3(CG_NB, (unsigned)nseg), TPB, 0, s>>>(seg, act, a, b, partial);
    k_cg_dot_final<<<(unsigned)((nseg + 63) / 64), 64, 0, s>>>(partial, act, nseg, out);
}
void launch_cg_nna(hipStream_t s, const Segment *seg, int64_t nseg, const int *act, const double *nna, const double *p,
                   double *ap)
{
    k_cg_nna<<<dim3(64, (unsigned)nseg), TPB, 0, s>>>(seg, act, nna, p, ap);
}
void launch_cg_axpy(hipStream_t s, const Segment *seg, int64_t nseg, const int *act, const double *alpha, double *x,
                    const double *p, double *resid, const double *ap)
{
    k_cg_axpy<<<dim3(64, (unsigned)nseg), TPB, 0, s>>>(seg, act, alpha, x, p, resid, ap);
}
void launch_cg_prec(hipStream_t s, const Segment *seg, int64_t nseg, const int *act, const double *prec,
                    const double *resid, double *zr)
{
    k_cg_prec<<<dim3(64, (unsigned)nseg), TPB, 0, s>>>(seg, act, prec, resid, zr);
}
void launch_cg_p(hipStream_t s, const Segment *seg, int64_t nseg, const int *act, const double *beta, const double *zr,
                 double *p)
{
    k_cg_p<<<dim3(64, (unsigned)nseg), TPB, 0, s>>>(seg, act, beta, zr, p);
}

}  // namespace madam
