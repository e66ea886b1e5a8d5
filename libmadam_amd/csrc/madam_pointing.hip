// Pointing expansion on the GPU: synthetic ring-scan attitude -> nested
// HEALPix pixel (int32) and (1, cos 2psi, sin 2psi) weights (float32).
//
// The reference receives pixels and weights from its caller
// (src/pointing.f90:13-14); its unused HEALPix routines
// (src/healpix_routines.f90:36-221, 407-455) are the model for the nested
// ang2pix.  Every step uses +, -, *, / and sqrt (plus the exact floor) with
// floating-point contraction disabled, so the pixels are bit-identical to the
// CPU expansion that performs the same steps in the same order.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>

#include "madam_internal.hpp"

#pragma clang fp contract(off)

namespace madam {

#define HALF_PI_D 1.5707963267948966192313216916398
#define PI_DD 3.1415926535897932384626433832795
#define SQRT3_D 1.7320508075688772935274463415059
#define TAN_PI_12_D 0.26794919243112270647255365849413

// sin and cos of 2 pi u (u in turns): exact reduction to an octant, Taylor
// polynomials to degree 19 / 20 on [0, pi/4]
__device__ void sincos_turns(double u, double *s, double *c)
{
    double f = u - floor(u);
    double q4 = f * 4.0;
    int quad = (int)q4;
    if (quad > 3) quad = 3;
    double r = q4 - (double)quad;
    int swap = 0;
    if (r > 0.5) { r = 1.0 - r; swap = 1; }
    double x = r * HALF_PI_D;
    double x2 = x * x;
    double sp = (1.0 / 121645100408832000.0);
    sp = -(1.0 / 355687428096000.0) + x2 * sp;
    sp = (1.0 / 1307674368000.0) + x2 * sp;
    sp = -(1.0 / 6227020800.0) + x2 * sp;
    sp = (1.0 / 39916800.0) + x2 * sp;
    sp = -(1.0 / 362880.0) + x2 * sp;
    sp = (1.0 / 5040.0) + x2 * sp;
    sp = -(1.0 / 120.0) + x2 * sp;
    sp = (1.0 / 6.0) + x2 * sp;
    double sn = x - x * (x2 * sp);
    double cp = (1.0 / 2432902008176640000.0);
    cp = -(1.0 / 6402373705728000.0) + x2 * cp;
    cp = (1.0 / 20922789888000.0) + x2 * cp;
    cp = -(1.0 / 87178291200.0) + x2 * cp;
    cp = (1.0 / 479001600.0) + x2 * cp;
    cp = -(1.0 / 3628800.0) + x2 * cp;
    cp = (1.0 / 40320.0) + x2 * cp;
    cp = -(1.0 / 720.0) + x2 * cp;
    cp = (1.0 / 24.0) + x2 * cp;
    cp = 0.5 - x2 * cp;
    double cs = 1.0 - x2 * cp;
    if (swap) { double t = sn; sn = cs; cs = t; }
    switch (quad) {
    case 0: *s = sn; *c = cs; break;
    case 1: *s = cs; *c = -sn; break;
    case 2: *s = -sn; *c = -cs; break;
    default: *s = -cs; *c = sn; break;
    }
}

__device__ double atan_unit(double t)
{
    double base = 0.0;
    if (t > TAN_PI_12_D) {
        t = (t * SQRT3_D - 1.0) / (t + SQRT3_D);
        base = PI_DD / 6.0;
    }
    double t2 = t * t;
    double p = (1.0 / 37.0);
    p = (1.0 / 35.0) - t2 * p;
    p = (1.0 / 33.0) - t2 * p;
    p = (1.0 / 31.0) - t2 * p;
    p = (1.0 / 29.0) - t2 * p;
    p = (1.0 / 27.0) - t2 * p;
    p = (1.0 / 25.0) - t2 * p;
    p = (1.0 / 23.0) - t2 * p;
    p = (1.0 / 21.0) - t2 * p;
    p = (1.0 / 19.0) - t2 * p;
    p = (1.0 / 17.0) - t2 * p;
    p = (1.0 / 15.0) - t2 * p;
    p = (1.0 / 13.0) - t2 * p;
    p = (1.0 / 11.0) - t2 * p;
    p = (1.0 / 9.0) - t2 * p;
    p = (1.0 / 7.0) - t2 * p;
    p = (1.0 / 5.0) - t2 * p;
    p = (1.0 / 3.0) - t2 * p;
    return base + (t - t * (t2 * p));
}

__device__ double atan2_pos(double y, double x)
{
    double ax = x < 0 ? -x : x, ay = y < 0 ? -y : y;
    if (ax == 0.0 && ay == 0.0) return 0.0;
    double a;
    if (ay <= ax) a = atan_unit(ay / ax);
    else a = HALF_PI_D - atan_unit(ax / ay);
    if (x < 0) a = PI_DD - a;
    if (y < 0) a = 2.0 * PI_DD - a;
    if (a >= 2.0 * PI_DD) a = 0.0;
    return a;
}

__device__ inline int64_t spread_bits(int64_t v)
{
    int64_t r = 0;
    for (int b = 0; b < 31; ++b) r |= ((v >> b) & 1) << (2 * b);
    return r;
}

__device__ int64_t vec2pix_nest(int64_t nside, double x, double y, double z)
{
    double rho2 = x * x + y * y;
    double za = z < 0 ? -z : z;
    double phi = atan2_pos(y, x);
    double tt = phi * (2.0 / PI_DD);
    if (tt >= 4.0) tt = 0.0;
    int64_t face, ix, iy;
    double dn = (double)nside;
    if (za <= 2.0 / 3.0) {
        double temp1 = dn * (0.5 + tt);
        double temp2 = dn * (z * 0.75);
        int64_t jp = (int64_t)(temp1 - temp2);
        int64_t jm = (int64_t)(temp1 + temp2);
        int64_t ifp = jp / nside, ifm = jm / nside;
        if (ifp == ifm) face = (ifp == 4 ? 0 : ifp) + 4;
        else if (ifp < ifm) face = ifp;
        else face = ifm + 8;
        ix = jm & (nside - 1);
        iy = nside - (jp & (nside - 1)) - 1;
    } else {
        int64_t ntt = (int64_t)tt;
        if (ntt > 3) ntt = 3;
        double tp = tt - (double)ntt;
        double one_minus = rho2 / (1.0 + za);
        double tmp = dn * sqrt(3.0 * one_minus);
        int64_t jp = (int64_t)(tp * tmp);
        int64_t jm = (int64_t)((1.0 - tp) * tmp);
        if (jp > nside - 1) jp = nside - 1;
        if (jm > nside - 1) jm = nside - 1;
        if (z >= 0) { face = ntt; ix = nside - jm - 1; iy = nside - jp - 1; }
        else { face = ntt + 8; ix = jp; iy = jm; }
    }
    return face * nside * nside + (spread_bits(ix) | (spread_bits(iy) << 1));
}

// grid over (det, sample); output pix[det * ld + k], w[(det * ld + k) * 3 + j]
__global__ void k_expand_pointing(ScanParams sp, const ScanDet *dets, int ndet, int64_t nside, int64_t i0, int64_t n,
                                  int64_t ld, int32_t *pix, float *w)
{
    int64_t tot = (int64_t)ndet * n;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        int d = (int)(t / n);
        int64_t k = t % n;
        ScanDet det = dets[d];
        double cb, sb, cpsi, spsi;
        sincos_turns(sp.opening_turns + det.open_offset_turns, &sb, &cb);
        sincos_turns(det.psi_turns, &spsi, &cpsi);
        double spin_samples = sp.fsample * sp.spin_period;
        double prec_samples = sp.fsample * sp.prec_period;
        double di = (double)(i0 + k);
        double us = di / spin_samples + det.spin_offset_turns;
        double ul = sp.lon0_turns + di / prec_samples;
        double sw, cw, sl, cl;
        sincos_turns(us, &sw, &cw);
        sincos_turns(ul, &sl, &cl);
        double ux = -(sw * sl), uy = sw * cl, uz = cw;
        double bx = cb * cl + sb * ux;
        double by = cb * sl + sb * uy;
        double bz = sb * uz;
        double dx = -(cw * sl), dy = cw * cl, dz = -sw;
        double nx = by * dz - bz * dy;
        double ny = bz * dx - bx * dz;
        double nz = bx * dy - by * dx;
        double ex = cpsi * dx + spsi * nx;
        double ey = cpsi * dy + spsi * ny;
        double ez = cpsi * dz + spsi * nz;
        double rho2 = bx * bx + by * by;
        double c2, s2;
        if (rho2 > 0.0) {
            double rho = sqrt(rho2);
            double eth_x = bz * bx / rho, eth_y = bz * by / rho, eth_z = -rho;
            double eph_x = -by / rho, eph_y = bx / rho;
            double u = ex * eth_x + ey * eth_y + ez * eth_z;
            double v = ex * eph_x + ey * eph_y;
            double nrm = u * u + v * v;
            c2 = (u * u - v * v) / nrm;
            s2 = -(2.0 * u * v) / nrm;
        } else {
            c2 = 1.0;
            s2 = 0.0;
        }
        int64_t o = (int64_t)d * ld + k;
        pix[o] = (int32_t)vec2pix_nest(nside, bx, by, bz);
        w[o * 3] = 1.0f;
        w[o * 3 + 1] = (float)c2;
        w[o * 3 + 2] = (float)s2;
    }
}

void launch_expand_pointing(hipStream_t s, ScanParams sp, const ScanDet *dets, int ndet, int64_t nside, int64_t i0,
                            int64_t n, int64_t ld, int32_t *pix, float *w)
{
    int64_t tot = (int64_t)ndet * n;
    int64_t g = (tot + 255) / 256;
    if (g > 65535 * 8) g = 65535 * 8;
    if (g < 1) g = 1;
    k_expand_pointing<<<(unsigned)g, 256, 0, s>>>(sp, dets, ndet, nside, i0, n, ld, pix, w);
}

// ---------------------------------------------------------------- the entry point that allocates nothing
//
// madam_expand_pointing stages the detector table in device memory it
// allocates per call.  The variant below stages it in `scratch`, a fixed
// array in the library's device image instead: each GPU the library runs on
// holds its own copy (NOALLOC_MAX_DETS * 24 bytes = 96 KiB) until the process
// ends.  Host code never passes `scratch` itself to a kernel (its host-side
// name is not a device address): the device address comes from
// hipGetSymbolAddress, once per device, with that device current.

static constexpr int NOALLOC_MAX_DETS = 4096;
static constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) ScanDet scratch[NOALLOC_MAX_DETS];

static std::mutex g_scratch_mu;
static void *g_scratch_dev[NOALLOC_MAX_DEVICES];   // scratch's device address, per device

// scratch's address on the current device (looked up on the device's first call)
static hipError_t scratch_on_current_device(void **out, int *dev_out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    *dev_out = dev;
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return hipSuccess;
}

}  // namespace madam

using namespace madam;

extern "C" __attribute__((visibility("default"))) int madam_expand_pointing_noalloc(
    const double *scan5, const double *dets3, int ndet, long nside, long i0, long n, int32_t *pix, float *weights,
    void *stream)
{
    char msg[256];
    if (ndet < 1 || ndet > NOALLOC_MAX_DETS) {
        snprintf(msg, sizeof msg, "madam_expand_pointing_noalloc: ndet = %d outside 1..%d (the staging array's size)",
                 ndet, NOALLOC_MAX_DETS);
        set_last_error(msg);
        return 1;
    }
    if (!scan5 || !dets3 || !pix || !weights || n < 0 || nside < 1 || (nside & (nside - 1)) || nside > (1L << 29)) {
        set_last_error("madam_expand_pointing_noalloc: null pointer, n < 0 or nside not a power of two");
        return 1;
    }
    void *sp = nullptr;
    int dev = -1;
    hipError_t e = scratch_on_current_device(&sp, &dev);
    if (e == hipSuccess) {
        hipStream_t s = (hipStream_t)stream;
        e = hipMemcpyAsync(sp, dets3, sizeof(ScanDet) * ndet, hipMemcpyHostToDevice, s);
        if (e == hipSuccess && n > 0) {
            ScanParams p{scan5[0], scan5[1], scan5[2], scan5[3], scan5[4]};
            launch_expand_pointing(s, p, (const ScanDet *)sp, ndet, nside, i0, n, n, pix, weights);
            e = hipGetLastError();
        }
    }
    if (e != hipSuccess) {
        snprintf(msg, sizeof msg, "madam_expand_pointing_noalloc: HIP error %d (%s) on device %d", (int)e,
                 hipGetErrorString(e), dev);
        set_last_error(msg);
        return 1;
    }
    return 0;
}
