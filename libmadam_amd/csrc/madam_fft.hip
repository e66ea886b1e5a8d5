// Batched FP64 real FFT convolution for the noise prior, and the noise filter.
//
// Reference: cinvmul -> convolve_pp -> convolve_interval
// (src/noise_routines.f90:745-913) on FFTW r2c/c2r plans
// (src/fourier_fftw_2003.f90:24-123), and build_filter / get_spectrum_interp
// / interpolate_psd (src/noise_routines.f90:130-190, 367-739).
//
// A real signal of length nof is packed as N = nof/2 complex points and
// transformed with a two-pass (four-step) Stockham-free Cooley-Tukey FFT,
// N = L1 * L2 with L1, L2 <= 1024, every sub-FFT done in LDS:
//   forward  pass A: per column n2, FFT over n1 (stride L2), twiddle W_N^(n2 k1)
//            pass B: per row k1, FFT over n2 (contiguous)        -> X[k1 + L1 k2]
//   spectral pass:   split into the real spectrum, multiply by the filter,
//                    merge back into the half-length inverse input
//   inverse  pass B': per row, inverse FFT; pass A': twiddle, inverse FFT per column
// All passes run in place on a [batch][N] double2 buffer.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

#include "madam_internal.hpp"

namespace madam {

#define F32(x) ((double)(float)(x))

static constexpr int FFT_TPB = 256;

__device__ inline double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cmulc(double2 a, double2 b)   // a * conj(b)
{
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ inline float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline float2 cmulc(float2 a, float2 b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// complex type of the fused passes' work buffer: double2 (FP64) or float2 (FP32)
template <typename C>
struct CxT;
template <>
struct CxT<double2> {
    using S = double;
    __device__ static double2 mk(double a, double b) { return make_double2(a, b); }
    __device__ static const double2 *sub1(const FftPlan &P) { return P.sub1; }
    __device__ static const double2 *sub2(const FftPlan &P) { return P.sub2; }
    __device__ static const double2 *na(const FftPlan &P) { return P.na; }
    __device__ static const double2 *nb(const FftPlan &P) { return P.nb; }
    __device__ static const double2 *fa(const FftPlan &P) { return P.fa; }
    __device__ static const double2 *fb(const FftPlan &P) { return P.fb; }
};
template <>
struct CxT<float2> {
    using S = float;
    __device__ static float2 mk(float a, float b) { return make_float2(a, b); }
    __device__ static const float2 *sub1(const FftPlan &P) { return P.sub1f; }
    __device__ static const float2 *sub2(const FftPlan &P) { return P.sub2f; }
    __device__ static const float2 *na(const FftPlan &P) { return P.naf; }
    __device__ static const float2 *nb(const FftPlan &P) { return P.nbf; }
    __device__ static const float2 *fa(const FftPlan &P) { return P.faf; }
    __device__ static const float2 *fb(const FftPlan &P) { return P.fbf; }
};

// W^m from a two-level table: a[m % 1024] * b[m / 1024]
template <typename C>
__device__ inline C tw2(const C *a, const C *b, int64_t m)
{
    return cmul(a[m & 1023], b[m >> 10]);
}

__device__ inline unsigned bitrev(unsigned v, int bits)
{
    return __brev(v) >> (32 - bits);
}

// In-place radix-2 FFT of `ncol` interleaved columns of length L held in LDS
// as s[row * ncol + col]; input already in bit-reversed row order.
// twL[j] = W_L^j for j < L / 2.
__device__ void lds_fft(double2 *s, int L, int logL, int ncol, const double2 *twL, bool inv)
{
    int nbf = (L / 2) * ncol;
    for (int st = 0; st < logL; ++st) {
        int half = 1 << st;
        for (int t = threadIdx.x; t < nbf; t += blockDim.x) {
            int col = t % ncol;
            int bf = t / ncol;
            int grp = bf >> st, j = bf & (half - 1);
            int i0 = grp * (2 * half) + j, i1 = i0 + half;
            double2 w = twL[j * (L / (2 * half))];
            if (inv) w.y = -w.y;
            double2 a = s[i0 * ncol + col], b = cmul(s[i1 * ncol + col], w);
            s[i0 * ncol + col] = make_double2(a.x + b.x, a.y + b.y);
            s[i1 * ncol + col] = make_double2(a.x - b.x, a.y - b.y);
        }
        __syncthreads();
    }
}

// Decimation-in-frequency form of lds_fft: natural-order input, output in
// bit-reversed row order (s[bitrev(k) * ncol + col] = X[k]).
__device__ void lds_fft_dif(double2 *s, int L, int logL, int ncol, const double2 *twL, bool inv)
{
    int nbf = (L / 2) * ncol;
    for (int st = logL - 1; st >= 0; --st) {
        int half = 1 << st;
        for (int t = threadIdx.x; t < nbf; t += blockDim.x) {
            int col = t % ncol;
            int bf = t / ncol;
            int grp = bf >> st, j = bf & (half - 1);
            int i0 = grp * (2 * half) + j, i1 = i0 + half;
            double2 w = twL[j * (L / (2 * half))];
            if (inv) w.y = -w.y;
            double2 a = s[i0 * ncol + col], b = s[i1 * ncol + col];
            s[i0 * ncol + col] = make_double2(a.x + b.x, a.y + b.y);
            s[i1 * ncol + col] = cmul(make_double2(a.x - b.x, a.y - b.y), w);
        }
        __syncthreads();
    }
}

// column pass; grid = (L2 / ncol, batch)
__global__ void k_fft_col(double2 *data, int64_t N, int L1, int logL1, int L2, int ncol, FftPlan P, int inv)
{
    extern __shared__ double2 s[];
    int64_t b = blockIdx.y;
    int c0 = blockIdx.x * ncol;
    double2 *d = data + b * N;
    int tot = L1 * ncol;
    if (!inv) {
        for (int t = threadIdx.x; t < tot; t += blockDim.x) {
            int col = t % ncol, n1 = t / ncol;
            s[bitrev(n1, logL1) * ncol + col] = d[(int64_t)n1 * L2 + c0 + col];
        }
        __syncthreads();
        lds_fft(s, L1, logL1, ncol, P.sub1, false);
        for (int t = threadIdx.x; t < tot; t += blockDim.x) {
            int col = t % ncol, k1 = t / ncol;
            int64_t n2 = c0 + col;
            double2 w = tw2(P.na, P.nb, (n2 * k1) & (N - 1))   /* N is a power of two */;
            d[(int64_t)k1 * L2 + n2] = cmul(s[k1 * ncol + col], w);
        }
    } else {
        for (int t = threadIdx.x; t < tot; t += blockDim.x) {
            int col = t % ncol, k1 = t / ncol;
            int64_t n2 = c0 + col;
            double2 w = tw2(P.na, P.nb, (n2 * k1) & (N - 1))   /* N is a power of two */;
            s[bitrev(k1, logL1) * ncol + col] = cmulc(d[(int64_t)k1 * L2 + n2], w);
        }
        __syncthreads();
        lds_fft(s, L1, logL1, ncol, P.sub1, true);
        for (int t = threadIdx.x; t < tot; t += blockDim.x) {
            int col = t % ncol, n1 = t / ncol;
            d[(int64_t)n1 * L2 + c0 + col] = s[n1 * ncol + col];
        }
    }
}

// row pass; grid = (L1 / nrow, batch); rows are contiguous runs of L2
__global__ void k_fft_row(double2 *data, int64_t N, int L1, int L2, int logL2, int nrow, FftPlan P, int inv)
{
    extern __shared__ double2 s[];
    int64_t b = blockIdx.y;
    int r0 = blockIdx.x * nrow;
    double2 *d = data + b * N;
    int tot = L2 * nrow;
    // s[n * nrow + row]
    for (int t = threadIdx.x; t < tot; t += blockDim.x) {
        int row = t / L2, n = t % L2;
        s[bitrev(n, logL2) * nrow + row] = d[(int64_t)(r0 + row) * L2 + n];
    }
    __syncthreads();
    lds_fft(s, L2, logL2, nrow, P.sub2, inv != 0);
    for (int t = threadIdx.x; t < tot; t += blockDim.x) {
        int row = t / L2, n = t % L2;
        d[(int64_t)(r0 + row) * L2 + n] = s[n * nrow + row];
    }
}

__device__ inline int64_t kidx(int64_t k, int L1, int L2)
{
    // X[k] with k = k1 + L1 * k2 lives at k1 * L2 + k2
    return (k % L1) * L2 + k / L1;
}

// real-spectrum split, filter multiply, merge; grid = (ceil((N/2+1)/TPB), batch)
__global__ void k_spectral(double2 *data, int64_t N, int L1, int L2, FftPlan P, const double *fc,
                           int64_t fc_stride, const int *ipsd_of_batch)
{
    int64_t b = blockIdx.y;
    double2 *d = data + b * N;
    const double *f = fc + (int64_t)ipsd_of_batch[b] * fc_stride;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k <= N / 2; k += (int64_t)gridDim.x * blockDim.x) {
        if (k == 0) {
            double2 a = d[0];
            double x0 = a.x + a.y, xn = a.x - a.y;
            double y0 = x0 * f[0], yn = xn * f[N];
            d[0] = make_double2(y0 + yn, y0 - yn);
            continue;
        }
        int64_t ik = kidx(k, L1, L2), im = kidx(N - k, L1, L2);
        double2 a = d[ik], bc = d[im];
        bc.y = -bc.y;   // conj(Z_{N-k})
        double2 E = make_double2(0.5 * (a.x + bc.x), 0.5 * (a.y + bc.y));
        double2 O = make_double2(0.5 * (a.y - bc.y), -0.5 * (a.x - bc.x));
        double2 W = tw2(P.fa, P.fb, k);
        double2 WO = cmul(W, O);
        double2 Xk = make_double2(E.x + WO.x, E.y + WO.y);
        double2 Xm = make_double2(E.x - WO.x, -(E.y - WO.y));   // X_{N-k} = conj(E - W O)
        double fk = f[k], fm = f[N - k];
        double2 Yk = make_double2(Xk.x * fk, Xk.y * fk);
        double2 Ym = make_double2(Xm.x * fm, Xm.y * fm);
        // z_k = (Yk + conj Ym) + i W^-k (Yk - conj Ym)
        double2 s1 = make_double2(Yk.x + Ym.x, Yk.y - Ym.y);
        double2 d1 = make_double2(Yk.x - Ym.x, Yk.y + Ym.y);
        double2 t1 = cmulc(d1, W);   // d1 * W^-k
        double2 zk = make_double2(s1.x - t1.y, s1.y + t1.x);
        d[ik] = zk;
        if (N - k != k) {
            // z_{N-k} = (Ym + conj Yk) - i W^k (Ym - conj Yk)
            double2 s2 = make_double2(Ym.x + Yk.x, Ym.y - Yk.y);
            double2 d2 = make_double2(Ym.x - Yk.x, Ym.y + Yk.y);
            double2 t2 = cmul(d2, W);
            double2 zm = make_double2(s2.x + t2.y, s2.y - t2.x);
            d[im] = zm;
        }
    }
}

// plain c2r of one filter row (no segment): imaginary parts ignored like FFTW
__global__ void k_c2r_prep(double2 *data, int64_t N, int L1, int L2, FftPlan P, const double *fc,
                           int64_t fc_stride)
{
    int64_t b = blockIdx.y;
    double2 *d = data + b * N;
    const double *Y = fc + b * fc_stride;   // real spectrum values
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < N; k += (int64_t)gridDim.x * blockDim.x) {
        double2 W = tw2(P.fa, P.fb, k <= N / 2 ? k : N - k);
        if (k > N / 2) { W.x = -W.x; }   // exp(-2 pi i k / nof) for k > N/2 from the mirror: -conj(W_{N-k})
        // with purely real Y: z_k = (Y_k + Y_{N-k}) + i W^-k (Y_k - Y_{N-k}), Y_N at index N
        double yk = Y[k], ym = Y[N - k];
        if (k == 0) { d[0] = make_double2(Y[0] + Y[N], Y[0] - Y[N]); continue; }
        double dr = yk - ym;
        double2 t = make_double2(dr * W.x, -dr * W.y);   // dr * conj(W)
        d[kidx(k, L1, L2)] = make_double2((yk + ym) - t.y, t.x);
    }
}

// Segment mean removal around the convolution (src/noise_routines.f90:840-850).
// Sums are deterministic: SEG_NB partial sums per segment, combined in order.
static constexpr int SEG_NB = 64;

// partial[i * SEG_NB + blk] = sum over the block's stride of scale * v[j], j < noba
template <typename V>
__global__ void k_seg_partial(const V *v, int64_t vstride, int use_b0, const Segment *seg, int64_t s0,
                              double scale, double *partial)
{
    int64_t i = blockIdx.y;
    Segment sg = seg[s0 + i];
    const V *x = use_b0 ? v + sg.b0 : v + i * vstride;
    double acc = 0;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < sg.noba; j += (int64_t)SEG_NB * blockDim.x)
        acc += (double)x[j] * scale;
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    __shared__ double sh[FFT_TPB / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int k = 0; k < FFT_TPB / 64; ++k) t += sh[k];
        partial[i * SEG_NB + blockIdx.x] = t;
    }
}

__device__ inline double seg_mean(const double *partial, int64_t i, int64_t noba)
{
    double t = 0;
    for (int k = 0; k < SEG_NB; ++k) t += partial[i * SEG_NB + k];
    return t / noba;
}

// work[i][j] = x[j] - mean for j < noba, 0 up to nof; grid (blocks, batch)
template <typename T>
__global__ void k_seg_load(const T *x, const Segment *seg, int64_t s0, int64_t nof, double *work,
                           const double *partial)
{
    int64_t i = blockIdx.y;
    Segment sg = seg[s0 + i];
    double mean = seg_mean(partial, i, sg.noba);
    const T *xs = x + sg.b0;
    double *w = work + i * nof;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < nof; j += (int64_t)gridDim.x * blockDim.x)
        w[j] = j < sg.noba ? (double)xs[j] - mean : 0.0;
}

// y[b0 + j] = work[i][j] / nof - mean(work / nof), j < noba
template <typename T>
__global__ void k_seg_store(const double *work, const Segment *seg, int64_t s0, int64_t nof, T *y,
                            const double *partial)
{
    int64_t i = blockIdx.y;
    Segment sg = seg[s0 + i];
    double inv = 1 / (double)nof;
    double mean = seg_mean(partial, i, sg.noba);
    const double *w = work + i * nof;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < sg.noba; j += (int64_t)gridDim.x * blockDim.x)
        y[sg.b0 + j] = (T)(w[j] * inv - mean);
}

__global__ void k_seg_ipsd(const Segment *seg, int64_t s0, int64_t nb, int *ipsd)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < nb) ipsd[i] = seg[s0 + i].ipsd;
}

static int ilog2(int64_t v)
{
    int r = 0;
    while ((1LL << r) < v) ++r;
    return r;
}

void fft_plan_init(FftPlan &p, int64_t nof, hipStream_t s)
{
    p.N = nof / 2;
    int logN = ilog2(p.N);
    // columns (stride L2, CONV_NCOL per block) no longer than the rows (contiguous, two per block)
    int l1 = logN / 2;
    int l2 = logN - l1;
    if (l2 > 11) { l2 = 11; l1 = logN - 11; }
    p.L1 = 1LL << l1;
    p.L2 = 1LL << l2;
    // exp(-2 pi i num / den), computed in long double and rounded once (to
    // FP64, and separately to FP32 for the FP32 tables)
    std::vector<long double> h;
    auto W = [&h](long double num, long double den) {
        long double a = -2.0L * 3.14159265358979323846264338327950288L * num / den;
        h.push_back(cosl(a));
        h.push_back(sinl(a));
    };
    int64_t n1 = std::max<int64_t>(p.L1 / 2, 1), n2 = std::max<int64_t>(p.L2 / 2, 1);
    int64_t nna = 1024, nnb = p.N / 1024 + 1;
    int64_t nfa = 1024, nfb = (p.N / 2) / 1024 + 1;
    int64_t tot = n1 + n2 + nna + nnb + nfa + nfb;
    h.reserve(2 * tot);
    for (int64_t j = 0; j < n1; ++j) W(j, p.L1);
    for (int64_t j = 0; j < n2; ++j) W(j, p.L2);
    for (int64_t j = 0; j < nna; ++j) W(j, p.N);
    for (int64_t j = 0; j < nnb; ++j) W(1024.0L * j, p.N);
    for (int64_t j = 0; j < nfa; ++j) W(j, nof);
    for (int64_t j = 0; j < nfb; ++j) W(1024.0L * j, nof);
    std::vector<double> hd(h.begin(), h.end());
    std::vector<float> hf(h.size());
    for (size_t i = 0; i < h.size(); ++i) hf[i] = (float)h[i];
    MADAM_CHECK(hipMalloc(&p.tw_mem, sizeof(double) * hd.size()));
    MADAM_CHECK(hipMemcpyAsync(p.tw_mem, hd.data(), sizeof(double) * hd.size(), hipMemcpyHostToDevice, s));
    MADAM_CHECK(hipMalloc(&p.tw32_mem, sizeof(float) * hf.size()));
    MADAM_CHECK(hipMemcpyAsync(p.tw32_mem, hf.data(), sizeof(float) * hf.size(), hipMemcpyHostToDevice, s));
    MADAM_CHECK(hipStreamSynchronize(s));   // hd and hf are freed on return
    const double2 *b = (const double2 *)p.tw_mem;
    p.sub1 = b;
    p.sub2 = p.sub1 + n1;
    p.na = p.sub2 + n2;
    p.nb = p.na + nna;
    p.fa = p.nb + nnb;
    p.fb = p.fa + nfa;
    const float2 *bf = (const float2 *)p.tw32_mem;
    p.sub1f = bf;
    p.sub2f = p.sub1f + n1;
    p.naf = p.sub2f + n2;
    p.nbf = p.naf + nna;
    p.faf = p.nbf + nnb;
    p.fbf = p.faf + nfa;
}

void fft_plan_free(FftPlan &p)
{
    if (p.tw_mem) (void)hipFree(p.tw_mem);
    if (p.tw32_mem) (void)hipFree(p.tw32_mem);
    p.tw_mem = nullptr;
    p.tw32_mem = nullptr;
    p.sub1 = p.sub2 = p.na = p.nb = p.fa = p.fb = nullptr;
    p.sub1f = p.sub2f = p.naf = p.nbf = p.faf = p.fbf = nullptr;
}

static void fft_forward(hipStream_t s, const FftPlan &p, double2 *d, int64_t batch)
{
    int logL1 = ilog2(p.L1), logL2 = ilog2(p.L2);
    int ncol = (int)(p.L2 < 4 ? p.L2 : 4);
    if (p.L1 * ncol > 4096) ncol = (int)(4096 / p.L1);
    if (ncol < 1) ncol = 1;
    dim3 gc((unsigned)(p.L2 / ncol), (unsigned)batch);
    k_fft_col<<<gc, FFT_TPB, sizeof(double2) * p.L1 * ncol, s>>>(d, p.N, (int)p.L1, logL1, (int)p.L2, ncol,
                                                                 p, 0);
    if (p.L2 > 1) {
        int nrow = (int)(4096 / p.L2);
        if (nrow > p.L1) nrow = (int)p.L1;
        if (nrow < 1) nrow = 1;
        dim3 gr((unsigned)(p.L1 / nrow), (unsigned)batch);
        k_fft_row<<<gr, FFT_TPB, sizeof(double2) * p.L2 * nrow, s>>>(d, p.N, (int)p.L1, (int)p.L2, logL2, nrow,
                                                                     p, 0);
    }
}

static void fft_inverse(hipStream_t s, const FftPlan &p, double2 *d, int64_t batch)
{
    int logL1 = ilog2(p.L1), logL2 = ilog2(p.L2);
    if (p.L2 > 1) {
        int nrow = (int)(4096 / p.L2);
        if (nrow > p.L1) nrow = (int)p.L1;
        if (nrow < 1) nrow = 1;
        dim3 gr((unsigned)(p.L1 / nrow), (unsigned)batch);
        k_fft_row<<<gr, FFT_TPB, sizeof(double2) * p.L2 * nrow, s>>>(d, p.N, (int)p.L1, (int)p.L2, logL2, nrow,
                                                                     p, 1);
    }
    int ncol = (int)(p.L2 < 4 ? p.L2 : 4);
    if (p.L1 * ncol > 4096) ncol = (int)(4096 / p.L1);
    if (ncol < 1) ncol = 1;
    dim3 gc((unsigned)(p.L2 / ncol), (unsigned)batch);
    k_fft_col<<<gc, FFT_TPB, sizeof(double2) * p.L1 * ncol, s>>>(d, p.N, (int)p.L1, logL1, (int)p.L2, ncol,
                                                                 p, 1);
}

// ---------------------------------------------------------------- fused-stage LDS FFTs
// Three radix-2 stages per LDS round trip: a work item holds the 8 elements
// base + m * h (m < 8) that those stages combine among themselves, applies the
// three butterfly levels in registers and writes them back.  The LDS layout is
// swizzled (lp: the low five bits of an element index xor the next five) so
// that the power-of-two strides of the stages spread over the banks; it
// permutes each group of 32 elements, so it needs no padding (it measured 3 %
// faster on the batched convolution than one padding element in eight).
__device__ inline int lp(int idx) { return idx ^ ((idx >> 5) & 31); }

// The twiddles of one work item of the fused-stage FFTs, from one table read:
// with b = W^(j / (h0 E)) (E = 2^T points per item, W = exp(-2 pi i), conjugated
// for the inverse), stage tt (half-length h0 2^tt) and butterfly r < 2^tt use
// W^((j + r h0) / (h0 2^(tt + 1))) = b^(2^(T - 1 - tt)) * W^(r / 2^(tt + 1)),
// the powers of b by repeated squaring and the second factors exact constants.
template <typename C>
__device__ inline C rot_const(C x, int num, int den, bool inv)
{
    // x * exp(-+ 2 pi i num / den), den in {2, 4, 8}, 0 <= num < den / 2
    using S = typename CxT<C>::S;
    const S h = (S)0.70710678118654752440;
    int k = num * (8 / den);   // in eighths of a turn: 0..3
    S sg = inv ? (S)1 : (S)-1;   // sign of the angle
    switch (k) {
    case 0: return x;
    case 1: {   // (h, sg h)
        return CxT<C>::mk(h * (x.x - sg * x.y), h * (x.y + sg * x.x));
    }
    case 2:     // sg i
        return CxT<C>::mk(-sg * x.y, sg * x.x);
    default: {  // (-h, sg h)
        return CxT<C>::mk(-h * (x.x + sg * x.y), h * (sg * x.x - x.y));
    }
    }
}

template <typename C>
__device__ inline void item_twiddles(C b, int T, bool inv, C w[3][4])
{
    if (inv) b.y = -b.y;
    C bt = b;
#pragma unroll
    for (int tt = 2; tt >= 0; --tt) {
        if (tt >= T) continue;
#pragma unroll
        for (int r = 0; r < (1 << tt); ++r) w[tt][r] = rot_const(bt, r, 2 << tt, inv);
        bt = cmul(bt, bt);
    }
}

// DIT, bit-reversed input order, natural output order; twL[j] = W_L^j (j < L / 2).
// 2^logncol interleaved transforms; every length is a power of two, so the
// index arithmetic is shifts and masks.
template <typename C>
__device__ void lds_fft8(C *s, int L, int logL, int logncol, const C *twL, bool inv)
{
    const int ncol = 1 << logncol;
    for (int st = 0; st < logL; st += 3) {
        int T = logL - st < 3 ? logL - st : 3;
        int h0 = 1 << st, E = 1 << T;
        int nitems = (L >> T) << logncol;
        for (int t = threadIdx.x; t < nitems; t += blockDim.x) {
            int col = t & (ncol - 1), it = t >> logncol;
            int j = it & (h0 - 1), grp = it >> st;
            int base = grp * (E * h0) + j;
            C a[8];
#pragma unroll
            for (int m = 0; m < 8; ++m)
                if (m < E) a[m] = s[lp((base + m * h0) * ncol + col)];
            C w[3][4];
            item_twiddles(twL[j << (logL - st - T)], T, inv, w);
#pragma unroll
            for (int tt = 0; tt < 3; ++tt) {
                if (tt >= T) break;
#pragma unroll
                for (int m0 = 0; m0 < 8; ++m0) {
                    if (m0 >= E || (m0 & (1 << tt))) continue;
                    int r = m0 & ((1 << tt) - 1);
                    C b = cmul(a[m0 + (1 << tt)], w[tt][r]), x = a[m0];
                    a[m0] = CxT<C>::mk(x.x + b.x, x.y + b.y);
                    a[m0 + (1 << tt)] = CxT<C>::mk(x.x - b.x, x.y - b.y);
                }
            }
#pragma unroll
            for (int m = 0; m < 8; ++m)
                if (m < E) s[lp((base + m * h0) * ncol + col)] = a[m];
        }
        __syncthreads();
    }
}

// DIF, natural input order, bit-reversed output order
template <typename C>
__device__ void lds_fft8_dif(C *s, int L, int logL, int logncol, const C *twL, bool inv)
{
    const int ncol = 1 << logncol;
    for (int hi = logL - 1; hi >= 0; hi -= 3) {
        int T = hi + 1 < 3 ? hi + 1 : 3;
        int lo = hi - T + 1;
        int h0 = 1 << lo, E = 1 << T;
        int nitems = (L >> T) << logncol;
        for (int t = threadIdx.x; t < nitems; t += blockDim.x) {
            int col = t & (ncol - 1), it = t >> logncol;
            int j = it & (h0 - 1), grp = it >> lo;
            int base = grp * (E * h0) + j;
            C a[8];
#pragma unroll
            for (int m = 0; m < 8; ++m)
                if (m < E) a[m] = s[lp((base + m * h0) * ncol + col)];
            C w[3][4];
            item_twiddles(twL[j << (logL - lo - T)], T, inv, w);
#pragma unroll
            for (int tt = 2; tt >= 0; --tt) {
                if (tt >= T) continue;
#pragma unroll
                for (int m0 = 0; m0 < 8; ++m0) {
                    if (m0 >= E || (m0 & (1 << tt))) continue;
                    int r = m0 & ((1 << tt) - 1);
                    C x = a[m0], y = a[m0 + (1 << tt)];
                    a[m0] = CxT<C>::mk(x.x + y.x, x.y + y.y);
                    a[m0 + (1 << tt)] = cmul(CxT<C>::mk(x.x - y.x, x.y - y.y), w[tt][r]);
                }
            }
#pragma unroll
            for (int m = 0; m < 8; ++m)
                if (m < E) s[lp((base + m * h0) * ncol + col)] = a[m];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- fused convolution
// Three passes instead of eight (load, 2 x column, 2 x row, spectral, store):
//   pass 1  forward column FFTs read x - mean straight from the baseline
//           vector (zero padding never leaves LDS), twiddle, store to work;
//   pass 2  one block per row pair (k1, L1 - k1): forward row FFTs, the
//           real-spectrum split / filter multiply / merge (every pair
//           {k, N - k} has one element in each row of the pair), inverse row
//           FFTs, store;
//   pass 3  inverse column FFTs write y = w / nof for the segment's
//           baselines directly, with deterministic per-block partial sums of
//           y for the mean, which a last pass subtracts.
static constexpr int CONV_NCOL = 8;
static_assert(CONV_NCOL == 8, "the column passes call the LDS FFTs with log2(CONV_NCOL) = 3");

// LDS elements for n points (lp maps [0, n) onto itself for n a multiple of
// 32, and onto [0, 32) below that)
__host__ __device__ inline int conv_lds_elems(int n) { return (n < 32 ? 32 : n) + 1; }
// The column passes hold L1 * CONV_NCOL points (74 KB of LDS at L1 = 512), so
// only two blocks fit on a CU: they run 512 threads, one 8-point work item each.
static constexpr int CONV_COL_TPB = 512;
static_assert(CONV_COL_TPB % CONV_NCOL == 0, "a column pass thread keeps one column");

// grid = (L2 / CONV_NCOL, batch)
// (C: the work type, double2 or float2; T: the type of x)
template <typename T, typename C>
__global__ void __launch_bounds__(CONV_COL_TPB) k_conv_col_fwd(const T *x, const Segment *seg, int64_t s0, const double *partial_in,
                               C *work, int64_t N, int L1, int logL1, int L2, FftPlan P)
{
    using S = typename CxT<C>::S;
    extern __shared__ __align__(16) unsigned char conv_lds[];
    C *s = reinterpret_cast<C *>(conv_lds);
    const int ncol = CONV_NCOL;
    int64_t b = blockIdx.y;
    // neighbouring column groups share the cache lines of x and work: keep
    // them on one XCD (gridDim.x is a multiple of 8, so every row of the grid
    // starts on XCD 0)
    int c0 = (int)xcd_block(blockIdx.x, gridDim.x) * ncol;
    Segment sg = seg[s0 + b];
    // the mean from the segment's SEG_NB partials: one per lane of the first
    // wave, a butterfly over the lanes (the same order in every block)
    __shared__ double smean;
    static_assert(SEG_NB == 64, "one partial per lane");
    if (threadIdx.x < 64) {
        double t = partial_in[b * SEG_NB + threadIdx.x];
        for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
        if (threadIdx.x == 0) smean = t / sg.noba;
    }
    __syncthreads();
    double mean = smean;
    const T *xs = x + sg.b0;
    int64_t noba = sg.noba;
    C *d = work + b * N;
    int tot = L1 * ncol;
    // the stage twiddles, read from LDS by every round of the FFT
    C *tw = s + conv_lds_elems(tot);
    for (int i = threadIdx.x; i < (L1 > 1 ? L1 / 2 : 1); i += blockDim.x) tw[i] = CxT<C>::sub1(P)[i];
    // (FP32 work: eight points per thread at a time, every load issued before
    // the LDS stores; FP64 keeps one at a time, the registers of eight cost
    // it more occupancy than the loads gain)
    constexpr int NB = sizeof(C) == 8 ? 8 : 1;
    for (int t0 = threadIdx.x; t0 < tot; t0 += NB * blockDim.x) {
        T a[NB], b2[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            int t = t0 + k * blockDim.x;
            int64_t jj = 2 * ((int64_t)(t / ncol) * L2 + c0 + t % ncol);
            a[k] = t < tot && jj < noba ? xs[jj] : (T)0;
            b2[k] = t < tot && jj + 1 < noba ? xs[jj + 1] : (T)0;
        }
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            int t = t0 + k * blockDim.x;
            if (t < tot) {
                int col = t % ncol, n1 = t / ncol;
                int64_t jj = 2 * ((int64_t)n1 * L2 + c0 + col);
                C v;
                v.x = jj < noba ? (S)((double)a[k] - mean) : (S)0;
                v.y = jj + 1 < noba ? (S)((double)b2[k] - mean) : (S)0;
                s[lp(bitrev(n1, logL1) * ncol + col)] = v;
            }
        }
    }
    __syncthreads();
    lds_fft8<C>(s, L1, logL1, 3, tw, false);   // CONV_NCOL = 8 columns
    // W_N^(n2 k1): a thread keeps its column and steps k1 by blockDim / ncol,
    // the twiddle by the matching power (two table reads per thread)
    {
        int col = threadIdx.x % ncol, kst = blockDim.x / ncol;
        int64_t n2 = c0 + col;
        C w = tw2(CxT<C>::na(P), CxT<C>::nb(P), (n2 * (int64_t)(threadIdx.x / ncol)) & (N - 1));
        C ws = tw2(CxT<C>::na(P), CxT<C>::nb(P), (n2 * kst) & (N - 1));   /* N is a power of two */
        for (int k1 = threadIdx.x / ncol; k1 < L1; k1 += kst) {
            d[(int64_t)k1 * L2 + n2] = cmul(s[lp(k1 * ncol + col)], w);
            w = cmul(w, ws);
        }
    }
}

// spectral step on one pair {k, N - k}: X[k] at s[ik], X[N - k] at s[im] (padded
// LDS indices), 0 < k <= N / 2
// fk = f[k], fm = f[N - k]
template <typename C>
__device__ inline void conv_pair(C *s, int ik, int im, int64_t k, int64_t N, const FftPlan &P, typename CxT<C>::S fk,
                                 typename CxT<C>::S fm)
{
    C a = s[ik], bc = s[im];
    bc.y = -bc.y;
    C E = CxT<C>::mk((typename CxT<C>::S)0.5 * (a.x + bc.x), (typename CxT<C>::S)0.5 * (a.y + bc.y));
    C O = CxT<C>::mk((typename CxT<C>::S)0.5 * (a.y - bc.y), -(typename CxT<C>::S)0.5 * (a.x - bc.x));
    C W = tw2(CxT<C>::fa(P), CxT<C>::fb(P), k);
    C WO = cmul(W, O);
    C Xk = CxT<C>::mk(E.x + WO.x, E.y + WO.y);
    C Xm = CxT<C>::mk(E.x - WO.x, -(E.y - WO.y));
    C Yk = CxT<C>::mk(Xk.x * fk, Xk.y * fk);
    C Ym = CxT<C>::mk(Xm.x * fm, Xm.y * fm);
    C s1 = CxT<C>::mk(Yk.x + Ym.x, Yk.y - Ym.y);
    C d1 = CxT<C>::mk(Yk.x - Ym.x, Yk.y + Ym.y);
    C t1 = cmulc(d1, W);
    s[ik] = CxT<C>::mk(s1.x - t1.y, s1.y + t1.x);
    if (ik != im) {
        C s2 = CxT<C>::mk(Ym.x + Yk.x, Ym.y - Yk.y);
        C d2 = CxT<C>::mk(Ym.x - Yk.x, Ym.y + Yk.y);
        C t2 = cmul(d2, W);
        s[im] = CxT<C>::mk(s2.x + t2.y, s2.y - t2.x);
    }
}

// grid = (L1 / 2, batch): block ra > 0 takes the row pair (ra, L1 - ra); block
// 0 takes rows 0 and L1 / 2, each paired with itself (L1 / 2 blocks per
// segment, a multiple of the blocks that fit the GPU at once, so the last
// round is not a nearly empty one).  The two rows are interleaved in LDS at
// k2 * 2 + r (padded).  fc holds the filter spectra in row order
// (fft_spectrum_rows): f[r + L1 k2] at [r * L2 + k2], f[N] at [N], so a row
// pair reads two contiguous rows.
template <typename C>
__global__ void k_conv_rows(C *work, int64_t N, int L1, int L2, int logL2, FftPlan P, const typename CxT<C>::S *fc,
                            int64_t fc_stride, const Segment *seg, int64_t s0)
{
    using S = typename CxT<C>::S;
    extern __shared__ __align__(16) unsigned char conv_lds[];
    C *s = reinterpret_cast<C *>(conv_lds);
    int64_t b = blockIdx.y;
    int ra = blockIdx.x, rb = ra == 0 ? L1 / 2 : L1 - ra;
    C *d = work + b * N;
    const S *f = fc + (int64_t)seg[s0 + b].ipsd * fc_stride;
    int tot = 2 * L2;
    // the stage twiddles: in LDS for FP32; FP64 reads the (cached) table, so
    // that its 32 KB of points alone set the blocks per CU (5, not 4)
    constexpr bool TWL = sizeof(C) == 8;
    const C *tw = CxT<C>::sub2(P);
    if (TWL) {
        C *twl = s + conv_lds_elems(tot);
        for (int i = threadIdx.x; i < (L2 > 1 ? L2 / 2 : 1); i += blockDim.x) twl[i] = tw[i];
        tw = twl;
    }
    // natural order in and out of LDS (no bank conflicts): the forward row
    // FFTs are DIF (bit-reversed output), the inverse ones DIT (bit-reversed
    // input), and the spectral step addresses the bit-reversed positions
    constexpr int NB = sizeof(C) == 8 ? 8 : 1;   // as in k_conv_col_fwd
    for (int t0 = threadIdx.x; t0 < tot; t0 += NB * blockDim.x) {
        C in[NB];   // every load issued before the LDS stores
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            int t = t0 + k * blockDim.x;
            if (t < tot) in[k] = d[(int64_t)((t >> logL2) == 0 ? ra : rb) * L2 + (t & (L2 - 1))];
        }
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            int t = t0 + k * blockDim.x;
            if (t < tot) s[lp((t & (L2 - 1)) * 2 + (t >> logL2))] = in[k];
        }
    }
    __syncthreads();
    lds_fft8_dif<C>(s, L2, logL2, 1, tw, false);   // 2 rows
    // X[ra + L1 k2] is at bitrev(k2) * 2; X[rb + L1 k2] at bitrev(k2) * 2 + 1
    if (ra == 0) {
        // row 0: pairs {k2, L2 - k2} (k2 = 0 holds X[0] and X[N / 2]);
        // row L1 / 2: pairs {k2, L2 - 1 - k2}
        const int h = L1 / 2;
        for (int t = threadIdx.x; t <= L2; t += blockDim.x) {
            if (t == 0) {
                C a = s[lp(0)];
                S x0 = a.x + a.y, xn = a.x - a.y;
                S y0 = x0 * f[0], yn = xn * f[N];   // f[0] at [0]
                s[lp(0)] = CxT<C>::mk(y0 + yn, y0 - yn);
            } else if (t <= L2 / 2) {
                int k2 = t;
                conv_pair(s, lp(bitrev(k2, logL2) * 2), lp(bitrev(L2 - k2, logL2) * 2), (int64_t)L1 * k2, N, P,
                          f[k2], f[L2 - k2]);
            } else {
                int k2 = t - (L2 / 2 + 1);
                conv_pair(s, lp(bitrev(k2, logL2) * 2 + 1), lp(bitrev(L2 - 1 - k2, logL2) * 2 + 1),
                          (int64_t)h + (int64_t)L1 * k2, N, P, f[(int64_t)h * L2 + k2],
                          f[(int64_t)h * L2 + L2 - 1 - k2]);
            }
        }
    } else {
        for (int k2 = threadIdx.x; k2 < L2; k2 += blockDim.x) {
            int64_t kA = (int64_t)ra + (int64_t)L1 * k2;
            int pp = lp(bitrev(k2, logL2) * 2), qq = lp(bitrev(L2 - 1 - k2, logL2) * 2 + 1);
            S fA = f[(int64_t)ra * L2 + k2], fB = f[(int64_t)rb * L2 + L2 - 1 - k2];   // f[kA], f[N - kA]
            if (kA <= N / 2) conv_pair(s, pp, qq, kA, N, P, fA, fB);
            else conv_pair(s, qq, pp, N - kA, N, P, fB, fA);
        }
    }
    __syncthreads();
    lds_fft8<C>(s, L2, logL2, 1, tw, true);
    for (int t = threadIdx.x; t < tot; t += blockDim.x) {
        int r = t >> logL2, n = t & (L2 - 1);
        int row = r == 0 ? ra : rb;
        d[(int64_t)row * L2 + n] = s[lp(n * 2 + r)];
    }
}

// grid = (L2 / CONV_NCOL, batch)
template <typename T, typename C>
__global__ void __launch_bounds__(CONV_COL_TPB) k_conv_col_inv(const C *work, const Segment *seg, int64_t s0, T *y, double *partial_out,
                               int64_t N, int L1, int logL1, int L2, FftPlan P)
{
    extern __shared__ __align__(16) unsigned char conv_lds[];
    C *s = reinterpret_cast<C *>(conv_lds);
    const int ncol = CONV_NCOL;
    int64_t b = blockIdx.y;
    int c0 = (int)xcd_block(blockIdx.x, gridDim.x) * ncol;   // as in k_conv_col_fwd
    Segment sg = seg[s0 + b];
    int64_t noba = sg.noba;
    const C *d = work + b * N;
    double inv = 1 / (double)(2 * N);
    int tot = L1 * ncol;
    C *tw = s + conv_lds_elems(tot);   // the stage twiddles in LDS
    for (int i = threadIdx.x; i < (L1 > 1 ? L1 / 2 : 1); i += blockDim.x) tw[i] = CxT<C>::sub1(P)[i];
    {
        int col = threadIdx.x % ncol, kst = blockDim.x / ncol;
        int64_t n2 = c0 + col;
        C w = tw2(CxT<C>::na(P), CxT<C>::nb(P), (n2 * (int64_t)(threadIdx.x / ncol)) & (N - 1));
        C ws = tw2(CxT<C>::na(P), CxT<C>::nb(P), (n2 * kst) & (N - 1));   /* N is a power of two */
        constexpr int NB = sizeof(C) == 8 ? 8 : 1;   // as in k_conv_col_fwd
        for (int k10 = threadIdx.x / ncol; k10 < L1; k10 += NB * kst) {
            C in[NB];   // every load issued before the twiddle products and LDS stores
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                int k1 = k10 + k * kst;
                if (k1 < L1) in[k] = d[(int64_t)k1 * L2 + n2];
            }
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                int k1 = k10 + k * kst;
                if (k1 < L1) {
                    s[lp(bitrev(k1, logL1) * ncol + col)] = cmulc(in[k], w);
                    w = cmul(w, ws);
                }
            }
        }
    }
    __syncthreads();
    lds_fft8<C>(s, L1, logL1, 3, tw, true);
    double acc = 0;
    T *ys = y + sg.b0;
    for (int t = threadIdx.x; t < tot; t += blockDim.x) {
        int col = t % ncol, n1 = t / ncol;
        int64_t jj = 2 * ((int64_t)n1 * L2 + c0 + col);
        C v = s[lp(n1 * ncol + col)];
        if (jj < noba) {
            double a = (double)v.x * inv;
            ys[jj] = (T)a;
            acc += a;
        }
        if (jj + 1 < noba) {
            double a = (double)v.y * inv;
            ys[jj + 1] = (T)a;
            acc += a;
        }
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    __shared__ double sh[CONV_COL_TPB / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int q = 0; q < CONV_COL_TPB / 64; ++q) t += sh[q];
        partial_out[b * gridDim.x + c0 / ncol] = t;
    }
}

// y[b0 + j] -= mean over the batch's segment b, the mean summed by the first
// wave from the segment's nblk partials (lane l adds partials l, l + 64, ...,
// then a butterfly over the lanes: the same order in every block)
template <typename T>
__global__ void k_conv_sub_mean(T *y, const Segment *seg, int64_t s0, const double *partial_out, int nblk)
{
    int64_t b = blockIdx.y;
    Segment sg = seg[s0 + b];
    __shared__ double smb;
    if (threadIdx.x < 64) {
        const double *po = partial_out + b * nblk;
        double t = 0;
        for (int k = threadIdx.x; k < nblk; k += 64) t += po[k];
        for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
        if (threadIdx.x == 0) smb = t / sg.noba;
    }
    __syncthreads();
    double mb = smb;
    T *ys = y + sg.b0;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < sg.noba; j += (int64_t)gridDim.x * blockDim.x)
        ys[j] = (T)((double)ys[j] - mb);
}

// fT[r * L2 + k2] = f[r + L1 * k2], fT[N] = f[N], per spectrum row
__global__ void k_spectrum_rows(const double *fc, int64_t nrow, int64_t stride, int64_t N, int L1, int L2,
                                double *fT)
{
    int64_t tot = nrow * (N + 1);
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        int64_t row = t / (N + 1), i = t % (N + 1);
        int64_t src = i == N ? N : (i % L2) * L1 + i / L2;
        fT[row * stride + i] = fc[row * stride + src];
    }
}

void fft_spectrum_rows(hipStream_t s, const FftPlan &p, const double *fc, int64_t nrow, int64_t stride, double *fT)
{
    int64_t tot = nrow * (p.N + 1);
    if (tot <= 0) return;
    k_spectrum_rows<<<(unsigned)std::min<int64_t>((tot + FFT_TPB - 1) / FFT_TPB, 65535 * 8), FFT_TPB, 0, s>>>(
        fc, nrow, stride, p.N, (int)p.L1, (int)p.L2, fT);
    MADAM_CHECK(hipGetLastError());
}

// Convolution of the batch's segments: the three fused passes when the plan
// allows them, otherwise the general eight-pass form (short FFTs).
static bool conv_fusable(const FftPlan &p)
{
    return p.L2 >= CONV_NCOL && p.L1 >= 2 && p.L1 * CONV_NCOL <= 8192 && p.L2 <= 2048;
}

template <typename T>
void fft_convolve_segments(hipStream_t s, const FftPlan &p, const Segment *dseg, int64_t nseg, const T *x, T *y,
                           const double *fc, const double *fc_rows, int64_t fc_stride, double *work, int64_t batch,
                           int64_t nseg_max, const Segment *hseg, const float *fc_rows32)
{
    (void)hseg;
    if (nseg > nseg_max)
        throw HipError(hipErrorInvalidValue, "fft_convolve_segments: more segments than the work buffer holds",
                       __FILE__, __LINE__);
    int64_t nof = 2 * p.N;
    int64_t npart = std::max(batch, nseg_max);
    int *ipsd = (int *)(work + batch * nof);                        // batch ints after the batch buffer
    double *partial = work + batch * nof + (batch + 1) / 2 + 1;    // npart * SEG_NB doubles
    int logL1 = ilog2(p.L1), logL2 = ilog2(p.L2);
    int ncb = (int)(p.L2 / CONV_NCOL);
    double *partial_out = partial + npart * SEG_NB;                 // npart * ncb doubles (every segment's)
    bool fused = conv_fusable(p);
    // the input means of every segment at once (one launch fills the GPU)
    if (fused)
        for (int64_t q0 = 0; q0 < nseg; q0 += 65535) {
            int64_t nq = std::min<int64_t>(nseg - q0, 65535);
            k_seg_partial<T><<<dim3(SEG_NB, (unsigned)nq), FFT_TPB, 0, s>>>(x, 0, 1, dseg, q0, 1.0, partial + q0 * SEG_NB);
        }
    for (int64_t s0 = 0; s0 < nseg; s0 += batch) {
        int64_t nb = nseg - s0 < batch ? nseg - s0 : batch;
        dim3 gp(SEG_NB, (unsigned)nb);
        if (fused) {
            bool f32 = std::is_same<T, float>::value && fc_rows32 != nullptr;
            dim3 gc((unsigned)ncb, (unsigned)nb), gr((unsigned)(p.L1 / 2), (unsigned)nb);
            // the points, then the stage twiddles (L / 2 of them)
            size_t col_elems = conv_lds_elems((int)(p.L1 * CONV_NCOL)) + std::max<int64_t>(p.L1 / 2, 1),
                   row_elems = conv_lds_elems((int)(2 * p.L2)) + std::max<int64_t>(p.L2 / 2, 1),
                   row_elems64 = conv_lds_elems((int)(2 * p.L2));   // (FP64 rows: no twiddles in LDS)
            if (f32) {
                // FP32 work: the same passes on float2, half the bytes and LDS
                float2 *w32 = (float2 *)work;
                k_conv_col_fwd<T, float2><<<gc, CONV_COL_TPB, sizeof(float2) * col_elems, s>>>(
                    x, dseg, s0, partial + s0 * SEG_NB, w32, p.N, (int)p.L1, logL1, (int)p.L2, p);
                k_conv_rows<float2><<<gr, FFT_TPB, sizeof(float2) * row_elems, s>>>(
                    w32, p.N, (int)p.L1, (int)p.L2, logL2, p, fc_rows32, fc_stride, dseg, s0);
                k_conv_col_inv<T, float2><<<gc, CONV_COL_TPB, sizeof(float2) * col_elems, s>>>(
                    w32, dseg, s0, y, partial_out + s0 * ncb, p.N, (int)p.L1, logL1, (int)p.L2, p);
            } else {
                k_conv_col_fwd<T, double2><<<gc, CONV_COL_TPB, sizeof(double2) * col_elems, s>>>(
                    x, dseg, s0, partial + s0 * SEG_NB, (double2 *)work, p.N, (int)p.L1, logL1, (int)p.L2, p);
                k_conv_rows<double2><<<gr, FFT_TPB, sizeof(double2) * row_elems64, s>>>(
                    (double2 *)work, p.N, (int)p.L1, (int)p.L2, logL2, p, fc_rows, fc_stride, dseg, s0);
                k_conv_col_inv<T, double2><<<gc, CONV_COL_TPB, sizeof(double2) * col_elems, s>>>(
                    (const double2 *)work, dseg, s0, y, partial_out + s0 * ncb, p.N, (int)p.L1, logL1, (int)p.L2, p);
            }
            continue;
        }
        k_seg_partial<T><<<gp, FFT_TPB, 0, s>>>(x, 0, 1, dseg, s0, 1.0, partial);
        unsigned gl = (unsigned)((nof + FFT_TPB * 4 - 1) / (FFT_TPB * 4));
        if (gl > 1024) gl = 1024;
        k_seg_load<T><<<dim3(gl, (unsigned)nb), FFT_TPB, 0, s>>>(x, dseg, s0, nof, work, partial);
        k_seg_ipsd<<<(unsigned)((nb + 255) / 256), 256, 0, s>>>(dseg, s0, nb, ipsd);
        fft_forward(s, p, (double2 *)work, nb);
        dim3 gs((unsigned)((p.N / 2 + 1 + FFT_TPB - 1) / FFT_TPB), (unsigned)nb);
        k_spectral<<<gs, FFT_TPB, 0, s>>>((double2 *)work, p.N, (int)p.L1, (int)p.L2, p, fc,
                                           fc_stride, ipsd);
        fft_inverse(s, p, (double2 *)work, nb);
        k_seg_partial<double><<<gp, FFT_TPB, 0, s>>>(work, nof, 0, dseg, s0, 1 / (double)nof, partial);
        k_seg_store<T><<<dim3(gl, (unsigned)nb), FFT_TPB, 0, s>>>(work, dseg, s0, nof, y, partial);
    }
    // the fused passes' outputs less their segment means, every segment in one
    // launch (the output y is read by nobody before this returns)
    if (fused) {
        unsigned gl = (unsigned)((nof / 2 + FFT_TPB * 16 - 1) / (FFT_TPB * 16));
        if (gl > 1024) gl = 1024;
        for (int64_t q0 = 0; q0 < nseg; q0 += 65535) {
            int64_t nq = std::min<int64_t>(nseg - q0, 65535);
            k_conv_sub_mean<T><<<dim3(gl, (unsigned)nq), FFT_TPB, 0, s>>>(y, dseg, q0, partial_out + q0 * ncb, ncb);
        }
    }
}

template void fft_convolve_segments<double>(hipStream_t, const FftPlan &, const Segment *, int64_t, const double *,
                                            double *, const double *, const double *, int64_t, double *, int64_t,
                                            int64_t, const Segment *, const float *);
template void fft_convolve_segments<float>(hipStream_t, const FftPlan &, const Segment *, int64_t, const float *,
                                           float *, const double *, const double *, int64_t, double *, int64_t,
                                           int64_t, const Segment *, const float *);

size_t fft_convolve_work_doubles(const FftPlan &p, int64_t batch, int64_t nseg_max)
{
    return (size_t)batch * 2 * p.N + (size_t)(batch + 1) / 2 + 1 + (size_t)std::max(batch, nseg_max) * SEG_NB +
           (size_t)std::max(batch, nseg_max) * (p.L2 / CONV_NCOL + 1) + 64;
}

__global__ void k_scale(double *x, double f, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) x[i] *= f;
}

void fft_c2r_rows(hipStream_t s, const FftPlan &p, const double *fc, int64_t nrows, double *out, double *work)
{
    (void)work;
    // out[row][nof] = c2r(fc[row][N+1]) / nof
    dim3 g((unsigned)((p.N + FFT_TPB - 1) / FFT_TPB), (unsigned)nrows);
    k_c2r_prep<<<g, FFT_TPB, 0, s>>>((double2 *)out, p.N, (int)p.L1, (int)p.L2, p, fc,
                                      p.N + 1);
    fft_inverse(s, p, (double2 *)out, nrows);
    int64_t n = nrows * 2 * p.N;
    k_scale<<<(unsigned)((n + 255) / 256 < 65535 ? (n + 255) / 256 : 65535), 256, 0, s>>>(out, 1 / (double)(2 * p.N), n);
}

// ---------------------------------------------------------------- noise filter

// interpolate_psd for one target frequency (src/noise_routines.f90:130-190),
// lf / lp are log(freq) / log(psd)
__device__ double interp_psd(const double *freq, const double *psd, const double *lf, const double *lp, int64_t n_in,
                             int64_t startbin, double tf)
{
    if (tf == 0) return 0.0;
    double flog = log(tf);
    // the reference walks j forward from startbin while lf[j+1] < flog, up to n_in - 2
    int64_t lo = startbin, hi = n_in - 2;
    if (lo > hi) lo = hi;
    // smallest j in [startbin, n_in-2] with lf[j+1] >= flog, else n_in-2
    int64_t a = lo, b = hi;
    while (a < b) {
        int64_t mid = (a + b) >> 1;
        if (lf[mid + 1] < flog) a = mid + 1;
        else b = mid;
    }
    int64_t j = a;
    if (psd[j] > 0 && psd[j + 1] > 0) {
        double r = (flog - lf[j]) / (lf[j + 1] - lf[j]);
        return exp(lp[j] * (1.0 - r) + lp[j + 1] * r);
    }
    double r = (tf - freq[j]) / (freq[j + 1] - freq[j]);
    return psd[j] * (1.0 - r) + psd[j + 1] * r;
}

__device__ double spec_value(const double *freq, const double *psd, const double *lf, const double *lp, int64_t n_in,
                             int64_t startbin, double f, double plateau, double margin, int *bad)
{
    double v = interp_psd(freq, psd, lf, lp, n_in, startbin, f);
    if (v > plateau * (1 + margin)) v = v - plateau;
    else v = v * margin;
    if (!(v > 0) && f != 0) atomicOr(bad, 1);
    return v;
}

// aspec / fcov for one PSD; work holds [freq | psd | lf | lp] (4 * n_in)
__global__ void k_build_filter(int64_t nof, double fsample, int nshort, double plateau, double margin, double sigma,
                               int64_t n_in, int64_t startbin, const double *tab, double *fcov, int *bad)
{
    const double *freq = tab, *psd = tab + n_in, *lf = tab + 2 * n_in, *lp = tab + 3 * n_in;
    double fa = fsample / nshort;
    int64_t nh = nof / 2 + 1;
    const double PI = 3.14159265358979323846;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nh; i += (int64_t)gridDim.x * blockDim.x) {
        double asp = 0;
        for (int k = 0; k <= 2; ++k) {
            double fi = k * fa + ((double)i * fa) / nof;
            double si = spec_value(freq, psd, lf, lp, n_in, startbin, fi, plateau, margin, bad);
            if (k == 0 && i == 0) si = 0.0;   // f(1) == 0: mean left unconstrained
            double x = PI * fi / fa;
            double gi;
            if (x < F32(1.e-30)) gi = 1;
            else { double sx = sin(x); gi = (sx * sx) / (x * x); }
            if (i == 0) {
                if (k == 0) asp = si;
                else asp = asp + 2 * si * gi;
            } else {
                int64_t m = nof - i;
                double fm = k * fa + ((double)m * fa) / nof;
                double sm = spec_value(freq, psd, lf, lp, n_in, startbin, fm, plateau, margin, bad);
                double xm = PI * fm / fa;
                double gm;
                if (xm < F32(1.e-30)) gm = 1;
                else { double sx = sin(xm); gm = (sx * sx) / (xm * xm); }
                asp = asp + si * gi + sm * gm;
            }
        }
        double v;
        if (fabs(asp) > sigma * F32(1e-30)) v = 1.0 / (fa * asp);
        else v = 0;
        if (i == 0) v = 0;
        fcov[i] = v;
    }
}

void launch_build_filter(hipStream_t s, int64_t nof, double fsample, int nshort, int radiometers, double sigma,
                         double plateau, int64_t n_in, int64_t startbin, const double *tab, double *fcov, int *bad)
{
    double margin = radiometers ? F32(1e-6) : F32(1e-3);
    int64_t nh = nof / 2 + 1;
    unsigned g = (unsigned)((nh + FFT_TPB - 1) / FFT_TPB);
    if (g > 65535) g = 65535;
    k_build_filter<<<g, FFT_TPB, 0, s>>>(nof, fsample, nshort, plateau, margin, sigma, n_in, startbin, tab, fcov,
                                         bad);
}

}  // namespace madam
