// Periodised orthogonal 2-D wavelet transforms (rwt/mdwt_r.c MDWT, rwt/midwt_r.c MIDWT)
// and the wavelet-l1 proximal step solved by ADMM (npg/admm.m) on gfx950.
//
// A transform level is two separable passes over the top-left block of the level
// (analysis: rows then columns; synthesis: columns then rows), each pass one launch
// that reads the block from one FP64 buffer and writes it to the other; regions outside
// the block keep the coarser levels' coefficients.  Each thread produces one pair of
// outputs (lowpass / highpass, or the even / odd synthesis samples) from the periodic
// neighbourhood of its input position; threads of a wave run along the contiguous
// index of the pass (columns for row passes, the column for column passes), so loads
// and stores coalesce.  The taps are kernel arguments.  All arithmetic is FP64 in the
// reference's summation order; images enter and leave as FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "ctprj_internal.h"

namespace ctprj {

constexpr int kWvMaxTaps = 32;
constexpr int kWvBlock = 256;

// Taps in the order of the sums (rwt/mdwt_r.c:85-88 fpsconv, rwt/midwt_r.c:87-92
// bpsconv): analysis lowpass lo[j] = h[j], highpass hi[j] = (-1)^j h[lh-1-j] applied to
// x[2p + j]; synthesis g0[k] = h[k], g1[k] = (-1)^k h[lh-1-k], even output from
// g*[lh-2-2j], odd from g*[lh-1-2j] on the coefficients at i + j - (lh/2 - 1).
struct WvTaps {
    int lh;
    double lo[kWvMaxTaps], hi[kWvMaxTaps];
    double e0[kWvMaxTaps / 2], e1[kWvMaxTaps / 2], o0[kWvMaxTaps / 2], o1[kWvMaxTaps / 2];
};

// Analysis along the contiguous index (rows pass): block bm x bn of images with row
// stride ld and image stride istride; thread (p, r, img) writes lowpass p and highpass
// bn/2 + p of row r.
__global__ __launch_bounds__(kWvBlock) void k_dwt_rows(WvTaps T, const double* __restrict__ src,
                                                       double* __restrict__ dst, int bm, int bn,
                                                       int ld, long long istride)
{
    const int half = bn >> 1;
    const long long t = (long long)blockIdx.x * kWvBlock + threadIdx.x;
    if (t >= (long long)half * bm) return;
    const int p = (int)(t % half), r = (int)(t / half);
    const long long base = blockIdx.y * istride + (long long)r * ld;
    double l = 0.0, h = 0.0;
    for (int j = 0; j < T.lh; ++j) {
        int k = 2 * p + j;
        while (k >= bn) k -= bn;
        const double v = src[base + k];
        l = __dadd_rn(l, __dmul_rn(v, T.lo[j]));
        h = __dadd_rn(h, __dmul_rn(v, T.hi[j]));
    }
    dst[base + p] = l;
    dst[base + half + p] = h;
}

// Analysis along columns: thread (c, p, img) writes rows p and bm/2 + p of column c.
__global__ __launch_bounds__(kWvBlock) void k_dwt_cols(WvTaps T, const double* __restrict__ src,
                                                       double* __restrict__ dst, int bm, int bn,
                                                       int ld, long long istride)
{
    const int half = bm >> 1;
    const long long t = (long long)blockIdx.x * kWvBlock + threadIdx.x;
    if (t >= (long long)half * bn) return;
    const int c = (int)(t % bn), p = (int)(t / bn);
    const long long base = blockIdx.y * istride + c;
    double l = 0.0, h = 0.0;
    for (int j = 0; j < T.lh; ++j) {
        int k = 2 * p + j;
        while (k >= bm) k -= bm;
        const double v = src[base + (long long)k * ld];
        l = __dadd_rn(l, __dmul_rn(v, T.lo[j]));
        h = __dadd_rn(h, __dmul_rn(v, T.hi[j]));
    }
    dst[base + (long long)p * ld] = l;
    dst[base + (long long)(half + p) * ld] = h;
}

// Synthesis along columns: thread (c, i, img) writes rows 2i and 2i+1 of column c from
// lowpass rows [0, M) and highpass rows [M, 2M), M = bm/2.
__global__ __launch_bounds__(kWvBlock) void k_idwt_cols(WvTaps T, const double* __restrict__ src,
                                                        double* __restrict__ dst, int bm, int bn,
                                                        int ld, long long istride)
{
    const int M = bm >> 1;
    const long long t = (long long)blockIdx.x * kWvBlock + threadIdx.x;
    if (t >= (long long)M * bn) return;
    const int c = (int)(t % bn), i = (int)(t / bn);
    const long long base = blockIdx.y * istride + c;
    const int nj = T.lh >> 1;
    double e = 0.0, o = 0.0;
    for (int j = 0; j < nj; ++j) {
        int k = i + j - (nj - 1);
        while (k < 0) k += M;
        const double lo = src[base + (long long)k * ld];
        const double hi = src[base + (long long)(M + k) * ld];
        e = __dadd_rn(__dadd_rn(e, __dmul_rn(lo, T.e0[j])), __dmul_rn(hi, T.e1[j]));
        o = __dadd_rn(__dadd_rn(o, __dmul_rn(lo, T.o0[j])), __dmul_rn(hi, T.o1[j]));
    }
    dst[base + (long long)(2 * i) * ld] = e;
    dst[base + (long long)(2 * i + 1) * ld] = o;
}

// Synthesis along rows: thread (i, r, img) writes columns 2i and 2i+1 of row r.
__global__ __launch_bounds__(kWvBlock) void k_idwt_rows(WvTaps T, const double* __restrict__ src,
                                                        double* __restrict__ dst, int bm, int bn,
                                                        int ld, long long istride)
{
    const int M = bn >> 1;
    const long long t = (long long)blockIdx.x * kWvBlock + threadIdx.x;
    if (t >= (long long)M * bm) return;
    const int i = (int)(t % M), r = (int)(t / M);
    const long long base = blockIdx.y * istride + (long long)r * ld;
    const int nj = T.lh >> 1;
    double e = 0.0, o = 0.0;
    for (int j = 0; j < nj; ++j) {
        int k = i + j - (nj - 1);
        while (k < 0) k += M;
        const double lo = src[base + k];
        const double hi = src[base + M + k];
        e = __dadd_rn(__dadd_rn(e, __dmul_rn(lo, T.e0[j])), __dmul_rn(hi, T.e1[j]));
        o = __dadd_rn(__dadd_rn(o, __dmul_rn(lo, T.o0[j])), __dmul_rn(hi, T.o1[j]));
    }
    dst[base + 2 * i] = e;
    dst[base + 2 * i + 1] = o;
}

// Support masks (utils/maskFunc.m with tomography/ConfigCT.m:85-86): an entry outside a
// mask (mask value 0) is dropped, i.e. read as zero; mask == NULL keeps every entry.
__device__ __forceinline__ double masked(const float* mask, long long i, double v)
{
    return (!mask || mask[i] != 0.0f) ? v : 0.0;
}

__global__ void k_f32_to_f64(const float* __restrict__ in, double* __restrict__ out, long long n,
                             double scale, const float* __restrict__ mask)
{
    const long long i = (long long)blockIdx.x * kWvBlock + threadIdx.x;
    if (i < n) out[i] = masked(mask, i, (double)in[i] * scale);
}

__global__ void k_f64_to_f32(const double* __restrict__ in, float* __restrict__ out, long long n,
                             const float* __restrict__ mask)
{
    const long long i = (long long)blockIdx.x * kWvBlock + threadIdx.x;
    if (i < n) out[i] = (float)masked(mask, i, in[i]);
}

// Per-group FP64 partial sums of up to two values, summed in a fixed LDS tree.
__device__ __forceinline__ void block_sum2(double a, double b, double* part)
{
    __shared__ double sa[kWvBlock], sb[kWvBlock];
    const int t = threadIdx.x;
    sa[t] = a;
    sb[t] = b;
    __syncthreads();
    for (int off = kWvBlock / 2; off > 0; off >>= 1) {
        if (t < off) {
            sa[t] += sa[t + off];
            sb[t] += sb[t + off];
        }
        __syncthreads();
    }
    if (t == 0) {
        part[2 * blockIdx.x] = sa[0];
        part[2 * blockIdx.x + 1] = sb[0];
    }
}

// sum |x| and sum x^2 of an FP64 array
__global__ __launch_bounds__(kWvBlock) void k_norms(const double* __restrict__ x, long long n,
                                                    const float* __restrict__ mask,
                                                    double* __restrict__ part)
{
    const long long i = (long long)blockIdx.x * kWvBlock + threadIdx.x;
    double a = 0.0, b = 0.0;
    if (i < n) {
        const double v = masked(mask, i, x[i]);
        a = fabs(v);
        b = v * v;
    }
    block_sum2(a, b, part);
}

// v = s + nu   (npg/admm.m:26, the argument of Psi)
__global__ void k_admm_sum(const double* __restrict__ s, const double* __restrict__ nu,
                           double* __restrict__ v, long long n)
{
    const long long i = (long long)blockIdx.x * kWvBlock + threadIdx.x;
    if (i < n) v[i] = s[i] + nu[i];
}

// alpha = max((a + rho Psi(s + nu)) / (1 + rho), 0)   (npg/admm.m:26)
__global__ void k_admm_alpha(const double* __restrict__ a, const double* __restrict__ psi,
                             double rho, double* __restrict__ alpha, long long n,
                             const float* __restrict__ mask)
{
    const long long i = (long long)blockIdx.x * kWvBlock + threadIdx.x;
    if (i < n)
        alpha[i] = fmax(__dadd_rn(a[i], __dmul_rn(rho, masked(mask, i, psi[i]))) / (1.0 + rho),
                        0.0);
}

// s = softThresh(Psit(alpha) - nu, u/rho); nu += s - Psit(alpha); partials of
// ||s - preS||^2 and ||s - Psit(alpha)||^2; preS = s   (npg/admm.m:27-33)
__global__ __launch_bounds__(kWvBlock) void k_admm_shrink(const double* __restrict__ pa,
                                                          double* __restrict__ s,
                                                          double* __restrict__ nu,
                                                          double* __restrict__ pre, double thr,
                                                          long long n,
                                                          const float* __restrict__ cmask,
                                                          double* __restrict__ part)
{
    const long long i = (long long)blockIdx.x * kWvBlock + threadIdx.x;
    double ds = 0.0, rs = 0.0;
    if (i < n) {
        const double p = masked(cmask, i, pa[i]);
        const double x = p - nu[i];
        const double t = fabs(thr);
        const double si = x > t ? x - t : (x < -t ? x + t : 0.0);
        s[i] = si;
        nu[i] = nu[i] + si - p;
        const double d = si - pre[i];
        const double r = si - p;
        ds = d * d;
        rs = r * r;
        pre[i] = si;
    }
    block_sum2(ds, rs, part);
}

// x = mask ? x : 0 in place
__global__ void k_f64_to_f64_masked(double* __restrict__ x, long long n,
                                    const float* __restrict__ mask)
{
    const long long i = (long long)blockIdx.x * kWvBlock + threadIdx.x;
    if (i < n) x[i] = masked(mask, i, x[i]);
}

// rho changes (npg/admm.m:43-49): nu *= f
__global__ void k_scale64(double* __restrict__ x, double f, long long n)
{
    const long long i = (long long)blockIdx.x * kWvBlock + threadIdx.x;
    if (i < n) x[i] *= f;
}

// out = max((a + rho Psi(s + nu)) / (1 + rho), 0) * scale   (npg/admm.m:52), FP32
__global__ void k_admm_final(const double* __restrict__ a, const double* __restrict__ psi,
                             double rho, double scale, float* __restrict__ out, long long n,
                             const float* __restrict__ mask)
{
    const long long i = (long long)blockIdx.x * kWvBlock + threadIdx.x;
    if (i < n)
        out[i] = (float)(fmax(__dadd_rn(a[i], __dmul_rn(rho, masked(mask, i, psi[i]))) /
                                  (1.0 + rho),
                              0.0) *
                         scale);
}

}  // namespace ctprj

using namespace ctprj;

struct ctprj_wavelet {
    int rows = 0, cols = 0, batch = 0, L = 0;
    long long n = 0;  // rows * cols * batch
    WvTaps T{};
    double *d_x = nullptr, *d_y = nullptr;                  // transform ping-pong
    double *d_a = nullptr, *d_s = nullptr, *d_nu = nullptr, *d_pre = nullptr, *d_pa = nullptr;
    double* d_part = nullptr;
    float* d_mask = nullptr;   // image support (NULL: every pixel)
    float* d_cmask = nullptr;  // coefficient support (NULL: every coefficient)
    long long blocks = 0;
    std::vector<double> part;
    ~ctprj_wavelet()
    {
        for (void* p : {(void*)d_x, (void*)d_y, (void*)d_a, (void*)d_s, (void*)d_nu,
                        (void*)d_pre, (void*)d_pa, (void*)d_part, (void*)d_mask, (void*)d_cmask})
            if (p) (void)hipFree(p);
    }
};

namespace {

unsigned grid1(long long n) { return (unsigned)((n + kWvBlock - 1) / kWvBlock); }

int last(const char* what)
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(e, what);
}

// forward transform of w->d_x (FP64, batch images) into w->d_x, using d_y as scratch
int dwt64(ctprj_wavelet* w, double* x, double* tmp, hipStream_t st)
{
    const long long is = (long long)w->rows * w->cols;
    for (int lev = 0; lev < w->L; ++lev) {
        const int bm = w->rows >> lev, bn = w->cols >> lev;
        dim3 g(grid1((long long)(bn / 2) * bm), (unsigned)w->batch);
        hipLaunchKernelGGL(k_dwt_rows, g, dim3(kWvBlock), 0, st, w->T, x, tmp, bm, bn, w->cols, is);
        dim3 g2(grid1((long long)(bm / 2) * bn), (unsigned)w->batch);
        hipLaunchKernelGGL(k_dwt_cols, g2, dim3(kWvBlock), 0, st, w->T, tmp, x, bm, bn, w->cols, is);
    }
    return last("dwt launch");
}

int idwt64(ctprj_wavelet* w, double* x, double* tmp, hipStream_t st)
{
    const long long is = (long long)w->rows * w->cols;
    for (int lev = w->L - 1; lev >= 0; --lev) {
        const int bm = w->rows >> lev, bn = w->cols >> lev;
        dim3 g(grid1((long long)(bm / 2) * bn), (unsigned)w->batch);
        hipLaunchKernelGGL(k_idwt_cols, g, dim3(kWvBlock), 0, st, w->T, x, tmp, bm, bn, w->cols, is);
        dim3 g2(grid1((long long)(bn / 2) * bm), (unsigned)w->batch);
        hipLaunchKernelGGL(k_idwt_rows, g2, dim3(kWvBlock), 0, st, w->T, tmp, x, bm, bn, w->cols, is);
    }
    return last("idwt launch");
}

// the two partial sums of the last reduction launch, summed on the host in order
int fetch2(ctprj_wavelet* w, hipStream_t st, double* a, double* b)
{
    hipError_t e = hipMemcpyAsync(w->part.data(), w->d_part, sizeof(double) * 2 * w->blocks,
                                  hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "wavelet reduction");
    double sa = 0.0, sb = 0.0;
    for (long long i = 0; i < w->blocks; ++i) {
        sa += w->part[2 * i];
        sb += w->part[2 * i + 1];
    }
    *a = sa;
    *b = sb;
    return 0;
}

}  // namespace

extern "C" int ctprj_wavelet_create(int rows, int cols, int batch, const double* h, int lh, int L,
                                    ctprj_wavelet** out)
{
    if (!h || !out) return fail(CTPRJ_EINVAL, "wavelet_create: NULL argument");
    *out = nullptr;
    if (lh < 2 || lh % 2 || lh > kWvMaxTaps)
        return fail(CTPRJ_EINVAL, "wavelet filter length %d must be even and in [2, %d]", lh,
                    kWvMaxTaps);
    if (L < 1 || L > 30 || rows < 2 || cols < 2 || batch < 1)
        return fail(CTPRJ_EINVAL, "wavelet_create: bad sizes (%d x %d x %d, L=%d)", rows, cols,
                    batch, L);
    if (rows % (1 << L) || cols % (1 << L))
        return fail(CTPRJ_EINVAL, "wavelet_create: %d x %d not divisible by 2^%d", rows, cols, L);
    const long long n = (long long)rows * cols * batch;
    if (n >= (1LL << 31) || batch > 65535)
        return fail(CTPRJ_EINVAL, "wavelet_create: %lld elements is too many", n);
    for (int i = 0; i < lh; ++i)
        if (!std::isfinite(h[i])) return fail(CTPRJ_EINVAL, "wavelet filter not finite");
    ctprj_wavelet* w = new (std::nothrow) ctprj_wavelet();
    if (!w) return fail(CTPRJ_ENOMEM, "host allocation");
    w->rows = rows;
    w->cols = cols;
    w->batch = batch;
    w->L = L;
    w->n = n;
    w->T.lh = lh;
    for (int j = 0; j < lh; ++j) {
        w->T.lo[j] = h[j];
        w->T.hi[j] = (j & 1 ? -1.0 : 1.0) * h[lh - 1 - j];
    }
    for (int j = 0; j < lh / 2; ++j) {
        const int ke = lh - 2 - 2 * j, ko = lh - 1 - 2 * j;
        w->T.e0[j] = h[ke];
        w->T.e1[j] = (ke & 1 ? -1.0 : 1.0) * h[lh - 1 - ke];
        w->T.o0[j] = h[ko];
        w->T.o1[j] = (ko & 1 ? -1.0 : 1.0) * h[lh - 1 - ko];
    }
    w->blocks = (n + kWvBlock - 1) / kWvBlock;
    w->part.assign((size_t)2 * w->blocks, 0.0);
    hipError_t e = hipSuccess;
    for (double** p : {&w->d_x, &w->d_y, &w->d_a, &w->d_s, &w->d_nu, &w->d_pre, &w->d_pa})
        if (e == hipSuccess) e = hipMalloc((void**)p, sizeof(double) * (size_t)n);
    if (e == hipSuccess) e = hipMalloc((void**)&w->d_part, sizeof(double) * 2 * (size_t)w->blocks);
    if (e != hipSuccess) {
        delete w;
        return fail(CTPRJ_ENOMEM, "wavelet workspace: %s", hipGetErrorString(e));
    }
    *out = w;
    return 0;
}

long long ctprj::wavelet_elems(const ctprj_wavelet* w) { return w ? w->n : 0; }

extern "C" int ctprj_wavelet_set_masks(ctprj_wavelet* w, const float* d_mask,
                                       const float* d_cmask)
{
    if (!w) return fail(CTPRJ_EINVAL, "wavelet_set_masks: NULL context");
    const size_t bytes = sizeof(float) * (size_t)w->n;
    float** dst[2] = {&w->d_mask, &w->d_cmask};
    const float* src[2] = {d_mask, d_cmask};
    for (int k = 0; k < 2; ++k) {
        if (*dst[k]) (void)hipFree(*dst[k]);
        *dst[k] = nullptr;
        if (!src[k]) continue;
        hipError_t e = hipMalloc((void**)dst[k], bytes);
        if (e == hipSuccess) e = hipMemcpy(*dst[k], src[k], bytes, hipMemcpyDeviceToDevice);
        if (e != hipSuccess) return hip_fail(e, "wavelet masks");
    }
    return 0;
}

extern "C" int ctprj_wavelet_destroy(ctprj_wavelet* w)
{
    delete w;
    return 0;
}

extern "C" int ctprj_dwt(ctprj_wavelet* w, const float* d_x, float* d_y, void* stream)
{
    if (!w || !d_x || !d_y) return fail(CTPRJ_EINVAL, "dwt: NULL argument");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_f32_to_f64, dim3(grid1(w->n)), dim3(kWvBlock), 0, st, d_x, w->d_x, w->n, 1.0,
                       w->d_mask);
    int rc = dwt64(w, w->d_x, w->d_y, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_f64_to_f32, dim3(grid1(w->n)), dim3(kWvBlock), 0, st, w->d_x, d_y, w->n,
                       w->d_cmask);
    return last("dwt output");
}

extern "C" int ctprj_idwt(ctprj_wavelet* w, const float* d_y, float* d_x, void* stream)
{
    if (!w || !d_x || !d_y) return fail(CTPRJ_EINVAL, "idwt: NULL argument");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_f32_to_f64, dim3(grid1(w->n)), dim3(kWvBlock), 0, st, d_y, w->d_x, w->n, 1.0,
                       w->d_cmask);
    int rc = idwt64(w, w->d_x, w->d_y, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_f64_to_f32, dim3(grid1(w->n)), dim3(kWvBlock), 0, st, w->d_x, d_x, w->n,
                       w->d_mask);
    return last("idwt output");
}

extern "C" int ctprj_wavelet_l1(ctprj_wavelet* w, const float* d_x, double* l1, void* stream)
{
    if (!w || !d_x || !l1) return fail(CTPRJ_EINVAL, "wavelet_l1: NULL argument");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_f32_to_f64, dim3(grid1(w->n)), dim3(kWvBlock), 0, st, d_x, w->d_x, w->n, 1.0,
                       w->d_mask);
    int rc = dwt64(w, w->d_x, w->d_y, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_norms, dim3(grid1(w->n)), dim3(kWvBlock), 0, st, w->d_x, w->n, w->d_cmask,
                       w->d_part);
    if ((rc = last("wavelet l1 launch"))) return rc;
    double sq;
    return fetch2(w, st, l1, &sq);
}

extern "C" int ctprj_prox_wavelet_admm(ctprj_wavelet* w, const float* d_a, double u,
                                       double relTol, int maxItr, const float* d_init,
                                       float* d_out, int* iterations, void* stream)
{
    if (!w || !d_a || !d_out) return fail(CTPRJ_EINVAL, "prox_wavelet_admm: NULL argument");
    if (!std::isfinite(u) || !std::isfinite(relTol) || maxItr < 0)
        return fail(CTPRJ_EINVAL, "prox_wavelet_admm: bad u, tolerance or maxItr");
    hipStream_t st = (hipStream_t)stream;
    const long long n = w->n;
    const unsigned g = grid1(n);
    relTol = std::min(1e-3, relTol);  // npg/admm.m:14
    int rc;
    // scale = ||a||  (npg/admm.m:17-18)
    hipLaunchKernelGGL(k_f32_to_f64, dim3(g), dim3(kWvBlock), 0, st, d_a, w->d_a, n, 1.0, w->d_mask);
    hipLaunchKernelGGL(k_norms, dim3(g), dim3(kWvBlock), 0, st, w->d_a, n, nullptr, w->d_part);
    if ((rc = last("admm scale launch"))) return rc;
    double l1, sq;
    if ((rc = fetch2(w, st, &l1, &sq))) return rc;
    const double scale = std::sqrt(sq);
    if (iterations) *iterations = 0;
    if (scale == 0.0) {
        hipError_t e = hipMemsetAsync(d_out, 0, sizeof(float) * (size_t)n, st);
        return e == hipSuccess ? 0 : hip_fail(e, "admm zero output");
    }
    const double inv = 1.0 / scale;
    u = u / scale;
    hipLaunchKernelGGL(k_scale64, dim3(g), dim3(kWvBlock), 0, st, w->d_a, inv, n);
    // preS = s = Psit(init / scale), nu = 0
    hipLaunchKernelGGL(k_f32_to_f64, dim3(g), dim3(kWvBlock), 0, st, d_init ? d_init : d_a,
                       w->d_pre, n, inv, w->d_mask);
    if ((rc = dwt64(w, w->d_pre, w->d_y, st))) return rc;
    if (w->d_cmask)
        hipLaunchKernelGGL(k_f64_to_f64_masked, dim3(g), dim3(kWvBlock), 0, st, w->d_pre, n,
                           w->d_cmask);
    hipError_t e = hipMemcpyAsync(w->d_s, w->d_pre, sizeof(double) * (size_t)n,
                                  hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(w->d_nu, 0, sizeof(double) * (size_t)n, st);
    if (e != hipSuccess) return hip_fail(e, "admm init");
    double rho = 1.0;
    int cnt = 0, itr = 0;
    for (;;) {
        ++itr;
        ++cnt;
        // alpha = max((a + rho Psi(s + nu)) / (1 + rho), 0) in d_x
        hipLaunchKernelGGL(k_admm_sum, dim3(g), dim3(kWvBlock), 0, st, w->d_s, w->d_nu, w->d_x, n);
        if ((rc = idwt64(w, w->d_x, w->d_y, st))) return rc;
        hipLaunchKernelGGL(k_admm_alpha, dim3(g), dim3(kWvBlock), 0, st, w->d_a, w->d_x, rho,
                           w->d_pa, n, w->d_mask);
        // Psit(alpha) in d_pa
        if ((rc = dwt64(w, w->d_pa, w->d_y, st))) return rc;
        hipLaunchKernelGGL(k_admm_shrink, dim3(g), dim3(kWvBlock), 0, st, w->d_pa, w->d_s, w->d_nu,
                           w->d_pre, u / rho, n, w->d_cmask, w->d_part);
        if ((rc = last("admm launch"))) return rc;
        double ds2, rs2;
        if ((rc = fetch2(w, st, &ds2, &rs2))) return rc;
        const double difS = std::sqrt(ds2), residual = std::sqrt(rs2);
        if (itr > maxItr) break;
        if (difS <= relTol && residual <= relTol) break;
        if (cnt > 20) {
            if (difS > 10 * residual) {
                rho = rho / 2;
                hipLaunchKernelGGL(k_scale64, dim3(g), dim3(kWvBlock), 0, st, w->d_nu, 2.0, n);
                cnt = 0;
            } else if (difS < residual / 10) {
                rho = rho * 2;
                hipLaunchKernelGGL(k_scale64, dim3(g), dim3(kWvBlock), 0, st, w->d_nu, 0.5, n);
                cnt = 0;
            }
        }
    }
    hipLaunchKernelGGL(k_admm_sum, dim3(g), dim3(kWvBlock), 0, st, w->d_s, w->d_nu, w->d_x, n);
    if ((rc = idwt64(w, w->d_x, w->d_y, st))) return rc;
    hipLaunchKernelGGL(k_admm_final, dim3(g), dim3(kWvBlock), 0, st, w->d_a, w->d_x, rho, scale,
                       d_out, n, w->d_mask);
    if (iterations) *iterations = itr;
    return last("admm output");
}
