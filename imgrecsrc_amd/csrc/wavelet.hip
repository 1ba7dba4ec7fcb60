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
// and stores coalesce.  Once a level's block fits in LDS, one work-group per image runs
// all coarser levels there.  The taps are kernel arguments.  All arithmetic is FP64 in
// the reference's summation order; images enter and leave as FP32.  The ADMM loop keeps
// its state (rho, counters, the stop flag) on the device and is read by the host once
// per chunk of iterations.
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
                                                       int ld, long long istride,
                                                       const int* __restrict__ stop)
{
    if (stop && *stop) return;  // uniform: a converged prox loop
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
                                                       int ld, long long istride,
                                                       const int* __restrict__ stop)
{
    if (stop && *stop) return;  // uniform: a converged prox loop
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
                                                        int ld, long long istride,
                                                        const int* __restrict__ stop)
{
    if (stop && *stop) return;  // uniform: a converged prox loop
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
                                                        int ld, long long istride,
                                                        const int* __restrict__ stop)
{
    if (stop && *stop) return;  // uniform: a converged prox loop
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

// The coarse levels of a transform in one work-group per image: once a level's block
// holds at most kWvSmall coefficients it stays in LDS (A; B the pass buffer) for every
// coarser level, with the passes, index arithmetic and FP64 sums of the four level
// kernels above, so the results are bit-identical to launching them level by level.
// The block is read from src and written to dst (which may be src).
constexpr int kWvSmall = 4096;
constexpr int kWvSmallThreads = 1024;

__device__ __forceinline__ void small_load(double* A, const double* img, int bm0, int bn0, int ld)
{
    for (int i = threadIdx.x; i < bm0 * bn0; i += kWvSmallThreads)
        A[i] = img[(long long)(i / bn0) * ld + i % bn0];
    __syncthreads();
}

__device__ __forceinline__ void small_store(const double* A, double* img, int bm0, int bn0, int ld)
{
    for (int i = threadIdx.x; i < bm0 * bn0; i += kWvSmallThreads)
        img[(long long)(i / bn0) * ld + i % bn0] = A[i];
}

// analysis of `levels` levels starting with the bm0 x bn0 block (row stride ld)
__global__ __launch_bounds__(kWvSmallThreads) void k_dwt_small(WvTaps T, const double* src,
                                                               double* dst, int bm0, int bn0,
                                                               int levels, int ld, long long istride,
                                                               const int* __restrict__ stop)
{
    if (stop && *stop) return;
    __shared__ double A[kWvSmall], B[kWvSmall];
    const long long ib = (long long)blockIdx.x * istride;
    small_load(A, src + ib, bm0, bn0, ld);
    for (int lev = 0; lev < levels; ++lev) {
        const int bm = bm0 >> lev, bn = bn0 >> lev;
        int half = bn >> 1;
        for (int t = threadIdx.x; t < half * bm; t += kWvSmallThreads) {  // rows: A -> B
            const int p = t % half, r = t / half;
            double l = 0.0, h = 0.0;
            for (int j = 0; j < T.lh; ++j) {
                int k = 2 * p + j;
                while (k >= bn) k -= bn;
                const double v = A[r * bn0 + k];
                l = __dadd_rn(l, __dmul_rn(v, T.lo[j]));
                h = __dadd_rn(h, __dmul_rn(v, T.hi[j]));
            }
            B[r * bn0 + p] = l;
            B[r * bn0 + half + p] = h;
        }
        __syncthreads();
        half = bm >> 1;
        for (int t = threadIdx.x; t < half * bn; t += kWvSmallThreads) {  // columns: B -> A
            const int c = t % bn, p = t / bn;
            double l = 0.0, h = 0.0;
            for (int j = 0; j < T.lh; ++j) {
                int k = 2 * p + j;
                while (k >= bm) k -= bm;
                const double v = B[k * bn0 + c];
                l = __dadd_rn(l, __dmul_rn(v, T.lo[j]));
                h = __dadd_rn(h, __dmul_rn(v, T.hi[j]));
            }
            A[p * bn0 + c] = l;
            A[(half + p) * bn0 + c] = h;
        }
        __syncthreads();
    }
    small_store(A, dst + ib, bm0, bn0, ld);
}

// synthesis of the `levels` coarsest levels, ending with the bm0 x bn0 block
__global__ __launch_bounds__(kWvSmallThreads) void k_idwt_small(WvTaps T, const double* src,
                                                                double* dst, int bm0, int bn0,
                                                                int levels, int ld, long long istride,
                                                                const int* __restrict__ stop)
{
    if (stop && *stop) return;
    __shared__ double A[kWvSmall], B[kWvSmall];
    const long long ib = (long long)blockIdx.x * istride;
    small_load(A, src + ib, bm0, bn0, ld);
    const int nj = T.lh >> 1;
    for (int lev = levels - 1; lev >= 0; --lev) {
        const int bm = bm0 >> lev, bn = bn0 >> lev;
        int M = bm >> 1;
        for (int t = threadIdx.x; t < M * bn; t += kWvSmallThreads) {  // columns: A -> B
            const int c = t % bn, i = t / bn;
            double e = 0.0, o = 0.0;
            for (int j = 0; j < nj; ++j) {
                int k = i + j - (nj - 1);
                while (k < 0) k += M;
                const double lo = A[k * bn0 + c];
                const double hi = A[(M + k) * bn0 + c];
                e = __dadd_rn(__dadd_rn(e, __dmul_rn(lo, T.e0[j])), __dmul_rn(hi, T.e1[j]));
                o = __dadd_rn(__dadd_rn(o, __dmul_rn(lo, T.o0[j])), __dmul_rn(hi, T.o1[j]));
            }
            B[(2 * i) * bn0 + c] = e;
            B[(2 * i + 1) * bn0 + c] = o;
        }
        __syncthreads();
        M = bn >> 1;
        for (int t = threadIdx.x; t < M * bm; t += kWvSmallThreads) {  // rows: B -> A
            const int i = t % M, r = t / M;
            double e = 0.0, o = 0.0;
            for (int j = 0; j < nj; ++j) {
                int k = i + j - (nj - 1);
                while (k < 0) k += M;
                const double lo = B[r * bn0 + k];
                const double hi = B[r * bn0 + M + k];
                e = __dadd_rn(__dadd_rn(e, __dmul_rn(lo, T.e0[j])), __dmul_rn(hi, T.e1[j]));
                o = __dadd_rn(__dadd_rn(o, __dmul_rn(lo, T.o0[j])), __dmul_rn(hi, T.o1[j]));
            }
            A[r * bn0 + 2 * i] = e;
            A[r * bn0 + 2 * i + 1] = o;
        }
        __syncthreads();
    }
    small_store(A, dst + ib, bm0, bn0, ld);
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

// ADMM loop state on the device (npg/admm.m:14-50): the loop's kernels read rho and the
// stop flag from here and k_admm_ctrl updates them, so the host only looks at the state
// once per chunk of iterations.  Kernels queued after the loop has stopped return at
// once.
struct AdmmCtl {
    double rho;     // penalty parameter
    double f;       // factor nu takes at the next iteration's start (rho change), else 1
    double scale;   // ||a||
    double inv;     // 1 / scale
    double u;       // the threshold weight divided by scale
    double relTol;  // stopping tolerance of this call
    int itr, cnt, done, zero;
    int maxItr;     // iteration bound of this call
    int pad;
};

// sum over the blocks of one of the two per-block partials (k = 0, 1) by one work-group
// of kWvBlock threads, in a fixed order
__device__ double block_total(const double* __restrict__ part, long long nblk, int k)
{
    __shared__ double sh[kWvBlock];
    const int t = threadIdx.x;
    double a = 0.0;
    for (long long b = t; b < nblk; b += kWvBlock) a += part[2 * b + k];
    sh[t] = a;
    __syncthreads();
    for (int off = kWvBlock / 2; off > 0; off >>= 1) {
        if (t < off) sh[t] += sh[t + off];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// scale = ||a|| from k_norms' partials; a zero input stops the loop before it starts
// (the output is then zero)   (npg/admm.m:17-18)
__global__ __launch_bounds__(kWvBlock) void k_admm_init(const double* __restrict__ part,
                                                        long long nblk, double u, double relTol,
                                                        int maxItr, AdmmCtl* __restrict__ c)
{
    const double sq = block_total(part, nblk, 1);
    if (threadIdx.x == 0) {
        c->relTol = relTol;
        c->maxItr = maxItr;
        const double scale = sqrt(sq);
        c->scale = scale;
        c->inv = 1.0 / scale;
        c->u = u / scale;
        c->rho = 1.0;
        c->f = 1.0;
        c->itr = 0;
        c->cnt = 0;
        c->zero = scale == 0.0;
        c->done = scale == 0.0;
    }
}

// x *= 1 / scale
__global__ void k_admm_scale_in(double* __restrict__ x, long long n, const AdmmCtl* __restrict__ c)
{
    if (c->zero) return;
    const double inv = c->inv;
    const long long i = (long long)blockIdx.x * kWvBlock + threadIdx.x;
    if (i < n) x[i] *= inv;
}

// out = masked(in / scale) in FP64 (the starting point)
__global__ void k_admm_load(const float* __restrict__ in, double* __restrict__ out, long long n,
                            const AdmmCtl* __restrict__ c, const float* __restrict__ mask)
{
    if (c->zero) return;
    const double inv = c->inv;
    const long long i = (long long)blockIdx.x * kWvBlock + threadIdx.x;
    if (i < n) out[i] = masked(mask, i, (double)in[i] * inv);
}

// s = preS, nu = 0
__global__ void k_admm_start(const double* __restrict__ pre, double* __restrict__ s,
                             double* __restrict__ nu, long long n)
{
    const long long i = (long long)blockIdx.x * kWvBlock + threadIdx.x;
    if (i < n) {
        s[i] = pre[i];
        nu[i] = 0.0;
    }
}

// v = s + nu   (npg/admm.m:26, the argument of Psi), after nu *= f of a rho change
// (npg/admm.m:43-49)
__global__ void k_admm_sum(const double* __restrict__ s, double* __restrict__ nu,
                           double* __restrict__ v, long long n, const AdmmCtl* __restrict__ c,
                           const int* __restrict__ stop)
{
    if (stop && *stop) return;
    const double f = c->f;
    const long long i = (long long)blockIdx.x * kWvBlock + threadIdx.x;
    if (i >= n) return;
    double w = nu[i];
    if (f != 1.0) {
        w *= f;
        nu[i] = w;
    }
    v[i] = s[i] + w;
}

// alpha = max((a + rho Psi(s + nu)) / (1 + rho), 0)   (npg/admm.m:26)
__global__ void k_admm_alpha(const double* __restrict__ a, const double* __restrict__ psi,
                             double* __restrict__ alpha, long long n,
                             const float* __restrict__ mask, const AdmmCtl* __restrict__ c,
                             const int* __restrict__ stop)
{
    if (stop && *stop) return;
    const double rho = c->rho;
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
                                                          double* __restrict__ pre, long long n,
                                                          const float* __restrict__ cmask,
                                                          double* __restrict__ part,
                                                          const AdmmCtl* __restrict__ c,
                                                          const int* __restrict__ stop)
{
    if (stop && *stop) return;
    const double t = fabs(c->u / c->rho);
    const long long i = (long long)blockIdx.x * kWvBlock + threadIdx.x;
    double ds = 0.0, rs = 0.0;
    if (i < n) {
        const double p = masked(cmask, i, pa[i]);
        const double x = p - nu[i];
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

// End of an iteration (npg/admm.m:35-50): difS = ||s - preS||, residual = ||s - Psit(alpha)||;
// stop after maxItr + 1 iterations or when both are within relTol; after more than 20
// iterations without a change, rho halves (nu doubles) when difS > 10 residual and
// doubles (nu halves) when difS < residual / 10.
__global__ __launch_bounds__(kWvBlock) void k_admm_ctrl(const double* __restrict__ part,
                                                        long long nblk, AdmmCtl* __restrict__ c)
{
    if (c->done) return;
    const double relTol = c->relTol;
    const int maxItr = c->maxItr;
    const double ds2 = block_total(part, nblk, 0);
    const double rs2 = block_total(part, nblk, 1);
    if (threadIdx.x != 0) return;
    const int itr = c->itr + 1;
    int cnt = c->cnt + 1;
    double rho = c->rho, f = 1.0;
    const double difS = sqrt(ds2), residual = sqrt(rs2);
    const bool done = itr > maxItr || (difS <= relTol && residual <= relTol);
    if (!done && cnt > 20) {
        if (difS > 10 * residual) {
            rho = rho / 2;
            f = 2.0;
            cnt = 0;
        } else if (difS < residual / 10) {
            rho = rho * 2;
            f = 0.5;
            cnt = 0;
        }
    }
    c->itr = itr;
    c->cnt = cnt;
    c->rho = rho;
    c->f = f;
    c->done = done;
}

// One transform level in one launch (images up to kWvLevelMax columns).  Out of place:
// the level reads its input block from one buffer and writes its three detail quadrants
// to the transform's output and the approximation (the next level's input) to a scratch
// buffer, so no work-group reads what another writes.  Sums, taps and index arithmetic
// are those of the two-pass kernels above.
constexpr int kWvLevelCols = 8;                       // columns per thread
constexpr int kWvLevelMax = kWvBlock * kWvLevelCols;  // 2048

// Analysis: work-group (p, img) forms the rows pass of input rows 2p + j (j < lh,
// periodic) one at a time in LDS and accumulates the columns pass of output rows p
// (lowpass) and bm/2 + p (highpass) in registers.  The approximation quadrant (rows
// < bm/2, columns < bn/2) goes to `ll`, the rest to dst.
__global__ __launch_bounds__(kWvBlock) void k_dwt_level(WvTaps T, const double* __restrict__ src,
                                                        double* __restrict__ dst,
                                                        double* __restrict__ ll, int bm, int bn,
                                                        int ld, long long istride,
                                                        const int* __restrict__ stop)
{
    if (stop && *stop) return;
    __shared__ double R[2][kWvLevelMax];
    const int p = blockIdx.x, tid = threadIdx.x, half = bn >> 1;
    const long long ib = (long long)blockIdx.y * istride;
    double accL[kWvLevelCols], accH[kWvLevelCols];
#pragma unroll
    for (int u = 0; u < kWvLevelCols; ++u) accL[u] = accH[u] = 0.0;
    for (int j = 0; j < T.lh; ++j) {
        int r = 2 * p + j;
        while (r >= bm) r -= bm;
        const double* row = src + ib + (long long)r * ld;
        double* Rj = R[j & 1];  // rewritten at j + 2, after every thread passed j + 1's barrier
        for (int q = tid; q < half; q += kWvBlock) {
            double l = 0.0, h = 0.0;
            for (int i = 0; i < T.lh; ++i) {
                int k = 2 * q + i;
                while (k >= bn) k -= bn;
                const double v = row[k];
                l = __dadd_rn(l, __dmul_rn(v, T.lo[i]));
                h = __dadd_rn(h, __dmul_rn(v, T.hi[i]));
            }
            Rj[q] = l;
            Rj[half + q] = h;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kWvLevelCols; ++u) {
            const int c = tid + u * kWvBlock;
            if (c < bn) {
                const double v = Rj[c];
                accL[u] = __dadd_rn(accL[u], __dmul_rn(v, T.lo[j]));
                accH[u] = __dadd_rn(accH[u], __dmul_rn(v, T.hi[j]));
            }
        }
    }
    const long long rl = ib + (long long)p * ld, rh = ib + (long long)((bm >> 1) + p) * ld;
#pragma unroll
    for (int u = 0; u < kWvLevelCols; ++u) {
        const int c = tid + u * kWvBlock;
        if (c >= bn) continue;
        (c < half ? ll : dst)[rl + c] = accL[u];
        dst[rh + c] = accH[u];
    }
}

// Synthesis: work-group (i, img) forms the columns pass of output rows 2i and 2i + 1 in
// LDS (the block's quadrant [0, bm/2) x [0, bn/2) is the coarser reconstruction `ll`, the
// rest the level's coefficients C) and then their rows pass.  ALPHA: the outputs are
// Psi(s + nu) of the ADMM and the kernel stores alpha = max((a + rho v) / (1 + rho), 0)
// instead (npg/admm.m:26).
template <bool ALPHA>
__global__ __launch_bounds__(kWvBlock) void k_idwt_level(WvTaps T, const double* __restrict__ C,
                                                         const double* __restrict__ ll,
                                                         double* __restrict__ dst, int bm, int bn,
                                                         int ld, long long istride,
                                                         const int* __restrict__ stop,
                                                         const double* __restrict__ a,
                                                         const float* __restrict__ mask,
                                                         const AdmmCtl* __restrict__ ctl)
{
    if (stop && *stop) return;
    __shared__ double E[kWvLevelMax], O[kWvLevelMax];
    const int i = blockIdx.x, tid = threadIdx.x, M = bm >> 1, half = bn >> 1, nj = T.lh >> 1;
    const long long ib = (long long)blockIdx.y * istride;
    for (int c = tid; c < bn; c += kWvBlock) {
        const double* lsrc = c < half ? ll : C;  // rows k < M
        double e = 0.0, o = 0.0;
        for (int j = 0; j < nj; ++j) {
            int k = i + j - (nj - 1);
            while (k < 0) k += M;
            const double lo = lsrc[ib + (long long)k * ld + c];
            const double hi = C[ib + (long long)(M + k) * ld + c];
            e = __dadd_rn(__dadd_rn(e, __dmul_rn(lo, T.e0[j])), __dmul_rn(hi, T.e1[j]));
            o = __dadd_rn(__dadd_rn(o, __dmul_rn(lo, T.o0[j])), __dmul_rn(hi, T.o1[j]));
        }
        E[c] = e;
        O[c] = o;
    }
    __syncthreads();
    const double rho = ALPHA ? ctl->rho : 0.0;
    const long long r0 = ib + (long long)(2 * i) * ld, r1 = r0 + ld;
    for (int q = tid; q < half; q += kWvBlock) {
        double e0 = 0.0, o0 = 0.0, e1 = 0.0, o1 = 0.0;
        for (int j = 0; j < nj; ++j) {
            int k = q + j - (nj - 1);
            while (k < 0) k += half;
            e0 = __dadd_rn(__dadd_rn(e0, __dmul_rn(E[k], T.e0[j])), __dmul_rn(E[half + k], T.e1[j]));
            o0 = __dadd_rn(__dadd_rn(o0, __dmul_rn(E[k], T.o0[j])), __dmul_rn(E[half + k], T.o1[j]));
            e1 = __dadd_rn(__dadd_rn(e1, __dmul_rn(O[k], T.e0[j])), __dmul_rn(O[half + k], T.e1[j]));
            o1 = __dadd_rn(__dadd_rn(o1, __dmul_rn(O[k], T.o0[j])), __dmul_rn(O[half + k], T.o1[j]));
        }
        const long long g[4] = {r0 + 2 * q, r0 + 2 * q + 1, r1 + 2 * q, r1 + 2 * q + 1};
        const double v[4] = {e0, o0, e1, o1};
#pragma unroll
        for (int t = 0; t < 4; ++t)
            dst[g[t]] = ALPHA ? fmax(__dadd_rn(a[g[t]], __dmul_rn(rho, masked(mask, g[t], v[t]))) /
                                         (1.0 + rho),
                                     0.0)
                              : v[t];
    }
}

// dst = src (a level-by-level transform's result moved to its output)
__global__ void k_copy64(const double* __restrict__ src, double* __restrict__ dst, long long n,
                         const int* __restrict__ stop)
{
    if (stop && *stop) return;
    const long long i = (long long)blockIdx.x * kWvBlock + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// x = mask ? x : 0 in place
__global__ void k_f64_to_f64_masked(double* __restrict__ x, long long n,
                                    const float* __restrict__ mask)
{
    const long long i = (long long)blockIdx.x * kWvBlock + threadIdx.x;
    if (i < n) x[i] = masked(mask, i, x[i]);
}

// out = max((a + rho Psi(s + nu)) / (1 + rho), 0) * scale   (npg/admm.m:52), FP32; zero for
// a zero input
__global__ void k_admm_final(const double* __restrict__ a, const double* __restrict__ psi,
                             float* __restrict__ out, long long n, const float* __restrict__ mask,
                             const AdmmCtl* __restrict__ c)
{
    const long long i = (long long)blockIdx.x * kWvBlock + threadIdx.x;
    if (i >= n) return;
    if (c->zero) {
        out[i] = 0.0f;
        return;
    }
    const double rho = c->rho, scale = c->scale;
    out[i] = (float)(fmax(__dadd_rn(a[i], __dmul_rn(rho, masked(mask, i, psi[i]))) / (1.0 + rho),
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
    AdmmCtl* d_ctl = nullptr;  // the prox loop's state
    float* d_mask = nullptr;   // image support (NULL: every pixel)
    hipStream_t cap_stream = nullptr;        // where the ADMM chunk is captured
    hipGraphExec_t chunk_exec = nullptr;     // the captured chunk (rebuilt with new masks)
    float* d_cmask = nullptr;  // coefficient support (NULL: every coefficient)
    long long blocks = 0;
    std::vector<double> part;
    ~ctprj_wavelet()
    {
        if (chunk_exec) (void)hipGraphExecDestroy(chunk_exec);
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        for (void* p : {(void*)d_x, (void*)d_y, (void*)d_a, (void*)d_s, (void*)d_nu,
                        (void*)d_pre, (void*)d_pa, (void*)d_part, (void*)d_ctl, (void*)d_mask,
                        (void*)d_cmask})
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

// first level whose block fits in LDS (w->L: none)
int small_level(const ctprj_wavelet* w)
{
    for (int lev = 0; lev < w->L; ++lev)
        if ((long long)(w->rows >> lev) * (w->cols >> lev) <= kWvSmall) return lev;
    return w->L;
}

// Level by level in place (x; tmp the pass buffer), two launches per level until the
// block fits in LDS; the kernels return at once while *stop is set (stop NULL: never).
int dwt64(ctprj_wavelet* w, double* x, double* tmp, hipStream_t st, const int* stop)
{
    const long long is = (long long)w->rows * w->cols;
    const int ls = small_level(w);
    for (int lev = 0; lev < ls; ++lev) {
        const int bm = w->rows >> lev, bn = w->cols >> lev;
        dim3 g(grid1((long long)(bn / 2) * bm), (unsigned)w->batch);
        hipLaunchKernelGGL(k_dwt_rows, g, dim3(kWvBlock), 0, st, w->T, x, tmp, bm, bn, w->cols, is,
                           stop);
        dim3 g2(grid1((long long)(bm / 2) * bn), (unsigned)w->batch);
        hipLaunchKernelGGL(k_dwt_cols, g2, dim3(kWvBlock), 0, st, w->T, tmp, x, bm, bn, w->cols, is,
                           stop);
    }
    if (ls < w->L)
        hipLaunchKernelGGL(k_dwt_small, dim3((unsigned)w->batch), dim3(kWvSmallThreads), 0, st,
                           w->T, x, x, w->rows >> ls, w->cols >> ls, w->L - ls, w->cols, is, stop);
    return last("dwt launch");
}

int idwt64(ctprj_wavelet* w, double* x, double* tmp, hipStream_t st, const int* stop)
{
    const long long is = (long long)w->rows * w->cols;
    const int ls = small_level(w);
    if (ls < w->L)
        hipLaunchKernelGGL(k_idwt_small, dim3((unsigned)w->batch), dim3(kWvSmallThreads), 0, st,
                           w->T, x, x, w->rows >> ls, w->cols >> ls, w->L - ls, w->cols, is, stop);
    for (int lev = ls - 1; lev >= 0; --lev) {
        const int bm = w->rows >> lev, bn = w->cols >> lev;
        dim3 g(grid1((long long)(bm / 2) * bn), (unsigned)w->batch);
        hipLaunchKernelGGL(k_idwt_cols, g, dim3(kWvBlock), 0, st, w->T, x, tmp, bm, bn, w->cols, is,
                           stop);
        dim3 g2(grid1((long long)(bn / 2) * bm), (unsigned)w->batch);
        hipLaunchKernelGGL(k_idwt_rows, g2, dim3(kWvBlock), 0, st, w->T, tmp, x, bm, bn, w->cols, is,
                           stop);
    }
    return last("idwt launch");
}

// Forward transform of src into dst with scratch t1; src is overwritten.  Up to
// kWvLevelMax columns one launch per level, the approximations alternating between t1
// and src, then the LDS-resident coarse levels; wider images level by level in place.
int dwt(ctprj_wavelet* w, double* src, double* dst, double* t1, hipStream_t st,
        const int* stop = nullptr)
{
    const long long is = (long long)w->rows * w->cols;
    if (w->cols > kWvLevelMax) {
        int rc = dwt64(w, src, t1, st, stop);
        if (rc) return rc;
        hipLaunchKernelGGL(k_copy64, dim3(grid1(w->n)), dim3(kWvBlock), 0, st, src, dst, w->n, stop);
        return last("dwt copy");
    }
    const int ls = small_level(w);
    const double* in = src;
    for (int lev = 0; lev < ls; ++lev) {
        const int bm = w->rows >> lev, bn = w->cols >> lev;
        double* llb = lev == w->L - 1 ? dst : ((lev & 1) ? src : t1);
        hipLaunchKernelGGL(k_dwt_level, dim3((unsigned)(bm / 2), (unsigned)w->batch), dim3(kWvBlock),
                           0, st, w->T, in, dst, llb, bm, bn, w->cols, is, stop);
        in = llb;
    }
    if (ls < w->L)
        hipLaunchKernelGGL(k_dwt_small, dim3((unsigned)w->batch), dim3(kWvSmallThreads), 0, st,
                           w->T, in, dst, w->rows >> ls, w->cols >> ls, w->L - ls, w->cols, is, stop);
    return last("dwt launch");
}

// Inverse transform of the coefficients C into dst with scratch t1; C may be overwritten.
// The coarse levels run in LDS, then one launch per level, the reconstructions
// alternating so that level 0 writes dst.  alpha (the ADMM's a, image mask and state):
// dst receives alpha = max((a + rho Psi(C)) / (1 + rho), 0) instead of Psi(C).
struct AlphaOut {
    const double* a;
    const float* mask;
    const AdmmCtl* ctl;
};

int idwt(ctprj_wavelet* w, double* C, double* dst, double* t1, hipStream_t st,
         const int* stop = nullptr, const AlphaOut* alpha = nullptr)
{
    const long long is = (long long)w->rows * w->cols;
    const unsigned g = grid1(w->n);
    if (w->cols > kWvLevelMax) {
        int rc = idwt64(w, C, t1, st, stop);
        if (rc) return rc;
        if (alpha)
            hipLaunchKernelGGL(k_admm_alpha, dim3(g), dim3(kWvBlock), 0, st, alpha->a, C, dst, w->n,
                               alpha->mask, alpha->ctl, stop);
        else
            hipLaunchKernelGGL(k_copy64, dim3(g), dim3(kWvBlock), 0, st, C, dst, w->n, stop);
        return last("idwt output");
    }
    const int ls = small_level(w);
    auto out = [&](int lev) { return (lev & 1) ? t1 : dst; };
    const double* llb = C;
    // the whole transform in LDS with the alpha output: Psi(C) in t1, alpha from it below
    double* small_out = (alpha && ls == 0) ? t1 : out(ls);
    if (ls < w->L) {
        hipLaunchKernelGGL(k_idwt_small, dim3((unsigned)w->batch), dim3(kWvSmallThreads), 0, st,
                           w->T, C, small_out, w->rows >> ls, w->cols >> ls, w->L - ls, w->cols, is,
                           stop);
        llb = small_out;
    }
    for (int lev = ls - 1; lev >= 0; --lev) {
        const int bm = w->rows >> lev, bn = w->cols >> lev;
        const dim3 gl((unsigned)(bm / 2), (unsigned)w->batch);
        if (lev == 0 && alpha)
            hipLaunchKernelGGL(k_idwt_level<true>, gl, dim3(kWvBlock), 0, st, w->T, C, llb, dst, bm,
                               bn, w->cols, is, stop, alpha->a, alpha->mask, alpha->ctl);
        else
            hipLaunchKernelGGL(k_idwt_level<false>, gl, dim3(kWvBlock), 0, st, w->T, C, llb,
                               out(lev), bm, bn, w->cols, is, stop, nullptr, nullptr, nullptr);
        llb = out(lev);
    }
    if (alpha && ls == 0)
        hipLaunchKernelGGL(k_admm_alpha, dim3(g), dim3(kWvBlock), 0, st, alpha->a, t1, dst, w->n,
                           alpha->mask, alpha->ctl, stop);
    return last("idwt launch");
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
    if (e == hipSuccess) e = hipMalloc((void**)&w->d_ctl, sizeof(AdmmCtl));
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
    if (w->chunk_exec) {  // it holds the old masks' addresses
        (void)hipGraphExecDestroy(w->chunk_exec);
        w->chunk_exec = nullptr;
    }
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
    int rc = dwt(w, w->d_x, w->d_y, w->d_pa, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_f64_to_f32, dim3(grid1(w->n)), dim3(kWvBlock), 0, st, w->d_y, d_y, w->n,
                       w->d_cmask);
    return last("dwt output");
}

extern "C" int ctprj_idwt(ctprj_wavelet* w, const float* d_y, float* d_x, void* stream)
{
    if (!w || !d_x || !d_y) return fail(CTPRJ_EINVAL, "idwt: NULL argument");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_f32_to_f64, dim3(grid1(w->n)), dim3(kWvBlock), 0, st, d_y, w->d_x, w->n, 1.0,
                       w->d_cmask);
    int rc = idwt(w, w->d_x, w->d_y, w->d_pa, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_f64_to_f32, dim3(grid1(w->n)), dim3(kWvBlock), 0, st, w->d_y, d_x, w->n,
                       w->d_mask);
    return last("idwt output");
}

long long ctprj::wavelet_l1_parts(const ctprj_wavelet* w) { return 2 * w->blocks; }

int ctprj::wavelet_l1_queue(ctprj_wavelet* w, const float* d_x, double* d_part, hipStream_t st)
{
    hipLaunchKernelGGL(k_f32_to_f64, dim3(grid1(w->n)), dim3(kWvBlock), 0, st, d_x, w->d_x, w->n, 1.0,
                       w->d_mask);
    int rc = dwt(w, w->d_x, w->d_y, w->d_pa, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_norms, dim3(grid1(w->n)), dim3(kWvBlock), 0, st, w->d_y, w->n, w->d_cmask,
                       d_part);
    return last("wavelet l1 launch");
}

double ctprj::wavelet_l1_total(const ctprj_wavelet* w, const double* part)
{
    double sa = 0.0;
    for (long long i = 0; i < w->blocks; ++i) sa += part[2 * i];
    return sa;
}

extern "C" int ctprj_wavelet_l1(ctprj_wavelet* w, const float* d_x, double* l1, void* stream)
{
    if (!w || !d_x || !l1) return fail(CTPRJ_EINVAL, "wavelet_l1: NULL argument");
    hipStream_t st = (hipStream_t)stream;
    int rc = wavelet_l1_queue(w, d_x, w->d_part, st);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(w->part.data(), w->d_part, sizeof(double) * 2 * w->blocks,
                                  hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "wavelet l1");
    *l1 = wavelet_l1_total(w, w->part.data());
    return 0;
}

namespace {

constexpr int kAdmmChunk = 3;  // ADMM iterations between two looks at the loop state

// kAdmmChunk iterations of npg/admm.m:25-50 on stream st.  Everything they read and
// write is held by the context (the call's tolerance and bound are in the loop state),
// so the sequence is captured once as a graph and replayed.
int admm_chunk(ctprj_wavelet* w, hipStream_t st)
{
    const long long n = w->n;
    const unsigned g = grid1(n);
    AdmmCtl* C = w->d_ctl;
    const int* stop = &C->done;
    const AlphaOut aout{w->d_a, w->d_mask, C};
    int rc;
    for (int k = 0; k < kAdmmChunk; ++k) {
        // alpha = max((a + rho Psi(s + nu)) / (1 + rho), 0) in d_pa
        hipLaunchKernelGGL(k_admm_sum, dim3(g), dim3(kWvBlock), 0, st, w->d_s, w->d_nu, w->d_x, n,
                           C, stop);
        if ((rc = idwt(w, w->d_x, w->d_pa, w->d_y, st, stop, &aout))) return rc;
        // Psit(alpha) in d_x, then the shrink and the stopping test
        if ((rc = dwt(w, w->d_pa, w->d_x, w->d_y, st, stop))) return rc;
        hipLaunchKernelGGL(k_admm_shrink, dim3(g), dim3(kWvBlock), 0, st, w->d_x, w->d_s, w->d_nu,
                           w->d_pre, n, w->d_cmask, w->d_part, C, stop);
        hipLaunchKernelGGL(k_admm_ctrl, dim3(1), dim3(kWvBlock), 0, st, w->d_part, w->blocks, C);
        if ((rc = last("admm launch"))) return rc;
    }
    return 0;
}

// The chunk as an instantiated graph, captured on the context's own stream (default
// priority) at first use.
int admm_graph(ctprj_wavelet* w)
{
    if (w->chunk_exec) return 0;
    hipError_t e = hipSuccess;
    if (!w->cap_stream) e = hipStreamCreateWithFlags(&w->cap_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(w->cap_stream, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) return hip_fail(e, "admm capture");
    const int rc = admm_chunk(w, w->cap_stream);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(w->cap_stream, &graph);
    if (!rc && e == hipSuccess) e = hipGraphInstantiate(&w->chunk_exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (rc) return rc;
    return e == hipSuccess ? 0 : hip_fail(e, "admm graph");
}

}  // namespace

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
    AdmmCtl* C = w->d_ctl;
    const int* stop = &C->done;
    // scale = ||a||, a /= scale, u /= scale  (npg/admm.m:17-18)
    hipLaunchKernelGGL(k_f32_to_f64, dim3(g), dim3(kWvBlock), 0, st, d_a, w->d_a, n, 1.0, w->d_mask);
    hipLaunchKernelGGL(k_norms, dim3(g), dim3(kWvBlock), 0, st, w->d_a, n, nullptr, w->d_part);
    hipLaunchKernelGGL(k_admm_init, dim3(1), dim3(kWvBlock), 0, st, w->d_part, w->blocks, u, relTol,
                       maxItr, C);
    hipLaunchKernelGGL(k_admm_scale_in, dim3(g), dim3(kWvBlock), 0, st, w->d_a, n, C);
    // preS = s = Psit(init / scale), nu = 0
    hipLaunchKernelGGL(k_admm_load, dim3(g), dim3(kWvBlock), 0, st, d_init ? d_init : d_a,
                       w->d_x, n, C, w->d_mask);
    int rc = dwt(w, w->d_x, w->d_pre, w->d_y, st, stop);
    if (rc) return rc;
    if (w->d_cmask)
        hipLaunchKernelGGL(k_f64_to_f64_masked, dim3(g), dim3(kWvBlock), 0, st, w->d_pre, n,
                           w->d_cmask);
    hipLaunchKernelGGL(k_admm_start, dim3(g), dim3(kWvBlock), 0, st, w->d_pre, w->d_s, w->d_nu, n);
    if ((rc = last("admm init launch"))) return rc;
    // iterations in chunks of kAdmmChunk (one graph launch), the host reading the state
    // after each chunk
    if ((rc = admm_graph(w))) return rc;
    AdmmCtl h{};
    // the stop comes after at most maxItr + 1 iterations
    const long long max_chunks = ((long long)maxItr + 1) / kAdmmChunk + 2;
    for (long long chunk = 0;; ++chunk) {
        if (chunk == max_chunks)
            return fail(CTPRJ_EHIP, "admm: the loop state did not stop");
        hipError_t e = hipGraphLaunch(w->chunk_exec, st);
        if (e != hipSuccess) return hip_fail(e, "admm chunk");
        e = hipMemcpyAsync(&h, C, sizeof(h), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e, "admm state");
        if (h.done) break;
    }
    hipLaunchKernelGGL(k_admm_sum, dim3(g), dim3(kWvBlock), 0, st, w->d_s, w->d_nu, w->d_x, n, C,
                       nullptr);
    if ((rc = idwt(w, w->d_x, w->d_pa, w->d_y, st))) return rc;
    hipLaunchKernelGGL(k_admm_final, dim3(g), dim3(kWvBlock), 0, st, w->d_a, w->d_pa, d_out, n,
                       w->d_mask, C);
    if (iterations) *iterations = h.itr;
    return last("admm output");
}
