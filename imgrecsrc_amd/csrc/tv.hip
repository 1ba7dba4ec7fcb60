// Total-variation proximal step on gfx950: argmin 0.5||D - X||^2 + lambda TV(D) over
// D >= 0 by the fast gradient projection on the dual (utils/denoise_bound_mod.m, as
// utils/TV.m:56-80 calls it), isotropic or l1 TV (utils/tlv.m), and the TV value.
//
// One iteration is two launches over the pixels: the primal step (C = X - lambda
// Lforward(R), D = max(C, 0), with the FP64 partial sums of ||D - Dold||^2, ||D||^2,
// ||C - D||^2 and ||C||^2) and the dual step (P = proj(R + Ltrans(D)/(8 lambda)), then
// the momentum R = P + (tk - 1)(P - Pold)/tkp1).  The dual fields P1 (vertical
// differences, rows 0..m-2) and P2 (horizontal, columns 0..n-2) are stored padded to
// m x n.  A one-work-group launch per iteration adds the four sums for the stopping test
// and the momentum restart and keeps the loop state on the device; the host reads it once
// per four iterations.  All arithmetic is FP64 in the reference's order; images are FP32.
#include <hip/hip_runtime.h>

#include <cmath>
#include <new>
#include <vector>

#include "ctprj_internal.h"

namespace ctprj {

constexpr int kTvBlock = 256;

__device__ __forceinline__ void tv_block_sum4(const double v[4], double* part)
{
    __shared__ double sh[4][kTvBlock];
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < 4; ++k) sh[k][t] = v[k];
    __syncthreads();
    for (int off = kTvBlock / 2; off > 0; off >>= 1) {
        if (t < off) {
#pragma unroll
            for (int k = 0; k < 4; ++k) sh[k][t] += sh[k][t + off];
        }
        __syncthreads();
    }
    if (t == 0) {
#pragma unroll
        for (int k = 0; k < 4; ++k) part[4 * (long long)blockIdx.x + k] = sh[k][0];
    }
}

// primal step (utils/denoise_bound_mod.m:118-119 with utils/Lforward.m)
// FGP loop state on the device (utils/denoise_bound_mod.m:95-150): the iteration's
// kernels read tk from here and k_tv_ctrl updates it, so the host looks at the state once
// per chunk of iterations; kernels queued after the stop return at once.
struct TvCtl {
    double t;     // tk of the next iteration (the last tkp1, or 1 after a restart)
    double fval;  // objective of the last iteration
    int i, count, done, pad;
};

// tkp1 from tk (utils/denoise_bound_mod.m:117), the same expression wherever it is formed
__device__ __forceinline__ double tv_next_t(double tk) { return (1.0 + sqrt(1.0 + 4.0 * tk * tk)) / 2.0; }

__global__ __launch_bounds__(kTvBlock) void k_tv_primal(const double* __restrict__ X,
                                                        double* __restrict__ D,
                                                        const double* __restrict__ R1,
                                                        const double* __restrict__ R2, int m,
                                                        int n, long long total, double lam,
                                                        double* __restrict__ part,
                                                        const TvCtl* __restrict__ c)
{
    if (c->done) return;  // uniform
    const long long t = (long long)blockIdx.x * kTvBlock + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (t < total) {
        const int j = (int)(t % n);
        const int i = (int)((t / n) % m);
        const double x = (i == 0) ? R1[t] : (i == m - 1 ? -R1[t - n] : R1[t] - R1[t - n]);
        const double y = (j == 0) ? R2[t] : (j == n - 1 ? -R2[t - 1] : R2[t] - R2[t - 1]);
        const double c = __dsub_rn(X[t], __dmul_rn(lam, x + y));
        const double d = fmax(c, 0.0);
        const double e = d - D[t], f = c - d;
        D[t] = d;
        v[0] = e * e;
        v[1] = d * d;
        v[2] = f * f;
        v[3] = c * c;
    }
    tv_block_sum4(v, part);
}

// dual step (utils/denoise_bound_mod.m:120-144 with utils/Ltrans.m)
__global__ __launch_bounds__(kTvBlock) void k_tv_dual(const double* __restrict__ D,
                                                      double* __restrict__ P1,
                                                      double* __restrict__ P2,
                                                      double* __restrict__ R1,
                                                      double* __restrict__ R2, int m, int n,
                                                      long long total, double lam, int iso,
                                                      const TvCtl* __restrict__ c)
{
    if (c->done) return;  // uniform
    const double tk = c->t, tkp1 = tv_next_t(tk);
    const long long t = (long long)blockIdx.x * kTvBlock + threadIdx.x;
    if (t >= total) return;
    const int j = (int)(t % n);
    const int i = (int)((t / n) % m);
    const bool has1 = i < m - 1, has2 = j < n - 1;
    const double d = D[t];
    const double l8 = 8.0 * lam;
    double p1 = has1 ? R1[t] + (d - D[t + n]) / l8 : 0.0;
    double p2 = has2 ? R2[t] + (d - D[t + 1]) / l8 : 0.0;
    if (iso) {
        const double a = sqrt(fmax(__dadd_rn(__dmul_rn(p1, p1), __dmul_rn(p2, p2)), 1.0));
        p1 = p1 / a;
        p2 = p2 / a;
    } else {
        p1 = fmax(fmin(p1, 1.0), -1.0);
        p2 = fmax(fmin(p2, 1.0), -1.0);
    }
    if (has1) {
        const double o = P1[t];
        P1[t] = p1;
        R1[t] = p1 + (tk - 1.0) * (p1 - o) / tkp1;
    }
    if (has2) {
        const double o = P2[t];
        P2[t] = p2;
        R2[t] = p2 + (tk - 1.0) * (p2 - o) / tkp1;
    }
}

// D = P = R = 0 and the loop state: tk = 1, no iteration yet (utils/denoise_bound_mod.m:60-86)
__global__ void k_tv_start(double* __restrict__ D, double* __restrict__ P1, double* __restrict__ P2,
                           double* __restrict__ R1, double* __restrict__ R2, long long total,
                           int maxIter, TvCtl* __restrict__ c)
{
    const long long t = (long long)blockIdx.x * kTvBlock + threadIdx.x;
    if (t < total) D[t] = P1[t] = P2[t] = R1[t] = R2[t] = 0.0;
    if (t == 0) {
        c->t = 1.0;
        c->fval = INFINITY;
        c->i = 0;
        c->count = 0;
        c->done = maxIter <= 0;
    }
}

// End of an iteration (utils/denoise_bound_mod.m:146-150): re = ||D - Dold|| / ||D||
// counts towards the stop below epsilon, fval = -||C - D||^2 + ||C||^2, a rising fval
// restarts the momentum (tkp1 = 1); stop after maxIter iterations or 5 in a row below
// epsilon.  The per-group partials are summed in a fixed order.
__global__ __launch_bounds__(kTvBlock) void k_tv_ctrl(const double* __restrict__ part,
                                                      long long nblk, double epsilon,
                                                      int maxIter, TvCtl* __restrict__ c)
{
    if (c->done) return;
    __shared__ double sh[4][kTvBlock];
    const int t = threadIdx.x;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (long long b = t; b < nblk; b += kTvBlock)
#pragma unroll
        for (int k = 0; k < 4; ++k) a[k] += part[4 * b + k];
#pragma unroll
    for (int k = 0; k < 4; ++k) sh[k][t] = a[k];
    __syncthreads();
    for (int off = kTvBlock / 2; off > 0; off >>= 1) {
        if (t < off) {
#pragma unroll
            for (int k = 0; k < 4; ++k) sh[k][t] += sh[k][t + off];
        }
        __syncthreads();
    }
    if (t != 0) return;
    const double re = sqrt(sh[0][0]) / sqrt(sh[1][0]);
    const int count = (re < epsilon) ? c->count + 1 : 0;
    const double fval = -sh[2][0] + sh[3][0];
    const double tkp1 = tv_next_t(c->t);
    const int i = c->i + 1;
    c->t = (fval > c->fval) ? 1.0 : tkp1;
    c->fval = fval;
    c->count = count;
    c->i = i;
    c->done = i >= maxIter || count >= 5;
}

// Support mask (utils/maskFunc.m): pixels where mask is 0 are zero; NULL keeps all.
__device__ __forceinline__ double tv_masked(const float* mask, long long i, double v)
{
    return (!mask || mask[i] != 0.0f) ? v : 0.0;
}

// TV of FP32 images (utils/tlv.m): iso sum sqrt(P1^2 + P2^2), l1 sum |P1| + |P2|
__global__ __launch_bounds__(kTvBlock) void k_tv_value(const float* __restrict__ X, int m, int n,
                                                       long long total, int iso,
                                                       const float* __restrict__ mask,
                                                       double* __restrict__ part)
{
    const long long t = (long long)blockIdx.x * kTvBlock + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (t < total) {
        const int j = (int)(t % n);
        const int i = (int)((t / n) % m);
        const double x = tv_masked(mask, t, X[t]);
        const double p1 = i < m - 1 ? x - tv_masked(mask, t + n, X[t + n]) : 0.0;
        const double p2 = j < n - 1 ? x - tv_masked(mask, t + 1, X[t + 1]) : 0.0;
        v[0] = iso ? sqrt(p1 * p1 + p2 * p2) : fabs(p1) + fabs(p2);
    }
    tv_block_sum4(v, part);
}

__global__ void k_tv_in(const float* __restrict__ in, double* __restrict__ out, long long n,
                        const float* __restrict__ mask)
{
    const long long i = (long long)blockIdx.x * kTvBlock + threadIdx.x;
    if (i < n) out[i] = tv_masked(mask, i, (double)in[i]);
}

__global__ void k_tv_out(const double* __restrict__ in, float* __restrict__ out, long long n,
                         const float* __restrict__ mask)
{
    const long long i = (long long)blockIdx.x * kTvBlock + threadIdx.x;
    if (i < n) out[i] = (float)tv_masked(mask, i, in[i]);
}

}  // namespace ctprj

using namespace ctprj;

struct ctprj_tv {
    int rows = 0, cols = 0, batch = 0, iso = 1;
    long long n = 0, blocks = 0;
    double *d_X = nullptr, *d_D = nullptr, *d_P1 = nullptr, *d_P2 = nullptr, *d_R1 = nullptr,
           *d_R2 = nullptr, *d_part = nullptr;
    TvCtl* d_ctl = nullptr;   // the FGP loop's state
    float* d_mask = nullptr;  // support (NULL: every pixel)
    std::vector<double> part;
    ~ctprj_tv()
    {
        for (void* p : {(void*)d_X, (void*)d_D, (void*)d_P1, (void*)d_P2, (void*)d_R1, (void*)d_R2,
                        (void*)d_part, (void*)d_ctl, (void*)d_mask})
            if (p) (void)hipFree(p);
    }
};

long long ctprj::tv_elems(const ctprj_tv* w) { return w ? w->n : 0; }

extern "C" int ctprj_tv_create(int rows, int cols, int batch, int iso, ctprj_tv** out)
{
    if (!out) return fail(CTPRJ_EINVAL, "tv_create: NULL argument");
    *out = nullptr;
    if (rows < 2 || cols < 2 || batch < 1)
        return fail(CTPRJ_EINVAL, "tv_create: bad sizes (%d x %d x %d)", rows, cols, batch);
    const long long n = (long long)rows * cols * batch;
    if (n >= (1LL << 31)) return fail(CTPRJ_EINVAL, "tv_create: %lld elements is too many", n);
    ctprj_tv* w = new (std::nothrow) ctprj_tv();
    if (!w) return fail(CTPRJ_ENOMEM, "host allocation");
    w->rows = rows;
    w->cols = cols;
    w->batch = batch;
    w->iso = iso ? 1 : 0;
    w->n = n;
    w->blocks = (n + kTvBlock - 1) / kTvBlock;
    w->part.assign((size_t)4 * w->blocks, 0.0);
    hipError_t e = hipSuccess;
    for (double** p : {&w->d_X, &w->d_D, &w->d_P1, &w->d_P2, &w->d_R1, &w->d_R2})
        if (e == hipSuccess) e = hipMalloc((void**)p, sizeof(double) * (size_t)n);
    if (e == hipSuccess) e = hipMalloc((void**)&w->d_part, sizeof(double) * 4 * (size_t)w->blocks);
    if (e == hipSuccess) e = hipMalloc((void**)&w->d_ctl, sizeof(TvCtl));
    if (e != hipSuccess) {
        delete w;
        return fail(CTPRJ_ENOMEM, "tv workspace: %s", hipGetErrorString(e));
    }
    *out = w;
    return 0;
}

extern "C" int ctprj_tv_set_mask(ctprj_tv* w, const float* d_mask)
{
    if (!w) return fail(CTPRJ_EINVAL, "tv_set_mask: NULL context");
    if (w->d_mask) (void)hipFree(w->d_mask);
    w->d_mask = nullptr;
    if (!d_mask) return 0;
    const size_t bytes = sizeof(float) * (size_t)w->n;
    hipError_t e = hipMalloc((void**)&w->d_mask, bytes);
    if (e == hipSuccess) e = hipMemcpy(w->d_mask, d_mask, bytes, hipMemcpyDeviceToDevice);
    return e == hipSuccess ? 0 : hip_fail(e, "tv mask");
}

extern "C" int ctprj_tv_destroy(ctprj_tv* w)
{
    delete w;
    return 0;
}

long long ctprj::tv_value_parts(const ctprj_tv* w) { return 4 * w->blocks; }

int ctprj::tv_value_queue(ctprj_tv* w, const float* d_x, double* d_part, hipStream_t st)
{
    hipLaunchKernelGGL(k_tv_value, dim3((unsigned)w->blocks), dim3(kTvBlock), 0, st, d_x, w->rows,
                       w->cols, w->n, w->iso, w->d_mask, d_part);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(e, "tv value launch");
}

double ctprj::tv_value_total(const ctprj_tv* w, const double* part)
{
    double s = 0.0;
    for (long long b = 0; b < w->blocks; ++b) s += part[4 * b];
    return s;
}

extern "C" int ctprj_tv_value(ctprj_tv* w, const float* d_x, double* tv, void* stream)
{
    if (!w || !d_x || !tv) return fail(CTPRJ_EINVAL, "tv_value: NULL argument");
    hipStream_t st = (hipStream_t)stream;
    int rc = tv_value_queue(w, d_x, w->d_part, st);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(w->part.data(), w->d_part, sizeof(double) * 4 * w->blocks,
                                  hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "tv value");
    *tv = tv_value_total(w, w->part.data());
    return 0;
}

extern "C" int ctprj_tv_denoise(ctprj_tv* w, const float* d_x, double lambda, double epsilon,
                                int maxIter, float* d_out, int* iterations, void* stream)
{
    if (!w || !d_x || !d_out) return fail(CTPRJ_EINVAL, "tv_denoise: NULL argument");
    if (!(lambda > 0.0) || !std::isfinite(lambda) || !std::isfinite(epsilon) || maxIter < 0)
        return fail(CTPRJ_EINVAL, "tv_denoise: lambda must be > 0, epsilon finite, maxIter >= 0");
    hipStream_t st = (hipStream_t)stream;
    const long long n = w->n;
    const unsigned g = (unsigned)w->blocks;
    hipLaunchKernelGGL(k_tv_in, dim3(g), dim3(kTvBlock), 0, st, d_x, w->d_X, n, w->d_mask);
    // D = 0, P = proj(Ltrans(0)) = 0, R = P   (utils/denoise_bound_mod.m:60-86)
    TvCtl* C = w->d_ctl;
    hipLaunchKernelGGL(k_tv_start, dim3(g), dim3(kTvBlock), 0, st, w->d_D, w->d_P1, w->d_P2,
                       w->d_R1, w->d_R2, n, maxIter, C);
    // iterations in chunks of kTvChunk, the host reading the state after each chunk
    constexpr int kTvChunk = 4;
    TvCtl h{};
    // the stop comes after at most maxIter iterations
    const long long max_chunks = (long long)maxIter / kTvChunk + 2;
    for (long long chunk = 0;; ++chunk) {
        if (chunk == max_chunks)
            return fail(CTPRJ_EHIP, "tv_denoise: the loop state did not stop");
        for (int k = 0; k < kTvChunk; ++k) {
            hipLaunchKernelGGL(k_tv_primal, dim3(g), dim3(kTvBlock), 0, st, w->d_X, w->d_D,
                               w->d_R1, w->d_R2, w->rows, w->cols, n, lambda, w->d_part, C);
            hipLaunchKernelGGL(k_tv_dual, dim3(g), dim3(kTvBlock), 0, st, w->d_D, w->d_P1,
                               w->d_P2, w->d_R1, w->d_R2, w->rows, w->cols, n, lambda, w->iso, C);
            hipLaunchKernelGGL(k_tv_ctrl, dim3(1), dim3(kTvBlock), 0, st, w->d_part, w->blocks,
                               epsilon, maxIter, C);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "tv iteration launch");
        e = hipMemcpyAsync(&h, C, sizeof(h), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e, "tv state");
        if (h.done) break;
    }
    hipLaunchKernelGGL(k_tv_out, dim3(g), dim3(kTvBlock), 0, st, w->d_D, d_out, n, w->d_mask);
    if (iterations) *iterations = h.i;
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(e, "tv output");
}
