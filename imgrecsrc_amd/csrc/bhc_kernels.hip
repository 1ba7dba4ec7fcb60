// Polychromatic beam-hardening model on gfx950: the Laplace transform of the
// mass-attenuation spectrum in the b1 / b0 / dis bases (bhc/Spline.m:122-302), the
// per-ray likelihood and gradient weights of bhc/poissLAlpha.m and bhc/gaussLAlpha.m,
// and the I-epsilon gradient / Hessian of bhc/poissLI.m and bhc/gaussLI.m.  Per-ray spectral values are FP32 and recomputed wherever needed (no
// ray x basis matrix is stored); the negative log-likelihood and every I-epsilon
// reduction are FP64 with a fixed summation order: per work-group partials, then one
// ordered pass.
#include "bhc_internal.h"

namespace ctprj {

// Moments M_n(z) = int_0^1 u^n exp(-z u) du, n = 0..3, of one spline segment.
// |z| <= 2: the series sum_k (-z)^k / (k! (n+k+1)) (18 terms, truncation < 3e-11);
// otherwise the recurrence M_0 = (1-e^-z)/z, M_n = (n M_{n-1} - e^-z)/z, whose error
// growth is at most n/|z| <= 1.5 per step there.
struct Moments {
    float m0, m1, m2, m3;
};

__device__ __forceinline__ Moments moments(float z)
{
    Moments M;
    if (fabsf(z) <= 2.0f) {
        float t = 1.0f;  // (-z)^k / k!
        float a0 = 1.0f, a1 = 0.5f, a2 = 1.0f / 3.0f, a3 = 0.25f;
#pragma unroll
        for (int k = 1; k <= 18; ++k) {
            t *= -z * (1.0f / (float)k);
            a0 = fmaf(t, 1.0f / (float)(k + 1), a0);
            a1 = fmaf(t, 1.0f / (float)(k + 2), a1);
            a2 = fmaf(t, 1.0f / (float)(k + 3), a2);
            a3 = fmaf(t, 1.0f / (float)(k + 4), a3);
        }
        M.m0 = a0;
        M.m1 = a1;
        M.m2 = a2;
        M.m3 = a3;
    } else {
        const float e = expf(-z);
        const float iz = 1.0f / z;
        M.m0 = (1.0f - e) * iz;
        M.m1 = (M.m0 - e) * iz;
        M.m2 = (2.0f * M.m1 - e) * iz;
        M.m3 = (3.0f * M.m2 - e) * iz;
    }
    return M;
}

// Spectrum basis values for one ray value s, by basis type; calls f(j, phi, dphi, ddphi)
// for j = 0..E-1 (derivatives in s); WANT = number of orders needed (1..3).
//
// b1: basis j is the hat on [k_j, k_{j+2}] peaking at k_{j+1}.  Substituting
// kappa = k_j + Delta_j u in phi_j(s) = int hat_j(kappa) exp(-kappa s) dkappa, with
// Delta_j = k_{j+1} - k_j, z_j = Delta_j s and w_j = Delta_j exp(-k_j s), gives per
// segment j = [k_j, k_{j+1}]
//   left half of basis j   : w_j M1                ,  -w_j (k_j M1 + Delta_j M2)
//                            w_j (k_j^2 M1 + 2 k_j Delta_j M2 + Delta_j^2 M3)
//   right half of basis j-1: w_j (M0-M1)           ,  -w_j (k_j (M0-M1) + Delta_j (M1-M2))
//                            w_j (k_j^2 (M0-M1) + 2 k_j Delta_j (M1-M2) + Delta_j^2 (M2-M3))
// for phi, phi' and phi'' (all moments at z_j).  These equal the closed forms of
// bhc/Spline.m:154-201 without their small-s cancellation.
// b0: basis j is the box [k_j, k_{j+1}] (bhc/Spline.m:205-262):
//   w_j M0,  -w_j (k_j M0 + Delta_j M1),  w_j (k_j^2 M0 + 2 k_j Delta_j M1 + Delta_j^2 M2).
// dis: basis j is the node k_{j+1} with weight (k_{j+2} - k_j)/2 (bhc/Spline.m:264-302):
//   dk exp(-k s),  -k phi,  k^2 phi.
template <int WANT, class F>
__device__ __forceinline__ void spline_basis(const BhcParams& P, float s, F&& f)
{
    if (P.basis == CTPRJ_BASIS_DIS) {
        for (int j = 0; j < P.E; ++j) {
            const float kj = P.knot[j + 1];
            const float p = P.delta[j] * expf(-kj * s);
            f(j, p, WANT >= 2 ? -kj * p : 0.0f, WANT >= 3 ? kj * kj * p : 0.0f);
        }
        return;
    }
    if (P.basis == CTPRJ_BASIS_B0) {
        for (int j = 0; j < P.E; ++j) {
            const float kj = P.knot[j], dj = P.delta[j];
            const float w = dj * expf(-kj * s);
            const Moments M = moments(dj * s);
            float d = 0.0f, dd = 0.0f;
            if (WANT >= 2) d = -w * fmaf(kj, M.m0, dj * M.m1);
            if (WANT >= 3) dd = w * fmaf(kj * kj, M.m0, fmaf(2.0f * kj * dj, M.m1, dj * dj * M.m2));
            f(j, w * M.m0, d, dd);
        }
        return;
    }
    float cp = 0.0f, cd = 0.0f, cdd = 0.0f;  // left half of the pending basis j-1
    for (int j = 0; j <= P.E; ++j) {
        const float kj = P.knot[j], dj = P.delta[j];
        const float w = dj * expf(-kj * s);
        const Moments M = moments(dj * s);
        const float c0 = M.m0 - M.m1;
        if (j >= 1) {  // right half of basis j-1
            float rp = w * c0, rd = 0.0f, rdd = 0.0f;
            if (WANT >= 2) rd = -w * fmaf(kj, c0, dj * (M.m1 - M.m2));
            if (WANT >= 3)
                rdd = w * fmaf(kj * kj, c0,
                               fmaf(2.0f * kj * dj, M.m1 - M.m2, dj * dj * (M.m2 - M.m3)));
            f(j - 1, cp + rp, cd + rd, cdd + rdd);
        }
        if (j < P.E) {  // left half of basis j
            cp = w * M.m1;
            if (WANT >= 2) cd = -w * fmaf(kj, M.m1, dj * M.m2);
            if (WANT >= 3) cdd = w * fmaf(kj * kj, M.m1, fmaf(2.0f * kj * dj, M.m2, dj * dj * M.m3));
        }
    }
}

// ----------------------------------------------------------------- polyIout ---

template <int WANT>
__global__ __launch_bounds__(kBhcBlock) void k_poly_iout(BhcParams P, const float* __restrict__ s,
                                                       long long count, float* __restrict__ BL,
                                                       float* __restrict__ sBL,
                                                       float* __restrict__ ssBL, int matrix)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float si = s[i];
    if (matrix) {
        float* row = BL + i * P.E;
        spline_basis<1>(P, si, [&](int j, float p, float, float) { row[j] = p; });
        return;
    }
    float b = 0.0f, bd = 0.0f, bdd = 0.0f;
    spline_basis<WANT>(P, si, [&](int j, float p, float d, float dd) {
        b = fmaf(P.Ie[j], p, b);
        bd = fmaf(P.Ie[j], d, bd);
        bdd = fmaf(P.Ie[j], dd, bdd);
    });
    if (BL) BL[i] = b;
    if (sBL) sBL[i] = bd;
    if (ssBL) ssBL[i] = bdd;
}

// --------------------------------------------------------- FP64 block reduce ---

__device__ __forceinline__ double block_sum(double v, double* sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = kBhcBlock / 2; off > 0; off >>= 1) {
        if (t < off) sh[t] += sh[t + off];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// Poisson NLL term of one ray, bhc/poissLAlpha.m:10-11 / bhc/poissLI.m:39:
// (BL - Imea) - Imea log(BL/Imea), with the limit 0 for Imea = 0.
__device__ __forceinline__ double nll_term(double bl, double im)
{
    return (im > 0.0) ? (bl - im) - im * log(bl / im) : bl;
}

// Gaussian (log-domain) residual of one ray, bhc/gaussLAlpha.m:10 / bhc/gaussLI.m:3:
// Err = log(BL/Imea); the NLL term is Err^2/2.
__device__ __forceinline__ double gauss_err(double bl, double im) { return log(bl / im); }

// ------------------------------------------------------------- alpha pass ---

// w may alias s (each thread reads s[i] before writing w[i]): no __restrict__ on them.
template <bool HESS>
__global__ __launch_bounds__(kBhcBlock) void k_bhc_alpha_ray(BhcParams P, const float* s,
                                                             const float* __restrict__ Imea,
                                                             long long count, float sscale,
                                                             float Ts, float* w,
                                                             float* __restrict__ hw,
                                                             double* __restrict__ partials)
{
    __shared__ double sh[kBhcBlock];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double f = 0.0;
    if (i < count) {
        const float si = sscale * s[i], im = Imea[i];
        float b = 0.0f, bd = 0.0f, bdd = 0.0f;
        spline_basis<HESS ? 3 : 2>(P, si, [&](int j, float p, float d, float dd) {
            b = fmaf(P.Ie[j], p, b);
            bd = fmaf(P.Ie[j], d, bd);
            if (HESS) bdd = fmaf(P.Ie[j], dd, bdd);
        });
        if (P.noise == CTPRJ_NOISE_GAUSSIAN) {
            const double err = gauss_err((double)b, (double)im);
            f = 0.5 * err * err;
            const float e = (float)err, q = bd / b;
            if (w) w[i] = Ts * (e * q);  // Phit(Err.*sBL./BL), bhc/gaussLAlpha.m:16
            if (HESS && hw) hw[i] = fmaf(1.0f - e, q * q, e * (bdd / b));  // :18
        } else {
            f = nll_term((double)b, (double)im);
            const float r = 1.0f - im / b;
            if (w) w[i] = Ts * (r * bd);  // Phit((1-Imea./BL).*sBL) with Phit = Ts*A^T
            if (HESS && hw) {
                const float q = bd / b;
                hw[i] = fmaf(q * q, im, bdd * r);  // bhc/poissLAlpha.m:15
            }
        }
    }
    const double tot = block_sum(f, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// ------------------------------------------------------------ I-epsilon pass ---

// Each work-group takes kBhcBlock rays: the basis values phi_ij go to LDS, then thread v
// forms value v of [H upper triangle (row-major j<=k), g, f] over the group's rays:
//   H_jk = sum_i phi_ij phi_ik q_i,  g_j = sum_i phi_ij r_i,  f = sum_i f_i,
//   Ir_i = sum_j Ie_j phi_ij (FP64), with
//   Poisson  (bhc/poissLI.m):  q = Imea/Ir^2,       r = 1 - Imea/Ir,  f_i = nll(Ir, Imea)
//   Gaussian (bhc/gaussLI.m):  q = (1 - Err)/Ir^2,  r = Err/Ir,       f_i = Err^2/2.
__global__ __launch_bounds__(kBhcBlock) void k_bhc_ie_ray(BhcParams P, const float* __restrict__ s,
                                                          const float* __restrict__ Imea,
                                                          long long count, float sscale,
                                                          double* __restrict__ partials)
{
    __shared__ float phi[kBhcBlock][CTPRJ_MAX_E + 1];
    __shared__ double q[kBhcBlock], r[kBhcBlock], ft[kBhcBlock];
    const int t = threadIdx.x;
    const long long i = (long long)blockIdx.x * kBhcBlock + t;
    const int E = P.E;
    if (i < count) {
        const float si = sscale * s[i];
        const double im = (double)Imea[i];
        double ir = 0.0;
        spline_basis<1>(P, si, [&](int j, float p, float, float) {
            phi[t][j] = p;
            ir = fma(P.Ied[j], (double)p, ir);
        });
        if (P.noise == CTPRJ_NOISE_GAUSSIAN) {  // bhc/gaussLI.m:3-10
            const double err = gauss_err(ir, im);
            q[t] = (1.0 - err) / (ir * ir);
            r[t] = err / ir;
            ft[t] = 0.5 * err * err;
        } else {
            q[t] = im / (ir * ir);
            r[t] = 1.0 - im / ir;
            ft[t] = nll_term(ir, im);
        }
    } else {
        for (int j = 0; j < E; ++j) phi[t][j] = 0.0f;
        q[t] = 0.0;
        r[t] = 0.0;
        ft[t] = 0.0;
    }
    __syncthreads();
    const int nh = E * (E + 1) / 2, nv = nh + E + 1;
    double* out = partials + (long long)blockIdx.x * nv;
    for (int v = t; v < nv; v += kBhcBlock) {
        double acc = 0.0;
        if (v < nh) {
            int j = 0, rem = v;  // row j of the upper triangle holds E-j entries
            while (rem >= E - j) { rem -= E - j; ++j; }
            const int k = j + rem;
            for (int m = 0; m < kBhcBlock; ++m)
                acc = fma((double)phi[m][j] * (double)phi[m][k], q[m], acc);
        } else if (v < nh + E) {
            const int j = v - nh;
            for (int m = 0; m < kBhcBlock; ++m) acc = fma((double)phi[m][j], r[m], acc);
        } else {
            for (int m = 0; m < kBhcBlock; ++m) acc += ft[m];
        }
        out[v] = acc;
    }
}

// Cost only of the I-epsilon pass: f_i of every ray with Ir_i formed exactly as in
// k_bhc_ie_ray (FP32 basis values, FP64 sum over j in order), one partial per group.
__global__ __launch_bounds__(kBhcBlock) void k_bhc_ie_cost(BhcParams P, const float* __restrict__ s,
                                                           const float* __restrict__ Imea,
                                                           long long count, float sscale,
                                                           double* __restrict__ partials)
{
    __shared__ double sh[kBhcBlock];
    const long long i = (long long)blockIdx.x * kBhcBlock + threadIdx.x;
    double f = 0.0;
    if (i < count) {
        const float si = sscale * s[i];
        const double im = (double)Imea[i];
        double ir = 0.0;
        spline_basis<1>(P, si, [&](int j, float p, float, float) {
            ir = fma(P.Ied[j], (double)p, ir);
        });
        if (P.noise == CTPRJ_NOISE_GAUSSIAN) {
            const double err = gauss_err(ir, im);
            f = 0.5 * err * err;
        } else {
            f = nll_term(ir, im);
        }
    }
    const double tot = block_sum(f, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// out[v] = sum over p of partials[p*nvals + v]: one work-group per value, each thread
// sums a fixed stride of partials in order, then a fixed LDS tree.
__global__ __launch_bounds__(kBhcBlock) void k_reduce_partials(const double* __restrict__ partials,
                                                               long long nparts, int nvals,
                                                               double* __restrict__ out)
{
    __shared__ double sh[kBhcBlock];
    const int v = blockIdx.x;
    double acc = 0.0;
    for (long long p = threadIdx.x; p < nparts; p += kBhcBlock) acc += partials[p * nvals + v];
    const double tot = block_sum(acc, sh);
    if (threadIdx.x == 0) out[v] = tot;
}

__global__ void k_unpack_ie(const double* __restrict__ red, int E, double* __restrict__ f,
                            double* __restrict__ g, double* __restrict__ H)
{
    const int nh = E * (E + 1) / 2;
    for (int v = threadIdx.x; v < nh + E + 1; v += blockDim.x) {
        if (v < nh) {
            if (!H) continue;
            int j = 0, rem = v;
            while (rem >= E - j) { rem -= E - j; ++j; }
            const int k = j + rem;
            H[j * E + k] = red[v];
            H[k * E + j] = red[v];
        } else if (v < nh + E) {
            if (g) g[v - nh] = red[v];
        } else if (f) {
            *f = red[v];
        }
    }
}

// -------------------------------------------------------------------- launchers ---

void launch_poly_iout(const BhcParams& P, const float* s, long long count, float* BL,
                      float* sBL, float* ssBL, bool matrix, hipStream_t st)
{
    const long long blocks = bhc_blocks(count);
    if (blocks <= 0) return;
    if (matrix || ssBL)
        hipLaunchKernelGGL(k_poly_iout<3>, dim3((unsigned)blocks), dim3(kBhcBlock), 0, st, P, s,
                           count, BL, sBL, ssBL, matrix ? 1 : 0);
    else if (sBL)
        hipLaunchKernelGGL(k_poly_iout<2>, dim3((unsigned)blocks), dim3(kBhcBlock), 0, st, P, s,
                           count, BL, sBL, ssBL, 0);
    else
        hipLaunchKernelGGL(k_poly_iout<1>, dim3((unsigned)blocks), dim3(kBhcBlock), 0, st, P, s,
                           count, BL, sBL, ssBL, 0);
}

void launch_bhc_alpha_ray(const BhcParams& P, const float* s, float sscale, const float* Imea,
                          long long count, float Ts, float* w, float* hw, double* partials,
                          hipStream_t st)
{
    const long long blocks = bhc_blocks(count);
    if (blocks <= 0) return;
    if (hw)
        hipLaunchKernelGGL(k_bhc_alpha_ray<true>, dim3((unsigned)blocks), dim3(kBhcBlock), 0, st,
                           P, s, Imea, count, sscale, Ts, w, hw, partials);
    else
        hipLaunchKernelGGL(k_bhc_alpha_ray<false>, dim3((unsigned)blocks), dim3(kBhcBlock), 0,
                           st, P, s, Imea, count, sscale, Ts, w, hw, partials);
}

void launch_bhc_ie_ray(const BhcParams& P, const float* s, float sscale, const float* Imea,
                       long long count, double* partials, hipStream_t st)
{
    const long long blocks = bhc_blocks(count);
    if (blocks <= 0) return;
    hipLaunchKernelGGL(k_bhc_ie_ray, dim3((unsigned)blocks), dim3(kBhcBlock), 0, st, P, s, Imea,
                       count, sscale, partials);
}

void launch_bhc_ie_cost(const BhcParams& P, const float* s, float sscale, const float* Imea,
                        long long count, double* partials, hipStream_t st)
{
    const long long blocks = bhc_blocks(count);
    if (blocks <= 0) return;
    hipLaunchKernelGGL(k_bhc_ie_cost, dim3((unsigned)blocks), dim3(kBhcBlock), 0, st, P, s, Imea,
                       count, sscale, partials);
}

void launch_reduce_partials(const double* partials, long long nparts, int nvals, double* out,
                            hipStream_t st)
{
    if (nvals <= 0) return;
    hipLaunchKernelGGL(k_reduce_partials, dim3(nvals), dim3(kBhcBlock), 0, st, partials, nparts,
                       nvals, out);
}

void launch_unpack_ie(const double* red, int E, double* f, double* g, double* H, hipStream_t st)
{
    hipLaunchKernelGGL(k_unpack_ie, dim3(1), dim3(256), 0, st, red, E, f, g, H);
}

}  // namespace ctprj
