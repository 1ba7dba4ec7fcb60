// Device functions of the polychromatic spectrum model shared by the beam-hardening
// kernels and the projector's fused forward store: the Laplace transform of the
// mass-attenuation spectrum in the b1 / b0 / dis bases (bhc/Spline.m:122-302) and the
// per-ray likelihood term and gradient weights of bhc/poissLAlpha.m and
// bhc/gaussLAlpha.m.
#pragma once

#include "bhc_internal.h"

namespace ctprj {

// Moments M_n(z) = int_0^1 u^n exp(-z u) du, n = 0..3, of one spline segment.
// |z| <= 2: the series sum_k c_nk z^k, c_nk = (-1)^k / (k! (n+k+1)), 14 terms in Horner
// form (truncation below 1.2e-8 of M_n; FP32 evaluation error below 6e-7 relative);
// otherwise the recurrence M_0 = (1-e^-z)/z, M_n = (n M_{n-1} - e^-z)/z, whose error
// growth is at most n/|z| <= 1.5 per step there.
constexpr int kMomentTerms = 14;

constexpr double moment_coef(int n, int k)
{
    double f = 1.0;
    for (int i = 2; i <= k; ++i) f *= i;
    return ((k & 1) ? -1.0 : 1.0) / (f * (n + k + 1));
}

struct MomentTable {
    float c[4][kMomentTerms];
};

constexpr MomentTable make_moment_table()
{
    MomentTable t{};
    for (int n = 0; n < 4; ++n)
        for (int k = 0; k < kMomentTerms; ++k) t.c[n][k] = (float)moment_coef(n, k);
    return t;
}

__constant__ constexpr MomentTable kMoment = make_moment_table();

struct Moments {
    float m0, m1, m2, m3;
};

// The first NM (1..4) moments by their first TERMS series terms (|z| <= 2: 14 terms;
// |z| <= 1: 11 terms, truncation below 2e-9 of M_n).
template <int NM, int TERMS = kMomentTerms>
__device__ __forceinline__ Moments moments_series(float z)
{
    float m[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int n = 0; n < NM; ++n) {
        float a = kMoment.c[n][TERMS - 1];
#pragma unroll
        for (int k = TERMS - 2; k >= 0; --k) a = fmaf(a, z, kMoment.c[n][k]);
        m[n] = a;
    }
    return Moments{m[0], m[1], m[2], m[3]};
}

__device__ __forceinline__ Moments moments(float z)
{
    if (fabsf(z) <= 2.0f) return moments_series<4>(z);
    Moments M;
    const float e = expf(-z);
    const float iz = 1.0f / z;
    M.m0 = (1.0f - e) * iz;
    M.m1 = (M.m0 - e) * iz;
    M.m2 = (2.0f * M.m1 - e) * iz;
    M.m3 = (3.0f * M.m2 - e) * iz;
    return M;
}

// Weighted moments W_n = w M_n(z) of b1 segment j, w = Delta_j e_j, z = Delta_j s, from
// the exponentials e_j = exp(-k_j s) and e_{j+1} = exp(-k_{j+1} s) of its two knots
// (e^-z = e_{j+1}/e_j): the series for small |z|, else the recurrence
// W_0 = Delta_j (e_j - e_{j+1})/z, W_n = (n W_{n-1} - Delta_j e_{j+1})/z, which needs no
// exponential of its own.  NM = moments needed (2..4).  Up to M_2 the series switches
// at |z| = 1 (11 terms; the recurrence then stays within 1.1e-6 of M_2), M_3 (second
// derivatives) keeps the switch at 2 (14 terms), where its recurrence error is 8e-7.
template <int NM>
__device__ __forceinline__ Moments weighted_moments(float z, float dj, float ej, float ej1)
{
    constexpr float kSwitch = NM >= 4 ? 2.0f : 1.0f;
    constexpr int kTerms = NM >= 4 ? kMomentTerms : 11;
    if (fabsf(z) <= kSwitch) {
        const Moments M = moments_series<NM, kTerms>(z);
        const float w = dj * ej;
        return Moments{w * M.m0, w * M.m1, w * M.m2, w * M.m3};
    }
    Moments W{0.0f, 0.0f, 0.0f, 0.0f};
    const float iz = 1.0f / z, de = dj * ej1;
    W.m0 = dj * (ej - ej1) * iz;
    W.m1 = (W.m0 - de) * iz;
    if (NM >= 3) W.m2 = (2.0f * W.m1 - de) * iz;
    if (NM >= 4) W.m3 = (3.0f * W.m2 - de) * iz;
    return W;
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
// bhc/Spline.m:154-201 without their small-s cancellation.  The products w_j M_n are
// formed by weighted_moments from the knots' exponentials, one exp per knot.
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
    float ej = expf(-P.knot[0] * s);
    for (int j = 0; j <= P.E; ++j) {
        const float kj = P.knot[j], dj = P.delta[j];
        const float ej1 = expf(-P.knot[j + 1] * s);
        const Moments W = weighted_moments<WANT + 1>(dj * s, dj, ej, ej1);
        ej = ej1;
        const float c0 = W.m0 - W.m1;
        if (j >= 1) {  // right half of basis j-1
            float rp = c0, rd = 0.0f, rdd = 0.0f;
            if (WANT >= 2) rd = -fmaf(kj, c0, dj * (W.m1 - W.m2));
            if (WANT >= 3)
                rdd = fmaf(kj * kj, c0, fmaf(2.0f * kj * dj, W.m1 - W.m2, dj * dj * (W.m2 - W.m3)));
            f(j - 1, cp + rp, cd + rd, cdd + rdd);
        }
        if (j < P.E) {  // left half of basis j
            cp = W.m1;
            if (WANT >= 2) cd = -fmaf(kj, W.m1, dj * W.m2);
            if (WANT >= 3) cdd = fmaf(kj * kj, W.m1, fmaf(2.0f * kj * dj, W.m2, dj * dj * W.m3));
        }
    }
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

// One ray of the alpha pass: spectrum sums BL, sBL (and ssBL) at the ray value si, the
// FP64 NLL term (returned) and the weights
//   Poisson  (bhc/poissLAlpha.m):  w = Ts (1 - Imea/BL) sBL,  hw = (sBL/BL)^2 Imea + ssBL (1 - Imea/BL)
//   Gaussian (bhc/gaussLAlpha.m):  w = Ts Err sBL/BL,         hw = (1 - Err)(sBL/BL)^2 + Err ssBL/BL
// with Err = log(BL/Imea).  *hw is written only when HESS.
template <bool HESS>
__device__ __forceinline__ double bhc_ray(const BhcParams& P, float si, float im, float Ts,
                                          float* w, float* hw)
{
    float b = 0.0f, bd = 0.0f, bdd = 0.0f;
    spline_basis<HESS ? 3 : 2>(P, si, [&](int j, float p, float d, float dd) {
        b = fmaf(P.Ie[j], p, b);
        bd = fmaf(P.Ie[j], d, bd);
        if (HESS) bdd = fmaf(P.Ie[j], dd, bdd);
    });
    if (P.noise == CTPRJ_NOISE_GAUSSIAN) {
        const double err = gauss_err((double)b, (double)im);
        const float e = (float)err, q = bd / b;
        *w = Ts * (e * q);  // Phit(Err.*sBL./BL), bhc/gaussLAlpha.m:16
        if (HESS) *hw = fmaf(1.0f - e, q * q, e * (bdd / b));  // :18
        return 0.5 * err * err;
    }
    const float r = 1.0f - im / b;
    *w = Ts * (r * bd);  // Phit((1-Imea./BL).*sBL) with Phit = Ts*A^T
    if (HESS) {
        const float q = bd / b;
        *hw = fmaf(q * q, im, bdd * r);  // bhc/poissLAlpha.m:15
    }
    return nll_term((double)b, (double)im);
}

}  // namespace ctprj
