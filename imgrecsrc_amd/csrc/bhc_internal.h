// Internal declarations of the beam-hardening kernels of libctprj.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/ctprj.h"

namespace ctprj {

// Kernel arguments of the spectrum model (bhc/Spline.m): E basis functions of the
// mass-attenuation spectrum and the noise model of the likelihood.
//   b1  (CTPRJ_BASIS_B1):  hats on the E+2 knots k_0..k_{E+1}, delta[j] = k_{j+1} - k_j
//   b0  (CTPRJ_BASIS_B0):  boxes on the E+1 knots k_0..k_E,     delta[j] = k_{j+1} - k_j
//   dis (CTPRJ_BASIS_DIS): nodes k_1..k_E of the E+2 knots,      delta[j] = (k_{j+2} - k_j)/2
struct BhcParams {
    int E;
    int basis;                      // CTPRJ_BASIS_*
    int noise;                      // CTPRJ_NOISE_*
    float knot[CTPRJ_MAX_E + 2];
    float delta[CTPRJ_MAX_E + 1];   // (rounded from double)
    float Ie[CTPRJ_MAX_E];          // basis coefficients (FP32 per-ray sums)
    double Ied[CTPRJ_MAX_E];        // basis coefficients (FP64 I-epsilon sums)
};

// The alpha pass fused into the symmetric forward projector's final store: each sinogram
// entry is stored as its weight w (ray value Ts times the ray sum) and its FP64 NLL term
// goes into the storing work-group's partial, partials[part_off + work-group].
struct BhcStore {
    BhcParams P;
    const float* Imea;  // measured intensities, plan sinogram layout
    double* partials;
    float Ts;
    int part_off;
};

// Validates the knots of `basis` (E+2, or E+1 for b0, increasing) and fills P.
int make_bhc_params(const double* knots, int E, const double* Ie, int basis, int noise,
                    BhcParams* P);

constexpr int kBhcBlock = 256;

// values reduced by the I-epsilon pass: E(E+1)/2 Hessian entries, E gradient, 1 NLL
__host__ __device__ inline int bhc_ie_nvals(int E) { return E * (E + 1) / 2 + E + 1; }

void launch_poly_iout(const BhcParams& P, const float* s, long long count, float* BL,
                      float* sBL, float* ssBL, bool matrix, hipStream_t st);
// per-ray alpha pass: gradient weight w (may alias s), Hessian weight hw, FP64 NLL
// partials (ray values are sscale * s[i]); Poisson: w = Ts (1 - Imea/BL) sBL,
// Gaussian: w = Ts log(BL/Imea) sBL/BL
void launch_bhc_alpha_ray(const BhcParams& P, const float* s, float sscale, const float* Imea,
                          long long count, float Ts, float* w, float* hw, double* partials,
                          hipStream_t st);
// per-ray I-epsilon pass: per work-group partial [H upper, g, f], bhc_ie_groups(count)
// = G work-groups, value v of group p at partials[v*G + c(p)], c a permutation of the
// groups; with phi (launch_bhc_ie_basis's matrix of the same s and spectrum basis) the
// basis values are read instead of evaluated, with the same results
void launch_bhc_ie_ray(const BhcParams& P, const float* s, float sscale, const float* Imea,
                       long long count, double* partials, hipStream_t st,
                       const float* phi = nullptr);
// cost only of the I-epsilon pass: one FP64 partial per work-group (phi as above)
void launch_bhc_ie_cost(const BhcParams& P, const float* s, float sscale, const float* Imea,
                        long long count, double* partials, hipStream_t st,
                        const float* phi = nullptr);
// the I-epsilon pass's basis matrix: value j of ray i at phi[j * count + i]
void launch_bhc_ie_basis(const BhcParams& P, const float* s, float sscale, long long count,
                         float* phi, hipStream_t st);
// largest basis matrix ctprj_ie_solve keeps for its passes (beyond it they evaluate the
// basis each time)
constexpr size_t kIeBasisCacheMax = (size_t)1 << 30;
// out[v] = sum_p partials[v*nparts + p] in a fixed order, v < nvals
void launch_reduce_partials(const double* partials, long long nparts, int nvals, double* out,
                            hipStream_t st);
// the I-epsilon pass's partials [H upper, g, f] reduced as launch_reduce_partials does,
// straight into f, g and H (full); NULL outputs are skipped
void launch_reduce_ie(const double* partials, long long nparts, int E, double* f, double* g,
                      double* H, hipStream_t st);

inline long long bhc_blocks(long long count) { return (count + kBhcBlock - 1) / kBhcBlock; }

// Work-groups of the I-epsilon pass: at most kIeMaxGroups, each taking the same number
// of kBhcBlock-ray tiles (but the last).
constexpr long long kIeMaxGroups = 4096;
inline long long bhc_ie_groups(long long count)
{
    const long long tiles = bhc_blocks(count);
    if (tiles <= 0) return 0;
    const long long per = (tiles + kIeMaxGroups - 1) / kIeMaxGroups;
    return (tiles + per - 1) / per;
}

// Doubles of reduction partials a plan of `rays` rays holds for its beam-hardening calls:
// one per kBhcBlock rays (NLL and cost passes) or the I-epsilon pass's groups x values.
inline long long bhc_red_doubles(long long rays)
{
    const long long a = bhc_blocks(rays), b = bhc_ie_groups(rays) * bhc_ie_nvals(CTPRJ_MAX_E);
    return a > b ? a : b;
}

}  // namespace ctprj
