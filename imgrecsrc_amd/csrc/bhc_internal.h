// Internal declarations of the beam-hardening kernels of libctprj.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/ctprj.h"

namespace ctprj {

// Kernel arguments of the spectrum model: E hat bases on the E+2 knots.
struct BhcParams {
    int E;
    float knot[CTPRJ_MAX_E + 2];    // k_0 .. k_{E+1}
    float delta[CTPRJ_MAX_E + 1];   // k_{j+1} - k_j (rounded from double)
    float Ie[CTPRJ_MAX_E];          // basis coefficients (FP32 per-ray sums)
    double Ied[CTPRJ_MAX_E];        // basis coefficients (FP64 I-epsilon sums)
};

constexpr int kBhcBlock = 256;

// values reduced by the I-epsilon pass: E(E+1)/2 Hessian entries, E gradient, 1 NLL
__host__ __device__ inline int bhc_ie_nvals(int E) { return E * (E + 1) / 2 + E + 1; }

void launch_b1_iout(const BhcParams& P, const float* s, long long count, float* BL, float* sBL,
                    float* ssBL, bool matrix, hipStream_t st);
// per-ray alpha pass: w = Ts (1 - Imea/BL) sBL (may alias s), hw, FP64 NLL partials
// (ray values are sscale * s[i])
void launch_bhc_alpha_ray(const BhcParams& P, const float* s, float sscale, const float* Imea,
                          long long count, float Ts, float* w, float* hw, double* partials,
                          hipStream_t st);
// per-ray I-epsilon pass: per work-group partial [H upper, g, f]
void launch_bhc_ie_ray(const BhcParams& P, const float* s, float sscale, const float* Imea,
                       long long count, double* partials, hipStream_t st);
// out[v] = sum_p partials[p*nvals + v] in a fixed order
void launch_reduce_partials(const double* partials, long long nparts, int nvals, double* out,
                            hipStream_t st);
// unpack [H upper, g, f] -> f, g, H (full)
void launch_unpack_ie(const double* red, int E, double* f, double* g, double* H, hipStream_t st);

inline long long bhc_blocks(long long count) { return (count + kBhcBlock - 1) / kBhcBlock; }

}  // namespace ctprj
