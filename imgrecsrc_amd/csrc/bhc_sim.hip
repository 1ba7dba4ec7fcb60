// Polychromatic measurement simulator (bhc/genBeamHarden.m:53-79) on gfx950.  For K
// materials with projections y_k = density_k * (A x_k), ray t measures
//   I(t) = sum_l exp(-Ts * sum_k y_k(t) kappa_k(l)) * iota_l * dEps_l
// (FP64, with the reference's evaluation order and no fused multiply-adds), Ts is
// reduced from 1 by factors of 0.8 until min I / max I >= scale, and the result is
// normalised by max I.  One thread per ray; the energy table is read uniformly by every
// lane of a wave.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "bhc_internal.h"
#include "ctprj_internal.h"

namespace ctprj {

constexpr int kSimBlock = 256;
// Ts = 0.8^j is zero in FP64 beyond j = 3340, where I is constant and min/max = 1, so the
// search always ends before this many trials for scale <= 1.
constexpr int kSimMaxTrials = 3400;

// tab: L rows of [kappa_0 .. kappa_{K-1}, iota, dEps] (K + 2 doubles per energy).
__global__ __launch_bounds__(kSimBlock) void k_sim_measure(const float* __restrict__ y,
                                                           long long count, int K,
                                                           const double* __restrict__ dens,
                                                           const double* __restrict__ tab,
                                                           int L, double ts,
                                                           double* __restrict__ I,
                                                           double* __restrict__ partials)
{
    __shared__ double smax[kSimBlock], smin[kSimBlock];
    const long long t = (long long)blockIdx.x * kSimBlock + threadIdx.x;
    double hi = -INFINITY, lo = INFINITY;
    if (t < count) {
        double yk[CTPRJ_SIM_MAX_K];
#pragma unroll
        for (int k = 0; k < CTPRJ_SIM_MAX_K; ++k)
            yk[k] = k < K ? __dmul_rn((double)y[(long long)k * count + t], dens[k]) : 0.0;
        const int row = K + 2;
        double acc = 0.0;
        for (int l = 0; l < L; ++l) {
            const double* e = tab + (long long)l * row;
            double pa = 0.0;
#pragma unroll
            for (int k = 0; k < CTPRJ_SIM_MAX_K; ++k)
                if (k < K) pa = __dadd_rn(pa, __dmul_rn(yk[k], e[k]));
            pa = __dmul_rn(pa, ts);
            acc = __dadd_rn(acc, __dmul_rn(__dmul_rn(exp(-pa), e[K]), e[K + 1]));
        }
        I[t] = acc;
        hi = lo = acc;
    }
    const int i = threadIdx.x;
    smax[i] = hi;
    smin[i] = lo;
    __syncthreads();
    for (int off = kSimBlock / 2; off > 0; off >>= 1) {
        if (i < off) {
            smax[i] = fmax(smax[i], smax[i + off]);
            smin[i] = fmin(smin[i], smin[i + off]);
        }
        __syncthreads();
    }
    if (i == 0) {
        partials[2 * blockIdx.x] = smax[0];
        partials[2 * blockIdx.x + 1] = smin[0];
    }
}

__global__ __launch_bounds__(kSimBlock) void k_sim_normalise(const double* __restrict__ I,
                                                             long long count, double m,
                                                             float* __restrict__ out)
{
    const long long t = (long long)blockIdx.x * kSimBlock + threadIdx.x;
    if (t < count) out[t] = (float)(I[t] / m);
}

static int sim_scratch(Plan* p, size_t bytes, char** base)
{
    if (p->sim_cap < bytes) {
        if (p->d_sim) (void)hipFree(p->d_sim);
        p->d_sim = nullptr;
        p->sim_cap = 0;
        hipError_t e = hipMalloc((void**)&p->d_sim, bytes);
        if (e != hipSuccess)
            return fail(CTPRJ_ENOMEM, "simulator workspace: %s", hipGetErrorString(e));
        p->sim_cap = bytes;
    }
    *base = (char*)p->d_sim;
    return 0;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Validates the spectrum, builds the energy table (dEps from bhc/genBeamHarden.m:53-54)
// and runs the Ts search on projections d_y (K x rays floats).
static int sim_run(Plan* p, const float* d_y, int K, const double* density, const double* kappa,
                   const double* epsilon, const double* iota, int L, double scale, float* d_Imea,
                   double* Ts, double* Imax, char* ws, hipStream_t st)
{
    const long long rays = p->sino_elems;
    const long long blocks = (rays + kSimBlock - 1) / kSimBlock;
    const int row = K + 2;
    double* d_I = (double*)ws;
    double* d_part = (double*)(ws + align256(sizeof(double) * rays));
    double* d_tab = (double*)((char*)d_part + align256(sizeof(double) * 2 * blocks));
    double* d_dens = d_tab + (size_t)row * L;

    std::vector<double> tab((size_t)row * L);
    for (int l = 0; l < L; ++l) {
        // energy bin edges: eps_1, midpoints, eps_L
        const double a = l == 0 ? epsilon[0] : (epsilon[l - 1] + epsilon[l]) / 2.0;
        const double b = l == L - 1 ? epsilon[L - 1] : (epsilon[l] + epsilon[l + 1]) / 2.0;
        for (int k = 0; k < K; ++k) tab[(size_t)l * row + k] = kappa[(size_t)k * L + l];
        tab[(size_t)l * row + K] = iota[l];
        tab[(size_t)l * row + K + 1] = b - a;
    }
    hipError_t e = hipMemcpyAsync(d_tab, tab.data(), sizeof(double) * tab.size(),
                                  hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_dens, density, sizeof(double) * K, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);  // tab is a host temporary
    if (e != hipSuccess) return hip_fail(e, "simulator table upload");

    std::vector<double> part((size_t)2 * blocks);
    double ts = 1.0, mx = 0.0;
    for (int trial = 0;; ++trial) {
        if (trial == kSimMaxTrials)
            return fail(CTPRJ_EINVAL, "gen_beam_harden: no Ts reached scale");
        hipLaunchKernelGGL(k_sim_measure, dim3((unsigned)blocks), dim3(kSimBlock), 0, st, d_y,
                           rays, K, d_dens, d_tab, L, ts, d_I, d_part);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(part.data(), d_part, sizeof(double) * part.size(),
                               hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e, "simulator pass");
        double mn = INFINITY;
        mx = -INFINITY;
        for (long long b = 0; b < blocks; ++b) {
            mx = std::fmax(mx, part[2 * b]);
            mn = std::fmin(mn, part[2 * b + 1]);
        }
        if (!(mx > 0.0) || !std::isfinite(mx))
            return fail(CTPRJ_EINVAL, "gen_beam_harden: max intensity %g (iota all zero?)", mx);
        if (mn / mx < scale) ts = ts * 0.8;  // bhc/genBeamHarden.m:71-72
        else break;
    }
    hipLaunchKernelGGL(k_sim_normalise, dim3((unsigned)blocks), dim3(kSimBlock), 0, st, d_I, rays,
                       mx, d_Imea);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "simulator normalise launch");
    *Ts = ts;
    *Imax = mx;
    return 0;
}

static int sim_check(const Plan* p, int K, const double* density, const double* kappa,
                     const double* epsilon, const double* iota, int L, double scale,
                     const float* d_Imea, const double* Ts, const double* Imax)
{
    if (!p || !density || !kappa || !epsilon || !iota || !d_Imea || !Ts || !Imax)
        return fail(CTPRJ_EINVAL, "gen_beam_harden: NULL argument");
    if (K < 1 || K > CTPRJ_SIM_MAX_K)
        return fail(CTPRJ_EINVAL, "gen_beam_harden: K=%d must be in [1, %d]", K, CTPRJ_SIM_MAX_K);
    if (L < 2) return fail(CTPRJ_EINVAL, "gen_beam_harden: need at least 2 energies (L=%d)", L);
    if (!(scale >= 0.0 && scale <= 1.0))
        return fail(CTPRJ_EINVAL, "gen_beam_harden: scale %g must be in [0, 1]", scale);
    for (int l = 0; l < L; ++l) {
        if (!std::isfinite(epsilon[l]) || !std::isfinite(iota[l]))
            return fail(CTPRJ_EINVAL, "gen_beam_harden: energy %d not finite", l);
        if (l && !(epsilon[l] > epsilon[l - 1]))
            return fail(CTPRJ_EINVAL, "gen_beam_harden: energies must increase (at %d)", l);
    }
    for (long long i = 0; i < (long long)K * L; ++i)
        if (!std::isfinite(kappa[i]))
            return fail(CTPRJ_EINVAL, "gen_beam_harden: kappa not finite");
    for (int k = 0; k < K; ++k)
        if (!std::isfinite(density[k]))
            return fail(CTPRJ_EINVAL, "gen_beam_harden: density not finite");
    return 0;
}

static size_t sim_bytes(const Plan* p, int K, int L, bool maps)
{
    const long long rays = p->sino_elems;
    const long long blocks = (rays + kSimBlock - 1) / kSimBlock;
    return align256(sizeof(double) * rays) + align256(sizeof(double) * 2 * blocks) +
           align256(sizeof(double) * ((size_t)(K + 2) * L + K)) +
           (maps ? align256(sizeof(float) * (size_t)K * rays) : 0);
}

}  // namespace ctprj

using namespace ctprj;

struct ctprj_plan : public Plan {};

extern "C" int ctprj_poly_measure(ctprj_plan* p, const float* d_y, int K, const double* density,
                                  const double* kappa, const double* epsilon, const double* iota,
                                  int L, double scale, float* d_Imea, double* Ts, double* Imax,
                                  void* stream)
{
    int rc = sim_check(p, K, density, kappa, epsilon, iota, L, scale, d_Imea, Ts, Imax);
    if (rc) return rc;
    if (!d_y) return fail(CTPRJ_EINVAL, "poly_measure: NULL projections");
    char* ws;
    if ((rc = sim_scratch(p, sim_bytes(p, K, L, false), &ws))) return rc;
    return sim_run(p, d_y, K, density, kappa, epsilon, iota, L, scale, d_Imea, Ts, Imax, ws,
                   (hipStream_t)stream);
}

extern "C" int ctprj_gen_beam_harden(ctprj_plan* p, const float* d_maps, int K,
                                     const double* density, const double* kappa,
                                     const double* epsilon, const double* iota, int L,
                                     double scale, float* d_Imea, double* Ts, double* Imax,
                                     void* stream)
{
    int rc = sim_check(p, K, density, kappa, epsilon, iota, L, scale, d_Imea, Ts, Imax);
    if (rc) return rc;
    if (!d_maps) return fail(CTPRJ_EINVAL, "gen_beam_harden: NULL density maps");
    char* ws;
    const size_t total = sim_bytes(p, K, L, true);
    if ((rc = sim_scratch(p, total, &ws))) return rc;
    float* d_y = (float*)(ws + sim_bytes(p, K, L, false));
    hipStream_t st = (hipStream_t)stream;
    for (int k = 0; k < K; ++k)  // y_k = A x_k (the density is applied in FP64 per ray)
        if ((rc = p->launch_forward(d_maps + (long long)k * p->img_elems,
                                    d_y + (long long)k * p->sino_elems, st, false)))
            return rc;
    return sim_run(p, d_y, K, density, kappa, epsilon, iota, L, scale, d_Imea, Ts, Imax, ws, st);
}
