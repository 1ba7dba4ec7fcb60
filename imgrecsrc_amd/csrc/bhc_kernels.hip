// Polychromatic beam-hardening model on gfx950: the Laplace transform of the
// mass-attenuation spectrum in the b1 / b0 / dis bases (bhc/Spline.m:122-302), the
// per-ray likelihood and gradient weights of bhc/poissLAlpha.m and bhc/gaussLAlpha.m,
// and the I-epsilon gradient / Hessian of bhc/poissLI.m and bhc/gaussLI.m.  Per-ray spectral values are FP32 and recomputed wherever needed (no
// ray x basis matrix is stored); the negative log-likelihood and every I-epsilon
// reduction are FP64 with a fixed summation order: per work-group partials, then one
// ordered pass.
#include "bhc_device.h"

namespace ctprj {

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

// Sum of v over a work-group of kBhcBlock threads: a butterfly over each wave's 64 lanes
// (register exchange, no LDS tree), then the waves' sums in wave order.  The summation
// order is fixed, so the result is deterministic.  Every thread returns the total.
__device__ __forceinline__ double block_sum(double v, double* sh)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int wave = (int)threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[wave] = v;
    __syncthreads();
    double r = sh[0];
#pragma unroll
    for (int w = 1; w < kBhcBlock / 64; ++w) r += sh[w];
    __syncthreads();
    return r;
}

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
        float wi, hwi;
        f = bhc_ray<HESS>(P, sscale * s[i], Imea[i], Ts, &wi, &hwi);
        if (w) w[i] = wi;
        if (HESS && hw) hw[i] = hwi;
    }
    const double tot = block_sum(f, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// ------------------------------------------------------------ I-epsilon pass ---

// A work-group takes tiles of kBhcBlock rays (tile = group, group + G, ...): the basis
// values phi_ij (FP32) and the ray weights q_i, r_i, f_i (FP64) of a tile go to LDS, and
// the group adds the tile's contribution to its sums [H upper triangle (row-major j<=k),
// g, f]:
//   H_jk = sum_i phi_ij phi_ik q_i,  g_j = sum_i phi_ij r_i,  f = sum_i f_i,
//   Ir_i = sum_j Ie_j phi_ij (FP64), with
//   Poisson  (bhc/poissLI.m):  q = Imea/Ir^2,       r = 1 - Imea/Ir,  f_i = nll(Ir, Imea)
//   Gaussian (bhc/gaussLI.m):  q = (1 - Err)/Ir^2,  r = Err/Ir,       f_i = Err^2/2.
// Register blocking: the outputs are cut into 4x4 blocks of H (upper block triangle)
// and 4-entry blocks of g (the last one also carries f); the S slots (a power of two at
// least the number of blocks) take L = kBhcBlock / S consecutive threads each, thread
// lane l of a slot summing the rays l, l + L, ... of every tile, so each LDS read of 4
// basis values feeds 4 or 16 FP64 FMAs.  At the end the L lanes of a slot add their sums
// by a butterfly inside the wave (a fixed order) and lane 0 writes them.  The basis rows
// are E values padded to an odd number of 16-byte units (the L lanes of a slot read L
// consecutive rows at the same column without bank conflicts), in dynamic LDS sized for
// E: at E = 20, 30 KiB per group, so five groups fit a CU.  The grid is capped
// (bhc_ie_groups), so the partials take at most kIeMaxGroups * nvals doubles whatever the
// ray count.
__host__ __device__ inline int ie_blocks(int E) { return (E + 3) / 4; }
__host__ __device__ inline int ie_slots(int E)
{
    const int nb = ie_blocks(E);
    return nb * (nb + 1) / 2 + nb;
}
// floats per LDS basis row: 4 per 4-block, an odd number of 16-byte units
__host__ __device__ inline int ie_row(int E) { return 4 * (ie_blocks(E) | 1); }
__host__ __device__ inline size_t ie_lds_bytes(int E)
{
    return sizeof(float) * kBhcBlock * (size_t)ie_row(E) + 3 * sizeof(double) * kBhcBlock;
}

// Five waves per SIMD (90 VGPRs instead of 104): five groups per CU, so the 1026 tiles of
// a 512^2 fan sinogram run in one round instead of two.
__global__ __launch_bounds__(kBhcBlock) __attribute__((amdgpu_waves_per_eu(5)))
void k_bhc_ie_ray(BhcParams P, const float* __restrict__ s, const float* __restrict__ Imea,
                  long long count, float sscale, double* __restrict__ partials)
{
    extern __shared__ __attribute__((aligned(16))) char ie_lds[];
    const int E = P.E, row = ie_row(E);
    float* phi = reinterpret_cast<float*>(ie_lds);  // [kBhcBlock][row]
    double* q = reinterpret_cast<double*>(ie_lds + sizeof(float) * kBhcBlock * row);
    double* r = q + kBhcBlock;
    double* ft = r + kBhcBlock;
    const int t = threadIdx.x;
    const int nb = ie_blocks(E), nhb = nb * (nb + 1) / 2, nslot = nhb + nb;
    int S = 4;  // at most 64 lanes per slot: a slot's lanes stay inside one wave
    while (S < nslot) S <<= 1;  // nslot <= 44 for E <= 32, so S <= 64
    const int L = kBhcBlock / S, slot = t / L, lane = t % L;
    int bi = 0, bj = 0;
    if (slot < nhb) {
        int rem = slot;  // block row bi of the upper triangle holds nb-bi blocks
        while (rem >= nb - bi) { rem -= nb - bi; ++bi; }
        bj = bi + rem;
    } else {
        bj = slot - nhb;
    }
    const bool last = slot >= nhb && bj == nb - 1;
    double acc[17];
#pragma unroll
    for (int v = 0; v < 17; ++v) acc[v] = 0.0;
    const long long ntiles = (count + kBhcBlock - 1) / kBhcBlock;
    for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const long long i = tile * kBhcBlock + t;
        float* my = phi + t * row;
        // spline_basis writes columns 0..E-1 of a ray's row; the padding columns feed only
        // block entries beyond E, which are never stored, so they are left as they are
        if (i < count) {
            const float si = sscale * s[i];
            const double im = (double)Imea[i];
            double ir = 0.0;
            spline_basis<1>(P, si, [&](int j, float p, float, float) {
                my[j] = p;
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
        } else {  // past the last ray: finite zeros (times a zero weight)
            for (int j = 0; j < row; ++j) my[j] = 0.0f;
            q[t] = 0.0;
            r[t] = 0.0;
            ft[t] = 0.0;
        }
        __syncthreads();
        if (slot < nhb) {
            for (int m = lane; m < kBhcBlock; m += L) {
                const float4 a = *reinterpret_cast<const float4*>(phi + m * row + 4 * bi);
                const float4 b = *reinterpret_cast<const float4*>(phi + m * row + 4 * bj);
                const double w = q[m];
                const double ad[4] = {(double)a.x * w, (double)a.y * w, (double)a.z * w,
                                      (double)a.w * w};
                const double bd[4] = {(double)b.x, (double)b.y, (double)b.z, (double)b.w};
#pragma unroll
                for (int x = 0; x < 4; ++x)
#pragma unroll
                    for (int y = 0; y < 4; ++y) acc[4 * x + y] = fma(ad[x], bd[y], acc[4 * x + y]);
            }
        } else if (slot < nslot) {
            for (int m = lane; m < kBhcBlock; m += L) {
                const float4 a = *reinterpret_cast<const float4*>(phi + m * row + 4 * bj);
                const double w = r[m];
                acc[0] = fma((double)a.x, w, acc[0]);
                acc[1] = fma((double)a.y, w, acc[1]);
                acc[2] = fma((double)a.z, w, acc[2]);
                acc[3] = fma((double)a.w, w, acc[3]);
                if (last) acc[16] += ft[m];
            }
        }
        __syncthreads();  // every basis value of the tile read
    }
    // the slot's L lane sums: a butterfly over its consecutive lanes (inside one wave)
    for (int off = L >> 1; off > 0; off >>= 1)
#pragma unroll
        for (int v = 0; v < 17; ++v) acc[v] += __shfl_xor(acc[v], off, 64);
    if (lane != 0 || slot >= nslot) return;
    const int nh = E * (E + 1) / 2, nv = nh + E + 1;
    double* out = partials + (long long)blockIdx.x * nv;
    if (slot < nhb) {
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                const int j = 4 * bi + x, k = 4 * bj + y;  // H_jk, stored for j <= k < E
                if (j <= k && k < E) out[j * E - j * (j - 1) / 2 + (k - j)] = acc[4 * x + y];
            }
    } else {
#pragma unroll
        for (int x = 0; x < 4; ++x)
            if (4 * bj + x < E) out[nh + 4 * bj + x] = acc[x];
        if (last) out[nv - 1] = acc[16];
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
    const long long groups = bhc_ie_groups(count);
    if (groups <= 0) return;
    hipLaunchKernelGGL(k_bhc_ie_ray, dim3((unsigned)groups), dim3(kBhcBlock), ie_lds_bytes(P.E), st,
                       P, s, Imea, count, sscale, partials);
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
