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
// values phi_ij (FP32) and the ray weights q_i, r_i (FP64) of a tile go to LDS, and the
// group adds the tile's contribution to its sums [H upper triangle (row-major j<=k), g,
// f]:
//   H_jk = sum_i phi_ij phi_ik q_i,  g_j = sum_i phi_ij r_i,  f = sum_i f_i,
//   Ir_i = sum_j Ie_j phi_ij (FP64), with
//   Poisson  (bhc/poissLI.m):  q = Imea/Ir^2,       r = 1 - Imea/Ir,  f_i = nll(Ir, Imea)
//   Gaussian (bhc/gaussLI.m):  q = (1 - Err)/Ir^2,  r = Err/Ir,       f_i = Err^2/2.
// Register blocking: H is cut into 4x4 blocks (upper block triangle); the S slots (a
// power of two at least the number of blocks, at least 4) take L = kBhcBlock / S
// consecutive lanes of one wave each, lane l of a slot summing the rays l, l + L, ... of
// every tile, so each LDS read of 4 + 4 basis values feeds 16 FP64 FMAs.  The slots of
// the diagonal blocks also sum g over their 4 entries (the other slots run the same
// instructions and drop those sums, so that no wave splits into two paths), and every
// thread sums f over its own rays (in LDS).  At E = 20 the 15 blocks take 15 slots of 16 lanes:
// each of the four waves holds four slots (the last one three) and runs 16 iterations
// per tile.  At the end the L lanes of a slot add their sums in the order of an XOR
// butterfly (offsets L/2, ..., 1), the offsets below 16 as DPP row shifts that leave the
// total in the slot's last lane, which writes it; f is added by block_sum.  The basis
// rows are E values padded to an odd number of 16-byte units (consecutive rows at one
// column spread over the LDS banks), in dynamic LDS sized for E: at E = 20, 26 KiB per
// group.  The grid is capped (bhc_ie_groups), so the partials take at most
// kIeMaxGroups * nvals doubles whatever the ray count; they are stored value-major (value
// v of group p at v * G + c(p), c ordering the groups XCD by XCD), so the reduction reads
// each value's G partials contiguously.
__host__ __device__ inline int ie_blocks(int E) { return (E + 3) / 4; }
// floats per LDS basis row: 4 per 4-block, an odd number of 16-byte units
__host__ __device__ inline int ie_row(int E) { return 4 * (ie_blocks(E) | 1); }
__host__ __device__ inline size_t ie_lds_bytes(int E)
{
    return sizeof(float) * kBhcBlock * (size_t)ie_row(E) + 3 * sizeof(double) * kBhcBlock;
}

// a + (a from the lane OFF below, in the same row of 16 lanes), for FP64 a; lanes with
// no such source add +0 (bound_ctrl; their sums are not used)
template <int OFF>
__device__ __forceinline__ double add_from_below(double a)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(a);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, 0x110 + OFF, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), 0x110 + OFF, 0xf, 0xf, true);
    return a + __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                                (unsigned)lo));
}

// q_i, r_i (stored in *qr) and f_i (returned) of one ray of the I-epsilon pass from Ir_i
// and Imea_i, with one FP64 division: Poisson q = (Imea/Ir)/Ir, r = 1 - Imea/Ir,
// f = (Ir - Imea) + Imea log(Imea/Ir) (bhc/poissLI.m; f = Ir where Imea = 0); Gaussian
// (bhc/gaussLI.m) q = (1 - Err)/Ir^2, r = Err/Ir, f = Err^2/2 with Err = log(Ir/Imea).
__device__ __forceinline__ double ie_ray_terms(int noise, double ir, double im, double2* qr)
{
    const double rinv = 1.0 / ir;
    if (noise == CTPRJ_NOISE_GAUSSIAN) {
        const double err = gauss_err(ir, im);
        if (qr) *qr = make_double2((1.0 - err) * rinv * rinv, err * rinv);
        return 0.5 * err * err;
    }
    const double ratio = im * rinv;
    if (qr) *qr = make_double2(ratio * rinv, 1.0 - ratio);
    return (im > 0.0) ? (ir - im) + im * log(ratio) : ir;
}

// One tile of the pass: the basis rows, q_i, r_i of rays tile * kBhcBlock + t into LDS,
// f_i added to the thread's LDS sum.  CACHED: the basis rows are read from phi_g (value
// j of ray i at phi_g[j * count + i], written by k_bhc_ie_basis with the same
// spline_basis) instead of evaluated, so every sum is the same bit for bit.
template <bool CACHED>
__device__ __forceinline__ void ie_tile_rows(const BhcParams& P, const float* __restrict__ s,
                                             const float* __restrict__ Imea,
                                             const float* __restrict__ phi_g, long long count,
                                             float sscale, long long tile, int row, float* phi,
                                             double2* qr, double* fs)
{
    const int t = threadIdx.x;
    const long long i = tile * kBhcBlock + t;
    float* my = phi + t * row;
    // spline_basis writes columns 0..E-1 of a ray's row; the padding columns feed only
    // block entries beyond E, which are never stored, so they are left as they are
    if (i < count) {
        const double im = (double)Imea[i];
        double ir = 0.0;
        if constexpr (CACHED) {
            for (int j = 0; j < P.E; ++j) {
                const float p = phi_g[j * count + i];
                my[j] = p;
                ir = fma(P.Ied[j], (double)p, ir);
            }
        } else {
            spline_basis<1>(P, sscale * s[i], [&](int j, float p, float, float) {
                my[j] = p;
                ir = fma(P.Ied[j], (double)p, ir);
            });
        }
        fs[t] += ie_ray_terms(P.noise, ir, im, &qr[t]);
    } else {  // past the last ray: finite zeros (times a zero weight)
        for (int j = 0; j < row; ++j) my[j] = 0.0f;
        qr[t] = make_double2(0.0, 0.0);
    }
}

// One tile's contribution to a slot lane's sums: block (bi, bj) of H in acc[0..15], g of
// block bj in acc[16..19].
__device__ __forceinline__ void ie_tile_sums(int lane, int L, int row, int bi, int bj,
                                             const float* phi, const double2* qr, double* acc)
{
    for (int m = lane; m < kBhcBlock; m += L) {
        const float4 a = *reinterpret_cast<const float4*>(phi + m * row + 4 * bi);
        const float4 b = *reinterpret_cast<const float4*>(phi + m * row + 4 * bj);
        const double2 w = qr[m];
        const double ad[4] = {(double)a.x * w.x, (double)a.y * w.x, (double)a.z * w.x,
                              (double)a.w * w.x};
        const double bd[4] = {(double)b.x, (double)b.y, (double)b.z, (double)b.w};
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) acc[4 * x + y] = fma(ad[x], bd[y], acc[4 * x + y]);
        // g of block bj (kept by the diagonal slots, where block bj is block bi)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[16 + y] = fma(bd[y], w.y, acc[16 + y]);
    }
}

// ONE: one tile per group (gridDim.x = tiles), so the sums start after the tile's basis
// rows and are not live across them; then five waves per SIMD fit without spills: five
// groups per CU, so the 1026 tiles of a 512^2 fan sinogram run in one round instead of
// two.  With several tiles per group (more than kIeMaxGroups tiles) the grid takes
// several rounds anyway, and four waves per SIMD leave room for the live sums.
template <bool ONE, bool CACHED>
__global__ __launch_bounds__(kBhcBlock) __attribute__((amdgpu_waves_per_eu(ONE ? 5 : 4)))
void k_bhc_ie_ray(BhcParams P, const float* __restrict__ s, const float* __restrict__ Imea,
                  const float* __restrict__ phi_g, long long count, float sscale,
                  double* __restrict__ partials)
{
    constexpr int kAcc = 20;  // 16 H entries, 4 g entries
    extern __shared__ __attribute__((aligned(16))) char ie_lds[];
    const int E = P.E, row = ie_row(E);
    float* phi = reinterpret_cast<float*>(ie_lds);  // [kBhcBlock][row]
    double2* qr = reinterpret_cast<double2*>(ie_lds + sizeof(float) * kBhcBlock * row);
    double* fs = reinterpret_cast<double*>(qr + kBhcBlock);  // each thread's f sum
    const int t = threadIdx.x;
    const int nb = ie_blocks(E), nhb = nb * (nb + 1) / 2;
    int S = 4;  // at most 64 lanes per slot: a slot's lanes stay inside one wave
    while (S < nhb) S <<= 1;  // nhb <= 36 for E <= 32, so S <= 64
    const int L = kBhcBlock / S, slot = t / L, lane = t % L;
    int bi = 0, bj = 0;
    if (slot < nhb) {
        int rem = slot;  // block row bi of the upper triangle holds nb-bi blocks
        while (rem >= nb - bi) { rem -= nb - bi; ++bi; }
        bj = bi + rem;
    }
    fs[t] = 0.0;
    double acc[kAcc];
    if constexpr (ONE) {
        ie_tile_rows<CACHED>(P, s, Imea, phi_g, count, sscale, blockIdx.x, row, phi, qr, fs);
        __syncthreads();
#pragma unroll
        for (int v = 0; v < kAcc; ++v) acc[v] = 0.0;
        if (slot < nhb) ie_tile_sums(lane, L, row, bi, bj, phi, qr, acc);
        __syncthreads();  // every basis value of the tile read
    } else {
#pragma unroll
        for (int v = 0; v < kAcc; ++v) acc[v] = 0.0;
        const long long ntiles = (count + kBhcBlock - 1) / kBhcBlock;
        for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
            ie_tile_rows<CACHED>(P, s, Imea, phi_g, count, sscale, tile, row, phi, qr, fs);
            __syncthreads();
            if (slot < nhb) ie_tile_sums(lane, L, row, bi, bj, phi, qr, acc);
            __syncthreads();  // every basis value of the tile read
        }
    }
    // f: every thread's sum over its rays, in a fixed order
    const double ftot = block_sum(fs[t], reinterpret_cast<double*>(ie_lds));
    // the slot's L lane sums, in the order of an XOR butterfly over its consecutive
    // lanes (inside one wave): offsets of 16 and more by shuffles, then row shifts of 8,
    // 4, 2, 1 towards the last lane.  The lanes a shift level reads have the offset's bit
    // set, so lane x - off is lane x ^ off, and the last lane ends with the butterfly's
    // value (the same in every lane of a butterfly).
    for (int off = L >> 1; off >= 16; off >>= 1)
#pragma unroll
        for (int v = 0; v < kAcc; ++v) acc[v] += __shfl_xor(acc[v], off, 64);
    if (L > 8) {
#pragma unroll
        for (int v = 0; v < kAcc; ++v) acc[v] = add_from_below<8>(acc[v]);
    }
    if (L > 4) {
#pragma unroll
        for (int v = 0; v < kAcc; ++v) acc[v] = add_from_below<4>(acc[v]);
    }
    if (L > 2) {
#pragma unroll
        for (int v = 0; v < kAcc; ++v) acc[v] = add_from_below<2>(acc[v]);
    }
    if (L > 1) {
#pragma unroll
        for (int v = 0; v < kAcc; ++v) acc[v] = add_from_below<1>(acc[v]);
    }
    const int nh = E * (E + 1) / 2, nv = nh + E + 1;
    const long long G = gridDim.x;
    // the group's column: groups go to the 8 XCDs round-robin (group b on XCD b % 8), and
    // the columns are ordered XCD by XCD, so the 128-byte lines of a value's row are
    // written from one XCD's L2 each and no line is written back from several
    const long long xb = blockIdx.x % 8, xq = G / 8, xr = G % 8;
    double* out = partials + (xb * xq + (xb < xr ? xb : xr) + blockIdx.x / 8);  // v at out[v * G]
    if (t == 0) out[(nv - 1) * G] = ftot;
    if (lane != L - 1 || slot >= nhb) return;
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            const int j = 4 * bi + x, k = 4 * bj + y;  // H_jk, stored for j <= k < E
            if (j <= k && k < E) out[(j * E - j * (j - 1) / 2 + (k - j)) * G] = acc[4 * x + y];
        }
    if (bi == bj) {
#pragma unroll
        for (int y = 0; y < 4; ++y)
            if (4 * bj + y < E) out[(nh + 4 * bj + y) * G] = acc[16 + y];
    }
}

// Cost only of the I-epsilon pass: f_i of every ray with Ir_i formed exactly as in
// k_bhc_ie_ray (FP32 basis values, FP64 sum over j in order), one partial per group;
// the basis values from phi_g when it is given.
__global__ __launch_bounds__(kBhcBlock) void k_bhc_ie_cost(BhcParams P, const float* __restrict__ s,
                                                           const float* __restrict__ Imea,
                                                           const float* __restrict__ phi_g,
                                                           long long count, float sscale,
                                                           double* __restrict__ partials)
{
    __shared__ double sh[kBhcBlock];
    const long long i = (long long)blockIdx.x * kBhcBlock + threadIdx.x;
    double f = 0.0;
    if (i < count) {
        const double im = (double)Imea[i];
        double ir = 0.0;
        if (phi_g) {
            for (int j = 0; j < P.E; ++j) ir = fma(P.Ied[j], (double)phi_g[j * count + i], ir);
        } else {
            spline_basis<1>(P, sscale * s[i], [&](int j, float p, float, float) {
                ir = fma(P.Ied[j], (double)p, ir);
            });
        }
        f = ie_ray_terms(P.noise, ir, im, nullptr);
    }
    const double tot = block_sum(f, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// Sum of value v = blockIdx.x over the nparts partials stored at partials[v * nparts + p]
// (thread-strided partial sums, then block_sum: a fixed order).
__device__ __forceinline__ double reduce_value(const double* __restrict__ partials, long long nparts,
                                               double* sh)
{
    const double* col = partials + (long long)blockIdx.x * nparts;
    double acc = 0.0;
    for (long long p = threadIdx.x; p < nparts; p += kBhcBlock) acc += col[p];
    return block_sum(acc, sh);
}

__global__ __launch_bounds__(kBhcBlock) void k_reduce_partials(const double* __restrict__ partials,
                                                               long long nparts,
                                                               double* __restrict__ out)
{
    __shared__ double sh[kBhcBlock];
    const double tot = reduce_value(partials, nparts, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = tot;
}

// The I-epsilon pass's reduction straight into f, g and H (full): value v of
// [H upper (row-major j <= k), g, f] goes to H[j][k] and H[k][j], g[v - nh] or *f
// (NULL outputs are skipped).
__global__ __launch_bounds__(kBhcBlock) void k_reduce_ie(const double* __restrict__ partials,
                                                         long long nparts, int E,
                                                         double* __restrict__ f,
                                                         double* __restrict__ g,
                                                         double* __restrict__ H)
{
    __shared__ double sh[kBhcBlock];
    const double tot = reduce_value(partials, nparts, sh);
    if (threadIdx.x != 0) return;
    const int v = blockIdx.x, nh = E * (E + 1) / 2;
    if (v < nh) {
        if (!H) return;
        int j = 0, rem = v;
        while (rem >= E - j) {
            rem -= E - j;
            ++j;
        }
        const int k = j + rem;
        H[j * E + k] = tot;
        H[k * E + j] = tot;
    } else if (v < nh + E) {
        if (g) g[v - nh] = tot;
    } else if (f) {
        *f = tot;
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
                       long long count, double* partials, hipStream_t st, const float* phi)
{
    const long long groups = bhc_ie_groups(count);
    if (groups <= 0) return;
    const bool one = groups == bhc_blocks(count);
    const size_t lds = ie_lds_bytes(P.E);
    const dim3 grid((unsigned)groups), block(kBhcBlock);
    if (phi) {
        if (one)
            hipLaunchKernelGGL((k_bhc_ie_ray<true, true>), grid, block, lds, st, P, s, Imea, phi,
                               count, sscale, partials);
        else
            hipLaunchKernelGGL((k_bhc_ie_ray<false, true>), grid, block, lds, st, P, s, Imea, phi,
                               count, sscale, partials);
    } else {
        if (one)
            hipLaunchKernelGGL((k_bhc_ie_ray<true, false>), grid, block, lds, st, P, s, Imea, phi,
                               count, sscale, partials);
        else
            hipLaunchKernelGGL((k_bhc_ie_ray<false, false>), grid, block, lds, st, P, s, Imea,
                               phi, count, sscale, partials);
    }
}

void launch_bhc_ie_cost(const BhcParams& P, const float* s, float sscale, const float* Imea,
                        long long count, double* partials, hipStream_t st, const float* phi)
{
    const long long blocks = bhc_blocks(count);
    if (blocks <= 0) return;
    hipLaunchKernelGGL(k_bhc_ie_cost, dim3((unsigned)blocks), dim3(kBhcBlock), 0, st, P, s, Imea,
                       phi, count, sscale, partials);
}

// The basis matrix of the I-epsilon pass, value j of ray i at phi[j * count + i] (so a
// wave's stores and loads of one j are contiguous), by the same spline_basis as the
// pass evaluates.
__global__ __launch_bounds__(kBhcBlock) void k_bhc_ie_basis(BhcParams P, const float* __restrict__ s,
                                                            long long count, float sscale,
                                                            float* __restrict__ phi)
{
    const long long i = (long long)blockIdx.x * kBhcBlock + threadIdx.x;
    if (i >= count) return;
    spline_basis<1>(P, sscale * s[i], [&](int j, float p, float, float) { phi[j * count + i] = p; });
}

void launch_bhc_ie_basis(const BhcParams& P, const float* s, float sscale, long long count,
                         float* phi, hipStream_t st)
{
    const long long blocks = bhc_blocks(count);
    if (blocks <= 0) return;
    hipLaunchKernelGGL(k_bhc_ie_basis, dim3((unsigned)blocks), dim3(kBhcBlock), 0, st, P, s, count,
                       sscale, phi);
}

void launch_reduce_partials(const double* partials, long long nparts, int nvals, double* out,
                            hipStream_t st)
{
    if (nvals <= 0) return;
    hipLaunchKernelGGL(k_reduce_partials, dim3(nvals), dim3(kBhcBlock), 0, st, partials, nparts,
                       out);
}

void launch_reduce_ie(const double* partials, long long nparts, int E, double* f, double* g,
                      double* H, hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce_ie, dim3(bhc_ie_nvals(E)), dim3(kBhcBlock), 0, st, partials, nparts,
                       E, f, g, H);
}


}  // namespace ctprj
