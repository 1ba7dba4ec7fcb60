// Projector kernels for gfx950: ray-driven forward gather, pixel-driven adjoint gather,
// FBP ramp filter, ray-table construction and image transpose.
//
// Work decomposition (all slices of a plan in one launch each):
//   forward : one work item = (slice, view, run of consecutive bins); one thread per
//             ray.  Views closer to the x axis (|cos| >= |sin|) walk image rows and
//             gather along the row; the others walk the rows of the transposed image,
//             so every gather is along a contiguous row.  Consecutive bins of a view are
//             adjacent lanes: neighbouring lanes read neighbouring pixels and the
//             sinogram store is one coalesced run per wave.
//   adjoint : one work item = (slice, pixel tile); one thread per pixel, gathering
//             over every view and the bins whose strip meets the pixel.  No atomics.
//
// Reference semantics: prj/cpuPrj.c:319-616 (forward), :46-317 (adjoint), :718-784 (FBP).
#include "ctprj_internal.h"

namespace ctprj {

// ------------------------------------------------------------------ ray table ---

// views[v] = (cosT, sinT, 0, 0); bins[b] = (cosB, sinB, d*t/sqrt(d^2+t^2), beamWidth*cosB)
// for fan slices (host-computed in the reference's float expressions).
__global__ void k_build_rays(const SliceDesc* __restrict__ slices, int nslices,
                             const float4* __restrict__ views, const float4* __restrict__ bins,
                             const long long* __restrict__ bin_off,
                             const long long* __restrict__ ray_prefix, Ray* __restrict__ rays)
{
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= ray_prefix[nslices]) return;
    int lo = 0, hi = nslices - 1;  // slice owning gid: last prefix <= gid
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ray_prefix[mid] <= gid) lo = mid; else hi = mid - 1;
    }
    const SliceDesc S = slices[lo];
    const long long local = gid - ray_prefix[lo];
    Ray r;
    if (S.fan) {
        const int v = (int)(local / S.W), b = (int)(local % S.W);
        const float4 T = views[S.view_off + v];
        const float4 B = bins[bin_off[lo] + b];
        // prj/cpuPrj.c:505-506; explicit roundings, no contraction
        r.cosR = __fsub_rn(__fmul_rn(B.x, T.x), __fmul_rn(B.y, T.y));
        r.sinR = __fadd_rn(__fmul_rn(B.y, T.x), __fmul_rn(B.x, T.y));
        r.dc = B.z;
        r.hbc = 0.5f * B.w;
    } else {
        const float4 T = views[S.view_off + local];
        r.cosR = T.x;
        r.sinR = T.y;
        r.dc = 0.0f;
        r.hbc = S.hbd;
    }
    const float c = fabsf(r.cosR), s = fabsf(r.sinR);
    const float hi_ = fmaxf(c, s), lo_ = fminf(c, s);
    r.h = 1.0f / (hi_ * S.er);
    r.b1 = 0.5f * (hi_ - lo_);
    r.s = lo_;
    r.k = lo_ > 0.0f ? 0.5f / lo_ : 0.0f;
    rays[S.ray_off + local] = r;
}

// -------------------------------------------------------------------- transpose ---

// out[j][i] = in[i][j] for every slice; 32x32 tiles through LDS.
__global__ __launch_bounds__(256) void k_transpose(const SliceDesc* __restrict__ slices,
                                                   const AdjItem* __restrict__ items,
                                                   const float* __restrict__ in,
                                                   float* __restrict__ out)
{
    __shared__ float tile[32][33];
    const AdjItem it = items[blockIdx.x];
    const SliceDesc S = slices[it.slice];
    const float* src = in + S.img_off;
    float* dst = out + S.img_off;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int k = 0; k < 32; k += 8) {
        const int r = it.row0 + ty + k, c = it.col0 + tx;
        if (r < S.n && c < S.n) tile[ty + k][tx] = src[(long long)r * S.n + c];
    }
    __syncthreads();
    for (int k = 0; k < 32; k += 8) {
        const int r = it.col0 + ty + k, c = it.row0 + tx;
        if (r < S.n && c < S.n) dst[(long long)r * S.n + c] = tile[tx][ty + k];
    }
}

// ---------------------------------------------------------------------- forward ---

template <bool FAN, bool COUNT>
__device__ __forceinline__ float forward_ray(const SliceDesc& S, const Ray& r, float cT, float sT,
                                             const float* __restrict__ img,
                                             const float* __restrict__ imgT)
{
    const bool rowmode = fabsf(cT) >= fabsf(sT);
    const RowBounds rb = row_bounds(r, cT, sT, FAN ? S.invd : 0.0f, rowmode);
    const int H = S.H, n = S.n, half = S.half;
    const float* base = (rowmode ? img : imgT) + S.img_off;
    // plain sums over blocks of 4 rows, compensated across blocks
    KahanAcc acc;
    float part = 0.0f;
    for (int v = -H; v <= H; ++v) {
        int lo, hi;
        row_range(rb, (float)v, H, lo, hi);
        const float* row = base + (long long)(v + half) * n + half;
        for (int u = lo; u <= hi; ++u) {
            const float x = rowmode ? (float)u : (float)v;
            const float y = rowmode ? (float)v : (float)u;
            const float dist = ray_dist(x, y, r.cosR, r.sinR, r.dc);
            const float hb = FAN ? r.hbc * fan_depth_scale(x, y, cT, sT, S.invd) : r.hbc;
            const float w = strip_weight(dist, hb, r);
            if (COUNT) part += (w != 0.0f) ? 1.0f : 0.0f;
            else part = fmaf(w, row[u], part);
        }
        if (((v + H) & 3) == 3 || v == H) {
            acc.add(part);
            part = 0.0f;
        }
    }
    return acc.s;
}

template <bool COUNT>
__global__ __launch_bounds__(256) void k_forward(const SliceDesc* __restrict__ slices,
                                                 const FwdItem* __restrict__ items,
                                                 const float4* __restrict__ views,
                                                 const Ray* __restrict__ rays,
                                                 const float* __restrict__ img,
                                                 const float* __restrict__ imgT,
                                                 float* __restrict__ sino)
{
    const FwdItem it = items[blockIdx.x];
    if (it.slice < 0) return;
    const SliceDesc S = slices[it.slice];
    const int b = it.bin0 + (int)threadIdx.x;
    if (b >= S.W) return;
    float* out = sino + S.sino_off + (long long)it.view * S.W + b;
    if (b < S.blo || b > S.bhi) {  // even detector: bin 0 carries no ray
        *out = 0.0f;
        return;
    }
    const float4 T = views[S.view_off + it.view];
    float val;
    if (S.fan) {
        const Ray r = rays[S.ray_off + (long long)it.view * S.W + b];
        val = forward_ray<true, COUNT>(S, r, T.x, T.y, img, imgT);
    } else {
        Ray r = rays[S.ray_off + it.view];
        r.dc = (float)(b - S.pC) * S.dSize;
        val = forward_ray<false, COUNT>(S, r, T.x, T.y, img, imgT);
    }
    *out = val;
}

// ---------------------------------------------------------------------- adjoint ---

template <bool FAN, bool FBP>
__device__ __forceinline__ float adjoint_pixel(const SliceDesc& S, int x, int y,
                                               const float4* __restrict__ views,
                                               const Ray* __restrict__ rays,
                                               const float* __restrict__ sino)
{
    const float fx = (float)x, fy = (float)y;
    const float* srow = sino + S.sino_off;
    // plain sums over blocks of 4 views, compensated across blocks
    KahanAcc acc;
    float vsum = 0.0f;
    for (int a = 0; a < S.np; ++a, srow += S.W) {
        const float4 T = views[S.view_off + a];
        const float cT = T.x, sT = T.y;
        int blo, bhi;
        float qd = 1.0f;
        if (FAN) {
            // perspective shadow of the pixel on the virtual detector (4 corners)
            const float oc = fmaf(fx, cT, fy * sT);
            const float qe = fmaf(fx, sT, fmaf(-fy, cT, S.d));
            const float hc = 0.5f * (cT + sT), hd = 0.5f * (cT - sT);
            // corners (x-.5,y-.5),(x+.5,y-.5),(x+.5,y+.5),(x-.5,y+.5)
            const float t0 = S.d * (oc - hc) * fast_rcp(qe + hd);
            const float t1 = S.d * (oc + hd) * fast_rcp(qe + hc);
            const float t2 = S.d * (oc + hc) * fast_rcp(qe - hd);
            const float t3 = S.d * (oc - hd) * fast_rcp(qe - hc);
            const float tl = fminf(fminf(t0, t1), fminf(t2, t3));
            const float tr = fmaxf(fmaxf(t0, t1), fmaxf(t2, t3));
            const float m = S.hbd + kSupportMargin * S.dSize;
            blo = (int)ceilf((tl - m) * S.invDSize + (float)S.pC);
            bhi = (int)floorf((tr + m) * S.invDSize + (float)S.pC);
            qd = fan_depth_scale(fx, fy, cT, sT, S.invd);
        } else {
            const Ray r = rays[S.ray_off + a];
            const float center = fmaf(fx, cT, fy * sT);
            const float ext = r.b1 + r.s + r.hbc + kSupportMargin;
            blo = (int)ceilf((center - ext) * S.invDSize + (float)S.pC);
            bhi = (int)floorf((center + ext) * S.invDSize + (float)S.pC);
        }
        blo = max(blo, S.blo);
        bhi = min(bhi, S.bhi);
        float part = 0.0f;
        if (FAN) {
            const Ray* rr = rays + S.ray_off + (long long)a * S.W;
            for (int b = blo; b <= bhi; ++b) {
                const Ray r = rr[b];
                const float dist = ray_dist(fx, fy, r.cosR, r.sinR, r.dc);
                const float w = strip_weight(dist, r.hbc * qd, r);
                part = fmaf(w, srow[b], part);
            }
            if (FBP) part *= 1.0f / (qd * qd);  // d^2/qe^2, prj/cpuPrj.c:258
        } else {
            const Ray r = rays[S.ray_off + a];
            for (int b = blo; b <= bhi; ++b) {
                const float dc = (float)(b - S.pC) * S.dSize;
                const float dist = ray_dist(fx, fy, r.cosR, r.sinR, dc);
                const float w = strip_weight(dist, r.hbc, r);
                part = fmaf(w, srow[b], part);
            }
        }
        vsum += part;
        if ((a & 3) == 3 || a == S.np - 1) {
            acc.add(vsum);
            vsum = 0.0f;
        }
    }
    return acc.s;
}

template <int TILE, bool FBP>
__global__ __launch_bounds__(TILE* TILE) void k_adjoint(const SliceDesc* __restrict__ slices,
                                                        const AdjItem* __restrict__ items,
                                                        const float4* __restrict__ views,
                                                        const Ray* __restrict__ rays,
                                                        const float* __restrict__ sino,
                                                        float* __restrict__ img)
{
    const AdjItem it = items[blockIdx.x];
    if (it.slice < 0) return;
    const SliceDesc S = slices[it.slice];
    const int row = it.row0 + (int)threadIdx.x / TILE;
    const int col = it.col0 + (int)threadIdx.x % TILE;
    if (row >= S.n || col >= S.n) return;
    const int x = col - S.half, y = row - S.half;
    float val = 0.0f;
    if (x >= -S.H && x <= S.H && y >= -S.H && y <= S.H) {
        if (S.fan) val = adjoint_pixel<true, FBP>(S, x, y, views, rays, sino);
        else val = adjoint_pixel<false, FBP>(S, x, y, views, rays, sino);
        if (FBP) val *= S.fbp_scale;
    }
    img[S.img_off + (long long)row * S.n + col] = val;
}

// --------------------------------------------------------------------------- FBP ---

// Fan cosine pre-weight (prj/cpuPrj.c:737-751) and Hann-windowed ramp filter
// (prj/utils.h:94-151) of one view per work-group.  The FFT filter of the reference is
// a circular convolution over 2W samples of the zero-padded view with
// g = 0.5 h + 0.25 h(.-1) + 0.25 h(.+1), h the periodic Ram-Lak kernel; g is tabulated
// on the host (2W floats per slice at kern_off).  out[i] = sum_j v[j] g[(i-j) mod 2W].
__global__ __launch_bounds__(256) void k_fbp_filter(const SliceDesc* __restrict__ slices,
                                                    const int2* __restrict__ items,
                                                    const float* __restrict__ kern,
                                                    const long long* __restrict__ kern_off,
                                                    const float* __restrict__ preweight,
                                                    const long long* __restrict__ prew_off,
                                                    const float* __restrict__ sino,
                                                    float* __restrict__ out)
{
    extern __shared__ float sh[];  // W weighted samples + 2W kernel taps
    const int2 it = items[blockIdx.x];
    const SliceDesc S = slices[it.x];
    const int W = S.W, L = 2 * W;
    float* v = sh;
    float* g = sh + W;
    const float* src = sino + S.sino_off + (long long)it.y * W;
    const float* pw = preweight + prew_off[it.x];
    const float* kk = kern + kern_off[it.x];
    for (int j = threadIdx.x; j < W; j += blockDim.x) v[j] = src[j] * (S.fan ? pw[j] : 1.0f);
    for (int j = threadIdx.x; j < L; j += blockDim.x) g[j] = kk[j];
    __syncthreads();
    float* dst = out + S.sino_off + (long long)it.y * W;
    for (int i = threadIdx.x; i < W; i += blockDim.x) {
        KahanAcc acc;
        float part = 0.0f;
        for (int j = 0; j < W; ++j) {
            int k = i - j;
            k += (k < 0) ? L : 0;
            part = fmaf(v[j], g[k], part);
            if ((j & 63) == 63) { acc.add(part); part = 0.0f; }
        }
        acc.add(part);
        dst[i] = acc.s;
    }
}

// ---------------------------------------------------------------------- helpers ---

__global__ void k_scale(float* __restrict__ p, long long n, float s)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] *= s;
}

// -------------------------------------------------------------------- launchers ---

void launch_build_rays(const SliceDesc* desc, int nslices, const float4* views,
                       const float4* bins, const long long* bin_off, const long long* ray_prefix,
                       Ray* rays, long long nray, hipStream_t st)
{
    const int threads = 256;
    const long long blocks = (nray + threads - 1) / threads;
    if (blocks > 0)
        hipLaunchKernelGGL(k_build_rays, dim3((unsigned)blocks), dim3(threads), 0, st, desc,
                           nslices, views, bins, bin_off, ray_prefix, rays);
}

void launch_transpose(int nitems, const SliceDesc* desc, const AdjItem* items, const float* in,
                      float* out, hipStream_t st)
{
    if (nitems > 0)
        hipLaunchKernelGGL(k_transpose, dim3(nitems), dim3(256), 0, st, desc, items, in, out);
}

void launch_forward_kernel(bool count, int block, int nitems, const SliceDesc* desc,
                           const FwdItem* items, const float4* views, const Ray* rays,
                           const float* img, const float* imgT, float* sino, hipStream_t st)
{
    if (count)
        hipLaunchKernelGGL(k_forward<true>, dim3(nitems), dim3(block), 0, st, desc, items,
                           views, rays, img, imgT, sino);
    else
        hipLaunchKernelGGL(k_forward<false>, dim3(nitems), dim3(block), 0, st, desc, items,
                           views, rays, img, imgT, sino);
}

void launch_adjoint_kernel(int tile, bool fbp, int nitems, const SliceDesc* desc,
                           const AdjItem* items, const float4* views, const Ray* rays,
                           const float* sino, float* img, hipStream_t st)
{
    if (tile == 8) {
        if (fbp)
            hipLaunchKernelGGL((k_adjoint<8, true>), dim3(nitems), dim3(64), 0, st, desc, items,
                               views, rays, sino, img);
        else
            hipLaunchKernelGGL((k_adjoint<8, false>), dim3(nitems), dim3(64), 0, st, desc, items,
                               views, rays, sino, img);
    } else {
        if (fbp)
            hipLaunchKernelGGL((k_adjoint<16, true>), dim3(nitems), dim3(256), 0, st, desc, items,
                               views, rays, sino, img);
        else
            hipLaunchKernelGGL((k_adjoint<16, false>), dim3(nitems), dim3(256), 0, st, desc,
                               items, views, rays, sino, img);
    }
}

void launch_fbp_filter(int nitems, size_t shmem, const SliceDesc* desc, const int2* items,
                       const float* kern, const long long* kern_off, const float* prew,
                       const long long* prew_off, const float* sino, float* out, hipStream_t st)
{
    hipLaunchKernelGGL(k_fbp_filter, dim3(nitems), dim3(256), shmem, st, desc, items, kern,
                       kern_off, prew, prew_off, sino, out);
}

void launch_scale(float* p, long long n, float s, hipStream_t st)
{
    const long long blocks = (n + 255) / 256;
    if (blocks > 0) hipLaunchKernelGGL(k_scale, dim3((unsigned)blocks), dim3(256), 0, st, p, n, s);
}

}  // namespace ctprj
