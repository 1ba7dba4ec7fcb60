// 8-fold symmetric projector kernels for gfx950.
//
// When prjFull is a multiple of 4 the view set is closed under the symmetry group D4 of
// the pixel grid (rotations by 90 degrees and reflections): with Q = prjFull/4,
//   R  (rotate +90):          pixel (x, y) -> (-y, x),  ray (a, b) -> (a + Q, b)
//   M  (reflect in the x axis): pixel (x, y) -> (x, -y), ray (a, b) -> (2Q - a, 2pC - b)
// (views mod 4Q; prj/cpuPrj.c:401-460 writes the same eight images).  Element
// e = k + 4m is R^k M^m.  Every weight is evaluated once, for a canonical view
// a0 in [0, Q/2] (0..45 degrees), and reused for the eight images:
//   A[(g a0, g b), g p] = W(a0, b, p)  for g in D4,
// which fixes the float value of every matrix entry, so forward and adjoint stay
// bitwise transposes.  Canonical views on a mirror line (a0 = 0, a0 = Q/2) use the four
// rotations only, otherwise each view would be counted twice.
// Parallel beam over half a turn (np <= 2Q, "fold"): ray (a + 2Q, b) is ray
// (a, 2pC - b); the adjoint uses the half group S = {R0, R1, M, R3 M} (which maps
// (0, 45) into [0, 180)) and folds, the forward takes half the detector with all eight
// images.  |dist| in the weight makes W(a0, b, p) == W(a0, 2pC - b, R^2 p) exactly.
#include "ctprj_internal.h"

namespace ctprj {

// (k, m) action on pixel coordinates: reflect first, then rotate k times.
__device__ __forceinline__ void apply_px(int k, int m, int x, int y, int& ox, int& oy)
{
    if (m) y = -y;
    switch (k & 3) {
        case 0: ox = x; oy = y; break;
        case 1: ox = -y; oy = x; break;
        case 2: ox = -x; oy = -y; break;
        default: ox = y; oy = -x; break;
    }
}

// (k, m) action on (canonical view a0, bin b) -> (view, sign, offset): bin = off + sign*b.
// Views beyond np read a zero row (valid = false).
__device__ __forceinline__ int map_view(const SliceDesc& S, int k, int m, int a0, int& sign,
                                        int& off)
{
    const int P = 4 * S.Q;
    int v = (m ? 2 * S.Q - a0 : a0) + (k & 3) * S.Q;
    v %= P;
    if (v < 0) v += P;
    sign = m ? -1 : 1;
    off = m ? 2 * S.pC : 0;
    if (S.fold && v >= 2 * S.Q) {
        v -= 2 * S.Q;
        sign = -sign;
        off = 2 * S.pC - off;
    }
    return v;
}

// Candidate bins of pixel (x, y) for a view (shared with the general adjoint's rule).
template <bool FAN>
__device__ __forceinline__ void pixel_bins(const SliceDesc& S, const Ray* rays_view, float fx,
                                           float fy, float cT, float sT, int& blo, int& bhi)
{
    if (FAN) {
        const float oc = fmaf(fx, cT, fy * sT);
        const float qe = fmaf(fx, sT, fmaf(-fy, cT, S.d));
        const float hc = 0.5f * (cT + sT), hd = 0.5f * (cT - sT);
        const float t0 = __fdividef(S.d * (oc - hc), qe + hd);
        const float t1 = __fdividef(S.d * (oc + hd), qe + hc);
        const float t2 = __fdividef(S.d * (oc + hc), qe - hd);
        const float t3 = __fdividef(S.d * (oc - hd), qe - hc);
        const float tl = fminf(fminf(t0, t1), fminf(t2, t3));
        const float tr = fmaxf(fmaxf(t0, t1), fmaxf(t2, t3));
        const float m = S.hbd + kSupportMargin * S.dSize;
        blo = (int)ceilf((tl - m) * S.invDSize + (float)S.pC);
        bhi = (int)floorf((tr + m) * S.invDSize + (float)S.pC);
    } else {
        const Ray r = rays_view[0];
        const float center = fmaf(fx, cT, fy * sT);
        const float ext = r.b1 + r.s + r.hbc + kSupportMargin;
        blo = (int)ceilf((center - ext) * S.invDSize + (float)S.pC);
        bhi = (int)floorf((center + ext) * S.invDSize + (float)S.pC);
    }
    blo = max(blo, S.blo);
    bhi = min(bhi, S.bhi);
}

// ------------------------------------------------------------------ forward ---

// Rows are summed in plain FP32 over blocks of kRowBlock rows, and the block sums are
// added with compensation: the long sum over rows stays at FP32 rounding of the result
// while the compensation costs a quarter of a per-row update.
constexpr int kRowBlock = 4;

template <bool FAN, int NACC>
__device__ __forceinline__ void fwd_sym_rows(const SliceDesc& S, const Ray& r, float cT, float sT,
                                             const float* __restrict__ img,
                                             const float* __restrict__ imgT, int vlo, int vhi,
                                             KahanAcc* acc)
{
    const RowBounds rb = row_bounds(r, cT, sT, FAN ? S.invd : 0.0f, true);
    const int H = S.H, n = S.n, half = S.half;
    const float* I = img + S.img_off;
    const float* IT = imgT + S.img_off;
    float part[NACC];
#pragma unroll
    for (int e = 0; e < NACC; ++e) part[e] = 0.0f;
    int inblock = 0;
    for (int v = vlo; v <= vhi; ++v) {
        int lo, hi;
        row_range(rb, (float)v, H, lo, hi);
        const float* rp = I + (long long)(v + half) * n + half;     // img row  v
        const float* rm = I + (long long)(half - v) * n + half;     // img row -v
        const float* tp = IT + (long long)(v + half) * n + half;    // imgT row  v
        const float* tm = IT + (long long)(half - v) * n + half;    // imgT row -v
        const float fy = (float)v;
        for (int u = lo; u <= hi; ++u) {
            const float fx = (float)u;
            const float dist = ray_dist(fx, fy, r.cosR, r.sinR, r.dc);
            const float hb = FAN ? r.hbc * fan_depth_scale(fx, fy, cT, sT, S.invd) : r.hbc;
            const float w = strip_weight(dist, hb, r);
            part[0] = fmaf(w, rp[u], part[0]);   // R0   ( x,  y)
            part[1] = fmaf(w, tm[u], part[1]);   // R1   (-y,  x)
            part[2] = fmaf(w, rm[-u], part[2]);  // R2   (-x, -y)
            part[3] = fmaf(w, tp[-u], part[3]);  // R3   ( y, -x)
            if (NACC == 8) {
                part[4] = fmaf(w, rm[u], part[4]);   // M     ( x, -y)
                part[5] = fmaf(w, tp[u], part[5]);   // R1M   ( y,  x)
                part[6] = fmaf(w, rp[-u], part[6]);  // R2M   (-x,  y)
                part[7] = fmaf(w, tm[-u], part[7]);  // R3M   (-y, -x)
            }
        }
        if (++inblock == kRowBlock || v == vhi) {
#pragma unroll
            for (int e = 0; e < NACC; ++e) {
                acc[e].add(part[e]);
                part[e] = 0.0f;
            }
            inblock = 0;
        }
    }
}

// One canonical ray run: NACC images (8, or the 4 rotations on a mirror line).
template <bool FAN, int NACC, int SEG>
__device__ __forceinline__ void fwd_sym_run(const SliceDesc& S, const FwdItem& it, float cT,
                                            float sT, int a0, const Ray* __restrict__ rays,
                                            const float* __restrict__ img,
                                            const float* __restrict__ imgT,
                                            float* __restrict__ sino, float (*red)[8][64])
{
    const int lane = (int)threadIdx.x & 63, seg = (int)threadIdx.x >> 6;
    const int b = it.bin0 + lane;
    const int bmax = S.fold ? S.pC : S.bhi;
    const bool active = b >= S.blo && b <= bmax;
    KahanAcc acc[NACC];
    if (active) {
        Ray r;
        if (FAN) r = rays[S.ray_off + (long long)it.view * S.W + b];
        else {
            r = rays[S.ray_off + it.view];
            r.dc = (float)(b - S.pC) * S.dSize;
        }
        const int nrow = 2 * S.H + 1, per = (nrow + SEG - 1) / SEG;
        const int vlo = -S.H + seg * per, vhi = min(S.H, vlo + per - 1);
        fwd_sym_rows<FAN, NACC>(S, r, cT, sT, img, imgT, vlo, vhi, acc);
    }
    if (SEG > 1) {
#pragma unroll
        for (int e = 0; e < NACC; ++e) red[seg][e][lane] = acc[e].s;
        __syncthreads();
        if (seg != 0) return;
#pragma unroll
        for (int e = 0; e < NACC; ++e) {
            float t = red[0][e][lane];
#pragma unroll
            for (int q = 1; q < SEG; ++q) t += red[q][e][lane];
            acc[e].s = t;
        }
    }
    float* out = sino + S.sino_off;
    if (b < S.blo) {  // even detector: bin 0 carries no ray in any view
#pragma unroll
        for (int e = 0; e < NACC; ++e) {
            int sg, off;
            const int v = map_view(S, e & 3, e >> 2, a0, sg, off);
            if (v < S.np) out[(long long)v * S.W] = 0.0f;
        }
        return;
    }
    if (!active) return;
#pragma unroll
    for (int e = 0; e < NACC; ++e) {
        int sg, off;
        const int v = map_view(S, e & 3, e >> 2, a0, sg, off);
        if (v < S.np) out[(long long)v * S.W + off + sg * b] = acc[e].s;
    }
}

// One work item: (slice, canonical view index, first bin of 64).  SEG waves share the
// 64 rays, each walking 1/SEG of the image rows; wave 0 sums the partials in order.
template <int SEG>
__global__ __launch_bounds__(64 * SEG) void k_forward_sym(const SliceDesc* __restrict__ slices,
                                                          const FwdItem* __restrict__ items,
                                                          const float4* __restrict__ views,
                                                          const Ray* __restrict__ rays,
                                                          const float* __restrict__ img,
                                                          const float* __restrict__ imgT,
                                                          float* __restrict__ sino)
{
    __shared__ float red[SEG][8][64];
    const FwdItem it = items[blockIdx.x];
    if (it.slice < 0) return;  // uniform over the work-group
    const SliceDesc S = slices[it.slice];
    const float4 T = views[S.view_off + it.view];
    const int a0 = __float_as_int(T.z);
    const bool mirror = (a0 == 0) || (2 * a0 == S.Q);
    // (fan, mirror) are uniform over the work-group: the barrier inside is reached by all
    if (S.fan) {
        if (mirror) fwd_sym_run<true, 4, SEG>(S, it, T.x, T.y, a0, rays, img, imgT, sino, red);
        else fwd_sym_run<true, 8, SEG>(S, it, T.x, T.y, a0, rays, img, imgT, sino, red);
    } else {
        if (mirror) fwd_sym_run<false, 4, SEG>(S, it, T.x, T.y, a0, rays, img, imgT, sino, red);
        else fwd_sym_run<false, 8, SEG>(S, it, T.x, T.y, a0, rays, img, imgT, sino, red);
    }
}

// ------------------------------------------------------------------ adjoint ---

// Orbit gather for one canonical view and one group element g' = (k2, m2): the weight
// W(a0, b, g'^-1 q0) is shared by the orbit pixels h q0 in HSET, which read view
// h g' (a0, b).  HSET 0: all eight h; 1: rotations; 2: reflections.  On a mirror view
// (a0 = 0 or Q/2) rotations g' pair with rotation h and reflections with reflection h,
// so that h g' is always a rotation: every entry then takes the forward's key
// (a0, b, R^-k q) and the two directions evaluate the same float weight.
template <bool FAN, bool FBP, int HSET>
__device__ __forceinline__ void adj_sym_elem(const SliceDesc& S, const Ray* __restrict__ rview,
                                             const float* __restrict__ sino,
                                             const float* __restrict__ zero_row, float cT,
                                             float sT, int a0, int k2, int m2, int x0, int y0,
                                             float* vpart)
{
    // p = g'^-1 q0; (R^k M)^-1 = R^k M, (R^k)^-1 = R^-k
    int px, py;
    if (m2) apply_px(k2, 1, x0, y0, px, py);
    else apply_px(4 - k2, 0, x0, y0, px, py);
    const float fx = (float)px, fy = (float)py;
    int blo, bhi;
    pixel_bins<FAN>(S, rview, fx, fy, cT, sT, blo, bhi);
    if (blo > bhi) return;
    const float* base[8];
    int sgn[8];
#pragma unroll
    for (int h = 0; h < 8; ++h) {  // element h g'
        const int k1 = h & 3, m1 = h >> 2;
        const int k = k1 + (m1 ? -k2 : k2), m = m1 ^ m2;
        int off;
        const int v = map_view(S, k, m, a0, sgn[h], off);
        base[h] = (v < S.np) ? sino + S.sino_off + (long long)v * S.W + off : zero_row;
        if (v >= S.np) sgn[h] = 1;
    }
    const float qd = FAN ? fan_depth_scale(fx, fy, cT, sT, S.invd) : 1.0f;
    float part[8];
#pragma unroll
    for (int h = 0; h < 8; ++h) part[h] = 0.0f;
    for (int b = blo; b <= bhi; ++b) {
        Ray r;
        if (FAN) r = rview[b];
        else {
            r = rview[0];
            r.dc = (float)(b - S.pC) * S.dSize;
        }
        const float dist = ray_dist(fx, fy, r.cosR, r.sinR, r.dc);
        const float w = strip_weight(dist, FAN ? r.hbc * qd : r.hbc, r);
#pragma unroll
        for (int h = 0; h < 8; ++h) {
            if ((HSET == 1 && h >= 4) || (HSET == 2 && h < 4)) continue;
            part[h] = fmaf(w, base[h][sgn[h] * b], part[h]);
        }
    }
    // FBP fan weight d^2/qe^2 (prj/cpuPrj.c:258); qe of g'^-1 q0 at a0 is the depth of
    // every orbit pixel at its own view
    const float fw = (FBP && FAN) ? 1.0f / (qd * qd) : 1.0f;
#pragma unroll
    for (int h = 0; h < 8; ++h) {
        if ((HSET == 1 && h >= 4) || (HSET == 2 && h < 4)) continue;
        vpart[h] += (FBP && FAN) ? part[h] * fw : part[h];
    }
}

// Canonical views i = i0, i0 + step, ... for the orbit of (x0, y0).
template <bool FAN, bool FBP>
__device__ __forceinline__ void adj_sym_pixel(const SliceDesc& S, const float4* __restrict__ views,
                                              const Ray* __restrict__ rays,
                                              const float* __restrict__ sino,
                                              const float* __restrict__ zero_row, int x0, int y0,
                                              int i0, int step, KahanAcc* acc)
{
    for (int i = i0; i < S.ncanon; i += step) {
        const float4 T = views[S.view_off + i];
        const int a0 = __float_as_int(T.z);
        const bool mirror = (a0 == 0) || (2 * a0 == S.Q);
        const Ray* rview = FAN ? rays + S.ray_off + (long long)i * S.W : rays + S.ray_off + i;
        float vpart[8];  // this view's sums; compensated once per view
#pragma unroll
        for (int h = 0; h < 8; ++h) vpart[h] = 0.0f;
#define CTPRJ_ELEM(HS, K, M) \
    adj_sym_elem<FAN, FBP, HS>(S, rview, sino, zero_row, T.x, T.y, a0, K, M, x0, y0, vpart)
        if (mirror) {
            // rotations g' -> rotation orbit pixels, reflections g' -> reflected ones
            const int nk = S.fold ? 2 : 4;
            for (int k = 0; k < nk; ++k) CTPRJ_ELEM(1, k, 0);
            for (int k = 0; k < nk; ++k) CTPRJ_ELEM(2, k, 1);
        } else if (S.fold) {
            // half-turn parallel scan: S = {R0, R1, M, R3 M} maps (0, 45) into [0, 180)
            CTPRJ_ELEM(0, 0, 0);
            CTPRJ_ELEM(0, 1, 0);
            CTPRJ_ELEM(0, 0, 1);
            CTPRJ_ELEM(0, 3, 1);
        } else {
            for (int k = 0; k < 4; ++k) CTPRJ_ELEM(0, k, 0);
            for (int k = 0; k < 4; ++k) CTPRJ_ELEM(0, k, 1);
        }
#undef CTPRJ_ELEM
#pragma unroll
        for (int h = 0; h < 8; ++h) acc[h].add(vpart[h]);
    }
}

// One work item: an 8x8 tile of the octant 0 <= y0 <= x0 <= H; each thread owns the
// orbit of (x0, y0) and writes its distinct pixels.  SEG waves split the canonical views
// (wave w takes i = w, w + SEG, ...); wave 0 sums their partials in order.
template <bool FBP, int SEG>
__global__ __launch_bounds__(64 * SEG) void k_adjoint_sym(const SliceDesc* __restrict__ slices,
                                                          const AdjItem* __restrict__ items,
                                                          const float4* __restrict__ views,
                                                          const Ray* __restrict__ rays,
                                                          const float* __restrict__ sino,
                                                          const float* __restrict__ zero_row,
                                                          float* __restrict__ img)
{
    __shared__ float red[SEG][8][64];
    const AdjItem it = items[blockIdx.x];
    if (it.slice < 0) return;  // uniform over the work-group
    const SliceDesc S = slices[it.slice];
    const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
    const int x0 = it.col0 + (lane & 7);
    const int y0 = it.row0 + (lane >> 3);
    const bool valid = y0 <= x0 && x0 <= S.H;
    KahanAcc acc[8];
    if (valid) {
        if (S.fan) adj_sym_pixel<true, FBP>(S, views, rays, sino, zero_row, x0, y0, wave, SEG, acc);
        else adj_sym_pixel<false, FBP>(S, views, rays, sino, zero_row, x0, y0, wave, SEG, acc);
    }
    if (SEG > 1) {
#pragma unroll
        for (int h = 0; h < 8; ++h) red[wave][h][lane] = acc[h].s;
        __syncthreads();
        if (wave != 0) return;
#pragma unroll
        for (int h = 0; h < 8; ++h) {
            float t = red[0][h][lane];
#pragma unroll
            for (int q = 1; q < SEG; ++q) t += red[q][h][lane];
            acc[h].s = t;
        }
    }
    if (!valid) return;
    if (FBP)
#pragma unroll
        for (int h = 0; h < 8; ++h) acc[h].s *= S.fbp_scale;
    float* out = img + S.img_off;
    const int n = S.n, half = S.half;
    const int nrot = (x0 == 0 && y0 == 0) ? 1 : 4;
    const bool refl = y0 > 0 && y0 < x0;
#pragma unroll
    for (int h = 0; h < 8; ++h) {
        if (h < 4 ? h >= nrot : !refl) continue;
        int x, y;
        apply_px(h & 3, h >> 2, x0, y0, x, y);
        out[(long long)(y + half) * n + x + half] = acc[h].s;
    }
}

// Even n: row 0 and column 0 lie outside the pixel grid the projector reads; the
// adjoint writes them as 0 (prj/cpuPrj.c:713-716).
__global__ void k_zero_border(const SliceDesc* __restrict__ slices, const int* __restrict__ ids,
                              float* __restrict__ img)
{
    const SliceDesc S = slices[ids[blockIdx.x]];
    float* out = img + S.img_off;
    for (int i = threadIdx.x; i < S.n; i += blockDim.x) {
        out[i] = 0.0f;
        out[(long long)i * S.n] = 0.0f;
    }
}

// ----------------------------------------------------------------- launchers ---

void launch_forward_sym(int seg, int nitems, const SliceDesc* desc, const FwdItem* items,
                        const float4* views, const Ray* rays, const float* img,
                        const float* imgT, float* sino, hipStream_t st)
{
    if (seg == 1)
        hipLaunchKernelGGL(k_forward_sym<1>, dim3(nitems), dim3(64), 0, st, desc, items, views,
                           rays, img, imgT, sino);
    else if (seg == 2)
        hipLaunchKernelGGL(k_forward_sym<2>, dim3(nitems), dim3(128), 0, st, desc, items, views,
                           rays, img, imgT, sino);
    else
        hipLaunchKernelGGL(k_forward_sym<4>, dim3(nitems), dim3(256), 0, st, desc, items, views,
                           rays, img, imgT, sino);
}

template <bool FBP>
static void launch_adjoint_sym_t(int seg, int nitems, const SliceDesc* desc, const AdjItem* items,
                                 const float4* views, const Ray* rays, const float* sino,
                                 const float* zero_row, float* img, hipStream_t st)
{
    switch (seg) {
        case 1:
            hipLaunchKernelGGL((k_adjoint_sym<FBP, 1>), dim3(nitems), dim3(64), 0, st, desc, items,
                               views, rays, sino, zero_row, img);
            break;
        case 2:
            hipLaunchKernelGGL((k_adjoint_sym<FBP, 2>), dim3(nitems), dim3(128), 0, st, desc,
                               items, views, rays, sino, zero_row, img);
            break;
        case 4:
            hipLaunchKernelGGL((k_adjoint_sym<FBP, 4>), dim3(nitems), dim3(256), 0, st, desc,
                               items, views, rays, sino, zero_row, img);
            break;
        default:
            hipLaunchKernelGGL((k_adjoint_sym<FBP, 8>), dim3(nitems), dim3(512), 0, st, desc,
                               items, views, rays, sino, zero_row, img);
            break;
    }
}

void launch_adjoint_sym(bool fbp, int seg, int nitems, const SliceDesc* desc,
                        const AdjItem* items, const float4* views, const Ray* rays,
                        const float* sino, const float* zero_row, float* img, hipStream_t st)
{
    if (fbp) launch_adjoint_sym_t<true>(seg, nitems, desc, items, views, rays, sino, zero_row, img, st);
    else launch_adjoint_sym_t<false>(seg, nitems, desc, items, views, rays, sino, zero_row, img, st);
}

void launch_zero_border(int nids, const SliceDesc* desc, const int* ids, float* img,
                        hipStream_t st)
{
    if (nids > 0) hipLaunchKernelGGL(k_zero_border, dim3(nids), dim3(256), 0, st, desc, ids, img);
}

}  // namespace ctprj
