// Symmetric projector kernels for gfx950.
//
// The view set of a scan over a full (or half) turn is closed under a symmetry group of
// the pixel grid.  With P = prjFull,
//   R  (rotate +90):          pixel (x, y) -> (-y, x),  ray (a, b) -> (a + P/4, b)
//   M  (reflect in the x axis): pixel (x, y) -> (x, -y), ray (a, b) -> (P/2 - a, 2pC - b)
// (views mod P; prj/cpuPrj.c:401-460 writes the same eight images for P % 4 == 0).
// Element e = k + 4m is R^k M^m.  A slice uses the largest group whose view maps are
// integral: D4 (all eight) for P % 4 == 0, V4 = {1, R^2, M, R^2 M} for even P, and
// Z2 = {1, R^2 M} (theta -> -theta) otherwise.  Every weight is evaluated once, for a
// canonical view a0 in [0, P/q] (q = 8, 4, 2), and reused for the group's images:
//   A[(g a0, g b), g p] = W(a0, b, p)  for g in the group,
// which fixes the float value of every matrix entry, so forward and adjoint stay
// bitwise transposes.  Canonical views on a mirror line (a0 = 0, a0 = P/q) use the
// rotations only, otherwise each view would be counted twice.
// Parallel beam over half a turn (np <= P/2, "fold"; D4 and V4): ray (a + P/2, b) is ray
// (a, 2pC - b); the adjoint uses a half set of elements that maps the canonical range
// into [0, 180) and folds, the forward takes half the detector with all the images.
// |dist| in the weight makes W(a0, b, p) == W(a0, 2pC - b, R^2 p) exactly.
#include "ctprj_internal.h"
#include "bhc_device.h"

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
// 4 v = (+-4 a0) + (k + 2m) P, a multiple of 4 for every element of the slice's group.
// Views beyond np read a zero row (the caller checks v < np).
__device__ __forceinline__ int map_view(const SliceDesc& S, int k, int m, int a0, int& sign,
                                        int& off)
{
    const int P = S.P;
    // a0 <= P/2, so v starts in [-P/2, 7P/2): reduce mod P without a division
    int v = ((m ? -4 * a0 : 4 * a0) + ((k + 2 * m) & 3) * P) >> 2;
    if (v < 0) v += P;
    if (v >= 2 * P) v -= 2 * P;
    if (v >= P) v -= P;
    sign = m ? -1 : 1;
    off = m ? 2 * S.pC : 0;
    if (S.fold && 2 * v >= P) {
        v -= P >> 1;
        sign = -sign;
        off = 2 * S.pC - off;
    }
    return v;
}

__device__ __forceinline__ bool mirror_view(const SliceDesc& S, int a0)
{
    return a0 == 0 || group_masks(S.sym).q * a0 == S.P;
}

// Candidate bins of pixel (x, y) for a view (shared with the general adjoint's rule).
template <bool FAN>
__device__ __forceinline__ void pixel_bins(const SliceDesc& S, const Ray* rays_view, float fx,
                                           float fy, float cT, float sT, int& blo, int& bhi)
{
    if (FAN) {
        // the pixel's shadow on the virtual detector lies within t_c +- ext: a corner
        // offset (do, dq) of the projected coordinates (|do|, |dq| <= rho) moves
        // t = d oc/qe by d |do qe - oc dq| / (qe (qe + dq)) <= d rho (qe + |oc|) / (qe (qe - rho))
        const float oc = fmaf(fx, cT, fy * sT);
        const float qe = fmaf(fx, sT, fmaf(-fy, cT, S.d));
        const float rho = 0.5f * (fabsf(cT) + fabsf(sT));
        const float iq = fast_rcp(qe), iqr = fast_rcp(qe - rho);
        const float tc = S.d * oc * iq;
        const float ext = S.d * rho * (qe + fabsf(oc)) * iq * iqr;
        const float m = ext + S.hbd + kSupportMargin * S.dSize;
        blo = (int)ceilf((tc - m) * S.invDSize + (float)S.pC);
        bhi = (int)floorf((tc + m) * S.invDSize + (float)S.pC);
        if (!(qe - rho > 0.0f)) {  // the source within a pixel: every bin is a candidate
            blo = S.blo;
            bhi = S.bhi;
        }
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

// The sum of a RedJob by one wave: lane l adds in[l], in[l + 64], ... in order, then a
// butterfly over the lanes; the order is fixed, so every caller gets the same value.
__device__ __forceinline__ void wave_sum_job(const RedJob& job)
{
    const int lane = (int)threadIdx.x & 63;
    double acc = 0.0;
    for (long long q = lane; q < job.n; q += 64) acc += job.in[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) *job.out = acc;
}

__global__ __launch_bounds__(64) void k_wave_sum(const RedJob job) { wave_sum_job(job); }

void launch_wave_sum(const RedJob& job, hipStream_t st)
{
    hipLaunchKernelGGL(k_wave_sum, dim3(1), dim3(64), 0, st, job);
}

// ------------------------------------------------------------------ forward ---

// Rows are summed in plain FP32 over blocks of kRowBlock rows, and the block sums are
// added with compensation: the long sum over rows stays at FP32 rounding of the result
// while the compensation costs a quarter of a per-row update.
constexpr int kRowBlock = 4;

// Raw buffer view of one slice's array: loads take a wave-uniform byte offset (a row) in
// an SGPR and a 32-bit per-lane byte offset (a column), so an address costs one integer
// add per lane instead of a 64-bit pointer update.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t slice_rsrc(const float* base, int elems)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), (short)0, elems * 4,
                                             kBufferWord3);
}

__device__ __forceinline__ float ldb(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// Four consecutive floats at a 4-byte aligned offset (one 16-byte load per lane).
struct Quad {
    float q[4];
};

// N (3 or 4) consecutive floats at a 4-byte aligned offset: one 12- or 16-byte load.
template <int N>
struct Win {
    float q[N];
};

template <int N>
__device__ __forceinline__ Win<N> ldbw(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    Win<N> w;
    if constexpr (N == 4) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
#pragma unroll
        for (int k = 0; k < 4; ++k) w.q[k] = __uint_as_float(v[k]);
    } else {
        const auto v = __builtin_amdgcn_raw_buffer_load_b96(r, voff, soff, 0);
#pragma unroll
        for (int k = 0; k < 3; ++k) w.q[k] = __uint_as_float(v[k]);
    }
    return w;
}

__device__ __forceinline__ Quad ldb4(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
    return Quad{{__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
                 __uint_as_float(v[3])}};
}

// One ray of canonical view a0 against every image g p of the pixels p it meets, for
// the elements g in MASK.  Rows mode walks image rows v = y with u = x inside; columns
// mode (canonical views steeper than 45 degrees) walks v = x with u = y inside.  Each
// image pixel g (x, y) is read from the image or its transpose at row +-v, column +-u,
// whichever makes u the contiguous index.  The pixels [lo, hi] a ray meets in one row
// are read as 4-pixel windows [s, s+3] (one 16-byte load per image instead of one load
// per pixel and image); windows are clamped into [-H, H] so they never leave the row
// (n >= 5), and pixels outside [lo, hi] or already taken by the previous window are
// skipped.
template <bool FAN, unsigned MASK, bool COLS>
__device__ __forceinline__ void fwd_sym_rows(const SliceDesc& S, const Ray& r, float cT, float sT,
                                             const float* __restrict__ img,
                                             const float* __restrict__ imgT, int vlo, int vhi,
                                             KahanAcc* acc)
{
    const RowBounds rb = row_bounds(r, cT, sT, FAN ? S.invd : 0.0f, !COLS);
    const int H = S.H, n = S.n, half = S.half;
    const __amdgpu_buffer_rsrc_t RI = slice_rsrc(img + S.img_off, n * n);
    const __amdgpu_buffer_rsrc_t RT = slice_rsrc(imgT + S.img_off, n * n);
    // rows mode: g (u, v) of element e at [RI|RT](row +-v, col +-u); columns mode swaps
    // the two arrays' roles
    const __amdgpu_buffer_rsrc_t RA = COLS ? RT : RI;
    const __amdgpu_buffer_rsrc_t RB = COLS ? RI : RT;
    float part[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) part[e] = 0.0f;
    int inblock = 0;
    for (int v = vlo; v <= vhi; ++v) {
        int lo, hi;
        row_range(rb, (float)v, H, lo, hi);
        const int op = (v + half) * n * 4;  // byte offset of row  v
        const int om = (half - v) * n * 4;  // byte offset of row -v
        const float fv = (float)v;
        for (int s0 = lo; s0 <= hi; s0 += 4) {
            const int s = min(s0, H - 3);
            const int cp = (half + s) * 4;      // window [s, s+3]
            const int cm = (half - s - 3) * 4;  // window [-(s+3), -s]
            // rows: e0 ( x, y) A[v][u]   e1 (-y, x) B[-v][u]  e2 (-x,-y) A[-v][-u]
            //       e3 ( y,-x) B[v][-u]  e4 ( x,-y) A[-v][u]  e5 ( y, x) B[v][u]
            //       e6 (-x, y) A[v][-u]  e7 (-y,-x) B[-v][-u]          (A = img, B = imgT)
            // cols: e0 A[v][u]  e1 B[v][-u]  e2 A[-v][-u]  e3 B[-v][u]
            //       e4 A[v][-u] e5 B[v][u]   e6 A[-v][u]   e7 B[-v][-u]  (A = imgT, B = img)
            Quad w8[8];
            if (MASK & 0x01) w8[0] = ldb4(RA, cp, op);
            if (MASK & 0x02) w8[1] = COLS ? ldb4(RB, cm, op) : ldb4(RB, cp, om);
            if (MASK & 0x04) w8[2] = ldb4(RA, cm, om);
            if (MASK & 0x08) w8[3] = COLS ? ldb4(RB, cp, om) : ldb4(RB, cm, op);
            if (MASK & 0x10) w8[4] = COLS ? ldb4(RA, cm, op) : ldb4(RA, cp, om);
            if (MASK & 0x20) w8[5] = ldb4(RB, cp, op);
            if (MASK & 0x40) w8[6] = COLS ? ldb4(RA, cp, om) : ldb4(RA, cm, op);
            if (MASK & 0x80) w8[7] = ldb4(RB, cm, om);
            // which images read column +u (window index j) and which -u (index 3 - j)
            constexpr unsigned PLUS = COLS ? 0x69u : 0x33u;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int u = s + j;
                if (u < s0 || u > hi) continue;
                const float fu = (float)u;
                const float fx = COLS ? fv : fu, fy = COLS ? fu : fv;
                const float dist = ray_dist(fx, fy, r.cosR, r.sinR, r.dc);
                const float hb = FAN ? r.hbc * fan_depth_scale(fx, fy, cT, sT, S.invd) : r.hbc;
                const float w = strip_weight(dist, hb, r);
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    if (MASK >> e & 1u)
                        part[e] = fmaf(w, w8[e].q[(PLUS >> e & 1u) ? j : 3 - j], part[e]);
            }
        }
        if (++inblock == kRowBlock || v == vhi) {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                if (!(MASK >> e & 1u)) continue;
                acc[e].add(part[e]);
                part[e] = 0.0f;
            }
            inblock = 0;
        }
    }
}

// Kernel argument of the forward launches: empty, or (BHC) the beam-hardening ray pass
// fused into the final store (ctprj_bhc_grad_alpha): every sinogram entry is stored
// once, as w of bhc_ray at Ts times its ray sum, and its FP64 NLL term goes into the
// storing work-group's partial at partials[part_off + work-group].
template <bool BHC>
struct FwdBhc {};
template <>
struct FwdBhc<true> : BhcStore {};

// Sum of v over the work-group (a multiple of 64 threads, at most 16 waves): a butterfly
// over each wave's lanes, then the waves' sums in wave order; every thread must call it.
__device__ __forceinline__ double group_sum(double v)
{
    __shared__ double sh[16];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int wave = (int)threadIdx.x >> 6, nw = (int)blockDim.x >> 6;
    __syncthreads();  // an earlier call's reads are done
    if ((threadIdx.x & 63) == 0) sh[wave] = v;
    __syncthreads();
    double r = sh[0];
    for (int w = 1; w < nw; ++w) r += sh[w];
    return r;
}

// The sum t of image e of canonical ray (it, lane) written to its view g_e a0; with BHC
// the entry's w instead, and the NLL term is returned (0 where nothing is stored).
template <bool BHC>
__device__ __forceinline__ double fwd_sym_store(const SliceDesc& S, const FwdItem& it, int a0,
                                                int e, int lane, float t, float* __restrict__ sino,
                                                const FwdBhc<BHC>& B)
{
    const int b = it.bin0 + lane;
    const int bmax = S.fold ? S.pC : S.bhi;
    const bool active = b >= S.blo && b <= bmax;
    float* out = sino + S.sino_off;
    int sg, off;
    const int v = map_view(S, e & 3, e >> 2, a0, sg, off);
    if (v >= S.np) return 0.0;
    // folded scans: images e and e ^ 2 (R^2 apart) meet at the centre bin, whose two sums
    // are equal up to rounding, and both reach bin 0 of an even detector; one writer
    // (k < 2) keeps the result deterministic and counts each entry once
    if (S.fold && (e & 2) && (b == S.pC || b < S.blo)) return 0.0;
    long long idx;
    float val = t;
    if (b < S.blo) {  // even detector: bin 0 carries no ray
        idx = (long long)v * S.W;
        val = 0.0f;
    } else if (active) {
        idx = (long long)v * S.W + off + sg * b;
    } else {
        return 0.0;
    }
    if constexpr (!BHC) {
        out[idx] = val;
        return 0.0;
    } else {
        float w, hw;
        const double f = bhc_ray<false>(B.P, B.Ts * val, B.Imea[S.sino_off + idx], B.Ts, &w, &hw);
        out[idx] = w;
        return f;
    }
}

// One canonical ray run over the images in MASK (the group, or its rotations on a
// mirror view).  The image rows are cut into nsplit * SEG segments: this work-group
// walks segments split * SEG + wave.  nsplit == 1 writes the sinogram; otherwise the
// work-group's sums go to part[e * 64 + lane] for k_forward_sym_combine.
template <bool FAN, unsigned MASK, bool COLS, int SEG, bool BHC>
__device__ __forceinline__ void fwd_sym_run(const SliceDesc& S, const FwdItem& it, float cT,
                                            float sT, int a0, const Ray* __restrict__ rays,
                                            const float* __restrict__ img,
                                            const float* __restrict__ imgT,
                                            float* __restrict__ sino, float (*red)[8][64],
                                            int split, int nsplit, float* __restrict__ part,
                                            float (*tb)[64])
{
    // the wave index is uniform over the wave: the row loop then runs on scalar registers
    const int lane = (int)threadIdx.x & 63;
    const int seg = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int b = it.bin0 + lane;
    const int bmax = S.fold ? S.pC : S.bhi;
    const bool active = b >= S.blo && b <= bmax;
    KahanAcc acc[8];
    if (active) {
        Ray r;
        if (FAN) r = rays[S.ray_off + (long long)it.view * S.W + b];
        else {
            r = rays[S.ray_off + it.view];
            r.dc = (float)(b - S.pC) * S.dSize;
        }
        const int nrow = 2 * S.H + 1, nseg = SEG * nsplit, per = (nrow + nseg - 1) / nseg;
        const int vlo = -S.H + (split * SEG + seg) * per, vhi = min(S.H, vlo + per - 1);
        fwd_sym_rows<FAN, MASK, COLS>(S, r, cT, sT, img, imgT, vlo, vhi, acc);
    }
    if (SEG > 1) {
#pragma unroll
        for (int e = 0; e < 8; ++e)
            if (MASK >> e & 1u) red[seg][e][lane] = acc[e].s;
        __syncthreads();
    }
    // wave w finalises the images e = w (mod SEG): the sum of the waves' partials in wave
    // order.  BHC: the sums wait in tb[e][lane] for the kernel's one fused store loop (a
    // single inlined copy of the spectrum sums)
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        if (!(MASK >> e & 1u) || (SEG > 1 && e % SEG != seg)) continue;
        float t = acc[e].s;
        if (SEG > 1) {
            t = red[0][e][lane];
#pragma unroll
            for (int q = 1; q < SEG; ++q) t += red[q][e][lane];
        }
        if (part) part[e * 64 + lane] = t;
        else if (BHC) tb[e][lane] = t;
        else fwd_sym_store<false>(S, it, a0, e, lane, t, sino, FwdBhc<false>{});
    }
}

template <bool FAN, int SEG, int GRP, bool BHC>
__device__ __forceinline__ void fwd_sym_group(const SliceDesc& S, const FwdItem& it, float cT,
                                              float sT, int a0, bool mirror,
                                              const Ray* __restrict__ rays,
                                              const float* __restrict__ img,
                                              const float* __restrict__ imgT,
                                              float* __restrict__ sino, float (*red)[8][64],
                                              int split, int nsplit, float* __restrict__ part,
                                              float (*tb)[64])
{
    // D4's canonical views lie in [0, 45] degrees and always walk rows; V4's reach 90 and
    // Z2's 180 degrees, and the steep ones walk columns
    constexpr GroupMasks GM = group_masks(GRP);
    const bool cols = GRP != kGrpD4 && fabsf(cT) < fabsf(sT);
#define CTPRJ_RUN(MASK, COLS)                                                                 \
    fwd_sym_run<FAN, MASK, COLS, SEG, BHC>(S, it, cT, sT, a0, rays, img, imgT, sino, red, split, \
                                           nsplit, part, tb)
    if (GRP != kGrpD4 && cols) {
        if (mirror) CTPRJ_RUN(GM.rot, GRP != kGrpD4);
        else CTPRJ_RUN(GM.all, GRP != kGrpD4);
    } else {
        if (mirror) CTPRJ_RUN(GM.rot, false);
        else CTPRJ_RUN(GM.all, false);
    }
#undef CTPRJ_RUN
}

// Work-group -> (split, item): the first nitems work-groups take split 0 of every item,
// the next nitems split 1, and so on.
__device__ __forceinline__ void split_item(int nitems, int& split, int& idx)
{
    split = (int)blockIdx.x / nitems;
    idx = (int)blockIdx.x - split * nitems;
}

// One work item: (slice, canonical view index, first bin of 64) of a slice under group
// GRP (one launch per group keeps each variant's registers to its own images).  SEG
// waves share the 64 rays, each walking 1/SEG of the work-group's image rows, and then
// share the final sums; with nsplit > 1, nsplit work-groups share the rows of an item
// and leave their sums in part (512 floats per work-group) for k_forward_sym_combine.
template <int SEG, int GRP, bool BHC>
__global__ __launch_bounds__(64 * SEG) void k_forward_sym(const SliceDesc* __restrict__ slices,
                                                          const FwdItem* __restrict__ items,
                                                          const float4* __restrict__ views,
                                                          const Ray* __restrict__ rays,
                                                          const float* __restrict__ img,
                                                          const float* __restrict__ imgT,
                                                          float* __restrict__ sino, int nitems,
                                                          int nsplit, float* __restrict__ part,
                                                          const FwdBhc<BHC> B)
{
    __shared__ float red[SEG][8][64];
    int split, idx;
    split_item(nitems, split, idx);
    const FwdItem it = items[idx];
    if (it.slice < 0) {  // uniform over the work-group: a padding item
        if constexpr (BHC)
            if (!part && threadIdx.x == 0) B.partials[B.part_off + blockIdx.x] = 0.0;
        return;
    }
    const SliceDesc S = slices[it.slice];
    const float4 T = views[S.view_off + it.view];
    const int a0 = __float_as_int(T.z);
    const bool mirror = a0 == 0 || group_masks(GRP).q * a0 == S.P;
    float* mine = part ? part + (long long)blockIdx.x * 512 : nullptr;
    // (fan, mirror, walk) are uniform over the work-group: the barrier inside is reached
    // by all of its waves
    __shared__ float tb[BHC ? 8 : 1][64];  // BHC: each lane's final sums, by image
    if (S.fan)
        fwd_sym_group<true, SEG, GRP, BHC>(S, it, T.x, T.y, a0, mirror, rays, img, imgT, sino,
                                           red, split, nsplit, mine, tb);
    else
        fwd_sym_group<false, SEG, GRP, BHC>(S, it, T.x, T.y, a0, mirror, rays, img, imgT, sino,
                                            red, split, nsplit, mine, tb);
    if constexpr (BHC) {
        if (!mine) {  // uniform: the fused store of this lane's images, then the NLL partial
            const int lane = (int)threadIdx.x & 63;
            const int seg = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
            const unsigned mask = mirror ? group_masks(GRP).rot : group_masks(GRP).all;
            double f = 0.0;
            // each lane reads back only its own entries: no barrier
#pragma unroll 1
            for (int e = seg; e < 8; e += SEG)
                if (mask >> e & 1u) f += fwd_sym_store<true>(S, it, a0, e, lane, tb[e][lane], sino, B);
            const double tot = group_sum(f);
            if (threadIdx.x == 0) B.partials[B.part_off + blockIdx.x] = tot;
        }
    }
}

// The nsplit partial sums of each canonical ray and image of an item, added in split
// order and written like the one-work-group path (512 threads: image e = thread / 64 of
// ray lane = thread % 64).
template <int GRP, bool BHC>
__global__ __launch_bounds__(512) void k_forward_sym_combine(const SliceDesc* __restrict__ slices,
                                                             const FwdItem* __restrict__ items,
                                                             const float4* __restrict__ views,
                                                             int nitems, int nsplit,
                                                             const float* __restrict__ part,
                                                             float* __restrict__ sino,
                                                             const FwdBhc<BHC> B)
{
    const FwdItem it = items[blockIdx.x];
    if (it.slice < 0) {  // uniform: a padding item
        if constexpr (BHC)
            if (threadIdx.x == 0) B.partials[B.part_off + blockIdx.x] = 0.0;
        return;
    }
    const SliceDesc S = slices[it.slice];
    const int a0 = __float_as_int(views[S.view_off + it.view].z);
    constexpr GroupMasks GM = group_masks(GRP);
    const unsigned mask = (a0 == 0 || GM.q * a0 == S.P) ? GM.rot : GM.all;
    const int e = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    double f = 0.0;
    if (mask >> e & 1u) {
        const float* p = part + (long long)blockIdx.x * 512 + threadIdx.x;
        float t = p[0];
        for (int k = 1; k < nsplit; ++k) t += p[(long long)k * nitems * 512];
        f = fwd_sym_store<BHC>(S, it, a0, e, lane, t, sino, B);
    }
    if constexpr (BHC) {  // the work-group's NLL partial (every thread takes part)
        const double tot = group_sum(f);
        if (threadIdx.x == 0) B.partials[B.part_off + blockIdx.x] = tot;
    }
}

// ------------------------------------------------------------------ adjoint ---

// Orbit gather for one canonical view and one group element g' = (k2, m2): the weight
// W(a0, b, g'^-1 q0) is shared by the orbit pixels h q0 for h in HMASK, which read view
// h g' (a0, b).  On a mirror view rotations g' pair with rotation h and reflections
// with reflection h, so that h g' is always a rotation: every entry then takes the
// forward's key (a0, b, R^-k q) and the two directions evaluate the same float weight.
// `tab` holds the view-row entries of (a0, g') for the eight h (wave-uniform).  The
// bins [blo, bhi] of the pixel are read as 4-bin windows (one 16-byte load per image),
// clamped into [0, W - 4] so they stay inside the row, reflected rows at 2pC - b.
template <bool FAN, bool FBP, unsigned HMASK, int WIN>
__device__ __forceinline__ void adj_sym_elem(const SliceDesc& S, const Ray* __restrict__ rview,
                                             __amdgpu_buffer_rsrc_t RS, __amdgpu_buffer_rsrc_t RZ,
                                             const int* __restrict__ tab, float cT, float sT,
                                             int k2, int m2, int x0, int y0, float* vpart)
{
    // p = g'^-1 q0; (R^k M)^-1 = R^k M, (R^k)^-1 = R^-k
    int px, py;
    if (m2) apply_px(k2, 1, x0, y0, px, py);
    else apply_px(4 - k2, 0, x0, y0, px, py);
    const float fx = (float)px, fy = (float)py;
    int blo, bhi;
    pixel_bins<FAN>(S, rview, fx, fy, cT, sT, blo, bhi);
    if (blo > bhi) return;
    __amdgpu_buffer_rsrc_t rs[8];
    int soff[8];
    bool flip[8];
#pragma unroll
    for (int h = 0; h < 8; ++h) {
        if (!(HMASK >> h & 1u)) continue;
        const int t = tab[h];
        const bool none = (t & kRowNone) != 0;
        rs[h] = none ? RZ : RS;
        soff[h] = none ? 0 : (t & ~(kRowFlip | kRowNone));
        flip[h] = (t & kRowFlip) != 0;
    }
    const float qd = FAN ? fan_depth_scale(fx, fy, cT, sT, S.invd) : 1.0f;
    float part[8];
#pragma unroll
    for (int h = 0; h < 8; ++h) part[h] = 0.0f;
    const int two_pc = 2 * S.pC, smax = S.W - WIN;
    for (int s0 = blo; s0 <= bhi; s0 += WIN) {
        const int s = min(s0, smax);
        // windows [s, s+WIN-1] and the reflected [2pC-s-WIN+1, 2pC-s]
        const int ob = s * 4, of = (two_pc - s - (WIN - 1)) * 4;
        Win<WIN> q[8];
#pragma unroll
        for (int h = 0; h < 8; ++h)
            if (HMASK >> h & 1u) q[h] = ldbw<WIN>(rs[h], flip[h] ? of : ob, soff[h]);
#pragma unroll
        for (int j = 0; j < WIN; ++j) {
            const int b = s + j;
            if (b < s0 || b > bhi) continue;
            Ray r;
            if (FAN) r = rview[b];
            else {
                r = rview[0];
                r.dc = (float)(b - S.pC) * S.dSize;
            }
            const float dist = ray_dist(fx, fy, r.cosR, r.sinR, r.dc);
            const float w = strip_weight(dist, FAN ? r.hbc * qd : r.hbc, r);
#pragma unroll
            for (int h = 0; h < 8; ++h)
                if (HMASK >> h & 1u)
                    part[h] = fmaf(w, flip[h] ? q[h].q[WIN - 1 - j] : q[h].q[j], part[h]);
        }
    }
    // FBP fan weight d^2/qe^2 (prj/cpuPrj.c:258); qe of g'^-1 q0 at a0 is the depth of
    // every orbit pixel at its own view
    const float fw = (FBP && FAN) ? 1.0f / (qd * qd) : 1.0f;
#pragma unroll
    for (int h = 0; h < 8; ++h)
        if (HMASK >> h & 1u) vpart[h] += (FBP && FAN) ? part[h] * fw : part[h];
}

// One fan-beam ray record through a buffer view of the slice's ray table (byte offsets:
// voff the bin, soff the view row).
__device__ __forceinline__ Ray ldray(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    const auto a = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
    const auto b = __builtin_amdgcn_raw_buffer_load_b128(r, voff + 16, soff, 0);
    Ray x;
    x.cosR = __uint_as_float(a[0]);
    x.sinR = __uint_as_float(a[1]);
    x.dc = __uint_as_float(a[2]);
    x.hbc = __uint_as_float(a[3]);
    x.h = __uint_as_float(b[0]);
    x.b1 = __uint_as_float(b[1]);
    x.s = __uint_as_float(b[2]);
    x.k = __uint_as_float(b[3]);
    return x;
}

// The same gather for a canonical view off the mirror lines and a compile-time element
// g' = E2: the reflection of every view row is fixed (elem_flip), so windows and bins are
// indexed without selects.  Sums go straight into the view's partials; views beyond np
// read the zero row (a wave-uniform choice of buffer), which adds exact zeros.  Fan beam:
// the rays of a window's first three bins are loaded together before the image windows
// (a pixel meets at most three bins when the detector pitch is at least the pixel's;
// the fourth is loaded where needed), RR being the slice's ray table and rrow the
// view's row.
template <bool FAN, bool FBP, unsigned HMASK, int E2, bool FOLD, int WIN>
__device__ __forceinline__ void adj_sym_elem_ct(const SliceDesc& S, const Ray* __restrict__ rview,
                                                __amdgpu_buffer_rsrc_t RS, __amdgpu_buffer_rsrc_t RZ,
                                                __amdgpu_buffer_rsrc_t RR, int rrow,
                                                const int* __restrict__ tab, float cT, float sT,
                                                int x0, int y0, float* vpart)
{
    constexpr int K2 = E2 & 3, M2 = E2 >> 2;
    int px, py;
    if (M2) apply_px(K2, 1, x0, y0, px, py);
    else apply_px(4 - K2, 0, x0, y0, px, py);
    const float fx = (float)px, fy = (float)py;
    int blo, bhi;
    pixel_bins<FAN>(S, rview, fx, fy, cT, sT, blo, bhi);
    if (blo > bhi) return;
    // rows of views beyond np read the zero row (wave-uniform resource selects)
    __amdgpu_buffer_rsrc_t rs[8];
    int soff[8];
#pragma unroll
    for (int h = 0; h < 8; ++h) {
        if (!(HMASK >> h & 1u)) continue;
        const int t = tab[h];
        const bool none = (t & kRowNone) != 0;
        rs[h] = none ? RZ : RS;
        soff[h] = none ? 0 : (t & ~(kRowFlip | kRowNone));
    }
    const float qd = FAN ? fan_depth_scale(fx, fy, cT, sT, S.invd) : 1.0f;
    // FBP fan weight d^2/qe^2 (prj/cpuPrj.c:258), folded into each weight
    const float fw = (FBP && FAN) ? 1.0f / (qd * qd) : 1.0f;
    const int two_pc = 2 * S.pC, smax = S.W - WIN;
    for (int s0 = blo; s0 <= bhi; s0 += WIN) {
        const int s = min(s0, smax);
        // windows [s, s+WIN-1] and the reflected [2pC-s-WIN+1, 2pC-s]
        const int ob = s * 4, of = (two_pc - s - (WIN - 1)) * 4;
        Ray rq[3];
        if (FAN) {
#pragma unroll
            for (int j = 0; j < 3; ++j) rq[j] = ldray(RR, (s + j) * (int)sizeof(Ray), rrow);
        }
        Win<WIN> q[8];
#pragma unroll
        for (int h = 0; h < 8; ++h)
            if (HMASK >> h & 1u) q[h] = ldbw<WIN>(rs[h], elem_flip(h, E2, FOLD) ? of : ob, soff[h]);
#pragma unroll
        for (int j = 0; j < WIN; ++j) {
            const int b = s + j;
            Ray r;
            if (FAN) r = j < 3 ? rq[j] : ldray(RR, b * (int)sizeof(Ray), rrow);
            else {
                r = rview[0];
                r.dc = (float)(b - S.pC) * S.dSize;
            }
            const float dist = ray_dist(fx, fy, r.cosR, r.sinR, r.dc);
            float w = strip_weight(dist, FAN ? r.hbc * qd : r.hbc, r);
            if (FBP && FAN) w *= fw;
            // bins of the window outside [s0, bhi] (taken by the previous window, or past
            // the pixel's candidates) add an exact zero
            w = (b >= s0 && b <= bhi) ? w : 0.0f;
#pragma unroll
            for (int h = 0; h < 8; ++h)
                if (HMASK >> h & 1u)
                    vpart[h] = fmaf(w, q[h].q[elem_flip(h, E2, FOLD) ? WIN - 1 - j : j], vpart[h]);
        }
    }
}

// ------------------------------------------------- LDS-staged adjoint gather ---
//
// The direct gather above has every lane load its own 3- or 4-bin windows of the eight
// sinogram rows and (fan beam) three 32-byte ray records per window, so neighbouring
// lanes fetch the same bins again through the vector-memory pipeline, the kernel's
// busiest unit.  The staged gather lets the wave fetch each (canonical view, element g')
// once: the bins [L0, L0 + kStageBins) that any of its 64 orbit pixels can read (the
// wave minimum of the pixels' first candidate bins, up to the maximum of their last
// candidate plus the window width), one load per sinogram row and one for the ray
// records, into a wave-private LDS area; the pixels then read their windows from LDS.
// Every pixel still forms the same weights from the same ray records and sums its bins
// in the same order, so the image is bit-identical.  A wave whose range exceeds
// kStageBins bins (fine detector pitch) stages it in chunks of kStageBins bins.
constexpr int kStageBins = 32;

struct AdjStage {
    float4 ray[2 * kStageBins];  // fan beam: bin k's record at ray[2k], ray[2k+1]
    // bin k's value in orbit image h's row at val[k][h]; the 48-byte stride spreads a
    // staging store's lanes (bins 4 q + j of rows h) over the LDS banks
    float val[kStageBins][12];
};

// Byte offset past the end of every slice's sinogram buffer (whose size is below 2^31
// bytes): a 16-byte raw buffer load there is out of range in every dword and reads 0.
constexpr int kStageNoRow = 0x7ffffff0;

// Wave minimum of lo and maximum of hi: prefix reductions inside each row of 16 lanes
// (DPP row shifts; lanes with no source read the operation's identity, which lets the
// compiler fold each shift into the min / max), then the four rows' last lanes.  The
// two chains are interleaved so that each DPP read finds its operand already written.
__device__ __forceinline__ void wave_range_i32(int lo, int hi, int& wlo, int& whi)
{
    lo = min(lo, __builtin_amdgcn_update_dpp(INT_MAX, lo, 0x111, 0xf, 0xf, false));
    hi = max(hi, __builtin_amdgcn_update_dpp(INT_MIN, hi, 0x111, 0xf, 0xf, false));
    lo = min(lo, __builtin_amdgcn_update_dpp(INT_MAX, lo, 0x112, 0xf, 0xf, false));
    hi = max(hi, __builtin_amdgcn_update_dpp(INT_MIN, hi, 0x112, 0xf, 0xf, false));
    lo = min(lo, __builtin_amdgcn_update_dpp(INT_MAX, lo, 0x114, 0xf, 0xf, false));
    hi = max(hi, __builtin_amdgcn_update_dpp(INT_MIN, hi, 0x114, 0xf, 0xf, false));
    lo = min(lo, __builtin_amdgcn_update_dpp(INT_MAX, lo, 0x118, 0xf, 0xf, false));
    hi = max(hi, __builtin_amdgcn_update_dpp(INT_MIN, hi, 0x118, 0xf, 0xf, false));
    const int a = __builtin_amdgcn_readlane(lo, 15), b = __builtin_amdgcn_readlane(lo, 31);
    const int c = __builtin_amdgcn_readlane(lo, 47), d = __builtin_amdgcn_readlane(lo, 63);
    const int e = __builtin_amdgcn_readlane(hi, 15), f = __builtin_amdgcn_readlane(hi, 31);
    const int g = __builtin_amdgcn_readlane(hi, 47), h = __builtin_amdgcn_readlane(hi, 63);
    wlo = min(min(a, b), min(c, d));
    whi = max(max(e, f), max(g, h));
}

// Orders this wave's LDS writes before its later reads (and reads before later writes).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The gather of adj_sym_elem_ct through the wave's LDS area.  Every lane of the wave
// takes part (the staging loads are the wave's); lanes with valid == false or an empty
// candidate range add nothing.
// One chunk of adj_sym_elem_lds: bins [L0, L0 + kStageBins) staged by the wave, then this
// lane's candidates blo..bhi inside the chunk summed.  Every lane of the wave takes part.
template <bool FAN, bool FBP, unsigned HMASK>
__device__ __forceinline__ void adj_sym_chunk(const SliceDesc& S, const Ray* __restrict__ rview,
                                              __amdgpu_buffer_rsrc_t RS, __amdgpu_buffer_rsrc_t RR,
                                              int rrow, int e2, int tl, float cT, float sT,
                                              float fx, float fy, int blo, int bhi, int L0,
                                              float* vpart, AdjStage* __restrict__ st)
{
    const int lane = (int)threadIdx.x & 63;
    // lane 8 ql + hl fetches bins L0 + 4 ql .. L0 + 4 ql + 3 of orbit image hl's row with
    // one 16-byte load, so one wave load brings the eight rows' kStageBins bins.  Reflected
    // rows read the reversed block at 2pC - b; its offsets may fall outside the row (into
    // a neighbouring row, or below the array, which reads 0) only for bins beyond the
    // detector or the pixels' candidates, whose weights are exact zeros.  Rows of views
    // beyond np, and images outside HMASK, give 0.
    const int hl = lane & 7, ql = lane >> 3;
    wave_lds_sync();  // the previous element's reads are done
    // lane hl's row entry tab[hl] = entry 8 e2 + hl of the view, which lane 8 e2 + hl
    // holds in tl (images outside the group have kRowNone entries)
    const int t = __builtin_amdgcn_ds_bpermute((8 * e2 + hl) << 2, tl);
    // the row's reflection from its entry (equal to elem_flip, checked when the table is
    // built): a per-view value keeps the store addresses below from being hoisted out of
    // the view loop as eight per-element sets
    const bool fl = (t & kRowFlip) != 0;
    const int b0 = L0 + 4 * ql;
    // a row without an entry reads past the buffer's end, which gives zeros (no branch)
    const int voff = (t & kRowNone) ? kStageNoRow
                                    : (t & ~(kRowFlip | kRowNone)) + 4 * (fl ? 2 * S.pC - b0 - 3 : b0);
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(RS, voff, 0, 0);
    const float x0 = __uint_as_float(v[0]), x1 = __uint_as_float(v[1]);
    const float x2 = __uint_as_float(v[2]), x3 = __uint_as_float(v[3]);
    float4 rv;
    if (FAN) {  // ray records of bins L0 .. L0 + kStageBins - 1: 16 bytes per lane
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(RR, L0 * (int)sizeof(Ray) + lane * 16,
                                                             rrow, 0);
        rv = make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
                         __uint_as_float(v[3]));
    }
    // xj is bin b0 + j, or b0 + 3 - j of a reflected row: reversed in registers, so that
    // the stores keep per-lane addresses
    st->val[4 * ql + 0][hl] = fl ? x3 : x0;
    st->val[4 * ql + 1][hl] = fl ? x2 : x1;
    st->val[4 * ql + 2][hl] = fl ? x1 : x2;
    st->val[4 * ql + 3][hl] = fl ? x0 : x3;
    if (FAN) st->ray[lane] = rv;
    wave_lds_sync();  // staged values visible to every lane
    if (blo > bhi) return;  // no candidate of this lane in the chunk
    const float qd = FAN ? fan_depth_scale(fx, fy, cT, sT, S.invd) : 1.0f;
    // FBP fan weight d^2/qe^2 (prj/cpuPrj.c:258), folded into each weight
    const float fw = (FBP && FAN) ? 1.0f / (qd * qd) : 1.0f;
    // exactly the pixel's candidate bins, in increasing order: the per-lane gathers add
    // the same products plus exact zeros (bins of a window past bhi), which change no sum
    // (a partial sum that starts at +0 never becomes -0)
    for (int b = blo; b <= bhi; ++b) {
        const int k = b - L0;
        Ray r;
        if (FAN) {
            const float4 a = st->ray[2 * k], c = st->ray[2 * k + 1];
            r.cosR = a.x;
            r.sinR = a.y;
            r.dc = a.z;
            r.hbc = a.w;
            r.h = c.x;
            r.b1 = c.y;
            r.s = c.z;
            r.k = c.w;
        } else {
            r = rview[0];
            r.dc = (float)(b - S.pC) * S.dSize;
        }
        const float dist = ray_dist(fx, fy, r.cosR, r.sinR, r.dc);
        float w = strip_weight(dist, FAN ? r.hbc * qd : r.hbc, r);
        if (FBP && FAN) w *= fw;
        const float4 lo = *reinterpret_cast<const float4*>(&st->val[k][0]);
        const float4 hi = *reinterpret_cast<const float4*>(&st->val[k][4]);
        const float q[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
        for (int h = 0; h < 8; ++h)
            if (HMASK >> h & 1u) vpart[h] = fmaf(w, q[h], vpart[h]);
    }
}

// The gather of adj_sym_elem_ct through the wave's LDS area, for element g' = e2 (a
// canonical view off the mirror lines) and the orbit images in HMASK.  Every lane of the
// wave takes part; lanes with valid == false or no candidate add nothing.  A range wider
// than kStageBins bins (fine detector pitch) is staged and summed in chunks of kStageBins
// bins, in increasing order, so the sums are bit-identical to the per-lane gathers'.
template <bool FAN, bool FBP, unsigned HMASK>
__device__ __forceinline__ void adj_sym_elem_lds(const SliceDesc& S, const Ray* __restrict__ rview,
                                                 __amdgpu_buffer_rsrc_t RS,
                                                 __amdgpu_buffer_rsrc_t RR, int rrow, int e2,
                                                 int tl, float cT, float sT, int x0, int y0,
                                                 bool valid, float* vpart,
                                                 AdjStage* __restrict__ st)
{
    const int k2 = e2 & 3, m2 = e2 >> 2;
    int px, py;
    if (m2) apply_px(k2, 1, x0, y0, px, py);
    else apply_px(4 - k2, 0, x0, y0, px, py);
    const float fx = (float)px, fy = (float)py;
    int blo, bhi;
    pixel_bins<FAN>(S, rview, fx, fy, cT, sT, blo, bhi);
    if (!valid || blo > bhi) {
        blo = INT_MAX;
        bhi = INT_MIN;
    }
    int L0, Lmax;
    wave_range_i32(blo, bhi, L0, Lmax);
    // wave-uniform; one chunk unless the range is wider than kStageBins bins
    for (int c0 = L0; c0 <= Lmax; c0 += kStageBins)
        adj_sym_chunk<FAN, FBP, HMASK>(S, rview, RS, RR, rrow, e2, tl, cT, sT, fx, fy,
                                       max(blo, c0), min(bhi, c0 + kStageBins - 1), c0, vpart,
                                       st);
}

// All elements of a canonical view off the mirror lines, unrolled at compile time.
// With STAGE the wave stages each element's bins in LDS (every lane takes part),
// otherwise the valid lanes gather directly.
template <bool FAN, bool FBP, int GRP, bool FOLD, int WIN, bool STAGE>
__device__ __forceinline__ void adj_sym_view_ct(const SliceDesc& S, const Ray* __restrict__ rview,
                                                __amdgpu_buffer_rsrc_t RS, __amdgpu_buffer_rsrc_t RZ,
                                                __amdgpu_buffer_rsrc_t RR, int rrow,
                                                const int* __restrict__ tv, float cT, float sT,
                                                int x0, int y0, bool valid, float* vpart,
                                                AdjStage* __restrict__ st)
{
    constexpr GroupMasks G = group_masks(GRP);
    constexpr unsigned GSET = FOLD ? G.fold : G.all;
    // the view's 64 row entries, entry l in lane l (the staged gathers pick theirs with a
    // lane permute)
    const int tl = STAGE ? tv[threadIdx.x & 63] : 0;
#define CTPRJ_E(E2)                                                                             \
    if constexpr ((GSET >> E2) & 1u) {                                                           \
        if (STAGE)                                                                               \
            adj_sym_elem_lds<FAN, FBP, G.all>(S, rview, RS, RR, rrow, E2, tl, cT, sT, x0, y0,     \
                                              valid, vpart, st);                                 \
        else if (valid)                                                                          \
            adj_sym_elem_ct<FAN, FBP, G.all, E2, FOLD, WIN>(S, rview, RS, RZ, RR, rrow,          \
                                                            tv + 8 * E2, cT, sT, x0, y0, vpart); \
    }
    CTPRJ_E(0) CTPRJ_E(1) CTPRJ_E(2) CTPRJ_E(3) CTPRJ_E(4) CTPRJ_E(5) CTPRJ_E(6) CTPRJ_E(7)
#undef CTPRJ_E
}

// Canonical views i = i0, i0 + step, ... for the orbit of (x0, y0) under group GRP.
// Every lane of the wave runs the view loop (the staged gather needs them all); lanes
// with valid == false add nothing.
template <bool FAN, bool FBP, int GRP, int WIN, bool STAGE>
__device__ __forceinline__ void adj_sym_pixel(const SliceDesc& S, const float4* __restrict__ views,
                                              const Ray* __restrict__ rays,
                                              const float* __restrict__ sino,
                                              const float* __restrict__ zero_row,
                                              const int* __restrict__ symtab, int x0, int y0,
                                              bool valid, int i0, int step, KahanAcc* acc,
                                              AdjStage* __restrict__ st)
{
    constexpr GroupMasks G = group_masks(GRP);
    constexpr unsigned REF = G.all & ~G.rot;
    const __amdgpu_buffer_rsrc_t RS = slice_rsrc(sino + S.sino_off, S.np * S.W);
    const __amdgpu_buffer_rsrc_t RZ = slice_rsrc(zero_row, S.W);
    // fan: the slice's ray table (ncanon rows of W records, below 2 GiB)
    const __amdgpu_buffer_rsrc_t RR = slice_rsrc(
        reinterpret_cast<const float*>(rays + S.ray_off),
        FAN ? S.ncanon * S.W * (int)(sizeof(Ray) / sizeof(float)) : 0);
    for (int i = i0; i < S.ncanon; i += step) {
        const float4 T = views[S.view_off + i];
        const int a0 = __float_as_int(T.z);
        const Ray* rview = FAN ? rays + S.ray_off + (long long)i * S.W : rays + S.ray_off + i;
        const int* tv = symtab + S.tab_off + (long long)i * 64;
        const int rrow = FAN ? i * S.W * (int)sizeof(Ray) : 0;
        float vpart[8];  // this view's sums; compensated once per view
#pragma unroll
        for (int h = 0; h < 8; ++h) vpart[h] = 0.0f;
        if (a0 == 0 || G.q * a0 == S.P) {
            // rotations g' -> rotation orbit pixels, reflections g' -> reflected ones
            const unsigned grot = S.fold ? G.mfold_rot : G.rot;
            const unsigned gref = S.fold ? G.mfold_ref : REF;
            if (valid) {
#pragma unroll 1
                for (int e = 0; e < 8; ++e)
                    if (grot >> e & 1u)
                        adj_sym_elem<FAN, FBP, G.rot, WIN>(S, rview, RS, RZ, tv + 8 * e, T.x, T.y,
                                                           e & 3, e >> 2, x0, y0, vpart);
#pragma unroll 1
                for (int e = 0; e < 8; ++e)
                    if (gref >> e & 1u)
                        adj_sym_elem<FAN, FBP, REF, WIN>(S, rview, RS, RZ, tv + 8 * e, T.x, T.y,
                                                         e & 3, e >> 2, x0, y0, vpart);
            }
        } else if (GRP != kGrpZ2 && S.fold) {
            // a half-turn parallel scan takes the half set that maps the canonical range
            // into [0, 180) and reads the rest folded
            adj_sym_view_ct<FAN, FBP, GRP, GRP != kGrpZ2, WIN, STAGE>(
                S, rview, RS, RZ, RR, rrow, tv, T.x, T.y, x0, y0, valid, vpart, st);
        } else {
            adj_sym_view_ct<FAN, FBP, GRP, false, WIN, STAGE>(S, rview, RS, RZ, RR, rrow, tv, T.x,
                                                              T.y, x0, y0, valid, vpart, st);
        }
#pragma unroll
        for (int h = 0; h < 8; ++h)
            if (G.all >> h & 1u) acc[h].add(vpart[h]);
    }
}

// Even n: row 0 and column 0 lie outside the pixel grid the projector reads; the adjoint
// writes them as 0 (prj/cpuPrj.c:713-716).  The work-group that stores a slice's marked
// tile (AdjItem.pad) does it; no orbit pixel lies there.
__device__ __forceinline__ void sym_zero_border(const SliceDesc& S, float* __restrict__ img)
{
    float* out = img + S.img_off;
    for (int i = threadIdx.x; i < S.n; i += blockDim.x) {
        out[i] = 0.0f;
        out[(long long)i * S.n] = 0.0f;
    }
}

// Orbit base pixels of a work item: an 8x8 tile of the fundamental domain of the group,
//   D4: 0 <= y0 <= x0 <= H,  V4: 0 <= x0, y0 <= H,  Z2: 0 <= x0 <= H, -H <= y0 <= H
// (tiles start inside it; the upper ends are checked here).
template <int GRP>
__device__ __forceinline__ bool adj_sym_base(const SliceDesc& S, const AdjItem& it, int lane,
                                             int& x0, int& y0)
{
    x0 = it.col0 + (lane & 7);
    y0 = it.row0 + (lane >> 3);
    return x0 <= S.H && y0 <= (GRP == kGrpD4 ? x0 : S.H);
}

// The sum t of orbit image h of (x0, y0) written to the image, unless an earlier element
// (in element order) reaches the same pixel: each distinct orbit pixel is written once.
template <bool FBP, int GRP>
__device__ __forceinline__ void adj_sym_store(const SliceDesc& S, int x0, int y0, int h, float t,
                                              float* __restrict__ img)
{
    constexpr GroupMasks G = group_masks(GRP);
    int ox, oy;
    apply_px(h & 3, h >> 2, x0, y0, ox, oy);
#pragma unroll
    for (int g = 0; g < 8; ++g) {
        if (g >= h) break;
        if (!(G.all >> g & 1u)) continue;
        int gx, gy;
        apply_px(g & 3, g >> 2, x0, y0, gx, gy);
        if (gx == ox && gy == oy) return;
    }
    img[S.img_off + (long long)(oy + S.half) * S.n + ox + S.half] = FBP ? t * S.fbp_scale : t;
}

// The canonical views are cut into nsplit * SEG strided sets: wave w of this work-group
// takes i = split * SEG + w + j * nsplit * SEG.  nsplit == 1 writes the image; otherwise
// the work-group's sums go to part[h * 64 + lane] for k_adjoint_sym_combine.
template <bool FBP, int SEG, int GRP, int WIN, bool STAGE>
__device__ __forceinline__ void adj_sym_tile(const SliceDesc& S, const AdjItem& it,
                                             const float4* __restrict__ views,
                                             const Ray* __restrict__ rays,
                                             const float* __restrict__ sino,
                                             const float* __restrict__ zero_row,
                                             const int* __restrict__ symtab,
                                             float* __restrict__ img, float (*red)[8][64],
                                             int split, int nsplit, float* __restrict__ part,
                                             AdjStage* __restrict__ stage)
{
    constexpr GroupMasks G = group_masks(GRP);
    // the wave index is uniform over the wave: the view loop then runs on scalar registers
    const int lane = (int)threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    int x0, y0;
    const bool valid = adj_sym_base<GRP>(S, it, lane, x0, y0);
    KahanAcc acc[8];
    if (STAGE || valid) {
        const int i0 = split * SEG + wave, step = SEG * nsplit;
        AdjStage* st = STAGE ? stage + wave : nullptr;
        if (S.fan)
            adj_sym_pixel<true, FBP, GRP, WIN, STAGE>(S, views, rays, sino, zero_row, symtab, x0,
                                                      y0, valid, i0, step, acc, st);
        else
            adj_sym_pixel<false, FBP, GRP, WIN, STAGE>(S, views, rays, sino, zero_row, symtab, x0,
                                                       y0, valid, i0, step, acc, st);
    }
    if (SEG > 1) {
#pragma unroll
        for (int h = 0; h < 8; ++h)
            if (G.all >> h & 1u) red[wave][h][lane] = acc[h].s;
        __syncthreads();
    }
    if (!part && it.pad) sym_zero_border(S, img);
    if (!valid) return;
    // wave w finalises the elements h = w (mod SEG): the sum of the waves' partials in
    // wave order
#pragma unroll
    for (int h = 0; h < 8; ++h) {
        if (!(G.all >> h & 1u) || (SEG > 1 && h % SEG != wave)) continue;
        float t = acc[h].s;
        if (SEG > 1) {
            t = red[0][h][lane];
#pragma unroll
            for (int q = 1; q < SEG; ++q) t += red[q][h][lane];
        }
        if (part) part[h * 64 + lane] = t;
        else adj_sym_store<FBP, GRP>(S, x0, y0, h, t, img);
    }
}

// One launch per symmetry group; SEG waves split the canonical views (wave w takes
// i = w, w + SEG, ...) and then share the final sums; with nsplit > 1, nsplit
// work-groups share the views of a tile and leave their sums in part (512 floats per
// work-group) for k_adjoint_sym_combine.  STAGE: the gathers off the mirror views go
// through each wave's LDS area (adj_sym_elem_lds).  The staged 4-bin-window variants ask
// for at least five waves per SIMD (96 VGPRs; D4 would take 99, four waves); the same
// request spills on the staged 3-bin D4 variants, which keep the compiler's own schedule
// (the window width matters there only on the mirror views' per-lane gathers).
template <bool FBP, int SEG, int GRP, int WIN, bool STAGE>
__global__ __launch_bounds__(64 * SEG)
__attribute__((amdgpu_waves_per_eu((STAGE && WIN == 4) ? 5 : 1))) void k_adjoint_sym(const SliceDesc* __restrict__ slices,
                                                          const AdjItem* __restrict__ items,
                                                          const float4* __restrict__ views,
                                                          const Ray* __restrict__ rays,
                                                          const float* __restrict__ sino,
                                                          const float* __restrict__ zero_row,
                                                          const int* __restrict__ symtab,
                                                          float* __restrict__ img, int nitems,
                                                          int nsplit, float* __restrict__ part,
                                                          const RedJob job)
{
    __shared__ float red[SEG][8][64];
    __shared__ AdjStage stage[STAGE ? SEG : 1];
    if (job.out && blockIdx.x == 0 && threadIdx.x < 64) wave_sum_job(job);  // on the side
    int split, idx;
    split_item(nitems, split, idx);
    const AdjItem it = items[idx];
    if (it.slice < 0) return;  // uniform over the work-group
    const SliceDesc S = slices[it.slice];
    adj_sym_tile<FBP, SEG, GRP, WIN, STAGE>(S, it, views, rays, sino, zero_row, symtab, img, red,
                                            split, nsplit,
                                            part ? part + (long long)blockIdx.x * 512 : nullptr,
                                            stage);
}

// The nsplit partial sums of each orbit image of a tile, added in split order and
// written like the one-work-group path (512 threads: orbit image h = thread / 64 of base
// pixel lane = thread % 64).
template <bool FBP, int GRP>
__global__ __launch_bounds__(512) void k_adjoint_sym_combine(const SliceDesc* __restrict__ slices,
                                                             const AdjItem* __restrict__ items,
                                                             int nitems, int nsplit,
                                                             const float* __restrict__ part,
                                                             float* __restrict__ img)
{
    const AdjItem it = items[blockIdx.x];
    if (it.slice < 0) return;
    const SliceDesc S = slices[it.slice];
    constexpr GroupMasks G = group_masks(GRP);
    if (it.pad) sym_zero_border(S, img);
    const int h = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    int x0, y0;
    if (!(G.all >> h & 1u) || !adj_sym_base<GRP>(S, it, lane, x0, y0)) return;
    const float* p = part + (long long)blockIdx.x * 512 + threadIdx.x;
    float t = p[0];
    for (int k = 1; k < nsplit; ++k) t += p[(long long)k * nitems * 512];
    adj_sym_store<FBP, GRP>(S, x0, y0, h, t, img);
}


// ----------------------------------------------------------------- launchers ---

template <int GRP, bool BHC>
static void launch_forward_sym_g(int seg, int nsplit, float* part, int nitems,
                                 const SliceDesc* desc, const FwdItem* items, const float4* views,
                                 const Ray* rays, const float* img, const float* imgT, float* sino,
                                 const FwdBhc<BHC>& B, hipStream_t st)
{
    float* pp = nsplit > 1 ? part : nullptr;
    const dim3 grid((unsigned)nitems * (unsigned)nsplit);
#define CTPRJ_LAUNCH(SEG)                                                                       \
    hipLaunchKernelGGL((k_forward_sym<SEG, GRP, BHC>), grid, dim3(64 * SEG), 0, st, desc, items, \
                       views, rays, img, imgT, sino, nitems, nsplit, pp, B)
    switch (seg) {
        case 1: CTPRJ_LAUNCH(1); break;
        case 2: CTPRJ_LAUNCH(2); break;
        case 4: CTPRJ_LAUNCH(4); break;
        case 8: CTPRJ_LAUNCH(8); break;
        default: CTPRJ_LAUNCH(16); break;
    }
#undef CTPRJ_LAUNCH
    if (pp)
        hipLaunchKernelGGL((k_forward_sym_combine<GRP, BHC>), dim3((unsigned)nitems), dim3(512), 0,
                           st, desc, items, views, nitems, nsplit, pp, sino, B);
}

template <bool BHC>
static void launch_forward_sym_b(int grp, int seg, int nsplit, float* part, int nitems,
                                 const SliceDesc* desc, const FwdItem* items, const float4* views,
                                 const Ray* rays, const float* img, const float* imgT, float* sino,
                                 const FwdBhc<BHC>& B, hipStream_t st)
{
    if (grp == kGrpD4)
        launch_forward_sym_g<kGrpD4, BHC>(seg, nsplit, part, nitems, desc, items, views, rays,
                                          img, imgT, sino, B, st);
    else if (grp == kGrpV4)
        launch_forward_sym_g<kGrpV4, BHC>(seg, nsplit, part, nitems, desc, items, views, rays,
                                          img, imgT, sino, B, st);
    else
        launch_forward_sym_g<kGrpZ2, BHC>(seg, nsplit, part, nitems, desc, items, views, rays,
                                          img, imgT, sino, B, st);
}

void launch_forward_sym(int grp, int seg, int nsplit, float* part, int nitems,
                        const SliceDesc* desc, const FwdItem* items, const float4* views,
                        const Ray* rays, const float* img, const float* imgT, float* sino,
                        const BhcStore* bhc, hipStream_t st)
{
    if (bhc) {
        FwdBhc<true> B;
        static_cast<BhcStore&>(B) = *bhc;
        launch_forward_sym_b<true>(grp, seg, nsplit, part, nitems, desc, items, views, rays, img,
                                   imgT, sino, B, st);
    } else {
        launch_forward_sym_b<false>(grp, seg, nsplit, part, nitems, desc, items, views, rays, img,
                                    imgT, sino, FwdBhc<false>{}, st);
    }
}

template <bool FBP, int GRP, int WIN>
static void launch_adjoint_sym_t(int seg, bool stage, int nsplit, float* part, int nitems,
                                 const SliceDesc* desc, const AdjItem* items, const float4* views,
                                 const Ray* rays, const float* sino, const float* zero_row,
                                 const int* symtab, float* img, const RedJob& job, hipStream_t st)
{
    float* pp = nsplit > 1 ? part : nullptr;
    const dim3 grid((unsigned)nitems * (unsigned)nsplit);
#define CTPRJ_LAUNCH(SEG, STAGE)                                                                 \
    hipLaunchKernelGGL((k_adjoint_sym<FBP, SEG, GRP, WIN, STAGE>), grid, dim3(64 * SEG), 0, st,  \
                       desc, items, views, rays, sino, zero_row, symtab, img, nitems, nsplit, pp, \
                       job)
    if (stage) {
        switch (seg) {
            case 1: CTPRJ_LAUNCH(1, true); break;
            case 2: CTPRJ_LAUNCH(2, true); break;
            case 4: CTPRJ_LAUNCH(4, true); break;
            default: CTPRJ_LAUNCH(8, true); break;
        }
    } else {
        switch (seg) {
            case 1: CTPRJ_LAUNCH(1, false); break;
            case 2: CTPRJ_LAUNCH(2, false); break;
            case 4: CTPRJ_LAUNCH(4, false); break;
            default: CTPRJ_LAUNCH(8, false); break;
        }
    }
#undef CTPRJ_LAUNCH
    if (pp)
        hipLaunchKernelGGL((k_adjoint_sym_combine<FBP, GRP>), dim3((unsigned)nitems), dim3(512), 0,
                           st, desc, items, nitems, nsplit, pp, img);
}

template <bool FBP, int WIN>
static void launch_adjoint_sym_f(int grp, int seg, bool stage, int nsplit, float* part, int nitems,
                                 const SliceDesc* desc, const AdjItem* items, const float4* views,
                                 const Ray* rays, const float* sino, const float* zero_row,
                                 const int* symtab, float* img, const RedJob& job, hipStream_t st)
{
    if (grp == kGrpD4)
        launch_adjoint_sym_t<FBP, kGrpD4, WIN>(seg, stage, nsplit, part, nitems, desc, items,
                                               views, rays, sino, zero_row, symtab, img, job, st);
    else if (grp == kGrpV4)
        launch_adjoint_sym_t<FBP, kGrpV4, WIN>(seg, stage, nsplit, part, nitems, desc, items,
                                               views, rays, sino, zero_row, symtab, img, job, st);
    else
        launch_adjoint_sym_t<FBP, kGrpZ2, WIN>(seg, stage, nsplit, part, nitems, desc, items,
                                               views, rays, sino, zero_row, symtab, img, job, st);
}

void launch_adjoint_sym(bool fbp, int grp, int seg, int win, bool stage, int nsplit, float* part,
                        int nitems, const SliceDesc* desc, const AdjItem* items,
                        const float4* views, const Ray* rays, const float* sino,
                        const float* zero_row, const int* symtab, float* img, const RedJob& job,
                        hipStream_t st)
{
    if (fbp && win == 3)
        launch_adjoint_sym_f<true, 3>(grp, seg, stage, nsplit, part, nitems, desc, items, views,
                                      rays, sino, zero_row, symtab, img, job, st);
    else if (fbp)
        launch_adjoint_sym_f<true, 4>(grp, seg, stage, nsplit, part, nitems, desc, items, views,
                                      rays, sino, zero_row, symtab, img, job, st);
    else if (win == 3)
        launch_adjoint_sym_f<false, 3>(grp, seg, stage, nsplit, part, nitems, desc, items, views,
                                       rays, sino, zero_row, symtab, img, job, st);
    else
        launch_adjoint_sym_f<false, 4>(grp, seg, stage, nsplit, part, nitems, desc, items, views,
                                       rays, sino, zero_row, symtab, img, job, st);
}


}  // namespace ctprj
