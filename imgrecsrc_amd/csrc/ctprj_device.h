// Device-side data layout and the pixel/strip footprint weight shared by every
// projector kernel.  Forward and adjoint evaluate a weight through the same inline
// functions on the same float inputs (per-ray table entries + integer pixel
// coordinates), so A and A^T are bitwise transposes of each other.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ctprj {

// Margin (pixels) added to every support bound.  Bounds only select candidates: the
// weight itself is exactly 0 outside the footprint, so a margin never changes a result.
constexpr float kSupportMargin = 0.02f;

// Word 3 of a raw (stride 0) buffer resource on the gfx9 family, gfx950 included: 32-bit
// data format, identity component selects.  Loads beyond num_records return 0.
constexpr int kBufferWord3 = 0x00020000;

// Symmetry group of a slice's view set used by the symmetric kernels (SliceDesc::sym).
// R rotates the pixel grid by 90 degrees, M reflects it in the x axis; element
// e = k + 4m is R^k M^m.  The group is the largest whose view maps are integral:
//   D4 (all eight elements) when prjFull % 4 == 0,
//   V4 = {1, R^2, M, R^2 M} when prjFull is even,
//   Z2 = {1, R^2 M} (theta -> -theta) for any prjFull.
enum : int { kGrpNone = 0, kGrpD4 = 1, kGrpV4 = 2, kGrpZ2 = 3 };

// View-row table of the symmetric adjoint: for canonical view i, element g' = e2 and
// orbit element h, entry [i][e2][h] locates the sinogram row of view h g' a0 as
// (byte offset of the row) | kRowFlip (bins read at 2pC - b) or kRowNone (view beyond np,
// or h g' outside the slice's group).
constexpr int kRowFlip = 1, kRowNone = 2;

// Element sets of a group as bit masks over e, and its canonical view range.
struct GroupMasks {
    unsigned all;        // the group's elements
    unsigned rot;        // its rotations
    unsigned fold;       // adjoint elements of a half-turn parallel scan, off a mirror view
    unsigned mfold_rot;  // ... on a mirror view, paired with rotation orbit pixels
    unsigned mfold_ref;  // ... on a mirror view, paired with reflected orbit pixels
    int q;               // canonical views 0 <= a0 <= prjFull/q; mirror views a0 = 0, prjFull/q
    int order;           // number of elements
};

// Whether view h g' a0 reads its bins reflected (at 2pC - b), for a canonical view a0
// off the mirror lines: element (k, m) = h g' reflects the bins when m = 1, and a
// half-turn parallel scan folds views in [P/2, P) onto [0, P/2) with one more
// reflection.  For 0 < a0 < P/q the folded views are, for every group,
//   m = 0: k >= 2;   m = 1: k = 1 or 2.
__host__ __device__ constexpr bool elem_flip(int h, int e2, bool fold)
{
    const int k1 = h & 3, m1 = h >> 2, k2 = e2 & 3, m2 = e2 >> 2;
    const int k = (k1 + (m1 ? -k2 : k2)) & 3, m = m1 ^ m2;
    const bool folded = m ? (k == 1 || k == 2) : (k >= 2);
    return (m != 0) != (fold && folded);
}

__host__ __device__ constexpr GroupMasks group_masks(int grp)
{
    // D4 fold set {1, R, M, R^3 M} maps (0, 45) degrees into [0, 180); V4's {1, M} maps
    // (0, 90) into [0, 180).  Z2 never folds (prjFull odd).
    return grp == kGrpD4   ? GroupMasks{0xFFu, 0x0Fu, 0x93u, 0x03u, 0x30u, 8, 8}
           : grp == kGrpV4 ? GroupMasks{0x55u, 0x05u, 0x11u, 0x01u, 0x10u, 4, 4}
           : grp == kGrpZ2 ? GroupMasks{0x41u, 0x01u, 0x00u, 0x00u, 0x00u, 2, 2}
                           : GroupMasks{0x01u, 0x01u, 0x00u, 0x00u, 0x00u, 1, 1};
}

// One slice of a plan.  Offsets are element offsets into the packed arrays.
struct SliceDesc {
    int n, W, np, fan;     // image edge, bins per view, views, 1 = fan beam
    int sym, fold;         // symmetry group (kGrp*); parallel beam folded over 180 deg
    int P, ncanon;         // prjFull; canonical views held in the view / ray tables
    int H;                 // (n-1)/2: pixel coordinates run over [-H, H]
    int half;              // n/2: coordinate -> index offset
    int pC;                // W/2: index of the bin at t = 0
    int blo, bhi;          // bins that carry rays: [pC-(W-1)/2, pC+(W-1)/2]
    float dSize, invDSize; // detector pitch
    float d, invd;         // source distance (fan), 0 for parallel
    float hbd;             // half beam width on the detector: dSize*effectiveRate/2
    float er;              // effectiveRate (outputs are divided by it)
    float fbp_scale;       // PI/np/dSize^2 (prj/cpuPrj.c:161,304)
    long long img_off, sino_off, view_off, ray_off;
    long long tab_off;     // symmetric slices: first entry of the view-row table
};

// Per-ray footprint constants.  Parallel beam: one entry per view (dc unused, the bin
// position is (b-pC)*dSize).  Fan beam: one entry per (view, bin).
struct __attribute__((aligned(16))) Ray {
    float cosR, sinR;  // unit normal of the ray line: dist = x*cosR + y*sinR - dc
    float dc;          // offset of the ray line from the rotation centre
    float hbc;         // half strip width at the isocentre, perpendicular to the ray
    float h;           // trapezoid height / effectiveRate: 1/(max(|c|,|s|)*er)
    float b1;          // plateau half width (max-min)/2
    float s;           // ramp width min(|c|,|s|)
    float k;           // 1/(2 s), 0 when s == 0
};

// Cumulative integral, divided by the height, of the unit-area trapezoid footprint of a
// pixel (the projection of the unit square on the detector normal) up to offset u:
//   0 for u <= -b2, ramp (u+b2)^2/(2s) up to -b1, plateau, ramp down, c for u >= b2,
// with b2 = b1 + s.  Written with clamps (v_med3_f32), no branches.
__device__ __forceinline__ float gcum(float u, float b1, float b2, float s, float b1x2,
                                      float k)
{
    const float u1 = __builtin_amdgcn_fmed3f(u + b2, 0.0f, s);
    const float u2 = __builtin_amdgcn_fmed3f(u + b1, 0.0f, b1x2);
    const float u3 = __builtin_amdgcn_fmed3f(u - b1, 0.0f, s);
    return (u2 + u3) + k * ((u1 - u3) * (u1 + u3));
}

// Area of overlap of the unit pixel centred `dist` from the ray line with the strip of
// half width hb (prj/utils.h:12-92 computes the same area by a 20-way case split).
// Evaluated at |dist| like the reference (the footprint is symmetric), which makes the
// float result of the pixel reflected through the ray identical bit for bit.
__device__ __forceinline__ float strip_weight(float dist, float hb, const Ray& r)
{
    const float ad = fabsf(dist);
    const float b2 = r.b1 + r.s;
    const float b1x2 = r.b1 + r.b1;
    const float gp = gcum(hb - ad, r.b1, b2, r.s, b1x2, r.k);
    const float gm = gcum(-hb - ad, r.b1, b2, r.s, b1x2, r.k);
    return r.h * (gp - gm);
}

// Hardware reciprocal (about 1 ulp): used only where a result selects candidates and
// is padded by kSupportMargin, never in a weight.
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Signed distance of pixel (x, y) from the ray line.
__device__ __forceinline__ float ray_dist(float x, float y, float cosR, float sinR, float dc)
{
    return fmaf(x, cosR, fmaf(y, sinR, -dc));
}

// Fan beam: strip width at the pixel relative to the isocentre width,
// 1 + (x sinT - y cosT)/d (prj/cpuPrj.c:549, 254).
__device__ __forceinline__ float fan_depth_scale(float x, float y, float cosT, float sinT,
                                                 float invd)
{
    return fmaf(fmaf(x, sinT, -(y * cosT)), invd, 1.0f);
}

// Candidate range along a row (or column) of the walk: the pixels whose footprint can
// meet the strip, |dist| < b2 + hb(x, y), solved as two linear inequalities in the
// inner coordinate u for the walk coordinate v:
//   u in [min(U(v)*iu, L(v)*il), max(...)],  U(v) = U0 - v*U1,  L(v) = L0 - v*L1.
struct RowBounds {
    float U0, U1, L0, L1, iu, il;
};

__device__ __forceinline__ RowBounds row_bounds(const Ray& r, float cT, float sT, float invd,
                                                bool rowmode)
{
    RowBounds rb;
    const float b2 = r.b1 + r.s;
    const float m = kSupportMargin;
    const float al = r.hbc * sT * invd;  // d(hb)/dx
    const float be = r.hbc * cT * invd;  // -d(hb)/dy
    rb.U0 = r.dc + b2 + r.hbc + m;
    rb.L0 = r.dc - b2 - r.hbc - m;
    if (rowmode) {  // u = x, v = y
        rb.U1 = r.sinR + be;
        rb.L1 = r.sinR - be;
        rb.iu = 1.0f / (r.cosR - al);
        rb.il = 1.0f / (r.cosR + al);
    } else {        // u = y, v = x
        rb.U1 = r.cosR - al;
        rb.L1 = r.cosR + al;
        rb.iu = 1.0f / (r.sinR + be);
        rb.il = 1.0f / (r.sinR - be);
    }
    return rb;
}

__device__ __forceinline__ void row_range(const RowBounds& rb, float v, int H, int& lo, int& hi)
{
    const float ua = fmaf(-v, rb.U1, rb.U0) * rb.iu;
    const float ub = fmaf(-v, rb.L1, rb.L0) * rb.il;
    const float fH = (float)H;
    lo = (int)ceilf(fmaxf(fminf(ua, ub), -fH - 1.0f));
    hi = (int)floorf(fminf(fmaxf(ua, ub), fH + 1.0f));
    lo = max(lo, -H);
    hi = min(hi, H);
}

// Compensated accumulation of a partial sum (Kahan): keeps the long sums over rows /
// views at FP32 rounding of the result instead of growing with the term count.
struct KahanAcc {
    float s = 0.0f, c = 0.0f;
    __device__ __forceinline__ void add(float v)
    {
        const float y = v - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

}  // namespace ctprj
