// Host side of libctprj: plans, geometry tables, work lists, launches and the
// reference-compatible single-slice entry points (prj_setup / prj_run).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <atomic>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "bhc_internal.h"
#include "ctprj_internal.h"

namespace ctprj {

// ------------------------------------------------------------------- errors ---

static thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char* what)
{
    return fail(CTPRJ_EHIP, "%s: %s", what, hipGetErrorString(e));
}

// ------------------------------------------------------------------ geometry ---

static constexpr float kRefPi = 3.14159265359f;  // prj/prj.h:9

// cos/sin of view a as the reference computes them: float theta = a*2*PI/prjFull
// (prj/cpuPrj.c:342), cos()/sin() of that float evaluated in double, stored as float.
static void view_trig(int a, int prjFull, float* c, float* s)
{
    float th = (float)(a * 2) * kRefPi;
    th = th / (float)prjFull;
    *c = (float)std::cos((double)th);
    *s = (float)std::sin((double)th);
}

// Per-bin fan constants in the reference's float expressions (prj/cpuPrj.c:492-507):
// t = (b-pC)*dSize, bq = sqrt(d*d+t*t), cosB = d/bq, sinB = t/bq, d*t/bq,
// beamWidth*cosB.
static void fan_bin(int b, const ctprj_geom& g, float out[4])
{
    const int pC = g.prjWidth / 2;
    const float t = (float)(b - pC) * g.dSize;
    const float d = g.d;
    const float dd = d * d, tt = t * t;
    const float bq = (float)std::sqrt((double)(dd + tt));
    const float cosB = d / bq, sinB = t / bq;
    const float dt = d * t;
    out[0] = cosB;
    out[1] = sinB;
    out[2] = dt / bq;
    out[3] = (g.dSize * g.effectiveRate) * cosB;
}

int validate_geom(const ctprj_geom& g, int idx)
{
    if (g.n < 1 || g.n > 65536) return fail(CTPRJ_EINVAL, "slice %d: n=%d out of range", idx, g.n);
    if (g.prjWidth < 1 || g.prjWidth > (1 << 20))
        return fail(CTPRJ_EINVAL, "slice %d: prjWidth=%d out of range", idx, g.prjWidth);
    if (g.np < 1 || g.np > (1 << 20)) return fail(CTPRJ_EINVAL, "slice %d: np=%d out of range", idx, g.np);
    if (g.prjFull < 1) return fail(CTPRJ_EINVAL, "slice %d: prjFull=%d must be >= 1", idx, g.prjFull);
    if (!(g.dSize > 0.0f) || !std::isfinite(g.dSize))
        return fail(CTPRJ_EINVAL, "slice %d: dSize must be > 0", idx);
    if (!(g.effectiveRate > 0.0f) || !(g.effectiveRate <= 4.0f))
        return fail(CTPRJ_EINVAL, "slice %d: effectiveRate must be in (0, 4]", idx);
    if (!std::isfinite(g.d)) return fail(CTPRJ_EINVAL, "slice %d: d not finite", idx);
    if (g.d > 0.0f) {
        // the source must lie outside the image disc (pixel corners at radius <= R)
        const double R = ((g.n - 1) / 2 + 0.5) * std::sqrt(2.0);
        if (g.d <= R + 1.0)
            return fail(CTPRJ_EINVAL, "slice %d: fan source distance d=%g lies inside the image "
                        "(needs d > %.1f)", idx, (double)g.d, R + 1.0);
        const double tmax = (g.prjWidth / 2 + 1) * (double)g.dSize;
        if (tmax > 2.0 * g.d)
            return fail(CTPRJ_EINVAL, "slice %d: detector half width %.1f too wide for d=%g", idx,
                        tmax, (double)g.d);
    }
    // the kernels address one slice's image and sinogram with 32-bit byte offsets
    const double sino = (double)g.np * g.prjWidth, img = (double)g.n * g.n;
    if (4.0 * sino >= 2147483648.0 || 4.0 * img >= 2147483648.0)
        return fail(CTPRJ_EINVAL, "slice %d: image or sinogram exceeds 2 GiB", idx);
    return 0;
}

// ------------------------------------------------------------------- device ---

static int check_device(int* dev_out)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n < 1) {
        (void)hipGetLastError();
        return fail(CTPRJ_ENODEV, "no HIP device visible (%s)",
                    e == hipSuccess ? "0 devices" : hipGetErrorString(e));
    }
    int dev = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return hip_fail(e, "hipGetDevice");
    hipDeviceProp_t p;
    if ((e = hipGetDeviceProperties(&p, dev)) != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
    if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
        return fail(CTPRJ_ENODEV, "device %d is %s; this library is built for gfx950", dev,
                    p.gcnArchName);
    *dev_out = dev;
    return 0;
}

// -------------------------------------------------------------------- plans ---

static void free_dev(void* p)
{
    if (p) (void)hipFree(p);
}

template <class T>
static void free_ptr(T*& p)
{
    free_dev((void*)p);
    p = nullptr;
}

void Tables::release()
{
    free_ptr(d_desc);
    free_ptr(d_views);
    free_ptr(d_rays);
    free_ptr(d_symtab);
    free_ptr(d_tr_items);
    free_ptr(d_fwd_items);
    free_ptr(d_symf_items);
    free_ptr(d_adj_items);
    free_ptr(d_syma_items);
}

void Tables::swap_geometry(Tables& o)
{
    std::swap(d_desc, o.d_desc);
    std::swap(d_views, o.d_views);
    std::swap(d_rays, o.d_rays);
    std::swap(d_symtab, o.d_symtab);
    std::swap(d_tr_items, o.d_tr_items);
    std::swap(n_tr_items, o.n_tr_items);
    std::swap(n_sym, o.n_sym);
    slice_sym.swap(o.slice_sym);
}

void Tables::swap_fwd(Tables& o)
{
    std::swap(d_fwd_items, o.d_fwd_items);
    std::swap(d_symf_items, o.d_symf_items);
    std::swap(n_fwd_items, o.n_fwd_items);
    std::swap(n_symf_items, o.n_symf_items);
    for (int g = 0; g < 4; ++g) {
        std::swap(symf_start[g], o.symf_start[g]);
        std::swap(symf_count[g], o.symf_count[g]);
    }
}

void Tables::swap_adj(Tables& o)
{
    std::swap(d_adj_items, o.d_adj_items);
    std::swap(d_syma_items, o.d_syma_items);
    std::swap(n_adj_items, o.n_adj_items);
    std::swap(n_syma_items, o.n_syma_items);
    for (int g = 0; g < 4; ++g) {
        std::swap(syma_start[g], o.syma_start[g]);
        std::swap(syma_count[g], o.syma_count[g]);
    }
}

Plan::~Plan()
{
    free_ptr(d_fbp_items);
    free_ptr(d_zero_row);
    free_ptr(d_fbp_kern);
    free_ptr(d_fbp_kern_off);
    free_ptr(d_fbp_prew);
    free_ptr(d_fbp_prew_off);
    free_ptr(d_imgT);
    free_ptr(d_sino_ws);
    free_ptr(d_red);
    free_ptr(d_red_out);
    free_ptr(d_sim);
    free_ptr(d_ie_phi);
    free_ptr(d_split);
}

// Failure injection for the error-path tests: when armed with k > 0, the k-th device
// allocation from then on fails as if the device were out of memory (once).
static std::atomic<int> g_fail_alloc{0};

static bool injected_alloc_failure()
{
    int k = g_fail_alloc.load();
    while (k > 0) {
        if (g_fail_alloc.compare_exchange_weak(k, k - 1)) return k == 1;
    }
    return false;
}

template <class T>
static int upload(T** dst, const std::vector<T>& src, const char* what)
{
    *dst = nullptr;
    if (src.empty()) return 0;
    hipError_t e = injected_alloc_failure() ? hipErrorOutOfMemory
                                            : hipMalloc((void**)dst, sizeof(T) * src.size());
    if (e != hipSuccess) return fail(CTPRJ_ENOMEM, "hipMalloc(%s, %zu B): %s", what,
                                     sizeof(T) * src.size(), hipGetErrorString(e));
    e = hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, what);
    return 0;
}

static int dev_alloc(void** p, size_t bytes, const char* what)
{
    hipError_t e = injected_alloc_failure() ? hipErrorOutOfMemory : hipMalloc(p, bytes ? bytes : 16);
    if (e != hipSuccess) return fail(CTPRJ_ENOMEM, "hipMalloc(%s, %zu B): %s", what, bytes,
                                     hipGetErrorString(e));
    return 0;
}

// Batch order.  The hardware hands work-group i to XCD i % 8, so the list is built as
// eight per-XCD queues interleaved (queue k at positions k, k + 8, ...).  Slices go to
// queues heaviest first (work n^2 np).  `spread` (the default, tune flag bit 3 clear):
// every slice is split into contiguous eighths, one per queue, which balances the XCDs
// whatever the per-slice cost.  Otherwise only slices above a quarter of one XCD's
// share are split and the others go whole to the least loaded queue, so each XCD's L2
// serves whole slices (uniform stacks).  Shorter queues are padded with empty items
// (slice = -1).
template <class Item>
static std::vector<Item> batch_order(const std::vector<Item>& in,
                                     const std::vector<ctprj_geom>& geoms, bool spread)
{
    std::vector<std::vector<Item>> per(geoms.size());
    for (const Item& it : in) per[it.slice].push_back(it);
    std::vector<int> order;
    double total = 0;
    auto work = [&](int s) { return (double)geoms[s].n * geoms[s].n * geoms[s].np; };
    for (int s = 0; s < (int)geoms.size(); ++s)
        if (!per[s].empty()) {
            order.push_back(s);
            total += work(s);
        }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work(a) > work(b); });
    std::vector<Item> q[8];
    double load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int s : order) {
        const std::vector<Item>& its = per[s];
        if (spread || work(s) > total / 32.0) {
            const size_t chunk = (its.size() + 7) / 8;
            for (int k = 0; k < 8; ++k) {
                const size_t a = std::min(its.size(), k * chunk), b = std::min(its.size(), a + chunk);
                q[k].insert(q[k].end(), its.begin() + a, its.begin() + b);
                load[k] += work(s) * (double)(b - a) / (double)its.size();
            }
        } else {
            const int k = (int)(std::min_element(load, load + 8) - load);
            q[k].insert(q[k].end(), its.begin(), its.end());
            load[k] += work(s);
        }
    }
    size_t L = 0;
    for (int k = 0; k < 8; ++k) L = std::max(L, q[k].size());
    std::vector<Item> out(8 * L);
    for (size_t i = 0; i < L; ++i)
        for (int k = 0; k < 8; ++k) {
            Item& o = out[8 * i + k];
            if (i < q[k].size()) o = q[k][i];
            else { o = Item{}; o.slice = -1; }
        }
    return out;
}

// Canonical views of the symmetric decomposition under group `grp`: a0 in [0, P/q]
// whose images (g a0, folded for a half-turn parallel scan) include a computed view.
// Returns the share of useful view evaluations (1 for a full or a half-turn scan).
static double symmetric_views(const ctprj_geom& g, int grp, bool fold, std::vector<int>* canon)
{
    const GroupMasks G = group_masks(grp);
    const int P = g.prjFull;
    long long slots = 0;
    for (int a0 = 0; G.q * a0 <= P; ++a0) {
        const bool mirror = (a0 == 0) || (G.q * a0 == P);
        const unsigned els = mirror ? G.rot : G.all;
        bool any = false;
        int count = 0;
        for (int e = 0; e < 8; ++e) {
            if (!(els >> e & 1u)) continue;
            const int k = e & 3, m = e >> 2;
            int v = ((m ? -4 * a0 : 4 * a0) + ((k + 2 * m) & 3) * P) / 4;
            v %= P;
            if (v < 0) v += P;
            if (fold && 2 * v >= P) v -= P / 2;
            if (v < g.np) any = true;
            ++count;
        }
        if (!any) continue;
        canon->push_back(a0);
        slots += fold ? count / 2 : count;
    }
    return slots > 0 ? (double)g.np / (double)slots : 0.0;
}

// Symmetry group of a slice under `flags` (bit 0: symmetric kernels on; bit 1: skip D4;
// bit 2: skip V4): the largest group whose view maps are integral and whose canonical
// evaluations are at least half useful, else kGrpNone (also for images under 5x5).
static int choose_group(const ctprj_geom& g, int flags, bool* fold, std::vector<int>* canon)
{
    canon->clear();
    *fold = false;
    // the symmetric kernels read 4-wide windows inside image rows (n >= 5) and sinogram
    // rows (prjWidth >= 4)
    if (!(flags & 1) || g.n < 5 || g.prjWidth < 4 || g.prjFull < 1 || g.np > g.prjFull)
        return kGrpNone;
    const bool fan = g.d > 0.0f;
    const int P = g.prjFull;
    const int cands[3] = {kGrpD4, kGrpV4, kGrpZ2};
    for (int grp : cands) {
        if (grp == kGrpD4 && (P % 4 || (flags & 2))) continue;
        if (grp == kGrpV4 && (P % 2 || (flags & 4))) continue;
        const bool fo = !fan && grp != kGrpZ2 && 2 * g.np <= P;
        std::vector<int> c;
        // the symmetric adjoint addresses a fan slice's ray table with 32-bit byte offsets
        if (fan && g.prjWidth * (double)sizeof(Ray) * (P / 2 + 2) >= 2147483648.0) break;
        if (symmetric_views(g, grp, fo, &c) >= 0.5) {
            *fold = fo;
            *canon = c;
            return grp;
        }
    }
    return kGrpNone;
}

// View of element (k, m) applied to canonical view a0 (the kernels' map_view), and
// whether its bins are reflected.
static int host_map_view(int P, bool fold, int k, int m, int a0, bool* flip)
{
    int v = ((m ? -4 * a0 : 4 * a0) + ((k + 2 * m) & 3) * P) / 4;
    v %= P;
    if (v < 0) v += P;
    bool f = m != 0;
    if (fold && 2 * v >= P) {
        v -= P / 2;
        f = !f;
    }
    *flip = f;
    return v;
}

int Plan::build_geometry(int flags, Tables& G) const
{
    std::vector<SliceDesc> desc(nslices);
    std::vector<float4> views, bins;
    std::vector<long long> bin_off(nslices), ray_prefix(nslices + 1, 0);
    std::vector<int> symtab;
    long long nview = 0, nray = 0;
    G.slice_sym.assign(nslices, 0);
    G.n_sym = 0;
    for (int s = 0; s < nslices; ++s) {
        const ctprj_geom& g = geoms[s];
        SliceDesc& D = desc[s];
        D = SliceDesc{};
        D.n = g.n;
        D.W = g.prjWidth;
        D.np = g.np;
        D.fan = g.d > 0.0f;
        D.H = (g.n - 1) / 2;
        D.half = g.n / 2;
        D.pC = g.prjWidth / 2;
        D.blo = D.pC - (g.prjWidth - 1) / 2;
        D.bhi = D.pC + (g.prjWidth - 1) / 2;
        D.dSize = g.dSize;
        D.invDSize = 1.0f / g.dSize;
        D.d = D.fan ? g.d : 0.0f;
        D.invd = D.fan ? 1.0f / g.d : 0.0f;
        D.hbd = 0.5f * (g.dSize * g.effectiveRate);
        D.er = g.effectiveRate;
        D.fbp_scale = kRefPi / (float)g.np / g.dSize / g.dSize;
        D.img_off = img_offsets[s];
        D.sino_off = sino_offsets[s];
        D.view_off = nview;
        D.ray_off = nray;
        std::vector<int> canon;
        bool fold = false;
        const int grp = choose_group(g, flags, &fold, &canon);
        const bool sym = grp != kGrpNone;
        if (sym) {
            D.sym = grp;
            D.fold = fold;
            D.P = g.prjFull;
            D.ncanon = (int)canon.size();
            D.tab_off = (long long)symtab.size();
            const unsigned all = group_masks(grp).all;
            for (int a0 : canon)
                for (int e2 = 0; e2 < 8; ++e2)
                    for (int h = 0; h < 8; ++h) {
                        if (!(all >> e2 & 1u) || !(all >> h & 1u)) {
                            symtab.push_back(kRowNone);
                            continue;
                        }
                        const int k1 = h & 3, m1 = h >> 2, k2 = e2 & 3, m2 = e2 >> 2;
                        bool flip = false;
                        const int v = host_map_view(g.prjFull, fold, k1 + (m1 ? -k2 : k2),
                                                    m1 ^ m2, a0, &flip);
                        // the kernels take the reflection of off-mirror views from
                        // elem_flip() at compile time
                        const bool mirror = a0 == 0 || group_masks(grp).q * a0 == g.prjFull;
                        if (!mirror && flip != elem_flip(h, e2, fold))
                            return fail(CTPRJ_EINVAL, "internal: reflection table mismatch "
                                        "(slice %d, view %d, elements %d/%d)", s, a0, e2, h);
                        symtab.push_back(v < g.np ? v * g.prjWidth * 4 + (flip ? kRowFlip : 0)
                                                  : kRowNone);
                    }
            for (int a0 : canon) {
                float c, si;
                view_trig(a0, g.prjFull, &c, &si);
                float tag;
                std::memcpy(&tag, &a0, sizeof(tag));
                views.push_back(make_float4(c, si, tag, 0.0f));
            }
            G.slice_sym[s] = grp;
            ++G.n_sym;
        } else {
            for (int a = 0; a < g.np; ++a) {
                float c, si;
                view_trig(a, g.prjFull, &c, &si);
                views.push_back(make_float4(c, si, 0.0f, 0.0f));
            }
        }
        const int ntab = sym ? D.ncanon : g.np;
        nview += ntab;
        bin_off[s] = (long long)bins.size();
        if (D.fan) {
            for (int b = 0; b < g.prjWidth; ++b) {
                float q[4];
                fan_bin(b, g, q);
                bins.push_back(make_float4(q[0], q[1], q[2], q[3]));
            }
            nray += (long long)ntab * g.prjWidth;
        } else {
            nray += ntab;
        }
        ray_prefix[s + 1] = nray;
    }
    // the transposed image serves the general forward's steep views, D4's odd rotations
    // and the column walks of V4 / Z2 canonical views beyond 45 degrees
    std::vector<AdjItem> tr;
    for (int s = 0; s < nslices; ++s) {
        for (int r = 0; r < geoms[s].n; r += 32)
            for (int c = 0; c < geoms[s].n; c += 32) tr.push_back(AdjItem{s, r, c, 0});
    }
    G.n_tr_items = (int)tr.size();
    float4* d_bins = nullptr;
    long long *d_bin_off = nullptr, *d_ray_prefix = nullptr;
    int rc;
    if ((rc = upload(&G.d_desc, desc, "slice table")) ||
        (rc = upload(&G.d_views, views, "view table")) ||
        (rc = upload(&d_bins, bins, "bin table")) || (rc = upload(&d_bin_off, bin_off, "bin offsets")) ||
        (rc = upload(&d_ray_prefix, ray_prefix, "ray offsets")) ||
        (rc = upload(&G.d_symtab, symtab, "symmetric view-row table")) ||
        (rc = upload(&G.d_tr_items, tr, "transpose work list")) ||
        (rc = dev_alloc((void**)&G.d_rays, sizeof(Ray) * (size_t)nray, "ray table"))) {
        free_dev(d_bins);
        free_dev(d_bin_off);
        free_dev(d_ray_prefix);
        return rc;
    }
    launch_build_rays(G.d_desc, nslices, G.d_views, d_bins, d_bin_off, d_ray_prefix, G.d_rays,
                      nray, 0);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    free_dev(d_bins);
    free_dev(d_bin_off);
    free_dev(d_ray_prefix);
    if (e != hipSuccess) return hip_fail(e, "building ray table");
    return 0;
}

// Concatenates per-group lists (each in batch order) and records the segments.
template <class Item>
static std::vector<Item> group_segments(std::vector<Item> (&by)[4],
                                        const std::vector<ctprj_geom>& geoms, bool spread,
                                        int* start, int* count)
{
    std::vector<Item> out;
    for (int g = 1; g < 4; ++g) {
        std::vector<Item> seg = batch_order(by[g], geoms, spread);
        start[g] = (int)out.size();
        count[g] = (int)seg.size();
        out.insert(out.end(), seg.begin(), seg.end());
    }
    return out;
}

int Plan::build_fwd_items(const ctprj_tune& t, Tables& G) const
{
    std::vector<FwdItem> items, by[4];
    const int B = t.fwd_block;
    for (int s = 0; s < nslices; ++s) {
        const ctprj_geom& g = geoms[s];
        if (G.slice_sym[s]) {
            const int pC = g.prjWidth / 2, bhi = pC + (g.prjWidth - 1) / 2;
            std::vector<int> canon;
            bool fold = false;
            choose_group(g, t.flags, &fold, &canon);
            const int bmax = fold ? pC : bhi;
            for (int i = 0; i < (int)canon.size(); ++i)
                for (int b0 = 0; b0 <= bmax; b0 += 64)
                    by[G.slice_sym[s]].push_back(FwdItem{s, i, b0, 0});
        } else {
            for (int a = 0; a < g.np; ++a)
                for (int b0 = 0; b0 < g.prjWidth; b0 += B) items.push_back(FwdItem{s, a, b0, 0});
        }
    }
    const bool spread = !(t.flags & 8);
    items = batch_order(items, geoms, spread);
    std::vector<FwdItem> sym = group_segments(by, geoms, spread, G.symf_start, G.symf_count);
    G.n_fwd_items = (int)items.size();
    G.n_symf_items = (int)sym.size();
    int rc = upload(&G.d_fwd_items, items, "forward work list");
    if (rc) return rc;
    return upload(&G.d_symf_items, sym, "symmetric forward work list");
}

int Plan::build_adj_items(const ctprj_tune& t, Tables& G) const
{
    std::vector<AdjItem> items, by[4];
    const int T = t.adj_tile;
    for (int s = 0; s < nslices; ++s) {
        const int n = geoms[s].n;
        const int grp = G.slice_sym[s];
        if (grp) {
            // 8x8 tiles of the group's fundamental domain of orbit base pixels; with even
            // n the first tile's work-group also zeroes row 0 and column 0 (pad = 1)
            const int H = (n - 1) / 2, nt = (H + 1 + 7) / 8;
            const size_t first = by[grp].size();
            if (grp == kGrpD4) {
                for (int ty = 0; ty < nt; ++ty)
                    for (int tx = ty; tx < nt; ++tx) by[grp].push_back(AdjItem{s, ty * 8, tx * 8, 0});
            } else if (grp == kGrpV4) {
                for (int ty = 0; ty < nt; ++ty)
                    for (int tx = 0; tx < nt; ++tx) by[grp].push_back(AdjItem{s, ty * 8, tx * 8, 0});
            } else {
                for (int y = -H; y <= H; y += 8)
                    for (int tx = 0; tx < nt; ++tx) by[grp].push_back(AdjItem{s, y, tx * 8, 0});
            }
            if (n % 2 == 0 && by[grp].size() > first) by[grp][first].pad = 1;
        } else {
            for (int r = 0; r < n; r += T)
                for (int c = 0; c < n; c += T) items.push_back(AdjItem{s, r, c, 0});
        }
    }
    const bool spread = !(t.flags & 8);
    items = batch_order(items, geoms, spread);
    std::vector<AdjItem> sym = group_segments(by, geoms, spread, G.syma_start, G.syma_count);
    G.n_adj_items = (int)items.size();
    G.n_syma_items = (int)sym.size();
    int rc = upload(&G.d_adj_items, items, "adjoint work list");
    if (rc) return rc;
    return upload(&G.d_syma_items, sym, "symmetric adjoint work list");
}

static int plan_new(const ctprj_geom* geoms, int nslices, int flags, Plan** out);

}  // namespace ctprj

using namespace ctprj;

struct ctprj_plan : public Plan {};

namespace ctprj {

static int plan_new(const ctprj_geom* geoms, int nslices, int flags, Plan** out)
{
    *out = nullptr;
    if (!geoms || nslices < 1) return fail(CTPRJ_EINVAL, "need at least one slice");
    for (int i = 0; i < nslices; ++i) {
        int rc = validate_geom(geoms[i], i);
        if (rc) return rc;
    }
    int dev = 0;
    int rc = check_device(&dev);
    if (rc) return rc;
    ctprj_plan* p = new (std::nothrow) ctprj_plan();
    if (!p) return fail(CTPRJ_ENOMEM, "host allocation");
    p->device = dev;
    p->nslices = nslices;
    p->geoms.assign(geoms, geoms + nslices);
    p->tune.fwd_block = 256;
    p->tune.adj_tile = 16;
    p->tune.flags = flags;
    p->tune.adj_seg = 8;
    p->tune.fwd_seg = 8;
    p->tune.adj_split = 1;
    p->tune.fwd_split = 1;
    std::vector<long long> kern_off(nslices), prew_off(nslices);
    std::vector<float> kern, prew;
    long long img = 0, sino = 0;
    int max_W = 0;
    for (int s = 0; s < nslices; ++s) {
        const ctprj_geom& g = geoms[s];
        p->img_offsets.push_back(img);
        p->sino_offsets.push_back(sino);
        img += (long long)g.n * g.n;
        sino += (long long)g.np * g.prjWidth;
        // FBP: filter taps (2W) and fan pre-weights (W)
        const int W = g.prjWidth, pC = W / 2;
        const long long L = 2LL * W;
        kern_off[s] = (long long)kern.size();
        auto ramlak = [&](long long k) -> double {
            k = ((k % L) + L) % L;
            if (k > L / 2) k = L - k;
            if (k == 0) return 0.25;
            if (k % 2 == 0) return 0.0;
            return -1.0 / (M_PI * M_PI * (double)k * (double)k);
        };
        for (long long k = 0; k < L; ++k)
            kern.push_back((float)(0.5 * ramlak(k) + 0.25 * ramlak(k - 1) + 0.25 * ramlak(k + 1)));
        prew_off[s] = (long long)prew.size();
        for (int b = 0; b < W; ++b) {
            double w = 1.0;
            if (g.d > 0.0f) {
                const double t = (double)(b - pC) * g.dSize;
                w = (double)g.d / std::sqrt((double)g.d * g.d + t * t);
            }
            prew.push_back((float)w);
        }
        max_W = std::max(max_W, W);
    }
    p->img_elems = img;
    p->sino_elems = sino;
    p->max_W = max_W;
    std::vector<int2> fbpi;
    for (int s = 0; s < nslices; ++s)
        for (int a = 0; a < geoms[s].np; ++a) fbpi.push_back(make_int2(s, a));
    p->n_fbp_items = (int)fbpi.size();
    std::vector<float> zeros((size_t)max_W + 1, 0.0f);
    p->red_cap = bhc_red_doubles(sino);
    if ((rc = upload(&p->d_fbp_items, fbpi, "fbp work list")) ||
        (rc = upload(&p->d_fbp_kern, kern, "fbp taps")) ||
        (rc = upload(&p->d_fbp_kern_off, kern_off, "fbp tap offsets")) ||
        (rc = upload(&p->d_fbp_prew, prew, "fbp preweights")) ||
        (rc = upload(&p->d_fbp_prew_off, prew_off, "fbp preweight offsets")) ||
        (rc = upload(&p->d_zero_row, zeros, "zero row")) ||
        (rc = dev_alloc((void**)&p->d_imgT, sizeof(float) * (size_t)img, "transposed image")) ||
        (rc = dev_alloc((void**)&p->d_sino_ws, sizeof(float) * (size_t)sino, "sinogram workspace")) ||
        // the beam-hardening passes' reduction partials and outputs: their calls
        // allocate nothing (and can be captured into a graph)
        (rc = dev_alloc((void**)&p->d_red, sizeof(double) * (size_t)p->red_cap,
                        "reduction workspace")) ||
        (rc = dev_alloc((void**)&p->d_red_out, sizeof(double) * (size_t)bhc_ie_nvals(CTPRJ_MAX_E),
                        "reduction output")) ||
        (rc = p->build_geometry(p->tune.flags, *p)) || (rc = p->build_fwd_items(p->tune, *p)) ||
        (rc = p->build_adj_items(p->tune, *p))) {
        delete p;
        return rc;
    }
    *out = p;
    return 0;
}

}  // namespace ctprj

extern "C" int ctprj_plan_create(const ctprj_geom* geoms, int nslices, ctprj_plan** out)
{
    if (!out) return fail(CTPRJ_EINVAL, "plan output pointer is NULL");
    Plan* p = nullptr;
    int rc = plan_new(geoms, nslices, 1, &p);
    *out = static_cast<ctprj_plan*>(p);
    return rc;
}

extern "C" int ctprj_plan_destroy(ctprj_plan* plan)
{
    delete plan;
    return 0;
}

extern "C" int ctprj_plan_layout(const ctprj_plan* p, long long* img_elems,
                                 long long* sino_elems, long long* img_offsets,
                                 long long* sino_offsets)
{
    if (!p) return fail(CTPRJ_EINVAL, "plan is NULL");
    if (img_elems) *img_elems = p->img_elems;
    if (sino_elems) *sino_elems = p->sino_elems;
    for (int s = 0; s < p->nslices; ++s) {
        if (img_offsets) img_offsets[s] = p->img_offsets[s];
        if (sino_offsets) sino_offsets[s] = p->sino_offsets[s];
    }
    return 0;
}

// ------------------------------------------------------------------ launches ---

namespace ctprj {

int Plan::launch_forward(const float* d_img, float* d_sino, hipStream_t st, bool count)
{
    if (count && n_sym) return fail(CTPRJ_EINVAL, "internal: count needs the general kernels");
    launch_transpose(n_tr_items, d_desc, d_tr_items, d_img, d_imgT, st);
    if (n_fwd_items)
        launch_forward_kernel(count, tune.fwd_block, n_fwd_items, d_desc, d_fwd_items, d_views,
                              d_rays, d_img, d_imgT, d_sino, st);
    for (int g = 1; g < 4; ++g)
        if (symf_count[g])
            launch_forward_sym(g, tune.fwd_seg, tune.fwd_split, d_split, symf_count[g], d_desc,
                               d_symf_items + symf_start[g], d_views, d_rays, d_img, d_imgT,
                               d_sino, nullptr, st);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(e, "forward launch");
}

long long Plan::bhc_fused_parts() const
{
    if (n_fwd_items) return -1;
    long long n = 0;
    for (int g = 1; g < 4; ++g) n += symf_count[g];
    return n;
}

int Plan::launch_forward_bhc(const float* d_img, float* d_w, BhcStore bhc, hipStream_t st)
{
    if (n_fwd_items) return fail(CTPRJ_EINVAL, "internal: fused store needs symmetric slices");
    launch_transpose(n_tr_items, d_desc, d_tr_items, d_img, d_imgT, st);
    bhc.part_off = 0;
    for (int g = 1; g < 4; ++g)
        if (symf_count[g]) {
            launch_forward_sym(g, tune.fwd_seg, tune.fwd_split, d_split, symf_count[g], d_desc,
                               d_symf_items + symf_start[g], d_views, d_rays, d_img, d_imgT,
                               d_w, &bhc, st);
            bhc.part_off += symf_count[g];
        }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(e, "forward launch");
}

int Plan::launch_adjoint(const float* d_sino, float* d_img, hipStream_t st, bool fbp,
                         const RedJob& job)
{
    if (job.out && !n_syma_items)
        return fail(CTPRJ_EINVAL, "internal: a side sum needs a symmetric adjoint launch");
    if (n_adj_items)
        launch_adjoint_kernel(tune.adj_tile, fbp, n_adj_items, d_desc, d_adj_items, d_views,
                              d_rays, d_sino, d_img, st);
    if (n_syma_items) {
        bool first = true;
        for (int g = 1; g < 4; ++g)
            if (syma_count[g]) {
                launch_adjoint_sym(fbp, g, tune.adj_seg, (tune.flags & 16) ? 3 : 4,
                                   (tune.flags & 32) != 0, tune.adj_split,
                                   d_split, syma_count[g], d_desc, d_syma_items + syma_start[g],
                                   d_views, d_rays, d_sino, d_zero_row, d_symtab, d_img,
                                   first ? job : RedJob{}, st);
                first = false;
            }
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(e, "adjoint launch");
}

int Plan::launch_fbp(const float* d_sino, float* d_img, hipStream_t st)
{
    const size_t shmem = sizeof(float) * 3 * (size_t)max_W;
    if (shmem > 160 * 1024) return fail(CTPRJ_EINVAL, "FBP: prjWidth %d too large", max_W);
    launch_fbp_filter(n_fbp_items, shmem, d_desc, d_fbp_items, d_fbp_kern, d_fbp_kern_off,
                      d_fbp_prew, d_fbp_prew_off, d_sino, d_sino_ws, st);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "fbp filter launch");
    return launch_adjoint(d_sino_ws, d_img, st, true);
}

}  // namespace ctprj

extern "C" int ctprj_forward(ctprj_plan* p, const float* d_img, float* d_sino, void* stream)
{
    if (!p || !d_img || !d_sino) return fail(CTPRJ_EINVAL, "ctprj_forward: NULL argument");
    return p->launch_forward(d_img, d_sino, (hipStream_t)stream, false);
}

extern "C" int ctprj_adjoint(ctprj_plan* p, const float* d_sino, float* d_img, void* stream)
{
    if (!p || !d_img || !d_sino) return fail(CTPRJ_EINVAL, "ctprj_adjoint: NULL argument");
    return p->launch_adjoint(d_sino, d_img, (hipStream_t)stream, false);
}

extern "C" int ctprj_fbp(ctprj_plan* p, const float* d_sino, float* d_img, void* stream)
{
    if (!p || !d_img || !d_sino) return fail(CTPRJ_EINVAL, "ctprj_fbp: NULL argument");
    return p->launch_fbp(d_sino, d_img, (hipStream_t)stream);
}

// nnz(A): the general kernels count the non-zero FP32 weights of every ray (a plan
// without the symmetric kernels is made for the count when needed).
extern "C" int ctprj_plan_nnz(ctprj_plan* p, long long* nnz, void* stream)
{
    if (!p || !nnz) return fail(CTPRJ_EINVAL, "ctprj_plan_nnz: NULL argument");
    hipStream_t st = (hipStream_t)stream;
    Plan* q = p;
    Plan* tmp = nullptr;
    if (p->n_sym) {
        int rc = plan_new(p->geoms.data(), p->nslices, 0, &tmp);
        if (rc) return rc;
        q = tmp;
    }
    float* d_img = nullptr;
    int rc = dev_alloc((void**)&d_img, sizeof(float) * (size_t)q->img_elems, "nnz image");
    if (rc) { delete static_cast<ctprj_plan*>(tmp); return rc; }
    hipError_t e = hipMemsetAsync(d_img, 0, sizeof(float) * (size_t)q->img_elems, st);
    if (e != hipSuccess) rc = hip_fail(e, "hipMemsetAsync");
    if (!rc) rc = q->launch_forward(d_img, q->d_sino_ws, st, true);
    std::vector<float> counts((size_t)q->sino_elems);
    if (!rc) {
        e = hipMemcpyAsync(counts.data(), q->d_sino_ws, sizeof(float) * counts.size(),
                           hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = hip_fail(e, "nnz readback");
    }
    free_dev(d_img);
    delete static_cast<ctprj_plan*>(tmp);
    if (rc) return rc;
    long long total = 0;
    for (float c : counts) total += (long long)c;
    *nnz = total;
    return 0;
}

extern "C" int ctprj_plan_modes(const ctprj_plan* p, int* modes)
{
    if (!p || !modes) return fail(CTPRJ_EINVAL, "NULL argument");
    for (int s = 0; s < p->nslices; ++s) modes[s] = p->slice_sym[s];
    return 0;
}

extern "C" int ctprj_plan_get_tune(const ctprj_plan* p, ctprj_tune* t)
{
    if (!p || !t) return fail(CTPRJ_EINVAL, "NULL argument");
    *t = p->tune;
    return 0;
}

extern "C" int ctprj_plan_set_tune(ctprj_plan* p, const ctprj_tune* t)
{
    if (!p || !t) return fail(CTPRJ_EINVAL, "NULL argument");
    if (t->fwd_block != 64 && t->fwd_block != 128 && t->fwd_block != 256)
        return fail(CTPRJ_EINVAL, "fwd_block must be 64, 128 or 256");
    if (t->fwd_seg != 1 && t->fwd_seg != 2 && t->fwd_seg != 4 && t->fwd_seg != 8 &&
        t->fwd_seg != 16)
        return fail(CTPRJ_EINVAL, "fwd_seg must be 1, 2, 4, 8 or 16");
    if (t->adj_tile != 8 && t->adj_tile != 16) return fail(CTPRJ_EINVAL, "adj_tile must be 8 or 16");
    if (t->flags & ~63) return fail(CTPRJ_EINVAL, "unknown tune flags 0x%x", t->flags);
    if (t->adj_seg != 1 && t->adj_seg != 2 && t->adj_seg != 4 && t->adj_seg != 8)
        return fail(CTPRJ_EINVAL, "adj_seg must be 1, 2, 4 or 8");
    if (t->adj_split < 1 || t->adj_split > 8 || t->fwd_split < 1 || t->fwd_split > 8)
        return fail(CTPRJ_EINVAL, "adj_split and fwd_split must be 1 to 8");
    // the parts this change rebuilds go into N; the plan keeps its tables until all of
    // them (and the split partial-sum space) exist
    ctprj_tune nt = *t;
    const bool geo = nt.flags != p->tune.flags;
    const bool fwd = geo || nt.fwd_block != p->tune.fwd_block;
    const bool adj = geo || nt.adj_tile != p->tune.adj_tile;
    Tables N;
    int rc = 0;
    if (geo) rc = p->build_geometry(nt.flags, N);
    if (geo && !rc) {
        // the work lists read the new slice groups
        if (!(rc = p->build_fwd_items(nt, N))) rc = p->build_adj_items(nt, N);
    } else if (!rc) {
        // same geometry: build the lists from the plan's slice groups
        N.slice_sym = p->slice_sym;
        if (fwd) rc = p->build_fwd_items(nt, N);
        if (adj && !rc) rc = p->build_adj_items(nt, N);
    }
    if (rc) return rc;  // N frees what it holds; the plan is unchanged
    const Tables& F = fwd ? N : *static_cast<const Tables*>(p);
    const Tables& A = adj ? N : *static_cast<const Tables*>(p);
    int split_rc = 0;
    size_t need = p->split_need(nt, F, A);
    if (need > kSplitMaxFloats) {
        split_rc = fail(CTPRJ_EINVAL, "split work items need %zu MiB of partial sums (at most %zu)",
                        need * sizeof(float) >> 20, kSplitMaxFloats * sizeof(float) >> 20);
        nt.adj_split = nt.fwd_split = 1;
        need = 0;
    }
    float* split = nullptr;
    if (need > p->split_cap && (rc = dev_alloc((void**)&split, sizeof(float) * need,
                                                "split partial sums")))
        return rc;
    // commit: swap the rebuilt parts in (N then frees the old ones)
    if (geo) p->swap_geometry(N);
    if (fwd) p->swap_fwd(N);
    if (adj) p->swap_adj(N);
    if (split) {
        free_ptr(p->d_split);
        p->d_split = split;
        p->split_cap = need;
    }
    p->tune = nt;
    return split_rc;
}

namespace ctprj {

// Partial-sum space of the split symmetric launches: per group launch nitems * nsplit
// work-groups of 512 floats; the launches of one call run in stream order and share it.
size_t Plan::split_need(const ctprj_tune& t, const Tables& F, const Tables& A) const
{
    size_t need = 0;
    for (int g = 1; g < 4; ++g) {
        if (t.fwd_split > 1) need = std::max(need, (size_t)F.symf_count[g] * t.fwd_split);
        if (t.adj_split > 1) need = std::max(need, (size_t)A.syma_count[g] * t.adj_split);
    }
    return need * 512;
}

}  // namespace ctprj

extern "C" int ctprj_debug_fail_alloc(int k)
{
    if (k < 0) return fail(CTPRJ_EINVAL, "ctprj_debug_fail_alloc: k must be >= 0");
    g_fail_alloc.store(k);
    return 0;
}

// Cache key of a plan: the device arch, then its slices' geometries in order, a run of
// k > 1 equal consecutive slices written once as "k*" + geometry.
static std::string tune_key(const ctprj_plan* p, const char* arch)
{
    std::ostringstream os;
    os << arch;
    const auto& gs = p->geoms;
    auto same = [](const ctprj_geom& a, const ctprj_geom& b) {
        return a.n == b.n && a.prjWidth == b.prjWidth && a.np == b.np && a.prjFull == b.prjFull &&
               a.dSize == b.dSize && a.effectiveRate == b.effectiveRate && a.d == b.d;
    };
    for (size_t i = 0; i < gs.size();) {
        size_t k = 1;
        while (i + k < gs.size() && same(gs[i], gs[i + k])) ++k;
        const ctprj_geom& g = gs[i];
        os << "|";
        if (k > 1) os << k << "*";
        os << g.n << "," << g.prjWidth << "," << g.np << "," << g.prjFull << "," << g.dSize << ","
           << g.effectiveRate << "," << g.d;
        i += k;
    }
    return os.str();
}

// Times the kernel shapes on this plan's geometry: general forward threads per group,
// symmetric forward waves and work-groups per ray run, general adjoint tile edge,
// symmetric adjoint waves and work-groups per tile for each bin window width (flag bit 4)
// and gather (LDS staging, bit 5), and (several slices) the work order of tune flag bit
// 3; keeps the symmetry flags (bits 0-2).
extern "C" int ctprj_plan_autotune(ctprj_plan* p, int reps, const char* cache_path, void* stream)
{
    if (!p) return fail(CTPRJ_EINVAL, "plan is NULL");
    if (reps < 1) reps = 3;
    hipStream_t st = (hipStream_t)stream;
    char arch[64] = "gfx950";
    ctprj_device_arch(arch, sizeof(arch));
    const std::string key = tune_key(p, arch) + "|f" + std::to_string(p->tune.flags & 7);
    if (cache_path) {
        std::ifstream f(cache_path);
        std::string line;
        while (std::getline(f, line)) {
            const size_t sp = line.rfind(' ');
            if (sp == std::string::npos || line.substr(0, sp) != key) continue;
            ctprj_tune t{};
            t.adj_split = t.fwd_split = 1;  // entries without the split fields
            if (std::sscanf(line.c_str() + sp + 1, "%d,%d,%d,%d,%d,%d,%d", &t.fwd_block,
                            &t.adj_tile, &t.flags, &t.adj_seg, &t.fwd_seg, &t.adj_split,
                            &t.fwd_split) >= 5)
                return ctprj_plan_set_tune(p, &t);
        }
    }
    float *d_img = nullptr, *d_sino = nullptr;
    int rc = dev_alloc((void**)&d_img, sizeof(float) * (size_t)p->img_elems, "autotune image");
    if (rc) return rc;
    if ((rc = dev_alloc((void**)&d_sino, sizeof(float) * (size_t)p->sino_elems, "autotune sino"))) {
        free_dev(d_img);
        return rc;
    }
    (void)hipMemsetAsync(d_img, 0, sizeof(float) * (size_t)p->img_elems, st);
    (void)hipMemsetAsync(d_sino, 0, sizeof(float) * (size_t)p->sino_elems, st);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    ctprj_tune best = p->tune;
    // average time of `reps` launches of one direction after one untimed launch
    auto time_dir = [&](bool fwd, float* out) -> int {
        float ms = 0;
        for (int r = -1; r < reps; ++r) {
            (void)hipEventRecord(e0, st);
            const int rc2 = fwd ? p->launch_forward(d_img, d_sino, st, false)
                                : p->launch_adjoint(d_sino, d_img, st, false);
            if (rc2) return rc2;
            (void)hipEventRecord(e1, st);
            (void)hipEventSynchronize(e1);
            float m = 0;
            (void)hipEventElapsedTime(&m, e0, e1);
            if (r >= 0) ms += m;
        }
        *out = ms;
        return 0;
    };
    // one knob at a time, the others held at the best found so far
    auto search = [&](int ctprj_tune::*knob, const int* vals, int nv, bool fwd) -> int {
        float best_ms = 1e30f;
        for (int i = 0; i < nv; ++i) {
            ctprj_tune t = best;
            t.*knob = vals[i];
            int rc2 = ctprj_plan_set_tune(p, &t);
            float ms = 0;
            if (!rc2) rc2 = time_dir(fwd, &ms);
            if (rc2) return rc2;
            if (ms < best_ms) { best_ms = ms; best.*knob = vals[i]; }
        }
        return ctprj_plan_set_tune(p, &best);
    };
    // waves per item and work-groups per item together (both cut an item's work); more
    // than one work-group per item only while the items are too few to fill the device
    // several times over
    auto search_split = [&](int ctprj_tune::*seg, int ctprj_tune::*split, const int* segs,
                            int ns, bool fwd, int nitems, float* out_ms = nullptr) -> int {
        float best_ms = 1e30f;
        const int splits[7] = {1, 2, 3, 4, 5, 6, 8};
        const int nsplits = nitems < 8192 ? 7 : 1;
        for (int i = 0; i < ns; ++i)
            for (int j = 0; j < nsplits; ++j) {
                const int k = splits[j];
                ctprj_tune t = best;
                t.*seg = segs[i];
                t.*split = k;
                if (ctprj_plan_set_tune(p, &t)) {
                    if (k > 1) continue;  // partial sums beyond the plan's cap
                    return CTPRJ_EINVAL;
                }
                float ms = 0;
                int rc2 = time_dir(fwd, &ms);
                if (rc2) return rc2;
                if (ms < best_ms) {
                    best_ms = ms;
                    best.*seg = segs[i];
                    best.*split = k;
                }
            }
        if (out_ms) *out_ms = best_ms;
        return ctprj_plan_set_tune(p, &best);
    };
    const int blocks[3] = {64, 128, 256}, tiles[2] = {8, 16};
    const int fsegs[5] = {1, 2, 4, 8, 16}, asegs[4] = {1, 2, 4, 8};
    if (!rc && p->nslices > 1) {
        // work order: every slice spread over the XCDs, or whole slices per XCD
        // (flag bit 3), timed on a forward + adjoint pair
        const int base = best.flags & ~8, pol[2] = {base, base | 8};
        float best_ms = 1e30f;
        for (int i = 0; i < 2 && !rc; ++i) {
            ctprj_tune t = best;
            t.flags = pol[i];
            float mf = 0, ma = 0;
            if (!(rc = ctprj_plan_set_tune(p, &t)) && !(rc = time_dir(true, &mf)) &&
                !(rc = time_dir(false, &ma)) && mf + ma < best_ms) {
                best_ms = mf + ma;
                best.flags = pol[i];
            }
        }
        if (!rc) rc = ctprj_plan_set_tune(p, &best);
    }
    if (!rc && p->n_fwd_items) rc = search(&ctprj_tune::fwd_block, blocks, 3, true);
    if (!rc && p->n_symf_items)
        rc = search_split(&ctprj_tune::fwd_seg, &ctprj_tune::fwd_split, fsegs, 5, true,
                          p->n_symf_items);
    if (!rc && p->n_adj_items) rc = search(&ctprj_tune::adj_tile, tiles, 2, false);
    if (!rc && p->n_syma_items) {
        // symmetric adjoint: bin windows 4 or 3 wide (flag bit 4) and gathers per lane
        // from L1 or staged per wave in LDS (bit 5), each variant with its best waves and
        // work-groups per tile (the variants differ in registers, so in the shape they
        // want)
        const ctprj_tune start = best;
        ctprj_tune winner = best;
        float win_ms = 1e30f;
        for (int v = 0; v < 4 && !rc; ++v) {
            best = start;
            best.flags = (start.flags & ~48) | ((v & 1) ? 16 : 0) | ((v & 2) ? 32 : 0);
            float ms = 1e30f;
            rc = search_split(&ctprj_tune::adj_seg, &ctprj_tune::adj_split, asegs, 4, false,
                              p->n_syma_items, &ms);
            if (!rc && ms < win_ms) {
                win_ms = ms;
                winner = best;
            }
        }
        best = winner;
        if (!rc) rc = ctprj_plan_set_tune(p, &best);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    free_dev(d_img);
    free_dev(d_sino);
    if (rc) return rc;
    rc = ctprj_plan_set_tune(p, &best);
    if (!rc && cache_path) {
        std::ofstream f(cache_path, std::ios::app);
        f << key << " " << best.fwd_block << "," << best.adj_tile << "," << best.flags << ","
          << best.adj_seg << "," << best.fwd_seg << "," << best.adj_split << ","
          << best.fwd_split << "\n";
    }
    return rc;
}

// ------------------------------------------------- reference-compatible API ---


static std::mutex g_ref_mu;
static ctprj_plan* g_ref_plan = nullptr;
static ctprj_geom g_ref_geom;
static float* g_ref_img = nullptr;
static float* g_ref_sino = nullptr;

extern "C" int prj_setup(int n, int prjWidth, int np, int prjFull, float dSize,
                         float effectiveRate, float d)
{
    std::lock_guard<std::mutex> lk(g_ref_mu);
    ctprj_geom g{n, prjWidth, np, prjFull, dSize, effectiveRate, d};
    ctprj_plan* p = nullptr;
    int rc = ctprj_plan_create(&g, 1, &p);
    if (rc) return rc;
    float *img = nullptr, *sino = nullptr;
    if ((rc = dev_alloc((void**)&img, sizeof(float) * (size_t)n * n, "prj image")) ||
        (rc = dev_alloc((void**)&sino, sizeof(float) * (size_t)np * prjWidth, "prj sinogram"))) {
        free_dev(img);
        ctprj_plan_destroy(p);
        return rc;
    }
    if (g_ref_plan) ctprj_plan_destroy(g_ref_plan);
    free_dev(g_ref_img);
    free_dev(g_ref_sino);
    g_ref_plan = p;
    g_ref_geom = g;
    g_ref_img = img;
    g_ref_sino = sino;
    return 0;
}

extern "C" int prj_get_geom(ctprj_geom* out)
{
    std::lock_guard<std::mutex> lk(g_ref_mu);
    if (!g_ref_plan) return fail(CTPRJ_ESTATE, "prj_setup has not been called");
    if (!out) return fail(CTPRJ_EINVAL, "NULL argument");
    *out = g_ref_geom;
    return 0;
}

extern "C" int prj_show(void)
{
    std::lock_guard<std::mutex> lk(g_ref_mu);
    if (!g_ref_plan) return fail(CTPRJ_ESTATE, "prj_setup has not been called");
    const ctprj_geom& g = g_ref_geom;
    std::printf("config.n=%d\nconfig.prjWidth=%d\nconfig.np=%d\nconfig.prjFull=%d\n"
                "config.dSize=%g\nconfig.effectiveRate=%g\nconfig.d=%g\n",
                g.n, g.prjWidth, g.np, g.prjFull, (double)g.dSize, (double)g.effectiveRate,
                (double)g.d);
    std::fflush(stdout);
    return 0;
}

extern "C" int prj_run(float* img, float* sino, int cmd)
{
    std::lock_guard<std::mutex> lk(g_ref_mu);
    if (!g_ref_plan) return fail(CTPRJ_ESTATE, "prj_setup has not been called");
    if (!img || !sino) return fail(CTPRJ_EINVAL, "prj_run: NULL buffer");
    const size_t ib = sizeof(float) * (size_t)g_ref_geom.n * g_ref_geom.n;
    const size_t sb = sizeof(float) * (size_t)g_ref_geom.np * g_ref_geom.prjWidth;
    hipError_t e;
    int rc = 0;
    if (cmd & CTPRJ_FWD_BIT) {
        if ((e = hipMemcpy(g_ref_img, img, ib, hipMemcpyHostToDevice)) != hipSuccess)
            return hip_fail(e, "prj_run H2D");
        rc = ctprj_forward(g_ref_plan, g_ref_img, g_ref_sino, nullptr);
        if (!rc && (e = hipMemcpy(sino, g_ref_sino, sb, hipMemcpyDeviceToHost)) != hipSuccess)
            rc = hip_fail(e, "prj_run D2H");
    } else if (cmd & (CTPRJ_BWD_BIT | CTPRJ_FBP_BIT)) {
        if ((e = hipMemcpy(g_ref_sino, sino, sb, hipMemcpyHostToDevice)) != hipSuccess)
            return hip_fail(e, "prj_run H2D");
        rc = (cmd & CTPRJ_BWD_BIT) ? ctprj_adjoint(g_ref_plan, g_ref_sino, g_ref_img, nullptr)
                                   : ctprj_fbp(g_ref_plan, g_ref_sino, g_ref_img, nullptr);
        if (!rc && (e = hipMemcpy(img, g_ref_img, ib, hipMemcpyDeviceToHost)) != hipSuccess)
            rc = hip_fail(e, "prj_run D2H");
    } else {
        return fail(CTPRJ_EINVAL, "prj_run: cmd %d has no FWD/BWD/FBP bit", cmd);
    }
    return rc;
}

// -------------------------------------------------------------------- misc ---

extern "C" const char* ctprj_last_error(void) { return g_err.c_str(); }

extern "C" const char* ctprj_version(void) { return "ctprj 0.1 gfx950"; }

extern "C" int ctprj_device_arch(char* buf, int buflen)
{
    if (!buf || buflen < 1) return fail(CTPRJ_EINVAL, "NULL buffer");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    hipDeviceProp_t p;
    if ((e = hipGetDeviceProperties(&p, dev)) != hipSuccess) return hip_fail(e, "properties");
    std::snprintf(buf, (size_t)buflen, "%s", p.gcnArchName);
    return 0;
}
