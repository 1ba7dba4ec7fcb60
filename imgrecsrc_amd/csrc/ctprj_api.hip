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
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

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
    const double sino = (double)g.np * g.prjWidth, img = (double)g.n * g.n;
    if (sino > 2e9 || img > 2e9) return fail(CTPRJ_EINVAL, "slice %d: too large", idx);
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

Plan::~Plan()
{
    free_dev(d_desc);
    free_dev(d_views);
    free_dev(d_rays);
    free_dev(d_fwd_items);
    free_dev(d_adj_items);
    free_dev(d_tr_items);
    free_dev(d_fbp_items);
    free_dev(d_fbp_kern);
    free_dev(d_fbp_kern_off);
    free_dev(d_fbp_prew);
    free_dev(d_fbp_prew_off);
    free_dev(d_imgT);
    free_dev(d_sino_ws);
    free_dev(d_red);
    free_dev(d_red_out);
}

template <class T>
static int upload(T** dst, const std::vector<T>& src, const char* what)
{
    if (src.empty()) {
        *dst = nullptr;
        return 0;
    }
    hipError_t e = hipMalloc((void**)dst, sizeof(T) * src.size());
    if (e != hipSuccess) return fail(CTPRJ_ENOMEM, "hipMalloc(%s, %zu B): %s", what,
                                     sizeof(T) * src.size(), hipGetErrorString(e));
    e = hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, what);
    return 0;
}

static int dev_alloc(void** p, size_t bytes, const char* what)
{
    hipError_t e = hipMalloc(p, bytes ? bytes : 16);
    if (e != hipSuccess) return fail(CTPRJ_ENOMEM, "hipMalloc(%s, %zu B): %s", what, bytes,
                                     hipGetErrorString(e));
    return 0;
}

// XCD-aware order: the hardware hands work-group i to XCD i % 8.  Lay the item list
// out so that XCD k runs a contiguous eighth of it (neighbouring views of one slice
// share that XCD's L2).  Pads with empty items (slice = -1) to a multiple of 8.
template <class Item>
static std::vector<Item> xcd_order(const std::vector<Item>& in)
{
    const size_t G = (in.size() + 7) / 8 * 8, per = G / 8;
    std::vector<Item> out(G);
    for (size_t i = 0; i < G; ++i) {
        const size_t src = (i % 8) * per + i / 8;
        if (src < in.size()) out[i] = in[src];
        else { out[i] = Item{}; out[i].slice = -1; }
    }
    return out;
}

int Plan::build_fwd_items()
{
    std::vector<FwdItem> items;
    const int B = tune.fwd_block;
    for (int s = 0; s < nslices; ++s) {
        const ctprj_geom& g = geoms[s];
        for (int a = 0; a < g.np; ++a)
            for (int b0 = 0; b0 < g.prjWidth; b0 += B) items.push_back(FwdItem{s, a, b0, 0});
    }
    items = xcd_order(items);
    free_dev(d_fwd_items);
    d_fwd_items = nullptr;
    n_fwd_items = (int)items.size();
    return upload(&d_fwd_items, items, "forward work list");
}

int Plan::build_adj_items()
{
    std::vector<AdjItem> items;
    const int T = tune.adj_tile;
    for (int s = 0; s < nslices; ++s) {
        const int n = geoms[s].n;
        for (int r = 0; r < n; r += T)
            for (int c = 0; c < n; c += T) items.push_back(AdjItem{s, r, c, 0});
    }
    items = xcd_order(items);
    free_dev(d_adj_items);
    d_adj_items = nullptr;
    n_adj_items = (int)items.size();
    return upload(&d_adj_items, items, "adjoint work list");
}

}  // namespace ctprj

using namespace ctprj;

struct ctprj_plan : public Plan {};

extern "C" int ctprj_plan_create(const ctprj_geom* geoms, int nslices, ctprj_plan** out)
{
    if (!out) return fail(CTPRJ_EINVAL, "plan output pointer is NULL");
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

    std::vector<SliceDesc> desc(nslices);
    std::vector<float4> views, bins;
    std::vector<long long> bin_off(nslices), ray_prefix(nslices + 1, 0);
    std::vector<long long> kern_off(nslices), prew_off(nslices);
    std::vector<float> kern, prew;
    long long img = 0, sino = 0, nview = 0, nray = 0;
    int max_n = 0, max_W = 0;
    for (int s = 0; s < nslices; ++s) {
        const ctprj_geom& g = geoms[s];
        SliceDesc& D = desc[s];
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
        D.img_off = img;
        D.sino_off = sino;
        D.view_off = nview;
        D.ray_off = nray;
        p->img_offsets.push_back(img);
        p->sino_offsets.push_back(sino);
        img += (long long)g.n * g.n;
        sino += (long long)g.np * g.prjWidth;
        for (int a = 0; a < g.np; ++a) {
            float c, si;
            view_trig(a, g.prjFull, &c, &si);
            views.push_back(make_float4(c, si, 0.0f, 0.0f));
        }
        nview += g.np;
        bin_off[s] = (long long)bins.size();
        if (D.fan) {
            for (int b = 0; b < g.prjWidth; ++b) {
                float q[4];
                fan_bin(b, g, q);
                bins.push_back(make_float4(q[0], q[1], q[2], q[3]));
            }
            nray += (long long)g.np * g.prjWidth;
        } else {
            nray += g.np;
        }
        ray_prefix[s + 1] = nray;
        // FBP: filter taps (2W) and fan pre-weights (W)
        const int W = g.prjWidth;
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
            if (D.fan) {
                const double t = (double)(b - D.pC) * g.dSize;
                w = (double)g.d / std::sqrt((double)g.d * g.d + t * t);
            }
            prew.push_back((float)w);
        }
        max_n = std::max(max_n, g.n);
        max_W = std::max(max_W, W);
    }
    p->img_elems = img;
    p->sino_elems = sino;
    p->max_W = max_W;
    p->tune.fwd_block = 64;
    p->tune.adj_tile = 16;
    p->tune.flags = 0;

    std::vector<AdjItem> tr;
    for (int s = 0; s < nslices; ++s)
        for (int r = 0; r < geoms[s].n; r += 32)
            for (int c = 0; c < geoms[s].n; c += 32) tr.push_back(AdjItem{s, r, c, 0});
    std::vector<int2> fbpi;
    for (int s = 0; s < nslices; ++s)
        for (int a = 0; a < geoms[s].np; ++a) fbpi.push_back(make_int2(s, a));
    p->n_tr_items = (int)tr.size();
    p->n_fbp_items = (int)fbpi.size();

    float4* d_bins = nullptr;
    long long *d_bin_off = nullptr, *d_ray_prefix = nullptr;
    if ((rc = upload(&p->d_desc, desc, "slice table")) ||
        (rc = upload(&p->d_views, views, "view table")) ||
        (rc = upload(&d_bins, bins, "bin table")) ||
        (rc = upload(&d_bin_off, bin_off, "bin offsets")) ||
        (rc = upload(&d_ray_prefix, ray_prefix, "ray offsets")) ||
        (rc = upload(&p->d_tr_items, tr, "transpose work list")) ||
        (rc = upload(&p->d_fbp_items, fbpi, "fbp work list")) ||
        (rc = upload(&p->d_fbp_kern, kern, "fbp taps")) ||
        (rc = upload(&p->d_fbp_kern_off, kern_off, "fbp tap offsets")) ||
        (rc = upload(&p->d_fbp_prew, prew, "fbp preweights")) ||
        (rc = upload(&p->d_fbp_prew_off, prew_off, "fbp preweight offsets")) ||
        (rc = dev_alloc((void**)&p->d_rays, sizeof(Ray) * (size_t)nray, "ray table")) ||
        (rc = dev_alloc((void**)&p->d_imgT, sizeof(float) * (size_t)img, "transposed image")) ||
        (rc = dev_alloc((void**)&p->d_sino_ws, sizeof(float) * (size_t)sino, "sinogram workspace")) ||
        (rc = p->build_fwd_items()) || (rc = p->build_adj_items())) {
        free_dev(d_bins);
        free_dev(d_bin_off);
        free_dev(d_ray_prefix);
        delete p;
        return rc;
    }
    launch_build_rays(p->d_desc, nslices, p->d_views, d_bins, d_bin_off, d_ray_prefix, p->d_rays,
                      nray, 0);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    free_dev(d_bins);
    free_dev(d_bin_off);
    free_dev(d_ray_prefix);
    if (e != hipSuccess) {
        rc = hip_fail(e, "building ray table");
        delete p;
        return rc;
    }
    *out = p;
    return 0;
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
    launch_transpose(n_tr_items, d_desc, d_tr_items, d_img, d_imgT, st);
    launch_forward_kernel(count, tune.fwd_block, n_fwd_items, d_desc, d_fwd_items, d_views,
                          d_rays, d_img, d_imgT, d_sino, st);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(e, "forward launch");
}

int Plan::launch_adjoint(const float* d_sino, float* d_img, hipStream_t st, bool fbp)
{
    launch_adjoint_kernel(tune.adj_tile, fbp, n_adj_items, d_desc, d_adj_items, d_views, d_rays,
                          d_sino, d_img, st);
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

extern "C" int ctprj_plan_nnz(ctprj_plan* p, long long* nnz, void* stream)
{
    if (!p || !nnz) return fail(CTPRJ_EINVAL, "ctprj_plan_nnz: NULL argument");
    hipStream_t st = (hipStream_t)stream;
    float* d_img = nullptr;
    int rc = dev_alloc((void**)&d_img, sizeof(float) * (size_t)p->img_elems, "nnz image");
    if (rc) return rc;
    hipError_t e = hipMemsetAsync(d_img, 0, sizeof(float) * (size_t)p->img_elems, st);
    if (e != hipSuccess) { free_dev(d_img); return hip_fail(e, "hipMemsetAsync"); }
    rc = p->launch_forward(d_img, p->d_sino_ws, st, true);
    std::vector<float> counts((size_t)p->sino_elems);
    if (!rc) {
        e = hipMemcpyAsync(counts.data(), p->d_sino_ws, sizeof(float) * counts.size(),
                           hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = hip_fail(e, "nnz readback");
    }
    free_dev(d_img);
    if (rc) return rc;
    long long total = 0;
    for (float c : counts) total += (long long)c;
    *nnz = total;
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
    if (t->adj_tile != 8 && t->adj_tile != 16) return fail(CTPRJ_EINVAL, "adj_tile must be 8 or 16");
    const bool fwd = t->fwd_block != p->tune.fwd_block, adj = t->adj_tile != p->tune.adj_tile;
    p->tune = *t;
    int rc = 0;
    if (fwd && (rc = p->build_fwd_items())) return rc;
    if (adj && (rc = p->build_adj_items())) return rc;
    return 0;
}

static std::string tune_key(const ctprj_plan* p, const char* arch)
{
    std::ostringstream os;
    os << arch;
    for (const ctprj_geom& g : p->geoms)
        os << "|" << g.n << "," << g.prjWidth << "," << g.np << "," << g.prjFull << ","
           << g.dSize << "," << g.effectiveRate << "," << g.d;
    return os.str();
}

extern "C" int ctprj_plan_autotune(ctprj_plan* p, int reps, const char* cache_path, void* stream)
{
    if (!p) return fail(CTPRJ_EINVAL, "plan is NULL");
    if (reps < 1) reps = 3;
    hipStream_t st = (hipStream_t)stream;
    char arch[64] = "gfx950";
    ctprj_device_arch(arch, sizeof(arch));
    const std::string key = tune_key(p, arch);
    if (cache_path) {
        std::ifstream f(cache_path);
        std::string line;
        while (std::getline(f, line)) {
            const size_t sp = line.rfind(' ');
            if (sp == std::string::npos || line.substr(0, sp) != key) continue;
            ctprj_tune t{};
            if (std::sscanf(line.c_str() + sp + 1, "%d,%d,%d", &t.fwd_block, &t.adj_tile, &t.flags) == 3)
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
    float best_f = 1e30f, best_a = 1e30f;
    const int blocks[3] = {64, 128, 256}, tiles[2] = {8, 16};
    for (int bi = 0; bi < 3 && !rc; ++bi) {
        ctprj_tune t = p->tune;
        t.fwd_block = blocks[bi];
        if ((rc = ctprj_plan_set_tune(p, &t))) break;
        float ms = 0;
        for (int r = -1; r < reps && !rc; ++r) {
            (void)hipEventRecord(e0, st);
            rc = p->launch_forward(d_img, d_sino, st, false);
            (void)hipEventRecord(e1, st);
            (void)hipEventSynchronize(e1);
            float m = 0;
            (void)hipEventElapsedTime(&m, e0, e1);
            if (r >= 0) ms += m;
        }
        if (ms < best_f) { best_f = ms; best.fwd_block = blocks[bi]; }
    }
    for (int ti = 0; ti < 2 && !rc; ++ti) {
        ctprj_tune t = p->tune;
        t.adj_tile = tiles[ti];
        if ((rc = ctprj_plan_set_tune(p, &t))) break;
        float ms = 0;
        for (int r = -1; r < reps && !rc; ++r) {
            (void)hipEventRecord(e0, st);
            rc = p->launch_adjoint(d_sino, d_img, st, false);
            (void)hipEventRecord(e1, st);
            (void)hipEventSynchronize(e1);
            float m = 0;
            (void)hipEventElapsedTime(&m, e0, e1);
            if (r >= 0) ms += m;
        }
        if (ms < best_a) { best_a = ms; best.adj_tile = tiles[ti]; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    free_dev(d_img);
    free_dev(d_sino);
    if (rc) return rc;
    rc = ctprj_plan_set_tune(p, &best);
    if (!rc && cache_path) {
        std::ofstream f(cache_path, std::ios::app);
        f << key << " " << best.fwd_block << "," << best.adj_tile << "," << best.flags << "\n";
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
