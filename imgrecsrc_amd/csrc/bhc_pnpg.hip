// On-device PNPG for the alpha step of the beam-hardening reconstruction: the projected
// Nesterov proximal-gradient method of npg/PNPG.m:50-167 (or its NPG variant,
// npg/NPG.m:45-148, the reference's default alphaStep) with the non-negativity prox
// (forcePositive, u = 0), on the polychromatic Poisson objective of bhc/poissLAlpha.m.
//
// Every decision of the method (momentum theta, the majorization test of
// utils/Utils.m:470-480, the step shrink / adaptive increase, the restart) is taken by
// a one-thread control kernel on device scalars, so a call enqueues a fixed kernel
// sequence with no host synchronisation and many calls can be captured in one graph.
// One call is one line-search trial: when the majorization test fails the iterate
// stays and the step 1/t shrinks (PNPG.m:77-93's retry happens on the next call); a
// restart (PNPG.m:103-112) likewise resets theta and retries.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "bhc_internal.h"
#include "ctprj_internal.h"

namespace ctprj {

// Device scalars of the method.
struct PnpgState {
    double cost;      // objective at alpha (obj.cost)
    double f_old;     // objective at xbar (this trial)
    double f_new;     // objective at the prox point (this trial)
    double dg;        // <newX - xbar, grad>
    double dd;        // ||newX - xbar||^2
    double t;         // inverse step (obj.t)
    double preT;      // t of the last accepted iteration
    double theta;     // momentum parameter
    double newTheta;  // this trial's theta
    double coef;      // (theta - 1)/newTheta
    int p;            // accepted iterations
    int trials;       // calls made
    int ppp;          // trials within the current iteration
    int cumu, cumuTol;
    int incStep;
    int accept;       // this trial accepted
    int restarts, rejects;
    int outer;        // NPG: outer iterations begun (NPG.m obj.p, retries included)
    int goodStep;     // NPG: no failed test in this iteration yet
    int moved;        // this trial's xbar differs from alpha (||xbar - alpha||_1 != 0)
    int keep;         // this accept keeps alpha ("force it to converge")
    double admmTol;   // non-negativity prox: obj.admmTol, whose retries PNPG.m:125-131 counts
};

// PNPG.m / NPG.m properties (defaults of the reference classes)
struct PnpgParams {
    double stepShrnk;  // 0.5: t <- t/stepShrnk on a failed test
    double stepIncre;  // 0.9: PNPG t <- t*stepIncre after cumuTol good steps
    double gamma, a;   // PNPG theta growth: 1/gamma + sqrt(a + B theta^2)
    int adaptive;      // adaptiveStep
    int restart;       // restart >= 0 enables restarts
    int incCumuTol;
    int npg;           // 1: the NPG variant (npg/NPG.m:45-148)
    int preSteps;      // NPG: iterations whose line search also grows the step
    int maxLs;         // failed tests that shrink the step before giving up (20)
};

// Start of a trial.  PNPG (PNPG.m:56-77): theta and xbar are recomputed each trial
// from B = t/preT.  NPG (NPG.m:48-66): theta = (1 + sqrt(1 + 4 theta^2))/2 and xbar once
// per outer iteration, the adaptive increase divides t by 4 (stepShrnk applied twice).
__host__ __device__ inline void pnpg_begin(PnpgState& S, const PnpgParams& P)
{
    if (S.ppp == 0) {
        S.incStep = 0;
        if (P.npg) {
            S.outer += 1;
            S.goodStep = 1;
            S.newTheta = (1.0 + sqrt(1.0 + 4.0 * S.theta * S.theta)) / 2.0;
            S.coef = (S.theta - 1.0) / S.newTheta;
        }
        if (P.adaptive && S.cumu >= S.cumuTol) {
            if (P.npg) {
                S.t = S.t * P.stepShrnk;
                S.t = S.t * P.stepShrnk;
            } else {
                S.t *= P.stepIncre;
            }
            S.cumu = 0;
            S.incStep = 1;
        }
    }
    S.ppp += 1;
    S.trials += 1;
    S.moved = 0;
    if (!P.npg) {
        const double B = S.t / S.preT;
        S.newTheta = (S.p == 0) ? 1.0 : 1.0 / P.gamma + sqrt(P.a + B * S.theta * S.theta);
        S.coef = (S.theta - 1.0) / S.newTheta;
    }
}

// Line-search outcome of a trial: 0 = another trial follows (shrunk or, NPG's first
// preSteps iterations, grown step), 1 = the test holds, 2 = given up (goodMM false).
// PNPG.m:83-103, NPG.m:71-96.
__host__ __device__ inline int pnpg_linesearch(PnpgState& S, const PnpgParams& P, bool holds)
{
    if (holds) {
        if (P.npg && S.outer <= P.preSteps && S.ppp < 18 && S.goodStep && S.t > 0) {
            S.t = S.t * P.stepShrnk;
            return 0;
        }
        return 1;
    }
    if (S.ppp <= P.maxLs && S.t > 0) {
        S.t = S.t / P.stepShrnk;
        S.goodStep = 0;
        if (S.incStep) {
            if (P.incCumuTol) S.cumuTol += 4;
            S.incStep = 0;
        }
        S.rejects += 1;
        return 0;
    }
    return 2;
}

// Accepting newX with objective newObj (PNPG.m:147-159, NPG.m:137-146).
__host__ __device__ inline void pnpg_accept(PnpgState& S, const PnpgParams& P, double newObj)
{
    S.accept = 1;
    S.cumu = (S.ppp == 1 && P.adaptive) ? S.cumu + 1 : 0;
    S.theta = S.newTheta;
    S.cost = newObj;
    S.preT = S.t;
    S.p += 1;
    S.ppp = 0;
}

__global__ void k_pnpg_begin(PnpgState* S, PnpgParams P)
{
    if (threadIdx.x | blockIdx.x) return;
    PnpgState s = *S;
    pnpg_begin(s, P);
    *S = s;
}

// xbar = max(alpha + coef (alpha - preAlpha), 0)   (PNPG.m:73-77)
// Support mask: alpha is the vector of the mask's pixels (utils/maskFunc.m), so pixels
// outside it (mask 0) stay zero; mask == NULL keeps every pixel.
__device__ __forceinline__ float pm(const float* mask, long long i, float v)
{
    return (!mask || mask[i] != 0.0f) ? v : 0.0f;
}

// moved[work-group] (when given) = whether any pixel of the group's xbar differs from
// alpha; k_pnpg_dot_reduce ORs the flags into S->moved (the restart test's
// ||xbar - alpha||_1 != 0, PNPG.m:104).  A flag per group instead of an atomic on one
// word per changed pixel: those serialised to ~14 ms on a 512 x 512^2 stack.
__global__ __launch_bounds__(256) void k_pnpg_xbar(const PnpgState* __restrict__ S,
                                                   const float* __restrict__ alpha,
                                                   const float* __restrict__ pre,
                                                   float* __restrict__ xbar, long long n,
                                                   const float* __restrict__ mask,
                                                   int* __restrict__ moved)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    int diff = 0;
    if (i < n) {
        const float c = (float)S->coef;
        const float a = alpha[i];
        const float x = pm(mask, i, fmaxf(fmaf(c, a - pre[i], a), 0.0f));
        xbar[i] = x;
        diff = x != a;
    }
    const int any = __syncthreads_or(diff);
    if (moved && threadIdx.x == 0) moved[blockIdx.x] = any;
}

// newX = max(xbar - grad/t, 0) and the FP64 partials of <newX-xbar, grad>, ||newX-xbar||^2
__global__ __launch_bounds__(256) void k_pnpg_prox(const PnpgState* __restrict__ S,
                                                   const float* __restrict__ xbar,
                                                   const float* __restrict__ grad,
                                                   float* __restrict__ newx, long long n,
                                                   const float* __restrict__ mask,
                                                   double* __restrict__ partials)
{
    __shared__ double sg[256], sd[256];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double dg = 0.0, dd = 0.0;
    if (i < n) {
        const float it = (float)(1.0 / S->t);
        const float g = pm(mask, i, grad[i]);
        const float x = fmaxf(fmaf(-it, g, xbar[i]), 0.0f);
        newx[i] = x;
        const double d = (double)x - (double)xbar[i];
        dg = d * (double)g;
        dd = d * d;
    }
    const int t = threadIdx.x;
    sg[t] = dg;
    sd[t] = dd;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off) {
            sg[t] += sg[t + off];
            sd[t] += sd[t + off];
        }
        __syncthreads();
    }
    if (t == 0) {
        partials[2 * (long long)blockIdx.x] = sg[0];
        partials[2 * (long long)blockIdx.x + 1] = sd[0];
    }
}

__global__ void k_pnpg_dot_reduce(const double* __restrict__ partials,
                                  const int* __restrict__ moved, long long nparts, PnpgState* S)
{
    __shared__ double sg[256], sd[256];
    const int t = threadIdx.x;
    double a = 0.0, b = 0.0;
    int m = 0;
    for (long long p = t; p < nparts; p += 256) {
        a += partials[2 * p];
        b += partials[2 * p + 1];
        m |= moved[p];
    }
    m = __syncthreads_or(m);
    sg[t] = a;
    sd[t] = b;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off) {
            sg[t] += sg[t + off];
            sd[t] += sd[t + off];
        }
        __syncthreads();
    }
    if (t == 0) {
        S->dg = sg[0];
        S->dd = sd[0];
        S->moved = m;
    }
}

// The non-negativity prox (customMax, bhc/beamhardenSpline.m:200-203,586-589) reports
// one inner iteration against the reference's maxInnerItr of 100.
constexpr int kNonnegInnerSearch = 1, kNonnegMaxInnerItr = 100;

// The decision after a finished line search with penalty u = 0 (PNPG.m:99-160 /
// NPG.m:99-146; goodMM = (ls == 1)), on the host and the device alike.  Returns 1 when
// the iterate is taken (keep: newX is alpha itself), 0 when the iteration is retried.
__host__ __device__ inline int pnpg_nonneg_decide(PnpgState& s, const PnpgParams& P, int ls)
{
    const bool goodMM = ls == 1;
    double newObj = s.f_new;
    s.keep = 0;
    if (newObj - s.cost > 0) {
        if (goodMM && s.moved && P.restart >= 0) {  // restart: drop the momentum, retry
            s.theta = 1.0;
            s.restarts += 1;
            s.ppp = 0;
            return 0;
        }
        if (!goodMM || s.f_old < newObj) {
            // reset() sets theta = 1 (its step re-estimate from the Hessian is not taken on
            // this graph-capturable path); then the admmTol schedule retries the
            // iteration until admmTol reaches 1e-6, and finally alpha is kept
            if (!goodMM) s.theta = 1.0;
            if (kNonnegInnerSearch < kNonnegMaxInnerItr && s.admmTol > 1e-6) {
                s.admmTol /= 10.0;
                s.ppp = 0;
                return 0;
            }
            newObj = s.cost;
            s.keep = 1;
        }
    }
    pnpg_accept(s, P, newObj);
    return 1;
}

// Majorization test f(newX) <= f(xbar) + <newX - xbar, g> + t/2 ||newX - xbar||^2
// (utils/Utils.m:470-480), then accept / restart / retry / keep (PNPG.m:79-160).
__global__ void k_pnpg_decide(PnpgState* S, PnpgParams P)
{
    if (threadIdx.x | blockIdx.x) return;
    PnpgState s = *S;
    const bool holds = (s.f_new - s.f_old) <= s.dg + 0.5 * s.t * s.dd;
    s.accept = 0;
    s.keep = 0;
    const int ls = pnpg_linesearch(s, P, holds);
    if (ls != 0) pnpg_nonneg_decide(s, P, ls);
    *S = s;
}

// wavelet prox: the ADMM input xbar - grad/t   (PNPG.m:78-79)
__global__ void k_pnpg_gradstep(const PnpgState* __restrict__ S, const float* __restrict__ xbar,
                                const float* __restrict__ grad, float* __restrict__ a,
                                long long n, const float* __restrict__ mask)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float it = (float)(1.0 / S->t);
    a[i] = fmaf(-it, pm(mask, i, grad[i]), xbar[i]);
}

// wavelet prox: FP64 partials of <newX - xbar, grad>, ||newX - xbar||^2,
// ||xbar - alpha||_1, ||alpha - newX||^2, ||newX||^2 (5 per work-group)
__global__ __launch_bounds__(256) void k_pnpg_wv_sums(const float* __restrict__ xbar,
                                                      const float* __restrict__ newx,
                                                      const float* __restrict__ grad,
                                                      const float* __restrict__ alpha, long long n,
                                                      double* __restrict__ partials)
{
    __shared__ double sh[5][256];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (i < n) {
        const double x = newx[i], xb = xbar[i], al = alpha[i];
        const double d = x - xb, e = al - x;
        v[0] = d * (double)grad[i];
        v[1] = d * d;
        v[2] = fabs(xb - al);
        v[3] = e * e;
        v[4] = x * x;
    }
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < 5; ++k) sh[k][t] = v[k];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off) {
#pragma unroll
            for (int k = 0; k < 5; ++k) sh[k][t] += sh[k][t + off];
        }
        __syncthreads();
    }
    if (t == 0) {
#pragma unroll
        for (int k = 0; k < 5; ++k) partials[5 * (long long)blockIdx.x + k] = sh[k][0];
    }
}

// on acceptance: preAlpha <- alpha, alpha <- newX   (PNPG.m:148-151)
// FP64 partial sums of sum_i w_i a_i^2 (w == NULL: sum a_i^2; mask zeroes a outside the
// support) per work-group of 256.
__global__ __launch_bounds__(256) void k_pnpg_sq_sum(const float* __restrict__ a,
                                                     const float* __restrict__ w, long long n,
                                                     const float* __restrict__ mask,
                                                     double* __restrict__ partials)
{
    __shared__ double sh[256];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        const double x = (double)pm(mask, i, a[i]);
        v = w ? (double)w[i] * x * x : x * x;
    }
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off) sh[t] += sh[t + off];
        __syncthreads();
    }
    if (t == 0) partials[blockIdx.x] = sh[0];
}

// a = a * mask (pixels outside the support set to zero)
__global__ void k_pnpg_mask(float* __restrict__ a, long long n, const float* __restrict__ mask)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = pm(mask, i, a[i]);
}

__global__ void k_pnpg_commit(const PnpgState* __restrict__ S, float* __restrict__ alpha,
                              float* __restrict__ pre, const float* __restrict__ newx, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !S->accept) return;
    const float a = alpha[i];
    pre[i] = a;
    if (!S->keep) alpha[i] = newx[i];
}

}  // namespace ctprj

using namespace ctprj;

struct ctprj_plan : public Plan {};

struct ctprj_pnpg {
    ctprj_plan* plan = nullptr;
    BhcParams bp{};
    PnpgParams pp{};
    float Ts = 1.0f;
    const float* d_Imea = nullptr;
    PnpgState* d_state = nullptr;
    float *d_pre = nullptr, *d_xbar = nullptr, *d_newx = nullptr, *d_grad = nullptr;
    float* d_s = nullptr;         // sinogram: projections, then gradient weights
    double* d_part = nullptr;     // prox partials (2 per work-group)
    int* d_moved = nullptr;       // per work-group: xbar differs from alpha
    double* d_ray_part = nullptr; // NLL partials
    long long nimg = 0, nsino = 0, prox_blocks = 0, ray_blocks = 0;
    // wavelet prox (PNPG.m properties admmTol, maxInnerItr, maxPossibleInnerItr)
    ctprj_wavelet* wv = nullptr;  // at most one of wv / tv is set
    ctprj_tv* tv = nullptr;
    double u = 0.0, admmTol = 1e-2, difAlpha = 1.0;
    int maxInnerItr = 100, maxPossibleInnerItr = 1000, innerSearch = 0;
    double* d_wpart = nullptr;    // 5 partials per work-group
    double* d_ppart = nullptr;    // partials of the two penalties of a trial
    long long ppart_cap = 0;      // doubles per penalty held by d_ppart
    float* d_mask = nullptr;      // support of alpha (NULL: every pixel)
    // host-driven trials: the state as last queued to d_state (f_old / f_new excepted,
    // which the trial's kernels write)
    PnpgState h_state{};
    bool h_valid = false;
    double t_reset = 0.0;         // last Hessian step estimate of reset()
    int resets = 0;
    ~ctprj_pnpg()
    {
        for (void* p : {(void*)d_state, (void*)d_pre, (void*)d_xbar, (void*)d_newx, (void*)d_grad,
                        (void*)d_s, (void*)d_part, (void*)d_moved, (void*)d_ray_part, (void*)d_wpart,
                        (void*)d_ppart, (void*)d_mask})
            if (p) (void)hipFree(p);
    }
};

static int pnpg_alloc(void** p, size_t bytes, const char* what)
{
    hipError_t e = hipMalloc(p, bytes ? bytes : 16);
    if (e != hipSuccess) return fail(CTPRJ_ENOMEM, "hipMalloc(%s): %s", what, hipGetErrorString(e));
    return 0;
}

extern "C" int ctprj_pnpg_create(ctprj_plan* plan, const float* d_Imea, const double* knots, int E,
                                 const double* Ie, float Ts, const float* d_alpha0, double t0,
                                 ctprj_pnpg** out)
{
    if (!plan || !d_Imea || !knots || !Ie || !d_alpha0 || !out)
        return fail(CTPRJ_EINVAL, "ctprj_pnpg_create: NULL argument");
    if (!(t0 > 0.0) || !std::isfinite(t0)) return fail(CTPRJ_EINVAL, "t0 must be > 0");
    *out = nullptr;
    BhcParams bp;
    int rc0 = make_bhc_params(knots, E, Ie, plan->bhc_basis, plan->bhc_noise, &bp);
    if (rc0) return rc0;
    ctprj_pnpg* s = new (std::nothrow) ctprj_pnpg();
    if (!s) return fail(CTPRJ_ENOMEM, "host allocation");
    s->plan = plan;
    s->d_Imea = d_Imea;
    s->Ts = Ts;
    s->bp = bp;
    s->pp = PnpgParams{0.5, 0.9, 2.0, 0.25, 1, 0, 1, 0, 2, 20};
    s->nimg = plan->img_elems;
    s->nsino = plan->sino_elems;
    s->prox_blocks = (s->nimg + 255) / 256;
    s->ray_blocks = bhc_blocks(s->nsino);
    int rc;
    if ((rc = pnpg_alloc((void**)&s->d_state, sizeof(PnpgState), "pnpg state")) ||
        (rc = pnpg_alloc((void**)&s->d_pre, sizeof(float) * s->nimg, "pnpg preAlpha")) ||
        (rc = pnpg_alloc((void**)&s->d_xbar, sizeof(float) * s->nimg, "pnpg xbar")) ||
        (rc = pnpg_alloc((void**)&s->d_newx, sizeof(float) * s->nimg, "pnpg newX")) ||
        (rc = pnpg_alloc((void**)&s->d_grad, sizeof(float) * s->nimg, "pnpg grad")) ||
        (rc = pnpg_alloc((void**)&s->d_s, sizeof(float) * s->nsino, "pnpg sinogram")) ||
        (rc = pnpg_alloc((void**)&s->d_part, sizeof(double) * 2 * s->prox_blocks, "pnpg partials")) ||
        (rc = pnpg_alloc((void**)&s->d_moved, sizeof(int) * s->prox_blocks, "pnpg moved flags")) ||
        (rc = pnpg_alloc((void**)&s->d_ray_part, sizeof(double) * s->ray_blocks, "pnpg NLL partials"))) {
        delete s;
        return rc;
    }
    PnpgState st{};
    st.t = t0;
    st.preT = t0;
    st.theta = 0.0;
    st.cumuTol = 4;
    st.cost = INFINITY;
    st.admmTol = 1e-2;  // PNPG.m / NPG.m default
    hipError_t e = hipMemcpy(s->d_state, &st, sizeof(st), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(s->d_pre, d_alpha0, sizeof(float) * s->nimg, hipMemcpyDeviceToDevice);
    if (e != hipSuccess) {
        delete s;
        return hip_fail(e, "pnpg init");
    }
    *out = s;
    return 0;
}

// objective at x (device double at dst): forward + FP64 NLL; with grad != NULL also the
// gradient weights and the adjoint (poissLAlpha)
static int pnpg_func(ctprj_pnpg* s, const float* x, double* dst, float* grad, hipStream_t st)
{
    Plan* p = s->plan;
    int rc;
    // plans whose slices all take the symmetric kernels: the ray pass runs in the forward's
    // store and the adjoint's first wave adds the NLL partials (as ctprj_bhc_grad_alpha
    // does; the same w, so the same gradient)
    const long long fused = p->bhc_fused_parts();
    if (fused >= 0 && fused <= p->red_cap) {
        if ((rc = p->launch_forward_bhc(x, s->d_s, BhcStore{s->bp, s->d_Imea, p->d_red, s->Ts, 0},
                                        st)))
            return rc;
        const RedJob job{p->d_red, fused, dst};
        if (grad) return p->launch_adjoint(s->d_s, grad, st, false, job);
        launch_wave_sum(job, st);
        hipError_t e = hipGetLastError();
        return e == hipSuccess ? 0 : hip_fail(e, "pnpg objective");
    }
    rc = p->launch_forward(x, s->d_s, st, false);
    if (rc) return rc;
    launch_bhc_alpha_ray(s->bp, s->d_s, s->Ts, s->d_Imea, s->nsino, s->Ts, grad ? s->d_s : nullptr,
                         nullptr, s->d_ray_part, st);
    launch_reduce_partials(s->d_ray_part, s->ray_blocks, 1, dst, st);
    if (grad) return p->launch_adjoint(s->d_s, grad, st, false);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(e, "pnpg objective");
}

// reset()'s step estimate stepSizeInit('hessian') (npg/PNPG.m:168-172,
// npg/Methods.m:108-115) at alpha: with g the (masked) gradient and Phi = Ts*A,
// t = hessian(g, 2)/||g||^2 = Ts^2 <A g, hw .* A g>/||g||^2, hw the Hessian ray weight of
// bhc/poissLAlpha.m:15 (or bhc/gaussLAlpha.m:18); NaN gives 1 and 0 gives eps as there.
// Uses d_xbar and d_s (free after the trial's decision) and the plan's sinogram
// workspace; synchronises.
static int pnpg_hessian_step(ctprj_pnpg* s, const float* d_alpha, double* t_out, hipStream_t st)
{
    Plan* p = s->plan;
    float* hw = p->d_sino_ws;
    int rc = p->launch_forward(d_alpha, s->d_s, st, false);
    if (rc) return rc;
    launch_bhc_alpha_ray(s->bp, s->d_s, s->Ts, s->d_Imea, s->nsino, s->Ts, s->d_s, hw,
                         s->d_ray_part, st);
    if ((rc = p->launch_adjoint(s->d_s, s->d_xbar, st, false))) return rc;
    const unsigned gi = (unsigned)((s->nimg + 255) / 256);
    if (s->d_mask)
        hipLaunchKernelGGL(k_pnpg_mask, dim3(gi), dim3(256), 0, st, s->d_xbar, s->nimg, s->d_mask);
    if ((rc = p->launch_forward(s->d_xbar, s->d_s, st, false))) return rc;
    // <A g, hw .* A g> over the rays, then ||g||^2 over the pixels, each to a device double
    double* sums = s->d_part;  // two doubles: the prox partials are free here
    hipLaunchKernelGGL(k_pnpg_sq_sum, dim3((unsigned)s->ray_blocks), dim3(256), 0, st, s->d_s, hw,
                       s->nsino, nullptr, s->d_ray_part);
    launch_reduce_partials(s->d_ray_part, s->ray_blocks, 1, sums, st);
    hipLaunchKernelGGL(k_pnpg_sq_sum, dim3((unsigned)s->prox_blocks), dim3(256), 0, st, s->d_xbar,
                       nullptr, s->nimg, nullptr, s->d_wpart);
    launch_reduce_partials(s->d_wpart, s->prox_blocks, 1, sums + 1, st);
    double h[2] = {0.0, 0.0};
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h, sums, sizeof(h), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "pnpg hessian step");
    double t = (double)s->Ts * (double)s->Ts * h[0] / h[1];
    if (std::isnan(t)) t = 1.0;
    if (t == 0.0) t = 2.220446049250313e-16;
    *t_out = t;
    return 0;
}

static int state_get(ctprj_pnpg* s, PnpgState* h, hipStream_t st)
{
    hipError_t e = hipMemcpyAsync(h, s->d_state, sizeof(*h), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e == hipSuccess ? 0 : hip_fail(e, "pnpg state read");
}

__global__ void k_pnpg_state_put(PnpgState* __restrict__ S, PnpgState h)
{
    if (threadIdx.x == 0) *S = h;
}

// Queues the host's state to the device (a kernel argument: no copy, no wait) and keeps
// it as the host-driven trials' copy.
static int state_put(ctprj_pnpg* s, const PnpgState* h, hipStream_t st)
{
    hipLaunchKernelGGL(k_pnpg_state_put, dim3(1), dim3(64), 0, st, s->d_state, *h);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "pnpg state write");
    s->h_state = *h;
    s->h_valid = true;
    return 0;
}

// The iterative prox of the solver: wavelet ADMM (npg/admm.m) or TV FGP
// (utils/TV.m:56-80), and its penalty ||W x||_1 or TV(x).
static int prox_solve(ctprj_pnpg* s, const float* a, double u, double tol, const float* init,
                      float* out, int* iters, hipStream_t st)
{
    if (s->wv) return ctprj_prox_wavelet_admm(s->wv, a, u, tol, s->maxInnerItr, init, out, iters, st);
    return ctprj_tv_denoise(s->tv, a, u, tol, s->maxInnerItr, out, iters, st);
}

// The penalty ||W x||_1 or TV(x) queued as partial sums (penalty_parts doubles), added
// on the host by penalty_total.
static long long penalty_parts(const ctprj_pnpg* s)
{
    return s->wv ? wavelet_l1_parts(s->wv) : tv_value_parts(s->tv);
}

static int penalty_queue(ctprj_pnpg* s, const float* x, double* d_part, hipStream_t st)
{
    return s->wv ? wavelet_l1_queue(s->wv, x, d_part, st) : tv_value_queue(s->tv, x, d_part, st);
}

static double penalty_total(const ctprj_pnpg* s, const double* part)
{
    return s->wv ? wavelet_l1_total(s->wv, part) : tv_value_total(s->tv, part);
}

// One trial with an iterative prox: the same sequence as the device trial, with the
// decisions of PNPG.m:64-160 taken on the host (the prox's inner loop needs it anyway).
static int pnpg_step_iterative(ctprj_pnpg* s, float* d_alpha, hipStream_t st)
{
    const PnpgParams& P = s->pp;
    PnpgState h;
    int rc = 0;
    if (s->h_valid) h = s->h_state;  // the device state is what the last trial queued
    else if ((rc = state_get(s, &h, st))) return rc;
    pnpg_begin(h, P);
    h.accept = 0;
    h.keep = 0;
    if ((rc = state_put(s, &h, st))) return rc;
    const unsigned gi = (unsigned)((s->nimg + 255) / 256);
    PnpgState* S = s->d_state;
    hipLaunchKernelGGL(k_pnpg_xbar, dim3(gi), dim3(256), 0, st, S, d_alpha, s->d_pre, s->d_xbar, s->nimg,
                       s->d_mask, nullptr);  // the trial's own sums test xbar != alpha
    if ((rc = pnpg_func(s, s->d_xbar, &S->f_old, s->d_grad, st))) return rc;
    hipLaunchKernelGGL(k_pnpg_gradstep, dim3(gi), dim3(256), 0, st, S, s->d_xbar, s->d_grad,
                       s->d_newx, s->nimg, s->d_mask);
    // newX = prox(xbar - grad/t, u/t, admmTol*difAlpha, maxInnerItr, alpha)  (PNPG.m:78-82)
    if ((rc = prox_solve(s, s->d_newx, s->u / h.t, s->admmTol * s->difAlpha, d_alpha,
                         s->d_newx, &s->innerSearch, st)))
        return rc;
    if ((rc = pnpg_func(s, s->d_newx, &S->f_new, nullptr, st))) return rc;
    hipLaunchKernelGGL(k_pnpg_wv_sums, dim3((unsigned)s->prox_blocks), dim3(256), 0, st, s->d_xbar,
                       s->d_newx, s->d_grad, d_alpha, s->nimg, s->d_wpart);
    // the penalties at newX and xbar (needed unless another trial follows) are queued
    // with the line-search sums, so that one read-back serves both
    const long long np = penalty_parts(s);
    if ((rc = penalty_queue(s, s->d_newx, s->d_ppart, st)) ||
        (rc = penalty_queue(s, s->d_xbar, s->d_ppart + np, st)))
        return rc;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "pnpg wavelet trial");
    std::vector<double> part((size_t)5 * s->prox_blocks), pen((size_t)(2 * np));
    e = hipMemcpyAsync(part.data(), s->d_wpart, sizeof(double) * part.size(),
                       hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(pen.data(), s->d_ppart, sizeof(double) * pen.size(),
                           hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return hip_fail(e, "pnpg wavelet sums");
    PnpgState r;
    if ((rc = state_get(s, &r, st))) return rc;  // f_old, f_new (synchronises)
    double sum[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (long long b = 0; b < s->prox_blocks; ++b)
        for (int k = 0; k < 5; ++k) sum[k] += part[5 * b + k];
    h.f_old = r.f_old;
    h.f_new = r.f_new;
    h.dg = sum[0];
    h.dd = sum[1];
    // majorization test of the smooth part (utils/Utils.m:470-480)
    const bool holds = (h.f_new - h.f_old) <= h.dg + 0.5 * h.t * h.dd;
    const int ls = pnpg_linesearch(h, P, holds);
    if (ls == 0) return state_put(s, &h, st);
    const bool goodMM = ls == 1;
    const double pen_new = penalty_total(s, pen.data()), pen_bar = penalty_total(s, pen.data() + np);
    double newObj = h.f_new + s->u * pen_new;
    const double objBar = h.f_old + s->u * pen_bar;
    bool keep = false;  // "force it to converge": newX = alpha (PNPG.m:141-144)
    if (newObj - h.cost > 0) {
        if (goodMM && sum[2] != 0.0 && P.restart >= 0) {  // PNPG.m:103-112
            h.theta = 1.0;
            h.restarts += 1;
            h.ppp = 0;
            return state_put(s, &h, st);
        } else if (!goodMM || objBar < newObj) {
            if (!goodMM) {  // reset(): theta = 1, t = min(t, Hessian step at alpha)
                h.theta = 1.0;
                double th = 0.0;
                if ((rc = pnpg_hessian_step(s, d_alpha, &th, st))) return rc;
                h.t = std::min(h.t, th);
                s->t_reset = th;
                s->resets += 1;
            }
            if (s->innerSearch < s->maxInnerItr && s->admmTol > 1e-6) {
                s->admmTol /= 10;
                h.ppp = 0;
                return state_put(s, &h, st);
            } else if (s->innerSearch >= s->maxInnerItr &&
                       s->maxInnerItr < s->maxPossibleInnerItr) {
                s->maxInnerItr *= 10;
                h.ppp = 0;
                return state_put(s, &h, st);
            }
            newObj = h.cost;
            keep = true;
        }
    }
    // accept (PNPG.m:147-160)
    pnpg_accept(h, P, newObj);
    s->difAlpha = keep ? 0.0 : std::sqrt(sum[3]) / std::max(std::sqrt(sum[4]), 1e-10);
    h.keep = keep ? 1 : 0;  // the commit then leaves alpha as it is
    if ((rc = state_put(s, &h, st))) return rc;
    hipLaunchKernelGGL(k_pnpg_commit, dim3(gi), dim3(256), 0, st, S, d_alpha, s->d_pre, s->d_newx, s->nimg);
    e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(e, "pnpg wavelet commit");
}

static int set_prox(ctprj_pnpg* s, ctprj_wavelet* w, ctprj_tv* tv, double u)
{
    if (!s) return fail(CTPRJ_EINVAL, "ctprj_pnpg_set_wavelet/tv: NULL solver");
    const long long held = w ? wavelet_elems(w) : (tv ? tv_elems(tv) : s->nimg);
    if (held != s->nimg)
        return fail(CTPRJ_EINVAL, "prox context holds %lld pixels, the plan %lld", held, s->nimg);
    if (!std::isfinite(u) || u < 0) return fail(CTPRJ_EINVAL, "u must be finite and >= 0");
    if (tv && !(u > 0)) return fail(CTPRJ_EINVAL, "the TV prox needs u > 0");
    PnpgState h;
    int rc = state_get(s, &h, nullptr);
    if (rc) return rc;
    if (h.trials) return fail(CTPRJ_ESTATE, "prox change after the first trial");
    if ((w || tv) && !s->d_wpart &&
        (rc = pnpg_alloc((void**)&s->d_wpart, sizeof(double) * 5 * s->prox_blocks, "pnpg sums")))
        return rc;
    const long long np = w ? wavelet_l1_parts(w) : (tv ? tv_value_parts(tv) : 0);
    if (np > s->ppart_cap) {
        if (s->d_ppart) (void)hipFree(s->d_ppart);
        s->d_ppart = nullptr;
        s->ppart_cap = 0;
        if ((rc = pnpg_alloc((void**)&s->d_ppart, sizeof(double) * 2 * np, "pnpg penalty sums")))
            return rc;
        s->ppart_cap = np;
    }
    s->wv = w;
    s->tv = tv;
    s->u = (w || tv) ? u : 0.0;
    return 0;
}

extern "C" int ctprj_pnpg_set_mask(ctprj_pnpg* s, const float* d_mask)
{
    if (!s) return fail(CTPRJ_EINVAL, "ctprj_pnpg_set_mask: NULL solver");
    if (s->d_mask) (void)hipFree(s->d_mask);
    s->d_mask = nullptr;
    if (!d_mask) return 0;
    const size_t bytes = sizeof(float) * (size_t)s->nimg;
    hipError_t e = hipMalloc((void**)&s->d_mask, bytes);
    if (e == hipSuccess) e = hipMemcpy(s->d_mask, d_mask, bytes, hipMemcpyDeviceToDevice);
    return e == hipSuccess ? 0 : hip_fail(e, "pnpg mask");
}

extern "C" int ctprj_pnpg_set_method(ctprj_pnpg* s, int npg, int preSteps)
{
    if (!s) return fail(CTPRJ_EINVAL, "ctprj_pnpg_set_method: NULL solver");
    if (preSteps < 0) return fail(CTPRJ_EINVAL, "preSteps must be >= 0");
    PnpgState h;
    int rc = state_get(s, &h, nullptr);
    if (rc) return rc;
    if (h.trials) return fail(CTPRJ_ESTATE, "method change after the first trial");
    s->pp.npg = npg ? 1 : 0;
    s->pp.preSteps = preSteps;
    return 0;
}

extern "C" int ctprj_pnpg_set_linesearch(ctprj_pnpg* s, int max_shrinks)
{
    if (!s) return fail(CTPRJ_EINVAL, "ctprj_pnpg_set_linesearch: NULL solver");
    if (max_shrinks < 0 || max_shrinks > 1000) return fail(CTPRJ_EINVAL, "max_shrinks must be 0 to 1000");
    s->pp.maxLs = max_shrinks;
    return 0;
}

extern "C" int ctprj_pnpg_set_wavelet(ctprj_pnpg* s, ctprj_wavelet* w, double u)
{
    return set_prox(s, w, nullptr, u);
}

extern "C" int ctprj_pnpg_set_tv(ctprj_pnpg* s, ctprj_tv* tv, double u)
{
    return set_prox(s, nullptr, tv, u);
}

extern "C" int ctprj_pnpg_step(ctprj_pnpg* s, float* d_alpha, void* stream)
{
    if (!s || !d_alpha) return fail(CTPRJ_EINVAL, "ctprj_pnpg_step: NULL argument");
    hipStream_t st = (hipStream_t)stream;
    if (s->wv || s->tv) return pnpg_step_iterative(s, d_alpha, st);
    const unsigned gi = (unsigned)((s->nimg + 255) / 256);
    PnpgState* S = s->d_state;
    hipLaunchKernelGGL(k_pnpg_begin, dim3(1), dim3(64), 0, st, S, s->pp);
    hipLaunchKernelGGL(k_pnpg_xbar, dim3(gi), dim3(256), 0, st, S, d_alpha, s->d_pre, s->d_xbar, s->nimg,
                       s->d_mask, s->d_moved);
    int rc = pnpg_func(s, s->d_xbar, &S->f_old, s->d_grad, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pnpg_prox, dim3((unsigned)s->prox_blocks), dim3(256), 0, st, S, s->d_xbar,
                       s->d_grad, s->d_newx, s->nimg, s->d_mask, s->d_part);
    hipLaunchKernelGGL(k_pnpg_dot_reduce, dim3(1), dim3(256), 0, st, s->d_part, s->d_moved,
                       s->prox_blocks, S);
    if ((rc = pnpg_func(s, s->d_newx, &S->f_new, nullptr, st))) return rc;
    hipLaunchKernelGGL(k_pnpg_decide, dim3(1), dim3(64), 0, st, S, s->pp);
    hipLaunchKernelGGL(k_pnpg_commit, dim3(gi), dim3(256), 0, st, S, d_alpha, s->d_pre, s->d_newx, s->nimg);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(e, "pnpg step");
}

extern "C" int ctprj_pnpg_stats(ctprj_pnpg* s, ctprj_pnpg_info* info, void* stream)
{
    if (!s || !info) return fail(CTPRJ_EINVAL, "ctprj_pnpg_stats: NULL argument");
    PnpgState st;
    hipStream_t hs = (hipStream_t)stream;
    hipError_t e = hipMemcpyAsync(&st, s->d_state, sizeof(st), hipMemcpyDeviceToHost, hs);
    if (e == hipSuccess) e = hipStreamSynchronize(hs);
    if (e != hipSuccess) return hip_fail(e, "pnpg stats");
    info->cost = st.cost;
    info->t = st.t;
    info->theta = st.theta;
    info->iterations = st.p;
    info->trials = st.trials;
    info->rejects = st.rejects;
    info->restarts = st.restarts;
    info->last_accepted = st.accept;
    info->difAlpha = s->difAlpha;
    info->admmTol = (s->wv || s->tv) ? s->admmTol : st.admmTol;
    info->innerSearch = s->innerSearch;
    info->maxInnerItr = s->maxInnerItr;
    info->t_reset = s->t_reset;
    info->resets = s->resets;
    return 0;
}

extern "C" int ctprj_debug_pnpg_decide(const double* in, int ls, int moved, double* out,
                                       int* code)
{
    if (!in || !out || !code) return fail(CTPRJ_EINVAL, "ctprj_debug_pnpg_decide: NULL argument");
    if (ls != 1 && ls != 2) return fail(CTPRJ_EINVAL, "ls must be 1 or 2");
    PnpgState s{};
    s.cost = in[0];
    s.f_old = in[1];
    s.f_new = in[2];
    s.theta = in[3];
    s.admmTol = in[4];
    s.t = s.preT = 1.0;
    s.newTheta = in[3];
    s.ppp = 1;
    s.moved = moved;
    PnpgParams P{0.5, 0.9, 2.0, 0.25, 1, 0, 1, 0, 2, 20};
    const int took = pnpg_nonneg_decide(s, P, ls);
    *code = took ? (s.keep ? 3 : 2) : (s.ppp == 0 && s.restarts ? 0 : 1);
    out[0] = s.cost;
    out[1] = s.theta;
    out[2] = s.admmTol;
    return 0;
}

extern "C" int ctprj_pnpg_destroy(ctprj_pnpg* s)
{
    delete s;
    return 0;
}

// ------------------------------------------------------------ snapshot / resume ---
//
// bhc/doBHC.m:78-91 runs the alpha step at five penalty levels, warm-starting each level
// from the previous one's alpha and Ie and saving the workspace after each.  A snapshot
// holds what a solver carries from one trial to the next: its device state, preAlpha
// and the host-side schedule (prox tolerance and inner-iteration limits, the reset()
// estimate, counters).  Loaded into a solver created with the same plan and arguments
// it continues bit for bit.  alpha (the caller's array), Ie and the mask are the
// caller's to keep; the wavelet and TV contexts carry nothing from one prox call to the
// next.  Layout: SnapHeader, PnpgState, SnapHost, float preAlpha[nimg].

namespace {

constexpr unsigned kSnapMagic = 0x474e504eu;  // "NPNG" little-endian
constexpr unsigned kSnapVersion = 1;

struct SnapHeader {
    unsigned magic, version;
    long long nimg, nsino;
    int E, prox;  // prox: 0 non-negativity, 1 wavelet, 2 TV
    int state_bytes, host_bytes;
};

struct SnapHost {
    PnpgParams pp;
    PnpgState h_state;
    double u, admmTol, difAlpha, t_reset;
    int maxInnerItr, maxPossibleInnerItr, innerSearch, h_valid, resets, pad;
};

long long snapshot_bytes(const ctprj_pnpg* s)
{
    return (long long)(sizeof(SnapHeader) + sizeof(PnpgState) + sizeof(SnapHost)) +
           (long long)sizeof(float) * s->nimg;
}

SnapHeader snapshot_header(const ctprj_pnpg* s)
{
    SnapHeader h{};
    h.magic = kSnapMagic;
    h.version = kSnapVersion;
    h.nimg = s->nimg;
    h.nsino = s->nsino;
    h.E = s->bp.E;
    h.prox = s->wv ? 1 : (s->tv ? 2 : 0);
    h.state_bytes = (int)sizeof(PnpgState);
    h.host_bytes = (int)sizeof(SnapHost);
    return h;
}

}  // namespace

extern "C" int ctprj_pnpg_snapshot_bytes(const ctprj_pnpg* s, long long* bytes)
{
    if (!s || !bytes) return fail(CTPRJ_EINVAL, "ctprj_pnpg_snapshot_bytes: NULL argument");
    *bytes = snapshot_bytes(s);
    return 0;
}

extern "C" int ctprj_pnpg_save(ctprj_pnpg* s, void* host, long long bytes, void* stream)
{
    if (!s || !host) return fail(CTPRJ_EINVAL, "ctprj_pnpg_save: NULL argument");
    if (bytes != snapshot_bytes(s))
        return fail(CTPRJ_EINVAL, "ctprj_pnpg_save: buffer of %lld bytes, the snapshot takes %lld",
                    bytes, snapshot_bytes(s));
    hipStream_t st = (hipStream_t)stream;
    char* b = static_cast<char*>(host);
    const SnapHeader h = snapshot_header(s);
    std::memcpy(b, &h, sizeof h);
    b += sizeof h;
    // the device state and preAlpha after every trial queued on the stream
    hipError_t e = hipMemcpyAsync(b, s->d_state, sizeof(PnpgState), hipMemcpyDeviceToHost, st);
    char* pre = b + sizeof(PnpgState) + sizeof(SnapHost);
    if (e == hipSuccess)
        e = hipMemcpyAsync(pre, s->d_pre, sizeof(float) * s->nimg, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "ctprj_pnpg_save");
    SnapHost hs{};
    hs.pp = s->pp;
    hs.h_state = s->h_state;
    hs.u = s->u;
    hs.admmTol = s->admmTol;
    hs.difAlpha = s->difAlpha;
    hs.t_reset = s->t_reset;
    hs.maxInnerItr = s->maxInnerItr;
    hs.maxPossibleInnerItr = s->maxPossibleInnerItr;
    hs.innerSearch = s->innerSearch;
    hs.h_valid = s->h_valid ? 1 : 0;
    hs.resets = s->resets;
    std::memcpy(b + sizeof(PnpgState), &hs, sizeof hs);
    return 0;
}

extern "C" int ctprj_pnpg_load(ctprj_pnpg* s, const void* host, long long bytes, void* stream)
{
    if (!s || !host) return fail(CTPRJ_EINVAL, "ctprj_pnpg_load: NULL argument");
    if (bytes != snapshot_bytes(s))
        return fail(CTPRJ_EINVAL, "ctprj_pnpg_load: %lld bytes, this solver's snapshot takes %lld",
                    bytes, snapshot_bytes(s));
    const char* b = static_cast<const char*>(host);
    SnapHeader h;
    std::memcpy(&h, b, sizeof h);
    const SnapHeader want = snapshot_header(s);
    if (h.magic != kSnapMagic || h.version != kSnapVersion)
        return fail(CTPRJ_EINVAL, "ctprj_pnpg_load: not a PNPG snapshot (version %u)", kSnapVersion);
    if (h.nimg != want.nimg || h.nsino != want.nsino || h.E != want.E ||
        h.state_bytes != want.state_bytes || h.host_bytes != want.host_bytes)
        return fail(CTPRJ_EINVAL,
                    "ctprj_pnpg_load: snapshot of %lld pixels / %lld rays / E = %d, solver of "
                    "%lld / %lld / %d", h.nimg, h.nsino, h.E, want.nimg, want.nsino, want.E);
    if (h.prox != want.prox)
        return fail(CTPRJ_EINVAL, "ctprj_pnpg_load: snapshot prox %d, solver prox %d (0 none, "
                    "1 wavelet, 2 TV)", h.prox, want.prox);
    b += sizeof h;
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipMemcpyAsync(s->d_state, b, sizeof(PnpgState), hipMemcpyHostToDevice, st);
    const char* pre = b + sizeof(PnpgState) + sizeof(SnapHost);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s->d_pre, pre, sizeof(float) * s->nimg, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);  // the caller may free `host`
    if (e != hipSuccess) return hip_fail(e, "ctprj_pnpg_load");
    SnapHost hs;
    std::memcpy(&hs, b + sizeof(PnpgState), sizeof hs);
    s->pp = hs.pp;
    s->h_state = hs.h_state;
    s->u = hs.u;
    s->admmTol = hs.admmTol;
    s->difAlpha = hs.difAlpha;
    s->t_reset = hs.t_reset;
    s->maxInnerItr = hs.maxInnerItr;
    s->maxPossibleInnerItr = hs.maxPossibleInnerItr;
    s->innerSearch = hs.innerSearch;
    s->h_valid = hs.h_valid != 0;
    s->resets = hs.resets;
    return 0;
}
