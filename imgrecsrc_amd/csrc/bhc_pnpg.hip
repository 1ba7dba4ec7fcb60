// On-device PNPG for the alpha step of the beam-hardening reconstruction: the projected
// Nesterov proximal-gradient method of npg/PNPG.m:50-167 with the non-negativity prox
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

#include <cmath>
#include <new>

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
};

// PNPG.m properties (defaults of the reference class)
struct PnpgParams {
    double stepShrnk;  // 0.5: t <- t/stepShrnk on a failed test
    double stepIncre;  // 0.9: t <- t*stepIncre after cumuTol good steps
    double gamma, a;   // theta growth: 1/gamma + sqrt(a + B theta^2)
    int adaptive;      // adaptiveStep
    int restart;       // restart >= 0 enables restarts
    int incCumuTol;
};

__global__ void k_pnpg_begin(PnpgState* S, PnpgParams P)
{
    if (threadIdx.x | blockIdx.x) return;
    if (S->ppp == 0) {  // first trial of an iteration (PNPG.m:58-63)
        S->incStep = 0;
        if (P.adaptive && S->cumu >= S->cumuTol) {
            S->t *= P.stepIncre;
            S->cumu = 0;
            S->incStep = 1;
        }
    }
    S->ppp += 1;
    S->trials += 1;
    const double B = S->t / S->preT;
    S->newTheta = (S->p == 0) ? 1.0 : 1.0 / P.gamma + sqrt(P.a + B * S->theta * S->theta);
    S->coef = (S->theta - 1.0) / S->newTheta;
}

// xbar = max(alpha + coef (alpha - preAlpha), 0)   (PNPG.m:73-77)
__global__ void k_pnpg_xbar(const PnpgState* __restrict__ S, const float* __restrict__ alpha,
                            const float* __restrict__ pre, float* __restrict__ xbar, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float c = (float)S->coef;
    xbar[i] = fmaxf(fmaf(c, alpha[i] - pre[i], alpha[i]), 0.0f);
}

// newX = max(xbar - grad/t, 0) and the FP64 partials of <newX-xbar, grad>, ||newX-xbar||^2
__global__ __launch_bounds__(256) void k_pnpg_prox(const PnpgState* __restrict__ S,
                                                   const float* __restrict__ xbar,
                                                   const float* __restrict__ grad,
                                                   float* __restrict__ newx, long long n,
                                                   double* __restrict__ partials)
{
    __shared__ double sg[256], sd[256];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double dg = 0.0, dd = 0.0;
    if (i < n) {
        const float it = (float)(1.0 / S->t);
        const float x = fmaxf(fmaf(-it, grad[i], xbar[i]), 0.0f);
        newx[i] = x;
        const double d = (double)x - (double)xbar[i];
        dg = d * (double)grad[i];
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

__global__ void k_pnpg_dot_reduce(const double* __restrict__ partials, long long nparts,
                                  PnpgState* S)
{
    __shared__ double sg[256], sd[256];
    const int t = threadIdx.x;
    double a = 0.0, b = 0.0;
    for (long long p = t; p < nparts; p += 256) {
        a += partials[2 * p];
        b += partials[2 * p + 1];
    }
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
    }
}

// Majorization test f(newX) <= f(xbar) + <newX - xbar, g> + t/2 ||newX - xbar||^2
// (utils/Utils.m:470-480), then accept / restart / shrink (PNPG.m:79-146).
__global__ void k_pnpg_decide(PnpgState* S, PnpgParams P)
{
    if (threadIdx.x | blockIdx.x) return;
    const bool holds = (S->f_new - S->f_old) <= S->dg + 0.5 * S->t * S->dd;
    S->accept = 0;
    if (!holds) {
        if (S->ppp <= 20 && S->t > 0) {
            S->t = S->t / P.stepShrnk;
            if (S->incStep) {
                if (P.incCumuTol) S->cumuTol += 4;
                S->incStep = 0;
            }
            S->rejects += 1;
            return;  // retry the same iteration on the next call
        }
        // PNPG.m:94-103 gives up on the line search: accept as it is (goodMM = false)
    } else if (S->p > 0 && S->f_new > S->cost && S->coef != 0.0 && P.restart >= 0) {
        S->theta = 1.0;  // restart: drop the momentum and retry (PNPG.m:103-112)
        S->restarts += 1;
        S->ppp = 0;
        return;
    }
    S->accept = 1;
    S->cumu = (S->ppp == 1 && P.adaptive) ? S->cumu + 1 : 0;
    S->theta = S->newTheta;
    S->cost = S->f_new;
    S->preT = S->t;
    S->p += 1;
    S->ppp = 0;
}

// on acceptance: preAlpha <- alpha, alpha <- newX   (PNPG.m:148-151)
__global__ void k_pnpg_commit(const PnpgState* __restrict__ S, float* __restrict__ alpha,
                              float* __restrict__ pre, const float* __restrict__ newx, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !S->accept) return;
    pre[i] = alpha[i];
    alpha[i] = newx[i];
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
    double* d_ray_part = nullptr; // NLL partials
    long long nimg = 0, nsino = 0, prox_blocks = 0, ray_blocks = 0;
    ~ctprj_pnpg()
    {
        for (void* p : {(void*)d_state, (void*)d_pre, (void*)d_xbar, (void*)d_newx, (void*)d_grad,
                        (void*)d_s, (void*)d_part, (void*)d_ray_part})
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
    s->pp = PnpgParams{0.5, 0.9, 2.0, 0.25, 1, 0, 1};
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
    int rc = p->launch_forward(x, s->d_s, st, false);
    if (rc) return rc;
    launch_bhc_alpha_ray(s->bp, s->d_s, s->Ts, s->d_Imea, s->nsino, s->Ts, grad ? s->d_s : nullptr,
                         nullptr, s->d_ray_part, st);
    launch_reduce_partials(s->d_ray_part, s->ray_blocks, 1, dst, st);
    if (grad) return p->launch_adjoint(s->d_s, grad, st, false);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(e, "pnpg objective");
}

extern "C" int ctprj_pnpg_step(ctprj_pnpg* s, float* d_alpha, void* stream)
{
    if (!s || !d_alpha) return fail(CTPRJ_EINVAL, "ctprj_pnpg_step: NULL argument");
    hipStream_t st = (hipStream_t)stream;
    const unsigned gi = (unsigned)((s->nimg + 255) / 256);
    PnpgState* S = s->d_state;
    hipLaunchKernelGGL(k_pnpg_begin, dim3(1), dim3(64), 0, st, S, s->pp);
    hipLaunchKernelGGL(k_pnpg_xbar, dim3(gi), dim3(256), 0, st, S, d_alpha, s->d_pre, s->d_xbar, s->nimg);
    int rc = pnpg_func(s, s->d_xbar, &S->f_old, s->d_grad, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pnpg_prox, dim3((unsigned)s->prox_blocks), dim3(256), 0, st, S, s->d_xbar,
                       s->d_grad, s->d_newx, s->nimg, s->d_part);
    hipLaunchKernelGGL(k_pnpg_dot_reduce, dim3(1), dim3(256), 0, st, s->d_part, s->prox_blocks, S);
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
    return 0;
}

extern "C" int ctprj_pnpg_destroy(ctprj_pnpg* s)
{
    delete s;
    return 0;
}
