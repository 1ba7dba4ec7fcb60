// Host entry points of the beam-hardening model (bhc/poissLAlpha.m, bhc/poissLI.m,
// bhc/gaussLAlpha.m, bhc/gaussLI.m, bhc/Spline.m polyIout) on top of a projector plan.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "bhc_internal.h"
#include "ctprj_internal.h"

namespace ctprj {

int make_bhc_params(const double* knots, int E, const double* Ie, int basis, int noise,
                    BhcParams* P)
{
    if (!knots) return fail(CTPRJ_EINVAL, "knots is NULL");
    if (E < 1 || E > CTPRJ_MAX_E) return fail(CTPRJ_EINVAL, "E=%d must be in [1, %d]", E, CTPRJ_MAX_E);
    if (basis != CTPRJ_BASIS_B1 && basis != CTPRJ_BASIS_B0 && basis != CTPRJ_BASIS_DIS)
        return fail(CTPRJ_EINVAL, "unknown spectrum basis %d", basis);
    if (noise != CTPRJ_NOISE_POISSON && noise != CTPRJ_NOISE_GAUSSIAN)
        return fail(CTPRJ_EINVAL, "unknown noise model %d", noise);
    const int nk = (basis == CTPRJ_BASIS_B0) ? E + 1 : E + 2;  // knots read
    for (int j = 0; j < nk; ++j)
        if (!std::isfinite(knots[j])) return fail(CTPRJ_EINVAL, "knot %d not finite", j);
    for (int j = 0; j + 1 < nk; ++j)
        if (!(knots[j + 1] > knots[j])) return fail(CTPRJ_EINVAL, "knots must increase (at %d)", j);
    *P = BhcParams{};
    P->E = E;
    P->basis = basis;
    P->noise = noise;
    for (int j = 0; j < nk; ++j) P->knot[j] = (float)knots[j];
    if (basis == CTPRJ_BASIS_DIS)
        for (int j = 0; j < E; ++j) P->delta[j] = (float)((knots[j + 2] - knots[j]) / 2.0);
    else
        for (int j = 0; j + 1 < nk; ++j) P->delta[j] = (float)(knots[j + 1] - knots[j]);
    for (int j = 0; j < E; ++j) {
        const double v = Ie ? Ie[j] : 0.0;
        if (!std::isfinite(v)) return fail(CTPRJ_EINVAL, "Ie[%d] not finite", j);
        P->Ie[j] = (float)v;
        P->Ied[j] = v;
    }
    return 0;
}

static int make_params(const Plan* p, const double* knots, int E, const double* Ie, BhcParams* P)
{
    return make_bhc_params(knots, E, Ie, p->bhc_basis, p->bhc_noise, P);
}

// The plan holds the reduction partials of every beam-hardening pass from its creation
// (bhc_red_doubles); a call never allocates.
static int check_red(const Plan* p, long long doubles)
{
    if (p->red_cap >= doubles && p->d_red && p->d_red_out) return 0;
    return fail(CTPRJ_ESTATE, "reduction workspace: %lld doubles needed, %lld held", doubles,
                p->red_cap);
}

static int last_launch(const char* what)
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : hip_fail(e, what);
}

}  // namespace ctprj

using namespace ctprj;

struct ctprj_plan : public Plan {};

extern "C" int ctprj_poly_iout(int basis, const double* knots, int E, const double* Ie,
                               const float* d_s, long long count, float* d_BL, float* d_sBL,
                               float* d_ssBL, void* stream)
{
    BhcParams P;
    int rc = make_bhc_params(knots, E, Ie, basis, CTPRJ_NOISE_POISSON, &P);
    if (rc) return rc;
    if (count < 0 || (count > 0 && !d_s)) return fail(CTPRJ_EINVAL, "ctprj_poly_iout: bad rays");
    const bool matrix = (Ie == nullptr);
    if (matrix && !d_BL) return fail(CTPRJ_EINVAL, "ctprj_poly_iout: matrix mode needs BL");
    launch_poly_iout(P, d_s, count, d_BL, matrix ? nullptr : d_sBL, matrix ? nullptr : d_ssBL,
                     matrix, (hipStream_t)stream);
    return last_launch("poly_iout launch");
}

extern "C" int ctprj_b1_iout(const double* knots, int E, const double* Ie, const float* d_s,
                             long long count, float* d_BL, float* d_sBL, float* d_ssBL,
                             void* stream)
{
    return ctprj_poly_iout(CTPRJ_BASIS_B1, knots, E, Ie, d_s, count, d_BL, d_sBL, d_ssBL, stream);
}

extern "C" int ctprj_plan_set_bhc_model(ctprj_plan* p, int basis, int noise)
{
    if (!p) return fail(CTPRJ_EINVAL, "plan is NULL");
    if (basis != CTPRJ_BASIS_B1 && basis != CTPRJ_BASIS_B0 && basis != CTPRJ_BASIS_DIS)
        return fail(CTPRJ_EINVAL, "unknown spectrum basis %d", basis);
    if (noise != CTPRJ_NOISE_POISSON && noise != CTPRJ_NOISE_GAUSSIAN)
        return fail(CTPRJ_EINVAL, "unknown noise model %d", noise);
    p->bhc_basis = basis;
    p->bhc_noise = noise;
    return 0;
}

extern "C" int ctprj_plan_get_bhc_model(const ctprj_plan* p, int* basis, int* noise)
{
    if (!p) return fail(CTPRJ_EINVAL, "plan is NULL");
    if (basis) *basis = p->bhc_basis;
    if (noise) *noise = p->bhc_noise;
    return 0;
}

extern "C" int ctprj_bhc_grad_alpha(ctprj_plan* p, const float* d_alpha, const float* d_Imea,
                                    const double* knots, int E, const double* Ie, float Ts,
                                    double* d_f, float* d_grad, float* d_hw, void* stream)
{
    if (!p || !d_alpha || !d_Imea || !d_f) return fail(CTPRJ_EINVAL, "bhc_grad_alpha: NULL argument");
    if (!Ie) return fail(CTPRJ_EINVAL, "bhc_grad_alpha: Ie is NULL");
    BhcParams P;
    int rc = make_params(p, knots, E, Ie, &P);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const long long rays = p->sino_elems, blocks = bhc_blocks(rays);
    if ((rc = check_red(p, blocks))) return rc;
    // all slices on the symmetric kernels and no Hessian weights: the ray pass runs in the
    // forward projector's final store (one NLL partial per forward work item)
    const long long fused = d_hw ? -1 : p->bhc_fused_parts();
    if (fused >= 0 && fused <= p->red_cap) {
        if ((rc = p->launch_forward_bhc(d_alpha, p->d_sino_ws, BhcStore{P, d_Imea, p->d_red, Ts, 0},
                                        st)))
            return rc;
        // the NLL partials are added by the adjoint's first wave, or (cost only) on their own
        const RedJob job{p->d_red, fused, d_f};
        if (d_grad) return p->launch_adjoint(p->d_sino_ws, d_grad, st, false, job);
        launch_wave_sum(job, st);
        return last_launch("bhc sum launch");
    }
    // s = A alpha  (Ts applied inside the ray pass)
    if ((rc = p->launch_forward(d_alpha, p->d_sino_ws, st, false))) return rc;
    float* w = d_grad ? p->d_sino_ws : nullptr;
    launch_bhc_alpha_ray(P, p->d_sino_ws, Ts, d_Imea, rays, Ts, w, d_hw, p->d_red, st);
    if ((rc = last_launch("bhc alpha ray launch"))) return rc;
    launch_reduce_partials(p->d_red, blocks, 1, d_f, st);
    if ((rc = last_launch("bhc reduce launch"))) return rc;
    if (d_grad) return p->launch_adjoint(p->d_sino_ws, d_grad, st, false);
    return 0;
}

extern "C" int ctprj_bhc_alpha_rays(ctprj_plan* p, const float* d_sraw, const float* d_Imea,
                                    const double* knots, int E, const double* Ie, float Ts,
                                    double* d_f, float* d_w, float* d_hw, void* stream)
{
    if (!p || !d_sraw || !d_Imea || !d_f) return fail(CTPRJ_EINVAL, "bhc_alpha_rays: NULL argument");
    if (!Ie) return fail(CTPRJ_EINVAL, "bhc_alpha_rays: Ie is NULL");
    BhcParams P;
    int rc = make_params(p, knots, E, Ie, &P);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const long long rays = p->sino_elems, blocks = bhc_blocks(rays);
    if ((rc = check_red(p, blocks))) return rc;
    launch_bhc_alpha_ray(P, d_sraw, Ts, d_Imea, rays, Ts, d_w, d_hw, p->d_red, st);
    if ((rc = last_launch("bhc alpha ray launch"))) return rc;
    launch_reduce_partials(p->d_red, blocks, 1, d_f, st);
    return last_launch("bhc reduce launch");
}

// ------------------------------------------------ the entry point that allocates nothing ---

namespace ctprj {

// Reduction partials of ctprj_bhc_grad_alpha_noalloc in a file-scope device array: one
// double per group of kBhcBlock rays, then the reduced value.  Each device holds its own
// copy until the process ends; every call on a device shares that device's copy.
constexpr long long kNoallocParts = 1LL << 16;  // groups per call: 16.7 M rays at most
static __device__ __align__(256) double scratch[kNoallocParts + 1];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices];  // device address of `scratch`, per device

// The current device's address of `scratch`, looked up once per device with that device
// current.  Host code never passes `scratch` itself: its name is a host-side stand-in.
static int noalloc_scratch(double** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(CTPRJ_EINVAL, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = static_cast<double*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace ctprj

extern "C" int ctprj_bhc_grad_alpha_noalloc(ctprj_plan* p, const float* d_alpha,
                                            const float* d_Imea, const double* knots, int E,
                                            const double* Ie, float Ts, double* d_f,
                                            float* d_grad, float* d_hw, void* stream)
{
    if (!p || !d_alpha || !d_Imea || !d_f)
        return fail(CTPRJ_EINVAL, "bhc_grad_alpha_noalloc: NULL argument");
    if (!Ie) return fail(CTPRJ_EINVAL, "bhc_grad_alpha_noalloc: Ie is NULL");
    BhcParams P;
    int rc = make_params(p, knots, E, Ie, &P);
    if (rc) return rc;
    const long long rays = p->sino_elems, blocks = bhc_blocks(rays);
    if (blocks > kNoallocParts)
        return fail(CTPRJ_EINVAL,
                    "bhc_grad_alpha_noalloc: %lld rays exceed the static workspace's %lld "
                    "(use ctprj_bhc_grad_alpha)", rays, kNoallocParts * kBhcBlock);
    double* part = nullptr;
    if ((rc = noalloc_scratch(&part))) return rc;
    hipStream_t st = (hipStream_t)stream;
    // the same launches as ctprj_bhc_grad_alpha (with the fused store when its partials fit
    // the static array, as they fit the plan's workspace there); every partial is written
    // before the reduction reads it, so nothing needs clearing
    const long long fused = d_hw ? -1 : p->bhc_fused_parts();
    if (fused >= 0 && fused <= kNoallocParts && fused <= p->red_cap) {
        if ((rc = p->launch_forward_bhc(d_alpha, p->d_sino_ws, BhcStore{P, d_Imea, part, Ts, 0},
                                        st)))
            return rc;
        const RedJob job{part, fused, d_f};
        if (d_grad) return p->launch_adjoint(p->d_sino_ws, d_grad, st, false, job);
        launch_wave_sum(job, st);
        return last_launch("bhc sum launch");
    }
    if ((rc = p->launch_forward(d_alpha, p->d_sino_ws, st, false))) return rc;
    float* w = d_grad ? p->d_sino_ws : nullptr;
    launch_bhc_alpha_ray(P, p->d_sino_ws, Ts, d_Imea, rays, Ts, w, d_hw, part, st);
    if ((rc = last_launch("bhc alpha ray launch"))) return rc;
    launch_reduce_partials(part, blocks, 1, d_f, st);
    if ((rc = last_launch("bhc reduce launch"))) return rc;
    if (d_grad) return p->launch_adjoint(p->d_sino_ws, d_grad, st, false);
    return 0;
}

extern "C" int ctprj_bhc_grad_ie(ctprj_plan* p, const float* d_alpha, const float* d_s,
                                 const float* d_Imea, const double* knots, int E,
                                 const double* Ie, float Ts, double* d_f, double* d_g,
                                 double* d_H, void* stream)
{
    if (!p || !d_Imea) return fail(CTPRJ_EINVAL, "bhc_grad_ie: NULL argument");
    if (!d_alpha && !d_s) return fail(CTPRJ_EINVAL, "bhc_grad_ie: need alpha or s");
    if (!Ie) return fail(CTPRJ_EINVAL, "bhc_grad_ie: Ie is NULL");
    BhcParams P;
    int rc = make_params(p, knots, E, Ie, &P);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const long long rays = p->sino_elems, groups = bhc_ie_groups(rays);
    const int nv = bhc_ie_nvals(E);
    if ((rc = check_red(p, groups * nv))) return rc;
    const float* s = d_s;
    if (!s) {
        if ((rc = p->launch_forward(d_alpha, p->d_sino_ws, st, false))) return rc;
        s = p->d_sino_ws;
    }
    launch_bhc_ie_ray(P, s, Ts, d_Imea, rays, p->d_red, st);
    if ((rc = last_launch("bhc ie ray launch"))) return rc;
    launch_reduce_ie(p->d_red, groups, E, d_f, d_g, d_H, st);
    return last_launch("bhc ie reduce launch");
}

// ------------------------------------------------------------- I-epsilon solve ---

namespace ctprj {

// Cholesky solve of the SPD system M x = r (n <= CTPRJ_MAX_E, FP64, in place on copies).
// Returns false when M is not numerically positive definite.
static bool chol_solve(std::vector<double> M, int n, std::vector<double> r, std::vector<double>* x)
{
    for (int j = 0; j < n; ++j) {
        double d = M[j * n + j];
        for (int k = 0; k < j; ++k) d -= M[j * n + k] * M[j * n + k];
        if (!(d > 0.0)) return false;
        const double l = std::sqrt(d);
        M[j * n + j] = l;
        for (int i = j + 1; i < n; ++i) {
            double v = M[i * n + j];
            for (int k = 0; k < j; ++k) v -= M[i * n + k] * M[j * n + k];
            M[i * n + j] = v / l;
        }
    }
    for (int i = 0; i < n; ++i) {
        double v = r[i];
        for (int k = 0; k < i; ++k) v -= M[i * n + k] * r[k];
        r[i] = v / M[i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double v = r[i];
        for (int k = i + 1; k < n; ++k) v -= M[k * n + i] * r[k];
        r[i] = v / M[i * n + i];
    }
    *x = r;
    return true;
}

}  // namespace ctprj

extern "C" int ctprj_ie_solve(ctprj_plan* p, const float* d_alpha, const float* d_sraw,
                              const float* d_Imea, const double* knots, int E, double* Ie,
                              float Ts, int maxit, double tol, ctprj_ie_info* info, void* stream)
{
    if (!p || !d_Imea || !Ie) return fail(CTPRJ_EINVAL, "ie_solve: NULL argument");
    if (!d_alpha && !d_sraw) return fail(CTPRJ_EINVAL, "ie_solve: need alpha or s");
    if (maxit < 1) return fail(CTPRJ_EINVAL, "ie_solve: maxit must be >= 1");
    BhcParams P;
    int rc = make_params(p, knots, E, Ie, &P);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const long long rays = p->sino_elems, blocks = bhc_blocks(rays), groups = bhc_ie_groups(rays);
    const int nv = bhc_ie_nvals(E);
    if ((rc = check_red(p, std::max(blocks, groups * nv)))) return rc;
    const float* s = d_sraw;
    if (!s) {  // the projection stays fixed over the I-epsilon step
        if ((rc = p->launch_forward(d_alpha, p->d_sino_ws, st, false))) return rc;
        s = p->d_sino_ws;
    }
    // the basis matrix of the fixed projection, formed once per call (as IeStep forms
    // A = polyIout(Phi(alpha), []) once, bhc/beamhardenSpline.m:447) in a buffer the plan
    // keeps for later calls, when it fits kIeBasisCacheMax; otherwise every pass
    // evaluates the basis
    const size_t phi_bytes = sizeof(float) * (size_t)rays * (size_t)E;
    const float* phi = nullptr;
    if (phi_bytes <= kIeBasisCacheMax) {
        if (p->ie_phi_cap < phi_bytes) {
            if (p->d_ie_phi) (void)hipFree(p->d_ie_phi);
            p->d_ie_phi = nullptr;
            p->ie_phi_cap = 0;
            if (hipMalloc((void**)&p->d_ie_phi, phi_bytes) == hipSuccess)
                p->ie_phi_cap = phi_bytes;
            else
                (void)hipGetLastError();  // a failed allocation is not the passes' error
        }
        if (p->d_ie_phi) {
            launch_bhc_ie_basis(P, s, Ts, rays, p->d_ie_phi, st);
            if ((rc = last_launch("ie basis launch"))) return rc;
            phi = p->d_ie_phi;
        }
    }
    std::vector<double> x(E), red(nv);
    for (int j = 0; j < E; ++j) x[j] = std::max(Ie[j], 0.0);
    int passes = 0;
    auto set_ie = [&](const std::vector<double>& v) {
        for (int j = 0; j < E; ++j) {
            P.Ie[j] = (float)v[j];
            P.Ied[j] = v[j];
        }
    };
    // f, g, H at x (FP64 ray sums on the device)
    auto full = [&](double* f, std::vector<double>* g, std::vector<double>* H) -> int {
        set_ie(x);
        launch_bhc_ie_ray(P, s, Ts, d_Imea, rays, p->d_red, st, phi);
        launch_reduce_partials(p->d_red, groups, nv, p->d_red_out, st);
        hipError_t e = hipMemcpyAsync(red.data(), p->d_red_out, sizeof(double) * nv,
                                      hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e, "ie_solve pass");
        ++passes;
        const int nh = E * (E + 1) / 2;
        g->assign(red.begin() + nh, red.begin() + nh + E);
        H->assign((size_t)E * E, 0.0);
        for (int v = 0, j = 0, k = 0; v < nh; ++v) {
            (*H)[j * E + k] = (*H)[k * E + j] = red[v];
            if (++k == E) k = ++j;
        }
        *f = red[nh + E];
        return 0;
    };
    auto cost = [&](const std::vector<double>& v, double* f) -> int {
        set_ie(v);
        launch_bhc_ie_cost(P, s, Ts, d_Imea, rays, p->d_red, st, phi);
        launch_reduce_partials(p->d_red, blocks, 1, p->d_red_out, st);
        hipError_t e = hipMemcpyAsync(f, p->d_red_out, sizeof(double), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(e, "ie_solve cost pass");
        ++passes;
        return 0;
    };
    double f = 0.0, dec = 0.0, lam = 1e-3;
    std::vector<double> g, H;
    int it = 0, converged = 0;
    if ((rc = full(&f, &g, &H))) return rc;
    for (it = 0; it < maxit; ++it) {
        // two-metric projected Newton (Bertsekas) with a Levenberg-Marquardt shift:
        // eps-active bounds take a diagonally scaled step, the free block a damped Newton
        // step; c_j = max(|H_jj|, 1e-6 max|H_kk|) keeps the scale positive where H is
        // indefinite (gaussLI's Hessian where Err > 1)
        double w2 = 0.0, xmax = 0.0, hmax = 0.0;
        for (int j = 0; j < E; ++j) {
            const double pg = x[j] - std::max(x[j] - g[j], 0.0);
            w2 += pg * pg;
            xmax = std::max(xmax, x[j]);
            hmax = std::max(hmax, std::fabs(H[j * E + j]));
        }
        const double eps = std::min(1e-2 * xmax, std::sqrt(w2));
        std::vector<double> c(E), d(E, 0.0);
        std::vector<int> F, A;
        for (int j = 0; j < E; ++j) {
            c[j] = std::max(std::fabs(H[j * E + j]), 1e-6 * std::max(hmax, 1e-300));
            if (x[j] <= eps && g[j] > 0.0) {
                A.push_back(j);
                d[j] = g[j] / c[j];
            } else {
                F.push_back(j);
            }
        }
        const int nf = (int)F.size();
        if (nf > 0) {
            std::vector<double> M((size_t)nf * nf), r(nf), sol;
            for (int a = 0; a < nf; ++a) {
                r[a] = g[F[a]];
                for (int b = 0; b < nf; ++b) M[a * nf + b] = H[F[a] * E + F[b]];
            }
            for (;;) {
                std::vector<double> Ml = M;
                for (int a = 0; a < nf; ++a) Ml[a * nf + a] += lam * c[F[a]];
                if (chol_solve(Ml, nf, r, &sol) || lam > 1e30) break;
                lam *= 10.0;
            }
            if (sol.size() == (size_t)nf)
                for (int a = 0; a < nf; ++a) d[F[a]] = sol[a];
        }
        double gdF = 0.0;
        for (int j : F) gdF += g[j] * d[j];
        dec = gdF;
        for (int j : A) dec += g[j] * std::min(x[j], d[j]);
        if (!(dec > tol * std::max(std::fabs(f), 1.0))) {
            converged = 1;
            break;
        }
        // backtracking on the projected path x(t) = max(x - t d, 0), halving, with
        // f(x(t)) <= f - 1e-4 (t g_F.d_F + g_A.(x_A - x_A(t)))
        double t = 1.0, fn = 0.0;
        bool ok = false;
        std::vector<double> xn(E);
        for (int trial = 0; trial < 30 && !ok; ++trial) {
            double ga = 0.0;
            for (int j = 0; j < E; ++j) xn[j] = std::max(x[j] - t * d[j], 0.0);
            for (int j : A) ga += g[j] * (x[j] - xn[j]);
            if ((rc = cost(xn, &fn))) return rc;
            ok = fn <= f - 1e-4 * (t * gdF + ga);
            if (!ok) t *= 0.5;
        }
        if (!ok) {  // stiffen the damping and retry from the same point
            lam *= 10.0;
            if (lam > 1e12) {
                converged = 1;
                break;
            }
            continue;
        }
        lam = (t == 1.0) ? std::max(lam / 3.0, 1e-12) : lam * 4.0;
        x = xn;
        if ((rc = full(&f, &g, &H))) return rc;
    }
    for (int j = 0; j < E; ++j) Ie[j] = x[j];
    if (info) {
        info->cost = f;
        info->decrement = dec;
        info->iterations = it;
        info->passes = passes;
        info->converged = converged;
    }
    return 0;
}
