// Host entry points of the beam-hardening model (bhc/poissLAlpha.m, bhc/poissLI.m,
// bhc/gaussLAlpha.m, bhc/gaussLI.m, bhc/Spline.m polyIout) on top of a projector plan.
#include <hip/hip_runtime.h>

#include <cmath>

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

static int ensure_red(Plan* p, long long doubles)
{
    if (p->red_cap >= doubles) return 0;
    if (p->d_red) (void)hipFree(p->d_red);
    p->d_red = nullptr;
    p->red_cap = 0;
    hipError_t e = hipMalloc((void**)&p->d_red, sizeof(double) * (size_t)doubles);
    if (e != hipSuccess) return fail(CTPRJ_ENOMEM, "reduction workspace: %s", hipGetErrorString(e));
    p->red_cap = doubles;
    if (!p->d_red_out) {
        e = hipMalloc((void**)&p->d_red_out, sizeof(double) * (size_t)bhc_ie_nvals(CTPRJ_MAX_E));
        if (e != hipSuccess) return fail(CTPRJ_ENOMEM, "reduction output: %s", hipGetErrorString(e));
    }
    return 0;
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
    if ((rc = ensure_red(p, blocks))) return rc;
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
    if ((rc = ensure_red(p, blocks))) return rc;
    launch_bhc_alpha_ray(P, d_sraw, Ts, d_Imea, rays, Ts, d_w, d_hw, p->d_red, st);
    if ((rc = last_launch("bhc alpha ray launch"))) return rc;
    launch_reduce_partials(p->d_red, blocks, 1, d_f, st);
    return last_launch("bhc reduce launch");
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
    const long long rays = p->sino_elems, blocks = bhc_blocks(rays);
    const int nv = bhc_ie_nvals(E);
    if ((rc = ensure_red(p, blocks * nv))) return rc;
    const float* s = d_s;
    if (!s) {
        if ((rc = p->launch_forward(d_alpha, p->d_sino_ws, st, false))) return rc;
        s = p->d_sino_ws;
    }
    launch_bhc_ie_ray(P, s, Ts, d_Imea, rays, p->d_red, st);
    if ((rc = last_launch("bhc ie ray launch"))) return rc;
    launch_reduce_partials(p->d_red, blocks, nv, p->d_red_out, st);
    if ((rc = last_launch("bhc ie reduce launch"))) return rc;
    launch_unpack_ie(p->d_red_out, E, d_f, d_g, d_H, st);
    return last_launch("bhc ie unpack launch");
}
