/*
 * libctprj -- MI355X (gfx950) X-ray CT projector, backprojector and polychromatic
 * Poisson beam-hardening likelihood.  Plain C ABI: pointers, sizes and int status codes.
 *
 * Conventions follow the reference projector (prj/prj.h:8-57, prj/cpuPrj.c):
 *   image  n x n float, row-major, pixel (x, y) at [(y + n/2) * n + x + n/2],
 *          x, y in [-(n-1)/2, (n-1)/2] (integer division); for even n, row 0 and
 *          column 0 are never read by the forward and are written as 0 by the adjoint;
 *   sinogram np x prjWidth float, view-major; view a at theta = a*2*PI/prjFull with
 *          PI = 3.14159265359f; bin b at t = (b - prjWidth/2) * dSize; for an even
 *          prjWidth, bin 0 is written as 0 by the forward and never read by the adjoint;
 *   beam   width dSize*effectiveRate; outputs are divided by effectiveRate;
 *          d > 0: fan beam, source at distance d from the rotation centre, virtual flat
 *          detector through the centre; d <= 0: parallel beam.
 *
 * All entry points return 0 on success or a negative CTPRJ_E* code; ctprj_last_error()
 * returns the message of the last failure on the calling thread.
 */
#ifndef CTPRJ_H
#define CTPRJ_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* command bits, prj/prj.h:30-32 */
#define CTPRJ_FWD_BIT (1 << 1)
#define CTPRJ_BWD_BIT (1 << 2)
#define CTPRJ_FBP_BIT (1 << 3)

/* status codes */
#define CTPRJ_OK 0
#define CTPRJ_EINVAL (-1)   /* bad argument / geometry */
#define CTPRJ_ENODEV (-2)   /* no usable gfx950 device */
#define CTPRJ_ENOMEM (-3)   /* device or host allocation failed */
#define CTPRJ_EHIP (-4)     /* HIP runtime error */
#define CTPRJ_ESTATE (-5)   /* call out of order (e.g. prj_run before prj_setup) */

/* Maximum spectrum basis size accepted by the beam-hardening entry points. */
#define CTPRJ_MAX_E 32

/* Geometry of one slice: the fields of struct prjConf, prj/prj.h:42-57. */
typedef struct ctprj_geom {
    int n;               /* image is n x n */
    int prjWidth;        /* detector bins per view */
    int np;              /* views computed */
    int prjFull;         /* views per 360 degrees */
    float dSize;         /* detector pitch, pixels */
    float effectiveRate; /* beam width / pitch */
    float d;             /* source-to-centre distance (pixels); <= 0 selects parallel */
} ctprj_geom;

typedef struct ctprj_plan ctprj_plan;

/* ------------------------------------------------------------------------------
 * Reference-compatible single-slice interface (host pointers, synchronous).
 * ------------------------------------------------------------------------------ */

/* Replaces setup() of prj/prj.h:60 / prj/cpuPrj.c:618-653 (also prj/gpuPrj.cu:794-865).
 * Configures the process-wide slice geometry used by prj_run. */
int prj_setup(int n, int prjWidth, int np, int prjFull, float dSize, float effectiveRate,
              float d);

/* Replaces showSetup() of prj/prj.h:61 / prj/cpuPrj.c:655-663: prints the geometry. */
int prj_show(void);

/* Replaces cpuPrj()/gpuPrj() of prj/prj.h:63-65 (prj/cpuPrj.c:679-806,
 * prj/gpuPrj.cu:887-1047).  cmd is CTPRJ_FWD_BIT (img -> sino), CTPRJ_BWD_BIT
 * (sino -> img) or CTPRJ_FBP_BIT (sino -> img).  Host pointers; copies to and from the
 * device and returns when the result is in host memory. */
int prj_run(float* img, float* sino, int cmd);

/* Reads back the geometry configured by prj_setup (the MEX 'config' echo,
 * prj/mPrj.c:160-170). */
int prj_get_geom(ctprj_geom* out);

/* ------------------------------------------------------------------------------
 * Plan (handle) interface: device pointers, caller's HIP stream, many slices per call.
 * No reference counterpart as one call; each slice behaves as cpuPrj on that slice.
 * ------------------------------------------------------------------------------ */

/* Plan for `nslices` independent slices (ragged sizes and mixed parallel/fan allowed).
 * Images of all slices are packed back to back in one float array, likewise sinograms;
 * see ctprj_plan_layout.  Allocates the plan's device tables and workspace on the
 * current HIP device, including the FP64 reduction partials of the beam-hardening calls
 * (ctprj_bhc_grad_alpha, _alpha_rays, _grad_ie, ctprj_ie_solve), which therefore never
 * allocate and can be captured in a HIP graph from their first call.  The calls of one
 * plan share its workspace (the transposed image, the filtered sinogram, partial sums):
 * calls on one plan that overlap on different streams race; use one stream per plan, or
 * one plan per stream. */
int ctprj_plan_create(const ctprj_geom* geoms, int nslices, ctprj_plan** plan);
int ctprj_plan_destroy(ctprj_plan* plan);

/* Total packed sizes (elements) and, when the arrays are non-NULL, per-slice element
 * offsets (nslices entries each). */
int ctprj_plan_layout(const ctprj_plan* plan, long long* img_elems, long long* sino_elems,
                      long long* img_offsets, long long* sino_offsets);

/* Number of non-zeros of the system matrix A of every slice, summed (the voxel-ray
 * updates of one application), computed on the device. */
int ctprj_plan_nnz(ctprj_plan* plan, long long* nnz, void* stream);

/* Per slice, the kernels it runs on: 0 general; symmetric, with the symmetry group of
 * its view set: 1 = 8-fold (prjFull % 4 == 0), 2 = 4-fold (prjFull even),
 * 3 = 2-fold (theta -> -theta, any prjFull).  A slice takes the first group that the
 * tune flags allow and whose symmetric evaluations land at least half on computed views.
 * `modes` holds nslices ints. */
int ctprj_plan_modes(const ctprj_plan* plan, int* modes);

/* Kernel shape knobs (the autotuner's search space). */
typedef struct ctprj_tune {
    int fwd_block;   /* general forward work-group: 64, 128 or 256 threads, one ray each */
    int adj_tile;    /* general adjoint pixel tile edge: 8 or 16 */
    int flags;       /* bit0: use the symmetric kernels where the geometry allows;
                        bit1: not the 8-fold group; bit2: not the 4-fold group;
                        bit3: work order keeps small slices whole on one XCD (else every
                        slice is spread over all XCDs);
                        bit4: the symmetric adjoint reads 3-bin windows (else 4);
                        bit5: the symmetric adjoint's waves stage the bins of each view
                        and element in LDS (else every lane loads its own windows) */
    int adj_seg;     /* symmetric adjoint: waves per 8x8 tile of orbit base pixels
                        splitting the views: 1, 2, 4 or 8 */
    int fwd_seg;     /* symmetric forward: waves per 64 canonical rays splitting the
                        image rows: 1, 2, 4, 8 or 16 */
    int adj_split;   /* symmetric adjoint: work-groups per tile splitting the views, 1-8;
                        above 1 their partial sums are added in a fixed order by a second
                        launch (more work-groups for small slices) */
    int fwd_split;   /* symmetric forward: work-groups per 64 canonical rays splitting the
                        image rows, 1-8, the same way */
} ctprj_tune;
int ctprj_plan_get_tune(const ctprj_plan* plan, ctprj_tune* tune);
/* Out-of-range fields are refused (CTPRJ_EINVAL) and change nothing.  Splits whose
 * partial sums would exceed the plan's 256 MiB are refused too; the other fields then
 * stay applied and both splits fall back to 1.  A change that fails part way (a device
 * allocation) changes nothing: the plan keeps its previous tables and shapes. */
int ctprj_plan_set_tune(ctprj_plan* plan, const ctprj_tune* tune);
/* Times every candidate shape on this plan's geometry (HIP events, `reps` runs each),
 * keeps the fastest and, if cache_path is non-NULL, records it in that text file keyed
 * by (geometry, device arch). */
int ctprj_plan_autotune(ctprj_plan* plan, int reps, const char* cache_path, void* stream);

/* sino = A img  (prj/cpuPrj.c:319-616 semantics), all slices in one launch. */
int ctprj_forward(ctprj_plan* plan, const float* d_img, float* d_sino, void* stream);
/* img = A^T sino (prj/cpuPrj.c:46-317 semantics), all slices in one launch. */
int ctprj_adjoint(ctprj_plan* plan, const float* d_sino, float* d_img, void* stream);
/* FBP command (prj/cpuPrj.c:718-784): fan pre-weight, ramp filter, weighted adjoint. */
int ctprj_fbp(ctprj_plan* plan, const float* d_sino, float* d_img, void* stream);

/* ------------------------------------------------------------------------------
 * Polychromatic beam-hardening model (bhc/poissLAlpha.m, bhc/poissLI.m,
 * bhc/gaussLAlpha.m, bhc/gaussLI.m, bhc/Spline.m polyIout).  Ie holds the E basis
 * coefficients of the mass-attenuation spectrum; `knots` the basis nodes:
 *   CTPRJ_BASIS_B1  (Spline 'b1', hats):  E+2 knots [kappa_1/q, kappa_1..kappa_E, kappa_E*q]
 *                   (bhc/beamhardenSpline.m:164-165);
 *   CTPRJ_BASIS_B0  (Spline 'b0', boxes): E+1 knots, box j = [k_j, k_{j+1}];
 *   CTPRJ_BASIS_DIS (Spline 'dis'):       E+2 knots, node j = k_{j+1} with weight
 *                   (k_{j+2} - k_j)/2 (bhc/Spline.m:271-272).
 * Ts is the operator scale (Phi = Ts*A, tomography/ConfigCT.m:118-119).  Host arrays for
 * knots/Ie (E <= CTPRJ_MAX_E); device arrays for images/sinograms.  NLL and I-epsilon
 * reductions are FP64; projections are FP32.
 * ------------------------------------------------------------------------------ */
enum { CTPRJ_BASIS_B1 = 0, CTPRJ_BASIS_B0 = 1, CTPRJ_BASIS_DIS = 2 };
enum { CTPRJ_NOISE_POISSON = 0, CTPRJ_NOISE_GAUSSIAN = 1 };

/* Spectrum sum of bhc/Spline.m (b1Iout :122-203, b0Iout :205-262, disIout :264-302) on
 * device: for every ray value s[i], BL[i] = sum_j Ie[j] phi_j(s[i]), sBL = d/ds BL,
 * ssBL = d2/ds2 BL.  Any of the outputs may be NULL.  With Ie == NULL the basis matrix
 * phi_j(s[i]) is written to BL as count x E (row-major) instead. */
int ctprj_poly_iout(int basis, const double* knots, int E, const double* Ie,
                    const float* d_s, long long count, float* d_BL, float* d_sBL,
                    float* d_ssBL, void* stream);
/* ctprj_poly_iout with CTPRJ_BASIS_B1 (bhc/Spline.m:122-203). */
int ctprj_b1_iout(const double* knots, int E, const double* Ie, const float* d_s,
                  long long count, float* d_BL, float* d_sBL, float* d_ssBL, void* stream);

/* Spectrum basis and noise model of the plan's beam-hardening calls below
 * (bhc/beamhardenSpline.m:60 opt.spectBasis, :155-162 opt.noiseType); a new plan starts
 * at CTPRJ_BASIS_B1 / CTPRJ_NOISE_POISSON.  Gaussian: Err = log(BL/Imea), f = Err'Err/2
 * (bhc/gaussLAlpha.m, bhc/gaussLI.m). */
int ctprj_plan_set_bhc_model(ctprj_plan* plan, int basis, int noise);
int ctprj_plan_get_bhc_model(const ctprj_plan* plan, int* basis, int* noise);

/* [f, g] = poissLAlpha(Imea, Ie, alpha, Phi, Phit, polyIout) (bhc/poissLAlpha.m:1-13):
 * s = Ts*A*alpha, f = sum(BL - Imea) - sum Imea*log(BL/Imea) (FP64, 0*log(.) = 0 for
 * Imea = 0), grad = Ts*A^T[(1 - Imea/BL) * sBL].  d_f: one device double.
 * d_grad may be NULL (cost only: one forward projection).  d_hw, if non-NULL, receives
 * the Hessian ray weight (sBL/BL)^2*Imea + ssBL*(1 - Imea/BL) (bhc/poissLAlpha.m:15).
 * Under CTPRJ_NOISE_GAUSSIAN this is gaussLAlpha (bhc/gaussLAlpha.m:1-21): f = Err'Err/2,
 * grad = Ts*A^T[Err * sBL/BL], Hessian weight (1 - Err)(sBL/BL)^2 + Err*ssBL/BL. */
int ctprj_bhc_grad_alpha(ctprj_plan* plan, const float* d_alpha, const float* d_Imea,
                         const double* knots, int E, const double* Ie, float Ts,
                         double* d_f, float* d_grad, float* d_hw, void* stream);

/* ctprj_bhc_grad_alpha for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Same arguments, results and launches as ctprj_bhc_grad_alpha, but the FP64
 * reduction partials live in a static device array of the library (512 KiB per device, held until the process ends) instead of the plan's
 * reduction workspace, so the call allocates nothing and leaves the plan's workspace
 * untouched; the plan's other buffers (created with the plan) serve as before.  At most 16,777,216 rays (np*prjWidth summed over the
 * plan's slices) per call: larger plans get CTPRJ_EINVAL.  Every call on one device
 * shares that array, so calls that overlap on one device (other streams or threads)
 * race and must be serialised by the caller. */
int ctprj_bhc_grad_alpha_noalloc(ctprj_plan* plan, const float* d_alpha, const float* d_Imea,
                                 const double* knots, int E, const double* Ie, float Ts,
                                 double* d_f, float* d_grad, float* d_hw, void* stream);

/* The per-ray part of poissLAlpha on a given projection: with s = Ts * d_sraw
 * (d_sraw = A*alpha, np*prjWidth per slice, packed), writes d_f (one device double) and,
 * when non-NULL, d_w = Ts*(1 - Imea/BL)*sBL (may alias d_sraw) and d_hw (Hessian ray
 * weight).  grad = A^T d_w is then ctprj_adjoint(d_w).  Lets a caller keep only the
 * projection between a cost and a gradient call: the spectral exponentials are
 * recomputed, never stored. */
int ctprj_bhc_alpha_rays(ctprj_plan* plan, const float* d_sraw, const float* d_Imea,
                         const double* knots, int E, const double* Ie, float Ts, double* d_f,
                         float* d_w, float* d_hw, void* stream);

/* [f, g, H] = poissLI(Imea, A, Ie) with A = polyIout(Phi(alpha), []) (bhc/poissLI.m,
 * bhc/beamhardenSpline.m:447): Ir = A*Ie, f as above, g = A^T(1 - Imea/Ir) (E doubles),
 * H = A^T diag(Imea/Ir^2) A (E x E doubles, row-major) -- FP64 sums, written to device
 * memory; the ray x basis matrix A is never formed.  Ray values are s = Ts * d_sraw with
 * d_sraw = A*alpha (packed); pass d_sraw = NULL together with d_alpha to project inside.
 * d_f / d_g / d_H may be NULL.  Under CTPRJ_NOISE_GAUSSIAN this is gaussLI
 * (bhc/gaussLI.m): f = Err'Err/2, g = A^T(Err/Ir), H = A^T diag((1 - Err)/Ir^2) A. */
int ctprj_bhc_grad_ie(ctprj_plan* plan, const float* d_alpha, const float* d_sraw,
                      const float* d_Imea, const double* knots, int E, const double* Ie,
                      float Ts, double* d_f, double* d_g, double* d_H, void* stream);

/* I-epsilon step: minimises the objective of ctprj_bhc_grad_ie over Ie >= 0 for a fixed
 * projection (the IeStep of bhc/beamhardenSpline.m:446-451; bhc/ActiveSet.m and
 * bhc/LBFGSB.m there).  Two-metric projected Newton with a Levenberg-Marquardt shift:
 * f, g, H are FP64 sums over all rays on the device (one pass per iterate, one
 * cost-only pass per backtracking trial); the E x E damped Cholesky solve, the
 * eps-active set and the backtracking (halving, sufficient decrease 1e-4) run on the
 * host in FP64.  The spectrum basis matrix is a discretised Laplace transform (condition
 * numbers near 1e18 for E = 20), so like the reference's capped IeStep a call is a fixed
 * budget of iterations.  Ie (host, E doubles) is the start point on entry and the
 * iterate on return.  Stops when the projected Newton decrement is at most
 * tol*max(|f|, 1), after maxit iterations, or when the damping exceeds 1e12.  Blocks
 * until done (one small device-to-host copy per pass).  Like IeStep, which forms
 * A = polyIout(Phi(alpha), []) once (bhc/beamhardenSpline.m:447), a call evaluates the
 * basis values of its projection once (rays x E floats) and its passes read them; the
 * plan keeps that buffer for later calls when it is at most 1 GiB (otherwise every pass
 * evaluates the basis; the results are the same either way). */
typedef struct ctprj_ie_info {
    double cost;       /* objective at the returned Ie */
    double decrement;  /* last Newton decrement g.d */
    int iterations;    /* Newton iterations taken */
    int passes;        /* ray passes (full and cost-only) */
    int converged;     /* 1 if the decrement or no-progress test stopped it */
} ctprj_ie_info;
int ctprj_ie_solve(ctprj_plan* plan, const float* d_alpha, const float* d_sraw,
                   const float* d_Imea, const double* knots, int E, double* Ie, float Ts,
                   int maxit, double tol, ctprj_ie_info* info, void* stream);

/* ------------------------------------------------------------------------------
 * Polychromatic measurement simulator (bhc/genBeamHarden.m:53-79).  For K materials
 * with projections y_k = density_k * (A x_k), ray t measures
 *   I(t) = sum_l exp(-Ts * sum_k y_k(t) kappa[k][l]) * iota[l] * dEps[l]
 * with dEps the widths of the energy bins [eps_1, midpoints, eps_L] (:53-54), in FP64
 * in the reference's evaluation order.  Ts starts at 1 and is multiplied by 0.8 until
 * min I / max I >= scale (:69-78; the reference's default scale is 1.1 * 2^-12);
 * d_Imea (sinogram layout, FP32) is then I / max I, *Ts the chosen Ts and *Imax = max I
 * (the reference divides iota by it).  One Ts and one max for all rays of the plan (one
 * slice per plan reproduces the reference's per-slice call).  kappa is K x L row-major
 * (material k's mass attenuation at the L energies, e.g. interpolated from the NIST
 * tables as bhc/genBeamHarden.m:19-29 does), epsilon strictly increasing, L >= 2,
 * 0 <= scale <= 1.  Allocates plan workspace on first use; blocks until Ts is found (one
 * small device-to-host copy per trial).
 * ------------------------------------------------------------------------------ */
#define CTPRJ_SIM_MAX_K 8
/* Ts search and normalisation on given per-material projections d_y (K sinograms of
 * the plan's size, contiguous; the density factors are applied inside, in FP64). */
int ctprj_poly_measure(ctprj_plan* plan, const float* d_y, int K, const double* density,
                       const double* kappa, const double* epsilon, const double* iota,
                       int L, double scale, float* d_Imea, double* Ts, double* Imax,
                       void* stream);
/* The same after projecting K density maps d_maps (K images of the plan's size,
 * contiguous) with the plan's forward projector. */
int ctprj_gen_beam_harden(ctprj_plan* plan, const float* d_maps, int K, const double* density,
                          const double* kappa, const double* epsilon, const double* iota,
                          int L, double scale, float* d_Imea, double* Ts, double* Imax,
                          void* stream);

/* ------------------------------------------------------------------------------
 * Orthogonal wavelet transforms and the wavelet-l1 proximal step: the sparsity prior
 * of the default BHC configuration (opt.proximal = 'wvltADMM',
 * bhc/beamhardenSpline.m:36,204-209), whose Psi / Psit are midwt / mdwt of the Rice
 * Wavelet Toolbox (tomography/ConfigCT.m:82-86).  A context holds `batch` images of
 * rows x cols (both divisible by 2^L; row-major, the MATLAB row index as row), the
 * scaling filter h (even length <= 32, e.g. daubcqf) and the level count L.  Images are
 * FP32 device arrays; the transforms and the ADMM state are FP64.  Create allocates
 * seven FP64 image-sized buffers.
 * ------------------------------------------------------------------------------ */
typedef struct ctprj_wavelet ctprj_wavelet;
int ctprj_wavelet_create(int rows, int cols, int batch, const double* h, int lh, int L,
                         ctprj_wavelet** out);
int ctprj_wavelet_destroy(ctprj_wavelet* w);
/* Support masks of Psi / Psit (tomography/ConfigCT.m:85-86, utils/maskFunc.m): d_mask
 * (image pixels, the reconstruction support opt.mask) and d_cmask (coefficients, maskk of
 * utils/wvltMask.m), float device arrays of the context's size, 0 = outside; NULL = no
 * mask.  Copied.  Afterwards ctprj_dwt reads masked pixels and returns masked
 * coefficients, ctprj_idwt the reverse, and the l1 value and the ADMM work on the masked
 * vectors exactly as the reference's closures. */
int ctprj_wavelet_set_masks(ctprj_wavelet* w, const float* d_mask, const float* d_cmask);
/* y = mdwt(x, h, L) (rwt/mdwt_r.c MDWT) of each image: periodised two-channel analysis,
 * rows then columns, L levels on the top-left block; lowpass first.  x may equal y. */
int ctprj_dwt(ctprj_wavelet* w, const float* d_x, float* d_y, void* stream);
/* x = midwt(y, h, L) (rwt/midwt_r.c MIDWT), the inverse.  y may equal x. */
int ctprj_idwt(ctprj_wavelet* w, const float* d_y, float* d_x, void* stream);
/* *l1 = ||mdwt(x)||_1 over all images (FP64; synchronises `stream`). */
int ctprj_wavelet_l1(ctprj_wavelet* w, const float* d_x, double* l1, void* stream);
/* x = argmin 0.5||x - a||^2 + I(x >= 0) + u ||mdwt(x)||_1 by the ADMM of npg/admm.m
 * (scaled by ||a||, adaptive penalty rho, stopping when both the change of the split
 * variable and the residual are at most min(relTol, 1e-3), or after maxItr + 1
 * iterations); d_init (may be NULL: a) starts the split variable.  *iterations gets the
 * iteration count (innerSearch).  The loop state stays on the device; three iterations
 * are one HIP graph (captured on the context's own stream at the first call, again
 * after new masks) and the host reads the state once per three iterations.  Calls on one
 * context share its buffers: they must not overlap on different streams. */
int ctprj_prox_wavelet_admm(ctprj_wavelet* w, const float* d_a, double u, double relTol,
                            int maxItr, const float* d_init, float* d_out, int* iterations,
                            void* stream);

/* ------------------------------------------------------------------------------
 * Total-variation proximal step (opt.proximal = 'tvl1' / 'tviso',
 * bhc/beamhardenSpline.m:215-224 -> utils/TV.m:56-80 -> utils/denoise_bound_mod.m): a
 * context holds `batch` images of rows x cols (rows, cols >= 2) and the TV type
 * (iso != 0: isotropic, else l1).  FP64 state (six image-sized buffers), FP32 images.
 * ------------------------------------------------------------------------------ */
typedef struct ctprj_tv ctprj_tv;
int ctprj_tv_create(int rows, int cols, int batch, int iso, ctprj_tv** out);
int ctprj_tv_destroy(ctprj_tv* w);
/* Support mask (opt.mask of TV.denoise, utils/TV.m:68-76): pixels outside it (0) are zero
 * in the input, the output and the TV value; NULL = no mask.  Copied. */
int ctprj_tv_set_mask(ctprj_tv* w, const float* d_mask);
/* *tv = TV(x) summed over the images (utils/tlv.m; synchronises `stream`). */
int ctprj_tv_value(ctprj_tv* w, const float* d_x, double* tv, void* stream);
/* x = argmin 0.5||x - a||^2 + lambda TV(x) over x >= 0 by FGP on the dual
 * (utils/denoise_bound_mod.m with the bounds [0, inf) of TV.denoise): stops after
 * maxIter iterations or when ||D - Dold|| / ||D|| < epsilon five times in a row; the
 * dual starts at zero (TV.denoise does not forward its init).  Norms and the stopping
 * test run over all images of the context together.  *iterations gets the count.
 * The loop state stays on the device; the host reads it once per four iterations. */
int ctprj_tv_denoise(ctprj_tv* w, const float* d_a, double lambda, double epsilon,
                     int maxIter, float* d_out, int* iterations, void* stream);

/* ------------------------------------------------------------------------------
 * On-device PNPG alpha step (npg/PNPG.m:50-167 with the non-negativity prox, on the
 * poissLAlpha objective, or gaussLAlpha under the plan's Gaussian model): no reference
 * counterpart as a C call -- the reference runs this loop in MATLAB
 * (bhc/beamhardenSpline.m:340-384 -> npg/NPG.m/PNPG.m).
 * ------------------------------------------------------------------------------ */
typedef struct ctprj_pnpg ctprj_pnpg;
typedef struct ctprj_pnpg_info {
    double cost;       /* objective at the current iterate */
    double t;          /* inverse step size */
    double theta;      /* momentum parameter */
    int iterations;    /* accepted iterations */
    int trials;        /* line-search trials (ctprj_pnpg_step calls) */
    int rejects;       /* trials rejected by the majorization test */
    int restarts;      /* momentum restarts */
    int last_accepted; /* 1 if the last trial was accepted */
    double difAlpha;   /* wavelet prox: relative change of the last accepted step */
    double admmTol;    /* wavelet prox: ADMM tolerance factor (npg/PNPG.m admmTol) */
    int innerSearch;   /* wavelet prox: ADMM iterations of the last trial */
    int maxInnerItr;   /* wavelet prox: ADMM iteration cap */
    double t_reset;    /* iterative priors: the Hessian step estimate of the last reset()
                          (npg/PNPG.m:168-172), 0 before any */
    int resets;        /* iterative priors: reset() calls (given-up line searches) */
} ctprj_pnpg_info;

/* Solver state over `plan` (all slices share one objective): Imea (device, packed like
 * sinograms), spectrum model (knots, E, Ie as in ctprj_bhc_grad_alpha), Ts, the start
 * point alpha0 (device image; also the first preAlpha) and the initial inverse step t0. */
int ctprj_pnpg_create(ctprj_plan* plan, const float* d_Imea, const double* knots, int E,
                      const double* Ie, float Ts, const float* d_alpha0, double t0,
                      ctprj_pnpg** out);
/* One line-search trial on d_alpha (updated in place when the trial is accepted).
 * Enqueues a fixed kernel sequence on `stream` with no host synchronisation and no
 * allocation: calls may be captured in a HIP graph and replayed. */
int ctprj_pnpg_step(ctprj_pnpg* solver, float* d_alpha, void* stream);
/* Failed majorization tests per iteration that shrink the step before the line search
 * gives up (goodMM = false, which resets theta and, with an iterative prior, the step
 * from the Hessian): the reference's 20 (npg/PNPG.m:85, npg/NPG.m:85) unless changed. */
int ctprj_pnpg_set_linesearch(ctprj_pnpg* solver, int max_shrinks);
/* Copies the device scalars to `info` (synchronises `stream`). */
int ctprj_pnpg_stats(ctprj_pnpg* solver, ctprj_pnpg_info* info, void* stream);
/* Selects the method before the first trial: npg = 0 PNPG (npg/PNPG.m, the default
 * here), npg = 1 NPG (npg/NPG.m:45-148, opt.alphaStep = 'NPG', the reference's default):
 * FISTA theta (1 + sqrt(1 + 4 theta^2))/2, xbar and its gradient once per iteration (a
 * trial that retries recomputes the same values), the adaptive increase t/4, and in
 * the first preSteps iterations (opt.preSteps = 2) the step grows while the test holds. */
int ctprj_pnpg_set_method(ctprj_pnpg* solver, int npg, int preSteps);
/* Support of alpha (the reference's alpha is the vector of opt.mask's pixels): pixels
 * outside d_mask (0) stay zero (xbar and the gradient are masked); NULL = no mask.
 * Copied.  A wavelet or TV prior should carry the same mask (ctprj_wavelet_set_masks,
 * ctprj_tv_set_mask). */
int ctprj_pnpg_set_mask(ctprj_pnpg* solver, const float* d_mask);
/* Switches the prox to the wavelet l1 penalty u ||mdwt(alpha)||_1 with non-negativity
 * (opt.proximal = 'wvltADMM', bhc/beamhardenSpline.m:204-209), solved per trial by
 * ctprj_prox_wavelet_admm with tolerance admmTol * difAlpha and the reference's
 * admmTol / maxInnerItr schedule (npg/PNPG.m:113-146); the objective then includes
 * u ||W alpha||_1.  `w` must hold exactly the plan's images (one batch image per slice
 * of equal size) and outlive the solver; w = NULL returns to the non-negativity prox.
 * Only before the first trial.  In this mode ctprj_pnpg_step synchronises `stream`
 * (the ADMM stopping tests run on the host) and cannot be captured in a graph.  On a
 * given-up line search theta restarts at 1 and t = min(t, the Hessian step estimate at
 * alpha), as the reference's reset() (npg/PNPG.m:168-172, npg/Methods.m:108-115). */
int ctprj_pnpg_set_wavelet(ctprj_pnpg* solver, ctprj_wavelet* w, double u);
/* The same with the TV prox (opt.proximal = 'tvl1' / 'tviso', penalty u TV(alpha),
 * utils/TV.m denoise with innerThresh = admmTol * difAlpha and maxInnerItr).  tv = NULL
 * returns to the non-negativity prox. */
int ctprj_pnpg_set_tv(ctprj_pnpg* solver, ctprj_tv* tv, double u);
int ctprj_pnpg_destroy(ctprj_pnpg* solver);
/* Snapshot and resume of a PNPG run (bhc/doBHC.m:78-91 saves its workspace after each
 * penalty level and warm-starts the next level from it).  A snapshot is an opaque host
 * blob of ctprj_pnpg_snapshot_bytes bytes holding the solver's iteration state (theta, t,
 * the counters, preAlpha, the prox's tolerance and inner-iteration schedule, the last
 * reset() estimate).  ctprj_pnpg_save fills it after every trial queued on `stream`
 * (synchronises); ctprj_pnpg_load restores it into a solver created with the same plan,
 * Imea, spectrum, Ts and prox (set_wavelet / set_tv / set_mask called as for the saved
 * one), after which ctprj_pnpg_step continues exactly as the saved solver would (bit for
 * bit).  alpha is the caller's array and is saved beside the blob.  A blob of another
 * size, version, image / ray count, E or prox gives CTPRJ_EINVAL.  No counterpart call
 * in the reference, which saves the MATLAB workspace (`save([filename '.mat'])`). */
int ctprj_pnpg_snapshot_bytes(const ctprj_pnpg* solver, long long* bytes);
int ctprj_pnpg_save(ctprj_pnpg* solver, void* host, long long bytes, void* stream);
int ctprj_pnpg_load(ctprj_pnpg* solver, const void* host, long long bytes, void* stream);

/* ------------------------------------------------------------------------------
 * Library state.
 * ------------------------------------------------------------------------------ */
const char* ctprj_last_error(void);
/* Device name / arch string of the current device (e.g. "gfx950"); 0 on success. */
int ctprj_device_arch(char* buf, int buflen);
/* Library version string. */
const char* ctprj_version(void);
/* Failure injection for tests of the error paths: with k > 0, the k-th device
 * allocation the library makes from now on fails as out of memory (once); 0 disarms.
 * No reference counterpart. */
int ctprj_debug_fail_alloc(int k);
/* The PNPG decision after a finished line search on the non-negativity path
 * (npg/PNPG.m:99-160 with u = 0), evaluated on the host by the same function the device
 * runs: in = {cost, f_old, f_new, theta (= this trial's newTheta), admmTol}, ls 1 (the
 * majorization test held) or 2 (given up), moved = ||xbar - alpha||_1 != 0.  code: 0
 * restart, 1 retry with admmTol/10, 2 take newX, 3 keep alpha; out = {cost, theta,
 * admmTol} after it.  For tests of the rule; no device work. */
int ctprj_debug_pnpg_decide(const double* in, int ls, int moved, double* out, int* code);

#ifdef __cplusplus
}
#endif
#endif /* CTPRJ_H */
