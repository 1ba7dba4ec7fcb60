"""Polychromatic beam-hardening model on the device.

Mirrors of the reference's MATLAB contract, computed by libctprj:

* ``Spline(sType, knots).polyIout(s, Ie)`` -> (BL, sBL, ssBL), sType 'b1' / 'b0' / 'dis'
  -- bhc/Spline.m:11-28,122-302
* ``poissLAlpha(Imea, Ie, alpha, proj, Ts, knots)`` -> (f, g, hw) -- bhc/poissLAlpha.m
* ``poissLI(Imea, Ie, proj, Ts, knots, alpha=|sraw=)`` -> (f, g, H) -- bhc/poissLI.m
* ``gaussLAlpha`` / ``gaussLI`` (same arguments) -- bhc/gaussLAlpha.m, bhc/gaussLI.m
* ``ie_solve`` -- the I-epsilon step of bhc/beamhardenSpline.m:446-451
* ``gen_beam_harden`` / ``poly_measure`` -- the measurement simulator bhc/genBeamHarden.m,
  with ``mass_atten_kappa`` (its table interpolation) and ``read_spectrum``
  (bhc/readSpectrum.m)

Every likelihood call takes ``basis='b1'|'b0'|'dis'`` (bhc/beamhardenSpline.m:60
opt.spectBasis); its knots are E+2 values for 'b1' and 'dis', E+1 for 'b0'.

``kappa_logspace`` / ``spline_knots`` reproduce the mass-attenuation grid and the
outer knots of bhc/beamhardenSpline.m:138-139,164-165; ``normalize_ie`` the
``Ie = Ie / polyIout(0, Ie)`` step of bhc/beamhardenSpline.m:182.
The FP64 values (NLL, the I-epsilon gradient and Hessian) come back as float64 tensors.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import MAX_E, IeInfo, check, dbl_array, lib
from .projector import Projector, _ptr, _stream_ptr


def kappa_logspace(E: int, lo: float = 0.03, hi: float = 30.0) -> np.ndarray:
    """E mass-attenuation nodes log-spaced over [lo, hi] (bhc/Spline.m:16 default range)."""
    return np.logspace(np.log10(lo), np.log10(hi), E)


def kappa_logspan(E: int, logspan: float = 3.0) -> np.ndarray:
    """bhc/beamhardenSpline.m:138-139: logspace(-floor(E/2)/(E-1)*logspan,
    floor(E/2-0.5)/(E-1)*logspan, E)."""
    return np.logspace(-np.floor(E / 2) / (E - 1) * logspan,
                       np.floor(E / 2 - 0.5) / (E - 1) * logspan, E)


def spline_knots(kappa) -> np.ndarray:
    """[kappa_1/q, kappa_1..kappa_E, kappa_E*q], q = kappa_2/kappa_1
    (bhc/beamhardenSpline.m:164-165)."""
    k = np.asarray(kappa, np.float64).ravel()
    q = k[1] / k[0]
    return np.concatenate([[k[0] / q], k, [k[-1] * q]])


def basis_at_zero(knots) -> np.ndarray:
    """phi_j(0) = (k_{j+2} - k_j)/2: the area of hat j."""
    k = np.asarray(knots, np.float64)
    return (k[2:] - k[:-2]) / 2.0


def normalize_ie(Ie, knots) -> np.ndarray:
    """Ie / polyIout(0, Ie) so that BL(0) = 1 (bhc/beamhardenSpline.m:182)."""
    Ie = np.asarray(Ie, np.float64)
    return Ie / float(basis_at_zero(knots) @ Ie)


BASIS = {"b1": 0, "b0": 1, "dis": 2}      # CTPRJ_BASIS_*
NOISE = {"poisson": 0, "gaussian": 1}     # CTPRJ_NOISE_*


def _basis_id(basis: str) -> int:
    try:
        return BASIS[basis.lower()]
    except KeyError:
        raise ValueError(f"unknown spectrum basis '{basis}' (b1, b0, dis)") from None


def _knots_ie(knots, Ie, basis: str = "b1"):
    """E (number of bases) and ctypes arrays of the knots and Ie of `basis`."""
    k = np.asarray(knots, np.float64).ravel()
    E = k.size - (1 if _basis_id(basis) == BASIS["b0"] else 2)
    if E < 1 or E > MAX_E:
        raise ValueError(f"{k.size} knots give {E} '{basis}' bases (need 1..{MAX_E})")
    kp, kh = dbl_array(k)
    if Ie is None:
        return E, kp, kh, None, None
    ie = np.asarray(Ie, np.float64).ravel()
    if ie.size != E:
        raise ValueError(f"Ie has {ie.size} entries, knots define {E} bases")
    ip, ih = dbl_array(ie)
    return E, kp, kh, ip, ih


def set_model(proj: Projector, basis: str = "b1", noise: str = "poisson") -> None:
    """Spectrum basis and noise model of the plan's beam-hardening calls."""
    check(lib().ctprj_plan_set_bhc_model(proj.handle, _basis_id(basis), NOISE[noise.lower()]))


def _f32(t: torch.Tensor, n: int, what: str) -> torch.Tensor:
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda" or t.dtype != torch.float32:
        raise TypeError(f"{what} must be a float32 CUDA tensor")
    if t.numel() != n:
        raise ValueError(f"{what} has {t.numel()} elements, expected {n}")
    return t.contiguous()


class Spline:
    """``Spline(sType, knots)`` of bhc/Spline.m with the polyIout closure on the device;
    sType 'b1' (hats, E+2 knots), 'b0' (boxes, E+1 knots) or 'dis' (nodes, E+2 knots)."""

    def __init__(self, sType: str, knots):
        self.sType = sType.lower()
        self.basis = _basis_id(self.sType)
        self.knots = np.asarray(knots, np.float64).ravel()

    def polyIout(self, s: torch.Tensor, Ie=None, want: int = 3):
        """Ie given: (BL, sBL, ssBL) per ray value.  Ie None: the count x E basis matrix."""
        n = s.numel()
        s = _f32(s.reshape(-1), n, "s")
        E, kp, _kh, ip, _ih = _knots_ie(self.knots, Ie, self.sType)
        if Ie is None:
            A = torch.empty((n, E), dtype=torch.float32, device=s.device)
            check(lib().ctprj_poly_iout(self.basis, kp, E, None, _ptr(s), n, _ptr(A), None,
                                        None, _stream_ptr()))
            return A
        outs = [torch.empty(n, dtype=torch.float32, device=s.device) for _ in range(want)]
        ptrs = [_ptr(o) for o in outs] + [None] * (3 - want)
        check(lib().ctprj_poly_iout(self.basis, kp, E, ip, _ptr(s), n, *ptrs, _stream_ptr()))
        return tuple(outs)


def _grad_alpha(noise, Imea, Ie, alpha, proj, Ts, knots, want_grad, want_hess, basis):
    alpha = _f32(alpha.reshape(-1), proj.img_elems, "alpha")
    Imea = _f32(Imea.reshape(-1), proj.sino_elems, "Imea")
    E, kp, _kh, ip, _ih = _knots_ie(knots, Ie, basis)
    set_model(proj, basis, noise)
    dev = alpha.device
    f = torch.empty(1, dtype=torch.float64, device=dev)  # written by the call
    g = torch.empty(proj.img_elems, dtype=torch.float32, device=dev) if want_grad else None
    hw = torch.empty(proj.sino_elems, dtype=torch.float32, device=dev) if want_hess else None
    check(lib().ctprj_bhc_grad_alpha(proj.handle, _ptr(alpha), _ptr(Imea), kp, E, ip,
                                     float(Ts), _ptr(f), _ptr(g) if g is not None else None,
                                     _ptr(hw) if hw is not None else None, _stream_ptr()))
    out = [f[0]]
    if want_grad:
        out.append(g)
    if want_hess:
        out.append(hw)
    return tuple(out) if len(out) > 1 else out[0]


def poissLAlpha(Imea: torch.Tensor, Ie, alpha: torch.Tensor, proj: Projector, Ts: float,
                knots, want_grad: bool = True, want_hess: bool = False, basis: str = "b1"):
    """[f, g(, hw)] of bhc/poissLAlpha.m with Phi = Ts*A, Phit = Ts*A^T and polyIout of
    `basis`.  f: float64 scalar tensor (FP64 reduction); g: float32 image; hw: Hessian
    ray weight."""
    return _grad_alpha("poisson", Imea, Ie, alpha, proj, Ts, knots, want_grad, want_hess, basis)


def gaussLAlpha(Imea: torch.Tensor, Ie, alpha: torch.Tensor, proj: Projector, Ts: float,
                knots, want_grad: bool = True, want_hess: bool = False, basis: str = "b1"):
    """[f, g(, hw)] of bhc/gaussLAlpha.m: Err = log(BL/Imea), f = Err'Err/2,
    g = Phit(Err.*sBL./BL), hw = (1-Err).*(sBL./BL).^2 + Err.*ssBL./BL."""
    return _grad_alpha("gaussian", Imea, Ie, alpha, proj, Ts, knots, want_grad, want_hess, basis)


def alpha_rays(proj: Projector, sraw: torch.Tensor, Imea: torch.Tensor, Ie, knots, Ts: float,
               want_w: bool, basis: str = "b1", noise: str = "poisson"):
    """Per-ray pass of poissLAlpha / gaussLAlpha on a stored projection sraw = A*alpha:
    (f, w or None)."""
    sraw = _f32(sraw.reshape(-1), proj.sino_elems, "sraw")
    Imea = _f32(Imea.reshape(-1), proj.sino_elems, "Imea")
    E, kp, _kh, ip, _ih = _knots_ie(knots, Ie, basis)
    set_model(proj, basis, noise)
    f = torch.empty(1, dtype=torch.float64, device=sraw.device)  # written by the call
    w = torch.empty_like(sraw) if want_w else None
    check(lib().ctprj_bhc_alpha_rays(proj.handle, _ptr(sraw), _ptr(Imea), kp, E, ip, float(Ts),
                                     _ptr(f), _ptr(w) if w is not None else None, None,
                                     _stream_ptr()))
    return f[0], w


def _grad_ie(noise, Imea, Ie, proj, Ts, knots, alpha, sraw, basis):
    if (alpha is None) == (sraw is None):
        raise ValueError("give exactly one of alpha and sraw")
    Imea = _f32(Imea.reshape(-1), proj.sino_elems, "Imea")
    E, kp, _kh, ip, _ih = _knots_ie(knots, Ie, basis)
    set_model(proj, basis, noise)
    dev = Imea.device
    # every entry is written by the call
    f = torch.empty(1, dtype=torch.float64, device=dev)
    g = torch.empty(E, dtype=torch.float64, device=dev)
    H = torch.empty((E, E), dtype=torch.float64, device=dev)
    a = _ptr(_f32(alpha.reshape(-1), proj.img_elems, "alpha")) if alpha is not None else None
    s = _ptr(_f32(sraw.reshape(-1), proj.sino_elems, "sraw")) if sraw is not None else None
    check(lib().ctprj_bhc_grad_ie(proj.handle, a, s, _ptr(Imea), kp, E, ip, float(Ts), _ptr(f),
                                  _ptr(g), _ptr(H), _stream_ptr()))
    return f[0], g, H


def poissLI(Imea: torch.Tensor, Ie, proj: Projector, Ts: float, knots,
            alpha: torch.Tensor | None = None, sraw: torch.Tensor | None = None,
            basis: str = "b1"):
    """[f, g, H] of bhc/poissLI.m with A = polyIout(Phi(alpha), []) never formed.

    Give alpha (projected inside) or sraw = A*alpha.  Returns float64 tensors."""
    return _grad_ie("poisson", Imea, Ie, proj, Ts, knots, alpha, sraw, basis)


def gaussLI(Imea: torch.Tensor, Ie, proj: Projector, Ts: float, knots,
            alpha: torch.Tensor | None = None, sraw: torch.Tensor | None = None,
            basis: str = "b1"):
    """[f, g, H] of bhc/gaussLI.m: Err = log(Ir/Imea), f = Err'Err/2, g = A'(Err./Ir),
    H = A' diag((1-Err)./Ir.^2) A."""
    return _grad_ie("gaussian", Imea, Ie, proj, Ts, knots, alpha, sraw, basis)


def ie_solve(Imea: torch.Tensor, Ie0, proj: Projector, Ts: float, knots,
             alpha: torch.Tensor | None = None, sraw: torch.Tensor | None = None,
             basis: str = "b1", noise: str = "poisson", maxit: int = 50, tol: float = 1e-12):
    """The I-epsilon step of bhc/beamhardenSpline.m:446-451 (IeStep = ActiveSet): minimise
    poissLI / gaussLI over Ie >= 0 for the projection of alpha (or sraw = A*alpha).
    Returns (Ie as a float64 numpy array, info dict)."""
    if (alpha is None) == (sraw is None):
        raise ValueError("give exactly one of alpha and sraw")
    Imea = _f32(Imea.reshape(-1), proj.sino_elems, "Imea")
    E, kp, _kh, ip, ih = _knots_ie(knots, Ie0, basis)
    set_model(proj, basis, noise)
    a = _ptr(_f32(alpha.reshape(-1), proj.img_elems, "alpha")) if alpha is not None else None
    s = _ptr(_f32(sraw.reshape(-1), proj.sino_elems, "sraw")) if sraw is not None else None
    info = IeInfo()
    check(lib().ctprj_ie_solve(proj.handle, a, s, _ptr(Imea), kp, E, ip, float(Ts), int(maxit),
                               float(tol), ctypes.byref(info), _stream_ptr()))
    return np.array(ih[:E], np.float64), {f: getattr(info, f) for f, _ in IeInfo._fields_}


# ------------------------------------------------------------------ simulator ---

def read_spectrum(path) -> tuple[np.ndarray, np.ndarray]:
    """(epsilon, iota) from a two-column text spectrum file, as bhc/readSpectrum.m reads
    the tungsten_<kV>kV_<ripple>_1Gy.data files (energy in keV, photons per bin)."""
    data = np.loadtxt(path, dtype=np.float64, ndmin=2)
    if data.shape[1] < 2:
        raise ValueError(f"{path}: expected two columns")
    return data[:, 0].copy(), data[:, 1].copy()


def mass_atten_kappa(energy_mev, mu_rho, epsilon) -> np.ndarray:
    """Mass attenuation at the spectrum energies epsilon (keV) from a table (energy in MeV,
    mass attenuation coefficient), as bhc/genBeamHarden.m:19-29 does with the NIST tables:
    the table points inside [min eps, max eps] plus one on each side, interpolated by a
    not-a-knot cubic spline (MATLAB interp1 'spline') in log energy / log attenuation."""
    from scipy.interpolate import CubicSpline

    e = np.asarray(energy_mev, np.float64).ravel() * 1e3
    mu = np.asarray(mu_rho, np.float64).ravel()
    eps = np.asarray(epsilon, np.float64).ravel()
    if e.shape != mu.shape:
        raise ValueError("energy and attenuation tables differ in length")
    inside = np.nonzero((e >= eps.min()) & (e <= eps.max()))[0]
    if inside.size == 0:
        raise ValueError("no table energy inside the spectrum's range")
    i1, i2 = inside[0] - 1, inside[-1] + 1
    if i1 < 0 or i2 >= e.size:
        raise ValueError("the table does not extend one point beyond the spectrum's range")
    x, y = np.log(e[i1:i2 + 1]), np.log(mu[i1:i2 + 1])
    if np.any(np.diff(x) <= 0):
        raise ValueError("table energies must increase inside the interpolation range")
    return np.exp(CubicSpline(x, y, bc_type="not-a-knot")(np.log(eps)))


def _sim_inputs(K, density, kappa, epsilon, iota):
    density = np.asarray(density, np.float64).ravel()
    eps = np.asarray(epsilon, np.float64).ravel()
    io = np.asarray(iota, np.float64).ravel()
    kap = np.asarray(kappa, np.float64).reshape(K, -1)
    if density.size != K or kap.shape[1] != eps.size or io.size != eps.size:
        raise ValueError("density (K), kappa (K x L), epsilon (L) and iota (L) disagree")
    return density, kap, eps, io


def _simulate(entry, proj, src, K, density, kappa, epsilon, iota, scale):
    density, kap, eps, io = _sim_inputs(K, density, kappa, epsilon, iota)
    dens_p, dens_h = dbl_array(density)
    kap_p, kap_h = dbl_array(kap.ravel())
    eps_p, eps_h = dbl_array(eps)
    io_p, io_h = dbl_array(io)
    Imea = torch.empty(proj.sino_elems, dtype=torch.float32, device=src.device)
    ts, imax = ctypes.c_double(), ctypes.c_double()
    check(entry(proj.handle, _ptr(src), K, dens_p, kap_p, eps_p, io_p, eps.size, float(scale),
                _ptr(Imea), ctypes.byref(ts), ctypes.byref(imax), _stream_ptr()))
    k_max = int(np.argmax(density))  # the first of the densest materials
    return {"Imea": Imea, "Ts": ts.value, "Imax": imax.value, "iota": io / imax.value,
            "epsilon": eps, "kappa": kap[k_max].copy(), "density": float(density[k_max])}


def gen_beam_harden(proj: Projector, maps, density, kappa, epsilon, iota,
                    scale: float = 2.0 ** -12 * 1.1) -> dict:
    """Polychromatic measurements of K materials (bhc/genBeamHarden.m): maps is a K x
    img_elems float32 CUDA tensor (or a list of K images), density K values, kappa K x L
    attenuations at the L energies epsilon, iota the spectrum there.  Returns the fields of
    the reference's (CTdata, args): Imea (normalised, FP32 sinogram), Ts, iota / max I,
    epsilon, and kappa / density of the densest material."""
    if isinstance(maps, (list, tuple)):
        maps = torch.stack([m.reshape(-1) for m in maps])
    K = max(1, maps.numel() // proj.img_elems)
    maps = _f32(maps.reshape(-1), K * proj.img_elems, "maps")
    return _simulate(lib().ctprj_gen_beam_harden, proj, maps, K, density, kappa, epsilon,
                     iota, scale)


def poly_measure(proj: Projector, y, density, kappa, epsilon, iota,
                 scale: float = 2.0 ** -12 * 1.1) -> dict:
    """gen_beam_harden on given projections y (K x sino_elems, y_k = A x_k before the
    density factor)."""
    if isinstance(y, (list, tuple)):
        y = torch.stack([t.reshape(-1) for t in y])
    K = max(1, y.numel() // proj.sino_elems)
    y = _f32(y.reshape(-1), K * proj.sino_elems, "y")
    return _simulate(lib().ctprj_poly_measure, proj, y, K, density, kappa, epsilon, iota, scale)


__all__ = ["gen_beam_harden", "poly_measure", "mass_atten_kappa", "read_spectrum", "ie_solve",
           "Spline", "poissLAlpha", "poissLI", "gaussLAlpha", "gaussLI", "alpha_rays",
           "set_model", "kappa_logspace", "kappa_logspan", "spline_knots", "basis_at_zero",
           "normalize_ie", "BASIS", "NOISE"]
