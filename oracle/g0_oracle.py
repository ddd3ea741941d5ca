"""g0 time tables for the oracle (TEST INFRASTRUCTURE ONLY).

Restates the reference's g0 tabulation (c++/keldy/impurity_oneband/model.cpp)
with scipy:
  flatband_analytic  make_flat_band_analytic (model.cpp:428-466): Ei closed form
  g0_real_axis       g^{<,>}(t) = (1/2pi) int e^{-i w t} g^{<,>}(w) dw with g(w) from
                     g0_model_omega (model.hpp:100-153, model.cpp:79-123), integrated on
                     the real axis by QUADPACK (scipy.integrate.quad, Fourier weights).
                     This is an independent route to the value the reference's
                     "contour" method (model.cpp:326-426) approximates.
"""
import numpy as np
import scipy.integrate as si
import scipy.special as sp


def time_mesh(time_max, nr_time_points):
    n = nr_time_points + (1 - nr_time_points % 2)  # model.cpp:87-90: made odd
    return np.linspace(-time_max, time_max, n)


def flatband_analytic_lesser(t, Gamma):
    """model.cpp:444-454"""
    t = np.asarray(t, dtype=np.float64)
    Gt = Gamma * t
    with np.errstate(all="ignore"):
        re = (np.exp(Gt) * sp.expi(-Gt) - np.exp(-Gt) * sp.expi(Gt)) / (2 * np.pi)
    re = np.where(t == 0.0, 0.0, re)
    im = 0.5 * np.exp(-np.abs(Gt))
    out = re + 1j * im
    return np.where(t == 0.0, 0.5j, out)


def flatband_analytic(time_max, nr_time_points, Gamma):
    t = time_mesh(time_max, nr_time_points)
    les = flatband_analytic_lesser(t, Gamma)
    return t, les, np.conj(les)


# --------------------------------------------------------------------------
def n_fermi(w, beta):
    """common.hpp:64-75 (real argument)."""
    if w == 0.0:
        return 0.5
    if np.isinf(beta):
        return 0.0 if w > 0 else 1.0
    if w > 0:
        y = np.exp(-beta * w)
        return y / (1.0 + y)
    return 1.0 / (np.exp(beta * w) + 1.0)


def hyb_R(w, bath, Gamma, D):
    """bath_hybrid_R_left (model.cpp:92-119), real frequency."""
    if bath == "flatband":
        return -0.5j * Gamma
    if bath == "semicircle":
        if abs(w) < D:
            return 0.5 * (Gamma / D) * (w - 1j * np.sqrt(D * D - w * w))
        sgn = -1.0 if w < 0 else 1.0
        return 0.5 * (Gamma / D) * (w - sgn * np.sqrt(w * w - D * D))
    if bath == "lorentzian":
        return Gamma * D / 2.0 / (w + 1j * D)
    raise ValueError(bath)


def g0_dot_omega(w, lesser, p):
    """g0_model_omega::g0_dot (model.hpp:100-112), real omega."""
    bath, G, D = p["bath_type"], p["Gamma"], p["half_bandwidth"]
    beta = np.inf if p["beta"] < 0 else p["beta"]
    hR = hyb_R(w, bath, G, D)
    hA = np.conj(hR)
    gR = 1.0 / (w - p["eps_d"] - 2 * hR)
    gA = 1.0 / (w - p["eps_d"] - 2 * hA)
    sign = 1 if lesser else -1
    muL, muR = -p["bias_V"] / 2, p["bias_V"] / 2
    out = -sign * gR * (n_fermi(sign * (w - muL), beta) * (hR - hA) +
                        n_fermi(sign * (w - muR), beta) * (hR - hA)) * gA
    return 1j * out.imag


def _fourier(h, t, epsabs, epsrel, finite=None):
    """(1/2pi) int e^{-i w t} i h(w) dw for a real function h."""
    if finite is not None:
        a, b = finite
        if t == 0.0:
            return 1j * si.quad(h, a, b, epsabs=epsabs, epsrel=epsrel, limit=2000)[0] / (2 * np.pi)
        c = si.quad(h, a, b, weight="cos", wvar=t, epsabs=epsabs, epsrel=epsrel, limit=2000)[0]
        s = si.quad(h, a, b, weight="sin", wvar=t, epsabs=epsabs, epsrel=epsrel, limit=2000)[0]
        return (s + 1j * c) / (2 * np.pi)
    c = s = 0.0
    if abs(t) < 1.0:
        for a, b in ((-np.inf, 0.0), (0.0, np.inf)):
            c += si.quad(lambda w: h(w) * np.cos(w * t), a, b, epsabs=epsabs, epsrel=epsrel, limit=2000)[0]
            s += si.quad(lambda w: h(w) * np.sin(w * t), a, b, epsabs=epsabs, epsrel=epsrel, limit=2000)[0]
        return (s + 1j * c) / (2 * np.pi)
    at = abs(t)
    for sgn in (1.0, -1.0):
        f = lambda w: h(sgn * w)
        c += si.quad(f, 0.0, np.inf, weight="cos", wvar=at, epsabs=epsabs, limlst=200)[0]
        s += sgn * si.quad(f, 0.0, np.inf, weight="sin", wvar=at, epsabs=epsabs, limlst=200)[0]
    s *= np.sign(t)
    return (s + 1j * c) / (2 * np.pi)


def g0_real_axis(t, lesser, p, epsabs=1e-13, epsrel=1e-11):
    """(1/2pi) int e^{-i w t} g(w) dw with g(w) = i h(w) purely imaginary.

    Flat band: the T=0, eps_d=0, V=0 spectral function of the same Gamma,
    2 Gamma theta(-+w)/(w^2+Gamma^2), is subtracted and added back through its
    closed-form transform (the Ei formula of make_flat_band_analytic), which
    leaves a remainder decaying like 1/w^3 for the quadrature.  Semicircle:
    compact support [-D, D].
    """
    h = lambda w: g0_dot_omega(w, lesser, p).imag
    G = p["Gamma"]
    if p["bath_type"] == "flatband":
        ref = flatband_analytic_lesser(t, G)
        if lesser:
            h_ref = lambda w: 2 * G / (w * w + G * G) if w < 0 else (G / (G * G) if w == 0 else 0.0)
        else:
            ref = np.conj(ref)
            h_ref = lambda w: -2 * G / (w * w + G * G) if w > 0 else (-G / (G * G) if w == 0 else 0.0)
        return complex(ref) + _fourier(lambda w: h(w) - h_ref(w), t, epsabs, epsrel)
    if p["bath_type"] == "semicircle":
        D = p["half_bandwidth"]
        return _fourier(h, t, epsabs, epsrel, finite=(-D, D))
    return _fourier(h, t, epsabs, epsrel)
