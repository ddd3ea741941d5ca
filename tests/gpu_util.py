"""Helpers shared by the GPU parity tests: warper descriptors built from the
oracle's restated warpers, so both sides consume the same tables."""
import numpy as np

from keldy_amd import capi
from oracle import oracle as O


def analytic_norm(kind, u_max, w, power=1):
    """The makers' normalisations, in long double as the reference computes them
    (product_1d_simple.hpp:560-630)."""
    L = np.longdouble
    if kind == "exponential":
        return float(-np.expm1(-L(u_max) / L(w)))
    if kind == "inverse":
        return float(np.log1p(L(u_max) / L(w)))
    if kind == "inverse_square":
        return float(L(u_max) / L(w + u_max))
    if kind == "inverse_power":
        um = L(u_max)
        return float((power - 1) * np.power(um + L(w), power - 1) / (np.power(um / L(w) + 1, power - 1) - 1))
    return 1.0


KIND = {"uniform": capi.KD_MAP_UNIFORM, "exponential": capi.KD_MAP_EXPONENTIAL,
        "inverse": capi.KD_MAP_INVERSE, "inverse_square": capi.KD_MAP_INVERSE_SQ,
        "inverse_power": capi.KD_MAP_INVERSE_POW}


def chain(spec):
    """spec: list of ("identity",) | ("plasma", t) | ("map", kind, u_max, w, power) (same map on
    every axis) in the order warpers are added.  Returns (oracle Train,
    (descs, maps, tables)) for the device."""
    o_ws, descs, maps = [], [], []
    for s in spec:
        if s[0] == "identity":
            o_ws.append(O.Warper(O.WARP_IDENTITY))
            descs.append(capi.WarperDesc(capi.KD_WARP_IDENTITY, 0, 0.0, 0))
        elif s[0] == "plasma":
            o_ws.append(O.plasma(s[1]))
            descs.append(capi.WarperDesc(capi.KD_WARP_PLASMA, 0, s[1], 0))
        else:
            _, kind, um, w, pw = s
            o_ws.append(O.product(O.Map1d(kind, um, w=w, power=pw)))
            m = capi.Map1d()
            m.kind, m.power, m.u_max, m.w = KIND[kind], pw, um, w
            m.norm = analytic_norm(kind, um, w, pw)
            descs.append(capi.WarperDesc(capi.KD_WARP_PRODUCT, 1, 0.0, len(maps)))
            maps.append(m)
    return O.Train(o_ws), (descs, maps, np.zeros(0))


def charge_ctx(g0, order, t, prec=capi.KD_PREC_FP64):
    ctx = capi.Context(0)
    ctx.set_g0_table(g0["lesser"], g0["greater"], g0["time_max"])
    ctx.set_integrand(order, (t, 0, 0), (t, 0, 1))
    ctx.set_precision(prec)
    return ctx


def oracle_charge(g0):
    tab = O.G0Table(g0["lesser"], g0["greater"], g0["time_max"])
    return tab
