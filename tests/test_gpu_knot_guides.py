"""Knot searches of the tabulated maps through their guides, against the oracle.

The device finds the knot interval of an inverse-CDF value (and of the mapped
time) from a guide built by kd_set_warper_chain: a uniform bucket index gives
the bounds of the interval, and a bisection inside them the interval itself.
The oracle runs the plain GSL bisection over the whole array
(gsl_interp_bsearch, behind warper_product_1d_simple_interp_nearest_t,
product_1d_simple.hpp:198-217).  Both must pick the same interval on every
point: the Jacobian 1/f with f = (f[j] + f[j+1]) / 2 and the mapped time (hence
the integrand value) both depend on it.

The knots are adversarial: exact duplicates in time (zero-width intervals),
a flat stretch of f (duplicated CDF knots), trailing zeros (n_inv < n_pts), a
cluster of 96 time knots inside one bucket (bisection steps inside the guide's
bounds).  The points put the inverse-CDF value exactly on every knot, one ulp
either side of every knot, halfway between knots, exactly on every bucket edge
of the guide and one ulp either side, and at l = 0 and 1.  A two-knot map (no
guide: plain bisection) runs beside them.

The device skips the time search where the interval record of the inverse-CDF
search already holds the mapped time (t[i] <= x < t[i+1]: the search would end
in i); a value one ulp below a CDF knot F[i+1] maps onto t[i+1] itself, so
those points take the time search.
"""
import numpy as np
import pytest

from keldy_amd import capi
from oracle import oracle as O
from tests import gpu_util as U

pytestmark = pytest.mark.gpu
T = 10.0


def _adversarial_knots(seed, n=257):
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(0.0, T, n - 98))
    t = np.concatenate([t, 3.0 + 1e-3 * np.sort(rng.uniform(size=96)), [0.0, T]])
    t = np.sort(t)
    for k in (20, 21, 150):  # exact duplicates (zero-width intervals)
        t[k + 1] = t[k]
    f = rng.uniform(0.2, 2.0, n)
    f[60:75] = 0.0    # flat CDF: duplicated F knots
    f[-6:] = 0.0      # trailing zeros: n_inv < n_pts
    return t, f


def _nearest(times, f1, tabs):
    """Oracle map and device descriptor of one nearest-interpolation map."""
    om, _ = O.nearest_map(times, f1)
    tk, fk, Fk = om.keep  # the normalised arrays the oracle map reads (t, f, F)
    m = capi.Map1d()
    m.kind, m.u_max, m.w, m.norm = capi.KD_MAP_NEAREST, float(tk[-1]), 1.0, 1.0
    m.n_pts, m.n_inv = len(tk), om.c.n_inv
    m.t_off = len(tabs)
    tabs.extend(tk)
    m.F_off = len(tabs)
    tabs.extend(Fk)
    m.f_off = len(tabs)
    tabs.extend(fk)
    return om, m, Fk[: om.c.n_inv]


def _edge_points(F):
    """l such that y = -l lies on the knots and one ulp either side, between
    them, and on the guide's bucket edges (2 n_inv buckets over [F[0], F[-1]])
    and one ulp either side."""
    n = len(F)
    G = 2 * n
    a, width = F[0], (F[-1] - F[0]) / G
    # one ulp below a knot F[i+1]: the mapped time rounds onto t[i+1], outside
    # the interval record [t[i], t[i+1]), so the device's time search runs
    y = [F, 0.5 * (F[:-1] + F[1:]), np.nextafter(F, -np.inf), np.nextafter(F, np.inf)]
    e = a + np.arange(G + 1) * width
    y += [e, np.nextafter(e, np.inf), np.nextafter(e, -np.inf)]
    l = np.clip(-np.concatenate(y), 0.0, 1.0)
    return np.concatenate([l, [0.0, 1.0, np.nextafter(0.0, 1.0), np.nextafter(1.0, 0.0)]])


def _time_search_points(om, l):
    """Points whose mapped time (FP64, as the device forms it) is not inside the
    interval record [t[i], t[i+1]) of their inverse-CDF interval i."""
    t, _, F = om.keep
    n_inv = om.c.n_inv
    y = -np.asarray(l)
    i = np.clip(np.searchsorted(F[:n_inv], y, side="right") - 1, 0, n_inv - 2)
    dx = F[i + 1] - F[i]
    with np.errstate(divide="ignore", invalid="ignore"):
        x = np.where(dx > 0, t[i] + (y - F[i]) / dx * (t[i + 1] - t[i]), 0.0)
    inside = (y >= F[0]) & (y <= F[n_inv - 1])
    return int(np.count_nonzero(inside & ~((t[i] <= x) & (x < t[i + 1]))))


@pytest.mark.parametrize("case", ["guided", "guided_and_two_knots"])
def test_nearest_knot_guides_vs_oracle(flat_g0, case):
    tabs = []
    t0, f0 = _adversarial_knots(7)
    om0, m0, F0 = _nearest(t0, f0, tabs)
    if case == "guided":
        t1, f1 = _adversarial_knots(8)
    else:
        t1, f1 = np.array([0.0, T]), np.array([1.0, 0.5])
    om1, m1, F1 = _nearest(t1, f1, tabs)
    assert om0.c.n_inv < om0.c.n_pts
    l0, l1 = _edge_points(F0), _edge_points(F1)
    n_search = _time_search_points(om0, l0)
    assert n_search > 0
    rng = np.random.default_rng(3)
    n = max(len(l0), len(l1))
    l = np.stack([rng.permutation(np.resize(l0, n)), rng.permutation(np.resize(l1, n))], axis=1)

    ctx = U.charge_ctx(flat_g0, 2, T)
    ctx.set_warper_chain([capi.WarperDesc(capi.KD_WARP_PRODUCT, 2, 0.0, 0)], [m0, m1], np.array(tabs))
    got, dom, jac = ctx.eval_warped_batch(l)

    train = O.Train([O.product([om0, om1])])
    f = O.charge_integrand(U.oracle_charge(flat_g0), T)
    res = [O.eval_warped(f, train, li) for li in l]
    ref = np.array([r[0] for r in res])
    rdom = np.array([r[1] for r in res])
    rjac = np.array([train.reverse(li)[1] for li in l])
    # y = 0 on the last inverse knot maps to the first zero interval of f: 1/f = inf there
    fin = np.isfinite(rjac)
    assert np.array_equal(np.isfinite(jac), fin) and np.count_nonzero(fin) > 0.99 * len(l)
    print(f"{case}: {len(l)} points ({len(l) - np.count_nonzero(fin)} with 1/f = inf, {n_search} of axis 0's "
          f"{len(l0)} with the time search), "
          f"max |jac - oracle| / |oracle| {np.max(np.abs(jac[fin] - rjac[fin]) / np.abs(rjac[fin])):.2e}")
    assert np.array_equal(dom, rdom)
    np.testing.assert_allclose(jac[fin], rjac[fin], rtol=1e-14, atol=0)
    err = np.max(np.abs(got[fin] - ref[fin])) / np.max(np.abs(ref[fin]))
    assert err <= 1e-10
