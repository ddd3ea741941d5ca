"""warper_projection_t (c++/keldy/warpers/projection.hpp:176-382): the host
construction against the reference's own known answers and the numpy
restatement in oracle/projection_oracle.py, the smoothing-width search against
the cost it minimises, and (GPU) the device application in a warper chain.

Known answers: projection_test.cpp:76-124 (twelve map values of a 100-bin,
100-sample projection at sigma = 0.0675, tolerance 5e-6) and :126-165 (the
projection and make_warper_from_proj agree under l -> 1 - l, tolerance 1e-3).
The twelve values are those of the pilot |sin(3(1 - l))| on [0, 1]: the
reference test builds its pilot through warper_product_1d_simple_interp_nearest_t
of a constant and ui_from_vi, and the values it states are reproduced (within
1.7e-6) by this construction with that reflected pilot, not with sin(3 l).
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import projection_oracle as P

SIGMA = 0.06754688329
REF_VALUES = [  # projection_test.cpp:112-123
    ("jacobian_forward", 0.3, 0.841834), ("jacobian_forward", 0.5, 0.977436),
    ("jacobian_forward", 0.9, 0.237891), ("ui_from_li", 0.3, 0.341224),
    ("ui_from_li", 0.5, 0.47522), ("ui_from_li", 0.9, 0.77635),
    ("li_from_ui", 0.3, 0.243646), ("li_from_ui", 0.5, 0.538089),
    ("li_from_ui", 0.9, 0.98085), ("jacobian_reverse", 0.3, 0.709755),
    ("jacobian_reverse", 0.5, 0.650334), ("jacobian_reverse", 0.9, 1.08937),
]


def _pilot(l):
    return complex(np.prod(np.sin(3.0 * (1.0 - np.asarray(l)))))


def _get(w, name, x):
    r = getattr(w, name)([x])
    return float(r[0]) if np.ndim(r) else float(r)


def test_reference_values_host_and_oracle():
    from keldy_amd import warpers as kw
    w = kw.WarperProjectionT(_pilot, 1, 100, 100, SIGMA, False)
    o = P.Projection(_pilot, 1, 100, 100, SIGMA)
    for name, x, ref in REF_VALUES:
        assert abs(_get(w, name, x) - ref) <= 5e-6, (name, x)
        assert abs(_get(o, name, x) - ref) <= 5e-6, (name, x)


@pytest.mark.parametrize("plant", ["unit_bins", "linear_smooth", "left_rect", "half_sigma"])
def test_oracle_planted_mistakes_miss_reference_values(plant):
    """Each plausible slip in the restatement moves some reference value by
    more than the 5e-6 tolerance: the known answers pin these steps."""
    o = P.Projection(_pilot, 1, 100, 100, SIGMA, plant=plant)
    assert max(abs(_get(o, n, x) - r) for n, x, r in REF_VALUES) > 5e-6


def test_oracle_akima_and_smoothing_vs_library():
    """The oracle's pieces against independent library routines: its Akima
    spline (GSL's end extension of the slopes) equals scipy's
    Akima1DInterpolator, which extends the slopes the same way, on ragged
    knots and is exact on a quadratic (the extension continues its linearly
    growing slopes); flat end slopes fail both.  Its local linear smoothing
    equals numpy's weighted polyfit over each kernel window."""
    from scipy.interpolate import Akima1DInterpolator
    rng = np.random.default_rng(0)
    x = np.cumsum(rng.uniform(0.1, 1.0, 30))
    y = np.sin(x) + 0.1 * rng.standard_normal(30)
    xs = np.linspace(x[0], x[-1], 500)
    ref = Akima1DInterpolator(x, y)(xs)
    got = [P.akima_eval(x, y, P.akima_coeffs(x, y), v) for v in xs]
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-14)
    bad = [P.akima_eval(x, y, P.akima_coeffs(x, y, flat_ends=True), v) for v in xs]
    assert np.max(np.abs(np.asarray(bad) - ref)) > 1e-3
    xu = np.linspace(0.0, 1.0, 11)
    q = lambda v: 3 * v * v - v + 0.5
    co = P.akima_coeffs(xu, q(xu))
    assert max(abs(P.akima_eval(xu, q(xu), co, v) - q(v)) for v in np.linspace(0, 1, 101)) < 1e-14
    co = P.akima_coeffs(xu, q(xu), flat_ends=True)
    assert max(abs(P.akima_eval(xu, q(xu), co, v) - q(v)) for v in np.linspace(0, 1, 101)) > 1e-4
    yv = np.exp(np.sin(np.arange(25) / 4.0)) * (1 + 0.1 * rng.standard_normal(25))
    yv[7] = 0.0  # a dead bin: weight zero
    s = 1.7
    H = int(3 * s)
    out = P.kernel_smoothing(yv, s)
    for i in range(25):
        j = np.array([k for k in range(max(0, i - H), min(25, i + H + 1)) if yv[k] > 0])
        w = np.exp(-((j - i) / s) ** 2 / 2)
        c1, c0 = np.polyfit(j.astype(float), np.log(yv[j]), 1, w=np.sqrt(w))
        assert abs(out[i] - np.exp(c0 + c1 * i)) <= 1e-12 * out[i]


def test_projection_and_product_maker_reflect():
    """projection_test.cpp:126-165: the projection warper and the product of
    1d maps fitted by make_warper_from_proj to the same pilot agree under
    l -> 1 - l (the one's CDF runs from 0, the other's from the far end)."""
    from keldy_amd import warpers as kw
    w1 = kw.WarperProjectionT(_pilot, 1, 5000, 100000, SIGMA, False)
    w2 = kw.make_warper_from_proj(_pilot, 1, 5000, 100000, SIGMA)
    for x in (0.3, 0.5, 0.9):
        assert abs(w1.ui_from_li([x])[0] - w2.ui_from_li([1 - x])[0]) <= 1e-3
        assert abs(w1.li_from_ui([x])[0] - (1 - w2.li_from_ui([x])[0])) <= 1e-3
        assert abs(w1.jacobian_reverse([x]) - w2.jacobian_reverse([1 - x])) <= 1e-3


def _samples(order, n, seed):
    """Pilot samples with structure, noise and a dead region (zero values
    below l = 0.04 on axis 0: bins with no positive data get zero weight)."""
    rng = np.random.default_rng(seed)
    l = O.Sobol(order, 0).points(n)
    v = np.prod(np.exp(-3 * l) * (1 + 0.3 * np.sin(7 * l)), axis=1) * (1 + 0.2 * rng.standard_normal(n))
    v = np.abs(v)
    v[l[:, 0] < 0.04] = 0.0
    return l, v


def test_host_tables_and_maps_vs_oracle():
    from keldy_amd import warpers as kw
    order, nb, n = 3, 40, 3000
    l, v = _samples(order, n, 1)
    w = kw.WarperProjectionT.from_samples(l, list(v), order, nb, 0.1, False)
    o = P.Projection(None, order, nb, n, 0.1, samples=(l, v))
    for a in range(order):
        np.testing.assert_allclose(w.get_xi(a), o.xi[a], rtol=1e-13, atol=0)
        np.testing.assert_allclose(w.get_fi(a), o.fn[a], rtol=1e-12, atol=0)
        np.testing.assert_allclose(w.get_fi_integrated(a), o.F[a], rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(w.get_fi_integrated_inverse(a), o.Finv[a], rtol=1e-12, atol=1e-15)
        assert np.all(np.diff(o.F[a]) > 0) and o.F[a][0] == 0.0 and abs(o.F[a][-1] - 1) < 1e-15
        counts = np.asarray(w.get_xi_counts(a))
        assert counts.sum() == n
    pts = O.Sobol(order, 0, True, True, 5).points(64)
    for p in pts:
        u, j = w.map_reverse(list(p))
        np.testing.assert_allclose(u, o.ui_from_li(p), rtol=1e-12, atol=1e-15)
        assert abs(j - o.jacobian_reverse(p)) <= 1e-12 * abs(j)
        assert abs(w.jacobian_forward(list(p)) - o.jacobian_forward(p)) <= 1e-12 * abs(w.jacobian_forward(list(p)))
        np.testing.assert_allclose(w.li_from_ui(list(p)), o.li_from_ui(p), rtol=1e-12, atol=1e-15)
        # the tabulated inverse CDF inverts the tabulated CDF up to the interpolation error
        np.testing.assert_allclose(w.li_from_ui(list(u)), p, atol=2e-3)


def test_optimized_sigma_minimises_leave_one_out_cost():
    """optimize_sigma (projection.hpp:82-127) is a bracketed Brent search from
    the given width: it returns a local minimum of the leave-one-out cost
    (none smaller on a fine grid around it, the cost grows either side) that
    is below the cost at the starting width, or the 0.3-bin floor."""
    from keldy_amd import warpers as kw
    order, nb = 2, 60
    l, v = _samples(order, 4000, 3)
    v[l[:, 0] < 0.04] = 1e-3  # keep every bin alive
    w = kw.WarperProjectionT.from_samples(l, list(v), order, nb, 0.1, True)
    size = (1 + 1 / (nb - 1)) / nb
    up = 1 / size
    s0 = 0.1 / size
    found = []
    for a in range(order):
        xi = np.asarray(w.get_xi(a))
        s = w.get_sigmas()[a] / size
        c = P.loo_cost(xi, s, up)
        assert c < P.loo_cost(xi, s0, up)
        if s > 0.3:
            grid = np.linspace(s / 1.5, min(s * 1.5, up), 201)
            assert c <= min(P.loo_cost(xi, g, up) for g in grid) * (1 + 1e-9)
            assert P.loo_cost(xi, s * 0.999, up) >= c and P.loo_cost(xi, s * 1.001, up) >= c
        else:
            assert s == 0.3 and P.loo_cost(xi, 0.31, up) > c
        found.append(s)
    assert found[0] == 0.3 and 2.0 < found[1] < 3.5  # one floor and one interior minimum
    # the smoothed tables follow from the found width as from a fixed one
    wf = kw.WarperProjectionT.from_samples(l, list(v), order, nb, w.get_sigmas()[1], False)
    np.testing.assert_allclose(wf.get_fi(1), w.get_fi(1), rtol=1e-12)


def test_train_composition_and_errors():
    from keldy_amd import warpers as kw
    t = 10.0
    l, v = _samples(3, 2000, 4)
    w = kw.WarperProjectionT.from_samples(l, list(v), 3, 30, 0.1, False)
    tr = kw.WarperTrainT()
    tr.emplace_back(kw.WarperPlasmaUvT(t))
    tr.emplace_back(w)
    p = [0.2, 0.7, 0.45]
    u, j = tr.map_reverse(p)
    x, j1 = w.map_reverse(p)
    u_ref, j2 = O.Train([O.plasma(t)]).reverse(np.asarray(x))
    np.testing.assert_allclose(u, u_ref, rtol=1e-13, atol=1e-14)
    assert abs(j - j1 * j2) <= 1e-13 * abs(j)
    with pytest.raises(Exception, match="axis"):
        w.get_fi(3)
    with pytest.raises(Exception, match="outside"):
        w.ui_from_li([1.5, 0.1, 0.1])
    with pytest.raises(Exception):
        kw.WarperProjectionT.from_samples(l, list(v), 3, 2, 0.1, False)  # fewer than 3 bins


@pytest.mark.gpu
def test_device_chain_with_projection_vs_oracle():
    """The projection fitted on the device (pilot = the integrator's warped
    integrand at the first unshifted Sobol points) and applied by the
    Keldysh-sum launch as the last warper of [plasma_uv(t), projection]:
    tables against the oracle's construction from the same pilot values,
    values and Jacobians against the oracle's chain at 1e-10 normwise."""
    from keldy_amd import common as kc
    from keldy_amd import impurity_oneband as kio
    from keldy_amd import warpers as kw
    params = dict(beta=-1.0, bias_V=0.0, eps_d=0.0, Gamma=1.0, time_max=30.0, nr_time_points_gf=10001,
                  alpha=0.0, bath_type="flatband", ft_method="analytic")
    model = kio.G0Model(kio.G0ModelOmega(params), False)
    t, order, nb, ns = 10.0, 4, 50, 4096
    c = kio.ComputeChargeQDirect(model, t, order, 0.0)
    c.warper.emplace_back(kw.WarperPlasmaUvT(t))
    proj = kw.WarperProjectionT(c, order, nb, ns, 0.1, False)
    g = kc.Sobol(order, 0)
    lp = np.array([g() for _ in range(ns)])
    pv, _, _ = c.evaluate_warped_integrand_batch(lp, False)
    o = P.Projection(None, order, nb, ns, 0.1, samples=(lp, np.abs(pv)))
    for a in range(order):
        np.testing.assert_allclose(proj.get_fi_integrated_inverse(a), o.Finv[a], rtol=1e-12, atol=1e-15)
    c.warper.emplace_back(proj)
    l = O.Sobol(order, 0, True, True, 7).points(128)
    got, dom, jac = c.evaluate_warped_integrand_batch(l)
    g0 = O.G0Table(model.g0_lesser[0, 0], model.g0_greater[0, 0], model.time_max)
    f = O.charge_integrand(g0, t)
    plasma = O.Train([O.plasma(t)])
    ref, rdom, rjac = [], [], []
    for p in l:
        x = o.ui_from_li(p)
        j1 = o.jacobian_reverse(p)
        v, d = O.eval_warped(f, plasma, x)
        ref.append(v * j1)
        rdom.append(d)
        rjac.append(j1 * plasma.reverse(x)[1])
    ref = np.array(ref)
    assert np.array_equal(dom, np.array(rdom)) and np.any(dom == 1)
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))
    np.testing.assert_allclose(jac, rjac, rtol=1e-11)
    c.run(4096)
    assert c.reduce_nr_points_run() == 4096
