"""Host-side parts of keldy_amd that run without a GPU: the C-ABI library's
exports, Sobol set-up, g0 tabulation, warpers and the pilot fit, each checked
against the oracle, the golden vectors or the reference's known answers."""
import os
import re

import numpy as np
import pytest

from oracle import g0_oracle
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_capi_exports_every_declared_symbol():
    from keldy_amd import capi
    hdr = open(os.path.join(ROOT, "include", "keldy_amd.h")).read()
    declared = set(re.findall(r"^(?:int|const char \*)\s*(kd_\w+)\s*\(", hdr, re.M))
    assert declared == set(capi.SYMBOLS)
    lib = capi.lib()
    for s in declared:
        assert hasattr(lib, s), s


def test_host_sobol_matches_reference_header():
    from keldy_amd.common import Sobol
    g = np.load(os.path.join(ROOT, "tests", "golden", "sobol_ref.npz"))
    for i in range(len([k for k in g.files if k.startswith("case")])):
        dim, seed, sh, sc, rs, n = (int(v) for v in g[f"case{i}"])
        s = Sobol(dim, seed, bool(sh), bool(sc), rs)
        pts = np.array([s() for _ in range(n)])
        assert np.array_equal(pts, g[f"pts{i}"]), f"case {i}"
    last = Sobol(2, 2**31 - 1)
    last()  # k = 2^31, the last point
    with pytest.raises(RuntimeError, match="More sample points"):
        last()


def test_host_sobol_tables_match_oracle():
    from keldy_amd.common import Sobol
    for args in ((5, 0, True, True, 3), (8, 7, False, True, 1), (3, 0, True, False, 0)):
        Csr, sh = Sobol(*args).generating_tables()
        oC, osh = O.Sobol(*args).tables()
        assert np.array_equal(np.array(Csr).reshape(args[0], 31), oC) and np.array_equal(sh, osh)


def test_compare_3way():
    from keldy_amd.common import ContourPtT, compare_3way

    def cp(t, k, s=0):
        c = ContourPtT()
        c.time, c.k_idx, c.timesplit_n = t, k, s
        return c
    assert compare_3way(cp(1.0, 0), cp(2.0, 0)) < 0
    assert compare_3way(cp(2.0, 1), cp(1.0, 1)) < 0
    assert compare_3way(cp(5.0, 0), cp(-5.0, 1)) < 0
    assert compare_3way(cp(1.0, 0, 0), cp(1.0, 0, 1)) < 0
    assert compare_3way(cp(1.0, 1, 0), cp(1.0, 1, 1)) > 0
    assert compare_3way(cp(1.0, 1, 2), cp(1.0, 1, 2)) == 0


# --------------------------------------------------------------------------
def _model(**kw):
    from keldy_amd import impurity_oneband as kio
    p = dict(beta=-1.0, bias_V=0.0, eps_d=0.0, Gamma=1.0, alpha=0.0, half_bandwidth=2.0,
             bath_type="flatband", time_max=10.0, nr_time_points_gf=10001, ft_method="analytic")
    p.update(kw)
    return kio.G0Model(kio.G0ModelOmega(p), False), p


def test_g0_analytic_vs_oracle():
    m, _ = _model()
    t, les, gre = g0_oracle.flatband_analytic(10.0, 10000, 1.0)
    np.testing.assert_allclose(m.time_mesh, t, rtol=0, atol=1e-13)
    assert np.max(np.abs(m.g0_lesser[0, 0] - les)) < 5e-15
    assert np.max(np.abs(m.g0_greater[0, 0] - gre)) < 5e-15


@pytest.mark.parametrize("kw", [dict(ft_method="contour"),
                                dict(ft_method="contour", eps_d=0.5, bias_V=0.2, beta=50.0),
                                dict(ft_method="contour", eps_d=-1.0, bias_V=2.0, beta=10.0, Gamma=1.5),
                                dict(ft_method="contour", bath_type="lorentzian", half_bandwidth=3.0, beta=20.0,
                                     bias_V=0.4)])
def test_g0_contour_vs_real_axis_oracle(kw):
    """contour tabulation (model.cpp:326-426) against an independent real-axis
    Fourier quadrature of the same g0(omega)."""
    m, p = _model(nr_time_points_gf=2001, **kw)
    mesh = m.time_mesh
    for k in (1000, 1001, 1100, 1500, 2000, 0, 400):
        t = mesh[k]
        assert abs(m.g0_lesser[0, 0, k] - g0_oracle.g0_real_axis(t, True, p)) < 1e-9, t
        assert abs(m.g0_greater[0, 0, k] - g0_oracle.g0_real_axis(t, False, p)) < 1e-9, t


def test_g0_semicircle_finite_contour_vs_oracle():
    m, p = _model(nr_time_points_gf=801, ft_method="contour", bath_type="semicircle", half_bandwidth=2.0,
                  eps_d=-0.5, bias_V=0.6, beta=-1.0)
    for k in (400, 401, 500, 800, 100):
        t = m.time_mesh[k]
        assert abs(m.g0_lesser[0, 0, k] - g0_oracle.g0_real_axis(t, True, p, epsabs=1e-12)) < 1e-8, t


@pytest.mark.parametrize("eps,ref", [(1.0, -0.09907315 + 0.06406786j), (-1.0, -0.2868307 + 0.05648988j)])
def test_g0_contour_reference_values(eps, ref):
    """model_flatband_test.cpp:170-214, tolerance 1e-7 as in the reference."""
    m, _ = _model(eps_d=eps, Gamma=1.5, ft_method="contour")
    assert abs(m.lesser_at(1.0) - ref) < 1e-7


def test_g0_symmetry_and_errors():
    from keldy_amd import impurity_oneband as kio
    m, _ = _model(ft_method="contour", eps_d=0.3, beta=5.0, bias_V=0.5)
    les = m.g0_lesser[0, 0]
    np.testing.assert_allclose(les, -np.conj(les[::-1]), rtol=0, atol=1e-15)  # model_flatband_test.cpp:235-238
    with pytest.raises(RuntimeError, match="Analytic flatband covers only"):
        _model(eps_d=0.2)
    with pytest.raises(RuntimeError, match="bath_type not defined"):
        kio.G0ModelOmega(dict(bath_type="cubic"))
    with pytest.raises(RuntimeError, match="singular"):
        kio.G0Model(kio.G0ModelOmega(dict(bath_type="flatband", ft_method="contour")), True)
    with pytest.raises(RuntimeError, match="not defined"):
        kio.G0Model(kio.G0ModelOmega(dict(bath_type="flatband", ft_method="laplace", time_max=1.0)), False)
    with pytest.raises(RuntimeError, match="unknown key"):
        kio.G0ModelOmega(dict(Gama=1.0))


def test_g0_keldysh_contour_adapter():
    """model.hpp:244-265 against the oracle's adapter, incl. alpha on equal points."""
    import ctypes
    from keldy_amd import impurity_oneband as kio
    m, _ = _model(alpha=0.3)
    g = kio.G0KeldyshContourT(m)
    og = O.G0Table(m.g0_lesser[0, 0], m.g0_greater[0, 0], 10.0, alpha=0.3)
    rng = np.random.default_rng(3)
    for _ in range(50):
        a = (rng.uniform(0, 10), 0, int(rng.integers(2)), int(rng.integers(3)))
        b = (rng.uniform(0, 10), 0, int(rng.integers(2)), int(rng.integers(3)))
        if rng.random() < 0.3:
            b = a
        internal = bool(rng.integers(2))
        v = g(kio.GfIndexT(*a), kio.GfIndexT(*b), internal)
        o = np.zeros(2)
        O.lib().ko_g0_eval(ctypes.byref(og.c), ctypes.byref(O.gf_index(*a)), ctypes.byref(O.gf_index(*b)),
                           int(internal), O._dp(o))
        assert abs(v - (o[0] + 1j * o[1])) <= 1e-15
    assert g(kio.GfIndexT(1.0, 0, 0), kio.GfIndexT(1.0, 1, 0)) == 0  # spin diagonal


# --------------------------------------------------------------------------
def _check_warper_1d(w, u_max, acc, negative):
    """The shape checks of warper_tests_common.hpp:38-150 (order 1)."""
    lst = [1e-9, 1e-4, 1e-3, 0.1, 0.2, 0.5, 0.7, 0.999, 0.9999, 1 - 1e-9]
    ls = list(lst)
    us = [x * u_max for x in lst]
    lo, hi = (1.0, 0.0) if negative else (0.0, 1.0)
    assert abs(w.li_from_ui([0.0])[0] - lo) < acc and abs(w.li_from_ui([u_max])[0] - hi) < acc
    prev = None
    for u in (reversed(us) if negative else us):
        l = w.li_from_ui([u])[0]
        assert -acc <= l <= 1 + acc
        if prev is not None:
            assert l >= prev
        prev = l
    prev = None
    for l in (reversed(ls) if negative else ls):
        u = w.ui_from_li([l])[0]
        assert -acc <= u <= u_max + acc
        if prev is not None:
            assert u > prev
        prev = u
    for u in us:
        assert abs(w.ui_from_li(w.li_from_ui([u]))[0] - u) < acc * max(1, u_max)
    for l in ls:
        assert abs(w.li_from_ui(w.ui_from_li([l]))[0] - l) < acc
        ui, jac = w.map_reverse([l])
        assert ui[0] == w.ui_from_li([l])[0] and jac == w.jacobian_reverse([l])


@pytest.mark.parametrize("name", ["exp", "inv", "inv_sq", "inv_pow2", "inv_pow3", "uniform", "nearest", "hybrid"])
def test_host_warpers_shape_and_oracle_parity(name):
    from keldy_amd import warpers as kw
    um = 10.0
    if name == "exp":
        w, ow, neg, acc = kw.make_product_1d_simple_exponential(um, 1.5), O.Map1d("exponential", um, w=1.5), True, 1e-9
    elif name == "inv":
        w, ow, neg, acc = kw.make_product_1d_simple_inverse(um, 1.5), O.Map1d("inverse", um, w=1.5), False, 1e-10
    elif name == "inv_sq":
        w, ow, neg, acc = (kw.make_product_1d_simple_inverse_square(um, 1.5), O.Map1d("inverse_square", um, w=1.5),
                           False, 1e-10)
    elif name.startswith("inv_pow"):
        pw = int(name[-1])
        w, ow, neg, acc = (kw.make_product_1d_simple_inverse_power(pw, um, 1.5),
                           O.Map1d("inverse_power", um, w=1.5, power=pw), False, 1e-9)
    elif name == "uniform":
        w, ow, neg, acc = kw.make_product_1d_simple_uniform(um), O.Map1d("uniform", um), False, 1e-14
    elif name == "nearest":
        w = kw.WarperProduct1dSimpleInterpNearestT(lambda u: np.exp(-u / 2.0), um, 1000)
        t, f1, F, cut = w.knots()
        ow, neg, acc = O.Map1d("nearest", um, t=np.array(t), f=np.array(f1), F=np.array(F), n_inv=cut + 1), True, 1e-9
    else:
        w = kw.WarperProduct1dSimpleInterpHybridT(lambda t: 1.0 * 1.0 * 1.0 / ((1.0 + t) * (1.0 + t) * (1.0 + t)),
                                                  um, 2000)
        # separate forward and inverse Steffen splines are inverse to ~1e-8 here
        ow, neg, acc = O.hybrid_map(3, 1.0, um, 2000), True, 1e-7
    _check_warper_1d(w, um, acc, neg)
    rng = np.random.default_rng(7)
    for l in rng.uniform(0, 1, 40):
        u_host, j_host = w.map_reverse([l])
        u_or, jr = ow.reverse(l)
        assert abs(u_host[0] - u_or) <= 1e-13 * max(1.0, abs(u_or))
        assert abs(j_host - jr) <= 1e-13 * abs(jr)


def test_product_1d_and_train_vs_oracle():
    from keldy_amd import warpers as kw
    t = 10.0
    tr = kw.WarperTrainT()
    tr.emplace_back(kw.WarperPlasmaUvT(t))
    tr.emplace_back(kw.make_product_1d_inverse_cube_alternate(4, t, 1.2))
    otr = O.Train([O.plasma(t), O.product([O.Map1d("inverse", t, w=1.2), O.Map1d("inverse_power", t, w=1.2, power=3),
                                           O.Map1d("inverse", t, w=1.2), O.Map1d("inverse_power", t, w=1.2, power=3)])])
    for l in O.Sobol(4, 0, True, True, 4).points(50):
        (u, j), (uo, jo) = tr.map_reverse(list(l)), otr.reverse(l)
        np.testing.assert_allclose(u, uo, rtol=1e-14, atol=1e-14)
        assert abs(j - jo) <= 1e-13 * abs(jo)
        # forward of reverse is the identity
        lf, jf = tr.map_forward(u)
        np.testing.assert_allclose(lf, l, atol=1e-10)
        assert abs(jf * j - 1) < 1e-9
    assert kw.vi_from_ui(10.0, [0.0, 2.0, 6.0, 1.0]) == [4.0, 4.0, 1.0, 1.0]  # warper_plasma_uv_test.cpp:35
    assert kw.ui_from_vi(10.0, [4.0, 4.0, 1.0, 1.0]) == [6.0, 2.0, 1.0, 0.0]


def test_callable_warper_domain_checks_and_device_refusal():
    from keldy_amd import warpers as kw
    tm = 1.5
    w = kw.WarperProduct1dSimpleT(lambda u: 1 / tm, lambda u: u / tm, lambda l: tm * l, tm)
    assert w.ui_from_li([0.5]) == [0.75] and w.jacobian_reverse([0.2, 0.3]) == pytest.approx(tm ** 2)
    with pytest.raises(RuntimeError, match="do not map domain boundaries"):
        kw.WarperProduct1dSimpleT(lambda u: 1.0, lambda u: 2 * u, lambda l: l, 1.0)


# --------------------------------------------------------------------------
def test_kernel_smoothing_vs_weighted_least_squares():
    """make_warper_from_proj.hpp:52-78: log-space local-linear regression with a
    Gaussian kernel, dead (<= 0) points excluded -- checked against numpy's
    weighted polynomial fit of each window."""
    from keldy_amd import warpers as kw
    rng = np.random.default_rng(5)
    y = np.exp(-np.linspace(0, 3, 60)) * (1 + 0.1 * rng.random(60))
    y[[7, 30, 31]] = 0.0
    sigma = 4.0
    out = np.array(kw.kernel_smoothing(list(y), sigma))
    N, H = len(y), max(min(int(3 * sigma), len(y)), 1)
    x = np.arange(N, dtype=float)
    for i in range(N):
        j = np.arange(max(0, i - H), min(N, i + H + 1))
        w = np.exp(-((j - i) / sigma) ** 2 / 2)
        keep = y[j] > 0
        c1, c0 = np.polyfit(x[j][keep], np.log(y[j][keep]), 1, w=np.sqrt(w[keep]))
        assert abs(out[i] - np.exp(c0 + c1 * i)) <= 1e-10 * out[i]


def test_pilot_fit_from_samples():
    from keldy_amd import warpers as kw
    from keldy_amd.common import Sobol
    g = Sobol(2, 0)
    l = np.array([g() for _ in range(2048)])
    vals = np.exp(-3 * l[:, 0]) * (1 + l[:, 1])
    w = kw.make_warper_from_proj_samples(l, vals, 2, 20, 0.1)
    assert w.size() == 2
    # the fitted bin densities 0.5 (f1[i] + f1[i+1]) along axis 0 decay like
    # exp(-3 l) (the knot values themselves alternate, make_warper_from_proj.hpp:146-148)
    t, f1, F, cut = w.get_1d_warper(0).knots()
    f1 = np.array(f1)
    dens = 0.5 * (f1[:-1] + f1[1:])
    assert dens[1] > dens[10] > dens[18] > 0
    assert abs(np.log(dens[2] / dens[17]) - 3 * (t[17] - t[2])) < 0.3


def test_dot_lead_g0_semicircle_kats():
    """model_semicircle_test.cpp Semicirc_contour: dot-dot, dot-lead and
    lead-dot g^<(t=1) at 1e-3 (values from ctint_keldysh), and the t <-> -t
    symmetries g^{<,>}_{ab}(t) = -conj(g^{<,>}_{ba}(-t)) (1e-8)."""
    kio = pytest.importorskip("keldy_amd.impurity_oneband")
    p = dict(beta=10.0, bias_V=2.0, eps_d=-1.0, Gamma=1.5, time_max=100.0, nr_time_points_gf=1000, alpha=0.0,
             half_bandwidth=2.0, bath_type="semicircle", ft_method="contour")
    m = kio.G0Model(kio.G0ModelOmega(p), True)
    tm = m.time_mesh
    k, km = int(np.argmin(abs(tm - 1.0))), int(np.argmin(abs(tm + 1.0)))
    assert tm[k] == pytest.approx(1.0, abs=1e-12)
    L, G = m.g0_lesser, m.g0_greater
    assert abs(L[0, 0, k] - (-0.23107349410223865 + 0.13213346415985422j)) < 1e-3
    assert abs(L[0, 1, k] - (0.17154101374667907 + 0.13375649387573293j)) < 1e-3
    assert abs(L[1, 0, k] - (-0.0066645142937036958 + 0.12215571972575601j)) < 1e-3
    for T in (L, G):
        assert abs(T[0, 0, k] + np.conj(T[0, 0, km])) < 1e-8
        assert abs(T[0, 1, k] + np.conj(T[1, 0, km])) < 1e-8
