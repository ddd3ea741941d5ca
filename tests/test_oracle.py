"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each check is independent of the oracle's own formulas: golden vectors from the
reference's compiled Sobol header (tests/golden/sobol_ref.npz), the C++
standard's known answer for std::mt19937_64, the reference tests' known values,
closed forms and brute force on small inputs.
"""
import itertools
import math
import os

import numpy as np
import pytest

from oracle import g0_oracle
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "sobol_ref.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def test_mt19937_64_known_answer(gold):
    # C++ standard [rand.predef]: the 10000th output of default-seeded mt19937_64
    assert O.mt64(5489, 10000)[-1] == 9981545732273789042
    assert O.mt64(0, 16) == [int(x) for x in gold["mt64_0"]]
    assert O.mt64(5489, 16) == [int(x) for x in gold["mt64_5489"]]


def test_sobol_matches_reference_header(gold):
    n_cases = len([k for k in gold.files if k.startswith("case")])
    for i in range(n_cases):
        dim, seed, sh, sc, rs, n = (int(v) for v in gold[f"case{i}"])
        pts = O.Sobol(dim, seed, bool(sh), bool(sc), rs).points(n)
        assert np.array_equal(pts, gold[f"pts{i}"]), f"case {i}"


def test_sobol_first_dims_are_textbook():
    # dim 1: van der Corput in Gray-code order; dim 2: the classic second Sobol
    # dimension with direction numbers m = 1, 3, 5, 15, 17, ...
    Cs = O.jk2008(2, 8)
    assert list(Cs[0]) == [1 << k for k in range(8)]
    assert list(Cs[1]) == [1, 3, 5, 15, 17, 51, 85, 255]
    pts = O.Sobol(1, 0).points(7)[:, 0]
    np.testing.assert_array_equal(pts, [0.5, 0.75, 0.25, 0.375, 0.875, 0.625, 0.125])


def test_sobol_exhaustion():
    s = O.Sobol(2, 2**31 - 1)
    s.next()  # k = 2^31 is the last point
    with pytest.raises(RuntimeError):
        s.next()


# --------------------------------------------------------------------------
def _cp(t, k, s=0):
    return O.gf_index(t, 0, k, s)


def test_compare_3way_reference_cases():
    """k_countour_pt.cpp: forward before backward; forward ordered by time,
    backward anti-ordered; equal points split by timesplit."""
    c3 = lambda a, b: O.lib().ko_compare_3way(a, b)
    import ctypes
    for t in (-1.0, 0.0, 1.0):
        for s in (-1, 0, 1):
            for k in (0, 1):
                a, b = _cp(t, k, s), _cp(t, k, s)
                assert c3(ctypes.byref(a), ctypes.byref(b)) == 0
    for t0, t1 in itertools.product((-2.0, 0.0, 2.0), repeat=2):
        a, b = _cp(t0, 0, 1), _cp(t1, 1, -1)
        assert c3(ctypes.byref(a), ctypes.byref(b)) < 0 < c3(ctypes.byref(b), ctypes.byref(a))
    a, b = _cp(1.0, 0), _cp(2.0, 0)
    assert c3(ctypes.byref(a), ctypes.byref(b)) < 0
    a, b = _cp(2.0, 1), _cp(1.0, 1)
    assert c3(ctypes.byref(a), ctypes.byref(b)) < 0
    a, b = _cp(1.0, 0, 0), _cp(1.0, 0, 1)
    assert c3(ctypes.byref(a), ctypes.byref(b)) < 0
    a, b = _cp(1.0, 1, 0), _cp(1.0, 1, 1)
    assert c3(ctypes.byref(a), ctypes.byref(b)) > 0


# --------------------------------------------------------------------------
def test_g0_flatband_analytic_known_values():
    # g^<(0) = i/2; g^> = conj(g^<); large-t tail Re g^< ~ -1/(pi Gamma t)
    g = g0_oracle.flatband_analytic_lesser(np.array([0.0, 50.0]), 1.0)
    assert g[0] == 0.5j
    assert abs(g[1].real + 1 / (np.pi * 50.0)) < 2e-5
    # odd real part, even imaginary part
    t = np.linspace(0.1, 9.9, 17)
    gp, gm = g0_oracle.flatband_analytic_lesser(t, 1.3), g0_oracle.flatband_analytic_lesser(-t, 1.3)
    np.testing.assert_allclose(gp.real, -gm.real, rtol=1e-14)
    np.testing.assert_allclose(gp.imag, gm.imag, rtol=1e-14)


@pytest.mark.parametrize("eps,ref", [(1.0, -0.09907315 + 0.06406786j), (-1.0, -0.2868307 + 0.05648988j)])
def test_g0_real_axis_reference_values(eps, ref):
    """model_flatband_test.cpp:170-214 (T=0, Gamma=1.5), tolerance 1e-7."""
    p = dict(beta=-1, bias_V=0.0, eps_d=eps, Gamma=1.5, half_bandwidth=2.0, bath_type="flatband")
    assert abs(g0_oracle.g0_real_axis(1.0, True, p) - ref) < 1e-7


def test_g0_real_axis_matches_analytic():
    p = dict(beta=-1, bias_V=0.0, eps_d=0.0, Gamma=1.0, half_bandwidth=2.0, bath_type="flatband")
    for t in (0.0, 0.3, 2.0, -1.7, 9.5):
        a = g0_oracle.flatband_analytic_lesser(t, 1.0)
        assert abs(g0_oracle.g0_real_axis(t, True, p) - a) < 1e-10
        assert abs(g0_oracle.g0_real_axis(t, False, p) - np.conj(a)) < 1e-10


# --------------------------------------------------------------------------
def _leibniz_det(m):
    n = m.shape[0]
    tot = 0j
    for perm in itertools.permutations(range(n)):
        inv = sum(1 for i in range(n) for j in range(i + 1, n) if perm[i] > perm[j])
        p = 1
        for i in range(n):
            p *= m[i, perm[i]]
        tot += (-1) ** inv * p
    return tot


def test_det_brute_force():
    rng = np.random.default_rng(1)
    for n in (1, 2, 3, 4, 5):
        for _ in range(5):
            m = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
            assert abs(O.det(m) - _leibniz_det(m)) <= 1e-12 * max(1, abs(_leibniz_det(m)))
    # exact pivoting case: zero leading entry
    m = np.array([[0, 1], [1, 0]], dtype=complex)
    assert O.det(m) == -1


@pytest.fixture(scope="module")
def flat20():
    t, les, gre = g0_oracle.flatband_analytic(30.0, 10000, 1.0)
    return O.G0Table(les, gre, 30.0)


@pytest.mark.parametrize("order", [1, 2])
def test_integrand_matches_explicit_wick(flat20, order):
    """wick_direct_test.cpp:61-206 (orders 1-2 against explicit contractions)."""
    f = O.charge_integrand(flat20, 20.0)
    rng = np.random.default_rng(order)
    for _ in range(20):
        times = rng.uniform(0, 20.0, size=order)
        v, d = f(times)
        w, e = f.eval_no_det(times)
        assert d == e == 1 and abs(v - w) <= 1e-14 * abs(w) + 1e-17


def test_integrand_brute_force_keldysh_sum(flat20):
    """Order 3: the 2^3 signed det*det sum from explicit Leibniz determinants
    of the Wick matrices built with the oracle's g0 adapter."""
    import ctypes
    f = O.charge_integrand(flat20, 20.0)
    n, t = 3, 20.0
    times = [7.5, 2.25, 13.0]
    A, B = O.gf_index(t, 0, 0, n), O.gf_index(t, 0, 1, n)

    def g(a, b, internal=1):
        o = np.zeros(2)
        O.lib().ko_g0_eval(ctypes.byref(flat20.c), ctypes.byref(a), ctypes.byref(b), internal, O._dp(o))
        return o[0] + 1j * o[1]

    tot, scale = 0j, 0.0
    for sig in itertools.product((0, 1), repeat=n):
        pts = [O.gf_index(times[i], 0, sig[i], i) for i in range(n)]
        m1 = np.zeros((n + 1, n + 1), complex)
        for i in range(n):
            for j in range(n):
                m1[i, j] = g(pts[i], pts[j])
            m1[i, n] = g(pts[i], B)
            m1[n, i] = g(A, pts[i])
        m2 = m1[:n, :n].copy()  # opposite spin: the same g0 (spin independent)
        term = (-1) ** sum(sig) * _leibniz_det(m1) * _leibniz_det(m2)
        tot += term
        scale += abs(term)
    v, _ = f(times)
    # the alternating sum cancels strongly: compare on the scale of its terms
    assert abs(v - tot) <= 1e-14 * scale


def test_charge_series_orders_1_to_3(flat20):
    """compute_obs_test.cpp:93-135 at the reference's 1e4 points, 1e-3 rel."""
    hz = [0.5, -0.15915494309189533577, 0.050660591821168885722, 0.0046743460474992114797]
    t = 20.0
    f = O.charge_integrand(flat20, t)
    tr = O.Train([O.plasma(t), O.product(O.Map1d("inverse_square", t, w=1.0))])
    for order in (1, 2, 3):
        s, npts, _ = O.run(f, tr, O.Sobol(order, 0), 10000, threads=4)
        q = -1j * (1j ** order) * s / npts
        assert abs(q - hz[order]) <= 1e-3 * abs(hz[order])
    # Q_1 = -1/(2 pi), Q_2 = 1/(2 pi^2) closed forms (the HZ values)
    assert math.isclose(hz[1], -1 / (2 * math.pi), rel_tol=1e-15)
    assert math.isclose(hz[2], 1 / (2 * math.pi ** 2), rel_tol=1e-15)


# --------------------------------------------------------------------------
def test_plasma_reference_vectors():
    """warper_plasma_uv_test.cpp:15-56"""
    tr = O.Train([O.plasma(10.0)])
    u, j = tr.reverse([4.0, 4.0, 1.0, 1.0])
    np.testing.assert_array_equal(u, [6.0, 2.0, 1.0, 0.0])
    assert j == 1.0
    u, _ = O.Train([O.plasma(1.0)]).reverse([0.2, 0.2, 0.4, 0.1])
    np.testing.assert_allclose(u, [0.8, 0.6, 0.2, 0.1], rtol=1e-15)


@pytest.mark.parametrize("kind,um,w,pw", [("exponential", 5.0, 2.0, 1), ("inverse", 10.0, 1.5, 1),
                                          ("inverse_square", 10.0, 1.5, 1), ("inverse_power", 10.0, 1.5, 3),
                                          ("uniform", 1.5, 1.0, 1)])
def test_maker_maps_closed_forms(kind, um, w, pw):
    """x = F^-1(l) satisfies the maker's CDF and 1/jacobian is its density
    (product_1d_simple_makers_test.cpp, product_1d_simple_test.cpp:10-75)."""
    def pdf(x):
        if kind == "exponential":
            return np.exp(-x / w) / (w * (1 - np.exp(-um / w)))
        if kind == "inverse":
            return 1 / ((w + x) * np.log1p(um / w))
        if kind == "inverse_square":
            return w * (w + um) / (um * (w + x) ** 2)
        if kind == "inverse_power":
            return (pw - 1) * (um + w) ** (pw - 1) / ((um / w + 1) ** (pw - 1) - 1) / (w + x) ** pw
        return 1 / um

    import scipy.integrate as si
    m = O.Map1d(kind, um, w=w, power=pw)
    for l in (1e-9, 0.1, 0.37, 0.5, 0.93, 1 - 1e-9):
        x, jr = m.reverse(l)
        assert 0 <= x <= um
        cdf = si.quad(pdf, 0, x, epsabs=1e-13, epsrel=1e-12)[0]
        # the exponential maker maps decreasingly (l = 1 at u = 0)
        target = 1 - l if kind == "exponential" else l
        assert abs(cdf - target) < 1e-9
        assert abs(1 / jr - pdf(x)) <= 1e-12 * pdf(x)


def test_nearest_map_identity_reference_case():
    """product_1d_simple_interp_test.cpp: constant f on [0, 1.5], l = (t_max - u)/t_max."""
    t = np.array([0., 0.5, 0.75, 1., 1.5])
    m, norm = O.nearest_map(t, np.ones(5))
    for l in (1e-9, 0.2, 0.5, 0.999):
        x, jr = m.reverse(l)
        assert abs(x - 1.5 * (1 - l)) < 1e-14 and abs(jr - 1.5) < 1e-14


def test_hybrid_map_matches_exact_cdf():
    """Steffen-spline CDF of w^3/(w+u)^3 against the exact inverse CDF."""
    w, um = 1.0, 10.0
    m = O.hybrid_map(3, w, um, 2000)
    c = 1 - (w / (w + um)) ** 2
    for l in (0.01, 0.3, 0.77, 0.99):
        x, jr = m.reverse(l)
        # decreasing map: F(u) from u_max; exact (1 - (w/(w+u))^2)/c = 1 - l
        exact = w / np.sqrt(1 - c * (1 - l)) - w
        assert abs(x - exact) < 1e-8 * um
