"""Pins the two oracle functions that no reference test fixes directly: the
g0 table evaluated between mesh points (TRIQS gf<retime> linear evaluation,
call sites model.hpp:253-263) and the Steffen spline of the hybrid map
(product_1d_simple.hpp:381-554, GSL's steffen interpolation).

The checks do not restate the oracle's formulas:
- off-grid g0 against numpy's own linear interpolation of the same table, and
  against the flat-band closed form (model.cpp:444-454) within the linear
  interpolation error bound dt^2/8 max|g''|, g'' from the closed form's
  derivatives;
- the Steffen spline by its defining properties: it interpolates, it is C^1,
  it preserves monotone data, it has zero slope at a data extremum, and on
  data whose neighbouring slopes differ by less than a factor 2 its knot slopes
  are those of the parabola through three points, so it reproduces a
  quadratic exactly on the interior intervals of a non-uniform grid.
Each check is shown to fail on an oracle built with a planted mistake
(keldy_oracle.c, KO_PLANT): an index off by one, swapped interpolation
weights, the limiter without its factor 1/2, and p_i with its weights swapped.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest
import scipy.special as sp

from oracle import g0_oracle
from oracle import oracle as O

SRC = os.path.join(os.path.dirname(O.__file__), "keldy_oracle.c")


@pytest.fixture(scope="module")
def variants(tmp_path_factory):
    """The restatement (0) and the planted mistakes 1-4, each its own library."""
    d = tmp_path_factory.mktemp("oracle_plants")
    libs = {0: O.lib()}
    for k in (1, 2, 3, 4):
        path = str(d / f"libko_plant{k}.so")
        subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-std=gnu11", "-fopenmp", f"-DKO_PLANT={k}", SRC,
                        "-o", path, "-lm"], check=True)
        libs[k] = O.load(path)
    return libs


# ---------------------------------------------------------------- g0 off the mesh
T_MAX, NPTS, GAMMA = 10.0, 10001, 1.0


@pytest.fixture(scope="module")
def flat_table():
    mesh, les, gre = g0_oracle.flatband_analytic(T_MAX, NPTS - 1, GAMMA)
    assert len(mesh) == NPTS
    return mesh, les, gre, O.G0Table(les, gre, T_MAX)


def _offgrid_times(seed, n=400):
    rng = np.random.default_rng(seed)
    dt = 2 * T_MAX / (NPTS - 1)
    k = rng.integers(3, NPTS - 4, n)
    # strictly between mesh points, away from t = 0's logarithmic singularity
    t = -T_MAX + (k + rng.uniform(0.05, 0.95, n)) * dt
    return t[np.abs(t) > 3 * dt]


def _g0(L, tab, t):
    """g0(t, 0) on the forward branch: greater for t > 0, lesser for t < 0."""
    out = np.zeros(2)
    a, b = O.gf_index(t, 0, 0), O.gf_index(0.0, 0, 0)
    L.ko_g0_eval(C.byref(tab.c), C.byref(a), C.byref(b), 1, out.ctypes.data_as(C.POINTER(C.c_double)))
    return complex(out[0], out[1])


def _interp_error(L, flat_table):
    mesh, les, gre, tab = flat_table
    err = 0.0
    for t in _offgrid_times(1):
        src = gre if t > 0 else les
        ref = np.interp(t, mesh, src.real) + 1j * np.interp(t, mesh, src.imag)
        err = max(err, abs(_g0(L, tab, t) - ref))
    return err


def _lesser_dd(t):
    """d^2/dt^2 of the flat-band lesser closed form
    g(t) = [e^{Gt} Ei(-Gt) - e^{-Gt} Ei(Gt)] / (2 pi) + (i/2) e^{-G|t|}:
    d/dt e^{Gt}Ei(-Gt) = G e^{Gt}Ei(-Gt) + 1/t and d/dt e^{-Gt}Ei(Gt) = -G e^{-Gt}Ei(Gt) + 1/t."""
    G = GAMMA
    re = (G * G * (np.exp(G * t) * sp.expi(-G * t) - np.exp(-G * t) * sp.expi(G * t)) + 2 * G / t) / (2 * np.pi)
    im = 0.5 * G * G * np.exp(-G * abs(t))
    return abs(re) + abs(im)


def _closed_form_excess(L, flat_table):
    """max over off-grid t of |oracle - closed form| / (dt^2/8 max|g''| on t's
    interval + 1e-14): <= 1 when the oracle interpolates linearly."""
    mesh, _, _, tab = flat_table
    dt = mesh[1] - mesh[0]
    worst = 0.0
    for t in _offgrid_times(2):
        i = int(np.floor((t + T_MAX) / dt))
        xs = np.linspace(mesh[i], mesh[i + 1], 9)
        bound = dt * dt / 8 * max(_lesser_dd(x) for x in xs) * 1.05 + 1e-14
        exact = complex(g0_oracle.flatband_analytic_lesser(t, GAMMA))
        if t > 0:  # the greater table: g^> = conj(g^<) for this model
            exact = exact.conjugate()
        worst = max(worst, abs(_g0(L, tab, t) - exact) / bound)
    return worst


def test_table_mesh_points_are_the_closed_form(flat_table):
    mesh, les, gre, _ = flat_table
    ref = g0_oracle.flatband_analytic_lesser(mesh, GAMMA)
    np.testing.assert_allclose(les, ref, rtol=0, atol=1e-15)
    np.testing.assert_array_equal(gre, np.conj(les))


def test_g0_offgrid_is_numpy_linear_interpolation(variants, flat_table):
    assert _interp_error(variants[0], flat_table) <= 1e-14


def test_g0_offgrid_within_linear_interpolation_bound(variants, flat_table):
    assert _closed_form_excess(variants[0], flat_table) <= 1.0


@pytest.mark.parametrize("plant", [1, 2])
def test_g0_offgrid_checks_catch_planted_mistakes(variants, flat_table, plant):
    assert _interp_error(variants[plant], flat_table) > 1e-6
    assert _closed_form_excess(variants[plant], flat_table) > 2.0


def test_g0_offgrid_two_orbital_tables(variants):
    """The dot-lead tables ([orbital][orbital][time]) index each block at its
    own offset: every block against numpy's interpolation of that block (random
    O(1) entries, so the rounding of the mesh position, ~1e3 eps, shows at 1e-13)."""
    rng = np.random.default_rng(5)
    n, tm = 2001, 20.0
    mesh = np.linspace(-tm, tm, n)
    les = rng.normal(size=(2, 2, n)) + 1j * rng.normal(size=(2, 2, n))
    gre = rng.normal(size=(2, 2, n)) + 1j * rng.normal(size=(2, 2, n))
    tab = O.G0Table(les, gre, tm)
    out = np.zeros(2)
    for _ in range(200):
        t = rng.uniform(-tm + 0.1, tm - 0.1)
        oa, ob = rng.integers(0, 2, 2)
        a, b = O.gf_index(t, 0, 0, 0, int(oa)), O.gf_index(0.0, 0, 0, 0, int(ob))
        variants[0].ko_g0_eval(C.byref(tab.c), C.byref(a), C.byref(b), 1,
                               out.ctypes.data_as(C.POINTER(C.c_double)))
        src = (gre if t > 0 else les)[oa, ob]
        ref = np.interp(t, mesh, src.real) + 1j * np.interp(t, mesh, src.imag)
        assert abs(complex(*out) - ref) <= 1e-12


# ---------------------------------------------------------------- Steffen spline
def _coeffs(L, x, y):
    x, y = np.ascontiguousarray(x, float), np.ascontiguousarray(y, float)
    co = np.zeros(4 * (len(x) - 1))
    dp = C.POINTER(C.c_double)
    L.ko_steffen_coeffs(x.ctypes.data_as(dp), y.ctypes.data_as(dp), len(x), co.ctypes.data_as(dp))
    return co.reshape(-1, 4)  # per interval: a, b, c, d of d + c h + b h^2 + a h^3


def _piece(co, x, i, xv):
    h = xv - x[i]
    a, b, c, d = co[i]
    return d + h * (c + h * (b + h * a))


def _nonuniform_grid(seed, n=40):
    rng = np.random.default_rng(seed)
    return np.concatenate([[0.0], np.cumsum(rng.uniform(0.2, 1.0, n - 1))]) / n


def _quadratic_error(L):
    """Interior intervals of y = 1 + 2x + 0.3x^2 on a non-uniform grid: the
    Steffen slopes equal the three-point parabola's, so the spline is the
    quadratic itself."""
    x = _nonuniform_grid(3)
    y = 1 + 2 * x + 0.3 * x * x
    co = _coeffs(L, x, y)
    err = 0.0
    for i in range(1, len(x) - 2):
        for f in (0.25, 0.5, 0.8):
            xv = x[i] + f * (x[i + 1] - x[i])
            err = max(err, abs(_piece(co, x, i, xv) - (1 + 2 * xv + 0.3 * xv * xv)))
    return err


def _monotone_violation(L):
    """Largest decrease of the spline over monotone (non-decreasing) data with
    steps of very different sizes, sampled densely on every interval."""
    rng = np.random.default_rng(11)
    worst = 0.0
    for _ in range(20):
        x = _nonuniform_grid(int(rng.integers(1 << 30)), 30)
        steps = rng.choice([0.0, 1e-3, 1.0, 50.0], size=len(x) - 1) * rng.uniform(0.5, 1.5, len(x) - 1)
        y = np.concatenate([[0.0], np.cumsum(steps)])
        co = _coeffs(L, x, y)
        for i in range(len(x) - 1):
            v = _piece(co, x, i, np.linspace(x[i], x[i + 1], 33))
            worst = max(worst, float(np.max(v[:-1] - v[1:])))
    return worst


def test_steffen_interpolates_and_is_c1(variants):
    x = _nonuniform_grid(7)
    y = np.sin(6 * x) + 0.2 * x
    co = _coeffs(variants[0], x, y)
    for i in range(len(x) - 1):
        h = x[i + 1] - x[i]
        a, b, c, d = co[i]
        assert d == y[i]
        assert abs(_piece(co, x, i, x[i + 1]) - y[i + 1]) <= 1e-13
        if i + 1 < len(x) - 1:
            assert abs(c + 2 * b * h + 3 * a * h * h - co[i + 1][2]) <= 1e-11 * (1 + abs(co[i + 1][2]))


def test_steffen_zero_slope_at_data_extremum(variants):
    x = _nonuniform_grid(8, 12)
    y = np.array([0, 1, 3, 4, 2, 1, 1.5, 3, 2.5, 2.0, 0.5, 0.0])
    co = _coeffs(variants[0], x, y)
    for i in range(1, len(x) - 1):
        if (y[i] - y[i - 1]) * (y[i + 1] - y[i]) <= 0:
            assert co[i][2] == 0.0


def test_steffen_reproduces_quadratic_and_preserves_monotonicity(variants):
    assert _quadratic_error(variants[0]) <= 1e-13
    assert _monotone_violation(variants[0]) <= 1e-12


@pytest.mark.parametrize("plant", [3, 4])
def test_steffen_checks_catch_planted_mistakes(variants, plant):
    assert _quadratic_error(variants[plant]) > 1e-8
