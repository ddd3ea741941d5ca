"""GPU parity of the HIP Keldysh-sum path against the CPU oracle, through the C-ABI.

Tolerances (FP64 path): the per-point values agree normwise,
max|gpu - oracle| <= 1e-10 * max|oracle| over a batch, and accumulated
estimates agree to 1e-10 relative.  The oracle sums the 2^n signed
det*det products with two pivoted LUs per configuration (wick_direct.cpp:110-149),
the device with one bordered LU, so bitwise equality is not expected.
"""
import numpy as np
import pytest

from keldy_amd import capi
from oracle import oracle as O
from tests import gpu_util as U

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _normwise(a, b):
    scale = max(np.max(np.abs(b)), 1e-300)
    return np.max(np.abs(a - b)) / scale


def _points(order, n, t, seed):
    rng = np.random.default_rng(seed)
    u = rng.uniform(0.0, t, size=(n, order))
    u[: n // 2] = -np.sort(-u[: n // 2], axis=1)  # ordered (simplex) half
    if n >= 8:
        u[3, 0] = -1e-3          # out of domain (wick_direct.cpp:53-57)
        u[5, :] = t              # vertices on the external time
        u[6, :] = 0.0            # all vertices at t = 0
        if order >= 2:
            u[7, 1] = u[7, 0]    # coincident vertices
    return u


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 6, 7, 8])
def test_eval_batch_fp64(flat_g0, order):
    t = 10.0
    n = 256 if order <= 5 else 48
    u = _points(order, n, t, seed=order)
    ctx = U.charge_ctx(flat_g0, order, t)
    got, dom = ctx.eval_batch(u)
    f = O.charge_integrand(U.oracle_charge(flat_g0), t)
    ref, rdom = f.batch(u)
    assert np.array_equal(dom, rdom)
    assert _normwise(got, ref) <= TOL


def test_eval_batch_order0(flat_g0):
    ctx = U.charge_ctx(flat_g0, 0, 10.0)
    got, dom = ctx.eval_batch(np.zeros((3, 0)))
    assert np.all(dom == 1)
    np.testing.assert_allclose(got, 0.5j, rtol=0, atol=1e-15)  # g^<(0) = i/2


@pytest.mark.parametrize("order", [2, 4, 6])
def test_eval_batch_fp32_lu(flat_g0, order):
    """FP32 LU with FP64 sum: error against the FP64 oracle, stated per order."""
    t = 10.0
    u = _points(order, 128, t, seed=10 + order)
    ctx = U.charge_ctx(flat_g0, order, t, prec=capi.KD_PREC_FP32_LU)
    got, _ = ctx.eval_batch(u)
    ref, _ = O.charge_integrand(U.oracle_charge(flat_g0), t).batch(u)
    assert _normwise(got, ref) <= 1e-4


@pytest.mark.parametrize("order", [1, 3, 5, 7])
def test_eval_warped_fp64(flat_g0, order):
    t = 10.0
    train, (descs, maps, tabs) = U.chain([("plasma", t), ("map", "inverse_square", t, 1.0, 1)])
    ctx = U.charge_ctx(flat_g0, order, t)
    ctx.set_warper_chain(descs, maps, tabs)
    l = O.Sobol(order, 0, True, True, 3).points(64)
    got, dom, jac = ctx.eval_warped_batch(l)
    f = O.charge_integrand(U.oracle_charge(flat_g0), t)
    ref = np.array([O.eval_warped(f, train, li)[0] for li in l])
    rdom = np.array([O.eval_warped(f, train, li)[1] for li in l])
    assert np.array_equal(dom, rdom)
    assert _normwise(got, ref) <= TOL


@pytest.mark.parametrize("order,spec", [
    (3, [("identity",)]),
    (4, [("plasma", 10.0), ("identity",), ("map", "exponential", 10.0, 2.0, 1)]),
    (5, [("plasma", 10.0), ("map", "inverse_power", 10.0, 1.5, 3)]),
    (6, [("map", "exponential", 10.0, 3.0, 1), ("identity",)]),
    (8, [("plasma", 10.0), ("map", "inverse_power", 10.0, 0.7, 4)]),
])
def test_eval_warped_identity_exponential_inverse_power(flat_g0, order, spec):
    """The identity warper (identity.hpp:30-49: u = l, J = 1, domain [0, 1]^n)
    and the exponential and inverse_power makers (product_1d_simple.hpp:560-630)
    on the device: values, domain flags and Jacobians against the oracle's
    chain on the same scrambled Sobol points."""
    t = 10.0
    train, (descs, maps, tabs) = U.chain(spec)
    ctx = U.charge_ctx(flat_g0, order, t)
    ctx.set_warper_chain(descs, maps, tabs)
    l = O.Sobol(order, 0, True, True, 11).points(64)
    got, dom, jac = ctx.eval_warped_batch(l)
    f = O.charge_integrand(U.oracle_charge(flat_g0), t)
    res = [O.eval_warped(f, train, li) for li in l]
    ref = np.array([r[0] for r in res])
    rdom = np.array([r[1] for r in res])
    rjac = np.array([train.reverse(li)[1] for li in l])
    assert np.array_equal(dom, rdom)
    assert np.any(dom == 1)
    assert _normwise(got, ref) <= TOL
    # a product of up to 8 per-axis factors from pow/log1p: device and libm differ by ulps
    np.testing.assert_allclose(jac, rjac, rtol=1e-12, atol=0)


def test_run_config1(flat_g0):
    """n=3, 2^16 scrambled Sobol points (seed 0), uniform product map on [0, 10]^3."""
    t, order, npts = 10.0, 3, 1 << 16
    train, (descs, maps, tabs) = U.chain([("map", "uniform", t, 1.0, 1)])
    sob = O.Sobol(order, 0, True, True, 0)
    Csr, shfr = sob.tables()
    ctx = U.charge_ctx(flat_g0, order, t)
    ctx.set_warper_chain(descs, maps, tabs)
    ctx.set_sobol(Csr, shfr, 0)
    acc = ctx.run(npts)
    f = O.charge_integrand(U.oracle_charge(flat_g0), t)
    s, npo, ndom = O.run(f, train, sob, npts, threads=4)
    assert acc.n_points == npo == npts
    assert acc.n_in_domain == ndom
    got = complex(acc.re, acc.im)
    assert abs(got - s) <= TOL * abs(s)
    assert ctx.sobol_state() == npts


@pytest.mark.parametrize("order", [4, 6])
def test_run_plasma_chain_and_continuation(flat_g0, order):
    t, npts = 10.0, 3000
    train, (descs, maps, tabs) = U.chain([("plasma", t), ("map", "inverse_square", t, 1.0, 1)])
    sob = O.Sobol(order, 0)
    Csr, shfr = sob.tables()
    ctx = U.charge_ctx(flat_g0, order, t)
    ctx.set_warper_chain(descs, maps, tabs)
    ctx.set_sobol(Csr, shfr, 0)
    ctx.run(1000)
    acc = ctx.run(npts - 1000)  # run() accumulates and continues the sequence
    f = O.charge_integrand(U.oracle_charge(flat_g0), t)
    s, npo, ndom = O.run(f, train, sob, npts, threads=4)
    assert acc.n_points == npts and acc.n_in_domain == ndom
    assert abs(complex(acc.re, acc.im) - s) <= TOL * abs(s)
