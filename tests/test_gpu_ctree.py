"""The time-ordered tree (KD_ALGO_CTREE, orders 6..10) against the oracle and
the other Keldysh-sum kernels.

The kernel sorts each point's vertex times and eliminates the vertices in
time order on vertex-level tables (one value per vertex pair and tree node);
the oracle (wick_direct.cpp:43-158) builds the two Wick matrices per
configuration in the caller's vertex order and runs two partial-pivoting LUs.
Agreement is normwise max|gpu - oracle| <= 1e-10 max|oracle| in FP64 (1e-4 for
the FP32 LU), on random points and the edge rows that stress the time order:
vertices already sorted and reverse-sorted, exact ties (which compare_3way
breaks by the vertex index), every vertex at the external time or at 0, and
clusters within 1e-4, 1e-7 and 1e-9 (nearly equal rows: small pivots that
must take the partial-pivoting FP64 path).
"""
import numpy as np
import pytest

from keldy_amd import capi
from oracle import oracle as O
from tests import gpu_util as U

pytestmark = pytest.mark.gpu
TOL = 1e-10
ORDERS = [6, 7, 8, 9, 10]


def _normwise(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _points(order, n, t, seed):
    rng = np.random.default_rng(seed)
    u = rng.uniform(0.0, t, size=(n, order))
    u[: n // 4] = -np.sort(-u[: n // 4], axis=1)         # reverse-sorted
    u[n // 4: n // 2] = np.sort(u[n // 4: n // 2], axis=1)  # sorted
    u[3, 0] = -1e-3                                       # out of the domain
    u[5, :] = t                                           # every vertex at the external time
    u[6, :] = 0.0
    u[7, 1] = u[7, 0]                                     # exact ties
    u[9, 2:5] = u[9, 0]
    u[10, -1] = t                                         # one vertex at the external time
    u[8, :] = 0.37 * t + 1e-4 * rng.uniform(size=order)   # a cluster
    return u


def _ctx(flat_g0, order, t, prec=capi.KD_PREC_FP64, alpha=0.0):
    ctx = capi.Context(0)
    ctx.set_g0_table(flat_g0["lesser"], flat_g0["greater"], flat_g0["time_max"], alpha=alpha)
    ctx.set_integrand(order, (t, 0, 0), (t, 0, 1))
    ctx.set_precision(prec)
    ctx.set_algorithm(capi.KD_ALGO_CTREE)
    assert ctx.algorithm() == capi.KD_ALGO_CTREE
    return ctx


@pytest.mark.parametrize("order", ORDERS)
def test_ctree_fp64_vs_oracle(flat_g0, order):
    t = 10.0
    u = _points(order, 48, t, 1200 + order)
    ref, rdom = O.charge_integrand(U.oracle_charge(flat_g0), t).batch(u)
    got, dom = _ctx(flat_g0, order, t).eval_batch(u)
    assert np.array_equal(dom, rdom)
    err = _normwise(got, ref)
    print(f"order {order}: time-ordered tree vs oracle, normwise {err:.2e}")
    assert err <= TOL


@pytest.mark.parametrize("order", ORDERS)
def test_ctree_fp32_lu_vs_oracle(flat_g0, order):
    """FP32 LU on the edge rows: the 1e-4 cluster rows carry the largest
    error of every FP32 kernel (1.4e-4 normwise for the per-configuration LU
    at order 8 on these rows), so the time-ordered tree is held to 2e-4 and to
    twice the per-configuration FP32 error on the same points."""
    t = 10.0
    u = _points(order, 48, t, 1210 + order)
    ref, _ = O.charge_integrand(U.oracle_charge(flat_g0), t).batch(u)
    got, _ = _ctx(flat_g0, order, t, prec=capi.KD_PREC_FP32_LU).eval_batch(u)
    cfg = U.charge_ctx(flat_g0, order, t, capi.KD_PREC_FP32_LU)
    cfg.set_algorithm(capi.KD_ALGO_CONFIG)
    err, err_cfg = _normwise(got, ref), _normwise(cfg.eval_batch(u)[0], ref)
    print(f"order {order}: time-ordered tree, FP32 LU, vs oracle normwise {err:.2e} "
          f"(per-configuration FP32 LU {err_cfg:.2e})")
    assert err <= 2e-4 and err <= 2 * err_cfg


@pytest.mark.parametrize("order", ORDERS)
def test_ctree_near_coincident_pairs_vs_oracle(flat_g0, order):
    """Vertices in pairs 1e-7 or 1e-9 apart (nearly equal rows): flagged
    configurations go through the partial-pivoting FP64 LU of the sorted
    vertex-level table."""
    t = 10.0
    rng = np.random.default_rng(1220 + order)
    rows = []
    for d in (1e-7, 1e-9):
        for _ in range(4):
            taus = rng.uniform(0.5, 9.5, (order + 1) // 2)
            x = np.empty(2 * len(taus))
            x[0::2], x[1::2] = taus, taus + d
            rows.append(rng.permutation(x[:order]))
    u = np.array(rows)
    ref, rdom = O.charge_integrand(U.oracle_charge(flat_g0), t).batch(u)
    got, dom = _ctx(flat_g0, order, t).eval_batch(u)
    assert np.array_equal(dom, rdom)
    err = _normwise(got, ref)
    print(f"order {order}: near-coincident pairs, normwise {err:.2e}")
    assert err <= TOL


@pytest.mark.parametrize("order", [6, 8, 10])
def test_ctree_alpha_half_all_fallback(flat_g0, order):
    """alpha = 1/2 at half filling: the diagonal D = g^<(0) - i/2 is exactly 0,
    so every configuration's first pivot fails the check and every
    configuration takes the pivoted FP64 path; compared on the alpha = 0
    integrand's scale (the values themselves are ~1e-20 of it)."""
    t = 10.0
    u = _points(order, 16, t, 1230 + order)
    u = u[np.all(u >= 0, axis=1)]
    tab = O.G0Table(flat_g0["lesser"], flat_g0["greater"], flat_g0["time_max"], alpha=0.5)
    ref, rdom = O.charge_integrand(tab, t).batch(u)
    scale = np.max(np.abs(O.charge_integrand(U.oracle_charge(flat_g0), t).batch(u)[0]))
    ctx = _ctx(flat_g0, order, t, alpha=0.5)
    got, dom = ctx.eval_batch(u)
    assert np.array_equal(dom, rdom)
    assert np.all(np.isfinite(got))
    assert np.max(np.abs(got - ref)) <= TOL * scale
    _, (descs, maps, tabs) = U.chain([("map", "uniform", t, 1.0, 1)])
    ctx.set_warper_chain(descs, maps, tabs)
    Csr, shfr = O.Sobol(order, 0, True, True, 1).tables()
    ctx.set_sobol(Csr, shfr, 0)
    acc = ctx.run(256)
    assert acc.n_in_domain == 256 and acc.n_fallback == acc.n_in_domain << order


@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("prec", [capi.KD_PREC_FP64, capi.KD_PREC_FP32_LU])
def test_ctree_run_matches_wsplit(flat_g0, order, prec):
    """The driver's accumulated sum, domain count and time bins through the
    time-ordered tree equal those of the wave-per-point split kernel on the
    same Sobol points (FP64: 1e-12 relative; FP32 LU: 1e-4), and the cutoff
    drops the same points."""
    t = 10.0
    a = _ctx(flat_g0, order, t, prec)
    b = U.charge_ctx(flat_g0, order, t, prec)
    b.set_algorithm(capi.KD_ALGO_WSPLIT)
    train, (descs, maps, tabs) = U.chain([("plasma", t), ("map", "inverse", t, 1.0, 1)])
    Csr, shfr = O.Sobol(order, 0, True, True, 11).tables()
    for c in (a, b):
        c.set_warper_chain(descs, maps, tabs)
        c.set_sobol(Csr, shfr, 0)
        c.set_time_bins(0.0, t, 7)
    m = 1 << 14
    ra, rb = a.run(m), b.run(m)
    tol = 1e-12 if prec == capi.KD_PREC_FP64 else 1e-4
    assert ra.n_points == rb.n_points == m and ra.n_in_domain == rb.n_in_domain
    za, zb = complex(ra.re, ra.im), complex(rb.re, rb.im)
    print(f"order {order} prec {prec}: ctree {za:.10e} wsplit {zb:.10e}; flops/point "
          f"{a.flops_per_point():.0f} vs {b.flops_per_point():.0f}")
    assert abs(za - zb) <= tol * abs(zb)
    (ba, ca, xa), (bb, cb, xb) = a.time_bins(), b.time_bins()
    assert np.array_equal(ca, cb) and xa == xb
    assert np.max(np.abs(ba - bb)) <= tol * 10 * np.max(np.abs(bb))
    # evaluation mode on warped points: values and Jacobians
    rng = np.random.default_rng(order)
    lpts = rng.uniform(size=(256, order))
    va, da = a.eval_warped_batch(lpts)[:2]
    vb, db = b.eval_warped_batch(lpts)[:2]
    assert np.array_equal(da, db)
    assert np.max(np.abs(va - vb)) <= tol * np.max(np.abs(vb))


@pytest.mark.parametrize("order", [6, 7, 8])
def test_ctree_cutoff_matches_wsplit(flat_g0, order):
    """The cutoff on the point's value (two points per wave at orders 6-7:
    each half sums and cuts its own point)."""
    t = 10.0
    u = _points(order, 64, t, 1240 + order)
    ref, _ = O.charge_integrand(U.oracle_charge(flat_g0), t).batch(u)
    a = np.sort(np.abs(ref[np.abs(ref) > 0]))
    cut = float(np.sqrt(a[len(a) // 2] * a[len(a) // 2 + 1]))  # between two values, not on one
    vals = []
    for algo in (capi.KD_ALGO_CTREE, capi.KD_ALGO_WSPLIT):
        ctx = capi.Context(0)
        ctx.set_g0_table(flat_g0["lesser"], flat_g0["greater"], flat_g0["time_max"])
        ctx.set_integrand(order, (t, 0, 0), (t, 0, 1), cutoff=cut)
        ctx.set_algorithm(algo)
        vals.append(ctx.eval_batch(u)[0])
    assert np.array_equal(vals[0] == 0, vals[1] == 0)
    assert np.max(np.abs(vals[0] - vals[1])) <= 1e-12 * np.max(np.abs(vals[1]))


def _dot_lead_tables():
    from keldy_amd import impurity_oneband as kio
    p = dict(beta=-1.0, bias_V=0.6, eps_d=-0.5, Gamma=1.0, time_max=20.0, nr_time_points_gf=2001, alpha=0.0,
             half_bandwidth=2.0, bath_type="semicircle", ft_method="contour")
    m = kio.G0Model(kio.G0ModelOmega(p), True)
    return m.g0_lesser, m.g0_greater, m.time_max


@pytest.mark.parametrize("order", ORDERS)
def test_ctree_current_dot_lead_vs_oracle(order):
    """The current's two-orbital externals (A on the lead orbital, compute_obs.hpp:67-80)
    at external times 5 and 20 (every vertex before it) and 3 (vertices after
    the external time too, where the border entries no longer depend on the
    vertex's Keldysh index)."""
    les, gre, tmax = _dot_lead_tables()
    tab = O.G0Table(les, gre, tmax)
    for t in (5.0, 20.0, 3.0):
        u = _points(order, 32, min(t, 10.0), 1250 + order)
        if t == 3.0:  # rows with vertices after the external time (their sum vanishes by causality)
            late = np.random.default_rng(order).uniform(0.0, 8.0, size=(8, order))
            u = np.concatenate([u, late])
        fo = O.Integrand(tab, O.gf_index(t, 0, 0, 0, 1), O.gf_index(t, 0, 1, 0, 0))
        ref, rdom = fo.batch(u)
        ctx = capi.Context(0)
        ctx.set_g0_table(les, gre, tmax)
        ctx.set_integrand(order, (t, 0, 0, 1), (t, 0, 1, 0))
        ctx.set_algorithm(capi.KD_ALGO_CTREE)
        got, dom = ctx.eval_batch(u)
        assert np.array_equal(dom, rdom)
        err = _normwise(got, ref)
        print(f"order {order}, current at t = {t}: normwise {err:.2e}")
        assert err <= TOL
