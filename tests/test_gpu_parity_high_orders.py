"""Oracle parity of the Keldysh sum at orders 9-12 for every algorithm and
launch shape the library can run there, the pivoted FP64 fallback under
adversarial inputs, and the time-resolved current (orders 9-10) per external
time.

The oracle (wick_direct.cpp:43-158: two partial-pivoting LUs per
configuration, 2^n configurations summed serially) is evaluated on the same
points.  Tolerance: normwise max|gpu - oracle| <= 1e-10 max|oracle| for every
FP64 path; 1e-4 for the FP32 LU (which keeps FP64 at order 10 and above,
so the FP64 bound applies there).
"""
import numpy as np
import pytest

from keldy_amd import capi
from oracle import oracle as O
from tests import gpu_util as U

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _normwise(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _points(order, n, t, seed):
    """Random vertex times with the edge rows of the lower-order test: an
    ordered half, one point out of the domain, every vertex at t, every vertex at
    0, two coinciding vertices, and a cluster of vertices within 1e-4 of each
    other (nearly equal rows of the Wick matrix)."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(0.0, t, size=(n, order))
    u[: n // 2] = -np.sort(-u[: n // 2], axis=1)
    u[3, 0] = -1e-3
    u[5, :] = t
    u[6, :] = 0.0
    u[7, 1] = u[7, 0]
    u[8, :] = 0.37 * t + 1e-4 * rng.uniform(size=order)
    return u


def _shapes(order):
    """(label, algorithm, lanes) for every Keldysh-sum kernel valid at this order."""
    out = [("config64", capi.KD_ALGO_CONFIG, 64), ("config32", capi.KD_ALGO_CONFIG, 32),
           ("config16", capi.KD_ALGO_CONFIG, 16), ("pair", capi.KD_ALGO_PAIR, 64),
           ("tree", capi.KD_ALGO_TREE, 64)]
    if order <= 10:
        out.append(("split", capi.KD_ALGO_SPLIT, 64))
        out.append(("wsplit", capi.KD_ALGO_WSPLIT, 64))
        out.append(("ctree", capi.KD_ALGO_CTREE, 64))
    return out


@pytest.fixture(scope="module")
def oracle_ref(flat_g0):
    cache = {}

    def get(order, n, seed):
        key = (order, n, seed)
        if key not in cache:
            u = _points(order, n, 10.0, seed)
            f = O.charge_integrand(U.oracle_charge(flat_g0), 10.0)
            cache[key] = (u,) + tuple(f.batch(u))
        return cache[key]

    return get


@pytest.mark.parametrize("order", [9, 10, 11, 12])
def test_eval_batch_fp64_every_shape_vs_oracle(flat_g0, oracle_ref, order):
    """Every algorithm (split, wave-per-point split, tree, paired, per-configuration at 64, 32 and 16
    lanes per point) against the oracle.  At orders 11-12 with 16 or 32 lanes a
    lane holds 128-256 configurations; the coinciding-vertex rows flag some of
    them for the pivoted fallback, which must recompute each exactly once."""
    n = 32 if order <= 10 else 16
    u, ref, rdom = oracle_ref(order, n, 900 + order)
    default = U.charge_ctx(flat_g0, order, 10.0)
    got, dom = default.eval_batch(u)
    assert np.array_equal(dom, rdom)
    assert _normwise(got, ref) <= TOL, f"default algorithm {default.algorithm()}"
    for label, algo, lanes in _shapes(order):
        ctx = U.charge_ctx(flat_g0, order, 10.0)
        ctx.set_algorithm(algo)
        ctx.set_launch_shape(lanes, 16)
        assert ctx.algorithm() == algo, label
        got, dom = ctx.eval_batch(u)
        assert np.array_equal(dom, rdom), label
        err = _normwise(got, ref)
        print(f"order {order} {label}: normwise error vs oracle {err:.2e}")
        assert err <= TOL, label


@pytest.mark.parametrize("order", [9, 10, 11, 12])
def test_eval_batch_fp32_lu_vs_oracle(flat_g0, oracle_ref, order):
    """FP32-LU mode: FP32 at order 9, FP64 kept from order 10 on."""
    n = 32 if order <= 10 else 16
    u, ref, _ = oracle_ref(order, n, 900 + order)
    ctx = U.charge_ctx(flat_g0, order, 10.0, prec=capi.KD_PREC_FP32_LU)
    got, _ = ctx.eval_batch(u)
    err = _normwise(got, ref)
    print(f"order {order} FP32-LU mode: normwise error vs oracle {err:.2e}")
    assert err <= (1e-4 if order < 10 else TOL)


@pytest.mark.parametrize("order", [8, 9, 10])
def test_near_coincident_vertex_pairs_vs_oracle(flat_g0, order):
    """Vertices in pairs 1e-7 or 1e-9 apart: the configurations that give both
    vertices of a pair the same Keldysh index have two nearly equal rows, an
    unpivoted pivot of ~delta that fails the pivot check, and go through the
    partial-pivoting FP64 LU; the sum stays well conditioned (the oracle and a
    40-digit evaluation of the same matrices agree to ~5e-12 at order 8).
    Every algorithm against the oracle at 1e-10."""
    t = 10.0
    rng = np.random.default_rng(960 + order)
    rows = []
    for d in (1e-7, 1e-9):
        for _ in range(4):
            taus = rng.uniform(0.5, 9.5, (order + 1) // 2)
            x = np.empty(2 * len(taus))
            x[0::2], x[1::2] = taus, taus + d
            rows.append(rng.permutation(x[:order]))
    u = np.array(rows)
    ref, rdom = O.charge_integrand(U.oracle_charge(flat_g0), t).batch(u)
    for label, algo, lanes in _shapes(order):
        ctx = U.charge_ctx(flat_g0, order, t)
        ctx.set_algorithm(algo)
        ctx.set_launch_shape(lanes, 16)
        got, dom = ctx.eval_batch(u)
        assert np.array_equal(dom, rdom), label
        err = _normwise(got, ref)
        print(f"order {order} {label}: near-coincident pairs, normwise error vs oracle {err:.2e}")
        assert err <= TOL, label


@pytest.mark.parametrize("order", [8, 9, 10])
def test_half_filling_alpha_half_all_fallback(flat_g0, order):
    """Keldysh shift alpha = 1/2 at half filling: every diagonal entry
    g^<(0) - i alpha is exactly 0 (model.hpp:244-265), so every configuration
    must fail the pivot check and take the partial-pivoting FP64 path: a run's
    fallback count is every configuration of every in-domain point.  The
    integrand then vanishes to ~1e-20 of its alpha = 0 size, below what any
    FP64 determinant resolves (the oracle's own values are off by tens of
    percent against a 40-digit evaluation), so values are compared on the
    alpha = 0 integrand's scale: device and oracle agree to 1e-10 of it and are
    finite."""
    t = 10.0
    g0 = dict(flat_g0)
    u = _points(order, 16, t, 950 + order)
    u = u[np.all(u >= 0, axis=1)]
    tab = O.G0Table(g0["lesser"], g0["greater"], g0["time_max"], alpha=0.5)
    ref, rdom = O.charge_integrand(tab, t).batch(u)
    scale = np.max(np.abs(O.charge_integrand(U.oracle_charge(flat_g0), t).batch(u)[0]))
    for label, algo, lanes in _shapes(order):
        ctx = capi.Context(0)
        ctx.set_g0_table(g0["lesser"], g0["greater"], g0["time_max"], alpha=0.5)
        ctx.set_integrand(order, (t, 0, 0), (t, 0, 1))
        ctx.set_algorithm(algo)
        ctx.set_launch_shape(lanes, 16)
        got, dom = ctx.eval_batch(u)
        assert np.array_equal(dom, rdom), label
        assert np.all(np.isfinite(got)), label
        assert np.max(np.abs(got - ref)) <= TOL * scale, label
        if label in ("split", "wsplit", "ctree", "config64", "tree"):
            # the uniform map on [0, t]^n: every point in the domain
            _, (descs, maps, tabs) = U.chain([("map", "uniform", t, 1.0, 1)])
            ctx.set_warper_chain(descs, maps, tabs)
            Csr, shfr = O.Sobol(order, 0, True, True, 1).tables()
            ctx.set_sobol(Csr, shfr, 0)
            acc = ctx.run(256)
            assert acc.n_in_domain == 256
            print(f"order {order} {label}, alpha=1/2: {acc.n_fallback} fallbacks for "
                  f"{acc.n_in_domain} in-domain points x {1 << order} configurations")
            assert acc.n_fallback == acc.n_in_domain << order


def _dot_lead_model():
    from keldy_amd import impurity_oneband as kio
    p = dict(beta=-1.0, bias_V=0.6, eps_d=-0.5, Gamma=1.0, time_max=20.0, nr_time_points_gf=2001, alpha=0.0,
             half_bandwidth=2.0, bath_type="semicircle", ft_method="contour")
    return kio.G0Model(kio.G0ModelOmega(p), True)


@pytest.mark.parametrize("order", [9, 10])
def test_time_resolved_current_vs_oracle_per_time(order):
    """The time-resolved current's workload (dot-lead g0 on [-20, 20], current
    externals {t, up, fwd, 0, orbital 1} / {t, up, bwd, 0, orbital 0},
    compute_obs.hpp:67-80) at three of its 128 external times: integrand
    values against the oracle at 1e-10 per time, then a short run per time
    whose fallback count is printed."""
    from keldy_amd import impurity_oneband as kio
    from keldy_amd import warpers as kw
    model = _dot_lead_model()
    tab = O.G0Table(model.g0_lesser, model.g0_greater, model.time_max)
    times = np.linspace(20.0 / 128, 20.0, 128)
    for i in (0, 63, 127):
        t = float(times[i])
        g0 = kio.G0KeldyshContourT(model)
        f = kio.IntegrandGDirect(g0, kio.GfIndexT(t, 0, 0, 0, 1), kio.GfIndexT(t, 0, 1, 0, 0), 0.0)
        u = _points(order, 24, t, 970 + i)
        got, dom = f.eval_batch(u)
        fo = O.Integrand(tab, O.gf_index(t, 0, 0, 0, 1), O.gf_index(t, 0, 1, 0, 0))
        ref, rdom = fo.batch(u)
        assert np.array_equal(dom, rdom)
        err = _normwise(got, ref)
        c = kio.ComputeCurrentJDirect(model, t, order, 0.0)
        c.warper.emplace_back(kw.WarperPlasmaUvT(t))
        c.warper.emplace_back(kw.make_product_1d_simple_inverse(t, 1.0))
        c.reset_rng("sobol", 0, True, True, 5)
        c.run(1 << 14)
        print(f"order {order}, t = {t:.4f}: normwise error vs oracle {err:.2e}; run of 2^14 points: "
              f"{c.reduce_nr_points_in_domain()} in domain, {c.reduce_nr_fallback()} configurations "
              f"recomputed by the pivoted FP64 LU")
        assert err <= TOL, f"t = {t}"
