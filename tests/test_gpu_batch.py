"""Batched uneven orders in one hipGraph launch, the launch-shape autotuner and
the FP32-LU precision modes, through the C-ABI.

- kd_run_batch over orders 1..N with uneven point counts gives, per order, the
  accumulator a separate kd_run gives on the same points (same kernels, same
  fixed-order reductions: equal to 1e-15 relative).
- kd_autotune leaves generator state and accumulator untouched; any launch
  shape gives the same sum up to FP64 reassociation (1e-12 relative).
- FP32 LU vs FP64 per order on the same points: relative error of the
  accumulated estimate <= 1e-4 (orders 1-8); KD_PREC_FP32_LU keeps FP64 from
  order 10 (identical to the FP64 run).
"""
import numpy as np
import pytest

from keldy_amd import capi
from oracle import oracle as O
from tests import gpu_util as U

pytestmark = pytest.mark.gpu


def _ctx(flat_g0, order, seed, prec=capi.KD_PREC_FP64, t=10.0):
    tr, (descs, maps, tab) = U.chain([("plasma", t), ("map", "inverse", t, 1.0, 1)])
    ctx = U.charge_ctx(flat_g0, order, t, prec)
    ctx.set_warper_chain(descs, maps, tab)
    s = O.Sobol(order, 0, True, True, seed)
    ctx.set_sobol(*s.tables(), 0)
    return ctx


def _rel(a, b):
    a, b = complex(a.re, a.im), complex(b.re, b.im)
    return abs(a - b) / max(abs(b), 1e-300)


def test_run_batch_matches_individual_runs(flat_g0):
    orders = list(range(1, 9))
    counts = [(1 << 18) // (n * n) for n in orders]
    batch = [_ctx(flat_g0, n, 40 + n) for n in orders]
    accs, ms = capi.run_batch(batch, counts)
    assert ms > 0
    for n, m, a, c in zip(orders, counts, accs, batch):
        ref = _ctx(flat_g0, n, 40 + n).run(m)
        assert a.n_points == ref.n_points == m and a.n_in_domain == ref.n_in_domain
        assert _rel(a, ref) <= 1e-15, f"order {n}"
        assert c.sobol_state() == m
    # continuation: a second batch call picks up where the first stopped
    accs2, _ = capi.run_batch(batch[:3], counts[:3])
    for n, m, a in zip(orders[:3], counts[:3], accs2):
        ref = _ctx(flat_g0, n, 40 + n)
        ref.run(m)
        r2 = ref.run(m)
        assert _rel(a, r2) <= 1e-15


def test_run_batch_errors(flat_g0):
    c = _ctx(flat_g0, 3, 1)
    with pytest.raises(capi.KdError, match="twice"):
        capi.run_batch([c, c], [64, 64])


@pytest.mark.parametrize("order", [3, 8])
def test_autotune(flat_g0, order):
    ctx = _ctx(flat_g0, order, 7)
    lanes, wpc, ms = ctx.autotune(1 << 16)
    # lanes == 0: the Schur-tree algorithm won (orders >= 6)
    assert lanes in (16, 32, 64) or lanes == 1 << order or (lanes == 0 and order >= 6)
    assert wpc in (16, 32, 64, 128) and ms > 0
    assert ctx.sobol_state() == 0 and ctx.accum().n_points == 0 and ctx.accum().re == 0.0
    a = ctx.run(1 << 16)
    ref = _ctx(flat_g0, order, 7)
    ref.set_launch_shape(64, 16)
    ref.set_algorithm(capi.KD_ALGO_CONFIG)
    r = ref.run(1 << 16)
    assert _rel(a, r) <= 1e-12


@pytest.mark.parametrize("order", [1, 2, 4, 6, 8, 10])
def test_fp32_error_per_order(flat_g0, order):
    m = max(1 << 12, (1 << 18) // (order * order))
    r64 = _ctx(flat_g0, order, 3).run(m)
    r32 = _ctx(flat_g0, order, 3, capi.KD_PREC_FP32_LU).run(m)
    rel = _rel(r32, r64)
    print(f"order {order}: FP32-LU relative error of the estimate {rel:.3e}")
    if order >= 10:
        assert rel <= 1e-15  # FP64 LU kept from order 10
        rall = _ctx(flat_g0, order, 3, capi.KD_PREC_FP32_LU_ALL).run(m)
        print(f"order {order}: FP32 at every order: {_rel(rall, r64):.3e}")
        assert _rel(rall, r64) <= 1e-2
    else:
        assert rel <= 1e-4


@pytest.mark.parametrize("order", [6, 7, 8, 9, 10])
@pytest.mark.parametrize("prec", [0, 1])
def test_tree_matches_per_configuration(flat_g0, order, prec):
    """The Schur-tree algorithm and the per-configuration LU on the same points:
    FP64 to 1e-12 normwise (reassociation of a cancelling 2^n-term sum), FP32
    to 1e-4."""
    t = 10.0
    rng = np.random.default_rng(order)
    u = rng.uniform(0.0, t, (256, order))
    u[:128] = -np.sort(-u[:128], axis=1)
    a = U.charge_ctx(flat_g0, order, t, prec)
    b = U.charge_ctx(flat_g0, order, t, prec)
    a.set_algorithm(capi.KD_ALGO_TREE)
    b.set_algorithm(capi.KD_ALGO_CONFIG)
    assert a.algorithm() == capi.KD_ALGO_TREE and b.algorithm() == capi.KD_ALGO_CONFIG
    va, da = a.eval_batch(u)
    vb, db = b.eval_batch(u)
    assert np.array_equal(da, db)
    tol = 1e-12 if prec == 0 else 1e-4
    assert np.max(np.abs(va - vb)) <= tol * np.max(np.abs(vb))
    fa, fb = a.flops_per_point(), b.flops_per_point()
    assert fa < fb
    print(f"order {order}: tree {fa:.0f} flops/point vs per-configuration {fb:.0f} ({fb / fa:.1f}x)")


@pytest.mark.parametrize("order,prec", [(7, 0), (8, 0), (10, 0), (12, 0), (8, 1)])
def test_pair_matches_per_configuration_and_oracle(flat_g0, order, prec):
    """The paired LU (the last vertex's two Keldysh indices share the
    elimination) against the per-configuration LU on the same points: FP64 to
    1e-12 normwise, FP32 to 1e-4; and against the oracle at order 8."""
    t = 10.0
    rng = np.random.default_rng(50 + order)
    n = 256 if order <= 10 else 64
    u = rng.uniform(0.0, t, (n, order))
    u[: n // 2] = -np.sort(-u[: n // 2], axis=1)
    u[3, :] = 0.0
    a = U.charge_ctx(flat_g0, order, t, prec)
    b = U.charge_ctx(flat_g0, order, t, prec)
    a.set_algorithm(capi.KD_ALGO_PAIR)
    b.set_algorithm(capi.KD_ALGO_CONFIG)
    assert a.algorithm() == capi.KD_ALGO_PAIR and b.algorithm() == capi.KD_ALGO_CONFIG
    va, da = a.eval_batch(u)
    vb, db = b.eval_batch(u)
    assert np.array_equal(da, db)
    tol = 1e-12 if prec == 0 else 1e-4
    assert np.max(np.abs(va - vb)) <= tol * np.max(np.abs(vb))
    fa, fb = a.flops_per_point(), b.flops_per_point()
    assert fa < fb
    print(f"order {order}: paired {fa:.0f} flops/point vs per-configuration {fb:.0f} ({fb / fa:.2f}x)")
    if order >= 8 and prec == 0:
        ref, rdom = O.charge_integrand(U.oracle_charge(flat_g0), t).batch(u[:32])
        assert np.array_equal(da[:32], rdom)
        assert np.max(np.abs(va[:32] - ref)) <= 1e-10 * np.max(np.abs(ref))
    # runs through the driver: the same accumulated sum
    s = O.Sobol(order, 0, True, True, 2)
    Csr, shfr = s.tables()
    for c in (a, b):
        c.set_sobol(Csr, shfr, 0)
    ra, rb = a.run(1 << 12), b.run(1 << 12)
    assert ra.n_in_domain == rb.n_in_domain
    assert abs(complex(ra.re, ra.im) - complex(rb.re, rb.im)) <= tol * abs(complex(rb.re, rb.im)) * 10


@pytest.mark.parametrize("algo", [capi.KD_ALGO_SPLIT, capi.KD_ALGO_WSPLIT], ids=["split", "wsplit"])
@pytest.mark.parametrize("order,prec", [(6, 0), (7, 0), (8, 0), (9, 0), (10, 0), (6, 1), (7, 1), (8, 1), (9, 1),
                                        (10, 2)])
def test_split_matches_per_configuration_and_oracle(flat_g0, order, prec, algo):
    """The split algorithm (Schur tree over the leading vertices, then one
    bordered LU per configuration) against the per-configuration LU on the same
    points: FP64 to 1e-12 normwise, FP32 to 1e-4 (1e-2 with FP32 at order 10,
    where both FP32 results carry the cancellation of 1024 terms); against the
    oracle at orders 8-10 (FP64, 32 points); the same domain flags; and the
    same accumulated sum through the driver, with time bins."""
    t = 10.0
    rng = np.random.default_rng(70 + order)
    n = 256 if order <= 9 else 96
    u = rng.uniform(0.0, t, (n, order))
    u[: n // 2] = -np.sort(-u[: n // 2], axis=1)
    u[3, :] = 0.0
    u[5, :] = u[5, 0]  # coinciding vertex times
    a = U.charge_ctx(flat_g0, order, t, prec)
    b = U.charge_ctx(flat_g0, order, t, prec)
    a.set_algorithm(algo)
    b.set_algorithm(capi.KD_ALGO_CONFIG)
    assert a.algorithm() == algo and b.algorithm() == capi.KD_ALGO_CONFIG
    va, da = a.eval_batch(u)
    vb, db = b.eval_batch(u)
    assert np.array_equal(da, db)
    tol = 1e-12 if prec == 0 else 1e-4 if prec == 1 else 1e-2
    assert np.max(np.abs(va - vb)) <= tol * np.max(np.abs(vb))
    fa, fb = a.flops_per_point(), b.flops_per_point()
    assert fa < fb
    print(f"order {order}: {'wsplit' if algo == capi.KD_ALGO_WSPLIT else 'split'} {fa:.0f} flops/point "
          f"vs per-configuration {fb:.0f} ({fb / fa:.2f}x)")
    if order >= 8 and prec == 0:
        ref, rdom = O.charge_integrand(U.oracle_charge(flat_g0), t).batch(u[:32])
        assert np.array_equal(da[:32], rdom)
        assert np.max(np.abs(va[:32] - ref)) <= 1e-10 * np.max(np.abs(ref))
    s = O.Sobol(order, 0, True, True, 4)
    Csr, shfr = s.tables()
    for c in (a, b):
        c.set_sobol(Csr, shfr, 0)
        c.set_time_bins(0.0, t, 7)
    m = 1 << 12
    ra, rb = a.run(m), b.run(m)
    assert ra.n_points == rb.n_points == m and ra.n_in_domain == rb.n_in_domain
    assert abs(complex(ra.re, ra.im) - complex(rb.re, rb.im)) <= tol * abs(complex(rb.re, rb.im)) * 10
    (ba, ca, xa), (bb, cb, xb) = a.time_bins(), b.time_bins()
    assert np.array_equal(ca, cb) and xa == xb
    assert np.max(np.abs(ba - bb)) <= tol * 10 * np.max(np.abs(bb))
