"""Time-resolved charge (compute_charge_Q_direct_time, compute_obs.hpp:85-101):
the Keldysh sum binned by the earliest vertex time on the GPU.

Checks (the reference's wick_direct_time_test.cpp:19-84 and binner semantics):
- the per-point integrand puts its value at min(times) and equals the plain
  integrand there (1e-16 there; here 1e-15 relative: same device evaluation);
- the bins of a run sum to the plain run's result on the same points (1e-13
  relative: bin sums reassociate the per-point values);
- bin contents equal the oracle's binning of its own per-point values (its
  warper chain and Keldysh sum at the same Sobol points)
  (1e-10 relative), counts of values added per bin match exactly, and points
  whose earliest time falls outside [0, t] are counted as dropped.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _params():
    return dict(beta=-1.0, bias_V=0.0, eps_d=0.0, Gamma=1.0, time_max=10.0, nr_time_points_gf=10001,
                alpha=0.0, bath_type="flatband", ft_method="analytic")


def test_integrand_time_point():
    from keldy_amd import impurity_oneband as kio
    m = kio.G0Model(kio.G0ModelOmega(_params()), False)
    g0 = kio.G0KeldyshContourT(m)
    A, B = kio.GfIndexT(5.0, 0, 0), kio.GfIndexT(5.0, 0, 0)
    ft = kio.IntegrandGDirectTime(g0, A, B, 0.0)
    fd = kio.IntegrandGDirect(g0, A, B, 0.0)
    res, dom = ft([2.5, 1.2])
    assert dom == 1 and len(res.data) == 1 and res.data[0][0] == 1.2
    v, _ = fd([2.5, 1.2])
    assert abs(res.data[0][1] - v) <= 1e-15 * abs(v)
    with pytest.raises(RuntimeError, match="Order 0"):
        ft([])


@pytest.mark.parametrize("order", [2, 4])
def test_binned_run_consistency(order):
    from keldy_amd import impurity_oneband as kio
    from keldy_amd import warpers as kw
    t, slices, n = 10.0, 10, 1 << 14
    m = kio.G0Model(kio.G0ModelOmega(_params()), False)
    cd = kio.ComputeChargeQDirect(m, t, order, 0.0)
    ct = kio.ComputeChargeQDirectTime(m, t, order, slices, 0.0)
    for c in (cd, ct):
        c.warper.emplace_back(kw.WarperPlasmaUvT(t))
        c.warper.emplace_back(kw.make_product_1d_simple_exponential(t, 1.0))
        c.run(n)
    b = ct.reduce_result()
    data = b.get_data()
    assert b.get_nr_bins() == slices and abs(b.get_bin_size() - t / slices) < 1e-15
    np.testing.assert_allclose(b.get_bin_coord(), (np.arange(slices) + 0.5) * t / slices, rtol=1e-15)
    total = cd.reduce_result()
    assert abs(data.sum() - total) <= 1e-13 * abs(total)
    # every point is counted once: in its bin, or dropped when its earliest time
    # is negative (the points outside the domain with this chain)
    added, dropped = b.get_nr_values_added().sum(), b.get_nr_values_dropped()
    assert added + dropped == n and dropped == n - cd.reduce_nr_points_in_domain()

    # the same binning by the oracle: its own warper chain and Keldysh sum at the
    # same Sobol points (wick_direct_time.cpp:42-163 bins by the earliest time)
    from tests import gpu_util as U
    l = O.Sobol(order, 0).points(n)
    train, _ = U.chain([("plasma", t), ("map", "exponential", t, 1.0, 1)])
    f = O.charge_integrand(O.G0Table(m.g0_lesser[0, 0], m.g0_greater[0, 0], m.time_max), t)
    vals = np.array([O.eval_warped(f, train, x)[0] for x in l])
    u = np.array([train.reverse(x)[0] for x in l])
    dvals, _, _ = cd.evaluate_warped_integrand_batch(l)
    assert np.max(np.abs(dvals - vals)) <= 1e-10 * np.max(np.abs(vals))
    tmin = u.min(axis=1)
    inb = (tmin >= 0.0) & (tmin <= t)
    idx = np.where(tmin == t, slices - 1, np.floor(np.where(inb, tmin, 0.0) / (t / slices)).astype(int))
    idx = np.minimum(idx, slices - 1)
    ref = np.zeros(slices, complex)
    np.add.at(ref, idx[inb], vals[inb])
    ref /= n
    assert np.max(np.abs(data - ref)) <= 1e-10 * np.max(np.abs(ref))
    np.testing.assert_array_equal(b.get_nr_values_added(), np.bincount(idx[inb], minlength=slices))
    assert b.get_nr_values_dropped() == int((~inb).sum())


def test_dropped_values():
    """keep_u_hypercube semantics: a chain whose earliest time can be negative
    (plasma_uv with uniform v on [0, 2t] per axis) drops those points."""
    from keldy_amd import impurity_oneband as kio
    from keldy_amd import warpers as kw
    t, n = 10.0, 1 << 12
    m = kio.G0Model(kio.G0ModelOmega(_params()), False)
    ct = kio.ComputeChargeQDirectTime(m, t, 2, 5, 0.0)
    ct.warper.emplace_back(kw.WarperPlasmaUvT(t))
    ct.warper.emplace_back(kw.make_product_1d_simple_uniform(2 * t))
    ct.run(n)
    b = ct.reduce_result()
    added, dropped = b.get_nr_values_added().sum(), b.get_nr_values_dropped()
    assert added + dropped == n and dropped > 0
    assert dropped == n - ct.reduce_nr_points_in_domain()
