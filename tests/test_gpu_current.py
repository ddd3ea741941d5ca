"""Current J with the dot-lead g0 (compute_current_J_direct, compute_obs.hpp:67-80)
on the GPU: the reference's known answer, parity of the two-orbital Keldysh
sum with the CPU oracle, and many external times batched in one hipGraph
launch at order 10 (the time-resolved current).

Tolerances: J_1 within 1e-5 of the exact value (compute_obs_test.cpp:183-206);
integrand values normwise 1e-10 against the oracle on the same two-orbital
tables; batched == individual runs to 1e-15 relative.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _semicircle(**kw):
    p = dict(beta=-1.0, bias_V=0.6, eps_d=-0.5, Gamma=1.0, time_max=100.0, nr_time_points_gf=5001, alpha=0.0,
             half_bandwidth=2.0, bath_type="semicircle", ft_method="contour")
    p.update(kw)
    return p


def test_current_J1_known_answer():
    """compute_obs_test.cpp ValueCurrent1: order 1, t = 100, [plasma_uv(100),
    inverse_square(100, 0.5)], 1e5 unshifted Sobol points;
    J_1 = -2 Re(i * result) = 0.041401946352562884 (analytic) within 1e-5."""
    from keldy_amd import impurity_oneband as kio
    from keldy_amd import warpers as kw
    c = kio.ComputeCurrentJDirect(_semicircle(), 100.0, 1, 0.0)
    c.warper.emplace_back(kw.WarperPlasmaUvT(100.0))
    c.warper.emplace_back(kw.make_product_1d_simple_inverse_square(100.0, 0.5))
    c.run(100000)
    J1 = -2 * (1j * c.reduce_result()).real
    assert abs(J1 - 0.041401946352562884) < 1e-5


@pytest.fixture(scope="module")
def dot_lead_model():
    from keldy_amd import impurity_oneband as kio
    return kio.G0Model(kio.G0ModelOmega(_semicircle(time_max=20.0, nr_time_points_gf=2001)), True)


@pytest.mark.parametrize("order", [1, 2, 3, 4, 6])
def test_current_integrand_vs_oracle(dot_lead_model, order):
    from keldy_amd import impurity_oneband as kio
    t = 10.0
    g0 = kio.G0KeldyshContourT(dot_lead_model)
    A, B = kio.GfIndexT(t, 0, 0, 0, 1), kio.GfIndexT(t, 0, 1, 0, 0)
    f = kio.IntegrandGDirect(g0, A, B, 0.0)
    rng = np.random.default_rng(order)
    u = rng.uniform(0.0, t, (64, order))
    got, dom = f.eval_batch(u)
    tab = O.G0Table(dot_lead_model.g0_lesser, dot_lead_model.g0_greater, dot_lead_model.time_max)
    fo = O.Integrand(tab, O.gf_index(t, 0, 0, 0, 1), O.gf_index(t, 0, 1, 0, 0))
    ref, rdom = fo.batch(u)
    assert np.array_equal(dom, rdom)
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))


def test_current_many_times_one_launch(dot_lead_model):
    """Time-resolved current at order 10: 128 external times in (0, 20], one
    integrator per time, all run as one hipGraph launch; equals running each
    integrator on its own (same kernels, fixed-order reductions)."""
    from keldy_amd import impurity_oneband as kio
    from keldy_amd import warpers as kw
    order, m = 10, 1 << 12
    times = np.linspace(20.0 / 128, 20.0, 128)

    def make(t):
        c = kio.ComputeCurrentJDirect(dot_lead_model, float(t), order, 0.0)
        c.warper.emplace_back(kw.WarperPlasmaUvT(float(t)))
        c.warper.emplace_back(kw.make_product_1d_simple_inverse(float(t), 1.0))
        c.reset_rng("sobol", 0, True, True, 7)
        return c

    batch = [make(t) for t in times]
    ms = kio.run_batch(batch, [m] * len(batch))
    assert ms > 0
    for i in (0, 63, 127):
        ref = make(times[i])
        ref.run(m)
        a, b = batch[i].reduce_result(), ref.reduce_result()
        assert abs(a - b) <= 1e-15 * abs(b)
        assert batch[i].reduce_nr_points_run() == m
    J = np.array([-2 * (1j * c.reduce_result()).real for c in batch])
    assert np.all(np.isfinite(J))
