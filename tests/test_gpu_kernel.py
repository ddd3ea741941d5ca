"""The kernel method on the GPU (integrand_g_kernel, wick_kernel.cpp:42-138;
compute_gf_kernel, compute_obs.hpp:105-117; single omega,
wick_kernel_single_omega.cpp:41-140) against the CPU oracle and the
reference's own tests.

Tolerances (FP64): per-configuration entries normwise 1e-11 against the
oracle (LU + zgetrs there, one unpivoted LU with a scaled back substitution
here); binned run totals normwise 1e-10 against the oracle on the same points
(bins are summed in another order); counts exact; the reference's regression
array within its 1e-6 and its particle-hole symmetry within its 1e-12.
"""
import numpy as np
import pytest

from keldy_amd import capi
from oracle import oracle as O
from tests import gpu_util as U

pytestmark = pytest.mark.gpu
UP, DOWN, FWD, BWD = 0, 1, 0, 1


def _normwise(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _points(order, n, t, seed):
    rng = np.random.default_rng(seed)
    u = rng.uniform(0.0, t, size=(n, order))
    u[: n // 2] = -np.sort(-u[: n // 2], axis=1)
    if n >= 8:
        u[3, 0] = -1e-3          # out of domain
        u[5, :] = t              # vertices on the external time
        u[6, :] = 0.0            # all vertices at 0
        if order >= 2:
            u[7, 1] = u[7, 0]    # coincident vertices
    return u


def _kernel_ctx(g0, order, X, omega=None):
    ctx = capi.Context(0)
    ctx.set_g0_table(g0["lesser"], g0["greater"], g0["time_max"])
    ctx.set_kernel_integrand(order, X, omega)
    return ctx


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 6, 7, 8])
def test_entries_match_oracle(flat_g0, order):
    t = 10.0
    X = capi.GfIndex(t, 0, 0, 0, 0)
    n = 64 if order <= 6 else 16
    u = _points(order, n, t, seed=100 + order)
    got, dom, _ = _kernel_ctx(flat_g0, order, X).eval_kernel_batch(u)
    K = O.KernelIntegrand(U.oracle_charge(flat_g0), O.gf_index(t, 0, 0))
    for p in range(n):
        ref, rdom = K.entries(u[p])
        assert dom[p] == rdom
        if rdom:
            assert _normwise(got[p], ref) <= 1e-11, p
        else:
            assert np.all(got[p] == 0)


def test_entries_order10_and_keep_u_false(flat_g0):
    t = 10.0
    X = capi.GfIndex(t, 0, 0, 0, 0)
    K = O.KernelIntegrand(U.oracle_charge(flat_g0), O.gf_index(t, 0, 0))
    u = _points(10, 4, t, seed=7)
    got, dom, _ = _kernel_ctx(flat_g0, 10, X).eval_kernel_batch(u)
    for p in range(4):
        ref, rdom = K.entries(u[p])
        assert dom[p] == rdom == 1
        assert _normwise(got[p], ref) <= 1e-11
    # a negative vertex time evaluated anyway (X at 9 keeps every g0 argument
    # inside the table's [-10, 10]; outside it the device yields NaN and the
    # host API throws, as the reference's g0 evaluation does)
    u = _points(3, 8, 9.0, seed=8)
    X9 = capi.GfIndex(9.0, 0, 0, 0, 0)
    got, dom, _ = _kernel_ctx(flat_g0, 3, X9).eval_kernel_batch(u, keep_u_hypercube=False)
    assert np.all(dom == 1)
    ref, _ = O.KernelIntegrand(U.oracle_charge(flat_g0), O.gf_index(9.0, 0, 0)).entries(u[3], keep_u_hypercube=False)
    assert _normwise(got[3], ref) <= 1e-11


@pytest.fixture(scope="module")
def dot_lead_default():
    from keldy_amd import impurity_oneband as kio
    return kio.G0Model(kio.G0ModelOmega({}), True)


def test_host_api_order1_order2_explicit(dot_lead_default):
    """wick_kernel_test.cpp:26-107 through IntegrandGKernel (reference names):
    sparse binner sizes 2 and 8, values equal to the explicit contractions."""
    from keldy_amd import impurity_oneband as kio
    m = dot_lead_default
    g = O.G0Table(m.g0_lesser, m.g0_greater, m.time_max)
    G = lambda a, b: O.g0_eval(g, a, b)
    P = lambda t, s, k: O.gf_index(t, s, k)
    X = P(5.0, UP, FWD)
    f = kio.IntegrandGKernel(kio.G0KeldyshContourT(m), kio.GfIndexT(5.0, UP, FWD))
    u = 2.0
    res, dom = f([u])
    assert dom == 1 and len(res.data) == 2
    dd = G(P(u, DOWN, FWD), P(u, DOWN, BWD))
    exp = {FWD: -G(P(u, UP, FWD), X) * dd, BWD: G(P(u, UP, BWD), X) * dd}
    for (t, k), v in res.data:
        assert t == u and abs(v - exp[k]) <= 1e-16
    v_ = 4.0
    res, dom = f([u, v_])
    assert dom == 1 and len(res.data) == 8
    sums = {}
    for (t, k), v in res.data:
        sums[(t, k)] = sums.get((t, k), 0) + v
    ref_sb, _ = O.KernelIntegrand(g, X)([u, v_])
    ref = {}
    for (t, k), v in ref_sb:
        ref[(t, k)] = ref.get((t, k), 0) + v
    assert sums.keys() == ref.keys()
    for key in ref:
        assert abs(sums[key] - ref[key]) <= 1e-15 + 1e-13 * abs(ref[key])
    assert f([1.0, -0.5])[1] == 0 and len(f([1.0, -0.5])[0].data) == 0


def _value_kernel_computer(n_bins=10):
    from keldy_amd import impurity_oneband as kio
    from keldy_amd import warpers as kw
    p = dict(beta=-1.0, bias_V=0.0, eps_d=0.0, Gamma=0.5, alpha=0.5, half_bandwidth=6.0, time_max=100.0,
             nr_time_points_gf=5001, bath_type="semicircle", ft_method="contour")
    model = kio.G0Model(kio.G0ModelOmega(p), False)
    c = kio.ComputeGfKernel(model, 100.0, 2, n_bins)
    c.warper.emplace_back(kw.WarperPlasmaUvT(100.0))
    c.warper.emplace_back(kw.make_product_1d_inverse_cube_alternate_interp_hybrid(2, 100.0, 1.0, 100000))
    return c, model


def test_value_kernel_regression_and_oracle():
    """compute_obs_kernel_test.cpp:12-60 (ValueKernel) through ComputeGfKernel;
    the same run on the oracle: bins normwise 1e-10, counts exact."""
    from tests.test_kernel_oracle import VALUE_KERNEL
    c, model = _value_kernel_computer()
    n = 1000000
    c.run(n)
    b = c.reduce_result()
    res = b.get_data() / b.get_bin_size()
    assert res.shape == (10, 2)
    assert np.max(np.abs(res[:, 0] - np.conj(res[:, 1]))) < 1e-12
    err = np.max(np.abs(res - VALUE_KERNEL))
    print(f"ValueKernel: max |result - reference array| = {err:.3e} (test bound 1e-6)")
    assert err < 1e-6
    g = O.G0Table(model.g0_lesser[0, 0], model.g0_greater[0, 0], model.time_max, 0.5)
    K = O.KernelIntegrand(g, O.gf_index(100.0, UP, FWD))
    tr = O.Train([O.plasma(100.0), O.product([O.hybrid_map(1, 1.0, 100.0, 100000),
                                              O.hybrid_map(3, 1.0, 100.0, 100000)])])
    bins, counts, dropped, ndom = O.kernel_run(K, tr, O.Sobol(2, 0), n, 100.0, 10)
    assert _normwise(b.get_data() * n, bins) <= 1e-10
    assert np.array_equal(b.get_nr_values_added(), counts.astype(np.int64))
    assert b.get_nr_values_dropped() == dropped and c.reduce_nr_points_in_domain() == ndom
    assert b.get_nr_bins(0) == 10 and b.get_nr_bins(1) == 2


@pytest.mark.parametrize("order", [4, 8])
def test_run_binned_vs_oracle(flat_g0, order):
    """kd_run with kernel bins against the oracle's compute_gf_kernel run on the
    same scrambled Sobol points and chain."""
    t = 10.0
    nb, n = 16, (1 << 14 if order == 4 else 1 << 10)
    train, (descs, maps, tabs) = U.chain([("plasma", t), ("map", "inverse", t, 1.0, 1)])
    ctx = _kernel_ctx(flat_g0, order, capi.GfIndex(t, 0, 0, 0, 0))
    ctx.set_warper_chain(descs, maps, tabs)
    ctx.set_kernel_bins(0.0, t, nb)
    s = O.Sobol(order, 0, True, True, 5)
    Csr, shfr = s.tables()
    ctx.set_sobol(Csr, shfr, 0)
    acc = ctx.run(n)
    vals, cnt, dropped = ctx.kernel_bins()
    K = O.KernelIntegrand(U.oracle_charge(flat_g0), O.gf_index(t, 0, 0))
    bins, counts, rdrop, ndom = O.kernel_run(K, train, O.Sobol(order, 0, True, True, 5), n, t, nb)
    assert acc.n_in_domain == ndom and dropped == rdrop
    assert np.array_equal(cnt, counts)
    assert _normwise(vals, bins) <= 1e-10
    # size-independent: the accumulator holds the total of every entry
    tot = complex(acc.re, acc.im)
    assert abs(tot - vals.sum()) <= 1e-12 * np.abs(vals).sum()
    assert cnt.sum() == (1 << order) * order * ndom - rdrop


def test_single_omega_host_api_vs_oracle():
    """kernel_single_omega test (wick_kernel_single_omega_test.cpp:8-32) inputs:
    flatband beta 1, X = (5, up, fwd), omega = 1, times {2, 0} and {4.9, 4.8};
    then ComputeGfKernelSingleOmega's run against the oracle's phase-weighted sum."""
    from keldy_amd import impurity_oneband as kio
    from keldy_amd import warpers as kw
    p = dict(beta=1.0, bias_V=0.0, eps_d=0.0, Gamma=1.0, time_max=10.0, nr_time_points_gf=1000, alpha=0.0,
             bath_type="flatband", ft_method="fft")
    m = kio.G0Model(kio.G0ModelOmega(p), False)
    g = O.G0Table(m.g0_lesser[0, 0], m.g0_greater[0, 0], m.time_max)
    f = kio.IntegrandGKernelSingleOmega(kio.G0KeldyshContourT(m), kio.GfIndexT(5.0, UP, FWD), 1.0)
    K = O.KernelIntegrand(g, O.gf_index(5.0, UP, FWD))
    for times in ([2.0, 0.0], [4.9, 4.8]):
        v, dom = f(times)
        r, rdom = K.omega(times, 1.0)
        assert dom == rdom == 1 and abs(v - r) <= 1e-13 * abs(r)
    c = kio.ComputeGfKernelSingleOmega(m, 5.0, 1.0, 3)
    c.warper.emplace_back(kw.WarperPlasmaUvT(5.0))
    c.run(4096)
    l = O.Sobol(3, 0).points(4096)
    tr = O.Train([O.plasma(5.0)])
    s = 0
    for li in l:
        u, j = tr.reverse(li)
        v, d = K.omega(u, 1.0)
        s += v * j
    assert abs(c.reduce_result() * 4096 - s) <= 1e-11 * abs(s)


def test_run_batch_kernel_times():
    """Several kernel computers (different external times) in one graph launch
    equal their own runs."""
    from keldy_amd import impurity_oneband as kio
    from keldy_amd import warpers as kw
    p = dict(beta=-1.0, bias_V=0.0, eps_d=0.0, Gamma=1.0, time_max=10.0, nr_time_points_gf=10001, alpha=0.0,
             bath_type="flatband", ft_method="analytic")
    m = kio.G0Model(kio.G0ModelOmega(p), False)

    def make(t):
        c = kio.ComputeGfKernel(m, t, 6, 20)
        c.warper.emplace_back(kw.WarperPlasmaUvT(t))
        c.warper.emplace_back(kw.make_product_1d_simple_inverse(t, 1.0))
        c.reset_rng("sobol", 0, True, True, 3)
        return c

    times = [2.0, 5.0, 10.0]
    batch = [make(t) for t in times]
    kio.run_batch(batch, [1 << 14] * 3)
    for i, t in enumerate(times):
        ref = make(t)
        ref.run(1 << 14)
        a, b = batch[i].reduce_result().get_data(), ref.reduce_result().get_data()
        assert _normwise(a, b) <= 1e-12
