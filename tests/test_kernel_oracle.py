"""The oracle's kernel method (integrand_g_kernel, wick_kernel.cpp:42-138;
first_row_expansion, det_expansion.hpp:44-74) pinned against the reference's
own tests and against what the mathematics fixes.

- first_row_expansion == brute-force cofactors (explicit minors), any size;
- orders 1 and 2 == the explicit Wick contractions of wick_kernel_test.cpp:26-107
  (per (vertex time, Keldysh index) sums, as that test compares them);
- compute_gf_kernel at order 2 on 10^6 Sobol points == the regression array of
  compute_obs_kernel_test.cpp:12-60 within its 1e-6, and lesser == conj(greater)
  (particle-hole symmetry) within its 1e-12;
- single-omega (wick_kernel_single_omega.cpp:41-140) == phase-weighted entries.
"""
import numpy as np
import pytest

from oracle import oracle as O

UP, DOWN, FWD, BWD = 0, 1, 0, 1


def test_first_row_expansion_is_the_cofactor_row():
    rng = np.random.default_rng(3)
    for m in range(1, 8):
        a = rng.normal(size=(m, m)) + 1j * rng.normal(size=(m, m))
        x, d = O.first_row_expansion(a)
        cof = np.array([(-1) ** j * (np.linalg.det(np.delete(np.delete(a, 0, 0), j, 1)) if m > 1 else 1.0)
                        for j in range(m)])
        assert np.max(np.abs(x - cof)) <= 1e-13 * np.max(np.abs(cof))
        assert abs(d - np.linalg.det(a)) <= 1e-13 * abs(d)
    # a zero in the (0, 0) corner needs the pivoting
    a = rng.normal(size=(4, 4)) + 0j
    a[0, 0] = 0.0
    x, _ = O.first_row_expansion(a)
    cof = np.array([(-1) ** j * np.linalg.det(np.delete(np.delete(a, 0, 0), j, 1)) for j in range(4)])
    assert np.max(np.abs(x - cof)) <= 1e-13 * np.max(np.abs(cof))
    with pytest.raises(RuntimeError):
        O.first_row_expansion(np.zeros((3, 3)))


@pytest.fixture(scope="module")
def dot_lead_default():
    """model_param_t defaults with the dot-lead g0 (wick_kernel_test.cpp:27-29)."""
    from keldy_amd import impurity_oneband as kio
    m = kio.G0Model(kio.G0ModelOmega({}), True)
    return O.G0Table(m.g0_lesser, m.g0_greater, m.time_max)


def _sums(entries, times):
    """per (vertex, Keldysh index) sums of the entries"""
    n = len(times)
    out = np.zeros((n, 2), dtype=complex)
    for c in range(1 << n):
        for i in range(n):
            out[i, (c >> i) & 1] += entries[c, i]
    return out


def test_order1_explicit(dot_lead_default):
    g = dot_lead_default
    X = O.gf_index(5.0, UP, FWD)
    u = 2.0
    K = O.KernelIntegrand(g, X)
    e, dom = K.entries([u])
    assert dom == 1 and e.shape == (2, 1)

    def pt(spin, k):
        return O.gf_index(u, spin, k)

    dd = O.g0_eval(g, pt(DOWN, FWD), pt(DOWN, BWD))
    assert abs(e[0, 0] - (-O.g0_eval(g, pt(UP, FWD), X) * dd)) <= 1e-16
    assert abs(e[1, 0] - O.g0_eval(g, pt(UP, BWD), X) * dd) <= 1e-16


def test_order2_explicit(dot_lead_default):
    g = dot_lead_default
    X = O.gf_index(5.0, UP, FWD)
    u, v = 2.0, 4.0
    G = lambda a, b: O.g0_eval(g, a, b)
    P = lambda t, s, k: O.gf_index(t, s, k)

    def det_down(a, b):
        return G(P(u, DOWN, FWD), P(u, DOWN, BWD)) * G(P(v, DOWN, FWD), P(v, DOWN, BWD)) \
            - G(P(v, DOWN, b), P(u, DOWN, a)) * G(P(u, DOWN, a), P(v, DOWN, b))

    def det_prod_u(a, b):
        s = 1 if a == b else -1
        return s * (G(P(u, UP, a), P(v, UP, b)) * G(P(v, UP, b), X) * det_down(a, b)
                    + G(P(v, UP, FWD), P(v, UP, BWD)) * G(P(u, UP, a), X) * G(P(v, DOWN, b), P(u, DOWN, a))
                    * G(P(u, DOWN, a), P(v, DOWN, b)))

    def det_prod_v(a, b):
        s = 1 if a == b else -1
        return -s * (-G(P(v, UP, b), P(u, UP, a)) * G(P(u, UP, a), X) * det_down(a, b)
                     - G(P(u, UP, FWD), P(u, UP, BWD)) * G(P(v, UP, b), X) * G(P(v, DOWN, b), P(u, DOWN, a))
                     * G(P(u, DOWN, a), P(v, DOWN, b)))

    exp = np.zeros((2, 2), dtype=complex)
    for a in (FWD, BWD):
        for b in (FWD, BWD):
            exp[0, a] += det_prod_u(a, b)
            exp[1, b] += det_prod_v(a, b)
    e, dom = O.KernelIntegrand(g, X).entries([u, v])
    assert dom == 1 and e.shape == (4, 2)
    got = _sums(e, [u, v])
    assert np.max(np.abs(got - exp)) <= 1e-16 + 1e-14 * np.max(np.abs(exp))
    # the reference's sparse-binner form: 8 entries in accumulate order
    sb, _ = O.KernelIntegrand(g, X)([u, v])
    assert len(sb) == 8 and sb[1][0] == (v, 0) and sb[3][0] == (v, 0) and sb[5][0] == (v, 1)


def test_domain_and_single_omega(dot_lead_default):
    g = dot_lead_default
    K = O.KernelIntegrand(g, O.gf_index(5.0, UP, FWD))
    assert K([1.0, -0.5]) == ([], 0)
    t = [3.0, 1.5, 0.5]
    e, _ = K.entries(t)
    for w in (0.0, 1.3):
        ref = sum(e[c, i] * np.exp(1j * w * t[i]) for c in range(8) for i in range(3))
        got, dom = K.omega(t, w)
        assert dom == 1 and abs(got - ref) <= 1e-14 * abs(ref)


def _value_kernel_params():
    return dict(beta=-1.0, bias_V=0.0, eps_d=0.0, Gamma=0.5, alpha=0.5, half_bandwidth=6.0, time_max=100.0,
                nr_time_points_gf=5001, bath_type="semicircle", ft_method="contour")


VALUE_KERNEL = np.array([
    [-3.67742471e-09 + 1.63164724e-04j, -3.67742341e-09 - 1.63164724e-04j],
    [5.50958865e-09 + 2.09202817e-05j, 5.50959042e-09 - 2.09202817e-05j],
    [-7.41242821e-09 + 2.68650657e-05j, -7.41242621e-09 - 2.68650657e-05j],
    [9.25484502e-09 + 3.59362761e-05j, 9.25484734e-09 - 3.59362761e-05j],
    [-1.08196801e-08 + 5.05717360e-05j, -1.08196774e-08 - 5.05717360e-05j],
    [1.16796193e-08 + 7.66168036e-05j, 1.16796227e-08 - 7.66168036e-05j],
    [-1.50691787e-08 + 1.30292898e-04j, -1.50691743e-08 - 1.30292898e-04j],
    [-9.34236787e-07 + 2.77177129e-04j, -9.34236781e-07 - 2.77177129e-04j],
    [-2.00377184e-04 + 1.06584119e-03j, -2.00377184e-04 - 1.06584119e-03j],
    [2.01317477e-04 - 1.85006394e-03j, 2.01317477e-04 + 1.85006394e-03j]])


def test_value_kernel_regression():
    """compute_obs_test ValueKernel (compute_obs_kernel_test.cpp:12-60): order 2,
    t = 100, 10 bins, [plasma_uv(100), inverse_cube_alternate_interp_hybrid(2,
    100, 1, 1e5)], 10^6 unshifted Sobol points; data / bin_size."""
    from keldy_amd import impurity_oneband as kio
    tmax = 100.0
    m = kio.G0Model(kio.G0ModelOmega(_value_kernel_params()), False)
    g = O.G0Table(m.g0_lesser[0, 0], m.g0_greater[0, 0], m.time_max, 0.5)
    K = O.KernelIntegrand(g, O.gf_index(tmax, UP, FWD))
    tr = O.Train([O.plasma(tmax), O.product([O.hybrid_map(1, 1.0, tmax, 100000),
                                             O.hybrid_map(3, 1.0, tmax, 100000)])])
    n = 1000000
    bins, counts, dropped, ndom = O.kernel_run(K, tr, O.Sobol(2, 0), n, tmax, 10)
    res = bins / n / (tmax / 10)
    assert np.max(np.abs(res[:, 0] - np.conj(res[:, 1]))) < 1e-12
    assert np.max(np.abs(res - VALUE_KERNEL)) < 1e-6
    assert dropped == 0 and np.all(counts[:, 0] == counts[:, 1]) and counts.sum() == 2 * 4 * ndom
