"""Host-side flow warper (WarperFlowT): same map as the oracle's definition,
invertible, Jacobians consistent; chain rules of the warper train."""
import numpy as np
import pytest

from oracle import flow_oracle as F
from tests.flow_util import flow_params

kw = pytest.importorskip("keldy_amd.warpers")


@pytest.mark.parametrize("n,layers", [(3, 2), (8, 6)])
def test_host_map_vs_oracle(n, layers):
    p = flow_params(n, layers, 13)
    w = kw.WarperFlowT(n, layers, p.astype(np.float32))
    rng = np.random.default_rng(2)
    l = rng.uniform(0.01, 0.99, (40, n))
    xo, ljo = F.forward(p, n, layers, l, emulate=True)
    for i in range(len(l)):
        x, j = w.map_reverse(list(l[i]))
        np.testing.assert_allclose(x, xo[i], rtol=0, atol=1e-6)
        assert abs(np.log(j) - ljo[i]) < 1e-5
        l2, jf = w.map_forward(x)
        np.testing.assert_allclose(l2, l[i], rtol=0, atol=1e-10)
        assert abs(j * jf - 1.0) < 1e-9


def test_flow_param_checks():
    with pytest.raises(RuntimeError, match="parameter count"):
        kw.WarperFlowT(3, 2, np.zeros(10, dtype=np.float32))
    with pytest.raises(RuntimeError, match="layers"):
        kw.WarperFlowT(3, 0, np.zeros(0, dtype=np.float32))
    w = kw.WarperFlowT(3, 2, np.zeros(F.n_params(3, 2), dtype=np.float32))
    assert w.order() == 3 and w.n_layers() == 2 and len(w.params()) == F.n_params(3, 2)
    # all-zero parameters: the identity map
    x, j = w.map_reverse([0.2, 0.5, 0.7])
    np.testing.assert_allclose(x, [0.2, 0.5, 0.7], atol=1e-14)
    assert abs(j - 1.0) < 1e-12


def test_flow_in_train_host_map():
    """The flow is applied first (last added): train.map_reverse == other(flow(l))."""
    n, layers = 3, 2
    p = flow_params(n, layers, 5).astype(np.float32)
    fw = kw.WarperFlowT(n, layers, p)
    tr = kw.WarperTrainT()
    tr.emplace_back(kw.WarperPlasmaUvT(10.0))
    tr.emplace_back(fw)
    l = [0.3, 0.6, 0.1]
    x, jf = fw.map_reverse(l)
    u, jt = tr.map_reverse(l)
    np.testing.assert_allclose(u, kw.WarperPlasmaUvT(10.0).map_reverse(x)[0], atol=1e-13)
    assert abs(jt - jf) < 1e-12 * abs(jf)


def test_conditioned_host_map_vs_oracle():
    """The time-conditioned flow on the host: T_c of the oracle at each value,
    with_cond moves to another member of the family, the parameter count
    carries the conditioning column."""
    n, layers = 6, 3
    p0 = flow_params(n, layers, 17)
    rng = np.random.default_rng(5)
    blocks, off = [], 0
    for k in range(layers):
        size = F.layer_size(n, k)
        blocks += [p0[off:off + size], rng.normal(0, 0.5, F.H)]
        off += size
    p = np.concatenate(blocks).astype(np.float32)
    with pytest.raises(RuntimeError, match="parameter count"):
        kw.WarperFlowT(n, layers, p)  # unconditioned count expected
    w = kw.WarperFlowT(n, layers, p, n_cond=1, cond=0.25)
    assert w.n_cond() == 1 and w.cond() == 0.25
    l = rng.uniform(0.01, 0.99, (20, n))
    for cv, ww in ((0.25, w), (0.9, w.with_cond(0.9))):
        xo, ljo = F.forward(p.astype(np.float64), n, layers, l, emulate=True, nc=1, cond=cv)
        for i in range(len(l)):
            x, j = ww.map_reverse(list(l[i]))
            np.testing.assert_allclose(x, xo[i], rtol=0, atol=1e-6)
            assert abs(np.log(j) - ljo[i]) < 1e-5
            l2, jf = ww.map_forward(x)
            np.testing.assert_allclose(l2, l[i], rtol=0, atol=1e-10)
