"""Pins the flow oracle (oracle/flow_oracle.py) by invariants, not by restating it:
inverse(forward(l)) == l, log J against a finite-difference Jacobian
determinant, log q == -log J at the preimage, the analytic gradient against
central differences and against torch autograd, and a planted mistake that
these checks catch."""
import numpy as np
import pytest

from oracle import flow_oracle as F
from tests.flow_util import flow_params


def _params(n, layers, seed):
    return flow_params(n, layers, seed)


@pytest.mark.parametrize("n,layers", [(2, 2), (3, 4), (8, 6)])
def test_inverse_of_forward(n, layers):
    p = _params(n, layers, 7)
    l = np.random.default_rng(3).uniform(0.02, 0.98, (200, n))
    x, logj = F.forward(p, n, layers, l)
    l2, logq = F.inverse(p, n, layers, x)
    np.testing.assert_allclose(l2, l, rtol=0, atol=1e-11)
    np.testing.assert_allclose(logq, -logj, rtol=0, atol=1e-9)


def test_identity_at_init():
    n, layers = 8, 6
    p = F.init_params(n, layers, 11)
    l = np.random.default_rng(0).uniform(0, 1, (50, n))
    x, logj = F.forward(p, n, layers, l)
    np.testing.assert_allclose(x, l, atol=1e-12)
    np.testing.assert_allclose(logj, 0.0, atol=1e-9)


@pytest.mark.parametrize("n,layers", [(2, 2), (4, 3), (8, 6)])
def test_logj_vs_finite_difference_determinant(n, layers):
    p = _params(n, layers, 5)
    rng = np.random.default_rng(9)
    h = 1e-6
    for _ in range(4):
        l = rng.uniform(0.1, 0.9, n)
        J = np.empty((n, n))
        for j in range(n):
            e = np.zeros(n)
            e[j] = h
            xp, _ = F.forward(p, n, layers, l + e)
            xm, _ = F.forward(p, n, layers, l - e)
            J[:, j] = (xp[0] - xm[0]) / (2 * h)
        _, logj = F.forward(p, n, layers, l)
        assert abs(np.log(abs(np.linalg.det(J))) - logj[0]) < 1e-6


def _fd_grad(p, n, layers, x, w, idx, h=1e-6):
    g = np.empty(len(idx))
    for q, i in enumerate(idx):
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        g[q] = (F.loss_and_grad(pp, n, layers, x, w)[0] - F.loss_and_grad(pm, n, layers, x, w)[0]) / (2 * h)
    return g


@pytest.mark.parametrize("n,layers", [(3, 2), (8, 6)])
def test_gradient_vs_central_differences(n, layers):
    p = _params(n, layers, 21)
    rng = np.random.default_rng(4)
    x = rng.uniform(0.05, 0.95, (64, n))
    w = rng.uniform(0.0, 2.0, 64)
    loss, g = F.loss_and_grad(p, n, layers, x, w)
    # the loss also equals the weighted mean of -log q from inverse()
    _, logq = F.inverse(p, n, layers, x)
    assert abs(loss - np.mean(w / w.mean() * -logq)) < 1e-12
    # every parameter block of every layer, sampled
    idx = []
    off = 0
    for k in range(layers):
        size = F.layer_size(n, k)
        idx += list(off + rng.choice(size, 12, replace=False))
        off += size
    gfd = _fd_grad(p, n, layers, x, w, idx)
    np.testing.assert_allclose(g[idx], gfd, rtol=2e-5, atol=2e-8)


def test_gradient_vs_torch_autograd():
    torch = pytest.importorskip("torch")
    n, layers = 5, 4
    p = _params(n, layers, 33)
    rng = np.random.default_rng(8)
    x = rng.uniform(0.05, 0.95, (40, n))
    w = rng.uniform(0.1, 1.0, 40)
    loss, g = F.loss_and_grad(p, n, layers, x, w)

    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    xt = torch.tensor(x, dtype=torch.float64)
    v = torch.log(xt) - torch.log1p(-xt)
    nlq = torch.sum(torch.log(xt) + torch.log1p(-xt), dim=1)
    off = 0
    blocks = []
    for k in range(layers):
        na, nb = F.n_a(n, k), F.n_b(n, k)
        d = {}
        for name, shape in (("W1", (F.H, na)), ("b1", (F.H,)), ("W2", (F.H, F.H)), ("b2", (F.H,)),
                            ("W3", (2 * nb, F.H)), ("b3", (2 * nb,))):
            size = int(np.prod(shape))
            d[name] = pt[off:off + size].reshape(shape)
            off += size
        blocks.append(d)
    for k in reversed(range(layers)):
        A, B = F.dims_a(n, k), F.dims_b(n, k)
        L = blocks[k]
        h1 = torch.relu(v[:, A] @ L["W1"].T + L["b1"])
        h2 = torch.relu(h1 @ L["W2"].T + L["b2"])
        o = h2 @ L["W3"].T + L["b3"]
        nb = len(B)
        s, t = F.s_scale(layers) * torch.tanh(o[:, :nb]), o[:, nb:]
        vB = (v[:, B] - t) * torch.exp(-s)
        cols = [None] * n
        for q, i in enumerate(A):
            cols[i] = v[:, i]
        for q, i in enumerate(B):
            cols[i] = vB[:, q]
        v = torch.stack(cols, dim=1)
        nlq = nlq + s.sum(dim=1)
    nlq = nlq - torch.sum(torch.log(torch.sigmoid(v)) + torch.log(torch.sigmoid(-v)), dim=1)
    wt = torch.tensor(w / w.mean(), dtype=torch.float64)
    lt = torch.mean(wt * nlq)
    lt.backward()
    assert abs(lt.item() - loss) < 1e-12
    np.testing.assert_allclose(g, pt.grad.numpy(), rtol=1e-10, atol=1e-13)


def test_planted_sign_error_is_caught(monkeypatch):
    """A plausible mistake in the backprop (wrong sign of the log-det term)
    must fail the central-difference check."""
    n, layers = 3, 2
    p = _params(n, layers, 2)
    rng = np.random.default_rng(1)
    x = rng.uniform(0.05, 0.95, (16, n))
    w = np.ones(16)
    orig = F.loss_and_grad

    def wrong(params, n_, layers_, x_, w_, emulate=False):
        loss, g = orig(params, n_, layers_, x_, w_, emulate)
        # flip the bias-gradient of the last layer's s outputs
        L = F.unpack(g, n_, layers_)[-1]
        L["b3"][: F.n_b(n_, layers_ - 1)] *= -1
        return loss, g

    _, g_bad = wrong(p, n, layers, x, w)
    off = F.n_params(n, layers) - 2 * F.n_b(n, layers - 1)
    idx = [off]
    gfd = _fd_grad(p, n, layers, x, w, idx)
    assert not np.allclose(g_bad[idx], gfd, rtol=1e-3)
    np.testing.assert_allclose(orig(p, n, layers, x, w)[1][idx], gfd, rtol=1e-5)


def test_bf16_rounding():
    v = np.array([1.0, 1.0078125, 1.00390625, 1.01171875, -3.14159265, 0.0], dtype=np.float64)
    r = F.bf16(v)
    # 7 stored mantissa bits: 1 + 2^-7 is representable; 1 + 2^-8 ties to even (1.0);
    # 1 + 3*2^-8 ties to even (1 + 2^-6)
    np.testing.assert_array_equal(r[:4], [1.0, 1.0078125, 1.0, 1.015625])
    assert abs(r[4] - (-3.140625)) < 1e-12 and r[5] == 0.0


# ---------------------------------------------------------------------------
# the time-conditioned flow (nc = 1): one extra input c to every coupling
# layer's conditioner, h1 = relu(W1 a + Wc c + b1)

def _cparams(n, layers, seed):
    """Conditioned parameters: the unconditioned ones of flow_params with a
    non-zero Wc appended to every layer."""
    p0 = flow_params(n, layers, seed)
    rng = np.random.default_rng(seed + 7)
    out, off = [], 0
    for k in range(layers):
        size = F.layer_size(n, k)
        out.append(p0[off:off + size])
        out.append(rng.normal(0, 0.5, F.H))
        off += size
    return np.concatenate(out).astype(np.float32).astype(np.float64)


def test_conditioned_layout_and_zero_wc_reduces_to_unconditioned():
    n, layers = 5, 3
    assert F.n_params(n, layers, 1) == F.n_params(n, layers) + layers * F.H
    p = _cparams(n, layers, 3)
    for L in F.unpack(p, n, layers, 1):
        L["Wc"][...] = 0.0
    p0 = np.concatenate([np.concatenate([L[k].ravel() for k in ("W1", "b1", "W2", "b2", "W3", "b3")])
                         for L in F.unpack(p, n, layers, 1)])
    l = np.random.default_rng(0).uniform(0.05, 0.95, (30, n))
    x1, j1 = F.forward(p, n, layers, l, nc=1, cond=0.7)
    x0, j0 = F.forward(p0, n, layers, l)
    np.testing.assert_array_equal(x1, x0)
    np.testing.assert_array_equal(j1, j0)


@pytest.mark.parametrize("n,layers", [(3, 2), (10, 4)])
def test_conditioned_inverse_and_dependence_on_c(n, layers):
    p = _cparams(n, layers, 5)
    rng = np.random.default_rng(2)
    l = rng.uniform(0.02, 0.98, (100, n))
    c = rng.uniform(0.0, 1.0, 100)
    x, logj = F.forward(p, n, layers, l, nc=1, cond=c)
    l2, logq = F.inverse(p, n, layers, x, nc=1, cond=c)
    np.testing.assert_allclose(l2, l, rtol=0, atol=1e-11)
    np.testing.assert_allclose(logq, -logj, rtol=0, atol=1e-9)
    # the map depends on c: another value moves the points
    xb, _ = F.forward(p, n, layers, l, nc=1, cond=c + 0.3)
    assert np.max(np.abs(xb - x)) > 1e-4


@pytest.mark.parametrize("n,layers", [(3, 2), (10, 4)])
def test_conditioned_gradient_vs_central_differences(n, layers):
    p = _cparams(n, layers, 21)
    rng = np.random.default_rng(4)
    x = rng.uniform(0.05, 0.95, (64, n))
    w = rng.uniform(0.0, 2.0, 64)
    c = rng.uniform(0.0, 1.0, 64)
    loss, g = F.loss_and_grad(p, n, layers, x, w, nc=1, cond=c)
    _, logq = F.inverse(p, n, layers, x, nc=1, cond=c)
    assert abs(loss - np.mean(w / w.mean() * -logq)) < 1e-12
    idx, off = [], 0
    for k in range(layers):
        size = F.layer_size(n, k, 1)
        idx += list(off + rng.choice(size - F.H, 8, replace=False))   # the unconditioned blocks
        idx += list(off + size - F.H + rng.choice(F.H, 8, replace=False))  # Wc
        off += size
    h = 1e-6
    gfd = []
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        gfd.append((F.loss_and_grad(pp, n, layers, x, w, nc=1, cond=c)[0]
                    - F.loss_and_grad(pm, n, layers, x, w, nc=1, cond=c)[0]) / (2 * h))
    np.testing.assert_allclose(g[idx], gfd, rtol=2e-5, atol=2e-8)


def test_conditioned_gradient_vs_torch_autograd():
    torch = pytest.importorskip("torch")
    n, layers = 5, 4
    p = _cparams(n, layers, 33)
    rng = np.random.default_rng(8)
    x = rng.uniform(0.05, 0.95, (40, n))
    w = rng.uniform(0.1, 1.0, 40)
    c = rng.uniform(0.0, 1.0, 40)
    loss, g = F.loss_and_grad(p, n, layers, x, w, nc=1, cond=c)
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    xt = torch.tensor(x, dtype=torch.float64)
    ct = torch.tensor(c, dtype=torch.float64)
    v = torch.log(xt) - torch.log1p(-xt)
    nlq = torch.sum(torch.log(xt) + torch.log1p(-xt), dim=1)
    off, blocks = 0, []
    for k in range(layers):
        na, nb = F.n_a(n, k), F.n_b(n, k)
        d = {}
        for name, shape in (("W1", (F.H, na)), ("b1", (F.H,)), ("W2", (F.H, F.H)), ("b2", (F.H,)),
                            ("W3", (2 * nb, F.H)), ("b3", (2 * nb,)), ("Wc", (F.H,))):
            size = int(np.prod(shape))
            d[name] = pt[off:off + size].reshape(shape)
            off += size
        blocks.append(d)
    for k in reversed(range(layers)):
        A, B = F.dims_a(n, k), F.dims_b(n, k)
        L = blocks[k]
        h1 = torch.relu(v[:, A] @ L["W1"].T + ct[:, None] * L["Wc"][None, :] + L["b1"])
        h2 = torch.relu(h1 @ L["W2"].T + L["b2"])
        o = h2 @ L["W3"].T + L["b3"]
        nb = len(B)
        s, t = F.s_scale(layers) * torch.tanh(o[:, :nb]), o[:, nb:]
        vB = (v[:, B] - t) * torch.exp(-s)
        cols = [None] * n
        for q, i in enumerate(A):
            cols[i] = v[:, i]
        for q, i in enumerate(B):
            cols[i] = vB[:, q]
        v = torch.stack(cols, dim=1)
        nlq = nlq + s.sum(dim=1)
    nlq = nlq - torch.sum(torch.log(torch.sigmoid(v)) + torch.log(torch.sigmoid(-v)), dim=1)
    lt = torch.mean(torch.tensor(w / w.mean(), dtype=torch.float64) * nlq)
    lt.backward()
    assert abs(lt.item() - loss) < 1e-12
    np.testing.assert_allclose(g, pt.grad.numpy(), rtol=1e-10, atol=1e-13)
