"""Affine-coupling flow warper: FP64 numpy restatement (TEST INFRASTRUCTURE ONLY).

The checker for the device flow kernels (keldy_amd/csrc/kernels/kd_flow.hip).
keldy has no learned warper; this map is the build's own design, so the oracle
follows the definition below, not a reference file.  It is pinned by
invertibility (inverse(forward(l)) == l), by the log-Jacobian against a finite
-difference Jacobian determinant, and by the analytic gradient against central
finite differences and against torch autograd (tests/test_flow_oracle.py).

Map T: l in (0,1)^n -> x in (0,1)^n.
    y = logit(clamp(l)); for k = 0 .. L-1 (coupling layer k):
        A_k = {i : (i + k) even} (conditioner), B_k = the other dims
        a = y[A_k];  h1 = relu(W1 a + b1);  h2 = relu(W2 h1 + b2);  o = W3 h2 + b3
        s = S tanh(o[:nB]);  t = o[nB:];   y[B_k] = y[B_k] * exp(s) + t
    S = 0.9 ln 2 / ceil(L/2) (float32), which keeps the compounded logit-space
    contraction of every coordinate above 1/2 and so the warped estimator's
    variance finite near the faces of the cube
    x = sigmoid(y)
    log J(l) = -sum log(l (1-l)) + sum_k sum s + sum log(x (1-x))
Parameter vector, per layer: W1[H][nA], b1[H], W2[H][H], b2[H], W3[2nB][H], b3[2nB],
and with one conditioning input c (nc = 1; the time-conditioned flow) Wc[H]:
the input layer becomes h1 = relu(W1 a + Wc c + b1), c the same for every
coupling layer (the flow is then a family of maps T_c, one per value of c).

With bf16=True the operands the device feeds to its bf16 MFMA GEMMs (h1, W2,
h2, W3) are rounded to bfloat16 (round to nearest even) first, as the device
does; everything else stays FP64 here (the device keeps the coordinates and
the Jacobian in FP64 and the network arithmetic in FP32).

Training loss (weighted maximum likelihood; w = |integrand| samples at x):
    loss = mean_p w~_p * (-log q(x_p)),  w~ = w / mean(w),
    -log q(x) = log J(T^{-1}(x)) = -sum log(l(1-l)) + sum_k sum s + sum log(x(1-x))
"""
import numpy as np

H = 64
CLAMP = 2.0 ** -40


def n_a(n, k):
    return (n + 1 - (k & 1)) // 2


def n_b(n, k):
    return n - n_a(n, k)


def dims_a(n, k):
    return [i for i in range(n) if (i + k) % 2 == 0]


def dims_b(n, k):
    return [i for i in range(n) if (i + k) % 2 == 1]


def layer_size(n, k, nc=0):
    na, nb = n_a(n, k), n_b(n, k)
    return H * na + H + H * H + H + 2 * nb * H + 2 * nb + nc * H


def n_params(n, layers, nc=0):
    return sum(layer_size(n, k, nc) for k in range(layers))


def s_scale(layers):
    return float(np.float32(0.62383246) / np.float32((layers + 1) // 2))


def unpack(params, n, layers, nc=0):
    """-> list of dicts W1, b1, W2, b2, W3, b3 (and Wc) (views into params)."""
    out, off = [], 0
    for k in range(layers):
        na, nb = n_a(n, k), n_b(n, k)
        d = {}
        shapes = [("W1", (H, na)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)),
                  ("W3", (2 * nb, H)), ("b3", (2 * nb,))] + ([("Wc", (H,))] if nc else [])
        for name, shape in shapes:
            size = int(np.prod(shape))
            d[name] = params[off:off + size].reshape(shape)
            off += size
        out.append(d)
    assert off == len(params)
    return out


def bf16(a):
    """Round float values to bfloat16 (nearest, ties to even), returned as float64."""
    a32 = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _net(L, a, emulate, S, c=None):
    """a: (P, nA), c: (P,) conditioning or None -> s, t (P, nB) and the
    intermediates for backprop."""
    z1 = a @ L["W1"].T + L["b1"]
    if "Wc" in L:
        z1 = z1 + c[:, None] * L["Wc"][None, :]
    h1 = np.maximum(z1, 0.0)
    W2 = L["W2"]
    W3 = L["W3"]
    if emulate:
        h1 = bf16(h1)
        W2 = bf16(W2)
    z2 = h1 @ W2.T + L["b2"]
    h2 = np.maximum(z2, 0.0)
    if emulate:
        h2 = bf16(h2)
        W3 = bf16(W3)
    o = h2 @ W3.T + L["b3"]
    nb = o.shape[1] // 2
    th = np.tanh(o[:, :nb])
    return S * th, o[:, nb:], dict(a=a, z1=z1, h1=h1, z2=z2, h2=h2, th=th)


def _logit(l):
    l = np.clip(l, CLAMP, 1.0 - CLAMP)
    return np.log(l) - np.log1p(-l), l


def _log_sig_prime(y):
    """log(sigmoid(y) (1 - sigmoid(y))) = -|y| - 2 log1p(exp(-|y|))"""
    ay = np.abs(y)
    return -ay - 2.0 * np.log1p(np.exp(-ay))


def _sigmoid(y):
    return 1.0 / (1.0 + np.exp(-y))


def _cond(cond, P, emulate):
    if cond is None:
        return None
    c = np.broadcast_to(np.asarray(cond, dtype=np.float64), (P,)).copy()
    return c.astype(np.float32).astype(np.float64) if emulate else c


def forward(params, n, layers, l, emulate=False, nc=0, cond=None):
    """T(l) (T_c(l) with nc = 1, cond a scalar or one value per point): returns
    x (P, n) and log J (P,)."""
    Ls = unpack(np.asarray(params, dtype=np.float64), n, layers, nc)
    y, lc = _logit(np.array(l, dtype=np.float64, ndmin=2))
    c = _cond(cond, y.shape[0], emulate)
    logj = -np.sum(np.log(lc) + np.log1p(-lc), axis=1)
    for k in range(layers):
        A, B = dims_a(n, k), dims_b(n, k)
        a = y[:, A].astype(np.float32).astype(np.float64) if emulate else y[:, A]
        s, t, _ = _net(Ls[k], a, emulate, s_scale(layers), c)
        y[:, B] = y[:, B] * np.exp(s) + t
        logj += s.sum(axis=1)
    logj += _log_sig_prime(y).sum(axis=1)
    return _sigmoid(y), logj


def inverse(params, n, layers, x, emulate=False, nc=0, cond=None):
    """T^{-1}(x): returns l (P, n) and log q(x) = -log J(T^{-1}(x))."""
    Ls = unpack(np.asarray(params, dtype=np.float64), n, layers, nc)
    v, xc = _logit(np.array(x, dtype=np.float64, ndmin=2))
    c = _cond(cond, v.shape[0], emulate)
    logq = -np.sum(np.log(xc) + np.log1p(-xc), axis=1)
    for k in reversed(range(layers)):
        A, B = dims_a(n, k), dims_b(n, k)
        a = v[:, A].astype(np.float32).astype(np.float64) if emulate else v[:, A]
        s, t, _ = _net(Ls[k], a, emulate, s_scale(layers), c)
        v[:, B] = (v[:, B] - t) * np.exp(-s)
        logq -= s.sum(axis=1)
    logq += _log_sig_prime(v).sum(axis=1)
    return _sigmoid(v), logq


def loss_and_grad(params, n, layers, x, w, emulate=False, nc=0, cond=None):
    """Weighted maximum-likelihood loss mean_p w~_p (-log q(x_p)), w~ = w / mean(w),
    and its gradient with respect to params (analytic backprop, FP64); with
    nc = 1, q = q_c at each point's conditioning value cond[p]."""
    params = np.asarray(params, dtype=np.float64)
    Ls = unpack(params, n, layers, nc)
    grad = np.zeros_like(params)
    G = unpack(grad, n, layers, nc)
    x = np.array(x, dtype=np.float64, ndmin=2)
    P = x.shape[0]
    cv = _cond(cond, P, emulate)
    wt = np.asarray(w, dtype=np.float64) / np.mean(w)
    v, xc = _logit(x)
    nlq = np.sum(np.log(xc) + np.log1p(-xc), axis=1)  # -log q, accumulated
    for k in reversed(range(layers)):
        A, B = dims_a(n, k), dims_b(n, k)
        a = v[:, A].astype(np.float32).astype(np.float64) if emulate else v[:, A]
        s, t, _ = _net(Ls[k], a, emulate, s_scale(layers), cv)
        v[:, B] = (v[:, B] - t) * np.exp(-s)
        nlq += s.sum(axis=1)
    nlq -= _log_sig_prime(v).sum(axis=1)
    loss = float(np.mean(wt * nlq))
    c = wt / P  # d loss / d (-log q_p)
    # d(-log q)/dv0 = -(1 - 2 sigmoid(v0))
    g = -(1.0 - 2.0 * _sigmoid(v)) * c[:, None]
    for k in range(layers):
        A, B = dims_a(n, k), dims_b(n, k)
        a = v[:, A].astype(np.float32).astype(np.float64) if emulate else v[:, A]
        s, t, I = _net(Ls[k], a, emulate, s_scale(layers), cv)
        vB = v[:, B]
        gB = g[:, B]
        es = np.exp(-s)
        d_s = -gB * vB + c[:, None]
        d_t = -gB * es
        S = s_scale(layers)
        d_os = d_s * (S - s * s / S)
        do = np.concatenate([d_os, d_t], axis=1)          # (P, 2nB)
        W3 = bf16(Ls[k]["W3"]) if emulate else Ls[k]["W3"]
        W2 = bf16(Ls[k]["W2"]) if emulate else Ls[k]["W2"]
        G[k]["W3"][...] = do.T @ I["h2"]
        G[k]["b3"][...] = do.sum(axis=0)
        dz2 = (do @ W3) * (I["z2"] > 0)
        G[k]["W2"][...] = dz2.T @ I["h1"]
        G[k]["b2"][...] = dz2.sum(axis=0)
        dz1 = (dz2 @ W2) * (I["z1"] > 0)
        G[k]["W1"][...] = dz1.T @ I["a"]
        G[k]["b1"][...] = dz1.sum(axis=0)
        if nc:
            G[k]["Wc"][...] = dz1.T @ cv
        da = dz1 @ Ls[k]["W1"]
        g_new = g.copy()
        g_new[:, B] = gB * es
        g_new[:, A] += da
        g = g_new
        # reconstruct v^{k+1} from v^k (the layer's forward direction)
        v = v.copy()
        v[:, B] = vB * np.exp(s) + t
    return loss, grad


def init_params(n, layers, seed, nc=0):
    """Deterministic initialisation: He-uniform W1, W2 (and Wc); zero biases;
    W3 = b3 = 0, so the untrained flow is the identity map."""
    rng = np.random.default_rng(seed)
    p = np.zeros(n_params(n, layers, nc))
    for k, L in enumerate(unpack(p, n, layers, nc)):
        na = max(n_a(n, k), 1)
        L["W1"][...] = rng.uniform(-1, 1, L["W1"].shape) * np.sqrt(6.0 / na)
        L["W2"][...] = rng.uniform(-1, 1, L["W2"].shape) * np.sqrt(6.0 / H)
        if nc:
            L["Wc"][...] = rng.uniform(-1, 1, L["Wc"].shape) * np.sqrt(6.0 / na)
    return p
