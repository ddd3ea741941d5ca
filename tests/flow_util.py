"""Shared helper of the flow tests: a flow with non-trivial parameters.

An untrained flow is the identity (W3 = b3 = 0); these parameters give every
term of the map and of the gradient a non-zero value while keeping the
coupling scales moderate (|s| well below its bound 2)."""
import numpy as np

from oracle import flow_oracle as F


def flow_params(n, layers, seed, scale=0.004):
    p = F.init_params(n, layers, seed)
    rng = np.random.default_rng(seed + 1)
    for L in F.unpack(p, n, layers):
        L["W3"][...] = rng.normal(0, scale, L["W3"].shape)
        L["b3"][...] = rng.normal(0, scale, L["b3"].shape)
        L["b1"][...] = rng.normal(0, 0.1, L["b1"].shape)
    return p.astype(np.float32).astype(np.float64)
