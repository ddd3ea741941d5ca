"""keldy.warpers counterpart (python/keldy/warpers_desc.py)."""
import numpy as np

from ._host import ext as _ext
from .common import Sobol

_w = _ext.warpers
WarperIdentityT = _w.WarperIdentityT
WarperPlasmaUvT = _w.WarperPlasmaUvT
WarperProduct1dSimpleT = _w.WarperProduct1dSimpleT
WarperProduct1dSimpleInterpNearestT = _w.WarperProduct1dSimpleInterpNearestT
WarperProduct1dSimpleInterpHybridT = _w.WarperProduct1dSimpleInterpHybridT
WarperProduct1dT = _w.WarperProduct1dT
WarperProduct1dInterpNearestT = _w.WarperProduct1dInterpNearestT
WarperProduct1dInterpHybridT = _w.WarperProduct1dInterpHybridT
WarperTrainT = _w.WarperTrainT
WarperFlowT = _w.WarperFlowT
WarperProjectionT = _w.WarperProjectionT
vi_from_ui = _w.vi_from_ui
ui_from_vi = _w.ui_from_vi
make_product_1d_simple_exponential = _w.make_product_1d_simple_exponential
make_product_1d_simple_inverse = _w.make_product_1d_simple_inverse
make_product_1d_simple_inverse_square = _w.make_product_1d_simple_inverse_square
make_product_1d_simple_inverse_power = _w.make_product_1d_simple_inverse_power
make_product_1d_simple_uniform = _w.make_product_1d_simple_uniform
make_product_1d_inverse_cube_alternate = _w.make_product_1d_inverse_cube_alternate
make_product_1d_inverse_cube_alternate_interp_nearest = _w.make_product_1d_inverse_cube_alternate_interp_nearest
make_product_1d_inverse_cube_alternate_interp_hybrid = _w.make_product_1d_inverse_cube_alternate_interp_hybrid
make_warper_from_proj_samples = _w.make_warper_from_proj_samples
kernel_smoothing = _w.kernel_smoothing


def make_warper_from_proj(warped_integrand, order, num_bins, nr_samples, sigma, keep_u_hypercube=False):
    """Pilot-fitted product warper (make_warper_from_proj.hpp:118-153).

    warped_integrand is either a callable l -> complex (evaluated once per
    pilot point, as in the reference) or an integrator object
    (ComputeChargeQDirect, ...): then the pilot evaluations of its warped
    integrand with its current warper chain AND the per-axis histograms run on
    the device (only the smoothing and the fit run on the host) -- without the
    positive-time restriction by default, as the reference's own recipe does
    (projection_test.py:86), so the g0 window must cover order * t.
    """
    device = getattr(warped_integrand, "make_warper_from_proj_device", None)
    if device is None:
        return _w.make_warper_from_proj(warped_integrand, order, num_bins, nr_samples, sigma)
    if order != warped_integrand.get_integrand_dimension():
        raise RuntimeError("order differs from the integrator's dimension")
    return device(num_bins, nr_samples, sigma, keep_u_hypercube)


__all__ = [n for n in dir() if n[0].isupper() or n.startswith(("make_", "vi_", "ui_", "kernel_"))]


def train_flow_warper(computer, n_layers=6, nr_samples=1 << 20, steps=2000, batch=1 << 14, lr=1e-3,
                      lr_final=1e-5, rng_seed_shift=1000003, init_seed=0, weight_power=1.0):
    """Learned coupling-flow warper (no counterpart in keldy).

    Collects nr_samples scrambled-Sobol points (seed shift rng_seed_shift) of
    `computer`'s warped integrand with its current chain, then trains the flow
    on the device (hand-written forward/backward/Adam kernels) to make its
    density proportional to |warped integrand|^weight_power.  Returns
    (WarperFlowT, per-step loss); append the warper last:
    computer.warper.emplace_back(w).
    """
    return computer.train_flow_warper(n_layers, nr_samples, steps, batch, lr, lr_final, rng_seed_shift, init_seed,
                                      weight_power)


def train_flow_warper_conditioned(computers, conds, n_layers=6, samples_per_computer=1 << 16, steps=2000,
                                  batch=1 << 12, lr=1e-3, lr_final=1e-5, rng_seed_shift=1000003, init_seed=0,
                                  weight_power=1.0):
    """Time-conditioned learned flow (no counterpart in keldy): one coupling
    flow with the conditioning value (e.g. the external time scaled to [0, 1])
    as an extra input of every layer, trained on the samples of all
    `computers` (computer i's tagged with conds[i], weights scaled to mean 1
    per computer).  Returns (WarperFlowT at conds[0], per-step loss); append
    flow.with_cond(conds[i]) last to computer i's warper train."""
    from keldy_amd import impurity_oneband as _io
    return _io.train_flow_warper_conditioned(list(computers), [float(c) for c in conds], n_layers,
                                             samples_per_computer, steps, batch, lr, lr_final, rng_seed_shift,
                                             init_seed, weight_power)
