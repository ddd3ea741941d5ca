"""Runs of several 2^21-point chunks, where the next chunk's points are
prepared on a second stream into a second record buffer while the current
chunk's Keldysh sum runs, against the same points run one chunk per call
(no overlap): the accumulated sum, the counts and the time bins are
bit-identical, each chunk's partial sums being reduced in the same fixed order
either way.  Four chunks (the last one partial) reuse both buffers; the cases
cover the one-point-per-wave tree (order 8), two points per wave with time
bins (order 6) and the learned-flow stage on the second stream (order 5), as
separate runs and as one batch captured into a hipGraph."""
import numpy as np
import pytest

from keldy_amd import capi
from oracle import oracle as O
from tests import gpu_util as U
from tests.flow_util import flow_params

pytestmark = pytest.mark.gpu
CH = 1 << 21


def _ctx(flat_g0, case, order, t):
    _, (descs, maps, tabs) = U.chain([("plasma", t), ("map", "inverse", t, 1.0, 1)])
    ctx = U.charge_ctx(flat_g0, order, t)
    if case == "flow":
        descs = list(descs) + [capi.WarperDesc(capi.KD_WARP_FLOW, 0, 0.0, 0)]
        ctx.flow_set_params(order, 4, flow_params(order, 4, 17))
    ctx.set_warper_chain(descs, maps, tabs)
    ctx.set_sobol(*O.Sobol(order, 0, True, True, 21).tables(), 0)
    if case == "bins":
        ctx.set_time_bins(0.0, t, 9)
    return ctx


@pytest.mark.parametrize("case,order", [("plain", 8), ("bins", 6), ("flow", 5)])
def test_multi_chunk_run_equals_chunk_by_chunk(flat_g0, case, order):
    t, n = 10.0, 3 * CH + 12345
    a, b = _ctx(flat_g0, case, order, t), _ctx(flat_g0, case, order, t)
    ra = a.run(n)
    for off in range(0, n, CH):
        rb = b.run(min(CH, n - off))
    print(f"{case}, order {order}: {n} points, in domain {ra.n_in_domain}, sum {ra.re:.12e} {ra.im:.12e}")
    assert (ra.re, ra.im, ra.n_points, ra.n_in_domain, ra.n_fallback) == \
        (rb.re, rb.im, rb.n_points, rb.n_in_domain, rb.n_fallback)
    assert ra.n_points == n and ra.n_in_domain > 0
    if case == "bins":
        (va, ca, xa), (vb, cb, xb) = a.time_bins(), b.time_bins()
        assert np.array_equal(va, vb) and np.array_equal(ca, cb) and xa == xb


def test_multi_chunk_batch_equals_separate_runs(flat_g0):
    """kd_run_batch captures the runs of several contexts into one hipGraph: a
    context of more than one chunk brings its second stream into the capture
    (the next chunk's preparation) and joins it back.  Each context's sums,
    counts and time bins equal those of a separate kd_run bit for bit."""
    t = 10.0
    cases = [("plain", 8, 2 * CH + 777), ("bins", 6, 2 * CH + 4096), ("flow", 5, CH + 1)]
    batch = [_ctx(flat_g0, case, order, t) for case, order, _ in cases]
    accs, ms = capi.run_batch(batch, [n for *_, n in cases])
    assert ms > 0
    for (case, order, n), ctx, a in zip(cases, batch, accs):
        ref = _ctx(flat_g0, case, order, t)
        r = ref.run(n)
        print(f"batch {case}, order {order}: {n} points, in domain {a.n_in_domain}, sum {a.re:.12e} {a.im:.12e}")
        assert (a.re, a.im, a.n_points, a.n_in_domain, a.n_fallback) == \
            (r.re, r.im, r.n_points, r.n_in_domain, r.n_fallback)
        assert a.n_points == n and a.n_in_domain > 0 and ctx.sobol_state() == n
        if case == "bins":
            (va, ca, xa), (vb, cb, xb) = ctx.time_bins(), ref.time_bins()
            assert np.array_equal(va, vb) and np.array_equal(ca, cb) and xa == xb
