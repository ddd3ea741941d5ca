"""kd_set_timing: the timing level changes what kd_last_timing reports and
nothing else (the accumulated sum and counts are the same at every level)."""
import numpy as np
import pytest

from keldy_amd import capi
from oracle import oracle as O
from tests import gpu_util as U

pytestmark = pytest.mark.gpu


def test_timing_levels_same_results(flat_g0):
    t, order = 10.0, 8
    _, (descs, maps, tabs) = U.chain([("plasma", t), ("map", "inverse", t, 1.0, 1)])
    Csr, shfr = O.Sobol(order, 0, True, True, 4).tables()
    res = {}
    for level in (capi.KD_TIMING_KERNELS, capi.KD_TIMING_TOTAL, capi.KD_TIMING_NONE):
        ctx = U.charge_ctx(flat_g0, order, t)
        ctx.set_warper_chain(descs, maps, tabs)
        ctx.set_sobol(Csr, shfr, 0)
        ctx.set_timing(level)
        acc = ctx.run(1 << 13)
        total, lu, prep, launches = ctx.last_timing()
        res[level] = (acc.re, acc.im, acc.n_in_domain, acc.n_fallback)
        print(f"level {level}: total {total:.4f} ms, lu {lu:.4f}, prep {prep:.4f}, launches {launches}")
        assert launches == 3
        if level == capi.KD_TIMING_KERNELS:
            assert total > 0 and lu > 0 and prep > 0 and total >= lu
        elif level == capi.KD_TIMING_TOTAL:
            assert total > 0 and lu == 0 and prep == 0
        else:
            assert total == 0 and lu == 0 and prep == 0
    assert res[capi.KD_TIMING_KERNELS] == res[capi.KD_TIMING_TOTAL] == res[capi.KD_TIMING_NONE]
    with pytest.raises(capi.KdError):
        ctx.set_timing(3)
