import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


def _have_gpu():
    """True when the library loads and sees a HIP device.  A library that is
    missing or fails to load is an error on a machine with a GPU driver
    (/dev/kfd), so a broken build cannot pass as a run of skipped tests."""
    try:
        from keldy_amd import capi
        capi.lib()
    except (ImportError, OSError) as e:
        if os.path.exists("/dev/kfd"):
            raise
        return f"library not loadable here ({e})"
    return capi.device_count() > 0 or "no HIP device"


def pytest_collection_modifyitems(config, items):
    if not any("gpu" in it.keywords for it in items):
        return
    have = _have_gpu()
    if have is True:
        return
    skip = pytest.mark.skip(reason=have)
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def flat_g0():
    """Flatband, Gamma=1, T=0, V=0, eps_d=0 analytic g0 on [-10, 10], 10001 points
    (the workload's model, make_flat_band_analytic, model.cpp:428-466)."""
    from oracle import g0_oracle
    t, les, gre = g0_oracle.flatband_analytic(10.0, 10000, 1.0)
    return {"time_max": 10.0, "lesser": les, "greater": gre}
