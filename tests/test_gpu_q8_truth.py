"""The order-8 charge against the committed high-precision value
(tests/golden/q8_truth.json: 16 replicas of 2^28 points of the same chain,
written by tests/golden/make_q8_truth.py): a fresh estimate from 8 replicas of
2^24 points on other scrambles agrees within 5 sigma (its own replica error
and the value's combined), and lies within the 1e-3 relative error the
benchmark targets."""
import json
import os

import numpy as np
import pytest

from tests.golden.make_q8_truth import ORDER, plain_chain_computer, replica_charges

pytestmark = pytest.mark.gpu


def test_q8_estimate_vs_committed_value():
    with open(os.path.join(os.path.dirname(__file__), "golden", "q8_truth.json")) as fh:
        truth = json.load(fh)
    assert truth["order"] == ORDER and truth["replicas"] * truth["points_per_replica"] >= 16 * (1 << 28)
    q = replica_charges(plain_chain_computer(), 8, 1 << 24, 2_000_000)
    err = np.hypot(q.std(ddof=1) / np.sqrt(len(q)), truth["charge_std_error"])
    print(f"\nQ_8(10): estimate {q.mean():.10e} +- {q.std(ddof=1) / np.sqrt(len(q)):.1e}, "
          f"committed {truth['charge_mean']:.10e} +- {truth['charge_std_error']:.1e}", flush=True)
    assert abs(q.mean() - truth["charge_mean"]) <= 5 * err
    assert abs(q.mean() / truth["charge_mean"] - 1) < 1e-3
