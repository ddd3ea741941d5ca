"""bench.py's contract pieces that need no GPU: the HBM traffic it reports is
read from the committed rocprofv3 --pmc summaries (FETCH_SIZE x 2.0 + WRITE_SIZE
x 1.0, KiB per dispatch), for the kernels and launch size they cover, and a
multi-GPU request exits with a message."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


@pytest.mark.parametrize("kernel", ["integrand_ctree_kernel<8,double>", "integrand_ctree_kernel<8,float>",
                                    "integrand_wsplit_kernel<8,double>", "integrand_wsplit_kernel<8,float>",
                                    "integrand_split_kernel<8,double>", "integrand_kernel<8,double>"])
def test_pmc_traffic_from_committed_summaries(kernel):
    t = bench.pmc_traffic(kernel, 1 << 21)
    assert t is not None, kernel
    assert t["bytes"] == t["fetch_bytes"] + t["write_bytes"]
    assert t["fetch_factor"] == 2.0 and t["write_factor"] == 1.0
    # the line the figure came from, recomputed from the summary text
    text = open(os.path.join(ROOT, t["source"])).read()
    line = next(l for l in text.splitlines() if l.split()[:2] == ["pmc-traffic", kernel])
    kv = dict(x.split("=", 1) for x in line.split()[2:])
    assert t["fetch_bytes"] == float(kv["fetch_kib"]) * 1024 * 2.0
    assert t["write_bytes"] == float(kv["write_kib"]) * 1024 * 1.0
    # the kernel reads at least its compacted records (times and Jacobian: 66 B per point at the
    # bench's domain fraction)
    assert t["bytes"] > 66 * (1 << 21)


def test_pmc_traffic_uncovered_is_none():
    assert bench.pmc_traffic("integrand_split_kernel<8,double>", 1 << 20) is None
    assert bench.pmc_traffic("integrand_tree_kernel<8,double>", 1 << 21) is None


def test_more_than_one_gpu_exits_with_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
