"""The bench's command-line contract, on the host: more than one GPU is refused with a JSON
message, and the counted-work helpers follow their stated per-unit figures."""
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "single-GPU" in msg["error"]


def test_workloads_listed():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0
    for w in ("population", "tissue", "ord2011", "serving", "tt06"):
        assert w in r.stdout


def test_counted_flops_of_the_bdf_statistics():
    import bench
    st = np.zeros((2, 8), dtype=np.int64)
    st[0, :5] = (10, 30, 1, 2, 12)         # steps, rhs evals, jacobians, setups, newton iterations
    st[0, 7] = 5                           # output points (one rhs each)
    st[1, :5] = (1, 1, 0, 0, 1)
    want = ((30 + 5) * bench.F_RHS + 2 * bench.F_SETUP + 12 * bench.F_NEWTON + 10 * bench.F_STEP
            + 1 * bench.F_RHS + 1 * bench.F_NEWTON + 1 * bench.F_STEP)
    assert bench.flops_of(st) == want
    assert bench.F_TT_STEP == 534
