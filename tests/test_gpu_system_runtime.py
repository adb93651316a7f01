"""GPU tier: the library on the system HIP runtime (/opt/rocm), without torch.

The test session imports torch, so its library calls run on torch's bundled
runtime; a child process without torch loads the system runtime and runs the
three summators (host and device entry points) against the FP64 oracle.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import json, sys
import numpy as np
sys.path[:0] = [sys.argv[1], sys.argv[1] + "/gstools-core_amd"]
import gstools_core, oracle
from gstools_core import _native, grid, workloads
assert "torch" not in sys.modules
w = workloads.c2(n=300_000)
md = w.modes
errs = {}
g = gstools_core.summate(md.cov_samples, md.z1, md.z2, w.pos)
errs["summate"] = float(np.abs(g[:5000] - oracle.summate(md.cov_samples, md.z1, md.z2,
                        np.ascontiguousarray(w.pos[:, :5000]))).max()) * md.scale
gi = gstools_core.summate_incompr(md.cov_samples, md.z1, md.z2, w.pos)
ri = oracle.summate_incompr(md.cov_samples, md.z1, md.z2, np.ascontiguousarray(w.pos[:, :5000]))
errs["summate_incompr"] = float(np.abs(gi[:, :5000] - ri).max()) * md.scale
axes = [np.arange(60.0), np.arange(50.0), np.arange(40.0)]
gg = grid.summate_grid(md.cov_samples, md.z1, md.z2, axes)
rg = oracle.summate(md.cov_samples, md.z1, md.z2, oracle.grid_points(axes)).reshape(gg.shape)
errs["grid"] = float(np.abs(gg - rg).max()) * md.scale
print(json.dumps({"runtime": _native.hip_runtime_path(), "errs": errs,
                  "torch": "torch" in sys.modules}))
"""


def test_library_on_the_system_runtime():
    env = dict(os.environ, GSC_HIP_RUNTIME="system")
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], capture_output=True, text=True,
                       env=env, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    print(res)
    assert not res["torch"]
    assert res["runtime"].startswith("/opt/rocm"), res["runtime"]
    for name, e in res["errs"].items():
        assert e <= 1e-5, (name, e)
