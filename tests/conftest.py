"""Shared test setup: import paths, the `gpu` marker and common fixtures."""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gstools-core_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")

# The GPU tests share device buffers with torch: importing it before the
# library's first call puts both on torch's HIP runtime (gstools_core binds to
# the runtime already loaded).  tests/test_gpu_system_runtime.py runs the
# library without torch, on the system runtime, in a child process.
try:
    import torch  # noqa: F401
except ImportError:  # the CPU tier does not need it
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs on the MI355X box)")
    config.addinivalue_line("markers", "slow: larger CPU oracle workloads")


@pytest.fixture(scope="session")
def field_rs():
    """The reference's own fixture and expected values (src/field.rs:268-427)."""
    with open(os.path.join(GOLDEN, "field_rs_fixture.json")) as f:
        d = json.load(f)
    return {k: (np.array(v, dtype=np.float64) if k != "source" else v) for k, v in d.items()}


def sigma_tol(m: int, rel: float = 1e-5) -> float:
    """Absolute tolerance on an UNSCALED field for max |err| <= rel * sigma.

    The reference returns unscaled sums; GSTools scales by sqrt(var / M) with
    var = 1, so an error bound of rel * sigma on the scaled field is
    rel * sqrt(M) on the unscaled one.
    """
    return rel * float(np.sqrt(max(m, 1)))


@pytest.fixture(scope="session")
def c5_subset():
    """The heavy-tailed 512^3 grid (M = 4000) and the FP64 oracle on a seeded
    10^6-point random subset of it (about 4e9 point-modes on the host)."""
    import oracle
    from gstools_core import workloads

    w = workloads.c5()
    md = w.modes
    idx = workloads.subset_indices(w.n_points, 1_000_000, seed=5)
    pts = workloads.grid_points(w.axes, idx)
    ref = oracle.summate(md.cov_samples, md.z1, md.z2, pts)
    return w, idx, pts, ref
