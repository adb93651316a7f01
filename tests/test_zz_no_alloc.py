"""GPU tier: gsc_summate_dev_noalloc (the entry point that allocates nothing)
against gsc_summate_dev (the same kernels on a caller workspace).

Both go through the C-ABI.  The two entry points run the same kernels on the
same inputs, so their outputs must agree bit for bit; the FP64 oracle checks
a subset at the project's 1e-5 sigma bar.  Two calls in a row, the second on
a smaller input, exercise the shared device array keeping the first call's
data.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle
from gstools_core import _native, workloads

pytestmark = pytest.mark.gpu


def _upload(a):
    return _native.DeviceBuffer.from_array(np.ascontiguousarray(a))


def _run_both(lib, w, variant):
    md = w.modes
    d, m, n = md.dim, md.m, w.n_points
    bufs = [_upload(md.cov_samples), _upload(md.z1), _upload(md.z2), _upload(w.pos)]
    out_a = _native.DeviceBuffer(8 * n)
    out_b = _native.DeviceBuffer(8 * n)
    opts = _native.make_opts(direct_variant=_native.DIRECT_VARIANTS[variant])
    nb = lib.gsc_workspace_bytes(d, m, n, ctypes.byref(opts))
    work = _native.DeviceBuffer(nb)
    args = (d, m, n, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr, n)
    _native.check(lib.gsc_summate_dev(*args, out_a.ptr, work.ptr, nb, ctypes.byref(opts), None))
    _native.check(lib.gsc_summate_dev_noalloc(*args, out_b.ptr, ctypes.byref(opts), None))
    _native.synchronize()
    a = out_a.to_array(np.float64, n)
    b = out_b.to_array(np.float64, n)
    for x in bufs + [out_a, out_b, work]:
        x.free()
    return a, b


@pytest.mark.parametrize("variant", ["local", "valu"])
def test_noalloc_matches_workspace_entry_point(variant):
    lib = _native.lib()
    # the larger call first, then a smaller one on the same device array
    for n, m, seed in ((2_000_000, 1000, 0), (70_001, 300, 1)):
        w = workloads.c2(n=n, m=m)
        a, b = _run_both(lib, w, variant)
        np.testing.assert_array_equal(b, a)
        idx = workloads.subset_indices(n, 20_000, seed=60 + seed)
        ref = oracle.summate(w.modes.cov_samples, w.modes.z1, w.modes.z2,
                             np.ascontiguousarray(w.pos[:, idx]))
        err = float(np.abs(b[idx] - ref).max()) * w.modes.scale
        print(f"noalloc {variant} n={n} m={m}: max|err|/sigma = {err:.3e}")
        assert err <= 1e-5


def test_noalloc_rejects_sizes_beyond_its_array():
    lib = _native.lib()
    opts = _native.make_opts()
    n = 2_000_000_000 // 64  # its local-variant workspace is far above the array
    assert lib.gsc_workspace_bytes(3, 1000, n, ctypes.byref(opts)) > (512 << 20)
    one = _native.DeviceBuffer(64)
    rc = lib.gsc_summate_dev_noalloc(3, 1000, n, one.ptr, one.ptr, one.ptr, one.ptr, n, one.ptr,
                                     ctypes.byref(opts), None)
    assert rc == _native.GSC_ERR_WORKSPACE
    assert "no-allocation" in _native.last_error()
    one.free()
