"""GPU tier: the kriging calculators (krige.rs:24-118) against the reference's
golden values and the FP64 oracle.

Tolerance: the reference's own tests allow 6 ulps (krige.rs:208, 220, 243);
the GPU sums in another order (tiles of 128 conditions, FMA), so random cases
are held to 1e-12 of the largest term of the sum, i.e. FP64 rounding only.
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

import oracle
import gstools_core
from conftest import GOLDEN
from gstools_core import _native

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def krige_rs():
    with open(os.path.join(GOLDEN, "krige_rs_fixture.json")) as f:
        d = json.load(f)
    return {k: (np.array(v, dtype=np.float64) if k != "source" else v) for k, v in d.items()}


def ulps(a, b):
    return np.abs(a - b) / np.spacing(np.abs(b))


def test_golden_values(krige_rs):
    m, v, c = krige_rs["krig_mat"], krige_rs["krig_vecs"], krige_rs["cond"]
    f, e = gstools_core.calc_field_krige_and_variance(m, v, c)
    assert np.max(ulps(f, krige_rs["field"])) <= 6
    assert np.max(ulps(e, krige_rs["error"])) <= 6
    f2 = gstools_core.calc_field_krige(m, v, c)
    assert np.max(ulps(f2, krige_rs["field"])) <= 6


def _scale(m, v, c):
    # largest term of the sums: |c_i| |fac_i| and |v_ij| |fac_i|
    fac = np.abs(m).T @ np.abs(v)
    return max(float(np.max(np.abs(c)[:, None] * fac, initial=0.0)),
               float(np.max(np.abs(v) * fac, initial=0.0)), 1e-300)


@pytest.mark.parametrize("n,N", [(1, 1), (3, 6), (63, 65), (64, 64), (65, 1000), (130, 63),
                                 (127, 129), (128, 128), (129, 300), (257, 2000),
                                 (200, 5000), (1, 777)])
def test_random_against_oracle(n, N):
    rng = np.random.default_rng(1000 * n + N)
    a = rng.normal(size=(n, n))
    m = a @ a.T / n + np.eye(n)
    v = rng.normal(size=(n, N))
    c = rng.normal(size=n)
    f, e = gstools_core.calc_field_krige_and_variance(m, v, c)
    rf, re = oracle.calc_field_krige_and_variance(m, v, c)
    tol = 1e-12 * n * _scale(m, v, c)
    assert np.abs(f - rf).max() <= tol
    assert np.abs(e - re).max() <= tol
    f2 = gstools_core.calc_field_krige(m, v, c)
    assert np.abs(f2 - rf).max() <= tol


def test_strided_views_and_reproducibility():
    rng = np.random.default_rng(7)
    m = rng.normal(size=(90, 90))
    v = rng.normal(size=(90, 3000))
    c = rng.normal(size=90)
    f, e = gstools_core.calc_field_krige_and_variance(m[::-1, ::-1], v[::-1, ::2], c[::-1])
    rf, re = oracle.calc_field_krige_and_variance(m[::-1, ::-1], v[::-1, ::2], c[::-1])
    tol = 1e-12 * 90 * _scale(m, v, c)
    assert np.abs(f - rf).max() <= tol and np.abs(e - re).max() <= tol
    f2, e2 = gstools_core.calc_field_krige_and_variance(m[::-1, ::-1], v[::-1, ::2], c[::-1])
    np.testing.assert_array_equal(f, f2)
    np.testing.assert_array_equal(e, e2)


def test_large_against_oracle_subset():
    rng = np.random.default_rng(8)
    n, N = 1000, 100_000
    a = rng.normal(size=(n, n)) / np.sqrt(n)
    m = a @ a.T
    v = rng.normal(size=(n, N))
    c = rng.normal(size=n)
    f, e = gstools_core.calc_field_krige_and_variance(m, v, c)
    sub = np.sort(rng.choice(N, 500, replace=False))
    rf, re = oracle.calc_field_krige_and_variance(m, np.ascontiguousarray(v[:, sub]), c)
    tol = 1e-12 * n * _scale(m, v[:, sub], c)
    assert np.abs(f[sub] - rf).max() <= tol
    assert np.abs(e[sub] - re).max() <= tol


def test_device_entry_point_matches_host():
    rng = np.random.default_rng(9)
    n, N = 150, 4000
    m = rng.normal(size=(n, n))
    v = rng.normal(size=(n, N))
    c = rng.normal(size=n)
    f, e = gstools_core.calc_field_krige_and_variance(m, v, c)
    lib = _native.lib()
    dm = _native.DeviceBuffer.from_array(m)
    dv = _native.DeviceBuffer.from_array(v)
    dc = _native.DeviceBuffer.from_array(c)
    out = _native.DeviceBuffer(16 * N)
    nb = lib.gsc_krige_workspace_bytes(n, N, 1)
    work = _native.DeviceBuffer(nb)
    opts = _native.make_opts()
    import ctypes
    _native.check(lib.gsc_krige_dev(n, N, dm.ptr, dv.ptr, N, dc.ptr, out.ptr, out.ptr + 8 * N,
                                    work.ptr, nb, ctypes.byref(opts), None))
    _native.synchronize()
    res = out.to_array(np.float64, 2 * N)
    np.testing.assert_array_equal(res[:N], f)
    np.testing.assert_array_equal(res[N:], e)
    # a workspace that is too small is refused
    with pytest.raises(ValueError):
        _native.check(lib.gsc_krige_dev(n, N, dm.ptr, dv.ptr, N, dc.ptr, out.ptr, out.ptr + 8 * N,
                                        work.ptr, nb - 8, ctypes.byref(opts), None))
    for b in (dm, dv, dc, out, work):
        b.free()


@pytest.mark.parametrize("with_error", [True, False])
def test_device_entry_point_row_pitch(with_error):
    # krig_vecs rows with a pitch above N (a view of a wider array): the same
    # results as the dense host call, bit for bit
    import ctypes
    rng = np.random.default_rng(10)
    n, N, pitch = 140, 1000, 1037
    m = rng.normal(size=(n, n))
    wide = rng.normal(size=(n, pitch))
    c = rng.normal(size=n)
    f, e = gstools_core.calc_field_krige_and_variance(m, np.ascontiguousarray(wide[:, :N]), c)
    if not with_error:
        f = gstools_core.calc_field_krige(m, np.ascontiguousarray(wide[:, :N]), c)
    lib = _native.lib()
    dm = _native.DeviceBuffer.from_array(m)
    dv = _native.DeviceBuffer.from_array(wide)
    dc = _native.DeviceBuffer.from_array(c)
    out = _native.DeviceBuffer(16 * N)
    nb = lib.gsc_krige_workspace_bytes(n, N, 1 if with_error else 0)
    work = _native.DeviceBuffer(nb)
    opts = _native.make_opts()
    _native.check(lib.gsc_krige_dev(n, N, dm.ptr, dv.ptr, pitch, dc.ptr, out.ptr,
                                    out.ptr + 8 * N if with_error else None, work.ptr, nb,
                                    ctypes.byref(opts), None))
    _native.synchronize()
    res = out.to_array(np.float64, 2 * N)
    np.testing.assert_array_equal(res[:N], f)
    if with_error:
        np.testing.assert_array_equal(res[N:], e)
    for b in (dm, dv, dc, out, work):
        b.free()


def test_empty_and_shape_errors():
    f, e = gstools_core.calc_field_krige_and_variance(np.zeros((0, 0)), np.zeros((0, 4)),
                                                      np.zeros(0))
    np.testing.assert_array_equal(f, np.zeros(4))
    np.testing.assert_array_equal(e, np.zeros(4))
    f, e = gstools_core.calc_field_krige_and_variance(np.eye(2), np.zeros((2, 0)), np.ones(2))
    assert f.shape == (0,) and e.shape == (0,)
    with pytest.raises(ValueError):
        gstools_core.calc_field_krige(np.eye(3), np.zeros((2, 4)), np.ones(3))
    with pytest.raises(ValueError):
        gstools_core.calc_field_krige(np.zeros((3, 2)), np.zeros((3, 4)), np.ones(3))
    with pytest.raises(ValueError):
        gstools_core.calc_field_krige(np.eye(3), np.zeros((3, 4)), np.ones(2))
