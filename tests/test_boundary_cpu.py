"""CPU tier: the C-ABI library and the Python drop-in's host logic (no GPU needed).

No compute call reaches a device here: argument and shape errors are raised
before the library touches HIP, and a valid call without a GPU must fail
loudly (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

import gstools_core
from gstools_core import _native, device, workloads

HEADER = os.path.join(ROOT, "include", "gstools_core_amd.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gsc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_header_symbol():
    lib = _native.lib()
    names = header_functions()
    assert "gsc_summate" in names and "gsc_summate_incompr" in names
    for n in names:
        assert hasattr(lib, n), f"{n} declared in the header but not exported"
    # the ctypes signature table covers exactly the header
    assert set(_native.SIGNATURES) == set(names)


def test_option_struct_mirrors_agree():
    # gsc_opts in the header, its ctypes mirror and the Rust crate-side mirror
    # (not compiled here) list the same fields, in order, with the same widths
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = re.search(r"typedef struct gsc_opts \{(.*?)\} gsc_opts;", text, flags=re.S).group(1)
    c_fields = re.findall(r"\b(int32_t|int64_t|double)\s+(\w+)\s*;", body)
    width = {"int32_t": 4, "int64_t": 8, "double": 8, "i32": 4, "i64": 8, "f64": 8}
    assert [n for _, n in c_fields] == [n for n, _ in _native.GscOpts._fields_]
    assert [width[t] for t, _ in c_fields] == [ctypes.sizeof(t) for _, t in
                                               _native.GscOpts._fields_]
    rust = open(os.path.join(ROOT, "gstools-core_amd", "rust", "src", "hip.rs")).read()
    rbody = re.search(r"pub struct GscOpts \{(.*?)\}", rust, flags=re.S).group(1)
    r_fields = re.findall(r"pub (\w+):\s*(\w+)", rbody)
    assert [n for n, _ in r_fields] == [n for _, n in c_fields]
    assert [width[t] for _, t in r_fields] == [width[t] for t, _ in c_fields]
    # every function the Rust extern block binds is declared in the header
    r_fns = set(re.findall(r"\bfn (gsc_\w+)\s*\(", rust))
    assert r_fns and r_fns <= set(header_functions())


def test_version_and_defaults():
    lib = _native.lib()
    assert lib.gsc_version().decode().startswith("1.1.0")
    o = _native.GscOpts()
    lib.gsc_opts_init(ctypes.byref(o))
    assert (o.out_dtype, o.path, o.precision, o.device, o.scale) == (0, 0, 0, -1, 1.0)
    assert (o.accum, o.direct_variant, o.blocks_per_cu, o.table_bytes) == (1, 3, 0, 0)
    assert gstools_core.__version__ == "1.1.0"


def test_structured_entry_point_signature():
    import inspect

    assert "summate_structured" in gstools_core.__all__
    ps = list(inspect.signature(gstools_core.summate_structured).parameters)
    assert ps[:5] == ["cov_samples", "z1", "z2", "axes", "num_threads"]
    with pytest.raises(ValueError):  # argument checks run before any device work
        gstools_core.summate_structured(np.zeros((3, 4)), np.zeros(4), np.zeros(4),
                                        [np.arange(3.0), np.arange(2.0)])


def test_workspace_bytes():
    valu = device.workspace_bytes(3, 1000, 10**6, variant="valu")
    assert valu >= 48 * 1000 and valu % 256 == 0
    assert device.workspace_bytes(3, 1000, 10, variant="valu") == valu  # no N term
    # the local variant adds the cell sort: at least a cell index and a slot per point
    loc = device.workspace_bytes(3, 1000, 10**6)
    assert loc >= valu + 8 * 10**6
    assert device.workspace_bytes(2, 1000, 10**6) >= valu + 8 * 10**6
    # the default variant ("auto") takes the local kernel from 256 modes up and
    # the FP64-phase VALU kernel (mode table only) below
    assert device.workspace_bytes(3, 255, 10**6) == device.workspace_bytes(3, 255, 10**6, "valu")
    assert device.workspace_bytes(3, 256, 10**6) == device.workspace_bytes(3, 256, 10**6, "local")
    # dimensions the local kernel does not serve need only the mode table
    assert device.workspace_bytes(1, 1000, 10**6) == device.workspace_bytes(1, 1000, 10**6, "valu")
    assert device.workspace_bytes(5, 10, 100) >= 7 * 8 * 10
    assert device.workspace_bytes(0, 10, 10) == 0


def test_krige_workspace_bytes():
    # the documented sizes: per-row-block partial sums of field and error
    # (blocks of 128 conditions), or w = mat cond for the field alone
    lib = _native.lib()
    for n, N in ((1, 1), (128, 10), (129, 10), (1000, 10**6)):
        assert lib.gsc_krige_workspace_bytes(n, N, 1) == 2 * 8 * (-(-n // 128)) * N
        assert lib.gsc_krige_workspace_bytes(n, N, 0) == 8 * n


def _ok_args(d=3, m=4, n=5):
    rng = np.random.default_rng(0)
    return rng.normal(size=(d, m)), rng.normal(size=m), rng.normal(size=m), rng.normal(size=(d, n))


@pytest.mark.parametrize("fn", [gstools_core.summate, gstools_core.summate_incompr])
def test_type_errors_like_pyo3(fn):
    cov, z1, z2, pos = _ok_args()
    with pytest.raises(TypeError):
        fn(cov.tolist(), z1, z2, pos)
    with pytest.raises(TypeError):
        fn(cov.astype(np.float32), z1, z2, pos)
    with pytest.raises(TypeError):
        fn(cov, z1[None, :], z2, pos)
    with pytest.raises(TypeError):
        fn(cov, z1, z2, pos.astype(">f8"))
    with pytest.raises(TypeError):
        fn(cov, z1, z2, pos, 2.0)
    with pytest.raises(OverflowError):
        fn(cov, z1, z2, pos, -1)


@pytest.mark.parametrize("fn", [gstools_core.summate, gstools_core.summate_incompr])
def test_shape_mismatch_raises_value_error(fn):
    cov, z1, z2, pos = _ok_args()
    with pytest.raises(ValueError):
        fn(cov[:2], z1, z2, pos)
    with pytest.raises(ValueError):
        fn(cov, z1[:3], z2, pos)
    with pytest.raises(ValueError):
        fn(cov, z1, z2[:3], pos)


def test_incompr_dimension_check():
    cov, z1, z2, pos = _ok_args(d=4)
    with pytest.raises(ValueError, match="two- and three-dimensional"):
        gstools_core.summate_incompr(cov, z1, z2, pos)
    cov, z1, z2, pos = _ok_args(d=1)
    with pytest.raises(ValueError):
        gstools_core.summate_incompr(cov, z1, z2, pos)


def test_fourier_shape_checks():
    cov, z1, z2, pos = _ok_args()
    with pytest.raises(ValueError):
        gstools_core.summate_fourier(np.ones(3), cov, z1, z2, pos)
    with pytest.raises(TypeError):
        gstools_core.summate_fourier([1.0] * 4, cov, z1, z2, pos)


def test_empty_point_set_needs_no_device():
    cov, z1, z2, _ = _ok_args()
    r = gstools_core.summate(cov, z1, z2, np.zeros((3, 0)))
    assert r.shape == (0,) and r.dtype == np.float64
    r = gstools_core.summate_incompr(cov, z1, z2, np.zeros((3, 0)))
    assert r.shape == (3, 0)


def test_capi_validation_before_device():
    lib = _native.lib()
    o = _native.make_opts()
    buf = np.zeros(16)
    p = ctypes.c_void_p(buf.ctypes.data)
    # incompr with d = 4: GSC_ERR_DIM, no device touched
    rc = lib.gsc_summate_incompr(4, 2, 2, p, 2, 1, p, 1, p, 1, p, 2, 1, p, 1, 4, ctypes.byref(o))
    assert rc == _native.GSC_ERR_DIM
    assert "two- and three-dimensional" in _native.last_error()
    # negative sizes, null pointers, bad dtype
    assert lib.gsc_summate(3, -1, 2, p, 1, 1, p, 1, p, 1, p, 1, 1, p, 1, None) == \
        _native.GSC_ERR_INVALID
    assert lib.gsc_summate(3, 2, 2, None, 1, 1, p, 1, p, 1, p, 1, 1, p, 1, None) == \
        _native.GSC_ERR_INVALID
    bad = _native.make_opts()
    bad.out_dtype = 7
    assert lib.gsc_summate(3, 2, 2, p, 2, 1, p, 1, p, 1, p, 2, 1, p, 1, ctypes.byref(bad)) == \
        _native.GSC_ERR_INVALID
    assert lib.gsc_summate(0, 2, 2, p, 2, 1, p, 1, p, 1, p, 2, 1, p, 1, None) == \
        _native.GSC_ERR_INVALID
    assert lib.gsc_summate(17, 2, 2, p, 2, 1, p, 1, p, 1, p, 2, 1, p, 1, None) == \
        _native.GSC_ERR_DIM
    # device form: workspace too small, pitch below N
    assert lib.gsc_summate_dev(3, 100, 10, p, p, p, p, 10, p, p, 16, None, None) == \
        _native.GSC_ERR_WORKSPACE
    assert lib.gsc_summate_dev(3, 1, 10, p, p, p, p, 5, p, p, 4096, None, None) == \
        _native.GSC_ERR_INVALID


def test_no_silent_cpu_fallback_without_gpu():
    if _native.lib().gsc_device_count() > 0:
        pytest.skip("a GPU is visible; the GPU tier covers this path")
    cov, z1, z2, pos = _ok_args()
    with pytest.raises(RuntimeError):
        gstools_core.summate(cov, z1, z2, pos)


def test_workload_generators_are_seeded_and_shaped():
    a = workloads.c1(n=1000)
    b = workloads.c1(n=1000)
    np.testing.assert_array_equal(a.modes.cov_samples, b.modes.cov_samples)
    np.testing.assert_array_equal(a.pos, b.pos)
    assert a.modes.cov_samples.shape == (2, 1000) and a.pos.shape == (2, 1000)
    assert 0.0 <= a.pos.min() and a.pos.max() <= 100.0
    # k ~ N(0, 1/l^2): sample std close to 1/l = 0.1
    assert abs(a.modes.cov_samples.std() - 0.1) < 0.01
    c2 = workloads.c2(n=10, m=1000)
    assert c2.modes.cov_samples.shape == (3, 1000)
    assert abs(c2.modes.cov_samples.std() - 0.05) < 0.005
    c5 = workloads.c5(side=8, m=4000)
    assert c5.n_points == 512
    assert np.abs(c5.modes.cov_samples).max() > 10.0  # heavy tail
    ens = workloads.c4(count=6)
    assert len(ens) == 6
    for w in ens:
        assert all(64 <= len(ax) <= 1024 for ax in w.axes)
        assert 5.0 <= w.modes.len_scale <= 50.0
    # different realisations draw different modes
    assert not np.array_equal(ens[0].modes.z1, ens[1].modes.z1)


def test_grid_point_expansion():
    axes = [np.arange(3.0), np.arange(4.0) * 2, np.arange(5.0) - 1]
    idx = np.array([0, 7, 59])
    pts = workloads.grid_points(axes, idx)
    full = np.stack(np.meshgrid(*axes, indexing="ij")).reshape(3, -1)
    np.testing.assert_array_equal(pts, full[:, idx])


def test_randmeth_argument_checks_need_no_device():
    from gstools_core.randmeth import IncomprRandMeth, RandMeth
    for kw in ({"model": "matern"}, {"dim": 4}, {"var": -1.0}, {"len_scale": 0.0},
               {"mode_no": 0}, {"nugget": -0.1}):
        args = {"model": "gaussian", "dim": 2}
        args.update(kw)
        with pytest.raises(ValueError):
            RandMeth(args.pop("model"), args.pop("dim"), **args)
    with pytest.raises(ValueError):
        IncomprRandMeth("gaussian", 1)
