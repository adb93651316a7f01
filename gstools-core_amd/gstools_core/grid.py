"""Structured-grid front-end of the summator (no counterpart in the reference).

GSTools evaluates a structured mesh by expanding it to a (d, N) point list and
calling ``summate`` (the reference has only point-list entry points,
src/lib.rs:33-65).  ``summate_grid`` takes the per-axis coordinates instead and
returns the same field on the C-order grid, computed either by the per-point
kernel on coordinates generated from indices, or by the factored contraction
on the MFMA cores (per-axis tables, split-precision GEMM).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from . import _array_arg, _threads_arg, _out_dtype, _ptr, _es  # noqa: F401

PATHS = {"auto": _native.GSC_PATH_AUTO, "direct": _native.GSC_PATH_DIRECT,
         "factored": _native.GSC_PATH_FACTORED}


def _axes_args(axes):
    if isinstance(axes, np.ndarray) and axes.ndim == 1:
        axes = [axes]
    axes = [np.asarray(a, dtype=np.float64) for a in axes]
    for a in axes:
        if a.ndim != 1:
            raise ValueError("every axis must be a 1-D coordinate array")
    return axes


def _opts(path, precision, out_code, scale, device, nt, accum="f32x2", variant="valu",
          tile="auto"):
    if path not in PATHS:
        raise ValueError(f"path must be one of {sorted(PATHS)}")
    if str(tile) not in _native.TILES:
        raise ValueError(f"tile must be one of {sorted(_native.TILES)}")
    if precision not in _native.PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(_native.PRECISIONS)}")
    if accum not in _native.ACCUM:
        raise ValueError(f"accum must be one of {sorted(_native.ACCUM)}")
    return _native.make_opts(out_dtype=out_code, scale=scale, device=device, num_threads=nt,
                             path=PATHS[path], precision=_native.PRECISIONS[precision],
                             accum=_native.ACCUM[accum],
                             direct_variant=_native.DIRECT_VARIANTS[variant],
                             tile=_native.TILES[str(tile)])


def grid_path(cov_samples, axes, path="auto", precision="auto") -> str:
    """The path ('direct' or 'factored') summate_grid takes for these sizes."""
    axes = _axes_args(axes)
    d = len(axes)
    shape = (ctypes.c_int64 * d)(*[len(a) for a in axes])
    o = _opts(path, precision, _native.GSC_F64, 1.0, -1, 0)
    r = _native.lib().gsc_grid_path(d, int(np.shape(cov_samples)[1]), shape, ctypes.byref(o))
    if r < 0:
        raise ValueError(_native.last_error() or "invalid grid arguments")
    return "direct" if r == _native.GSC_PATH_DIRECT else "factored"


FIELDS = {"scalar": _native.GSC_FIELD_SCALAR, "incompr": _native.GSC_FIELD_INCOMPR,
          "fourier": _native.GSC_FIELD_FOURIER}


def _grid_call(field, sf, cov_samples, z1, z2, axes, num_threads, path, precision, out_dtype,
               scale, device, accum, variant, tile="auto"):
    cov = _array_arg("modes" if field == "fourier" else "cov_samples", cov_samples, 2)
    z1 = _array_arg("z1", z1, 1)
    z2 = _array_arg("z2", z2, 1)
    nt = _threads_arg(num_threads)
    axes = _axes_args(axes)
    d = len(axes)
    if cov.shape[0] != d:
        raise ValueError(f"cov_samples has {cov.shape[0]} rows but {d} axes were given")
    if cov.shape[1] != z1.shape[0] or cov.shape[1] != z2.shape[0]:
        raise ValueError("cov_samples, z1 and z2 disagree on the number of modes")
    if d < 1 or d > 3:
        raise ValueError("summate_grid supports 1 to 3 axes")
    if field == "incompr" and d not in (2, 3):
        raise ValueError("Only two- and three-dimensional problems are supported.")
    ssf = 0
    if field == "fourier":
        sf = _array_arg("spectrum_factor", sf, 1)
        if sf.shape[0] != cov.shape[1]:
            raise ValueError(f"modes has {cov.shape[1]} columns but spectrum_factor has "
                             f"{sf.shape[0]}")
        (ssf,) = _es(sf)
    dt, code = _out_dtype(out_dtype)
    shape = tuple(len(a) for a in axes)
    out = np.empty(((d,) if field == "incompr" else ()) + shape, dtype=dt)
    if out.size == 0:
        return out
    o = _opts(path, precision, code, scale, device, nt, accum, variant, tile)
    c_axes = (ctypes.c_void_p * d)(*[a.ctypes.data for a in axes])
    c_as = (ctypes.c_int64 * d)(*[a.strides[0] // 8 for a in axes])
    c_shape = (ctypes.c_int64 * d)(*shape)
    (csd, csm), (s1,), (s2,) = _es(cov), _es(z1), _es(z2)
    rc = _native.lib().gsc_summate_grid_field(
        FIELDS[field], d, cov.shape[1], _ptr(sf) if field == "fourier" else None, ssf, _ptr(cov),
        csd, csm, _ptr(z1), s1, _ptr(z2), s2, c_axes, c_as, c_shape, _ptr(out), ctypes.byref(o))
    _native.check(rc)
    return out


def summate_grid(cov_samples, z1, z2, axes, num_threads=None, *, path="auto", precision="auto",
                 out_dtype=np.float64, scale=1.0, device=-1, accum="f32x2", variant="valu",
                 tile="auto"):
    """Field on the grid axes[0] x ... x axes[d-1] (d in 1..3), C order, shape (len(ax) ...).

    Equal to ``summate(cov_samples, z1, z2, pos)`` with pos the expanded grid
    (``np.meshgrid(*axes, indexing='ij')`` flattened), within the factored
    precision's error (fp16x3 default: below 1e-5 sigma).
    """
    return _grid_call("scalar", None, cov_samples, z1, z2, axes, num_threads, path, precision,
                      out_dtype, scale, device, accum, variant, tile)


def summate_incompr_grid(cov_samples, z1, z2, axes, num_threads=None, *, path="auto",
                         precision="auto", out_dtype=np.float64, scale=1.0, device=-1,
                         accum="f32x2", variant="valu", tile="auto"):
    """Incompressible field (field.rs:97-182) on the grid axes[0] x .. x
    axes[d-1], d in {2, 3}: shape (d, len(axes[0]), ...), C order.  Equal to
    ``summate_incompr`` on the expanded grid; the factored path folds the
    projector into the last axis' table (one contraction for all components)."""
    return _grid_call("incompr", None, cov_samples, z1, z2, axes, num_threads, path, precision,
                      out_dtype, scale, device, accum, variant, tile)


def summate_fourier_grid(spectrum_factor, modes, z1, z2, axes, num_threads=None, *,
                         path="auto", precision="auto", out_dtype=np.float64, scale=1.0,
                         device=-1, accum="f32x2", variant="valu", tile="auto"):
    """Fourier-method field (field.rs:219-249) on the grid axes[0] x .. x
    axes[d-1]: shape (len(axes[0]), ...), C order; equal to ``summate_fourier``
    on the expanded grid."""
    return _grid_call("fourier", spectrum_factor, modes, z1, z2, axes, num_threads, path,
                      precision, out_dtype, scale, device, accum, variant, tile)


def grid_workspace_bytes(dim, m, shape, path="auto", precision="auto", field="scalar",
                         variant="valu", tile="auto") -> int:
    c_shape = (ctypes.c_int64 * dim)(*shape)
    o = _opts(path, precision, _native.GSC_F64, 1.0, -1, 0, variant=variant, tile=tile)
    return int(_native.lib().gsc_grid_field_workspace_bytes(FIELDS[field], dim, m, c_shape,
                                                            ctypes.byref(o)))


def summate_grid_dev(dim, m, cov, z1, z2, axes_ptrs, shape, out, work, work_bytes, *,
                     path="auto", precision="auto", out_f32=False, scale=1.0, stream=0,
                     accum="f32x2", variant="valu", field="scalar", sf=None, tile="auto"):
    """Device form: every pointer is a device address (int); see
    gsc_summate_grid_field_dev (field "scalar", "incompr" or "fourier"; sf:
    the spectrum factors of the Fourier field)."""
    o = _opts(path, precision, _native.GSC_F32 if out_f32 else _native.GSC_F64, scale, -1, 0,
              accum, variant, tile)
    c_axes = (ctypes.c_void_p * dim)(*axes_ptrs)
    c_shape = (ctypes.c_int64 * dim)(*shape)
    rc = _native.lib().gsc_summate_grid_field_dev(FIELDS[field], dim, m, sf, cov, z1, z2, c_axes,
                                                  c_shape, out, work, work_bytes,
                                                  ctypes.byref(o), ctypes.c_void_p(stream or None))
    _native.check(rc)
