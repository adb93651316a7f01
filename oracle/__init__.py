"""CPU oracle for the randomization-method summators -- TEST INFRASTRUCTURE ONLY.

A ctypes front-end to ``oracle/summator_oracle.c`` (an FP64 restatement of the
reference's ``src/field.rs`` summators, see that file's header for the
line-by-line correspondence).  Only ``tests/``, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` may import this module, and only as the
checker or the timed CPU baseline.  The product library never calls it.

Functions take numpy float64 arrays of any layout (strides are passed through)
and return new float64 arrays laid out like the reference's results:
``summate`` -> (N,), ``summate_incompr`` -> (d, N) Fortran order
(field.rs:166-174 builds an (N, d) array and reverses its axes).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_summator.so")
_lib = None

_i64 = ctypes.c_int64
_int = ctypes.c_int
_dp = ctypes.c_void_p


def build() -> None:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    lib.gsc_oracle_summate.argtypes = [_int, _i64, _i64, _dp, _i64, _i64, _dp, _i64, _dp, _i64,
                                       _dp, _i64, _i64, _dp, _i64, _int]
    lib.gsc_oracle_summate.restype = None
    lib.gsc_oracle_summate_fourier.argtypes = [_int, _i64, _i64, _dp, _i64, _dp, _i64, _i64,
                                               _dp, _i64, _dp, _i64, _dp, _i64, _i64, _dp, _i64,
                                               _int]
    lib.gsc_oracle_summate_fourier.restype = None
    lib.gsc_oracle_summate_incompr.argtypes = [_int, _i64, _i64, _dp, _i64, _i64, _dp, _i64, _dp,
                                               _i64, _dp, _i64, _i64, _dp, _i64, _i64, _int, _int]
    lib.gsc_oracle_summate_incompr.restype = _int
    lib.gsc_oracle_summate_incompr_modechunk.argtypes = [_int, _i64, _i64, _dp, _i64, _i64, _dp,
                                                         _i64, _dp, _i64, _dp, _i64, _i64, _dp,
                                                         _i64, _i64, _i64, _int]
    lib.gsc_oracle_summate_incompr_modechunk.restype = _int
    lib.gsc_oracle_krige.argtypes = [_i64, _i64, _dp, _i64, _i64, _dp, _i64, _i64, _dp, _i64,
                                     _dp, _i64, _dp, _i64]
    lib.gsc_oracle_krige.restype = None
    lib.gsc_oracle_max_threads.argtypes = []
    lib.gsc_oracle_max_threads.restype = _int
    _lib = lib
    return lib


def _view(a, ndim):
    a = np.asarray(a)
    if a.dtype != np.float64:
        raise TypeError("oracle expects float64 arrays")
    if a.ndim != ndim:
        raise ValueError(f"oracle expects a {ndim}-d array")
    es = a.itemsize
    return a, ctypes.c_void_p(a.ctypes.data), tuple(s // es for s in a.strides)


def _check(cov, z1, z2, pos):
    if cov.shape[0] != pos.shape[0] or cov.shape[1] != z1.shape[0] or cov.shape[1] != z2.shape[0]:
        raise ValueError("shape mismatch")


def max_threads() -> int:
    return int(_load().gsc_oracle_max_threads())


def summate(cov_samples, z1, z2, pos, num_threads: int = 0) -> np.ndarray:
    """FP64 restatement of field.rs:37-65 (summator)."""
    lib = _load()
    cov, pc, (csd, csm) = _view(cov_samples, 2)
    z1, p1, (s1,) = _view(z1, 1)
    z2, p2, (s2,) = _view(z2, 1)
    pos, pp, (psd, psp) = _view(pos, 2)
    _check(cov, z1, z2, pos)
    n = pos.shape[1]
    out = np.empty(n, dtype=np.float64)
    lib.gsc_oracle_summate(cov.shape[0], cov.shape[1], n, pc, csd, csm, p1, s1, p2, s2,
                           pp, psd, psp, ctypes.c_void_p(out.ctypes.data), 1, int(num_threads or 0))
    return out


def summate_fourier(spectrum_factor, modes, z1, z2, pos, num_threads: int = 0) -> np.ndarray:
    """FP64 restatement of field.rs:219-249 (summator_fourier)."""
    lib = _load()
    sf, psf, (ssf,) = _view(spectrum_factor, 1)
    md, pm, (msd, msm) = _view(modes, 2)
    z1, p1, (s1,) = _view(z1, 1)
    z2, p2, (s2,) = _view(z2, 1)
    pos, pp, (psd, psp) = _view(pos, 2)
    _check(md, z1, z2, pos)
    if sf.shape[0] != md.shape[1]:
        raise ValueError("shape mismatch")
    n = pos.shape[1]
    out = np.empty(n, dtype=np.float64)
    lib.gsc_oracle_summate_fourier(md.shape[0], md.shape[1], n, psf, ssf, pm, msd, msm, p1, s1,
                                   p2, s2, pp, psd, psp, ctypes.c_void_p(out.ctypes.data), 1,
                                   int(num_threads or 0))
    return out


def summate_incompr(cov_samples, z1, z2, pos, num_threads: int = 0, variant: int = 0) -> np.ndarray:
    """FP64 restatement of field.rs:97-182 (summator_incompr), modes in index order.

    variant 0 is the reference's association (field.rs:138-151); variant 1 the
    projector form p_i = delta_i0 - k_i k_0 / |k|^2.
    """
    lib = _load()
    cov, pc, (csd, csm) = _view(cov_samples, 2)
    z1, p1, (s1,) = _view(z1, 1)
    z2, p2, (s2,) = _view(z2, 1)
    pos, pp, (psd, psp) = _view(pos, 2)
    _check(cov, z1, z2, pos)
    d, n = pos.shape
    out = np.empty((d, n), dtype=np.float64, order="F")
    rc = lib.gsc_oracle_summate_incompr(d, cov.shape[1], n, pc, csd, csm, p1, s1, p2, s2, pp,
                                        psd, psp, ctypes.c_void_p(out.ctypes.data), 1, d,
                                        int(variant), int(num_threads or 0))
    if rc != 0:
        raise ValueError("Only two- and three-dimensional problems are supported.")
    return out


def summate_incompr_modechunk(cov_samples, z1, z2, pos, num_threads: int = 0,
                              min_len: int = 100) -> np.ndarray:
    """summator_incompr in the reference's mode-outer chunked order (field.rs:130-163)."""
    lib = _load()
    cov, pc, (csd, csm) = _view(cov_samples, 2)
    z1, p1, (s1,) = _view(z1, 1)
    z2, p2, (s2,) = _view(z2, 1)
    pos, pp, (psd, psp) = _view(pos, 2)
    _check(cov, z1, z2, pos)
    d, n = pos.shape
    out = np.empty((d, n), dtype=np.float64, order="F")
    rc = lib.gsc_oracle_summate_incompr_modechunk(d, cov.shape[1], n, pc, csd, csm, p1, s1, p2,
                                                  s2, pp, psd, psp,
                                                  ctypes.c_void_p(out.ctypes.data), 1, d,
                                                  int(min_len), int(num_threads or 0))
    if rc == -1:
        raise ValueError("Only two- and three-dimensional problems are supported.")
    if rc != 0:
        raise MemoryError("oracle accumulators could not be allocated")
    return out


def grid_points(axes) -> np.ndarray:
    """Expand per-axis coordinates into a (d, Nx*Ny[*Nz]) point list, C order (last axis fastest)."""
    mesh = np.meshgrid(*[np.asarray(a, dtype=np.float64) for a in axes], indexing="ij")
    return np.stack([m.ravel() for m in mesh])


def calc_field_krige_and_variance(krige_mat, krig_vecs, cond):
    """(field, error) of krige.rs:24-70 (sequential condition order)."""
    m, mp, (m0, m1) = _view(krige_mat, 2)
    v, vp, (v0, v1) = _view(krig_vecs, 2)
    c, cp, (cs,) = _view(cond, 1)
    n, N = v.shape
    if m.shape != (n, n) or c.shape[0] != n:
        raise ValueError("shape mismatch")
    field = np.zeros(N)
    error = np.zeros(N)
    _load().gsc_oracle_krige(n, N, mp, m0, m1, vp, v0, v1, cp, cs,
                             ctypes.c_void_p(field.ctypes.data), 1,
                             ctypes.c_void_p(error.ctypes.data), 1)
    return field, error


def calc_field_krige(krige_mat, krig_vecs, cond):
    """field of krige.rs:88-118."""
    return calc_field_krige_and_variance(krige_mat, krig_vecs, cond)[0]
