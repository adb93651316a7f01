"""gstools_core -- MI355X (gfx950) drop-in for GSTools-Core's randomization summators.

Same module name, function names, positional/keyword parameters, defaults,
return shapes and dtypes as the reference PyO3 module (src/lib.rs:29-84):

    summate(cov_samples, z1, z2, pos, num_threads=None)          -> float64 (N,)
    summate_incompr(cov_samples, z1, z2, pos, num_threads=None)  -> float64 (d, N), F order
    summate_fourier(spectrum_factor, modes, z1, z2, pos, num_threads=None) -> float64 (N,)

The sums run on the GPU through the C-ABI of libgstools_core_amd.so; there is
no CPU fallback.  Deliberate differences from the reference:
  * shape mismatches and unsupported dimensions raise ValueError instead of
    aborting the process (field.rs:44-46, 104-106, 180 panic under
    panic = "abort");
  * the GIL is released while the device works (ctypes drops it);
  * keyword-only extensions: out_dtype (np.float64 / np.float32), scale
    (multiplies the result, e.g. sqrt(var / M)), device (HIP ordinal),
    variant ("auto", the default: "local" from 256 modes up, "valu" below;
    "local": points sorted into spatial cells, FP64 phase at each cell centre
    and split-fp16 (scalar fields) or split-bf16 (incompressible fields)
    matrix-core phases inside the cell, 2-D and 3-D;
    "valu": FP64 phases on the vector ALUs;
    "mfma": FP64 phases in matrix-core tiles, d <= 3).
num_threads is accepted with the reference's meaning (worker threads); here
it bounds the host threads that stage strided inputs.

Extensions: summate_structured, summate_incompr_structured and
summate_fourier_structured (the three fields on a structured grid, here or
in gstools_core.grid with their workspace forms), gstools_core.ensemble.Ensemble
(many realisations in one graph-captured plan), gstools_core.randmeth
(RandMeth / IncomprRandMeth caller glue), and the kriging calculators
calc_field_krige_and_variance / calc_field_krige (krige.rs:24-118).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native

__version__ = "1.1.0"

__all__ = ["summate", "summate_incompr", "summate_fourier", "summate_structured",
           "summate_incompr_structured", "summate_fourier_structured",
           "calc_field_krige_and_variance", "calc_field_krige", "__version__"]


def _array_arg(name: str, a, ndim: int):
    """Argument extraction as PyReadonlyArray{1,2}<f64> does it (lib.rs:37-40)."""
    if not isinstance(a, np.ndarray):
        raise TypeError(f"argument '{name}': '{type(a).__name__}' object cannot be converted "
                        f"to 'PyArray<f64, Ix{ndim}>'")
    if a.dtype != np.dtype(np.float64) or not a.dtype.isnative:
        raise TypeError(f"argument '{name}': type mismatch: from={a.dtype}, to=float64")
    if a.ndim != ndim:
        raise TypeError(f"argument '{name}': dimensionality mismatch: from={a.ndim}, to={ndim}")
    # element strides must be whole elements; odd byte views are copied
    if any(s % 8 for s in a.strides) or a.ctypes.data % 8:
        a = np.ascontiguousarray(a)
    return a


def _threads_arg(num_threads) -> int:
    """Option<usize> extraction (lib.rs:41): None, or a non-negative int."""
    if num_threads is None:
        return 0
    if isinstance(num_threads, (bool, np.bool_)) or not isinstance(num_threads, (int, np.integer)):
        raise TypeError(f"argument 'num_threads': '{type(num_threads).__name__}' object cannot be "
                        "interpreted as an integer")
    if num_threads < 0:
        raise OverflowError("argument 'num_threads': can't convert negative int to unsigned")
    return min(int(num_threads), 1 << 30)


def _variant(variant) -> int:
    if variant not in _native.DIRECT_VARIANTS:
        raise ValueError(f"variant must be one of {sorted(_native.DIRECT_VARIANTS)}")
    return _native.DIRECT_VARIANTS[variant]


def _out_dtype(out_dtype):
    dt = np.dtype(out_dtype)
    if dt == np.float64:
        return dt, _native.GSC_F64
    if dt == np.float32:
        return dt, _native.GSC_F32
    raise ValueError("out_dtype must be float64 or float32")


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def _es(a: np.ndarray):
    return tuple(s // a.itemsize for s in a.strides)


def _check_modes(cov, z1, z2, pos):
    if cov.shape[0] != pos.shape[0]:
        raise ValueError(f"cov_samples has {cov.shape[0]} rows but pos has {pos.shape[0]} "
                         "(spatial dimensions differ)")
    if cov.shape[1] != z1.shape[0] or cov.shape[1] != z2.shape[0]:
        raise ValueError(f"cov_samples has {cov.shape[1]} modes but z1 has {z1.shape[0]} and "
                         f"z2 has {z2.shape[0]}")


def summate(cov_samples, z1, z2, pos, num_threads=None, *, out_dtype=np.float64, scale=1.0,
            device=-1, variant="auto"):
    """Scalar randomization-method field (field.rs:37-65, lib.rs:33-48).

    u(x_j) = sum_m z1[m] cos(k_m . x_j) + z2[m] sin(k_m . x_j), k_m = cov_samples[:, m].
    """
    cov = _array_arg("cov_samples", cov_samples, 2)
    z1 = _array_arg("z1", z1, 1)
    z2 = _array_arg("z2", z2, 1)
    pos = _array_arg("pos", pos, 2)
    nt = _threads_arg(num_threads)
    _check_modes(cov, z1, z2, pos)
    dt, code = _out_dtype(out_dtype)
    d, n = pos.shape
    out = np.empty(n, dtype=dt)
    if n == 0:
        return out
    opts = _native.make_opts(out_dtype=code, scale=scale, device=device, num_threads=nt,
                             direct_variant=_variant(variant))
    (csd, csm), (s1,), (s2,), (psd, psp) = _es(cov), _es(z1), _es(z2), _es(pos)
    rc = _native.lib().gsc_summate(d, cov.shape[1], n, _ptr(cov), csd, csm, _ptr(z1), s1,
                                   _ptr(z2), s2, _ptr(pos), psd, psp, _ptr(out), 1,
                                   ctypes.byref(opts))
    _native.check(rc)
    return out


def summate_incompr(cov_samples, z1, z2, pos, num_threads=None, *, out_dtype=np.float64,
                    scale=1.0, device=-1, variant="auto"):
    """Incompressible vector field (field.rs:97-182, lib.rs:50-65); d must be 2 or 3.

    Returns a (d, N) array in Fortran order, like the reference (field.rs:166-174).
    """
    cov = _array_arg("cov_samples", cov_samples, 2)
    z1 = _array_arg("z1", z1, 1)
    z2 = _array_arg("z2", z2, 1)
    pos = _array_arg("pos", pos, 2)
    nt = _threads_arg(num_threads)
    _check_modes(cov, z1, z2, pos)
    d, n = pos.shape
    if d not in (2, 3):
        raise ValueError("Only two- and three-dimensional problems are supported.")
    dt, code = _out_dtype(out_dtype)
    out = np.empty((d, n), dtype=dt, order="F")
    if n == 0:
        return out
    opts = _native.make_opts(out_dtype=code, scale=scale, device=device, num_threads=nt,
                             direct_variant=_variant(variant))
    (csd, csm), (s1,), (s2,), (psd, psp) = _es(cov), _es(z1), _es(z2), _es(pos)
    osd, osp = _es(out)
    rc = _native.lib().gsc_summate_incompr(d, cov.shape[1], n, _ptr(cov), csd, csm, _ptr(z1), s1,
                                           _ptr(z2), s2, _ptr(pos), psd, psp, _ptr(out), osd, osp,
                                           ctypes.byref(opts))
    _native.check(rc)
    return out


def summate_fourier(spectrum_factor, modes, z1, z2, pos, num_threads=None, *,
                    out_dtype=np.float64, scale=1.0, device=-1, variant="auto"):
    """Fourier-method field (field.rs:219-249, lib.rs:67-84).

    u(x_j) = sum_m sf[m] (z1[m] cos(k_m . x_j) + z2[m] sin(k_m . x_j)), k_m = modes[:, m].
    """
    sf = _array_arg("spectrum_factor", spectrum_factor, 1)
    md = _array_arg("modes", modes, 2)
    z1 = _array_arg("z1", z1, 1)
    z2 = _array_arg("z2", z2, 1)
    pos = _array_arg("pos", pos, 2)
    nt = _threads_arg(num_threads)
    _check_modes(md, z1, z2, pos)
    if sf.shape[0] != md.shape[1]:
        raise ValueError(f"modes has {md.shape[1]} columns but spectrum_factor has {sf.shape[0]}")
    dt, code = _out_dtype(out_dtype)
    d, n = pos.shape
    out = np.empty(n, dtype=dt)
    if n == 0:
        return out
    opts = _native.make_opts(out_dtype=code, scale=scale, device=device, num_threads=nt,
                             direct_variant=_variant(variant))
    (ssf,), (msd, msm), (s1,), (s2,), (psd, psp) = _es(sf), _es(md), _es(z1), _es(z2), _es(pos)
    rc = _native.lib().gsc_summate_fourier(d, md.shape[1], n, _ptr(sf), ssf, _ptr(md), msd, msm,
                                           _ptr(z1), s1, _ptr(z2), s2, _ptr(pos), psd, psp,
                                           _ptr(out), 1, ctypes.byref(opts))
    _native.check(rc)
    return out


def _krige_args(krige_mat, krig_vecs, cond, num_threads):
    m = _array_arg("krige_mat", krige_mat, 2)
    v = _array_arg("krig_vecs", krig_vecs, 2)
    c = _array_arg("cond", cond, 1)
    nt = _threads_arg(num_threads)
    # the reference asserts these (krige.rs:31-33, 80-82) and aborts
    if m.shape[0] != m.shape[1]:
        raise ValueError("krige_mat must be square")
    if m.shape[0] != v.shape[0]:
        raise ValueError("krige_mat and krig_vecs disagree on the number of conditions")
    if m.shape[0] != c.shape[0]:
        raise ValueError("krige_mat and cond disagree on the number of conditions")
    return m, v, c, nt


def _krige(m, v, c, nt, device, with_error):
    n, N = v.shape
    field = np.empty(N)
    error = np.empty(N) if with_error else None
    if N == 0:
        return field, error
    opts = _native.make_opts(device=device, num_threads=nt)
    (m0, m1), (v0, v1), (cs,) = _es(m), _es(v), _es(c)
    rc = _native.lib().gsc_krige(n, N, _ptr(m), m0, m1, _ptr(v), v0, v1, _ptr(c), cs,
                                 _ptr(field), 1, _ptr(error) if with_error else None, 1,
                                 ctypes.byref(opts))
    _native.check(rc)
    return field, error


def calc_field_krige_and_variance(krige_mat, krig_vecs, cond, num_threads=None, *, device=-1):
    """Kriging field and error variance (krige.rs:24-70, lib.rs:86-103).

    field[j] = sum_i cond[i] (krige_mat[:, i] . krig_vecs[:, j]),
    error[j] = sum_i krig_vecs[i, j] (krige_mat[:, i] . krig_vecs[:, j]).
    """
    m, v, c, nt = _krige_args(krige_mat, krig_vecs, cond, num_threads)
    return _krige(m, v, c, nt, device, True)


def calc_field_krige(krige_mat, krig_vecs, cond, num_threads=None, *, device=-1):
    """Kriging field alone (krige.rs:88-118, lib.rs:105-118)."""
    m, v, c, nt = _krige_args(krige_mat, krig_vecs, cond, num_threads)
    return _krige(m, v, c, nt, device, False)[0]



def summate_structured(cov_samples, z1, z2, axes, num_threads=None, **options):
    """``summate`` on the structured grid axes[0] x ... x axes[d-1] (d in 1..3),
    without expanding the grid into points: the field in C order, shape
    (len(axes[0]), ...).  The keyword-only options (path, precision,
    out_dtype, scale, device, accum, variant) are those of
    ``gstools_core.grid.summate_grid``, which this calls.
    """
    from .grid import summate_grid

    return summate_grid(cov_samples, z1, z2, axes, num_threads, **options)


def summate_incompr_structured(cov_samples, z1, z2, axes, num_threads=None, **options):
    """``summate_incompr`` on the structured grid axes[0] x ... x axes[d-1]
    (d in 2..3) without expanding it: shape (d, len(axes[0]), ...), C order.
    Options as for ``summate_structured``."""
    from .grid import summate_incompr_grid

    return summate_incompr_grid(cov_samples, z1, z2, axes, num_threads, **options)


def summate_fourier_structured(spectrum_factor, modes, z1, z2, axes, num_threads=None,
                               **options):
    """``summate_fourier`` on the structured grid axes[0] x ... x axes[d-1]
    without expanding it: the field in C order, shape (len(axes[0]), ...).
    Options as for ``summate_structured``."""
    from .grid import summate_fourier_grid

    return summate_fourier_grid(spectrum_factor, modes, z1, z2, axes, num_threads, **options)

