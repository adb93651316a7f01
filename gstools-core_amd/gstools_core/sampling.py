"""Seeded spectral-mode sampling on the device (the caller side of the summators).

GSTools' RandMeth draws the wave vectors k from the covariance model's spectral
density and z1, z2 ~ N(0, 1) on the host, then calls ``summate``
(gen_benchmark_inputs.py:17-29).  ``sample_modes`` draws the same quantities
on the GPU from a seed (gsc_sample_modes):

* ``"gaussian"``:    k ~ N(0, len_scale^-2 I), the spectral law of the
  covariance exp(-r^2 / (2 len_scale^2)) (GSTools' Gaussian model, whose
  default rescale gives exp(-pi r^2 / (4 l^2)), corresponds to
  len_scale = l sqrt(2 / pi); gstools_core.randmeth applies that);
* ``"exponential"``: k = g / (len_scale |w|), g ~ N(0, I), w ~ N(0, 1)
  (isotropic multivariate t with one degree of freedom), the spectral law of
  the covariance exp(-r / len_scale).

The random bits are Philox4x64-10 keyed (seed, stream), identical word for word
to ``numpy.random.Philox(key=np.array([seed, stream], np.uint64)).random_raw()``; mode m uses words
8m .. 8m + 7 (four Box-Muller pairs: g_d = n_d, z1 = n_4, z2 = n_5, w = n_6), so
the first modes of a longer draw equal a shorter draw with the same seed.  The
stream is not GSTools' (whose numpy RandomState sequence depends on its own
seeding); it has the same distributions.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native

MODELS = {"gaussian": 0, "exponential": 1}


def _model(model: str) -> int:
    if model not in MODELS:
        raise ValueError(f"model must be one of {sorted(MODELS)}")
    return MODELS[model]


def sample_modes(model: str, dim: int, m: int, len_scale: float, seed: int, stream: int = 0,
                 device: int = -1):
    """Returns (cov_samples (dim, m), z1 (m), z2 (m)) as float64 numpy arrays."""
    cov = np.empty((dim, m), dtype=np.float64)
    z1 = np.empty(m, dtype=np.float64)
    z2 = np.empty(m, dtype=np.float64)
    opts = _native.make_opts(device=device)
    rc = _native.lib().gsc_sample_modes(_model(model), int(dim), int(m), float(len_scale),
                                        int(seed) & 0xFFFFFFFFFFFFFFFF,
                                        int(stream) & 0xFFFFFFFFFFFFFFFF, cov.ctypes.data,
                                        z1.ctypes.data, z2.ctypes.data, ctypes.byref(opts))
    _native.check(rc)
    return cov, z1, z2


def sample_modes_dev(model: str, dim: int, m: int, len_scale: float, seed: int, cov: int,
                     z1: int, z2: int, *, stream_id: int = 0, stream: int = 0) -> None:
    """Device form: cov (dim, m) row-major, z1, z2 are device pointers; the
    launch is enqueued on the HIP stream `stream` (0: null stream)."""
    opts = _native.make_opts()
    rc = _native.lib().gsc_sample_modes_dev(_model(model), int(dim), int(m), float(len_scale),
                                            int(seed) & 0xFFFFFFFFFFFFFFFF,
                                            int(stream_id) & 0xFFFFFFFFFFFFFFFF, cov, z1, z2,
                                            ctypes.byref(opts), ctypes.c_void_p(stream or None))
    _native.check(rc)
