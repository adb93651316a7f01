"""CPU oracle of the seeded mode sampler -- TEST INFRASTRUCTURE ONLY.

Restates gstools_core's on-device sampler (csrc/sampling.hip) in numpy so the
tests can check the device draws value by value.  Only ``tests/`` may import it.

The reference has no sampler: GSTools' RandMeth (the caller, referenced at
gen_benchmark_inputs.py:17-29 and not part of /root/reference) draws the modes
on the host.  The distributions follow the benchmark's rules: k ~ N(0, l^-2 I)
(Gaussian model, the spectral law of the covariance exp(-r^2 / (2 l^2))) and
k = g / (l |w|) (exponential model, isotropic multivariate t with one degree
of freedom, the spectral law of exp(-r / l)), z1, z2 ~ N(0, 1).

What pins each part:
* the random words: ``numpy.random.Philox`` is Philox4x64-10 (Random123); the
  tests also restate the cipher in pure Python and compare word for word;
* Box-Muller and the model transforms: closed forms, checked by distribution
  tests (Kolmogorov-Smirnov against the normal and Cauchy laws) in the tests.
"""
from __future__ import annotations

import numpy as np

MODELS = {"gaussian": 0, "exponential": 1}


def philox_words(seed: int, stream: int, n_words: int) -> np.ndarray:
    """The first n_words 64-bit words of Philox4x64-10 keyed (seed, stream)."""
    key = np.array([int(seed) & 0xFFFFFFFFFFFFFFFF, int(stream) & 0xFFFFFFFFFFFFFFFF],
                   dtype=np.uint64)  # numpy reads a list holding 2^64 - 1 through float64
    bg = np.random.Philox(key=key)
    return bg.random_raw(n_words).astype(np.uint64)


def sincospi(x: np.ndarray):
    """(sin(pi x), cos(pi x)) with an exact reduction to |r| <= 1/4 of a half
    turn, so values near the zeros keep full relative accuracy (as the
    device's sincospi does; sin(2 pi u) computed as np.sin(2 * np.pi * u)
    would not)."""
    q = np.rint(2.0 * x)
    r = x - 0.5 * q  # exact
    s, c = np.sin(np.pi * r), np.cos(np.pi * r)
    k = q.astype(np.int64) & 3
    sin = np.select([k == 0, k == 1, k == 2], [s, c, -s], -c)
    cos = np.select([k == 0, k == 1, k == 2], [c, -s, -c], s)
    return sin, cos


def normals(seed: int, stream: int, m: int) -> np.ndarray:
    """(m, 8) standard normals: row i from words 8i .. 8i + 7 by Box-Muller."""
    w = philox_words(seed, stream, 8 * m).reshape(m, 8)
    u = (w >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    rho = np.sqrt(-2.0 * np.log(1.0 - u[:, 0::2]))
    sin, cos = sincospi(2.0 * u[:, 1::2])
    n = np.empty((m, 8))
    n[:, 0::2] = rho * cos
    n[:, 1::2] = rho * sin
    return n


def sample_modes(model: str, dim: int, m: int, len_scale: float, seed: int, stream: int = 0):
    """(cov_samples (dim, m), z1, z2) as drawn by gsc_sample_modes."""
    n = normals(seed, stream, m)
    if model == "gaussian":
        scale = np.full(m, 1.0 / len_scale)
    elif model == "exponential":
        scale = 1.0 / (len_scale * np.maximum(np.abs(n[:, 6]), 2.0 ** -1022))
    else:
        raise ValueError(f"model must be one of {sorted(MODELS)}")
    cov = (n[:, :dim] * scale[:, None]).T.copy()
    return cov, n[:, 4].copy(), n[:, 5].copy()
