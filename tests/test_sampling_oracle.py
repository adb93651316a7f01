"""CPU tier: the oracle of the seeded mode sampler (oracle/sampling.py).

Pins the random words against an independent pure-Python restatement of
Philox4x64-10 (Salmon et al., SC 2011; the constants of Random123), and the
transforms against the laws they must produce.
"""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy import stats

from oracle import sampling as osamp

_M0, _M1 = 0xD2E7470EE14C6C93, 0xCA5A826395121157
_W0, _W1 = 0x9E3779B97F4A7C15, 0xBB67AE8584CAA73B
_MASK = (1 << 64) - 1


def _philox4x64_10(ctr, key):
    c, k = list(ctr), list(key)
    for _ in range(10):
        p0, p1 = _M0 * c[0], _M1 * c[2]
        c = [(p1 >> 64) ^ c[1] ^ k[0], p1 & _MASK, (p0 >> 64) ^ c[3] ^ k[1], p0 & _MASK]
        k = [(k[0] + _W0) & _MASK, (k[1] + _W1) & _MASK]
    return c


@pytest.mark.parametrize("seed,stream", [(0, 0), (19031977, 0), (2**64 - 1, 5), (123, 2**63)])
def test_words_are_philox4x64_10_of_counter_b_plus_1(seed, stream):
    words = osamp.philox_words(seed, stream, 24)
    for b in range(6):
        ref = _philox4x64_10([b + 1, 0, 0, 0], [seed, stream])
        assert [int(x) for x in words[4 * b:4 * b + 4]] == ref


def test_sincospi_reduction():
    x = np.concatenate([np.linspace(0, 2, 10001), [0.5, 1.0, 1.5, 0.25, 1e-300, 2 - 2**-52]])
    s, c = osamp.sincospi(x)
    # np.sin(np.pi * x) itself is off by up to ulp(pi x) ~ 4.4e-16 near x = 2
    np.testing.assert_allclose(s, np.sin(np.pi * x), atol=2e-15)
    np.testing.assert_allclose(c, np.cos(np.pi * x), atol=2e-15)
    assert s[-1] == pytest.approx(-np.pi * 2.0 ** -52, rel=1e-15)
    # exact zeros and ones at the quarter turns
    s, c = osamp.sincospi(np.array([0.0, 0.5, 1.0, 1.5]))
    np.testing.assert_array_equal(s, [0.0, 1.0, 0.0, -1.0])
    np.testing.assert_array_equal(np.abs(c), [1.0, 0.0, 1.0, 0.0])
    # near a zero of cos the relative accuracy stays: cos(pi (1/2 - e)) = sin(pi e)
    e = 2.0 ** -40
    assert osamp.sincospi(np.array([0.5 - e]))[1][0] == pytest.approx(np.sin(np.pi * e), rel=1e-15)


def test_mode_m_does_not_depend_on_m_total():
    a = osamp.sample_modes("gaussian", 3, 50, 2.0, 7)
    b = osamp.sample_modes("gaussian", 3, 200, 2.0, 7)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y[..., :50])


def test_box_muller_of_known_uniforms():
    # one block by hand: u = (w >> 11) 2^-53, pairs (u0, u1) -> rho (cos, sin)(2 pi u1)
    w = osamp.philox_words(3, 0, 8)
    u = [(int(x) >> 11) * 2.0 ** -53 for x in w]
    n = osamp.normals(3, 0, 1)[0]
    for i in range(4):
        rho = math.sqrt(-2.0 * math.log(1.0 - u[2 * i]))
        assert n[2 * i] == pytest.approx(rho * math.cos(2 * math.pi * u[2 * i + 1]),
                                         rel=1e-14, abs=1e-15)
        assert n[2 * i + 1] == pytest.approx(rho * math.sin(2 * math.pi * u[2 * i + 1]),
                                             rel=1e-14, abs=1e-15)


def test_gaussian_model_laws():
    cov, z1, z2 = osamp.sample_modes("gaussian", 3, 200_000, 4.0, 11)
    for x in (cov[0] * 4.0, cov[1] * 4.0, cov[2] * 4.0, z1, z2):
        assert stats.kstest(x, "norm").pvalue > 1e-3
    # independence of the components (correlations near 0)
    c = np.corrcoef(np.vstack([cov, z1, z2]))
    assert np.max(np.abs(c - np.eye(5))) < 0.01


def test_exponential_model_is_multivariate_t_one_dof():
    ls = 8.0
    cov, z1, z2 = osamp.sample_modes("exponential", 3, 200_000, ls, 12)
    # each component of l k is standard Cauchy (t with 1 dof)
    for d in range(3):
        assert stats.kstest(cov[d] * ls, "cauchy").pvalue > 1e-3
    # |l k|^2 / 3 is F(3, 1) distributed
    f = np.sum((cov * ls) ** 2, axis=0) / 3.0
    assert stats.kstest(f, stats.f(3, 1).cdf).pvalue > 1e-3
    assert stats.kstest(z1, "norm").pvalue > 1e-3


def test_streams_are_independent():
    a = osamp.sample_modes("gaussian", 2, 100_000, 1.0, 5, stream=0)[1]
    b = osamp.sample_modes("gaussian", 2, 100_000, 1.0, 5, stream=1)[1]
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.02
    assert not np.array_equal(a, b)
