"""GPU tier: on-device seeded mode sampling and seeded ensembles.

The device sampler must reproduce the oracle (oracle/sampling.py: the same
Philox4x64-10 words, Box-Muller and model transforms in numpy) to FP64 libm
accuracy (1e-13 relative), keep mode m independent of M, give the stated laws
on large draws, and a seeded ensemble must equal the explicit ensemble built
from the same draws.
"""
from __future__ import annotations

import numpy as np
import pytest
from scipy import stats

import oracle
from oracle import sampling as osamp
from gstools_core import _native
from gstools_core.ensemble import Ensemble
from gstools_core.sampling import sample_modes, sample_modes_dev

pytestmark = pytest.mark.gpu


def _close(a, b, rtol=1e-13):
    # relative to the array's scale (values near 0 compare in absolute terms)
    scale = max(1.0, float(np.max(np.abs(b)))) if np.size(b) else 1.0
    return float(np.max(np.abs(a - b))) / scale <= rtol if np.size(b) else True


@pytest.mark.parametrize("model", ["gaussian", "exponential"])
@pytest.mark.parametrize("dim,m,seed,stream", [(1, 1, 0, 0), (2, 1000, 19031977, 0),
                                               (3, 4097, 2**64 - 1, 3), (3, 257, 42, 2**63)])
def test_device_sampler_matches_oracle(model, dim, m, seed, stream):
    cov, z1, z2 = sample_modes(model, dim, m, 7.5, seed, stream)
    rc, r1, r2 = osamp.sample_modes(model, dim, m, 7.5, seed, stream)
    assert _close(z1, r1) and _close(z2, r2)
    # exponential wave vectors are heavy-tailed: compare per mode
    rel = np.abs(cov - rc) / np.maximum(np.abs(rc).max(axis=0), 1e-300)
    assert float(rel.max()) <= 1e-13


def test_device_entry_point_and_prefix_property():
    m = 3000
    buf = _native.DeviceBuffer(8 * 5 * m)
    sample_modes_dev("gaussian", 3, m, 2.0, 99, buf.ptr, buf.ptr + 8 * 3 * m, buf.ptr + 8 * 4 * m)
    _native.synchronize()
    flat = buf.to_array(np.float64, 5 * m)
    cov, z1, z2 = sample_modes("gaussian", 3, m, 2.0, 99)
    np.testing.assert_array_equal(flat[:3 * m].reshape(3, m), cov)
    np.testing.assert_array_equal(flat[3 * m:4 * m], z1)
    np.testing.assert_array_equal(flat[4 * m:], z2)
    c2, a2, b2 = sample_modes("gaussian", 3, 100, 2.0, 99)
    np.testing.assert_array_equal(c2, cov[:, :100])
    np.testing.assert_array_equal(a2, z1[:100])
    buf.free()


def test_laws_on_device():
    cov, z1, z2 = sample_modes("gaussian", 3, 1_000_000, 4.0, 5)
    for x in (cov[0] * 4.0, cov[2] * 4.0, z1, z2):
        assert stats.kstest(x, "norm").pvalue > 1e-3
    cov, z1, _ = sample_modes("exponential", 3, 1_000_000, 8.0, 6)
    for d in range(3):
        assert stats.kstest(cov[d] * 8.0, "cauchy").pvalue > 1e-3
    assert stats.kstest(z1, "norm").pvalue > 1e-3


def test_bad_arguments_raise():
    with pytest.raises(ValueError):
        sample_modes("matern", 2, 10, 1.0, 0)
    with pytest.raises(ValueError):
        sample_modes("gaussian", 4, 10, 1.0, 0)
    with pytest.raises(ValueError):
        sample_modes("gaussian", 2, 10, 0.0, 0)


def _explicit(model, grids, ms, ls, seeds, **kw):
    reals = []
    for g, m, l, s in zip(grids, ms, ls, seeds):
        cov, z1, z2 = sample_modes(model, len(g), m, l, s)
        reals.append((cov, z1, z2, g))
    return Ensemble(reals, **kw), reals


@pytest.mark.parametrize("model", ["gaussian", "exponential"])
def test_seeded_ensemble_equals_explicit_and_reseeds(model):
    rng = np.random.default_rng(77)
    grids = [[np.arange(nx) * 1.0, np.arange(ny) * 1.0]
             for nx, ny in ((64, 80), (200, 130), (17, 300), (256, 256), (1, 50))]
    ms = [300, 1000, 64, 1000, 10]
    ls = list(rng.uniform(5.0, 50.0, size=len(grids)))
    seeds = [11, 12, 13, 2**64 - 5, 15]
    kw = dict(out_dtype=np.float64, precision="fp16x3")
    ens = Ensemble.seeded(model, grids, ms, ls, seeds, **kw)
    ref, reals = _explicit(model, grids, ms, ls, seeds, **kw)
    ens.run()
    ref.run()
    a, b = ens.results(), ref.results()
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    # one realisation against the FP64 oracle
    cov, z1, z2, g = reals[1]
    pts = oracle.grid_points(g)
    sub = np.random.default_rng(3).choice(pts.shape[1], 2000, replace=False)
    want = oracle.summate(cov, z1, z2, pts[:, sub])
    assert np.abs(a[1].ravel()[sub] - want).max() / np.sqrt(ms[1]) <= 1e-5
    # new seeds: the next runs draw new modes; replays repeat them exactly
    seeds2 = [s + 1000 for s in seeds]
    ens.reseed(seeds2)
    ens.run()
    c = ens.results()
    ens.run()
    d = ens.results()
    ref2, _ = _explicit(model, grids, ms, ls, seeds2, **kw)
    ref2.run()
    e = ref2.results()
    for x, y, z in zip(c, d, e):
        np.testing.assert_array_equal(x, y)
        np.testing.assert_array_equal(x, z)
    assert not np.array_equal(c[1], a[1])
    for o in (ens, ref, ref2):
        o.close()
