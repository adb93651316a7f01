"""GPU tier: the randomization-method generators (gstools_core.randmeth).

The caller contract of the summators (GSTools' RandMeth / IncomprRandMeth):
field = sqrt(var / M) * summate(...) (+ nugget noise), and for incompressible
fields U e_1 + U sqrt(var / M) * summate_incompr(...).  GSTools is not
available, so these are pinned by the formulas (exact comparisons with the
summators) and by statistics of the generated fields.
"""
from __future__ import annotations

import numpy as np
import pytest

import gstools_core
from gstools_core import grid
from gstools_core.randmeth import IncomprRandMeth, RandMeth

pytestmark = pytest.mark.gpu


def test_scalar_field_is_scaled_sum():
    rm = RandMeth("gaussian", 2, var=2.5, len_scale=10.0, mode_no=800, seed=4)
    pos = np.random.default_rng(1).uniform(0, 300, size=(2, 50_000))
    f = rm(pos)
    s = gstools_core.summate(rm.cov_samples, rm.z_1, rm.z_2, pos)
    np.testing.assert_array_equal(f, s * np.sqrt(2.5 / 800))
    # same seed, same field; a new seed, new modes
    np.testing.assert_array_equal(RandMeth("gaussian", 2, var=2.5, len_scale=10.0,
                                           mode_no=800, seed=4)(pos), f)
    rm.reset_seed(5)
    assert not np.array_equal(rm(pos), f)


def test_structured_mesh_uses_the_grid_path():
    rm = RandMeth("exponential", 3, var=1.0, len_scale=8.0, mode_no=500, seed=9)
    axes = [np.arange(40) * 1.0, np.arange(33) * 0.5, np.arange(20) * 2.0]
    f = rm(axes, mesh_type="structured")
    assert f.shape == (40, 33, 20)
    g = grid.summate_grid(rm.cov_samples, rm.z_1, rm.z_2, axes, scale=rm.scale)
    np.testing.assert_array_equal(f, g)
    pts = np.stack([a.ravel() for a in np.meshgrid(*axes, indexing="ij")])
    u = rm(pts)
    assert np.abs(f.ravel() - u).max() <= 1e-5  # sigma = sqrt(var) = 1


def test_field_statistics():
    var = 2.5
    rm = RandMeth("gaussian", 2, var=var, len_scale=10.0, mode_no=1000, seed=21)
    pos = np.random.default_rng(2).uniform(0, 2000, size=(2, 1_000_000))
    f = rm(pos)
    assert abs(f.mean()) < 0.1 * np.sqrt(var)
    assert abs(f.var() / var - 1.0) < 0.1
    # spatial variance of one realisation: (var / M) sum (z1^2 + z2^2) / 2
    want = var / rm.mode_no * 0.5 * np.sum(rm.z_1 ** 2 + rm.z_2 ** 2)
    assert abs(f.var() / want - 1.0) < 0.05


def test_nugget_adds_white_noise():
    pos = np.random.default_rng(3).uniform(0, 100, size=(3, 400_000))
    a = RandMeth("gaussian", 3, len_scale=5.0, mode_no=200, seed=8, nugget=0.0)(pos)
    b = RandMeth("gaussian", 3, len_scale=5.0, mode_no=200, seed=8, nugget=0.5)(pos)
    noise = b - a
    assert abs(noise.var() / 0.5 - 1.0) < 0.02
    assert abs(noise.mean()) < 0.01
    c = RandMeth("gaussian", 3, len_scale=5.0, mode_no=200, seed=8, nugget=0.5)(pos,
                                                                               add_nugget=False)
    np.testing.assert_array_equal(a, c)


@pytest.mark.parametrize("dim", [2, 3])
def test_incompressible_field(dim):
    u0 = 2.0
    rm = IncomprRandMeth("gaussian", dim, mean_velocity=u0, var=1.0, len_scale=10.0,
                         mode_no=1000, seed=30 + dim)
    pos = np.random.default_rng(4).uniform(0, 1500, size=(dim, 300_000))
    f = rm(pos)
    s = gstools_core.summate_incompr(rm.cov_samples, rm.z_1, rm.z_2, pos)
    want = u0 * np.sqrt(1.0 / 1000) * s
    want[0] += u0
    np.testing.assert_allclose(f, want, rtol=0, atol=1e-12)
    assert abs(f[0].mean() - u0) < 0.1 * u0
    for c in range(1, dim):
        assert abs(f[c].mean()) < 0.1 * u0
    # divergence-free: central differences on a fine structured mesh
    h = 0.01
    axes = [np.arange(3) * h + 37.0 for _ in range(dim)]
    g = rm(axes, mesh_type="structured")
    centre = (1,) * dim
    div = 0.0
    for c in range(dim):
        hi = list(centre)
        lo = list(centre)
        hi[c], lo[c] = 2, 0
        div += (g[(c,) + tuple(hi)] - g[(c,) + tuple(lo)]) / (2 * h)
    grad = abs(g[(0,) + (2,) * dim] - g[(0,) + (0,) * dim]) / (2 * h) + 1e-3
    assert abs(div) < 1e-3 * max(1.0, grad)


def test_bad_arguments():
    with pytest.raises(ValueError):
        IncomprRandMeth("gaussian", 1)
    rm = RandMeth("gaussian", 2, mode_no=10, seed=1)
    with pytest.raises(ValueError):
        rm(np.zeros((3, 5)))
    with pytest.raises(ValueError):
        rm([np.arange(3.0)], mesh_type="structured")
    with pytest.raises(ValueError):
        rm(np.zeros((2, 5)), mesh_type="hexagonal")


@pytest.mark.parametrize("model,rescale", [("gaussian", True), ("gaussian", False),
                                           ("exponential", True)])
def test_covariance_of_the_drawn_modes(model, rescale):
    # E[cos(k . r)] over the spectral law is the model's correlation at lag r:
    # GSTools' Gaussian exp(-pi r^2 / (4 l^2)) (rescale), the sampler's
    # exp(-r^2 / (2 l^2)) (no rescale), the exponential exp(-r / l); averaged
    # over M = 20000 drawn modes, along each axis
    ls = 10.0
    rm = RandMeth(model, 3, len_scale=ls, mode_no=20_000, seed=77, rescale=rescale)
    for r in (0.25 * ls, ls, 2.0 * ls):
        if model == "exponential":
            want = np.exp(-r / ls)
        elif rescale:
            want = np.exp(-np.pi * r * r / (4.0 * ls * ls))
        else:
            want = np.exp(-r * r / (2.0 * ls * ls))
        for axis in range(3):
            got = float(np.mean(np.cos(rm.cov_samples[axis] * r)))
            assert abs(got - want) < 0.03, (model, rescale, r, axis, got, want)


@pytest.mark.parametrize("dim", [2, 3])
def test_incompressible_structured_mesh_uses_the_grid_path(dim):
    rm = IncomprRandMeth("gaussian", dim, mean_velocity=1.5, len_scale=6.0, mode_no=400,
                         seed=90 + dim)
    axes = [np.arange(n) * 0.7 for n in ((50, 37) if dim == 2 else (20, 24, 33))]
    f = rm(axes, mesh_type="structured")
    assert f.shape == (dim,) + tuple(len(a) for a in axes)
    g = grid.summate_incompr_grid(rm.cov_samples, rm.z_1, rm.z_2, axes,
                                  scale=rm.mean_u * rm.scale)
    g[0] += rm.mean_u
    np.testing.assert_array_equal(f, g)
    pts = np.stack([a.ravel() for a in np.meshgrid(*axes, indexing="ij")])
    u = rm(pts)
    assert np.abs(f.reshape(dim, -1) - u).max() <= 1e-5 * rm.mean_u
