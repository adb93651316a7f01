"""GPU tier: structured-grid summation (direct-on-grid and factored MFMA paths).

The oracle for a grid is the FP64 point-list oracle on the expanded grid
points (the product identity is exact in real arithmetic).  Tolerances per
split-precision format, on the field scaled by sqrt(1/M) (sigma = 1):
  fp16x3, bf16x6 and the direct path: 1e-5 (the project's target);
  bf16x3: 5e-5 (its 16-bit significand sits at the edge of 1e-5, reported);
  fp16x1: 5e-3, bf16x1: 5e-2 (single products, for speed only).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import sigma_tol
import gstools_core
from gstools_core import grid, workloads

pytestmark = pytest.mark.gpu

TOL = {"fp16x3": 1e-5, "bf16x6": 1e-5, "auto": 1e-5, "bf16x3": 5e-5, "fp16x1": 5e-3,
       "bf16x1": 5e-2}


def _scaled_err(a, b, m):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b))) / np.sqrt(m)


def _modes(d, m, ls, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(0, 1.0 / ls, size=(d, m)), rng.normal(size=m), rng.normal(size=m)


# (3, 300) and (300, 700) take axis 1 as the GEMM row axis (it pads less to
# 256-row tiles) and are stored transposed; (700, 300) keeps axis 0
@pytest.mark.parametrize("shape", [(37, 50), (130, 129), (5, 7, 200), (33, 17, 129), (16, 16, 16),
                                   (64, 1, 40), (3, 300), (300, 700), (700, 300)])
@pytest.mark.parametrize("precision", ["fp16x3", "bf16x3", "bf16x6", "fp16x1", "bf16x1"])
def test_factored_small_grids(shape, precision):
    d = len(shape)
    cov, z1, z2 = _modes(d, 150, 4.0, 1000 + sum(shape))
    axes = [np.arange(n) * 0.75 - 3.0 for n in shape]
    ref = oracle.summate(cov, z1, z2, oracle.grid_points(axes)).reshape(shape)
    res = {}
    for tile in ("256", "128"):
        for accum in ("f32x2", "f64"):
            g = grid.summate_grid(cov, z1, z2, axes, path="factored", precision=precision,
                                  tile=tile, accum=accum)
            assert g.shape == shape
            err = _scaled_err(g, ref, 150)
            print(f"{shape} {precision} tile {tile} {accum}: {err:.3e}")
            assert err <= TOL[precision]
            res[tile, accum] = g
    # both tile shapes against each other (they agree bit for bit when the
    # plans keep the same 2-D orientation, within the format's error otherwise)
    for accum in ("f32x2", "f64"):
        assert _scaled_err(res["128", accum], res["256", accum], 150) <= 2 * TOL[precision]


@pytest.mark.parametrize("shape", [(1000,), (37, 50), (5, 7, 200), (2, 3, 4)])
def test_direct_grid_equals_point_list_kernel(shape):
    d = len(shape)
    cov, z1, z2 = _modes(d, 300, 6.0, 7 + d)
    axes = [np.linspace(-10, 40, n) for n in shape]
    g = grid.summate_grid(cov, z1, z2, axes, path="direct")
    pts = oracle.grid_points(axes)
    p = gstools_core.summate(cov, z1, z2, pts, variant="valu")
    np.testing.assert_array_equal(g.ravel(), p)  # same kernel arithmetic, bit for bit
    ref = oracle.summate(cov, z1, z2, pts)
    assert _scaled_err(g.ravel(), ref, 300) <= 1e-5
    # the default point-list kernel (cell-local phases) agrees within tolerance
    loc = gstools_core.summate(cov, z1, z2, pts)
    assert _scaled_err(loc, ref, 300) <= 1e-5


def test_auto_path_and_empty_axes():
    cov, z1, z2 = _modes(3, 10, 2.0, 3)
    assert grid.grid_path(cov, [np.arange(64.0)] * 3) == "factored"
    assert grid.grid_path(cov, [np.arange(8.0)] * 3) == "direct"
    assert grid.grid_path(cov[:1], [np.arange(64.0)]) == "direct"
    r = grid.summate_grid(cov, z1, z2, [np.arange(4.0), np.zeros(0), np.arange(3.0)])
    assert r.shape == (4, 0, 3)


def test_strided_axes_and_f32_output():
    cov, z1, z2 = _modes(2, 200, 5.0, 4)
    big = np.arange(300.0)
    axes = [big[::3], big[::-2][:70]]
    g = grid.summate_grid(cov, z1, z2, axes, out_dtype=np.float32, scale=0.5)
    assert g.dtype == np.float32
    ref = oracle.summate(cov, z1, z2, oracle.grid_points([np.ascontiguousarray(a) for a in axes]))
    assert _scaled_err(g.ravel() / 0.5, ref, 200) <= 1e-5


def test_summate_structured_is_the_grid_entry_point():
    cov, z1, z2 = _modes(3, 300, 4.0, 6)
    axes = [np.arange(40.0), np.arange(0.0, 30.0, 0.5), np.arange(33.0)]
    a = gstools_core.summate_structured(cov, z1, z2, axes)
    b = grid.summate_grid(cov, z1, z2, axes)
    np.testing.assert_array_equal(a, b)
    f = gstools_core.summate_structured(cov, z1, z2, axes, out_dtype=np.float32, path="direct")
    ref = oracle.summate(cov, z1, z2, oracle.grid_points(axes))
    assert f.dtype == np.float32 and _scaled_err(f.ravel(), ref, 300) <= 1e-5


def test_c3_grid_256_factored_and_direct():
    w = workloads.c3()
    md = w.modes
    idx = workloads.subset_indices(w.n_points, 1_000_000, seed=3)
    pts = workloads.grid_points(w.axes, idx)
    ref = oracle.summate(md.cov_samples, md.z1, md.z2, pts)
    # every format and accumulation the benchmark times at this size
    cases = [(p, a) for p in ("fp16x3", "bf16x6", "bf16x3") for a in ("f64", "f32x2")]
    cases += [("fp16x3", "f32"), ("fp16x1", "f32")]
    for prec, acc in cases:
        g = grid.summate_grid(md.cov_samples, md.z1, md.z2, w.axes, path="factored",
                              precision=prec, out_dtype=np.float32, accum=acc)
        err = _scaled_err(g.ravel()[idx], ref, md.m)
        print(f"c3 factored {prec} accum {acc}: max|err|/sigma = {err:.3e}")
        assert err <= TOL[prec] * (2 if acc == "f32" and prec == "fp16x3" else 1)
    g = grid.summate_grid(md.cov_samples, md.z1, md.z2, w.axes, path="direct",
                          out_dtype=np.float32)
    err = _scaled_err(g.ravel()[idx], ref, md.m)
    print(f"c3 direct: max|err|/sigma = {err:.3e}")
    assert err <= 1e-5


def test_c5_grid_512_heavy_tailed(c5_subset):
    # every split format the benchmark times, on the seeded 10^6-point subset;
    # bf16x3 is reported against its own (looser) bound: its 16-bit
    # significand does not reach 1e-5 sigma
    w, idx, _, ref = c5_subset
    md = w.modes
    cases = (("fp16x3", "f64", 1e-5), ("fp16x3", "f32x2", 1e-5), ("fp16x3", "f32", 2e-5),
             ("bf16x6", "f32x2", 1e-5), ("bf16x3", "f32x2", TOL["bf16x3"]))
    for prec, acc, tol in cases:
        g = grid.summate_grid(md.cov_samples, md.z1, md.z2, w.axes, path="factored",
                              precision=prec, out_dtype=np.float32, accum=acc)
        err = _scaled_err(g.ravel()[idx], ref, md.m)
        print(f"c5 factored {prec} accum {acc}: max|err|/sigma = {err:.3e} on {idx.size} points")
        assert err <= tol
        del g
    g = grid.summate_grid(md.cov_samples, md.z1, md.z2, w.axes, path="direct",
                          out_dtype=np.float32)
    err = _scaled_err(g.ravel()[idx], ref, md.m)
    print(f"c5 direct: max|err|/sigma = {err:.3e}")
    assert err <= 1e-5


def test_factored_zero_modes_after_a_nonempty_call():
    # M = 0 is a zero field on every path; an earlier call must leave nothing
    # behind in the library's reused device buffers
    cov, z1, z2 = _modes(2, 300, 4.0, 9)
    axes = [np.arange(64.0), np.arange(64.0)]
    grid.summate_grid(cov, z1, z2, axes, path="factored")
    for path in ("factored", "auto", "direct"):
        r = grid.summate_grid(np.zeros((2, 0)), np.zeros(0), np.zeros(0), axes, path=path)
        np.testing.assert_array_equal(r, np.zeros((64, 64)))
    r = grid.summate_grid(np.zeros((3, 0)), np.zeros(0), np.zeros(0), [np.arange(40.0)] * 3,
                          path="factored", out_dtype=np.float32)
    np.testing.assert_array_equal(r, np.zeros((40, 40, 40), np.float32))


def test_c4_realisations_sample():
    ens = workloads.c4(count=512)
    for w in (ens[0], ens[1], ens[511]):
        md = w.modes
        g = grid.summate_grid(md.cov_samples, md.z1, md.z2, w.axes, out_dtype=np.float32)
        idx = workloads.subset_indices(w.n_points, 10_000, seed=4)
        ref = oracle.summate(md.cov_samples, md.z1, md.z2, workloads.grid_points(w.axes, idx))
        err = _scaled_err(g.ravel()[idx], ref, md.m)
        print(f"{w.name} {g.shape}: {err:.3e}")
        assert err <= 1e-5


# -- the incompressible and Fourier fields on grids ---------------------------------

@pytest.mark.parametrize("shape", [(37, 50), (300, 40), (3, 300), (5, 7, 200), (33, 17, 129),
                                   (16, 16, 16)])
@pytest.mark.parametrize("precision", ["fp16x3", "bf16x6"])
def test_incompr_grid_factored(shape, precision):
    d = len(shape)
    cov, z1, z2 = _modes(d, 150, 4.0, 2000 + sum(shape))
    axes = [np.arange(n) * 0.75 - 3.0 for n in shape]
    g = grid.summate_incompr_grid(cov, z1, z2, axes, path="factored", precision=precision)
    assert g.shape == (d,) + shape
    ref = oracle.summate_incompr(cov, z1, z2, oracle.grid_points(axes))
    err = _scaled_err(g.reshape(d, -1), ref, 150)
    print(f"incompr {shape} {precision}: {err:.3e}")
    assert err <= TOL[precision]


@pytest.mark.parametrize("shape", [(37, 50), (5, 7, 200), (2, 3, 4)])
@pytest.mark.parametrize("variant", ["valu", "local"])
def test_incompr_grid_direct(shape, variant):
    d = len(shape)
    cov, z1, z2 = _modes(d, 200, 5.0, 3000 + sum(shape))
    axes = [np.linspace(-10, 40, n) for n in shape]
    g = grid.summate_incompr_grid(cov, z1, z2, axes, path="direct", variant=variant,
                                  out_dtype=np.float32, scale=0.5)
    assert g.shape == (d,) + shape and g.dtype == np.float32
    pts = oracle.grid_points(axes)
    p = gstools_core.summate_incompr(cov, z1, z2, pts, variant=variant, scale=0.5,
                                     out_dtype=np.float32)
    np.testing.assert_array_equal(g.reshape(d, -1), p)  # same kernels on the expanded grid
    ref = oracle.summate_incompr(cov, z1, z2, pts)
    assert _scaled_err(g.reshape(d, -1) / 0.5, ref, 200) <= 1e-5


def test_incompr_grid_is_divergence_free_and_k0_nan():
    # the factored path keeps sum_i d_i u_i = 0 (central differences on a fine
    # mesh) and, like the reference (field.rs:138), gives NaN for k = 0
    cov, z1, z2 = _modes(3, 300, 3.0, 41)
    h = 0.05
    axes = [np.arange(32) * h + 5.0 for _ in range(3)]
    g = grid.summate_incompr_grid(cov, z1, z2, axes, path="factored")
    div = ((g[0, 2:, 1:-1, 1:-1] - g[0, :-2, 1:-1, 1:-1]) +
           (g[1, 1:-1, 2:, 1:-1] - g[1, 1:-1, :-2, 1:-1]) +
           (g[2, 1:-1, 1:-1, 2:] - g[2, 1:-1, 1:-1, :-2])) / (2 * h)
    grad = np.abs(g[0, 2:] - g[0, :-2]).max() / (2 * h)
    assert np.abs(div).max() < 2e-3 * grad
    cov0 = cov.copy()
    cov0[:, 7] = 0.0
    g0 = grid.summate_incompr_grid(cov0, z1, z2, [np.arange(20.0)] * 2 + [np.arange(20.0)],
                                   path="factored")
    assert np.isnan(g0).all()


@pytest.mark.parametrize("shape", [(37, 50), (3, 300), (5, 7, 200), (1000,)])
def test_fourier_grid(shape):
    d = len(shape)
    rng = np.random.default_rng(sum(shape))
    side = 6
    ax = (np.arange(side) - (side - 1) / 2.0) * (2 * np.pi / 60.0)
    kg = np.stack(np.meshgrid(*([ax] * d), indexing="ij")).reshape(d, -1)
    m = kg.shape[1]
    sf = np.exp(-(kg ** 2).sum(axis=0) * 25.0)
    z1, z2 = rng.normal(size=m), rng.normal(size=m)
    axes = [np.arange(n) * 0.9 for n in shape]
    pts = oracle.grid_points(axes)
    ref = oracle.summate_fourier(sf, kg, z1, z2, pts)
    for path in ("auto", "direct", "factored") if d > 1 else ("auto", "direct"):
        g = gstools_core.summate_fourier_structured(sf, kg, z1, z2, axes, path=path)
        assert g.shape == shape
        err = float(np.abs(g.ravel() - ref).max())
        print(f"fourier {shape} {path}: {err / sf.max() / np.sqrt(m):.3e}")
        assert err <= sigma_tol(m) * sf.max()


def test_c3_incompr_grid_256():
    # the 256^3 configuration as an incompressible field: factored (projector
    # folded into the last axis' table) against the FP64 oracle on 10^6 points
    w = workloads.c3()
    md = w.modes
    idx = workloads.subset_indices(w.n_points, 1_000_000, seed=13)
    pts = workloads.grid_points(w.axes, idx)
    ref = oracle.summate_incompr(md.cov_samples, md.z1, md.z2, pts)
    g = grid.summate_incompr_grid(md.cov_samples, md.z1, md.z2, w.axes, path="factored",
                                  precision="fp16x3", out_dtype=np.float32)
    err = _scaled_err(g.reshape(3, -1)[:, idx], ref, md.m)
    print(f"c3 incompr factored fp16x3: max|err|/sigma = {err:.3e} on {idx.size} points")
    assert err <= 1e-5


def test_product_form_single_call_2d():
    # a 2-D grid on uniform axes (np.arange / np.linspace) of several tiles
    # takes the product form on the host-array entry point (S per tile, one
    # tile of A and C); the device entry point, which does not see the axes on
    # the host, keeps the whole-grid tables.  Both meet the bar, and they
    # differ (the product form rounds other table entries), which shows that
    # the product path ran.  Irregular axes take the whole-grid tables on both.
    import torch
    from gstools_core import device  # noqa: F401  (loads the library)
    dev = torch.device("cuda:0")
    for k, axes in enumerate(([np.arange(600.0) * 0.7 - 3.0, np.linspace(-5.0, 80.0, 700)],
                              [np.sort(np.random.default_rng(5).uniform(0, 400, 600)),
                               np.linspace(-5.0, 80.0, 700)])):
        cov, z1, z2 = _modes(2, 500, 6.0, 21)
        h = grid.summate_grid(cov, z1, z2, axes, path="factored", precision="fp16x3")
        shape = [len(a) for a in axes]
        nb = grid.grid_workspace_bytes(2, 500, shape, path="factored", precision="fp16x3")
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
        tc, t1, t2, tax = t(cov), t(z1), t(z2), [t(a) for a in axes]
        out = torch.empty(shape, dtype=torch.float64, device=dev)
        work = torch.empty(nb, dtype=torch.uint8, device=dev)
        grid.summate_grid_dev(2, 500, tc.data_ptr(), t1.data_ptr(), t2.data_ptr(),
                              [a.data_ptr() for a in tax], shape, out.data_ptr(),
                              work.data_ptr(), nb, path="factored", precision="fp16x3")
        torch.cuda.synchronize()
        d = out.cpu().numpy()
        idx = np.random.default_rng(9).choice(h.size, 20000, replace=False)
        ref = oracle.summate(cov, z1, z2, oracle.grid_points(axes)[:, idx])
        eh, ed = _scaled_err(h.ravel()[idx], ref, 500), _scaled_err(d.ravel()[idx], ref, 500)
        print(f"2-D grid host {eh:.3e} device {ed:.3e} sigma")
        assert eh <= 1e-5 and ed <= 1e-5
        if k == 0:
            assert not np.array_equal(h, d)
        else:
            np.testing.assert_array_equal(h, d)
