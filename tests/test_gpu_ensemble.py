"""GPU tier: ensembles (one graph-captured plan over many realisations) and the tuner.

Each realisation of an ensemble must equal the single-grid evaluation of the
same path (same kernels, same arithmetic) and the FP64 oracle within the
path's tolerance (1e-5 sigma for the direct path and fp16x3).  The graph is
replayed several times with eager library work in between, and every replay
is checked (graph replays must not depend on state left by earlier ones).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import gstools_core
from gstools_core import _native, grid, workloads
from gstools_core.ensemble import Ensemble

pytestmark = pytest.mark.gpu


def _reals(dim, shapes, ms, seed, jitter=False):
    # uniform axes (2-D factored realisations take the product form), or,
    # with jitter (a bool or one per realisation), irregular ones (tables of
    # the whole grid, as single-grid calls)
    rng = np.random.default_rng(seed)
    out = []
    for i, (sh, m) in enumerate(zip(shapes, ms)):
        ls = rng.uniform(3.0, 12.0)
        cov = rng.normal(0, 1.0 / ls, size=(dim, m))
        z1, z2 = rng.normal(size=m), rng.normal(size=m)
        axes = [np.arange(n) * rng.uniform(0.5, 1.5) + rng.uniform(-5, 5) for n in sh]
        jit = jitter[i] if isinstance(jitter, (list, tuple)) else jitter
        if jit:
            axes = [a + rng.uniform(-0.2, 0.2, size=a.size) for a in axes]
        out.append((cov, z1, z2, axes))
    return out


def _tol(dim, m, jitter):
    # a realisation is evaluated like its single-grid (host-array) call: the
    # same plan (the product form on uniform 2-D axes when it is smaller, the
    # whole-grid tables otherwise) and the same tile arithmetic, so they agree
    # to rounding of the output conversion
    return 1e-9 * np.sqrt(max(m, 1))


SHAPES_2D = [(8, 9), (64, 100), (130, 257), (1, 40), (200, 3), (77, 129), (256, 256)]
MS_2D = [50, 300, 1000, 17, 64, 513, 1000]


@pytest.mark.parametrize("jitter", [False, True])
@pytest.mark.parametrize("mode", ["auto", "direct", "factored", "mixed"])
def test_ensemble_2d_matches_single_grid_calls(mode, jitter):
    reals = _reals(2, SHAPES_2D, MS_2D, 10, jitter)
    if mode == "auto":
        paths = None
    elif mode == "mixed":
        paths = ["direct" if i % 2 else "factored" for i in range(len(reals))]
    else:
        paths = [mode] * len(reals)
    ens = Ensemble(reals, paths=paths, out_dtype=np.float64, precision="fp16x3")
    info = ens.info
    assert info["factored"] + info["direct"] == len(reals)
    ens.run()
    res = ens.results()
    for i, (cov, z1, z2, axes) in enumerate(reals):
        p = paths[i] if paths else grid.grid_path(cov, axes)
        if mode == "factored" and min(len(a) for a in axes) < 1:
            p = "direct"
        single = grid.summate_grid(cov, z1, z2, axes, path=p, precision="fp16x3")
        tol = _tol(2, cov.shape[1], jitter or p == "direct")
        np.testing.assert_allclose(res[i], single, rtol=0, atol=tol)
        ref = oracle.summate(cov, z1, z2, oracle.grid_points(axes)).reshape(single.shape)
        assert np.abs(res[i] - ref).max() / np.sqrt(cov.shape[1]) <= 1e-5
    ens.close()


def test_ensemble_3d_and_scales():
    shapes = [(16, 20, 33), (40, 8, 129), (5, 5, 5), (64, 64, 64)]
    ms = [100, 257, 30, 1000]
    reals = _reals(3, shapes, ms, 11)
    scales = [1.0 / np.sqrt(m) for m in ms]
    ens = Ensemble(reals, scales=scales, out_dtype=np.float32)
    ens.run()
    res = ens.results()
    for r, (cov, z1, z2, axes), s in zip(res, reals, scales):
        ref = oracle.summate(cov, z1, z2, oracle.grid_points(axes)).reshape(r.shape) * s
        assert r.dtype == np.float32
        assert np.abs(r - ref).max() <= 1e-5 + 2e-7 * np.abs(ref).max()
    ens.close()


def test_graph_replays_with_eager_work_between():
    reals = _reals(2, SHAPES_2D[:5], MS_2D[:5], 12)
    ens = Ensemble(reals, out_dtype=np.float64)
    ens.run()
    first = [r.copy() for r in ens.results()]
    rng = np.random.default_rng(0)
    for k in range(4):
        # eager library work and synchronisation between replays
        cov = rng.normal(size=(3, 40))
        gstools_core.summate(cov, rng.normal(size=40), rng.normal(size=40),
                             rng.normal(size=(3, 1000)))
        grid.summate_grid(reals[2][0], reals[2][1], reals[2][2], reals[2][3])
        ens.run()
        again = ens.results()
        for a, b in zip(first, again):
            np.testing.assert_array_equal(a, b)
    ens.close()


@pytest.mark.parametrize("dim,jitter", [(2, False), (2, True), (3, False)])
def test_many_realisations_match_single_grid_calls(dim, jitter):
    # a larger plan (48 factored realisations of uneven 2-D / 3-D grids): every
    # realisation must equal its single-grid evaluation, on every replay
    rng = np.random.default_rng(40 + dim)
    count = 48
    if dim == 2:
        shapes = [tuple(int(v) for v in rng.integers(17, 300, size=2)) for _ in range(count)]
    else:
        shapes = [tuple(int(v) for v in rng.integers(17, 48, size=3)) for _ in range(count)]
    ms = [int(v) for v in rng.integers(16, 400, size=count)]
    reals = _reals(dim, shapes, ms, 50 + dim, jitter)
    ens = Ensemble(reals, paths=["factored"] * count, out_dtype=np.float64, precision="fp16x3")
    assert ens.info["factored"] == count
    singles = [grid.summate_grid(cov, z1, z2, axes, path="factored", precision="fp16x3")
               for cov, z1, z2, axes in reals]
    worst = [0.0, 0.0]
    for rep in range(3):
        ens.run()
        res = ens.results()
        for i, (x, y) in enumerate(zip(res, singles)):
            np.testing.assert_allclose(x, y, rtol=0, atol=_tol(dim, ms[i], jitter),
                                       err_msg=f"replay {rep} realisation {i}")
            if rep == 0:
                cov, z1, z2, axes = reals[i]
                ref = oracle.summate(cov, z1, z2, oracle.grid_points(axes)).reshape(x.shape)
                e = [np.abs(x - ref).max() / np.sqrt(ms[i]), np.abs(y - ref).max() / np.sqrt(ms[i])]
                assert e[0] <= 1e-5, f"realisation {i}: {e[0]:.3e} sigma"
                worst = [max(a, b) for a, b in zip(worst, e)]
        grid.summate_grid(reals[0][0], reals[0][1], reals[0][2], reals[0][3])
    print(f"dim {dim} jitter {jitter}: max|err|/sigma ensemble {worst[0]:.3e}, "
          f"single calls {worst[1]:.3e}")
    ens.close()


@pytest.mark.parametrize("dim,jitter", [(2, False), (2, True), (3, False)])
def test_table_batches_equal_one_region(dim, jitter):
    # a bound on the tables held at once (table_bytes) splits the factored
    # realisations into batches that build and use their tables in turn in
    # one region: a smaller workspace, the same values bit for bit, on every
    # replay, with direct realisations in the same plan
    rng = np.random.default_rng(70 + dim)
    count = 40
    if dim == 2:
        shapes = [tuple(int(v) for v in rng.integers(17, 300, size=2)) for _ in range(count)]
    else:
        shapes = [tuple(int(v) for v in rng.integers(17, 48, size=3)) for _ in range(count)]
    ms = [int(v) for v in rng.integers(16, 400, size=count)]
    paths = ["direct" if i % 7 == 3 else "factored" for i in range(count)]
    reals = _reals(dim, shapes, ms, 80 + dim, jitter)
    kw = dict(paths=paths, out_dtype=np.float64, precision="fp16x3")
    whole = Ensemble(reals, **kw)
    whole.run()
    ref = [x.copy() for x in whole.results()]
    ws_whole = whole.info["workspace_bytes"]
    whole.close()
    bounded = Ensemble(reals, table_bytes=ws_whole // 5, **kw)
    assert bounded.info["workspace_bytes"] < ws_whole // 2
    for rep in range(3):
        bounded.run()
        for i, (x, y) in enumerate(zip(bounded.results(), ref)):
            np.testing.assert_array_equal(x, y, err_msg=f"replay {rep} realisation {i}")
        grid.summate_grid(reals[0][0], reals[0][1], reals[0][2], reals[0][3])
    print(f"dim {dim}: workspace {ws_whole / 1e6:.1f} MB -> "
          f"{bounded.info['workspace_bytes'] / 1e6:.1f} MB")
    bounded.close()
    with pytest.raises(ValueError):
        Ensemble(reals, table_bytes=-1, **kw)


def test_fp64_flush_plan_two_blocks_per_cu():
    # the FP64-flush GEMM runs two persistent blocks per CU by default; every
    # realisation equals its single-grid evaluation with the same accumulation
    rng = np.random.default_rng(77)
    count = 20
    shapes = [tuple(int(v) for v in rng.integers(30, 260, size=2)) for _ in range(count)]
    ms = [int(v) for v in rng.integers(16, 600, size=count)]
    reals = _reals(2, shapes, ms, 78)
    ens = Ensemble(reals, paths=["factored"] * count, out_dtype=np.float64, precision="fp16x3",
                   accum="f64")
    singles = [grid.summate_grid(cov, z1, z2, axes, path="factored", precision="fp16x3",
                                 accum="f64") for cov, z1, z2, axes in reals]
    for rep in range(2):
        ens.run()
        for i, (x, y) in enumerate(zip(ens.results(), singles)):
            np.testing.assert_allclose(x, y, rtol=0, atol=_tol(2, ms[i], False),
                                       err_msg=f"replay {rep} realisation {i}")
    ens.close()


@pytest.mark.parametrize("accum", ["f32x2", "f32"])
def test_c4_all_realisations_in_one_plan(accum):
    # the plan the benchmark times: all 512 realisations, fp16x3, FP32 output,
    # each checked against the FP64 oracle on 10^4 seeded points; with the
    # two-level accumulation (the library's) and with one FP32 sum (a tuner
    # candidate at M <= 1024)
    ens_w = workloads.c4(count=512)
    reals = [(w.modes.cov_samples, w.modes.z1, w.modes.z2, w.axes) for w in ens_w]
    scales = [w.modes.scale for w in ens_w]
    ens = Ensemble(reals, scales=scales, out_dtype=np.float32, precision="fp16x3", accum=accum)
    assert ens.info["factored"] == 512
    ens.run()
    res = ens.results()
    worst, n_checked = 0.0, 0
    for r, (w, out) in enumerate(zip(ens_w, res)):
        idx = workloads.subset_indices(w.n_points, 10_000, seed=41 + r)
        ref = oracle.summate(w.modes.cov_samples, w.modes.z1, w.modes.z2,
                             workloads.grid_points(w.axes, idx)) * w.modes.scale
        e = float(np.abs(out.ravel()[idx].astype(np.float64) - ref).max())
        assert e <= 1e-5, f"realisation {r} {out.shape}: {e:.3e}"
        worst = max(worst, e)
        n_checked += idx.size
    print(f"c4 all 512 realisations, accumulation {accum} ({n_checked} points) "
          f"max|err|/sigma = {worst:.3e}")
    ens.close()


def test_ensemble_zero_mode_realisation():
    # a realisation without modes is a zero grid (taken by the direct kernel),
    # next to factored ones
    reals = _reals(2, [(64, 64), (64, 64), (100, 40)], [300, 0, 200], 14)
    reals[1] = (np.zeros((2, 0)), np.zeros(0), np.zeros(0), reals[1][3])
    ens = Ensemble(reals, paths=["factored"] * 3, out_dtype=np.float64)
    ens.run()
    res = ens.results()
    np.testing.assert_array_equal(res[1], np.zeros((64, 64)))
    for i in (0, 2):
        cov, z1, z2, axes = reals[i]
        ref = oracle.summate(cov, z1, z2, oracle.grid_points(axes)).reshape(res[i].shape)
        assert np.abs(res[i] - ref).max() / np.sqrt(cov.shape[1]) <= 1e-5
    ens.close()


def test_tuner_measures_chooses_and_caches(tmp_path):
    from gstools_core.tuner import Tuner, padded_work
    t = Tuner(path=str(tmp_path / "tune.json"), repeats=3)
    big = t.choose(2, (300, 400), 1000)
    small = t.choose(2, (8, 9), 50)
    assert t.measured == 2
    for c in (big, small):
        assert c["path"] in ("direct", "factored") and c["ms"] > 0
        if c["path"] == "factored":
            assert c["precision"] in ("fp16x3", "bf16x6") and c["tile"] in ("256", "128")
    print("chosen", big, small)
    assert big["path"] == "factored"  # a 512 x 512 class at M = 1024 favours the MFMA path
    table = t.class_times(2, (300, 400), 1000)
    # direct + 2 formats x 3 accumulations (f32 joins at M <= 1024) x 2 tiles
    assert len(table["ms"]) == 13 and table["best"] in table["ms"]
    assert len(t.class_times(2, (8, 9), 50)["ms"]) == 13
    assert "f32" not in t.accums_for(2048) and "f32" in t.accums_for(1024)
    # a fixed format is honoured; a second tuner reads the cache without measuring
    c = t.choose(2, (300, 400), 1000, precision="bf16x6", accum="f64")
    assert c["path"] == "direct" or (c["precision"], c["accum"]) == ("bf16x6", "f64")
    t2 = Tuner(path=str(tmp_path / "tune.json"))
    assert t2.choose(2, (300, 400), 1000) == big and t2.measured == 0
    # padding model: a grid inside a class never has more padded work than it
    for key in table["ms"]:
        assert padded_work((300, 400), 1000, key) <= padded_work((512, 512), 1024, key)
    reals = _reals(2, [(300, 400), (8, 9), (64, 70)], [1000, 50, 300], 13)
    ens = Ensemble(reals, tune=True, tuner=t2, out_dtype=np.float64)
    assert {"precision", "accum", "paths", "tiles", "plan_ms", "kept"} <= set(ens.tuned)
    # the kept plan is the fastest of the three timed as whole plans
    assert ens.tuned["plan_ms"][ens.tuned["kept"]] == min(ens.tuned["plan_ms"].values())
    ens.run()
    for r, (cov, z1, z2, axes) in zip(ens.results(), reals):
        ref = oracle.summate(cov, z1, z2, oracle.grid_points(axes)).reshape(r.shape)
        assert np.abs(r - ref).max() / np.sqrt(cov.shape[1]) <= 1e-5
    ens.close()


def test_mixed_tile_plan_matches_single_grid_calls():
    # one plan with both GEMM tile shapes (two persistent launches) and direct
    # realisations: each realisation equals its single-grid call with its tile
    rng = np.random.default_rng(88)
    count = 16
    shapes = [tuple(int(v) for v in rng.integers(20, 300, size=2)) for _ in range(count)]
    ms = [int(v) for v in rng.integers(16, 400, size=count)]
    jitter = [i % 2 == 1 for i in range(count)]  # both operand forms in both tile shapes
    reals = _reals(2, shapes, ms, 89, jitter)
    tiles = ["128" if i % 3 else "256" for i in range(count)]
    paths = ["direct" if i % 5 == 4 else "factored" for i in range(count)]
    ens = Ensemble(reals, paths=paths, tiles=tiles, out_dtype=np.float64, precision="fp16x3")
    for rep in range(2):
        ens.run()
        for i, (x, (cov, z1, z2, axes)) in enumerate(zip(ens.results(), reals)):
            y = grid.summate_grid(cov, z1, z2, axes, path=paths[i], precision="fp16x3",
                                  tile=tiles[i])
            tol = _tol(2, ms[i], jitter[i] or paths[i] == "direct")
            np.testing.assert_allclose(x, y, rtol=0, atol=tol,
                                       err_msg=f"replay {rep} realisation {i}")
            ref = oracle.summate(cov, z1, z2, oracle.grid_points(axes)).reshape(x.shape)
            assert np.abs(x - ref).max() / np.sqrt(ms[i]) <= 1e-5
    ens.close()


def test_device_buffers_roundtrip():
    a = np.arange(1000, dtype=np.float64)
    b = _native.DeviceBuffer.from_array(a)
    np.testing.assert_array_equal(b.to_array(np.float64, 1000), a)
    b.free()


@pytest.mark.parametrize("dim", [2, 3])
def test_ensemble_128_row_tiles_match_single_grid_calls(dim):
    rng = np.random.default_rng(60 + dim)
    count = 24
    if dim == 2:
        shapes = [tuple(int(v) for v in rng.integers(17, 400, size=2)) for _ in range(count)]
    else:
        shapes = [tuple(int(v) for v in rng.integers(17, 60, size=3)) for _ in range(count)]
    ms = [int(v) for v in rng.integers(16, 500, size=count)]
    reals = _reals(dim, shapes, ms, 70 + dim)
    for accum in ("f32x2", "f64"):
        ens = Ensemble(reals, paths=["factored"] * count, out_dtype=np.float64,
                       precision="fp16x3", tile="128", accum=accum)
        singles = [grid.summate_grid(cov, z1, z2, axes, path="factored", precision="fp16x3",
                                     tile="128", accum=accum) for cov, z1, z2, axes in reals]
        for rep in range(2):
            ens.run()
            for i, (x, y) in enumerate(zip(ens.results(), singles)):
                np.testing.assert_allclose(x, y, rtol=0, atol=_tol(dim, ms[i], False),
                                           err_msg=f"{accum} replay {rep} realisation {i}")
        ens.close()
