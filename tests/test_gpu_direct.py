"""GPU tier: the direct summators (HIP kernels) against the FP64 oracle.

Tolerance (stated by the project's accuracy target): max |err| <= 1e-5 sigma
on the field scaled by sqrt(var / M), var = 1; on the reference's unscaled
output that is 1e-5 * sqrt(M) (conftest.sigma_tol).  For the incompressible
field the max runs over all components.

Large configurations are computed in full on the GPU and checked against the
oracle on a seeded random subset of points; the subset sizes keep the oracle
to a few seconds on the host.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import sigma_tol

import oracle
import gstools_core
from gstools_core import workloads

pytestmark = pytest.mark.gpu


def _err(a, b):
    return float(np.max(np.abs(np.asarray(a, dtype=np.float64) - b))) if np.size(b) else 0.0


def test_device_visible():
    from gstools_core import _native
    assert _native.lib().gsc_device_count() >= 1


# -- the reference's golden vectors ----------------------------------------------

def test_golden_summate(field_rs):
    r = gstools_core.summate(field_rs["cov_samples"], field_rs["z1"], field_rs["z2"],
                             field_rs["pos"])
    assert r.shape == (8,) and r.dtype == np.float64
    assert _err(r, field_rs["summate"]) <= sigma_tol(10)


def test_golden_incompr(field_rs):
    r = gstools_core.summate_incompr(field_rs["cov_samples"], field_rs["z1"], field_rs["z2"],
                                     field_rs["pos"])
    assert r.shape == (3, 8) and r.flags.f_contiguous
    assert _err(r, field_rs["summate_incompr"]) <= sigma_tol(10)


def test_golden_fourier(field_rs):
    r = gstools_core.summate_fourier(field_rs["spectrum_factor"], field_rs["cov_samples"],
                                     field_rs["z1"], field_rs["z2"], field_rs["pos"])
    assert _err(r, field_rs["summate_fourier"]) <= sigma_tol(10) * np.abs(
        field_rs["spectrum_factor"]).max()


# -- benchmark configurations -----------------------------------------------------

def test_c1_full_against_oracle():
    w = workloads.c1()
    md = w.modes
    gpu = gstools_core.summate(md.cov_samples, md.z1, md.z2, w.pos)
    ref = oracle.summate(md.cov_samples, md.z1, md.z2, w.pos)
    err = _err(gpu, ref) * md.scale
    print(f"c1 max|err|/sigma = {err:.3e}")
    assert err <= 1e-5


def test_c2_subset_against_oracle():
    w = workloads.c2()
    md = w.modes
    gpu = gstools_core.summate(md.cov_samples, md.z1, md.z2, w.pos)
    idx = workloads.subset_indices(w.n_points, 1_000_000, seed=2)
    ref = oracle.summate(md.cov_samples, md.z1, md.z2, w.pos[:, idx])
    err = _err(gpu[idx], ref) * md.scale
    print(f"c2 max|err|/sigma = {err:.3e} on {idx.size} points")
    assert err <= 1e-5
    # the scaled field has unit variance (statistical sanity of the whole output)
    assert abs(float(np.var(gpu * md.scale)) - 1.0) < 0.1


def test_c2_incompr_subset_against_oracle():
    w = workloads.c2(n=4_000_000)
    md = w.modes
    gpu = gstools_core.summate_incompr(md.cov_samples, md.z1, md.z2, w.pos)
    assert gpu.shape == (3, 4_000_000)
    idx = workloads.subset_indices(w.n_points, 1_000_000, seed=3)
    ref = oracle.summate_incompr(md.cov_samples, md.z1, md.z2, w.pos[:, idx])
    err = _err(gpu[:, idx], ref) * md.scale
    print(f"c2i max|err|/sigma = {err:.3e} on {idx.size} points")
    assert err <= 1e-5


def test_heavy_tailed_large_phase(c5_subset):
    # exponential-model modes of the 512^3 heavy-tailed case on 10^6 scattered
    # points of that grid: |k . x| reaches ~1e5 rad; FP64 range reduction
    w, idx, pts, ref = c5_subset
    md = w.modes
    for variant in ("local", "valu"):
        gpu = gstools_core.summate(md.cov_samples, md.z1, md.z2, pts, variant=variant)
        err = _err(gpu, ref) * md.scale
        print(f"c5 point list ({variant}) max|err|/sigma = {err:.3e} on {idx.size} points")
        assert err <= 1e-5


def test_fourier_random_against_oracle():
    rng = np.random.default_rng(21)
    modes = rng.normal(0, 0.2, size=(2, 500))
    sf = rng.uniform(0.0, 1.5, size=500)
    z1, z2 = rng.normal(size=500), rng.normal(size=500)
    pos = rng.uniform(0, 100, size=(2, 30_000))
    gpu = gstools_core.summate_fourier(sf, modes, z1, z2, pos)
    ref = oracle.summate_fourier(sf, modes, z1, z2, pos)
    assert _err(gpu, ref) <= sigma_tol(500) * sf.max()


# -- edge cases and layouts -------------------------------------------------------

@pytest.mark.parametrize("d,m,n", [(1, 37, 1), (1, 200, 5000), (2, 1, 2049), (2, 33, 2047),
                                   (3, 64, 2048), (3, 1000, 2049), (4, 50, 777), (7, 20, 300),
                                   (16, 5, 64)])
def test_dimensions_and_ragged_sizes(d, m, n):
    rng = np.random.default_rng(100 + d * 7 + n)
    cov = rng.normal(0, 0.5, size=(d, m))
    z1, z2 = rng.normal(size=m), rng.normal(size=m)
    pos = rng.uniform(-40, 40, size=(d, n))
    gpu = gstools_core.summate(cov, z1, z2, pos)
    ref = oracle.summate(cov, z1, z2, pos)
    assert _err(gpu, ref) <= sigma_tol(m)


def test_zero_modes_gives_zero():
    rng = np.random.default_rng(22)
    r = gstools_core.summate(np.zeros((3, 0)), np.zeros(0), np.zeros(0),
                             rng.normal(size=(3, 100)))
    np.testing.assert_array_equal(r, np.zeros(100))
    r = gstools_core.summate_incompr(np.zeros((2, 0)), np.zeros(0), np.zeros(0),
                                     rng.normal(size=(2, 10)))
    np.testing.assert_array_equal(r, np.zeros((2, 10)))


def test_zero_wavevector_incompr_is_nan_like_reference():
    # k = 0 -> k_2 = 0/0 = NaN in the reference (field.rs:138)
    r = gstools_core.summate_incompr(np.zeros((2, 1)), np.ones(1), np.ones(1), np.ones((2, 3)))
    assert np.isnan(r).all()


def test_strided_and_reversed_inputs():
    rng = np.random.default_rng(23)
    cov = rng.normal(0, 0.3, size=(3, 300))
    z1, z2 = rng.normal(size=300), rng.normal(size=300)
    pos = rng.uniform(0, 60, size=(3, 5000))
    dense = gstools_core.summate(cov, z1, z2, pos)
    np.testing.assert_array_equal(gstools_core.summate(np.asfortranarray(cov), z1, z2,
                                                       np.asfortranarray(pos)), dense)
    rev = gstools_core.summate(cov[:, ::-1], z1[::-1], z2[::-1], pos[:, ::-1])
    assert _err(rev[::-1], dense) <= sigma_tol(300) * 0.1
    big = np.zeros((3, 10000))
    big[:, 1::2] = pos
    np.testing.assert_array_equal(gstools_core.summate(cov, z1, z2, big[:, 1::2]), dense)
    inc = gstools_core.summate_incompr(cov, z1, z2, pos)
    inc2 = gstools_core.summate_incompr(cov, z1, z2, np.asfortranarray(pos))
    np.testing.assert_array_equal(inc, inc2)


def test_float32_output_and_scale():
    w = workloads.c1(n=20_000)
    md = w.modes
    r64 = gstools_core.summate(md.cov_samples, md.z1, md.z2, w.pos)
    r32 = gstools_core.summate(md.cov_samples, md.z1, md.z2, w.pos, out_dtype=np.float32,
                               scale=md.scale)
    assert r32.dtype == np.float32
    np.testing.assert_allclose(r32, r64 * md.scale, atol=1e-6)


def test_reflection_symmetry_at_scale():
    # u(-x; z1, z2) = u(x; z1, -z2) exactly in the mathematics; on the GPU the
    # two sides share every rounding except the phase sign
    w = workloads.c2(n=1_000_000)
    md = w.modes
    a = gstools_core.summate(md.cov_samples, md.z1, md.z2, -w.pos)
    b = gstools_core.summate(md.cov_samples, md.z1, -md.z2, w.pos)
    assert _err(a, b) * md.scale <= 2e-6


def test_device_entry_point_matches_host_entry_point():
    torch = pytest.importorskip("torch")
    from gstools_core import device
    w = workloads.c1(n=50_000)
    md = w.modes
    host = gstools_core.summate(md.cov_samples, md.z1, md.z2, w.pos)
    dev = torch.device("cuda:0")
    cov = torch.from_numpy(np.ascontiguousarray(md.cov_samples)).to(dev)
    z1 = torch.from_numpy(md.z1).to(dev)
    z2 = torch.from_numpy(md.z2).to(dev)
    pos = torch.from_numpy(np.ascontiguousarray(w.pos)).to(dev)
    out = torch.empty(w.n_points, dtype=torch.float64, device=dev)
    nb = device.workspace_bytes(2, md.m, w.n_points)
    work = torch.empty(nb, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    device.summate_dev(2, md.m, w.n_points, cov.data_ptr(), z1.data_ptr(), z2.data_ptr(),
                       pos.data_ptr(), w.n_points, out.data_ptr(), work.data_ptr(), nb,
                       stream=stream)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), host)
    # incompressible device form, (d, N) row-major output
    outi = torch.empty((2, w.n_points), dtype=torch.float64, device=dev)
    device.summate_incompr_dev(2, md.m, w.n_points, cov.data_ptr(), z1.data_ptr(), z2.data_ptr(),
                               pos.data_ptr(), w.n_points, outi.data_ptr(), w.n_points,
                               work.data_ptr(), nb, stream=stream)
    torch.cuda.synchronize()
    hosti = gstools_core.summate_incompr(md.cov_samples, md.z1, md.z2, w.pos)
    np.testing.assert_array_equal(outi.cpu().numpy(), hosti)


@pytest.mark.parametrize("variant", ["local", "valu"])
def test_device_entry_points_with_row_pitch(variant):
    # positions and incompressible output rows with pitches above N (views of
    # wider arrays): the same values as the dense host calls, bit for bit
    torch = pytest.importorskip("torch")
    from gstools_core import device
    w = workloads.c2(n=30_000)
    md = w.modes
    n, ld, old = w.n_points, w.n_points + 301, w.n_points + 77
    dev = torch.device("cuda:0")
    cov = torch.from_numpy(np.ascontiguousarray(md.cov_samples)).to(dev)
    z1 = torch.from_numpy(md.z1).to(dev)
    z2 = torch.from_numpy(md.z2).to(dev)
    wide = torch.full((3, ld), float("nan"), dtype=torch.float64, device=dev)
    wide[:, :n] = torch.from_numpy(np.ascontiguousarray(w.pos)).to(dev)
    nb = device.workspace_bytes(3, md.m, n, variant=variant)
    work = torch.empty(nb, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    out = torch.empty(n, dtype=torch.float64, device=dev)
    device.summate_dev(3, md.m, n, cov.data_ptr(), z1.data_ptr(), z2.data_ptr(),
                       wide.data_ptr(), ld, out.data_ptr(), work.data_ptr(), nb,
                       stream=stream, variant=variant)
    outi = torch.full((3, old), -7.0, dtype=torch.float64, device=dev)
    device.summate_incompr_dev(3, md.m, n, cov.data_ptr(), z1.data_ptr(), z2.data_ptr(),
                               wide.data_ptr(), ld, outi.data_ptr(), old, work.data_ptr(), nb,
                               stream=stream, variant=variant)
    torch.cuda.synchronize()
    host = gstools_core.summate(md.cov_samples, md.z1, md.z2, w.pos, variant=variant)
    hosti = gstools_core.summate_incompr(md.cov_samples, md.z1, md.z2, w.pos, variant=variant)
    np.testing.assert_array_equal(out.cpu().numpy(), host)
    oi = outi.cpu().numpy()
    np.testing.assert_array_equal(oi[:, :n], hosti)
    assert np.all(oi[:, n:] == -7.0)  # nothing written past N in a row


def test_committed_oracle_vectors():
    import os
    from conftest import GOLDEN
    with np.load(os.path.join(GOLDEN, "oracle_vectors.npz")) as f:
        v = {k: f[k] for k in f.files}
    for p in ("c1", "c2", "c5"):
        g = gstools_core.summate(v[f"{p}_cov"], v[f"{p}_z1"], v[f"{p}_z2"], v[f"{p}_pos"])
        m = v[f"{p}_cov"].shape[1]
        err = _err(g, v[f"{p}_out"])
        print(f"{p}: max|err|/sigma = {err / np.sqrt(m):.3e}")
        assert err <= sigma_tol(m)
    g = gstools_core.summate_incompr(v["c2_cov"], v["c2_z1"], v["c2_z2"], v["c2_pos"])
    assert _err(g, v["c2i_out"]) <= sigma_tol(1000)
    g = gstools_core.summate_fourier(v["f_sf"], v["f_modes"], v["f_z1"], v["f_z2"], v["f_pos"])
    assert _err(g, v["f_out"]) <= sigma_tol(400) * v["f_sf"].max()


@pytest.mark.parametrize("variant", ["mfma", "valu", "local"])
def test_direct_variants_against_oracle(variant):
    # every phase engine (matrix-core FP64 tiles, VALU FP64 FMA chains, cell-local
    # split-bf16 matrix-core phases), scalar, incompressible and Fourier
    for d, m, n, seed in ((1, 300, 3001, 1), (2, 1000, 20_000, 2), (3, 1000, 20_000, 3),
                          (3, 17, 4099, 4), (2, 1, 77, 5)):
        rng = np.random.default_rng(500 + seed)
        cov = rng.normal(0, 0.05, size=(d, m))
        z1, z2 = rng.normal(size=m), rng.normal(size=m)
        pos = rng.uniform(0, 400, size=(d, n))
        g = gstools_core.summate(cov, z1, z2, pos, variant=variant)
        ref = oracle.summate(cov, z1, z2, pos)
        err = _err(g, ref) / np.sqrt(m)
        print(f"{variant} d={d} m={m} n={n}: {err:.3e}")
        assert err <= 1e-5
        if d in (2, 3):
            gi = gstools_core.summate_incompr(cov, z1, z2, pos, variant=variant)
            refi = oracle.summate_incompr(cov, z1, z2, pos)
            assert _err(gi, refi) / np.sqrt(m) <= 1e-5
    sf = np.abs(rng.normal(size=m))
    g = gstools_core.summate_fourier(sf, cov, z1, z2, pos, variant=variant)
    assert _err(g, oracle.summate_fourier(sf, cov, z1, z2, pos)) <= sigma_tol(m) * sf.max()


def test_direct_variants_agree_at_scale():
    # every one of the 1e7 points of the C2 workload: the cell-local kernel
    # (fp16 split local phases) against the FP64-phase VALU kernel; the FP64
    # matrix-core tiles against the VALU kernel on the first 2e6
    w = workloads.c2()
    md = w.modes
    b = gstools_core.summate(md.cov_samples, md.z1, md.z2, w.pos, variant="valu")
    c = gstools_core.summate(md.cov_samples, md.z1, md.z2, w.pos, variant="local")
    err = _err(c, b) * md.scale
    print(f"local vs valu at all {w.n_points} points: {err:.3e} sigma")
    assert err <= 5e-6
    head = np.ascontiguousarray(w.pos[:, :2_000_000])
    a = gstools_core.summate(md.cov_samples, md.z1, md.z2, head, variant="mfma")
    assert _err(a, b[:2_000_000]) * md.scale <= 2e-6
    wi = workloads.c2(n=4_000_000)
    bi = gstools_core.summate_incompr(md.cov_samples, md.z1, md.z2, wi.pos, variant="valu")
    ci = gstools_core.summate_incompr(md.cov_samples, md.z1, md.z2, wi.pos, variant="local")
    err = _err(ci, bi) * md.scale
    print(f"incompr local vs valu at all 4e6 points: {err:.3e} sigma")
    assert err <= 5e-6


@pytest.mark.parametrize("num_threads", [1, 3, 16])
def test_num_threads_on_strided_host_inputs(num_threads):
    # num_threads (field.rs:48-51, lib.rs:41) sizes the host pool that gathers
    # strided inputs; every pool size gives the default call's result bit for bit
    w = workloads.c2(n=1_000_000)
    md = w.modes
    pos_f = np.asfortranarray(w.pos)          # (3, N) with point stride 3
    pos_r = w.pos[:, ::-1]                    # negative point stride
    base = gstools_core.summate(md.cov_samples, md.z1, md.z2, w.pos)
    g = gstools_core.summate(md.cov_samples, md.z1, md.z2, pos_f, num_threads=num_threads)
    np.testing.assert_array_equal(g, base)
    g = gstools_core.summate(md.cov_samples, md.z1, md.z2, pos_r, num_threads=num_threads)
    np.testing.assert_array_equal(g, base[::-1])
    gi = gstools_core.summate_incompr(md.cov_samples, md.z1, md.z2, pos_f,
                                      num_threads=num_threads)
    np.testing.assert_array_equal(gi, gstools_core.summate_incompr(md.cov_samples, md.z1, md.z2,
                                                                   w.pos))
    ref = oracle.summate(md.cov_samples, md.z1, md.z2, np.ascontiguousarray(w.pos[:, :20_000]))
    assert _err(g[::-1][:20_000], ref) * md.scale <= 1e-5


@pytest.mark.parametrize("scale", [1e-35, 1e30])
def test_local_extreme_coordinate_scales(scale):
    # the same phases with positions scaled by `scale` and wavevectors by
    # 1 / scale: the power-of-two scaling of the fp16 phase product takes
    # exponents near -110 and +100 (the FP32 point records hold offsets
    # down to about 1e-37); both against the FP64 VALU kernel
    rng = np.random.default_rng(91)
    m, n = 500, 50_000
    cov = rng.normal(0, 0.03, size=(3, m)) / scale
    z1, z2 = rng.normal(size=m), rng.normal(size=m)
    pos = rng.uniform(0, 400, size=(3, n)) * scale
    a = gstools_core.summate(cov, z1, z2, pos, variant="local")
    b = gstools_core.summate(cov, z1, z2, pos, variant="valu")
    assert np.all(np.isfinite(a))
    assert _err(a, b) / np.sqrt(m) <= 1e-5
    ai = gstools_core.summate_incompr(cov, z1, z2, pos, variant="local")
    bi = gstools_core.summate_incompr(cov, z1, z2, pos, variant="valu")
    assert np.all(np.isfinite(ai))
    assert _err(ai, bi) / np.sqrt(m) <= 1e-5


def test_local_without_cell_r0_table():
    # 16384 cells x (4000 + 1024) modes of r0 words (scalar) and x (4000 +
    # 64) modes of r0 parts (incompressible) exceed the 256 MB table cap, so
    # every wave evaluates its cell's r0 itself; 1000 modes keep the
    # precomputed tables.  Both against the FP64 VALU kernels.
    rng = np.random.default_rng(77)
    pos = rng.uniform(0, 500, size=(3, 1_100_000))
    for m in (4000, 1000):
        cov = rng.normal(0, 0.03, size=(3, m))
        z1, z2 = rng.normal(size=m), rng.normal(size=m)
        a = gstools_core.summate(cov, z1, z2, pos, variant="local")
        b = gstools_core.summate(cov, z1, z2, pos, variant="valu")
        err = _err(a, b) / np.sqrt(m)
        print(f"local vs valu, m={m}: {err:.3e} sigma")
        assert err <= 1e-5
        ai = gstools_core.summate_incompr(cov, z1, z2, pos, variant="local")
        bi = gstools_core.summate_incompr(cov, z1, z2, pos, variant="valu")
        assert np.all(np.isfinite(ai))
        err = _err(ai, bi) / np.sqrt(m)
        print(f"incompressible local vs valu, m={m}: {err:.3e} sigma")
        assert err <= 1e-5


# -- the cell-local (binned) kernel ---------------------------------------------

def _rand_modes(rng, d, m, s=0.05):
    return rng.normal(0, s, size=(d, m)), rng.normal(size=m), rng.normal(size=m)


@pytest.mark.parametrize("n", [1, 15, 16, 63, 64, 65, 1000, 4097])
def test_local_partial_waves(n):
    rng = np.random.default_rng(700 + n)
    for d in (2, 3):
        cov, z1, z2 = _rand_modes(rng, d, 300)
        pos = rng.uniform(-50, 250, size=(d, n))
        g = gstools_core.summate(cov, z1, z2, pos, variant="local")
        assert _err(g, oracle.summate(cov, z1, z2, pos)) <= sigma_tol(300)
        gi = gstools_core.summate_incompr(cov, z1, z2, pos, variant="local")
        assert _err(gi, oracle.summate_incompr(cov, z1, z2, pos)) <= sigma_tol(300)


def test_local_degenerate_point_sets():
    rng = np.random.default_rng(31)
    cov, z1, z2 = _rand_modes(rng, 3, 500, 0.2)
    cases = {
        "identical": np.full((3, 3000), 7.25),
        "line": np.stack([np.linspace(0, 100, 3000), np.full(3000, 3.0), np.full(3000, -2.0)]),
        "plane": np.stack([rng.uniform(0, 50, 3000), rng.uniform(0, 50, 3000), np.zeros(3000)]),
        "two clusters": np.concatenate([rng.normal(0, 1, size=(3, 1500)),
                                        rng.normal(1e4, 1, size=(3, 1500))], axis=1),
        "far offset": rng.uniform(1e6, 1e6 + 30, size=(3, 3000)),
    }
    for name, pos in cases.items():
        g = gstools_core.summate(cov, z1, z2, pos, variant="local")
        err = _err(g, oracle.summate(cov, z1, z2, pos)) / np.sqrt(500)
        print(f"{name}: {err:.3e}")
        assert err <= 1e-5, name


def test_local_planar_slice_keeps_modes_smooth():
    # a constant-z slice of a 3-D model with a short length scale: the
    # dimension without extent adds nothing to a mode's local-phase bound, so
    # the modes stay on the matrix-core path.  (With a unit cell width there,
    # every mode with |kappa_z| > 2 rev/unit, about half of them here, went to
    # the FP64 fallback and the summation kernel ran ~1.4x longer than on a
    # slab of tiny but non-zero thickness.)
    import ctypes
    import torch
    from gstools_core import _native, device
    rng = np.random.default_rng(17)
    n, m = 1_000_000, 2000
    xy = [rng.uniform(0, 20, n), rng.uniform(0, 20, n)]
    planar = np.stack(xy + [np.full(n, 3.5)])
    slab = np.stack(xy + [3.5 + rng.uniform(0, 1e-3, n)])
    cov = rng.normal(0, 1 / 0.05, size=(3, m))
    z1, z2 = rng.normal(size=m), rng.normal(size=m)
    idx = rng.choice(n, 5000, replace=False)
    ref = oracle.summate(cov, z1, z2, planar[:, idx])
    g = gstools_core.summate(cov, z1, z2, planar)
    assert _err(g[idx], ref) / np.sqrt(m) <= 1e-5
    gi = gstools_core.summate_incompr(cov, z1, z2, planar)
    refi = oracle.summate_incompr(cov, z1, z2, planar[:, idx])
    assert _err(gi[:, idx], refi) / np.sqrt(m) <= 1e-5
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    tc, t1, t2 = t(cov), t(z1), t(z2)
    out = torch.empty(n, dtype=torch.float64, device=dev)
    nb = device.workspace_bytes(3, m, n)
    work = torch.empty(nb, dtype=torch.uint8, device=dev)
    kernel_ms = {}
    for name, pos in (("planar", planar), ("slab", slab)):
        tp = t(pos)
        ph = (ctypes.c_double * 3)()
        _native.check(_native.lib().gsc_direct_phase_times(
            _native.GSC_FIELD_SCALAR, 3, m, n, None, tc.data_ptr(), t1.data_ptr(),
            t2.data_ptr(), tp.data_ptr(), n, out.data_ptr(), n, work.data_ptr(), nb,
            ctypes.byref(_native.make_opts()), None, 5, ph))
        kernel_ms[name] = ph[1]
    print(f"summation kernel: planar {kernel_ms['planar']:.3f} ms, "
          f"slab {kernel_ms['slab']:.3f} ms")
    assert kernel_ms["planar"] < 1.2 * kernel_ms["slab"]


def test_local_non_finite_points_only_affect_themselves():
    rng = np.random.default_rng(32)
    cov, z1, z2 = _rand_modes(rng, 2, 200)
    pos = rng.uniform(0, 100, size=(2, 5000))
    pos[0, 10] = np.nan
    pos[1, 20] = np.inf
    pos[0, 30] = -np.inf
    g = gstools_core.summate(cov, z1, z2, pos, variant="local")
    ref = oracle.summate(cov, z1, z2, pos)
    bad = np.array([10, 20, 30])
    assert np.isnan(g[bad]).all() and np.isnan(ref[bad]).all()
    ok = np.setdiff1d(np.arange(5000), bad)
    assert _err(g[ok], ref[ok]) <= sigma_tol(200)
    gi = gstools_core.summate_incompr(cov, z1, z2, pos, variant="local")
    refi = oracle.summate_incompr(cov, z1, z2, pos)
    assert np.isnan(gi[:, bad]).all() and np.isnan(refi[:, bad]).all()
    assert _err(gi[:, ok], refi[:, ok]) <= sigma_tol(200)


def test_local_is_bitwise_reproducible_and_order_free():
    # a point's result depends only on its cell (fixed by the bounding box) and
    # the mode order, never on the slot the counting sort gives it
    w = workloads.c2(n=300_000)
    md = w.modes
    a = gstools_core.summate(md.cov_samples, md.z1, md.z2, w.pos)
    b = gstools_core.summate(md.cov_samples, md.z1, md.z2, w.pos)
    np.testing.assert_array_equal(a, b)
    perm = np.random.default_rng(33).permutation(w.n_points)
    c = gstools_core.summate(md.cov_samples, md.z1, md.z2, w.pos[:, perm])
    np.testing.assert_array_equal(c, a[perm])
    ai = gstools_core.summate_incompr(md.cov_samples, md.z1, md.z2, w.pos)
    ci = gstools_core.summate_incompr(md.cov_samples, md.z1, md.z2, w.pos[:, perm])
    np.testing.assert_array_equal(ci, ai[:, perm])


def test_local_rough_modes_keep_fp64_phase():
    # a wide spectrum: most modes are too fast for the FP32 local phase and take
    # the FP64 path; a few slow ones stay on the matrix cores
    rng = np.random.default_rng(34)
    m = 400
    cov = rng.normal(0, 0.05, size=(3, m))
    cov[:, ::3] *= 400.0
    z1, z2 = rng.normal(size=m), rng.normal(size=m)
    pos = rng.uniform(0, 1000, size=(3, 20_000))
    g = gstools_core.summate(cov, z1, z2, pos, variant="local")
    err = _err(g, oracle.summate(cov, z1, z2, pos)) / np.sqrt(m)
    print(f"mixed rough/smooth: {err:.3e}")
    assert err <= 1e-5
    gi = gstools_core.summate_incompr(cov, z1, z2, pos, variant="local")
    err = _err(gi, oracle.summate_incompr(cov, z1, z2, pos)) / np.sqrt(m)
    print(f"mixed rough/smooth, incompressible: {err:.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("n", [3000, 1_000_000])
def test_incompr_output_layouts_through_the_c_abi(n):
    # the reference's Fortran-order (d, N) result is written by a device
    # transpose and one flat copy; C order and a reversed view take the
    # strided scatter; small calls use one pageable copy, large ones the
    # pinned chunks.  All layouts hold the same values.
    import ctypes
    from gstools_core import _native
    w = workloads.c2(n=n)
    md = w.modes
    cov = np.ascontiguousarray(md.cov_samples)
    opts = _native.make_opts()
    lib = _native.lib()

    def call(out, sd, sp):
        rc = lib.gsc_summate_incompr(3, md.m, n, cov.ctypes.data, md.m, 1, md.z1.ctypes.data, 1,
                                     md.z2.ctypes.data, 1, w.pos.ctypes.data, n, 1,
                                     out.ctypes.data, sd, sp, ctypes.byref(opts))
        _native.check(rc)

    f = np.empty((3, n), order="F")
    call(f, 1, 3)
    c = np.empty((3, n))
    call(c, n, 1)
    np.testing.assert_array_equal(c, f)
    big = np.zeros((3, 2 * n))
    view = big[:, ::-2]  # negative point stride into a larger buffer
    rc = lib.gsc_summate_incompr(3, md.m, n, cov.ctypes.data, md.m, 1, md.z1.ctypes.data, 1,
                                 md.z2.ctypes.data, 1, w.pos.ctypes.data, n, 1,
                                 view.ctypes.data, 2 * n, -2, ctypes.byref(opts))
    _native.check(rc)
    np.testing.assert_array_equal(view, f)
    assert not big[:, 0::2].any()
    np.testing.assert_array_equal(gstools_core.summate_incompr(md.cov_samples, md.z1, md.z2,
                                                               w.pos), f)


def test_roofline_calibrations_are_consistent():
    # The sine issue rate and the local kernels' instruction-mix ceilings:
    # rates near the rated 19.66 T sines/s (any clock from 1 GHz), the mixes
    # below the bare sine rate, the three-component mix below the scalar one.
    import ctypes
    from gstools_core import _native
    lib = _native.lib()
    s, sf = ctypes.c_double(), ctypes.c_double()
    _native.check(lib.gsc_calibrate_sine(1, ctypes.byref(s), ctypes.byref(sf)))
    ms, mv = ctypes.c_double(), ctypes.c_double()
    _native.check(lib.gsc_calibrate_local_mix(1, ctypes.byref(ms), ctypes.byref(mv)))
    assert 8e12 < s.value < 2.0e13
    assert sf.value < s.value
    assert 2e12 < mv.value < ms.value < s.value
    assert lib.gsc_calibrate_local_mix(0, ctypes.byref(ms), ctypes.byref(mv)) != 0


@pytest.mark.parametrize("m", [255, 256])
def test_auto_variant_picks_by_mode_count(m):
    # the default variant runs the FP64-phase VALU kernel below 256 modes and
    # the local kernel from 256 up, bit for bit as when asked by name
    rng = np.random.default_rng(90 + m)
    for d in (2, 3):
        cov = rng.normal(0, 0.2, size=(d, m))
        z1, z2 = rng.normal(size=m), rng.normal(size=m)
        pos = rng.uniform(0, 50, size=(d, 5000))
        want = "local" if m >= 256 else "valu"
        np.testing.assert_array_equal(gstools_core.summate(cov, z1, z2, pos),
                                      gstools_core.summate(cov, z1, z2, pos, variant=want))
        np.testing.assert_array_equal(gstools_core.summate_incompr(cov, z1, z2, pos),
                                      gstools_core.summate_incompr(cov, z1, z2, pos, variant=want))
