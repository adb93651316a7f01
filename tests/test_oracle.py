"""Pins the CPU oracle (oracle/summator_oracle.c) before it is trusted as the checker.

Checks, in order of strength:
  1. the reference's own golden vectors (src/field.rs:359-429, 334-357), with
     the reference's own comparison rules (assert_eq! / assert_ulps_eq! 6 ulps);
  2. closed forms (M = 1, k = 0, x = 0) against numpy/math evaluated
     independently;
  3. invariants (linearity in z, x -> -x symmetry, mode permutation, the
     incompressible field's divergence and projector);
  4. brute force with mpmath at large phases (the heavy-tailed case), which
     the reference's tests do not reach;
  5. an independent numpy formulation (matrix product phases) at moderate size.
Each is chosen so that a plausible slip (swapped cos/sin, a dropped term, a
transposed operand, a wrong projector index) fails at least one of them.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle

EPS = np.finfo(np.float64).eps


def ulps_eq(a, b, max_ulps=6):
    """approx::ulps_eq! for f64 (approx 0.5): abs-diff <= EPSILON, else same sign and ulps."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    close = np.abs(a - b) <= EPS
    same_sign = np.signbit(a) == np.signbit(b)
    ulps = np.abs(a.view(np.int64) - b.view(np.int64))
    return np.all(close | (same_sign & (ulps <= max_ulps)))


# -- 1. reference golden vectors ------------------------------------------------

def test_golden_summate_bit_exact(field_rs):
    r = oracle.summate(field_rs["cov_samples"], field_rs["z1"], field_rs["z2"], field_rs["pos"])
    np.testing.assert_array_equal(r, field_rs["summate"])  # field.rs:363 assert_eq!


def test_golden_fourier_bit_exact(field_rs):
    r = oracle.summate_fourier(field_rs["spectrum_factor"], field_rs["cov_samples"],
                               field_rs["z1"], field_rs["z2"], field_rs["pos"])
    np.testing.assert_array_equal(r, field_rs["summate_fourier"])  # field.rs:337


def test_golden_incompr_reference_rule(field_rs):
    # variant 0 = the association written at field.rs:138-151
    r = oracle.summate_incompr(field_rs["cov_samples"], field_rs["z1"], field_rs["z2"],
                               field_rs["pos"], variant=0)
    assert r.shape == (3, 8) and r.flags.f_contiguous
    assert ulps_eq(r, field_rs["summate_incompr"], 6)  # field.rs:388 assert_ulps_eq!


def test_golden_incompr_projector_bit_exact(field_rs):
    r = oracle.summate_incompr(field_rs["cov_samples"], field_rs["z1"], field_rs["z2"],
                               field_rs["pos"], variant=1)
    np.testing.assert_array_equal(r, field_rs["summate_incompr"])


def test_golden_incompr_modechunk(field_rs):
    r = oracle.summate_incompr_modechunk(field_rs["cov_samples"], field_rs["z1"], field_rs["z2"],
                                         field_rs["pos"])
    assert ulps_eq(r, field_rs["summate_incompr"], 6)


def test_ulps_rule_rejects_real_mistakes(field_rs):
    # the comparison itself must be able to fail: 8 ulps on a value of magnitude
    # > 1 (below 1 the EPSILON shortcut of approx absorbs a few ulps)
    exp = field_rs["summate_incompr"].copy()
    bad = exp.copy()
    for _ in range(8):
        bad[0, 1] = np.nextafter(bad[0, 1], np.inf)
    assert abs(exp[0, 1]) > 1.0
    assert not ulps_eq(bad, exp, 6)


# -- 2. closed forms ------------------------------------------------------------

def test_single_mode_closed_form():
    rng = np.random.default_rng(3)
    k = rng.normal(size=(3, 1))
    z1, z2 = rng.normal(size=1), rng.normal(size=1)
    pos = rng.uniform(-20, 20, size=(3, 50))
    r = oracle.summate(k, z1, z2, pos)
    ph = [((0.0 + k[0, 0] * pos[0, j]) + k[1, 0] * pos[1, j]) + k[2, 0] * pos[2, j]
          for j in range(50)]
    exp = [z1[0] * math.cos(p) + z2[0] * math.sin(p) for p in ph]
    np.testing.assert_array_equal(r, exp)


def test_zero_wavevector_and_origin():
    rng = np.random.default_rng(4)
    z1, z2 = rng.normal(size=17), rng.normal(size=17)
    pos = rng.uniform(-5, 5, size=(2, 9))
    r = oracle.summate(np.zeros((2, 17)), z1, z2, pos)
    np.testing.assert_allclose(r, np.full(9, z1.sum()), rtol=1e-14)
    k = rng.normal(size=(2, 17))
    r0 = oracle.summate(k, z1, z2, np.zeros((2, 3)))
    np.testing.assert_allclose(r0, np.full(3, z1.sum()), rtol=1e-14)


def test_empty_inputs():
    rng = np.random.default_rng(5)
    pos = rng.uniform(size=(3, 4))
    np.testing.assert_array_equal(oracle.summate(np.zeros((3, 0)), np.zeros(0), np.zeros(0), pos),
                                  np.zeros(4))
    assert oracle.summate(rng.normal(size=(3, 5)), rng.normal(size=5), rng.normal(size=5),
                          np.zeros((3, 0))).shape == (0,)


# -- 3. invariants --------------------------------------------------------------

@pytest.fixture(scope="module")
def rand_case():
    rng = np.random.default_rng(11)
    k = rng.normal(0, 0.3, size=(3, 64))
    return k, rng.normal(size=64), rng.normal(size=64), rng.uniform(-30, 30, size=(3, 200))


def test_linearity_in_weights(rand_case):
    k, z1, z2, pos = rand_case
    rng = np.random.default_rng(12)
    w1, w2 = rng.normal(size=64), rng.normal(size=64)
    lhs = oracle.summate(k, 2.0 * z1 - 3.0 * w1, 2.0 * z2 - 3.0 * w2, pos)
    rhs = 2.0 * oracle.summate(k, z1, z2, pos) - 3.0 * oracle.summate(k, w1, w2, pos)
    np.testing.assert_allclose(lhs, rhs, atol=1e-11)
    # cos part and sin part separately: z2 = 0 gives an even function of x
    even = oracle.summate(k, z1, np.zeros(64), pos)
    np.testing.assert_allclose(even, oracle.summate(k, z1, np.zeros(64), -pos), atol=1e-12)


def test_reflection_symmetry(rand_case):
    k, z1, z2, pos = rand_case
    np.testing.assert_allclose(oracle.summate(k, z1, z2, -pos), oracle.summate(k, z1, -z2, pos),
                               atol=1e-12)


def test_mode_permutation(rand_case):
    k, z1, z2, pos = rand_case
    perm = np.random.default_rng(13).permutation(64)
    np.testing.assert_allclose(oracle.summate(k[:, perm], z1[perm], z2[perm], pos),
                               oracle.summate(k, z1, z2, pos), atol=1e-11)


def test_strided_views_match_dense(rand_case):
    k, z1, z2, pos = rand_case
    dense = oracle.summate(k, z1, z2, pos)
    pos_f = np.asfortranarray(pos)
    np.testing.assert_array_equal(oracle.summate(k, z1, z2, pos_f), dense)
    rev = oracle.summate(k[:, ::-1], z1[::-1], z2[::-1], pos[:, ::-1])
    np.testing.assert_allclose(rev[::-1], dense, atol=1e-11)
    big = np.zeros((3, 400))
    big[:, ::2] = pos
    np.testing.assert_array_equal(oracle.summate(k, z1, z2, big[:, ::2]), dense)


def test_incompr_projector_properties(rand_case):
    k, z1, z2, pos = rand_case
    # k parallel to e_0: projector first column vanishes -> component 0 is 0
    kx = np.zeros_like(k)
    kx[0] = k[0]
    u = oracle.summate_incompr(kx, z1, z2, pos)
    np.testing.assert_allclose(u[0], 0.0, atol=1e-12)
    np.testing.assert_allclose(u[1:], 0.0, atol=1e-12)
    # k perpendicular to e_0: projector = e_0 -> u_0 equals the scalar field
    kp = k.copy()
    kp[0] = 0.0
    u = oracle.summate_incompr(kp, z1, z2, pos)
    np.testing.assert_allclose(u[0], oracle.summate(kp, z1, z2, pos), atol=1e-12)
    np.testing.assert_allclose(u[1:], 0.0, atol=1e-12)
    # both associations agree to rounding
    np.testing.assert_allclose(oracle.summate_incompr(k, z1, z2, pos, variant=0),
                               oracle.summate_incompr(k, z1, z2, pos, variant=1), atol=1e-12)


def test_incompr_divergence_free(rand_case):
    k, z1, z2, _ = rand_case
    rng = np.random.default_rng(14)
    x0 = rng.uniform(-10, 10, size=(3, 20))
    h = 1e-4
    div = np.zeros(20)
    for i in range(3):
        e = np.zeros((3, 1))
        e[i] = h
        up = oracle.summate_incompr(k, z1, z2, x0 + e)[i]
        dn = oracle.summate_incompr(k, z1, z2, x0 - e)[i]
        div += (up - dn) / (2 * h)
    scale = np.abs(oracle.summate_incompr(k, z1, z2, x0)).max() * np.abs(k).max()
    assert np.abs(div).max() < 1e-6 * max(scale, 1.0)


def test_incompr_rejects_other_dims():
    with pytest.raises(ValueError):
        oracle.summate_incompr(np.ones((4, 2)), np.ones(2), np.ones(2), np.ones((4, 3)))
    with pytest.raises(ValueError):
        oracle.summate_incompr(np.ones((1, 2)), np.ones(2), np.ones(2), np.ones((1, 3)))


def test_modechunk_matches_point_parallel():
    rng = np.random.default_rng(15)
    k = rng.normal(0, 0.2, size=(3, 450))
    z1, z2 = rng.normal(size=450), rng.normal(size=450)
    pos = rng.uniform(0, 50, size=(3, 300))
    a = oracle.summate_incompr(k, z1, z2, pos)
    b = oracle.summate_incompr_modechunk(k, z1, z2, pos, min_len=100)
    np.testing.assert_allclose(a, b, atol=1e-11)


# -- 4. large phases: libm vs mpmath -------------------------------------------

def test_large_phase_against_mpmath():
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 40
    rng = np.random.default_rng(16)
    # |k . x| up to ~7e5 rad, as in the heavy-tailed 512^3 case
    k = np.array([[913.0], [-401.5], [77.25]])
    z1, z2 = np.array([0.7]), np.array([-1.3])
    pos = rng.uniform(0, 511, size=(3, 40))
    r = oracle.summate(k, z1, z2, pos)
    for j in range(40):
        ph = ((0.0 + k[0, 0] * pos[0, j]) + k[1, 0] * pos[1, j]) + k[2, 0] * pos[2, j]
        c = float(mpmath.cos(mpmath.mpf(ph)))
        s = float(mpmath.sin(mpmath.mpf(ph)))
        exp = z1[0] * c + z2[0] * s
        assert abs(r[j] - exp) <= 4 * EPS * (abs(z1[0]) + abs(z2[0]))


# -- 5. independent numpy formulation -------------------------------------------

def test_against_numpy_matrix_form():
    rng = np.random.default_rng(17)
    k = rng.normal(0, 0.1, size=(2, 300))
    z1, z2 = rng.normal(size=300), rng.normal(size=300)
    pos = rng.uniform(0, 100, size=(2, 500))
    ph = k.T @ pos  # (M, N)
    exp = (z1[:, None] * np.cos(ph) + z2[:, None] * np.sin(ph)).sum(axis=0)
    np.testing.assert_allclose(oracle.summate(k, z1, z2, pos), exp, atol=1e-10)
    kk = (k * k).sum(axis=0)
    p = np.stack([1.0 - k[0] * k[0] / kk, -k[1] * k[0] / kk])  # (2, M)
    z12 = z1[:, None] * np.cos(ph) + z2[:, None] * np.sin(ph)
    np.testing.assert_allclose(oracle.summate_incompr(k, z1, z2, pos), p @ z12, atol=1e-10)
    sf = rng.uniform(0.1, 2.0, size=300)
    np.testing.assert_allclose(oracle.summate_fourier(sf, k, z1, z2, pos), sf @ z12, atol=1e-10)


# -- 6. committed oracle vectors (tests/golden/make_vectors.py) -----------------

@pytest.fixture(scope="module")
def vectors():
    import os
    from conftest import GOLDEN
    with np.load(os.path.join(GOLDEN, "oracle_vectors.npz")) as f:
        return {k: f[k] for k in f.files}


def test_fixture_vectors_reproduce_bit_exact(vectors):
    v = vectors
    np.testing.assert_array_equal(oracle.summate(v["c1_cov"], v["c1_z1"], v["c1_z2"], v["c1_pos"]),
                                  v["c1_out"])
    np.testing.assert_array_equal(oracle.summate(v["c2_cov"], v["c2_z1"], v["c2_z2"], v["c2_pos"]),
                                  v["c2_out"])
    np.testing.assert_array_equal(
        oracle.summate_incompr(v["c2_cov"], v["c2_z1"], v["c2_z2"], v["c2_pos"]), v["c2i_out"])
    np.testing.assert_array_equal(oracle.summate(v["c5_cov"], v["c5_z1"], v["c5_z2"], v["c5_pos"]),
                                  v["c5_out"])
    np.testing.assert_array_equal(
        oracle.summate_fourier(v["f_sf"], v["f_modes"], v["f_z1"], v["f_z2"], v["f_pos"]),
        v["f_out"])


def test_fixture_inputs_match_workload_generators(vectors):
    from gstools_core import workloads
    w = workloads.c1(n=1500)
    np.testing.assert_array_equal(w.modes.cov_samples, vectors["c1_cov"])
    np.testing.assert_array_equal(w.pos, vectors["c1_pos"])
    w = workloads.c2(n=1000)
    np.testing.assert_array_equal(w.modes.z2, vectors["c2_z2"])
    np.testing.assert_array_equal(w.pos, vectors["c2_pos"])
