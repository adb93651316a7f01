"""CPU tier: the kriging oracle (oracle/krige_oracle.c) against the reference's
golden values (src/krige.rs:134-243) and the closed forms of the calculators."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN


@pytest.fixture(scope="module")
def krige_rs():
    with open(os.path.join(GOLDEN, "krige_rs_fixture.json")) as f:
        d = json.load(f)
    return {k: (np.array(v, dtype=np.float64) if k != "source" else v) for k, v in d.items()}


def ulps(a, b):
    return np.abs(a - b) / np.spacing(np.abs(b))


def test_golden_values(krige_rs):
    f, e = oracle.calc_field_krige_and_variance(krige_rs["krig_mat"], krige_rs["krig_vecs"],
                                                krige_rs["cond"])
    # the reference's own rule is 6 ulps; the sequential order reproduces them exactly
    assert np.max(ulps(f, krige_rs["field"])) == 0
    assert np.max(ulps(e, krige_rs["error"])) == 0
    f2 = oracle.calc_field_krige(krige_rs["krig_mat"], krige_rs["krig_vecs"], krige_rs["cond"])
    np.testing.assert_array_equal(f2, f)


def test_closed_forms():
    rng = np.random.default_rng(5)
    n, N = 7, 30
    v = rng.normal(size=(n, N))
    c = rng.normal(size=n)
    # identity: field = cond . v, error = |v|^2
    f, e = oracle.calc_field_krige_and_variance(np.eye(n), v, c)
    np.testing.assert_allclose(f, c @ v, rtol=1e-14)
    np.testing.assert_allclose(e, np.sum(v * v, axis=0), rtol=1e-14)
    # general: field = (M c)^T v, error = diag(v^T M v); a transposed matrix
    # (the calculators contract with the columns of M) changes both
    M = rng.normal(size=(n, n))
    f, e = oracle.calc_field_krige_and_variance(M, v, c)
    np.testing.assert_allclose(f, (M @ c) @ v, rtol=1e-12)
    np.testing.assert_allclose(e, np.einsum("kj,ki,ij->j", v, M, v), rtol=1e-12)
    ft, _ = oracle.calc_field_krige_and_variance(np.ascontiguousarray(M.T), v, c)
    assert not np.allclose(ft, f)


def test_strided_and_empty():
    rng = np.random.default_rng(6)
    M = rng.normal(size=(9, 9))
    v = rng.normal(size=(9, 40))
    c = rng.normal(size=9)
    f, e = oracle.calc_field_krige_and_variance(M[::-1, ::-1].copy(), v[::-1].copy(), c[::-1].copy())
    f2, e2 = oracle.calc_field_krige_and_variance(M[::-1, ::-1], v[::-1], c[::-1])
    np.testing.assert_array_equal(f, f2)
    np.testing.assert_array_equal(e, e2)
    f0, e0 = oracle.calc_field_krige_and_variance(np.zeros((0, 0)), np.zeros((0, 5)), np.zeros(0))
    np.testing.assert_array_equal(f0, np.zeros(5))
    np.testing.assert_array_equal(e0, np.zeros(5))
