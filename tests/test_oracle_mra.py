"""Pins the oracle's filter bank and codec (mra/Filters.h:202-397,
encode_scale.cuh, encode_details.cuh, decode_scale_children.cuh) against the
mathematics: closed forms of every literal, orthogonality of the 12x12
analysis operator, vanishing moments of the details, perfect reconstruction,
and the reference's own known-answer children (unit_tests_mra.cu:7-66)
evaluated in 50-digit arithmetic (tests/golden/mra_known_answers.json)."""
import json
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mra_known_answers.json")
PERM = [0, 2, 1, 3]


@pytest.fixture(scope="module")
def ka():
    with open(GOLDEN) as f:
        return json.load(f)


def closed_filters(ka):
    return np.array([[[[float(v) for v in row] for row in k] for k in fam] for fam in ka["filters"]])


def analysis_matrix(F):
    """12x12 M with out = M @ children(flat c*3+m) / 2."""
    M = np.zeros((12, 12))
    for fam in range(4):
        for i in range(3):
            for k in range(4):
                for j in range(3):
                    M[3 * fam + i, 3 * PERM[k] + j] = F[fam, k, i, j]
    return M


def test_literals_match_closed_forms(ka):
    lit = oracle.filters(np.float64)
    cf = closed_filters(ka)
    # the literals are written with 16 digits after the point (some rounded,
    # some truncated): each lies within one unit of the 16th digit, plus the
    # rounding of both to binary64
    assert np.all(np.abs(lit - cf) <= 1e-16 + 2 * np.spacing(np.abs(cf))), np.max(np.abs(lit - cf))
    # zero entries are exact zeros
    assert np.array_equal(lit == 0.0, cf == 0.0)


def test_fp32_table_is_rounded_fp64_table():
    assert np.array_equal(oracle.filters(np.float32), oracle.filters(np.float64).astype(np.float32))


def test_analysis_operator_is_orthogonal(ka):
    M = analysis_matrix(closed_filters(ka))
    assert np.allclose(M @ M.T, np.eye(12), atol=1e-15, rtol=0)
    Ml = analysis_matrix(oracle.filters(np.float64))
    assert np.allclose(Ml @ Ml.T, np.eye(12), atol=4e-16 * 12, rtol=0)


def linear_children(a, b, c):
    """Children modes (orthonormal Legendre: mean, sqrt(3)*xi, sqrt(3)*eta) of
    u = a + b*sqrt(3)*xi + c*sqrt(3)*eta on the parent square [-1,1]^2."""
    ch = np.zeros((4, 3))
    for k in range(4):
        sx = 1 if k & 1 else -1
        sy = 1 if k & 2 else -1
        ch[k] = [a + sx * b * np.sqrt(3) / 2 + sy * c * np.sqrt(3) / 2, b / 2, c / 2]
    return ch


@pytest.mark.parametrize("abc", [(1.0, 0.0, 0.0), (0.3, -2.0, 0.7), (-5.0, 1.5, -3.25)])
def test_linear_fields_encode_exactly_with_zero_details(abc):
    out = oracle.encode_mw(linear_children(*abc))
    assert np.allclose(out[:3], abc, atol=1e-14, rtol=0)
    assert np.allclose(out[3:], 0.0, atol=1e-14, rtol=0)


def test_perfect_reconstruction_fp64_and_fp32():
    rng = np.random.default_rng(7)
    for _ in range(200):
        x = rng.normal(size=12) * 10 ** rng.uniform(-3, 3)
        y = oracle.decode_mw(oracle.encode_mw(x))
        assert np.allclose(y, x, atol=8 * np.spacing(np.max(np.abs(x))), rtol=0)
        xf = x.astype(np.float32)
        yf = oracle.decode_mw(oracle.encode_mw(xf, np.float32), np.float32)
        assert np.allclose(yf, xf, atol=16 * np.spacing(np.float32(np.max(np.abs(xf)))), rtol=0)
    for _ in range(50):
        c = rng.normal(size=4)
        assert np.allclose(oracle.decode_hw(oracle.encode_hw(c)), c, atol=8 * np.spacing(np.max(np.abs(c))), rtol=0)


def test_known_answer_children(ka):
    mw_in = np.array([[float(v) for v in r] for r in ka["mw_children"]])
    mw_exp = np.array([float(v) for v in ka["mw_encoded"]])
    got = oracle.encode_mw(mw_in)
    assert np.allclose(got, mw_exp, atol=8 * np.spacing(10.0), rtol=0), got - mw_exp
    hw_in = np.array([float(v) for v in ka["hw_children"]])
    hw_exp = np.array([float(v) for v in ka["hw_encoded"]])
    assert np.allclose(oracle.encode_hw(hw_in), hw_exp, atol=4e-16 * 4, rtol=0)
    # Haar: scale = mean * 2, alpha/beta/gamma are the y, x and xy differences
    assert np.allclose(hw_exp, [2.5, -1.0, -0.5, 0.0], atol=1e-15)


def test_mean_is_average_of_children():
    rng = np.random.default_rng(3)
    ch = rng.normal(size=(4, 3))
    assert abs(oracle.encode_mw(ch)[0] - ch[:, 0].mean()) < 1e-15


def test_product_filter_table_matches_oracle():
    import mwdg2
    for prec, dt in ((8, np.float64), (4, np.float32)):
        assert np.array_equal(mwdg2.filter_table(prec), oracle.filters(dt))
