"""Pins the oracle's LGL basis against exact sympy fixtures and SBP identities."""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def load(N):
    with open(os.path.join(GOLDEN, f"basis_N{N}.json")) as fh:
        g = json.load(fh)
    return {k: (np.array(v, dtype=np.float64) if k != "N" else v) for k, v in g.items()}


@pytest.mark.parametrize("N", [2, 3, 4, 5, 6])
def test_basis_matches_exact_fixture(oracle, N):
    g = load(N)
    b = oracle.basis(N)
    for key in ("nodes", "weights", "lhat_minus", "lhat_plus"):
        np.testing.assert_allclose(b[key], g[key], rtol=0, atol=2e-15 * max(1, np.abs(g[key]).max()), err_msg=key)
    for key in ("D", "Dhat", "Vinv"):
        scale = np.abs(g[key]).max()
        np.testing.assert_allclose(b[key], g[key], rtol=0, atol=1e-14 * scale, err_msg=key)
    # D_split: the corner entries are 2 D_11 + 1/w_1 = 0 exactly in exact
    # arithmetic, round-off in floating point.
    np.testing.assert_allclose(b["Dsplit"], g["Dsplit"], rtol=0, atol=2e-14 * np.abs(g["Dsplit"]).max())


@pytest.mark.parametrize("N", [2, 3, 4, 5, 6])
def test_sbp_identities(oracle, N):
    b = oracle.basis(N)
    x, w, D = b["nodes"], b["weights"], b["D"]
    # exact differentiation of polynomials up to degree N-1
    for k in range(N):
        np.testing.assert_allclose(D @ x**k, k * x ** max(k - 1, 0) if k else 0 * x, atol=1e-12)
    # summation by parts: M D + D^T M = B
    M = np.diag(w)
    B = np.zeros((N, N))
    B[0, 0], B[-1, -1] = -1, 1
    np.testing.assert_allclose(M @ D + D.T @ M, B, atol=1e-13)
    # quadrature exact to degree 2N-3
    for k in range(2 * N - 2):
        exact = (1 - (-1) ** (k + 1)) / (k + 1)
        assert abs(w @ x**k - exact) < 1e-14
    # lhat(+-1) multiplied by w is the unit vector at the end nodes
    np.testing.assert_allclose(b["lhat_minus"] * w, np.eye(N)[0], atol=1e-15)
    np.testing.assert_allclose(b["lhat_plus"] * w, np.eye(N)[-1], atol=1e-15)


def test_interpolation_matrix_is_exact_for_polynomials(oracle):
    b4 = oracle.basis(4)
    b7 = oracle.basis(7)
    V = oracle.interpolation_matrix(b4["nodes"], b7["nodes"])
    for k in range(4):
        np.testing.assert_allclose(V @ b4["nodes"] ** k, b7["nodes"] ** k, atol=1e-14)
