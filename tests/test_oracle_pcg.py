"""Pins the oracle's generic PCG (octree-pipelined-pcg.ts:305-594) on the
reference's own known-answer cases (tests/octree-pipelined-pcg.test.ts),
with expected solutions recomputed independently by LAPACK."""
import json
import math
import os

import numpy as np
import pytest

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "pcg_kats.json")))


def test_spd5_matches_lapack(oracle):
    g = GOLD["spd5"]
    A = np.array(g["matrix"], float)
    b = np.array(g["rhs"], float)
    x, info = oracle.dense_pcg(A, b, jacobi=True, max_iter=g["max_iterations"], rel_tol=g["relative_tolerance"])
    assert info["status"] == 0
    assert info["iterations"] <= 5
    assert info["reductions"] == max(1, 2 * info["iterations"])
    ref = np.linalg.solve(A, b)
    assert np.max(np.abs(x - ref)) < g["solution_tolerance"]


def test_poisson17_manufactured(oracle):
    g = GOLD["poisson17"]
    n = g["count"]
    fw = [0.75 + ((f * 13) % 11) / 9 for f in range(n + 1)]
    A = np.zeros((n, n))
    for r in range(n):
        A[r, r] = fw[r] + fw[r + 1] + g["diagonal_shift"]
        if r > 0:
            A[r, r - 1] = -fw[r]
        if r + 1 < n:
            A[r, r + 1] = -fw[r + 1]
    sol = np.array([math.sin((r + 1) * 0.37) + 0.2 * math.cos((r + 1) * 0.91) for r in range(n)])
    b = A @ sol
    x, info = oracle.dense_pcg(A, b, jacobi=True, max_iter=g["max_iterations"], rel_tol=g["relative_tolerance"])
    assert info["status"] == 0
    assert np.max(np.abs(x - sol)) < g["solution_tolerance"]
    assert info["reductions"] == 2 * info["iterations"]


def test_exact_initial_guess_zero_iterations(oracle):
    g = GOLD["exact_initial_guess"]
    A = np.array(g["matrix"], float)
    sol = np.array(g["solution"], float)
    x, info = oracle.dense_pcg(A, A @ sol, x0=sol, jacobi=False, rel_tol=g["relative_tolerance"])
    assert info["status"] == 0
    assert info["iterations"] == g["iterations"]
    assert info["reductions"] == g["reductions"]
    assert info["residual_norm"] == 0.0


def test_indefinite_operator_rejected(oracle):
    g = GOLD["indefinite"]
    x, info = oracle.dense_pcg(np.array(g["matrix"], float), np.array(g["rhs"], float), jacobi=False)
    assert info["status"] == 13  # operator-is-not-positive
    assert info["iterations"] == 0


def test_nonsymmetric_operator_rejected(oracle):
    A = np.array([[4.0, -1.0], [0.0, 3.0]])
    x, info = oracle.dense_pcg(A, np.array([1.0, 1.0]), jacobi=False)
    assert info["status"] == 12  # operator-is-not-symmetric


def test_exhausted_budget(oracle):
    g = GOLD["exhausted"]
    A = np.array(g["matrix"], float)
    x, info = oracle.dense_pcg(A, np.array(g["rhs"], float), jacobi=False, max_iter=g["max_iterations"],
                               rel_tol=g["relative_tolerance"])
    assert info["status"] == 5  # maximum-iterations
    assert info["iterations"] == g["iterations"]
    assert info["reductions"] == g["reductions"]
    assert info["residual_norm"] > g["relative_tolerance"] * np.linalg.norm(g["rhs"])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_spd_against_lapack(oracle, seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((30, 30))
    A = M @ M.T + 30 * np.eye(30)
    b = rng.standard_normal(30)
    x, info = oracle.dense_pcg(A, b, jacobi=True, max_iter=200, rel_tol=1e-13)
    assert info["status"] == 0
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=0, atol=1e-10)


def test_dots_are_compensated(oracle):
    # 1e16 + 1 - 1e16 loses the 1 in naive summation; both compensated dots keep it
    a = np.array([1e16, 1.0, -1e16, 1.0])
    b = np.ones(4)
    lib = oracle.lib()
    assert lib.orc_dot_neumaier(a, b, 4) == 2.0
    assert lib.orc_dot_kahan(np.array([1.0, 1e-16, 1e-16, 1e-16]), np.ones(4), 4) > 1.0
