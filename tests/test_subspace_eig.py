"""The Davidson solves' subspace eigensolver (host code, no GPU) against
LAPACK (numpy.linalg.eigh), closed forms and invariants: the 2x2 closed
form, a diagonal matrix (already converged), exact degeneracies, graded
and random symmetric matrices up to the 64 x 64 subspace limit."""
import numpy as np
import pytest


def _check(A, tol=1e-12):
    import sbdx
    E, V = sbdx.symmetric_eig(A)
    ref = np.linalg.eigvalsh(A)
    scale = max(1.0, np.max(np.abs(ref)))
    assert np.all(np.diff(E) >= -1e-14 * scale)
    assert np.max(np.abs(E - ref)) < tol * scale
    assert np.max(np.abs(V.T @ V - np.eye(len(E)))) < 1e-12
    assert np.max(np.abs(A @ V - V * E)) < tol * scale * 10


def test_two_by_two_closed_form():
    import sbdx
    a, b, c = 1.5, -0.25, -3.0
    E, V = sbdx.symmetric_eig(np.array([[a, b], [b, c]]))
    m, r = 0.5 * (a + c), np.hypot(0.5 * (a - c), b)
    assert abs(E[0] - (m - r)) < 1e-15 and abs(E[1] - (m + r)) < 1e-15


def test_diagonal_and_degenerate():
    _check(np.diag([3.0, -1.0, 2.0, -1.0, 0.0]))
    Q, _ = np.linalg.qr(np.random.default_rng(1).standard_normal((6, 6)))
    _check(Q @ np.diag([1.0, 1.0, 1.0, -2.0, -2.0, 5.0]) @ Q.T)


@pytest.mark.parametrize("n", [1, 3, 12, 24, 40, 64])
def test_random_symmetric(n):
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, n))
    _check(X + X.T)


def test_davidson_like_graded():
    # a subspace matrix of a converging Davidson solve: one dominant low
    # Ritz value, tiny couplings, values ~ -300
    rng = np.random.default_rng(3)
    n = 24
    A = np.diag(-306.0 + np.sort(rng.uniform(0, 5, n)))
    off = rng.standard_normal((n, n)) * np.logspace(-1, -12, n)[None, :]
    A = A + np.triu(off, 1) + np.triu(off, 1).T
    _check(A, tol=1e-13)
