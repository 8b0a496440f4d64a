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


# The block solver's generalized Rayleigh-Ritz step H y = E S y against
# LAPACK (scipy.linalg.eigh with b=S): S near the identity at the level an
# FP32-stored basis leaves (1e-7), a general SPD S, the identity (reduces to
# the symmetric case), and a singular S, which must be refused.
def _gcheck(H, S, tol=1e-11):
    import scipy.linalg
    import sbdx
    E, Y = sbdx.symmetric_geig(H, S)
    ref = scipy.linalg.eigh(H, S, eigvals_only=True)
    scale = max(1.0, np.max(np.abs(ref)))
    assert np.max(np.abs(E - ref)) < tol * scale
    assert np.max(np.abs(Y.T @ S @ Y - np.eye(len(E)))) < 1e-10
    assert np.max(np.abs(H @ Y - S @ Y * E)) < tol * scale * 10


@pytest.mark.parametrize("n", [1, 5, 24, 64])
def test_generalized_near_identity_overlap(n):
    rng = np.random.default_rng(10 + n)
    X = rng.standard_normal((n, n))
    H = -300.0 * np.eye(n) + X + X.T
    F = rng.standard_normal((n, n)) * 1e-7
    _gcheck(H, np.eye(n) + F + F.T)


def test_generalized_spd_overlap_and_identity():
    import sbdx
    rng = np.random.default_rng(5)
    n = 12
    X = rng.standard_normal((n, n))
    H = X + X.T
    B = rng.standard_normal((n, n))
    _gcheck(H, B @ B.T + n * np.eye(n))
    E, Y = sbdx.symmetric_geig(H, np.eye(n))
    E0, _ = sbdx.symmetric_eig(H)
    assert np.max(np.abs(E - E0)) < 1e-12


def test_generalized_singular_overlap_refused():
    import sbdx
    S = np.ones((3, 3))
    with pytest.raises(sbdx.SbdxError):
        sbdx.symmetric_geig(np.eye(3), S)
