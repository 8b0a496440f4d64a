"""The bound behind the two-sector solve's abandonment rule (CPU).

For a unit vector u = sum c_i v_i over the eigenvectors of a symmetric H,
with Rayleigh quotient theta and residual r = H u - theta u:
|r|^2 = sum |c_i|^2 (lambda_i - theta)^2 >= |c_1|^2 (theta - lambda_1)^2,
so theta - lambda_1 <= |r| / |c_1|.  The thick solver abandons a sector
once theta - E_other > 10 |r|; were the sector's lowest eigenvalue lambda_1
below E_other, that test could pass only with |c_1| < 1/10.  Both facts are
checked here on random symmetric matrices and random vectors, including
vectors built to have a small overlap with the lowest state."""
import numpy as np


def _ritz(H, u):
    u = u / np.linalg.norm(u)
    theta = float(u @ H @ u)
    r = H @ u - theta * u
    return theta, float(np.linalg.norm(r)), u


def test_rayleigh_residual_bound():
    rng = np.random.default_rng(0)
    for trial in range(200):
        n = int(rng.integers(2, 40))
        A = rng.standard_normal((n, n))
        H = (A + A.T) / 2
        lam, V = np.linalg.eigh(H)
        u = rng.standard_normal(n)
        if trial % 2:
            # small overlap with the lowest state
            u = u - (V[:, 0] @ u) * V[:, 0] + 1e-3 * rng.standard_normal() * V[:, 0]
        theta, rn, u = _ritz(H, u)
        c1 = abs(V[:, 0] @ u)
        assert theta - lam[0] <= rn / max(c1, 1e-300) * (1 + 1e-10) + 1e-12


def test_abandon_test_needs_small_overlap():
    """Whenever theta - E > 10 |r| holds although lambda_1 < E, the overlap
    with the lowest state is below 1/10."""
    rng = np.random.default_rng(1)
    fired = 0
    for trial in range(4000):
        n = int(rng.integers(3, 30))
        A = rng.standard_normal((n, n))
        H = (A + A.T) / 2
        lam, V = np.linalg.eigh(H)
        # the other sector's result E somewhere above lambda_1
        E = lam[0] + abs(rng.standard_normal()) * (lam[-1] - lam[0]) * 0.5
        # vectors near an excited state, so that the test can fire
        k = int(rng.integers(1, n))
        u = V[:, k] + 10 ** rng.uniform(-4, -0.5) * rng.standard_normal(n)
        theta, rn, u = _ritz(H, u)
        if theta - E > 10 * rn and lam[0] < E:
            fired += 1
            assert abs(V[:, 0] @ u) < 0.1
    assert fired > 50  # the premise is exercised, not vacuous
