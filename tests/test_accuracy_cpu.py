"""The per-band eigenvalue error estimate (qimpy_amd.accuracy) against the
exact eigenvalues of dense Hermitian matrices: Rayleigh-Ritz on perturbed
eigenvector subspaces, with isolated, near-degenerate (1e-5 Ha) and
degenerate eigenvalues; the estimate must bound the true error of every
wanted Ritz value. A plausible mistake (the residual instead of its square,
the gap to the own Ritz value, no cluster merge) fails these cases."""
import numpy as np
import pytest

from qimpy_amd.accuracy import eig_error_estimate


def _ritz(H, lam, V, m, eps, rng):
    n = H.shape[0]
    X = V[:, :m] + eps * (rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m)))
    Q, _ = np.linalg.qr(X)
    th, W = np.linalg.eigh(Q.conj().T @ H @ Q)
    Y = Q @ W
    return th, np.linalg.norm(H @ Y - Y * th, axis=0)


@pytest.mark.parametrize("kind", ["isolated", "near_degenerate", "degenerate"])
def test_estimate_bounds_error(kind):
    rng = np.random.default_rng({"isolated": 1, "near_degenerate": 2, "degenerate": 3}[kind])
    worst, tight = 0.0, []
    for _ in range(60):
        n, m, want = 80, 14, 10
        A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        lam, V = np.linalg.eigh(0.5 * (A + A.conj().T))
        if kind == "near_degenerate":
            lam[6] = lam[5] + 1e-5
        elif kind == "degenerate":
            lam[3] = lam[2] = lam[4]
        H = (V * lam) @ V.conj().T
        th, r = _ritz(H, lam, V, m, 10 ** rng.uniform(-7, -3), rng)
        est = eig_error_estimate(th, r, want)
        err = np.abs(th[:want] - lam[:want])
        worst = max(worst, float((err / est).max()))
        tight.append(float(np.median(est / np.maximum(r[:want], 1e-300))))
    assert worst <= 1.0
    assert np.median(tight) < 0.1  # far tighter than the residual bound |theta - lambda| <= r


def test_no_upper_neighbour_uses_residual_bound():
    th = np.array([0.1, 0.5])
    r = np.array([1e-4, 2e-4])
    est = eig_error_estimate(th, r, 2)
    assert est[1] == pytest.approx(2e-4)  # last Ritz value: only |theta - lambda| <= r
    assert est[0] == pytest.approx(1e-8 / (0.4 - 2e-4))


def test_cluster_uses_block_residual_and_outer_gap():
    th = np.array([0.0, 1.0, 1.0 + 1e-4, 2.0])
    r = np.array([1e-6, 3e-5, 4e-5, 1e-6])
    est = eig_error_estimate(th, r, 3)
    rc2 = 3e-5**2 + 4e-5**2
    gap = min(1.0 - 1e-6, 2.0 - (1.0 + 1e-4) - 1e-6)
    assert est[1] == pytest.approx(rc2 / gap) and est[2] == pytest.approx(rc2 / gap)
