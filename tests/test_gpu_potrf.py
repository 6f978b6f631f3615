"""GPU parity of the grid-wide Cholesky factor inverse (qmx_potrf_inv), which
gives the orthonormalising matrix U = L^-H of ortho_matrix and eighg
(math/_linalg.py:110-183) on the GPU.

Against LAPACK (numpy / torch on the CPU, FP64): X = L^-1 to 1e-11 relative
(scaled by the condition number of L), U^H O U = 1 to 1e-10; the first
non-positive pivot reported like torch.linalg.cholesky_ex (info = k + 1),
including a NaN input; sizes from 1 to 2048 (every panel width of the
kernel), a batch, and eighg_robust's fall-back to canonical
orthogonalisation on an indefinite overlap.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

C128 = torch.complex128


def _spd(n, seed, cond=1e3):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    ev = np.geomspace(1.0, cond, n)
    O = (Q * ev) @ Q.conj().T
    return 0.5 * (O + O.conj().T)


@pytest.mark.parametrize("n", [1, 2, 5, 13, 32, 100, 282, 564, 769, 1024, 1300, 2048])
def test_inverse_factor(gpu, n):
    from qimpy_amd.linalg import cholesky_inverse
    O = _spd(n, n)
    print(f"[potrf] n={n}", flush=True)
    X, info = cholesky_inverse(torch.tensor(O, device="cuda"))
    assert int(info.item()) == 0
    X = X.cpu().numpy()
    L = np.linalg.cholesky(O)
    Xr = np.linalg.inv(L)
    err = np.abs(X - Xr).max() / np.abs(Xr).max()
    U = X.conj().T
    orth = np.abs(U.conj().T @ O @ U - np.eye(n)).max()
    print(f"    |X - L^-1| {err:.1e}, |U^H O U - 1| {orth:.1e}", flush=True)
    assert np.abs(np.triu(X, 1)).max() == 0.0
    assert err < 1e-11 * max(1.0, np.sqrt(1e3))
    assert orth < 1e-10


def test_batch_and_ortho_matrix(gpu):
    from qimpy_amd.linalg import ortho_matrix
    n = 300
    Os = np.stack([_spd(n, 40 + b) for b in range(3)])
    print("[potrf] batch of 3 via ortho_matrix", flush=True)
    U = ortho_matrix(torch.tensor(Os, device="cuda")).cpu().numpy()
    for b in range(3):
        Xr = np.linalg.inv(np.linalg.cholesky(Os[b]))
        assert np.abs(U[b] - Xr.conj().T).max() < 1e-10 * np.abs(Xr).max()


@pytest.mark.parametrize("n,bad", [(40, 0), (40, 17), (600, 333), (1100, 1099)])
def test_not_positive_definite(gpu, n, bad):
    """A negative eigen-direction placed so that leading minor `bad` + 1
    is the first one that is not positive definite."""
    from qimpy_amd.linalg import cholesky_inverse
    O = _spd(n, 7)
    O[bad, :] = 0.0
    O[:, bad] = 0.0
    O[bad, bad] = -1.0
    ref = int(torch.linalg.cholesky_ex(torch.tensor(O)).info.item())
    print(f"[potrf] n={n} not positive definite at {bad} (reference info {ref})", flush=True)
    _, info = cholesky_inverse(torch.tensor(O, device="cuda"))
    assert int(info.item()) == ref == bad + 1


def test_nonfinite(gpu):
    from qimpy_amd.linalg import cholesky_inverse
    O = _spd(50, 3)
    O[20, 10] = np.nan
    print("[potrf] non-finite input", flush=True)
    _, info = cholesky_inverse(torch.tensor(O, device="cuda"))
    assert int(info.item()) > 0


def test_eighg_robust_fallback(gpu):
    """An indefinite overlap (as float32 rounding can leave one) takes the
    canonical path; a positive definite one the Cholesky path; both give
    H V = O V E on the retained span."""
    from qimpy_amd import linalg
    n = 400
    rng = np.random.default_rng(1)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    H = 0.5 * (A + A.conj().T)
    O = _spd(n, 9)
    before = linalg.n_canonical_fallbacks
    E, V = linalg.eighg_robust(torch.tensor(H, device="cuda"), torch.tensor(O, device="cuda"))
    assert linalg.n_canonical_fallbacks == before
    E, V = E.cpu().numpy(), V.cpu().numpy()
    assert np.abs(H @ V - (O @ V) * E[None, :]).max() < 1e-8 * np.abs(E).max()
    Ob = O.copy()
    Ob[5, 5] = -1e-3
    linalg.eighg_robust(torch.tensor(H, device="cuda"), torch.tensor(Ob, device="cuda"))
    assert linalg.n_canonical_fallbacks == before + 1
