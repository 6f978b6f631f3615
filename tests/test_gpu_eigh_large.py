"""GPU parity of the grid-wide FP64 eigensolver (qmx_heevx_large), which
solves the subspace problems above n = 256: eigh(U^H H U) inside eighg
(math/_linalg.py:110-183) for the expanded Davidson subspace of the 64-atom
cell (2 x 282 = 564) and the CheFSI Rayleigh-Ritz and Davidson start of the
256-atom cell (1024, 2048).

Against LAPACK (numpy.linalg.eigvalsh, FP64): every eigenvalue to 1e-12 of
the spectral radius; the wanted eigenvectors with residual |HV - VE| below
1e-11 |H| and orthonormality to 1e-11. Cases: random spectra with rows in
LDS (n <= 1280) and in the work area (1300, 2048), all or the lower half of
the eigenvectors, exactly degenerate clusters, rows decoupled at 1e6 (the
dropped directions of a canonical start), several launch shapes of the
reduction, and a non-finite input (info = 1, NaN results, the grid drains).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

C128 = torch.complex128


def _hermitian(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return 0.5 * (A + A.conj().T)


def _unitary(n, rng):
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    return Q


def _check(H, E, V, w, tol_e=1e-12, tol_r=1e-11, scale=None):
    ref = np.linalg.eigvalsh(H)
    if scale is None:
        scale = max(np.abs(ref).max(), 1e-300)
    E = E.cpu().numpy()
    V = V.cpu().numpy()[:, :w]
    err_e = np.abs(E - ref).max() / scale
    res = np.abs(H @ V - V * E[None, :w]).max() / scale
    orth = np.abs(V.conj().T @ V - np.eye(w)).max()
    print(f"    eig {err_e:.1e}  residual {res:.1e}  orthonormality {orth:.1e}", flush=True)
    assert err_e <= tol_e
    assert res <= tol_r
    assert orth < 1e-11
    return V


@pytest.mark.parametrize("n,half", [(257, False), (300, True), (564, True), (564, False),
                                    (1024, False), (1300, True), (2048, True)])
def test_random(gpu, n, half):
    from qimpy_amd.linalg import eigh_large
    w = n // 2 if half else n
    print(f"[eigh_large] random n={n} n_want={w}", flush=True)
    H = _hermitian(n, n)
    E, V, info = eigh_large(torch.tensor(H, device="cuda"), w)
    assert int(info.item()) == 0
    assert V.shape == (n, w)
    _check(H, E, V, w)


@pytest.mark.parametrize("mult", [2, 8, 32])
def test_degenerate(gpu, mult):
    from qimpy_amd.linalg import eigh_large
    n = 564
    rng = np.random.default_rng(mult)
    ev = np.repeat(rng.standard_normal((n + mult - 1) // mult), mult)[:n]
    Q = _unitary(n, rng)
    H = (Q * ev) @ Q.conj().T
    H = 0.5 * (H + H.conj().T)
    print(f"[eigh_large] degenerate n={n} multiplicity {mult}", flush=True)
    E, V, info = eigh_large(torch.tensor(H, device="cuda"), n // 2)
    assert int(info.item()) == 0
    _check(H, E, V, n // 2)


def test_decoupled_shift(gpu):
    """Rows decoupled at 1e6 above a coupled block with close pairs (gaps 1e-7
    and 1e-4) and one exactly degenerate pair: the lower half of the
    spectrum to the norm of the coupled block, no weight on the decoupled
    rows."""
    from qimpy_amd.linalg import eigh_large
    rng = np.random.default_rng(5)
    nb, n = 500, 564
    ev = np.sort(rng.standard_normal(nb)) * 10.0
    ev[1], ev[11], ev[21] = ev[0], ev[10] + 1e-7, ev[20] + 1e-4
    Q = _unitary(nb, rng)
    H = np.zeros((n, n), complex)
    H[:nb, :nb] = (Q * ev) @ Q.conj().T
    H[np.arange(nb, n), np.arange(nb, n)] = 1e6
    w = 282
    print("[eigh_large] decoupled shift n=564", flush=True)
    E, V, info = eigh_large(torch.tensor(H, device="cuda"), w)
    assert int(info.item()) == 0
    ref = np.linalg.eigvalsh(H)
    scale = np.abs(ev).max()
    E, V = E.cpu().numpy(), V.cpu().numpy()
    assert np.abs(E[:w] - ref[:w]).max() <= 1e-9 * scale
    assert np.abs(H @ V - V * E[None, :w]).max() <= 1e-9 * scale
    assert np.abs(V.conj().T @ V - np.eye(w)).max() < 1e-11
    assert np.abs(V[nb:]).max() < 1e-12


@pytest.mark.parametrize("groups", [16, 64, 256])
def test_launch_shapes(gpu, groups):
    """The reduction's grid size changes only the rounding."""
    from qimpy_amd import linalg
    n = 700
    H = _hermitian(n, 17)
    print(f"[eigh_large] n={n} with {groups} workgroups", flush=True)
    old = linalg.LARGE_GROUPS
    linalg.LARGE_GROUPS = groups
    try:
        E, V, info = linalg.eigh_large(torch.tensor(H, device="cuda"))
    finally:
        linalg.LARGE_GROUPS = old
    assert int(info.item()) == 0
    _check(H, E, V, n)


def test_batch_and_dispatch(gpu):
    """A batch of 3 through linalg.eigh (one call per matrix), V as returned
    to the caller (a column-major view) usable in products."""
    from qimpy_amd.linalg import eigh
    n = 400
    Hs = np.stack([_hermitian(n, 100 + b) for b in range(3)])
    print(f"[eigh_large] batch of 3 at n={n}", flush=True)
    E, V = eigh(torch.tensor(Hs, device="cuda"))
    assert E.shape == (3, n) and V.shape == (3, n, n)
    for b in range(3):
        _check(Hs[b], E[b], V[b], n)
    R = torch.tensor(Hs, device="cuda") @ V - V * E[:, None, :]
    assert R.abs().max().item() < 1e-11 * np.abs(Hs).sum(-1).max()


def test_eighg_n_want(gpu):
    """eighg with n_want: the generalized eigenvectors of the n_want lowest
    eigenvalues satisfy H V = O V E and V^H O V = 1."""
    from qimpy_amd.linalg import eighg
    n, w = 600, 300
    rng = np.random.default_rng(3)
    H = _hermitian(n, 3)
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    O = X @ X.conj().T / n + np.eye(n)
    print(f"[eigh_large] eighg n={n} n_want={w}", flush=True)
    E, V = eighg(torch.tensor(H, device="cuda"), torch.tensor(O, device="cuda"), n_want=w)
    E, V = E.cpu().numpy(), V.cpu().numpy()[:, :w]
    Oi = np.linalg.cholesky(O)
    ref = np.linalg.eigvalsh(np.linalg.solve(Oi, np.linalg.solve(Oi, H).conj().T).conj().T)
    scale = np.abs(ref).max()
    assert np.abs(E - ref).max() < 1e-10 * scale
    assert np.abs(H @ V - (O @ V) * E[None, :w]).max() < 1e-9 * scale
    assert np.abs(V.conj().T @ O @ V - np.eye(w)).max() < 1e-10


def test_nonfinite(gpu):
    """A NaN gives info = 1 and NaN eigenvalues and vectors; the grid drains."""
    from qimpy_amd.linalg import eigh_large
    n = 300
    H = torch.eye(n, dtype=C128)
    H[7, 3] = float("nan")
    print("[eigh_large] non-finite input", flush=True)
    E, V, info = eigh_large(H.cuda(), 10)
    assert int(info.item()) == 1
    assert torch.isnan(E).all() and torch.isnan(V.real).all()
