"""Subspace linear algebra in FP64 on the GPU (reference: math/_linalg.py).

``eighg`` solves H V = O V E by Cholesky orthonormalisation
U = (L^-1)^H (ortho_matrix, _linalg.py:110-141) and eigh(U^H H U) with
V <- U V (_linalg.py:143-183). Whatever the precision of the wavefunctions,
these small dense problems always run in complex128 on the device.
"""
from __future__ import annotations

from typing import Protocol, TypeVar, Generic

import torch

from . import rc

T = TypeVar("T")


class Waitable(Protocol[T]):
    def wait(self) -> T: ...


class Waitless(Generic[T]):
    """A result that is available immediately (mpi/_waitable.py:20-25)."""

    def __init__(self, result: T):
        self.result = result

    def wait(self) -> T:
        return self.result


def abs_squared(x: torch.Tensor) -> torch.Tensor:
    if x.is_complex():
        return torch.view_as_real(x).square().sum(dim=-1)
    return x.square()


def dagger(x: torch.Tensor) -> torch.Tensor:
    return x.conj().transpose(-2, -1)


def _fp64(x: torch.Tensor) -> torch.Tensor:
    if x.is_complex():
        return x.to(torch.complex128)
    return x.to(torch.float64)


def ortho_matrix(O: torch.Tensor, use_cholesky: bool = True) -> torch.Tensor:
    assert O.shape[-2] == O.shape[-1]
    O = _fp64(O)
    if use_cholesky:
        identity = torch.eye(O.shape[-1], device=O.device, dtype=O.dtype).view(
            (1,) * (O.dim() - 2) + O.shape[-2:]
        )
        L = torch.linalg.cholesky(O)
        return dagger(torch.linalg.solve_triangular(L, identity.expand_as(L), upper=False))
    lbda, V = torch.linalg.eigh(O)
    return V @ ((1.0 / torch.sqrt(lbda))[..., None] * dagger(V))


def eighg(H, O: torch.Tensor, use_cholesky: bool = True):
    """E (ascending) and V with H V = O V E; H may be a Waitable."""
    rc.compute_stream_wait_current()
    if rc.compute_stream is not None:
        with torch.cuda.stream(rc.compute_stream):
            U = ortho_matrix(O, use_cholesky)
    else:
        U = ortho_matrix(O, use_cholesky)
    Hres = _fp64(H.wait() if hasattr(H, "wait") else H)
    rc.current_stream_wait_compute()
    E, V = torch.linalg.eigh(dagger(U) @ (Hres @ U))
    return E, U @ V


def eighg_canonical(H, O: torch.Tensor, rtol: float = 1e-4, shift: float = 1e6):
    """H V = O V E by canonical orthogonalisation, for overlaps that rounding
    has made indefinite or near-singular (subspaces assembled in float32).

    The FP64 eigenvectors W, lambda of O give U = W lambda^-1/2 on the
    directions with lambda > rtol * max(lambda); the others are dropped by
    zeroing their columns of U and shifting their diagonal of U^H H U by
    `shift` Ha so that they sort above every physical eigenvalue. Batch
    shapes stay fixed (no host synchronisation). In exact arithmetic the
    retained span and hence the Ritz values are those of Cholesky eighg.
    """
    O = _fp64(O)
    Hres = _fp64(H.wait() if hasattr(H, "wait") else H)
    lam, W = torch.linalg.eigh(O)
    keep = lam > rtol * lam[..., -1:]
    s = torch.where(keep, lam.clamp_min(1e-300).rsqrt(), torch.zeros_like(lam))
    U = W * s[..., None, :].to(W.dtype)
    Hr = dagger(U) @ (Hres @ U)
    Hr = Hr + torch.diag_embed(torch.where(keep, 0.0, shift).to(Hr.dtype))
    E, V = torch.linalg.eigh(Hr)
    return E, U @ V


#: Number of eighg_robust calls that needed the canonical path.
n_canonical_fallbacks = 0


def eighg_robust(H, O: torch.Tensor):
    """Cholesky eighg when the overlap factorises, canonical orthogonalisation
    (eighg_canonical) when float32 rounding has made it indefinite."""
    global n_canonical_fallbacks
    O = _fp64(O)
    L, info = torch.linalg.cholesky_ex(O)
    if bool((info != 0).any()):
        n_canonical_fallbacks += 1
        return eighg_canonical(H, O)
    identity = torch.eye(O.shape[-1], device=O.device, dtype=O.dtype).expand_as(L)
    U = dagger(torch.linalg.solve_triangular(L, identity, upper=False))
    Hres = _fp64(H.wait() if hasattr(H, "wait") else H)
    E, V = torch.linalg.eigh(dagger(U) @ (Hres @ U))
    return E, U @ V
