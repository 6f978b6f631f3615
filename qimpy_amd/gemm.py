"""Band-space GEMMs of the subspace algebra.

Two shapes cover the path (reference: _wavefunction_dot.py):
- ``dot(A, B)``: (..., M, N) = conj(A)(..., M, G) @ B(..., N, G)^T — overlaps
  C^H C, C^H HC and projections beta^H C (_dot, :66-84);
- ``rotate(M, C)``: (..., Nout, G) = M(..., Nin, Nout)^T @ C(..., Nin, G) —
  subspace rotations C V and the nonlocal beta (D P) (_matmul, :168-185).
Both operands keep the basis index G contiguous, so every call is a plain
batched library GEMM on the device, in the precision of its inputs.
"""
from __future__ import annotations

import torch


#: Target length of each K-chunk of the split-K overlap GEMM.
SPLIT_K_CHUNK = 512


def dot(A: torch.Tensor, B: torch.Tensor) -> torch.Tensor:
    """conj(A) @ B^T with the long basis axis split into chunks that run as one
    batched GEMM and are then summed (a plain GEMM with a (<=1024)^2 output
    and K = 10^4..10^5 leaves most of the device idle)."""
    G = A.shape[-1]
    s = G // SPLIT_K_CHUNK
    if s < 2:
        return torch.matmul(A.conj(), B.transpose(-2, -1))
    Gs = (G // s) * s
    batch = torch.broadcast_shapes(A.shape[:-2], B.shape[:-2])
    A = A.expand(batch + A.shape[-2:])
    B = B.expand(batch + B.shape[-2:])
    a = A[..., :Gs].unflatten(-1, (s, Gs // s)).transpose(-3, -2)  # (..., s, M, Gs/s)
    b = B[..., :Gs].unflatten(-1, (s, Gs // s)).transpose(-3, -2)
    out = torch.matmul(a.conj(), b.transpose(-2, -1)).sum(dim=-3)
    if Gs < G:
        out += torch.matmul(A[..., Gs:].conj(), B[..., Gs:].transpose(-2, -1))
    return out


def rotate(M: torch.Tensor, C: torch.Tensor) -> torch.Tensor:
    if M.dtype != C.dtype:
        M = M.to(C.dtype)
    return torch.matmul(M.transpose(-2, -1), C)
