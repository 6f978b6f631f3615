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


def dot(A: torch.Tensor, B: torch.Tensor) -> torch.Tensor:
    return torch.matmul(A.conj(), B.transpose(-2, -1))


def rotate(M: torch.Tensor, C: torch.Tensor) -> torch.Tensor:
    if M.dtype != C.dtype:
        M = M.to(C.dtype)
    return torch.matmul(M.transpose(-2, -1), C)
