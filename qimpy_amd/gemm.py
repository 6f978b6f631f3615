"""Band-space GEMMs of the subspace algebra.

Two shapes cover the path (reference: _wavefunction_dot.py):
- ``dot(A, B)``: (..., M, N) = conj(A)(..., M, G) @ B(..., N, G)^T — overlaps
  C^H C, C^H HC and projections beta^H C (_dot, :66-84). Long basis axis,
  small output: ``dot128`` runs it on the library's split-K MFMA kernel
  (qmx_dot) with an FP64 result;
- ``rotate(M, C)``: (..., Nout, G) = M(..., Nin, Nout)^T @ C(..., Nin, G) —
  subspace rotations C V and the nonlocal beta (D P) (_matmul, :168-185), a
  bandwidth-bound batched library GEMM in the precision of C.
Both operands keep the basis index G contiguous.
"""
from __future__ import annotations

import torch

from . import _lib, rc

_DTYPE = {torch.complex64: 0, torch.complex128: 1}


def dot128(A: torch.Tensor, B: torch.Tensor, out: torch.Tensor | None = None,
           accumulate: bool = False) -> torch.Tensor:
    """conj(A) @ B^T over the last axis, batched over the leading axes, as
    complex128 (qmx_dot). A: (..., M, G), B: (..., N, G), both with unit
    stride along G and one uniform stride per batch (any leading shape that
    flattens to a single batch axis)."""
    if A.dtype != B.dtype or A.dtype not in _DTYPE:
        raise TypeError(f"dot128: complex64 or complex128 operands of one type, not {A.dtype}/{B.dtype}")
    batch = torch.broadcast_shapes(A.shape[:-2], B.shape[:-2])
    M, G = A.shape[-2:]
    N = B.shape[-2]
    if B.shape[-1] != G:
        raise ValueError("dot128: basis lengths differ")
    A, B = _resolved(A), _resolved(B)  # the kernel reads raw memory
    A3 = _as_batched(A.expand(batch + A.shape[-2:]))
    B3 = _as_batched(B.expand(batch + B.shape[-2:]))
    nb = A3.shape[0]
    if out is None:
        out = torch.empty(batch + (M, N), dtype=torch.complex128, device=A.device)
        accumulate = False
    o3 = out.reshape(nb, M, N)
    if o3.data_ptr() != out.data_ptr() or o3.stride(-1) != 1:
        raise ValueError("dot128: out must be a contiguous (..., M, N) complex128 tensor")
    code = _DTYPE[A.dtype]
    need = int(_lib.lib().qmx_dot_workspace_bytes(code, nb, M, N, G))
    ws = torch.empty(max(need, 1), dtype=torch.uint8, device=A.device)
    _lib.call("qmx_dot", code, A3.data_ptr(), A3.stride(1), A3.stride(0), B3.data_ptr(),
              B3.stride(1), B3.stride(0), nb, M, N, G, o3.data_ptr(), o3.stride(1), o3.stride(0),
              int(accumulate), ws.data_ptr(), need, rc.stream_ptr())
    return out


def _resolved(X: torch.Tensor) -> torch.Tensor:
    """X with torch's lazy conjugate and negative bits applied (a copy only
    when one is set), so its memory holds the values it represents."""
    if X.is_conj():
        X = X.resolve_conj()
    if X.is_neg():
        X = X.resolve_neg()
    return X


def _as_batched(X: torch.Tensor) -> torch.Tensor:
    """(..., M, G) -> (batch, M, G) view with unit G stride, copying only
    when the leading axes do not fold into one uniformly strided axis."""
    if X.stride(-1) != 1:
        X = X.contiguous()
    M, G = X.shape[-2:]
    lead = X.shape[:-2]
    if len(lead) == 0:
        return X.reshape(1, M, G) if X.is_contiguous() else X.unsqueeze(0)
    try:
        return X.view(-1, M, G)
    except RuntimeError:
        return X.contiguous().view(-1, M, G)


def dot(A: torch.Tensor, B: torch.Tensor) -> torch.Tensor:
    """conj(A) @ B^T in the precision of the operands (the reference's
    _dot returns the wavefunction dtype), computed by dot128."""
    return dot128(A, B).to(A.dtype)


def rotate(M: torch.Tensor, C: torch.Tensor) -> torch.Tensor:
    if M.dtype != C.dtype:
        M = M.to(C.dtype)
    return torch.matmul(M.transpose(-2, -1), C)
