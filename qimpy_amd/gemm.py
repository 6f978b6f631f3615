"""Band-space GEMMs of the subspace algebra.

Two shapes cover the path (reference: _wavefunction_dot.py):
- ``dot(A, B)``: (..., M, N) = conj(A)(..., M, G) @ B(..., N, G)^T — overlaps
  C^H C, C^H HC and projections beta^H C (_dot, :66-84). Long basis axis,
  small output: ``dot128`` runs it on the library's split-K MFMA kernel
  (qmx_dot) with an FP64 result;
- ``rotate(M, C)``: (..., Nout, G) = M(..., Nin, Nout)^T @ C(..., Nin, G) —
  subspace rotations C V, the nonlocal beta (D P) and Gram-Schmidt updates
  (_matmul, :168-185; _hamiltonian.py:21-24) on the library's MFMA kernel
  (qmx_gemm): FP32 MFMA for complex64, FP64 MFMA for complex128, or BF16 x 3
  split precision for complex64 when enabled (``set_bf16x3``).
Both operands keep the basis index G contiguous.
"""
from __future__ import annotations

from typing import Optional

import torch

from . import _lib, rc

_DTYPE = {torch.complex64: 0, torch.complex128: 1}


#: Shapes launched while a tuning recorder is active (qimpy_amd.autotune).
recorder: Optional[dict] = None


def dot128(A: torch.Tensor, B: torch.Tensor, out: torch.Tensor | None = None,
           accumulate: bool = False) -> torch.Tensor:
    """conj(A) @ B^T over the last axis, batched over the leading axes, as
    complex128 (qmx_dot). A: (..., M, G), B: (..., N, G), both with unit
    stride along G and one uniform stride per batch (any leading shape that
    flattens to a single batch axis)."""
    if A.dtype != B.dtype or A.dtype not in _DTYPE:
        raise TypeError(f"dot128: complex64 or complex128 operands of one type, not {A.dtype}/{B.dtype}")
    batch = _bshape(A.shape[:-2], B.shape[:-2])
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
    if recorder is not None:
        recorder.setdefault(("dot", code, nb, M, N, G), 0)
        recorder[("dot", code, nb, M, N, G)] += 1
    need = int(_lib.lib().qmx_dot_workspace_bytes(code, nb, M, N, G))
    ws = torch.empty(max(need, 1), dtype=torch.uint8, device=A.device)
    _lib.call("qmx_dot", code, A3.data_ptr(), A3.stride(1), A3.stride(0), B3.data_ptr(),
              B3.stride(1), B3.stride(0), nb, M, N, G, o3.data_ptr(), o3.stride(1), o3.stride(0),
              int(accumulate), ws.data_ptr(), need, rc.stream_ptr())
    return out


def _bshape(a: torch.Size, b: torch.Size) -> torch.Size:
    """Broadcast of two leading shapes (the common equal case without
    torch.broadcast_shapes, which costs tens of microseconds per call)."""
    if a == b:
        return a
    return torch.broadcast_shapes(a, b)


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


#: Arithmetic of complex64 rotations: 0 FP32 MFMA, 1 BF16 x 3 split precision.
_variant_c64 = 0


def set_bf16x3(on: bool) -> None:
    """Run complex64 rotations as BF16 x 3 split-precision products (three
    BF16 MFMAs per real product, FP32 accumulation) instead of FP32 MFMA."""
    global _variant_c64
    _variant_c64 = 1 if on else 0


def bf16x3_enabled() -> bool:
    return bool(_variant_c64)


def _batch_view(X: torch.Tensor, batch: tuple) -> tuple[torch.Tensor, int]:
    """X broadcast to the leading dims `batch`, with the one batch stride
    that walks them in row-major order (0 for a single batch or an operand
    broadcast along every batch axis); a contiguous copy when the leading
    axes do not fold into one uniformly strided axis."""
    X = X.expand(batch + X.shape[-2:])
    axes = [i for i in range(len(batch)) if batch[i] != 1]
    if not axes:
        return X, 0
    step = X.stride(axes[-1])
    expect = step
    for i in reversed(axes):
        if X.stride(i) != expect:
            X = X.contiguous()
            return X, X.shape[-2] * X.shape[-1]
        expect *= batch[i]
    return X, step


def inplace_rows_max(dtype: torch.dtype) -> int:
    """Most output rows of an in-place rotation (out is C): one block of the
    rotation kernel's output rows."""
    if dtype == torch.complex128:
        return 80
    return 96 if _variant_c64 else 128


def rotate(M: torch.Tensor, C: torch.Tensor, out: Optional[torch.Tensor] = None,
           alpha: float = 1.0, beta: float = 0.0) -> torch.Tensor:
    """out = alpha M^T @ C + beta out over the last two axes, batched over the
    leading axes (broadcast), on qmx_gemm. M: (..., Nin, Nout) complex (any
    precision, any strides with one unit-stride axis); C: (..., Nin, G)
    complex64 or complex128 with unit stride along G; out: (..., Nout, G) of
    C's type with unit stride along G (allocated when None; beta is then 0).
    out must not overlap M; it may be C itself (a rotation in place, for
    Nout = Nin <= inplace_rows_max) and must not overlap C otherwise."""
    if C.dtype not in _DTYPE:
        raise TypeError(f"rotate: complex64 or complex128 operand, not {C.dtype}")
    if not M.is_complex():
        M = M.to(C.dtype)
    M, C = _resolved(M), _resolved(C)
    Nin, Nout = M.shape[-2:]
    if C.shape[-2] != Nin:
        raise ValueError(f"rotate: M {tuple(M.shape)} does not match C {tuple(C.shape)}")
    G = C.shape[-1]
    batch = _bshape(M.shape[:-2], C.shape[:-2])
    if out is None:
        out = torch.empty(batch + (Nout, G), dtype=C.dtype, device=C.device)
        beta = 0.0
    elif out.shape != batch + (Nout, G) or out.dtype != C.dtype:
        raise ValueError(f"rotate: out {tuple(out.shape)} {out.dtype} != {batch + (Nout, G)}")
    inplace = out is not None and out.data_ptr() == C.data_ptr()
    if inplace and (out.stride() != C.stride() or out.shape != C.shape or C.stride(-1) != 1
                    or Nout > inplace_rows_max(C.dtype)):
        raise ValueError("rotate: in place needs out = C (same view, unit stride along G, "
                         f"at most {inplace_rows_max(C.dtype)} rows)")
    if C.stride(-1) != 1:
        C = C.contiguous()
    if M.stride(-1) == 1:
        trans, ldm = 1, M.stride(-2)  # op(M)[i][k] = M[k][i], i contiguous
    elif M.stride(-2) == 1:
        trans, ldm = 0, M.stride(-1)
    else:
        M = M.contiguous()
        trans, ldm = 1, M.stride(-2)
    if out.stride(-1) != 1:
        raise ValueError("rotate: out needs unit stride along G")
    nb = 1
    for d in batch:
        nb *= int(d)
    Mb, sm = _batch_view(M, batch)
    Cb, sw = _batch_view(C, batch)
    ob, so = _batch_view(out, batch)
    if ob.data_ptr() != out.data_ptr() or (nb > 1 and so == 0):
        raise ValueError("rotate: out must have one uniform batch stride")
    if Mb is not M:  # a copy changed its strides
        trans, ldm = (1, Mb.stride(-2)) if Mb.stride(-1) == 1 else (0, Mb.stride(-1))
    variant = _variant_c64 if C.dtype == torch.complex64 else 0
    if recorder is not None:
        mcode = _DTYPE[Mb.dtype] if Mb.dtype in _DTYPE else 1
        k = ("rotate", _DTYPE[C.dtype], mcode, variant, nb, Nout, Nin, G)
        recorder[k] = recorder.get(k, 0) + 1
    _lib.call("qmx_gemm", _DTYPE[C.dtype], _DTYPE[Mb.dtype] if Mb.dtype in _DTYPE else 1, variant,
              trans, nb, Nout, Nin, G, float(alpha), Mb.data_ptr(), ldm, sm, Cb.data_ptr(),
              Cb.stride(-2), sw, float(beta), out.data_ptr(), out.stride(-2), so, rc.stream_ptr())
    return out


def matmul(A: torch.Tensor, B: torch.Tensor, out: Optional[torch.Tensor] = None,
           alpha: float = 1.0, beta: float = 0.0) -> torch.Tensor:
    """out = alpha A @ B + beta out for small batched complex matrices (the
    subspace products of a Davidson step: S H S, S W, Loewdin factors) on
    the rotation kernel (qmx_gemm): A (..., m, k) of any strides with one
    unit-stride axis, B (..., k, n) with unit stride along n, out (..., m, n)
    of B's type. A @ B is rotate's (A^T)^T @ B."""
    return rotate(A.transpose(-2, -1), B, out=out, alpha=alpha, beta=beta)

