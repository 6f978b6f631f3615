"""Subspace linear algebra in FP64 on the GPU (reference: math/_linalg.py).

``eighg`` solves H V = O V E by Cholesky orthonormalisation
U = (L^-1)^H (ortho_matrix, _linalg.py:110-141) and eigh(U^H H U) with
V <- U V (_linalg.py:143-183). Whatever the precision of the wavefunctions,
these small dense problems always run in complex128 on the device.
"""
from __future__ import annotations

from typing import Generic, Optional, Protocol, Sequence, TypeVar

import numpy as np
import torch

from . import _lib, rc

T = TypeVar("T")


class Waitable(Protocol[T]):
    def wait(self) -> T: ...


class Waitless(Generic[T]):
    """A result that is available immediately (mpi/_waitable.py:20-25)."""

    def __init__(self, result: T):
        self.result = result

    def wait(self) -> T:
        return self.result


#: Size limits of the batched Jacobi (qmx_heevj_batched) and Householder + QL
#: (qmx_heev_small_batched) kernels.
JACOBI_MAX_N = 96
SMALL_QL_MAX_N = 256
#: Largest n of the grid-wide eigensolver (qmx_heevx_large), used above
#: SMALL_QL_MAX_N; larger matrices go to the library eigensolver.
LARGE_MAX_N = 2048
#: Workgroups of its reduction (0: the library's choice from n).
LARGE_GROUPS = 0
#: Largest n of the grid-wide Cholesky factor inverse (qmx_potrf_inv), and
#: its workgroups (0: the library's choice from n).
POTRF_MAX_N = 2048
POTRF_GROUPS = 0


_large_ws: dict = {}
_potrf_ws: dict = {}


#: Batched on-chip eigensolvers are latency-bound per launch but run all
#: matrices of a batch concurrently; the library one runs them in turn. Cost
#: models (ms, measured on MI355X) pick the cheapest for (n, batch).
def _cost_jacobi(n: int) -> float:
    return 0.8 * (n / 36.0) ** 2.4 if n <= JACOBI_MAX_N else float("inf")


def _cost_ql(n: int) -> float:
    return 1.1 * (n / 36.0) ** 1.95 if n <= SMALL_QL_MAX_N else float("inf")


def _cost_bisect(n: int) -> float:
    return 0.5 * (n / 36.0) ** 1.6 if n <= SMALL_QL_MAX_N else float("inf")


def _cost_library(n: int, n_mat: int) -> float:
    return n_mat * (0.35 + 0.02 * n)


def small_method(n: int, n_mat: int) -> str:
    """'bisect', 'jacobi', 'ql' or 'library' for a batch of n_mat matrices of
    size n."""
    costs = {"bisect": _cost_bisect(n), "jacobi": _cost_jacobi(n), "ql": _cost_ql(n),
             "library": _cost_library(n, n_mat)}
    return min(costs, key=costs.get)


def eigh_small(H: torch.Tensor, method: str = "ql") -> tuple[torch.Tensor, torch.Tensor]:
    """Batched on-chip eigensolve of (..., n, n) complex128 on the GPU:
    method "bisect" (Householder + bisection + inverse iteration, n <= 256),
    "ql" (Householder + QL, n <= 256) or "jacobi" (n <= 96)."""
    n = H.shape[-1]
    batch = H.shape[:-2]
    if method == "bisect":
        (res,), info = eigh_grouped([H])
        eigh_small.last_info = info
        return res
    Hf = _fp64(H).reshape((-1, n, n)).contiguous()
    B = Hf.shape[0]
    E = torch.empty((B, n), dtype=torch.float64, device=H.device)
    V = torch.empty((B, n, n), dtype=torch.complex128, device=H.device)
    if method == "jacobi":
        need = int(_lib.lib().qmx_heevj_workspace_bytes(n, B))
        ws = torch.empty(max(need, 1), dtype=torch.uint8, device=H.device)
        _lib.call("qmx_heevj_batched", Hf.data_ptr(), n, n, n * n, B, E.data_ptr(), V.data_ptr(),
                  n, n * n, ws.data_ptr(), need, 0, rc.stream_ptr())
    else:
        need = int(_lib.lib().qmx_heev_small_workspace_bytes(n, B))
        ws = torch.empty(max(need, 1), dtype=torch.uint8, device=H.device)
        info = torch.zeros(B, dtype=torch.int32, device=H.device)
        _lib.call("qmx_heev_small_batched", Hf.data_ptr(), n, n, n * n, B, E.data_ptr(),
                  V.data_ptr(), n, n * n, info.data_ptr(), ws.data_ptr(), need, rc.stream_ptr())
        # QL non-convergence is checked lazily by callers that synchronise anyway.
        eigh_small.last_info = info
    return E.reshape(batch + (n,)), V.reshape(batch + (n, n))


_HEEV_DESC = np.dtype([("A", np.uint64), ("E", np.uint64), ("V", np.uint64), ("info", np.uint64),
                       ("work", np.uint64), ("lda", np.int64), ("ldv", np.int64), ("n", np.int32),
                       ("n_want", np.int32)])


def eigh_grouped(mats: list, n_want: Optional[Sequence[int]] = None,
                 method: str = "bisect") -> tuple[list, torch.Tensor]:
    """Eigendecompositions of several batches of Hermitian complex128
    matrices (each (..., n_i, n_i), n_i <= SMALL_QL_MAX_N, on the GPU)
    without a host synchronisation. Returns [(E, V)] in the input order and
    an int32 tensor per matrix, nonzero where the eigensolver failed.

    method "bisect" (qmx_heevx_grouped): one launch for all matrices;
    n_want[i] (default all) eigenvectors of the lowest eigenvalues of batch
    i are computed, in the first n_want[i] columns of V (the others are left
    unwritten); E holds every eigenvalue. method "ql"
    (qmx_heev_small_grouped): all eigenvectors, one launch per accumulator
    placement (n <= 128 in LDS, above in the workspace)."""
    dev = rc.device
    flat = [_fp64(H).reshape((-1,) + tuple(H.shape[-2:])).contiguous() for H in mats]
    want = [int(H.shape[-1]) for H in flat] if n_want is None else [
        int(w) if w and 0 < int(w) <= H.shape[-1] else int(H.shape[-1]) for w, H in zip(n_want, flat)]
    outs = []
    for H in flat:
        B, n = H.shape[0], H.shape[-1]
        if n > SMALL_QL_MAX_N:
            raise ValueError(f"eigh_grouped: n = {n} > {SMALL_QL_MAX_N}")
        outs.append((torch.empty((B, n), dtype=torch.float64, device=dev),
                     torch.empty((B, n, n), dtype=torch.complex128, device=dev)))
    count = sum(H.shape[0] for H in flat)
    info = torch.zeros(max(count, 1), dtype=torch.int32, device=dev)
    if count == 0:
        return [(E.reshape(H.shape[:-1]), V.reshape(H.shape)) for H, (E, V) in zip(mats, outs)], info
    bisect = method == "bisect"
    work_of = {}
    total = 0
    for H, w in zip(flat, want):
        n = H.shape[-1]
        if (n, w) not in work_of:
            work_of[(n, w)] = int(_lib.lib().qmx_heevx_work_bytes(n, w) if bisect
                                  else _lib.lib().qmx_heev_small_work_bytes(n))
        total += work_of[(n, w)] * H.shape[0]
    ws = torch.empty(total, dtype=torch.uint8, device=dev)
    # One descriptor per matrix, built per batch with array arithmetic.
    parts = {True: [], False: []}
    off, i_info = 0, 0
    ws0, info0 = ws.data_ptr(), info.data_ptr()
    for H, (E, V), w in zip(flat, outs, want):
        B, n = H.shape[0], H.shape[-1]
        b = np.arange(B, dtype=np.uint64)
        d = np.empty(B, dtype=_HEEV_DESC)
        d["A"] = np.uint64(H.data_ptr()) + b * np.uint64(n * n * 16)
        d["E"] = np.uint64(E.data_ptr()) + b * np.uint64(n * 8)
        d["V"] = np.uint64(V.data_ptr()) + b * np.uint64(n * n * 16)
        d["info"] = np.uint64(info0 + 4 * i_info) + b * np.uint64(4)
        d["work"] = np.uint64(ws0 + off) + b * np.uint64(work_of[(n, w)])
        d["lda"], d["ldv"], d["n"], d["n_want"] = n, n, n, w
        parts[bisect or n <= 128].append(d)
        off += work_of[(n, w)] * B
        i_info += B
    keep = []
    for small, rows in parts.items():
        if not rows:
            continue
        arr = np.concatenate(rows)
        d = torch.from_numpy(arr.view(np.uint8)).pin_memory().to(dev, non_blocking=True)
        keep.append(d)
        n_max = int(arr["n"].max())
        _lib.call("qmx_heevx_grouped" if bisect else "qmx_heev_small_grouped", d.data_ptr(),
                  len(arr), n_max, rc.stream_ptr())
    return [(E.reshape(H.shape[:-1]), V.reshape(H.shape)) for H, (E, V) in zip(mats, outs)], info


def eigh_nosync(H: torch.Tensor):
    """(E, V, info) like eigh for complex128 (..., n, n) on the GPU, without
    a host synchronisation: info is None or an int32 tensor per matrix,
    nonzero where the QL iteration did not converge or the input was not
    finite, for the caller to check at its next synchronisation point."""
    n = H.shape[-1]
    n_mat = 1
    for b in H.shape[:-2]:
        n_mat *= int(b)
    if H.is_cuda and 0 < n <= SMALL_QL_MAX_N:
        method = small_method(n, n_mat)
        if method != "library":
            E, V = eigh_small(H, method)
            return E, V, (eigh_small.last_info if method in ("ql", "bisect") else None)
    if H.is_cuda and SMALL_QL_MAX_N < n <= LARGE_MAX_N:
        return eigh_large(H)
    E, V = eigh(H)
    return E, V, None


def eigh_large(H: torch.Tensor, n_want: Optional[int] = None):
    """(E, V, info) of complex128 (..., n, n), 2 <= n <= LARGE_MAX_N, on the
    GPU by qmx_heevx_large, one matrix per call: all n eigenvalues
    (ascending) and the eigenvectors of the n_want lowest (all by default)
    in the columns of V, a column-major (n, n_want) view per matrix; info
    (int32 per matrix) is 1 for a non-finite input (E and V NaN). Reads the
    lower triangle."""
    n = H.shape[-1]
    nw = n if n_want is None else max(1, min(int(n_want), n))
    batch = H.shape[:-2]
    Hf = _fp64(H).reshape((-1, n, n))
    if Hf.stride(-1) != 1 or Hf.stride(-2) != n:
        Hf = Hf.contiguous()
    need = int(_lib.lib().qmx_heevx_large_work_bytes(n, nw))
    key = (H.device, rc.stream_ptr())  # one workspace per stream: streams may overlap
    ws = _large_ws.get(key)
    if ws is None or ws.numel() < need:
        ws = torch.empty(need, dtype=torch.uint8, device=H.device)
        _large_ws[key] = ws
    B = Hf.shape[0]
    E = torch.empty((B, n), dtype=torch.float64, device=H.device)
    Vt = torch.empty((B, nw, n), dtype=torch.complex128, device=H.device)
    info = torch.empty(B, dtype=torch.int32, device=H.device)
    for b in range(B):
        _lib.call("qmx_heevx_large", Hf[b].data_ptr(), n, n, nw, E[b].data_ptr(), Vt[b].data_ptr(),
                  1, n, info[b:].data_ptr(), ws.data_ptr(), need, LARGE_GROUPS, rc.stream_ptr())
    return E.reshape(batch + (n,)), Vt.transpose(-1, -2).reshape(batch + (n, nw)), info


def eigh(H: torch.Tensor, n_want: Optional[int] = None) -> tuple[torch.Tensor, torch.Tensor]:
    """Hermitian eigendecomposition in FP64 (lower triangle, ascending), like
    torch.linalg.eigh. On the GPU: n <= SMALL_QL_MAX_N runs the whole batch
    in one batched launch (bisection, Householder + QL or Jacobi) where the
    cost model says it beats the library; SMALL_QL_MAX_N < n <= LARGE_MAX_N
    runs the grid-wide solver qmx_heevx_large, which also takes n_want, the
    number of lowest eigenvectors wanted (V then has n_want columns; smaller
    n return all n); larger n use the library eigensolver."""
    H = _fp64(H)
    n = H.shape[-1]
    if not H.is_cuda or n == 0:
        return torch.linalg.eigh(H)
    batch = H.shape[:-2]
    n_mat = 1
    for b in batch:
        n_mat *= int(b)
    if n <= SMALL_QL_MAX_N:
        method = small_method(n, n_mat)
        if method != "library":
            E, V = eigh_small(H, method)
            if method in ("ql", "bisect") and bool((eigh_small.last_info != 0).any()):
                return torch.linalg.eigh(H)  # QL did not converge / non-finite input
            return E, V
        return torch.linalg.eigh(H)
    if n <= LARGE_MAX_N:
        E, V, _ = eigh_large(H, n_want)
        return E, V
    return torch.linalg.eigh(H)


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


def cholesky_inverse(O: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """(X, info) for Hermitian positive definite complex128 O (..., n, n) on
    the GPU, n <= POTRF_MAX_N: X = L^-1 with O = L L^H (lower triangular,
    from the lower triangle of O, as torch.linalg.cholesky), by
    qmx_potrf_inv one matrix per call; info (int32 per matrix) is 0 or k + 1
    for the first non-positive pivot k (X then not meaningful), as
    torch.linalg.cholesky_ex."""
    n = O.shape[-1]
    batch = O.shape[:-2]
    Of = O.to(torch.complex128).reshape((-1, n, n))
    if Of.stride(-1) != 1 or Of.stride(-2) != n:
        Of = Of.contiguous()
    need = int(_lib.lib().qmx_potrf_inv_work_bytes(n))
    key = (O.device, rc.stream_ptr())  # one workspace per stream: streams may overlap
    ws = _potrf_ws.get(key)
    if ws is None or ws.numel() < need:
        ws = torch.empty(need, dtype=torch.uint8, device=O.device)
        _potrf_ws[key] = ws
    X = torch.empty_like(Of)
    info = torch.empty(Of.shape[0], dtype=torch.int32, device=O.device)
    for b in range(Of.shape[0]):
        _lib.call("qmx_potrf_inv", Of[b].data_ptr(), n, n, X[b].data_ptr(), info[b:].data_ptr(),
                  ws.data_ptr(), need, POTRF_GROUPS, rc.stream_ptr())
    return X.reshape(batch + (n, n)), info.reshape(batch)


def _own_cholesky(O: torch.Tensor) -> bool:
    return O.is_cuda and O.is_complex() and 0 < O.shape[-1] <= POTRF_MAX_N


def ortho_matrix(O: torch.Tensor, use_cholesky: bool = True) -> torch.Tensor:
    """U with U^H O U = 1: L^-H of the Cholesky factor O = L L^H (use_cholesky,
    as math/_linalg.py:110-141; on the GPU by qmx_potrf_inv), else the
    symmetric O^-1/2."""
    assert O.shape[-2] == O.shape[-1]
    O = _fp64(O)
    if use_cholesky:
        if _own_cholesky(O):
            X, info = cholesky_inverse(O)
            if bool((info != 0).any()):
                raise torch.linalg.LinAlgError(
                    f"ortho_matrix: the overlap is not positive definite (info {info.tolist()})")
            return dagger(X)
        identity = torch.eye(O.shape[-1], device=O.device, dtype=O.dtype).view(
            (1,) * (O.dim() - 2) + O.shape[-2:]
        )
        L = torch.linalg.cholesky(O)
        return dagger(torch.linalg.solve_triangular(L, identity.expand_as(L), upper=False))
    lbda, V = eigh(O)
    return V @ ((1.0 / torch.sqrt(lbda))[..., None] * dagger(V))


def eighg(H, O: torch.Tensor, use_cholesky: bool = True, n_want: Optional[int] = None):
    """E (ascending) and V with H V = O V E; H may be a Waitable. n_want
    (optional): only the eigenvectors of the n_want lowest eigenvalues are
    needed (V may then have just those columns)."""
    rc.compute_stream_wait_current()
    if rc.compute_stream is not None:
        with torch.cuda.stream(rc.compute_stream):
            U = ortho_matrix(O, use_cholesky)
    else:
        U = ortho_matrix(O, use_cholesky)
    Hres = _fp64(H.wait() if hasattr(H, "wait") else H)
    rc.current_stream_wait_compute()
    E, V = eigh(dagger(U) @ (Hres @ U), n_want)
    return E, U @ V


def eighg_canonical(H, O: torch.Tensor, rtol: float = 1e-4, shift: float = 1e6,
                    n_want: Optional[int] = None):
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
    lam, W = eigh(O)
    keep = lam > rtol * lam[..., -1:]
    s = torch.where(keep, lam.clamp_min(1e-300).rsqrt(), torch.zeros_like(lam))
    U = W * s[..., None, :].to(W.dtype)
    Hr = dagger(U) @ (Hres @ U)
    Hr = Hr + torch.diag_embed(torch.where(keep, 0.0, shift).to(Hr.dtype))
    E, V = eigh(Hr, n_want)
    return E, U @ V


#: Number of eighg_robust calls that needed the canonical path.
n_canonical_fallbacks = 0


def eighg_robust(H, O: torch.Tensor, n_want: Optional[int] = None):
    """Cholesky eighg when the overlap factorises, canonical orthogonalisation
    (eighg_canonical) when float32 rounding has made it indefinite."""
    global n_canonical_fallbacks
    O = _fp64(O)
    if _own_cholesky(O):
        X, info = cholesky_inverse(O)
        if bool((info != 0).any()):
            n_canonical_fallbacks += 1
            return eighg_canonical(H, O, n_want=n_want)
        U = dagger(X)
    else:
        L, info = torch.linalg.cholesky_ex(O)
        if bool((info != 0).any()):
            n_canonical_fallbacks += 1
            return eighg_canonical(H, O, n_want=n_want)
        identity = torch.eye(O.shape[-1], device=O.device, dtype=O.dtype).expand_as(L)
        U = dagger(torch.linalg.solve_triangular(L, identity, upper=False))
    Hres = _fp64(H.wait() if hasattr(H, "wait") else H)
    E, V = eigh(dagger(U) @ (Hres @ U), n_want)
    return E, U @ V
