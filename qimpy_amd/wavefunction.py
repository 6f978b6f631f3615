"""Electronic wavefunctions (reference: dft/electrons/_wavefunction*.py).

``coeff`` is (n_spins, nk, n_bands, n_spinor, n_basis) on the GPU, in
complex128 (the reference's only type) or complex64 (the mixed-precision
form). Projections ``_proj`` (n_spins, nk, n_projectors, n_bands) are cached
and invalidated by ``Ions.beta_version`` exactly as in the reference: they
propagate through ``@`` (_wavefunction_dot.py:188-190), scalar ``*`` /
band scalings (_wavefunction_arithmetic.py:9-42), ``+``/``-`` when both
operands hold them (:45-109), slicing (_wavefunction_slice.py:9-39) and
``cat`` (:42-70). One process: ``band_division`` is always None, the basis is
never split and every collective of the reference is a no-op.
"""
from __future__ import annotations

from typing import Any, Optional, Union

import numpy as np
import torch

from . import _lib, gemm, rc
from .basis import DTYPE_CODE, Basis
from .linalg import Waitless, abs_squared, ortho_matrix


class Wavefunction:
    basis: Basis
    coeff: torch.Tensor
    _proj: Optional[torch.Tensor]
    _proj_version: int
    band_division: None

    def __init__(
        self,
        basis: Basis,
        *,
        coeff: Optional[torch.Tensor] = None,
        band_division=None,
        n_bands: int = 0,
        n_spins: int = 0,
        n_spinor: int = 0,
        dtype: torch.dtype = torch.complex128,
    ) -> None:
        if band_division is not None:
            raise NotImplementedError("band-split wavefunctions need more than one process")
        self.basis = basis
        self.band_division = None
        self._proj_invalidate()
        if coeff is None:
            assert n_bands
            n_spins = n_spins or basis.n_spins
            n_spinor = n_spinor or basis.n_spinor
            self.coeff = torch.zeros(
                (n_spins, basis.nk, n_bands, n_spinor, basis.n_tot), dtype=dtype,
                device=rc.device,
            )
            ions = basis.ions
            if ions is not None:
                self._proj = torch.zeros(
                    (n_spins, basis.nk, ions.n_projectors, n_bands), dtype=dtype,
                    device=rc.device,
                )
                self._proj_version = ions.beta_version
        else:
            self.coeff = coeff

    # ------------------------------------------------------------ projections
    @property
    def proj(self) -> torch.Tensor:
        if self._proj_is_valid():
            return self._proj
        ions = self.basis.ions
        beta = ions.beta_coeff(self.coeff.dtype)
        self._proj = gemm.dot(beta.flatten(-2), self.coeff.flatten(-2))
        self._proj_version = ions.beta_version
        return self._proj

    def _proj_invalidate(self) -> None:
        self._proj = None
        self._proj_version = -1

    def _proj_is_valid(self) -> bool:
        ions = self.basis.ions
        return ions is not None and self._proj_version == ions.beta_version

    # ------------------------------------------------------------ basics
    def zeros_like(self, n_bands: int = 0) -> "Wavefunction":
        return Wavefunction(
            self.basis, n_bands=n_bands or self.coeff.shape[2], n_spins=self.coeff.shape[0],
            n_spinor=self.coeff.shape[3], dtype=self.coeff.dtype,
        )

    def clone(self) -> "Wavefunction":
        result = Wavefunction(self.basis, coeff=self.coeff.clone().detach())
        if self._proj_is_valid():
            result._proj = self._proj.clone().detach()
            result._proj_version = self._proj_version
        return result

    def to(self, dtype: torch.dtype) -> "Wavefunction":
        """Copy in another precision (projections are recomputed on demand)."""
        return Wavefunction(self.basis, coeff=self.coeff.to(dtype))

    def n_bands(self) -> int:
        return self.coeff.shape[2]

    @property
    def spinorial(self) -> bool:
        return self.coeff.shape[-2] != 1

    @property
    def non_spinor(self) -> "Wavefunction":
        if self.spinorial:
            return Wavefunction(self.basis, coeff=self.coeff.flatten(2, 3).unsqueeze(3))
        return self

    def _rows(self) -> torch.Tensor:
        """coeff as contiguous (S*K, B, G) rows (X = 1)."""
        c = self.coeff
        if not c.is_contiguous():
            c = c.contiguous()
            self.coeff = c
        return c

    def band_norm(self) -> torch.Tensor:
        c = self._rows()
        S, K, B, X, G = c.shape
        out = torch.empty((S, K, B), dtype=torch.float64, device=c.device)
        _lib.call(
            "qmx_band_reduce", DTYPE_CODE[c.dtype], c.data_ptr(), X * G, B * X * G, S * K, B,
            X * G, None, 0, out.data_ptr(), None, rc.stream_ptr(),
        )
        return out.sqrt()

    def band_ke(self) -> torch.Tensor:
        c = self._rows()
        S, K, B, X, G = c.shape
        if S != 1:
            return torch.einsum("skbxg, kg -> skb", abs_squared(c).to(torch.float64),
                                self.basis.get_ke())
        norm2 = torch.empty((S, K, B), dtype=torch.float64, device=c.device)
        kin = torch.empty_like(norm2)
        ke = self.basis.get_ke()
        _lib.call(
            "qmx_band_reduce", DTYPE_CODE[c.dtype], c.data_ptr(), X * G, B * X * G, K, B, G,
            ke.data_ptr(), ke.stride(0), norm2.data_ptr(), kin.data_ptr(), rc.stream_ptr(),
        )
        return kin

    def norm(self) -> float:
        return float(np.sqrt(abs_squared(self.coeff).sum().item()))

    def constrain(self) -> None:
        """Zero padded coefficients (_wavefunction.py:264-275)."""
        self._scale_constrain(None)

    def scale_constrain_(self, scale: torch.Tensor) -> None:
        """coeff *= scale[s, k, b] and zero the padding, in one pass."""
        self._scale_constrain(scale.to(torch.float64).contiguous())
        if self._proj_is_valid():
            self._proj *= scale.to(self._proj.real.dtype)[:, :, None, :]

    def _scale_constrain(self, scale: Optional[torch.Tensor]) -> None:
        c = self._rows()
        S, K, B, X, G = c.shape
        if X != 1:
            raise NotImplementedError("spinorial wavefunctions are not on this path")
        for s in range(S):
            cs = c[s]
            _lib.call(
                "qmx_scale_constrain", DTYPE_CODE[c.dtype], cs.data_ptr(), cs.stride(1),
                cs.stride(0), K, B, G, None if scale is None else scale[s].data_ptr(),
                self.basis.n_i32.data_ptr(), rc.stream_ptr(),
            )

    # ------------------------------------------------------------ randomize
    def randomize(self, seed: int = 0, b_start: int = 0, b_stop: Optional[int] = None) -> None:
        """Bandwidth-limited xorshift random bands (_wavefunction_init.py:27-125)."""
        c = self._rows()
        S, K, B, X, G = c.shape
        b_stop = b_stop if b_stop else B
        if b_start >= b_stop:
            return
        basis = self.basis
        ke = basis.get_ke()
        for s in range(S):
            cs = c[s]
            # spin enters the state as n_max * nk_tot * i_spin (:85-93):
            _lib.call(
                "qmx_randomize", DTYPE_CODE[c.dtype], cs.data_ptr(), cs.stride(1), cs.stride(0),
                K, G, basis.n_max, K * s, basis.n_i32.data_ptr(), ke.data_ptr(), ke.stride(0),
                seed, b_start, b_stop, rc.stream_ptr(),
            )
        self._proj_invalidate()

    def randomize_selected(self, i_spin, i_k, i_band, seed: int) -> None:
        """Randomize selected bands (_wavefunction_init.py:128-188)."""
        basis = self.basis
        n_bands = self.n_bands()
        dev = self.coeff.device
        G = basis.n_tot
        i_basis = torch.arange(G, device=dev)
        x = (
            (1 + seed)
            + i_basis.view(1, -1)
            + basis.n_max * (i_band.view(-1, 1) + n_bands * (i_k.view(-1, 1) + basis.nk * i_spin.view(-1, 1)))
        ).to(torch.int64)

        def randn(x):
            u = []
            for _ in range(2):
                x ^= x << 13
                x ^= x >> 7
                x ^= x << 17
                u.append((0.5**64) * x.to(torch.float64) + 0.5)
            return torch.sqrt(-2.0 * torch.log(u[0])) * torch.exp((2j * np.pi) * u[1])

        randn(x)
        vals = randn(x)
        ke = basis.get_ke()[i_k]
        vals = vals * (1.0 / (1.0 + ((4.0 / 3) * ke) ** 6))
        self.coeff[i_spin, i_k, i_band, 0] = vals.to(self.coeff.dtype)
        self.constrain()  # the padding of every band, as the reference (:176-177)
        self._proj_invalidate()

    # ------------------------------------------------------------ algebra
    def dot(self, other: "Wavefunction"):
        if not isinstance(other, Wavefunction):
            return NotImplemented
        assert self.basis is other.basis
        assert self.coeff.shape[-1] == other.coeff.shape[-1]
        if self.spinorial != other.spinorial:
            raise NotImplementedError("one-sided spinorial products are not on this path")
        return Waitless(gemm.dot(self.coeff.flatten(-2), other.coeff.flatten(-2)))

    __xor__ = dot

    def dot_O(self, other: "Wavefunction"):
        return self.dot(other)  # norm-conserving: O = 1 (_wavefunction_dot.py:110-126)

    def overlap(self) -> "Wavefunction":
        return self

    def matmul(self, mat: torch.Tensor) -> "Wavefunction":
        if not isinstance(mat, torch.Tensor):
            return NotImplemented
        n_bands_in, n_spinor, n_basis = self.coeff.shape[-3:]
        if mat.shape[-2] != n_bands_in:
            raise ValueError(f"matrix rows {mat.shape[-2]} != bands {n_bands_in}")
        C_out = gemm.rotate(mat, self.coeff.flatten(-2))
        C_out = C_out.view(C_out.shape[:-2] + (mat.shape[-1], n_spinor, n_basis))
        result = Wavefunction(self.basis, coeff=C_out)
        if self._proj_is_valid():
            result._proj = self._proj @ mat.to(self._proj.dtype)
            result._proj_version = self._proj_version
        return result

    __matmul__ = matmul

    def orthonormalize(self) -> "Wavefunction":
        return self @ ortho_matrix(self.dot_O(self).wait(), use_cholesky=True)

    # arithmetic (_wavefunction_arithmetic.py)
    @staticmethod
    def _scale_kind(scale) -> tuple[bool, bool]:
        if isinstance(scale, (float, int)) and not isinstance(scale, bool):
            return True, True
        if isinstance(scale, torch.Tensor) and scale.dim() == 5:
            return True, tuple(scale.shape[-2:]) == (1, 1)
        return False, False

    def __mul__(self, scale: Union[float, torch.Tensor]):
        ok, band = self._scale_kind(scale)
        if not ok:
            return NotImplemented
        s = scale.to(self.coeff.real.dtype) if isinstance(scale, torch.Tensor) and not scale.is_complex() else scale
        result = Wavefunction(self.basis, coeff=self.coeff * s)
        if band and self._proj_is_valid():
            ps = s if not isinstance(s, torch.Tensor) else s[..., 0, 0][:, :, None, :]
            result._proj = self._proj * ps
            result._proj_version = self._proj_version
        return result

    __rmul__ = __mul__

    def __imul__(self, scale: Union[float, torch.Tensor]):
        ok, band = self._scale_kind(scale)
        if not ok:
            return NotImplemented
        s = scale.to(self.coeff.real.dtype) if isinstance(scale, torch.Tensor) and not scale.is_complex() else scale
        self.coeff *= s
        if band and self._proj_is_valid():
            ps = s if not isinstance(s, torch.Tensor) else s[..., 0, 0][:, :, None, :]
            self._proj *= ps
        else:
            self._proj_invalidate()
        return self

    def __add__(self, other: "Wavefunction"):
        if not isinstance(other, Wavefunction):
            return NotImplemented
        assert self.basis is other.basis
        result = Wavefunction(self.basis, coeff=self.coeff + other.coeff)
        if self._proj_is_valid() and other._proj_is_valid():
            result._proj = self._proj + other._proj
            result._proj_version = self._proj_version
        return result

    def __iadd__(self, other: "Wavefunction"):
        if not isinstance(other, Wavefunction):
            return NotImplemented
        assert self.basis is other.basis
        self.coeff += other.coeff
        if self._proj_is_valid() and other._proj_is_valid():
            self._proj += other._proj
        else:
            self._proj_invalidate()
        return self

    def __sub__(self, other: "Wavefunction"):
        if not isinstance(other, Wavefunction):
            return NotImplemented
        assert self.basis is other.basis
        result = Wavefunction(self.basis, coeff=self.coeff - other.coeff)
        if self._proj_is_valid() and other._proj_is_valid():
            result._proj = self._proj - other._proj
            result._proj_version = self._proj_version
        return result

    def __isub__(self, other: "Wavefunction"):
        if not isinstance(other, Wavefunction):
            return NotImplemented
        assert self.basis is other.basis
        self.coeff -= other.coeff
        if self._proj_is_valid() and other._proj_is_valid():
            self._proj -= other._proj
        else:
            self._proj_invalidate()
        return self

    # slicing (_wavefunction_slice.py)
    @staticmethod
    def _proj_index(index: Any) -> Any:
        if isinstance(index, tuple) and len(index) > 2:
            return index[:2] + (slice(None),) + index[2:]
        return index

    def __getitem__(self, index: Any) -> "Wavefunction":
        result = Wavefunction(self.basis, coeff=self.coeff[index])
        if self._proj_is_valid():
            result._proj = self._proj[self._proj_index(index)]
            result._proj_version = self._proj_version
        return result

    def __setitem__(self, index: Any, other: "Wavefunction") -> None:
        self.coeff[index] = other.coeff
        if self._proj_is_valid() and other._proj_is_valid():
            self._proj[self._proj_index(index)] = other._proj
        else:
            self._proj_invalidate()

    def cat(self, other: "Wavefunction", dim: int = 2, clear: bool = False) -> "Wavefunction":
        result = Wavefunction(self.basis, coeff=torch.cat((self.coeff, other.coeff), dim=dim))
        both = self._proj_is_valid() and other._proj_is_valid()
        if both:
            # _proj is (S, K, P, B): coefficient dims >= 2 sit one axis later.
            # (_wavefunction_slice.py:70 concatenates along `dim` itself, i.e.
            # along projectors for bands, which fails whenever both operands
            # cache projections, as CheFSI's C0.cat(C) does on one process.)
            result._proj = torch.cat((self._proj, other._proj), dim=dim + 1 if dim >= 2 else dim)
            result._proj_version = self._proj_version
        if clear:
            del self.coeff
            del other.coeff
            if both:
                self._proj_invalidate()
                other._proj_invalidate()
        return result
