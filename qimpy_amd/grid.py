"""Wavefunction FFT grid (reference: grid/_grid.py, grid/_fft.py).

``Grid.fft`` / ``Grid.ifft`` keep QimPy's conventions: the forward transform
exp(-i G r) is scaled by 1/N (``norm="forward"``) so G = 0 is the mean, and the
inverse is unscaled (_grid.py:220-268). Complex boxes use the batched c2c
Stockham transform of libqmx (``qmx_fft3d``); real input is transformed as
complex and truncated to the half space (shapeH), and half-space input is
extended by Hermitian symmetry before the inverse (the r2c/c2r dispatch of
_fft.py:228-337). One process, so the grid is never split (_basis.py:163).
"""
from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

from . import _lib, rc
from .systems import signed_freqs

_DT = {torch.complex64: _lib.QMX_C64, torch.complex128: _lib.QMX_C128}


def fft3d_inplace(data: torch.Tensor, direction: int, scale: float) -> torch.Tensor:
    """Transform the last three (contiguous) axes of `data` in place."""
    if data.dtype not in _DT:
        raise TypeError(f"qmx_fft3d needs complex64/complex128, not {data.dtype}")
    if not data.is_cuda:
        raise RuntimeError("qmx_fft3d operates on GPU tensors")
    assert data.is_contiguous()
    shape = np.ascontiguousarray(data.shape[-3:], dtype=np.int32)
    batch = int(np.prod(data.shape[:-3], dtype=np.int64))
    _lib.call(
        "qmx_fft3d", _DT[data.dtype], shape.ctypes.data_as(_lib.c_i32p), data.data_ptr(),
        batch, direction, float(scale), rc.stream_ptr(),
    )
    return data


class Grid:
    def __init__(self, shape: Sequence[int], Rbasis=None):
        self.shape = tuple(int(s) for s in shape)
        self.shapeH = (self.shape[0], self.shape[1], 1 + self.shape[2] // 2)
        self.Rbasis = None if Rbasis is None else np.asarray(Rbasis, dtype=np.float64)

    @property
    def N(self) -> int:
        return int(np.prod(self.shape))

    def get_mesh(self, space: str = "G") -> torch.Tensor:
        """Integer mesh (shape + (3,)): 'R' real space, 'G' signed frequencies,
        'H' half reciprocal space (_grid.py:160-182, _fft.py:65-82)."""
        if space == "R":
            m1 = [np.arange(s) for s in self.shape]
        else:
            m1 = [signed_freqs(s) for s in self.shape]
            if space == "H":
                m1[2] = m1[2][: self.shapeH[2]]
        mesh = np.stack(np.meshgrid(*m1, indexing="ij"), -1)
        return torch.tensor(mesh, device=rc.device)

    def fft(self, v: torch.Tensor) -> torch.Tensor:
        if v.shape[-3:] != self.shape:
            raise ValueError(f"fft input box {tuple(v.shape[-3:])} != grid {self.shape}")
        if any(s == 0 for s in v.shape[:-3]):
            return v if v.is_complex() else v.new_zeros(v.shape[:-3] + self.shapeH)
        if v.is_complex():
            return fft3d_inplace(v.clone(memory_format=torch.contiguous_format), -1, 1.0 / self.N)
        ctype = torch.complex128 if v.dtype == torch.float64 else torch.complex64
        out = fft3d_inplace(v.to(ctype).contiguous(), -1, 1.0 / self.N)
        return out[..., : self.shapeH[2]].contiguous()

    def ifft(self, v: torch.Tensor) -> torch.Tensor:
        if not v.is_complex():
            raise TypeError("ifft input must be complex")
        if v.shape[-1] == self.shape[2] and v.shape[-3:] == self.shape:
            if any(s == 0 for s in v.shape[:-3]):
                return v
            return fft3d_inplace(v.clone(memory_format=torch.contiguous_format), +1, 1.0)
        if v.shape[-3:] != self.shapeH:
            raise ValueError(f"ifft input {tuple(v.shape[-3:])} matches neither shape nor shapeH")
        rtype = torch.float64 if v.dtype == torch.complex128 else torch.float32
        if any(s == 0 for s in v.shape[:-3]):
            return torch.zeros(v.shape[:-3] + self.shape, dtype=rtype, device=v.device)
        full = self._hermitian_extend(v)
        return fft3d_inplace(full, +1, 1.0).real.contiguous()

    def _hermitian_extend(self, vH: torch.Tensor) -> torch.Tensor:
        n0, n1, n2 = self.shape
        full = torch.zeros(vH.shape[:-3] + self.shape, dtype=vH.dtype, device=vH.device)
        nh = self.shapeH[2]
        full[..., :nh] = vH
        # v(-G) = conj(v(G)) for the missing z >= nh entries:
        iz = torch.arange(nh, n2, device=vH.device)
        src_z = (-iz) % n2
        i0 = (-torch.arange(n0, device=vH.device)) % n0
        i1 = (-torch.arange(n1, device=vH.device)) % n1
        src = vH[..., src_z]  # (..., n0, n1, len)
        src = src[..., i0, :, :][..., :, i1, :]
        full[..., nh:] = src.conj()
        return full
