"""Plane-wave basis and its operators (reference: dft/electrons/_basis.py,
_basis_ops.py).

The sphere is the set of box plane waves with ``KE(G+k) < ke_cutoff``
(_basis.py:167-177); ``fft_index`` lists their linear box indices in
increasing order, padded with the first out-of-sphere indices to the common
length ``n_tot = n_max`` (_basis.py:183-204, one basis process so no extra
padding). ``apply_potential`` and ``apply_ke`` run on the sparse-FFT HIP
kernels through ``qmx_apply_local``; nothing here has a CPU fallback.
"""
from __future__ import annotations

import ctypes as ct
from typing import Optional

import numpy as np
import torch

from . import _lib, rc
from .grid import Grid
from .systems import signed_freqs

DTYPE_CODE = {torch.complex64: _lib.QMX_C64, torch.complex128: _lib.QMX_C128}
REAL_OF = {torch.complex64: torch.float32, torch.complex128: torch.float64}

#: Largest workspace the local apply allocates; bands beyond it are blocked.
WORKSPACE_CAP = 1 << 30


class Basis:
    """Plane waves of one cell at a set of k-points, on one GPU."""

    def __init__(
        self,
        *,
        Rbasis: np.ndarray,
        k: np.ndarray,
        wk: np.ndarray,
        ke_cutoff: float,
        shape,
        ions=None,
        n_spins: int = 1,
        n_spinor: int = 1,
    ):
        rc.require_gpu()
        if n_spinor not in (1, 2):
            raise ValueError(f"n_spinor must be 1 or 2, not {n_spinor}")
        self.Rbasis = np.asarray(Rbasis, dtype=np.float64)
        self.Gbasis = 2 * np.pi * np.linalg.inv(self.Rbasis.T)  # lattice/_lattice.py:217
        self.volume = float(abs(np.linalg.det(self.Rbasis)))
        self.k_np = np.asarray(k, dtype=np.float64).reshape(-1, 3)
        self.wk_np = np.asarray(wk, dtype=np.float64)
        self.n_spins, self.n_spinor = n_spins, n_spinor
        self.ke_cutoff = float(ke_cutoff)
        self.ions = ions
        self.grid = Grid(shape, self.Rbasis)
        self.real_wavefunctions = False
        dev = rc.device
        self.k = torch.tensor(self.k_np, device=dev)
        self.wk = torch.tensor(self.wk_np, device=dev)
        w_spin = 2 // (n_spins * n_spinor)  # _basis.py:137
        self.w_sk = w_spin * self.wk.view(1, -1, 1)

        # Sphere selection and ordering on the host in float64:
        mesh = np.stack(
            np.meshgrid(*[signed_freqs(s) for s in self.grid.shape], indexing="ij"), -1
        ).reshape(-1, 3)
        N = mesh.shape[0]
        nk = self.k_np.shape[0]
        within = np.empty((nk, N), dtype=bool)
        for ik in range(nk):
            Gk = (mesh + self.k_np[ik]) @ self.Gbasis.T
            within[ik] = 0.5 * np.einsum("gi,gi->g", Gk, Gk) < self.ke_cutoff
        self.n_np = within.sum(axis=1).astype(np.int64)
        self.n_min, self.n_max = int(self.n_np.min()), int(self.n_np.max())
        self.n_tot = self.n_max
        self.n_avg = float((self.n_np * self.wk_np).sum())
        assert self.n_tot <= N
        keys = np.where(within, 0, N) + np.arange(N)[None]
        self.fft_index_np = np.argsort(keys, axis=1, kind="stable")[:, : self.n_tot]
        iG = mesh[self.fft_index_np]  # (nk, n_tot, 3)
        self.iG = torch.tensor(iG, device=dev)
        self.fft_index = torch.tensor(self.fft_index_np, device=dev)
        self.n = torch.tensor(self.n_np, device=dev)
        self.n_i32 = self.n.to(torch.int32)
        g_range = torch.arange(self.n_tot, device=dev)
        pad = torch.where(g_range[None, :] >= self.n[:, None])
        self.pad_index = (slice(None), pad[0], slice(None), slice(None), pad[1])
        self.pad_index_mine = self.pad_index
        self.mine = slice(0, self.n_tot)
        Gk = (self.iG.to(torch.float64) + self.k[:, None, :]) @ torch.tensor(
            self.Gbasis.T, device=dev
        )
        self._ke = {torch.float64: 0.5 * Gk.square().sum(dim=-1)}  # get_ke :256-270
        self._ke[torch.float32] = self._ke[torch.float64].to(torch.float32)
        self._gk = {torch.float64: Gk.contiguous()}  # Cartesian G + k (get_gradient / i)
        self._gk[torch.float32] = self._gk[torch.float64].to(torch.float32)

        # Sparse-FFT plan (index tables on the device):
        geom, keep = _lib.make_geom(self.grid.shape, self.n_np, self.fft_index_np)
        self._plan = _lib.create_plan(geom)
        info = np.zeros(2, dtype=np.int64)
        _lib.call("qmx_plan_info", self._plan, info.ctypes.data_as(_lib.c_i64p))
        self.n_active_planes, self.n_active_columns = int(info[0]), int(info[1])
        self._workspace: Optional[torch.Tensor] = None
        self._V_cache: dict = {}

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            try:
                _lib.lib().qmx_plan_destroy(plan)
            except Exception:  # pragma: no cover - interpreter shutdown
                pass
            self._plan = None

    @property
    def nk(self) -> int:
        return int(self.k_np.shape[0])

    @property
    def n_ideal(self) -> float:
        return ((2.0 * self.ke_cutoff) ** 1.5) * self.volume / (6 * np.pi**2)

    def get_ke(self, basis_slice: slice = slice(None), dtype=torch.float64) -> torch.Tensor:
        """KE of each plane wave (nk x len(slice)), cached per precision."""
        return self._ke[dtype][:, basis_slice]

    def set_local_path(self, path: int) -> None:
        """Kernel sequence of the local apply (qmx_plan_set_local_path): 0
        the fused plane pass where a plane fits LDS, else the separate
        y, x and y passes; 1 always the separate passes."""
        _lib.call("qmx_plan_set_local_path", self._plan, int(path))

    def local_path(self, dtype: torch.dtype) -> str:
        """'plane' or 'separate': the kernel sequence the apply takes."""
        code = int(_lib.lib().qmx_plan_local_path(self._plan, DTYPE_CODE[dtype]))
        return {2: "plane", 1: "separate"}[code]

    # ---------------------------------------------------------------- operators
    def workspace(self, dtype: torch.dtype, n_bands: int) -> tuple[torch.Tensor, int]:
        need = int(_lib.lib().qmx_plan_workspace_bytes(self._plan, DTYPE_CODE[dtype], n_bands))
        one = int(_lib.lib().qmx_plan_workspace_bytes(self._plan, DTYPE_CODE[dtype], 1))
        size = max(one, min(need, WORKSPACE_CAP))
        if self._workspace is None or self._workspace.numel() < size:
            self._workspace = torch.empty(size, dtype=torch.uint8, device=rc.device)
        return self._workspace, int(self._workspace.numel())

    def potential_box(self, V, dtype: torch.dtype) -> torch.Tensor:
        """V(r) on the box, contiguous, in the real type matching `dtype`."""
        data = V.data if hasattr(V, "data") and not isinstance(V, torch.Tensor) else V
        if not isinstance(data, torch.Tensor):
            data = torch.as_tensor(np.asarray(data), device=rc.device)
        if data.dim() == 4 and data.shape[0] == 1:
            data = data[0]
        if tuple(data.shape) != self.grid.shape:
            raise ValueError(f"potential shape {tuple(data.shape)} != box {self.grid.shape}")
        rtype = REAL_OF[dtype]
        key = (data.data_ptr(), data._version, rtype)
        cached = self._V_cache.get(key)
        if cached is None:
            cached = data.to(device=rc.device, dtype=rtype).contiguous()
            if len(self._V_cache) >= 4:  # keep the latest few (V_ks and V_tau alternate)
                self._V_cache.pop(next(iter(self._V_cache)))
            self._V_cache[key] = cached
        return cached

    def apply_local(self, coeff: torch.Tensor, V, with_ke: bool, out: Optional[torch.Tensor] = None,
                    accumulate: bool = False) -> torch.Tensor:
        """out = [KE C] + V C on (S, K, B, X, G) coefficients via qmx_apply_local."""
        if coeff.dtype not in DTYPE_CODE:
            raise TypeError(f"unsupported wavefunction dtype {coeff.dtype}")
        S, K, B, X, G = coeff.shape
        if K != self.nk or G != self.n_tot or X != 1:
            raise ValueError(f"coefficients {tuple(coeff.shape)} do not match the basis")
        if coeff.stride(-1) != 1 or coeff.stride(2) < G or (K > 1 and coeff.stride(1) < G):
            coeff = coeff.contiguous()
        if out is None:
            out = torch.empty((S, K, B, X, G), dtype=coeff.dtype, device=coeff.device)
        Vbox = self.potential_box(V, coeff.dtype)
        ke = self.get_ke(dtype=REAL_OF[coeff.dtype]) if with_ke else None
        ws, ws_bytes = self.workspace(coeff.dtype, B)
        code = DTYPE_CODE[coeff.dtype]
        stream = rc.stream_ptr()
        esize = coeff.element_size()
        for s in range(S):
            cs, os_ = coeff[s], out[s]
            _lib.call(
                "qmx_apply_local", self._plan, code, cs.data_ptr(), cs.stride(1), cs.stride(0),
                os_.data_ptr(), os_.stride(1), os_.stride(0), Vbox.data_ptr(),
                ke.data_ptr() if ke is not None else None, ke.stride(0) if ke is not None else 0,
                B, int(accumulate), ws.data_ptr(), ws_bytes, stream,
            )
        del esize
        return out

    def gradient(self, coeff: torch.Tensor, i_dir: int, alpha: float = 1.0,
                 out: Optional[torch.Tensor] = None, accumulate: bool = False) -> torch.Tensor:
        """out = [out +] alpha i (G + k)_dir C on (S, K, B, X, G) coefficients
        (qmx_apply_gradient)."""
        if coeff.dtype not in DTYPE_CODE:
            raise TypeError(f"unsupported wavefunction dtype {coeff.dtype}")
        S, K, B, X, G = coeff.shape
        if K != self.nk or G != self.n_tot:
            raise ValueError(f"coefficients {tuple(coeff.shape)} do not match the basis")
        if coeff.stride(-1) != 1 or coeff.stride(3) < G or coeff.stride(2) < X * G:
            coeff = coeff.contiguous()
        if out is None:
            out = torch.empty((S, K, B, X, G), dtype=coeff.dtype, device=coeff.device)
        gk = self._gk[REAL_OF[coeff.dtype]]
        code = DTYPE_CODE[coeff.dtype]
        for s in range(S):
            for x in range(X):
                cs, os_ = coeff[s, :, :, x], out[s, :, :, x]
                _lib.call("qmx_apply_gradient", code, cs.data_ptr(), cs.stride(1), cs.stride(0),
                          K, B, G, gk.data_ptr(), int(i_dir), float(alpha), os_.data_ptr(),
                          os_.stride(1), os_.stride(0), int(accumulate), rc.stream_ptr())
        return out

    def apply_gradient(self, C, i_dir: int):
        """Basis.apply_gradient (_basis_ops.py:15-26): i (G + k)_dir C."""
        from .wavefunction import Wavefunction

        return Wavefunction(self, coeff=self.gradient(C.coeff, i_dir),
                            band_division=C.band_division)

    def apply_tau(self, V_tau, coeff: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        """out -= 1/2 sum_dir grad_dir(V_tau grad_dir C), the meta-GGA term of
        H (_hamiltonian.py:13-20): per direction one gradient, one local apply
        and one accumulating gradient scaled by -1/2."""
        tmp = None
        for i_dir in range(3):
            tmp = self.gradient(coeff, i_dir, out=tmp)
            tmp = self.apply_local(tmp, V_tau, with_ke=False, out=tmp)
            self.gradient(tmp, i_dir, alpha=-0.5, out=out, accumulate=True)
        return out

    def apply_ke(self, C):
        from .wavefunction import Wavefunction

        coeff = C.coeff * self.get_ke(dtype=REAL_OF[C.coeff.dtype])[None, :, None, None, :]
        return Wavefunction(self, coeff=coeff, band_division=C.band_division)

    def apply_potential(self, V, C):
        from .wavefunction import Wavefunction

        out = self.apply_local(C.coeff, V, with_ke=False)
        return Wavefunction(self, coeff=out, band_division=C.band_division)

    def apply_ke_potential(self, V, C):
        """Fused T C + V C (one pass over the sphere, KE folded into the gather)."""
        from .wavefunction import Wavefunction

        out = self.apply_local(C.coeff, V, with_ke=True)
        return Wavefunction(self, coeff=out, band_division=C.band_division)

    def collect_density(self, C, f: torch.Tensor, need_Mvec: bool = False) -> torch.Tensor:
        """Electron density of `C` with fillings `f` (S, K, B) on the box, as
        (n, S0, S1, S2) float64 (Basis.collect_density, dft/electrons/
        _basis_ops.py:207-287; prefactor f w_sk / volume at :229):
        - without need_Mvec, [n_tot] for one spin channel, [n_tot, M_z] for
          two; spinor components are summed (the prefactor tiled over them,
          :231-234, accum_norm_ over k, bands and components at :338);
        - with need_Mvec (spinorial C), [n_tot, M_x, M_y, M_z] from the
          component densities rho_0, rho_1 (accum_norm_ at :329) and the
          cross term X = sum f psi_1 conj(psi_0) (accum_prod_ at :330-336):
          n_tot = rho_0 + rho_1, M_x = 2 Re X, M_y = 2 Im X,
          M_z = rho_0 - rho_1 (:277-284).
        One qmx_collect_density launch sequence per spin and component, one
        qmx_collect_density_cross per spin."""
        coeff = C.coeff if hasattr(C, "coeff") else C
        if coeff.dtype not in DTYPE_CODE:
            raise TypeError(f"unsupported wavefunction dtype {coeff.dtype}")
        S, K, B, X, G = coeff.shape
        if K != self.nk or G != self.n_tot or X not in (1, 2):
            raise ValueError(f"coefficients {tuple(coeff.shape)} do not match the basis")
        if need_Mvec and X != 2:
            raise ValueError("need_Mvec needs spinorial wavefunctions (n_spinor = 2)")
        f = torch.as_tensor(f, dtype=torch.float64, device=coeff.device)
        if tuple(f.shape) != (S, K, B):
            raise ValueError(f"fillings {tuple(f.shape)} != {(S, K, B)}")
        if (coeff.stride(-1) != 1 or coeff.stride(3) < G or coeff.stride(2) < X * G
                or (K > 1 and coeff.stride(1) < G)):
            coeff = coeff.contiguous()
        prefac = (f * (self.w_sk.to(torch.float64) / self.volume)).contiguous()  # (S, K, B)
        code = DTYPE_CODE[coeff.dtype]
        lib = _lib.lib()

        def workspace(fn):
            need, one = int(fn(self._plan, code, B)), int(fn(self._plan, code, 1))
            return torch.empty(max(one, min(need, WORKSPACE_CAP)), dtype=torch.uint8,
                               device=coeff.device)

        ws = workspace(lib.qmx_density_workspace_bytes)
        box = self.grid.shape
        n_diag = 2 if need_Mvec else S
        rho_diag = torch.empty((n_diag,) + box, dtype=torch.float64, device=coeff.device)
        for s in range(S):
            for x in range(X):
                cs = coeff[s, :, :, x]
                # need_Mvec: component x summed over spins; else spin s over components
                out, acc = (rho_diag[x], s > 0) if need_Mvec else (rho_diag[s], x > 0)
                _lib.call("qmx_collect_density", self._plan, 0, K, code, cs.data_ptr(),
                          cs.stride(1), cs.stride(0), B, prefac[s].data_ptr(), prefac.stride(1),
                          out.data_ptr(), int(acc), ws.data_ptr(), ws.numel(), rc.stream_ptr())
        if not need_Mvec:
            density = torch.empty_like(rho_diag)
            density[0] = rho_diag.sum(dim=0)  # n_tot
            if S == 2:
                density[1] = rho_diag[0] - rho_diag[1]  # M_z
            return density
        del ws
        ws = workspace(lib.qmx_density_cross_workspace_bytes)
        cross = torch.empty((2,) + box, dtype=torch.float64, device=coeff.device)  # Re, Im
        for s in range(S):
            c0, c1 = coeff[s, :, :, 0], coeff[s, :, :, 1]
            _lib.call("qmx_collect_density_cross", self._plan, 0, K, code, c0.data_ptr(),
                      c1.data_ptr(), c0.stride(1), c0.stride(0), B, prefac[s].data_ptr(),
                      prefac.stride(1), cross.data_ptr(), int(s > 0), ws.data_ptr(), ws.numel(),
                      rc.stream_ptr())
        density = torch.empty((4,) + box, dtype=torch.float64, device=coeff.device)
        density[0] = rho_diag.sum(dim=0)  # n_tot
        density[1] = 2.0 * cross[0]  # M_x
        density[2] = 2.0 * cross[1]  # M_y
        density[3] = rho_diag[0] - rho_diag[1]  # M_z
        return density

    @classmethod
    def from_spec(cls, spec, ions=None) -> "Basis":
        return cls(Rbasis=spec.Rbasis, k=spec.k, wk=spec.wk, ke_cutoff=spec.ke_cutoff,
                   shape=spec.shape, ions=ions)
