"""Write reference_vectors.npz: input/output vectors of the reference's own
wavefunction-initialisation and preconditioner functions, for the checks of
the oracle that no closed form can make (which of the two uniforms sets the
Box-Muller modulus, the seed layout of randomize_selected, the exp(-x) floor
of the preconditioner).

The reference package cannot be imported in this image (its __init__ needs
mpi4py and h5py, which are absent), so this script compiles only the
functions it needs out of the reference's source files as they lie under
/root/reference (decorators dropped) and calls them on small seeded inputs,
with plain attribute objects in place of the Basis/Wavefunction arguments
they read. No reference source is stored: the fixture holds the inputs and
outputs only. Run in a container that has /root/reference:

    python tests/golden/make_reference_vectors.py

Functions (reference files under src/qimpy/dft/electrons/):
- _randn (_wavefunction_init.py:11-24);
- _randomize (_wavefunction_init.py:27-125), basis-split branch;
- _randomize_selected (_wavefunction_init.py:128-188);
- Davidson.precondition (_davidson.py:92-99).
"""
from __future__ import annotations

import ast
import os
from types import SimpleNamespace
from typing import Optional

import numpy as np
import torch

REF = "/root/reference/src/qimpy/dft/electrons"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_vectors.npz")


def load_functions(path: str, names: list, cls: Optional[str] = None) -> dict:
    """Compile the top-level functions (or methods of class `cls`) `names`
    of a reference source file, without decorators, in a namespace holding
    torch and numpy."""
    with open(path) as f:
        tree = ast.parse(f.read())
    body = tree.body
    if cls is not None:
        body = next(n for n in body if isinstance(n, ast.ClassDef) and n.name == cls).body
    defs = []
    for node in body:
        if isinstance(node, ast.FunctionDef) and node.name in names:
            node.decorator_list = []
            defs.append(node)
    if len(defs) != len(names):
        raise RuntimeError(f"{path}: found {[d.name for d in defs]}, wanted {names}")
    mod = ast.Module(body=defs, type_ignores=[])
    ns = {"torch": torch, "np": np, "Optional": Optional, "electrons": SimpleNamespace(),
          "Wavefunction": object, "__builtins__": __builtins__}
    exec(compile(mod, path, "exec"), ns)
    return {n: ns[n] for n in names}


def fake_basis(K: int, G: int, n: np.ndarray, KE: torch.Tensor) -> SimpleNamespace:
    g = torch.arange(G)
    pad = torch.where(g[None, :] >= torch.tensor(n)[:, None])
    return SimpleNamespace(
        n_spinor=1, n_spins=1, n_max=int(n.max()), n_tot=G, real_wavefunctions=False,
        kpoints=SimpleNamespace(division=SimpleNamespace(i_start=0, i_stop=K, n_tot=K)),
        division=SimpleNamespace(i_start=0, i_stop=G),
        pad_index_mine=(slice(None), pad[0], slice(None), slice(None), pad[1]),
        mine=slice(0, G), get_ke=lambda sl=slice(None): KE[:, sl])


def main():
    torch.set_default_dtype(torch.float64)  # as the reference's rc.py:54
    init = load_functions(os.path.join(REF, "_wavefunction_init.py"),
                          ["_randn", "_randomize", "_randomize_selected"])
    dav = load_functions(os.path.join(REF, "_davidson.py"), ["precondition"], cls="Davidson")
    rng = np.random.default_rng(2024)
    out = {}

    # _randn: three consecutive draws from a set of int64 states.
    x0 = np.concatenate([np.array([1, 2, 3, -1, -2, 2**62, -(2**62), 2**63 - 1, -(2**63)]),
                         rng.integers(-(2**63), 2**63 - 1, size=55, dtype=np.int64)]).astype(np.int64)
    x = torch.from_numpy(x0.copy())
    z = [init["_randn"](x).numpy() for _ in range(3)]
    out.update(randn_x0=x0, randn_z=np.stack(z), randn_x3=x.numpy())

    # A small basis: K k-points, padded length G, in-sphere counts n.
    K, G, B = 3, 37, 6
    n = np.array([37, 33, 29])
    KE = torch.from_numpy(rng.uniform(0.0, 3.0, size=(K, G)))
    basis = fake_basis(K, G, n, KE)
    out.update(basis_n=n, basis_KE=KE.numpy())

    # _randomize over band windows [b_start, b_stop) and seeds:
    for seed, b_start, b_stop in ((0, 0, B), (7, 2, 5), (3, 4, B)):
        coeff = torch.zeros((1, K, B, 1, G), dtype=torch.complex128)
        me = SimpleNamespace(basis=basis, band_division=None, coeff=coeff)
        init["_randomize"](me, seed=seed, b_start=b_start, b_stop=b_stop)
        out[f"randomize_{seed}_{b_start}_{b_stop}"] = coeff[0, :, :, 0].numpy()

    # _randomize_selected on a few (k, band) pairs, seeded by an iteration:
    i_k = torch.tensor([0, 1, 1, 2])
    i_band = torch.tensor([5, 0, 3, 2])
    i_spin = torch.zeros_like(i_k)
    coeff = torch.from_numpy(rng.standard_normal((1, K, B, 1, G)) + 0j)
    before = coeff.clone()
    me = SimpleNamespace(basis=basis, band_division=None, coeff=coeff, n_bands=lambda: B)
    init["_randomize_selected"](me, i_spin, i_k, i_band, 17)
    out.update(sel_i_k=i_k.numpy(), sel_i_band=i_band.numpy(), sel_seed=np.array(17),
               sel_before=before[0, :, :, 0].numpy(), sel_after=coeff[0, :, :, 0].numpy())

    # Davidson.precondition on residuals Cerr with per-band KEref:
    Cerr = torch.from_numpy(rng.standard_normal((1, K, B, 1, G)) + 1j * rng.standard_normal(
        (1, K, B, 1, G)))
    KEref = torch.from_numpy(rng.uniform(0.05, 2.0, size=(1, K, B)))
    KEref[0, 0, 0] = 1e-3  # x large: the exp(-x) floor is irrelevant
    KE_p = KE.clone()
    KE_p[:, :3] = 0.0  # x = 0: the floor keeps the division finite
    me = SimpleNamespace(electrons=SimpleNamespace(basis=fake_basis(K, G, n, KE_p)))
    pre = dav["precondition"](me, Cerr, KEref)
    out.update(pre_KE=KE_p.numpy(), pre_KEref=KEref[0].numpy(), pre_Cerr=Cerr[0, :, :, 0].numpy(),
               pre_out=pre[0, :, :, 0].numpy())
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}: {sorted(out)}")


if __name__ == "__main__":
    main()
