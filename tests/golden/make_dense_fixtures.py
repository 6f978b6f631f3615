"""Write dense_cells.npz: exact band energies of cells drawn from the uneven
batched distribution, for the GPU tests of the batched solver and the
request server (tests/test_gpu_configs.py).

The fixture is data: for each selected cell its workload id and cell id
(the seeded generator qimpy_amd.systems.make_cfg2_cell rebuilds the cell),
and the lowest n_bands eigenvalues of every k-point from the oracle's dense
brute force (oracle.dense: H(G, G') assembled on the in-sphere plane waves,
then LAPACK). Cells are taken in id order among those small enough for a
dense solve (box <= MAX_BOX, at most MAX_K k-points), so the test workloads
are the leading members of the distribution the benchmark draws, filtered
by size.

Run from the repository root: python tests/golden/make_dense_fixtures.py
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np
import scipy.linalg

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.basis import Basis  # noqa: E402
from oracle.dense import dense_hamiltonian  # noqa: E402
from oracle.hamiltonian import projectors  # noqa: E402
from qimpy_amd import systems  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dense_cells.npz")
#: (workload id, number of cells, largest box, most k-points)
SELECTIONS = ((2, 12, 36, 4), (3, 20, 30, 4))


def select(cfg: int, count: int, max_box: int, max_k: int) -> list[int]:
    ids, i = [], 0
    while len(ids) < count:
        spec = systems.make_cfg2_cell(i, cfg=cfg)
        if spec.shape[0] <= max_box and len(spec.k) <= max_k:
            ids.append(i)
        i += 1
    return ids


def dense_eigs(spec) -> np.ndarray:
    ob = Basis(spec)
    beta = projectors(ob)
    out = []
    for ik in range(len(ob.k)):
        H = dense_hamiltonian(ob, ik, beta=beta)
        out.append(scipy.linalg.eigh(H, eigvals_only=True, subset_by_index=(0, spec.n_bands - 1),
                                     driver="evr"))
    return np.array(out)


def main():
    data = {}
    for cfg, count, max_box, max_k in SELECTIONS:
        ids = select(cfg, count, max_box, max_k)
        data[f"cfg{cfg}_ids"] = np.array(ids, dtype=np.int64)
        for i in ids:
            spec = systems.make_cfg2_cell(i, cfg=cfg)
            t0 = time.perf_counter()
            data[f"cfg{cfg}_{i}"] = dense_eigs(spec)
            print(f"cfg{cfg} cell {i}: box {spec.shape[0]}, {len(spec.k)} k, {spec.n_bands} bands,"
                  f" {time.perf_counter() - t0:.1f} s", flush=True)
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
