"""FP64 CPU oracle for the plane-wave band solve — TEST INFRASTRUCTURE ONLY.

This package restates the reference algorithm (QimPy, dft/electrons and
math/_linalg.py) in numpy/scipy, complex128 throughout, for one spin channel
and no spinor components (S = X = 1; arrays are (K, B, G)). It is the
checker that the HIP path in ``qimpy_amd`` is compared against and the CPU
baseline that the benchmark times beside the GPU. Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import it;
the product path never does.

How each function is pinned (beyond retyping the formula):

- ``basis.Basis``: sphere counts and ``fft_index`` order against an
  independent brute-force enumeration of in-sphere plane waves, sorted by
  linear box index (tests/test_oracle.py).
- ``fft`` conventions: against a direct O(N^2) DFT on a tiny box, plus the
  Parseval and integral identities of the reference's grid/test_field.py.
- ``hamiltonian.apply_potential``: against the brute-force convolution
  sum_G' V~(G-G') psi(G'), against V = const, and Hermiticity.
- ``hamiltonian`` (full H): against an explicitly assembled dense H
  (``dense.dense_hamiltonian``), whose eigenvalues from LAPACK pin Davidson
  and CheFSI end to end.
- ``linalg.eighg``: H V = O V E, V^H O V = I, against scipy.linalg.eigh(H, O).
- ``init.randomize``: integer xorshift states bit-exact against an
  independent pure-Python big-integer implementation of the same recurrence.
- ``density.collect_density``: against the closed forms of one and two plane
  waves, the cell integral sum f w (Parseval) and a direct O(N^2) inverse
  DFT on a tiny box.
- ``davidson.davidson`` / ``chefsi.chefsi``: converged eigenvalues against
  the dense brute force to 1e-9 Ha; the Chebyshev recurrence against the
  closed-form Chebyshev polynomial on a diagonal H.
"""
