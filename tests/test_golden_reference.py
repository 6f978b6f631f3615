"""The oracle against vectors written by the reference's own functions
(tests/golden/reference_vectors.npz, made by tests/golden/
make_reference_vectors.py): Box-Muller and the xorshift states, randomize
over band windows, randomize_selected and the Davidson preconditioner.
These pin the conventions no closed form fixes: which uniform sets the
modulus, the seed layout per (k, band) of randomize_selected, the number of
discarded draws, and the exp(-x) floor of the preconditioner. Bit-level
equality is expected for the integer states; float64 results to 1e-14."""
import os
from types import SimpleNamespace

import numpy as np
import pytest

from oracle.davidson import precondition
from oracle.init import randn, randomize, randomize_selected

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors.npz")


@pytest.fixture(scope="module")
def vec():
    return np.load(PATH)


def _basis(vec, KE=None):
    n = vec["basis_n"]
    KE = vec["basis_KE"] if KE is None else KE
    G = KE.shape[1]
    return SimpleNamespace(n_max=int(n.max()), KE=KE,
                           pad_mask=np.arange(G)[None, :] >= n[:, None])


def test_randn_matches_reference(vec):
    x = vec["randn_x0"].copy()
    with np.errstate(over="ignore"):
        z = np.stack([randn(x) for _ in range(3)])
    assert np.array_equal(x, vec["randn_x3"])  # integer states bit-exact
    assert np.abs(z - vec["randn_z"]).max() < 1e-14 * np.abs(vec["randn_z"]).max()


def test_randn_swapped_uniforms_fail(vec):
    """A Box-Muller with the two uniforms exchanged (a plausible slip) does
    not reproduce the reference vectors."""
    from oracle.init import xorshift_draw
    x = vec["randn_x0"].copy()
    with np.errstate(over="ignore"):
        u0, u1 = xorshift_draw(x), xorshift_draw(x)
    z = np.sqrt(-2.0 * np.log(u1)) * np.exp(2j * np.pi * u0)
    assert np.abs(z - vec["randn_z"][0]).max() > 1e-3


@pytest.mark.parametrize("seed,b_start,b_stop", [(0, 0, 6), (7, 2, 5), (3, 4, 6)])
def test_randomize_windows_match_reference(vec, seed, b_start, b_stop):
    b = _basis(vec)
    ref = vec[f"randomize_{seed}_{b_start}_{b_stop}"]
    C = np.zeros_like(ref)
    randomize(b, C, seed=seed, b_start=b_start, b_stop=b_stop)
    assert np.abs(C - ref).max() < 1e-14 * max(np.abs(ref).max(), 1.0)


def test_randomize_selected_matches_reference(vec):
    b = _basis(vec)
    C = vec["sel_before"].copy()
    randomize_selected(b, C, vec["sel_i_k"], vec["sel_i_band"], int(vec["sel_seed"]))
    ref = vec["sel_after"]
    assert np.abs(C - ref).max() < 1e-14 * np.abs(ref).max()
    # the untouched bands keep their in-sphere values; the padding of every
    # band is zero afterwards (the reference constrains the whole array):
    touched = np.zeros(C.shape[:2], bool)
    touched[vec["sel_i_k"], vec["sel_i_band"]] = True
    inside = ~np.broadcast_to(b.pad_mask[:, None, :], C.shape)
    keep = inside & ~touched[:, :, None]
    assert np.array_equal(C[keep], vec["sel_before"][keep])
    assert not np.any(C[~inside]) and np.any(vec["sel_before"][~inside])


def test_precondition_matches_reference(vec):
    b = _basis(vec, KE=vec["pre_KE"])
    out = precondition(b, vec["pre_Cerr"], vec["pre_KEref"])
    ref = vec["pre_out"]
    assert np.all(np.isfinite(ref))
    assert np.abs(out - ref).max() < 1e-14 * np.abs(ref).max()
