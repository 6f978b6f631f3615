"""Host logic of the batched cell solver, checked against the oracle on CPU:
sphere tables (fft_index order and G-vectors), projectors, band classes."""
import numpy as np
import pytest
import torch

from oracle.basis import Basis as OBasis
from oracle.hamiltonian import projectors as o_projectors
from qimpy_amd import systems
from qimpy_amd.batched import CellJob, choose_band_classes, sphere_tables
from qimpy_amd.ions import build_projectors


@pytest.mark.parametrize("spec", [
    systems.small_cell(),
    systems.small_cell(shape=(10, 12, 15), kmesh=(2, 1, 1), seed=5),
    systems.make_cfg2_cell(3),
], ids=["small", "small_b", "cfg2_3"])
def test_sphere_tables_match_oracle_basis(spec):
    ob = OBasis(spec)
    tables = sphere_tables(spec.Rbasis, spec.k, spec.ke_cutoff, spec.shape)
    assert len(tables) == len(spec.k)
    for ik, (lin, iG) in enumerate(tables):
        n = int(ob.n[ik])
        assert len(lin) == n
        np.testing.assert_array_equal(lin, ob.fft_index[ik, :n])
        np.testing.assert_array_equal(iG, ob.iG[ik, :n])


def test_cell_job_tables_and_guard():
    spec = systems.small_cell(seed=9)
    ob = OBasis(spec)
    job = CellJob(spec, job_id=4)
    np.testing.assert_array_equal(job.n, ob.n)
    assert job.n_tot == ob.n_tot and job.K == len(spec.k) and job.P == spec.n_projectors
    for ik in range(job.K):
        n = int(ob.n[ik])
        np.testing.assert_array_equal(job.fft_index[ik, :n], ob.fft_index[ik, :n])
        assert not job.fft_index[ik, n:].any()
    too_many = systems.small_cell(n_bands=40)
    with pytest.raises(ValueError, match="exceeds"):
        CellJob(too_many)


def test_projectors_match_oracle():
    spec = systems.small_cell(seed=21, n_atoms=3)
    ob = OBasis(spec)
    got = build_projectors(torch.tensor(spec.positions), spec.proj_rc, spec.proj_l,
                           torch.tensor(ob.iG), torch.tensor(ob.k), spec.Gbasis, spec.volume,
                           torch.tensor(ob.n)).numpy()
    ref = o_projectors(ob)
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() < 1e-15 * max(1.0, np.abs(ref).max()) + 1e-16


def test_band_classes():
    demand = [9, 10, 12, 17, 20, 35, 36, 40, 66, 71]
    for overhead in (0.0, 5.0, 12.0, 1e6):
        classes = choose_band_classes(demand, overhead_bands=overhead)
        assert classes == sorted(classes) and all(c % 8 == 0 for c in classes)
        assert max(classes) >= max(demand)
    # no overhead: every demand gets its own granule; huge overhead: one class
    assert choose_band_classes(demand, overhead_bands=0.0) == [16, 24, 40, 72]
    assert choose_band_classes(demand, overhead_bands=1e6) == [72]
    assert choose_band_classes([]) == []


def test_quintic_coeff_matches_oracle():
    """The host spline fit of the projector radials (ions.quintic_coeff)
    equals the oracle's restatement of get_coeff, batched and single, and
    refuses fewer than four samples."""
    from oracle.projectors import quintic_coeff as o_coeff
    from qimpy_amd.ions import quintic_coeff

    rng = np.random.default_rng(2)
    G = np.arange(60) * 0.02
    samples = np.stack([np.exp(-G**2 * a) * np.cos(b * G) for a, b in rng.uniform(0.5, 3, (4, 2))])
    assert np.abs(quintic_coeff(samples) - o_coeff(samples)).max() < 1e-13
    assert np.abs(quintic_coeff(samples[1]) - o_coeff(samples)[:, 1]).max() < 1e-13
    with pytest.raises(ValueError):
        quintic_coeff(samples[:, :3])
